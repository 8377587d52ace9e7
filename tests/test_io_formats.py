"""Output formats: the VTK writer against a value-by-value rendering of the
reference's file layout (io_utils/vtk_writer.py:51-86), 2D and 3D, with and
without a solid mask."""
import os

import numpy as np

from io_utils.vtk_writer import write_vtk


def loop_render(rho, u, solid, dx, step, t):
    """The file text written one value at a time, in the order and format of
    the reference's writer."""
    if u.shape[0] == 2:
        nx, ny = rho.shape
        rho = rho.reshape(nx, ny, 1)
        ux, uy = u[0].reshape(nx, ny, 1), u[1].reshape(nx, ny, 1)
        uz = np.zeros_like(ux)
        solid = solid.reshape(nx, ny, 1) if solid is not None else None
    else:
        ux, uy, uz = u
    nx, ny, nz = rho.shape
    out = ["# vtk DataFile Version 3.0\n", f"LBM output step={step} t={t:.6e}\n", "ASCII\n",
           "DATASET STRUCTURED_POINTS\n", f"DIMENSIONS {nx} {ny} {nz}\n", "ORIGIN 0 0 0\n",
           f"SPACING {dx} {dx} {dx}\n", f"POINT_DATA {nx * ny * nz}\n",
           "SCALARS density float\n", "LOOKUP_TABLE default\n"]
    out += [f"{v:.6e}\n" for v in rho.ravel()]
    out.append("VECTORS velocity float\n")
    for i in range(nx):
        for j in range(ny):
            for k in range(nz):
                out.append(f"{ux[i, j, k]:.6e} {uy[i, j, k]:.6e} {uz[i, j, k]:.6e}\n")
    if solid is not None:
        out += ["SCALARS solid int\n", "LOOKUP_TABLE default\n"]
        out += [f"{int(v)}\n" for v in solid.ravel()]
    return "".join(out)


def test_vtk_matches_reference_layout(tmp_path):
    rng = np.random.default_rng(3)
    for shape, with_solid in (((5, 4, 3), True), ((6, 5), False), ((4, 3, 2), False)):
        rho = rng.standard_normal(shape).astype(np.float32) * 1e-3 + 1.0
        u = (rng.standard_normal((len(shape),) + shape) * 1e-2).astype(np.float32)
        u.flat[0] = 0.0
        u.flat[1] = -0.0
        solid = rng.random(shape) < 0.3 if with_solid else None
        p = tmp_path / f"f{len(shape)}{int(with_solid)}.vtk"
        write_vtk(str(p), rho, u, solid, dx=1.0, step=40, t=40.0)
        assert p.read_text() == loop_render(rho, u, solid, 1.0, 40, 40.0)


def test_snapshot_keys(tmp_path):
    """save_snapshot writes the reference's sweep snapshot: file name
    <label>_step<step:05d>.npz and the keys, shapes and dtypes of
    scripts/run_thesis_sweep.py:225-241 (phi_3d, u_3d, solid_3d, the two
    mid-plane phi slices, the solid slice, step, nx, ny, nz)."""
    import torch
    from lbm.snapshots import save_snapshot

    class Fields:                       # the attributes save_snapshot reads
        pass
    nx, ny, nz = 6, 5, 4
    rng = np.random.default_rng(3)
    f = Fields()
    phi = rng.random((nx, ny, nz)).astype(np.float32)
    u = rng.standard_normal((3, nx, ny, nz)).astype(np.float32)
    solid = rng.random((nx, ny, nz)) < 0.2
    f.phi, f.u, f.solid = torch.from_numpy(phi), torch.from_numpy(u), torch.from_numpy(solid)
    path = save_snapshot(f, str(tmp_path / "snaps"), "ridge_th162", 250)
    assert os.path.basename(path) == "ridge_th162_step00250.npz"
    z = np.load(path)
    assert sorted(z.files) == sorted(["phi_3d", "u_3d", "solid_3d", "xz_slice", "yz_slice",
                                      "xz_solid", "step", "nx", "ny", "nz"])
    np.testing.assert_array_equal(z["phi_3d"], phi)
    np.testing.assert_array_equal(z["u_3d"], u)
    np.testing.assert_array_equal(z["solid_3d"], solid)
    assert z["phi_3d"].dtype == np.float32 and z["u_3d"].dtype == np.float32
    assert z["solid_3d"].dtype == np.bool_
    np.testing.assert_array_equal(z["xz_slice"], phi[:, ny // 2, :])
    np.testing.assert_array_equal(z["yz_slice"], phi[nx // 2, :, :])
    np.testing.assert_array_equal(z["xz_solid"], solid[:, ny // 2, :])
    assert (int(z["step"]), int(z["nx"]), int(z["ny"]), int(z["nz"])) == (250, nx, ny, nz)
