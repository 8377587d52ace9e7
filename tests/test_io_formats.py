"""Output formats: the VTK writer against a value-by-value rendering of the
reference's file layout (io_utils/vtk_writer.py:51-86), 2D and 3D, with and
without a solid mask."""
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
    from lbm.snapshots import snapshot_arrays  # noqa: F401  (import check)
    import importlib
    m = importlib.import_module("lbm.snapshots")
    assert hasattr(m, "save_snapshot")
