"""Sweep snapshots in the reference's format (scripts/run_thesis_sweep.py:
218-241): a compressed .npz per (label, step) with the full 3D fields and the
two mid-plane phi slices."""
import os

import numpy as np


def snapshot_arrays(solver, step):
    """The snapshot dictionary of a 3D solver at `step`."""
    phi = solver.phi.detach().cpu().numpy()
    solid = solver.solid.cpu().numpy()
    u = solver.u.detach().cpu().numpy()
    nx, ny, nz = phi.shape
    return dict(phi_3d=phi, u_3d=u, solid_3d=solid,
                xz_slice=phi[:, ny // 2, :], yz_slice=phi[nx // 2, :, :],
                xz_solid=solid[:, ny // 2, :], step=step, nx=nx, ny=ny, nz=nz)


def save_snapshot(solver, snapshot_dir, label, step):
    """Write <snapshot_dir>/<label>_step<step:05d>.npz; returns the path."""
    os.makedirs(snapshot_dir, exist_ok=True)
    path = os.path.join(snapshot_dir, f"{label}_step{step:05d}.npz")
    np.savez_compressed(path, **snapshot_arrays(solver, step))
    return path
