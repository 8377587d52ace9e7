"""Shared helpers: rebuild golden-fixture cases for the oracle and the HIP solver."""
import json
import os

import numpy as np

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def case_names():
    return sorted(n[:-4] for n in os.listdir(GOLD)
                  if n.endswith(".npz") and not n.startswith(("mms", "cfg2")))


def load(name):
    z = np.load(os.path.join(GOLD, name + ".npz"), allow_pickle=False)
    return z, json.loads(str(z["meta"]))


def config_kwargs(m):
    nx, ny, nz = m["grid"]
    kw = dict(nx=nx, ny=ny, nz=nz, rho_l=1.0, rho_g=1.0 / 828.0, sigma=0.01424,
              xi=4.0, tau_l=0.53, tau_g=0.53, theta_eq=m["theta"])
    if m["tau_h"] is not None:
        kw["tau_h"] = m["tau_h"]
    if "g_force" in m:
        kw["g_force"] = tuple(m["g_force"]) + (0.0,) * (3 - len(m["g_force"]))
    return kw


def inject_event(m):
    """(step after which the case injects a droplet, inject_droplet args) or None."""
    inj = m.get("inject")
    if not inj:
        return None
    return inj["after"], (tuple(inj["center"]), float(inj["radius"]), float(inj["xi"]),
                          tuple(inj["u"]))


def solver_kwargs(m, z=None, device=None):
    """Solver keyword arguments of a case; a per-node amplification field
    ("field" in the metadata) comes from the fixture, as a numpy array for
    the oracle or a torch tensor on `device` for the solver."""
    kw = dict(m["kw"])
    kw.setdefault("boundary_relax", 0.3)
    if kw.get("geo_amplification") == "field":
        if z is None:
            raise ValueError("this case needs its fixture for the amplification field")
        amp = np.asarray(z["amp_field"])
        if device is not None:
            import torch
            amp = torch.from_numpy(amp).to(device)
        kw["geo_amplification"] = amp
    return kw


def rel(a, b, mask=None, floor=0.0):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    if mask is not None:
        a, b = a[..., mask], b[..., mask]
    den = max(np.linalg.norm(b), floor * np.sqrt(max(b.size, 1)), 1e-300)
    return float(np.linalg.norm(a - b) / den)
