"""ctypes binding of liblbm_amd.so (the C ABI in include/lbm_amd.h).

The library is built in-tree by ``python pytorch-lbm_amd/build.py`` (or
``__graft_entry__.build()``).  torch is imported first so that the HIP runtime
torch ships is the one the library binds to (same SONAME), which lets device
pointers and streams pass between the two.  There is no fallback: if the
library is missing the import fails.
"""
import ctypes
import os

import torch  # noqa: F401  (loads the HIP runtime the library links against)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblbm_amd.so")
# an alternative build of the same library (e.g. one compiled with other
# tuning macros) can be selected for measurements
LIB_PATH = os.environ.get("LBM_AMD_LIB", LIB_PATH)

if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"{LIB_PATH} is missing: build it with `python pytorch-lbm_amd/build.py` "
        "(hipcc, gfx950). The solver has no CPU fallback.")

lib = ctypes.CDLL(LIB_PATH)

LBM_F32, LBM_F64, LBM_BF16, LBM_F16 = 0, 1, 2, 3
FRAC_NONE, FRAC_F32, FRAC_F64 = 0, 1, 2

FIELD = dict(phi=0, rho=1, u=2, P=3, mu=4, f=5, g=6, phi_raw=7,
             is_boundary=16, near_wall=17, wall_weight=18, wall_normal=19,
             partial=20, solid=21)


class LbmParams(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("precision", ctypes.c_int),
        ("rho_l", ctypes.c_double), ("rho_g", ctypes.c_double),
        ("sigma", ctypes.c_double), ("xi", ctypes.c_double),
        ("beta", ctypes.c_double), ("kappa", ctypes.c_double),
        ("tau_l", ctypes.c_double), ("tau_g", ctypes.c_double),
        ("tau_h", ctypes.c_double),
        ("theta_eq", ctypes.c_double), ("phi_c", ctypes.c_double),
        ("g_force", ctypes.c_double * 3),
        ("stab_mode", ctypes.c_int),
        ("boundary_relax", ctypes.c_double),
        ("sharpen_reduce", ctypes.c_double),
        ("mu_wall_scale", ctypes.c_double),
        ("ac_scale", ctypes.c_double),
        ("ghost_scale", ctypes.c_double),
        ("geometric_wetting", ctypes.c_int),
        ("geo_amplification", ctypes.c_double),
        ("direct_wetting", ctypes.c_int),
        ("interface_mode", ctypes.c_int),
        ("mu_drive", ctypes.c_double),
    ]


class LbmCaseSpec(ctypes.Structure):
    _fields_ = [
        ("p", LbmParams),
        ("substrate", ctypes.c_int),
        ("R_star", ctypes.c_double), ("R_d", ctypes.c_double),
        ("center", ctypes.c_double * 3),
        ("radius", ctypes.c_double),
        ("u_impact", ctypes.c_double * 3),
        ("steps", ctypes.c_longlong),
    ]


_vp = ctypes.c_void_p
_SIGS = {
    "lbm_params_default": (None, [ctypes.POINTER(LbmParams), ctypes.c_int, ctypes.c_int,
                                  ctypes.c_int]),
    "lbm_create": (ctypes.c_int, [ctypes.POINTER(LbmParams), ctypes.POINTER(_vp)]),
    "lbm_destroy": (ctypes.c_int, [_vp]),
    "lbm_set_geometry": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int, _vp]),
    "lbm_set_amplification_field": (ctypes.c_int, [_vp, _vp, _vp]),
    "lbm_init_fields": (ctypes.c_int, [_vp, _vp, _vp, _vp]),
    "lbm_step": (ctypes.c_int, [_vp, ctypes.c_int, _vp]),
    "lbm_get_field": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp]),
    "lbm_field_bytes": (ctypes.c_size_t, [_vp, ctypes.c_int]),
    "lbm_grad_lap": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp]),
    "lbm_grad_lap_noalloc": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp]),
    "lbm_get_mass": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_double),
                                    ctypes.POINTER(ctypes.c_double), _vp]),
    "lbm_set_mass_target": (ctypes.c_int, [_vp, ctypes.c_double]),
    "lbm_step_count": (ctypes.c_longlong, [_vp]),
    "lbm_inject_droplet": (ctypes.c_int, [_vp, _vp, ctypes.POINTER(ctypes.c_double), _vp]),
    "lbm_measure": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_double), _vp]),
    "lbm_measure_async": (ctypes.c_int, [_vp, _vp, _vp]),
    "lbm_contact_radius": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_double,
                                          ctypes.POINTER(ctypes.c_double), _vp]),
    "lbm_base_radius": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_double,
                                       ctypes.POINTER(ctypes.c_double), _vp]),
    "lbm_last_error": (ctypes.c_char_p, []),
    "lbm_graph_capture": (ctypes.c_int, [_vp, ctypes.c_int]),
    "lbm_check_division": (ctypes.c_int, [ctypes.c_float, ctypes.POINTER(ctypes.c_ulonglong),
                                          _vp]),
    "lbm_graph_launch": (ctypes.c_int, [_vp, _vp]),
    "lbm_hbm_stream": (ctypes.c_int, [ctypes.c_size_t, ctypes.c_int, _vp,
                                      ctypes.POINTER(ctypes.c_double)]),
    "lbm_set_block_sizes": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int]),
    "lbm_set_occupancy": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int]),
    "lbm_get_launch": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int)]),
    "lbm_set_split": (ctypes.c_int, [_vp, ctypes.c_int]),
    "lbm_set_fused_tile": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "lbm_get_fused_tile": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int)]),
    "lbm_state_bytes": (ctypes.c_size_t, [_vp]),
    "lbm_save_state": (ctypes.c_int, [_vp, _vp, _vp]),
    "lbm_load_state": (ctypes.c_int, [_vp, _vp, _vp]),
    "lbm_case_footprint": (ctypes.c_size_t, [ctypes.POINTER(LbmParams)]),
    "lbm_batch_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_size_t, ctypes.POINTER(_vp)]),
    "lbm_batch_destroy": (ctypes.c_int, [_vp]),
    "lbm_batch_add": (ctypes.c_int, [_vp, ctypes.POINTER(LbmCaseSpec), _vp,
                                     ctypes.POINTER(ctypes.c_int)]),
    "lbm_batch_add_fields": (ctypes.c_int, [_vp, ctypes.POINTER(LbmCaseSpec), _vp, _vp,
                                            ctypes.c_int, _vp, _vp, _vp,
                                            ctypes.POINTER(ctypes.c_int)]),
    "lbm_batch_remove": (ctypes.c_int, [_vp, ctypes.c_int]),
    "lbm_batch_solver": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(_vp)]),
    "lbm_batch_load_state": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp]),
    "lbm_batch_status": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_longlong),
                                        ctypes.POINTER(ctypes.c_longlong)]),
    "lbm_batch_bytes_used": (ctypes.c_size_t, [_vp]),
    "lbm_batch_step": (ctypes.c_int, [_vp, ctypes.c_int, _vp]),
    "lbm_batch_set_block_sizes": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int]),
    "lbm_batch_set_split": (ctypes.c_int, [_vp, ctypes.c_int]),
    "lbm_batch_set_occupancy": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int]),
    "lbm_batch_get_launch": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int)]),
    "lbm_batch_autotune": (ctypes.c_int, [_vp, ctypes.c_int, _vp, ctypes.POINTER(ctypes.c_int),
                                          ctypes.POINTER(ctypes.c_int)]),
    "lbm_batch_graph_capture": (ctypes.c_int, [_vp, ctypes.c_int, _vp]),
    "lbm_batch_graph_launch": (ctypes.c_int, [_vp, _vp]),
    "lbm_autotune": (ctypes.c_int, [_vp, ctypes.c_int, _vp, ctypes.POINTER(ctypes.c_int),
                                    ctypes.POINTER(ctypes.c_int)]),
    "lbm_profile_steps": (ctypes.c_int, [_vp, ctypes.c_int, _vp,
                                         ctypes.POINTER(ctypes.c_double)]),
}
for _name, (_res, _args) in _SIGS.items():
    _fn = getattr(lib, _name)
    _fn.restype = _res
    _fn.argtypes = _args


class LbmError(RuntimeError):
    pass


def check(rc, what):
    if rc != 0:
        msg = lib.lbm_last_error().decode(errors="replace")
        raise LbmError(f"{what} failed ({rc}): {msg}")


def stream_ptr(device=None):
    """The current torch stream on `device` as a hipStream_t."""
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)
