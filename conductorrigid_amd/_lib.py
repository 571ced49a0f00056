"""ctypes binding of the HIP library libconductorrigid.so (C ABI in include/conductorrigid.h).

The library is loaded from this package directory only; there is no fallback:
if it is missing or fails to load, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libconductorrigid.so")

CR_OK = 0
CR_ERR_INVALID = 1
CR_ERR_DEVICE = 2
CR_ERR_NOT_CONVERGED = 3
CR_ERR_STATE = 4
CR_PRECISION_FP64 = 0
CR_PRECISION_FP32_BULK = 1
CR_X0_PREVIOUS = 0
CR_X0_ZERO = 1
CR_X0_CHARGES = 2
CR_MLP_IN = 176
CR_MLP_TRUNK = 168
CR_MLP_HIDDEN = 128
CR_MLP_OUT = 4
CR_MLP_PARAMS = 39684
CR_PRECOND_DIAGONAL = 0
CR_PRECOND_NEARFIELD_BLOCK = 1
CR_NUM_STAGES = 10
STAGE_NAMES = ["mesh_clear", "spread", "fft_z_r2c", "fft_y_fwd", "fft_x_scale", "fft_y_inv", "fft_z_c2r",
               "contract", "real_space", "constraint_rows"]

vp = C.c_void_p


class CrParams(C.Structure):
    _fields_ = [("L", C.c_double * 3), ("xi", C.c_double), ("errortol", C.c_double), ("rc", C.c_double),
                ("mesh", C.c_int * 3), ("P", C.c_int), ("precision", C.c_int), ("restart", C.c_int),
                ("max_iter", C.c_int), ("gmres_tol", C.c_double), ("preconditioner", C.c_int)]


class CrResolved(C.Structure):
    _fields_ = [("rc", C.c_double), ("kcut", C.c_double), ("mesh", C.c_int * 3), ("P", C.c_int),
                ("h", C.c_double * 3), ("eta", C.c_double * 3), ("N_table", C.c_int), ("drtable", C.c_double),
                ("selfterm", C.c_double), ("prefac", C.c_double), ("xiterm", C.c_double), ("quadW", C.c_double),
                ("precision", C.c_int), ("restart", C.c_int), ("max_iter", C.c_int), ("gmres_tol", C.c_double)]

    def as_dict(self) -> dict:
        out = {}
        for k, _ in self._fields_:
            v = getattr(self, k)
            out[k] = list(v) if isinstance(v, C.Array) else v
        return out


class CrSystem(C.Structure):
    _fields_ = [("N", C.c_int), ("pos", vp), ("tag", vp), ("body_of_tag", vp), ("body_charge", vp),
                ("N_rigid", C.c_int), ("N_isolated", C.c_int), ("Nmax", C.c_int), ("rigid_membership", vp),
                ("rigid_members", vp), ("rigid_size", vp), ("rigid_com", vp), ("nlist", vp), ("head_list", vp),
                ("n_neigh", vp), ("L", C.c_double * 3)]


class CrSystemF32(C.Structure):
    _fields_ = CrSystem._fields_


class CrTuning(C.Structure):
    _fields_ = [("tile", C.c_int * 3), ("lds_bytes", C.c_int), ("spread_contract_ms", C.c_double),
                ("candidates", C.c_int)]


class CrSolveStats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("matvecs", C.c_int), ("restarts", C.c_int), ("converged", C.c_int),
                ("resid", C.c_double), ("tolerance", C.c_double), ("true_resid", C.c_double),
                ("true_relres", C.c_double), ("fallback", C.c_int)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


# every symbol the header declares, with (restype, argtypes)
_SIGS = {
    "cr_create": (C.c_int, [C.POINTER(vp), C.c_int]),
    "cr_destroy": (None, [vp]),
    "cr_last_error": (C.c_char_p, []),
    "cr_set_params": (C.c_int, [vp, C.POINTER(CrParams), C.c_int, C.c_int]),
    "cr_get_resolved": (C.c_int, [vp, C.POINTER(CrResolved)]),
    "cr_update_field": (C.c_int, [vp, C.POINTER(C.c_double)]),
    "cr_set_box": (C.c_int, [vp, C.POINTER(C.c_double), vp]),
    "cr_compute_force_f32": (C.c_int, [vp, C.POINTER(CrSystemF32), vp, vp, vp, C.c_int, C.POINTER(CrSolveStats),
                                       vp]),
    "cr_zero_force": (C.c_int, [vp, C.c_int, vp, vp]),
    "cr_compute_force": (C.c_int, [vp, C.POINTER(CrSystem), vp, vp, vp, C.c_int, C.POINTER(CrSolveStats), vp]),
    "cr_matvec": (C.c_int, [vp, C.POINTER(CrSystem), vp, vp, vp]),
    "cr_precondition": (C.c_int, [vp, C.POINTER(CrSystem), vp, vp, vp]),
    "cr_objective": (C.c_int, [vp, C.POINTER(CrSystem), vp, vp]),
    "cr_spread": (C.c_int, [vp, C.POINTER(CrSystem), vp, vp, vp]),
    "cr_recip_grid": (C.c_int, [vp, vp, vp, vp]),
    "cr_fft_c2c": (C.c_int, [vp, vp, vp, C.c_int, vp]),
    "cr_contract": (C.c_int, [vp, C.POINTER(CrSystem), vp, vp]),
    "cr_contract_force": (C.c_int, [vp, C.POINTER(CrSystem), vp, vp, vp]),
    "cr_real_space": (C.c_int, [vp, C.POINTER(CrSystem), vp, vp, vp]),
    "cr_real_space_force": (C.c_int, [vp, C.POINTER(CrSystem), vp, vp, vp]),
    "cr_profile_matvec": (C.c_int, [vp, C.POINTER(CrSystem), vp, C.c_int, C.POINTER(C.c_double), C.c_int, vp]),
    "cr_bd_step": (C.c_int, [vp, C.POINTER(CrSystem), vp, vp, vp, vp, vp, C.c_double, C.c_double, C.c_double,
                             C.c_double, C.c_double, vp]),
    "cr_autotune": (C.c_int, [vp, C.POINTER(CrSystem), C.POINTER(CrTuning), vp]),
    "cr_set_tiles": (C.c_int, [vp, C.POINTER(C.c_int), C.c_int]),
    "cr_set_residual_check": (C.c_int, [vp, C.c_int]),
    "cr_mlp_create": (C.c_int, [C.POINTER(vp), C.c_int, C.c_ulonglong]),
    "cr_mlp_destroy": (None, [vp]),
    "cr_mlp_set_params": (C.c_int, [vp, vp]),
    "cr_mlp_get_params": (C.c_int, [vp, vp]),
    "cr_mlp_features": (C.c_int, [vp, C.POINTER(CrSystem), vp, vp]),
    "cr_mlp_targets": (C.c_int, [vp, C.POINTER(CrSystem), vp, vp, vp]),
    "cr_mlp_charges": (C.c_int, [vp, C.POINTER(CrSystem), vp, vp, vp]),
    "cr_mlp_predict": (C.c_int, [vp, C.c_int, vp, vp, vp]),
    "cr_mlp_gradient": (C.c_int, [vp, C.c_int, vp, vp, vp, vp, vp]),
    "cr_mlp_train_step": (C.c_int, [vp, C.c_int, vp, vp, C.c_float, vp, vp]),
    "cr_build_nlist": (C.c_int, [vp, C.c_int, vp, C.POINTER(C.c_double), C.c_double, vp, vp, vp, C.c_long,
                                 C.POINTER(C.c_long), vp]),
    "cr_build_nlist_noalloc": (C.c_int, [vp, C.c_int, vp, C.POINTER(C.c_double), C.c_double, vp, vp, vp, C.c_long,
                                         C.POINTER(C.c_long), vp]),
    "cr_batch_create": (C.c_int, [C.POINTER(vp), C.c_int]),
    "cr_batch_destroy": (None, [vp]),
    "cr_batch_compute_force": (C.c_int, [vp, C.c_int, C.POINTER(vp), C.POINTER(CrSystem), C.POINTER(vp),
                                         C.POINTER(vp), C.POINTER(vp), C.c_int, C.c_int, C.POINTER(CrSolveStats),
                                         vp]),
}

_lib = None


class CrError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"conductorrigid error {code}: {msg}")
        self.code = code


def load():
    """Load libconductorrigid.so (raises if it is missing: there is no fallback path)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run conductorrigid_amd.build.build() first")
        lib = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _lib = lib
    return _lib


def check(code: int) -> None:
    if code != CR_OK:
        raise CrError(code, load().cr_last_error().decode())


def symbols() -> list[str]:
    return list(_SIGS)
