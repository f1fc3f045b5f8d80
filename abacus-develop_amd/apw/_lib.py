"""ctypes binding of libapw.so (the C ABI declared in include/apw.h).

The library is the HIP implementation; nothing here computes.  Loading fails loudly when
the shared library is missing (build it with ``python abacus-develop_amd/build.py``).
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(HERE), "lib", "libapw.so")

APW_OK = 0
APW_FP64 = 0
APW_FP32_MIXED = 1
APW_FP32_MIXED_X32 = 2

STATUS = {-1: "APW_EINVAL", -2: "APW_EHIP", -3: "APW_ENOMEM", -4: "APW_EBREAKDOWN", -5: "APW_EEIG",
          -6: "APW_ECALLBACK"}


class ApwError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


class HamDesc(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("npw", C.c_int), ("npwx", C.c_int),
                ("ig2ixyz", C.POINTER(C.c_int)), ("gk2", C.POINTER(C.c_double)), ("tpiba2", C.c_double),
                ("veff", C.POINTER(C.c_double)), ("nkb", C.c_int), ("vkb", C.POINTER(C.c_double)),
                ("ntype", C.c_int), ("na", C.POINTER(C.c_int)), ("nh", C.POINTER(C.c_int)), ("nhm", C.c_int),
                ("deeq", C.POINTER(C.c_double)), ("precision", C.c_int)]


class DiagStats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("ntry", C.c_int), ("notconv", C.c_int), ("hpsi_calls", C.c_int),
                ("hpsi_bands", C.c_longlong), ("fp64_switch_iter", C.c_int), ("max_residual", C.c_double),
                ("t_hpsi_s", C.c_double), ("t_total_s", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


HPSI_CB = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int)
SPSI_CB = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int)
HPSI_CB_C64 = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int)

_vp, _i, _d, _ch = C.c_void_p, C.c_int, C.c_double, C.c_char
PROTOTYPES = {
    "apw_last_error": (C.c_char_p, []),
    "apw_ctx_create": (_i, [C.POINTER(_vp), _i, _vp]),
    "apw_ctx_destroy": (_i, [_vp]),
    "apw_ctx_synchronize": (_i, [_vp]),
    "apw_ctx_stream": (_vp, [_vp]),
    "apw_ctx_set_profiling": (_i, [_vp, _i]),
    "apw_ctx_set_eig_single_wg": (_i, [_vp, _i]),
    "apw_ctx_profile_count": (_i, [_vp]),
    "apw_ctx_profile_entry": (_i, [_vp, _i, C.c_char_p, _i, C.POINTER(C.c_longlong), C.POINTER(_d), C.POINTER(_d),
                                   C.POINTER(_d)]),
    "apw_ham_create": (_i, [_vp, C.POINTER(HamDesc), C.POINTER(_vp)]),
    "apw_ham_destroy": (_i, [_vp]),
    "apw_ham_set_precision": (_i, [_vp, _i]),
    "apw_hpsi": (_i, [_vp, _vp, _i, _vp, _i, _i]),
    "apw_act_ekinetic": (_i, [_vp, _i, _i, _vp, _vp]),
    "apw_act_veff": (_i, [_vp, _i, _i, _vp, _vp]),
    "apw_act_nonlocal": (_i, [_vp, _i, _i, _vp, _vp]),
    "apw_recip_to_real": (_i, [_vp, _vp, _vp, _i, _d]),
    "apw_real_to_recip": (_i, [_vp, _vp, _vp, _i, _d]),
    "apw_david_create": (_i, [_vp, C.POINTER(_d), _i, _i, _i, C.POINTER(_vp)]),
    "apw_david_destroy": (_i, [_vp]),
    "apw_david_diag": (_i, [_vp, _vp, HPSI_CB, SPSI_CB, _vp, _i, _vp, C.POINTER(_d), _d, _i, _i, _i, _i,
                            C.POINTER(DiagStats)]),
    "apw_david_diag_c64": (_i, [_vp, HPSI_CB_C64, _vp, _i, _vp, C.POINTER(_d), _d, _i, _i, _i,
                                C.POINTER(DiagStats)]),
    "apw_david_diag_batched": (_i, [_i, C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_i), C.POINTER(_vp),
                                    C.POINTER(_vp), _d, _i, _i, _i, _i, C.POINTER(DiagStats), C.POINTER(_i)]),
    "apw_ekinetic_pw_op": (_i, [_vp, _i, _i, _i, _d, _vp, _vp, _vp]),
    "apw_veff_pw_op": (_i, [_vp, _i, _vp, _vp]),
    "apw_nonlocal_pw_op": (_i, [_vp, _i, _i, _i, C.POINTER(_i), C.POINTER(_i), _i, _i, _i, _i, _i, _vp, _vp, _vp]),
    "apw_set_3d_fft_box_op": (_i, [_vp, _i, _vp, _vp, _vp]),
    "apw_set_recip_to_real_output_op": (_i, [_vp, _i, _i, _d, _vp, _vp]),
    "apw_set_real_to_recip_output_op": (_i, [_vp, _i, _i, _i, _d, _vp, _vp, _vp]),
    "apw_fft3d": (_i, [_vp, _i, _i, _i, _vp, _i]),
    "apw_zgemm": (_i, [_vp, _ch, _ch, _i, _i, _i, C.POINTER(_d), _vp, _i, _vp, _i, C.POINTER(_d), _vp, _i]),
    "apw_ham_set_vkb": (_i, [_vp, _vp, _i]),
    "apw_cal_ylm_real_op": (_i, [_vp, _i, _i, _d, _d, _d, _d, _d, _vp, _vp, _vp]),
    "apw_cal_sk_op": (_i, [_vp] + [_i] * 14 + [_vp, _vp, _vp, _d, _vp, _vp, _vp, _vp, _vp, _vp]),
    "apw_cal_vnl_op": (_i, [_vp, _i, _i, _i, _i, _i, _i, _vp, _vp, _vp, _d, _d, _d, _d] + [_vp] * 9),
    "apw_getvnl": (_i, [_vp, _i, _i, _i, _i, _i, _i, _vp, _vp, _vp, _d, _d] + [_vp] * 7),
    "apw_write_wfc_pw": (_i, [_vp, C.c_char_p, _i, _i, _vp, _d, _i, _d, _d, _d, _vp, _i, _i, _i, _i, _vp, _vp, _i]),
    "apw_read_wfc_pw": (_i, [_vp, C.c_char_p, _i, _i, _i, _vp, _d, _d, _vp, _i, _i, _i, _i, _vp, _vp, _i]),
    "apw_dngvx_op": (_i, [_vp, _i, _i, _vp, _vp, _i, _vp, _vp]),
    "apw_dngvd_op": (_i, [_vp, _i, _i, _vp, _vp, _vp, _vp]),
    "apw_dav_subspace_create": (_i, [_vp, C.POINTER(_d), _i, _i, _i, _d, _i, _i, C.POINTER(_vp)]),
    "apw_dav_subspace_destroy": (_i, [_vp]),
    "apw_dav_subspace_diag": (_i, [_vp, _vp, HPSI_CB, _vp, _i, _vp, C.POINTER(_d), C.POINTER(_i), _i,
                                   C.POINTER(DiagStats)]),
    "apw_psi_to_rho": (_i, [_vp, _vp, _i, _i, C.POINTER(_d), _d, _vp, _i]),
    "apw_elecstate_pw_op": (_i, [_vp, _i, _i, _d, C.POINTER(_vp), _vp]),
    "apw_elecstate_pw_op_nc": (_i, [_vp, _i, _i, _i, _d, C.POINTER(_vp), _vp, _vp]),
    "apw_zgemm_x32": (_i, [_vp, _ch, _ch, _i, _i, _i, C.POINTER(_d), _vp, _i, _vp, _i, C.POINTER(_d), _vp, _i]),
    "apw_cgemm": (_i, [_vp, _ch, _ch, _i, _i, _i, C.POINTER(C.c_float), _vp, _i, _vp, _i, C.POINTER(C.c_float), _vp, _i]),
    "apw_dnevx_op": (_i, [_vp, _i, _i, _vp, _i, _vp, _vp]),
    "apw_probe_mfma_layout": (_i, [_vp, _vp]),
    "apw_tune_set_mode": (_i, [_i]),
    "apw_tune_load": (_i, [C.c_char_p]),
    "apw_tune_save": (_i, [C.c_char_p]),
    "apw_tune_clear": (_i, []),
    "apw_tune_stats": (_i, [C.POINTER(C.c_longlong)] * 4),
    "apw_tune_entry": (_i, [_i, C.c_char_p, _i, C.POINTER(_i), C.POINTER(_i), C.POINTER(_d)]),
    "apw_zgemm_noalloc": (_i, [_vp, _ch, _ch, _i, _i, _i, C.POINTER(_d), _vp, _i, _vp, _i, C.POINTER(_d), _vp, _i]),
}

_LIB = None


def load(path: str = LIB_PATH) -> C.CDLL:
    """Load libapw.so (once) and declare the prototypes."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} not built: run `python abacus-develop_amd/build.py`")
    lib = C.CDLL(path)
    for name, (res, args) in PROTOTYPES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = lib
    return lib


def check(rc: int) -> int:
    if rc < 0:
        raise ApwError(rc, load().apw_last_error().decode(errors="replace"))
    return rc
