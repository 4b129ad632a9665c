"""ctypes binding of libsincdvr.so (the C-ABI declared in include/sincdvr.h).

Loading fails loudly: there is no CPU fallback for any operator.
"""
import ctypes
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libsincdvr.so")

SDV_OK = 0
ERRORS = {-1: "SDV_ERR_ARG", -2: "SDV_ERR_HIP", -3: "SDV_ERR_UNSUPPORTED", -4: "SDV_ERR_NOMEM",
          -5: "SDV_ERR_NODEVICE"}

# Every symbol include/sincdvr.h declares.
EXPORTS = [
    "sdv_config_default", "sdv_init", "sdv_free", "sdv_last_error", "sdv_set_precision",
    "sdv_op_tinv", "sdv_op_tinv_dev", "sdv_op_tinv_noalloc_dev",
    "sdv_call_twoe_matel00", "sdv_call_twoe_matel00_dev",
    "sdv_mult_ke", "sdv_mult_ke_dev",
    "sdv_mult_reducedpot", "sdv_mult_reducedpot_dev",
    "sdv_get_reducedpot", "sdv_get_reducedpot_dev",
    "sdv_twoereduced", "sdv_twoereduced_dev", "sdv_twoe_convolutions", "sdv_profile", "sdv_profile_read", "sdv_nuclear_potential",
    "sdv_get_kernel_octant", "sdv_get_ke_stencil", "sdv_circulant_edge",
    "sdv_sparseops_matel", "sdv_sparseops_matel_dev", "sdv_pot_matel", "sdv_pot_matel_dev",
    "sdv_actreduced00", "sdv_actreduced00_dev",
    "sdv_circulant_sizes", "sdv_set_circulant", "sdv_set_pair_batch",
    "sdv_batch_create", "sdv_batch_set_schedule", "sdv_batch_run", "sdv_batch_capture", "sdv_batch_stats",
    "sdv_batch_free", "sdv_batch_profile", "sdv_batch_profile_read",
    "sdv_server_create", "sdv_server_submit", "sdv_server_poll", "sdv_server_result", "sdv_server_copy_result",
    "sdv_server_release",
    "sdv_server_stats", "sdv_server_free",
]
BATCH_FUSED, BATCH_GROUPED = 0, 1
REQ_COULOMB, REQ_KE = 0, 1
STAGES = ["xfwd", "yz", "xinv", "matel_assemble", "matel_gemm", "matel_reduce", "ke"]


class SdvConfig(ctypes.Structure):
    _fields_ = [
        ("numpoints", ctypes.c_int),
        ("spacing", ctypes.c_double),
        ("toothnbig", ctypes.c_int),
        ("toothnsmall", ctypes.c_int),
        ("numspf", ctypes.c_int),
        ("cnormflag", ctypes.c_int),
        ("scalingflag", ctypes.c_int),
        ("scalingtheta", ctypes.c_double),
        ("scalingstretch", ctypes.c_double),
        ("scalingdistance", ctypes.c_double),
        ("smoothness", ctypes.c_double),
        ("notwoflag", ctypes.c_int),
        ("debugflag", ctypes.c_int),
        ("precision", ctypes.c_int),
        ("maskflag", ctypes.c_int),
        ("masknumpoints", ctypes.c_int),
        ("velflag", ctypes.c_int),
    ]


class SdvBatchSet(ctypes.Structure):
    _fields_ = [
        ("ctx", ctypes.c_void_p),
        ("numspf", ctypes.c_int),
        ("inspfs", ctypes.c_void_p),
        ("twoereduced", ctypes.c_void_p),
    ]


class SdvError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


_lib = None


def load():
    """Load the HIP library; raise if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"libsincdvr.so not built at {LIB_PATH}; run __graft_entry__.build()")
    lib = ctypes.CDLL(LIB_PATH)
    vp, i, d = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    P = ctypes.POINTER
    lib.sdv_config_default.argtypes = [P(SdvConfig)]
    lib.sdv_config_default.restype = None
    lib.sdv_init.argtypes = [P(SdvConfig), P(vp)]
    lib.sdv_free.argtypes = [vp]
    lib.sdv_last_error.restype = ctypes.c_char_p
    lib.sdv_set_precision.argtypes = [vp, i]
    lib.sdv_op_tinv.argtypes = [vp, vp, vp, i, i]
    lib.sdv_op_tinv_dev.argtypes = [vp, vp, vp, i, i, vp]
    lib.sdv_op_tinv_noalloc_dev.argtypes = [vp, vp, vp, i, i, vp]
    lib.sdv_call_twoe_matel00.argtypes = [vp, i, i, vp, vp, vp, vp]
    lib.sdv_call_twoe_matel00_dev.argtypes = [vp, i, i, vp, vp, vp, vp, vp]
    lib.sdv_mult_ke.argtypes = [vp, vp, vp, i]
    lib.sdv_mult_ke_dev.argtypes = [vp, vp, vp, i, vp]
    lib.sdv_mult_reducedpot.argtypes = [vp, i, i, vp, vp, vp]
    lib.sdv_mult_reducedpot_dev.argtypes = [vp, i, i, vp, vp, vp, vp]
    lib.sdv_get_reducedpot.argtypes = [vp, vp, vp, vp]
    lib.sdv_get_reducedpot_dev.argtypes = [vp, vp, vp, vp, vp]
    lib.sdv_twoereduced.argtypes = [vp, i, i, vp, vp, vp]
    lib.sdv_twoereduced_dev.argtypes = [vp, i, i, vp, vp, vp, vp]
    lib.sdv_twoe_convolutions.argtypes = [vp, i, i, i]
    lib.sdv_profile.argtypes = [vp, i]
    lib.sdv_nuclear_potential.argtypes = [vp, i, vp, vp, i, vp]
    lib.sdv_profile_read.argtypes = [vp, vp, vp]
    lib.sdv_get_kernel_octant.argtypes = [vp, vp]
    lib.sdv_get_ke_stencil.argtypes = [vp, vp]
    lib.sdv_circulant_edge.argtypes = [vp]
    lib.sdv_sparseops_matel.argtypes = [vp, i, i, vp, vp, vp]
    lib.sdv_sparseops_matel_dev.argtypes = [vp, i, i, vp, vp, vp, vp]
    lib.sdv_pot_matel.argtypes = [vp, i, i, vp, vp, vp, vp]
    lib.sdv_pot_matel_dev.argtypes = [vp, i, i, vp, vp, vp, vp, vp]
    lib.sdv_actreduced00.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, i, i, d, d, vp]
    lib.sdv_actreduced00_dev.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, i, i, d, d, vp, vp]
    lib.sdv_circulant_sizes.argtypes = [vp, i]
    lib.sdv_set_circulant.argtypes = [vp, i]
    lib.sdv_set_pair_batch.argtypes = [vp, i]
    lib.sdv_batch_create.argtypes = [i, P(SdvBatchSet), i, P(vp)]
    lib.sdv_batch_set_schedule.argtypes = [vp, i, i, i]
    lib.sdv_batch_run.argtypes = [vp, vp]
    lib.sdv_batch_capture.argtypes = [vp, vp]
    lib.sdv_batch_stats.argtypes = [vp, P(ctypes.c_longlong), P(ctypes.c_longlong)]
    lib.sdv_batch_free.argtypes = [vp]
    lib.sdv_batch_profile.argtypes = [vp, i]
    ll = ctypes.c_longlong
    lib.sdv_server_create.argtypes = [ll, ll, i, P(vp)]
    lib.sdv_server_submit.argtypes = [vp, vp, i, vp, P(ll)]
    lib.sdv_server_poll.argtypes = [vp, i, P(i)]
    lib.sdv_server_result.argtypes = [vp, ll, P(vp), P(i), P(d), P(d)]
    lib.sdv_server_copy_result.argtypes = [vp, ll, vp, vp]
    lib.sdv_server_release.argtypes = [vp, ll]
    lib.sdv_server_stats.argtypes = [vp, P(ll), P(ll), P(ll), P(ll), P(ll)]
    lib.sdv_server_free.argtypes = [vp]
    lib.sdv_batch_profile_read.argtypes = [vp, vp, vp]
    for name in EXPORTS:
        getattr(lib, name).restype = getattr(lib, name).restype if name in (
            "sdv_last_error", "sdv_config_default") else ctypes.c_int
    _lib = lib
    return lib


def check(rc):
    if rc != SDV_OK:
        raise SdvError(rc, load().sdv_last_error().decode())
    return rc
