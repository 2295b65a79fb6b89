"""ctypes binding of libth2d.so (the C ABI declared in include/th2d.h).

The shared library is built in-tree by ``__graft_entry__.build()`` (or
``python -m thirring2d_amd.build``).  Loading fails loudly when it is
missing: there is no CPU fallback for any entry point.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libth2d.so")

D = C.c_double
I = C.c_int
L = C.c_long
P = C.c_void_p
U64 = C.c_ulonglong
F = C.c_float

TH2D_OK = 0
TH2D_EARG = -1
TH2D_EHIP = -2
TH2D_ENOCONV = -3
TH2D_EBREAKDOWN = -4
TH2D_ESINGULAR = -5
TH2D_ENOMEM = -6
TH2D_ELAUNCH = -7
TH2D_ESYNC = -8

PREC_F64 = 0
PREC_F32 = 1
PREC_MIXED = 2


class Geom(C.Structure):
    _fields_ = [("nt", I), ("nx", I), ("anti_t", I), ("anti_x", I), ("m", D), ("mu", D)]


class CgParams(C.Structure):
    _fields_ = [("tol_rel", D), ("tol_abs_rr", D), ("max_iter", I), ("precision", I),
                ("delta", D)]


class CgInfo(C.Structure):
    _fields_ = [("iterations", I), ("restarts", I), ("status", I), ("true_rel_resid", D),
                ("rr", D)]


class HmcParams(C.Structure):
    _fields_ = [("g", D), ("nf", I), ("nsteps", I), ("traj_len", D), ("phys_norm", I),
                ("cg", CgParams)]


class HmcResult(C.Structure):
    _fields_ = [("dS", D), ("acts", D * 8), ("accepted", I), ("cg_iterations", I),
                ("cg_solves", I), ("max_true_rel_resid", D), ("status", I)]


class BagGeom(C.Structure):
    _fields_ = [("nt", I), ("nx", I), ("bc", I), ("with_mass_monomers", I), ("m", D),
                ("mu", D)]


class BagMove(C.Structure):
    _fields_ = [("no", I), ("nr", I), ("occ", I * 2), ("fr", I * 2), ("vals", I * 2),
                ("accepted", I), ("pad", I), ("ratio", D)]


class Batch(C.Structure):
    _fields_ = [("count", I), ("max_n", I), ("off", P), ("n", P), ("poff", P)]


class GemmJob(C.Structure):
    _fields_ = [("a", L), ("b", L), ("c", L), ("lda", I), ("ldb", I), ("ldc", I),
                ("M", I), ("N", I), ("K", I)]


# name -> (restype, argtypes); every symbol declared in include/*.h
SIGNATURES = {
    # th2d.h
    "th2d_ctx_create": (I, [C.POINTER(P), P]),
    "th2d_ctx_destroy": (I, [P]),
    "th2d_strerror": (C.c_char_p, [I]),
    "th2d_device_count": (I, [C.POINTER(I)]),
    "th2d_device_props": (I, [C.POINTER(I), C.POINTER(I), C.POINTER(C.c_size_t), C.c_char_p, I]),
    "th2d_malloc": (I, [C.POINTER(P), C.c_size_t]),
    "th2d_free": (I, [P]),
    "th2d_memset": (I, [P, P, I, C.c_size_t]),
    "th2d_h2d": (I, [P, P, P, C.c_size_t]),
    "th2d_d2h": (I, [P, P, P, C.c_size_t]),
    "th2d_d2d": (I, [P, P, P, C.c_size_t]),
    "th2d_sync": (I, [P]),
    "th2d_timer_start": (I, [P]),
    "th2d_timer_stop": (I, [P, C.POINTER(F)]),
    "th2d_kernel_timing": (I, [P, I]),
    "th2d_kernel_time": (I, [P, C.POINTER(D), C.POINTER(L), C.POINTER(D)]),
    "th2d_vec_to_eo": (I, [P, C.POINTER(Geom), P, P]),
    "th2d_vec_from_eo": (I, [P, C.POINTER(Geom), P, P]),
    "th2d_links_to_eo": (I, [P, C.POINTER(Geom), P, P]),
    "th2d_links_from_eo": (I, [P, C.POINTER(Geom), P, P]),
    "th2d_vec_f64_to_f32": (I, [P, P, P, L]),
    "th2d_vec_f32_to_f64": (I, [P, P, P, L]),
    "th2d_fold_links": (I, [P, C.POINTER(Geom), P, P, P]),
    "th2d_gauge_step": (I, [P, C.POINTER(Geom), P, P, D, P, P]),
    "th2d_dslash": (I, [P, C.POINTER(Geom), P, P, P, I, I, I]),
    "th2d_dslash_batched": (I, [P, I, C.POINTER(Geom), P, P, P, I]),
    "th2d_mdagm": (I, [P, C.POINTER(Geom), P, P, P, P]),
    "th2d_cg_mdagm": (I, [P, C.POINTER(Geom), P, P, P, P, C.POINTER(CgParams),
                          C.POINTER(CgInfo)]),
    "th2d_force": (I, [P, C.POINTER(Geom), P, P, P, P, P, D, D, I, P]),
    "th2d_dot": (I, [P, L, P, P, C.POINTER(D), C.POINTER(D)]),
    "th2d_norm2": (I, [P, L, P, C.POINTER(D)]),
    "th2d_gauge_action": (I, [P, C.POINTER(Geom), P, D, I, C.POINTER(D)]),
    "th2d_hmc_trajectory": (I, [P, C.POINTER(Geom), C.POINTER(HmcParams), P, P, P, P, D, U64,
                                U64, C.POINTER(HmcResult)]),
    "th2d_philox4x32_10": (I, [P, L, P, P]),
    "th2d_set_hmc_graph": (I, [P, I]),
    "th2d_random_links": (I, [P, C.POINTER(Geom), P, U64]),
    "th2d_puregauge_sweep": (I, [P, C.POINTER(Geom), P, P, D, I]),
    "th2d_puregauge_sweeps": (I, [P, C.POINTER(Geom), P, D, I, I, U64, U64]),
    "th2d_set_cg_tile": (I, [P, I, I]),
    "th2d_cg_last_tile": (I, [P, C.POINTER(I), C.POINTER(I)]),
    "th2d_set_cg_autotune": (I, [P, I]),
    "th2d_set_cg_resident": (I, [P, I]),
    "th2d_set_cg_graph": (I, [P, I]),
    "th2d_set_cg_reliable": (I, [P, I]),
    "th2d_cg_last_resident": (I, [P, C.POINTER(I)]),
    "th2d_set_cg_resident_rows": (I, [P, I]),
    "th2d_set_cg_spin_limit": (I, [P, L]),
    "th2d_set_cg_small": (I, [P, I]),
    "th2d_cg_sync_fallbacks": (I, [P, C.POINTER(L)]),
    "th2d_set_dslash_variant": (I, [I]),
    "th2d_set_cg_max_blocks": (I, [P, I]),
    "th2d_set_cg_chunk": (I, [P, I]),
    "th2d_cg_mdagm_batched": (I, [P, I, C.POINTER(Geom), C.POINTER(P), C.POINTER(P),
                                  C.POINTER(P), C.POINTER(P), C.POINTER(CgParams),
                                  C.POINTER(CgInfo)]),
    "th2d_hmc_trajectory_batched": (I, [P, I, C.POINTER(Geom), C.POINTER(HmcParams),
                                        C.POINTER(P), C.POINTER(P), C.POINTER(P), C.POINTER(P),
                                        C.POINTER(D), U64, C.POINTER(U64),
                                        C.POINTER(HmcResult)]),
    "th2d_fermion_matrix": (I, [P, C.POINTER(Geom), P, P]),
    "th2d_zdet": (I, [P, I, P, C.POINTER(D), C.POINTER(D)]),
    # th2d_bag.h
    "th2d_bag_apply": (I, [P, C.POINTER(BagGeom), P, P, P, I, I]),
    "th2d_bag_cg": (I, [P, C.POINTER(BagGeom), P, P, P, I, D, I, C.POINTER(I), C.POINTER(I)]),
    "th2d_lu_batched": (I, [P, C.POINTER(Batch), P, I, P, P]),
    "th2d_lu_logdet": (I, [P, C.POINTER(Batch), P, I, P, P, P, P]),
    "th2d_lu_logdet_refined": (I, [P, C.POINTER(Batch), P, P, P, P, P, P]),
    "th2d_lu_inverse": (I, [P, C.POINTER(Batch), P, I, P, P]),
    "th2d_lu_inverse_noalloc": (I, [P, C.POINTER(Batch), P, I, P, P]),
    "th2d_gemm_batched": (I, [P, P, I, P, I, I, I, D]),
    "th2d_set_dense_mfma": (I, [I]),
    "th2d_set_lu_panel": (I, [I]),
    "th2d_set_lu_lookahead": (I, [I]),
    "th2d_copy_f64_to_f32": (I, [P, P, P, L]),
    "th2d_bag_build_masked": (I, [P, P, P, P, C.POINTER(Batch), P, I]),
    "th2d_bag_create": (I, [P, C.POINTER(BagGeom), C.POINTER(P)]),
    "th2d_bag_destroy": (I, [P]),
    "th2d_bag_set_field": (I, [P, P]),
    "th2d_bag_refresh": (I, [P]),
    "th2d_bag_logdet": (I, [P, C.POINTER(D), C.POINTER(I), C.POINTER(D), C.POINTER(I)]),
    "th2d_bag_ratio": (I, [P, P, I, P, I, C.POINTER(D)]),
    "th2d_bag_accept": (I, [P, P, I, P, P, I]),
    "th2d_bag_views": (I, [P, C.POINTER(P), C.POINTER(P), C.POINTER(I)]),
    "th2d_bag_calc_dinv": (I, [P, C.POINTER(BagGeom), P]),
    "th2d_bag_small_weights": (I, [P, C.POINTER(BagGeom), P, P, L, P]),
    "th2d_bag_brute_enumerate": (I, [P, C.POINTER(BagGeom), P, P, C.POINTER(L)]),
    "th2d_bag_matrix_element": (I, [C.POINTER(BagGeom), I, I, I, I, D, C.POINTER(D)]),
    "th2d_bag_determinant": (I, [P, C.POINTER(BagGeom), P, C.POINTER(D), C.POINTER(I)]),
    "th2d_bag_batch_create": (I, [P, I, P, P, C.POINTER(P)]),
    "th2d_bag_batch_destroy": (I, [P]),
    "th2d_bag_batch_sweep": (I, [P, I, U64, U64, P]),
    "th2d_bag_batch_refresh": (I, [P]),
    "th2d_bag_batch_state": (I, [P, P, P, P, P]),
}

_lib = None


class Th2dError(RuntimeError):
    def __init__(self, code: int, where: str):
        self.code = code
        msg = lib().th2d_strerror(code).decode() if _lib is not None else str(code)
        super().__init__(f"{where}: th2d status {code} ({msg})")


def lib():
    """Load libth2d.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                "(the HIP extension is required; there is no CPU fallback)")
        lib_ = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib_, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib_
    return _lib


def check(code: int, where: str) -> int:
    if code != TH2D_OK:
        raise Th2dError(code, where)
    return code
