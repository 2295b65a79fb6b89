"""Host parts of the compatibility libraries (no GPU): the reference's MT
stream of libthirring_hmc.so against the oracle (which is pinned bitwise to
the reference's own mersenne_inline.c and its golden draws)."""
import ctypes as C

import numpy as np

from jagged import load


def _draws(lib, n):
    mi = C.c_int.in_dll(lib, "mersenne_i")
    arr = (C.c_double * 624).in_dll(lib, "mersenne_array")
    gen = lib.mersenne_generate
    gen.restype = C.c_double
    out = []
    for _ in range(n):
        if mi.value > 0:
            mi.value -= 1
            out.append(arr[mi.value])
        else:
            out.append(gen())
    return np.array(out)


def test_compat_mt_stream_matches_oracle(oracle):
    lib = load("libthirring_hmc.so")
    for seed in (2001, 1002, 7):
        lib.seed_mersenne(C.c_long(seed))
        a = _draws(lib, 2000)
        b = oracle.Rng(seed).draws(2000)
        np.testing.assert_array_equal(a, b)


def test_compat_weak_globals_writable():
    lib = load("libthirring_hmc.so")
    for name, val in (("m", 0.25), ("g", 0.7), ("mu", 0.1)):
        C.c_double.in_dll(lib, name).value = val
        assert C.c_double.in_dll(lib, name).value == val
    bag = load("libthirring_bag.so")
    C.c_int.in_dll(bag, "max_changes").value = 5
    assert C.c_int.in_dll(bag, "current_sign").value == 1
