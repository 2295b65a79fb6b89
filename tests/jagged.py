"""The reference's jagged arrays over numpy storage, for calling the
compatibility libraries (include/thirring_hmc.h, include/thirring_bag.h)
through ctypes: vectors v[t][x], fields F[t][x][dir], configurations
field[t][x], with rows one element longer than nx as alloc_vector makes
them (hmc.c:105-112)."""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def load(name):
    from thirring2d_amd import _native
    _native.lib()
    return C.CDLL(os.path.join(ROOT, "thirring2d_amd", name))


class Rows:
    """2-level jagged array: ptr is a T** over storage[nt][nx+1]."""

    def __init__(self, nt, nx, dtype, nat=None):
        self.nt, self.nx = nt, nx
        self.a = np.zeros((nt, nx + 1), dtype=dtype)
        if nat is not None:
            self.a[:, :nx] = np.asarray(nat).reshape(nt, nx)
        self.rows = (C.c_void_p * nt)(*[self.a[t].ctypes.data for t in range(nt)])

    @property
    def ptr(self):
        return C.cast(self.rows, C.c_void_p)

    def natural(self):
        return self.a[:, :self.nx].reshape(-1).copy()


class Links:
    """3-level jagged array F[t][x][dir] (double ***) over storage[nt][nx+1][2]."""

    def __init__(self, nt, nx, nat=None):
        self.nt, self.nx = nt, nx
        self.a = np.zeros((nt, nx + 1, 2))
        if nat is not None:
            self.a[:, :nx, :] = np.asarray(nat).reshape(nt, nx, 2)
        self.cols = [(C.c_void_p * (nx + 1))(*[self.a[t, x].ctypes.data for x in range(nx + 1)])
                     for t in range(nt)]
        self.rows = (C.c_void_p * nt)(*[C.cast(c, C.c_void_p).value for c in self.cols])

    @property
    def ptr(self):
        return C.cast(self.rows, C.c_void_p)

    def natural(self):
        return self.a[:, :self.nx, :].reshape(-1, 2).copy()
