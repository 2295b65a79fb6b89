"""The allocation-free LU inverse (th2d_lu_inverse_noalloc), called through
the C ABI: its results equal th2d_lu_inverse's bit for bit on two calls in
a row, the second with a smaller batch (the static scratch keeps the first
call's contents), in FP64 and FP32; a batch beyond the scratch is refused."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _inverses(Bm, N, mats, prec):
    db = Bm.DenseBatch(mats, prec).lu()
    lib = N.lib()
    X0 = Bm.DeviceArray(db.A.nbytes)
    X1 = Bm.DeviceArray(db.A.nbytes)
    N.check(lib.th2d_lu_inverse(db.ctx.ptr, C.byref(db.batch), db.A.p, prec, db.piv.p, X0.p),
            "inverse")
    N.check(lib.th2d_lu_inverse_noalloc(db.ctx.ptr, C.byref(db.batch), db.A.p, prec, db.piv.p,
                                        X1.p), "inverse_noalloc")
    n = db.A.nbytes // np.dtype(db.dtype).itemsize
    return db, X0.download((n,), db.dtype), X1.download((n,), db.dtype)


def test_noalloc_inverse_matches_allocating_inverse():
    from thirring2d_amd import _native as N
    from thirring2d_amd import bag as Bm
    rng = np.random.default_rng(17)
    for ns in ([50, 130, 257], [20, 33]):  # the second call is smaller
        mats = [rng.normal(size=(n, n)) + 3 * np.eye(n) for n in ns]
        for prec in (N.PREC_F64, N.PREC_F32):
            db, x0, x1 = _inverses(Bm, N, mats, prec)
            assert np.array_equal(x0, x1)
            if prec == N.PREC_F64:
                for o, n, M in zip(db.off, db.ns, mats):
                    Xi = x1[o:o + n * n].reshape(n, n)
                    assert np.max(np.abs(Xi @ M - np.eye(n))) <= 1e-10


def test_noalloc_inverse_refuses_batches_beyond_scratch():
    from thirring2d_amd import _native as N
    from thirring2d_amd import bag as Bm
    db = Bm.DenseBatch([np.eye(4)], N.PREC_F64).lu()
    big = N.Batch(1, 4194305, db._off.p, db._n.p, db._poff.p)  # refused before any launch
    X = Bm.DeviceArray(db.A.nbytes)
    rc = N.lib().th2d_lu_inverse_noalloc(db.ctx.ptr, C.byref(big), db.A.p, N.PREC_F64, db.piv.p,
                                         X.p)
    assert rc == N.TH2D_EARG
