"""Host-side mirror of the reference's fermion-bag interface over libth2d.

The reference keeps the bag configuration ``field``, the propagator
``Dinv`` and the background inverse in globals (fermionbag.c,
fluctuation_determinant.c) and exposes ``fM``/``fM_transpose``,
``cg_propagator`` (vec_ops.c:96-321), ``calc_Dinv`` (fermion_matrix.c:
191-351), ``determinant_mu`` (full_determinant.c:144-186) and the ratio /
accept hooks (fluctuation_determinant.c:305-871).  ``BagLattice`` and
``BagState`` hold the same state as objects; arrays are numpy in natural
site order (s = t*nx + x) and every numerical step runs in libth2d.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N
from .hmc import _SHUTDOWN, Context, DeviceArray

BC_ANTISYMMETRIC = 0
BC_OPENX = 1

# field values (Thirring.h:34-39): empty, monomer, link tails 2+dir, heads 4+dir
EMPTY, MONOMER = 0, 1


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class BagLattice:
    """Lattice of the fermion-bag family (Thirring.h:14-50, runtime values)."""

    def __init__(self, nt, nx, bc=BC_ANTISYMMETRIC, with_mass_monomers=0, m=0.1, mu=0.0):
        self.geom = N.BagGeom(nt, nx, bc, with_mass_monomers, m, mu)
        self.nt, self.nx, self.V = nt, nx, nt * nx
        self.ctx = Context.get()

    def _field_dev(self, field):
        return DeviceArray.from_host(_i32(field))

    def fM(self, field, psi, transpose=False):
        """chi = fM psi (vec_ops.c:96-131) or fM^T psi (:135-172); psi may
        hold several vectors (rows)."""
        psi = _f64(psi)
        nrhs = psi.size // self.V
        f = self._field_dev(field)
        src = DeviceArray.from_host(psi)
        dst = DeviceArray(psi.nbytes)
        N.check(N.lib().th2d_bag_apply(self.ctx.ptr, C.byref(self.geom), f.p, dst.p, src.p,
                                       1 if transpose else 0, nrhs), "th2d_bag_apply")
        return dst.download(psi.shape, np.float64)

    def fM_transpose(self, field, psi):
        return self.fM(field, psi, transpose=True)

    def cg_propagator(self, field, src, tol_abs_rr=1e-30, max_iter=10000):
        """Solve fM x = src (cg_propagator, vec_ops.c:311-321); returns
        (x, iterations per RHS, status per RHS)."""
        src = _f64(src)
        nrhs = src.size // self.V
        f = self._field_dev(field)
        s = DeviceArray.from_host(src)
        x = DeviceArray(src.nbytes)
        its = (C.c_int * nrhs)()
        sts = (C.c_int * nrhs)()
        rc = N.lib().th2d_bag_cg(self.ctx.ptr, C.byref(self.geom), f.p, x.p, s.p, nrhs,
                                 tol_abs_rr, max_iter, its, sts)
        if rc not in (N.TH2D_OK, N.TH2D_ENOCONV, N.TH2D_EBREAKDOWN):
            N.check(rc, "th2d_bag_cg")
        return x.download(src.shape, np.float64), list(its), list(sts)

    def calc_Dinv(self):
        """Dinv[e*V/2 + o] = (D^-1)_{e,o}, Dinv[V^2/4 + o*V/2 + e] =
        (D^-1)_{o,e} (fermion_matrix.c:191-351)."""
        h = self.V // 2
        d = DeviceArray(16 * h * h)
        N.check(N.lib().th2d_bag_calc_dinv(self.ctx.ptr, C.byref(self.geom), d.p),
                "th2d_bag_calc_dinv")
        return d.download((2 * h * h,), np.float64)

    def small_weights(self, Dinv, configs):
        """det_e det_o of configurations (rows of uint8 fields) from Dinv
        (tests/thirring_brute.c:273-345)."""
        cfg = np.ascontiguousarray(configs, dtype=np.uint8)
        count = cfg.shape[0]
        dD = DeviceArray.from_host(_f64(Dinv))
        dc = DeviceArray.from_host(cfg)
        dw = DeviceArray(8 * max(count, 1))
        N.check(N.lib().th2d_bag_small_weights(self.ctx.ptr, C.byref(self.geom), dD.p, dc.p,
                                               count, dw.p), "th2d_bag_small_weights")
        return dw.download((count,), np.float64)

    def brute_enumerate(self):
        """Z[nl, nm], |Z|[nl, nm] over every configuration (thirring_brute)."""
        h = self.V // 2
        Z = np.zeros((h + 1, self.V + 1))
        Za = np.zeros_like(Z)
        n = C.c_long()
        N.check(N.lib().th2d_bag_brute_enumerate(self.ctx.ptr, C.byref(self.geom),
                                                 Z.ctypes.data_as(C.c_void_p),
                                                 Za.ctypes.data_as(C.c_void_p), C.byref(n)),
                "th2d_bag_brute_enumerate")
        return Z, Za, n.value


class DenseBatch:
    """Square matrices of uneven orders packed into one device buffer."""

    def __init__(self, mats=None, precision=N.PREC_F64, orders=None):
        """From host matrices, or uninitialised device storage for the
        given orders."""
        self.ctx = Context.get()
        self.prec = precision
        self.dtype = np.float64 if precision == N.PREC_F64 else np.float32
        self.ns = ([int(n) for n in orders] if mats is None
                   else [int(np.asarray(m).shape[0]) for m in mats])
        off = np.cumsum([0] + [n * n for n in self.ns])[:-1].astype(np.int64)
        poff = np.cumsum([0] + self.ns)[:-1].astype(np.int64)
        self.off, self.poff = off, poff
        if mats is None:
            self.A = DeviceArray(np.dtype(self.dtype).itemsize * sum(n * n for n in self.ns))
        else:
            flat = np.concatenate([np.asarray(m, dtype=self.dtype).ravel() for m in mats])
            self.A = DeviceArray.from_host(flat)
        self._off = DeviceArray.from_host(off)
        self._poff = DeviceArray.from_host(poff)
        self._n = DeviceArray.from_host(np.array(self.ns, dtype=np.int32))
        self.count = len(self.ns)
        self.batch = N.Batch(self.count, max(self.ns), self._off.p, self._n.p, self._poff.p)
        self.piv = DeviceArray(4 * max(1, sum(self.ns)))
        self.info = DeviceArray(4 * self.count)

    def lu(self):
        N.check(N.lib().th2d_lu_batched(self.ctx.ptr, C.byref(self.batch), self.A.p, self.prec,
                                        self.piv.p, self.info.p), "th2d_lu_batched")
        return self

    def logdet(self):
        la = DeviceArray(8 * self.count)
        sg = DeviceArray(4 * self.count)
        N.check(N.lib().th2d_lu_logdet(self.ctx.ptr, C.byref(self.batch), self.A.p, self.prec,
                                       self.piv.p, self.info.p, la.p, sg.p), "th2d_lu_logdet")
        return la.download((self.count,), np.float64), sg.download((self.count,), np.int32)

    def logdet_refined(self, original: "DenseBatch"):
        """log|det| of these FP32 factors corrected in FP64 against the FP64
        matrices `original` (th2d_lu_logdet_refined)."""
        la = DeviceArray(8 * self.count)
        sg = DeviceArray(4 * self.count)
        N.check(N.lib().th2d_lu_logdet_refined(self.ctx.ptr, C.byref(self.batch), original.A.p,
                                               self.A.p, self.piv.p, self.info.p, la.p, sg.p),
                "th2d_lu_logdet_refined")
        return la.download((self.count,), np.float64), sg.download((self.count,), np.int32)

    def inverse(self):
        X = DeviceArray(self.A.nbytes)
        N.check(N.lib().th2d_lu_inverse(self.ctx.ptr, C.byref(self.batch), self.A.p, self.prec,
                                        self.piv.p, X.p), "th2d_lu_inverse")
        flat = X.download((self.A.nbytes // np.dtype(self.dtype).itemsize,), self.dtype)
        return [flat[o:o + n * n].reshape(n, n) for o, n in zip(self.off, self.ns)]

    def matrices(self):
        flat = self.A.download((self.A.nbytes // np.dtype(self.dtype).itemsize,), self.dtype)
        return [flat[o:o + n * n].reshape(n, n) for o, n in zip(self.off, self.ns)]

    def infos(self):
        return self.info.download((self.count,), np.int32)


class BagState:
    """Determinant-ratio state of the fermion-bag update: B = (D_FF)^-1 over
    the free sites and log|det D_FF| (update_background /
    extended_determinant / accept hooks, fluctuation_determinant.c)."""

    def __init__(self, lat: BagLattice):
        self.lat = lat
        self.ptr = C.c_void_p()
        N.check(N.lib().th2d_bag_create(lat.ctx.ptr, C.byref(lat.geom), C.byref(self.ptr)),
                "th2d_bag_create")

    def __del__(self):
        try:
            if self.ptr and not _SHUTDOWN[0]:
                N.lib().th2d_bag_destroy(self.ptr)
                self.ptr = C.c_void_p()
        except Exception:
            pass

    def set_field(self, field):
        f = _i32(field)
        N.check(N.lib().th2d_bag_set_field(self.ptr, f.ctypes.data_as(C.c_void_p)),
                "th2d_bag_set_field")

    def refresh(self):
        N.check(N.lib().th2d_bag_refresh(self.ptr), "th2d_bag_refresh")

    def logdet(self):
        la, sg, lad, sgd = C.c_double(), C.c_int(), C.c_double(), C.c_int()
        N.check(N.lib().th2d_bag_logdet(self.ptr, C.byref(la), C.byref(sg), C.byref(lad),
                                        C.byref(sgd)), "th2d_bag_logdet")
        return la.value, sg.value, lad.value, sgd.value

    def ratio(self, occupy=(), free=()):
        o, f = _i32(list(occupy)), _i32(list(free))
        r = C.c_double()
        N.check(N.lib().th2d_bag_ratio(self.ptr, o.ctypes.data_as(C.c_void_p), len(o),
                                       f.ctypes.data_as(C.c_void_p), len(f), C.byref(r)),
                "th2d_bag_ratio")
        return r.value

    def accept(self, occupy=(), values=(), free=()):
        o, v, f = _i32(list(occupy)), _i32(list(values)), _i32(list(free))
        N.check(N.lib().th2d_bag_accept(self.ptr, o.ctypes.data_as(C.c_void_p), len(o),
                                        v.ctypes.data_as(C.c_void_p),
                                        f.ctypes.data_as(C.c_void_p), len(f)),
                "th2d_bag_accept")


class BagBatch:
    """Many bag configurations advanced by Metropolis proposals in lockstep
    on the device (th2d_bag_batch_*): the update loop of fermionbag.c:
    284-379 over the determinant ratios of fluctuation_determinant.c for a
    batch of matrices of any sizes."""

    def __init__(self, lattices, fields):
        self.ctx = Context.get()
        self.lats = list(lattices)
        self.count = len(self.lats)
        self.Vs = [l.V for l in self.lats]
        geoms = (N.BagGeom * self.count)(*[l.geom for l in self.lats])
        flat = np.concatenate([_i32(f) for f in fields])
        self.ptr = C.c_void_p()
        N.check(N.lib().th2d_bag_batch_create(self.ctx.ptr, self.count, geoms,
                                              flat.ctypes.data_as(C.c_void_p),
                                              C.byref(self.ptr)), "th2d_bag_batch_create")

    def __del__(self):
        try:
            if self.ptr and not _SHUTDOWN[0]:
                N.lib().th2d_bag_batch_destroy(self.ptr)
                self.ptr = C.c_void_p()
        except Exception:
            pass

    def sweep(self, nsteps, seed=0, counter=0, trace=False):
        """nsteps proposals per matrix; with trace, the proposals as an
        (nsteps, count) array of BagMove records."""
        tr = DeviceArray(C.sizeof(N.BagMove) * nsteps * self.count) if trace else None
        N.check(N.lib().th2d_bag_batch_sweep(self.ptr, nsteps, seed, counter,
                                             tr.p if tr else None), "th2d_bag_batch_sweep")
        if not trace:
            return None
        raw = tr.download((C.sizeof(N.BagMove) * nsteps * self.count,), np.uint8)
        recs = (N.BagMove * (nsteps * self.count)).from_buffer_copy(raw.tobytes())
        return [list(recs[k * self.count:(k + 1) * self.count]) for k in range(nsteps)]

    def refresh(self):
        N.check(N.lib().th2d_bag_batch_refresh(self.ptr), "th2d_bag_batch_refresh")

    def state(self):
        """(log|det D_FF|, sign, fields, accepted) per matrix."""
        la = np.empty(self.count)
        sg = np.empty(self.count, np.int32)
        fl = np.empty(sum(self.Vs), np.int32)
        acc = np.empty(self.count, np.int64)
        N.check(N.lib().th2d_bag_batch_state(self.ptr, la.ctypes.data_as(C.c_void_p),
                                             sg.ctypes.data_as(C.c_void_p),
                                             fl.ctypes.data_as(C.c_void_p),
                                             acc.ctypes.data_as(C.c_void_p)),
                "th2d_bag_batch_state")
        off = np.cumsum([0] + self.Vs)
        return la, sg, [fl[off[k]:off[k + 1]] for k in range(self.count)], acc


def build_masked_batch(lattices, fields, precision=N.PREC_F64) -> DenseBatch:
    """Dense masked matrices (identity rows/columns at occupied sites,
    fM_index elements between free sites; vec_ops.c:107-131) of several
    configurations on lattices of any sizes, built on the device in one
    launch (th2d_bag_build_masked)."""
    ctx = Context.get()
    geoms = (N.BagGeom * len(lattices))(*[l.geom for l in lattices])
    dgeoms = DeviceArray(C.sizeof(geoms))
    N.check(N.lib().th2d_h2d(ctx.ptr, dgeoms.p, C.cast(geoms, C.c_void_p), C.sizeof(geoms)),
            "th2d_h2d")
    flat = np.concatenate([_i32(f) for f in fields])
    foff = np.cumsum([0] + [len(f) for f in fields])[:-1].astype(np.int64)
    dF, dFo = DeviceArray.from_host(flat), DeviceArray.from_host(foff)
    db = DenseBatch(None, precision, orders=[l.V for l in lattices])
    N.check(N.lib().th2d_bag_build_masked(ctx.ptr, dgeoms.p, dF.p, dFo.p, C.byref(db.batch),
                                          db.A.p, precision), "th2d_bag_build_masked")
    db._keep = (dgeoms, dF, dFo)
    return db
