"""Host-side mirror of the reference's HMC interface (hmc.c) over libth2d.

The reference keeps the lattice in globals and compile-time constants and
exposes ``fm_mul``, ``fm_conjugate_mul``, ``fmdm_invert_cg``,
``momentum_step``, ``gauge_step``, ``update_gauge`` ... (hmc.c:73-746).
``StaggeredLattice`` holds the same state as an object: the link angles
live on the device in eo order, the methods take and return numpy arrays in
the reference's natural site order (s = t*nx + x) and every numerical step
runs in the HIP kernels of libth2d.
"""
from __future__ import annotations

import atexit
import ctypes as C

import numpy as np

from . import _native as N

# Set at interpreter exit, after the context is destroyed: device memory
# still referenced by live objects is then left to the process's end
# rather than freed from finalizers while the runtime shuts down.
_SHUTDOWN = [False]


class Context:
    """One libth2d context (stream + scratch) per process and device."""

    _inst = None

    def __init__(self):
        self.ptr = C.c_void_p()
        N.check(N.lib().th2d_ctx_create(C.byref(self.ptr), None), "th2d_ctx_create")

    @classmethod
    def get(cls) -> "Context":
        if cls._inst is None:
            cls._inst = Context()
        return cls._inst

    def sync(self):
        N.check(N.lib().th2d_sync(self.ptr), "th2d_sync")


@atexit.register
def _teardown():
    """Drain and destroy the context before the interpreter finalizes, so
    that no HIP call runs during the runtime's own teardown."""
    ctx = Context._inst
    if ctx is not None:
        try:
            N.lib().th2d_sync(ctx.ptr)
            N.lib().th2d_ctx_destroy(ctx.ptr)
        except Exception:
            pass
        Context._inst = None
    _SHUTDOWN[0] = True


class DeviceArray:
    """A device allocation with numpy upload/download."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        self.ptr = C.c_void_p()
        N.check(N.lib().th2d_malloc(C.byref(self.ptr), max(self.nbytes, 1)), "th2d_malloc")

    def __del__(self):
        try:
            if self.ptr and not _SHUTDOWN[0]:
                N.lib().th2d_free(self.ptr)
                self.ptr = C.c_void_p()
        except Exception:
            pass

    @classmethod
    def from_host(cls, arr: np.ndarray) -> "DeviceArray":
        arr = np.ascontiguousarray(arr)
        d = cls(arr.nbytes)
        d.upload(arr)
        return d

    def upload(self, arr: np.ndarray):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        N.check(N.lib().th2d_h2d(Context.get().ptr, self.ptr, arr.ctypes.data_as(C.c_void_p),
                                 arr.nbytes), "th2d_h2d")

    def download(self, shape, dtype) -> np.ndarray:
        out = np.empty(shape, dtype=dtype)
        assert out.nbytes <= self.nbytes
        N.check(N.lib().th2d_d2h(Context.get().ptr, out.ctypes.data_as(C.c_void_p), self.ptr,
                                 out.nbytes), "th2d_d2h")
        return out

    @property
    def p(self):
        return self.ptr


def cg_params(tol_rel=1e-10, tol_abs_rr=0.0, max_iter=100000, precision=N.PREC_F64,
              delta=1e-4) -> N.CgParams:
    return N.CgParams(tol_rel, tol_abs_rr, max_iter, precision, delta)


class StaggeredLattice:
    """U(1) staggered fermions on an nt x nx lattice (hmc.c)."""

    def __init__(self, nt, nx, m=0.1, g=0.5, mu=0.0, anti_t=1, anti_x=1, nf=2):
        if nt % 2 or nx % 2:
            raise ValueError("lattice extents must be even")
        self.nt, self.nx, self.V = nt, nx, nt * nx
        self.g, self.nf = g, nf
        self.geom = N.Geom(nt, nx, anti_t, anti_x, m, mu)
        self.ctx = Context.get()
        V = self.V
        self.A = DeviceArray(16 * V)        # eo [dir][V] doubles
        self.K = DeviceArray(32 * V)        # folded links FP64
        self.K32 = DeviceArray(16 * V)      # folded links FP32
        self._stage = DeviceArray(16 * V)
        self._a = DeviceArray(16 * V)
        self._b = DeviceArray(16 * V)
        self._c = DeviceArray(16 * V)
        self.set_links(np.zeros((V, 2)))

    # -------------------------------------------------------------- fields
    @property
    def m(self):
        return self.geom.m

    @property
    def mu(self):
        return self.geom.mu

    def _g(self):
        return C.byref(self.geom)

    def set_links(self, A_nat: np.ndarray):
        """Upload A[t][x][dir] (natural order, shape (V, 2)) and refold."""
        A_nat = np.ascontiguousarray(A_nat, dtype=np.float64).reshape(self.V, 2)
        self._stage.upload(A_nat)
        L = N.lib()
        N.check(L.th2d_links_to_eo(self.ctx.ptr, self._g(), self.A.p, self._stage.p), "links")
        self.refold()

    def refold(self):
        N.check(N.lib().th2d_fold_links(self.ctx.ptr, self._g(), self.K.p, self.K32.p, self.A.p),
                "fold")

    def get_links(self) -> np.ndarray:
        N.check(N.lib().th2d_links_from_eo(self.ctx.ptr, self._g(), self._stage.p, self.A.p),
                "links")
        return self._stage.download((self.V, 2), np.float64)

    def random_links(self, seed: int):
        N.check(N.lib().th2d_random_links(self.ctx.ptr, self._g(), self.A.p, seed), "rand")
        self.refold()

    def thermalize_pure_gauge(self, sweeps=100, seed=0, counter=0, from_zero=True):
        """A = 0, then `sweeps` pure-gauge Metropolis sweeps with device
        draws (hmc.c:927-929 with update_puregauge_hb, hmc.c:82-93)."""
        if from_zero:
            self.set_links(np.zeros((self.V, 2)))
        N.check(N.lib().th2d_puregauge_sweeps(self.ctx.ptr, self._g(), self.A.p, self.g,
                                              self.nf, sweeps, seed, counter), "puregauge")
        self.refold()

    def to_eo(self, v: np.ndarray, dst: DeviceArray):
        self._stage.upload(np.ascontiguousarray(v, dtype=np.complex128))
        N.check(N.lib().th2d_vec_to_eo(self.ctx.ptr, self._g(), dst.p, self._stage.p), "to_eo")

    def from_eo(self, src: DeviceArray) -> np.ndarray:
        N.check(N.lib().th2d_vec_from_eo(self.ctx.ptr, self._g(), self._stage.p, src.p), "from_eo")
        return self._stage.download(self.V, np.complex128)

    # ------------------------------------------------------------ operator
    def fm_mul(self, v: np.ndarray) -> np.ndarray:
        """chi = M v (hmc.c:123-184)."""
        return self._apply(v, 0)

    def fm_conjugate_mul(self, v: np.ndarray) -> np.ndarray:
        """chi = M^dag v: the true adjoint (hmc.c:188-249 is meant to be it)."""
        return self._apply(v, 1)

    def _apply(self, v, dagger, precision=N.PREC_F64):
        self.to_eo(v, self._a)
        L = N.lib()
        if precision == N.PREC_F32:
            n = self.V
            N.check(L.th2d_vec_f64_to_f32(self.ctx.ptr, self._c.p, self._a.p, n), "f32")
            N.check(L.th2d_dslash(self.ctx.ptr, self._g(), self._b.p, self._c.p, self.K32.p,
                                  dagger, N.PREC_F32, 1), "dslash")
            N.check(L.th2d_vec_f32_to_f64(self.ctx.ptr, self._a.p, self._b.p, n), "f64")
            return self.from_eo(self._a)
        N.check(L.th2d_dslash(self.ctx.ptr, self._g(), self._b.p, self._a.p, self.K.p, dagger,
                              N.PREC_F64, 1), "dslash")
        return self.from_eo(self._b)

    def apply_f32(self, v, dagger=0):
        return self._apply(v, dagger, N.PREC_F32)

    # -------------------------------------------------------------- solver
    def fmdm_invert_cg(self, b: np.ndarray, params: N.CgParams | None = None):
        """x = (M^dag M)^{-1} b (hmc.c:341-404); returns (x, CgInfo)."""
        params = params or cg_params()
        self.to_eo(b, self._a)
        info = N.CgInfo()
        rc = N.lib().th2d_cg_mdagm(self.ctx.ptr, self._g(), self._b.p, self._a.p, self.K.p,
                                   self.K32.p, C.byref(params), C.byref(info))
        if rc not in (N.TH2D_OK, N.TH2D_ENOCONV, N.TH2D_EBREAKDOWN):
            N.check(rc, "th2d_cg_mdagm")
        return self.from_eo(self._b), info

    # --------------------------------------------------------------- force
    def force(self, chi: np.ndarray, st: np.ndarray | None = None) -> np.ndarray:
        """dS/dA (natural (V,2)) given chi = (M^dag M)^{-1} psi."""
        self.to_eo(chi, self._a)
        stp = None
        if st is not None:
            self.to_eo(st, self._c)
            stp = self._c.p
        N.check(N.lib().th2d_force(self.ctx.ptr, self._g(), None, self.A.p, self.K.p, self._a.p,
                                   stp, 0.0, self.g, self.nf, self._b.p), "force")
        N.check(N.lib().th2d_links_from_eo(self.ctx.ptr, self._g(), self._stage.p, self._b.p),
                "links")
        return self._stage.download((self.V, 2), np.float64)

    def momentum_step(self, mom, psi, st, eps, params=None):
        """mom - eps dS/dA with chi solved on the device (hmc.c:504-665)."""
        chi, info = self.fmdm_invert_cg(psi, params)
        F = self.force(chi, st)
        return np.asarray(mom) - eps * F, F, info

    def gauge_step(self, mom, eps):
        """A += 2 eps mom on the device (hmc.c:97-101)."""
        mom = np.ascontiguousarray(mom, dtype=np.float64).reshape(self.V, 2)
        self._stage.upload(mom)
        L = N.lib()
        N.check(L.th2d_links_to_eo(self.ctx.ptr, self._g(), self._a.p, self._stage.p), "links")
        N.check(L.th2d_gauge_step(self.ctx.ptr, self._g(), self.A.p, self._a.p, eps, self.K.p,
                                  self.K32.p), "gauge_step")

    def calc_gauge_action(self) -> float:
        out = C.c_double()
        N.check(N.lib().th2d_gauge_action(self.ctx.ptr, self._g(), self.A.p, self.g, self.nf,
                                          C.byref(out)), "gauge_action")
        return out.value

    # ----------------------------------------------------------- trajectory
    def update_gauge(self, nsteps=10, traj_len=1.0, params=None, phys_norm=0, noise=None,
                     u_accept=None, seed=0, counter=0) -> N.HmcResult:
        """One HMC trajectory (hmc.c:671-746) on the device-resident field.

        noise = (xi, mom, st) natural-order arrays from the reference's MT
        stream, or None for the device generator keyed by (seed, counter).
        """
        params = params or cg_params()
        hp = N.HmcParams(self.g, self.nf, nsteps, traj_len, phys_norm, params)
        res = N.HmcResult()
        if noise is not None:
            xi, mom, st = (np.ascontiguousarray(a) for a in noise)
            xi = xi.astype(np.complex128); st = st.astype(np.complex128)
            mom = mom.astype(np.float64)
            args = [xi.ctypes.data_as(C.c_void_p), mom.ctypes.data_as(C.c_void_p),
                    st.ctypes.data_as(C.c_void_p)]
        else:
            args = [None, None, None]
        u = 0.5 if u_accept is None else u_accept
        rc = N.lib().th2d_hmc_trajectory(self.ctx.ptr, self._g(), C.byref(hp), self.A.p, *args,
                                         u, seed, counter, C.byref(res))
        if rc not in (N.TH2D_OK, N.TH2D_ENOCONV, N.TH2D_EBREAKDOWN):
            N.check(rc, "th2d_hmc_trajectory")
        self.refold()
        return res


class ChainBatch:
    """Independent chains advanced in lockstep: every inversion of all
    chains is one batched solve (th2d_hmc_trajectory_batched).  Chains may
    have different lattice sizes, couplings and masses."""

    def __init__(self, lattices):
        self.lats = list(lattices)
        self.ctx = Context.get()
        n = len(self.lats)
        self._geoms = (N.Geom * n)(*[l.geom for l in self.lats])

    def __len__(self):
        return len(self.lats)

    def update_gauge(self, nsteps=10, traj_len=1.0, params=None, phys_norm=0, noise=None,
                     u_accept=None, seed=0, counters=None):
        """One trajectory per chain; noise = list of (xi, mom, st) or None."""
        n = len(self.lats)
        params = params or cg_params()
        hps = (N.HmcParams * n)(*[N.HmcParams(l.g, l.nf, nsteps, traj_len, phys_norm, params)
                                  for l in self.lats])
        As = (C.c_void_p * n)(*[l.A.p for l in self.lats])
        u = (C.c_double * n)(*(u_accept if u_accept is not None else [0.5] * n))
        ctr = (C.c_ulonglong * n)(*(counters if counters is not None else [0] * n))
        res = (N.HmcResult * n)()
        keep = []
        if noise is not None:
            ptrs = [[], [], []]
            for (xi, mom, st) in noise:
                arrs = (np.ascontiguousarray(xi, dtype=np.complex128),
                        np.ascontiguousarray(mom, dtype=np.float64),
                        np.ascontiguousarray(st, dtype=np.complex128))
                keep.append(arrs)
                for q in range(3):
                    ptrs[q].append(arrs[q].ctypes.data_as(C.c_void_p))
            xs, ms, ss = ((C.c_void_p * n)(*p) for p in ptrs)
        else:
            xs = ms = ss = None
        rc = N.lib().th2d_hmc_trajectory_batched(self.ctx.ptr, n, self._geoms, hps, As, xs, ms,
                                                 ss, u, seed, ctr, res)
        if rc not in (N.TH2D_OK, N.TH2D_ENOCONV, N.TH2D_EBREAKDOWN):
            N.check(rc, "th2d_hmc_trajectory_batched")
        for l in self.lats:
            l.refold()
        return list(res)

    def fmdm_invert_cg(self, bs, params=None):
        """Batched x_k = (M_k^dag M_k)^{-1} b_k; returns ([x_k], [CgInfo])."""
        n = len(self.lats)
        params = params or cg_params()
        xs_d = [DeviceArray(16 * l.V) for l in self.lats]
        bs_d = [DeviceArray(16 * l.V) for l in self.lats]
        for l, b, bd in zip(self.lats, bs, bs_d):
            l.to_eo(b, bd)
        xs = (C.c_void_p * n)(*[d.p for d in xs_d])
        bp = (C.c_void_p * n)(*[d.p for d in bs_d])
        ks = (C.c_void_p * n)(*[l.K.p for l in self.lats])
        k32 = (C.c_void_p * n)(*[l.K32.p for l in self.lats])
        infos = (N.CgInfo * n)()
        rc = N.lib().th2d_cg_mdagm_batched(self.ctx.ptr, n, self._geoms, xs, bp, ks, k32,
                                           C.byref(params), infos)
        if rc not in (N.TH2D_OK, N.TH2D_ENOCONV, N.TH2D_EBREAKDOWN):
            N.check(rc, "th2d_cg_mdagm_batched")
        return [l.from_eo(d) for l, d in zip(self.lats, xs_d)], list(infos)
