"""Python binding of the CPU oracle (test infrastructure only).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's cpu_baseline leg
import this module; the product (``thirring2d_amd``) never does.

Arrays follow the reference layout: site s = t*nx + x, complex vectors
as complex128[V], link fields as float64[V, 2] (A[t][x][dir]).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

D = C.c_double
I = C.c_int
P = C.c_void_p


class Lattice(C.Structure):
    _fields_ = [("nt", I), ("nx", I), ("anti_t", I), ("anti_x", I),
                ("m", D), ("g", D), ("mu", D), ("nf", I)]


class BagLattice(C.Structure):
    _fields_ = [("nt", I), ("nx", I), ("bc", I), ("with_mass_monomers", I),
                ("m", D), ("mu", D)]


class MT(C.Structure):
    _fields_ = [("mt", C.c_uint32 * 624), ("out", D * 624), ("i", I)]


BC_ANTISYMMETRIC = 0
BC_OPENX = 1


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        sig = {
            "orc_mt_seed": (None, [P, C.c_long]),
            "orc_mt_next": (D, [P]),
            "orc_mt_skip": (None, [P, C.c_long]),
            "orc_gauss_pair": (None, [P, P, P]),
            "orc_philox4x32_10": (None, [P, P, P]),
            "orc_philox_gauss2": (None, [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, P, P]),
            "orc_philox_uniform2": (None, [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, P, P]),
            "orc_fm_mul": (None, [P, P, P, P]),
            "orc_fm_dag_mul": (None, [P, P, P, P]),
            "orc_fermion_matrix": (None, [P, P, P]),
            "orc_zlogdet": (I, [I, P, P, P]),
            "orc_cg_mdm": (I, [P, P, P, P, D, D, I, P]),
            "orc_random_pseudofermion": (D, [P, P, P, P, P, I]),
            "orc_stochastic_vector": (None, [P, P, P, I]),
            "orc_random_momentum": (D, [P, P, P, I]),
            "orc_gauge_action": (D, [P, P]),
            "orc_pseudofermion_action": (D, [P, P, P, D, D, I, P]),
            "orc_stochastic_md_action": (D, [P, P, P]),
            "orc_gauge_step": (None, [P, P, P, D]),
            "orc_momentum_step": (I, [P, P, P, P, P, D, D, D, I, P]),
            "orc_update_gauge": (I, [P, P, P, I, D, D, D, I, I, P, P, P]),
            "orc_puregauge_sweep": (None, [P, P, P]),
            "orc_fM_index": (D, [P, I, I, I, I, D]),
            "orc_fM": (None, [P, P, P, P]),
            "orc_fM_transpose": (None, [P, P, P, P]),
            "orc_cg_propagator": (I, [P, P, P, P, D, I, P]),
            "orc_dlogdet": (I, [I, P, P, P]),
            "orc_dinverse": (I, [I, P]),
            "orc_determinant_mu": (I, [P, P, D, P, P]),
            "orc_reduced_matrix": (I, [P, P, D, P]),
            "orc_calc_Dinv": (I, [P, P]),
            "orc_bag_weight": (I, [P, P, P, P, P, P]),
            "orc_brute_enumerate": (C.c_long, [P, P, P]),
            "orc_worm": (C.c_long, [P, P, P, D, I, P, P, P, P]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def ref_mt():
    """The reference's own mersenne_inline.c, compiled unmodified into
    oracle/_ref (None when it has not been built)."""
    global _REF
    if _REF is None:
        path = os.path.join(_HERE, "_ref", "libref_mt.so")
        if not os.path.exists(path):
            return None
        R = C.CDLL(path)
        R.seed_mersenne.restype = None
        R.seed_mersenne.argtypes = [C.c_long]
        R.mersenne_generate.restype = D
        R.mersenne_generate.argtypes = []
        _REF = R
    return _REF


def ref_mt_draws(seed: int, n: int, skip: int = 0) -> np.ndarray:
    """n draws of the reference generator, replaying the mersenne() macro
    (mersenne.h:11) over its exported state."""
    R = ref_mt()
    if R is None:
        raise RuntimeError("oracle/_ref/libref_mt.so not built")
    idx = C.c_int.in_dll(R, "mersenne_i")
    arr = (D * 624).in_dll(R, "mersenne_array")
    R.seed_mersenne(seed)

    def nxt():
        if idx.value > 0:
            idx.value -= 1
            return arr[idx.value]
        return R.mersenne_generate()

    for _ in range(skip):
        nxt()
    return np.array([nxt() for _ in range(n)])


def _p(a: np.ndarray):
    return a.ctypes.data_as(P)


class Rng:
    """MT19937 stream of the reference (seed, reverse consumption)."""

    def __init__(self, seed: int, warmup: int = 0):
        self.st = MT()
        lib().orc_mt_seed(C.byref(self.st), seed)
        if warmup:
            lib().orc_mt_skip(C.byref(self.st), warmup)

    def next(self) -> float:
        return lib().orc_mt_next(C.byref(self.st))

    def draws(self, n: int) -> np.ndarray:
        return np.array([self.next() for _ in range(n)])

    def gauss_pairs(self, n: int) -> np.ndarray:
        """n Box-Muller pairs (orc_gauss_pair) as an (n, 2) array."""
        out = np.empty((n, 2))
        c, s = D(), D()
        f = lib().orc_gauss_pair
        for k in range(n):
            f(C.byref(self.st), C.byref(c), C.byref(s))
            out[k, 0] = c.value
            out[k, 1] = s.value
        return out

    @property
    def ref(self):
        return C.byref(self.st)


def philox4x32_10(ctr_key: np.ndarray) -> np.ndarray:
    """Philox4x32-10 blocks of an (n, 6) uint32 array of (4 counter words,
    2 key words); returns (n, 4) uint32."""
    ck = np.ascontiguousarray(ctr_key, dtype=np.uint32).reshape(-1, 6)
    out = np.empty((ck.shape[0], 4), np.uint32)
    f = lib().orc_philox4x32_10
    for j in range(ck.shape[0]):
        row = ck[j]
        f(_p(row[:4].copy()), _p(row[4:].copy()), _p(out[j]))
    return out


def philox_gauss(seed: int, stream: int, n: int, ctr: int = 0) -> np.ndarray:
    """The device noise mapping: n complex N(0,1) draws for (seed, stream),
    indices 0..n-1 at counter ctr."""
    out = np.empty(n, np.complex128)
    re, im = D(), D()
    f = lib().orc_philox_gauss2
    for i in range(n):
        f(seed, stream, i, ctr, C.byref(re), C.byref(im))
        out[i] = complex(re.value, im.value)
    return out


def philox_uniform(seed: int, stream: int, n: int, ctr: int = 0) -> np.ndarray:
    """(n, 2) uniform pairs of the device generator for (seed, stream),
    indices 0..n-1 at counter ctr."""
    out = np.empty((n, 2))
    a, b = D(), D()
    f = lib().orc_philox_uniform2
    for i in range(n):
        f(seed, stream, i, ctr, C.byref(a), C.byref(b))
        out[i] = (a.value, b.value)
    return out


def lattice(nt, nx, m=0.1, g=0.5, mu=0.0, anti_t=1, anti_x=1, nf=2) -> Lattice:
    return Lattice(nt, nx, anti_t, anti_x, m, g, mu, nf)


def _cv(v):
    v = np.ascontiguousarray(v, dtype=np.complex128)
    return v


def fm_mul(L: Lattice, A, v):
    v = _cv(v); A = np.ascontiguousarray(A, dtype=np.float64)
    out = np.empty_like(v)
    lib().orc_fm_mul(C.byref(L), _p(A), _p(v), _p(out))
    return out


def fm_dag_mul(L: Lattice, A, v):
    v = _cv(v); A = np.ascontiguousarray(A, dtype=np.float64)
    out = np.empty_like(v)
    lib().orc_fm_dag_mul(C.byref(L), _p(A), _p(v), _p(out))
    return out


def fermion_matrix(L: Lattice, A):
    V = L.nt * L.nx
    A = np.ascontiguousarray(A, dtype=np.float64)
    M = np.empty(V * V, dtype=np.complex128)
    lib().orc_fermion_matrix(C.byref(L), _p(A), _p(M))
    return M.reshape(V, V, order="F")  # column-major: M[row + col*V]


def zlogdet(M):
    M = np.asfortranarray(M, dtype=np.complex128).copy(order="F")
    la, ar = D(), D()
    info = lib().orc_zlogdet(M.shape[0], _p(M), C.byref(la), C.byref(ar))
    return la.value, ar.value, info


def cg_mdm(L: Lattice, A, b, tol_abs_rr=0.0, tol_rel=1e-12, max_iter=100000):
    b = _cv(b); A = np.ascontiguousarray(A, dtype=np.float64)
    x = np.empty_like(b)
    st = I()
    it = lib().orc_cg_mdm(C.byref(L), _p(A), _p(b), _p(x), tol_abs_rr, tol_rel, max_iter,
                          C.byref(st))
    return x, it, st.value


def random_pseudofermion(L, rng: Rng, A, phys_norm=0):
    V = L.nt * L.nx
    A = np.ascontiguousarray(A, dtype=np.float64)
    xi = np.empty(V, np.complex128); psi = np.empty(V, np.complex128)
    S = lib().orc_random_pseudofermion(C.byref(L), rng.ref, _p(A), _p(xi), _p(psi), phys_norm)
    return S, xi, psi


def stochastic_vector(L, rng: Rng, phys_norm=0):
    v = np.empty(L.nt * L.nx, np.complex128)
    lib().orc_stochastic_vector(C.byref(L), rng.ref, _p(v), phys_norm)
    return v


def random_momentum(L, rng: Rng, phys_norm=0):
    mom = np.empty((L.nt * L.nx, 2), np.float64)
    S = lib().orc_random_momentum(C.byref(L), rng.ref, _p(mom), phys_norm)
    return S, mom


def gauge_action(L, A):
    A = np.ascontiguousarray(A, dtype=np.float64)
    return lib().orc_gauge_action(C.byref(L), _p(A))


def pseudofermion_action(L, A, psi, tol_abs_rr=0.0, tol_rel=1e-12, max_iter=100000):
    A = np.ascontiguousarray(A, dtype=np.float64); psi = _cv(psi)
    it = I()
    S = lib().orc_pseudofermion_action(C.byref(L), _p(A), _p(psi), tol_abs_rr, tol_rel,
                                       max_iter, C.byref(it))
    return S, it.value


def stochastic_md_action(L, A, st):
    A = np.ascontiguousarray(A, dtype=np.float64); st = _cv(st)
    return lib().orc_stochastic_md_action(C.byref(L), _p(A), _p(st))


def gauge_step(L, A, mom, eps):
    A = np.array(A, dtype=np.float64, order="C")
    mom = np.ascontiguousarray(mom, dtype=np.float64)
    lib().orc_gauge_step(C.byref(L), _p(A), _p(mom), eps)
    return A


def momentum_step(L, mom, A, psi, st, eps, tol_abs_rr=0.0, tol_rel=1e-12, max_iter=100000):
    """Returns (new mom, force dS/dA, cg iterations)."""
    mom = np.array(mom, dtype=np.float64, order="C")
    A = np.ascontiguousarray(A, dtype=np.float64)
    psi = _cv(psi); st = _cv(st)
    F = np.empty_like(mom)
    it = lib().orc_momentum_step(C.byref(L), _p(mom), _p(A), _p(psi), _p(st), eps, tol_abs_rr,
                                 tol_rel, max_iter, _p(F))
    return mom, F, it


def update_gauge(L, rng: Rng, A, nsteps=10, traj_len=1.0, tol_abs_rr=0.0, tol_rel=1e-10,
                 max_iter=100000, phys_norm=0):
    """One trajectory; returns (accepted, dS, acts[8], cg_iters, A_new)."""
    A = np.array(A, dtype=np.float64, order="C")
    dS = D(); acts = np.empty(8); its = I()
    acc = lib().orc_update_gauge(C.byref(L), rng.ref, _p(A), nsteps, traj_len, tol_abs_rr,
                                 tol_rel, max_iter, phys_norm, C.byref(dS), _p(acts),
                                 C.byref(its))
    return acc, dS.value, acts, its.value, A


def puregauge_sweep(L, rng: Rng, A):
    A = np.array(A, dtype=np.float64, order="C")
    lib().orc_puregauge_sweep(C.byref(L), rng.ref, _p(A))
    return A


def hmc_chain(nt, nx, m, g, seed, ntraj, nsteps, burn=0, sweeps=100, tol_rel=1e-10):
    """A Markov chain as hmc.c's main runs it (hmc.c:927-939): A = 0,
    `sweeps` pure-gauge sweeps, then update_gauge trajectories; returns an
    (ntraj, 3) array of (accepted, dH, gauge action after the trajectory)
    for the trajectories after `burn`."""
    L = lattice(nt, nx, m=m, g=g)
    rng = Rng(seed, warmup=543210)
    A = np.zeros((nt * nx, 2))
    for _ in range(sweeps):
        A = puregauge_sweep(L, rng, A)
    out = []
    for k in range(burn + ntraj):
        acc, dS, _, _, A = update_gauge(L, rng, A, nsteps=nsteps, tol_rel=tol_rel)
        if k >= burn:
            out.append((acc, dS, gauge_action(L, A)))
    return np.array(out)


# ---------------------------------------------------------------- bag ---
def bag_lattice(nt, nx, bc=BC_ANTISYMMETRIC, with_mass_monomers=0, m=0.1, mu=0.0):
    return BagLattice(nt, nx, bc, with_mass_monomers, m, mu)


def _field(f):
    return np.ascontiguousarray(f, dtype=np.int32)


def fM_index(B, t1, x1, t2, x2, mu=None):
    return lib().orc_fM_index(C.byref(B), t1, x1, t2, x2, B.mu if mu is None else mu)


def fM(B, field, psi, transpose=False):
    field = _field(field); psi = np.ascontiguousarray(psi, dtype=np.float64)
    chi = np.empty_like(psi)
    f = lib().orc_fM_transpose if transpose else lib().orc_fM
    f(C.byref(B), _p(field), _p(chi), _p(psi))
    return chi


def cg_propagator(B, field, src, tol_abs_rr=1e-30, max_iter=10000):
    field = _field(field); src = np.ascontiguousarray(src, dtype=np.float64)
    x = np.empty_like(src); st = I()
    it = lib().orc_cg_propagator(C.byref(B), _p(field), _p(x), _p(src), tol_abs_rr, max_iter,
                                 C.byref(st))
    return x, it, st.value


def dlogdet(M):
    M = np.array(M, dtype=np.float64, order="C")
    la, sg = D(), I()
    info = lib().orc_dlogdet(M.shape[0], _p(M), C.byref(la), C.byref(sg))
    return la.value, sg.value, info


def dinverse(M):
    M = np.array(M, dtype=np.float64, order="C")
    info = lib().orc_dinverse(M.shape[0], _p(M))
    return M, info


def determinant_mu(B, field, mu=None):
    field = _field(field)
    la, sg = D(), I()
    n = lib().orc_determinant_mu(C.byref(B), _p(field), B.mu if mu is None else mu,
                                 C.byref(la), C.byref(sg))
    return la.value, sg.value, n


def reduced_matrix(B, field, mu=None):
    """The fM_index matrix restricted to the free sites (natural order)."""
    field = _field(field)
    n = int(np.count_nonzero(field == 0))
    M = np.empty(n * n, np.float64)
    lib().orc_reduced_matrix(C.byref(B), _p(field), B.mu if mu is None else mu, _p(M))
    return M.reshape(n, n)


def calc_Dinv(B):
    h = B.nt * B.nx // 2
    Dinv = np.empty(2 * h * h, np.float64)
    info = lib().orc_calc_Dinv(C.byref(B), _p(Dinv))
    if info:
        raise RuntimeError(f"orc_calc_Dinv: singular ({info})")
    return Dinv


def bag_weight(B, field, Dinv):
    field = _field(field); Dinv = np.ascontiguousarray(Dinv, dtype=np.float64)
    la, sg, n = D(), I(), I()
    r = lib().orc_bag_weight(C.byref(B), _p(field), _p(Dinv), C.byref(la), C.byref(sg),
                             C.byref(n))
    return la.value, sg.value, n.value, r


def brute_enumerate(B):
    V = B.nt * B.nx
    Z = np.zeros((V // 2 + 1, V + 1)); Za = np.zeros_like(Z)
    n = lib().orc_brute_enumerate(C.byref(B), _p(Z), _p(Za))
    return Z, Za, n


def worm(B, field, rng: Rng, U, n_attempts=20, n_links=None, n_monomers=0, sign=1):
    """The worm susceptibility (fermionbag.c:450-590); returns (steps, value,
    field, n_links, n_monomers, sign)."""
    f = np.array(field, dtype=np.int32)
    nl = I(int(np.count_nonzero((f == 2) | (f == 3))) if n_links is None else n_links)
    nm, sg, val = I(n_monomers), I(sign), D()
    steps = lib().orc_worm(C.byref(B), _p(f), rng.ref, U, n_attempts, C.byref(nl), C.byref(nm),
                           C.byref(sg), C.byref(val))
    return steps, val.value, f, nl.value, nm.value, sg.value
