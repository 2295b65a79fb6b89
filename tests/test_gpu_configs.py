"""GPU tests at the sizes of the benchmark configurations.

C2: the Dslash with 8 right-hand sides sharing the link reads (hmc.c:123-184
applied to 8 vectors): FP64 <= 1e-12 and FP32 <= 2e-6 relative to the
oracle.
C3: the fermion force at 128 x 128, m = 0.1, g = 0.8 (momentum_step,
hmc.c:504-665): FP64 solve (1e-14) against the oracle's force <= 1e-10;
central finite differences (h = 1e-6) on 64 random links; the mixed
(FP32 inner) solve at 1e-10 against the FP64 force, bound stated below.
C4: one batched mixed-precision trajectory of chains of sides
{32, 48, 64, 96, 128} against the FP64 batched path on identical inputs
(update_gauge, hmc.c:671-746), bound stated below.
C5: a 48 x 48 bag matrix (N = 2304) through about 100 rank-1 and rank-2
proposals and accepts (fluctuation_determinant.c:664-871): ratios and the
log-determinant against full FP64 LU recomputes <= 1e-9.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.fixture(scope="module")
def hmc():
    from thirring2d_amd import hmc as H
    return H


@pytest.fixture(scope="module")
def native():
    from thirring2d_amd import _native as N
    return N


# ------------------------------------------------------------------ C2 --
@pytest.mark.parametrize("nt,nx", [(256, 256), (16, 24)])
def test_dslash_eight_rhs(hmc, native, oracle, nt, nx):
    N = native
    V = nt * nx
    rng = np.random.default_rng(2002)
    A = rng.normal(size=(V, 2))
    lat = hmc.StaggeredLattice(nt, nx, m=0.05, g=0.5)
    lat.set_links(A)
    L = oracle.lattice(nt, nx, m=0.05, g=0.5)
    nrhs = 8
    vs = rng.normal(size=(nrhs, V)) + 1j * rng.normal(size=(nrhs, V))
    ctx = hmc.Context.get().ptr
    nat = hmc.DeviceArray.from_host(vs)
    ein = hmc.DeviceArray(16 * V * nrhs)
    eout = hmc.DeviceArray(16 * V * nrhs)
    e32i = hmc.DeviceArray(8 * V * nrhs)
    e32o = hmc.DeviceArray(8 * V * nrhs)
    back = hmc.DeviceArray(16 * V * nrhs)
    lib = N.lib()
    for k in range(nrhs):
        N.check(lib.th2d_vec_to_eo(ctx, C.byref(lat.geom), ein.p.value + 16 * V * k,
                                   nat.p.value + 16 * V * k), "to_eo")
    N.check(lib.th2d_vec_f64_to_f32(ctx, e32i.p, ein.p, V * nrhs), "f32")

    def natural(buf):
        for k in range(nrhs):
            N.check(lib.th2d_vec_from_eo(ctx, C.byref(lat.geom), back.p.value + 16 * V * k,
                                         buf.p.value + 16 * V * k), "from_eo")
        return back.download((nrhs, V), np.complex128)

    for dag in (0, 1):
        ref = np.stack([(oracle.fm_dag_mul if dag else oracle.fm_mul)(L, A, v) for v in vs])
        N.check(lib.th2d_dslash(ctx, C.byref(lat.geom), eout.p, ein.p, lat.K.p, dag,
                                N.PREC_F64, nrhs), "dslash f64")
        assert _rel(natural(eout), ref) <= 1e-12
        N.check(lib.th2d_dslash(ctx, C.byref(lat.geom), e32o.p, e32i.p, lat.K32.p, dag,
                                N.PREC_F32, nrhs), "dslash f32")
        N.check(lib.th2d_vec_f32_to_f64(ctx, eout.p, e32o.p, V * nrhs), "f64")
        assert _rel(natural(eout), ref) <= 2e-6


def test_dslash_batched_matches_oracle(hmc, native, oracle):
    """th2d_dslash_batched: M and M^dag of lattices of different sizes,
    masses and chemical potentials in one launch, each against the oracle's
    fm_mul / true adjoint (hmc.c:123-184) to 1e-12; calls with the same
    chains again reuse the uploaded table."""
    N = native
    rng = np.random.default_rng(77)
    dims = [(16, 16, 0.1, 0.0), (8, 24, 0.3, 0.2), (64, 32, 0.05, 0.0), (6, 6, 0.5, -0.1)]
    ctx = hmc.Context.get().ptr
    lib = N.lib()
    lats, orcs, As, vs, ein, eout = [], [], [], [], [], []
    for nt, nx, m, mu in dims:
        V = nt * nx
        A = rng.normal(size=(V, 2))
        lat = hmc.StaggeredLattice(nt, nx, m=m, g=0.5, mu=mu)
        lat.set_links(A)
        v = rng.normal(size=V) + 1j * rng.normal(size=V)
        e = hmc.DeviceArray(16 * V)
        lat.to_eo(v, e)
        lats.append(lat)
        orcs.append(oracle.lattice(nt, nx, m=m, g=0.5, mu=mu))
        As.append(A)
        vs.append(v)
        ein.append(e)
        eout.append(hmc.DeviceArray(16 * V))
    n = len(dims)
    geoms = (N.Geom * n)(*[l.geom for l in lats])
    P = C.c_void_p * n
    outs, ins, Ks = P(*[e.p for e in eout]), P(*[e.p for e in ein]), P(*[l.K.p for l in lats])
    for dag in (0, 1, 1, 0):
        N.check(lib.th2d_dslash_batched(ctx, n, geoms, outs, ins, Ks, dag), "dslash_batched")
        for k in range(n):
            f = oracle.fm_dag_mul if dag else oracle.fm_mul
            ref = f(orcs[k], As[k], vs[k])
            assert _rel(lats[k].from_eo(eout[k]), ref) <= 1e-12, (dims[k], dag)
    bad = P(*[e.p for e in ein])
    assert lib.th2d_dslash_batched(ctx, n, geoms, bad, ins, Ks, 0) == N.TH2D_EARG


# ------------------------------------------------------------------ C3 --
C3 = dict(nt=128, nx=128, m=0.1, g=0.8)


@pytest.fixture(scope="module")
def c3(hmc, oracle):
    nt, nx = C3["nt"], C3["nx"]
    V = nt * nx
    A = np.random.default_rng(1003).normal(size=(V, 2))
    r = np.random.default_rng(2003)
    xi = r.normal(size=V) + 1j * r.normal(size=V)
    L = oracle.lattice(nt, nx, m=C3["m"], g=C3["g"])
    psi = oracle.fm_dag_mul(L, A, xi)
    lat = hmc.StaggeredLattice(nt, nx, m=C3["m"], g=C3["g"])
    lat.set_links(A)
    return lat, L, A, psi


def test_c3_force_matches_oracle(hmc, oracle, c3):
    """All 32,768 links: FP64 solve to 1e-14, force <= 1e-10 relative to the
    oracle's momentum_step force."""
    lat, L, A, psi = c3
    V = L.nt * L.nx
    _, F_ref, _ = oracle.momentum_step(L, np.zeros((V, 2)), A, psi, np.zeros(V, complex), 0.1,
                                       tol_rel=1e-14)
    chi, info = lat.fmdm_invert_cg(psi, hmc.cg_params(tol_rel=1e-14))
    assert info.status == 0
    F = lat.force(chi)
    assert _rel(F, F_ref) <= 1e-10


def test_c3_force_finite_differences(hmc, oracle, c3):
    """Central differences (h = 1e-6) of S = S_g + psi^dag (M^dag M)^-1 psi on
    64 random links (seed 3003).  S_pf is evaluated in the stationary form
    2 Re psi^dag x - |M x|^2, whose error is second order in the solver's
    residual, with x from the device FP64 solve at 1e-13 and M applied by the
    oracle; S_g by the oracle.  Bound: 1e-6 max(1, |F|) + 2e-5 (the round-off
    of S ~ 1e5 over 2h)."""
    lat, L, A, psi = c3
    V = L.nt * L.nx
    chi, _ = lat.fmdm_invert_cg(psi, hmc.cg_params(tol_rel=1e-14))
    F = lat.force(chi)
    pick = np.random.default_rng(3003)
    sites = pick.integers(0, V, size=64)
    dirs = pick.integers(0, 2, size=64)
    p = hmc.cg_params(tol_rel=1e-13)
    h = 1e-6

    def action(Ap):
        lat.set_links(Ap)
        x, info = lat.fmdm_invert_cg(psi, p)
        assert info.status == 0
        Mx = oracle.fm_mul(L, Ap, x)
        return 2 * np.vdot(psi, x).real - np.vdot(Mx, Mx).real + oracle.gauge_action(L, Ap)

    try:
        for s, d in zip(sites, dirs):
            Sp = []
            for sg in (1, -1):
                Ap = A.copy()
                Ap[s, d] += sg * h
                Sp.append(action(Ap))
            fd = (Sp[0] - Sp[1]) / (2 * h)
            assert abs(F[s, d] - fd) <= 1e-6 * max(1.0, abs(fd)) + 2e-5, (s, d, F[s, d], fd)
    finally:
        lat.set_links(A)


def test_c3_force_mixed_solve(hmc, native, c3):
    """The force from the mixed-precision solve (FP32 inner CG, FP64 true
    residual <= 1e-10) against the FP64 one (1e-14).  The solution's
    relative error is at most kappa * 1e-10 with kappa = (m^2 + 4)/m^2 = 401,
    and the force is bilinear in chi and M chi: bound 2 kappa 1e-10 = 8e-8
    (max |dF| / max |F|)."""
    lat, L, A, psi = c3
    chi64, i64 = lat.fmdm_invert_cg(psi, hmc.cg_params(tol_rel=1e-14))
    chi32, i32 = lat.fmdm_invert_cg(psi, hmc.cg_params(tol_rel=1e-10,
                                                       precision=native.PREC_MIXED))
    assert i64.status == 0 and i32.status == 0 and i32.true_rel_resid <= 1e-10
    F64, F32 = lat.force(chi64), lat.force(chi32)
    kappa = (C3["m"] ** 2 + 4) / C3["m"] ** 2
    err = _rel(F32, F64)
    assert err <= 2 * kappa * 1e-10, err


# ------------------------------------------------------------------ C4 --
def test_c4_batched_mixed_trajectory_vs_fp64(hmc, native):
    """Ten chains, two each of sides 32, 48, 64, 96, 128, g uniform in
    [0.2, 1.0] (seed 4004), m = 0.1, A ~ N(0,1), device noise: one batched
    trajectory of 40 x dt = 0.025 with mixed-precision solves (the 96^2
    chains on the step kernel, the others resident) and one with FP64 solves
    on identical inputs.  Each solve stops at a true residual tau = 1e-10;
    the action psi^dag (M^dag M)^-1 psi = |xi|^2 is then off by at most
    |M^-1 xi| |r| <= tau |xi|^2 sqrt(m^2 + 4) / m, and dH holds two such
    actions: per chain |dH_mixed - dH_fp64| <= 4 tau S_pf0 sqrt(m^2+4)/m
    + 1e-8 |dH| (S_pf0 = |xi|^2)."""
    N = native
    rng = np.random.default_rng(4004)
    sides = [32, 48, 64, 96, 128] * 2
    gs = rng.uniform(0.2, 1.0, size=len(sides))
    m, tau = 0.1, 1e-10
    out = {}
    for prec in (N.PREC_F64, N.PREC_MIXED):
        lats = []
        for k, (L_, g) in enumerate(zip(sides, gs)):
            lat = hmc.StaggeredLattice(L_, L_, m=m, g=float(g))
            lat.random_links(1004 + k)
            lats.append(lat)
        batch = hmc.ChainBatch(lats)
        res = batch.update_gauge(nsteps=40, traj_len=1.0, seed=4004,
                                 counters=list(range(len(sides))),
                                 params=hmc.cg_params(tol_rel=tau, precision=prec))
        for r in res:
            assert r.status == 0 and r.max_true_rel_resid <= tau
        out[prec] = res
    for r64, r32, L_ in zip(out[N.PREC_F64], out[N.PREC_MIXED], sides):
        assert np.isfinite(r64.dS)
        assert r64.acts[1] == r32.acts[1]  # same noise
        bound = 4 * tau * r64.acts[1] * np.sqrt(m * m + 4) / m + 1e-8 * abs(r64.dS)
        assert abs(r32.dS - r64.dS) <= bound, (L_, r32.dS, r64.dS, bound)


# ------------------------------------------------------------------ C5 --
def _bag_field(nt, nx, rho, rng):
    V = nt * nx
    f = np.zeros(V, np.int32)
    for s in rng.permutation(V):
        if f[s] or rng.random() > rho:
            continue
        t, x = divmod(int(s), nx)
        d = int(rng.integers(2))
        t2, x2 = ((t + 1) % nt, x) if d == 0 else (t, (x + 1) % nx)
        s2 = t2 * nx + x2
        if not f[s2]:
            f[s], f[s2] = 2 + d, 4 + d
    free = np.flatnonzero(f == 0)
    for s in rng.choice(free, size=6, replace=False):
        f[s] = 1
    return f


def _c5_propose(rng, f, nt, nx):
    """Rank-1 (add or remove a monomer) and rank-2 (add or remove a link,
    move a monomer) changes, as the reference's worm and link moves make
    them (fermionbag.c:284-379)."""
    V = nt * nx
    s = int(rng.integers(V))
    kind = rng.random()
    if kind < 0.3:
        if f[s] == 0:
            return [s], [1], []
        if f[s] == 1:
            return [], [], [s]
        return None
    t, x = divmod(s, nx)
    d = int(rng.integers(2))
    t2, x2 = ((t + 1) % nt, x) if d == 0 else (t, (x + 1) % nx)
    s2 = t2 * nx + x2
    if f[s] == 0 and f[s2] == 0:
        return [s, s2], [2 + d, 4 + d], []
    if f[s] == 2 + d and f[s2] == 4 + d:
        return [], [], [s, s2]
    if f[s] == 0 and f[s2] == 1:
        return [s], [1], [s2]
    return None


def test_c5_rank_updates_at_2304(oracle):
    """L = 48 (N = 2304), m = 0.1, occupation about 0.3: B = (D_FF)^-1 on
    the device through about 100 proposals.  The initial and final
    log|det D_FF| agree with an FP64 LU of the oracle's reduced matrix
    (numpy slogdet) to 1e-9 relative; 12 sampled ratios agree with the ratio
    of two such full recomputes to 1e-9 relative; the accepted ratios
    multiply up to the change of the determinant."""
    from thirring2d_amd import bag as Bm
    nt = nx = 48
    m = 0.1
    rng = np.random.default_rng(5005)
    f = _bag_field(nt, nx, 0.3, rng)
    lat = Bm.BagLattice(nt, nx, m=m)
    B = oracle.bag_lattice(nt, nx, m=m)
    st = Bm.BagState(lat)
    st.set_field(f)

    def full(field):
        s, l = np.linalg.slogdet(oracle.reduced_matrix(B, field))
        return l, int(s)

    l0, s0 = full(f)
    la, sg, _, _ = st.logdet()
    assert sg == s0 and abs(la - l0) <= 1e-9 * abs(l0)
    tried = accepted = checked = 0
    log_acc, sign_acc = 0.0, 1
    while tried < 100:
        prop = _c5_propose(rng, f, nt, nx)
        if prop is None:
            continue
        occ, vals, fr = prop
        r = st.ratio(occ, fr)
        tried += 1
        fn = f.copy()
        fn[fr] = 0
        fn[occ] = vals
        if checked < 12 and tried % 8 == 0:
            lo, so = full(f)
            ln, sn = full(fn)
            want = so * sn * np.exp(ln - lo)
            assert abs(r - want) <= 1e-9 * abs(want), (occ, fr, r, want)
            checked += 1
        if r != 0.0 and rng.random() < min(1.0, abs(r)):
            st.accept(occ, vals, fr)
            f = fn
            accepted += 1
            log_acc += np.log(abs(r))
            sign_acc *= int(np.sign(r))
    assert checked == 12 and accepted >= 30
    l1, s1 = full(f)
    la, sg, _, _ = st.logdet()
    assert sg == s1 and abs(la - l1) <= 1e-9 * abs(l1)
    assert abs((l1 - l0) - log_acc) <= 1e-9 * abs(l1) and s1 * s0 == sign_acc
