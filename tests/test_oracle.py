"""Pins of the CPU oracle (no GPU).

Each check is independent of the oracle's own formulas: the reference's
compiled RNG, the brute-force values the reference's program printed, closed forms,
identities of the operator and finite differences.
"""
import json
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def test_mt_matches_reference_build(oracle):
    """MT19937 stream equals the reference's mersenne_inline.c, compiled
    unmodified into oracle/_ref (bitwise), for even, odd and negative seeds."""
    if oracle.ref_mt() is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    for seed in (2001, 1001, 54357345, -7, 0):
        a = oracle.Rng(seed).draws(1500)
        b = oracle.ref_mt_draws(seed, 1500)
        np.testing.assert_array_equal(a, b)


def test_mt_golden(oracle):
    """Draws after the reference's warm-up of 543210 (hmc.c:877) against the
    committed vectors written from the reference build."""
    with open(os.path.join(HERE, "golden", "mt_draws.json")) as f:
        gold = json.load(f)
    for seed, draws in gold["draws"].items():
        r = oracle.Rng(int(seed), warmup=gold["warmup"])
        np.testing.assert_array_equal(r.draws(len(draws)), np.array(draws))


def _field(nt, nx, rng, A=True):
    V = nt * nx
    return rng.normal(size=(V, 2)) if A else None


def test_adjoint_identity(oracle):
    rng = np.random.default_rng(0)
    for mu in (0.0, 0.3):
        L = oracle.lattice(8, 12, m=0.1, mu=mu)
        A = rng.normal(size=(96, 2))
        v = rng.normal(size=96) + 1j * rng.normal(size=96)
        w = rng.normal(size=96) + 1j * rng.normal(size=96)
        lhs = np.vdot(w, oracle.fm_mul(L, A, v))
        rhs = np.vdot(oracle.fm_dag_mul(L, A, w), v)
        assert abs(lhs - rhs) < 1e-12 * abs(lhs)


def test_adjoint_is_2m_minus_M_at_mu0(oracle):
    rng = np.random.default_rng(1)
    L = oracle.lattice(16, 16, m=0.07)
    A = rng.normal(size=(256, 2))
    v = rng.normal(size=256) + 1j * rng.normal(size=256)
    d = oracle.fm_dag_mul(L, A, v) - (2 * L.m * v - oracle.fm_mul(L, A, v))
    assert np.abs(d).max() < 1e-14


def test_dense_matrix_matches_apply(oracle):
    rng = np.random.default_rng(2)
    L = oracle.lattice(8, 8, m=0.2, mu=0.1)
    A = rng.normal(size=(64, 2))
    v = rng.normal(size=64) + 1j * rng.normal(size=64)
    M = oracle.fermion_matrix(L, A)
    assert np.abs(M @ v - oracle.fm_mul(L, A, v)).max() < 1e-13
    assert np.abs(M.conj().T @ v - oracle.fm_dag_mul(L, A, v)).max() < 1e-13


@pytest.mark.parametrize("m", [0.05, 0.1, 0.5])
def test_free_field_logdet(oracle, m):
    """A = 0, mu = 0, antiperiodic t and x: log|det M| =
    1/2 sum_p log(m^2 + sin^2 p_t + sin^2 p_x), p = 2 pi (n + 1/2) / L."""
    n = 16
    L = oracle.lattice(n, n, m=m)
    la, _, info = oracle.zlogdet(oracle.fermion_matrix(L, np.zeros((n * n, 2))))
    p = 2 * np.pi * (np.arange(n) + 0.5) / n
    cf = 0.5 * np.sum(np.log(m * m + np.sin(p)[:, None] ** 2 + np.sin(p)[None, :] ** 2))
    assert info == 0 and abs(la - cf) < 1e-12 * abs(cf) + 1e-12


def test_gauge_covariance(oracle):
    """A_mu(x) -> A_mu(x) + phi(x) - phi(x+mu), psi -> e^{i phi} psi leaves
    M covariant: M' (e^{i phi} v) = e^{i phi} (M v)."""
    rng = np.random.default_rng(3)
    nt = nx = 8
    L = oracle.lattice(nt, nx, m=0.3)
    A = rng.normal(size=(64, 2))
    phi = rng.normal(size=64)
    A2 = A.copy()
    for t in range(nt):
        for x in range(nx):
            s = t * nx + x
            A2[s, 0] += phi[s] - phi[((t + 1) % nt) * nx + x]
            A2[s, 1] += phi[s] - phi[t * nx + (x + 1) % nx]
    v = rng.normal(size=64) + 1j * rng.normal(size=64)
    lhs = oracle.fm_mul(L, A2, np.exp(1j * phi) * v)
    rhs = np.exp(1j * phi) * oracle.fm_mul(L, A, v)
    assert np.abs(lhs - rhs).max() < 1e-13


def test_cg_true_residual_and_fourier(oracle):
    rng = np.random.default_rng(4)
    L = oracle.lattice(8, 8, m=0.1)
    A = rng.normal(size=(64, 2))
    b = rng.normal(size=64) + 1j * rng.normal(size=64)
    x, it, st = oracle.cg_mdm(L, A, b, tol_rel=1e-13)
    r = b - oracle.fm_dag_mul(L, A, oracle.fm_mul(L, A, x))
    assert st == 0 and np.linalg.norm(r) <= 1e-13 * np.linalg.norm(b)
    # A = 0: M^dag M is diagonal in momentum space with m^2 + sin^2 + sin^2
    # on antiperiodic momenta; check against the dense solve
    L0 = oracle.lattice(8, 8, m=0.1)
    M = oracle.fermion_matrix(L0, np.zeros((64, 2)))
    x0, _, _ = oracle.cg_mdm(L0, np.zeros((64, 2)), b, tol_rel=1e-14)
    assert np.abs(x0 - np.linalg.solve(M.conj().T @ M, b)).max() < 1e-11
    ev = np.linalg.eigvalsh(M.conj().T @ M)
    p = 2 * np.pi * (np.arange(8) + 0.5) / 8
    cf = np.sort((0.01 + np.sin(p)[:, None] ** 2 + np.sin(p)[None, :] ** 2).ravel())
    np.testing.assert_allclose(ev, cf, atol=1e-12)


def test_force_finite_differences(oracle):
    rng = np.random.default_rng(5)
    L = oracle.lattice(8, 8, m=0.1, g=0.8, mu=0.2)
    A = rng.normal(size=(64, 2))
    w = rng.normal(size=64) + 1j * rng.normal(size=64)
    st = rng.normal(size=64) + 1j * rng.normal(size=64)
    psi = oracle.fm_dag_mul(L, A, w)
    _, F, _ = oracle.momentum_step(L, np.zeros((64, 2)), A, psi, st, 0.1, tol_rel=1e-14)

    def S(Ap):
        return (oracle.pseudofermion_action(L, Ap, psi, tol_rel=1e-14)[0]
                + oracle.gauge_action(L, Ap) + oracle.stochastic_md_action(L, Ap, st))
    h = 1e-6
    for s, d in [(0, 0), (7, 1), (63, 0), (56, 0), (63, 1), (20, 1)]:
        Ap, Am = A.copy(), A.copy()
        Ap[s, d] += h
        Am[s, d] -= h
        fd = (S(Ap) - S(Am)) / (2 * h)
        assert abs(F[s, d] - fd) <= 1e-6 * max(1.0, abs(fd))


def test_pseudofermion_action_equals_noise_norm(oracle):
    """psi = M^dag xi  =>  psi^dag (M^dag M)^{-1} psi = |xi|^2."""
    L = oracle.lattice(8, 8, m=0.1)
    A = np.random.default_rng(6).normal(size=(64, 2))
    rng = oracle.Rng(99)
    S0, xi, psi = oracle.random_pseudofermion(L, rng, A)
    S1, _ = oracle.pseudofermion_action(L, A, psi, tol_rel=1e-14)
    assert abs(S0 - S1) < 1e-10 * S0


def test_leapfrog_reversible_and_second_order(oracle):
    """The MD integrator: dH shrinks ~ dt^2 (hmc.c:708-716)."""
    L = oracle.lattice(8, 8, m=0.5, g=1.0)
    A0 = np.random.default_rng(7).normal(size=(64, 2)) * 0.3
    dH = []
    for ns in (20, 40):
        rng = oracle.Rng(5)
        _, d, _, _, _ = oracle.update_gauge(L, rng, A0, nsteps=ns, tol_rel=1e-13)
        dH.append(abs(d))
    assert dH[1] < dH[0] / 2.5


def test_fM_transpose_is_transpose(oracle):
    rng = np.random.default_rng(8)
    for bc in (oracle.BC_ANTISYMMETRIC, oracle.BC_OPENX):
        B = oracle.bag_lattice(6, 8, bc=bc, m=0.2, mu=0.3)
        field = (rng.random(48) < 0.3).astype(np.int32)
        u, v = rng.normal(size=48), rng.normal(size=48)
        assert abs(u @ oracle.fM(B, field, v) - v @ oracle.fM(B, field, u, transpose=True)) < 1e-12


def test_fM_matches_fM_index(oracle):
    """Masked apply = identity rows for occupied sites, fM_index elsewhere."""
    rng = np.random.default_rng(9)
    for bc in (oracle.BC_ANTISYMMETRIC, oracle.BC_OPENX):
        B = oracle.bag_lattice(6, 6, bc=bc, m=0.2, mu=0.2)
        field = (rng.random(36) < 0.3).astype(np.int32)
        M = np.zeros((36, 36))
        for i in range(36):
            for j in range(36):
                if field[i] or field[j]:
                    M[i, j] = float(i == j)
                else:
                    M[i, j] = oracle.fM_index(B, i // 6, i % 6, j // 6, j % 6)
        v = rng.normal(size=36)
        assert np.abs(M @ v - oracle.fM(B, field, v)).max() < 1e-14


def test_cg_propagator_solves(oracle):
    rng = np.random.default_rng(10)
    B = oracle.bag_lattice(8, 8, m=0.3)
    field = np.zeros(64, np.int32)
    field[[3, 4, 20, 21]] = 1
    src = rng.normal(size=64)
    x, it, st = oracle.cg_propagator(B, field, src, tol_abs_rr=1e-28)
    assert st == 0
    assert np.abs(oracle.fM(B, field, x) - src).max() < 1e-12


def test_dlogdet_against_numpy(oracle):
    rng = np.random.default_rng(11)
    for n in (1, 5, 40):
        M = rng.normal(size=(n, n))
        la, sg, info = oracle.dlogdet(M)
        s2, l2 = np.linalg.slogdet(M)
        assert info == 0 and sg == s2 and abs(la - l2) < 1e-12 * max(1, abs(l2))


def test_calc_dinv_blocks(oracle):
    B = oracle.bag_lattice(6, 6, with_mass_monomers=1, m=0.0, mu=0.1)
    D = np.array([[oracle.fM_index(B, i // 6, i % 6, j // 6, j % 6) for j in range(36)]
                  for i in range(36)])
    X = np.linalg.inv(D)
    Dinv = oracle.calc_Dinv(B)
    h = 18
    for i in range(36):
        for j in range(36):
            pi, pj = (i // 6 + i % 6) % 2, (j // 6 + j % 6) % 2
            if pi == 0 and pj == 1:
                assert abs(Dinv[(i // 2) * h + j // 2] - X[i, j]) < 1e-12
            if pi == 1 and pj == 0:
                assert abs(Dinv[h * h + (i // 2) * h + j // 2] - X[i, j]) < 1e-12


def test_bag_weight_is_jacobi_ratio(oracle):
    """det_e det_o of the occupied-site propagator blocks
    (fluctuation_determinant.c:203-237) equals det(D_FF) / det(D)
    (determinant_mu over the free sites, full_determinant.c:144-186)."""
    rng = np.random.default_rng(12)
    B = oracle.bag_lattice(6, 6, with_mass_monomers=1, m=0.0, mu=0.1)
    Dinv = oracle.calc_Dinv(B)
    laD, sD, _ = oracle.determinant_mu(B, np.zeros(36, np.int32))
    checked = 0
    for _ in range(200):
        field = np.zeros(36, np.int32)
        # random pairs of neighbouring sites (one even, one odd) occupied
        for _ in range(rng.integers(1, 5)):
            s = int(rng.integers(36))
            t, x = divmod(s, 6)
            s2 = t * 6 + (x + 1) % 6
            if field[s] == 0 and field[s2] == 0:
                field[s] = field[s2] = 1
        la, sg, n, r = oracle.bag_weight(B, field, Dinv)
        lf, sf, _ = oracle.determinant_mu(B, field)
        if r != 0 or sg == 0 or sf == 0:
            continue
        assert sg == sf * sD
        assert abs((la - (lf - laD))) < 1e-10
        checked += 1
    assert checked > 50


def test_brute_table_matches_reference_run(oracle):
    """Exact enumeration on 4x4 OPENX, m=0, U=1, mu=0.1 (tests/parameter):
    the values the reference's tests/thirring_brute.c printed when it was
    compiled and run on this lattice: Z(0,0) = 1, Z(0,2) = 9.96905 (|Z| 10.0567),
    Z(1,0) = 9.3976, Z(2,2) = 629.076, 45 non-zero cells."""
    B = oracle.bag_lattice(4, 4, bc=oracle.BC_OPENX, with_mass_monomers=1, m=0.0, mu=0.1)
    Z, Za, n = oracle.brute_enumerate(B)
    assert Z[0, 0] == 1.0
    assert f"{Z[0, 2]:g}" == "9.96905" and f"{Za[0, 2]:g}" == "10.0567"
    assert f"{Z[1, 0]:g}" == "9.3976"
    assert f"{Z[2, 2]:g}" == "629.076"
    assert int((Z != 0).sum()) == 45


def test_bag_matrix_is_hmc_matrix_at_zero_field(oracle):
    """fM_index on an antisymmetric lattice (fermion_matrix.c:28-64) is the
    HMC operator of hmc.c:269-310 at A = 0, whose log|det| the free-field
    closed form pins (test_free_field_logdet)."""
    nt, nx, m, mu = 6, 8, 0.3, 0.2
    B = oracle.bag_lattice(nt, nx, bc=oracle.BC_ANTISYMMETRIC, m=m, mu=mu)
    V = nt * nx
    D = np.array([[oracle.fM_index(B, i // nx, i % nx, j // nx, j % nx) for j in range(V)]
                  for i in range(V)])
    L = oracle.lattice(nt, nx, m=m, mu=mu)
    M = oracle.fermion_matrix(L, np.zeros((V, 2)))
    assert np.max(np.abs(M.imag)) == 0.0
    assert np.max(np.abs(D - M.real)) <= 1e-15


def test_puregauge_sweep_samples_link_distribution(oracle):
    """Metropolis with uniform proposals samples exp(beta cos A) per link
    (hmc.c:82-93), beta = Nf/g: <cos A> = I1(beta)/I0(beta)."""
    from scipy.special import i0, i1
    nt, nx, g = 8, 8, 0.8
    L = oracle.lattice(nt, nx, g=g)
    rng = oracle.Rng(77)
    A = np.zeros((nt * nx, 2))
    acc = []
    for k in range(600):
        A = oracle.puregauge_sweep(L, rng, A)
        if k >= 100:
            acc.append(np.mean(np.cos(A)))
    beta = 2.0 / g
    want = i1(beta) / i0(beta)
    err = np.std(acc) / np.sqrt(len(acc) / 5.0)  # sweeps are correlated
    assert abs(np.mean(acc) - want) <= 5 * err + 1e-3


def test_gauss_pair_moments(oracle):
    """orc_gauss_pair (Box-Muller over MT pairs, hmc.c:424-429) draws two
    independent N(0,1): over 2e5 pairs the mean, variance, fourth moment and
    the covariance of the two outputs lie within 5 sigma of N(0,1)'s, and a
    Kolmogorov-Smirnov test against the normal CDF passes.  A slip such as
    sqrt(-log x1) (variance 1/2) or a shared angle fails it."""
    from scipy import stats
    n = 200_000
    z = oracle.Rng(2001, warmup=543210).gauss_pairs(n)
    for k in range(2):
        x = z[:, k]
        assert abs(x.mean()) < 5 / np.sqrt(n)
        assert abs((x * x).mean() - 1) < 5 * np.sqrt(2 / n)
        assert abs((x ** 4).mean() - 3) < 5 * np.sqrt(96 / n)
        assert stats.kstest(x, "norm").pvalue > 1e-4
    assert abs((z[:, 0] * z[:, 1]).mean()) < 5 / np.sqrt(n)
    # |z|^2 of a pair is exponential with mean 2 (the reference's S = |xi|^2)
    r2 = (z * z).sum(axis=1)
    assert abs(r2.mean() - 2) < 5 * 2 / np.sqrt(n)


def test_zlogdet_phase_matches_numpy(oracle):
    """orc_zlogdet (determinant, hmc.c:314-336, as log|det| and arg det)
    against numpy's complex slogdet: the modulus and the phase, on random
    complex matrices and on the fermion matrix at A != 0 and mu != 0."""
    rng = np.random.default_rng(12)
    mats = [rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)) for n in (1, 2, 5, 17, 40)]
    for nt, nx, mu in ((4, 4, 0.0), (6, 8, 0.3), (8, 8, -0.2)):
        L = oracle.lattice(nt, nx, m=0.2, mu=mu)
        mats.append(oracle.fermion_matrix(L, rng.normal(size=(nt * nx, 2))))
    for M in mats:
        la, ph, info = oracle.zlogdet(M)
        s, l = np.linalg.slogdet(M)
        assert info == 0
        assert abs(la - l) <= 1e-11 * max(1.0, abs(l))
        assert abs(np.angle(s * np.exp(-1j * ph))) <= 1e-10
    # a permutation with one transposition has det -1: phase pi
    P = np.eye(6, dtype=complex)[[1, 0, 2, 3, 4, 5]]
    la, ph, info = oracle.zlogdet(P)
    assert info == 0 and abs(la) < 1e-15 and abs(abs(ph) - np.pi) < 1e-15


# Random123's published known-answer vectors for philox4x32 with 10 rounds
# (counter words, key words, output words).
PHILOX_KAT = [
    ([0x00000000, 0x00000000, 0x00000000, 0x00000000], [0x00000000, 0x00000000],
     [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    ([0xffffffff, 0xffffffff, 0xffffffff, 0xffffffff], [0xffffffff, 0xffffffff],
     [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
]


def test_philox_known_answers(oracle):
    ck = np.array([c + k for c, k, _ in PHILOX_KAT], dtype=np.uint32)
    out = oracle.philox4x32_10(ck)
    np.testing.assert_array_equal(out, np.array([o for _, _, o in PHILOX_KAT], np.uint32))


def test_philox_gauss_moments(oracle):
    """The device noise mapping (one Philox block -> two 53-bit uniforms ->
    Box-Muller) restated in the oracle: N(0,1) moments within 5 sigma and a
    KS test, Re and Im uncorrelated, streams and counters independent."""
    from scipy import stats
    n = 100_000
    z = oracle.philox_gauss(1002, 1, n)
    w = oracle.philox_gauss(1002, 2, n)
    for x in (z.real, z.imag, w.real):
        assert abs(x.mean()) < 5 / np.sqrt(n)
        assert abs((x * x).mean() - 1) < 5 * np.sqrt(2 / n)
        assert abs((x ** 4).mean() - 3) < 5 * np.sqrt(96 / n)
        assert stats.kstest(x, "norm").pvalue > 1e-4
    assert abs((z.real * z.imag).mean()) < 5 / np.sqrt(n)
    assert abs((z.real * w.real).mean()) < 5 / np.sqrt(n)
    z1 = oracle.philox_gauss(1002, 1, 1000, ctr=1)
    assert abs(np.mean(z[:1000].real * z1.real)) < 5 / np.sqrt(1000)


def test_reduced_matrix_matches_fM(oracle):
    """orc_reduced_matrix: column j is fM applied to the unit vector of free
    site j, read at the free sites (vec_ops.c:96-131 drops hops into
    occupied sites, so the free block of fM is the reduced matrix)."""
    rng = np.random.default_rng(31)
    for nt, nx, bc, wmm in ((4, 6, 0, 0), (6, 6, 1, 1), (8, 4, 0, 1)):
        B = oracle.bag_lattice(nt, nx, bc=bc, with_mass_monomers=wmm, m=0.3, mu=0.2)
        V = nt * nx
        f = np.zeros(V, np.int32)
        f[rng.choice(V, size=V // 4, replace=False)] = 1
        free = np.flatnonzero(f == 0)
        M = oracle.reduced_matrix(B, f)
        for j, s in enumerate(free):
            e = np.zeros(V)
            e[s] = 1.0
            assert np.array_equal(oracle.fM(B, f, e)[free], M[:, j])
        la, sg, n = oracle.determinant_mu(B, f)
        s_np, l_np = np.linalg.slogdet(M)
        assert n == len(free) and sg == int(s_np)
        if s_np != 0:
            assert abs(la - l_np) < 1e-12 * max(1, abs(l_np))


def test_philox_uniform_moments(oracle):
    """The device uniform pairs: mean 1/2, variance 1/12, strictly inside
    (0,1), the two members uncorrelated, a KS test against U(0,1)."""
    from scipy import stats
    n = 100_000
    u = oracle.philox_uniform(77, 0x50475357, n, ctr=3)
    assert u.min() > 0.0 and u.max() < 1.0
    for k in range(2):
        assert abs(u[:, k].mean() - 0.5) < 5 * np.sqrt(1 / 12 / n)
        assert abs(u[:, k].var() - 1 / 12) < 5 * np.sqrt(1 / 180 / n)
        assert stats.kstest(u[:, k], "uniform").pvalue > 1e-4
    assert abs(np.corrcoef(u[:, 0], u[:, 1])[0, 1]) < 5 / np.sqrt(n)


def _bag_config(nt, nx, rng, rho=0.35, monomer_pairs=2):
    V = nt * nx
    f = np.zeros(V, np.int32)
    for s in rng.permutation(V):
        if f[s] or rng.random() > rho:
            continue
        t, x = divmod(int(s), nx)
        d = int(rng.integers(2))
        s2 = ((t + 1) % nt) * nx + x if d == 0 else t * nx + (x + 1) % nx
        if not f[s2]:
            f[s], f[s2] = 2 + d, 4 + d
    for _ in range(monomer_pairs):
        free = [s for s in range(V) if f[s] == 0 and f[((s // nx + 1) % nt) * nx + s % nx] == 0]
        if free:
            s = free[int(rng.integers(len(free)))]
            f[s] = f[((s // nx + 1) % nt) * nx + s % nx] = 1
    return f


def _legal(f, nt, nx):
    """Every link value has its partner (tail 2+d at s, head 4+d at s+d)."""
    V = nt * nx
    for s in range(V):
        t, x = divmod(s, nx)
        v = f[s]
        if v in (2, 3):
            d = v - 2
            s2 = ((t + 1) % nt) * nx + x if d == 0 else t * nx + (x + 1) % nx
            if f[s2] != v + 2:
                return False
        elif v in (4, 5):
            d = v - 4
            s2 = ((t - 1) % nt) * nx + x if d == 0 else t * nx + (x - 1) % nx
            if f[s2] != v - 2:
                return False
        elif v not in (0, 1):
            return False
    return True


@pytest.mark.parametrize("wmm", [1, 0])
def test_worm_invariants(oracle, wmm):
    """orc_worm (fermionbag.c:450-590): after each measurement every link has
    its two ends and no source monomer is left (every worm closes), the
    monomers are conserved, n_links counts the links, the sign equals the
    sign of det(D_FF) det(D) of the new configuration, and the occupied set
    only changes through accepted link insertions and removals (their
    weights U det ratios)."""
    rng = np.random.default_rng(77)
    nt = nx = 6
    V = nt * nx
    B = oracle.bag_lattice(nt, nx, with_mass_monomers=wmm, m=0.0 if wmm else 0.3)
    la_d, sg_d, _ = oracle.determinant_mu(B, np.zeros(V, np.int32))
    f = _bag_config(nt, nx, rng)
    r = oracle.Rng(55)
    _, s0, _ = oracle.determinant_mu(B, f)
    sign = s0 * sg_d
    nm0 = int(np.count_nonzero(f == 1))
    changed = False
    for _ in range(6):
        steps, val, g, nl, nm, sign = oracle.worm(B, f, r, U=1.3, n_attempts=5,
                                                  n_monomers=nm0, sign=sign)
        assert _legal(g, nt, nx)
        assert nl == np.count_nonzero((g == 2) | (g == 3)) and nm == nm0
        assert np.count_nonzero(g == 1) == nm0
        _, sg, _ = oracle.determinant_mu(B, g)
        assert sign == sg * sg_d
        assert steps != 0
        changed |= not np.array_equal(g != 0, f != 0)
        f = g
    # occupations change only with mass monomers (new links / removals)
    assert changed == bool(wmm)
