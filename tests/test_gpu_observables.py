"""Chain observables of the device HMC (mixed precision, device noise)
against the oracle's chains (the reference's update_gauge, MT noise), at
16 x 16, m = 0.1, g = 0.5, 40 x dt = 0.025 (hmc.c:671-746, 927-939).

Eight independent chains on each side, each started from A = 0 with 100
pure-gauge sweeps, 10 trajectories of burn-in, then 30 measured: the
acceptance rate, <S_g> and <exp(-dH)> (per-chain means, so autocorrelation
within a chain does not enter the error) agree within 4 combined standard
errors.  The oracle chains run in 8 CPU processes.

The reference draws momenta and pseudofermion noise with twice the variance
its Hamiltonian implies (p ~ N(0,1) with H_kin = sum p^2, hmc.c:490-493 and
723; Re, Im xi ~ N(0,1) with S = |xi|^2, hmc.c:425-429), so its chains need
not satisfy Creutz's identity <exp(-dH)> = 1.  With the physical
normalisation (phys_norm = 1: variances 1/2) the device chains do, within
5 standard errors."""
import multiprocessing as mp

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NT = NX = 16
M, G, NSTEPS, BURN, NTRAJ, NCH = 0.1, 0.5, 40, 10, 30, 8


def _oracle_chain(seed):
    import oracle as O
    return O.hmc_chain(NT, NX, M, G, seed, NTRAJ, NSTEPS, burn=BURN)


def _compare(a, b):
    ma, mb = a.mean(), b.mean()
    err = np.sqrt(a.var(ddof=1) / len(a) + b.var(ddof=1) / len(b))
    return abs(ma - mb), err


def test_chain_observables_agree_with_oracle():
    from thirring2d_amd import _native as N
    from thirring2d_amd import hmc as H
    with mp.get_context("spawn").Pool(NCH) as pool:
        pending = pool.map_async(_oracle_chain, [3001 + 17 * k for k in range(NCH)])
        lats = []
        for k in range(NCH):
            lat = H.StaggeredLattice(NT, NX, m=M, g=G)
            lat.thermalize_pure_gauge(100, seed=600 + k)
            lats.append(lat)
        batch = H.ChainBatch(lats)
        p = H.cg_params(tol_rel=1e-10, precision=N.PREC_MIXED)
        rng = np.random.default_rng(4321)
        gpu = []
        for it in range(BURN + NTRAJ):
            res = batch.update_gauge(nsteps=NSTEPS, params=p, seed=97, counters=[it] * NCH,
                                     u_accept=list(rng.random(NCH)))
            if it >= BURN:
                gpu.append([(r.accepted, r.dS, l.calc_gauge_action()) for r, l in zip(res, lats)])
        orc = np.stack(pending.get(timeout=600), axis=1)  # (NTRAJ, NCH, 3)
    gpu = np.array(gpu)
    for k, name in ((0, "acceptance"), (2, "S_g")):
        d, err = _compare(gpu[:, :, k].mean(axis=0), orc[:, :, k].mean(axis=0))
        assert d <= 4 * err, (name, gpu[:, :, k].mean(), orc[:, :, k].mean(), err)
    d, err = _compare(np.exp(-gpu[:, :, 1]).mean(axis=0), np.exp(-orc[:, :, 1]).mean(axis=0))
    assert d <= 4 * err, ("exp(-dH)", d, err)
    assert 0.5 < gpu[:, :, 0].mean() < 1.0


def test_creutz_identity_with_physical_noise():
    from thirring2d_amd import _native as N
    from thirring2d_amd import hmc as H
    lats = []
    for k in range(NCH):
        lat = H.StaggeredLattice(NT, NX, m=M, g=G)
        lat.thermalize_pure_gauge(100, seed=800 + k)
        lats.append(lat)
    batch = H.ChainBatch(lats)
    p = H.cg_params(tol_rel=1e-10, precision=N.PREC_MIXED)
    rng = np.random.default_rng(99)
    e = []
    for it in range(BURN + NTRAJ):
        res = batch.update_gauge(nsteps=NSTEPS, params=p, phys_norm=1, seed=31,
                                 counters=[it] * NCH, u_accept=list(rng.random(NCH)))
        if it >= BURN:
            e += [np.exp(-r.dS) for r in res]
    e = np.array(e)
    assert abs(e.mean() - 1) <= 5 * e.std(ddof=1) / np.sqrt(e.size), e.mean()
