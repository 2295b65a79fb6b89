"""GPU checks of the device noise generator (Philox4x32-10 in field.hip).

The device generator supplies the Gaussian link fields (th2d_random_links)
and every pseudofermion, momentum and stochastic field of a trajectory run
on device noise.  Pinned here by: Random123's published known-answer
vectors, bitwise agreement of the raw blocks with the oracle's restatement
on random counters and keys, the Gaussian mapping against the oracle's
restatement, and N(0,1) moments of the fields the trajectories draw.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KAT = [
    ([0x00000000, 0x00000000, 0x00000000, 0x00000000], [0x00000000, 0x00000000],
     [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    ([0xffffffff, 0xffffffff, 0xffffffff, 0xffffffff], [0xffffffff, 0xffffffff],
     [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
]


@pytest.fixture(scope="module")
def hmc():
    from thirring2d_amd import hmc as H
    return H


def _device_blocks(H, ck):
    from thirring2d_amd import _native as N
    ck = np.ascontiguousarray(ck, dtype=np.uint32)
    n = ck.shape[0]
    din = H.DeviceArray.from_host(ck)
    dout = H.DeviceArray(16 * n)
    N.check(N.lib().th2d_philox4x32_10(H.Context.get().ptr, n, din.p, dout.p), "philox")
    return dout.download((n, 4), np.uint32)


def test_device_philox_known_answers(hmc, oracle):
    ck = np.array([c + k for c, k, _ in KAT], dtype=np.uint32)
    np.testing.assert_array_equal(_device_blocks(hmc, ck),
                                  np.array([o for _, _, o in KAT], np.uint32))
    rng = np.random.default_rng(5)
    ck = rng.integers(0, 2 ** 32, size=(2048, 6), dtype=np.uint64).astype(np.uint32)
    np.testing.assert_array_equal(_device_blocks(hmc, ck), oracle.philox4x32_10(ck))


def test_random_links_match_oracle_mapping(hmc, oracle):
    """th2d_random_links: the raw eo buffer holds the oracle's Gaussian pairs
    of stream 0x4c4e4b53 in order (reals 2i, 2i+1 = pair i)."""
    nt, nx, seed = 16, 24, 1003
    V = nt * nx
    lat = hmc.StaggeredLattice(nt, nx)
    lat.random_links(seed)
    raw = lat.A.download(2 * V, np.float64)
    z = oracle.philox_gauss(seed, 0x4c4e4b53, V)
    ref = np.column_stack([z.real, z.imag]).ravel()
    assert np.max(np.abs(raw - ref)) <= 1e-13


def test_random_links_moments(hmc):
    """A ~ N(0,1) at 512 x 512 (2^19 angles): moments within 5 sigma."""
    from scipy import stats
    lat = hmc.StaggeredLattice(512, 512)
    lat.random_links(1002)
    a = lat.get_links().ravel()
    n = a.size
    assert abs(a.mean()) < 5 / np.sqrt(n)
    assert abs((a * a).mean() - 1) < 5 * np.sqrt(2 / n)
    assert abs((a ** 4).mean() - 3) < 5 * np.sqrt(96 / n)
    assert stats.kstest(a[:200_000], "norm").pvalue > 1e-4
    b = a.reshape(-1, 2)
    assert abs((b[:, 0] * b[:, 1]).mean()) < 5 / np.sqrt(n / 2)


@pytest.mark.parametrize("phys_norm", [0, 1])
def test_trajectory_noise_moments(hmc, phys_norm):
    """The start-of-trajectory actions of a device-noise trajectory are sums
    of the drawn fields (hmc.c:418-499, 695-699): S_pf0 = |xi|^2,
    S_mom0 = sum p^2 over the 2V links and, at mu = 0, S_md0 = m |st|^2.
    Each of Re and Im ~ N(0, s^2) (s = 1 as the reference, s^2 = 1/2 with
    phys_norm), so E S_pf0 = E S_mom0 = 2 V s^2 and E S_md0 = 2 m V s^2,
    each with standard deviation 2 sqrt(V) s^2 (S_md0: times m).  Checked to
    5 sigma over 4 counters at 128 x 128; different counters draw different
    fields."""
    from thirring2d_amd import _native as N
    nt = nx = 128
    V = nt * nx
    m = 0.1
    s2 = 0.5 if phys_norm else 1.0
    lat = hmc.StaggeredLattice(nt, nx, m=m, g=0.8)
    lat.random_links(1003)
    A0 = lat.get_links()
    acts = []
    for ctr in range(4):
        lat.set_links(A0)
        res = lat.update_gauge(nsteps=1, traj_len=0.01, phys_norm=phys_norm, seed=2003,
                               counter=ctr, params=hmc.cg_params(tol_rel=1e-8,
                                                                 precision=N.PREC_MIXED))
        assert res.status == 0
        acts.append(np.array(res.acts))
    acts = np.array(acts)
    sd = 2 * np.sqrt(V) * s2 / np.sqrt(4)
    assert abs(acts[:, 1].mean() - 2 * V * s2) < 5 * sd
    assert abs(acts[:, 3].mean() - 2 * V * s2) < 5 * sd
    assert abs(acts[:, 2].mean() - 2 * m * V * s2) < 5 * m * sd
    assert len(set(acts[:, 1])) == 4 and len(set(acts[:, 3])) == 4


def _eo_link(nt, nx, s, d):
    t, x = divmod(s, nx)
    V = nt * nx
    return d * V + ((t + x) & 1) * (V // 2) + (t * nx + x) // 2


def test_puregauge_sweeps_match_host_draw_sweep(hmc, oracle):
    """th2d_puregauge_sweeps (device draws) against th2d_puregauge_sweep
    (update_puregauge_hb, hmc.c:82-93, draws supplied) given the same draws,
    restated by the oracle for every link: identical fields after one
    sweep."""
    from thirring2d_amd import _native as N
    nt, nx, seed, ctr = 8, 12, 91, 5
    V = nt * nx
    A0 = np.random.default_rng(8).normal(size=(V, 2))
    a = hmc.StaggeredLattice(nt, nx, g=0.5)
    a.set_links(A0)
    N.check(N.lib().th2d_puregauge_sweeps(a.ctx.ptr, a._g(), a.A.p, 0.5, 2, 1, seed, ctr), "sw")
    U = oracle.philox_uniform(seed, 0x50475357, 2 * V, ctr=ctr)
    draws = np.empty(4 * V)
    for s in range(V):
        for d in range(2):
            l = _eo_link(nt, nx, s, d)
            draws[4 * s + 2 * d], draws[4 * s + 2 * d + 1] = U[l]
    b = hmc.StaggeredLattice(nt, nx, g=0.5)
    b.set_links(A0)
    dd = hmc.DeviceArray.from_host(draws)
    N.check(N.lib().th2d_puregauge_sweep(b.ctx.ptr, b._g(), b.A.p, dd.p, 0.5, 2), "sweep")
    Aa, Ab = a.get_links(), b.get_links()
    assert np.array_equal(Aa, Ab)
    assert 0.1 < np.mean(Aa != A0) < 0.9  # some proposals accepted, some not


def test_puregauge_thermalization_distribution(hmc):
    """100 device sweeps from A = 0 (hmc.c:927-929): every link follows
    exp(beta cos A), beta = Nf/g, so <cos A> = I1(beta)/I0(beta) within
    5 sigma over the 32,768 links of 128^2 (g = 0.5 and g = 1.0)."""
    from scipy.special import i0, i1
    for g in (0.5, 1.0):
        lat = hmc.StaggeredLattice(128, 128, g=g)
        lat.thermalize_pure_gauge(sweeps=100, seed=11)
        c = np.cos(lat.get_links()).ravel()
        beta = 2 / g
        want = i1(beta) / i0(beta)
        var = 0.5 * (1 + (i0(beta) - 2 * i1(beta) / beta) / i0(beta)) - want ** 2
        assert abs(c.mean() - want) < 5 * np.sqrt(var / c.size), (g, c.mean(), want)
