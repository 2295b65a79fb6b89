"""The batched device bag driver (th2d_bag_batch_*): many configurations
advanced by Metropolis proposals on the device (fermionbag.c:284-379 with
the ratios and accept hooks of fluctuation_determinant.c:305-871).

Every traced proposal is replayed on the host: its ratio equals the ratio of
full determinants of the oracle's reduced matrices (determinant_mu,
full_determinant.c:144-186) to 1e-9 relative; the fields after the sweep
are the replayed ones; log|det D_FF| and its sign equal a full recompute to
1e-9; the accept counters equal the traced decisions.  At the C5 sizes
(up to 48 x 48, N = 2304) sampled ratios and the final log-determinants are
checked the same way."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _field(nt, nx, rho, rng, bc=0):
    V = nt * nx
    f = np.zeros(V, np.int32)
    for s in rng.permutation(V):
        if f[s] or rng.random() > rho:
            continue
        t, x = divmod(int(s), nx)
        d = int(rng.integers(2))
        if d == 1 and bc == 1 and x == nx - 1:
            continue
        t2, x2 = ((t + 1) % nt, x) if d == 0 else (t, (x + 1) % nx)
        s2 = t2 * nx + x2
        if not f[s2]:
            f[s], f[s2] = 2 + d, 4 + d
    free = np.flatnonzero(f == 0)
    f[rng.choice(free, size=min(3, len(free)), replace=False)] = 1
    return f


def _apply(f, mv):
    g = f.copy()
    for a in range(mv.nr):
        g[mv.fr[a]] = 0
    for a in range(mv.no):
        g[mv.occ[a]] = mv.vals[a]
    return g


def test_bag_batch_replays_against_oracle(oracle):
    from thirring2d_amd import bag as Bm
    rng = np.random.default_rng(606)
    dims = [(6, 6, 0), (8, 6, 1), (8, 8, 0), (6, 10, 1), (10, 8, 0), (4, 6, 0)]
    lats = [Bm.BagLattice(nt, nx, bc=bc, m=0.3, mu=0.1) for nt, nx, bc in dims]
    Bs = [oracle.bag_lattice(nt, nx, bc=bc, m=0.3, mu=0.1) for nt, nx, bc in dims]
    fields = [_field(nt, nx, 0.3, rng, bc) for nt, nx, bc in dims]
    batch = Bm.BagBatch(lats, fields)
    la0, sg0, f0, acc0 = batch.state()
    for B, f, l, s in zip(Bs, fields, la0, sg0):
        lo, so, _ = oracle.determinant_mu(B, f)
        assert so == s and abs(l - lo) <= 1e-10 * max(1.0, abs(lo))
    steps = 80
    trace = batch.sweep(steps, seed=99, counter=0, trace=True)
    la, sg, fl, acc = batch.state()
    cur = [f.copy() for f in fields]
    moves = accepted = 0
    for step in trace:
        for b, mv in enumerate(step):
            if mv.no + mv.nr == 0:
                assert mv.ratio == 0.0 and mv.accepted == 0
                continue
            new = _apply(cur[b], mv)
            lo, so, _ = oracle.determinant_mu(Bs[b], cur[b])
            ln, sn, _ = oracle.determinant_mu(Bs[b], new)
            want = 0.0 if sn == 0 else so * sn * np.exp(ln - lo)
            if want == 0.0:
                assert abs(mv.ratio) <= 1e-8
            else:
                assert abs(mv.ratio - want) <= 1e-9 * abs(want), (b, mv.no, mv.nr, mv.ratio, want)
            moves += 1
            if mv.accepted:
                assert mv.ratio != 0.0
                cur[b] = new
                accepted += 1
    assert moves > 200 and accepted > 60
    assert sum(acc) - sum(acc0) == accepted
    for B, f, g, l, s in zip(Bs, cur, fl, la, sg):
        assert np.array_equal(f, g)
        lo, so, _ = oracle.determinant_mu(B, f)
        assert so == s and abs(l - lo) <= 1e-9 * max(1.0, abs(lo))
    # a refresh from scratch agrees with the updated state
    batch.refresh()
    la2, sg2, _, _ = batch.state()
    assert np.all(sg2 == sg) and np.max(np.abs(la2 - la) / np.maximum(1.0, np.abs(la))) <= 1e-9


def test_bag_batch_c5_sizes(oracle):
    """Four matrices of 48^2, 40^2, 32^2, 16^2 (m = 0.1, occupation ~0.3),
    100 proposals each: 12 sampled ratios and the final log|det D_FF|
    against FP64 LU of the oracle's reduced matrices (numpy slogdet)."""
    from thirring2d_amd import bag as Bm
    rng = np.random.default_rng(5005)
    Ls = [48, 40, 32, 16]
    lats = [Bm.BagLattice(L, L, m=0.1) for L in Ls]
    Bs = [oracle.bag_lattice(L, L, m=0.1) for L in Ls]
    fields = [_field(L, L, 0.3, rng) for L in Ls]
    batch = Bm.BagBatch(lats, fields)
    trace = batch.sweep(100, seed=7, counter=11, trace=True)
    la, sg, fl, acc = batch.state()

    def full(B, f):
        s, l = np.linalg.slogdet(oracle.reduced_matrix(B, f))
        return l, int(s)

    cur = [f.copy() for f in fields]
    checked = nmoves = 0
    for k, step in enumerate(trace):
        for b, mv in enumerate(step):
            if mv.no + mv.nr == 0:
                continue
            new = _apply(cur[b], mv)
            nmoves += 1
            if checked < 12 and nmoves % 10 == 0:
                lo, so = full(Bs[b], cur[b])
                ln, sn = full(Bs[b], new)
                want = so * sn * np.exp(ln - lo)
                assert abs(mv.ratio - want) <= 1e-9 * abs(want), (Ls[b], k, mv.ratio, want)
                checked += 1
            if mv.accepted:
                cur[b] = new
    assert checked == 12
    for B, f, g, l, s in zip(Bs, cur, fl, la, sg):
        assert np.array_equal(f, g)
        lo, so = full(B, f)
        assert so == s and abs(l - lo) <= 1e-9 * abs(lo)
    assert np.all(acc > 20)
