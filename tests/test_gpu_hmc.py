"""GPU parity of the HMC hot path (hmc.c) against the CPU oracle.

Tolerances: FP64 operator and force <= 1e-12 relative (max |diff| / max
|ref|); FP32 operator <= 2e-6 relative; solver results are compared through
the true residual (<= the requested tolerance) and against the oracle's
solution at the solver tolerance.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SIZES = [(4, 4), (8, 8), (16, 16), (8, 32), (32, 8), (48, 48), (64, 32)]


def _rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.fixture(scope="module")
def hmc():
    from thirring2d_amd import hmc as H
    return H


def _setup(H, O, nt, nx, seed, m=0.1, g=0.5, mu=0.0, anti_t=1, anti_x=1):
    rng = np.random.default_rng(seed)
    V = nt * nx
    A = rng.normal(size=(V, 2))
    lat = H.StaggeredLattice(nt, nx, m=m, g=g, mu=mu, anti_t=anti_t, anti_x=anti_x)
    lat.set_links(A)
    L = O.lattice(nt, nx, m=m, g=g, mu=mu, anti_t=anti_t, anti_x=anti_x)
    v = rng.normal(size=V) + 1j * rng.normal(size=V)
    return lat, L, A, v, rng


@pytest.mark.parametrize("nt,nx", SIZES)
@pytest.mark.parametrize("mu", [0.0, 0.3])
def test_fm_mul_and_adjoint(hmc, oracle, nt, nx, mu):
    lat, L, A, v, _ = _setup(hmc, oracle, nt, nx, 11 + nt + nx, mu=mu)
    assert _rel(lat.fm_mul(v), oracle.fm_mul(L, A, v)) <= 1e-12
    assert _rel(lat.fm_conjugate_mul(v), oracle.fm_dag_mul(L, A, v)) <= 1e-12


@pytest.mark.parametrize("anti_t,anti_x", [(1, 1), (1, 0), (0, 0)])
def test_boundary_conditions(hmc, oracle, anti_t, anti_x):
    lat, L, A, v, _ = _setup(hmc, oracle, 8, 12, 5, anti_t=anti_t, anti_x=anti_x)
    assert _rel(lat.fm_mul(v), oracle.fm_mul(L, A, v)) <= 1e-12


def test_links_round_trip(hmc, oracle):
    lat, L, A, v, _ = _setup(hmc, oracle, 16, 8, 3)
    np.testing.assert_array_equal(lat.get_links(), A)


@pytest.mark.parametrize("nt,nx", [(8, 8), (32, 32), (64, 64)])
def test_dslash_fp32(hmc, oracle, nt, nx):
    lat, L, A, v, _ = _setup(hmc, oracle, nt, nx, 7)
    for dag, ref in ((0, oracle.fm_mul(L, A, v)), (1, oracle.fm_dag_mul(L, A, v))):
        assert _rel(lat.apply_f32(v, dag), ref) <= 2e-6


def test_dslash_lds_variant_matches_direct(hmc, oracle):
    from thirring2d_amd import _native as N
    lat, L, A, v, _ = _setup(hmc, oracle, 36, 70, 9)
    ref = oracle.fm_mul(L, A, v)
    try:
        N.lib().th2d_set_dslash_variant(1)
        out_lds = lat.fm_mul(v)
        outd_lds = lat.fm_conjugate_mul(v)
    finally:
        N.lib().th2d_set_dslash_variant(0)
    assert _rel(out_lds, ref) <= 1e-12
    assert _rel(outd_lds, oracle.fm_dag_mul(L, A, v)) <= 1e-12


@pytest.mark.parametrize("nt,nx", [(4, 4), (8, 8), (16, 16), (24, 40), (64, 64)])
def test_cg_fp64(hmc, oracle, nt, nx):
    lat, L, A, b, _ = _setup(hmc, oracle, nt, nx, 21, m=0.1)
    x, info = lat.fmdm_invert_cg(b, hmc.cg_params(tol_rel=1e-12))
    assert info.status == 0
    r = b - oracle.fm_dag_mul(L, A, oracle.fm_mul(L, A, x))
    assert np.linalg.norm(r) / np.linalg.norm(b) <= 1e-12
    assert info.true_rel_resid <= 1e-12
    xo, it, st = oracle.cg_mdm(L, A, b, tol_rel=1e-13)
    assert st == 0
    assert _rel(x, xo) <= 1e-9
    assert abs(info.iterations - it) <= max(3, it // 10)


@pytest.mark.parametrize("nt,nx", [(16, 16), (64, 64), (128, 128)])
def test_cg_mixed(hmc, oracle, nt, nx):
    from thirring2d_amd import _native as N
    lat, L, A, b, _ = _setup(hmc, oracle, nt, nx, 31, m=0.05)
    x, info = lat.fmdm_invert_cg(b, hmc.cg_params(tol_rel=1e-10, precision=N.PREC_MIXED))
    assert info.status == 0
    r = b - oracle.fm_dag_mul(L, A, oracle.fm_mul(L, A, x))
    rel = np.linalg.norm(r) / np.linalg.norm(b)
    assert rel <= 1e-10
    assert abs(rel - info.true_rel_resid) <= 1e-3 * rel + 1e-15


def test_cg_reference_criterion(hmc, oracle):
    """Absolute stop r.r < 1e-30 on the recursive residual (hmc.c:34,381)."""
    lat, L, A, b, _ = _setup(hmc, oracle, 16, 16, 41, m=0.1)
    x, info = lat.fmdm_invert_cg(b, hmc.cg_params(tol_rel=0.0, tol_abs_rr=1e-30))
    xo, it, st = oracle.cg_mdm(L, A, b, tol_abs_rr=1e-30, tol_rel=0.0)
    assert info.status == 0 and st == 0
    assert info.rr < 1e-30
    assert _rel(x, xo) <= 1e-12
    assert abs(info.iterations - it) <= 3


CG_SHAPES = [(16, 64), (8, 64), (16, 32), (8, 32), (16, 16), (8, 16), (8, 8), (4, 8), (4, 4),
             (2, 4), (2, 2)]


@pytest.mark.parametrize("prec", [0, 2])
@pytest.mark.parametrize("tile", CG_SHAPES)
def test_cg_tiles(hmc, oracle, tile, prec):
    """Every compiled tile shape solves the same system, on a lattice that is
    not a multiple of the larger tiles (partial tiles at the edges)."""
    import ctypes as C
    from thirring2d_amd import _native as N
    lat, L, A, b, _ = _setup(hmc, oracle, 40, 72, 51, m=0.2)
    assert N.lib().th2d_set_cg_tile(lat.ctx.ptr, *tile) == 0
    try:
        x, info = lat.fmdm_invert_cg(b, hmc.cg_params(tol_rel=1e-12 if prec == 0 else 1e-10,
                                                       precision=prec))
        tt, tx = C.c_int(), C.c_int()
        N.lib().th2d_cg_last_tile(lat.ctx.ptr, C.byref(tt), C.byref(tx))
    finally:
        N.lib().th2d_set_cg_tile(lat.ctx.ptr, 0, 0)
    assert (tt.value, tx.value) == tile
    r = b - oracle.fm_dag_mul(L, A, oracle.fm_mul(L, A, x))
    assert np.linalg.norm(r) / np.linalg.norm(b) <= (1e-12 if prec == 0 else 1e-10)


def test_cg_tile_rejected(hmc):
    from thirring2d_amd import _native as N
    ctx = hmc.Context.get()
    assert N.lib().th2d_set_cg_tile(ctx.ptr, 12, 12) == N.TH2D_EARG
    assert N.lib().th2d_set_cg_tile(ctx.ptr, 0, 0) == 0


@pytest.mark.parametrize("n,prec", [(256, 0), (256, 2), (512, 2)])
def test_cg_large_lattice(hmc, oracle, n, prec):
    """Regression: the solver's scratch block holds eight vectors; with seven
    the partial sums and the stop flag overran the allocation by one vector,
    which only left the allocator's padding at large lattices."""
    from thirring2d_amd import _native as N
    lat, L, A, b, _ = _setup(hmc, oracle, n, n, 91, m=0.2)
    x, info = lat.fmdm_invert_cg(b, hmc.cg_params(tol_rel=1e-10, precision=prec))
    assert info.status == 0
    r = b - oracle.fm_dag_mul(L, A, oracle.fm_mul(L, A, x))
    assert np.linalg.norm(r) / np.linalg.norm(b) <= 1e-10


def test_cg_several_tiles_per_block(hmc, oracle):
    """Fewer workgroups than tiles: links are restaged per tile and step."""
    from thirring2d_amd import _native as N
    lat, L, A, b, _ = _setup(hmc, oracle, 64, 64, 53, m=0.2)
    N.lib().th2d_set_cg_tile(lat.ctx.ptr, 8, 8)
    N.lib().th2d_set_cg_max_blocks(lat.ctx.ptr, 7)
    try:
        x, info = lat.fmdm_invert_cg(b, hmc.cg_params(tol_rel=1e-12))
    finally:
        N.lib().th2d_set_cg_tile(lat.ctx.ptr, 0, 0)
        N.lib().th2d_set_cg_max_blocks(lat.ctx.ptr, 0)
    r = b - oracle.fm_dag_mul(L, A, oracle.fm_mul(L, A, x))
    assert np.linalg.norm(r) / np.linalg.norm(b) <= 1e-12


@pytest.mark.parametrize("mu", [0.0, 0.25])
def test_force_parity(hmc, oracle, mu):
    lat, L, A, v, rng = _setup(hmc, oracle, 16, 16, 61, m=0.1, g=0.8, mu=mu)
    psi = oracle.fm_dag_mul(L, A, v)
    st = rng.normal(size=L.nt * L.nx) + 1j * rng.normal(size=L.nt * L.nx)
    _, F_ref, _ = oracle.momentum_step(L, np.zeros((256, 2)), A, psi, st, 0.1, tol_rel=1e-14)
    chi, info = lat.fmdm_invert_cg(psi, hmc.cg_params(tol_rel=1e-14))
    F = lat.force(chi, st)
    assert _rel(F, F_ref) <= 1e-10


def test_force_finite_differences(hmc, oracle):
    """Central differences h = 1e-6 of S_g + S_pf on links incl. wrap-around."""
    lat, L, A, v, rng = _setup(hmc, oracle, 12, 12, 71, m=0.1, g=0.8)
    psi = oracle.fm_dag_mul(L, A, v)
    chi, _ = lat.fmdm_invert_cg(psi, hmc.cg_params(tol_rel=1e-14))
    F = lat.force(chi)
    h = 1e-6
    links = [(0, 0), (11, 1), (143, 0), (132, 0), (143, 1), (70, 1), (5, 0), (99, 1)]
    for s, d in links:
        Sp = []
        for sg in (1, -1):
            Ap = A.copy()
            Ap[s, d] += sg * h
            Sp.append(oracle.pseudofermion_action(L, Ap, psi, tol_rel=1e-14)[0]
                      + oracle.gauge_action(L, Ap))
        fd = (Sp[0] - Sp[1]) / (2 * h)
        assert abs(F[s, d] - fd) <= 1e-6 * max(1.0, abs(fd))


def test_gauge_step_and_action(hmc, oracle):
    lat, L, A, v, rng = _setup(hmc, oracle, 16, 8, 81)
    mom = rng.normal(size=(128, 2))
    lat.gauge_step(mom, 0.05)
    np.testing.assert_allclose(lat.get_links(), oracle.gauge_step(L, A, mom, 0.05), rtol=0,
                               atol=1e-15)
    assert abs(lat.calc_gauge_action() - oracle.gauge_action(L, lat.get_links())) <= 1e-11


@pytest.mark.parametrize("nsteps", [10, 40])
def test_trajectory_parity(hmc, oracle, nsteps):
    """Same MT noise, same field: dH, the actions and the accept decision of
    three consecutive trajectories agree with the oracle (16x16, m=0.1, g=0.5)."""
    nt = nx = 16
    V = nt * nx
    L = oracle.lattice(nt, nx, m=0.1, g=0.5)
    A0 = np.random.default_rng(1001).normal(size=(V, 2))
    lat = hmc.StaggeredLattice(nt, nx, m=0.1, g=0.5)
    lat.set_links(A0)
    r_or = oracle.Rng(2001, warmup=543210)
    r_gpu = oracle.Rng(2001, warmup=543210)
    A = A0.copy()
    for _ in range(3):
        # noise for the device in the reference draw order
        xi = np.array([complex(*_pair(r_gpu)) for _ in range(V)])
        mom = np.array([_pair(r_gpu) for _ in range(V)])
        st = np.array([complex(*_pair(r_gpu)) for _ in range(V)])
        u = r_gpu.next()
        res = lat.update_gauge(nsteps=nsteps, noise=(xi, mom, st), u_accept=u,
                               params=hmc.cg_params(tol_rel=1e-12))
        acc, dS, acts, its, A = oracle.update_gauge(L, r_or, A, nsteps=nsteps, tol_rel=1e-12)
        assert res.status == 0
        # nsteps = 10 (dt = 0.1) is beyond the leapfrog's stability limit at
        # m = 0.1 (dH ~ 1e5): round-off differences grow along the
        # trajectory, so the bound is looser there than in the stable case.
        tol = 1e-6 if nsteps == 10 else 1e-9
        assert abs(res.dS - dS) <= tol * max(1.0, abs(dS))
        np.testing.assert_allclose(np.array(res.acts), acts, rtol=tol, atol=tol)
        assert res.accepted == acc
        np.testing.assert_allclose(lat.get_links(), A, rtol=0, atol=1e-9)


def _pair(r):
    x1, x2 = r.next(), r.next()
    rad = np.sqrt(-2 * np.log(x1))
    return rad * np.cos(2 * np.pi * x2), rad * np.sin(2 * np.pi * x2)


def test_trajectory_mixed_vs_fp64(hmc):
    """Mixed-precision trajectory against the FP64 one on the same field and
    device noise, in the stable regime (64x64, m = 0.1, 40 x dt = 0.025,
    residual 1e-10): |dH_mixed - dH_fp64| <= 1e-5 + 1e-6 |dH|."""
    from thirring2d_amd import _native as N
    out = []
    for prec in (N.PREC_F64, N.PREC_MIXED):
        lat = hmc.StaggeredLattice(64, 64, m=0.1, g=0.5)
        lat.random_links(1002)
        res = lat.update_gauge(nsteps=40, params=hmc.cg_params(tol_rel=1e-10, precision=prec),
                               seed=77, counter=0)
        assert res.status == 0
        assert res.max_true_rel_resid <= 1e-10
        out.append(res.dS)
    assert abs(out[0]) < 10.0
    assert abs(out[0] - out[1]) <= 1e-5 + 1e-6 * abs(out[0])


def test_batched_cg_uneven(hmc, oracle):
    """One batched solve of lattices of different sizes and masses: every
    chain reaches its own true residual (FP64 and mixed)."""
    from thirring2d_amd import _native as N
    dims = [(16, 16), (8, 12), (24, 16), (32, 32)]
    setups = [_setup(hmc, oracle, nt, nx, 100 + k, m=0.1 + 0.05 * k)
              for k, (nt, nx) in enumerate(dims)]
    batch = hmc.ChainBatch([s[0] for s in setups])
    for prec, tol in ((N.PREC_F64, 1e-12), (N.PREC_MIXED, 1e-10)):
        xs, infos = batch.fmdm_invert_cg([s[3] for s in setups],
                                         hmc.cg_params(tol_rel=tol, precision=prec))
        for (lat, L, A, b, _), x, inf in zip(setups, xs, infos):
            assert inf.status == 0
            r = b - oracle.fm_dag_mul(L, A, oracle.fm_mul(L, A, x))
            assert np.linalg.norm(r) / np.linalg.norm(b) <= tol


def _last_resident(ctx):
    import ctypes as C
    from thirring2d_amd import _native as N
    on = C.c_int(-1)
    assert N.lib().th2d_cg_last_resident(ctx.ptr, C.byref(on)) == 0
    return on.value


@pytest.mark.parametrize("mu1", [0.0, 0.25])
def test_cg_resident_matches_step_kernel(hmc, oracle, mu1):
    """The resident FP32 solver (one cooperative launch, vectors in
    registers, boundary layers exchanged through memory) and the step kernel
    solve the same mixed-precision batch: lattices smaller than a region
    (one clipped region that exchanges with itself) and lattices of several
    full regions; both reach the true residual, with iteration counts within
    10%.  With
    mu1 != 0 one chain has a chemical potential (the kernel instance with
    the e^{+-mu} factors) and a larger mass; M is then not normal, and the
    FP32 passes of the two kernels round differently enough that only the
    true residual is compared for that chain."""
    from thirring2d_amd import _native as N
    dims = [(128, 128), (64, 64), (40, 48), (16, 16), (64, 192), (8, 60)]
    setups = [_setup(hmc, oracle, nt, nx, 300 + k,
                     m=(0.3 if k == 1 and mu1 else 0.05 + 0.03 * k),
                     mu=mu1 if k == 1 else 0.0)
              for k, (nt, nx) in enumerate(dims)]
    batch = hmc.ChainBatch([s[0] for s in setups])
    ctx = setups[0][0].ctx
    runs = {}
    N.lib().th2d_set_cg_reliable(ctx.ptr, 0)  # both kernels restart every pass
    try:
        for on in (1, 0):
            assert N.lib().th2d_set_cg_resident(ctx.ptr, on) == 0
            xs, infos = batch.fmdm_invert_cg([s[3] for s in setups],
                                             hmc.cg_params(tol_rel=1e-10, precision=N.PREC_MIXED))
            assert _last_resident(ctx) == on
            runs[on] = (xs, infos)
    finally:
        N.lib().th2d_set_cg_resident(ctx.ptr, 1)
        N.lib().th2d_set_cg_reliable(ctx.ptr, 0)
    for k, (lat, L, A, b, _) in enumerate(setups):
        for on in (1, 0):
            x, inf = runs[on][0][k], runs[on][1][k]
            assert inf.status == 0
            r = b - oracle.fm_dag_mul(L, A, oracle.fm_mul(L, A, x))
            assert np.linalg.norm(r) / np.linalg.norm(b) <= 1e-10
        if k == 1 and mu1:
            continue
        it1, it0 = runs[1][1][k].iterations, runs[0][1][k].iterations
        assert abs(it1 - it0) <= max(3, it0 // 10)


@pytest.mark.parametrize("case", ["one_256", "mixed"])
def test_cg_resident_half_rows(hmc, oracle, case):
    """Resident regions of 32 x 64 sites (th2d_set_cg_resident_rows): a
    single 256^2 chain (32 regions on one XCD), and a batch of 128^2, 64 x
    192, 32 x 64 and 64^2 with a chemical potential on one chain (the
    kernel instance with the e^{+-mu} factors).  With 32 rows, with 64 and
    with the automatic choice (32 here: the doubled regions still fit)
    every solve reaches the true residual against the oracle, and the
    iteration counts of the two shapes agree within 10%."""
    import ctypes as C
    from thirring2d_amd import _native as N
    dims = [(256, 256)] if case == "one_256" else [(128, 128), (64, 192), (32, 64), (64, 64)]
    setups = [_setup(hmc, oracle, nt, nx, 700 + k, m=0.05 + 0.02 * k,
                     mu=0.2 if (case == "mixed" and k == 2) else 0.0)
              for k, (nt, nx) in enumerate(dims)]
    batch = hmc.ChainBatch([s[0] for s in setups])
    ctx = setups[0][0].ctx
    lib = N.lib()
    runs = {}
    tt, tx = C.c_int(0), C.c_int(0)
    lib.th2d_set_cg_reliable(ctx.ptr, 0)
    try:
        for rows in (64, 32, 0):
            assert lib.th2d_set_cg_resident_rows(ctx.ptr, rows) == 0
            xs, infos = batch.fmdm_invert_cg([s[3] for s in setups],
                                             hmc.cg_params(tol_rel=1e-10, precision=N.PREC_MIXED))
            assert _last_resident(ctx) == 1
            assert lib.th2d_cg_last_tile(ctx.ptr, C.byref(tt), C.byref(tx)) == 0
            assert (tt.value, tx.value) == ((64, 64) if rows == 64 else (32, 64)), rows
            runs[rows] = (xs, infos)
    finally:
        lib.th2d_set_cg_resident_rows(ctx.ptr, 0)
    assert lib.th2d_set_cg_resident_rows(ctx.ptr, 48) == N.TH2D_EARG
    for k, (lat, L, A, b, _) in enumerate(setups):
        for rows in (64, 32, 0):
            x, inf = runs[rows][0][k], runs[rows][1][k]
            assert inf.status == 0
            r = b - oracle.fm_dag_mul(L, A, oracle.fm_mul(L, A, x))
            assert np.linalg.norm(r) / np.linalg.norm(b) <= 1e-10, (dims[k], rows)
        it64, it32 = runs[64][1][k].iterations, runs[32][1][k].iterations
        assert abs(it64 - it32) <= max(3, it64 // 10), (dims[k], it64, it32)


@pytest.mark.parametrize("m", [0.2, 0.05])
def test_cg_reliable_updates(hmc, oracle, m):
    """Mixed solves with reliable updates (the FP32 Krylov sequence continues
    across FP64 passes) and with restarts reach the same true residual; on a
    well-conditioned system (m = 0.2) the updates save iterations (at
    m = 0.05 they may cost some: the FP32 residual drifts within a pass)."""
    from thirring2d_amd import _native as N
    dims = [(128, 128), (64, 128), (128, 64)]
    setups = [_setup(hmc, oracle, nt, nx, 600 + k, m=m) for k, (nt, nx) in enumerate(dims)]
    batch = hmc.ChainBatch([s[0] for s in setups])
    ctx = setups[0][0].ctx
    its = {}
    try:
        for rel in (0, 1):
            assert N.lib().th2d_set_cg_reliable(ctx.ptr, rel) == 0
            xs, infos = batch.fmdm_invert_cg([s[3] for s in setups],
                                             hmc.cg_params(tol_rel=1e-10, precision=N.PREC_MIXED))
            assert _last_resident(ctx) == 1
            for (lat, L, A, b, _), x, inf in zip(setups, xs, infos):
                assert inf.status == 0
                r = b - oracle.fm_dag_mul(L, A, oracle.fm_mul(L, A, x))
                assert np.linalg.norm(r) / np.linalg.norm(b) <= 1e-10
            its[rel] = sum(inf.iterations for inf in infos)
    finally:
        N.lib().th2d_set_cg_reliable(ctx.ptr, 0)
    if m >= 0.2:
        assert its[1] <= its[0]


def test_cg_resident_full_device(hmc, oracle):
    """16 lattices of 256^2 are 256 regions, one per CU: the whole batch
    runs resident and every chain reaches its true residual."""
    from thirring2d_amd import _native as N
    n = 16
    setups = [_setup(hmc, oracle, 256, 256, 400 + k, m=0.05) for k in range(n)]
    batch = hmc.ChainBatch([s[0] for s in setups])
    xs, infos = batch.fmdm_invert_cg([s[3] for s in setups],
                                     hmc.cg_params(tol_rel=1e-10, precision=N.PREC_MIXED))
    assert _last_resident(setups[0][0].ctx) == 1
    for k in range(n):
        assert infos[k].status == 0
        assert infos[k].true_rel_resid <= 1e-10
    for k in (0, n - 1):
        lat, L, A, b, _ = setups[k]
        r = b - oracle.fm_dag_mul(L, A, oracle.fm_mul(L, A, xs[k]))
        assert np.linalg.norm(r) / np.linalg.norm(b) <= 1e-10


def test_cg_small_lattice_kernel(hmc, oracle):
    """FP64 solves of lattices up to 2048 sites run the whole CG in one
    workgroup (th2d_set_cg_small): a batch of uneven small lattices (1 to 2
    sites per thread, 256 and 1024 threads) reaches the true residual, the
    solutions agree with the oracle's, and the step kernel (small path off)
    solves the same batch to the same tolerance with iteration counts within
    2%."""
    from thirring2d_amd import _native as N
    dims = [(16, 16), (8, 8), (24, 40), (32, 64), (6, 10), (32, 16)]
    setups = [_setup(hmc, oracle, nt, nx, 300 + k, m=0.1 + 0.03 * k)
              for k, (nt, nx) in enumerate(dims)]
    batch = hmc.ChainBatch([s[0] for s in setups])
    ctx = setups[0][0].ctx
    runs = {}
    try:
        for on in (1, 0):
            N.check(N.lib().th2d_set_cg_small(ctx.ptr, on), "small")
            xs, infos = batch.fmdm_invert_cg([s[3] for s in setups],
                                             hmc.cg_params(tol_rel=1e-12))
            assert _last_resident(ctx) == (3 if on else 0)
            runs[on] = (xs, infos)
    finally:
        N.lib().th2d_set_cg_small(ctx.ptr, 1)
    for k, (lat, L, A, b, _) in enumerate(setups):
        inf = runs[1][1][k]
        assert inf.status == 0 and inf.true_rel_resid <= 1e-12
        xo, _, _ = oracle.cg_mdm(L, A, b, tol_rel=1e-13)
        assert _rel(runs[1][0][k], xo) <= 1e-9
        it1, it0 = inf.iterations, runs[0][1][k].iterations
        assert abs(it1 - it0) <= max(2, it0 // 50)


def test_md_step_graph_matches_eager(hmc):
    """Small-lattice FP64 trajectories with every MD step a replay of one
    captured graph (th2d_set_hmc_graph) give bitwise the results of the
    step-by-step launches: dH, the actions, accept decisions, solver
    statistics and fields, over two consecutive trajectories."""
    from thirring2d_amd import _native as N
    dims = [(16, 16), (8, 12), (32, 16), (24, 24)]
    out = {}
    ctx = hmc.Context.get()
    try:
        for on in (1, 0):
            N.check(N.lib().th2d_set_hmc_graph(ctx.ptr, on), "graph")
            lats = []
            for k, (nt, nx) in enumerate(dims):
                lat = hmc.StaggeredLattice(nt, nx, m=0.2, g=0.6)
                lat.random_links(70 + k)
                lats.append(lat)
            batch = hmc.ChainBatch(lats)
            rec = []
            for t in range(2):
                res = batch.update_gauge(nsteps=12, params=hmc.cg_params(tol_rel=1e-11), seed=9,
                                         counters=[t] * len(dims), u_accept=[0.3] * len(dims))
                rec.append([(r.dS, tuple(r.acts), r.accepted, r.cg_iterations, r.cg_solves,
                             r.max_true_rel_resid, r.status) for r in res])
            out[on] = (rec, [l.get_links() for l in lats])
    finally:
        N.lib().th2d_set_hmc_graph(ctx.ptr, 1)
    assert out[1][0] == out[0][0]
    for a, b in zip(out[1][1], out[0][1]):
        assert np.array_equal(a, b)
    assert all(r[6] == 0 and r[4] == 13 for tr in out[1][0] for r in tr)


def test_cg_resident_wait_timeout_falls_back(hmc, oracle):
    """A bounded wait of the resident solver that gives up (forced from the
    host with a spin limit of 0 polls: every workgroup still exits normally)
    makes the solve fall back to the step kernel from the same right-hand
    sides (a pass whose resident solve was left in flight takes a zero
    correction and is repeated): the solve and a batched trajectory both
    finish with every chain at its true residual, and the fallbacks are
    counted.  With the default limit the resident kernel runs again."""
    import ctypes as C
    from thirring2d_amd import _native as N
    setups = [_setup(hmc, oracle, 128, 128, 900 + k, m=0.1) for k in range(4)]
    batch = hmc.ChainBatch([s[0] for s in setups])
    ctx = setups[0][0].ctx
    p = hmc.cg_params(tol_rel=1e-10, precision=N.PREC_MIXED)
    n0, n1 = C.c_long(), C.c_long()
    N.check(N.lib().th2d_cg_sync_fallbacks(ctx.ptr, C.byref(n0)), "fallbacks")
    N.check(N.lib().th2d_set_cg_spin_limit(ctx.ptr, 0), "spin")
    try:
        xs, infos = batch.fmdm_invert_cg([s[3] for s in setups], p)
        assert _last_resident(ctx) == 0
        for (lat, L, A, b, _), x, inf in zip(setups, xs, infos):
            assert inf.status == 0 and inf.true_rel_resid <= 1e-10
            r = b - oracle.fm_dag_mul(L, A, oracle.fm_mul(L, A, x))
            assert np.linalg.norm(r) / np.linalg.norm(b) <= 1e-10
        res = batch.update_gauge(nsteps=2, traj_len=0.1, seed=5, counters=[0, 1, 2, 3], params=p)
        for r in res:
            assert r.status == 0 and r.max_true_rel_resid <= 1e-10
        N.check(N.lib().th2d_cg_sync_fallbacks(ctx.ptr, C.byref(n1)), "fallbacks")
        assert n1.value - n0.value >= 4  # the solve and the trajectory's 3 solves
    finally:
        N.check(N.lib().th2d_set_cg_spin_limit(ctx.ptr, -1), "spin")
    xs, infos = batch.fmdm_invert_cg([s[3] for s in setups], p)
    assert _last_resident(ctx) == 1
    assert all(inf.status == 0 and inf.true_rel_resid <= 1e-10 for inf in infos)


def test_cg_resident_groups(hmc, oracle):
    """A batch with more regions than the device holds resident workgroups
    (12 x 256^2 + 20 x 128^2 = 272 regions) is solved by the resident kernel
    in two launches, one per group of chains; every chain reaches the true
    residual, with the step kernel's iteration counts within 10%."""
    from thirring2d_amd import _native as N
    dims = [(256, 256)] * 12 + [(128, 128)] * 20
    setups = [_setup(hmc, oracle, nt, nx, 700 + k, m=0.1 + 0.002 * k)
              for k, (nt, nx) in enumerate(dims)]
    batch = hmc.ChainBatch([s[0] for s in setups])
    ctx = setups[0][0].ctx
    runs = {}
    N.lib().th2d_set_cg_reliable(ctx.ptr, 0)
    try:
        for on in (1, 0):
            assert N.lib().th2d_set_cg_resident(ctx.ptr, on) == 0
            xs, infos = batch.fmdm_invert_cg([s[3] for s in setups],
                                             hmc.cg_params(tol_rel=1e-10, precision=N.PREC_MIXED))
            assert _last_resident(ctx) == on
            runs[on] = (xs, infos)
    finally:
        N.lib().th2d_set_cg_resident(ctx.ptr, 1)
    for k, (lat, L, A, b, _) in enumerate(setups):
        inf = runs[1][1][k]
        assert inf.status == 0 and inf.true_rel_resid <= 1e-10
        if k % 7 == 0:  # the oracle's residual on a sample of the chains
            x = runs[1][0][k]
            r = b - oracle.fm_dag_mul(L, A, oracle.fm_mul(L, A, x))
            assert np.linalg.norm(r) / np.linalg.norm(b) <= 1e-10
        it1, it0 = inf.iterations, runs[0][1][k].iterations
        assert abs(it1 - it0) <= max(3, it0 // 10)


def test_cg_resident_splits_clipped_regions(hmc, oracle):
    """A batch with lattices cut into several regions with a clipped edge
    region (96^2: regions of 64 and 32 sites a side) is split: those chains
    in the step kernel, the others (128^2, 32^2, 64^2) in the resident
    kernel.  Every chain reaches the true residual, with the iteration
    counts of the step kernel alone within 10%."""
    from thirring2d_amd import _native as N
    dims = [(96, 96), (128, 128), (32, 32), (96, 96), (64, 64)]
    setups = [_setup(hmc, oracle, nt, nx, 600 + k, m=0.1) for k, (nt, nx) in enumerate(dims)]
    batch = hmc.ChainBatch([s[0] for s in setups])
    ctx = setups[0][0].ctx
    runs = {}
    try:
        for on in (1, 0):
            assert N.lib().th2d_set_cg_resident(ctx.ptr, on) == 0
            xs, infos = batch.fmdm_invert_cg([s[3] for s in setups],
                                             hmc.cg_params(tol_rel=1e-10, precision=N.PREC_MIXED))
            assert _last_resident(ctx) == (2 if on else 0)
            runs[on] = (xs, infos)
    finally:
        N.lib().th2d_set_cg_resident(ctx.ptr, 1)
    for k, (lat, L, A, b, _) in enumerate(setups):
        x, inf = runs[1][0][k], runs[1][1][k]
        assert inf.status == 0
        r = b - oracle.fm_dag_mul(L, A, oracle.fm_mul(L, A, x))
        assert np.linalg.norm(r) / np.linalg.norm(b) <= 1e-10
        it1, it0 = inf.iterations, runs[0][1][k].iterations
        assert abs(it1 - it0) <= max(3, it0 // 10)


def test_cg_resident_declines_clipped_regions(hmc, oracle):
    """A batch of only lattices cut into several regions with a clipped
    edge region (96^2) is solved by the step kernel."""
    from thirring2d_amd import _native as N
    setups = [_setup(hmc, oracle, nt, nx, 650 + k, m=0.1)
              for k, (nt, nx) in enumerate([(96, 96), (96, 160)])]
    batch = hmc.ChainBatch([s[0] for s in setups])
    xs, infos = batch.fmdm_invert_cg([s[3] for s in setups],
                                     hmc.cg_params(tol_rel=1e-10, precision=N.PREC_MIXED))
    assert _last_resident(setups[0][0].ctx) == 0
    for (lat, L, A, b, _), x, inf in zip(setups, xs, infos):
        assert inf.status == 0
        r = b - oracle.fm_dag_mul(L, A, oracle.fm_mul(L, A, x))
        assert np.linalg.norm(r) / np.linalg.norm(b) <= 1e-10


@pytest.mark.parametrize("prec", ["f64", "mixed"])
def test_cg_step_graph_matches_eager(hmc, oracle, prec):
    """The step kernel's launches replayed from a captured graph and issued
    one by one give bitwise the same solutions and iteration counts, over
    three solves in a row with other work (the FP64 true-residual passes)
    between the replays; the graph is captured once and reused."""
    from thirring2d_amd import _native as N
    dims = [(96, 96), (40, 56), (16, 16), (96, 160)]
    setups = [_setup(hmc, oracle, nt, nx, 900 + k, m=0.1) for k, (nt, nx) in enumerate(dims)]
    batch = hmc.ChainBatch([s[0] for s in setups])
    ctx = setups[0][0].ctx
    precision = N.PREC_F64 if prec == "f64" else N.PREC_MIXED
    rng = np.random.default_rng(901)
    rhs = [[rng.normal(size=nt * nx) + 1j * rng.normal(size=nt * nx) for nt, nx in dims]
           for _ in range(3)]
    out = {}
    try:
        for g in (1, 0):
            assert N.lib().th2d_set_cg_graph(ctx.ptr, g) == 0
            out[g] = [batch.fmdm_invert_cg(b, hmc.cg_params(tol_rel=1e-10, precision=precision))
                      for b in rhs]
    finally:
        N.lib().th2d_set_cg_graph(ctx.ptr, 1)
    for (xs1, inf1), (xs0, inf0) in zip(out[1], out[0]):
        for x1, x0, i1, i0 in zip(xs1, xs0, inf1, inf0):
            assert i1.status == 0 and i0.status == 0
            assert i1.iterations == i0.iterations
            assert np.array_equal(x1, x0)
    # and the solutions are solutions
    for (lat, L, A, _, _), x, b in zip(setups, out[1][2][0], rhs[2]):
        r = b - oracle.fm_dag_mul(L, A, oracle.fm_mul(L, A, x))
        assert np.linalg.norm(r) / np.linalg.norm(b) <= 1e-10


def test_cg_resident_falls_back_when_too_large(hmc, oracle):
    """A batch with a lattice larger than an XCD's resident workgroups
    (512^2 = 64 regions) runs in the step kernel."""
    from thirring2d_amd import _native as N
    setups = [_setup(hmc, oracle, 512, 512, 500 + k, m=0.2) for k in range(5)]
    batch = hmc.ChainBatch([s[0] for s in setups])
    xs, infos = batch.fmdm_invert_cg([s[3] for s in setups],
                                     hmc.cg_params(tol_rel=1e-10, precision=N.PREC_MIXED))
    assert _last_resident(setups[0][0].ctx) == 0
    for inf in infos:
        assert inf.status == 0 and inf.true_rel_resid <= 1e-10


def test_batched_trajectory_matches_oracle(hmc, oracle):
    """Chains of different sizes and couplings in one batched trajectory,
    each with its own MT noise: every chain equals the oracle's trajectory
    (stable step, 20 x dt = 0.05)."""
    dims = [(16, 16), (8, 12), (12, 8)]
    gs = [0.5, 0.8, 0.3]
    lats, Ls, A0s, noise, us = [], [], [], [], []
    orc_rngs = []
    for k, ((nt, nx), g) in enumerate(zip(dims, gs)):
        V = nt * nx
        A0 = np.random.default_rng(300 + k).normal(size=(V, 2))
        lat = hmc.StaggeredLattice(nt, nx, m=0.2, g=g)
        lat.set_links(A0)
        lats.append(lat)
        Ls.append(oracle.lattice(nt, nx, m=0.2, g=g))
        A0s.append(A0)
        r = oracle.Rng(5000 + k, warmup=100)
        xi = np.array([complex(*_pair(r)) for _ in range(V)])
        mom = np.array([_pair(r) for _ in range(V)])
        st = np.array([complex(*_pair(r)) for _ in range(V)])
        us.append(r.next())
        noise.append((xi, mom, st))
        orc_rngs.append(oracle.Rng(5000 + k, warmup=100))
    batch = hmc.ChainBatch(lats)
    res = batch.update_gauge(nsteps=20, traj_len=1.0, noise=noise, u_accept=us,
                             params=hmc.cg_params(tol_rel=1e-12))
    for k in range(3):
        acc, dS, acts, its, A1 = oracle.update_gauge(Ls[k], orc_rngs[k], A0s[k], nsteps=20,
                                                     tol_rel=1e-12)
        assert res[k].status == 0
        assert abs(res[k].dS - dS) <= 1e-8 * max(1.0, abs(dS))
        assert res[k].accepted == acc
        np.testing.assert_allclose(lats[k].get_links(), A1, rtol=0, atol=1e-9)
