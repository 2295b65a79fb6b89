"""GPU parity of the fermion-bag and dense-determinant path against the CPU
oracle (vec_ops.c, fermion_matrix.c, full_determinant.c,
fluctuation_determinant.c, tests/thirring_brute.c).

Tolerances: masked operator bit-exact-level (<= 1e-14 absolute); bag CG
solution <= 1e-10 absolute at the reference stop r.r < 1e-26; FP64 LU
log|det| <= 1e-11 relative and identical sign, FP64 inverse <= 1e-10
relative (max |diff| / max |ref|); FP32 LU log|det| within 8 n eps32
absolute of the FP64 value (stated mixed-precision bound) and identical
sign; Dinv <= 1e-12; bag weights and Z table <= 1e-10 relative; rank-k
ratios against full LU recomputes <= 1e-9 relative.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def bag():
    from thirring2d_amd import bag as Bm
    return Bm


@pytest.fixture(scope="module")
def native():
    from thirring2d_amd import _native as N
    return N


def _random_field(nt, nx, rho, rng, bc=0, monomers=2):
    """Random legal links (tail 2+dir, head 4+dir) and monomers."""
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
    for s in rng.choice(free, size=min(monomers, len(free)), replace=False):
        f[s] = 1
    return f


@pytest.mark.parametrize("bc", [0, 1])
@pytest.mark.parametrize("wmm", [0, 1])
@pytest.mark.parametrize("nt,nx", [(4, 4), (8, 6), (16, 12)])
def test_bag_apply(bag, oracle, bc, wmm, nt, nx):
    rng = np.random.default_rng(100 + nt + 3 * bc + wmm)
    lat = bag.BagLattice(nt, nx, bc=bc, with_mass_monomers=wmm, m=0.3, mu=0.2)
    B = oracle.bag_lattice(nt, nx, bc=bc, with_mass_monomers=wmm, m=0.3, mu=0.2)
    f = _random_field(nt, nx, 0.3, rng, bc)
    psi = rng.normal(size=(3, nt * nx))
    for tr in (False, True):
        got = lat.fM(f, psi, transpose=tr)
        want = np.stack([oracle.fM(B, f, p, transpose=tr) for p in psi])
        assert np.max(np.abs(got - want)) <= 1e-14


@pytest.mark.parametrize("bc", [0, 1])
def test_bag_cg(bag, oracle, bc):
    rng = np.random.default_rng(7 + bc)
    nt, nx = 8, 10
    lat = bag.BagLattice(nt, nx, bc=bc, m=0.2, mu=0.1)
    B = oracle.bag_lattice(nt, nx, bc=bc, m=0.2, mu=0.1)
    f = _random_field(nt, nx, 0.25, rng, bc)
    src = rng.normal(size=(2, nt * nx))
    x, its, sts = lat.cg_propagator(f, src, tol_abs_rr=1e-26)
    for r in range(2):
        xo, it, st = oracle.cg_propagator(B, f, src[r], tol_abs_rr=1e-26)
        assert sts[r] == 0 and st == 0
        assert abs(its[r] - it) <= 2
        assert np.max(np.abs(x[r] - xo)) <= 1e-10
        # it solves fM x = src
        assert np.max(np.abs(oracle.fM(B, f, x[r]) - src[r])) <= 1e-10


def _mats(rng, ns, shift=0.0):
    return [rng.normal(size=(n, n)) + shift * np.eye(n) for n in ns]


NS = [1, 5, 31, 32, 33, 64, 100, 257]


@pytest.mark.parametrize("mfma", [1, 0])
def test_lu_logdet_f64(bag, native, oracle, mfma):
    rng = np.random.default_rng(11)
    mats = _mats(rng, NS)
    native.lib().th2d_set_dense_mfma(mfma)
    try:
        db = bag.DenseBatch(mats, native.PREC_F64).lu()
        la, sg = db.logdet()
    finally:
        native.lib().th2d_set_dense_mfma(1)
    assert np.all(db.infos() == 0)
    for M, l, s in zip(mats, la, sg):
        s_np, l_np = np.linalg.slogdet(M)
        lo, so, _ = oracle.dlogdet(M)
        assert s == int(s_np) == so
        assert abs(l - l_np) <= 1e-11 * max(1.0, abs(l_np))
        assert abs(l - lo) <= 1e-11 * max(1.0, abs(lo))


def test_lu_mfma_matches_valu(bag, native):
    """The MFMA 16x16x4 operand/result layout: MFMA and vector-ALU trailing
    updates give the same factors to rounding."""
    rng = np.random.default_rng(12)
    mats = _mats(rng, [70, 130])
    out = {}
    for mfma in (1, 0):
        native.lib().th2d_set_dense_mfma(mfma)
        try:
            for prec in (native.PREC_F64, native.PREC_F32):
                out[(mfma, prec)] = bag.DenseBatch(mats, prec).lu().matrices()
        finally:
            native.lib().th2d_set_dense_mfma(1)
    for prec, tol in ((native.PREC_F64, 1e-12), (native.PREC_F32, 1e-4)):
        for a, b in zip(out[(1, prec)], out[(0, prec)]):
            assert np.max(np.abs(a - b)) / np.max(np.abs(b)) <= tol


@pytest.mark.parametrize("reg", [1, 0])
def test_lu_panels_match_numpy(bag, native, reg):
    """Both LU panel factorisations (16 columns in registers with the row
    interchanges applied afterwards; 32 columns with whole-row swaps as they
    go) give numpy's log|det| and sign, on orders from 2 to 2304 in one
    batch."""
    rng = np.random.default_rng(21)
    ns = [2, 16, 17, 48, 300, 1024, 1025, 2304]
    mats = _mats(rng, ns)
    native.lib().th2d_set_lu_panel(reg)
    try:
        db = bag.DenseBatch(mats, native.PREC_F64).lu()
        la, sg = db.logdet()
    finally:
        native.lib().th2d_set_lu_panel(1)
    assert np.all(db.infos() == 0)
    for M, l, s in zip(mats, la, sg):
        s_np, l_np = np.linalg.slogdet(M)
        assert s == int(s_np)
        assert abs(l - l_np) <= 1e-10 * max(1.0, abs(l_np))


@pytest.mark.parametrize("reg", [1, 0])
def test_lu_factors_match_lapack(bag, native, reg):
    """The pivots equal LAPACK's partial-pivot choices (scipy's getrf: the
    first row of largest |a|), and L and U agree with LAPACK's to rounding,
    for both panels, on orders either side of the 16- and 128-column block
    edges and of the panel's thread counts, with a matrix of entries in
    {-1/2, 0, 1/2, 1} full of exact ties."""
    sl = pytest.importorskip("scipy.linalg")
    rng = np.random.default_rng(23)
    ns = [1, 5, 16, 17, 127, 128, 129, 300, 769, 1537, 2304]
    mats = _mats(rng, ns) + [rng.choice([-0.5, 0.0, 0.5, 1.0], size=(500, 500))]
    native.lib().th2d_set_lu_panel(reg)
    try:
        db = bag.DenseBatch(mats, native.PREC_F64).lu()
    finally:
        native.lib().th2d_set_lu_panel(1)
    piv = db.piv.download((sum(db.ns),), np.int32)
    o = 0
    for M, F in zip(mats, db.matrices()):
        n = M.shape[0]
        lu, pv = sl.lu_factor(M)
        assert np.array_equal(piv[o:o + n], pv), n
        assert np.max(np.abs(F - lu)) <= 1e-11 * np.max(np.abs(lu)), n
        o += n


def test_lu_lookahead_bitwise(bag, native):
    """The look-ahead LU (panels on their own CUs beside the previous
    block's trailing update, interchanges outside a block deferred) gives
    bitwise the factors and pivots of the in-order LU, FP64 and FP32, on
    uneven orders across block edges."""
    rng = np.random.default_rng(21)
    mats = _mats(rng, [2304, 129, 300, 1000, 64, 1537], shift=0.5)
    out = {}
    try:
        for la in (1, 0):
            native.lib().th2d_set_lu_lookahead(la)
            for prec in (native.PREC_F64, native.PREC_F32):
                db = bag.DenseBatch(mats, prec).lu()
                out[(la, prec)] = (db.matrices(), db.piv.download((sum(db.ns),), np.int32))
    finally:
        native.lib().th2d_set_lu_lookahead(1)
    for prec in (native.PREC_F64, native.PREC_F32):
        (m1, p1), (m0, p0) = out[(1, prec)], out[(0, prec)]
        assert np.array_equal(p1, p0)
        for a, b in zip(m1, m0):
            assert np.array_equal(a, b)


def test_lu_logdet_f32_bound(bag, native):
    """FP32-MFMA LU: log|det| within the stated bound of the FP64 value, on
    orders up to 2304 (several 128-column blocks, every panel width).

    The FP32 factors are those of A + dA with |dA| <= gamma_n |L||U|, so
    log|det| moves by tr(A^-1 dA), of order n eps32 for these shifted
    Gaussian matrices (growth |U|/|A| <= 20).  LAPACK's sgetrf on the same
    matrices is off by at most 1.1 n eps32; the bound is 8 n eps32
    (2.2e-3 at n = 2304)."""
    rng = np.random.default_rng(13)
    ns = [16, 64, 100, 256, 400, 1500, 2304]
    mats = _mats(rng, ns, shift=4.0)
    la64, sg64 = bag.DenseBatch(mats, native.PREC_F64).lu().logdet()
    la32, sg32 = bag.DenseBatch(mats, native.PREC_F32).lu().logdet()
    eps32 = float(np.finfo(np.float32).eps)
    for n, a, b, s1, s2 in zip(ns, la32, la64, sg32, sg64):
        assert s1 == s2
        assert abs(a - b) <= 8 * n * eps32, (n, abs(a - b))


def test_lu_logdet_f32_refined(bag, native, oracle):
    """FP32-MFMA LU with the FP64 correction of its log-determinant
    (th2d_lu_logdet_refined) against FP64 LU: the first-order correction
    tr(B^-1 E) leaves the second-order term, stated as
    5e-11 |log det| + 1e-10 (measured: 1.1e-11 |log det| at order 1500;
    scipy's FP32 LU of the same matrices with the same correction is off by
    at most 2.6e-12 |log det|); at least 1000x closer than the uncorrected
    FP32 value where that is off by > 1e-6.
    Also on masked bag matrices of 16^2 .. 48^2."""
    rng = np.random.default_rng(13)
    ns = [16, 64, 100, 256, 400, 1500, 2304]
    mats = _mats(rng, ns, shift=4.0)
    rng2 = np.random.default_rng(51)
    for L in (16, 32, 48):
        B = oracle.bag_lattice(L, L, m=0.1)
        f = np.zeros(L * L, np.int32)
        f[rng2.choice(L * L, size=L * L // 5, replace=False)] = 1
        M = oracle.reduced_matrix(B, np.zeros(L * L, np.int32))
        occ = np.flatnonzero(f)
        M[occ, :] = 0.0
        M[:, occ] = 0.0
        M[occ, occ] = 1.0
        mats.append(M)
    d64 = bag.DenseBatch(mats, native.PREC_F64)
    la64, sg64 = bag.DenseBatch(mats, native.PREC_F64).lu().logdet()
    f32 = bag.DenseBatch(mats, native.PREC_F32).lu()
    la32, sg32 = f32.logdet()
    lar, sgr = f32.logdet_refined(d64)
    for k, M in enumerate(mats):
        assert sgr[k] == sg64[k] == sg32[k]
        err, err32 = abs(lar[k] - la64[k]), abs(la32[k] - la64[k])
        assert err <= 5e-11 * abs(la64[k]) + 1e-10, (M.shape[0], err, err32)
        if err32 > 1e-6:
            assert err * 1000 <= err32, (M.shape[0], err, err32)


def test_lu_singular(bag, native):
    M = np.ones((40, 40))
    M[:, 3] = 0.0
    db = bag.DenseBatch([np.eye(3) * 2.0, M], native.PREC_F64).lu()
    la, sg = db.logdet()
    assert db.infos()[0] == 0 and db.infos()[1] > 0
    assert abs(la[0] - 3 * np.log(2.0)) < 1e-14 and sg[0] == 1
    assert la[1] == -np.inf and sg[1] == 0


@pytest.mark.parametrize("prec", [0, 1])
def test_lu_inverse(bag, native, prec):
    rng = np.random.default_rng(14)
    mats = _mats(rng, [3, 33, 97, 160], shift=3.0)
    inv = bag.DenseBatch(mats, prec).lu().inverse()
    tol = 1e-10 if prec == native.PREC_F64 else 5e-4
    for M, X in zip(mats, inv):
        ref = np.linalg.inv(M)
        assert np.max(np.abs(X - ref)) / np.max(np.abs(ref)) <= tol


def test_gemm_batched(bag, native):
    import ctypes as C
    from thirring2d_amd.hmc import Context, DeviceArray
    rng = np.random.default_rng(15)
    shapes = [(70, 50, 33), (5, 130, 17), (64, 64, 64)]
    blocks, jobs, off = [], [], 0
    for (M, Nn, K) in shapes:
        a, b, c = rng.normal(size=(M, K)), rng.normal(size=(K, Nn)), rng.normal(size=(M, Nn))
        jobs.append((off, off + a.size, off + a.size + b.size, K, Nn, Nn, M, Nn, K))
        blocks += [a.ravel(), b.ravel(), c.ravel()]
        off += a.size + b.size + c.size
    base = np.concatenate(blocks)
    for mfma in (1, 0):
        native.lib().th2d_set_dense_mfma(mfma)
        try:
            d = DeviceArray.from_host(base)
            J = (native.GemmJob * len(jobs))(*[native.GemmJob(*j) for j in jobs])
            dj = DeviceArray(C.sizeof(J))
            native.check(native.lib().th2d_h2d(Context.get().ptr, dj.p, C.cast(J, C.c_void_p),
                                               C.sizeof(J)), "h2d")
            native.check(native.lib().th2d_gemm_batched(Context.get().ptr, d.p, native.PREC_F64,
                                                        dj.p, len(jobs), 70, 130, -0.5), "gemm")
            out = d.download(base.shape, np.float64)
        finally:
            native.lib().th2d_set_dense_mfma(1)
        for (a0, b0, c0, lda, ldb, ldc, M, Nn, K) in jobs:
            A = base[a0:a0 + M * K].reshape(M, K)
            Bm = base[b0:b0 + K * Nn].reshape(K, Nn)
            Cc = base[c0:c0 + M * Nn].reshape(M, Nn)
            assert np.allclose(out[c0:c0 + M * Nn].reshape(M, Nn), Cc - 0.5 * A @ Bm,
                               rtol=1e-12, atol=1e-12)


def test_build_masked_matches_operator(bag, native, oracle):
    """The dense masked matrix acts as fM (identity rows/columns on occupied
    sites, vec_ops.c:107-131) for a batch of configurations."""
    import ctypes as C
    from thirring2d_amd.hmc import Context, DeviceArray
    rng = np.random.default_rng(16)
    dims = [(4, 4, 0), (6, 8, 1), (10, 6, 0)]
    geoms, fields, foff, ns, offs = [], [], [], [], []
    o = fo = 0
    for (nt, nx, bc) in dims:
        geoms.append(native.BagGeom(nt, nx, bc, 0, 0.4, 0.15))
        f = _random_field(nt, nx, 0.3, rng, bc)
        fields.append(f)
        foff.append(fo)
        fo += f.size
        ns.append(nt * nx)
        offs.append(o)
        o += (nt * nx) ** 2
    G = (native.BagGeom * len(geoms))(*geoms)
    dG = DeviceArray(C.sizeof(G))
    native.lib().th2d_h2d(Context.get().ptr, dG.p, C.cast(G, C.c_void_p), C.sizeof(G))
    dF = DeviceArray.from_host(np.concatenate(fields).astype(np.int32))
    dFo = DeviceArray.from_host(np.array(foff, np.int64))
    db = bag.DenseBatch([np.zeros((n, n)) for n in ns], native.PREC_F64)
    native.check(native.lib().th2d_bag_build_masked(Context.get().ptr, dG.p, dF.p, dFo.p,
                                                    C.byref(db.batch), db.A.p, native.PREC_F64),
                 "build_masked")
    for (nt, nx, bc), f, M in zip(dims, fields, db.matrices()):
        B = oracle.bag_lattice(nt, nx, bc=bc, m=0.4, mu=0.15)
        v = rng.normal(size=nt * nx)
        assert np.max(np.abs(M @ v - oracle.fM(B, f, v))) <= 1e-14
        assert np.max(np.abs(M.T @ v - oracle.fM(B, f, v, transpose=True))) <= 1e-14


@pytest.mark.parametrize("nt,nx,bc", [(4, 4, 0), (4, 6, 1), (8, 8, 0), (6, 10, 1)])
def test_calc_dinv(bag, oracle, nt, nx, bc):
    lat = bag.BagLattice(nt, nx, bc=bc, with_mass_monomers=1, m=0.0, mu=0.1)
    B = oracle.bag_lattice(nt, nx, bc=bc, with_mass_monomers=1, m=0.0, mu=0.1)
    assert np.max(np.abs(lat.calc_Dinv() - oracle.calc_Dinv(B))) <= 1e-12


@pytest.mark.parametrize("nt,nx,bc", [(4, 4, 0), (4, 6, 1), (4, 8, 0)])
def test_small_weights(bag, oracle, nt, nx, bc):
    rng = np.random.default_rng(17 + nx)
    lat = bag.BagLattice(nt, nx, bc=bc, with_mass_monomers=1, m=0.0, mu=0.1)
    B = oracle.bag_lattice(nt, nx, bc=bc, with_mass_monomers=1, m=0.0, mu=0.1)
    Dinv = oracle.calc_Dinv(B)
    cfgs, want = [], []
    for k in range(300):
        f = _random_field(nt, nx, 0.1 + 0.2 * (k % 4), rng, bc, monomers=k % 4)
        cfgs.append(f.astype(np.uint8))
        nl, nm = int(np.sum((f == 2) | (f == 3))), int(np.sum(f == 1))
        la, sg, n, r = oracle.bag_weight(B, f, Dinv)
        want.append(1.0 if nl + nm == 0 else (0.0 if (nm % 2 or r) else sg * np.exp(la)))
    got = lat.small_weights(Dinv, np.array(cfgs))
    want = np.array(want)
    assert np.count_nonzero(want) > 50
    assert np.max(np.abs(got - want) / np.maximum(1.0, np.abs(want))) <= 1e-10


def test_brute_table_matches_reference_run(bag, oracle):
    """4x4 OPENX, m=0, mu=0.1 (tests/parameter): the reference's recorded
    Z(0,2) = 9.96905 (|Z| 10.0567), Z(1,0) = 9.3976, Z(2,2) = 629.076 and
    45 non-zero cells, and the oracle's whole table."""
    lat = bag.BagLattice(4, 4, bc=bag.BC_OPENX, with_mass_monomers=1, m=0.0, mu=0.1)
    Z, Za, n = lat.brute_enumerate()
    B = oracle.bag_lattice(4, 4, bc=oracle.BC_OPENX, with_mass_monomers=1, m=0.0, mu=0.1)
    Zo, Zao, no = oracle.brute_enumerate(B)
    assert n == no
    assert Z[0, 0] == 1.0
    assert f"{Z[0, 2]:g}" == "9.96905" and f"{Za[0, 2]:g}" == "10.0567"
    assert f"{Z[1, 0]:g}" == "9.3976"
    assert f"{Z[2, 2]:g}" == "629.076"
    assert int((Z != 0).sum()) == 45
    scale = np.maximum(1.0, np.abs(Zao))
    assert np.max(np.abs(Z - Zo) / scale) <= 1e-10
    assert np.max(np.abs(Za - Zao) / scale) <= 1e-10


def _propose(rng, f, nt, nx, bc):
    s = int(rng.integers(nt * nx))
    t, x = divmod(s, nx)
    d = int(rng.integers(2))
    if d == 1 and bc == 1 and x == nx - 1:
        return None
    t2, x2 = ((t + 1) % nt, x) if d == 0 else (t, (x + 1) % nx)
    s2 = t2 * nx + x2
    if f[s] == 0 and f[s2] == 0:
        return [s, s2], [2 + d, 4 + d], []
    if f[s] == 2 + d and f[s2] == 4 + d:
        return [], [], [s, s2]
    if f[s] == 0 and f[s2] == 1 and rng.random() < 0.5:  # monomer hop onto s
        return [s], [1], [s2]
    return None


@pytest.mark.parametrize("nt,nx,bc,m", [(6, 6, 0, 0.0), (8, 6, 1, 0.0), (6, 8, 0, 0.3)])
def test_bag_ratio_sequence(bag, oracle, nt, nx, bc, m):
    """Rank-k determinant ratios and accepted updates of B = (D_FF)^-1
    against full LU recomputes (the principle of check_det,
    fluctuation_determinant.c:880-970)."""
    rng = np.random.default_rng(18 + nt + bc)
    wmm = 1 if m == 0.0 else 0
    lat = bag.BagLattice(nt, nx, bc=bc, with_mass_monomers=wmm, m=m)
    B = oracle.bag_lattice(nt, nx, bc=bc, with_mass_monomers=wmm, m=m)
    st = bag.BagState(lat)
    f = np.zeros(nt * nx, np.int32)
    if m > 0:
        f[[1, 8]] = 1
        st.set_field(f)
    la0, s0, _ = oracle.determinant_mu(B, f)
    la, sg, _, _ = st.logdet()
    assert abs(la - la0) <= 1e-10 and sg == s0
    tried = accepted = 0
    for _ in range(300):
        prop = _propose(rng, f, nt, nx, bc)
        if prop is None:
            continue
        occ, vals, fr = prop
        r = st.ratio(occ, fr)
        fn = f.copy()
        fn[fr] = 0
        fn[occ] = vals
        lo, so, _ = oracle.determinant_mu(B, f)
        ln, sn, _ = oracle.determinant_mu(B, fn)
        want = 0.0 if sn == 0 else so * sn * np.exp(ln - lo)
        if want == 0.0:
            assert abs(r) <= 1e-8
        else:
            assert abs(r - want) <= 1e-9 * abs(want)
        tried += 1
        if abs(r) > 1e-6 and rng.random() < 0.7:
            st.accept(occ, vals, fr)
            f = fn
            accepted += 1
            la, sg, _, _ = st.logdet()
            lt, stt, _ = oracle.determinant_mu(B, f)
            assert abs(la - lt) <= 1e-9 * max(1.0, abs(lt)) and sg == stt
    assert tried > 40 and accepted > 15
    # a refresh from scratch agrees with the updated state
    la, sg, _, _ = st.logdet()
    st.refresh()
    la2, sg2, _, _ = st.logdet()
    assert abs(la - la2) <= 1e-9 * max(1.0, abs(la2)) and sg == sg2
