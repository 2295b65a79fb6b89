"""The reference's own function signatures (include/thirring_hmc.h,
include/thirring_bag.h) called the way the reference's programs call them,
with jagged arrays and globals, against the CPU oracle.

Tolerances as the native tests: operator <= 1e-12 relative, solver
solutions at the requested tolerance, determinants 1e-10 relative in
log|det| and 1e-10 in phase, bag ratios 1e-9 relative; the random streams
and accept decisions are identical.
"""
import ctypes as C
import os

import numpy as np
import pytest

from jagged import Links, Rows, load

pytestmark = pytest.mark.gpu

P, D, I = C.c_void_p, C.c_double, C.c_int


class Cplx(C.Structure):
    """double complex returned by value (two SSE eightbytes, as a struct of
    two doubles in the x86-64 calling convention)."""
    _fields_ = [("re", C.c_double), ("im", C.c_double)]


def _rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.fixture(scope="module")
def hmc_lib():
    lib = load("libthirring_hmc.so")
    for n in ("fm_mul", "fm_conjugate_mul", "fmdm_mul", "fmdm_invert_cg", "fm_invert_cg",
              "stochastic_vector", "fermion_matrix", "update_gauge", "update_puregauge_hb"):
        getattr(lib, n).restype = None
    for n in ("calc_gauge_action", "gauge_step", "random_pseudofermion", "pseudofermion_action",
              "stochastic_md_action", "random_momentum", "momentum_step"):
        getattr(lib, n).restype = D
    lib.gauge_step.argtypes = [P, P, D]
    lib.momentum_step.argtypes = [P, P, P, P, D]
    lib.determinant.restype = Cplx
    lib.thirring_hmc_set_cg.argtypes = [D, D, I, I]
    lib.thirring_hmc_set_md.argtypes = [I, D]
    lib.thirring_hmc_last_trajectory.argtypes = [C.POINTER(I), C.POINTER(D), C.POINTER(I)]
    lib.seed_mersenne.argtypes = [C.c_long]
    lib.mersenne_generate.restype = D
    return lib


def _globals(lib, m, g, mu):
    C.c_double.in_dll(lib, "m").value = m
    C.c_double.in_dll(lib, "g").value = g
    C.c_double.in_dll(lib, "mu").value = mu


def _setup(lib, oracle, nt, nx, seed, m=0.1, g=0.5, mu=0.0):
    rng = np.random.default_rng(seed)
    V = nt * nx
    A = rng.normal(size=(V, 2))
    _globals(lib, m, g, mu)
    assert lib.thirring_hmc_init(nt, nx) == 0
    L = oracle.lattice(nt, nx, m=m, g=g, mu=mu)
    v = rng.normal(size=V) + 1j * rng.normal(size=V)
    return L, A, v


@pytest.mark.parametrize("nt,nx,mu", [(8, 8, 0.0), (12, 16, 0.3)])
def test_fm_mul_signatures(hmc_lib, oracle, nt, nx, mu):
    L, A, v = _setup(hmc_lib, oracle, nt, nx, 1 + nt, mu=mu)
    F = Links(nt, nx, A)
    vin, vout = Rows(nt, nx, np.complex128, v), Rows(nt, nx, np.complex128)
    hmc_lib.fm_mul(vin.ptr, vout.ptr, F.ptr)
    assert _rel(vout.natural(), oracle.fm_mul(L, A, v)) <= 1e-12
    hmc_lib.fm_conjugate_mul(vin.ptr, vout.ptr, F.ptr)
    assert _rel(vout.natural(), oracle.fm_dag_mul(L, A, v)) <= 1e-12
    hmc_lib.fmdm_mul(vin.ptr, vout.ptr, F.ptr)
    want = oracle.fm_dag_mul(L, A, oracle.fm_mul(L, A, v))
    assert _rel(vout.natural(), want) <= 1e-12
    assert hmc_lib.thirring_hmc_last_status() == 0


def test_invert_signatures(hmc_lib, oracle):
    nt, nx = 16, 16
    L, A, v = _setup(hmc_lib, oracle, nt, nx, 3)
    F = Links(nt, nx, A)
    vin, vout = Rows(nt, nx, np.complex128, v), Rows(nt, nx, np.complex128)
    hmc_lib.fmdm_invert_cg(vin.ptr, vout.ptr, F.ptr)  # reference stop: r.r < 1e-30
    x = vout.natural()
    xo, it, st = oracle.cg_mdm(L, A, v, tol_abs_rr=1e-30, tol_rel=0.0)
    assert _rel(x, xo) <= 1e-12
    hmc_lib.fm_invert_cg(vin.ptr, vout.ptr, F.ptr)  # M^-1 v
    assert _rel(oracle.fm_mul(L, A, vout.natural()), v) <= 1e-10
    assert hmc_lib.thirring_hmc_last_status() == 0


def test_fermion_matrix_and_determinant(hmc_lib, oracle):
    nt, nx = 6, 8
    L, A, v = _setup(hmc_lib, oracle, nt, nx, 5, mu=0.2)
    F = Links(nt, nx, A)
    A_glob = C.c_void_p.in_dll(hmc_lib, "A")
    A_glob.value = F.ptr.value
    try:
        V = nt * nx
        M = np.zeros(V * V, np.complex128)
        hmc_lib.fermion_matrix(M.ctypes.data_as(P))
        Mo = oracle.fermion_matrix(L, A)
        assert _rel(M.reshape(V, V, order="F"), Mo) <= 1e-14
        det = hmc_lib.determinant()
        z = complex(det.re, det.im)
        la, ph, info = oracle.zlogdet(Mo)
        assert info == 0
        assert abs(np.log(abs(z)) - la) <= 1e-10 * max(1.0, abs(la))
        assert abs(np.angle(z * np.exp(-1j * ph))) <= 1e-10
        s, l = np.linalg.slogdet(Mo)
        assert abs(l - la) <= 1e-10 * max(1.0, abs(l))
        assert abs(np.angle(z / abs(z) * np.conj(s))) <= 1e-10
    finally:
        A_glob.value = None


def test_random_fields_follow_the_reference_stream(hmc_lib, oracle):
    """random_pseudofermion, random_momentum, stochastic_vector draw the
    reference's MT stream in its order and normalisation (Re, Im ~ N(0,1))."""
    nt, nx = 8, 8
    L, A, _ = _setup(hmc_lib, oracle, nt, nx, 7)
    F = Links(nt, nx, A)
    hmc_lib.seed_mersenne(2001)
    rng = oracle.Rng(2001)
    v = Rows(nt, nx, np.complex128)
    S = hmc_lib.random_pseudofermion(v.ptr, F.ptr)
    So, xi, psi = oracle.random_pseudofermion(L, rng, A)
    assert abs(S - So) <= 1e-14 * So
    assert _rel(v.natural(), psi) <= 1e-12
    mom = Links(nt, nx)
    Sm = hmc_lib.random_momentum(mom.ptr)
    Smo, momo = oracle.random_momentum(L, rng)
    assert abs(Sm - Smo) <= 1e-14 * Smo and np.array_equal(mom.natural(), momo)
    st = Rows(nt, nx, np.complex128)
    hmc_lib.stochastic_vector(st.ptr)
    assert np.array_equal(st.natural(), oracle.stochastic_vector(L, rng))


def test_actions_and_steps(hmc_lib, oracle):
    nt, nx = 8, 12
    L, A, v = _setup(hmc_lib, oracle, nt, nx, 9, mu=0.25)
    hmc_lib.thirring_hmc_set_cg(1e-12, 0.0, 100000, 0)
    try:
        F = Links(nt, nx, A)
        psi = Rows(nt, nx, np.complex128, v)
        rng = np.random.default_rng(10)
        stv = rng.normal(size=nt * nx) + 1j * rng.normal(size=nt * nx)
        st = Rows(nt, nx, np.complex128, stv)
        assert abs(hmc_lib.calc_gauge_action(F.ptr) - oracle.gauge_action(L, A)) <= 1e-12 * abs(
            oracle.gauge_action(L, A))
        Spf = hmc_lib.pseudofermion_action(psi.ptr, F.ptr)
        Spfo, _ = oracle.pseudofermion_action(L, A, v, tol_rel=1e-12)
        assert abs(Spf - Spfo) <= 1e-10 * abs(Spfo)
        Smd = hmc_lib.stochastic_md_action(st.ptr, F.ptr)
        assert abs(Smd - oracle.stochastic_md_action(L, A, stv)) <= 1e-12 * abs(Smd)
        mom0 = rng.normal(size=(nt * nx, 2))
        mom = Links(nt, nx, mom0)
        hmc_lib.momentum_step(mom.ptr, F.ptr, psi.ptr, st.ptr, 0.05)
        momo, _, _ = oracle.momentum_step(L, mom0, A, v, stv, 0.05, tol_rel=1e-12)
        assert _rel(mom.natural(), momo) <= 1e-9
        hmc_lib.gauge_step(F.ptr, mom.ptr, 0.05)
        assert _rel(F.natural(), oracle.gauge_step(L, A, mom.natural(), 0.05)) <= 1e-14
        assert hmc_lib.thirring_hmc_last_status() == 0
    finally:
        hmc_lib.thirring_hmc_set_cg(0.0, 1e-30, 100000, 0)


def test_update_gauge_matches_oracle(hmc_lib, oracle):
    """update_gauge draws the reference's stream (pseudofermion, momentum,
    stochastic vector, accept) and runs the trajectory on the GPU: same
    dH, accept decisions and fields as the oracle, trajectory by trajectory."""
    nt, nx = 8, 8
    L, A, _ = _setup(hmc_lib, oracle, nt, nx, 11, m=0.3, g=0.5)
    hmc_lib.thirring_hmc_set_cg(1e-12, 0.0, 100000, 0)
    hmc_lib.thirring_hmc_set_md(20, 1.0)
    try:
        F = Links(nt, nx, A)
        hmc_lib.seed_mersenne(4001)
        rng = oracle.Rng(4001)
        Ao = A.copy()
        n_acc = 0
        for _ in range(4):
            hmc_lib.update_gauge(F.ptr)
            acc, dH, its = I(), D(), I()
            hmc_lib.thirring_hmc_last_trajectory(C.byref(acc), C.byref(dH), C.byref(its))
            acco, dSo, _, _, Ao = oracle.update_gauge(L, rng, Ao, nsteps=20, tol_rel=1e-12)
            assert acc.value == acco
            assert abs(dH.value - dSo) <= 1e-7 * max(1.0, abs(dSo))
            assert _rel(F.natural(), Ao) <= 1e-9
            n_acc += acco
        assert hmc_lib.thirring_hmc_last_status() == 0
        assert n_acc >= 1
    finally:
        hmc_lib.thirring_hmc_set_cg(0.0, 1e-30, 100000, 0)
        hmc_lib.thirring_hmc_set_md(10, 1.0)


def test_hmc_checkpoint_resumes_bitwise(hmc_lib, oracle, tmp_path):
    """thirring_hmc_write_checkpoint after two trajectories, then a third;
    a fresh start (other field, other seed) that reads the checkpoint
    repeats the third trajectory bit for bit: dH, accept and field.  The MT
    state in the file equals the stream's (the next draws of a generator
    restored from it match the oracle's stream)."""
    nt, nx = 16, 16
    L, A, _ = _setup(hmc_lib, oracle, nt, nx, 31, m=0.1, g=0.5)
    hmc_lib.thirring_hmc_set_cg(1e-12, 0.0, 100000, 0)
    hmc_lib.thirring_hmc_set_md(40, 1.0)
    hmc_lib.thirring_hmc_write_checkpoint.argtypes = [C.c_char_p, P, C.c_long]
    hmc_lib.thirring_hmc_read_checkpoint.argtypes = [C.c_char_p, P, C.POINTER(C.c_long)]
    path = str(tmp_path / "hmc.ckpt").encode()

    def traj(F):
        hmc_lib.update_gauge(F.ptr)
        acc, dH, its = I(), D(), I()
        hmc_lib.thirring_hmc_last_trajectory(C.byref(acc), C.byref(dH), C.byref(its))
        assert hmc_lib.thirring_hmc_last_status() == 0
        return acc.value, dH.value

    try:
        F = Links(nt, nx, A)
        hmc_lib.seed_mersenne(2001)
        for _ in range(2):
            traj(F)
        assert hmc_lib.thirring_hmc_write_checkpoint(path, F.ptr, 2) == 0
        acc3, dH3 = traj(F)
        A3 = F.natural()
        # the file's stream: 3 x 2V Gaussian pairs + 1 accept draw per
        # trajectory consumed from seed 2001 so far, twice
        rng = oracle.Rng(2001)
        rng.draws(2 * (3 * 2 * nt * nx + 1))
        # start over from somewhere else, then resume
        _globals(hmc_lib, 0.7, 0.9, 0.0)
        G = Links(nt, nx, np.zeros((nt * nx, 2)))
        hmc_lib.seed_mersenne(77)
        traj(G)
        n = C.c_long(-1)
        assert hmc_lib.thirring_hmc_read_checkpoint(path, G.ptr, C.byref(n)) == 0
        assert n.value == 2
        assert C.c_double.in_dll(hmc_lib, "m").value == 0.1
        acc, dH = traj(G)
        assert acc == acc3 and dH == dH3
        assert np.array_equal(G.natural(), A3)
        # and the generator continues the reference stream draw for draw
        rng.draws(3 * 2 * nt * nx + 1)
        mi = C.c_int.in_dll(hmc_lib, "mersenne_i")
        arr = (C.c_double * 624).in_dll(hmc_lib, "mersenne_array")
        nxt = []
        for _ in range(700):
            if mi.value > 0:
                mi.value -= 1
                nxt.append(arr[mi.value])
            else:
                nxt.append(hmc_lib.mersenne_generate())
        assert np.array_equal(np.array(nxt), rng.draws(700))
        assert hmc_lib.thirring_hmc_read_checkpoint(b"/nonexistent/x", G.ptr, C.byref(n)) != 0
    finally:
        hmc_lib.thirring_hmc_set_cg(0.0, 1e-30, 100000, 0)
        hmc_lib.thirring_hmc_set_md(10, 1.0)


def test_puregauge_matches_oracle(hmc_lib, oracle):
    nt, nx = 8, 10
    L, A, _ = _setup(hmc_lib, oracle, nt, nx, 13)
    F = Links(nt, nx, A)
    hmc_lib.seed_mersenne(5)
    rng = oracle.Rng(5)
    Ao = A
    for _ in range(3):
        hmc_lib.update_puregauge_hb(F.ptr)
        Ao = oracle.puregauge_sweep(L, rng, Ao)
    assert _rel(F.natural(), Ao) <= 1e-15


# ------------------------------------------------------------- bag family --
@pytest.fixture(scope="module")
def bag_lib():
    lib = load("libthirring_bag.so")
    for n in ("fM_index", "determinant", "determinant_mu", "det_added_link", "det_removed_link",
              "det_moved_monomer"):
        getattr(lib, n).restype = D
    lib.fM_index.argtypes = [I, I, I, I, D]
    lib.determinant_mu.argtypes = [D]
    for n in ("fM", "fM_transpose", "cg_propagator", "calc_Dinv", "update_background",
              "new_link", "new_monomer", "removed_link", "removed_monomer",
              "moved_source_monomer", "not_moved_source_monomer"):
        getattr(lib, n).restype = None
    return lib


def _bag_setup(lib, oracle, nt, nx, bc, wmm, m, mu, f):
    C.c_double.in_dll(lib, "m").value = m
    C.c_double.in_dll(lib, "mu").value = mu
    assert lib.thirring_bag_init(nt, nx, bc, wmm) == 0
    field = Rows(nt, nx, np.int32, f)
    C.c_void_p.in_dll(lib, "field").value = field.ptr.value
    B = oracle.bag_lattice(nt, nx, bc=bc, with_mass_monomers=wmm, m=m, mu=mu)
    return field, B


def test_bag_operator_signatures(bag_lib, oracle):
    nt, nx = 8, 6
    rng = np.random.default_rng(21)
    f = np.zeros(nt * nx, np.int32)
    f[[3, 4]] = [2 + 1, 4 + 1]
    f[[17, 29]] = 1
    field, B = _bag_setup(bag_lib, oracle, nt, nx, 0, 0, 0.3, 0.2, f)
    try:
        psi = rng.normal(size=nt * nx)
        pin, out = Rows(nt, nx, np.float64, psi), Rows(nt, nx, np.float64)
        bag_lib.fM(out.ptr, pin.ptr)  # (out, in) as the reference
        assert np.max(np.abs(out.natural() - oracle.fM(B, f, psi))) <= 1e-14
        bag_lib.fM_transpose(out.ptr, pin.ptr)
        assert np.max(np.abs(out.natural() - oracle.fM(B, f, psi, transpose=True))) <= 1e-14
        bag_lib.cg_propagator(out.ptr, pin.ptr)
        xo, it, st = oracle.cg_propagator(B, f, psi, tol_abs_rr=1e-30)
        assert np.max(np.abs(out.natural() - xo)) <= 1e-10
        for (t1, x1, t2, x2) in [(0, 0, 1, 0), (2, 3, 2, 4), (7, 5, 0, 5), (3, 0, 3, 5), (4, 4, 4, 4)]:
            assert bag_lib.fM_index(t1, x1, t2, x2, 0.2) == oracle.fM_index(B, t1, x1, t2, x2, 0.2)
        la, sg, _ = oracle.determinant_mu(B, f, 0.2)
        assert abs(np.log(abs(bag_lib.determinant())) - la) <= 1e-10 * max(1.0, abs(la))
        assert np.sign(bag_lib.determinant()) == sg
        assert bag_lib.thirring_bag_last_status() == 0
    finally:
        C.c_void_p.in_dll(bag_lib, "field").value = None


def test_bag_calc_dinv_signature(bag_lib, oracle, tmp_path, monkeypatch):
    """calc_Dinv equals the oracle's propagator, writes the reference's
    cache file (fermion_matrix.c:206-220,342-346: named for the lattice and
    mu, V*V/2 doubles), and a second call reads that file instead of
    computing (a changed file shows up in Dinv)."""
    monkeypatch.chdir(tmp_path)
    nt, nx = 6, 6
    field, B = _bag_setup(bag_lib, oracle, nt, nx, 0, 1, 0.0, 0.1, np.zeros(36, np.int32))
    h = nt * nx // 2
    buf = np.zeros(nt * nx * nt * nx)
    g = C.c_void_p.in_dll(bag_lib, "Dinv")
    g.value = buf.ctypes.data
    try:
        bag_lib.calc_Dinv()
        ref = oracle.calc_Dinv(B)
        assert np.max(np.abs(buf[:2 * h * h] - ref)) <= 1e-12
        cache = tmp_path / "free_propagator_eo_T6X6_mu0.100000"
        assert cache.exists() and cache.stat().st_size == 8 * 2 * h * h
        assert np.array_equal(np.fromfile(cache, dtype=np.float64), buf[:2 * h * h])
        changed = 2.0 * buf[:2 * h * h]
        changed.tofile(cache)
        buf[:] = 0.0
        bag_lib.calc_Dinv()
        assert np.array_equal(buf[:2 * h * h], changed)
        assert bag_lib.thirring_bag_last_status() == 0
    finally:
        g.value = None
        C.c_void_p.in_dll(bag_lib, "field").value = None


@pytest.mark.parametrize("bc", [0, 1])
def test_bag_calc_dinv_without_mass_monomers(bag_lib, oracle, tmp_path, monkeypatch, bc):
    """Without mass monomers calc_Dinv (fermion_matrix.c:355-425) gives the
    full V x V inverse in natural order, Dinv[i1*V + i2] = (D^-1)_{i1,i2},
    against the inverse of the oracle's fM_index matrix <= 1e-12; the cache
    file is named for the lattice, m and mu (runs at different masses never
    share it) and holds all V*V doubles; a short file is ignored."""
    monkeypatch.chdir(tmp_path)
    nt, nx = 6, 8
    V = nt * nx
    m, mu = 0.3, 0.1
    field, B = _bag_setup(bag_lib, oracle, nt, nx, bc, 0, m, mu, np.zeros(V, np.int32))
    buf = np.zeros(V * V)
    g = C.c_void_p.in_dll(bag_lib, "Dinv")
    g.value = buf.ctypes.data
    try:
        D_full = oracle.reduced_matrix(B, np.zeros(V, np.int32))
        ref = np.linalg.inv(D_full).ravel()
        name = f"free_propagator_T{nt}X{nx}_m{m:.6f}_mu{mu:.6f}" + ("_open" if bc else "")
        cache = tmp_path / name
        np.zeros(V * V // 2).tofile(cache)  # a half-length file: ignored
        bag_lib.calc_Dinv()
        assert np.max(np.abs(buf - ref)) <= 1e-12 * np.max(np.abs(ref))
        assert cache.stat().st_size == 8 * V * V
        assert np.array_equal(np.fromfile(cache, dtype=np.float64), buf)
        changed = 2.0 * buf
        changed.tofile(cache)
        buf[:] = 0.0
        bag_lib.calc_Dinv()
        assert np.array_equal(buf, changed)
        assert bag_lib.thirring_bag_last_status() == 0
        # another mass is another file
        C.c_double.in_dll(bag_lib, "m").value = 0.5
        bag_lib.calc_Dinv()
        assert (tmp_path / name.replace("m0.300000", "m0.500000")).exists()
        assert not np.array_equal(buf, changed)
    finally:
        g.value = None
        C.c_void_p.in_dll(bag_lib, "field").value = None


def test_config_checkpoint(bag_lib, oracle, tmp_path, monkeypatch):
    """write_config / read_config (fermionbag.c:120-161): the file is named
    for the lattice, U, m and mu and holds NT*NX int32 at t*NX + x; reading
    it back restores the field."""
    monkeypatch.chdir(tmp_path)
    nt, nx = 8, 6
    rng = np.random.default_rng(31)
    f = rng.integers(0, 7, size=nt * nx).astype(np.int32)
    C.c_double.in_dll(bag_lib, "U").value = 0.5
    field, B = _bag_setup(bag_lib, oracle, nt, nx, 0, 1, 0.1, 0.0, f)
    try:
        bag_lib.write_config()
        path = tmp_path / "config_checkpoint_T8X6_U0.5m0.1mu0"
        assert path.exists()
        assert np.array_equal(np.fromfile(path, dtype=np.int32), f)
        other = Rows(nt, nx, np.int32, np.zeros(nt * nx, np.int32))
        C.c_void_p.in_dll(bag_lib, "field").value = other.ptr.value
        bag_lib.read_config(str(path).encode())
        assert np.array_equal(other.natural(), f)
        assert bag_lib.thirring_bag_last_status() == 0
    finally:
        C.c_void_p.in_dll(bag_lib, "field").value = None


def test_read_config_missing_file_exits(tmp_path):
    """A configuration file that cannot be opened ends the program with the
    reference's message and status 1 (fermionbag.c:149-155)."""
    import subprocess
    import sys
    code = (
        "import ctypes as C, numpy as np, sys\n"
        f"sys.path.insert(0, {os.path.dirname(os.path.dirname(os.path.abspath(__file__)))!r})\n"
        f"sys.path.insert(0, {os.path.dirname(os.path.abspath(__file__))!r})\n"
        "from jagged import load, Rows\n"
        "lib = load('libthirring_bag.so')\n"
        "assert lib.thirring_bag_init(4, 4, 0, 1) == 0\n"
        "f = Rows(4, 4, np.int32, np.zeros(16, np.int32))\n"
        "C.c_void_p.in_dll(lib, 'field').value = f.ptr.value\n"
        "lib.read_config(b'no_such_configuration')\n"
        "print('returned')\n")
    r = subprocess.run([sys.executable, "-c", code], cwd=tmp_path, capture_output=True, text=True,
                       timeout=240)
    assert r.returncode == 1, r.stdout + r.stderr
    assert "Could not read configuration" in r.stdout and "returned" not in r.stdout


def test_bag_update_sequence(bag_lib, oracle):
    """det_added_link / det_removed_link / det_moved_monomer and the accept
    hooks, driven like fermionbag.c's add_link / remove_link / worm, against
    full determinants of the oracle; the hooks keep field, n_monomers,
    n_links and current_sign as the reference does."""
    nt, nx = 6, 6
    V = nt * nx
    f = np.zeros(V, np.int32)
    field, B = _bag_setup(bag_lib, oracle, nt, nx, 0, 0, 0.4, 0.0, f)
    C.c_int.in_dll(bag_lib, "max_changes").value = 7
    nmono = C.c_int.in_dll(bag_lib, "n_monomers")
    nlinks = C.c_int.in_dll(bag_lib, "n_links")
    nmono.value = nlinks.value = 0
    want_mono = want_links = 0
    sign = C.c_int.in_dll(bag_lib, "current_sign")
    try:
        bag_lib.update_background()
        rng = np.random.default_rng(23)
        tried = acc = 0
        kinds = set()

        def det(ff):
            la, sg, _ = oracle.determinant_mu(B, ff)
            return la, sg

        la_d, sg_d = det(np.zeros(V, np.int32))
        for _ in range(400):
            cur = field.natural().astype(np.int32)
            s = int(rng.integers(V))
            t, x = divmod(s, nx)
            nu = int(rng.integers(2))
            t2, x2 = ((t + 1) % nt, x) if nu == 0 else (t, (x + 1) % nx)
            s2 = t2 * nx + x2
            la0, sg0 = det(cur)
            new = cur.copy()
            if cur[s] == 0 and cur[s2] == 0:
                kind = int(rng.integers(2))
                r = bag_lib.det_added_link(t, x, t2, x2)
                new[s], new[s2] = (2 + nu, (nu + 2) % 4 + 2) if kind == 0 else (1, 1)
            elif cur[s] == 2 + nu and cur[s2] == (nu + 2) % 4 + 2:
                kind = 2
                r = bag_lib.det_removed_link(t, x, t2, x2)
                new[s] = new[s2] = 0
            elif cur[s] == 1 and cur[s2] == 1:
                kind = 4
                r = bag_lib.det_removed_link(t, x, t2, x2)
                new[s] = new[s2] = 0
            elif cur[s] in (1, 6) and rng.random() < 0.5:
                # a monomer moves by two sites, to a site of its parity
                dt, dx = [(2, 0), (0, 2), (1, 1), (1, -1)][int(rng.integers(4))]
                t2, x2 = (t + dt) % nt, (x + dx) % nx
                s2 = t2 * nx + x2
                if cur[s2] != 0:
                    continue
                kind = 3
                r = bag_lib.det_moved_monomer(t, x, t2, x2)
                new[s], new[s2] = 0, 6
            else:
                continue
            la1, sg1 = det(new)
            want = np.exp(la1 - la0)
            assert abs(r - want) <= 1e-9 * want
            tried += 1
            if rng.random() < 0.6:
                if kind == 0:
                    bag_lib.new_link(t, x, nu)  # n_links is the caller's (fermionbag.c)
                elif kind == 1:
                    bag_lib.new_monomer(t, x, t2, x2)
                    want_mono += 2
                elif kind == 2:
                    bag_lib.removed_link(t, x, nu)
                    want_links -= 1
                elif kind == 4:
                    bag_lib.removed_monomer(t, x, nu)
                    want_mono -= 2
                else:
                    bag_lib.moved_source_monomer(t, x, t2, x2)
                kinds.add(kind)
                acc += 1
                assert np.array_equal(field.natural(), new)
                assert sign.value == sg1 * sg_d
        assert bag_lib.thirring_bag_last_status() == 0
        assert tried > 40 and acc > 20 and kinds == {0, 1, 2, 3, 4}
        assert nmono.value == want_mono and nlinks.value == want_links
    finally:
        C.c_void_p.in_dll(bag_lib, "field").value = None
        C.c_int.in_dll(bag_lib, "max_changes").value = 0


@pytest.mark.parametrize("wmm,m", [(1, 0.0), (0, 0.3)])
def test_worm_susceptibility_matches_oracle(bag_lib, oracle, wmm, m):
    """thirring_bag_worm (measure_susceptibility, fermionbag.c:450-590) on
    the program's field with the device ratios and hooks, against the
    oracle's worm with full determinants: the same MT stream gives the same
    summed signs, printed value, final field, n_links, n_monomers and
    current_sign, measurement after measurement."""
    nt = nx = 6
    V = nt * nx
    rng = np.random.default_rng(91 + wmm)
    Bo = oracle.bag_lattice(nt, nx, with_mass_monomers=wmm, m=m)
    while True:  # a configuration with det D_FF != 0 (m = 0 allows zeros)
        f0 = np.zeros(V, np.int32)
        for s in rng.permutation(V):
            if f0[s] or rng.random() > 0.35:
                continue
            t, x = divmod(int(s), nx)
            d = int(rng.integers(2))
            s2 = ((t + 1) % nt) * nx + x if d == 0 else t * nx + (x + 1) % nx
            if not f0[s2]:
                f0[s], f0[s2] = 2 + d, 4 + d
        if oracle.determinant_mu(Bo, f0)[1] != 0:
            break
    field, B = _bag_setup(bag_lib, oracle, nt, nx, 0, wmm, m, 0.0, f0)
    U = 1.3
    C.c_double.in_dll(bag_lib, "U").value = U
    nl = C.c_int.in_dll(bag_lib, "n_links")
    nm = C.c_int.in_dll(bag_lib, "n_monomers")
    sign = C.c_int.in_dll(bag_lib, "current_sign")
    nl.value = int(np.count_nonzero((f0 == 2) | (f0 == 3)))
    nm.value = 0
    bag_lib.thirring_bag_worm.restype = C.c_long
    bag_lib.thirring_bag_worm.argtypes = [I, C.POINTER(D)]
    bag_lib.seed_mersenne.argtypes = [C.c_long]
    try:
        bag_lib.update_background()
        bag_lib.seed_mersenne(55)
        r = oracle.Rng(55)
        f, o_sign, o_nl, o_nm = f0.copy(), sign.value, nl.value, 0
        _, s_ff, _ = oracle.determinant_mu(B, f0)
        _, s_d, _ = oracle.determinant_mu(B, np.zeros(V, np.int32))
        assert o_sign == s_ff * s_d
        for _ in range(4):
            v = D()
            steps = bag_lib.thirring_bag_worm(5, C.byref(v))
            o_steps, o_val, f, o_nl, o_nm, o_sign = oracle.worm(
                B, f, r, U, n_attempts=5, n_links=o_nl, n_monomers=o_nm, sign=o_sign)
            assert bag_lib.thirring_bag_last_status() == 0
            assert steps == o_steps and abs(v.value - o_val) <= 1e-15 * max(1.0, abs(o_val))
            assert np.array_equal(field.natural(), f)
            assert (nl.value, nm.value, sign.value) == (o_nl, o_nm, o_sign)
    finally:
        C.c_void_p.in_dll(bag_lib, "field").value = None
