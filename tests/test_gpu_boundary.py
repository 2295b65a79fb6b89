"""The drop-in boundary as a C program uses it: tests/native/program_hmc.c
and program_bag.c define the reference's globals, counters, hooks and their
own MT19937 stream, and are linked with gcc against the compat libraries
exactly as INTEGRATION.md shows (-lthirring_hmc -lth2d, -lthirring_bag
-lth2d).  The program's definitions must replace the libraries' weak ones:
the library reads the program's m, draws from the program's generator,
fills the program's Dinv and field, calls its update_linklists and writes
its current_sign.  Results are checked against the oracle."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "thirring2d_amd")


def _build_and_run(tmp_path, name, lib, args=()):
    src = os.path.join(ROOT, "tests", "native", name + ".c")
    exe = str(tmp_path / name)
    subprocess.run(["gcc", "-std=gnu11", "-O2", "-Wall", "-I" + os.path.join(ROOT, "include"),
                    "-o", exe, src, "-L" + PKG, "-l" + lib, "-lth2d", "-Wl,-rpath," + PKG, "-lm"],
                   check=True)
    p = subprocess.run([exe, *args], cwd=tmp_path, stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=120)
    assert p.returncode == 0, p.stdout
    out = {}
    for line in p.stdout.splitlines():
        if "=" in line:
            k, v = line.split("=", 1)
            out.setdefault(k, []).append(v.split())
    return out


def test_program_owned_hmc_globals_and_stream(tmp_path, oracle):
    out = _build_and_run(tmp_path, "program_hmc", "thirring_hmc", [str(tmp_path / "c.ckpt")])
    nt = nx = 8
    V = nt * nx
    r0, r1 = map(float, out["mass_response"][0])
    # (M delta_0)(0) = m: the library follows the program's m
    assert abs(r0 - 0.3) <= 1e-15 and abs(r1 - 0.7) <= 1e-15
    # update_gauge drew from the program's generator, in the reference order
    assert int(out["generate_calls"][0][0]) > 0
    assert int(out["status"][0][0]) == 0
    t, x = np.divmod(np.arange(V), nx)
    A = np.stack([0.3 * np.sin(1.0 + t + 2.0 * x), 0.3 * np.cos(2.0 * t - x)], axis=1)
    L = oracle.lattice(nt, nx, m=0.3, g=0.5)
    rng = oracle.Rng(4001)
    for k, (kk, acc, dH) in enumerate(out["traj"]):
        acco, dSo, _, _, A = oracle.update_gauge(L, rng, A, nsteps=20, tol_rel=1e-12)
        assert int(acc) == acco
        assert abs(float(dH) - dSo) <= 1e-7 * max(1.0, abs(dSo))
    Ag = np.array(out["A"][0], dtype=float).reshape(V, 2)
    assert np.max(np.abs(Ag - A)) <= 1e-9
    # the checkpoint: written; read back with the program's own generator,
    # whose state words are not the library's to set
    assert int(out["ckpt_write"][0][0]) == 0
    rd, n = out["ckpt_read"][0]
    assert int(rd) == 1 and int(n) == 3


def test_program_owned_bag_state(tmp_path, oracle):
    out = _build_and_run(tmp_path, "program_bag", "thirring_bag")
    nt = nx = 6
    V = nt * nx
    B = oracle.bag_lattice(nt, nx, m=0.4)
    empty = np.zeros(V, np.int32)
    Dfull = np.linalg.inv(oracle.reduced_matrix(B, empty))
    dinv = np.array(out["dinv"][0], dtype=float)
    assert np.max(np.abs(dinv - Dfull.ravel()[:8])) <= 1e-12 * np.max(np.abs(Dfull))
    f0 = empty.copy()
    f0[2 * nx + 3], f0[3 * nx + 3] = 2, 4
    psi = np.sin(1.0 + np.arange(V))
    assert np.max(np.abs(np.array(out["fM"][0], dtype=float) - oracle.fM(B, f0, psi))) <= 1e-14

    def det(f):
        la, sg, _ = oracle.determinant_mu(B, f)
        return la, sg

    f1 = f0.copy()
    f1[0], f1[1] = 3, 5  # link in direction x from (0,0) to (0,1)
    f2 = f1.copy()
    f2[4 * nx + 1], f2[5 * nx + 1] = 2, 4
    (l0, _), (l1, _), (l2, s2) = det(f0), det(f1), det(f2)
    a, b = map(float, out["added"][0])
    r = float(out["removed"][0][0])
    assert abs(a - np.exp(l1 - l0)) <= 1e-9 * np.exp(l1 - l0)
    assert abs(r - np.exp(l0 - l1)) <= 1e-9 * np.exp(l0 - l1)
    assert abs(b - np.exp(l2 - l1)) <= 1e-9 * np.exp(l2 - l1)
    # the program's hooks and state
    assert int(out["linklists"][0][0]) >= 2
    _, sd = det(empty)
    assert int(out["sign"][0][0]) == s2 * sd
    assert np.array_equal(np.array(out["field"][0], dtype=int), f2)
    assert int(out["status"][0][0]) == 0
