"""Benchmark: HMC trajectories per second at 256x256 (mixed precision and
FP64, CG true relative residual <= 1e-10), with the Dslash bandwidth.

Workload (one GPU): 256x256 lattice, antiperiodic t and x, m = 0.05,
g = 0.5.  Every chain starts as hmc.c's main does (hmc.c:927-929): A = 0,
then 100 pure-gauge Metropolis sweeps (device draws), then the warm-up
trajectories.  A trajectory is 128 leapfrog steps of dt = 1/128 (the
reference's 10 x 0.1, hmc.c:708-709, is beyond the leapfrog's stability
limit at this mass and volume: it is timed as a separate, flagged leg),
with pseudofermion / momentum / stochastic noise from the device generator
and every CG solve to a true relative residual of 1e-10.  One step = one
trajectory of every chain; accepted trajectories move the chains.

Prints one JSON line.  `value` is the aggregate trajectories per second of
a batch of independent 256^2 chains advanced in lockstep (every inversion
of all chains is one batched solve) with mixed precision (FP32 CG, FP64
true-residual restarts); the batched FP64 path and single chains in FP64
and mixed precision are timed beside it.  `roofline` is for the dominant
kernel, timed with HIP events on its stream: the resident FP32 solver (one
cooperative launch per inner solve), whose bound is its FP32 arithmetic
and per-iteration synchronisation, not HBM (its vectors stay in
registers).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
# measured HBM bytes per launch of the dominant kernel (FETCH_SIZE + WRITE_SIZE,
# copy-calibrated): k_cg_resident at 16 thermalized chains, 128 x dt = 1/128
# (profiles/k_cg_resident_hmc256_16chains_mixed.md), k_cg_step at 32 chains
# (profiles/k_cg_step_hmc256_32chains_mixed.md)
# key: (lattice size, chains, resident solver)
TRAFFIC_BYTES = {(256, 32, False): 199.6e6, (256, 16, True): 165.1e6}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--chains", type=int, default=16)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-dslash", action="store_true")
    ap.add_argument("--dslash-sweep", action="store_true",
                    help="only the Dslash size sweep 256^2..4096^2 (FP64, FP32; 1 and 8 RHS at 256^2)")
    ap.add_argument("--dense-only", action="store_true",
                    help="only the dense leg: batched LU of 64 bag matrices of order 256..2304 "
                         "(FP64, FP32), then --c5-proposals bag proposals")
    ap.add_argument("--only-headline", action="store_true",
                    help="time the batched mixed-precision leg only (for profiling)")
    ap.add_argument("--nsteps", type=int, default=128,
                    help="leapfrog steps per trajectory of the headline (dt = 1/nsteps)")
    ap.add_argument("--ref-md-steps", type=int, default=3,
                    help="trajectories of the reference integrator leg, 10 x dt = 0.1 (0 = skip)")
    ap.add_argument("--no-c2", action="store_true", help="skip the 8-RHS solve leg")
    ap.add_argument("--no-c3", action="store_true", help="skip the 128^2 force leg")
    ap.add_argument("--oracle-apply", action="store_true",
                    help="time the oracle's fm_mul at 256^2 on this host and exit")
    ap.add_argument("--c4-steps", type=int, default=50,
                    help="trajectories per chain of the 256-chain uneven batch (0 = skip)")
    ap.add_argument("--c5-proposals", type=int, default=150,
                    help="bag proposals per matrix in the dense leg (0 = skip the leg)")
    ap.add_argument("--c1-steps", type=int, default=100,
                    help="16^2 FP64 trajectories (0 = skip)")
    ap.add_argument("--only-legs", default="", metavar="LIST",
                    help="run only these legs, comma separated, and exit: c1, c4, "
                         "bag-driver (for profiling)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="run a plain device copy of MIB MiB five times and exit (scale of "
                         "the FETCH_SIZE / WRITE_SIZE counters)")
    return ap.parse_args()


def copy_rate(mib):
    """Five device copies of MIB MiB, then twenty timed with HIP events:
    (ms per copy, read+write GB/s), the rate a plain one-pass stream reaches
    on this GPU."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(20):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / 20
    del src, dst
    torch.cuda.empty_cache()
    return ms, 2 * n * 4 / ms / 1e6


def copy_calibration(mib):
    """The copies of copy_rate: the scale of the FETCH_SIZE and WRITE_SIZE
    counters, and the copy rate."""
    ms, gbs = copy_rate(mib)
    print(json.dumps({"copy_calibration_bytes": mib * (1 << 20), "launches": 25,
                      "timed_launches": 20, "ms_per_copy": ms, "GB/s_read_plus_write": gbs}),
          flush=True)


def run_batch(H, N, lats, gs, precision, steps, warmup, nsteps=10, ctr0=0):
    """Time `steps` trajectories of every chain of `lats` (one batched call
    per step, accept draws from the host: accepted trajectories move the
    chains).  Returns (seconds, per-trajectory records, kernel ms, kernel
    launches, kernel bytes)."""
    import ctypes as C
    batch = H.ChainBatch(lats)
    for l, g in zip(lats, gs):
        l.g = g
    params = H.cg_params(tol_rel=1e-10, precision=precision, max_iter=200000)
    L = N.lib()
    nch = len(lats)
    ctr = ctr0
    for _ in range(warmup):
        u = list(np.random.default_rng(10_000 + ctr).random(nch))
        batch.update_gauge(nsteps=nsteps, traj_len=1.0, params=params, seed=4242,
                           counters=[ctr] * nch, u_accept=u)
        ctr += 1
    L.th2d_kernel_timing(batch.ctx.ptr, 1)
    batch.ctx.sync()
    t0 = time.perf_counter()
    res = []
    for _ in range(steps):
        u = list(np.random.default_rng(ctr).random(nch))
        rr = batch.update_gauge(nsteps=nsteps, traj_len=1.0, params=params, seed=4242,
                                counters=[ctr] * nch, u_accept=u)
        ctr += 1
        res += [(r.dS, r.accepted, r.cg_iterations, r.cg_solves, r.max_true_rel_resid, r.status,
                 r.acts[1]) for r in rr]
    batch.ctx.sync()
    dt = time.perf_counter() - t0
    ms, n, by = C.c_double(), C.c_long(), C.c_double()
    L.th2d_kernel_time(batch.ctx.ptr, C.byref(ms), C.byref(n), C.byref(by))
    L.th2d_kernel_timing(batch.ctx.ptr, 0)
    return dt, res, ms.value, n.value, by.value


def thermalized(H, size, chains, m=0.05, g=0.5, seed=1002):
    """Chains started as hmc.c:927-929: A = 0, 100 pure-gauge sweeps."""
    lats = []
    for c in range(chains):
        lat = H.StaggeredLattice(size, size, m=m, g=g)
        lat.thermalize_pure_gauge(100, seed=seed + c)
        lats.append(lat)
    return lats


def run_hmc(H, N, size, precision, steps, warmup, chains=1, m=0.05, g=0.5, nsteps=128,
            lats=None):
    lats = lats or thermalized(H, size, chains, m, g)
    return run_batch(H, N, lats, [g] * len(lats), precision, steps, warmup, nsteps=nsteps)


def copy_fields(H, lats):
    """Fresh lattices holding the same fields."""
    out = []
    for l in lats:
        n = H.StaggeredLattice(l.nt, l.nx, m=l.m, g=l.g)
        n.set_links(l.get_links())
        out.append(n)
    return out


def mixed_vs_fp64(H, N, lats, nsteps, tau=1e-10):
    """One trajectory of every chain with mixed and with FP64 solves from
    the same fields and noise: per-chain |dH_mixed - dH_fp64| against the
    bound 4 tau |xi|^2 sqrt(m^2+4)/m (two actions, each off by at most
    |M^-1 xi| |r|)."""
    out = {}
    for prec in (N.PREC_MIXED, N.PREC_F64):
        b = H.ChainBatch(copy_fields(H, lats))
        out[prec] = b.update_gauge(nsteps=nsteps, params=H.cg_params(tol_rel=tau, precision=prec,
                                                                     max_iter=200000),
                                   seed=777, counters=[5] * len(lats),
                                   u_accept=[0.5] * len(lats))
    m = lats[0].m
    d = [abs(a.dS - b.dS) for a, b in zip(out[N.PREC_MIXED], out[N.PREC_F64])]
    bound = [4 * tau * b.acts[1] * np.sqrt(m * m + 4) / m for b in out[N.PREC_F64]]
    return {"chains": len(lats), "max_abs_dH_diff": float(max(d)),
            "max_diff_over_bound": float(max(x / y for x, y in zip(d, bound))),
            "dH_fp64": [float(r.dS) for r in out[N.PREC_F64]],
            "dH_mixed": [float(r.dS) for r in out[N.PREC_MIXED]],
            "bound": "4 tau |xi|^2 sqrt(m^2+4)/m per chain, tau = 1e-10"}


def c4_batch(H, start):
    """The C4 chains: 256 lattices, L uniform on {32,48,64,96,128} (seed
    4004), g ~ U[0.2,1], m = 0.1; A ~ N(0,1) per chain (start "gauss", the
    configuration as specified) or A = 0 plus 100 pure-gauge sweeps as
    hmc.c:927-929 starts (start "thermal")."""
    rng = np.random.default_rng(4004)
    Ls = rng.choice([32, 48, 64, 96, 128], size=256)
    gs = rng.uniform(0.2, 1.0, size=256)
    lats = []
    for c, (Lc, gc) in enumerate(zip(Ls, gs)):
        lat = H.StaggeredLattice(int(Lc), int(Lc), m=0.1, g=float(gc))
        if start == "gauss":
            lat.random_links(40040 + c)
        else:
            lat.thermalize_pure_gauge(100, seed=40040 + c)
        lats.append(lat)
    return Ls, gs, lats


def c4_run(H, N, lats, gs, Ls, steps, nsteps):
    dt, res, kms, kn, kby = run_batch(H, N, lats, list(gs), N.PREC_MIXED, steps, 1,
                                      nsteps=nsteps)
    dH = np.array([r[0] for r in res])
    return {"value": 256 * steps / dt, "unit": "traj/s", "trajectories": 256 * steps,
            "nsteps": nsteps, "dt": 1.0 / nsteps,
            "acceptance": float(np.mean([r[1] for r in res])),
            "exp_minus_dH_mean": float(np.mean(np.exp(-np.minimum(dH, 50.0)))),
            "dH_median": float(np.median(dH)),
            "max_true_rel_resid": max(r[4] for r in res),
            "cg_kernel_GBs": kby / (kms * 1e-3) / 1e9 if kms > 0 else 0.0}


def c4_leg(H, N, steps, nsteps=64):
    """C4: 256 chains, mixed CG to 1e-10, autotuned tiles, `steps`
    trajectories each.
    - as specified: A ~ N(0,1), 10 x dt = 0.1 (hmc.c:708-709; beyond the
      leapfrog stability limit for most chains, so nearly every trajectory
      is rejected and the fields stay put);
    - stable: thermalized starts, nsteps x dt = 1/nsteps;
    - one trajectory of every stable chain with mixed and with FP64 solves
      from identical fields and noise: per-chain |dH_mixed - dH_fp64|
      against 4 tau |xi|^2 sqrt(m^2+4)/m."""
    Ls, gs, lats = c4_batch(H, "gauss")
    spec = c4_run(H, N, lats, gs, Ls, steps, 10)
    del lats
    Ls, gs, lats = c4_batch(H, "thermal")
    stable = c4_run(H, N, lats, gs, Ls, steps, nsteps)
    cmp = mixed_vs_fp64(H, N, lats, nsteps)
    out = {"value": stable["value"], "unit": "traj/s", "chains": 256,
           "sites": int(np.sum(Ls.astype(np.int64) ** 2)),
           "sizes": "L in {32,48,64,96,128}, seed 4004",
           "as_specified_A_gauss_10x0.1": spec, "stable_thermalized": stable,
           "mixed_vs_fp64_dH": {k: cmp[k] for k in ("chains", "max_abs_dH_diff",
                                                     "max_diff_over_bound", "bound")}}
    return out


def c1_leg(H, N, steps):
    """16x16 as the reference runs it (hmc.c:671-746): m = 0.1, g = 0.5,
    A ~ N(0,1) (seed 1001), FP64, 10 x dt = 0.1; and the stable variant
    with 40 steps."""
    out = {}
    for nsteps, tag in ((10, "nsteps10"), (40, "nsteps40")):
        lat = H.StaggeredLattice(16, 16, m=0.1, g=0.5)
        lat.random_links(1001)
        dt, res, *_ = run_batch(H, N, [lat], [0.5], N.PREC_F64, steps, 2, nsteps=nsteps)
        out[tag] = {"value": steps / dt, "unit": "traj/s", "trajectories": steps,
                    "acceptance": float(np.mean([r[1] for r in res])),
                    "dH_median": float(np.median([r[0] for r in res])),
                    "max_true_rel_resid": max(r[4] for r in res)}
    out["reference_cpu_1core_O3_nsteps10"] = 23.4
    return out


def c2_leg(H, N):
    """256^2, m = 0.05, g = 0.5, A ~ N(0,1) (seed 1002), 8 Gaussian
    right-hand sides (seed 2002): the 8 solves of (M^dag M) x = b as one
    batched solve, FP64 and mixed (FP32 inner CG, FP64 true-residual
    restarts); per RHS the true residual and the iterations."""
    import ctypes as C
    size, nrhs = 256, 8
    V = size * size
    rng = np.random.default_rng(2002)
    bs = [rng.normal(size=V) + 1j * rng.normal(size=V) for _ in range(nrhs)]
    lat = H.StaggeredLattice(size, size, m=0.05, g=0.5)
    lat.random_links(1002)
    out = {"lattice": [size, size], "rhs": nrhs, "m": 0.05}
    L = N.lib()
    ctx = H.Context.get().ptr
    bd = [H.DeviceArray(16 * V) for _ in range(nrhs)]
    xd = [H.DeviceArray(16 * V) for _ in range(nrhs)]
    for b, d in zip(bs, bd):
        lat.to_eo(b, d)
    geoms = (N.Geom * nrhs)(*([lat.geom] * nrhs))
    xp = (C.c_void_p * nrhs)(*[d.p for d in xd])
    bp = (C.c_void_p * nrhs)(*[d.p for d in bd])
    kp = (C.c_void_p * nrhs)(*([lat.K.p] * nrhs))
    k32 = (C.c_void_p * nrhs)(*([lat.K32.p] * nrhs))
    for prec, tag in ((N.PREC_F64, "fp64"), (N.PREC_MIXED, "mixed")):
        p = H.cg_params(tol_rel=1e-10, precision=prec, max_iter=200000)
        infos = (N.CgInfo * nrhs)()
        for timed in (0, 1):  # warm-up, then timed
            if timed:
                L.th2d_timer_start(ctx)
            N.check(L.th2d_cg_mdagm_batched(ctx, nrhs, geoms, xp, bp, kp, k32, C.byref(p), infos),
                    "cg batched")
        ms = C.c_float()
        L.th2d_timer_stop(ctx, C.byref(ms))
        out[tag] = {"ms_8_rhs": ms.value,
                    "true_rel_resid": [float(i.true_rel_resid) for i in infos],
                    "iterations": [int(i.iterations) for i in infos],
                    "status": [int(i.status) for i in infos]}
    out["mixed_over_fp64_speed"] = out["fp64"]["ms_8_rhs"] / out["mixed"]["ms_8_rhs"]
    return out


def c3_leg(H, N):
    """Force at 128^2, m = 0.1, g = 0.8 (momentum_step, hmc.c:504-665) on
    all 32,768 links, psi = M^dag xi: the force from the FP64 solve (1e-14)
    against central differences (h = 1e-6) of S = S_g + psi^dag (M^dag M)^-1
    psi on 64 random links (seed 3003), and the force from the mixed solve
    (1e-10) against the FP64 one.  S_pf is evaluated in the stationary form
    2 Re psi^dag x - |M x|^2 (error second order in the solver's residual).
    Everything runs on the device."""
    nt = nx = 128
    V = nt * nx
    A = np.random.default_rng(1003).normal(size=(V, 2))
    r = np.random.default_rng(2003)
    xi = r.normal(size=V) + 1j * r.normal(size=V)
    lat = H.StaggeredLattice(nt, nx, m=0.1, g=0.8)
    lat.set_links(A)
    psi = lat.fm_conjugate_mul(xi)
    chi64, i64 = lat.fmdm_invert_cg(psi, H.cg_params(tol_rel=1e-14))
    chi32, i32 = lat.fmdm_invert_cg(psi, H.cg_params(tol_rel=1e-10, precision=N.PREC_MIXED))
    F64, F32 = lat.force(chi64), lat.force(chi32)
    pick = np.random.default_rng(3003)
    sites = pick.integers(0, V, size=64)
    dirs = pick.integers(0, 2, size=64)
    p13 = H.cg_params(tol_rel=1e-13)
    h = 1e-6

    def action(Ap):
        lat.set_links(Ap)
        x, _ = lat.fmdm_invert_cg(psi, p13)
        Mx = lat.fm_mul(x)
        return 2 * np.vdot(psi, x).real - np.vdot(Mx, Mx).real + lat.calc_gauge_action()

    fd_err = []
    for s_, d in zip(sites, dirs):
        Sp = []
        for sg in (1, -1):
            Ap = A.copy()
            Ap[s_, d] += sg * h
            Sp.append(action(Ap))
        fd = (Sp[0] - Sp[1]) / (2 * h)
        fd_err.append(abs(F64[s_, d] - fd) / max(1.0, abs(fd)))
    lat.set_links(A)
    d32 = np.abs(F32 - F64) / np.max(np.abs(F64))
    return {"links": 2 * V,
            "fp64_vs_fd_64_links": {"max_rel": float(np.max(fd_err)),
                                    "rms_rel": float(np.sqrt(np.mean(np.square(fd_err)))),
                                    "h": h, "note": "|F - FD| / max(1, |FD|)"},
            "mixed_vs_fp64": {"max_rel": float(np.max(d32)),
                              "rms_rel": float(np.sqrt(np.mean(np.square(d32)))),
                              "bound": 2 * (0.01 + 4) / 0.01 * 1e-10},
            "cg_iterations_fp64": int(i64.iterations), "cg_iterations_mixed": int(i32.iterations)}


def _bag_field(rng, L, rho):
    """Random legal links and monomer pairs up to occupied fraction rho."""
    V = L * L
    f = np.zeros(V, np.int32)
    target = int(rho * V)
    tries = 0
    while np.count_nonzero(f) < target and tries < 20 * V:
        tries += 1
        s = int(rng.integers(V))
        t, x = divmod(s, L)
        d = int(rng.integers(2))
        s2 = ((t + 1) % L) * L + x if d == 0 else t * L + (x + 1) % L
        if f[s] or f[s2]:
            continue
        if rng.random() < 0.8:
            f[s], f[s2] = 2 + d, 4 + d
        else:
            f[s] = f[s2] = 1
    return f


def c5_leg(N, proposals, device_steps=1000):
    """64 masked bag matrices, L uniform on even {16..48} (seed 5005),
    occupancy rho ~ U[0.1, 0.5], m = 0.1: batched LU log-det in FP64 and
    FP32 on MFMA, then rank-1/rank-2 ratio+update proposals."""
    from thirring2d_amd import bag as Bm
    from thirring2d_amd.hmc import Context
    import ctypes as C
    Ls, lats, fields = c5_batch()
    ctx = Context.get()
    lib = N.lib()
    out = {"matrices": 64, "orders": [int(L) ** 2 for L in Ls][:4] + ["..."],
           "max_order": int(max(Ls)) ** 2}
    flops = float(sum((2.0 / 3.0) * float(L) ** 6 for L in Ls))
    la = {}
    for prec, peak, tag in ((N.PREC_F64, 78.6, "f64"), (N.PREC_F32, 157.3, "f32")):
        db = Bm.build_masked_batch(lats, fields, prec)
        db.lu()  # warm-up
        db = Bm.build_masked_batch(lats, fields, prec)
        ctx.sync()
        lib.th2d_timer_start(ctx.ptr)
        db.lu()
        ms = C.c_float()
        lib.th2d_timer_stop(ctx.ptr, C.byref(ms))
        la[tag], sg = db.logdet()
        tf = flops / (ms.value * 1e-3) / 1e12
        out["lu_" + tag] = {"ms": ms.value, "TFLOP/s": tf, "peak": peak, "mfma_frac": tf / peak}
    # the FP32 factors with the FP64 correction of their log-determinant
    db64 = Bm.build_masked_batch(lats, fields, N.PREC_F64)
    db32 = Bm.build_masked_batch(lats, fields, N.PREC_F32)
    db32.lu()
    db32.logdet_refined(db64)  # warm-up
    ctx.sync()
    lib.th2d_timer_start(ctx.ptr)
    lar, _ = db32.logdet_refined(db64)
    ms = C.c_float()
    lib.th2d_timer_stop(ctx.ptr, C.byref(ms))
    out["lu_f32_refined"] = {"correction_ms": ms.value,
                             "total_ms": out["lu_f32"]["ms"] + ms.value,
                             "logdet_max_abs_err_vs_f64": float(np.max(np.abs(lar - la["f64"]))),
                             "logdet_max_rel_err_vs_f64": float(np.max(np.abs(lar - la["f64"]) /
                                                                       np.abs(la["f64"])))}
    out["f32_logdet_max_abs_err_vs_f64"] = float(np.max(np.abs(la["f32"] - la["f64"])))
    out["f32_logdet_max_rel_err_vs_f64"] = float(np.max(np.abs(la["f32"] - la["f64"]) /
                                                        np.abs(la["f64"])))
    # rank-1/rank-2 updates on the device-resident inverse of a subset
    prng = np.random.default_rng(5006)
    n_up = n_acc = 0
    t_up = 0.0
    for k in range(8):
        L = int(Ls[k])
        st = Bm.BagState(lats[k])
        st.set_field(fields[k])
        f = fields[k].copy()
        t0 = time.perf_counter()
        for _ in range(proposals):
            s = int(prng.integers(L * L))
            t, x = divmod(s, L)
            d = int(prng.integers(2))
            s2 = ((t + 1) % L) * L + x if d == 0 else t * L + (x + 1) % L
            if f[s] == 0 and f[s2] == 0:
                occ, vals, fr = [s, s2], [2 + d, 4 + d], []
            elif f[s] == 2 + d and f[s2] == 4 + d:
                occ, vals, fr = [], [], [s, s2]
            elif f[s] == 1 and f[s2] == 0:
                occ, vals, fr = [s2], [1], [s]
            else:
                continue
            r = st.ratio(occ, fr)
            n_up += 1
            if abs(r) > 1e-8 and prng.random() < min(1.0, abs(r)):
                st.accept(occ, vals, fr)
                f[fr] = 0
                f[occ] = vals
                n_acc += 1
        t_up += time.perf_counter() - t0
        del st
    out["woodbury_host_driven"] = {"proposals": n_up, "accepted": n_acc, "seconds": t_up,
                                   "proposals_per_s": n_up / t_up if t_up > 0 else 0.0,
                                   "matrices": 8}
    out["device_driver"] = bag_driver_leg(N, lats, fields, device_steps)
    return out


def c5_batch():
    """The C5 configurations: 64 bag lattices, L uniform on even {16..48}
    (seed 5005), m = 0.1, occupancy rho ~ U[0.1, 0.5]."""
    from thirring2d_amd import bag as Bm
    rng = np.random.default_rng(5005)
    Ls = rng.choice(np.arange(16, 49, 2), size=64)
    lats = [Bm.BagLattice(int(L), int(L), m=0.1) for L in Ls]
    fields = [_bag_field(rng, int(L), rng.uniform(0.1, 0.5)) for L in Ls]
    return Ls, lats, fields


def only_legs(names):
    """The legs named in `names`, alone, as one JSON line."""
    from thirring2d_amd import _native as N
    from thirring2d_amd import hmc as H
    out = {}
    for name in names:
        if name == "c1":
            out["c1_16x16_fp64"] = c1_leg(H, N, 100)
        elif name == "c4":
            N.lib().th2d_set_cg_autotune(H.Context.get().ptr, 1)
            out["c4_256_chains_uneven"] = c4_leg(H, N, 50)
        elif name == "bag-driver":
            _, lats, fields = c5_batch()
            out["bag_driver"] = bag_driver_leg(N, lats, fields, 1000)
        else:
            raise SystemExit(f"bench.py: unknown leg {name!r}")
    print(json.dumps(out), flush=True)


def bag_driver_leg(N, lats, fields, steps):
    """All 64 matrices through `steps` proposals each on the device
    (th2d_bag_batch_sweep: one proposal per matrix per step, no host
    synchronisation between steps), then the drift of log|det D_FF| against
    a refresh from scratch."""
    from thirring2d_amd import bag as Bm
    import ctypes as C
    batch = Bm.BagBatch(lats, fields)
    batch.sweep(20, seed=5, counter=0)  # warm-up
    _, _, _, acc0 = batch.state()
    lib = N.lib()
    lib.th2d_timer_start(batch.ctx.ptr)
    moves = batch.sweep(steps, seed=5, counter=1000, trace=True)
    ms = C.c_float()
    lib.th2d_timer_stop(batch.ctx.ptr, C.byref(ms))
    la, sg, _, acc = batch.state()
    batch.refresh()
    la2, sg2, _, _ = batch.state()
    n = len(lats) * steps
    # an accepted change rewrites B once per pass (freeing, then occupying):
    # V^2 doubles read and written
    vv = [float(l.V) ** 2 for l in lats]
    nbytes = sum(16.0 * vv[b] * ((mv.nr > 0) + (mv.no > 0))
                 for step in moves for b, mv in enumerate(step) if mv.accepted)
    return {"matrices": len(lats), "proposals": n, "accepted": int(np.sum(acc - acc0)),
            "ms": ms.value, "proposals_per_s": n / (ms.value * 1e-3),
            "B_update_bytes": nbytes, "B_update_GBs_whole_sweep": nbytes / (ms.value * 1e-3) / 1e9,
            "logdet_drift_max_abs": float(np.max(np.abs(la - la2))),
            "sign_changes_on_refresh": int(np.sum(sg != sg2))}


def dslash_bw(H, N, size, nrhs, precision, reps=50):
    """Algorithmic GB/s of th2d_dslash: per site 2 links + nrhs x (in + out)."""
    import ctypes as C
    lat = H.StaggeredLattice(size, size, m=0.05, g=0.5)
    lat.random_links(1002)
    V = size * size
    cb = 16 if precision == N.PREC_F64 else 8
    src = H.DeviceArray(cb * V * nrhs)
    dst = H.DeviceArray(cb * V * nrhs)
    L = N.lib()
    L.th2d_memset(lat.ctx.ptr, src.p, 0, cb * V * nrhs)
    K = lat.K.p if precision == N.PREC_F64 else lat.K32.p
    for _ in range(3):
        L.th2d_dslash(lat.ctx.ptr, lat._g(), dst.p, src.p, K, 0, precision, nrhs)
    L.th2d_timer_start(lat.ctx.ptr)
    for _ in range(reps):
        L.th2d_dslash(lat.ctx.ptr, lat._g(), dst.p, src.p, K, 0, precision, nrhs)
    ms = C.c_float()
    L.th2d_timer_stop(lat.ctx.ptr, C.byref(ms))
    per = ms.value / reps
    bytes_ = V * (2 * cb + nrhs * 2 * cb)
    return per, bytes_ / (per * 1e-3) / 1e9


def dslash_batched_bw(H, N, size, chains, reps=50):
    """Algorithmic GB/s of th2d_dslash_batched (FP64): `chains` lattices of
    size^2, each with its own links, in one launch; per site one vector in,
    one out and two links."""
    import ctypes as C
    V = size * size
    lats, ins, outs = [], [], []
    L = N.lib()
    for c in range(chains):
        lat = H.StaggeredLattice(size, size, m=0.05, g=0.5)
        lat.random_links(1002 + c)
        src, dst = H.DeviceArray(16 * V), H.DeviceArray(16 * V)
        L.th2d_memset(lat.ctx.ptr, src.p, 0, 16 * V)
        lats.append(lat)
        ins.append(src)
        outs.append(dst)
    geoms = (N.Geom * chains)(*[l.geom for l in lats])
    P = C.c_void_p * chains
    po, pi, pk = P(*[d.p for d in outs]), P(*[d.p for d in ins]), P(*[l.K.p for l in lats])
    ctx = lats[0].ctx.ptr
    for _ in range(3):
        N.check(L.th2d_dslash_batched(ctx, chains, geoms, po, pi, pk, 0), "dslash_batched")
    L.th2d_timer_start(ctx)
    for _ in range(reps):
        L.th2d_dslash_batched(ctx, chains, geoms, po, pi, pk, 0)
    ms = C.c_float()
    L.th2d_timer_stop(ctx, C.byref(ms))
    per = ms.value / reps
    return per, chains * V * 64 / (per * 1e-3) / 1e9


def dslash_sweep():
    """Algorithmic bandwidth of th2d_dslash from 256^2 to 4096^2 (FP64 and
    FP32, one RHS; 8 RHS at 256^2): the sizes above ~2048^2 (FP64 vectors
    and links of 200 MB and more) leave the 256 MB last-level cache."""
    from thirring2d_amd import _native as N
    from thirring2d_amd import hmc as H
    out = {}
    for size in (256, 512, 1024, 2048, 4096):
        for prec, tag in ((N.PREC_F64, "f64"), (N.PREC_F32, "f32")):
            for nrhs in ((1, 8) if size == 256 else (1,)):
                ms, gbs = dslash_bw(H, N, size, nrhs, prec)
                cb = 16 if prec == N.PREC_F64 else 8
                out[f"{size}_{tag}_{nrhs}rhs"] = {
                    "ms": ms, "GB/s": gbs, "frac": gbs / HBM_PEAK_GBS,
                    "bytes": size * size * (2 * cb + nrhs * 2 * cb)}
    print(json.dumps({"dslash_sweep": out, "unit": "GB/s algorithmic", "peak": HBM_PEAK_GBS}),
          flush=True)


# The reference's own fm_mul at 256^2 (-O3, one core) took 5.58 ms per
# apply in this build's development container; the oracle's restatement
# took ORACLE_APPLY_MS_DEV on the same host (measured with
# `python bench.py --oracle-apply`): their ratio carries the reference's
# speed over to the host this benchmark runs on.
REF_APPLY_MS_DEV = 5.58
ORACLE_APPLY_MS_DEV = 7.35


def oracle_apply_ms(size=256, reps=20):
    """Part of the CPU baseline: the oracle's fm_mul at size^2 on one host
    core, ms per apply (the calibration ORACLE_APPLY_MS_DEV above)."""
    import oracle as O
    V = size * size
    L = O.lattice(size, size, m=0.05, g=0.5)
    rng = np.random.default_rng(1002)
    A = rng.normal(size=(V, 2))
    v = rng.normal(size=V) + 1j * rng.normal(size=V)
    O.fm_mul(L, A, v)
    t0 = time.perf_counter()
    for _ in range(reps):
        O.fm_mul(L, A, v)
    return 1e3 * (time.perf_counter() - t0) / reps


def cpu_baseline(iters_per_traj, size, n_it=1200):
    """The oracle's CG iteration at size^2 on one host core, scaled to a
    trajectory by the CG iterations per trajectory of the single-chain FP64
    run on the GPU."""
    import oracle as O
    V = size * size
    L = O.lattice(size, size, m=0.05, g=0.5)
    rng = np.random.default_rng(1002)
    A = rng.normal(size=(V, 2))
    b = rng.normal(size=V) + 1j * rng.normal(size=V)
    t0 = time.perf_counter()
    O.cg_mdm(L, A, b, tol_abs_rr=0.0, tol_rel=1e-300, max_iter=n_it + 1)
    dt = time.perf_counter() - t0
    per_it = dt / n_it
    out = {"value": 1.0 / (per_it * iters_per_traj), "unit": "traj/s", "cores": 1,
           "kind": "port",
           "sample": f"{n_it} CG iterations of the C oracle on {size}^2 "
                     f"({dt:.1f} s, 1 thread), scaled by {iters_per_traj:.0f} CG iterations "
                     "per trajectory (single-chain FP64 run on the GPU)"}
    if ORACLE_APPLY_MS_DEV:
        out["reference_over_port_apply"] = REF_APPLY_MS_DEV / ORACLE_APPLY_MS_DEV
    return out


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs on one GPU; multi-GPU runs are not part of it "
              "(independent chains and bag matrices shard across GPUs without communication)",
              file=sys.stderr)
        sys.exit(2)
    if args.oracle_apply:
        print(json.dumps({"oracle_fm_mul_ms_256": oracle_apply_ms()}), flush=True)
        return
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.dslash_sweep:
        dslash_sweep()
        return
    if args.only_legs:
        only_legs([x for x in args.only_legs.split(",") if x])
        return
    if args.dense_only:
        from thirring2d_amd import _native as N
        print(json.dumps({"c5_bag_dense": c5_leg(N, args.c5_proposals)}), flush=True)
        return
    import ctypes as C
    from thirring2d_amd import _native as N
    from thirring2d_amd import hmc as H

    size, nch, nsteps = args.size, args.chains, args.nsteps
    start = thermalized(H, size, nch)
    # headline: batched mixed precision (its warm-up trajectories move the
    # chains from the pure-gauge start)
    lats = copy_fields(H, start)
    dt_mx, res_mx, kms, kn, kbytes = run_hmc(H, N, size, N.PREC_MIXED, args.steps, args.warmup,
                                             nsteps=nsteps, lats=lats)
    on = C.c_int(0)
    N.check(N.lib().th2d_cg_last_resident(H.Context.get().ptr, C.byref(on)), "cg_last_resident")
    resident = bool(on.value)
    tps = nch * args.steps / dt_mx
    acc = float(np.mean([r[1] for r in res_mx]))
    dH = np.array([r[0] for r in res_mx])
    # the solver kernel: 88 flop per site and iteration (two operator
    # applications of 34 flop and the fused vector updates), FP32 peak
    site_iters = kbytes / 120.0
    tflops = site_iters * 88.0 / (kms * 1e-3) / 1e12 if kms > 0 else 0.0
    avg_ms = kms / max(kn, 1)
    traffic = TRAFFIC_BYTES.get((size, nch, resident))
    roofline = {"bound": "mfma", "achieved": tflops, "peak": FP32_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": tflops / FP32_PEAK_TFLOPS, "traffic": traffic,
                "kernel": "k_cg_resident<float>" if resident else "k_cg_step<float>",
                "launches": kn, "avg_launch_ms": avg_ms,
                "flop_per_site_iteration": 88,
                "hbm_GBs_measured": (traffic / (avg_ms * 1e-3) / 1e9) if traffic else None,
                "note": ("FP32 vector-ALU arithmetic (the FP32 peak, 157.3 TFLOP/s, is the same "
                         "for the vector ALUs and the matrix cores); the vectors and the "
                         "links of each wave's rows stay in registers (the halo in LDS), so "
                         "HBM sees `traffic` bytes per launch; the rest of each iteration "
                         "is the synchronisation of a chain's regions" if resident else
                         "streaming CG: every iteration reads and writes the vectors")}
    if args.only_headline:
        print(json.dumps({"metric": "batched mixed HMC (profiling run)", "value": tps,
                          "unit": "traj/s", "chains": nch, "nsteps": nsteps, "acceptance": acc,
                          "kernel_avg_ms": avg_ms, "launches": kn, "resident": resident,
                          "roofline": roofline}), flush=True)
        return
    small = max(2, args.steps // 4)
    after = lats  # the chains after the timed trajectories
    dt_b64, res_b64, *_ = run_hmc(H, N, size, N.PREC_F64, small, 0, nsteps=nsteps,
                                  lats=copy_fields(H, after))
    dt_1mx, res_1mx, *_ = run_hmc(H, N, size, N.PREC_MIXED, small, 0, nsteps=nsteps,
                                  lats=copy_fields(H, after[:1]))
    dt_164, res_164, *_ = run_hmc(H, N, size, N.PREC_F64, small, 0, nsteps=nsteps,
                                  lats=copy_fields(H, after[:1]))
    tps_b64 = nch * small / dt_b64
    tps_1mx = small / dt_1mx
    tps_164 = small / dt_164
    its_164 = float(np.mean([r[2] for r in res_164]))

    out = {
        "metric": "HMC traj/s at 256^2 (mixed precision, batched chains, CG true residual <= 1e-10)",
        "value": tps,
        "unit": "traj/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * dt_mx / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic: A = 0 + 100 pure-gauge sweeps (device Philox, seeds 1002+chain), "
                "device Philox noise",
        "config": {"workload": f"HMC {size}x{size}, m=0.05, g=0.5, {nsteps} x dt=1/{nsteps}, "
                               f"CG 1e-10, {nch} chains per batched step",
                   "lattice": [size, size], "chains": nch, "nsteps": nsteps,
                   "dt": 1.0 / nsteps, "tol_rel": 1e-10},
        "acceptance": acc,
        "exp_minus_dH_mean": float(np.mean(np.exp(-dH))),
        "dH_median": float(np.median(dH)),
        "batched_fp64": {"value": tps_b64, "unit": "traj/s", "chains": nch,
                         "acceptance": float(np.mean([r[1] for r in res_b64]))},
        "single_chain_mixed": {"value": tps_1mx, "unit": "traj/s"},
        "single_chain_fp64": {"value": tps_164, "unit": "traj/s", "cg_iters_per_traj": its_164},
        "batched_mixed_over_single_fp64": tps / tps_164,
        "batched_mixed_over_batched_fp64": tps / tps_b64,
        "single_mixed_over_single_fp64": tps_1mx / tps_164,
        "cg_iters_per_traj": float(np.mean([r[2] for r in res_mx])),
        "max_true_rel_resid": max(r[4] for r in res_mx + res_b64 + res_1mx + res_164),
        "roofline": roofline,
    }
    if acc == 0.0:
        out["flag"] = "acceptance 0: the timed trajectories were all rejected"
    out["mixed_vs_fp64_dH"] = mixed_vs_fp64(H, N, after, nsteps)
    if args.ref_md_steps:
        dt_r, res_r, *_ = run_hmc(H, N, size, N.PREC_MIXED, args.ref_md_steps, 1, nsteps=10,
                                  lats=copy_fields(H, start))
        out["reference_md_10x0.1"] = {
            "value": nch * args.ref_md_steps / dt_r, "unit": "traj/s", "chains": nch,
            "acceptance": float(np.mean([r[1] for r in res_r])),
            "dH_median": float(np.median([r[0] for r in res_r])),
            "flag": "dt = 0.1 (hmc.c:708-709) is beyond the leapfrog stability limit at "
                    "m = 0.05 on 256^2: every trajectory is rejected"}
    if not args.no_dslash:
        d = {}
        for (sz, nrhs, prec, tag) in [(256, 1, N.PREC_F64, "256_f64"), (256, 8, N.PREC_F64,
                                      "256x8rhs_f64"), (256, 8, N.PREC_F32, "256x8rhs_f32"),
                                      (4096, 1, N.PREC_F64, "4096_f64"),
                                      (4096, 1, N.PREC_F32, "4096_f32")]:
            ms, gbs = dslash_bw(H, N, sz, nrhs, prec)
            d[tag] = {"ms": ms, "GB/s": gbs, "frac": gbs / HBM_PEAK_GBS}
        ms, gbs = dslash_batched_bw(H, N, 256, 16)
        d["256x16chains_f64"] = {"ms": ms, "GB/s": gbs, "frac": gbs / HBM_PEAK_GBS,
                                 "note": "th2d_dslash_batched: 16 lattices of 256^2, "
                                         "each with its own links, in one launch"}
        out["dslash"] = d
    if args.c1_steps:
        out["c1_16x16_fp64"] = c1_leg(H, N, args.c1_steps)
    if not args.no_c2:
        out["c2_256_8rhs"] = c2_leg(H, N)
    if not args.no_c3:
        out["c3_force_128"] = c3_leg(H, N)
    if args.c4_steps:
        N.lib().th2d_set_cg_autotune(H.Context.get().ptr, 1)
        out["c4_256_chains_uneven"] = c4_leg(H, N, args.c4_steps)
    if args.c5_proposals:
        out["c5_bag_dense"] = c5_leg(N, args.c5_proposals)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(its_164, size)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
