"""Benchmark: HMC trajectories per second at 256x256 (mixed precision and
FP64, CG true relative residual <= 1e-10), with the Dslash bandwidth.

Workload (one GPU): 256x256 lattice, antiperiodic t and x, m = 0.05,
g = 0.5, link angles A ~ N(0,1) (device generator, seed 1002), 10 leapfrog
steps of dt = 0.1 per trajectory, pseudofermion / momentum / stochastic
noise from the device generator, every CG solve to a true relative
residual of 1e-10.  One step = one trajectory.

Prints one JSON line.  `value` is the aggregate trajectories per second of
a batch of independent 256^2 chains advanced in lockstep (every inversion
of all chains is one batched solve) with mixed precision (FP32 CG, FP64
true-residual restarts); the batched FP64 path and single chains in FP64
and mixed precision are timed beside it.  `roofline` is for the dominant
kernel, the CG step kernel, timed with HIP events on its stream.
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
# measured HBM bytes per k_cg_step launch (profiles/k_cg_step_hmc256_32chains_mixed.md)
TRAFFIC_BYTES = {(256, 32): 199.6e6}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--chains", type=int, default=32)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-dslash", action="store_true")
    ap.add_argument("--only-headline", action="store_true",
                    help="time the batched mixed-precision leg only (for profiling)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="run a plain device copy of MIB MiB five times and exit (scale of "
                         "the FETCH_SIZE / WRITE_SIZE counters)")
    return ap.parse_args()


def copy_calibration(mib):
    import torch
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": n * 4, "launches": 5}), flush=True)


def run_hmc(H, N, size, precision, steps, warmup, chains=1, seed_links=1002):
    """Time `steps` trajectories of `chains` independent chains (one batched
    call per step).  Returns (seconds, per-trajectory records, kernel ms,
    kernel launches, kernel bytes)."""
    import ctypes as C
    lats = []
    for c in range(chains):
        lat = H.StaggeredLattice(size, size, m=0.05, g=0.5)
        lat.random_links(seed_links + c)
        lats.append(lat)
    batch = H.ChainBatch(lats)
    params = H.cg_params(tol_rel=1e-10, precision=precision, max_iter=200000)
    L = N.lib()
    ctr = 0
    for _ in range(warmup):
        batch.update_gauge(nsteps=10, traj_len=1.0, params=params, seed=4242,
                           counters=[ctr] * chains)
        ctr += 1
    L.th2d_kernel_timing(batch.ctx.ptr, 1)
    batch.ctx.sync()
    t0 = time.perf_counter()
    res = []
    for _ in range(steps):
        rr = batch.update_gauge(nsteps=10, traj_len=1.0, params=params, seed=4242,
                                counters=[ctr] * chains, u_accept=list(np.random.default_rng(ctr).random(chains)))
        ctr += 1
        res += [(r.dS, r.accepted, r.cg_iterations, r.cg_solves, r.max_true_rel_resid, r.status)
                for r in rr]
    batch.ctx.sync()
    dt = time.perf_counter() - t0
    ms, n, by = C.c_double(), C.c_long(), C.c_double()
    L.th2d_kernel_time(batch.ctx.ptr, C.byref(ms), C.byref(n), C.byref(by))
    L.th2d_kernel_timing(batch.ctx.ptr, 0)
    return dt, res, ms.value, n.value, by.value


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


def cpu_baseline(iters_per_traj, size):
    """The oracle's CG iteration at size^2 on one host core, scaled to a
    trajectory by the measured GPU iteration count per trajectory."""
    import oracle as O
    V = size * size
    L = O.lattice(size, size, m=0.05, g=0.5)
    rng = np.random.default_rng(1002)
    A = rng.normal(size=(V, 2))
    b = rng.normal(size=V) + 1j * rng.normal(size=V)
    n_it = 300
    t0 = time.perf_counter()
    O.cg_mdm(L, A, b, tol_abs_rr=0.0, tol_rel=1e-300, max_iter=n_it + 1)
    dt = time.perf_counter() - t0
    per_it = dt / n_it
    return {"value": 1.0 / (per_it * iters_per_traj), "unit": "traj/s", "cores": 1,
            "kind": "port",
            "sample": f"{n_it} CG iterations of the C oracle on {size}^2 "
                      f"({dt:.1f} s, 1 thread), scaled by {iters_per_traj:.0f} CG iterations "
                      "per trajectory measured on the GPU run"}


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs on one GPU; multi-GPU runs are not part of it "
              "(independent chains shard trivially, see DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    from thirring2d_amd import _native as N
    from thirring2d_amd import hmc as H

    size, nch = args.size, args.chains
    # headline: batched mixed precision
    dt_mx, res_mx, kms, kn, kbytes = run_hmc(H, N, size, N.PREC_MIXED, args.steps, args.warmup,
                                             chains=nch)
    if args.only_headline:
        print(json.dumps({"metric": "batched mixed HMC (profiling run)", "value": nch * args.steps / dt_mx,
                          "unit": "traj/s", "chains": nch, "kernel_avg_ms": kms / max(kn, 1),
                          "launches": kn}), flush=True)
        return
    small = max(2, args.steps // 4)
    dt_b64, res_b64, *_ = run_hmc(H, N, size, N.PREC_F64, small, 1, chains=nch)
    dt_1mx, res_1mx, *_ = run_hmc(H, N, size, N.PREC_MIXED, small, 1, chains=1)
    dt_164, res_164, *_ = run_hmc(H, N, size, N.PREC_F64, small, 1, chains=1)
    tps = nch * args.steps / dt_mx
    tps_b64 = nch * small / dt_b64
    tps_1mx = small / dt_1mx
    tps_164 = small / dt_164
    its_164 = float(np.mean([r[2] for r in res_164]))
    achieved = kbytes / (kms * 1e-3) / 1e9 if kms > 0 else 0.0

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
        "data": "synthetic: A~N(0,1) seeds 1002+chain, device Philox noise",
        "config": {"workload": f"HMC {size}x{size}, m=0.05, g=0.5, 10 x dt=0.1, CG 1e-10, "
                               f"{nch} chains per batched step",
                   "lattice": [size, size], "chains": nch, "nsteps": 10, "dt": 0.1,
                   "tol_rel": 1e-10},
        "batched_fp64": {"value": tps_b64, "unit": "traj/s", "chains": nch},
        "single_chain_mixed": {"value": tps_1mx, "unit": "traj/s"},
        "single_chain_fp64": {"value": tps_164, "unit": "traj/s", "cg_iters_per_traj": its_164},
        "batched_mixed_over_single_fp64": tps / tps_164,
        "cg_iters_per_traj": float(np.mean([r[2] for r in res_mx])),
        "max_true_rel_resid": max(r[4] for r in res_mx + res_b64 + res_1mx + res_164),
        "acceptance": float(np.mean([r[1] for r in res_mx])),
        "dH_median": float(np.median([r[0] for r in res_mx])),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     # HBM bytes per launch of this kernel at this workload, from the
                     # FETCH_SIZE / WRITE_SIZE passes in profiles/ (copy-calibrated)
                     "traffic": TRAFFIC_BYTES.get((size, nch)),
                     "kernel": "k_cg_step<float> (CG iteration, 120 B/site algorithmic)",
                     "launches": kn, "avg_launch_ms": kms / max(kn, 1)},
    }
    if not args.no_dslash:
        d = {}
        for (sz, nrhs, prec, tag) in [(256, 1, N.PREC_F64, "256_f64"), (256, 8, N.PREC_F64,
                                      "256x8rhs_f64"), (256, 8, N.PREC_F32, "256x8rhs_f32"),
                                      (4096, 1, N.PREC_F64, "4096_f64"),
                                      (4096, 1, N.PREC_F32, "4096_f32")]:
            ms, gbs = dslash_bw(H, N, sz, nrhs, prec)
            d[tag] = {"ms": ms, "GB/s": gbs, "frac": gbs / HBM_PEAK_GBS}
        out["dslash"] = d
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(its_164, size)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
