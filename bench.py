"""Benchmark: DQMC sweeps/s at N=256, L=80, FP64 on one MI355X.

Workload (the reference's headline configuration): 16x16 periodic square
Hubbard model (N=256), t=1, t'=-0.25, U=6, mu=-1, beta=8, dt=0.1 (L=80),
n_matmul=10, delay rank 32 requested (the N <= 256 Metropolis kernel keeps 16
register-resident delay slots; the rank changes only the rounding of G, not
the Markov chain), equal-time measurements every 8 slices,
unequal-time G(t,0)/G(0,t) + susceptibilities every 2 sweeps, real FP64.
A "step" is one full sweep (L slices of Metropolis updates + wraps/recalcs +
measurements) of every chain in the batch.  `value` is the aggregate
sweeps/s of the batched engine (chains x steps / seconds); the plain
single-chain form is timed beside it.  Inputs are generated on the host by
the package's generator (seeded, synthetic), uploaded before timing.

Prints one JSON line: metric/value/unit, timing, config, the roofline of the
dominant kernel (the N x N x N batched GEMM, timed with HIP events on the
engine stream) and the CPU baseline (the oracle on the host cores).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "dqmc-dev_amd"))

PEAK_FP64_TFLOPS = 78.6  # MI355X dense FP64 (matrix = vector)
# HBM bytes per 256 x 256 x 256 matrix product of the dominant GEMM, from the
# FETCH_SIZE / WRITE_SIZE counter passes scaled by a 512 MiB copy
# (profiles/k_gemm_N256_L80_256chains.md): 1.0012 x the 3 x 512 KiB it must move.
GEMM_HBM_BYTES_PER_MATRIX_N256 = 1574796


def workload(nchains, seed=7):
    from dqmc_amd import gen
    return gen.create_batch(Nfiles=nchains, seed=seed, Nx=16, Ny=16, tp=-0.25, U=6.0, mu=-1.0,
                            dt=0.1, L=80, n_matmul=10, n_delay=32, period_eqlt=8,
                            period_uneqlt=2, n_sweep_warm=0, n_sweep_meas=100000)


def flop_model(N, L, F, n_matmul, accepts_per_sweep, n_delay, q_eq, ue_fraction):
    """Algorithmic FP64 flops per sweep per chain (both spins): standard
    LAPACK/BLAS op counts applied to the reference's call sequence -- wraps
    4(L-F) N^3, recalc F[2(n-1) + Q(F)] N^3, updates, eq and ue measurements."""
    N3 = float(N) ** 3
    E = 1 + (F - 1) // 2
    s0 = 1 + (F - 1) % 2
    QF = 2 * (F % 2 == 0) + 4 / 3 + ((F - s0) // 2) * 19 / 3 + 14 / 3
    per_spin = 4 * (L - F) + F * (2 * (n_matmul - 1) + QF) + 4 * q_eq
    a = accepts_per_sweep / (N * L)
    per_spin += L * (2 * a + 2 * a * (n_delay - 1) / N)
    bsofi = {1: 2.0 / 3 * 4, 2: 30.0, 3: 77.0, 4: 138.7}.get(E, 138.7 * (E / 4) ** 3)
    ue = 2 * E + bsofi + 8 * (L - E) + 12 * L
    per_spin += ue * ue_fraction
    return 2 * per_spin * N3


def cpu_baseline(seconds_budget=20.0):
    """The oracle (CPU restatement, MKL sequential, 2 OpenMP threads -- one per
    spin -- as the reference's design point) on a bounded sample of the same
    workload: one chain, a few sweeps including one unequal-time sweep."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    sim = workload(1)[0]
    # sample: 8 sweeps (four with the unequal-time measurement)
    t0 = time.time()
    out = oracle.run(sim, max_sweeps=8)
    dt = time.time() - t0
    return {"value": out["sweeps_run"] / dt, "unit": "sweeps/s", "cores": 2, "kind": "port",
            "sample": "1 chain x %d sweeps (every 2nd with unequal-time), incl. initial G; %.1f s"
                      % (out["sweeps_run"], dt)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--chains", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-single", action="store_true")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MB",
                    help="first run one device-to-device copy of MB MiB (a known-size kernel that "
                         "calibrates the FETCH_SIZE / WRITE_SIZE counters of a profiler pass)")
    ap.add_argument("--gemm-tile", type=int, default=0,
                    help="GEMM C tile: 0 = autotune, 1 = 64x64, 2 = 128x128, 3 = 128x64")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (chains shard trivially across GPUs, "
              "but multi-GPU runs are not part of it); use --gpus 1", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    from dqmc_amd import engine

    if args.copy_probe > 0:
        import torch
        src = torch.ones(args.copy_probe << 17, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    tiles = {1: "64x64", 2: "128x128", 3: "128x64"}

    def timed(nch, steps, warmup, time_gemm):
        sims = workload(nch)
        eng = engine.Engine(sims, device=0, time_gemm=time_gemm, gemm_tile=args.gemm_tile)
        tile = eng.gemm_tile()
        eng.run(0)  # initial G outside the timed region
        eng.run(warmup)
        acc0 = eng.accepts()
        eng.gemm_stats()
        t0 = time.perf_counter()
        eng.run(steps)  # synchronises at the end
        dt = time.perf_counter() - t0
        acc = eng.accepts() - acc0
        gstats = eng.gemm_stats()
        eng.close()
        return dt, acc, gstats, tile

    nch = args.chains
    dt, acc, (g_ms, g_n, g_flops), tile = timed(nch, args.steps, args.warmup, True)
    value = nch * args.steps / dt
    N, L, F, n_matmul = 256, 80, 8, 10
    acc_per_sweep = acc / (nch * args.steps)
    fl = flop_model(N, L, F, n_matmul, acc_per_sweep, 32, L // 8, 0.5)
    model_tflops = fl * value / 1e12
    single = None
    if not args.no_single:
        dts, accs, _, _ = timed(1, max(2, args.steps // 2), 1, False)
        single = max(2, args.steps // 2) / dts
    achieved = g_flops / (g_ms * 1e-3) / 1e12 if g_ms > 0 else 0.0
    line = {
        "metric": "DQMC sweeps/s at N=256 L=80 FP64 (batched chains, aggregate)",
        "value": round(value, 4),
        "unit": "sweeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (package generator, seed 7 + jump per chain)",
        "config": {"workload": "16x16 square Hubbard N=256 L=80 t'=-0.25 U=6 mu=-1 beta=8 "
                               "n_matmul=10 delay=32 (16 register slots) eq every 8 slices, "
                               "uneq every 2 sweeps",
                   "chains": nch, "parallelism": "1 GPU, chains batched"},
        "single_chain_sweeps_per_s": None if single is None else round(single, 4),
        "batched_over_single": None if single is None else round(value / single, 3),
        "model_gflop_per_sweep": round(fl / 1e9, 2),
        "model_tflops": round(model_tflops, 3),
        "model_frac_fp64_peak": round(model_tflops / PEAK_FP64_TFLOPS, 4),
        "accepts_per_sweep": round(acc_per_sweep, 1),
        "roofline": {"bound": "mfma",
                     "kernel": "k_gemm<double,N,N> tile %s (N x N x N, batch 2*chains)" % tiles[tile],
                     "achieved": round(achieved, 3), "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / PEAK_FP64_TFLOPS, 4),
                     "traffic": GEMM_HBM_BYTES_PER_MATRIX_N256 * 2 * nch,
                     "avg_launch_ms": round(g_ms, 5), "launches": g_n},
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline()
    print(json.dumps(line))


if __name__ == "__main__":
    main()
