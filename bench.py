"""Benchmark: DQMC sweeps/s at N=256, L=80, FP64 on one MI355X.

Workload (the reference's headline configuration): 16x16 periodic square
Hubbard model (N=256), t=1, t'=-0.25, U=6, mu=-1, beta=8, dt=0.1 (L=80),
n_matmul=10, delay rank 32 (the N <= 256 Metropolis kernel keeps the rank-32
A factor in registers and B in LDS; the rank changes only the rounding of G,
not the Markov chain), equal-time measurements every 8 slices,
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
# HBM bytes per 256 x 256 x 256 matrix product of the dominant GEMM (k_gemm_w),
# from the FETCH_SIZE / WRITE_SIZE counter passes scaled by a 512 MiB copy
# (profiles/k_gemm_N256_L80_256chains.md): 1.0010 x the 3 x 512 KiB it must move.
GEMM_HBM_BYTES_PER_MATRIX_N256 = 1574365


# optional measurement groups of the reference (src/meas.c) on top of the
# default sets, for profiling the contraction kernels (--all-meas).  Left
# out: local JQ (the reference defines it for trans_sym = 0 files only) and
# gen_suscept (N^2 bins per slice: 168 MB of accumulators per chain at
# N = 256); the complex workload also leaves out nematic, which the
# reference defines without Peierls phases.
ALL_MEAS = dict(meas_bond_corr=1, meas_thermal=1, meas_2bond_corr=1, meas_energy_corr=1,
                meas_nematic_corr=1, meas_chiral=1)


def workload(nchains, seed=7, **extra):
    from dqmc_amd import gen
    return gen.create_batch(Nfiles=nchains, seed=seed, Nx=16, Ny=16, tp=-0.25, U=6.0, mu=-1.0,
                            dt=0.1, L=80, n_matmul=10, n_delay=32, period_eqlt=8,
                            period_uneqlt=2, n_sweep_warm=0, n_sweep_meas=100000, **extra)


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


def _model(flops_per_sweep, sweeps_per_s, prefix=""):
    """Model FLOP rate of a line: algorithmic flops per chain-sweep (flop_model)
    times chain-sweeps/s, and its fraction of the FP64 matrix peak."""
    tf = flops_per_sweep * sweeps_per_s / 1e12
    return {prefix + "model_gflop_per_sweep": round(flops_per_sweep / 1e9, 4),
            prefix + "model_tflops": round(tf, 4),
            prefix + "model_frac_fp64_peak": round(tf / PEAK_FP64_TFLOPS, 5)}


_CPU_CHAIN = r"""
import os, sys, time, json
sys.path[:0] = [os.path.join(%(root)r, "dqmc-dev_amd"), os.path.join(%(root)r, "oracle"), %(root)r]
import bench, oracle
sim = bench.workload(%(nproc)d)[%(k)d]
while time.time() < %(start)r:  # all chains start together
    time.sleep(0.01)
t0 = time.time()
oracle.run(sim, max_sweeps=0)  # the initial B/C/G build alone (setup)
t1 = time.time()
out = oracle.run(sim, max_sweeps=%(sweeps)d)
t2 = time.time()
# sweeps / (wall - setup), as the reference's own sweep rate is defined
print(json.dumps({"sweeps": out["sweeps_run"], "seconds": (t2 - t1) - (t1 - t0)}))
"""


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(sweeps=4, cpu_share=16):
    """The oracle (CPU restatement, MKL sequential, 2 OpenMP threads per chain
    -- one per spin, the reference's design point) as the reference's stack
    mode runs it: independent chains in concurrent processes, one per two of
    the host's cores (the box's CPU share, at most the cores visible), each a
    bounded sample of `sweeps` sweeps of the same workload (every second one
    with the unequal-time measurement, plus the initial G).  Started as child
    processes that begin their sweeps at a common time; the aggregate is the
    sum of the chains' rates, each measured while all of them run."""
    import json
    import subprocess
    try:
        ncpu = len(os.sched_getaffinity(0))
    except AttributeError:
        ncpu = os.cpu_count() or 2
    nproc = max(1, min(ncpu, cpu_share) // 2)
    env = dict(os.environ, OMP_NUM_THREADS="2", MKL_NUM_THREADS="1")
    t0 = time.time()
    start = t0 + 8.0  # after every child has imported numpy and built its input
    procs = [subprocess.Popen([sys.executable, "-c", _CPU_CHAIN % dict(root=ROOT, nproc=nproc, k=k,
                                                                       sweeps=sweeps, start=start)],
                              stdout=subprocess.PIPE, env=env, text=True)
             for k in range(nproc)]
    res = []
    for pr in procs:
        outp, _ = pr.communicate()
        if pr.returncode != 0:
            raise RuntimeError("CPU baseline process failed")
        res.append(json.loads(outp.strip().splitlines()[-1]))
    rates = [r["sweeps"] / r["seconds"] for r in res]
    single = sum(rates) / len(rates)
    return {"value": sum(rates), "unit": "sweeps/s", "cores": 2 * nproc, "kind": "port",
            "sample": "%d chains x %d sweeps (every 2nd with unequal-time; the initial G build "
                      "timed apart and subtracted) in %d concurrent processes of 2 OpenMP threads "
                      "(stack mode), %.1f s each; per chain %.3f sweeps/s"
                      % (nproc, sweeps, nproc, max(r["seconds"] for r in res), single),
            "single_chain_value": round(single, 4),
            "cpu_model": _cpu_model(), "host_cpus_visible": ncpu,
            "blas": "Intel MKL 2021.4 LP64 sequential (conda libmkl_intel_lp64/sequential/core), "
                    "the BLAS/LAPACK the reference's build/Makefile links"}


# --------------------------------------------------------------------------
# the other configurations (parity workloads of the test suite; each prints
# its own JSON line when selected with --config)
# --------------------------------------------------------------------------
def _run_engines(groups, steps, warmup, threads):
    """Run each engine (one per shape group) for warmup + steps sweeps, the
    engines concurrently from `threads` host threads (each engine has its own
    HIP stream); returns (seconds for the timed sweeps, total chain-sweeps)."""
    import threading

    def pool(fn, what):
        todo = list(groups)
        lock = threading.Lock()
        done = [0]

        def worker():
            while True:
                with lock:
                    if not todo:
                        return
                    eng = todo.pop()
                fn(eng)
                with lock:
                    done[0] += 1
                    if len(groups) > 1 and done[0] % 16 == 0:
                        print("%s: %d/%d engines" % (what, done[0], len(groups)),
                              file=sys.stderr, flush=True)

        ths = [threading.Thread(target=worker) for _ in range(min(threads, len(groups)))]
        for t in ths:
            t.start()
        for t in ths:
            t.join()

    pool(lambda eng: (eng.run(0), eng.run(warmup)), "warmup")  # initial G + warmup, untimed
    t0 = time.perf_counter()
    pool(lambda eng: eng.run(steps), "timed")
    dt = time.perf_counter() - t0
    return dt, steps * sum(e.nc for e in groups)


def _line(metric, value, unit, steps, warmup, dt, workload_desc, extra, nchains):
    out = {"metric": metric, "value": round(value, 4), "unit": unit, "n_gpus": 1,
           "steps": steps, "warmup": warmup, "ms_per_step": round(dt / steps * 1e3, 3),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic (package generator)",
           "config": {"workload": workload_desc, "chains": nchains,
                      "parallelism": "1 GPU, chains batched"}}
    out.update(extra)
    return out


def bench_config1(args):
    """4x4 N=16, t'=0, U=4, mu=0, beta=4 (L=40), n_matmul=8, delay 8, seed 1234,
    100 warmup + 400 measurement sweeps, equal-time measurements."""
    from dqmc_amd import engine, gen
    kw = dict(Nx=4, Ny=4, tp=0.0, U=4.0, mu=0.0, dt=0.1, L=40, n_matmul=8, n_delay=8,
              period_eqlt=8, n_sweep_warm=100, n_sweep_meas=400)
    sim = gen.create_batch(Nfiles=1, seed=1234, **kw)
    eng = engine.Engine(sim, graphs=not args.no_graphs)
    dt1, n1 = _run_engines([eng], 400, 100, 1)
    eng.close()
    nb = args.chains if args.chains != 256 else 1024
    sims = gen.create_batch(Nfiles=nb, seed=1234, **kw)
    engb = engine.Engine(sims, graphs=not args.no_graphs)
    dtb, nbs = _run_engines([engb], args.steps * 25, args.warmup, 1)
    engb.close()
    # F = 5, eq every 8 slices (5 per sweep), no unequal-time
    fl = flop_model(16, 40, 5, 8, 0.5 * 16 * 40, 8, 5, 0.0)
    return _line("DQMC sweeps/s, 4x4 Hubbard N=16 L=40 FP64 (single chain)", n1 / dt1,
                 "sweeps/s", 400, 100, dt1,
                 "4x4 square Hubbard N=16 L=40 U=4 mu=0 beta=4 n_matmul=8 delay=8 seed 1234, "
                 "100 warmup + 400 measurement sweeps, eq every 8 slices",
                 dict({"batched_chains": nb, "batched_sweeps_per_s": round(nbs / dtb, 2),
                       "model_acceptance_assumed": 0.5},
                      **_model(fl, n1 / dt1), **_model(fl, nbs / dtb, "batched_")), 1)


def bench_config3(args):
    """Complex Peierls 12x12 N=144, nflux=1, t'=0, U=4, mu=-0.5, beta=6 (L=60),
    n_matmul=10, delay 16, complex128, eq + unequal-time with the bond
    (current-current jj, jsjs, kk, ksks) correlators."""
    from dqmc_amd import engine, gen
    kw = dict(Nx=12, Ny=12, tp=0.0, U=4.0, mu=-0.5, dt=0.1, L=60, nflux=1, alpha=1.0,
              n_matmul=10, n_delay=16, period_eqlt=10, period_uneqlt=2, n_sweep_warm=0,
              n_sweep_meas=100000, meas_bond_corr=1)
    if args.all_bond_sets:  # also the thermal and two-hop (2-bond) correlators
        kw.update(meas_thermal=1, meas_2bond_corr=1)
    if args.all_meas:
        kw.update({k: v for k, v in ALL_MEAS.items() if k != "meas_nematic_corr"})
    eng = engine.Engine(gen.create_batch(Nfiles=1, seed=3, **kw), graphs=not args.no_graphs)
    dt1, n1 = _run_engines([eng], max(2, args.steps), 1, 1)
    eng.close()
    nb = args.chains  # 256: one chain (one Metropolis workgroup) per CU
    engb = engine.Engine(gen.create_batch(Nfiles=nb, seed=3, **kw), graphs=not args.no_graphs)
    dtb, nbs = _run_engines([engb], args.steps, args.warmup, 1)
    engb.close()
    return _line("DQMC sweeps/s, complex Peierls 12x12 N=144 L=60 (batched chains, aggregate)",
                 nbs / dtb, "sweeps/s", args.steps, args.warmup, dtb,
                 "12x12 square Hubbard, nflux=1, Peierls phases in the Landau gauge (alpha=1; "
                 "at t'=0 gauge-equivalent to the reference generator's alpha=1/2), complex128, "
                 "N=144 L=60 U=4 mu=-0.5 beta=6 n_matmul=10 delay=16, eq every 10 slices, "
                 "uneq + bond correlators every 2 sweeps"
                 + (" (+ thermal and 2-bond sets)" if args.all_bond_sets else "")
                 + (" (+ energy and chiral groups)" if args.all_meas else ""),
                 dict({"single_chain_sweeps_per_s": round(n1 / dt1, 4),
                       "batched_over_single": round(nbs / dtb / (n1 / dt1), 3), "dtype": "c128",
                       "model_acceptance_assumed": 0.5},
                      # complex arithmetic: 4 real flops per complex multiply-add pair;
                      # F = 6, eq every 10 slices, unequal-time every 2 sweeps (bond
                      # correlators not counted)
                      **_model(4 * flop_model(144, 60, 6, 10, 0.5 * 144 * 60, 16, 6, 0.5),
                               nbs / dtb)), nb)


def bench_config4(args):
    """128 chains of uneven sizes: side 6..20, beta in [2, 10] (L = 20..100),
    U in {2, 4, 8}, mu in [-1.5, 0], t'=0, seed 99 + chain (host draws from
    numpy default_rng(99)).  Chains go to five size buckets (sides 6-8, 9-11,
    12-14, 15-17, 18-20), each one engine padded to its largest N and L
    (dqmc_amd.buckets: inert padding sites, identity padding slices).  Each
    bucket's n_matmul is autotuned before the timed region
    (buckets.tune_n_matmul: the fastest of 5, 8, 10, 16, 20 whose wrap drift
    stays below 1e-3, the drift of the reference's own headline setting;
    --no-tune: 10, or 8 for N > 200), with a partial last
    group where it does not divide a chain's L; then the delay rank
    (buckets.tune_delay_rank: the fastest compiled rank of 4, 8, 16, 32;
    --no-tune: each file's n_delay = min(32, N), as compiled); eq
    measurements every n_matmul slices.  The engines run concurrently, each
    sweep one HIP-graph launch per engine (captured before the threads start)."""
    from dqmc_amd import buckets, gen
    nch = args.chains if args.chains != 256 else 128
    r = __import__("numpy").random.default_rng(99)
    edges = ((6, 8), (9, 11), (12, 14), (15, 17), (18, 20))
    groups = {}
    for c in range(nch):
        side = int(r.integers(6, 21))
        L = int(round(float(r.uniform(2.0, 10.0)) / 0.1))
        U = float(r.choice([2.0, 4.0, 8.0]))
        mu = float(r.uniform(-1.5, 0.0))
        N = side * side
        # the chain's own file must satisfy the reference's L % n_matmul
        # (and L % period_eqlt) rule; the bucket engine regroups it
        nm = max(d for d in range(1, 11) if L % d == 0)
        kw = dict(Nx=side, Ny=side, tp=0.0, U=U, mu=mu, dt=0.1, L=L, n_matmul=nm,
                  n_delay=min(32, N), period_eqlt=nm, n_sweep_warm=0, n_sweep_meas=100000)
        sim = gen.create_batch(Nfiles=1, seed=99 + c, **kw)[0]
        bk = next(i for i, (lo, hi) in enumerate(edges) if lo <= side <= hi)
        groups.setdefault(bk, []).append(sim)
    engs, tuned = [], []
    for bk in sorted(groups):
        npad = edges[bk][1] ** 2
        nm = 10 if npad < 200 else 8
        if not args.no_tune:
            nm, table = buckets.tune_n_matmul(groups[bk], graphs=not args.no_graphs)
            tuned.append({"bucket": bk, "n_matmul": nm,
                          "trials": [[n, round(t, 4), d] for n, t, d in table]})
            print("config 4: bucket %d n_matmul trials (n, s/sweep, drift): %s -> %d"
                  % (bk, table, nm), file=sys.stderr, flush=True)
        kd = 0
        if not args.no_tune:
            kd, ktable = buckets.tune_delay_rank(groups[bk], n_matmul=nm,
                                                 graphs=not args.no_graphs)
            tuned[-1].update(delay_rank=kd, rank_trials=[[k, round(t, 4)] for k, t in ktable])
            print("config 4: bucket %d delay rank trials (rank, s/sweep): %s -> %d"
                  % (bk, ktable, kd), file=sys.stderr, flush=True)
        engs.append(buckets.PaddedBatch(groups[bk], nm, graphs=not args.no_graphs,
                                        delay_rank=kd))
        print("config 4: bucket %d: %d chains padded to N=%d L=%d, n_matmul %d, delay rank %d" %
              (bk, len(groups[bk]), engs[-1].N, engs[-1].L, nm, engs[-1].engine.delay_rank()),
              file=sys.stderr, flush=True)
    if not args.no_graphs:
        # capture each engine's sweep graph here, one engine at a time (stream
        # capture fails while another host thread makes HIP calls); every sweep
        # has the same launch pattern (equal-time measurements, no warm-up, no
        # unequal-time), so the concurrent sweeps below only replay graphs
        for e in engs:
            e.run(0)
            e.run(1)
    dt, n = _run_engines(engs, args.steps, args.warmup, len(engs))
    # each chain's own work (its N, L, the bucket's n_matmul, eq every n_matmul
    # slices); padding is overhead, not counted
    fl = 0.0
    for e in engs:
        for Nc, Lc in zip(e.Nc, e.Lc):
            Fc = -(-Lc // e.n_matmul)
            fl += flop_model(Nc, Lc, Fc, e.n_matmul, 0.5 * Nc * Lc, e.engine.delay_rank(),
                             Lc // e.n_matmul, 0.0)
    fl /= nch
    blist = [[e.nc, e.N, e.L, e.n_matmul, e.engine.delay_rank()] for e in engs]
    for e in engs:
        e.close()
    return _line("DQMC sweeps/s, 128 uneven real chains (aggregate)", n / dt, "sweeps/s",
                 args.steps, args.warmup, dt,
                 "%d real chains, side 6..20, L 20..100, U in {2,4,8}, mu in [-1.5,0], "
                 "seed 99+chain; %d size buckets (one padded engine each), run concurrently"
                 % (nch, len(engs)),
                 dict({"buckets": blist,  # chains, N, L, n_matmul, delay rank
                       "tuning": tuned, "model_acceptance_assumed": 0.5},
                      **_model(fl, n / dt)), nch)


def bench_config5(args):
    """Adjoint: 10x10 N=100, t'=-0.2, U=4, mu=-0.3, beta=5 (L=50), 64 fixed
    HS fields from seed 2024, UDT checkpoints every 10 slices: forward
    log|det M_up M_dn| and G, backward gradients for t, t', mu, lambda and the
    full hopping matrix K."""
    import numpy as np
    from dqmc_amd import adjoint
    from dqmc_amd import gen
    nf = 64
    # the 64 fields as create_batch(seed=2024) draws them (the validated ones)
    fields = np.array([s["state"]["hs"] for s in gen.create_batch(
        Nfiles=nf, seed=2024, Nx=10, Ny=10, L=50, n_matmul=10, period_eqlt=10, n_sweep_warm=0,
        n_sweep_meas=1)])
    prob = adjoint.AdjointProblem(10, 10, 50, 0.1, fields, n_matmul=10)
    K = prob.hopping(1.0, -0.2)
    lam, _ = adjoint.lambda_of_U(4.0, 0.1)
    adjoint.logdet_and_grads(prob, K, -0.3, lam)  # warm (engine build, kernels)
    reps = max(1, args.steps)
    t0 = time.perf_counter()
    for _ in range(reps):
        f, gK, gmu, glam, d = adjoint.logdet_and_grads(prob, K, -0.3, lam)
        gt = float(np.sum(gK * prob.t_pattern))
        gtp = float(np.sum(gK * prob.tp_pattern))
    dt = time.perf_counter() - t0
    prob.close()
    return _line("DQMC adjoint evaluations/s, 10x10 N=100 L=50 (forward + gradients per HS field)",
                 nf * reps / dt, "fields/s", reps, 1, dt,
                 "10x10 square Hubbard N=100 L=50 t'=-0.2 U=4 mu=-0.3 beta=5, 64 HS fields "
                 "(seed 2024), checkpoints every 10 slices; log|det M| and dlogdet/d{t,t',mu,"
                 "lambda,K}", {"grad_t_sum": round(gt, 6), "grad_tp_sum": round(gtp, 6)}, nf)


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
                    help="GEMM C tile: 0 = autotune, 1 = 64x64, 2 = 128x64 (K step 16), 3 = 128x64 (K step 8), "
                         "4 = 128x64 on the 16x16x4 form")
    ap.add_argument("--config", type=int, default=2, choices=(1, 2, 3, 4, 5),
                    help="workload: 2 = the headline N=256 L=80 batch (default); 1 = 4x4 "
                         "single chain; 3 = complex Peierls N=144; 4 = 128 uneven chains; "
                         "5 = adjoint gradients")
    ap.add_argument("--no-graphs", action="store_true",
                    help="launch the sweeps kernel by kernel instead of replaying HIP graphs "
                         "(for profilers: rocprofv3 --kernel-trace crashed on graph launches)")
    ap.add_argument("--all-bond-sets", action="store_true",
                    help="--config 3: also measure the thermal and two-hop bond correlators")
    ap.add_argument("--all-meas", action="store_true",
                    help="--config 2 / 3: the optional measurement groups on (profiling the "
                         "contraction kernels; not the headline workload)")
    ap.add_argument("--no-tune", action="store_true",
                    help="--config 4: fixed n_matmul per bucket instead of the autotuner")
    ap.add_argument("--threads", type=int, default=8,
                    help="host threads driving the per-shape engines of --config 4")
    ap.add_argument("--engines", type=int, default=2,
                    help="--config 2: split the chains over this many engines (own HIP streams, "
                         "driven from one host thread each): while one engine's Metropolis "
                         "kernel holds one CU per chain, the other's products use the rest")
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

    if args.config != 2:
        fn = {1: bench_config1, 3: bench_config3, 4: bench_config4, 5: bench_config5}[args.config]
        print(json.dumps(fn(args)))
        return

    tiles = {1: "k_gemm<double,N,N> 64x64 (4x4x4 form)", 2: "k_gemm<double,N,N> 128x64/16 (4x4x4)",
             3: "k_gemm<double,N,N> 128x64/8 (4x4x4)",
             4: "k_gemm_w 128x64/16 (16x16x4 form, fragment-order LDS)"}

    def timed(nch, steps, warmup, time_gemm):
        sims = workload(nch, **(ALL_MEAS if args.all_meas else {}))
        eng = engine.Engine(sims, device=0, time_gemm=time_gemm, gemm_tile=args.gemm_tile,
                            graphs=not args.no_graphs)
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

    def timed_split(nch, ne, steps, warmup):
        """The chains split over ne engines run concurrently, one host thread
        each.  Every engine replays the graphs of both sweep patterns (with and
        without the unequal-time measurement), captured before the threads
        start: stream capture fails while another thread makes HIP calls."""
        import threading
        per = [nch // ne + (1 if i < nch % ne else 0) for i in range(ne)]
        sims = workload(nch, **(ALL_MEAS if args.all_meas else {}))
        engs, k = [], 0
        for n in per:
            eng = engine.Engine(sims[k:k + n], device=0, gemm_tile=args.gemm_tile,
                                graphs=not args.no_graphs)
            k += n
            eng.run(0)  # initial G outside the timed region
            eng.run(warmup)
            engs.append(eng)
        tile = engs[0].gemm_tile()
        acc0 = sum(e.accepts() for e in engs)
        ths = [threading.Thread(target=e.run, args=(steps,)) for e in engs]
        t0 = time.perf_counter()
        for t in ths:
            t.start()
        for t in ths:
            t.join()
        dt = time.perf_counter() - t0
        acc = sum(e.accepts() for e in engs) - acc0
        for e in engs:
            e.close()
        return dt, acc, tile

    nch = args.chains
    # the production path: every sweep one HIP-graph launch per engine (eager
    # with --no-graphs); the chains in args.engines engines run concurrently
    ne = max(1, min(args.engines, nch))
    if ne > 1:
        # two warmup sweeps at least, so that both sweep patterns are captured
        # before the engines run concurrently
        warm = max(args.warmup, 2)
        dt, acc, tile = timed_split(nch, ne, args.steps, warm)
    else:
        warm = args.warmup
        dt, acc, _, tile = timed(nch, args.steps, args.warmup, False)
    value = nch * args.steps / dt
    # the same workload launched kernel by kernel with a HIP event pair around
    # every N x N x N GEMM on the engine stream: the dominant kernel's average
    # duration for the roofline (the event pairs cost a little throughput)
    dte, _, (g_ms, g_n, g_flops), _ = timed(nch, args.steps, 0, True)
    N, L, F, n_matmul = 256, 80, 8, 10
    acc_per_sweep = acc / (nch * args.steps)
    fl = flop_model(N, L, F, n_matmul, acc_per_sweep, 32, L // 8, 0.5)
    model_tflops = fl * value / 1e12
    single = None
    if not args.no_single:
        dts, accs, _, _ = timed(1, max(2, args.steps // 2), 1, False)
        single = max(2, args.steps // 2) / dts
    achieved = g_flops / (g_ms * 1e-3) / 1e12 if g_ms > 0 else 0.0
    # sustained FP64 MFMA rate of a register-only loop (no memory traffic):
    # the ceiling of the instruction forms on this device, beside the
    # published peak
    loop = {"16x16x4": round(engine.probe_mfma(0), 2), "4x4x4": round(engine.probe_mfma(1), 2)}
    line = {
        "metric": "DQMC sweeps/s at N=256 L=80 FP64 (batched chains, aggregate)",
        "value": round(value, 4),
        "unit": "sweeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": warm,
        "ms_per_step": round(dt / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (package generator, seed 7 + jump per chain)",
        "config": {"workload": "16x16 square Hubbard N=256 L=80 t'=-0.25 U=6 mu=-1 beta=8 "
                               "n_matmul=10 delay rank 32 (A in registers, B in LDS) eq every 8 slices, "
                               "uneq every 2 sweeps"
                               + (" + bond, thermal, 2-bond, energy, nematic and chiral groups "
                                  "(profiling run, not the headline workload)"
                                  if args.all_meas else ""),
                   "chains": nch, "parallelism": "1 GPU, chains batched",
                   "engines": ne},
        "graphs": not args.no_graphs,
        "eager_timed_gemms_sweeps_per_s": round(nch * args.steps / dte, 4),
        "single_chain_sweeps_per_s": None if single is None else round(single, 4),
        "batched_over_single": None if single is None else round(value / single, 3),
        "model_gflop_per_sweep": round(fl / 1e9, 2),
        "model_tflops": round(model_tflops, 3),
        "model_frac_fp64_peak": round(model_tflops / PEAK_FP64_TFLOPS, 4),
        "accepts_per_sweep": round(acc_per_sweep, 1),
        "roofline": {"bound": "mfma",
                     "kernel": "%s (N x N x N, batch 2*chains)" % tiles[tile],
                     "achieved": round(achieved, 3), "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / PEAK_FP64_TFLOPS, 4),
                     "traffic": GEMM_HBM_BYTES_PER_MATRIX_N256 * 2 * nch,
                     "traffic_source": "FETCH_SIZE + WRITE_SIZE counter passes of this kernel "
                                       "(profiles/k_gemm_N256_L80_256chains.md), per launch",
                     "avg_launch_ms": round(g_ms, 5), "launches": g_n,
                     "timing": "HIP events on the engine stream around every launch, eager pass",
                     "mfma_loop_tflops": loop},
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline()
    print(json.dumps(line))


if __name__ == "__main__":
    main()
