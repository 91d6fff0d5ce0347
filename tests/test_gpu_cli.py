"""The process seam on the GPU: dqmc_1 / dqmc_stack over .h5 files, checked
against the oracle run of the same file (src/main_1.c, src/main_stack.c,
src/dqmc.c:513-671 semantics: checkpoints, time limit, resume, batching)."""
import os
import subprocess

import numpy as np
import pytest

from dqmc_amd import gen, h5
import oracle

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "dqmc-dev_amd", "bin")


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _run(args, timeout=120):
    print("run:", " ".join(args), flush=True)
    return subprocess.run(args, capture_output=True, text=True, timeout=timeout)


def _logs(files):
    out = ""
    for fn in files:
        if os.path.exists(fn + ".log"):
            out += "\n--- %s.log\n%s" % (fn, open(fn + ".log").read()[-1500:])
    return out


def _check_against_oracle(fn, sim, eq_keys=("density", "double_occ", "nn", "zz", "pair_sw"),
                          ue_keys=(), eq_tol=1e-9):
    r = h5.read_sim(fn)
    out = oracle.run(sim)
    assert np.array_equal(r["state"]["hs"], out["state"]["hs"]), fn
    assert [int(v) for v in r["state"]["rng"]] == [int(v) for v in out["state"]["rng"]], fn
    assert int(r["state"]["sweep"]) == int(sim["params"]["n_sweep"])
    assert int(r["state"]["partial_write"]) == 0
    assert int(r["meas_eqlt"]["n_sample"]) == int(out["meas_eqlt"]["n_sample"])
    for k in eq_keys:
        assert rel(r["meas_eqlt"][k], out["meas_eqlt"][k]) < eq_tol, (fn, k)
    for k in ue_keys:
        assert rel(r["meas_uneqlt"][k], out["meas_uneqlt"][k]) < 1e-8, (fn, k)
    return r


def test_dqmc_1_golden_file_with_checkpoints(tmp_path, golden):
    sim = gen.create_batch(seed=1234, n_sweep_warm=20, n_sweep_meas=20, checkpoint_every=7)[0]
    fn = str(tmp_path / "golden.h5")
    h5.write_sim(fn, sim)
    r = _run([os.path.join(BIN, "dqmc_1"), "-l", fn + ".log", fn])
    log = open(fn + ".log").read()
    assert r.returncode == 0, log
    assert "checkpointing every 7 sweeps" in log and "sim_data_save() succeeded" in log
    out = _check_against_oracle(fn, sim)
    assert np.array_equal(out["state"]["hs"], np.array(golden["hs_end"]["square"]))


@pytest.mark.parametrize("geometry", ["honeycomb", "kagome"])
def test_dqmc_1_other_lattices_with_unequal_time(tmp_path, geometry):
    """A two- and a three-orbital lattice through the command-line driver, with
    flux (complex path) and unequal-time measurements, against the oracle."""
    sim = gen.create_batch(seed=71, geometry=geometry, Nx=4, Ny=4, U=4.0, mu=-0.2, L=20,
                           n_matmul=5, period_eqlt=5, period_uneqlt=2, nflux=1,
                           n_sweep_warm=4, n_sweep_meas=6)[0]
    fn = str(tmp_path / (geometry + ".h5"))
    h5.write_sim(fn, sim)
    r = _run([os.path.join(BIN, "dqmc_1"), "-l", fn + ".log", fn])
    assert r.returncode == 0, open(fn + ".log").read()
    _check_against_oracle(fn, sim, ue_keys=("gt0", "nn", "xx", "zz", "pair_sw"))


def test_dqmc_stack_batches_two_shapes(tmp_path):
    a = gen.create_batch(Nfiles=3, seed=31, Nx=4, Ny=4, U=5.0, mu=-0.4, L=20, n_matmul=5,
                         period_eqlt=5, period_uneqlt=2, n_sweep_warm=2, n_sweep_meas=6)
    b = gen.create_batch(Nfiles=2, seed=32, Nx=4, Ny=6, U=4.0, mu=-0.2, L=24, n_matmul=6,
                         period_eqlt=6, nflux=1, n_sweep_warm=2, n_sweep_meas=5)
    fa = h5.create_batch_files(str(tmp_path / "a"), a)
    fb = h5.create_batch_files(str(tmp_path / "b"), b)
    stack = str(tmp_path / "stack")
    with open(stack, "w") as f:
        f.write("\n".join(fa + fb) + "\n")
    r = _run([os.path.join(BIN, "dqmc_stack"), "-B", "8", stack])
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count("completed:") == 5, r.stdout + _logs(fa + fb)
    assert open(stack).read().strip() == ""
    for fn, s in zip(fa, a):
        _check_against_oracle(fn, s, ue_keys=("gt0", "nn", "zz"))
    for fn, s in zip(fb, b):
        _check_against_oracle(fn, s)


def test_time_limit_interrupts_and_resume_matches_oracle(tmp_path):
    sim = gen.create_batch(seed=41, Nx=4, Ny=4, U=4.0, mu=-0.3, L=40, n_matmul=8, period_eqlt=8,
                           n_sweep_warm=10, n_sweep_meas=40000, checkpoint_every=10)[0]
    fn = str(tmp_path / "long.h5")
    h5.write_sim(fn, sim)
    # 40010 sweeps cannot finish in 2 s: the limit stops the run (src/sig.c:70-75)
    r = _run([os.path.join(BIN, "dqmc_1"), "-t", "2", "-l", fn + ".log", fn])
    log = open(fn + ".log").read()
    assert r.returncode == 1, log  # WRAP_INCOMPLETE
    assert "dqmc() interrupted" in log
    ck = h5.read_sim(fn)
    s0 = int(ck["state"]["sweep"])
    assert s0 % 10 == 0 and int(ck["state"]["partial_write"]) == 0
    # resume from the checkpoint for 25 more sweeps (n_sweep edited in the file)
    n_end = s0 + 25
    with h5.File(fn, "r+") as f:
        f.overwrite("/params/n_sweep", np.int32(n_end))
    resumed = h5.read_sim(fn)
    r = _run([os.path.join(BIN, "dqmc_1"), "-l", fn + ".log", fn])
    assert r.returncode == 0, open(fn + ".log").read()
    _check_against_oracle(fn, resumed)


def test_time_limit_without_checkpoints_stops_within_a_few_sweeps(tmp_path):
    """checkpoint_every = 0: the whole run is one dqmc_engine_run call.  The
    time limit must still end it within a few sweeps (the engine queues a
    sweep only after the one two before it has finished), not after every
    sweep the host could queue.  10^5 sweeps of an 8x8 chain take minutes."""
    import time
    sim = gen.create_batch(seed=43, Nx=8, Ny=8, U=4.0, mu=-0.3, L=40, n_matmul=8, period_eqlt=8,
                           n_sweep_warm=0, n_sweep_meas=100000, checkpoint_every=0)[0]
    fn = str(tmp_path / "nockpt.h5")
    h5.write_sim(fn, sim)
    t0 = time.time()
    r = _run([os.path.join(BIN, "dqmc_1"), "-t", "3", "-l", fn + ".log", fn], timeout=90)
    wall = time.time() - t0
    log = open(fn + ".log").read()
    print("wall %.1f s with a 3 s limit" % wall, flush=True)
    assert r.returncode == 1, log  # WRAP_INCOMPLETE
    assert "dqmc() interrupted" in log and "time limit" in log
    assert wall < 3 + 15, wall
    assert int(h5.read_sim(fn)["state"]["sweep"]) == 0  # nothing saved after an interruption


def test_dqmc_stack_pads_uneven_sizes(tmp_path):
    # files of uneven N and L with one n_matmul / period_eqlt share a padded
    # engine per element type; each must end as the oracle run of its own file:
    # the same field and RNG state bit for bit, checkpoints included.  The
    # accumulators are held to 1e-8 relative: at U = 6 with ten slices between
    # recomputations the wraps amplify rounding ~exp(2 lambda 9) ~ 1e6 times,
    # and sums of products of G entries then differ by ~1e-9 between any two
    # correct evaluation orders (the padded G is factored at another size)
    real = [gen.create_batch(seed=50 + k, Nx=nx, Ny=ny, U=U, mu=-0.3, L=L, n_matmul=10,
                             period_eqlt=10, n_sweep_warm=2, n_sweep_meas=5,
                             checkpoint_every=3)[0]
            for k, (nx, ny, L, U) in enumerate([(4, 4, 20, 4.0), (5, 5, 30, 6.0), (6, 4, 10, 3.0)])]
    cplx = [gen.create_batch(seed=60 + k, Nx=nx, Ny=ny, U=4.0, mu=-0.2, L=L, n_matmul=10,
                             period_eqlt=10, nflux=1, n_sweep_warm=2, n_sweep_meas=4)[0]
            for k, (nx, ny, L) in enumerate([(4, 4, 20), (5, 4, 30)])]
    files = []
    for k, s in enumerate(real + cplx):
        fn = str(tmp_path / ("u%d.h5" % k))
        h5.write_sim(fn, s)
        files.append(fn)
    stack = str(tmp_path / "stack")
    with open(stack, "w") as f:
        f.write("\n".join(files) + "\n")
    r = _run([os.path.join(BIN, "dqmc_stack"), "-B", "8", stack])
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count("completed:") == len(files), r.stdout + _logs(files)
    for fn in files:
        assert "padded from" in open(fn + ".log").read(), fn
    for fn, s in zip(files, real + cplx):
        print("file", fn, "N", int(s["params"]["N"]), "L", int(s["params"]["L"]), flush=True)
        _check_against_oracle(fn, s, eq_keys=("density", "double_occ", "g00", "nn", "xx", "zz",
                                              "pair_sw"), eq_tol=1e-8)


def test_dqmc_stack_pads_uneven_sizes_with_unequal_time(tmp_path):
    # as above with unequal-time measurements every sweep after warm-up: the
    # padded engine's identity slices follow each file's last slice, so the
    # file's G(t,0), G(0,t), G(t,t) are its own; the accumulators of its
    # slices and bins match the oracle run of the file (calc_ue_g's 1e-8)
    real = [gen.create_batch(seed=90 + k, Nx=nx, Ny=ny, U=4.0, mu=-0.3, L=L, n_matmul=10,
                             period_eqlt=5, period_uneqlt=1, n_sweep_warm=1, n_sweep_meas=3)[0]
            for k, (nx, ny, L) in enumerate([(4, 4, 20), (5, 5, 30), (3, 4, 10)])]
    cplx = [gen.create_batch(seed=95 + k, Nx=nx, Ny=ny, U=4.0, mu=-0.2, L=L, n_matmul=10,
                             period_eqlt=5, period_uneqlt=1, nflux=1, n_sweep_warm=1,
                             n_sweep_meas=3)[0]
            for k, (nx, ny, L) in enumerate([(4, 4, 20), (5, 3, 30)])]
    files = []
    for k, s in enumerate(real + cplx):
        fn = str(tmp_path / ("w%d.h5" % k))
        h5.write_sim(fn, s)
        files.append(fn)
    stack = str(tmp_path / "stack")
    with open(stack, "w") as f:
        f.write("\n".join(files) + "\n")
    r = _run([os.path.join(BIN, "dqmc_stack"), "-B", "8", stack])
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count("completed:") == len(files), r.stdout + _logs(files)
    for fn in files:
        assert "padded from" in open(fn + ".log").read(), fn
    for fn, s in zip(files, real + cplx):
        print("file", fn, "N", int(s["params"]["N"]), "L", int(s["params"]["L"]), flush=True)
        _check_against_oracle(fn, s, eq_keys=("density", "double_occ", "nn", "zz"), eq_tol=1e-8,
                              ue_keys=("gt0", "gt0_u", "nn", "xx", "zz", "pair_sw"))


def test_dqmc_1_profile_table(tmp_path):
    """-p prints the reference's profile table (src/prof.c:12-50) into the log."""
    sim = gen.create_batch(seed=44, Nx=4, Ny=4, U=4.0, mu=-0.3, L=20, n_matmul=5, period_eqlt=5,
                           period_uneqlt=2, n_sweep_warm=2, n_sweep_meas=4, checkpoint_every=3)[0]
    fn = str(tmp_path / "prof.h5")
    h5.write_sim(fn, sim)
    r = _run([os.path.join(BIN, "dqmc_1"), "-p", "-l", fn + ".log", fn])
    log = open(fn + ".log").read()
    assert r.returncode == 0, log
    assert "_% of all_|___total (s)_|___us per call_|___# calls" in log
    rows = {ln.split("|")[0].strip(): ln for ln in log.splitlines() if ln.count("|") == 4}
    for k in ("wall", "updates", "recalc", "wrap", "half_wrap", "meas_eq", "calc_o", "bsofi",
              "expand_g", "meas_uneq", "checkpoint"):
        assert k in rows, (k, log[-2000:])
    assert int(rows["checkpoint"].split("|")[4]) == 2  # saves before sweeps 0 and 3
    _check_against_oracle(fn, sim, ue_keys=("gt0", "nn", "zz"))
