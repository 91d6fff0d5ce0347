"""The file and process seams on the CPU: .h5 layout round trip (the layout
util/gen_1band_unified_hub.py:929-1160 writes and src/data.c:345-789 reads),
dqmc_1 / dqmc_stack command lines in dry-run mode, and the stack file's LIFO
pop order (src/main_stack.c:120-205)."""
import os
import subprocess

import numpy as np
import pytest

from dqmc_amd import gen, h5

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "dqmc-dev_amd", "bin")


def _same(a, b):
    if isinstance(b, str):
        return a == b
    if isinstance(b, (int, float)):
        return np.asarray(a).shape == () and a == b
    b = np.asarray(b)
    return a.dtype == b.dtype and a.shape == b.shape and np.array_equal(a, b)


@pytest.mark.parametrize("geometry", ["square", "honeycomb", "kagome"])
@pytest.mark.parametrize("nflux", [0, 1])
def test_h5_roundtrip_batch_layout(tmp_path, nflux, geometry):
    # bond measurements are defined on the square lattice only
    extra = dict(meas_bond_corr=1, meas_energy_corr=1) if geometry == "square" else {}
    sims = gen.create_batch(Nfiles=3, seed=5, geometry=geometry, Nx=4, Ny=4, L=20, n_matmul=5,
                            period_eqlt=5, period_uneqlt=1, nflux=nflux, **extra)
    files = h5.create_batch_files(str(tmp_path / "sim"), sims)
    assert os.path.exists(str(tmp_path / "sim.h5.params"))
    for fn, s in zip(files, sims):
        r = h5.read_sim(fn)
        for g in ("metadata", "params", "state", "meas_eqlt", "meas_uneqlt"):
            assert set(r[g]) == set(s[g]), g
            for k, v in s[g].items():
                assert _same(r[g][k], v), (fn, g, k)
    # complex accumulators are the compound {"r","i"} type, real ones float64
    want = np.complex128 if nflux else np.float64
    assert h5.read_sim(files[0])["params"]["exp_Ku"].dtype == want


def test_dqmc_1_dry_run(tmp_path):
    s = gen.create_batch(Nfiles=1, seed=3, Nx=4, Ny=4, L=20, n_matmul=5, period_eqlt=5)[0]
    fn = str(tmp_path / "a.h5")
    h5.write_sim(fn, s)
    log = str(tmp_path / "a.log")
    r = subprocess.run([os.path.join(BIN, "dqmc_1"), "-n", "-l", log, fn], capture_output=True)
    assert r.returncode == 0
    text = open(log).read()
    assert "minimum RAM requirement" in text and "toggled dry run" in text
    assert h5.read_sim(fn)["state"]["sweep"] == 0  # untouched


def test_dqmc_1_rejects_partial_write(tmp_path):
    s = gen.create_batch(Nfiles=1, seed=3, Nx=4, Ny=4, L=20, n_matmul=5, period_eqlt=5)[0]
    fn = str(tmp_path / "a.h5")
    h5.write_sim(fn, s)
    with h5.File(fn, "r+") as f:
        f.overwrite("/state/partial_write", 1)
    r = subprocess.run([os.path.join(BIN, "dqmc_1"), fn], capture_output=True, text=True)
    assert r.returncode == 255  # WRAP_FAIL
    assert "sim_data_read_alloc() failed: -2" in r.stdout  # CORRUPT_FAIL (src/data.c:11)


def test_dqmc_stack_dry_run_pops_lifo(tmp_path):
    sims = gen.create_batch(Nfiles=3, seed=4, Nx=4, Ny=4, L=20, n_matmul=5, period_eqlt=5)
    files = h5.create_batch_files(str(tmp_path / "s"), sims)
    stack = str(tmp_path / "stack")
    with open(stack, "w") as f:
        f.write("\n".join(files) + "\n\n")
    r = subprocess.run([os.path.join(BIN, "dqmc_stack"), "-n", "-B", "2", stack],
                       capture_output=True, text=True)
    assert r.returncode == 0
    started = [ln.split("starting:")[1].strip() for ln in r.stdout.splitlines() if "starting:" in ln]
    assert started == files[::-1]  # last line first
    assert open(stack).read().strip() == ""
    assert not os.path.exists(stack + "~")  # lock released
    for fn in files:
        assert os.path.exists(fn + ".log")


def test_dqmc_stack_missing_stack_file(tmp_path):
    r = subprocess.run([os.path.join(BIN, "dqmc_stack"), str(tmp_path / "nope")],
                       capture_output=True, text=True)
    assert r.returncode == 1  # EXIT_FAILURE (src/main_stack.c:425-429)
