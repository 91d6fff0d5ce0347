"""The C-ABI library loads and exports every entry point include/dqmc_hip.h
declares (no GPU needed: nothing is called but dqmc_version / dqmc_acc_len)."""
import os
import re

import numpy as np

from dqmc_amd import capi, engine, gen

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "dqmc_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(dqmc_[a-z_0-9]+)\s*\(", src)))


def test_all_symbols_exported():
    L = engine.lib()
    names = declared()
    assert "dqmc_engine_run" in names and "dqmc_engine_create" in names
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert b"gfx950" in L.dqmc_version()


def test_acc_len_matches_generator_sizes():
    L = engine.lib()
    sim = gen.create_batch(seed=1, Nx=4, Ny=4, tp=-0.2, L=20, n_matmul=5, period_eqlt=5,
                           period_uneqlt=1, meas_energy_corr=1, meas_bond_corr=1)[0]
    ps, keep = capi.make_params(sim["params"])
    for ue, names, grp in ((0, capi.EQ_ACC, sim["meas_eqlt"]), (1, capi.UE_ACC, sim["meas_uneqlt"])):
        for k, name in enumerate(names):
            n = L.dqmc_acc_len(ps, ue, k)
            assert n == (grp[name].size if name in grp else 0), (ue, name)
