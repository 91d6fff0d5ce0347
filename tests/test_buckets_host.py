"""Host side of the size buckets (dqmc_amd.buckets.pad_sim): the padded
parameters describe the same chain plus inert sites, and padding sites and
pairs belong to no measurement bin."""
import numpy as np
import pytest

from dqmc_amd import buckets, gen


def _sim(side=4, L=20, nm=10, **kw):
    return gen.create_batch(Nfiles=1, seed=5, Nx=side, Ny=side, L=L, n_matmul=nm,
                            period_eqlt=nm, U=4.0, mu=-0.3, **kw)[0]


def test_padding_keeps_the_chain_block_and_adds_identity():
    s = _sim()
    q = buckets.pad_sim(s, 25, 30, 10, 2, 17)
    p, pp = s["params"], q["params"]
    assert int(pp["N"]) == 25 and int(pp["L"]) == 30 and int(pp["F"]) == 3
    for k in ("exp_Ku", "inv_exp_Kd", "exp_halfKu"):
        a = pp[k]
        assert np.array_equal(a[:16, :16], p[k])
        assert np.array_equal(a[16:, 16:], np.eye(9))
        assert not a[:16, 16:].any() and not a[16:, :16].any()
    # inert sites: no coupling to the field, never flipped
    assert np.all(pp["exp_lambda"][:, 16:] == 1.0) and np.all(pp["del"][:, 16:] == 0.0)
    assert np.array_equal(pp["exp_lambda"][:, :16], np.asarray(p["exp_lambda"]).reshape(2, 16))
    # the chain's bins are unchanged, padding sites and pairs belong to none
    assert np.array_equal(pp["map_i"][:16], np.asarray(p["map_i"]).reshape(16))
    assert np.all(pp["map_i"][16:] == -1)
    mij = pp["map_ij"]
    assert np.array_equal(mij[:16, :16], np.asarray(p["map_ij"]).reshape(16, 16))
    assert np.all(mij[16:, :] == -1) and np.all(mij[:, 16:] == -1)
    assert q["state"]["hs"].shape == (30, 25)
    assert np.array_equal(q["state"]["hs"][:20, :16], s["state"]["hs"])
    assert q["meas_eqlt"]["nn"].shape == (17,) and q["meas_eqlt"]["density"].shape == (2,)


def test_padding_refuses_what_needs_equal_sizes():
    with pytest.raises(ValueError):
        buckets.pad_sim(_sim(meas_bond_corr=1), 25, 30, 10, 2, 17)
    with pytest.raises(ValueError):
        buckets.pad_sim(_sim(), 9, 30, 10, 2, 17)


def test_padding_keeps_the_chains_equal_time_cadence():
    s = gen.create_batch(Nfiles=1, seed=5, Nx=4, Ny=4, L=20, n_matmul=10, period_eqlt=5,
                         U=4.0)[0]
    q = buckets.pad_sim(s, 25, 30, 10, 2, 17)
    assert int(q["params"]["period_eqlt"]) == 5 and int(q["params"]["n_matmul"]) == 10


def test_padding_places_unequal_time_accumulators_by_slice_and_bin():
    s = _sim(period_uneqlt=2)
    rs = np.random.RandomState(0)
    for k in buckets._UE_IJ:
        s["meas_uneqlt"][k] = rs.randn(*np.shape(s["meas_uneqlt"][k]))
    q = buckets.pad_sim(s, 25, 30, 10, 2, 17)
    for k in buckets._UE_IJ:
        a = q["meas_uneqlt"][k].reshape(30, 17)
        assert np.array_equal(a[:20, :16], s["meas_uneqlt"][k].reshape(20, 16))
        assert not a[20:, :].any() and not a[:, 16:].any()
    assert int(q["params"]["period_uneqlt"]) == 2
