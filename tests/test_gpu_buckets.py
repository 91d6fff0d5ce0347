"""GPU: chains of uneven size padded into one engine (dqmc_amd.buckets) run as
they run alone.

A chain whose L is a multiple of the engine's n_matmul has the same random
stream, recalculation points and measurement slices (each chain at its own
period_eqlt) as in an engine of its own: its HS field and RNG state
must match bit for bit, G and the equal-time accumulators to 1e-10 relative.
A chain whose last group is partial (L = 23 in groups of 10) is compared with
its own engine at n_matmul = 1: the same Markov chain, stabilised at other
slices, so the fields agree and G agrees to 1e-9.
"""
import numpy as np
import pytest

from dqmc_amd import buckets, engine, gen

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(np.asarray(b)).max(), 1e-300)


def _sim(side, L, nm, seed, U=4.0, mu=-0.5):
    return gen.create_batch(Nfiles=1, seed=seed, Nx=side, Ny=side, L=L, n_matmul=nm,
                            period_eqlt=nm, U=U, mu=mu, tp=0.0, n_sweep_warm=1,
                            n_sweep_meas=100)[0]


def _alone(sim, sweeps):
    e = engine.Engine([sim])
    e.run(sweeps)
    out = (e.get_state(0), e.get_meas(0)[0], e.green(0, 0), e.green(0, 1), e.accepts())
    e.close()
    return out


def test_padded_chains_match_their_own_engines():
    shapes = [(4, 20, 10, 11, 4.0), (6, 30, 10, 12, 8.0), (5, 10, 10, 13, 2.0), (6, 20, 10, 14, 4.0)]
    sims = [_sim(s, L, nm, seed, U=U) for s, L, nm, seed, U in shapes]
    sweeps = 4
    pb = buckets.PaddedBatch(sims, 10)
    assert (pb.N, pb.L) == (36, 30)
    pb.run(sweeps)
    for c, sim in enumerate(sims):
        st, meq, (gu, pu), (gd, pd), _ = _alone(sim, sweeps)
        s2 = pb.state(c)
        print("chain", c, "shape", shapes[c][:2], flush=True)
        assert int(s2["sweep"]) == int(st["sweep"]) == sweeps
        assert np.array_equal(s2["rng"], st["rng"])
        assert np.array_equal(s2["hs"], st["hs"])
        g2u, p2u = pb.green(c, 0)
        g2d, p2d = pb.green(c, 1)
        assert rel(g2u, gu) < 1e-10 and rel(g2d, gd) < 1e-10
        assert abs(p2u - pu) < 1e-10
        m2 = pb.meas(c)
        assert int(m2["n_sample"]) == int(meq["n_sample"]) > 0
        for k in ("density", "double_occ", "g00", "nn", "xx", "zz", "pair_sw"):
            assert m2[k].shape == meq[k].shape
            assert rel(m2[k], meq[k]) < 1e-10, k
    pb.close()


def test_padded_chains_keep_their_own_equal_time_cadence():
    """Chains with different period_eqlt share one padded engine; each is
    measured at its own cadence (src/dqmc.c:361), so its sample count and
    accumulators are those of an engine of its own."""
    sims = [gen.create_batch(Nfiles=1, seed=70 + k, Nx=side, Ny=side, L=L, n_matmul=10,
                             period_eqlt=pe, U=4.0, mu=-0.4, tp=0.0, n_sweep_warm=1,
                             n_sweep_meas=100)[0]
            for k, (side, L, pe) in enumerate([(4, 20, 5), (5, 30, 10), (4, 30, 3), (6, 10, 2)])]
    sweeps = 3
    pb = buckets.PaddedBatch(sims, 10)
    pb.run(sweeps)
    for c, sim in enumerate(sims):
        st, meq, _, _, _ = _alone(sim, sweeps)
        m2 = pb.meas(c)
        L, pe = int(sim["params"]["L"]), int(sim["params"]["period_eqlt"])
        assert int(m2["n_sample"]) == int(meq["n_sample"]) == (sweeps - 1) * (L // pe), c
        assert np.array_equal(pb.state(c)["hs"], st["hs"])
        for k in ("density", "double_occ", "nn", "zz"):
            assert rel(m2[k], meq[k]) < 1e-10, (c, k)
    pb.close()


def test_partial_last_group():
    a = _sim(5, 23, 1, 21)
    b = _sim(4, 30, 10, 22)
    sweeps = 3
    pb = buckets.PaddedBatch([a, b], 10)
    assert pb.L == 30
    pb.run(sweeps)
    st, _, (gu, _), (gd, _), _ = _alone(a, sweeps)
    s2 = pb.state(0)
    assert np.array_equal(s2["rng"], st["rng"])
    assert np.array_equal(s2["hs"], st["hs"])
    assert rel(pb.green(0, 0)[0], gu) < 1e-9 and rel(pb.green(0, 1)[0], gd) < 1e-9
    pb.close()


@pytest.mark.parametrize("nflux", [0, 1])
def test_padded_chains_measure_unequal_time_as_alone(nflux):
    """Unequal-time measurements on chains of uneven N and L in one engine:
    fields and RNG bit for bit, the unequal-time accumulators of the chain's
    own slices and bins to 1e-8 (calc_ue_g's bar; the padded engine groups
    the slices into blocks of another count, which moves only rounding)."""
    shapes = [(4, 20, 81), (5, 30, 82), (3, 10, 83)]
    sims = [gen.create_batch(Nfiles=1, seed=seed, Nx=side, Ny=side, L=L, n_matmul=10,
                             period_eqlt=5, period_uneqlt=1, U=4.0, mu=-0.3, nflux=nflux,
                             n_sweep_warm=1, n_sweep_meas=100)[0]
            for side, L, seed in shapes]
    sweeps = 3
    pb = buckets.PaddedBatch(sims, 10)
    try:
        pb.run(sweeps)
        for c, sim in enumerate(sims):
            print("chain", c, "shape", shapes[c][:2], flush=True)
            e = engine.Engine([sim])
            try:
                e.run(sweeps)
                st = e.get_state(0)
                _, ue = e.get_meas(0)
            finally:
                e.close()
            s2 = pb.state(c)
            assert np.array_equal(s2["rng"], st["rng"]) and np.array_equal(s2["hs"], st["hs"])
            u2 = pb.meas_uneqlt(c)
            assert int(u2["n_sample"]) == int(ue["n_sample"]) == sweeps - 1
            for k in ("gt0", "gt0_u", "gt0_d", "nn", "xx", "zz", "pair_sw"):
                assert u2[k].shape == ue[k].shape, k
                assert rel(u2[k], ue[k]) < 1e-8, (c, k)
    finally:
        pb.close()


def test_padded_chains_share_period_uneqlt():
    a = gen.create_batch(Nfiles=1, seed=3, Nx=4, Ny=4, L=20, n_matmul=10, period_eqlt=10,
                         period_uneqlt=1)[0]
    b = gen.create_batch(Nfiles=1, seed=4, Nx=3, Ny=3, L=20, n_matmul=10, period_eqlt=10,
                         period_uneqlt=2)[0]
    with pytest.raises(ValueError):
        buckets.PaddedBatch([a, b], 10)


def test_wrap_drift_check_leaves_the_chain_unchanged():
    # the CHECK_G_WRP comparison only reads G: the chain is the same bit for
    # bit; the drift (wrap rounding, ~1e-8..1e-7 at n_matmul = 10, U = 4 on
    # 6x6) is small, grows with the interval, and reading it resets it
    sims = [_sim(6, 40, 10, 31 + k) for k in range(3)]
    a = engine.Engine(sims, check_drift=True)
    b = engine.Engine(sims)
    try:
        a.run(3)
        b.run(3)
        for c in range(3):
            sa, sb = a.get_state(c), b.get_state(c)
            assert np.array_equal(sa["hs"], sb["hs"]) and np.array_equal(sa["rng"], sb["rng"])
            assert np.array_equal(a.green(c, 0)[0], b.green(c, 0)[0])
        d = a.drift()
        print("drift", d, flush=True)
        assert d.shape == (3,) and np.all(d > 0) and np.all(d < 1e-5)
        assert np.all(a.drift() == 0.0)
    finally:
        a.close()
        b.close()
    wide = engine.Engine([_sim(6, 40, 20, 31, U=8.0)], check_drift=True)
    narrow = engine.Engine([_sim(6, 40, 5, 31, U=8.0)], check_drift=True)
    try:
        wide.run(2)
        narrow.run(2)
        dw, dn = wide.drift()[0], narrow.drift()[0]
        print("drift n=20 %.3g, n=5 %.3g" % (dw, dn), flush=True)
        assert dw > dn > 0
    finally:
        wide.close()
        narrow.close()


def test_n_matmul_autotuner_picks_a_stable_interval():
    sims = [_sim(4, 20, 10, 41), _sim(5, 40, 10, 42), _sim(6, 30, 10, 43)]
    pick, table = buckets.tune_n_matmul(sims, candidates=(5, 10, 20, 40), sweeps=1)
    print("table", table, "pick", pick, flush=True)
    assert [n for n, _, _ in table] == [5, 10, 20]  # 40 > the shortest L
    assert all(t > 0 and d > 0 for _, t, d in table)
    ok = [(n, t) for n, t, d in table if d <= 1e-3]
    assert pick == (min(ok, key=lambda x: x[1])[0] if ok else 5)


def test_regrouped_interval_matches_the_oracle_at_that_interval():
    """A bucket run at an n_matmul other than the files' (the autotuner's
    choice): each chain is the oracle run of its own file with n_matmul (and
    period_eqlt, the cadence the reference ties to it) set to the bucket's --
    the field and RNG bit for bit, G and the accumulators to rounding."""
    import oracle
    sims = [_sim(4, 20, 10, 51), _sim(5, 40, 10, 52), _sim(6, 30, 10, 53)]
    sweeps = 4
    for nm in (5, 10):
        pb = buckets.PaddedBatch(sims, nm)
        try:
            pb.run(sweeps)
            for c, sim in enumerate(sims):
                ref = gen.create_batch(Nfiles=1, seed=51 + c, Nx=[4, 5, 6][c], Ny=[4, 5, 6][c],
                                       L=[20, 40, 30][c], n_matmul=nm, period_eqlt=10, U=4.0,
                                       mu=-0.5, tp=0.0, n_sweep_warm=1, n_sweep_meas=sweeps - 1)[0]
                assert np.array_equal(ref["state"]["hs"], sim["state"]["hs"])
                out = oracle.run(ref)
                st = pb.state(c)
                print("n_matmul", nm, "chain", c, flush=True)
                assert np.array_equal(st["hs"], out["state"]["hs"])
                assert [int(v) for v in st["rng"]] == [int(v) for v in out["state"]["rng"]]
                m = pb.meas(c)
                assert int(m["n_sample"]) == int(out["meas_eqlt"]["n_sample"])
                for k in ("density", "double_occ", "nn", "xx", "zz", "pair_sw"):
                    assert rel(m[k], out["meas_eqlt"][k]) < 1e-9, (nm, c, k)
        finally:
            pb.close()


def test_delay_rank_autotuner_times_each_compiled_rank():
    sims = [_sim(9, 20, 10, 44), _sim(10, 40, 10, 45)]  # N = 81, 100: ranks 8, 16, 32
    hs0 = [np.array(s["state"]["hs"]) for s in sims]
    pick, table = buckets.tune_delay_rank(sims, n_matmul=10, candidates=(4, 8, 12, 16, 32), sweeps=1)
    print("table", table, "pick", pick, flush=True)
    assert [k for k, _ in table] == [8, 16, 32]  # 4 -> 8 and 12 -> 8 share a variant
    assert all(t > 0 for _, t in table)
    assert pick == min(table, key=lambda r: r[1])[0]
    assert all(np.array_equal(np.asarray(s["state"]["hs"]), h) for s, h in zip(sims, hs0))
    # chains of one shape at their own interval run unpadded
    one = [_sim(10, 20, 10, 46), _sim(10, 20, 10, 47)]
    pick1, table1 = buckets.tune_delay_rank(one, candidates=(16, 32), sweeps=1)
    assert [k for k, _ in table1] == [16, 32] and pick1 in (16, 32)
