"""The benchmark configurations on the GPU against the oracle, at their full
shapes, plus the complex-path pins (zero phases, gauge change).

Tolerances: HS fields, RNG states and acceptance counts bit-exact; equal-time
G 1e-10 relative max-norm; accumulators 1e-9 (equal time) and 1e-8
(unequal time) relative to the largest entry, except at N = 256, beta = 8,
U = 6, where the reference algorithm itself is less exact than that (see
test_config2_full_size_with_unequal_time) and both sides are refereed."""
import numpy as np
import pytest

from dqmc_amd import engine, gen
import oracle
import simtools

pytestmark = pytest.mark.gpu


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def _continue(sim, out):
    """The simulation dict after an oracle run, for the next chunk."""
    return dict(sim, state=out["state"], meas_eqlt=out["meas_eqlt"],
                meas_uneqlt=out["meas_uneqlt"] or sim["meas_uneqlt"])


def jackknife(s, sx):
    """Jackknife mean and error of sum(sx)/sum(s) over i.i.d. bins (first
    index), the estimator of util/util.py:31-46."""
    n = s.shape[0]
    S, SX = s.sum(0), sx.sum(0)
    full = (SX / S).real
    jk = ((SX - sx) / (S - s)[:, None]).real
    m = jk.mean(0)
    return n * full - (n - 1) * m, np.sqrt((n - 1) * ((jk - m) ** 2).mean(0))


def test_config1_exact_trajectory_and_binned_observables():
    """4x4, t'=0, U=4, mu=0, beta=4 (L=40), n_matmul=8, delay 8, seed 1234,
    100 warm-up + 400 measurement sweeps, equal-time measurements every 8
    slices.  The HS field and the RNG state are bit-exact after every chunk;
    density, double occupancy and zz, binned in 10 bins of 40 sweeps, agree
    within 2 sigma of their jackknife errors (and in fact to rounding)."""
    sim = gen.create_batch(seed=1234, Nx=4, Ny=4, tp=0.0, U=4.0, mu=0.0, dt=0.1, L=40,
                           n_matmul=8, n_delay=8, period_eqlt=8, n_sweep_warm=100,
                           n_sweep_meas=400)[0]
    eng = engine.Engine([sim])
    try:
        assert eng.delay_rank() == 8
        ref = sim
        chunks = [100] + [40] * 10
        snaps = {"gpu": [], "ref": []}
        for k in chunks:
            eng.run(k)
            out = oracle.run(ref, max_sweeps=k)
            ref = _continue(ref, out)
            st = eng.get_state(0)
            assert np.array_equal(st["hs"], out["state"]["hs"])
            assert np.array_equal(st["rng"], out["state"]["rng"])
            eq, _ = eng.get_meas(0)
            snaps["gpu"].append(eq)
            snaps["ref"].append(out["meas_eqlt"])
        assert eng.accepts() > 0
        eq, req = snaps["gpu"][-1], snaps["ref"][-1]
        assert eq["n_sample"] == req["n_sample"] == 400 * 5
        for k in ("density", "double_occ", "zz", "nn", "xx", "g00", "pair_sw"):
            assert rel(eq[k], req[k]) < 1e-9, k
        for k in ("density", "double_occ", "zz"):
            res = {}
            for who in ("gpu", "ref"):
                sn = snaps[who]
                s = np.array([float(np.real(sn[i + 1]["sign"]) - np.real(sn[i]["sign"]))
                              for i in range(10)])
                sx = np.array([np.asarray(sn[i + 1][k]) - np.asarray(sn[i][k]) for i in range(10)])
                res[who] = jackknife(s, sx)
            (mg, eg), (mr, er) = res["gpu"], res["ref"]
            assert np.all(eg > 0)
            assert np.all(np.abs(mg - mr) <= 2 * np.sqrt(eg ** 2 + er ** 2)), k
        # half filling on the bipartite lattice: density 1 per configuration
        np.testing.assert_allclose(np.asarray(eq["density"]) / np.real(eq["sign"]), 1.0,
                                   atol=1e-10)
    finally:
        eng.close()


def test_config2_full_size_with_unequal_time():
    """16x16 (N=256), t'=-0.25, U=6, mu=-1, beta=8 (L=80), n_matmul=10, delay
    32, seed 7, equal-time every 8 slices, unequal-time every 2 sweeps: two
    sweeps, the first with the unequal-time measurement.

    At this shape the reference algorithm is not exact to 1e-10: its wraps
    drift by ~1e-4 per sweep (CHECK_G_WRP) and its unequal-time expansion over
    blocks of L/E = 20 slices loses up to ~3e-4 (relative to the largest
    entry) in the accumulators.  So (a) the field and the random stream must
    be bit-exact; (b) the equal-time accumulators must agree within 1e-6;
    (c) the unequal-time accumulators of the first sweep are refereed by an
    independent stable evaluation of every block (simtools.ue_referee: UDV
    with a QR after every slice) at 16 slices, and the GPU must be no
    less accurate there than the reference algorithm (the oracle)."""
    sim = gen.create_1(init_rng=gen.splitmix64_seed(7), Nx=16, Ny=16, tp=-0.25, U=6.0, mu=-1.0,
                       dt=0.1, L=80, n_matmul=10, n_delay=32, period_eqlt=8, period_uneqlt=2,
                       n_sweep_warm=0, n_sweep_meas=10)
    eng = engine.Engine([sim])
    try:
        eng.run(1)
        out1 = oracle.run(sim, max_sweeps=1)
        st = eng.get_state(0)
        assert np.array_equal(st["hs"], out1["state"]["hs"])
        _, ue = eng.get_meas(0)
        assert ue["n_sample"] == out1["meas_uneqlt"]["n_sample"] == 1
        taus = [0, 1, 10, 11, 19, 20, 21, 30, 31, 40, 50, 51, 60, 70, 71, 79]
        s1 = dict(sim, state=dict(sim["state"], hs=st["hs"]))
        ref = simtools.ue_referee_accumulators(s1, out1["phase"], taus)
        nij = int(sim["params"]["num_ij"])
        e_gpu = max(simtools.slice_error(ue[k], ref[k], nij, taus) for k in ref if k != "n_sample"
                    and k != "sign")
        e_ora = max(simtools.slice_error(out1["meas_uneqlt"][k], ref[k], nij, taus) for k in ref
                    if k != "n_sample" and k != "sign")
        print("config 2 unequal-time accumulators against the stable referee: GPU %.2e, "
              "reference algorithm %.2e" % (e_gpu, e_ora), flush=True)
        assert e_gpu <= max(e_ora, 1e-8) and e_gpu < 1e-5
        eng.run(1)
        out = oracle.run(sim, max_sweeps=2)
        st = eng.get_state(0)
        assert np.array_equal(st["hs"], out["state"]["hs"])
        assert np.array_equal(st["rng"], out["state"]["rng"])
        eq, _ = eng.get_meas(0)
        assert eq["n_sample"] == out["meas_eqlt"]["n_sample"] == 20
        worst = max((rel(eq[k], v), k) for k, v in out["meas_eqlt"].items() if k != "n_sample")
        print("config 2 equal-time accumulators against the oracle: %.2e (%s)" % worst, flush=True)
        assert worst[0] < 1e-6
        s2 = dict(sim, state=dict(sim["state"], hs=st["hs"]))
        for spin in (0, 1):
            G, _ = eng.green(0, spin)
            Go, _ = oracle.calc_eq_g(s2, spin, 0)
            print("config 2 spin %d: |G_gpu - G_oracle| / |G| = %.2e" % (spin, rel(G, Go)),
                  flush=True)
            assert rel(G, Go) < 1e-8
    finally:
        eng.close()


def _config3(**kw):
    base = dict(Nx=12, Ny=12, nflux=1, tp=0.0, U=4.0, mu=-0.5, dt=0.1, L=60, n_matmul=10,
                n_delay=16, period_eqlt=10, period_uneqlt=1, n_sweep_warm=0, n_sweep_meas=2,
                meas_bond_corr=1, meas_thermal=1, meas_2bond_corr=1)
    base.update(kw)
    return gen.create_batch(seed=3, **base)[0]


def test_config3_complex_N144_calc_eq_g():
    """Complex 12x12 (N=144) is the largest complex order of the resident
    pivoted QR (k_qrp_res, 128 < N <= 144)."""
    sim = _config3()
    eng = engine.Engine([sim])
    try:
        for spin in (0, 1):
            for f in (0, 5):
                G, ph = eng.calc_eq_g(0, spin, f)
                Go, pho = oracle.calc_eq_g(sim, spin, f)
                assert rel(G, Go) < 1e-10, (spin, f)
                assert abs(ph - pho) < 1e-10
    finally:
        eng.close()


def test_config3_complex_N144_sweeps_with_all_bond_sets():
    """Config 3's shape with the bond (jj), thermal and 2-bond correlators
    measured at every sweep.  After one sweep the unequal-time accumulators
    of every group are refereed at 16 slices by the stable evaluation
    (simtools.ue_referee): the reference algorithm (the oracle) is within
    ~5e-9 of it here, and the GPU must be within 5x of the oracle's error
    (or 2e-9) and below 1e-7.  After two sweeps the field and the random
    stream are bit-exact, equal-time accumulators agree within 1e-9 and
    unequal-time ones within 1e-7 of the oracle's."""
    sim = _config3()
    eng = engine.Engine([sim])
    try:
        eng.run(1)
        out1 = oracle.run(sim, max_sweeps=1)
        st = eng.get_state(0)
        assert np.array_equal(st["hs"], out1["state"]["hs"])
        _, ue = eng.get_meas(0)
        taus = [0, 1, 10, 11, 12, 19, 20, 21, 30, 39, 40, 41, 50, 51, 52, 59]
        s1 = dict(sim, state=dict(sim["state"], hs=st["hs"]))
        ref = simtools.ue_referee_accumulators(s1, out1["phase"], taus)
        L = int(sim["params"]["L"])
        report = []
        for k, r in ref.items():
            if k in ("n_sample", "sign"):
                continue
            n = np.asarray(r).size // L
            eg = simtools.slice_error(ue[k], r, n, taus)
            eo = simtools.slice_error(out1["meas_uneqlt"][k], r, n, taus)
            report.append("%s %.1e/%.1e" % (k, eg, eo))
            assert eg <= max(5 * eo, 2e-9) and eg < 1e-7, (k, eg, eo)
        print("config 3 unequal-time accumulators against the stable referee (GPU/oracle): "
              + ", ".join(report), flush=True)
        eng.run(1)
        out = oracle.run(sim)
        st = eng.get_state(0)
        assert np.array_equal(st["hs"], out["state"]["hs"])
        assert np.array_equal(st["rng"], out["state"]["rng"])
        eq, ue = eng.get_meas(0)
        checked = 0
        for mine, ref, tol in ((eq, out["meas_eqlt"], 1e-9), (ue, out["meas_uneqlt"], 1e-7)):
            for k, v in ref.items():
                if k == "n_sample":
                    continue
                assert rel(mine[k], v) < tol, (k, rel(mine[k], v))
                checked += 1
        assert checked >= 30
    finally:
        eng.close()


def test_complex_engine_on_zero_phases_reproduces_the_real_golden(golden):
    """The complex kernels (complex Metropolis ratio and phase, complex
    delayed flush, complex GEMMs and QRs) on the golden file with complex
    dtype and zero imaginary parts: the reference's hs_end["square"]."""
    sim = simtools.as_complex(gen.create_batch(seed=1234, n_sweep_warm=20, n_sweep_meas=20)[0])
    eng = engine.Engine([sim])
    try:
        assert eng.cplx
        eng.run(40)
        assert np.array_equal(eng.get_state(0)["hs"], np.array(golden["hs_end"]["square"]))
    finally:
        eng.close()


def test_complex_engine_is_gauge_invariant():
    """alpha = 1/2 and alpha = 1 Peierls gauges at t' = 0 (related by a
    site-diagonal unitary): identical trajectories on the GPU, equal to the
    oracle's, and equal gauge-invariant accumulators."""
    kw = dict(Nx=4, Ny=6, nflux=1, tp=0.0, U=4.0, mu=-0.4, L=20, n_matmul=5, period_eqlt=5,
              period_uneqlt=2, n_sweep_warm=2, n_sweep_meas=6, meas_bond_corr=1,
              meas_energy_corr=1, meas_thermal=1, meas_2bond_corr=1)
    sa = gen.create_batch(seed=77, alpha=0.5, **kw)[0]
    sb = gen.create_batch(seed=77, alpha=1.0, **kw)[0]
    ea, eb = engine.Engine([sa]), engine.Engine([sb])
    try:
        ea.run(8)
        eb.run(8)
        ob = oracle.run(sb)
        ha, hb = ea.get_state(0)["hs"], eb.get_state(0)["hs"]
        assert np.array_equal(ha, hb) and np.array_equal(hb, ob["state"]["hs"])
        qa, ua = ea.get_meas(0)
        qb, ub = eb.get_meas(0)
        checked = 0
        for A, B in ((qa, qb), (ua, ub)):
            for k in B:
                if k == "n_sample" or k in simtools.GAUGE_VARIANT:
                    continue
                assert rel(A[k], B[k]) < 1e-9, k
                checked += 1
        assert checked > 20
        for k in ("jj", "kk", "j2j2", "jnjn", "zz"):
            assert rel(ub[k], ob["meas_uneqlt"][k]) < 1e-8, k
    finally:
        ea.close()
        eb.close()
