"""The oracle against what the mathematics fixes (closed forms at U=0,
brute-force products at small beta, invariants).  These pin calc_eq_g,
calc_ue_g and the measurement formulas where the reference ships no data."""
import numpy as np
import pytest
from scipy.linalg import expm

from dqmc_amd import gen
import oracle


def _sim(**kw):
    base = dict(Nx=4, Ny=4, U=4.0, mu=-0.3, tp=-0.1, L=20, n_matmul=5, period_eqlt=5,
                n_sweep_warm=1, n_sweep_meas=3)
    base.update(kw)
    return gen.create_batch(seed=3, **base)[0]


def _brute_G(sim, spin, f=0):
    """(I + B_{l-1} ... B_l)^-1 at slice boundary f*n_matmul, products in
    plain numpy (the C engine reads numpy arrays transposed, so use A.T)."""
    p = sim["params"]
    N, L, n = int(p["N"]), int(p["L"]), int(p["n_matmul"])
    expK = (p["exp_Kd"] if spin else p["exp_Ku"]).T
    el = p["exp_lambda"]
    hs = sim["state"]["hs"]
    l0 = f * n
    P = np.eye(N, dtype=expK.dtype)
    for q in range(L):
        l = (l0 + q) % L
        s = (1 - hs[l]) if spin else hs[l]
        B = expK * el[s, np.arange(N)][None, :]
        P = B @ P
    return np.linalg.inv(np.eye(N) + P), P


@pytest.mark.parametrize("nflux", [0, 1])
def test_calc_eq_g_free_closed_form(nflux):
    sim = _sim(U=0.0, nflux=nflux)
    p = sim["params"]
    L, dt = 20, 0.1
    for spin in (0, 1):
        K = (p["Kd"] if spin else p["Ku"]).T  # transposed reading of the raw buffer
        G_exact = np.linalg.inv(np.eye(16) + expm(-L * dt * K))
        G, ph = oracle.calc_eq_g(sim, spin, 0)
        assert np.abs(G - G_exact).max() < 1e-12
        assert abs(ph - 1.0) < 1e-12  # det(I + e^{-beta K}) > 0


@pytest.mark.parametrize("nflux", [0, 1])
@pytest.mark.parametrize("f", [0, 2])
def test_calc_eq_g_brute_force_and_phase(nflux, f):
    sim = _sim(nflux=nflux, L=20, dt=0.05)
    for spin in (0, 1):
        G_b, P = _brute_G(sim, spin, f)
        G, ph = oracle.calc_eq_g(sim, spin, f)
        assert np.abs(G - G_b).max() / np.abs(G_b).max() < 1e-9
        d = np.linalg.det(np.eye(16) + P)
        assert abs(ph - d / abs(d)) < 1e-9


def test_calc_ue_g_free_closed_form():
    sim = _sim(U=0.0, period_uneqlt=1, L=20, n_matmul=5)
    p = sim["params"]
    dt, L = 0.1, 20
    K = p["Ku"].T
    G0 = np.linalg.inv(np.eye(16) + expm(-L * dt * K))
    G0t, Gtt, Gt0 = oracle.calc_ue_g(sim, 0)
    for t in range(L):
        et = expm(-t * dt * K)
        # G(t,0) = e^{-tK} G(0), G(t,t) = G(0), and G(0,t) = -(I - G(0)) e^{tK}
        # for t > 0 (at t = 0 the block is the equal-time G itself)
        assert np.abs(Gt0[t] - et @ G0).max() < 1e-10
        assert np.abs(Gtt[t] - G0).max() < 1e-10
        want = G0 if t == 0 else -(np.eye(16) - G0) @ np.linalg.inv(et)
        assert np.abs(G0t[t] - want).max() < 1e-10


def test_calc_ue_g_tau0_equals_equal_time_G():
    # CHECK_G_UE of src/dqmc.c:405-408
    sim = _sim(period_uneqlt=1, L=40, n_matmul=8)
    for spin in (0, 1):
        G, _ = oracle.calc_eq_g(sim, spin, 0)
        G0t, Gtt, Gt0 = oracle.calc_ue_g(sim, spin)
        assert np.abs(Gtt[0] - G).max() < 1e-9
        assert np.abs(Gt0[0] - G).max() < 1e-9


def test_measure_eqlt_free_wick_identities():
    # At U=0 the up and down G are equal: zz = xx for every distance, double
    # occupancy = n_up * n_dn, pair_sw(0) = sum_i G_ii^2 / N.
    sim = _sim(U=0.0)
    G, _ = oracle.calc_eq_g(sim, 0, 0)
    m = oracle.measure_eqlt(sim, 1.0, G, G)
    assert np.abs(m["zz"] - m["xx"]).max() < 1e-12
    n = 1 - np.diag(G)
    assert abs(m["double_occ"][0] - np.mean(n * n)) < 1e-12
    assert abs(m["density"][0] - np.mean(2 * n)) < 1e-12
    assert abs(m["pair_sw"][0] - np.mean(np.diag(G) ** 2)) < 1e-12
    # g00(r): translation average of G(i+r, i)
    r = 1
    want = np.mean([G[(i % 4 + r) % 4 + 4 * (i // 4), i] for i in range(16)])
    assert abs(m["g00"][r] - want) < 1e-12


def test_measure_eqlt_particle_hole_half_filling():
    # mu=0, t'=0 on the bipartite square lattice: G_dn = I - D G_up^T D per
    # configuration, so density is exactly 1 and double_occ(up) = double_occ(dn)
    sim = _sim(mu=0.0, tp=0.0, U=4.0)
    gu, ph_u = oracle.calc_eq_g(sim, 0, 0)
    gd, ph_d = oracle.calc_eq_g(sim, 1, 0)
    D = np.diag([(-1.0) ** ((i % 4) + (i // 4)) for i in range(16)])
    assert np.abs(gd - (np.eye(16) - D @ gu.T @ D)).max() < 1e-9
    m = oracle.measure_eqlt(sim, 1.0, gu, gd)
    assert abs(m["density"][0] - 1.0) < 1e-12


@pytest.mark.parametrize("U", [0.0, 4.0])
def test_stable_unequal_time_referee(U):
    """The referee used at full size (tests/simtools.ue_referee: UDV forms with
    a QR after every slice, sums inverted with the scale splitting) agrees
    with calc_ue_g where that is exact (small N, short blocks), and at U = 0
    with the closed form G(t,0) = e^{-tK} (I + e^{-beta K})^-1."""
    import simtools
    sim = _sim(U=U, period_uneqlt=1, L=40, n_matmul=8)
    taus = [0, 1, 7, 8, 9, 20, 39]
    for spin in (0, 1):
        G0t, Gtt, Gt0 = oracle.calc_ue_g(sim, spin)
        gt0, g0t, gtt = simtools.ue_referee(sim, spin, taus)
        for t in taus:
            assert np.abs(gt0[t] - Gt0[t]).max() < 1e-10 * np.abs(Gt0[t]).max()
            assert np.abs(gtt[t] - Gtt[t]).max() < 1e-10 * np.abs(Gtt[t]).max()
            if t:
                assert np.abs(g0t[t] - G0t[t]).max() < 1e-10 * np.abs(G0t[t]).max()
        if U == 0.0:
            K = (sim["params"]["Kd"] if spin else sim["params"]["Ku"]).T
            G0 = np.linalg.inv(np.eye(16) + expm(-4.0 * K))
            for t in taus:
                want = expm(-0.1 * t * K) @ G0
                assert np.abs(gt0[t] - want).max() < 1e-10 * np.abs(want).max()
