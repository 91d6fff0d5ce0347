"""Adjoint of log|det M| on the GPU (dqmc_engine_adjoint + dqmc_amd.adjoint).

(1) The engine's per-field pieces -- log|det M_s|, W = sum_l e^{V_l}(I - G_l)
    e^{-V_l}, sum_l tr G_l and the lambda term -- against explicit products on
    a 4x4 lattice (brute force, fine at beta = 2).
(2) The differentiation workload: 10x10 square lattice, t' = -0.2, U = 4,
    mu = -0.3, beta = 5 (dt = 0.1, L = 50), UDT checkpoints every 10 slices,
    64 fixed fields drawn as create_batch(seed=2024) draws them.  Gradients
    with respect to mu, lambda, t, t' and entries of the full K against
    central finite differences (h = 1e-6), and the tr(G dM) identity
    d/dmu = -tr(d/dK) = dt (2 N L - sum tr G).
(3) The torch.autograd.Functions.
(4) Complex (Peierls, nflux = 1) K on a 4x4 lattice: engine pieces against
    brute force, gradients against central differences in the real and the
    imaginary part of K entries, and the tr(G dM) identity.
(5) The adjoint of a G-based observable, O = sum_s tr(X G_s)
    (dqmc_engine_adjoint_green): engine pieces against dense unequal-time
    products, the density's mu derivative at lambda = 0 against the free
    closed form, and at the differentiation workload d<n>/d{mu, lambda, t, t',
    K} against central differences (h = 1e-6; 1e-4 for single K entries,
    where the density's rounding dominates a 1e-6 quotient)."""
import numpy as np
import pytest

from dqmc_amd import adjoint, gen

pytestmark = pytest.mark.gpu


def _fields(Nx, Ny, L, nf, seed):
    sims = gen.create_batch(Nfiles=nf, seed=seed, Nx=Nx, Ny=Ny, L=L, n_matmul=10, period_eqlt=10,
                            n_sweep_warm=0, n_sweep_meas=1)
    return np.array([s["state"]["hs"] for s in sims])


def test_engine_pieces_match_brute_force():
    from test_adjoint_host import _brute
    Nx, Ny, L = 4, 4, 20
    prob = adjoint.AdjointProblem(Nx, Ny, L, 0.1, _fields(Nx, Ny, L, 3, 11), n_matmul=10)
    K = prob.hopping(1.0, -0.2)
    mu, lam = -0.3, adjoint.lambda_of_U(4.0, 0.1)[0]
    f, d = prob.run(K, mu, lam)
    for k in range(3):
        b, _ = _brute(prob, K, mu, lam, prob.fields[k])
        assert abs(f[k] - b["logdet"]) < 1e-10 * abs(b["logdet"])
        W = d["W"][k].sum(axis=0)
        assert np.max(np.abs(W - b["W"])) < 1e-9 * np.max(np.abs(b["W"]))
        assert abs(d["trG"][k].sum() - b["trG"]) < 1e-9 * abs(b["trG"])
        assert abs(d["sG"][k].sum() - b["sG"]) < 1e-9 * max(1.0, abs(b["sG"]))


def test_set_params_matches_a_fresh_engine():
    # an evaluation after dqmc_engine_set_params equals one on an engine built
    # with those parameters, bit for bit
    Nx, Ny, L = 4, 4, 20
    fields = _fields(Nx, Ny, L, 3, 12)
    a = adjoint.AdjointProblem(Nx, Ny, L, 0.1, fields, n_matmul=10)
    b = adjoint.AdjointProblem(Nx, Ny, L, 0.1, fields, n_matmul=10)
    try:
        lam = adjoint.lambda_of_U(4.0, 0.1)[0]
        a.run(a.hopping(1.0, -0.2), -0.3, lam)  # builds a's engine
        f2, d2 = a.run(a.hopping(0.9, 0.1), 0.2, 1.1 * lam)  # replaces its parameters
        g2, e2 = b.run(b.hopping(0.9, 0.1), 0.2, 1.1 * lam)  # a fresh engine
        assert np.array_equal(f2, g2)
        for k in ("W", "trG", "sG", "logdet"):
            assert np.array_equal(d2[k], e2[k]), k
    finally:
        a.close()
        b.close()


@pytest.fixture(scope="module")
def cfg5():
    Nx = Ny = 10
    L = 50
    prob = adjoint.AdjointProblem(Nx, Ny, L, 0.1, _fields(Nx, Ny, L, 64, 2024), n_matmul=10)
    return prob, -0.3, adjoint.lambda_of_U(4.0, 0.1)[0]


def test_gradients_match_central_differences(cfg5):
    prob, mu, lam = cfg5
    t, tp = 1.0, -0.2
    K = prob.hopping(t, tp)
    f0, gK, gmu, glam, d = adjoint.logdet_and_grads(prob, K, mu, lam)
    F = lambda K_, mu_, lam_: prob.run(K_, mu_, lam_)[0].sum()
    h = 1e-6
    # mu and the tr(G dM) identity
    fd = (F(K, mu + h, lam) - F(K, mu - h, lam)) / (2 * h)
    assert abs(gmu.sum() - fd) < 1e-6 * abs(fd)
    assert abs(gmu.sum() + np.trace(gK)) < 1e-9 * abs(gmu.sum())
    # lambda
    fd = (F(K, mu, lam + h) - F(K, mu, lam - h)) / (2 * h)
    assert abs(glam.sum() - fd) < 1e-6 * max(abs(fd), 1.0)
    # t, t' (projections of the K gradient)
    for pat, v in ((prob.t_pattern, "t"), (prob.tp_pattern, "t'")):
        g = float(np.sum(gK * pat))
        fd = (F(K + h * pat, mu, lam) - F(K - h * pat, mu, lam)) / (2 * h)
        assert abs(g - fd) < 1e-6 * abs(fd), v
    # individual entries of the full K
    for (i, j) in [(0, 1), (5, 5), (17, 42)]:
        E = np.zeros_like(K)
        E[i, j] = 1.0
        fd = (F(K + h * E, mu, lam) - F(K - h * E, mu, lam)) / (2 * h)
        assert abs(gK[i, j] - fd) < 1e-6 * max(abs(fd), np.max(np.abs(gK))), (i, j, gK[i, j], fd)


def test_torch_autograd_function(cfg5):
    torch = pytest.importorskip("torch")
    prob, mu, lam = cfg5
    K = torch.tensor(prob.hopping(1.0, -0.2), dtype=torch.float64, requires_grad=True)
    m = torch.tensor(mu, dtype=torch.float64, requires_grad=True)
    la = torch.tensor(lam, dtype=torch.float64, requires_grad=True)
    w = torch.linspace(0.5, 1.5, 64, dtype=torch.float64)
    y = (adjoint.LogDet.apply(K, m, la, prob) * w).sum()
    y.backward()
    f0, gK, gmu, glam, d = adjoint.logdet_and_grads(prob, K.detach().numpy(), mu, lam)
    # weighted gradient from the per-field pieces
    Ww = np.tensordot(w.numpy(), d["W"].sum(axis=1), axes=(0, 0))
    gKw = prob.grad_K_from(K.detach().numpy(), mu, Ww, d["E"])
    assert np.max(np.abs(K.grad.numpy() - gKw)) < 1e-12 * np.max(np.abs(gKw))
    assert abs(float(m.grad) - float(np.dot(w.numpy(), gmu))) < 1e-9 * abs(float(m.grad))
    assert abs(float(la.grad) - float(np.dot(w.numpy(), glam))) < 1e-9 * max(1.0, abs(float(la.grad)))


def _fd_entries(F, K, entries, h=1e-6):
    from test_adjoint_host import _fd_entries as fd
    return fd(F, K, entries, h)


def test_complex_engine_pieces_and_gradients():
    from test_adjoint_host import _brute
    Nx, Ny, L = 4, 4, 20
    prob = adjoint.AdjointProblem(Nx, Ny, L, 0.1, _fields(Nx, Ny, L, 3, 13), n_matmul=10, nflux=1)
    try:
        K = prob.hopping(1.0, -0.2)
        assert np.iscomplexobj(K)
        mu, lam = -0.3, adjoint.lambda_of_U(4.0, 0.1)[0]
        f, gK, gmu, glam, d = adjoint.logdet_and_grads(prob, K, mu, lam)
        for k in range(3):
            b, _ = _brute(prob, K, mu, lam, prob.fields[k])
            assert abs(f[k] - b["logdet"]) < 1e-10 * abs(b["logdet"])
            W = d["W"][k].sum(axis=0)
            assert np.max(np.abs(W - b["W"])) < 1e-9 * np.max(np.abs(b["W"]))
            assert abs(d["trG"][k].sum() - b["trG"]) < 1e-9 * abs(b["trG"])
        F = lambda K_, mu_=mu, lam_=lam: prob.run(K_, mu_, lam_)[0].sum()
        h = 1e-6
        fd = (F(K, mu + h) - F(K, mu - h)) / (2 * h)
        assert abs(gmu.sum() - fd) < 1e-6 * abs(fd)
        assert abs(gmu.sum() + np.trace(gK).real) < 1e-9 * abs(gmu.sum())
        fd = (F(K, mu, lam + h) - F(K, mu, lam - h)) / (2 * h)
        assert abs(glam.sum() - fd) < 1e-6 * max(abs(fd), 1.0)
        for pat, v in ((prob.t_pattern, "t"), (prob.tp_pattern, "t'")):
            g = np.real(np.sum(np.conj(gK) * pat))
            fd = (F(K + h * pat) - F(K - h * pat)) / (2 * h)
            assert abs(g - fd) < 1e-6 * abs(fd), v
        ent = [(0, 1), (1, 0), (5, 5), (3, 12)]
        scale = np.max(np.abs(gK))
        for (i, j), fd in zip(ent, _fd_entries(F, K, ent)):
            assert abs(gK[i, j] - fd) < 1e-6 * max(abs(fd), scale), (i, j, gK[i, j], fd)
    finally:
        prob.close()


@pytest.mark.parametrize("nflux", [0, 1])
def test_green_adjoint_pieces_match_dense_products(nflux):
    from test_adjoint_host import _brute_green
    Nx, Ny, L = 4, 4, 20
    prob = adjoint.AdjointProblem(Nx, Ny, L, 0.1, _fields(Nx, Ny, L, 3, 14), n_matmul=5, nflux=nflux)
    try:
        rs = np.random.RandomState(3)
        K = prob.hopping(1.0, -0.2)
        mu, lam = -0.3, adjoint.lambda_of_U(4.0, 0.1)[0]
        X = rs.randn(16, 16) + (1j * rs.randn(16, 16) if nflux else 0)
        O, d = prob.run_green(K, mu, lam, X)
        # O uses the equal-time G (1e-10); the Z_l sums inherit the accuracy of
        # calc_ue_g's blocks, which the engine holds to 1e-8 against the
        # oracle (measured here: 7.5e-9 relative on W)
        for k in range(3):
            b, _ = _brute_green(prob, K, mu, lam, prob.fields[k], X)
            assert abs(O[k] - b["obs"]) < 1e-10 * max(1.0, abs(b["obs"]))
            W = d["W"][k].sum(axis=0)
            assert np.max(np.abs(W - b["W"])) < 5e-8 * np.max(np.abs(b["W"]))
            # trZ and sZ are sums with cancellation: bar relative to sum |Z_ii|
            assert abs(d["trZ"][k].sum() - b["trZ"]) < 5e-8 * b["dscale"]
            assert abs(d["sZ"][k].sum() - b["sZ"]) < 5e-8 * b["dscale"]
    finally:
        prob.close()


def test_free_compressibility_on_the_engine():
    """lambda = 0: dn/dmu = (2/N) beta sum_k f_k (1 - f_k) at the
    differentiation workload's lattice (10x10, t' = -0.2, beta = 5)."""
    Nx = Ny = 10
    L = 50
    prob = adjoint.AdjointProblem(Nx, Ny, L, 0.1, _fields(Nx, Ny, L, 2, 15), n_matmul=10)
    try:
        K = prob.hopping(1.0, -0.2)
        mu = -0.3
        n, gK, gmu, glam, d = adjoint.green_and_grads(prob, K, mu, 0.0, adjoint.density_X(prob.N))
        beta = L * 0.1
        f = 1.0 / (1.0 + np.exp(beta * (np.linalg.eigvalsh(K) - mu)))
        assert np.all(np.abs(2 + n - 2 * f.mean()) < 1e-10)
        want = 2.0 / prob.N * beta * np.sum(f * (1 - f))
        assert np.all(np.abs(gmu - want) < 1e-9 * want), (gmu, want)
        # the K route: d/dmu = -tr(d/dK)
        assert abs(gmu.sum() + np.trace(gK)) < 1e-9 * abs(gmu.sum())
    finally:
        prob.close()


def test_density_gradients_match_central_differences(cfg5):
    prob, mu, lam = cfg5
    K = prob.hopping(1.0, -0.2)
    X = adjoint.density_X(prob.N)
    n, gK, gmu, glam, d = adjoint.green_and_grads(prob, K, mu, lam, X)
    assert np.all((n + 2 > 0.5) & (n + 2 < 1.5))
    F = lambda K_, mu_=mu, lam_=lam: prob.run_green(K_, mu_, lam_, X)[0].sum()
    h = 1e-6
    fd = (F(K, mu + h) - F(K, mu - h)) / (2 * h)
    assert abs(gmu.sum() - fd) < 1e-6 * abs(fd), (gmu.sum(), fd)
    assert abs(gmu.sum() + np.trace(gK)) < 1e-9 * abs(gmu.sum())
    fd = (F(K, mu, lam + h) - F(K, mu, lam - h)) / (2 * h)
    assert abs(glam.sum() - fd) < 1e-6 * abs(fd), (glam.sum(), fd)
    for pat, v in ((prob.t_pattern, "t"), (prob.tp_pattern, "t'")):
        g = float(np.sum(gK * pat))
        fd = (F(K + h * pat) - F(K - h * pat)) / (2 * h)
        assert abs(g - fd) < 1e-6 * abs(fd), (v, g, fd)
    # single entries of K: the density's derivative there is ~0.04, and at
    # h = 1e-6 the rounding of the 64 densities leaves ~3e-7 of noise in the
    # difference quotient (measured: h = 1e-4, 1e-5, 1e-6 give -0.0439084775,
    # -0.0439084722, -0.0439082584 against the adjoint's -0.0439084787 for
    # K[17, 42]); h = 1e-4 keeps the truncation error at ~1e-9
    scale = np.max(np.abs(gK))
    ent = [(0, 1), (5, 5), (17, 42)]
    for (i, j), fd in zip(ent, _fd_entries(F, K, ent, h=1e-4)):
        assert abs(gK[i, j] - fd) < 1e-6 * max(abs(fd), scale), (i, j, gK[i, j], fd)


def test_torch_green_trace(cfg5):
    torch = pytest.importorskip("torch")
    prob, mu, lam = cfg5
    K = torch.tensor(prob.hopping(1.0, -0.2), dtype=torch.float64, requires_grad=True)
    m = torch.tensor(mu, dtype=torch.float64, requires_grad=True)
    la = torch.tensor(lam, dtype=torch.float64, requires_grad=True)
    X = adjoint.density_X(prob.N)
    w = torch.linspace(0.5, 1.5, 64, dtype=torch.float64)
    (adjoint.GreenTrace.apply(K, m, la, X, prob) * w).sum().backward()
    O, d = prob.run_green(K.detach().numpy(), mu, lam, X)
    Ww = np.tensordot(w.numpy(), d["W"].sum(axis=1), axes=(0, 0))
    gKw = prob.grad_K_from(K.detach().numpy(), mu, Ww, d["E"])
    assert np.max(np.abs(K.grad.numpy() - gKw)) < 1e-12 * np.max(np.abs(gKw))
    gmu = prob.dt * d["trZ"].sum(axis=1).real
    assert abs(float(m.grad) - float(np.dot(w.numpy(), gmu))) < 1e-9 * abs(float(m.grad))
    glam = d["sZ"].sum(axis=1).real
    assert abs(float(la.grad) - float(np.dot(w.numpy(), glam))) < 1e-9 * abs(float(la.grad))
