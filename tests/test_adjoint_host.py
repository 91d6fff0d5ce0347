"""Host side of the adjoint (dqmc_amd.adjoint): the chain rule from the
engine's W = sum_l e^{V_l} (I - G_l) e^{-V_l} to d log det M / dK, d/dmu and
d/dlambda, checked on a lattice small enough for brute force.  Here W, tr G
and the lambda term are computed in numpy from explicit products (no GPU);
the gradients must match central finite differences of
log|det(I + B_{L-1}..B_0)| itself."""
import numpy as np
from scipy.linalg import expm

from dqmc_amd import adjoint, gen


def _brute(prob, K, mu, lam, field):
    N, L, dt = prob.N, prob.L, prob.dt
    E = expm(-dt * (K - mu * np.eye(N)))
    out = {"logdet": 0.0, "W": np.zeros((N, N)), "trG": 0.0, "sG": 0.0}
    for sigma in (1, -1):
        V = [np.exp(sigma * lam * (2 * field[l] - 1)) for l in range(L)]
        B = [E * V[l][None, :] for l in range(L)]  # E diag(V_l)
        P = np.eye(N)
        for l in range(L):
            P = B[l] @ P
        out["logdet"] += np.linalg.slogdet(np.eye(N) + P)[1]
        for l in range(L):
            # G_l: B_l at the right end: (I + B_{l-1}..B_0 B_{L-1}..B_l)^-1
            Q = np.eye(N)
            for k in list(range(l, L)) + list(range(0, l)):
                Q = B[k] @ Q
            G = np.linalg.inv(np.eye(N) + Q)
            out["W"] += (V[l][:, None] * (np.eye(N) - G)) / V[l][None, :]
            out["trG"] += np.trace(G)
            out["sG"] += np.sum(sigma * (2 * field[l] - 1) * (1 - np.diag(G)))
    return out, E


def test_host_gradients_match_finite_differences():
    rs = np.random.RandomState(4)
    prob = adjoint.AdjointProblem(2, 3, L=6, dt=0.1, fields=rs.randint(0, 2, (1, 6, 6)), n_matmul=2)
    K = prob.hopping(1.0, -0.2) + 0.05 * rs.randn(6, 6)  # a general (non-symmetric) K
    mu, lam = -0.3, 0.6
    field = prob.fields[0]
    b, E = _brute(prob, K, mu, lam, field)
    gK = prob.grad_K_from(K, mu, b["W"], E)
    gmu = prob.dt * (2 * prob.N * prob.L - b["trG"])
    glam = b["sG"]
    f = lambda K_, mu_, lam_: _brute(prob, K_, mu_, lam_, field)[0]["logdet"]
    h = 1e-6
    for (i, j) in [(0, 1), (2, 2), (5, 0), (3, 4)]:
        Kp, Km = K.copy(), K.copy()
        Kp[i, j] += h
        Km[i, j] -= h
        fd = (f(Kp, mu, lam) - f(Km, mu, lam)) / (2 * h)
        assert abs(fd - gK[i, j]) < 1e-7 * max(1.0, abs(fd)), (i, j, fd, gK[i, j])
    fd_mu = (f(K, mu + h, lam) - f(K, mu - h, lam)) / (2 * h)
    fd_lam = (f(K, mu, lam + h) - f(K, mu, lam - h)) / (2 * h)
    assert abs(fd_mu - gmu) < 1e-7 * abs(fd_mu)
    assert abs(fd_lam - glam) < 1e-7 * max(1.0, abs(fd_lam))
    # mu enters through K - mu I: the two routes agree (tr(G dM) identity)
    assert abs(gmu + np.trace(gK)) < 1e-9 * abs(gmu)


def test_lambda_of_U_derivative():
    for U in (2.0, 4.0, 8.0):
        lam, dl = adjoint.lambda_of_U(U, 0.1)
        h = 1e-6
        fd = (adjoint.lambda_of_U(U + h, 0.1)[0] - adjoint.lambda_of_U(U - h, 0.1)[0]) / (2 * h)
        assert abs(fd - dl) < 1e-8 * abs(dl)
        # cosh(lambda) = e^{U dt / 2} (gen_1band_unified_hub.py:925-927)
        assert abs(np.cosh(lam) - np.exp(0.5 * U * 0.1)) < 1e-12


def test_symmetric_exponential_and_frechet_match_scipy():
    """The eigendecomposition route (symmetric K) against scipy's expm and
    expm_frechet (Al-Mohy and Higham), on a lattice hopping matrix whose
    spectrum is degenerate (exact ties in the divided differences) and on a
    random symmetric one."""
    from scipy.linalg import expm_frechet
    rs = np.random.RandomState(7)
    K_lat = gen.hopping_square(4, 4, t=1.0, tp=-0.25)[0].real
    B = rs.randn(16, 16)
    for K in (K_lat, 0.5 * (B + B.T)):
        A = -0.1 * (K - 0.3 * np.eye(16))
        s = adjoint._SymExp(A)
        for c in (1.0, -1.0, 0.5, -0.5):
            assert np.abs(s.exp(c) - expm(c * A)).max() < 1e-13
        Z = rs.randn(16, 16)
        ref = expm_frechet(A, Z, compute_expm=False)
        assert np.abs(s.frechet(Z) - ref).max() < 1e-12 * np.abs(ref).max()


def test_symmetric_K_gradient_matches_finite_differences():
    """grad_K_from on a symmetric K (eigendecomposition route) against central
    differences of the brute-force log det."""
    rs = np.random.RandomState(5)
    prob = adjoint.AdjointProblem(2, 3, L=6, dt=0.1, fields=rs.randint(0, 2, (1, 6, 6)), n_matmul=2)
    K = prob.hopping(1.0, -0.2)
    mu, lam = 0.2, 0.5
    field = prob.fields[0]
    b, E = _brute(prob, K, mu, lam, field)
    gK = prob.grad_K_from(K, mu, b["W"], E, np.linalg.inv(E))
    f = lambda K_: _brute(prob, K_, mu, lam, field)[0]["logdet"]
    h = 1e-6
    for (i, j) in [(0, 1), (1, 0), (2, 2), (5, 3)]:
        Kp, Km = K.copy(), K.copy()
        Kp[i, j] += h
        Km[i, j] -= h
        fd = (f(Kp) - f(Km)) / (2 * h)
        assert abs(fd - gK[i, j]) < 1e-7 * max(1.0, abs(fd)), (i, j, fd, gK[i, j])
