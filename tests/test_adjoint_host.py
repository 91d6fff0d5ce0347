"""Host side of the adjoint (dqmc_amd.adjoint): the chain rule from the
engine's W = sum_l e^{V_l} (I - G_l) e^{-V_l} to d log det M / dK, d/dmu and
d/dlambda, checked on a lattice small enough for brute force.  Here W, tr G
and the lambda term are computed in numpy from explicit products (no GPU);
the gradients must match central finite differences of
log|det(I + B_{L-1}..B_0)| itself."""
import numpy as np
from scipy.linalg import expm

from dqmc_amd import adjoint, gen


def _bmats(prob, K, mu, lam, field, sigma):
    N, L, dt = prob.N, prob.L, prob.dt
    E = expm(-dt * (K - mu * np.eye(N)))
    V = [np.exp(sigma * lam * (2 * field[l] - 1)) for l in range(L)]
    return E, V, [E * V[l][None, :] for l in range(L)]  # B_l = E diag(V_l)


def _brute(prob, K, mu, lam, field):
    N, L = prob.N, prob.L
    dt_ = np.result_type(K, np.float64)
    out = {"logdet": 0.0, "W": np.zeros((N, N), dtype=dt_), "trG": 0.0, "sG": 0.0}
    for sigma in (1, -1):
        E, V, B = _bmats(prob, K, mu, lam, field, sigma)
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


def _brute_green(prob, K, mu, lam, field, X):
    """O = sum_s tr(X G_s) from dense products, and the adjoint pieces from
    dense unequal-time Green's functions G(t_l,0) = R_l G, G(0,t_l) =
    -(I - G) R_l^-1, R_l = B_{l-1}..B_0."""
    N, L = prob.N, prob.L
    dt_ = np.result_type(K, np.float64)
    # dscale: sum of |(Z_l)_ii|, the condition scale of trZ and sZ
    out = {"obs": 0.0, "W": np.zeros((N, N), dtype=dt_), "trZ": 0.0, "sZ": 0.0, "dscale": 0.0}
    I = np.eye(N)
    for sigma in (1, -1):
        E, V, B = _bmats(prob, K, mu, lam, field, sigma)
        R = [I]
        for l in range(L):
            R.append(B[l] @ R[-1])
        G = np.linalg.inv(I + R[L])
        out["obs"] += np.trace(X @ G)
        for l in range(L):
            Z = (R[l] @ G) @ X @ (-(I - G) @ np.linalg.inv(R[l]))
            out["W"] += (V[l][:, None] * Z) / V[l][None, :]
            out["trZ"] += np.trace(Z)
            out["sZ"] += np.sum(sigma * (2 * field[l] - 1) * np.diag(Z))
            out["dscale"] += np.sum(np.abs(np.diag(Z)))
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


def _fd_entries(f, K, entries, h=1e-6):
    """Central differences of f in the real and (complex K) imaginary part of
    single entries: d/dRe K_ij + i d/dIm K_ij."""
    out = []
    for (i, j) in entries:
        dK = np.zeros_like(K)
        dK[i, j] = h
        g = (f(K + dK) - f(K - dK)) / (2 * h)
        if np.iscomplexobj(K):
            g = g + 1j * (f(K + 1j * dK) - f(K - 1j * dK)) / (2 * h)
        out.append(g)
    return out


def test_hermitian_exponential_and_frechet_match_scipy():
    """The eigendecomposition route on the Peierls (complex Hermitian) hopping
    matrix and a random Hermitian one against scipy."""
    from scipy.linalg import expm_frechet
    rs = np.random.RandomState(8)
    K_lat = gen.hopping_square(4, 4, t=1.0, tp=-0.25, nflux=1)[0]
    B = rs.randn(16, 16) + 1j * rs.randn(16, 16)
    for K in (K_lat, 0.5 * (B + B.conj().T)):
        assert np.iscomplexobj(K) and np.allclose(K, K.conj().T)
        A = -0.1 * (K - 0.3 * np.eye(16))
        s = adjoint._SymExp(A)
        for c in (1.0, -1.0, 0.5):
            assert np.abs(s.exp(c) - expm(c * A)).max() < 1e-13
        Z = rs.randn(16, 16) + 1j * rs.randn(16, 16)
        ref = expm_frechet(A, Z, compute_expm=False)
        assert np.abs(s.frechet(Z) - ref).max() < 1e-12 * np.abs(ref).max()


def test_complex_K_gradient_matches_finite_differences():
    """Peierls-phase K (nflux = 1): the gradient d/dRe K + i d/dIm K of
    log|det M| from the chain rule against central differences in both parts,
    on the Hermitian K (eigendecomposition route) and on a non-Hermitian one
    (scipy's Frechet derivative), and the tr(G dM) identity
    d/dmu = -Re tr(dK gradient) = dt (2 N L - Re sum tr G)."""
    rs = np.random.RandomState(9)
    prob = adjoint.AdjointProblem(3, 3, L=6, dt=0.1, fields=rs.randint(0, 2, (1, 6, 9)), n_matmul=2,
                                  nflux=1)
    field = prob.fields[0]
    mu, lam = -0.2, 0.55
    K0 = prob.hopping(1.0, -0.2)
    assert np.iscomplexobj(K0) and np.abs(K0.imag).max() > 0.1
    Kn = K0 + 0.05 * (rs.randn(9, 9) + 1j * rs.randn(9, 9))
    for K in (K0, Kn):
        b, E = _brute(prob, K, mu, lam, field)
        gK = prob.grad_K_from(K, mu, b["W"], E)
        f = lambda K_: _brute(prob, K_, mu, lam, field)[0]["logdet"]
        ent = [(0, 1), (1, 0), (2, 2), (5, 3), (8, 1)]
        for (i, j), fd in zip(ent, _fd_entries(f, K, ent)):
            assert abs(fd - gK[i, j]) < 1e-7 * max(1.0, abs(fd)), (i, j, fd, gK[i, j])
        gmu = prob.dt * (2 * prob.N * prob.L - np.real(b["trG"]))
        fd_mu = (_brute(prob, K, mu + 1e-6, lam, field)[0]["logdet"]
                 - _brute(prob, K, mu - 1e-6, lam, field)[0]["logdet"]) / 2e-6
        assert abs(fd_mu - gmu) < 1e-7 * abs(fd_mu)
        assert abs(gmu + np.trace(gK).real) < 1e-9 * abs(gmu)
        # the t direction carries the phases: d/dt = Re sum conj(gK) o pattern
        pat = prob.t_pattern
        g_t = np.real(np.sum(np.conj(gK) * pat))
        fd_t = (f(K + 1e-6 * pat) - f(K - 1e-6 * pat)) / 2e-6
        assert abs(g_t - fd_t) < 1e-7 * abs(fd_t)


def test_tr_G_dM_identity_complex():
    """d log|det M| = Re tr(M^-1 dM) for every direction the chain rule
    covers: the chain-rule gradient contracted with a random complex dK
    equals Re tr(G dM) with dM formed from the exact derivative of the
    product (Frechet derivatives of exp per slice)."""
    from scipy.linalg import expm_frechet
    rs = np.random.RandomState(10)
    prob = adjoint.AdjointProblem(3, 3, L=5, dt=0.1, fields=rs.randint(0, 2, (1, 5, 9)), n_matmul=1,
                                  nflux=1)
    field = prob.fields[0]
    mu, lam = 0.1, 0.4
    K = prob.hopping(1.0, -0.2)
    n = prob.N
    dK = rs.randn(n, n) + 1j * rs.randn(n, n)
    b, E = _brute(prob, K, mu, lam, field)
    gK = prob.grad_K_from(K, mu, b["W"], E)
    lhs = np.real(np.sum(np.conj(gK) * dK))
    A = -prob.dt * (K - mu * np.eye(n))
    dE = expm_frechet(A, -prob.dt * dK, compute_expm=False)
    rhs = 0.0
    for sigma in (1, -1):
        _, V, B = _bmats(prob, K, mu, lam, field, sigma)
        dB = [dE * V[l][None, :] for l in range(prob.L)]
        P = np.eye(n)
        dP = np.zeros((n, n), dtype=complex)
        for l in range(prob.L):
            dP = B[l] @ dP + dB[l] @ P
            P = B[l] @ P
        rhs += np.real(np.trace(np.linalg.inv(np.eye(n) + P) @ dP))
    assert abs(lhs - rhs) < 1e-10 * abs(rhs)


def test_green_observable_gradients_match_finite_differences():
    """O = sum_s tr(X G_s) (random X; real and Peierls K): the unequal-time
    pieces Z_l with the log-determinant's chain rule against central
    differences of O itself in K entries, mu and lambda."""
    for nflux in (0, 1):
        rs = np.random.RandomState(11 + nflux)
        prob = adjoint.AdjointProblem(3, 3, L=6, dt=0.1, fields=rs.randint(0, 2, (1, 6, 9)),
                                      n_matmul=2, nflux=nflux)
        field = prob.fields[0]
        mu, lam = -0.25, 0.5
        K = prob.hopping(1.0, -0.2)
        X = rs.randn(9, 9) + (1j * rs.randn(9, 9) if nflux else 0)
        b, E = _brute_green(prob, K, mu, lam, field, X)
        O = lambda K_, mu_=mu, lam_=lam: np.real(_brute_green(prob, K_, mu_, lam_, field, X)[0]["obs"])
        gK = prob.grad_K_from(K, mu, b["W"], E)
        ent = [(0, 1), (3, 3), (8, 2)]
        for (i, j), fd in zip(ent, _fd_entries(O, K, ent)):
            assert abs(fd - gK[i, j]) < 1e-7 * max(1.0, abs(fd)), (nflux, i, j, fd, gK[i, j])
        h = 1e-6
        fd_mu = (O(K, mu + h) - O(K, mu - h)) / (2 * h)
        assert abs(fd_mu - prob.dt * np.real(b["trZ"])) < 1e-7 * max(1.0, abs(fd_mu))
        fd_lam = (O(K, mu, lam + h) - O(K, mu, lam - h)) / (2 * h)
        assert abs(fd_lam - np.real(b["sZ"])) < 1e-7 * max(1.0, abs(fd_lam))


def test_free_compressibility_closed_form():
    """At lambda = 0 every B_l is E and G = (I + e^{-beta K_s})^-1 is the
    Fermi function of K - mu: the density's mu derivative from the Z_l pieces
    equals (2/N) beta sum_k f_k (1 - f_k) over the eigenvalues of K."""
    prob = adjoint.AdjointProblem(3, 3, L=20, dt=0.1, fields=np.zeros((1, 20, 9), dtype=np.int32),
                                  n_matmul=4)
    K = prob.hopping(1.0, -0.2)
    mu = 0.3
    X = adjoint.density_X(prob.N)
    b, _ = _brute_green(prob, K, mu, 0.0, prob.fields[0], X)
    beta = prob.L * prob.dt
    f = 1.0 / (1.0 + np.exp(beta * (np.linalg.eigvalsh(K) - mu)))
    assert abs(2 + b["obs"] - 2 * f.mean()) < 1e-12
    want = 2.0 / prob.N * beta * np.sum(f * (1 - f))
    assert abs(prob.dt * b["trZ"] - want) < 1e-12 * want
