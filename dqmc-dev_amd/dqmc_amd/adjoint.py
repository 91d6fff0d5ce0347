"""Differentiable log|det M| and G-based observables of the Hubbard DQMC
weight for fixed HS fields.

No counterpart in the reference (it computes only the phase of det M,
src/greens.c:180-191).  For a batch of fixed fields s_{l,i} = +-1 the weight
of field f is det M_up det M_dn with
    M_s = I + B_{L-1} ... B_0,   B_l = exp(-dt K_s) diag(exp(sigma_s lambda s_l)),
    K_s = K - mu I,  sigma_up = +1, sigma_dn = -1,
K the hopping matrix -- the generator's kij, real, or Hermitian with Peierls
phases when the lattice is threaded by flux -- and lambda the HS coupling
(cosh-form e^lambda = e^{U dt/2} + sqrt(e^{U dt} - 1), gen_1band_unified_hub.py
:925-927).  The GPU engine returns per field and spin log|det M_s| and the
adjoint sums W_s, sum_l tr G_l and the lambda term (dqmc_engine_adjoint); the
gradient with respect to K follows from d log|det M| = Re tr(W E^-1 dE) with
E = expm(A), A = -dt K_s, and the adjoint of the exponential's Frechet
derivative, tr(Z L(A, X)) = tr(L(A, Z) X):
    d log|det| / dK = -dt conj(L(A, W E^-1))^T
(for complex K the gradient is d/dRe K + i d/dIm K, torch's convention; for
real K it is -dt L(A, W E^-1)^T).  Gradients with respect to t, t' (d K / dt =
the nearest-neighbour hopping pattern with its phases), mu (= -Re tr of the K
gradient, also dt (2 N L - Re sum tr G)) and U (through lambda(U)) are
projections of these.

A G-based observable O = sum_s tr(X_s G_s) (G at slice 0; X = -I/N for both
spins gives the density minus 2) has the same chain rule with W replaced by
the engine's W_X = sum_l e^{V_l} Z_l e^{-V_l}, Z_l = G(t_l,0) X G(0,t_l)
(dqmc_engine_adjoint_green): d O / d mu = dt Re sum_l tr Z_l, d O / d lambda =
Re sum_l sum_i sigma s_li (Z_l)_ii.

`LogDet` and `GreenTrace` are the torch.autograd.Functions; `AdjointProblem`
holds the fields and the lattice.  The engine's matrices are column-major views
of the raw buffers (a numpy array A is seen as A^T), so exponentials are
passed transposed to make the engine use exp(-dt K) itself.
"""
import ctypes as C

import numpy as np
from scipy.linalg import expm, expm_frechet

from . import engine as _eng
from . import gen


class _SymExp:
    """exp(c A) and the Frechet derivative of exp at A for a Hermitian (real
    symmetric) A, from one eigendecomposition A = V diag(a) V^H:
    exp(c A) = V diag(e^{c a}) V^H, and (Daleckii-Krein) L(A, Z) =
    V (Phi o V^H Z V) V^H with the divided differences
    Phi_ij = (e^{a_i} - e^{a_j}) / (a_i - a_j), e^{a_i} on ties."""

    def __init__(self, A):
        self.A = A
        self.a, self.V = np.linalg.eigh(A)
        self.Vh = self.V.conj().T

    def exp(self, c=1.0):
        return (self.V * np.exp(c * self.a)) @ self.Vh

    def frechet(self, Z):
        a, V, Vh = self.a, self.V, self.Vh
        d = a[:, None] - a[None, :]
        small = np.abs(d) < 1e-8
        dd = np.where(small, 1.0, d)
        # e^{a_j} expm1(d) / d, its Taylor form for near ties
        phi = np.exp(a)[None, :] * np.where(small, 1.0 + 0.5 * d + d * d / 6.0, np.expm1(dd) / dd)
        return V @ (phi * (Vh @ Z @ V)) @ Vh


def lambda_of_U(U, dt):
    """HS coupling of the reference's cosh form and d lambda / d U."""
    a = np.exp(0.5 * U * dt) + np.sqrt(np.expm1(U * dt))
    lam = np.log(a)
    da = 0.5 * dt * np.exp(0.5 * U * dt) + 0.5 * dt * np.exp(U * dt) / np.sqrt(np.expm1(U * dt))
    return lam, da / a


class AdjointProblem:
    """A batch of fixed HS fields on one lattice (square, periodic; complex
    hopping with nflux flux quanta through the torus).

    fields: int32 array [nf][L][N] of 0/1 (the file's /state/hs layout)."""

    def __init__(self, Nx, Ny, L, dt, fields, n_matmul=10, device=0, nflux=0):
        self.Nx, self.Ny, self.N, self.L, self.dt = Nx, Ny, Nx * Ny, L, dt
        self.n_matmul = n_matmul
        self.nflux = nflux
        self.cplx = nflux != 0
        self.fields = np.asarray(fields, dtype=np.int32).reshape(-1, L, Nx * Ny)
        self.device = device
        # template simulation (maps etc.); its physical content is replaced below
        self._tmpl = gen.create_1(init_rng=gen.splitmix64_seed(1), Nx=Nx, Ny=Ny, L=L, dt=dt,
                                  n_matmul=n_matmul, period_eqlt=n_matmul, n_sweep_warm=0,
                                  n_sweep_meas=1, nflux=nflux)
        real = (lambda a: a) if self.cplx else (lambda a: a.real)
        self.t_pattern = real(gen.hopping_square(Nx, Ny, t=1.0, tp=0.0, nflux=nflux)[0])
        self.tp_pattern = real(gen.hopping_square(Nx, Ny, t=0.0, tp=1.0, nflux=nflux)[0])
        self._dtype = np.complex128 if self.cplx else np.float64
        self._eng = None  # built on the first evaluation, then reused
        self._sym = None  # eigendecomposition of -dt K_s of the last evaluation

    def close(self):
        if self._eng is not None:
            self._eng.close()
            self._eng = None

    def hopping(self, t=1.0, tp=0.0):
        return t * self.t_pattern + tp * self.tp_pattern

    def _sims(self, K, mu, lam):
        N, dt = self.N, self.dt
        Ks = np.asarray(K, dtype=self._dtype) - mu * np.eye(N)
        # a Hermitian hopping matrix (every lattice the generator builds):
        # one eigendecomposition gives the four exponentials and, in the
        # chain rule, the Frechet derivative
        self._sym = _SymExp(-dt * Ks) if np.array_equal(Ks, Ks.conj().T) else None
        ex = self._sym.exp if self._sym is not None else (lambda c: expm(-c * dt * Ks))
        E = ex(1.0)
        self._Einv = ex(-1.0)  # E^-1, reused by the chain rule
        base = dict(self._tmpl["params"])
        base["exp_Ku"] = base["exp_Kd"] = np.ascontiguousarray(E.T)
        base["inv_exp_Ku"] = base["inv_exp_Kd"] = np.ascontiguousarray(self._Einv.T)
        base["exp_halfKu"] = base["exp_halfKd"] = np.ascontiguousarray(ex(0.5).T)
        base["inv_exp_halfKu"] = base["inv_exp_halfKd"] = np.ascontiguousarray(ex(-0.5).T)
        el = np.exp(lam) * np.ones(N)
        base["exp_lambda"] = np.array((1.0 / el, el))
        base["del"] = np.array((el ** 2 - 1.0, el ** -2 - 1.0))
        sims = []
        for f in self.fields:
            st = dict(self._tmpl["state"])
            st["hs"] = f.copy()
            sims.append({"params": base, "state": st, "meas_eqlt": self._tmpl["meas_eqlt"],
                         "meas_uneqlt": {}, "metadata": self._tmpl["metadata"]})
        return sims, E

    def _engine(self, K, mu, lam):
        sims, E = self._sims(K, mu, lam)
        if self._eng is None:
            # the fields are fixed: one engine for every evaluation, its
            # parameters replaced when K, mu or lambda change
            self._eng = _eng.Engine(sims, device=self.device, gemm_tile=1)
        else:  # every field shares the parameters: one struct for all chains
            self._eng.set_params(range(len(sims)), sims[0]["params"])
        return self._eng, E

    def run(self, K, mu, lam):
        """Forward + adjoint sums on the GPU.  Returns (logdet [nf] summed over
        spins, dict of per-field pieces)."""
        eng, E = self._engine(K, mu, lam)
        nf, N = len(self.fields), self.N
        lib = eng.L
        dt_ = self._dtype
        logdet = np.zeros(2 * nf)
        W = np.zeros(2 * nf * N * N, dtype=dt_)
        trG = np.zeros(2 * nf, dtype=dt_)
        sG = np.zeros(2 * nf, dtype=dt_)
        dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
        lib.dqmc_engine_adjoint.argtypes = [C.c_void_p] + [C.POINTER(C.c_double)] * 4
        rc = lib.dqmc_engine_adjoint(eng.h, dp(logdet), dp(W), dp(trG), dp(sG))
        if rc != 0:
            raise _eng.DqmcError("dqmc_engine_adjoint: %s" % lib.dqmc_last_error().decode())
        # column-major buffers -> (row, col) matrices
        Wm = W.reshape(nf, 2, N, N).transpose(0, 1, 3, 2)
        return logdet.reshape(nf, 2).sum(axis=1), {
            "logdet": logdet.reshape(nf, 2), "W": Wm, "trG": trG.reshape(nf, 2),
            "sG": sG.reshape(nf, 2), "E": E, "Einv": self._Einv}

    def run_green(self, K, mu, lam, X):
        """O_f = sum_s tr(X_s G_s) for every field and its adjoint sums.  X:
        one N x N matrix for both spins, or a pair (X_up, X_dn).  Returns
        (O [nf], dict of per-field, per-spin pieces: obs, W, trZ, sZ, E, Einv)."""
        eng, E = self._engine(K, mu, lam)
        nf, N = len(self.fields), self.N
        lib = eng.L
        dt_ = self._dtype
        Xs = np.asarray(X, dtype=dt_)
        if Xs.ndim == 2:
            Xs = np.stack([Xs, Xs])
        # column-major per spin: the transpose's C-order buffer
        Xc = np.ascontiguousarray(Xs.transpose(0, 2, 1))
        obs = np.zeros(2 * nf, dtype=dt_)
        W = np.zeros(2 * nf * N * N, dtype=dt_)
        trZ = np.zeros(2 * nf, dtype=dt_)
        sZ = np.zeros(2 * nf, dtype=dt_)
        dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
        lib.dqmc_engine_adjoint_green.argtypes = [C.c_void_p] + [C.POINTER(C.c_double)] * 5
        rc = lib.dqmc_engine_adjoint_green(eng.h, dp(Xc), dp(obs), dp(W), dp(trZ), dp(sZ))
        if rc != 0:
            raise _eng.DqmcError("dqmc_engine_adjoint_green: %s" % lib.dqmc_last_error().decode())
        Wm = W.reshape(nf, 2, N, N).transpose(0, 1, 3, 2)
        O = obs.reshape(nf, 2).sum(axis=1)
        return (O if self.cplx else O.real), {
            "obs": obs.reshape(nf, 2), "W": Wm, "trZ": trZ.reshape(nf, 2),
            "sZ": sZ.reshape(nf, 2), "E": E, "Einv": self._Einv}

    def grad_K_from(self, K, mu, W_weighted, E, Einv=None):
        """d/dK of sum_f w_f sum_s Re(log det or observable), given
        sum_f w_f W_f,s summed over s (both spins share K_s = K - mu I here);
        Einv: E^-1 if at hand.  Complex K: d/dRe K + i d/dIm K."""
        dt = self.dt
        A = -dt * (np.asarray(K, dtype=self._dtype) - mu * np.eye(self.N))
        Z = W_weighted @ (np.linalg.inv(E) if Einv is None else Einv)
        if np.array_equal(A, A.conj().T):
            # the decomposition of the last evaluation when it is of this A
            sym = self._sym if self._sym is not None and np.array_equal(self._sym.A, A) else _SymExp(A)
            Lf = sym.frechet(Z)
        else:
            Lf = expm_frechet(A, Z, compute_expm=False)
        g = -dt * Lf.T
        return g.conj() if np.iscomplexobj(g) else g


def logdet_and_grads(prob, K, mu, lam):
    """Per-field log|det M_up M_dn| and the gradients of their SUM with respect
    to K (full matrix), mu and lambda (per field for mu and lambda)."""
    f, d = prob.run(K, mu, lam)
    Wsum = d["W"].sum(axis=(0, 1))
    gK = prob.grad_K_from(K, mu, Wsum, d["E"], d.get("Einv"))
    N, L, dt = prob.N, prob.L, prob.dt
    gmu = dt * (2 * N * L - d["trG"].sum(axis=1).real)
    glam = d["sG"].sum(axis=1).real
    return f, gK, gmu, glam, d


def green_and_grads(prob, K, mu, lam, X):
    """Per-field O_f = Re sum_s tr(X_s G_s) and the gradients of their SUM with
    respect to K (full matrix), mu and lambda (per field for mu and lambda)."""
    O, d = prob.run_green(K, mu, lam, X)
    Wsum = d["W"].sum(axis=(0, 1))
    gK = prob.grad_K_from(K, mu, Wsum, d["E"], d.get("Einv"))
    gmu = prob.dt * d["trZ"].sum(axis=1).real
    glam = d["sZ"].sum(axis=1).real
    return np.real(O), gK, gmu, glam, d


def density_X(N):
    """X of the density: n = 2 + sum_s tr(X G_s) with X = -I/N (n = (1/N)
    sum_s sum_i (1 - G_s,ii))."""
    return -np.eye(N) / N


try:
    import torch

    def _weighted_grads(prob, ctx_K, ctx_mu, d, g, trkey, skey, trsign, troff):
        w = g.detach().cpu().numpy().astype(np.float64)
        Ww = np.tensordot(w, d["W"].sum(axis=1), axes=(0, 0))
        gK = prob.grad_K_from(ctx_K, ctx_mu, Ww, d["E"], d.get("Einv"))
        N, L, dt = prob.N, prob.L, prob.dt
        gmu = float(np.dot(w, dt * (troff * 2 * N * L + trsign * d[trkey].sum(axis=1).real)))
        glam = float(np.dot(w, d[skey].sum(axis=1).real))
        return (torch.as_tensor(gK), torch.tensor(gmu, dtype=torch.float64),
                torch.tensor(glam, dtype=torch.float64))

    class LogDet(torch.autograd.Function):
        """y[f] = log|det M_up M_dn| of field f; differentiable in K (N x N,
        real or complex Hermitian), mu and lam (0-d tensors).  `prob` carries
        the fields and lattice."""

        @staticmethod
        def forward(ctx, K, mu, lam, prob):
            Kn = K.detach().cpu().numpy()
            f, d = prob.run(Kn, float(mu), float(lam))
            ctx.prob, ctx.K, ctx.mu, ctx.d = prob, Kn, float(mu), d
            return torch.as_tensor(f, dtype=torch.float64)

        @staticmethod
        def backward(ctx, g):
            gK, gmu, glam = _weighted_grads(ctx.prob, ctx.K, ctx.mu, ctx.d, g, "trG", "sG", -1, 1)
            return gK, gmu, glam, None

    class GreenTrace(torch.autograd.Function):
        """y[f] = Re sum_s tr(X_s G_s) of field f (G at slice 0, X fixed: an
        N x N matrix for both spins or a pair); differentiable in K, mu and
        lam like LogDet."""

        @staticmethod
        def forward(ctx, K, mu, lam, X, prob):
            Kn = K.detach().cpu().numpy()
            Xn = X.detach().cpu().numpy() if torch.is_tensor(X) else np.asarray(X)
            O, d = prob.run_green(Kn, float(mu), float(lam), Xn)
            ctx.prob, ctx.K, ctx.mu, ctx.d = prob, Kn, float(mu), d
            return torch.as_tensor(np.real(O), dtype=torch.float64)

        @staticmethod
        def backward(ctx, g):
            gK, gmu, glam = _weighted_grads(ctx.prob, ctx.K, ctx.mu, ctx.d, g, "trZ", "sZ", 1, 0)
            return gK, gmu, glam, None, None
except ImportError:  # torch is optional for the numpy interface
    LogDet = GreenTrace = None
