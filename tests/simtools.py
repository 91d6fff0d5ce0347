"""Helpers shared by the tests (not part of the package)."""
import numpy as np

from dqmc_amd import capi

_NUM_PARAMS = ("peierlsu", "peierlsd", "pp_u", "pp_d", "ppr_u", "ppr_d", "exp_Ku", "exp_Kd",
               "inv_exp_Ku", "inv_exp_Kd", "exp_halfKu", "exp_halfKd", "inv_exp_halfKu",
               "inv_exp_halfKd", "Ku", "Kd")


def as_complex(sim):
    """The same simulation as a complex (USE_CPLX) file: every `num` dataset
    complex128 with zero imaginary parts, the field and the random state
    unchanged."""
    p = dict(sim["params"])
    for k in _NUM_PARAMS:
        p[k] = np.asarray(p[k]).astype(np.complex128)
    out = {"params": p, "state": dict(sim["state"]), "metadata": dict(sim.get("metadata", {}))}
    for g in ("meas_eqlt", "meas_uneqlt"):
        out[g] = {k: (v.astype(np.complex128) if k != "n_sample" else v.copy())
                  for k, v in sim[g].items()}
    assert capi.is_complex(out)
    return out


# Accumulators that differ between the alpha = 1/2 and alpha = 1 gauges of
# the generator: the pair fields pick up e^{2i theta}; and g00 / gt0 weight
# G[i,j] with the straight-line Peierls phase of every site pair, which for
# pairs whose path wraps both torus directions differs by pi between the two
# gauges (the phase construction of util/tight_binding.py, reproduced by the
# generator) -- the bond phases, and so K and the trajectory, are covariant.
GAUGE_VARIANT = {"pair_sw", "pair_bb", "pair_b2b2", "g00", "g00_u", "g00_d", "gt0", "gt0_u",
                 "gt0_d"}


def _bmats(sim, spin):
    """B_l = exp_K diag(e^{+-lambda s_l}) of one spin for every slice (the raw
    buffer reading of the reference, src/dqmc.c:29-43), as numpy matrices."""
    p = sim["params"]
    N, L = int(p["N"]), int(p["L"])
    expK = (p["exp_Kd"] if spin else p["exp_Ku"]).T
    el = p["exp_lambda"]
    hs = sim["state"]["hs"]
    return [expK * el[(1 - hs[l]) if spin else hs[l], np.arange(N)][None, :] for l in range(L)]


def _qrp(X):
    """Q, R, piv with X[:, piv] = Q R: columns sorted by decreasing norm, then
    Householder QR (the pre-pivoting of src/greens.c:136-150)."""
    piv = np.argsort(-np.linalg.norm(X, axis=0), kind="stable")
    Q, R = np.linalg.qr(X[:, piv])
    return Q, R, piv


def _udv_left(U, D, V, B):
    """UDV of B (U D V): B U D = Q R P^T."""
    Q, R, piv = _qrp((B @ U) * D[None, :])
    d = np.abs(np.diag(R)).astype(float)
    d[d == 0] = 1.0
    T = R / d[:, None]
    Pt = np.zeros_like(T)
    Pt[:, piv] = T
    return Q, d, Pt @ V


def _udv_right(U, D, V, B):
    """UDV of (U D V) B: QR of (D V B)^T keeps the column scale on the left."""
    X = (D[:, None] * (V @ B)).T
    Q, R, piv = _qrp(X)
    d = np.abs(np.diag(R)).astype(float)
    d[d == 0] = 1.0
    T = R / d[:, None]
    Pt = np.zeros_like(T)
    Pt[:, piv] = T
    # (D V B) = (Q R P^T)^T = P R^T Q^T = (Pt^T) diag(d) Q^T with Pt^T lower-unit-ish
    return U @ Pt.T, d, Q.T


def _inv_sum(UA, DA, VA, UB, DB, VB):
    """(UA DA VA + UB DB VB)^-1 with the scale splitting of stable DQMC:
    D = Db Ds, Db = max(D, 1), Ds = min(D, 1)."""
    Dab, Das = np.maximum(DA, 1.0), np.minimum(DA, 1.0)
    Dbb, Dbs = np.maximum(DB, 1.0), np.minimum(DB, 1.0)
    iVB = np.linalg.inv(VB)
    iUA = np.linalg.inv(UA)
    M = (Das[:, None] * (VA @ iVB)) / Dbb[None, :] + (iUA @ UB) * (Dbs[None, :] / Dab[:, None])
    return (iVB / Dbb[None, :]) @ np.linalg.solve(M, iUA / Dab[:, None])


def ue_referee(sim, spin, taus):
    """G(t,0), G(0,t), G(t,t) at the slices `taus` by a stable evaluation that
    is independent of calc_ue_g (src/greens.c:214-494): UDV forms of
    B(t,0) = B_{t-1}..B_0 and B(beta,t) = B_{L-1}..B_t with a QR (columns
    sorted by norm first) after every slice, then
      G(t,0) = [B(t,0)^-1 + B(beta,t)]^-1,
      G(0,t) = -[B(t,0) + B(beta,t)^-1]^-1   (t > 0),
      G(t,t) = [I + B(t,0) B(beta,t)]^-1,
    each sum inverted with the scale splitting.  Returns dicts t -> matrix."""
    Bs = _bmats(sim, spin)
    L = len(Bs)
    N = Bs[0].shape[0]
    I = np.eye(N, dtype=Bs[0].dtype)
    left = {0: (I, np.ones(N), I)}
    U, D, V = I, np.ones(N), I
    for t in range(1, L + 1):
        U, D, V = _udv_left(U, D, V, Bs[t - 1])
        left[t] = (U, D, V)
    right = {L: (I, np.ones(N), I)}
    U, D, V = I, np.ones(N), I
    for t in range(L - 1, -1, -1):
        U, D, V = _udv_right(U, D, V, Bs[t])
        right[t] = (U, D, V)
    gt0, g0t, gtt = {}, {}, {}
    for t in taus:
        U1, D1, V1 = left[t]
        U2, D2, V2 = right[t]
        # B(t,0)^-1 = V1^-1 D1^-1 U1^-1
        gt0[t] = _inv_sum(np.linalg.inv(V1), 1.0 / D1, np.linalg.inv(U1), U2, D2, V2)
        if t > 0:
            g0t[t] = -_inv_sum(U1, D1, V1, np.linalg.inv(V2), 1.0 / D2, np.linalg.inv(U2))
        # I + B(t,0) B(beta,t) = I + U1 (D1 V1 U2 D2) V2: fold the middle by one more UDV
        Um, Dm, Vm = _udv_left(I, np.ones(N), I, (D1[:, None] * (V1 @ U2)) * D2[None, :])
        gtt[t] = _inv_sum(I, np.ones(N), I, U1 @ Um, Dm, Vm @ V2)
    return gt0, g0t, gtt


def ue_referee_accumulators(sim, phase, taus):
    """The unequal-time accumulators of one measurement (src/dqmc.c:389-458:
    calc_ue_g, the half-wrap of every block, measure_uneqlt) with the blocks
    taken from ue_referee at the slices `taus` (0 included), through the
    oracle's measure_uneqlt.  Only the accumulator slices t in `taus` are
    meaningful.  Returns the accumulator dict."""
    import oracle
    p = sim["params"]
    N, L = int(p["N"]), int(p["L"])
    cplx = np.iscomplexobj(p["exp_Ku"])
    dt = np.complex128 if cplx else np.float64
    taus = sorted(set(taus) | {0})
    arrs = []
    for spin in (0, 1):
        gt0, g0t, gtt = ue_referee(sim, spin, taus)
        H = (p["inv_exp_halfKd"] if spin else p["inv_exp_halfKu"]).T
        Hi = (p["exp_halfKd"] if spin else p["exp_halfKu"]).T
        A = [np.zeros((L, N, N), dtype=dt) for _ in range(3)]
        for t in taus:
            g = {0: g0t.get(t, gtt[0]) if t > 0 else gtt[0], 1: gtt[t], 2: gt0[t]}
            for k in range(3):
                A[k][t] = H @ g[k] @ Hi
        arrs.append(A)
    (u0t, utt, ut0), (d0t, dtt, dt0) = arrs
    return oracle.measure_uneqlt(sim, phase, u0t, utt, ut0, d0t, dtt, dt0)


def slice_error(acc, ref, n_per_t, taus):
    """max |acc - ref| over the accumulator slices t in taus, relative to the
    largest |ref| there."""
    a, r = np.asarray(acc), np.asarray(ref)
    idx = np.concatenate([np.arange(t * n_per_t, (t + 1) * n_per_t) for t in taus])
    return np.abs(a[idx] - r[idx]).max() / max(np.abs(r[idx]).max(), 1e-300)
