"""Pins the oracle's term-by-term restatement of the optional unequal-time
contractions (src/meas.c:409-471, 860-1365: thermal, 2-bond, energy kv/kn)
against an independent derivation: per-spin Wick factorisation of the same
operators on random Green's functions.

Operators (per spin; the reference's conventions):
  hop H = p c+_x c_y, <H> = -p G[y,x] (no delta term, also for a two-hop bond
  that returns to its start site); density <n_x> = 1 - G[x,x];
  connected <c+_a c_b (t) c+_c c_d (0)> = (dt d_ad - G0t[d,a]) Gt0[b,c];
  J = F - B, K = F + B on a bond, N = n_x0 + n_x1, jn = sum_s J_s N_-s.
A mistake in either form (a dropped term, a swapped index or phase) breaks
the agreement, which holds to rounding."""
import numpy as np
import pytest

from dqmc_amd import gen
import oracle


def _wick(sim, phase, G0t, Gtt, Gt0, L):
    p = sim["params"]
    N = int(p["N"])
    cplx = p["exp_Ku"].dtype.kind == "c"
    nb, nb2 = int(p["num_b"]), int(p["num_b2"])
    bonds, b2s = p["bonds"].reshape(-1), p["bond2s"].reshape(-1)
    P = [p["peierlsu"], p["peierlsd"]]
    ppf = [p["pp_u"].reshape(-1), p["pp_d"].reshape(-1)]
    ppr = [p["ppr_u"].reshape(-1), p["ppr_d"].reshape(-1)]
    dtp = np.complex128 if cplx else np.float64
    sizes = {k: int(p[n]) for k, n in (("bb", "num_bb"), ("bs", "num_bs"), ("bb2", "num_bb2"),
                                       ("b2b", "num_b2b"), ("b2b2", "num_b2b2"))}
    out = {k: np.zeros(sizes[s] * L, dtype=dtp) for k, s in (
        ("jnj", "bb"), ("jjn", "bb"), ("jnjn", "bb"), ("j2jn", "bb2"), ("j2j", "bb2"),
        ("jnj2", "b2b"), ("jj2", "b2b"), ("j2j2", "b2b2"), ("k2k2", "b2b2"), ("kv", "bs"),
        ("kn", "bs"))}

    def ph(s, a, b):  # C's p[a + N b] on the raw buffer = numpy P[b, a]
        return P[s][b, a] if cplx else 1.0

    for t in range(L):
        dt = 1.0 if t == 0 else 0.0

        def E(s, ops, tau):
            G = Gtt[s][t] if tau else Gtt[s][0]
            return sum((1.0 - G[x, x]) if c == "n" else -c * G[y, x] for x, y, c in ops)

        def Cn(s, A, B):
            g0t, gt0 = G0t[s][t], Gt0[s][t]
            tot = 0.0
            for xa, ya, ca in A:
                for xb, yb, cb in B:
                    wa = 1.0 if ca == "n" else ca
                    wb = 1.0 if cb == "n" else cb
                    tot += wa * wb * ((dt if xa == yb else 0.0) - g0t[yb, xa]) * gt0[ya, xb]
            return tot

        def J1(s, b, sg=-1):
            i0, i1 = bonds[b], bonds[b + nb]
            return [(i0, i1, ph(s, i0, i1)), (i1, i0, sg * ph(s, i1, i0))]

        def J2(s, c, sg=-1):
            i0, i2 = b2s[c], b2s[c + nb2]
            return [(i0, i2, ppf[s][c]), (i2, i0, sg * ppr[s][c])]

        def Nb(b):
            return [(bonds[b], bonds[b], "n"), (bonds[b + nb], bonds[b + nb], "n")]

        def tot2(A, B):  # <sum_s A_s (t) sum_s B_s (0)>
            return (E(0, A[0], 1) + E(1, A[1], 1)) * (E(0, B[0], 0) + E(1, B[1], 0)) + \
                Cn(0, A[0], B[0]) + Cn(1, A[1], B[1])

        def jn_left(A, An, B):  # <sum_s A_s An_-s (t) sum_s B_s (0)>
            r = 0.0
            for s in (0, 1):
                o = 1 - s
                r += (E(s, A[s], 1) * E(s, B[s], 0) + Cn(s, A[s], B[s])) * E(o, An[o], 1)
                r += E(s, A[s], 1) * (E(o, An[o], 1) * E(o, B[o], 0) + Cn(o, An[o], B[o]))
            return r

        def jn_right(A, B, Bn):  # <sum_s A_s (t) sum_s B_s Bn_-s (0)>
            r = 0.0
            for s in (0, 1):
                o = 1 - s
                r += (E(s, A[s], 1) * E(s, B[s], 0) + Cn(s, A[s], B[s])) * E(o, Bn[o], 0)
                r += E(s, B[s], 0) * (E(o, A[o], 1) * E(o, Bn[o], 0) + Cn(o, A[o], Bn[o]))
            return r

        for c in range(nb):
            Bj = [J1(s, c) for s in (0, 1)]
            Bn = [Nb(c)] * 2
            for b in range(nb):
                o = p["map_bb"][c][b] + sizes["bb"] * t
                pre = phase / int(p["degen_bb"])
                Aj = [J1(s, b) for s in (0, 1)]
                An = [Nb(b)] * 2
                out["jnj"][o] += pre * jn_left(Aj, An, Bj)
                out["jjn"][o] += pre * jn_right(Aj, Bj, Bn)
                r = 0.0
                for s in (0, 1):
                    w = 1 - s
                    r += (E(s, Aj[s], 1) * E(s, Bj[s], 0) + Cn(s, Aj[s], Bj[s])) * \
                         (E(w, An[w], 1) * E(w, Bn[w], 0) + Cn(w, An[w], Bn[w]))
                    r += (E(s, Aj[s], 1) * E(s, Bn[s], 0) + Cn(s, Aj[s], Bn[s])) * \
                         (E(w, An[w], 1) * E(w, Bj[w], 0) + Cn(w, An[w], Bj[w]))
                out["jnjn"][o] += pre * r
            for b in range(nb2):
                o = p["map_bb2"][c][b] + sizes["bb2"] * t
                pre = phase / int(p["degen_bb2"])
                A2 = [J2(s, b) for s in (0, 1)]
                out["j2j"][o] += pre * tot2(A2, Bj)
                out["j2jn"][o] += pre * jn_right(A2, Bj, Bn)
        for c in range(nb2):
            B2 = [J2(s, c) for s in (0, 1)]
            K2 = [J2(s, c, +1) for s in (0, 1)]
            for b in range(nb2):
                o = p["map_b2b2"][c][b] + sizes["b2b2"] * t
                pre = phase / int(p["degen_b2b2"])
                out["j2j2"][o] += pre * tot2([J2(s, b) for s in (0, 1)], B2)
                out["k2k2"][o] += pre * tot2([J2(s, b, +1) for s in (0, 1)], K2)
            for b in range(nb):
                o = p["map_b2b"][c][b] + sizes["b2b"] * t
                pre = phase / int(p["degen_b2b"])
                Aj = [J1(s, b) for s in (0, 1)]
                out["jj2"][o] += pre * tot2(Aj, B2)
                out["jnj2"][o] += pre * jn_left(Aj, [Nb(b)] * 2, B2)
        for j in range(N):
            nj = [(j, j, "n")]
            for b in range(nb):
                o = p["map_bs"][j][b] + sizes["bs"] * t
                pre = phase / int(p["degen_bs"])
                K = [J1(s, b, +1) for s in (0, 1)]
                kv = 0.0
                for s in (0, 1):
                    kv += (E(s, K[s], 1) * E(s, nj, 0) + Cn(s, K[s], nj)) * E(1 - s, nj, 0)
                out["kv"][o] += pre * kv
                out["kn"][o] += pre * tot2(K, [nj, nj])
    return out


@pytest.mark.parametrize("nflux", [0, 1])
def test_optional_contractions_match_wick_factorisation(nflux):
    tog = dict(meas_thermal=1, meas_2bond_corr=1, meas_energy_corr=1)
    # 2 x 3 lattice (real): some two-hop bonds return to their start site
    Nx = 2 if nflux == 0 else 3
    sim = gen.create_batch(seed=3, Nx=Nx, Ny=3, L=2, n_matmul=2, period_eqlt=2, period_uneqlt=1,
                           nflux=nflux, tp=-0.3, **tog)[0]
    N, L = int(sim["params"]["N"]), 2
    rs = np.random.RandomState(1)

    def rnd():
        a = rs.randn(L, N, N) * 0.3
        return a + 1j * rs.randn(L, N, N) * 0.3 if nflux else a

    Gu0t, Gutt, Gut0, Gd0t, Gdtt, Gdt0 = (rnd() for _ in range(6))
    phase = (0.6 + 0.8j) if nflux else -1.0
    ref = oracle.measure_uneqlt(sim, phase, Gu0t, Gutt, Gut0, Gd0t, Gdtt, Gdt0)
    mine = _wick(sim, phase, [Gu0t, Gd0t], [Gutt, Gdtt], [Gut0, Gdt0], L)
    for k, v in mine.items():
        r = np.asarray(ref[k])
        assert np.max(np.abs(r)) > 0.1, k
        assert np.max(np.abs(r - v)) <= 1e-13 * np.max(np.abs(r)), k
