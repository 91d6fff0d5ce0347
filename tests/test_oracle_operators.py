"""Pins the oracle's measurement formulas (src/meas.c:66-1383, restated term by
term in oracle/dqmc_oracle.c and oracle/meas_ext.inc) against an independent
derivation: each accumulator is written as the expectation of the physical
operator it measures, and that expectation is evaluated by Wick's theorem
(the recursive Pfaffian expansion over elementary fermion operators) on
random Green's functions, real and complex.

Contractions, for operators at time t placed left of those at time 0
(G(t,0)[i,j] = <c_i(t) c+_j(0)>, the reference's arrays):
  <c_b(t) c+_a(t)>  = Gtt[b,a]        <c+_a(t) c_b(t)>  = d_ab - Gtt[b,a]
  <c_b(0) c+_a(0)>  = G00[b,a]        <c+_a(0) c_b(0)>  = d_ab - G00[b,a]
  <c_b(t) c+_a(0)>  = Gt0[b,a]        <c+_a(t) c_b(0)>  = d_t0 d_ab - G0t[b,a]
with G00 = Gtt at t = 0 (src/meas.c:746-747).  Equal-time groups are the
same with every array equal to G and t = 0.  The reference's own conventions
on Kronecker deltas are kept as switches: a hop c+_x c_y measures no delta
even when x = y (a two-hop bond that returns to its start site), except in
the equal-time j2 of local JQ; the nematic correlators drop every delta
between different bilinears.

A mistake in either form (a dropped term, a swapped index or spin, a
misplaced Peierls phase, a wrong bin) breaks the agreement, which holds to
rounding."""
import itertools

import numpy as np
import pytest

from dqmc_amd import gen
import oracle

_SELF = {}
_bid = itertools.count()


class Ctx:
    def __init__(self, G0t, Gtt, Gt0, G00, t0, cross_delta=True):
        self.G0t, self.Gtt, self.Gt0, self.G00 = G0t, Gtt, Gt0, G00
        self.dt = 1.0 if t0 else 0.0
        self.cross = cross_delta


def _contract(p, q, x):
    """<p q> for elementary operators (dag, site, spin, tau, bilinear id), p left of q."""
    if p[2] != q[2] or p[0] == q[0]:
        return 0.0
    s = p[2]
    if p[0]:  # <c+_a c_b>
        a, b = p[1], q[1]
        if p[3] == q[3]:
            G, d = (x.Gtt if p[3] == "t" else x.G00), 1.0
        elif p[3] == "t":
            G, d = x.G0t, x.dt
        else:
            raise ValueError("operators at time 0 must stand right of those at t")
        keep = _SELF[p[4]] if (p[4] is not None and p[4] == q[4]) else x.cross
        return (d if (keep and a == b) else 0.0) - G[s][b, a]
    b, a = p[1], q[1]  # <c_b c+_a>
    if p[3] == q[3]:
        G = x.Gtt if p[3] == "t" else x.G00
    elif p[3] == "t":
        G = x.Gt0
    else:
        raise ValueError("operators at time 0 must stand right of those at t")
    return G[s][b, a]


def wick(ops, x):
    """Expectation of the operator product (left to right) by the Pfaffian expansion."""
    if not ops:
        return 1.0
    if len(ops) % 2:
        return 0.0
    first, tot = ops[0], 0.0
    for q in range(1, len(ops)):
        c = _contract(first, ops[q], x)
        if c != 0.0:
            tot += (-1) ** (q - 1) * c * wick(ops[1:q] + ops[q + 1:], x)
    return tot


def ev(terms, x):
    return sum(c * wick(list(o), x) for c, o in terms)


def mul(*As):
    out = [(1.0, ())]
    for A in As:
        out = [(c1 * c2, o1 + o2) for c1, o1 in out for c2, o2 in A]
    return out


def add(*As, w=None):
    w = w or [1.0] * len(As)
    return [(wi * c, o) for wi, A in zip(w, As) for c, o in A]


def bil(a, b, s1, s2, tau, coef=1.0, selfdelta=True):
    """coef c+_{a s1} c_{b s2}"""
    k = next(_bid)
    _SELF[k] = selfdelta
    return [(coef, ((1, a, s1, tau, k), (0, b, s2, tau, k)))]


class Ops:
    """The reference's operators on one lattice (phases from the raw buffers:
    a hop c+_a c_b carries p[a + N b], i.e. numpy P[b, a])."""

    def __init__(self, p):
        self.p = p
        self.N = int(p["N"])
        self.cplx = p["exp_Ku"].dtype.kind == "c"
        self.P = [p["peierlsu"], p["peierlsd"]]
        self.nb, self.nb2 = int(p["num_b"]), int(p["num_b2"])
        self.bonds = p["bonds"].reshape(-1)
        self.b2s = p["bond2s"].reshape(-1)
        self.ppf = [p["pp_u"].reshape(-1), p["pp_d"].reshape(-1)]
        self.ppr = [p["ppr_u"].reshape(-1), p["ppr_d"].reshape(-1)]

    def ph(self, s, a, b):
        return self.P[s][b, a] if self.cplx else 1.0

    def n(self, i, s, tau):
        return bil(i, i, s, s, tau)

    def ntot(self, i, tau):
        return add(self.n(i, 0, tau), self.n(i, 1, tau))

    def sz(self, i, tau):
        return add(self.n(i, 0, tau), self.n(i, 1, tau), w=[0.5, -0.5])

    def sx(self, i, tau):
        return add(bil(i, i, 0, 1, tau), bil(i, i, 1, 0, tau), w=[0.5, 0.5])

    def spin(self, a, i, tau):
        sig = [np.array([[0, 1], [1, 0]]), np.array([[0, -1j], [1j, 0]]),
               np.array([[1, 0], [0, -1]])][a]
        return add(*[bil(i, i, s1, s2, tau, 0.5 * sig[s1, s2])
                     for s1 in (0, 1) for s2 in (0, 1) if sig[s1, s2] != 0])

    def v(self, i, tau):
        return mul(self.n(i, 0, tau), self.n(i, 1, tau))

    def pair(self, i, tau):  # c_{i up} c_{i dn}
        return [(1.0, ((0, i, 0, tau, None), (0, i, 1, tau, None)))]

    def pair_dag(self, j, tau):  # c+_{j dn} c+_{j up}
        return [(1.0, ((1, j, 1, tau, None), (1, j, 0, tau, None)))]

    def bond(self, b):
        return self.bonds[b], self.bonds[b + self.nb]

    def hop(self, a, b, s, tau, coef=1.0, selfdelta=False):
        return bil(a, b, s, s, tau, coef * self.ph(s, a, b), selfdelta)

    def J(self, b, s, tau, sg=-1):  # J (sg = -1) or K (sg = +1) of bond b, spin s
        i0, i1 = self.bond(b)
        return add(self.hop(i0, i1, s, tau), self.hop(i1, i0, s, tau, sg))

    def J2(self, c, s, tau, sg=-1, selfdelta=False):
        i0, i2 = self.b2s[c], self.b2s[c + self.nb2]
        return add(bil(i0, i2, s, s, tau, self.ppf[s][c], selfdelta),
                   bil(i2, i0, s, s, tau, sg * self.ppr[s][c], selfdelta))


def _rand(rs, shape, cplx, scale=0.3):
    a = rs.randn(*shape) * scale
    return a + 1j * rs.randn(*shape) * scale if cplx else a


def _close(ref, mine, name):
    ref, mine = np.asarray(ref), np.asarray(mine)
    assert np.max(np.abs(ref)) > 1e-3, name
    assert np.max(np.abs(ref - mine)) <= 1e-12 * np.max(np.abs(ref)), name


# ---------------------------------------------------------------- equal time
def _eqlt_expected(sim, phase, gu, gd):
    p = sim["params"]
    O = Ops(p)
    N = O.N
    G = [gu, gd]
    x = Ctx(G, G, G, G, True)
    dt = np.complex128 if O.cplx else np.float64
    num = lambda k: np.zeros(int(p[k]), dtype=dt)
    out = {k: num("num_i") for k in ("density", "density_u", "density_d", "double_occ")}
    for k in ("g00", "g00_u", "g00_d", "nn", "xx", "zz", "pair_sw", "vv", "vn"):
        out[k] = num("num_ij")
    for i in range(N):
        r = p["map_i"][i]
        pre = phase / int(p["degen_i"])
        nu, nd = ev(O.n(i, 0, "t"), x), ev(O.n(i, 1, "t"), x)
        out["density"][r] += pre * (nu + nd)
        out["density_u"][r] += pre * nu
        out["density_d"][r] += pre * nd
        out["double_occ"][r] += pre * ev(O.v(i, "t"), x)
    for j in range(N):
        for i in range(N):
            r = p["map_ij"][j][i]
            pre = phase / int(p["degen_ij"])
            # <c_i c+_j> carrying the phase of the hop c+_j c_i
            gs = [ev([(1.0, ((0, i, s, "t", None), (1, j, s, "0", None)))], x) * O.ph(s, j, i)
                  for s in (0, 1)]
            out["g00"][r] += 0.5 * pre * (gs[0] + gs[1])
            out["g00_u"][r] += pre * gs[0]
            out["g00_d"][r] += pre * gs[1]
            out["nn"][r] += pre * ev(mul(O.ntot(i, "t"), O.ntot(j, "0")), x)
            out["xx"][r] += pre * ev(mul(O.sx(i, "t"), O.sx(j, "0")), x)
            out["zz"][r] += pre * ev(mul(O.sz(i, "t"), O.sz(j, "0")), x)
            out["pair_sw"][r] += pre * ev(mul(O.pair(i, "t"), O.pair_dag(j, "0")), x)
            out["vv"][r] += pre * ev(mul(O.v(i, "t"), O.v(j, "0")), x)
            out["vn"][r] += pre * ev(mul(O.v(i, "t"), O.ntot(j, "0")), x)
    nb = O.nb
    out["kv"], out["kn"], out["kk"] = num("num_bs"), num("num_bs"), num("num_bb")
    for j in range(N):
        for b in range(nb):
            r = p["map_bs"][j][b]
            pre = phase / int(p["degen_bs"])
            K = add(O.J(b, 0, "t", +1), O.J(b, 1, "t", +1))
            out["kv"][r] += pre * ev(mul(K, O.v(j, "0")), x)
            out["kn"][r] += pre * ev(mul(K, O.ntot(j, "0")), x)
    for c in range(nb):
        Kc = add(O.J(c, 0, "0", +1), O.J(c, 1, "0", +1))
        for b in range(nb):
            r = p["map_bb"][c][b]
            Kb = add(O.J(b, 0, "t", +1), O.J(b, 1, "t", +1))
            out["kk"][r] += phase / int(p["degen_bb"]) * ev(mul(Kb, Kc), x)
    if int(p["meas_chiral"]):
        out["chi"] = num("num_plaq_accum")
        eps = [((0, 1, 2), 1), ((1, 2, 0), 1), ((2, 0, 1), 1), ((0, 2, 1), -1), ((2, 1, 0), -1),
               ((1, 0, 2), -1)]
        for a in range(int(p["num_plaq"])):
            i0, i1, i2 = (p["plaqs"][k][a] for k in range(3))
            # 4i <S_i0 . (S_i1 x S_i2)>: the normalisation of the reference's expansion
            chi = sum(sg * ev(mul(O.spin(u, i0, "t"), O.spin(v, i1, "t"), O.spin(w, i2, "t")), x)
                      for (u, v, w), sg in eps)
            val = 4j * chi
            out["chi"][p["map_plaq"][a]] += phase / int(p["degen_plaq"]) * (val if O.cplx else val.real)
    if int(p["meas_gen_suscept"]):
        nij = int(p["num_ij"])
        for k in ("uuuu", "dddd", "uudd", "dduu"):
            out[k] = np.zeros(nij * nij, dtype=dt)
        _gen_suscept(p, phase / (int(p["Nx"]) * int(p["Ny"])) ** 2, x, out, 0, nij)
    if int(p["meas_local_JQ"]):
        out["j"], out["jn"], out["j2"] = num("num_b_accum"), num("num_b_accum"), num("num_b2_accum")
        for c in range(nb):
            i0, i1 = O.bond(c)
            r = p["map_b"][c]
            pre = phase / int(p["degen_b"])
            Jb = [O.J(c, s, "t") for s in (0, 1)]
            Nb = [add(O.n(i0, s, "t"), O.n(i1, s, "t")) for s in (0, 1)]
            out["j"][r] += pre * ev(add(Jb[0], Jb[1]), x)
            out["jn"][r] += pre * ev(add(mul(Jb[0], Nb[1]), mul(Jb[1], Nb[0])), x)
        for c in range(O.nb2):
            out["j2"][p["map_b2"][c]] += phase / int(p["degen_b2"]) * ev(
                add(O.J2(c, 0, "t", selfdelta=True), O.J2(c, 1, "t", selfdelta=True)), x)
    return out


def _gen_suscept(p, pre, x, out, t, nij):
    """The four spin labels of the connected <c+_i1 c_j1 (t) c+_i2 c_j2 (0)>:
    <c+_i1(t) c_j2(0)> of the first spin times <c_j1(t) c+_i2(0)> of the second,
    binned by the displacement of each contraction (i1, j2) and (i2, j1)."""
    N = int(p["N"])
    labels = {"uuuu": (0, 0), "dddd": (1, 1), "uudd": (0, 1), "dduu": (1, 0)}
    for j1 in range(N):
        for i1 in range(N):
            for j2 in range(N):
                for i2 in range(N):
                    r = p["map_ij"][j2][i1] + nij * p["map_ij"][j1][i2] + nij * nij * t
                    for k, (s1, s2) in labels.items():
                        a = wick([(1, i1, s1, "t", None), (0, j2, s1, "0", None)], x)
                        b = wick([(0, j1, s2, "t", None), (1, i2, s2, "0", None)], x)
                        out[k][r] += pre * a * b


@pytest.mark.parametrize("nflux", [0, 1])
def test_equal_time_groups_are_operator_expectations(nflux):
    sim = gen.create_batch(seed=5, Nx=3, Ny=3, L=2, n_matmul=2, period_eqlt=2, tp=-0.3,
                           nflux=nflux, meas_energy_corr=1, meas_chiral=1,
                           meas_gen_suscept=1)[0]
    rs = np.random.RandomState(2)
    cplx = nflux != 0
    N = int(sim["params"]["N"])
    gu, gd = _rand(rs, (N, N), cplx), _rand(rs, (N, N), cplx)
    phase = (0.6 - 0.8j) if cplx else -1.0
    ref = oracle.measure_eqlt(sim, phase, gu, gd)
    mine = _eqlt_expected(sim, phase, gu, gd)
    for k, v in mine.items():
        _close(ref[k], v, k)
    assert set(mine) >= {"nn", "xx", "zz", "pair_sw", "kk", "kv", "kn", "vv", "vn", "chi",
                         "uuuu", "dddd", "uudd", "dduu"}


@pytest.mark.parametrize("nflux", [0, 1])
def test_local_jq_is_an_operator_expectation(nflux):
    sim = gen.create_batch(seed=6, Nx=3, Ny=3, L=2, n_matmul=2, period_eqlt=2, tp=-0.3,
                           nflux=nflux, trans_sym=0, meas_local_JQ=1)[0]
    rs = np.random.RandomState(3)
    cplx = nflux != 0
    N = int(sim["params"]["N"])
    gu, gd = _rand(rs, (N, N), cplx), _rand(rs, (N, N), cplx)
    # make some diagonal entries larger so the deltas of the two-hop bonds matter
    phase = (0.28 + 0.96j) if cplx else 1.0
    ref = oracle.measure_eqlt(sim, phase, gu, gd)
    mine = _eqlt_expected(sim, phase, gu, gd)
    for k in ("j", "jn", "j2", "density", "nn", "g00"):
        _close(ref[k], mine[k], k)


def test_two_hop_bonds_return_to_their_start_site_on_narrow_lattices():
    """On Nx = 2 some two-hop bonds start and end on one site: the equal-time
    j2 keeps the delta there, the unequal-time two-bond maps drop it."""
    sim = gen.create_batch(seed=6, Nx=2, Ny=3, L=2, n_matmul=2, period_eqlt=2, tp=-0.3,
                           trans_sym=0, meas_local_JQ=1)[0]
    b2 = sim["params"]["bond2s"]
    assert np.any(b2[0] == b2[1])
    rs = np.random.RandomState(4)
    N = int(sim["params"]["N"])
    gu, gd = _rand(rs, (N, N), False), _rand(rs, (N, N), False)
    ref = oracle.measure_eqlt(sim, 1.0, gu, gd)
    _close(ref["j2"], _eqlt_expected(sim, 1.0, gu, gd)["j2"], "j2")


# -------------------------------------------------------------- unequal time
def _uneqlt_expected(sim, phase, G0t, Gtt, Gt0, L, groups):
    p = sim["params"]
    O = Ops(p)
    N, nb = O.N, O.nb
    dt = np.complex128 if O.cplx else np.float64
    sz = lambda k: np.zeros(int(p[k]) * L, dtype=dt)
    out = {}
    if "default" in groups:
        for k in ("gt0", "gt0_u", "gt0_d", "nn", "xx", "zz", "pair_sw", "vv", "vn"):
            out[k] = sz("num_ij")
    if "bond" in groups:
        for k in ("pair_bb", "jj", "jsjs", "kk", "ksks"):
            out[k] = sz("num_bb")
    if "pair_bb_only" in groups:
        out["pair_bb"] = sz("num_bb")
    if "nematic" in groups:
        out["nem_nnnn"], out["nem_ssss"] = sz("num_bb"), sz("num_bb")
    if "gensus" in groups:
        nij = int(p["num_ij"])
        for k in ("uuuu", "dddd", "uudd", "dduu"):
            out[k] = np.zeros(nij * nij * L, dtype=dt)
    for t in range(L):
        x = Ctx([a[t] for a in G0t], [a[t] for a in Gtt], [a[t] for a in Gt0],
                [a[0] for a in Gtt], t == 0)
        if "default" in groups:
            for j in range(N):
                for i in range(N):
                    r = p["map_ij"][j][i] + int(p["num_ij"]) * t
                    pre = phase / int(p["degen_ij"])
                    gs = [ev([(1.0, ((0, i, s, "t", None), (1, j, s, "0", None)))], x) * O.ph(s, j, i)
                          for s in (0, 1)]
                    out["gt0"][r] += 0.5 * pre * (gs[0] + gs[1])
                    out["gt0_u"][r] += pre * gs[0]
                    out["gt0_d"][r] += pre * gs[1]
                    out["nn"][r] += pre * ev(mul(O.ntot(i, "t"), O.ntot(j, "0")), x)
                    out["xx"][r] += pre * ev(mul(O.sx(i, "t"), O.sx(j, "0")), x)
                    out["zz"][r] += pre * ev(mul(O.sz(i, "t"), O.sz(j, "0")), x)
                    out["pair_sw"][r] += pre * ev(mul(O.pair(i, "t"), O.pair_dag(j, "0")), x)
                    out["vv"][r] += pre * ev(mul(O.v(i, "t"), O.v(j, "0")), x)
                    out["vn"][r] += pre * ev(mul(O.v(i, "t"), O.ntot(j, "0")), x)
        if "bond" in groups or "pair_bb_only" in groups:
            for c in range(nb):
                j0, j1 = O.bond(c)
                # singlet bond pair c+_{j1 dn} c+_{j0 up} + c+_{j0 dn} c+_{j1 up}, and its partner
                Dc = [(1.0, ((1, j1, 1, "0", None), (1, j0, 0, "0", None))),
                      (1.0, ((1, j0, 1, "0", None), (1, j1, 0, "0", None)))]
                Jc = [O.J(c, s, "0") for s in (0, 1)]
                Kc = [O.J(c, s, "0", +1) for s in (0, 1)]
                for b in range(nb):
                    i0, i1 = O.bond(b)
                    r = p["map_bb"][c][b] + int(p["num_bb"]) * t
                    pre = phase / int(p["degen_bb"])
                    Db = [(1.0, ((0, i0, 0, "t", None), (0, i1, 1, "t", None))),
                          (1.0, ((0, i1, 0, "t", None), (0, i0, 1, "t", None)))]
                    out["pair_bb"][r] += 0.5 * pre * ev(mul(Db, Dc), x)
                    if "bond" not in groups:
                        continue
                    Jb = [O.J(b, s, "t") for s in (0, 1)]
                    Kb = [O.J(b, s, "t", +1) for s in (0, 1)]
                    out["jj"][r] += pre * ev(mul(add(*Jb), add(*Jc)), x)
                    out["jsjs"][r] += pre * ev(mul(add(*Jb, w=[1, -1]), add(*Jc, w=[1, -1])), x)
                    out["kk"][r] += pre * ev(mul(add(*Kb), add(*Kc)), x)
                    out["ksks"][r] += pre * ev(mul(add(*Kb, w=[1, -1]), add(*Kc, w=[1, -1])), x)
        if "nematic" in groups:
            xn = Ctx(x.G0t, x.Gtt, x.Gt0, x.G00, t == 0, cross_delta=False)
            for c in range(2 * N):  # the first two bond types (NEM_BONDS, src/meas.c:16)
                j0, j1 = O.bond(c)
                for b in range(2 * N):
                    i0, i1 = O.bond(b)
                    r = p["map_bb"][c][b] + int(p["num_bb"]) * t
                    pre = phase / int(p["degen_bb"])
                    nnnn = ssss = 0.0
                    for s in itertools.product((0, 1), repeat=4):
                        val = ev(mul(O.n(i0, s[0], "t"), O.n(i1, s[1], "t"), O.n(j0, s[2], "0"),
                                     O.n(j1, s[3], "0")), xn)
                        nnnn += val
                        ssss += (-1) ** sum(s) * val
                    out["nem_nnnn"][r] += pre * nnnn
                    out["nem_ssss"][r] += pre * ssss
        if "gensus" in groups:
            nij = int(p["num_ij"])
            _gen_suscept(p, phase / (int(p["Nx"]) * int(p["Ny"])) ** 2, x, out, t, nij)
    return out


def _ue_inputs(N, L, cplx, seed):
    """Random G(0,t), G(t,t), G(t,0) per spin, independent for t > 0; at t = 0
    all three are the one equal-time G(0,0), as calc_ue_g returns them."""
    rs = np.random.RandomState(seed)
    G0t, Gtt, Gt0 = [[_rand(rs, (L, N, N), cplx) for _ in range(2)] for _ in range(3)]
    for s in (0, 1):
        G0t[s][0] = Gtt[s][0]
        Gt0[s][0] = Gtt[s][0]
    return G0t, Gtt, Gt0


@pytest.mark.parametrize("nflux", [0, 1])
def test_unequal_time_default_bond_and_gensus_sets_are_operator_expectations(nflux):
    L = 2
    sim = gen.create_batch(seed=7, Nx=3, Ny=3, L=L, n_matmul=2, period_eqlt=2, period_uneqlt=1,
                           tp=-0.3, nflux=nflux, meas_bond_corr=1, meas_energy_corr=1,
                           meas_gen_suscept=1)[0]
    cplx = nflux != 0
    N = int(sim["params"]["N"])
    G0t, Gtt, Gt0 = _ue_inputs(N, L, cplx, 11)
    phase = (0.6 + 0.8j) if cplx else -1.0
    ref = oracle.measure_uneqlt(sim, phase, G0t[0], Gtt[0], Gt0[0], G0t[1], Gtt[1], Gt0[1])
    mine = _uneqlt_expected(sim, phase, G0t, Gtt, Gt0, L, {"default", "bond", "gensus"})
    for k, v in mine.items():
        _close(ref[k], v, k)


def test_unequal_time_bond_set_without_translation_symmetry():
    """Every bond pair in its own bin (trans_sym = 0): the bin of each pair is pinned."""
    L = 2
    sim = gen.create_batch(seed=8, Nx=3, Ny=3, L=L, n_matmul=2, period_eqlt=2, period_uneqlt=1,
                           tp=-0.3, nflux=1, trans_sym=0, meas_bond_corr=1)[0]
    N = int(sim["params"]["N"])
    G0t, Gtt, Gt0 = _ue_inputs(N, L, True, 12)
    ref = oracle.measure_uneqlt(sim, 1.0, G0t[0], Gtt[0], Gt0[0], G0t[1], Gtt[1], Gt0[1])
    mine = _uneqlt_expected(sim, 1.0, G0t, Gtt, Gt0, L, {"bond"})
    for k, v in mine.items():
        _close(ref[k], v, k)


def test_unequal_time_pair_bb_only():
    L = 2
    sim = gen.create_batch(seed=9, Nx=3, Ny=3, L=L, n_matmul=2, period_eqlt=2, period_uneqlt=1,
                           tp=-0.3, trans_sym=0, meas_pair_bb_only=1)[0]
    N = int(sim["params"]["N"])
    G0t, Gtt, Gt0 = _ue_inputs(N, L, False, 13)
    ref = oracle.measure_uneqlt(sim, -1.0, G0t[0], Gtt[0], Gt0[0], G0t[1], Gtt[1], Gt0[1])
    mine = _uneqlt_expected(sim, -1.0, G0t, Gtt, Gt0, L, {"pair_bb_only"})
    _close(ref["pair_bb"], mine["pair_bb"], "pair_bb")


def test_unequal_time_nematic_is_an_operator_expectation():
    """nem_nnnn / nem_ssss: <n_i0 n_i1 (t) n_j0 n_j1 (0)> summed over the 16
    spin labels (with the sign of the spins for ssss), every delta between
    different bilinears dropped as the reference does (src/meas.c:576-581);
    real builds only (src/meas.c:1369-1375)."""
    L = 2
    sim = gen.create_batch(seed=10, Nx=3, Ny=2, L=L, n_matmul=2, period_eqlt=2, period_uneqlt=1,
                           tp=-0.3, trans_sym=0, meas_nematic_corr=1)[0]
    N = int(sim["params"]["N"])
    G0t, Gtt, Gt0 = _ue_inputs(N, L, False, 14)
    ref = oracle.measure_uneqlt(sim, 1.0, G0t[0], Gtt[0], Gt0[0], G0t[1], Gtt[1], Gt0[1])
    mine = _uneqlt_expected(sim, 1.0, G0t, Gtt, Gt0, L, {"nematic"})
    for k in ("nem_nnnn", "nem_ssss"):
        _close(ref[k], mine[k], k)
