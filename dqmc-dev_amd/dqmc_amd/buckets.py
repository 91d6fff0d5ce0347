"""Chains of uneven size in one engine: padding to a size bucket.

No counterpart in the reference, which runs every chain as its own process
(``src/main_stack.c``); one engine per distinct (N, L) would leave the GPU
idle behind single-chain launches.  Here the chains of a bucket share one
engine of the bucket's N and L:

* sites ``>= N_c`` of chain c are inert: their blocks of every exponential are
  the identity, ``exp_lambda = 1`` and ``del = 0`` (never flipped, no random
  draw), so G keeps the chain's own block and a constant ``1/2`` block;
* slices ``>= L_c`` are identity slices: no update, no wrap, no measurement
  (the engine skips them, ``dqmc_tuning.chain_L``); the chain's groups of
  ``n_matmul`` slices follow the engine's, so its last group may be partial;
* the padding sites and pairs map to no measurement bin (-1), and the bins
  beyond a chain's own count (up to the largest chain's) stay empty; both
  are dropped again by :meth:`PaddedBatch.meas`;
* unequal-time measurements (the default set, one period_uneqlt for the
  bucket): the identity slices follow the chain's last slice, so its G(t,0),
  G(0,t), G(t,t) for t < L_c are its own; the accumulator slices t >= L_c
  and the padding bins are dropped by :meth:`PaddedBatch.meas_uneqlt`.

Every chain keeps its own ``period_eqlt`` (the engine measures each padded
chain at its own cadence), so its accumulators collect the samples a run of
its own file collects.  A chain whose L is a multiple of the engine's
``n_matmul`` runs exactly as it would alone (same random stream,
recalculation points and measurement slices); at another ``n_matmul`` only
the recalculation points move, which changes G by rounding.  Supported: the
default equal- and unequal-time measurements; the optional groups need equal
sizes.
"""
import numpy as np

from . import engine as _eng

_NUM_I = ("density", "density_u", "density_d", "double_occ")
_NUM_IJ = ("g00", "g00_u", "g00_d", "nn", "xx", "zz", "pair_sw")
_UE_IJ = ("gt0", "gt0_u", "gt0_d", "nn", "xx", "zz", "pair_sw")  # [L][num_ij] each
_EXPS = ("exp_Ku", "exp_Kd", "inv_exp_Ku", "inv_exp_Kd", "exp_halfKu", "exp_halfKd",
         "inv_exp_halfKu", "inv_exp_halfKd")
_TOGGLES = ("meas_bond_corr", "meas_thermal", "meas_2bond_corr", "meas_energy_corr",
            "meas_nematic_corr", "meas_chiral", "meas_local_JQ", "meas_pair_bb_only",
            "meas_gen_suscept")
_OPTIONAL_COUNTS = ("num_b", "num_bb", "num_bs", "num_b2", "num_bb2", "num_b2b", "num_b2b2",
                    "num_plaq", "num_plaq_accum", "num_b_accum", "num_b2_accum")


def _int(x):
    return int(np.asarray(x).reshape(()))


def _block(a, n_pad, fill_diag):
    a = np.asarray(a)
    out = np.zeros((n_pad, n_pad), dtype=a.dtype)
    n = a.shape[0]
    out[:n, :n] = a
    if fill_diag is not None:
        out[np.arange(n, n_pad), np.arange(n, n_pad)] = fill_diag
    return out


def pad_sim(sim, N_pad, L_pad, n_matmul, num_i_pad, num_ij_pad):
    """Padded copy of a simulation dict (params, state, meas_eqlt,
    meas_uneqlt); the chain's own period_eqlt is kept."""
    p = sim["params"]
    N, L = _int(p["N"]), _int(p["L"])
    for t in _TOGGLES:
        if _int(p.get(t, 0)):
            raise ValueError("%s needs chains of equal size" % t)
    if N > N_pad or L > L_pad or L_pad % n_matmul:
        raise ValueError("bad padding N=%d->%d L=%d->%d n_matmul=%d" % (N, N_pad, L, L_pad, n_matmul))
    q = dict(p)
    for k in _EXPS:
        q[k] = _block(p[k], N_pad, 1.0)
    for k in ("peierlsu", "peierlsd"):
        q[k] = _block(p[k], N_pad, None)
        q[k][N:, :] = 1.0
        q[k][:, N:] = 1.0
    el = np.ones((2, N_pad))
    el[:, :N] = np.asarray(p["exp_lambda"]).reshape(2, N)
    dl = np.zeros((2, N_pad))
    dl[:, :N] = np.asarray(p["del"]).reshape(2, N)
    q["exp_lambda"], q["del"] = el, dl
    mi = np.full(N_pad, -1, dtype=np.int32)
    mi[:N] = np.asarray(p["map_i"]).reshape(N)
    mij = np.full((N_pad, N_pad), -1, dtype=np.int32)
    mij[:N, :N] = np.asarray(p["map_ij"]).reshape(N, N)
    q["map_i"], q["map_ij"] = mi, mij
    q.update(N=np.int32(N_pad), L=np.int32(L_pad), F=np.int32(L_pad // n_matmul),
             n_matmul=np.int32(n_matmul), num_i=np.int32(num_i_pad), num_ij=np.int32(num_ij_pad))
    for k in _OPTIONAL_COUNTS:
        q[k] = np.int32(0)
    st = dict(sim["state"])
    hs = np.zeros((L_pad, N_pad), dtype=np.int32)
    hs[:L, :N] = np.asarray(sim["state"]["hs"]).reshape(L, N)
    st["hs"] = hs
    meq = {}
    for k, v in sim["meas_eqlt"].items():
        v = np.asarray(v)
        if k in _NUM_I or k in _NUM_IJ:
            out = np.zeros(num_i_pad if k in _NUM_I else num_ij_pad, dtype=v.dtype)
            out[:v.shape[0]] = v
            meq[k] = out
        else:
            meq[k] = v.copy()
    mue = {}
    if _int(p.get("period_uneqlt", 0)) > 0:
        nij = _int(p["num_ij"])
        for k, v in sim["meas_uneqlt"].items():
            v = np.asarray(v)
            if k in _UE_IJ:
                out = np.zeros((L_pad, num_ij_pad), dtype=v.dtype)
                out[:L, :nij] = v.reshape(L, nij)
                mue[k] = out.reshape(-1)
            else:
                mue[k] = v.copy()
    return {"params": q, "state": st, "meas_eqlt": meq, "meas_uneqlt": mue,
            "metadata": sim.get("metadata", {})}


class PaddedBatch:
    """One engine for chains of uneven N and L (see the module docstring).

    sims: simulation dicts of one dtype; n_matmul: the engine's group size
    (the last group of a chain whose L it does not divide is partial);
    graphs, check_drift: as for :class:`dqmc_amd.engine.Engine`."""

    def __init__(self, sims, n_matmul, device=0, gemm_tile=0, graphs=True, check_drift=False,
                 delay_rank=0):
        self.sims = sims
        self.Nc = [_int(s["params"]["N"]) for s in sims]
        self.Lc = [_int(s["params"]["L"]) for s in sims]
        self.N = max(self.Nc)
        self.L = max(-(-l // n_matmul) * n_matmul for l in self.Lc)
        self.n_matmul = n_matmul
        self.num_i = [_int(s["params"]["num_i"]) for s in sims]
        self.num_ij = [_int(s["params"]["num_ij"]) for s in sims]
        if len({_int(s["params"].get("period_uneqlt", 0)) for s in sims}) != 1:
            raise ValueError("the chains of a bucket share period_uneqlt")
        ni, nij = max(self.num_i), max(self.num_ij)
        self.padded = [pad_sim(s, self.N, self.L, n_matmul, ni, nij) for s in sims]
        self.engine = _eng.Engine(self.padded, device=device, gemm_tile=gemm_tile, graphs=graphs,
                                  delay_rank=delay_rank,
                                  check_drift=check_drift,
                                  chain_L=self.Lc, chain_N=self.Nc)

    @property
    def nc(self):
        return len(self.sims)

    def run(self, n_sweeps, stop=None):
        return self.engine.run(n_sweeps, stop)

    def close(self):
        self.engine.close()

    def state(self, c):
        """Chain c's state in its own shape."""
        st = self.engine.get_state(c)
        st["hs"] = st["hs"][:self.Lc[c], :self.Nc[c]].copy()
        return st

    def meas(self, c):
        """Chain c's equal-time accumulators in its own bins."""
        eq, _ = self.engine.get_meas(c)
        out = {}
        for k, v in eq.items():
            if k in _NUM_I:
                out[k] = v[:self.num_i[c]].copy()
            elif k in _NUM_IJ:
                out[k] = v[:self.num_ij[c]].copy()
            else:
                out[k] = v
        return out

    def meas_uneqlt(self, c):
        """Chain c's unequal-time accumulators in its own slices and bins."""
        _, ue = self.engine.get_meas(c)
        out = {}
        nij_pad = max(self.num_ij)
        for k, v in ue.items():
            if k in _UE_IJ:
                out[k] = v.reshape(self.L, nij_pad)[:self.Lc[c], :self.num_ij[c]].reshape(-1).copy()
            else:
                out[k] = v
        return out

    def green(self, c, spin):
        """Chain c's current equal-time G (its own block) and phase."""
        G, ph = self.engine.green(c, spin)
        n = self.Nc[c]
        return G[:n, :n].copy(), ph


def tune_n_matmul(sims, candidates=(5, 8, 10, 16, 20), tol=1e-3, sweeps=2, device=0,
                  graphs=True):
    """Stabilization interval for one bucket of chains (no counterpart in the
    reference, where n_matmul is an input).

    For every candidate n no larger than the shortest chain's L: a trial
    engine times `sweeps` sweeps after a first one, and a second trial engine
    with the wrap drift check (the reference's CHECK_G_WRP,
    ``src/dqmc.c:277-287``) runs two sweeps (the drift grows as the field
    leaves its random start) and then reports the largest
    |G_wrapped - G_recomputed| of one more sweep over the chains.  Candidates
    whose drift exceeds `tol` (absolute on G) are dropped and the fastest
    remaining one is returned, with the table of (n, seconds per sweep,
    drift).  If none passes, the smallest candidate is returned.  The default
    tolerance is the drift of the reference's own headline setting
    (n_matmul = 10 at U = 6, beta = 8, 16x16: 1e-4 to 1e-3 once equilibrated,
    for the reference algorithm and this engine alike).  The trial chains
    are copies: `sims` is unchanged."""
    import time
    lmin = min(_int(s["params"]["L"]) for s in sims)
    cands = [n for n in candidates if n <= lmin] or [min(candidates)]
    table = []
    for n in cands:
        pb = PaddedBatch(sims, n, device=device, graphs=graphs)
        try:
            pb.run(1)
            t0 = time.perf_counter()
            pb.run(sweeps)
            dt = (time.perf_counter() - t0) / sweeps
        finally:
            pb.close()
        pb = PaddedBatch(sims, n, device=device, check_drift=True, graphs=graphs)
        try:
            pb.run(2)
            pb.engine.drift()
            pb.run(1)
            drift = float(np.max(pb.engine.drift()))
        finally:
            pb.close()
        table.append((n, dt, drift))
    ok = [r for r in table if r[2] <= tol]
    pick = min(ok, key=lambda r: r[1])[0] if ok else min(cands)
    return pick, table


def tune_delay_rank(sims, n_matmul=None, candidates=(4, 8, 16, 32), sweeps=2, device=0,
                    graphs=True):
    """Delay rank of the Metropolis kernel for one bucket of chains (the
    reference takes n_delay from the input file, ``src/updates.c:7``).

    Every candidate is rounded to a compiled variant of the engine's kernel
    (Engine.delay_rank); each distinct rank gets a trial engine on copies of
    the chains (`sims` is unchanged) that times `sweeps` sweeps after a first
    one.  The rank changes only the rounding of G, not the chain, so the
    fastest one is returned, with the table of (rank, seconds per sweep).
    n_matmul: the bucket's stabilization interval (default: the chains' own;
    chains of one shape at their own interval run unpadded, with every
    measurement they carry)."""
    import time
    own = {(_int(s["params"]["N"]), _int(s["params"]["L"]), _int(s["params"]["n_matmul"]))
           for s in sims}
    if n_matmul is None:
        n_matmul = _int(sims[0]["params"]["n_matmul"])
    plain = len(own) == 1 and next(iter(own))[2] == n_matmul
    table, seen = [], set()
    for want in candidates:
        if plain:
            eng = _eng.Engine(sims, device=device, graphs=graphs, delay_rank=want)
        else:
            eng = PaddedBatch(sims, n_matmul, device=device, graphs=graphs,
                              delay_rank=want).engine
        try:
            k = eng.delay_rank()
            if k in seen:
                continue
            seen.add(k)
            eng.run(1)
            t0 = time.perf_counter()
            eng.run(sweeps)
            table.append((k, (time.perf_counter() - t0) / sweeps))
        finally:
            eng.close()
    return min(table, key=lambda r: r[1])[0], table
