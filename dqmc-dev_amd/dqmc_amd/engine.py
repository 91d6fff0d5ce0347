"""Python binding of libdqmc_hip.so (C ABI in include/dqmc_hip.h).

The engine is the GPU replacement of the reference's `dqmc()` sweep engine
(src/dqmc.c:84-511); this module only marshals simulation dicts (the .h5
groups as numpy arrays) across the C ABI.  It never falls back to a CPU path:
if the shared library or a HIP device is missing, it raises.
"""
import ctypes as C
import os

import numpy as np

from . import capi

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libdqmc_hip.so")
_lib = None


class DqmcTuning(C.Structure):
    _fields_ = [("delay_rank", C.c_int), ("flags", C.c_int), ("gemm_tile", C.c_int),
                ("chain_L", C.POINTER(C.c_int)), ("chain_N", C.POINTER(C.c_int))]


def probe_mfma(form):
    """Sustained FP64 matrix-core TFLOP/s of the current device (form 0:
    16x16x4 MFMA, 1: 4x4x4 MFMA), from a register-only loop."""
    out = C.c_double(0.0)
    if lib().dqmc_probe_mfma(int(form), C.byref(out)) != 0:
        raise RuntimeError(lib().dqmc_last_error().decode())
    return out.value


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libdqmc_hip.so not built (%s); run __graft_entry__.build()" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        P = C.POINTER
        L.dqmc_engine_create.argtypes = [P(C.c_void_p), P(P(capi.DqmcParams)), C.c_int, C.c_int,
                                         P(DqmcTuning)]
        L.dqmc_engine_create.restype = C.c_int
        L.dqmc_engine_destroy.argtypes = [C.c_void_p]
        L.dqmc_engine_destroy.restype = None
        L.dqmc_engine_set_state.argtypes = [C.c_void_p, C.c_int, P(C.c_uint64), C.c_int, P(C.c_int32)]
        L.dqmc_engine_get_state.argtypes = [C.c_void_p, C.c_int, P(C.c_uint64), P(C.c_int),
                                            P(C.c_int32)]
        L.dqmc_engine_set_meas.argtypes = [C.c_void_p, C.c_int, P(capi.DqmcMeas), P(capi.DqmcMeas)]
        L.dqmc_engine_get_meas.argtypes = [C.c_void_p, C.c_int, P(capi.DqmcMeas), P(capi.DqmcMeas)]
        L.dqmc_engine_run.argtypes = [C.c_void_p, C.c_int, P(C.c_int)]
        L.dqmc_engine_green.argtypes = [C.c_void_p, C.c_int, C.c_int, P(C.c_double), P(C.c_double)]
        L.dqmc_engine_calc_eq_g.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, P(C.c_double),
                                            P(C.c_double)]
        L.dqmc_engine_calc_ue_g.argtypes = [C.c_void_p, C.c_int, C.c_int, P(C.c_double),
                                            P(C.c_double), P(C.c_double)]
        L.dqmc_engine_accepts.argtypes = [C.c_void_p]
        L.dqmc_engine_accepts.restype = C.c_int64
        L.dqmc_engine_gemm_stats.argtypes = [C.c_void_p, P(C.c_double), P(C.c_int64), P(C.c_double)]
        L.dqmc_last_error.restype = C.c_char_p
        L.dqmc_version.restype = C.c_char_p
        L.dqmc_acc_len.argtypes = [P(capi.DqmcParams), C.c_int, C.c_int]
        L.dqmc_acc_len.restype = C.c_long
        L.dqmc_engine_gemm_tile.argtypes = [C.c_void_p]
        L.dqmc_engine_gemm_tile.restype = C.c_int
        L.dqmc_engine_delay_rank.argtypes = [C.c_void_p]
        L.dqmc_engine_delay_rank.restype = C.c_int
        L.dqmc_probe_mfma.argtypes = [C.c_int, P(C.c_double)]
        L.dqmc_probe_mfma.restype = C.c_int
        L.dqmc_engine_drift.argtypes = [C.c_void_p, P(C.c_double)]
        L.dqmc_engine_drift.restype = C.c_int
        L.dqmc_engine_set_params.argtypes = [C.c_void_p, C.c_int, P(capi.DqmcParams)]
        L.dqmc_engine_set_params.restype = C.c_int
        L.dqmc_engine_profile.argtypes = [C.c_void_p, P(C.c_char_p), P(C.c_double), P(C.c_int64),
                                          C.c_int]
        L.dqmc_engine_profile.restype = C.c_int
        L.dqmc_engine_slice_accepts.argtypes = [C.c_void_p, C.c_int, P(C.c_int64)]
        L.dqmc_engine_slice_accepts.restype = C.c_int
        for f in ("set_state", "get_state", "set_meas", "get_meas", "run", "green", "calc_eq_g",
                  "calc_ue_g", "gemm_stats"):
            getattr(L, "dqmc_engine_" + f).restype = C.c_int
        _lib = L
    return _lib


class DqmcError(RuntimeError):
    pass


def _check(rc, what):
    if rc < 0:
        raise DqmcError("%s failed (%d): %s" % (what, rc, lib().dqmc_last_error().decode()))
    return rc


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Engine:
    """A batch of chains of one shape on one GPU.

    sims: list of simulation dicts ({"params", "state", "meas_eqlt",
    "meas_uneqlt"}) -- all with the same N, L, n_matmul, dtype and
    measurement set.  State and accumulators are uploaded at construction.

    graphs: capture each sweep into a HIP graph (one per measurement pattern)
    and replay it.  HIP's stream capture fails while other host threads make
    HIP calls, so engines driven concurrently from several threads need
    graphs=False.  Engines that time their GEMMs (time_gemm) run eagerly.

    check_drift: at every recalculation also wrap G and compare it with the
    recomputed one (the reference's CHECK_G_WRP, src/dqmc.c:277-287); read
    with :meth:`drift`."""

    def __init__(self, sims, device=0, delay_rank=0, time_gemm=False, gemm_tile=0, chain_L=None,
                 chain_N=None, graphs=True, check_drift=False, profile=False):
        self.L = lib()
        self.sims = sims
        self.cplx = capi.is_complex(sims[0])
        self.N = int(sims[0]["params"]["N"])
        self.Lt = int(sims[0]["params"]["L"])
        self._keep = []
        structs = []
        for s in sims:
            ps, keep = capi.make_params(s["params"])
            self._keep.append((ps, keep))
            structs.append(C.pointer(ps))
        arr = (C.POINTER(capi.DqmcParams) * len(sims))(*structs)
        self._arr = arr
        # own sizes of chains padded to the common N, L (dqmc_amd.buckets)
        self._cl = None if chain_L is None else (C.c_int * len(sims))(*chain_L)
        self._cn = None if chain_N is None else (C.c_int * len(sims))(*chain_N)
        flags = ((1 if time_gemm else 0) | (0 if graphs else 2) | (4 if check_drift else 0) |
                 (8 if profile else 0))
        tune = DqmcTuning(delay_rank, flags, gemm_tile,
                          C.cast(self._cl, C.POINTER(C.c_int)) if self._cl is not None else None,
                          C.cast(self._cn, C.POINTER(C.c_int)) if self._cn is not None else None)
        h = C.c_void_p()
        _check(self.L.dqmc_engine_create(C.byref(h), arr, len(sims), device, C.byref(tune)),
               "dqmc_engine_create")
        self.h = h
        for c, s in enumerate(sims):
            self.set_state(c, s["state"])
            self.set_meas(c, s["meas_eqlt"], s.get("meas_uneqlt") or None)

    def drift(self):
        """Per chain, the largest |G_wrapped - G_recomputed| over both spins
        and all recalculations since the previous call (needs check_drift)."""
        out = np.zeros(len(self.sims))
        _check(self.L.dqmc_engine_drift(self.h, out.ctypes.data_as(C.POINTER(C.c_double))),
               "dqmc_engine_drift")
        return out

    def close(self):
        if getattr(self, "h", None):
            self.L.dqmc_engine_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def nc(self):
        return len(self.sims)

    def set_params(self, c, params):
        """Replace chain c's numeric parameters (exponentials, exp_lambda, del,
        Peierls phases) with those of a params dict of the same shape; c may
        be a sequence of chains that all get these parameters (every chain:
        one upload, copied on the device)."""
        ps, keep = capi.make_params(params)
        cs = [int(k) for k in (c if hasattr(c, "__iter__") else (c,))]
        every = sorted(cs) == list(range(len(self.sims)))
        for k in ([-1] if every else cs):
            _check(self.L.dqmc_engine_set_params(self.h, k, C.byref(ps)),
                   "dqmc_engine_set_params")
        for k in cs:
            self._keep[k] = (ps, keep)

    def set_state(self, c, st):
        rng = np.ascontiguousarray(st["rng"], dtype=np.uint64)
        hs = np.ascontiguousarray(st["hs"], dtype=np.int32)
        _check(self.L.dqmc_engine_set_state(self.h, c, rng.ctypes.data_as(C.POINTER(C.c_uint64)),
                                            int(st["sweep"]),
                                            hs.ctypes.data_as(C.POINTER(C.c_int32))),
               "set_state")

    def get_state(self, c):
        rng = np.zeros(17, dtype=np.uint64)
        hs = np.zeros((self.Lt, self.N), dtype=np.int32)
        sw = C.c_int(0)
        _check(self.L.dqmc_engine_get_state(self.h, c, rng.ctypes.data_as(C.POINTER(C.c_uint64)),
                                            C.byref(sw), hs.ctypes.data_as(C.POINTER(C.c_int32))),
               "get_state")
        return {"rng": rng, "sweep": np.array(sw.value, dtype=np.int32), "hs": hs}

    def set_meas(self, c, eq, ue=None):
        me, _ = capi.make_meas(eq, capi.EQ_ACC, self.cplx)
        if ue:
            mu, _u = capi.make_meas(ue, capi.UE_ACC, self.cplx)
            self._tmp = _u
        mp = C.byref(mu) if ue else None
        self._tmp2 = _
        _check(self.L.dqmc_engine_set_meas(self.h, c, C.byref(me), mp), "set_meas")

    def get_meas(self, c):
        s = self.sims[c]
        me, ae = capi.make_meas(s["meas_eqlt"], capi.EQ_ACC, self.cplx)
        ue = s.get("meas_uneqlt") or None
        if ue:
            mu, au = capi.make_meas(ue, capi.UE_ACC, self.cplx)
        _check(self.L.dqmc_engine_get_meas(self.h, c, C.byref(me), C.byref(mu) if ue else None),
               "get_meas")
        out_e = capi.meas_to_dict(me, ae, self.cplx)
        out_u = capi.meas_to_dict(mu, au, self.cplx) if ue else {}
        return out_e, out_u

    def run(self, n_sweeps, stop=None):
        flag = stop if stop is not None else C.c_int(0)
        return _check(self.L.dqmc_engine_run(self.h, n_sweeps, C.byref(flag)), "run")

    def _mat(self, n=1):
        dt = np.complex128 if self.cplx else np.float64
        return np.zeros((n, self.N, self.N), dtype=dt)

    def green(self, c, spin):
        """Current equal-time G of (chain, spin) as a (row, col) matrix and the phase."""
        G = self._mat()
        ph = np.zeros(2)
        _check(self.L.dqmc_engine_green(self.h, c, spin, _dp(G), _dp(ph)), "green")
        return G[0].T.copy(), (complex(ph[0], ph[1]) if self.cplx else ph[0])

    def calc_eq_g(self, c, spin, f=0):
        G = self._mat()
        ph = np.zeros(2)
        _check(self.L.dqmc_engine_calc_eq_g(self.h, c, spin, f, _dp(G), _dp(ph)), "calc_eq_g")
        return G[0].T.copy(), (complex(ph[0], ph[1]) if self.cplx else ph[0])

    def calc_ue_g(self, c, spin):
        arrs = [self._mat(self.Lt) for _ in range(3)]
        _check(self.L.dqmc_engine_calc_ue_g(self.h, c, spin, *[_dp(a) for a in arrs]), "calc_ue_g")
        return tuple(a.transpose(0, 2, 1).copy() for a in arrs)

    def accepts(self):
        return int(self.L.dqmc_engine_accepts(self.h))

    def slice_accepts(self, c):
        """Accepted flips of chain c per slice, summed over all sweeps run."""
        out = np.zeros(self.Lt, dtype=np.int64)
        _check(self.L.dqmc_engine_slice_accepts(self.h, c, out.ctypes.data_as(C.POINTER(C.c_int64))),
               "dqmc_engine_slice_accepts")
        return out

    def profile(self):
        """{section: (ms, calls)} in the reference's profile order (needs profile=True)."""
        names = (C.c_char_p * 16)()
        ms = np.zeros(16)
        calls = np.zeros(16, dtype=np.int64)
        n = _check(self.L.dqmc_engine_profile(self.h, names, ms.ctypes.data_as(C.POINTER(C.c_double)),
                                              calls.ctypes.data_as(C.POINTER(C.c_int64)), 16),
                   "dqmc_engine_profile")
        return {names[k].decode(): (float(ms[k]), int(calls[k])) for k in range(n)}

    def gemm_tile(self):
        """GEMM C-tile shape in use (1 = 64x64, 2 = 128x64 with K steps of 16,
        3 = 128x64 with K steps of 8)."""
        return self.L.dqmc_engine_gemm_tile(self.h)

    def delay_rank(self):
        """Delay rank of the Metropolis kernel in use (the params' n_delay or
        the requested rank, rounded to a compiled variant)."""
        return self.L.dqmc_engine_delay_rank(self.h)

    def gemm_stats(self):
        ms = C.c_double(0)
        n = C.c_int64(0)
        fl = C.c_double(0)
        _check(self.L.dqmc_engine_gemm_stats(self.h, C.byref(ms), C.byref(n), C.byref(fl)),
               "gemm_stats")
        return ms.value, n.value, fl.value

    def result_sims(self):
        """Simulation dicts with the engine's current state and accumulators."""
        out = []
        for c, s in enumerate(self.sims):
            st = dict(s["state"])
            st.update(self.get_state(c))
            eq, ue = self.get_meas(c)
            out.append({"params": s["params"], "state": st, "meas_eqlt": eq, "meas_uneqlt": ue})
        return out
