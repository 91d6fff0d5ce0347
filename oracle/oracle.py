"""ctypes binding of the CPU oracle (oracle/lib/liboracle_{r,c}.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The engine never imports it.
"""
import ctypes as C
import os
import sys

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_PKG = os.path.join(os.path.dirname(_HERE), "dqmc-dev_amd")
if _PKG not in sys.path:
    sys.path.insert(0, _PKG)

from dqmc_amd import capi  # noqa: E402

_libs = {}




def lib(cplx):
    key = "c" if cplx else "r"
    if key not in _libs:
        path = os.path.join(_HERE, "lib", "liboracle_%s.so" % key)
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle` (%s missing)" % path)
        # RTLD_GLOBAL: MKL's CPU-dispatch libraries (libmkl_avx512.so, ...) resolve
        # symbols of libmkl_core / libmkl_sequential from the global scope.
        L = C.CDLL(path, mode=C.RTLD_GLOBAL)
        pre = "oracle_%s_" % key
        P = C.POINTER
        f = getattr(L, pre + "dqmc")
        f.argtypes = [P(capi.DqmcParams), P(C.c_uint64), P(C.c_int), P(C.c_int32),
                      P(capi.DqmcMeas), P(capi.DqmcMeas), C.c_int, P(C.c_int64), P(C.c_double)]
        f.restype = C.c_int
        f = getattr(L, pre + "calc_eq_g")
        f.argtypes = [P(capi.DqmcParams), P(C.c_int32), C.c_int, C.c_int, P(C.c_double),
                      P(C.c_double)]
        f.restype = C.c_int
        f = getattr(L, pre + "calc_eq_g_fine")
        f.argtypes = [P(capi.DqmcParams), P(C.c_int32), C.c_int, C.c_int, P(C.c_double),
                      P(C.c_double)]
        f.restype = C.c_int
        f = getattr(L, pre + "calc_ue_g")
        f.argtypes = [P(capi.DqmcParams), P(C.c_int32), C.c_int, P(C.c_double), P(C.c_double),
                      P(C.c_double)]
        f.restype = C.c_int
        f = getattr(L, pre + "measure_eqlt")
        f.argtypes = [P(capi.DqmcParams), P(C.c_double), P(C.c_double), P(C.c_double),
                      P(capi.DqmcMeas)]
        f.restype = None
        f = getattr(L, pre + "measure_uneqlt")
        f.argtypes = [P(capi.DqmcParams), P(C.c_double)] + [P(C.c_double)] * 6 + [P(capi.DqmcMeas)]
        f.restype = None
        f = getattr(L, pre + "rand_uint")
        f.argtypes = [P(C.c_uint64), C.c_int, P(C.c_uint64)]
        f.restype = None
        f = getattr(L, pre + "shuffle")
        f.argtypes = [P(C.c_uint64), C.c_int, P(C.c_int32)]
        f.restype = None
        _libs[key] = (L, pre)
    return _libs[key]


def _fn(cplx, name):
    L, pre = lib(cplx)
    return getattr(L, pre + name)


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def rand_uint(state, n):
    st = np.ascontiguousarray(state, dtype=np.uint64).copy()
    out = np.zeros(n, dtype=np.uint64)
    _fn(False, "rand_uint")(st.ctypes.data_as(C.POINTER(C.c_uint64)), n,
                            out.ctypes.data_as(C.POINTER(C.c_uint64)))
    return out, st


def shuffle(state, n):
    st = np.ascontiguousarray(state, dtype=np.uint64).copy()
    a = np.zeros(n, dtype=np.int32)
    _fn(False, "shuffle")(st.ctypes.data_as(C.POINTER(C.c_uint64)), n,
                          a.ctypes.data_as(C.POINTER(C.c_int32)))
    return a, st


def run(sim, max_sweeps=None):
    """Run dqmc() on a simulation dict (not modified); returns a new dict with
    updated state / meas_eqlt / meas_uneqlt plus 'n_acc' and 'phase'."""
    cplx = capi.is_complex(sim)
    ps, keep = capi.make_params(sim["params"])
    st = sim["state"]
    rng = np.ascontiguousarray(st["rng"], dtype=np.uint64).copy()
    sweep = C.c_int(int(st["sweep"]))
    hs = np.ascontiguousarray(st["hs"], dtype=np.int32).copy()
    meq, aeq = capi.make_meas(sim["meas_eqlt"], capi.EQ_ACC, cplx)
    mue, aue = capi.make_meas(sim.get("meas_uneqlt") or {}, capi.UE_ACC, cplx)
    nacc = C.c_int64(0)
    ph = np.zeros(2)
    n = max_sweeps if max_sweeps is not None else int(sim["params"]["n_sweep"])
    rc = _fn(cplx, "dqmc")(C.byref(ps), rng.ctypes.data_as(C.POINTER(C.c_uint64)),
                           C.byref(sweep), hs.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(meq),
                           C.byref(mue), n, C.byref(nacc), _dp(ph))
    if rc < 0:
        raise MemoryError("oracle allocation failed")
    out = {"params": sim["params"],
           "state": dict(st, rng=rng, sweep=np.array(sweep.value, dtype=np.int32), hs=hs),
           "meas_eqlt": capi.meas_to_dict(meq, aeq, cplx),
           "meas_uneqlt": capi.meas_to_dict(mue, aue, cplx) if aue else {},
           "n_acc": int(nacc.value), "phase": complex(ph[0], ph[1]) if cplx else ph[0],
           "sweeps_run": rc}
    return out


def calc_eq_g(sim, spin, f=0, hs=None):
    cplx = capi.is_complex(sim)
    ps, keep = capi.make_params(sim["params"])
    N = int(sim["params"]["N"])
    hs = np.ascontiguousarray(sim["state"]["hs"] if hs is None else hs, dtype=np.int32)
    G = np.zeros((N, N), dtype=np.complex128 if cplx else np.float64)  # column-major view: G.T
    ph = np.zeros(2)
    _fn(cplx, "calc_eq_g")(C.byref(ps), hs.ctypes.data_as(C.POINTER(C.c_int32)), spin, f, _dp(G),
                           _dp(ph))
    return G.T.copy(), (complex(ph[0], ph[1]) if cplx else ph[0])


def calc_eq_g_fine(sim, spin, l=0, hs=None):
    """G at slice boundary l with a QR after every slice (CHECK_G_ACC)."""
    cplx = capi.is_complex(sim)
    ps, keep = capi.make_params(sim["params"])
    N = int(sim["params"]["N"])
    hs = np.ascontiguousarray(sim["state"]["hs"] if hs is None else hs, dtype=np.int32)
    G = np.zeros((N, N), dtype=np.complex128 if cplx else np.float64)
    ph = np.zeros(2)
    _fn(cplx, "calc_eq_g_fine")(C.byref(ps), hs.ctypes.data_as(C.POINTER(C.c_int32)), spin, l,
                                _dp(G), _dp(ph))
    return G.T.copy(), (complex(ph[0], ph[1]) if cplx else ph[0])


def calc_ue_g(sim, spin, hs=None):
    """Returns (G0t, Gtt, Gt0) as arrays [L, N, N] of ordinary (row, col) matrices."""
    cplx = capi.is_complex(sim)
    ps, keep = capi.make_params(sim["params"])
    N = int(sim["params"]["N"])
    L = int(sim["params"]["L"])
    hs = np.ascontiguousarray(sim["state"]["hs"] if hs is None else hs, dtype=np.int32)
    dt = np.complex128 if cplx else np.float64
    arrs = [np.zeros((L, N, N), dtype=dt) for _ in range(3)]
    _fn(cplx, "calc_ue_g")(C.byref(ps), hs.ctypes.data_as(C.POINTER(C.c_int32)), spin,
                           *[_dp(a) for a in arrs])
    return tuple(a.transpose(0, 2, 1).copy() for a in arrs)


def measure_eqlt(sim, phase, gu, gd, group=None):
    """One measure_eqlt on (row, col) matrices gu, gd; returns the accumulator dict."""
    cplx = capi.is_complex(sim)
    ps, keep = capi.make_params(sim["params"])
    m, arrs = capi.make_meas(group if group is not None else sim["meas_eqlt"], capi.EQ_ACC, cplx)
    ph = np.array([complex(phase).real, complex(phase).imag])
    dt = np.complex128 if cplx else np.float64
    gu_c = np.ascontiguousarray(np.asarray(gu, dtype=dt).T)
    gd_c = np.ascontiguousarray(np.asarray(gd, dtype=dt).T)
    _fn(cplx, "measure_eqlt")(C.byref(ps), _dp(ph), _dp(gu_c), _dp(gd_c), C.byref(m))
    return capi.meas_to_dict(m, arrs, cplx)


def measure_uneqlt(sim, phase, Gu0t, Gutt, Gut0, Gd0t, Gdtt, Gdt0, group=None):
    cplx = capi.is_complex(sim)
    ps, keep = capi.make_params(sim["params"])
    m, arrs = capi.make_meas(group if group is not None else sim["meas_uneqlt"], capi.UE_ACC,
                             cplx)
    ph = np.array([complex(phase).real, complex(phase).imag])
    dt = np.complex128 if cplx else np.float64
    cs = [np.ascontiguousarray(np.asarray(a, dtype=dt).transpose(0, 2, 1))
          for a in (Gu0t, Gutt, Gut0, Gd0t, Gdtt, Gdt0)]
    _fn(cplx, "measure_uneqlt")(C.byref(ps), _dp(ph), *[_dp(a) for a in cs], C.byref(m))
    return capi.meas_to_dict(m, arrs, cplx)
