"""dqmc_calc_eq_g_noalloc (calc_eq_g on caller-held group products, work space
in the library's static device array) against dqmc_engine_calc_eq_g (the same
stabilized G, computed by an engine in its own buffers from the same field).

Two calls in a row on one GPU, the second with a smaller input (fewer chains
and a smaller N, so the static work space is reused at other strides), for
the real and the complex build.  Bar: G within 1e-10 relative max-norm and
the phase within 1e-10, the equal-time G bar of the parity tests.  The group
products are formed in numpy, so they differ from the engine's own GEMM
products by rounding only.

Device buffers for the inputs and outputs come from the HIP runtime that the
engine library itself loaded (hipMalloc through ctypes)."""
import ctypes as C

import numpy as np
import pytest

from dqmc_amd import engine, gen

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip():
    engine.lib()
    path = None
    with open("/proc/self/maps") as f:
        for line in f:
            parts = line.split()
            if len(parts) >= 6 and "libamdhip64.so" in parts[-1]:
                if path is None or "/opt/rocm" in parts[-1]:
                    path = parts[-1]
    assert path, "libamdhip64 not loaded by the engine library"
    h = C.CDLL(path)
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
    h.hipStreamSynchronize.argtypes = [C.c_void_p]
    h.hipStreamDestroy.argtypes = [C.c_void_p]
    for f in ("hipMalloc", "hipFree", "hipMemcpy", "hipStreamCreate", "hipStreamSynchronize",
              "hipStreamDestroy", "hipDeviceSynchronize"):
        getattr(h, f).restype = C.c_int
    return h


def _group_products(sims):
    """[2 chains][F][N][N] raw (column-major) buffers of C_f = B_{(f+1)n-1}..B_{fn},
    B_l = expK diag(exp_lambda[s_l]) as the reference builds it (src/dqmc.c:29-43)."""
    p = sims[0]["params"]
    N, L, n = int(p["N"]), int(p["L"]), int(p["n_matmul"])
    F = L // n
    cplx = p["exp_Ku"].dtype.kind == "c"
    out = np.zeros((2 * len(sims), F, N, N), dtype=np.complex128 if cplx else np.float64)
    el = p["exp_lambda"]
    for c, s in enumerate(sims):
        hs = s["state"]["hs"]
        for spin in (0, 1):
            expK = (p["exp_Kd"] if spin else p["exp_Ku"]).T  # raw buffer reading
            for f in range(F):
                P = np.eye(N, dtype=out.dtype)
                for l in range(f * n, (f + 1) * n):
                    sl = (1 - hs[l]) if spin else hs[l]
                    P = (expK * el[sl, np.arange(N)][None, :]) @ P
                out[2 * c + spin, f] = P.T  # column-major storage
    return out


def _noalloc(h, L, Cm, N, F, nc, f, stream):
    cplx = Cm.dtype.kind == "c"
    es = 16 if cplx else 8
    bufs = [C.c_void_p() for _ in range(3)]
    sizes = [Cm.nbytes, 2 * nc * N * N * es, 2 * nc * es]
    for b, sz in zip(bufs, sizes):
        assert h.hipMalloc(C.byref(b), sz) == 0
    try:
        src = np.ascontiguousarray(Cm)
        assert h.hipMemcpy(bufs[0], src.ctypes.data, sizes[0], H2D) == 0
        dp = lambda b: C.cast(b, C.POINTER(C.c_double))
        rc = L.dqmc_calc_eq_g_noalloc(int(cplx), N, F, nc, f, dp(bufs[0]), dp(bufs[1]),
                                      dp(bufs[2]), stream)
        assert rc == 0, L.dqmc_last_error().decode()
        assert (h.hipStreamSynchronize(stream) if stream else h.hipDeviceSynchronize()) == 0
        G = np.zeros((2 * nc, N, N), dtype=Cm.dtype)
        ph = np.zeros(2 * nc, dtype=Cm.dtype)
        assert h.hipMemcpy(G.ctypes.data, bufs[1], sizes[1], D2H) == 0
        assert h.hipMemcpy(ph.ctypes.data, bufs[2], sizes[2], D2H) == 0
        return G.transpose(0, 2, 1), ph
    finally:
        for b in bufs:
            h.hipFree(b)


def _check_call(h, L, nflux, Nx, nc, Lt, nm, f, seed, stream):
    sims = gen.create_batch(Nfiles=nc, seed=seed, Nx=Nx, Ny=Nx, nflux=nflux, U=4.0, mu=-0.3,
                            tp=-0.1, L=Lt, n_matmul=nm, period_eqlt=nm)
    N, F = Nx * Nx, Lt // nm
    print("noalloc call: nflux", nflux, "N", N, "chains", nc, "f", f, flush=True)
    Cm = _group_products(sims)
    G, ph = _noalloc(h, L, Cm, N, F, nc, f, stream)
    eng = engine.Engine(sims)
    try:
        for c in range(nc):
            for spin in (0, 1):
                Ge, phe = eng.calc_eq_g(c, spin, f)
                b = 2 * c + spin
                err = np.abs(G[b] - Ge).max() / np.abs(Ge).max()
                assert err < 1e-10, (c, spin, err)
                assert abs(ph[b] - phe) < 1e-10, (c, spin, ph[b], phe)
    finally:
        eng.close()


@pytest.mark.parametrize("nflux", [0, 1])
def test_noalloc_calc_eq_g_matches_engine_twice(nflux):
    h = _hip()
    L = engine.lib()
    P = C.POINTER(C.c_double)
    L.dqmc_calc_eq_g_noalloc.argtypes = [C.c_int] * 5 + [P, P, P, C.c_void_p]
    L.dqmc_calc_eq_g_noalloc.restype = C.c_int
    st = C.c_void_p()
    assert h.hipStreamCreate(C.byref(st)) == 0
    try:
        # first call: 3 chains of N = 64 on a created stream, a cyclic product
        _check_call(h, L, nflux, 8, 3, 40, 8, 4, 17, st)
        # second call: fewer chains, smaller N, on the null stream
        _check_call(h, L, nflux, 6, 2, 20, 5, 1, 29, None)
    finally:
        h.hipStreamDestroy(st)


def test_noalloc_rejects_requests_beyond_the_scratch():
    L = engine.lib()
    P = C.POINTER(C.c_double)
    L.dqmc_calc_eq_g_noalloc.argtypes = [C.c_int] * 5 + [P, P, P, C.c_void_p]
    L.dqmc_calc_eq_g_noalloc.restype = C.c_int
    dummy = (C.c_double * 1)()
    p = C.cast(dummy, P)
    # 6 real chains of N = 256 need about 70 MiB of work space: refused on the
    # host before any device work (the pointers are never dereferenced)
    rc = L.dqmc_calc_eq_g_noalloc(0, 256, 8, 6, 0, p, p, p, None)
    assert rc == -10
    assert b"work space" in L.dqmc_last_error()
