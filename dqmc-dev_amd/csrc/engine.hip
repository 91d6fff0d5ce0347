// engine.hip -- the batched DQMC engine behind the C ABI of include/dqmc_hip.h.
//
// Replaces `static int dqmc(struct sim_data *)` (src/dqmc.c:84-511) for a
// batch of independent chains of one shape.  All chains advance in lock step:
// every launch covers all chains and both spins (batch index b = 2*chain +
// spin), so small-N work still fills the GPU when many chains are present.
// Per sweep and slice l the launch sequence mirrors src/dqmc.c:247-386:
//   k_update (Metropolis + delayed updates, src/updates.c:7-84)
//   recalc slice:  C_f from the HS scalings (GEMM chain) + stabilized G
//                  (QRP -> [GEMM, pre-pivot, QR, GEMM]* -> LU solve + phase)
//   other slices:  wrap G <- B_l G B_l^-1 (2 GEMMs with fused HS scalings)
//   eq measurement: half-wrap (2 GEMMs) + contraction kernel
// and once per measured sweep the unequal-time block (src/dqmc.c:389-458).
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <chrono>
#include <mutex>
#include <string>
#include <vector>

#include "dq_common.hpp"
#include "dq_factor.hpp"
#include "dq_gemm.hpp"
#include "dq_meas.hpp"
#include "dq_meas2.hpp"
#include "dq_adjoint.hpp"
#include "dq_qrp.hpp"
#include "dq_qr_blocked.hpp"
#include "dq_rng.hpp"
#include "dq_ue.hpp"
#include "dq_update.hpp"
#include "dq_update2.hpp"
#include "dq_update3.hpp"
#include "dqmc_hip.h"

static thread_local std::string g_err;

static int set_err(int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIPCHK(x)                                                                        \
    do {                                                                                 \
        hipError_t e_ = (x);                                                             \
        if (e_ != hipSuccess)                                                            \
            return set_err(DQMC_ERR_HIP, "%s:%d %s: %s", __FILE__, __LINE__, #x,         \
                           hipGetErrorString(e_));                                       \
    } while (0)

#define LCHK()                                                                          \
    do {                                                                                \
        hipError_t e_ = hipGetLastError();                                              \
        if (e_ != hipSuccess)                                                           \
            return set_err(DQMC_ERR_HIP, "%s:%d launch: %s", __FILE__, __LINE__,         \
                           hipGetErrorString(e_));                                      \
    } while (0)

#define RC(x)                 \
    do {                      \
        int rc_ = (x);        \
        if (rc_) return rc_;  \
    } while (0)

extern "C" const char* dqmc_last_error(void) { return g_err.c_str(); }
extern "C" const char* dqmc_version(void) { return "dqmc-mi355x 0.1 (gfx950, FP64 MFMA)"; }

// --------------------------------------------------------------------------
// accumulator sizes (src/data.c:441-517)
// --------------------------------------------------------------------------
extern "C" long dqmc_acc_len(const dqmc_params* p, int ue, int k)
{
    const long L = p->L, nij = p->num_ij;
    if (!ue) {
        switch (k) {
        case DQ_EQ_density: case DQ_EQ_density_u: case DQ_EQ_density_d: case DQ_EQ_double_occ:
            return p->num_i;
        case DQ_EQ_g00: case DQ_EQ_g00_u: case DQ_EQ_g00_d: case DQ_EQ_nn: case DQ_EQ_xx:
        case DQ_EQ_zz: case DQ_EQ_pair_sw:
            return nij;
        case DQ_EQ_uuuu: case DQ_EQ_dddd: case DQ_EQ_uudd: case DQ_EQ_dduu:
            return p->meas_gen_suscept ? nij * nij : 0;
        case DQ_EQ_chi: return p->meas_chiral ? p->num_plaq_accum : 0;
        case DQ_EQ_kk: return p->meas_energy_corr ? p->num_bb : 0;
        case DQ_EQ_kv: case DQ_EQ_kn: return p->meas_energy_corr ? p->num_bs : 0;
        case DQ_EQ_vv: case DQ_EQ_vn: return p->meas_energy_corr ? nij : 0;
        case DQ_EQ_j: case DQ_EQ_jn: return p->meas_local_JQ ? p->num_b_accum : 0;
        case DQ_EQ_j2: return p->meas_local_JQ ? p->num_b2_accum : 0;
        }
        return 0;
    }
    if (p->period_uneqlt <= 0) return 0;
    switch (k) {
    case DQ_UE_gt0: case DQ_UE_gt0_u: case DQ_UE_gt0_d: case DQ_UE_nn: case DQ_UE_xx:
    case DQ_UE_zz: case DQ_UE_pair_sw:
        return nij * L;
    case DQ_UE_uuuu: case DQ_UE_dddd: case DQ_UE_uudd: case DQ_UE_dduu:
        return p->meas_gen_suscept ? nij * nij * L : 0;
    case DQ_UE_pair_bb:
        return (p->meas_bond_corr || p->meas_pair_bb_only) ? (long)p->num_bb * L : 0;
    case DQ_UE_jj: case DQ_UE_jsjs: case DQ_UE_kk: case DQ_UE_ksks:
        return p->meas_bond_corr ? (long)p->num_bb * L : 0;
    case DQ_UE_jnjn: case DQ_UE_jnj: case DQ_UE_jjn:
        return p->meas_thermal ? (long)p->num_bb * L : 0;
    case DQ_UE_j2jn: return p->meas_thermal ? (long)p->num_b2b * L : 0;
    case DQ_UE_jnj2: return p->meas_thermal ? (long)p->num_bb2 * L : 0;
    case DQ_UE_pair_b2b2: case DQ_UE_j2j2: case DQ_UE_js2js2: case DQ_UE_k2k2: case DQ_UE_ks2ks2:
        return p->meas_2bond_corr ? (long)p->num_b2b2 * L : 0;
    case DQ_UE_j2j: return p->meas_2bond_corr ? (long)p->num_b2b * L : 0;
    case DQ_UE_jj2: return p->meas_2bond_corr ? (long)p->num_bb2 * L : 0;
    case DQ_UE_kv: case DQ_UE_kn: return p->meas_energy_corr ? (long)p->num_bs * L : 0;
    case DQ_UE_vv: case DQ_UE_vn: return p->meas_energy_corr ? nij * L : 0;
    case DQ_UE_nem_nnnn: case DQ_UE_nem_ssss:
        return p->meas_nematic_corr ? (long)p->num_bb * L : 0;
    }
    return 0;
}

// accumulators computed on the GPU: all of them, except the nematic pair in
// complex builds, which the reference refuses too (src/meas.c:1369-1375)
static bool eq_supported(int k, int cplx)
{
    (void)cplx;
    return k >= 0 && k < DQ_EQ_COUNT;
}
static bool ue_supported(int k, int cplx)
{
    if (cplx && (k == DQ_UE_nem_nnnn || k == DQ_UE_nem_ssss)) return false;
    return k >= 0 && k < DQ_UE_COUNT;
}

// --------------------------------------------------------------------------
// engine state
// --------------------------------------------------------------------------
enum {
    PS_wall, PS_updates, PS_calcb, PS_multb, PS_recalc, PS_wrap, PS_half_wrap, PS_meas_eq,
    PS_calc_o, PS_bsofi, PS_expand_g, PS_meas_uneq, PS_meas_uneq_sub, PS_checkpoint, PS_COUNT
};

struct ProfPending {
    int s;
    hipEvent_t a, b;
};

struct dqmc_engine {
    int nc = 0, N = 0, L = 0, F = 0, n_matmul = 0, cplx = 0, kd = 16, device = 0;
    int period_eqlt = 0, period_uneqlt = 0, energy = 0;
    int num_i = 0, num_ij = 0;
    size_t esz = 8;
    hipStream_t st = nullptr;
    std::vector<void*> allocs;
    std::vector<int> sweep, n_sweep, n_sweep_warm;  // host mirrors
    bool stale = true;
    void *expK = 0, *iexpK = 0, *ehalf = 0, *iehalf = 0;  // [B][NN] T
    void* Bthi = 0;  // [B][L][NN] T, half-wrapped B_l^-1 (unequal-time only)
    double *exp_lambda = 0, *del = 0;                       // [nc][2N]
    int* hs = 0;                                            // [nc][L N]
    uint64_t* rng = 0;                                      // [nc][17]
    void *G = 0, *C = 0, *Ab = 0, *Bb = 0, *W = 0;          // G [B][NN]; C, W [B][F][NN]
    double *escale = 0, *iscale = 0;                        // [B][L N]
    void *X = 0, *Y = 0, *Y2 = 0, *Z = 0, *Q = 0, *Tm = 0, *Rs = 0, *Tp = 0, *Mm = 0, *T1 = 0;
    void* tau = 0;                                          // [B][N] T
    int *jpvt = 0, *perm = 0, *ipiv = 0;                    // [B][N]
    double *dvec = 0, *vvec = 0;                            // [B][N]
    void *phb = 0, *phb2 = 0, *phase = 0;                   // [B], [B], [nc] T
    unsigned long long* nacc = 0;                           // [nc]
    unsigned long long* nacc_sl = 0;                        // [nc][L]
    uint64_t* raw = 0;                                      // [nc][2 L N] draws of a sweep
    int *order = 0, *draws = 0;                             // [nc][L][N], [nc]
    double* unif = 0;                                       // [nc][L][N]
    int *active = 0, *flag_eq = 0, *flag_ue = 0;            // [nc]
    int *d_sweep = 0, *d_nsweep = 0, *d_warm = 0;           // [nc]
    long eq_off[DQMC_MAX_ACC], ue_off[DQMC_MAX_ACC];
    long eq_len[DQMC_MAX_ACC], ue_len[DQMC_MAX_ACC];
    long eq_total = 0, ue_total = 0;
    long *d_eq_off = 0, *d_ue_off = 0;
    void *acc_eq = 0, *acc_ue = 0, *sign_eq = 0, *sign_ue = 0;
    int *ns_eq = 0, *ns_ue = 0;
    int *rp_i = 0, *ix_i = 0, *rp_ij = 0, *ix_ij = 0, *degen = 0;
    int *rp_bb = 0, *ix_bb = 0, *rp_bs = 0, *ix_bs = 0, *bonds = 0;
    int num_b = 0, num_bb = 0, num_bs = 0, bond = 0;
    void *peierlsu = 0, *peierlsd = 0;
    // unequal time
    void *G0t = 0, *Gtt = 0, *Gt0 = 0;                      // [B][L][NN]
    void *Gred = 0, *Wa = 0, *Wb = 0, *Qb = 0, *taub = 0, *T2 = 0;
    int gemm_adiv = 1, gemm_bdiv = 1;  // one-shot batch divisors for the next gemm()
    int gemm_tile = GT_64;             // GEMM C-tile shape (GT_*)
    // optional measurement groups (dq_meas2.hpp)
    int thermal = 0, twobond = 0, nematic = 0, gensus = 0, chiral = 0, localjq = 0, pbbonly = 0;
    int Nx = 0, Ny = 0, num_b2 = 0, num_bb2 = 0, num_b2b = 0, num_b2b2 = 0, num_plaq = 0;
    int num_plaq_acc = 0, num_b_acc = 0, num_b2_acc = 0;
    int *rp2[PM_COUNT] = {}, *ix2[PM_COUNT] = {};
    long srp2[PM_COUNT] = {}, six2[PM_COUNT] = {};
    int *degen2 = 0, *bond2s = 0, *plaqs = 0;
    void *pp = 0, *gs = 0;
    void *btab = 0, *bph = 0, *b2tab = 0;  // bond-pair tables of the unequal-time measurement
    // adjoint outputs (allocated on first use)
    void *adjW = 0, *adjT = 0;
    double* adjL = 0;
    void *adjX = 0, *adjZ = 0, *adjO = 0;  // observable adjoint: X per (chain, spin), Z_l, tr(X G)
    int* invij = 0;      // [c][j][num_ij] row of column j in bin r (when ij_percol)
    bool ij_percol = false;
    // chains of uneven size (padded to N, L): own slice / site counts
    bool ragged = false;
    std::vector<int> Lc, Nc;
    int *d_Lc = 0, *d_Nc = 0;
    int* flag_sl = 0;    // [c] per-slice equal-time measurement flags
    // equal-time cadence per chain (padded engines: each chain keeps its
    // file's period_eqlt) and the slices at which any chain measures
    std::vector<int> pe;
    int* d_pe = 0;
    std::vector<char> meas_at;
    // blocked QR workspace: explicit V, compact-WY T factors, GEMM scratch, dets
    void *Vb = 0, *Tb = 0, *Wq1 = 0, *Wq2 = 0, *detq = 0, *detm = 0;
    long sVb = 0, sTb = 0, sWq = 0;
    // GEMM timing (bench)
    bool time_gemm = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> gemm_events;
    // one sweep captured per measurement pattern (bit 0: equal-time, bit 1:
    // unequal-time) and replayed; off while the GEMMs are timed.  Captured in
    // relaxed mode: engines driven from other host threads keep making
    // synchronous calls while one of them captures.
    bool use_graph = true;
    hipGraph_t graph[4] = {};
    hipGraphExec_t gexec[4] = {};
    // the host queues sweep s only after sweep s - 2 has finished on the
    // device, so a stop flag set while the sweeps run is seen within two
    // sweeps (the reference polls before every sweep, src/dqmc.c:224-230)
    hipEvent_t done[2] = {};
    // CHECK_G_WRP (src/dqmc.c:277-287): at every recalculation the wrapped G
    // (into Gw) against the recomputed one; per chain the largest |difference|
    bool check_drift = false;
    void* Gw = 0;
    unsigned long long* d_drift = 0;  // [nc] bits of a non-negative double
    double gemm_ms = 0.0, gemm_flops = 0.0;
    long long gemm_launches = 0;
    // per-section profile (dqmc_tuning.flags bit 3)
    bool prof = false;
    std::vector<hipEvent_t> ev_pool;
    std::vector<ProfPending> prof_pending;
    double prof_ms[PS_COUNT] = {};
    long long prof_calls[PS_COUNT] = {};
};

template <typename T>
static T* dp(void* p) { return (T*)p; }

// Profile sections: the reference's table (src/prof.h:8-22) in its order.
static const char* const g_prof_names[PS_COUNT] = {
    "wall", "updates", "calcb", "multb", "recalc", "wrap", "half_wrap", "meas_eq",
    "calc_o", "bsofi", "expand_g", "meas_uneq", "meas_uneq_sub", "checkpoint"};

// A section of device work on the engine stream, timed by a HIP event pair
// when the engine profiles (eager launches only); a no-op otherwise.
static hipEvent_t prof_event(dqmc_engine* e)
{
    hipEvent_t ev = nullptr;
    if (!e->ev_pool.empty()) {
        ev = e->ev_pool.back();
        e->ev_pool.pop_back();
    } else if (hipEventCreate(&ev) != hipSuccess) {
        ev = nullptr;
    }
    return ev;
}

struct ProfSec {
    dqmc_engine* e;
    int s;
    hipEvent_t a = nullptr;
    ProfSec(dqmc_engine* e_, int s_) : e(e_), s(s_)
    {
        if (e->prof && (a = prof_event(e))) (void)hipEventRecord(a, e->st);
    }
    // ends the section (at the latest when it goes out of scope)
    void stop()
    {
        if (!a) return;
        hipEvent_t b = prof_event(e);
        if (b) {
            (void)hipEventRecord(b, e->st);
            e->prof_pending.push_back({s, a, b});
        } else {
            e->ev_pool.push_back(a);
        }
        a = nullptr;
    }
    ~ProfSec() { stop(); }
};

// adds the pending sections' durations (ms) to the totals
static void prof_collect(dqmc_engine* e)
{
    if (e->prof_pending.empty()) return;
    (void)hipStreamSynchronize(e->st);
    for (auto& p : e->prof_pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
            e->prof_ms[p.s] += ms;
            e->prof_calls[p.s]++;
        }
        e->ev_pool.push_back(p.a);
        e->ev_pool.push_back(p.b);
    }
    e->prof_pending.clear();
}

static int dalloc(dqmc_engine* e, void** p, size_t bytes)
{
    *p = nullptr;
    if (bytes == 0) bytes = 256;
    hipError_t err = hipMalloc(p, bytes);
    if (err != hipSuccess)
        return set_err(DQMC_ERR_NOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(err));
    e->allocs.push_back(*p);
    // zero on the engine stream and wait: later set-up copies use the null
    // stream, which does not order against a non-blocking stream
    err = hipMemsetAsync(*p, 0, bytes, e->st);
    if (err == hipSuccess) err = hipStreamSynchronize(e->st);
    if (err != hipSuccess) return set_err(DQMC_ERR_HIP, "hipMemset: %s", hipGetErrorString(err));
    return 0;
}

#define DALLOC(ptr, bytes) RC(dalloc(e, (void**)&(ptr), (size_t)(bytes)))

static inline unsigned nblk(long total, int tpb = 256)
{
    long b = (total + tpb - 1) / tpb;
    if (b > 65535L * 16) b = 65535L * 16;
    return (unsigned)(b < 1 ? 1 : b);
}

// --------------------------------------------------------------------------
// GEMM helper
// --------------------------------------------------------------------------
template <typename T>
static int gemm(dqmc_engine* e, int opa, int opb, int M, int N, int K, T alpha, const T* A, int lda,
                long sA, const T* B, int ldb, long sB, T beta, T* C, int ldc, long sC, int batch,
                const double* r = nullptr, long sr = 0, const double* u = nullptr, long su = 0,
                const double* w = nullptr, long sw = 0, int id_l = -1, int id_side = 0,
                int a_upper = 0)
{
    GemmArgs<T> g;
    g.a_upper = a_upper;
    // B_l of an identity slice (l >= the chain's own L): the product by it is
    // a copy of the other operand (batch entry b belongs to chain b / 2)
    g.idLc = (e->ragged && id_l >= 0) ? e->d_Lc : nullptr;
    g.idl = id_l;
    g.idside = id_side;
    g.M = M; g.N = N; g.K = K;
    g.A = A; g.lda = lda; g.sA = sA;
    g.B = B; g.ldb = ldb; g.sB = sB;
    g.C = C; g.ldc = ldc; g.sC = sC;
    g.r = r; g.sr = sr; g.u = u; g.su = su; g.w = w; g.sw = sw;
    g.alpha = alpha; g.beta = beta;
    g.beta_zero = (re(beta) == 0.0 && im(beta) == 0.0);
    g.adiv = e->gemm_adiv > 0 ? e->gemm_adiv : 1;
    g.bdiv = e->gemm_bdiv > 0 ? e->gemm_bdiv : 1;
    e->gemm_adiv = e->gemm_bdiv = 1;
    hipEvent_t a0 = 0, a1 = 0;
    const bool timed = e->time_gemm && M == e->N && N == e->N && K == e->N && batch == 2 * e->nc;
    if (timed) {
        (void)hipEventCreate(&a0);
        (void)hipEventCreate(&a1);
        (void)hipEventRecord(a0, e->st);
    }
    hipError_t err = gemm_launch<T>(opa, opb, g, batch, e->st, e->gemm_tile);
    if (timed) {
        (void)hipEventRecord(a1, e->st);
        e->gemm_events.push_back({a0, a1});
        e->gemm_flops = 2.0 * M * N * (double)K * batch * (is_cplx<T>::value ? 4 : 1);
    }
    if (err != hipSuccess) return set_err(DQMC_ERR_HIP, "gemm launch: %s", hipGetErrorString(err));
    return 0;
}

template <typename T>
__global__ void k_copy_block(int N, const T* src, long ssrc, int lds, T alpha, T* dst, long sdst,
                             int ldd, int batch)
{
    EW_IB(N);
    EW_FOR_J(N) dst[b * sdst + i + (long)ldd * j] = alpha * src[b * ssrc + i + (long)lds * j];
}

// --------------------------------------------------------------------------
// B-product chains
// --------------------------------------------------------------------------
// C_f = B_{(f+1)n-1} ... B_{fn} for groups f0..f0+nf-1 of every (chain, spin)
// (mul_seq over B, src/greens.c:8-41, called from src/dqmc.c:199-201, 273-274).
template <typename T>
static int build_C(dqmc_engine* e, int f0, int nf)
{
    const int N = e->N, n = e->n_matmul, B = 2 * e->nc, F = e->F;
    const long NN = (long)N * N, LN = (long)e->L * N;
    T* C = dp<T>(e->C);
    T* W = dp<T>(e->W);
    const T one = from_re<T>(1.0), zero = from_re<T>(0.0);
    for (int f = f0; f < f0 + nf; f++) {
        const int l0 = f * n;
        T* bufs[2] = {C + f * NN, W + f * NN};
        int cur = ((n - 1) % 2 == 0) ? 0 : 1;  // the last product lands in C_f
        {
            ProfSec ps(e, PS_calcb);
            k_colscale<T><<<ew_grid(N, N, B), ew_threads(N), 0, e->st>>>(N, dp<T>(e->expK), NN, 1,
                                                          e->escale + (long)N * l0, LN, bufs[cur],
                                                          (long)F * NN, B,
                                                          e->ragged ? e->d_Lc : nullptr, l0);
            LCHK();
        }
        ProfSec ps(e, PS_multb);
        for (int s = 1; s < n; s++) {
            RC(gemm<T>(e, OP_N, OP_N, N, N, N, one, dp<T>(e->expK), N, NN, bufs[cur], N,
                       (long)F * NN, zero, bufs[1 - cur], N, (long)F * NN, B, nullptr, 0,
                       e->escale + (long)N * (l0 + s), LN, nullptr, 0, l0 + s, 1));
            cur = 1 - cur;
        }
    }
    return 0;
}

// out[b] = alpha * C[maxp1-1] ... C[min] (indices mod F), as mul_seq over the
// C_f array (src/greens.c:8-41).  Products of more than two factors use Rs/Tp.
template <typename T>
static int mul_seq_C(dqmc_engine* e, int min, int maxp1, T alpha, T* out, int ldo, long so)
{
    const int N = e->N, F = e->F, B = 2 * e->nc;
    const long NN = (long)N * N, sC = (long)F * NN;
    const T* C = dp<T>(e->C);
    const int n_mul = (min == maxp1) ? F : (F + maxp1 - min) % F;
    if (n_mul == 1) {
        k_copy_block<T><<<ew_grid(N, N, B), ew_threads(N), 0, e->st>>>(N, C + min * NN, sC, N, alpha, out, so,
                                                        ldo, B);
        LCHK();
        return 0;
    }
    const T one = from_re<T>(1.0), zero = from_re<T>(0.0);
    T* tmp[2] = {dp<T>(e->Rs), dp<T>(e->Tp)};
    const T* prev = C + min * NN;
    long sprev = sC;
    int l = (min + 1) % F;
    for (int step = 1; step < n_mul; step++, l = (l + 1) % F) {
        const bool last = (step == n_mul - 1);
        T* dst = last ? out : tmp[step & 1];
        RC(gemm<T>(e, OP_N, OP_N, N, N, N, last ? alpha : one, C + l * NN, N, sC, prev, N, sprev,
                   zero, dst, last ? ldo : N, last ? so : NN, B));
        prev = dst;
        sprev = last ? so : NN;
    }
    return 0;
}

// --------------------------------------------------------------------------
// blocked Householder QR and blocked upper-triangular solve (dq_qr_blocked.hpp)
// --------------------------------------------------------------------------
template <typename T, int QM, int NCW, int NW>
static void launch_panel(dqmc_engine* e, int batch, int m, int k0, int nbp, T* A, int lda, long sA,
                         T* tau, long stau, T* V, int ldv, long sV, T* Tk, long sT, T* det, int init)
{
    k_qr_panel<T, QM, NCW, NW><<<batch, 64 * NW, 0, e->st>>>(m, k0, nbp, A, lda, sA, tau, stau, V,
                                                             ldv, sV, Tk, sT, det, init);
}

// A[b] (m x n, lda) <- R (upper) + reflectors (below); tau[b][0..min(m,n));
// Q[b] (m x m, ldq), if non-null, <- H_0 ... H_{k-1}; det[b], if non-null,
// <- prod det(H_i).  Panels of 64 (real) / 32 (complex) columns.
// Q = H_0 H_1 ... H_{kmax-1} (m x m, ldq) from the explicit reflectors in the
// V workspace (ld m) and the compact-WY T factors of their NB-column panels in
// the T workspace: Q <- (I - V_p T_p V_p^H) Q, panels from the last.
template <typename T>
static int accumulate_q(dqmc_engine* e, int m, int kmax, int NB, T* Q, int ldq, long sQ, int batch)
{
    T* V = dp<T>(e->Vb);
    const long sV = e->sVb;
    const int ldv = m;
    T* Tb = dp<T>(e->Tb);
    const long sT = e->sTb;
    T* W1 = dp<T>(e->Wq1);
    T* W2 = dp<T>(e->Wq2);
    const long sW = e->sWq;
    const T one = from_re<T>(1.0), mone = from_re<T>(-1.0), zero = from_re<T>(0.0);
    const int np = (kmax + NB - 1) / NB;
    k_set_identity<T><<<ew_grid(m, m, batch), ew_threads(m), 0, e->st>>>(m, Q, sQ, ldq, 0, batch);
    LCHK();
    for (int p = np - 1; p >= 0; p--) {
        const int k0 = p * NB;
        const int nbp = (kmax - k0 < NB) ? kmax - k0 : NB;
        const int mp = m - k0;
        T* Tk = Tb + (long)p * NB * NB;
        T* Vk = V + k0 + (long)ldv * k0;
        T* Qs = Q + k0 + (long)ldq * k0;
        RC(gemm<T>(e, OP_C, OP_N, nbp, mp, mp, one, Vk, ldv, sV, Qs, ldq, sQ, zero, W1, NB, sW,
                   batch));
        RC(gemm<T>(e, OP_N, OP_N, nbp, mp, nbp, one, Tk, NB, sT, W1, NB, sW, zero, W2, NB, sW,
                   batch));
        RC(gemm<T>(e, OP_N, OP_N, mp, mp, nbp, mone, Vk, ldv, sV, W2, NB, sW, one, Qs, ldq, sQ,
                   batch));
    }
    return 0;
}

template <typename T>
static int qr_blocked(dqmc_engine* e, int m, int n, T* A, int lda, long sA, T* tau, long stau,
                      T* Q, int ldq, long sQ, T* det, int batch)
{
    constexpr bool Z = is_cplx<T>::value;
    if (m > 1024) return set_err(DQMC_ERR_UNSUPPORTED, "QR of %d rows", m);
    const int qm = m <= 256 ? 4 : (m <= 512 ? 8 : 16);
    // panel width 64 (real) / 32 (complex) up to 512 rows, half that beyond;
    // real panels of up to 256 rows in 8 waves of 8 columns (two workgroups
    // per CU), the others in 16 waves (a lane's panel share and the rest of
    // the step within 128 VGPRs)
    const int NB = qm == 16 ? (Z ? 16 : 32) : (Z ? 32 : 64);
    const int kmax = m < n ? m : n;
    T* V = dp<T>(e->Vb);
    const long sV = e->sVb;
    const int ldv = m;
    T* Tb = dp<T>(e->Tb);
    const long sT = e->sTb;
    T* W1 = dp<T>(e->Wq1);
    T* W2 = dp<T>(e->Wq2);
    const long sW = e->sWq;
    const T one = from_re<T>(1.0), mone = from_re<T>(-1.0), zero = from_re<T>(0.0);
    int np = 0;
    for (int k0 = 0; k0 < kmax; k0 += NB, np++) {
        const int nbp = (kmax - k0 < NB) ? kmax - k0 : NB;
        T* Tk = Tb + (long)np * NB * NB;
        if constexpr (Z) {
            if (qm == 4) launch_panel<T, 4, 2, 16>(e, batch, m, k0, nbp, A, lda, sA, tau, stau, V, ldv, sV, Tk, sT, det, k0 == 0);
            else if (qm == 8) launch_panel<T, 8, 2, 16>(e, batch, m, k0, nbp, A, lda, sA, tau, stau, V, ldv, sV, Tk, sT, det, k0 == 0);
            else launch_panel<T, 16, 1, 16>(e, batch, m, k0, nbp, A, lda, sA, tau, stau, V, ldv, sV, Tk, sT, det, k0 == 0);
        } else {
            if (qm == 4) launch_panel<T, 4, 8, 8>(e, batch, m, k0, nbp, A, lda, sA, tau, stau, V, ldv, sV, Tk, sT, det, k0 == 0);
            else if (qm == 8) launch_panel<T, 8, 4, 16>(e, batch, m, k0, nbp, A, lda, sA, tau, stau, V, ldv, sV, Tk, sT, det, k0 == 0);
            else launch_panel<T, 16, 2, 16>(e, batch, m, k0, nbp, A, lda, sA, tau, stau, V, ldv, sV, Tk, sT, det, k0 == 0);
        }
        LCHK();
        const int nt = n - k0 - nbp;
        if (nt > 0) {
            const int mp = m - k0;
            T* At = A + k0 + (long)lda * (k0 + nbp);
            T* Vk = V + k0 + (long)ldv * k0;
            RC(gemm<T>(e, OP_C, OP_N, nbp, nt, mp, one, Vk, ldv, sV, At, lda, sA, zero, W1, NB, sW,
                       batch));
            RC(gemm<T>(e, OP_C, OP_N, nbp, nt, nbp, one, Tk, NB, sT, W1, NB, sW, zero, W2, NB, sW,
                       batch));
            RC(gemm<T>(e, OP_N, OP_N, mp, nt, nbp, mone, Vk, ldv, sV, W2, NB, sW, one, At, lda, sA,
                       batch));
        }
    }
    if (Q) RC(accumulate_q<T>(e, m, kmax, NB, Q, ldq, sQ, batch));
    return 0;
}

template <typename T>
static int q_from_reflectors(dqmc_engine* e, int n, const T* tau, T* Q, long sQ, int batch);

// Q column-pivoted QR of the first product with the matrix resident on one CU
// (k_qrp_res), when it fits: real n <= 256, complex n <= 144.
template <typename T>
static bool qrp_resident_fits(int n)
{
    return is_cplx<T>::value ? n <= 144 : n <= 256;
}

template <typename T>
static int qrp_resident(dqmc_engine* e, int n, T* A, long sA, T* tau, int* jpvt, T* Q, long sQ,
                        T* det, int batch)
{
    constexpr bool Z = is_cplx<T>::value;
    T* V = dp<T>(e->Vb);
    const long sV = e->sVb;
#define QRP(R, SR, SL, CB)                                                                  \
    k_qrp_res<T, R, SR, SL, CB><<<batch, QRP_THREADS, 0, e->st>>>(n, n, A, n, sA, tau, n, jpvt, \
                                                                  n, V, n, sV, det)
    if constexpr (Z) {
        if (n <= 64) QRP(1, 8, 0, 2);
        else if (n <= 128) QRP(2, 14, 2, 2);
        else if (n <= 144) QRP(3, 14, 4, 2);
        else return set_err(DQMC_ERR_UNSUPPORTED, "resident pivoted QR of order %d", n);
    } else {
        if (n <= 64) QRP(1, 8, 0, 4);
        else if (n <= 128) QRP(2, 16, 0, 4);
        else if (n <= 192) QRP(3, 24, 0, 4);
        else if (n <= 256) QRP(4, 24, 8, 8);
        else return set_err(DQMC_ERR_UNSUPPORTED, "resident pivoted QR of order %d", n);
    }
#undef QRP
    LCHK();
    return q_from_reflectors<T>(e, n, tau, Q, sQ, batch);
}

// Explicit Q = H_0 ... H_{n-1} from the unit-lower reflectors in the V
// workspace (ld n) and tau: compact-WY factors of 64-column panels from their
// Gram matrices (one GEMM + k_larft_gram), then the backward three-GEMM
// accumulation.
template <typename T>
static int q_from_reflectors(dqmc_engine* e, int n, const T* tau, T* Q, long sQ, int batch)
{
    T* V = dp<T>(e->Vb);
    const long sV = e->sVb;
    const int NB = 64;
    const int np = (n + NB - 1) / NB;
    T* Tb = dp<T>(e->Tb);
    const long sT = e->sTb;
    const T one = from_re<T>(1.0), zero = from_re<T>(0.0);
    for (int p = 0; p < np; p++) {
        const int k0 = p * NB;
        const int nbp = (n - k0 < NB) ? n - k0 : NB;
        T* Vk = V + k0 + (long)n * k0;
        RC(gemm<T>(e, OP_C, OP_N, nbp, nbp, n - k0, one, Vk, n, sV, Vk, n, sV, zero,
                   Tb + (long)p * NB * NB, NB, sT, batch));
    }
    k_larft_gram<T><<<dim3(np, batch), 512, 0, e->st>>>(NB, n, tau, n, Tb, sT);
    LCHK();
    return accumulate_q<T>(e, n, n, NB, Q, n, sQ, batch);
}

// X[b] (n x ncols, ldx) <- U[b]^-1 X[b] for upper-triangular U (ldu): 64-row
// blocks from the bottom, GEMM updates off the diagonal, LDS solves on it.
template <typename T>
static int trsm_upper_blocked(dqmc_engine* e, int n, int ncols, const T* U, int ldu, long sU, T* X,
                              int ldx, long sX, int batch)
{
    if (n <= 0 || ncols <= 0) return 0;
    const int BS = 64;
    const T one = from_re<T>(1.0), mone = from_re<T>(-1.0);
    const int nblocks = (n + BS - 1) / BS;
    for (int kb = nblocks - 1; kb >= 0; kb--) {
        const int k0 = kb * BS;
        const int nk = (n - k0 < BS) ? n - k0 : BS;
        if (k0 + nk < n)
            RC(gemm<T>(e, OP_N, OP_N, nk, ncols, n - k0 - nk, mone, U + k0 + (long)ldu * (k0 + nk),
                       ldu, sU, X + k0 + nk, ldx, sX, one, X + k0, ldx, sX, batch));
        if constexpr (!is_cplx<T>::value) {
            if (nk == 64) {  // one thread per right-hand side
                k_trsm_diag64_cols<<<dim3((ncols + TRSM_COLS - 1) / TRSM_COLS, batch), TRSM_COLS,
                                     0, e->st>>>(
                    ncols, U + k0 + (long)ldu * k0, sU, ldu, X + k0, sX, ldx);
                LCHK();
                continue;
            }
        }
        dim3 grid((ncols + 4 * LU_NC - 1) / (4 * LU_NC), batch);
        k_trsm_diag64<T><<<grid, 256, 0, e->st>>>(nk, ncols, U + k0 + (long)ldu * k0, sU, ldu,
                                                  X + k0, sX, ldx);
        LCHK();
    }
    return 0;
}

// --------------------------------------------------------------------------
// stabilized equal-time G (calc_eq_g, src/greens.c:69-208) for all b:
// Gout[b] <- (I + C_{l-1} ... C_l)^-1, phase_b[b] <- phase of det(I + ...)
// --------------------------------------------------------------------------
template <typename T>
static int launch_getrs(dqmc_engine* e, T* Gout, T* Vlast, T* phase_b)
{
    const int N = e->N, B = 2 * e->nc;
    const long NN = (long)N * N;
    const int waves_per_blk = 4;
    dim3 grid((N + waves_per_blk * LU_NC - 1) / (waves_per_blk * LU_NC), B);
#define GRS(R)                                                                                  \
    k_getrs_phase<T, R><<<grid, 256, 0, e->st>>>(N, dp<T>(e->Mm), Gout, NN, e->ipiv, N, e->vvec, \
                                                 N, Vlast, NN, dp<T>(e->tau), N, phase_b)
    if (N <= 64) GRS(1);
    else if (N <= 128) GRS(2);
    else if (N <= 256) GRS(4);
    else if (N <= 512) GRS(8);
    else GRS(16);
#undef GRS
    LCHK();
    return 0;
}

template <typename T>
static int calc_eq_g(dqmc_engine* e, int l, T* Gout, T* phase_b)
{
    const int N = e->N, F = e->F, B = 2 * e->nc;
    const long NN = (long)N * N;
    const int n_mul = 2;  // N_MUL, src/greens.h:5
    T *X = dp<T>(e->X), *Y = dp<T>(e->Y), *Y2 = dp<T>(e->Y2), *Z = dp<T>(e->Z);
    T *Q = dp<T>(e->Q), *Tm = dp<T>(e->Tm), *Rs = dp<T>(e->Rs), *Tp = dp<T>(e->Tp);
    T* Mm = dp<T>(e->Mm);
    T* tau = dp<T>(e->tau);
    T* detq = dp<T>(e->detq);
    T* detm = dp<T>(e->detm);
    const T one = from_re<T>(1.0), zero = from_re<T>(0.0);
    int m = (l + 1 + (F - 1) % n_mul) % F;
    RC(mul_seq_C<T>(e, l, m, one, X, N, NN));
    // (1) column-pivoted QR of the first product (xgeqp3)
    if (qrp_resident_fits<T>(N)) {  // also forms det(Q) = prod det(H_j)
        RC(qrp_resident<T>(e, N, X, NN, tau, e->jpvt, Q, NN, detq, B));
    } else {
        // streaming pivoted QR; Q from the reflectors by compact-WY GEMMs
        if (N <= 256)
            k_qr<T, true, 4><<<B, QR_THREADS, 0, e->st>>>(N, N, X, N, NN, tau, N, e->jpvt, N, nullptr, N, NN);
        else if (N <= 512)
            k_qr<T, true, 8><<<B, QR_THREADS, 0, e->st>>>(N, N, X, N, NN, tau, N, e->jpvt, N, nullptr, N, NN);
        else
            k_qr<T, true, 16><<<B, QR_THREADS, 0, e->st>>>(N, N, X, N, NN, tau, N, e->jpvt, N, nullptr, N, NN);
        LCHK();
        k_extract_v<T><<<dim3(N, B), 256, 0, e->st>>>(N, X, NN, dp<T>(e->Vb), e->sVb);
        LCHK();
        RC(q_from_reflectors<T>(e, N, tau, Q, NN, B));
        k_reflector_det<T><<<B, 64, 0, e->st>>>(N, N, X, N, NN, tau, N, detq);
        LCHK();
    }
    k_eqg_diag<T><<<dim3((N + 255) / 256, B), 256, 0, e->st>>>(N, X, NN, e->dvec, N);
    LCHK();
    k_eqg_init_T<T><<<dim3((N + 255) / 256, (N + EQG_COLS - 1) / EQG_COLS, B), 256, 0, e->st>>>(
        N, X, NN, e->jpvt, N, e->dvec, N, Tm, NN);
    LCHK();
    while (m != l) {
        const int next = (m + n_mul) % F;
        RC(mul_seq_C<T>(e, m, next, one, Y, N, NN));
        m = next;
        // (3a) Y2 = Y * Q * diag(d)   (xunmqr R,N + column scaling)
        RC(gemm<T>(e, OP_N, OP_N, N, N, N, one, Y, N, NN, Q, N, NN, zero, Y2, N, NN, B, nullptr,
                   0, nullptr, 0, e->dvec, N));
        // (3b) pre-pivot by column norms, (3c) blocked QR with explicit Q
        k_prepivot<T><<<B, 1024, 0, e->st>>>(N, Y2, NN, Z, NN, e->perm, N);
        LCHK();
        RC(qr_blocked<T>(e, N, N, Z, N, NN, tau, N, Q, N, NN, detq, B));
        // (3d) T = diag(1/d) R * P^T T   (xtrmm L,U)
        k_eqg_diag<T><<<dim3((N + 255) / 256, B), 256, 0, e->st>>>(N, Z, NN, e->dvec, N);
        LCHK();
        k_eqg_post_qr<T><<<dim3((N + 255) / 256, (N + EQG_COLS - 1) / EQG_COLS, B), 256, 0,
                           e->st>>>(N, Z, NN, e->perm, N, Tm, NN, e->dvec, N, Rs, Tp, NN);
        LCHK();
        // Rs is upper triangular: its product skips the zero blocks
        RC(gemm<T>(e, OP_N, OP_N, N, N, N, one, Rs, N, NN, Tp, N, NN, zero, Tm, N, NN, B, nullptr,
                   0, nullptr, 0, nullptr, 0, -1, 0, 1));
    }
    // Eq. 2.12: G = M'^-1 diag(v) Q^H with M' = diag(ds) T + diag(v) Q^H, solved
    // stably through the QR of M' (M' = Q_M R_M, G = R_M^-1 Q_M^H RHS)
    k_eqg_final_prep<T><<<dim3((N + 31) / 32, (N + 31) / 32, B), dim3(32, 8), 0, e->st>>>(
        N, Q, NN, e->dvec, N, Tm, NN, Mm, Gout, NN, e->vvec, N);
    LCHK();
    RC(qr_blocked<T>(e, N, N, Mm, N, NN, tau, N, X, N, NN, detm, B));
    RC(gemm<T>(e, OP_C, OP_N, N, N, N, one, X, N, NN, Gout, N, NN, zero, Y, N, NN, B));
    RC(trsm_upper_blocked<T>(e, N, N, Mm, N, NN, Y, N, NN, B));
    k_copy_block<T><<<ew_grid(N, N, B), ew_threads(N), 0, e->st>>>(N, Y, NN, N, one, Gout, NN, N, B);
    LCHK();
    k_eqg_phase<T><<<B, 64, 0, e->st>>>(N, detq, detm, e->vvec, N, Mm, NN, phase_b);
    LCHK();
    return 0;
}

// --------------------------------------------------------------------------
// sweep bookkeeping on the device (graph-friendly: no host round trips)
// --------------------------------------------------------------------------
__global__ void k_sweep_flags(int nc, int period_uneqlt, const int* sweep, const int* nsweep,
                              const int* warm, int* active, int* flag_eq, int* flag_ue)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    const int s = sweep[c];
    const int act = s < nsweep[c];
    active[c] = act;
    flag_eq[c] = act && s >= warm[c];
    flag_ue[c] = act && s >= warm[c] && period_uneqlt > 0 && (s % period_uneqlt) == 0;
}

// equal-time measurement flags of slice l for chains padded to a common L,
// each at its own cadence pe[c] (src/dqmc.c:361: (l + 1) % period_eqlt == 0)
__global__ void k_flag_slice(int nc, const int* flag_eq, const int* Lc, const int* pe, int l,
                             int* out)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < nc) out[c] = flag_eq[c] && l < Lc[c] && pe[c] > 0 && (l + 1) % pe[c] == 0;
}

__global__ void k_sweep_advance(int nc, int* sweep, const int* active)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < nc) sweep[c] += active[c];
}

// --------------------------------------------------------------------------
// typed engine operations
// --------------------------------------------------------------------------
template <typename T>
static int init_green(dqmc_engine* e)
{
    const int nc = e->nc;
    k_scales<<<nblk((long)nc * e->L * e->N), 256, 0, e->st>>>(e->N, e->L, nc, e->hs, e->exp_lambda,
                                                               e->escale, e->iscale);
    LCHK();
    RC(build_C<T>(e, 0, e->F));
    RC(calc_eq_g<T>(e, 0, dp<T>(e->G), dp<T>(e->phb)));
    // all chains: phase = phase_up * phase_dn (src/dqmc.c:219)
    std::vector<int> ones(nc, 1);
    HIPCHK(hipMemcpyAsync(e->active, ones.data(), nc * sizeof(int), hipMemcpyHostToDevice, e->st));
    k_phase_combine<T><<<(nc + 255) / 256, 256, 0, e->st>>>(nc, dp<T>(e->phb), dp<T>(e->phase),
                                                             e->active);
    LCHK();
    HIPCHK(hipStreamSynchronize(e->st));
    e->stale = false;
    return 0;
}

// Adjoint pass for the current (fixed) fields of every chain: log|det M| per
// (chain, spin) and the sums W, tr G, lambda term of dq_adjoint.hpp.  The
// slices are walked like a sweep without updates: wraps, and a fresh
// stabilized G from the stored group products every n_matmul slices.
template <typename T>
static int adjoint_t(dqmc_engine* e, double* logdet, double* Wout, double* trG, double* sGl)
{
    const int N = e->N, L = e->L, nc = e->nc, B = 2 * nc, n = e->n_matmul, F = e->F;
    const long NN = (long)N * N, LN = (long)L * N;
    const T one = from_re<T>(1.0), zero = from_re<T>(0.0);
    if (!e->adjW) {
        DALLOC(e->adjW, B * NN * e->esz);
        DALLOC(e->adjT, 2L * B * e->esz);
        DALLOC(e->adjL, B * 8);
    }
    T* W = dp<T>(e->adjW);
    T* tr = dp<T>(e->adjT);
    T* sg = tr + B;
    HIPCHK(hipMemsetAsync(W, 0, B * NN * sizeof(T), e->st));
    HIPCHK(hipMemsetAsync(tr, 0, 2L * B * sizeof(T), e->st));
    T* G = dp<T>(e->G);
    T* T1 = dp<T>(e->T1);
    k_scales<<<nblk((long)nc * L * N), 256, 0, e->st>>>(N, L, nc, e->hs, e->exp_lambda, e->escale,
                                                        e->iscale);
    LCHK();
    RC(build_C<T>(e, 0, F));
    RC(calc_eq_g<T>(e, 0, G, dp<T>(e->phb)));
    k_eqg_logdet<T><<<B, 64, 0, e->st>>>(N, e->vvec, N, dp<T>(e->Mm), NN, e->adjL);
    LCHK();
    for (int l = 0; l < L; l++) {
        k_adj_accum_W<T><<<nblk(B * NN), 256, 0, e->st>>>(N, G, NN, e->escale, e->iscale, LN, l, W,
                                                           NN, B, 1.0, -1.0);
        LCHK();
        k_adj_accum_diag<T><<<B, 64, 0, e->st>>>(N, L, G, NN, e->hs, l, tr, sg, 1.0, -1.0);
        LCHK();
        if ((l + 1) % n == 0) {
            RC(calc_eq_g<T>(e, ((l / n) + 1) % F, G, dp<T>(e->phb)));
        } else {
            RC(gemm<T>(e, OP_N, OP_N, N, N, N, one, G, N, NN, dp<T>(e->iexpK), N, NN, zero, T1, N, NN,
                       B, nullptr, 0, e->iscale + (long)N * l, LN));
            RC(gemm<T>(e, OP_N, OP_N, N, N, N, one, dp<T>(e->expK), N, NN, T1, N, NN, zero, G, N, NN,
                       B, nullptr, 0, e->escale + (long)N * l, LN));
        }
    }
    HIPCHK(hipStreamSynchronize(e->st));
    if (logdet) HIPCHK(hipMemcpy(logdet, e->adjL, B * 8, hipMemcpyDeviceToHost));
    if (Wout) HIPCHK(hipMemcpy(Wout, W, B * NN * sizeof(T), hipMemcpyDeviceToHost));
    if (trG) HIPCHK(hipMemcpy(trG, tr, B * sizeof(T), hipMemcpyDeviceToHost));
    if (sGl) HIPCHK(hipMemcpy(sGl, sg, B * sizeof(T), hipMemcpyDeviceToHost));
    e->stale = true;  // G was walked through all slices; the next run starts fresh
    return 0;
}

template <typename T>
static MeasArgs<T> meas_args(dqmc_engine* e, int ue)
{
    MeasArgs<T> a;
    memset(&a, 0, sizeof a);
    a.N = e->N; a.L = e->L; a.num_i = e->num_i; a.num_ij = e->num_ij;
    a.energy = e->energy;
    a.phase = dp<T>(e->phase);
    a.rp_i = e->rp_i; a.ix_i = e->ix_i; a.srp_i = e->num_i + 1; a.six_i = e->N;
    a.rp_ij = e->rp_ij; a.ix_ij = e->ix_ij; a.srp_ij = e->num_ij + 1; a.six_ij = (long)e->N * e->N;
    a.degen = e->degen;
    a.num_b = e->num_b; a.num_bb = e->num_bb; a.num_bs = e->num_bs; a.bond = e->bond;
    a.bonds = e->bonds; a.sbonds = 2L * e->num_b;
    a.rp_bb = e->rp_bb; a.ix_bb = e->ix_bb; a.srp_bb = e->num_bb + 1; a.six_bb = (long)e->num_b * e->num_b;
    a.rp_bs = e->rp_bs; a.ix_bs = e->ix_bs; a.srp_bs = e->num_bs + 1; a.six_bs = (long)e->num_b * e->N;
    a.peierlsu = dp<T>(e->peierlsu); a.peierlsd = dp<T>(e->peierlsd);
    a.sP = is_cplx<T>::value ? (long)e->N * e->N : 0;
    a.thermal = e->thermal; a.twobond = e->twobond; a.pbbonly = e->pbbonly;
    a.chiral = e->chiral; a.localjq = e->localjq;
    a.Nx = e->Nx; a.Ny = e->Ny; a.num_b2 = e->num_b2; a.num_bb2 = e->num_bb2;
    a.num_b2b = e->num_b2b; a.num_b2b2 = e->num_b2b2; a.num_plaq = e->num_plaq;
    a.num_plaq_acc = e->num_plaq_acc; a.num_b_acc = e->num_b_acc; a.num_b2_acc = e->num_b2_acc;
    for (int m = 0; m < PM_COUNT; m++) {
        a.rp2[m] = e->rp2[m]; a.ix2[m] = e->ix2[m]; a.srp2[m] = e->srp2[m]; a.six2[m] = e->six2[m];
    }
    a.degen2 = e->degen2;
    a.bond2s = e->bond2s; a.sbond2s = 2L * e->num_b2;
    a.pp = dp<T>(e->pp); a.spp = 4L * e->num_b2;
    a.plaqs = e->plaqs; a.splaqs = 3L * e->num_plaq;
    a.gs = dp<T>(e->gs); a.sgs = (long)e->L * 4 * e->num_ij;
    a.invij = e->invij; a.sinv = (long)e->N * e->num_ij;
    if (ue) {
        a.flag = e->flag_ue; a.acc = dp<T>(e->acc_ue); a.sacc = e->ue_total; a.off = e->d_ue_off;
        a.n_sample = e->ns_ue; a.sign = dp<T>(e->sign_ue);
    } else {
        a.flag = e->flag_eq; a.acc = dp<T>(e->acc_eq); a.sacc = e->eq_total; a.off = e->d_eq_off;
        a.n_sample = e->ns_eq; a.sign = dp<T>(e->sign_eq);
    }
    return a;
}

// equal-time optional groups on the half-wrapped G of ma: gen_suscept
// (factorised, dq_meas2.hpp), chiral and local JQ
template <typename T>
static int meas_eq_extra(dqmc_engine* e, const MeasArgs<T>& ma)
{
    const int nc = e->nc;
    if (e->gensus) {
        const long nij = e->num_ij;
        k_gensus_ab<T><<<dim3((nij + 255) / 256, 1, nc), 256, 0, e->st>>>(ma, 0);
        LCHK();
        k_gensus_outer<T><<<dim3((nij * nij + 255) / 256, 1, nc), 256, 0, e->st>>>(ma, 0);
        LCHK();
    }
    if (e->chiral || e->localjq) {
        const int nb = std::max(e->chiral ? e->num_plaq_acc : 0,
                                e->localjq ? std::max(e->num_b_acc, e->num_b2_acc) : 0);
        k_meas_eq_extra<T><<<dim3((nb + 255) / 256, nc), 256, 0, e->st>>>(ma);
        LCHK();
    }
    return 0;
}

// unequal-time Green's functions and measurements
#include "dq_ue_host.hpp"

// Delay ranks of the compiled Metropolis variants (powers of two between the
// bounds below):
//   N <= 64, registers:                  real 8..16, complex 4..8
//   N <= 256, A in registers, B in LDS:  real 8..32, complex 8..16
//   N <= 512:                            real 8..16 (512 threads, B in LDS),
//                                        complex 4 (registers, 2 rows a thread)
// A request takes the largest compiled rank not above it, the smallest below
// them all; N > 512 runs the global-memory kernel at any rank <= N.
static int upd_rank(int N, bool cplx, int want)
{
    if (N > UPD2_MAXN) return std::max(1, std::min(want, N));
    int lo, hi;
    if (N <= 64) { lo = cplx ? 4 : 8; hi = cplx ? 8 : 16; }
    else if (N <= UPD3_THREADS) { lo = 8; hi = cplx ? 16 : 32; }
    else { lo = cplx ? 4 : 8; hi = cplx ? 4 : 16; }
    int k = lo;
    while (2 * k <= hi && 2 * k <= want) k *= 2;
    return k;
}

// Metropolis sweep of slice l for every chain at the engine's delay rank
// (upd_rank).  The flush work per acceptance does not depend on the rank
// (rounds cost ~rank, there are ~1/rank of them); the gemv grows with it.
template <typename T>
static int launch_update(dqmc_engine* e, int l)
{
    const int N = e->N, nc = e->nc;
    const long NN = (long)N * N;
    if (N <= UPD2_MAXN) {
        Upd2Args<T> a;
        a.N = N; a.L = e->L; a.l = l;
        a.G = dp<T>(e->G); a.sG = NN;
        a.del = e->del; a.exp_lambda = e->exp_lambda; a.hs = e->hs;
        a.order = e->order; a.unif = e->unif;
        a.phase = dp<T>(e->phase); a.escale = e->escale; a.iscale = e->iscale;
        a.n_acc = e->nacc; a.n_acc_sl = e->nacc_sl; a.active = e->active;
        a.Lc = e->ragged ? e->d_Lc : nullptr;
        constexpr bool Z = is_cplx<T>::value;
        const int kd = e->kd;
        // small N: the register-only kernel (its flushes are cheap there); up
        // to N = 256 B in LDS and the flush on the matrix cores; up to N = 512
        // the same with 512 threads (real; complex keeps the register kernel)
        if (N <= 64) {
            if (kd == (Z ? 8 : 16)) k_update_reg<T, 1, Z ? 8 : 16><<<nc, UPD2_THREADS, 0, e->st>>>(a);
            else k_update_reg<T, 1, Z ? 4 : 8><<<nc, UPD2_THREADS, 0, e->st>>>(a);
        } else if (N <= UPD3_THREADS) {
            if (!Z && kd == 32) k_update_lds<T, 1, Z ? 16 : 32><<<nc, UPD3_THREADS, 0, e->st>>>(a);
            else if (kd == 16) k_update_lds<T, 1, 16><<<nc, UPD3_THREADS, 0, e->st>>>(a);
            else k_update_lds<T, 1, 8><<<nc, UPD3_THREADS, 0, e->st>>>(a);
        } else if constexpr (!Z) {
            if (kd == 16) k_update_lds<T, 1, 16, 512><<<nc, 512, 0, e->st>>>(a);
            else k_update_lds<T, 1, 8, 512><<<nc, 512, 0, e->st>>>(a);
        } else {
            k_update_reg<T, 2, 4><<<nc, UPD2_THREADS, 0, e->st>>>(a);
        }
        LCHK();
        return 0;
    }
    UpdArgs<T> ua;
    ua.N = N; ua.L = e->L; ua.l = l; ua.kd = e->kd;
    ua.G = dp<T>(e->G); ua.sG = NN;
    ua.Ab = dp<T>(e->Ab); ua.Bb = dp<T>(e->Bb); ua.sAB = (long)N * e->kd;
    ua.del = e->del; ua.exp_lambda = e->exp_lambda; ua.hs = e->hs;
    ua.order = e->order; ua.unif = e->unif;
    ua.phase = dp<T>(e->phase); ua.escale = e->escale; ua.iscale = e->iscale;
    ua.n_acc = e->nacc; ua.n_acc_sl = e->nacc_sl; ua.active = e->active;
    ua.Lc = e->ragged ? e->d_Lc : nullptr;
    k_update<T><<<nc, UPD_THREADS, 0, e->st>>>(ua);
    LCHK();
    return 0;
}

// d[b / 2] <- max(d[b / 2], max |Gw[b] - G[b]|) for active chains; the bits of
// a non-negative double order as the values, so the maximum is an integer one
template <typename T>
__global__ __launch_bounds__(256) void k_drift(int N, const T* Gw, const T* G, long NN,
                                               const int* active, unsigned long long* d)
{
    const int b = blockIdx.y;
    if (!active[b >> 1]) return;
    const T* x = Gw + b * NN;
    const T* y = G + b * NN;
    double m = 0.0;
    for (long e = blockIdx.x * 256L + threadIdx.x; e < NN; e += (long)gridDim.x * 256)
        m = fmax(m, absv(x[e] - y[e]));
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(d + (b >> 1), (unsigned long long)__double_as_longlong(m));
}

// One sweep of every chain (device work only: kernel launches on the engine
// stream, so that it can be captured into a HIP graph): flags, the sweep's
// random stream, L slices of updates with wraps or recalculations, the
// equal-time measurements and, if due, the unequal-time ones.
template <typename T>
static int sweep_body(dqmc_engine* e, bool any_eq, bool any_ue)
{
    const int N = e->N, L = e->L, nc = e->nc, B = 2 * nc, n = e->n_matmul, F = e->F;
    const long NN = (long)N * N, LN = (long)L * N;
    const T one = from_re<T>(1.0), zero = from_re<T>(0.0);
    T* G = dp<T>(e->G);
    T* T1 = dp<T>(e->T1);
    k_sweep_flags<<<(nc + 255) / 256, 256, 0, e->st>>>(nc, e->period_uneqlt, e->d_sweep,
                                                        e->d_nsweep, e->d_warm, e->active,
                                                        e->flag_eq, e->flag_ue);
    LCHK();
    {
        // the sweep's random stream: site orders and uniforms (dq_rng.hpp)
        ProfSec ps(e, PS_updates);
        k_rng_raw<<<nc, 64, 0, e->st>>>(N, L, e->draws, e->rng, e->raw, 2L * L * N, e->active,
                                        e->d_Lc, e->d_Nc);
        LCHK();
        k_rng_orders<<<nc, 256, 0, e->st>>>(N, L, e->raw, 2L * L * N, e->del, e->draws, e->order,
                                            e->unif, e->active, e->d_Lc, e->d_Nc);
        LCHK();
    }
    for (int l = 0; l < L; l++) {
        {
            ProfSec ps(e, PS_updates);
            RC(launch_update<T>(e, l));
        }
        const int f = l / n;
        const bool recalc = ((l + 1) % n == 0);
        if (recalc) {
            if (e->check_drift) {  // the wrap this slice would have had
                RC(gemm<T>(e, OP_N, OP_N, N, N, N, one, G, N, NN, dp<T>(e->iexpK), N, NN, zero, T1,
                           N, NN, B, nullptr, 0, e->iscale + (long)N * l, LN, nullptr, 0, l, 2));
                RC(gemm<T>(e, OP_N, OP_N, N, N, N, one, dp<T>(e->expK), N, NN, T1, N, NN, zero,
                           dp<T>(e->Gw), N, NN, B, nullptr, 0, e->escale + (long)N * l, LN, nullptr,
                           0, l, 1));
            }
            RC(build_C<T>(e, f, 1));
            ProfSec ps(e, PS_recalc);
            RC(calc_eq_g<T>(e, (f + 1) % F, G, dp<T>(e->phb)));
            if (e->check_drift) {
                k_drift<T><<<dim3(nblk(NN, 256 * 16), B), 256, 0, e->st>>>(N, dp<T>(e->Gw), G, NN,
                                                                          e->active, e->d_drift);
                LCHK();
            }
            k_phase_combine<T><<<(nc + 255) / 256, 256, 0, e->st>>>(nc, dp<T>(e->phb),
                                                                     dp<T>(e->phase), e->active);
            LCHK();
        } else {
            // G <- B_l G B_l^-1 = expK diag(e_l) [G diag(e'_l) invexpK]  (src/dqmc.c:294-295)
            ProfSec ps(e, PS_wrap);
            RC(gemm<T>(e, OP_N, OP_N, N, N, N, one, G, N, NN, dp<T>(e->iexpK), N, NN, zero, T1,
                       N, NN, B, nullptr, 0, e->iscale + (long)N * l, LN, nullptr, 0, l, 2));
            RC(gemm<T>(e, OP_N, OP_N, N, N, N, one, dp<T>(e->expK), N, NN, T1, N, NN, zero, G,
                       N, NN, B, nullptr, 0, e->escale + (long)N * l, LN, nullptr, 0, l, 1));
        }
        if (any_eq && e->meas_at[l]) {
            {
                // half-wrap e^{dtK/2} G e^{-dtK/2} (src/dqmc.c:364-380), into Y2
                ProfSec ps(e, PS_half_wrap);
                RC(gemm<T>(e, OP_N, OP_N, N, N, N, one, G, N, NN, dp<T>(e->ehalf), N, NN, zero,
                           T1, N, NN, B));
                RC(gemm<T>(e, OP_N, OP_N, N, N, N, one, dp<T>(e->iehalf), N, NN, T1, N, NN, zero,
                           dp<T>(e->Y2), N, NN, B));
            }
            ProfSec ps(e, PS_meas_eq);
            MeasArgs<T> ma = meas_args<T>(e, 0);
            ma.G = dp<T>(e->Y2);
            ma.sG = NN;
            if (e->ragged) {  // chains past their own L, or not due at l, do not measure
                k_flag_slice<<<(nc + 255) / 256, 256, 0, e->st>>>(nc, e->flag_eq, e->d_Lc, e->d_pe,
                                                                   l, e->flag_sl);
                LCHK();
                ma.flag = e->flag_sl;
            }
            if (e->ij_percol && e->num_i <= 256) {
                k_meas_eq_strip<T, is_cplx<T>::value ? 8 : 16><<<nc, 256, 0, e->st>>>(ma);
            } else {
                const int nb = (std::max(e->num_i, e->num_ij) + 255) / 256;
                k_meas_eq<T><<<dim3(nb, nc), 256, 0, e->st>>>(ma);
            }
            LCHK();
            if (e->energy) {
                const int nb2 = (std::max(e->num_bs, e->num_bb) + 255) / 256;
                k_meas_eq_energy<T><<<dim3(nb2, nc), 256, 0, e->st>>>(ma);
                LCHK();
            }
            RC(meas_eq_extra<T>(e, ma));
        }
    }
    if (any_ue) RC(ue_measure<T>(e));
    k_sweep_advance<<<(nc + 255) / 256, 256, 0, e->st>>>(nc, e->d_sweep, e->active);
    LCHK();
    return 0;
}

template <typename T>
static int run_t(dqmc_engine* e, int n_sweeps, const volatile int* stop);

template <typename T>
static int run_prof(dqmc_engine* e, int n_sweeps, const volatile int* stop)
{
    if (!e->prof) return run_t<T>(e, n_sweeps, stop);
    const auto t0 = std::chrono::steady_clock::now();
    const int rc = run_t<T>(e, n_sweeps, stop);
    prof_collect(e);
    e->prof_ms[PS_wall] +=
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    e->prof_calls[PS_wall]++;
    return rc;
}

template <typename T>
static int run_t(dqmc_engine* e, int n_sweeps, const volatile int* stop)
{
    const int nc = e->nc;
    if (e->stale) RC(init_green<T>(e));
    for (int s = 0; s < n_sweeps; s++) {
        if (s >= 2) HIPCHK(hipEventSynchronize(e->done[s & 1]));  // sweep s - 2 has ended
        if (stop && *stop) {
            HIPCHK(hipStreamSynchronize(e->st));
            return DQMC_INTERRUPTED;
        }
        bool any_active = false, any_eq = false, any_ue = false;
        for (int c = 0; c < nc; c++) {
            const bool act = e->sweep[c] < e->n_sweep[c];
            any_active |= act;
            any_eq |= act && e->sweep[c] >= e->n_sweep_warm[c] && e->pe[c] > 0;
            any_ue |= act && e->sweep[c] >= e->n_sweep_warm[c] && e->period_uneqlt > 0 &&
                      e->sweep[c] % e->period_uneqlt == 0;
        }
        if (!any_active) break;
        const int key = (any_eq ? 1 : 0) | (any_ue ? 2 : 0);
        if (e->use_graph && !e->time_gemm && !e->prof) {
            if (!e->gexec[key]) {
                HIPCHK(hipStreamBeginCapture(e->st, hipStreamCaptureModeRelaxed));
                const int rc = sweep_body<T>(e, any_eq, any_ue);
                hipGraph_t g = nullptr;
                const hipError_t ce = hipStreamEndCapture(e->st, &g);
                if (rc) {
                    if (g) (void)hipGraphDestroy(g);
                    return rc;
                }
                if (ce != hipSuccess)
                    return set_err(DQMC_ERR_HIP, "sweep capture: %s", hipGetErrorString(ce));
                hipGraphExec_t x = nullptr;
                const hipError_t ie = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
                if (ie != hipSuccess) {
                    (void)hipGraphDestroy(g);
                    return set_err(DQMC_ERR_HIP, "sweep graph: %s", hipGetErrorString(ie));
                }
                e->graph[key] = g;
                e->gexec[key] = x;
            }
            HIPCHK(hipGraphLaunch(e->gexec[key], e->st));
        } else {
            RC(sweep_body<T>(e, any_eq, any_ue));
        }
        HIPCHK(hipEventRecord(e->done[s & 1], e->st));
        for (int c = 0; c < nc; c++)
            if (e->sweep[c] < e->n_sweep[c]) e->sweep[c]++;
    }
    HIPCHK(hipStreamSynchronize(e->st));
    return 0;
}

// --------------------------------------------------------------------------
// C ABI
// --------------------------------------------------------------------------
// chains of one engine: equal shapes and measurement sets; the equal-time
// cadence may differ between chains padded to a common size (padded)
static int same_shape(const dqmc_params* a, const dqmc_params* b, bool padded)
{
    return a->N == b->N && a->L == b->L && a->F == b->F && a->n_matmul == b->n_matmul &&
           a->is_complex == b->is_complex && (padded || a->period_eqlt == b->period_eqlt) &&
           a->period_uneqlt == b->period_uneqlt && a->num_i == b->num_i &&
           a->num_ij == b->num_ij && a->meas_energy_corr == b->meas_energy_corr &&
           a->meas_bond_corr == b->meas_bond_corr && a->meas_thermal == b->meas_thermal &&
           a->meas_2bond_corr == b->meas_2bond_corr && a->meas_chiral == b->meas_chiral &&
           a->meas_gen_suscept == b->meas_gen_suscept && a->meas_local_JQ == b->meas_local_JQ &&
           a->meas_nematic_corr == b->meas_nematic_corr &&
           a->meas_pair_bb_only == b->meas_pair_bb_only && a->num_b == b->num_b &&
           a->num_bb == b->num_bb && a->num_bs == b->num_bs;
}

// Bin lists of a site / pair map.  Entries outside [0, nbins) belong to no
// bin (the padding sites of chains padded to a common N map to -1).
static void csr_of_map(const int32_t* map, int n_in, int nbins, std::vector<int>& rp,
                       std::vector<int>& ix)
{
    rp.assign(nbins + 1, 0);
    for (int q = 0; q < n_in; q++)
        if (map[q] >= 0 && map[q] < nbins) rp[map[q] + 1]++;
    for (int r = 0; r < nbins; r++) rp[r + 1] += rp[r];
    std::vector<int> fill(rp.begin(), rp.end() - 1);
    ix.assign(n_in, 0);
    for (int q = 0; q < n_in; q++)  // increasing q within a bin
        if (map[q] >= 0 && map[q] < nbins) ix[fill[map[q]]++] = q;
}

// chain c's numeric parameters (the exponentials, exp_lambda, Delta and the
// Peierls phases) onto the engine stream
static int upload_numeric(dqmc_engine* e, int c, const dqmc_params* p)
{
    const long N = e->N, NN = N * N;
    const size_t mb = NN * e->esz;
    const void* src[8] = {p->exp_Ku, p->exp_Kd, p->inv_exp_Ku, p->inv_exp_Kd,
                          p->exp_halfKu, p->exp_halfKd, p->inv_exp_halfKu, p->inv_exp_halfKd};
    void* dst[4] = {e->expK, e->iexpK, e->ehalf, e->iehalf};
    for (int k = 0; k < 8; k++)
        HIPCHK(hipMemcpyAsync((char*)dst[k / 2] + (2 * c + (k & 1)) * mb, src[k], mb,
                              hipMemcpyHostToDevice, e->st));
    HIPCHK(hipMemcpyAsync(e->exp_lambda + c * 2L * N, p->exp_lambda, 2L * N * 8, hipMemcpyHostToDevice, e->st));
    HIPCHK(hipMemcpyAsync(e->del + c * 2L * N, p->del, 2L * N * 8, hipMemcpyHostToDevice, e->st));
    if (e->cplx) {
        HIPCHK(hipMemcpyAsync((char*)e->peierlsu + c * mb, p->peierlsu, mb, hipMemcpyHostToDevice, e->st));
        HIPCHK(hipMemcpyAsync((char*)e->peierlsd + c * mb, p->peierlsd, mb, hipMemcpyHostToDevice, e->st));
    }
    const long nb2 = e->num_b2;  // phases of the two-hop bonds
    if (nb2 > 0 && p->pp_u && p->ppr_u && p->pp_d && p->ppr_d) {
        const double* src[4] = {p->pp_u, p->ppr_u, p->pp_d, p->ppr_d};
        for (int q = 0; q < 4; q++)
            HIPCHK(hipMemcpyAsync((char*)e->pp + ((long)c * 4 + q) * nb2 * e->esz, src[q],
                                  nb2 * e->esz, hipMemcpyHostToDevice, e->st));
    }
    return 0;
}

static int engine_setup(dqmc_engine* e, const dqmc_params* const* P)
{
    const dqmc_params* p0 = P[0];
    const int nc = e->nc, N = e->N, L = e->L, F = e->F, B = 2 * nc;
    const long NN = (long)N * N;
    const size_t es = e->esz;
    // check the measurement set against what the GPU computes
    for (int k = 0; k < DQ_EQ_COUNT; k++)
        if (dqmc_acc_len(p0, 0, k) > 0 && !eq_supported(k, e->cplx))
            return set_err(DQMC_ERR_UNSUPPORTED,
                           "equal-time accumulator %d is enabled but not implemented on the GPU", k);
    for (int k = 0; k < DQ_UE_COUNT; k++)
        if (dqmc_acc_len(p0, 1, k) > 0 && !ue_supported(k, e->cplx))
            return set_err(DQMC_ERR_UNSUPPORTED,
                           "unequal-time accumulator %d is not available in complex builds "
                           "(peierls phase not implemented for nematic measurements)",
                           k);
    DALLOC(e->expK, B * NN * es);
    DALLOC(e->iexpK, B * NN * es);
    DALLOC(e->ehalf, B * NN * es);
    DALLOC(e->iehalf, B * NN * es);
    DALLOC(e->exp_lambda, nc * 2L * N * 8);
    DALLOC(e->del, nc * 2L * N * 8);
    DALLOC(e->hs, (long)nc * L * N * 4);
    DALLOC(e->rng, nc * 17 * 8);
    DALLOC(e->G, B * NN * es);
    DALLOC(e->C, B * F * NN * es);
    DALLOC(e->W, B * F * NN * es);
    DALLOC(e->Ab, (long)B * N * e->kd * es);
    DALLOC(e->Bb, (long)B * N * e->kd * es);
    DALLOC(e->escale, (long)B * L * N * 8);
    DALLOC(e->iscale, (long)B * L * N * 8);
    void** work[] = {&e->X, &e->Y, &e->Y2, &e->Z, &e->Q, &e->Tm, &e->Rs, &e->Tp, &e->Mm, &e->T1};
    for (void** w : work) DALLOC(*w, B * NN * es);
    DALLOC(e->tau, (long)B * 2 * N * es);
    e->sVb = (e->period_uneqlt > 0 ? 4L : 1L) * NN;
    e->sTb = 2L * N * 64 + 64 * 64;
    e->sWq = 64L * 2 * N;
    DALLOC(e->Vb, B * e->sVb * es);
    DALLOC(e->Tb, B * e->sTb * es);
    DALLOC(e->Wq1, B * e->sWq * es);
    DALLOC(e->Wq2, B * e->sWq * es);
    DALLOC(e->detq, B * es);
    DALLOC(e->detm, B * es);
    DALLOC(e->jpvt, (long)B * N * 4);
    DALLOC(e->perm, (long)B * N * 4);
    DALLOC(e->ipiv, (long)B * N * 4);
    DALLOC(e->dvec, (long)B * N * 8);
    DALLOC(e->vvec, (long)B * N * 8);
    DALLOC(e->phb, B * es);
    DALLOC(e->phb2, B * es);
    DALLOC(e->phase, nc * es);
    DALLOC(e->nacc, nc * 8);
    DALLOC(e->nacc_sl, (long)nc * L * 8);
    if (e->check_drift) {
        DALLOC(e->Gw, B * NN * es);
        DALLOC(e->d_drift, nc * 8);
    }
    DALLOC(e->raw, (long)nc * 2 * L * N * 8);
    DALLOC(e->order, (long)nc * L * N * 4);
    DALLOC(e->unif, (long)nc * L * N * 8);
    DALLOC(e->draws, nc * 4);
    DALLOC(e->active, nc * 4);
    DALLOC(e->flag_eq, nc * 4);
    DALLOC(e->flag_ue, nc * 4);
    DALLOC(e->d_sweep, nc * 4);
    DALLOC(e->d_nsweep, nc * 4);
    DALLOC(e->d_warm, nc * 4);
    // accumulators
    e->eq_total = 0;
    for (int k = 0; k < DQMC_MAX_ACC; k++) {
        e->eq_len[k] = k < DQ_EQ_COUNT ? dqmc_acc_len(p0, 0, k) : 0;
        e->eq_off[k] = e->eq_len[k] > 0 ? e->eq_total : 0;
        e->eq_total += e->eq_len[k];
        e->ue_len[k] = k < DQ_UE_COUNT ? dqmc_acc_len(p0, 1, k) : 0;
        e->ue_off[k] = e->ue_len[k] > 0 ? e->ue_total : 0;
        e->ue_total += e->ue_len[k];
    }
    DALLOC(e->d_eq_off, DQMC_MAX_ACC * 8);
    DALLOC(e->d_ue_off, DQMC_MAX_ACC * 8);
    HIPCHK(hipMemcpyAsync(e->d_eq_off, e->eq_off, sizeof e->eq_off, hipMemcpyHostToDevice, e->st));
    HIPCHK(hipMemcpyAsync(e->d_ue_off, e->ue_off, sizeof e->ue_off, hipMemcpyHostToDevice, e->st));
    DALLOC(e->acc_eq, (long)nc * e->eq_total * es);
    DALLOC(e->acc_ue, (long)nc * (e->ue_total ? e->ue_total : 1) * es);
    DALLOC(e->sign_eq, nc * es);
    DALLOC(e->sign_ue, nc * es);
    DALLOC(e->ns_eq, nc * 4);
    DALLOC(e->ns_ue, nc * 4);
    DALLOC(e->rp_i, (long)nc * (e->num_i + 1) * 4);
    DALLOC(e->ix_i, (long)nc * N * 4);
    DALLOC(e->rp_ij, (long)nc * (e->num_ij + 1) * 4);
    DALLOC(e->ix_ij, (long)nc * NN * 4);
    DALLOC(e->degen, nc * 4 * 4);
    DALLOC(e->bonds, (long)nc * 2 * (e->num_b > 0 ? e->num_b : 1) * 4);
    DALLOC(e->rp_bb, (long)nc * (e->num_bb + 1) * 4);
    DALLOC(e->ix_bb, (long)nc * ((long)e->num_b * e->num_b > 0 ? (long)e->num_b * e->num_b : 1) * 4);
    DALLOC(e->rp_bs, (long)nc * (e->num_bs + 1) * 4);
    DALLOC(e->ix_bs, (long)nc * ((long)e->num_b * N > 0 ? (long)e->num_b * N : 1) * 4);
    DALLOC(e->peierlsu, (e->cplx ? nc * NN : 1) * es);
    DALLOC(e->peierlsd, (e->cplx ? nc * NN : 1) * es);
    // optional groups: which extra maps are needed, their input counts and bins
    const bool ueon = e->period_uneqlt > 0;
    const long nb1 = e->num_b, nb2 = e->num_b2;
    const bool need[PM_COUNT] = {ueon && (e->thermal || e->twobond), ueon && (e->thermal || e->twobond),
                                 ueon && e->twobond, (bool)e->chiral, (bool)e->localjq, (bool)e->localjq};
    const long nin[PM_COUNT] = {nb2 * nb1, nb1 * nb2, nb2 * nb2, e->num_plaq, nb1, nb2};
    const int nbin[PM_COUNT] = {e->num_bb2, e->num_b2b, e->num_b2b2, e->num_plaq_acc, e->num_b_acc,
                                e->num_b2_acc};
    const int32_t* const* pmaps[PM_COUNT] = {&p0->map_bb2, &p0->map_b2b, &p0->map_b2b2, &p0->map_plaq,
                                             &p0->map_b, &p0->map_b2};
    (void)pmaps;
    for (int m = 0; m < PM_COUNT; m++) {
        if (!need[m]) continue;
        if (nbin[m] <= 0 || nin[m] <= 0)
            return set_err(DQMC_ERR_ARG, "optional measurement map %d is empty", m);
        e->srp2[m] = nbin[m] + 1L;
        e->six2[m] = nin[m];
        DALLOC(e->rp2[m], (long)nc * e->srp2[m] * 4);
        DALLOC(e->ix2[m], (long)nc * e->six2[m] * 4);
    }
    DALLOC(e->degen2, (long)nc * PM_COUNT * 4);
    // does every column j of map_ij feed every bin exactly once (translation
    // symmetry)?  Then the strip kernel can walk columns with an inverse table.
    e->ij_percol = e->num_ij == N && N <= UES_NMAX;
    for (int c = 0; c < nc && e->ij_percol; c++) {
        std::vector<char> seen(N);
        for (int j = 0; j < N && e->ij_percol; j++) {
            std::fill(seen.begin(), seen.end(), 0);
            for (int i = 0; i < N; i++) {
                const int r = P[c]->map_ij[i + (long)N * j];
                if (r < 0 || r >= N || seen[r]) { e->ij_percol = false; break; }
                seen[r] = 1;
            }
        }
    }
    if (e->ij_percol) DALLOC(e->invij, (long)nc * N * N * 4);
    DALLOC(e->bond2s, (long)nc * 2 * (nb2 > 0 ? nb2 : 1) * 4);
    DALLOC(e->pp, (long)nc * 4 * (nb2 > 0 ? nb2 : 1) * es);
    DALLOC(e->plaqs, (long)nc * 3 * (e->num_plaq > 0 ? e->num_plaq : 1) * 4);
    if (e->gensus) DALLOC(e->gs, (long)nc * L * 4 * e->num_ij * es);
    if (e->period_uneqlt > 0) {
        const long LNN = (long)L * NN;
        DALLOC(e->G0t, B * LNN * es);
        DALLOC(e->Gtt, B * LNN * es);
        DALLOC(e->Gt0, B * LNN * es);
        RC(ue_alloc(e));
        if ((e->bond || e->thermal || e->pbbonly || e->twobond || e->energy) && e->num_b > 0) {
            DALLOC(e->btab, B * L * 3L * e->num_b * es);
            DALLOC(e->bph, (long)nc * 4 * e->num_b * es);
        }
        if ((e->thermal || e->twobond) && e->num_b2 > 0)
            DALLOC(e->b2tab, B * L * 3L * e->num_b2 * es);
    }
    // per-chain parameters
    std::vector<int> rp, ix, deg(4 * nc);
    for (int c = 0; c < nc; c++) {
        const dqmc_params* p = P[c];
        const size_t mb = NN * es;
        (void)mb;
        RC(upload_numeric(e, c, p));
        csr_of_map(p->map_i, N, e->num_i, rp, ix);
        HIPCHK(hipMemcpy(e->rp_i + c * (e->num_i + 1L), rp.data(), rp.size() * 4, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(e->ix_i + c * (long)N, ix.data(), ix.size() * 4, hipMemcpyHostToDevice));
        if (e->ij_percol) {
            std::vector<int> inv((size_t)N * N);
            for (int j = 0; j < N; j++)
                for (int i = 0; i < N; i++) inv[(size_t)j * N + p->map_ij[i + (long)N * j]] = i;
            HIPCHK(hipMemcpy(e->invij + c * NN, inv.data(), NN * 4, hipMemcpyHostToDevice));
        }
        csr_of_map(p->map_ij, (int)NN, e->num_ij, rp, ix);
        HIPCHK(hipMemcpy(e->rp_ij + c * (e->num_ij + 1L), rp.data(), rp.size() * 4, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(e->ix_ij + c * NN, ix.data(), ix.size() * 4, hipMemcpyHostToDevice));
        if (e->num_b > 0) {
            const long nb = e->num_b;
            HIPCHK(hipMemcpy(e->bonds + c * 2 * nb, p->bonds, 2 * nb * 4, hipMemcpyHostToDevice));
            csr_of_map(p->map_bb, (int)(nb * nb), e->num_bb, rp, ix);
            HIPCHK(hipMemcpy(e->rp_bb + c * (e->num_bb + 1L), rp.data(), rp.size() * 4, hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(e->ix_bb + c * nb * nb, ix.data(), ix.size() * 4, hipMemcpyHostToDevice));
            csr_of_map(p->map_bs, (int)(nb * N), e->num_bs, rp, ix);
            HIPCHK(hipMemcpy(e->rp_bs + c * (e->num_bs + 1L), rp.data(), rp.size() * 4, hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(e->ix_bs + c * nb * N, ix.data(), ix.size() * 4, hipMemcpyHostToDevice));
        }
        {
            const int32_t* maps[PM_COUNT] = {p->map_bb2, p->map_b2b, p->map_b2b2, p->map_plaq,
                                             p->map_b, p->map_b2};
            for (int m = 0; m < PM_COUNT; m++) {
                if (!need[m]) continue;
                if (!maps[m]) return set_err(DQMC_ERR_ARG, "chain %d: optional map %d missing", c, m);
                csr_of_map(maps[m], (int)nin[m], nbin[m], rp, ix);
                HIPCHK(hipMemcpy(e->rp2[m] + c * e->srp2[m], rp.data(), rp.size() * 4, hipMemcpyHostToDevice));
                HIPCHK(hipMemcpy(e->ix2[m] + c * e->six2[m], ix.data(), ix.size() * 4, hipMemcpyHostToDevice));
            }
            const int d2[PM_COUNT] = {p->degen_bb2, p->degen_b2b, p->degen_b2b2, p->degen_plaq,
                                      p->degen_b, p->degen_b2};
            HIPCHK(hipMemcpy(e->degen2 + c * PM_COUNT, d2, sizeof d2, hipMemcpyHostToDevice));
            if (nb2 > 0 && p->bond2s)
                HIPCHK(hipMemcpy(e->bond2s + c * 2 * nb2, p->bond2s, 2 * nb2 * 4, hipMemcpyHostToDevice));
            // (the two-hop bond phases pp came with upload_numeric)
            if (e->num_plaq > 0 && p->plaqs)
                HIPCHK(hipMemcpy(e->plaqs + c * 3L * e->num_plaq, p->plaqs, 3L * e->num_plaq * 4,
                                 hipMemcpyHostToDevice));
        }
        deg[4 * c] = p->degen_i;
        deg[4 * c + 1] = p->degen_ij;
        deg[4 * c + 2] = p->degen_bb;
        deg[4 * c + 3] = p->degen_bs;
        e->n_sweep[c] = p->n_sweep;
        e->n_sweep_warm[c] = p->n_sweep_warm;
    }
    HIPCHK(hipMemcpy(e->degen, deg.data(), deg.size() * 4, hipMemcpyHostToDevice));
    // sites that consume a uniform (Delta's not both zero, src/updates.c:21)
    std::vector<int> nd(nc);
    for (int c = 0; c < nc; c++) {
        int n = 0;
        for (int i = 0; i < N; i++) n += !(P[c]->del[i] == 0.0 && P[c]->del[i + N] == 0.0);
        nd[c] = n;
    }
    HIPCHK(hipMemcpy(e->draws, nd.data(), nc * 4, hipMemcpyHostToDevice));
    if (e->ragged) {
        DALLOC(e->d_Lc, nc * 4);
        DALLOC(e->d_Nc, nc * 4);
        DALLOC(e->flag_sl, nc * 4);
        DALLOC(e->d_pe, nc * 4);
        HIPCHK(hipMemcpy(e->d_pe, e->pe.data(), nc * 4, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(e->d_Lc, e->Lc.data(), nc * 4, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(e->d_Nc, e->Nc.data(), nc * 4, hipMemcpyHostToDevice));
    }
    HIPCHK(hipMemcpy(e->d_nsweep, e->n_sweep.data(), nc * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(e->d_warm, e->n_sweep_warm.data(), nc * 4, hipMemcpyHostToDevice));
    HIPCHK(hipStreamSynchronize(e->st));
    return 0;
}

// Times every GEMM tile shape on the engine's dominant product (N x N x N over
// all 2 n_chains matrices: the wraps and the B-chain products) and keeps the
// fastest.  Operands are the engine's own buffers; their contents are
// irrelevant and the output buffer is scratch at this point.
template <typename T>
static int autotune_gemm(dqmc_engine* e)
{
    const int N = e->N, B = 2 * e->nc;
    const long NN = (long)N * N;
    const int cand[4] = {GT_64, GT_128x64, GT_128, GT_W};
    const int ncand = is_cplx<T>::value ? 2 : (N % 128 == 0 ? 4 : 3);
    hipEvent_t a0, a1;
    HIPCHK(hipEventCreate(&a0));
    HIPCHK(hipEventCreate(&a1));
    float best = 1e30f;
    int pick = GT_64;
    int rc = 0;
    for (int c = 0; c < ncand && !rc; c++) {
        e->gemm_tile = cand[c];
        const int reps = 3;
        for (int r = 0; r <= reps && !rc; r++) {
            if (r == 1) (void)hipEventRecord(a0, e->st);
            rc = gemm<T>(e, OP_N, OP_N, N, N, N, from_re<T>(1.0), dp<T>(e->expK), N, NN,
                         dp<T>(e->G), N, NN, from_re<T>(0.0), dp<T>(e->X), N, NN, B);
        }
        if (rc) break;
        (void)hipEventRecord(a1, e->st);
        if (hipEventSynchronize(a1) != hipSuccess) { rc = set_err(DQMC_ERR_HIP, "autotune sync"); break; }
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, a0, a1);
        if (ms < best) { best = ms; pick = cand[c]; }
    }
    (void)hipEventDestroy(a0);
    (void)hipEventDestroy(a1);
    e->gemm_tile = pick;
    return rc;
}

// d[c n + i] = d[i] for chains c = 1 .. nc-1 (n doubles per chain)
__global__ __launch_bounds__(256) void k_bcast_chain0(double* d, long n)
{
    const long c = blockIdx.y + 1;
    for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256)
        d[c * n + i] = d[i];
}
__global__ __launch_bounds__(256) void k_fill_i32(int* d, int n, int v)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) d[i] = v;
}

extern "C" int dqmc_engine_set_params(dqmc_engine* e, int c, const dqmc_params* p)
{
    if (!e || !p || c < -1 || c >= e->nc) return set_err(DQMC_ERR_ARG, "bad arguments");
    if (p->N != e->N || p->L != e->L || p->n_matmul != e->n_matmul ||
        (p->is_complex != 0) != (e->cplx != 0))
        return set_err(DQMC_ERR_ARG, "parameters of another shape than the engine's");
    if (e->ragged) return set_err(DQMC_ERR_UNSUPPORTED, "padded chains keep their parameters");
    HIPCHK(hipSetDevice(e->device));
    HIPCHK(hipStreamSynchronize(e->st));
    const bool all = c < 0;
    RC(upload_numeric(e, all ? 0 : c, p));
    if (all && e->nc > 1) {  // one upload, then chain 0's blocks to every chain
        const long N = e->N, NN = N * N, ed = e->esz / 8;
        struct { void* d; long n; } arr[] = {
            {e->expK, 2 * NN * ed}, {e->iexpK, 2 * NN * ed}, {e->ehalf, 2 * NN * ed},
            {e->iehalf, 2 * NN * ed}, {e->exp_lambda, 2 * N}, {e->del, 2 * N},
            {e->cplx ? e->peierlsu : nullptr, NN * ed}, {e->cplx ? e->peierlsd : nullptr, NN * ed},
            {e->num_b2 > 0 && p->pp_u ? e->pp : nullptr, 4L * e->num_b2 * ed}};
        for (auto& a : arr) {
            if (!a.d) continue;
            const long nb = std::min<long>(nblk(a.n, 256), 64);
            k_bcast_chain0<<<dim3(nb, e->nc - 1), 256, 0, e->st>>>((double*)a.d, a.n);
            LCHK();
        }
    }
    // sites that consume a uniform (Delta's not both zero, src/updates.c:21)
    int nd = 0;
    for (int i = 0; i < e->N; i++) nd += !(p->del[i] == 0.0 && p->del[i + e->N] == 0.0);
    if (all) {
        k_fill_i32<<<nblk(e->nc, 256), 256, 0, e->st>>>(e->draws, e->nc, nd);
        LCHK();
    } else {
        HIPCHK(hipMemcpyAsync(e->draws + c, &nd, 4, hipMemcpyHostToDevice, e->st));
    }
    HIPCHK(hipStreamSynchronize(e->st));
    e->stale = true;  // G is recomputed before the next sweep
    return 0;
}

extern "C" int dqmc_engine_adjoint(dqmc_engine* e, double* logdet, double* W, double* trG,
                                   double* sG)
{
    if (!e) return set_err(DQMC_ERR_ARG, "null engine");
    if (e->ragged) return set_err(DQMC_ERR_UNSUPPORTED, "the adjoint needs chains of equal size");
    HIPCHK(hipSetDevice(e->device));
    return e->cplx ? adjoint_t<zd>(e, logdet, W, trG, sG) : adjoint_t<double>(e, logdet, W, trG, sG);
}

// The unequal-time buffers of an engine whose files measure no unequal-time
// quantities (allocated on first use by the observable adjoint): the G blocks,
// bsofi's work space, and Householder vectors for its 2N-row panels (4 N^2 per
// matrix where calc_eq_g alone needs N^2).  The sweep graphs captured with the
// smaller vector block are dropped; that block stays allocated until destroy.
static int ue_ensure(dqmc_engine* e)
{
    if (e->G0t) return 0;
    HIPCHK(hipStreamSynchronize(e->st));  // no graph of this engine still running
    const long NN = (long)e->N * e->N, LNN = (long)e->L * NN, B = 2L * e->nc;
    DALLOC(e->G0t, B * LNN * e->esz);
    DALLOC(e->Gtt, B * LNN * e->esz);
    DALLOC(e->Gt0, B * LNN * e->esz);
    RC(ue_alloc(e));
    if (e->sVb < 4 * NN) {
        for (int k = 0; k < 4; k++) {
            if (e->gexec[k]) (void)hipGraphExecDestroy(e->gexec[k]);
            if (e->graph[k]) (void)hipGraphDestroy(e->graph[k]);
            e->gexec[k] = nullptr;
            e->graph[k] = nullptr;
        }
        e->sVb = 4 * NN;
        DALLOC(e->Vb, B * e->sVb * e->esz);
    }
    return 0;
}

// Observable adjoint: O[b] = tr(X_s G) at slice 0 and, with the unequal-time
// Green's functions of calc_ue_g in the plain frame, the sums
// W_X = sum_l e^{V_l} Z_l e^{-V_l}, sum_l tr Z_l and sum_l sum_i sigma s_li
// (Z_l)_ii with Z_l = G(t_l,0) X G^(0,t_l) (dq_adjoint.hpp).  Xh: the two
// spins' X, column-major, shared by every chain.
template <typename T>
static int adjoint_green_t(dqmc_engine* e, const double* Xh, double* obs, double* Wout,
                           double* trZ, double* sZ)
{
    const int N = e->N, L = e->L, nc = e->nc, B = 2 * nc;
    const long NN = (long)N * N, LN = (long)L * N, LNN = (long)L * NN;
    const T one = from_re<T>(1.0), zero = from_re<T>(0.0);
    RC(ue_ensure(e));
    if (!e->adjW) {
        DALLOC(e->adjW, B * NN * e->esz);
        DALLOC(e->adjT, 2L * B * e->esz);
        DALLOC(e->adjL, B * 8);
    }
    if (!e->adjX) {
        DALLOC(e->adjX, B * NN * e->esz);
        DALLOC(e->adjZ, B * NN * e->esz);
        DALLOC(e->adjO, B * e->esz);
    }
    {
        std::vector<T> xb((size_t)B * NN);
        for (int b = 0; b < B; b++)
            memcpy(xb.data() + (size_t)b * NN, (const char*)Xh + (size_t)(b & 1) * NN * sizeof(T),
                   NN * sizeof(T));
        HIPCHK(hipMemcpyAsync(e->adjX, xb.data(), xb.size() * sizeof(T), hipMemcpyHostToDevice,
                              e->st));
        HIPCHK(hipStreamSynchronize(e->st));  // xb leaves scope
    }
    T* X = dp<T>(e->adjX);
    T* Z = dp<T>(e->adjZ);
    T* W = dp<T>(e->adjW);
    T* tr = dp<T>(e->adjT);
    T* sg = tr + B;
    T* T1 = dp<T>(e->T1);
    T* G0t = dp<T>(e->G0t);
    T* Gt0 = dp<T>(e->Gt0);
    k_scales<<<nblk((long)nc * L * N), 256, 0, e->st>>>(N, L, nc, e->hs, e->exp_lambda, e->escale,
                                                        e->iscale);
    LCHK();
    RC(build_C<T>(e, 0, e->F));
    RC(ue_calc<T>(e, false));
    k_adj_trace_prod<T><<<B, 64, 0, e->st>>>(N, X, NN, dp<T>(e->Gtt), LNN, dp<T>(e->adjO));
    LCHK();
    // G^(0, 0) = G - I
    k_adj_add_diag<T><<<nblk((long)N * B), 256, 0, e->st>>>(N, G0t, LNN, -1.0, B);
    LCHK();
    {
        const long w = (long)(sizeof(T) / 8);  // doubles per element
        k_zero<<<nblk(B * NN * w), 256, 0, e->st>>>(B * NN * w, (double*)W);
        LCHK();
        k_zero<<<1, 256, 0, e->st>>>(2L * B * w, (double*)tr);
        LCHK();
    }
    for (int l = 0; l < L; l++) {
        // T1 = X G^(0, t_l), Z = G(t_l, 0) T1
        RC(gemm<T>(e, OP_N, OP_N, N, N, N, one, X, N, NN, G0t + (long)l * NN, N, LNN, zero, T1, N,
                   NN, B));
        RC(gemm<T>(e, OP_N, OP_N, N, N, N, one, Gt0 + (long)l * NN, N, LNN, T1, N, NN, zero, Z, N,
                   NN, B));
        k_adj_accum_W<T><<<nblk(B * NN), 256, 0, e->st>>>(N, Z, NN, e->escale, e->iscale, LN, l, W,
                                                           NN, B, 0.0, 1.0);
        LCHK();
        k_adj_accum_diag<T><<<B, 64, 0, e->st>>>(N, L, Z, NN, e->hs, l, tr, sg, 0.0, 1.0);
        LCHK();
    }
    HIPCHK(hipStreamSynchronize(e->st));
    if (obs) HIPCHK(hipMemcpy(obs, e->adjO, B * sizeof(T), hipMemcpyDeviceToHost));
    if (Wout) HIPCHK(hipMemcpy(Wout, W, B * NN * sizeof(T), hipMemcpyDeviceToHost));
    if (trZ) HIPCHK(hipMemcpy(trZ, tr, B * sizeof(T), hipMemcpyDeviceToHost));
    if (sZ) HIPCHK(hipMemcpy(sZ, sg, B * sizeof(T), hipMemcpyDeviceToHost));
    e->stale = true;
    return 0;
}

extern "C" int dqmc_engine_adjoint_green(dqmc_engine* e, const double* X, double* obs, double* W,
                                         double* trZ, double* sZ)
{
    if (!e || !X) return set_err(DQMC_ERR_ARG, "null engine or X");
    if (e->ragged) return set_err(DQMC_ERR_UNSUPPORTED, "the adjoint needs chains of equal size");
    HIPCHK(hipSetDevice(e->device));
    return e->cplx ? adjoint_green_t<zd>(e, X, obs, W, trZ, sZ)
                   : adjoint_green_t<double>(e, X, obs, W, trZ, sZ);
}

extern "C" int dqmc_engine_gemm_tile(dqmc_engine* e)
{
    return e ? e->gemm_tile : 0;
}

extern "C" int dqmc_engine_delay_rank(dqmc_engine* e)
{
    return e ? e->kd : 0;
}

extern "C" int dqmc_engine_create(dqmc_engine** out, const dqmc_params* const* P, int n_chains,
                                  int device, const dqmc_tuning* tune)
{
    if (!out || !P || n_chains <= 0) return set_err(DQMC_ERR_ARG, "bad arguments");
    *out = nullptr;
    const dqmc_params* p0 = P[0];
    if (p0->N <= 0 || p0->N > QR_MAXM || p0->L <= 0 || p0->n_matmul <= 0 ||
        p0->L % p0->n_matmul != 0 || p0->F != p0->L / p0->n_matmul)
        return set_err(DQMC_ERR_ARG, "bad shape N=%d L=%d n_matmul=%d F=%d", p0->N, p0->L,
                       p0->n_matmul, p0->F);
    if (p0->N > UPD_MAXN) return set_err(DQMC_ERR_ARG, "N=%d exceeds %d", p0->N, UPD_MAXN);
    const bool padded = tune && (tune->chain_L || tune->chain_N);
    for (int c = 1; c < n_chains; c++)
        if (!same_shape(p0, P[c], padded))
            return set_err(DQMC_ERR_ARG, "chain %d has a different shape than chain 0", c);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return set_err(DQMC_ERR_HIP, "no HIP device available");
    if (device < 0 || device >= ndev) return set_err(DQMC_ERR_ARG, "bad device %d", device);
    HIPCHK(hipSetDevice(device));
    dqmc_engine* e = new dqmc_engine();
    e->nc = n_chains; e->N = p0->N; e->L = p0->L; e->F = p0->F; e->n_matmul = p0->n_matmul;
    e->cplx = p0->is_complex; e->device = device;
    e->period_eqlt = p0->period_eqlt; e->period_uneqlt = p0->period_uneqlt;
    e->energy = p0->meas_energy_corr;
    e->num_i = p0->num_i; e->num_ij = p0->num_ij;
    e->num_b = p0->num_b; e->num_bb = p0->num_bb; e->num_bs = p0->num_bs;
    e->bond = p0->meas_bond_corr;
    e->thermal = p0->meas_thermal; e->twobond = p0->meas_2bond_corr;
    e->nematic = p0->meas_nematic_corr; e->gensus = p0->meas_gen_suscept;
    e->chiral = p0->meas_chiral; e->localjq = p0->meas_local_JQ;
    e->pbbonly = p0->meas_pair_bb_only;
    e->Nx = p0->Nx; e->Ny = p0->Ny; e->num_b2 = p0->num_b2; e->num_bb2 = p0->num_bb2;
    e->num_b2b = p0->num_b2b; e->num_b2b2 = p0->num_b2b2; e->num_plaq = p0->num_plaq;
    e->num_plaq_acc = p0->num_plaq_accum; e->num_b_acc = p0->num_b_accum;
    e->num_b2_acc = p0->num_b2_accum;
    e->esz = e->cplx ? 16 : 8;
    // the file's n_delay (the flush at k == n_delay, src/updates.c:58), or the
    // tuning's rank, rounded to a compiled variant
    e->kd = upd_rank(e->N, e->cplx,
                     (tune && tune->delay_rank > 0) ? tune->delay_rank
                                                    : (p0->n_delay > 0 ? p0->n_delay : 1 << 30));
    e->time_gemm = tune && (tune->flags & 1);
    e->use_graph = !(tune && (tune->flags & 2));
    e->check_drift = tune && (tune->flags & 4);
    e->prof = tune && (tune->flags & 8);
    if (tune && (tune->chain_L || tune->chain_N)) {
        // chains of uneven size, padded by the caller to N and L
        e->ragged = true;
        e->Lc.resize(n_chains);
        e->Nc.resize(n_chains);
        for (int c = 0; c < n_chains; c++) {
            e->Lc[c] = tune->chain_L ? tune->chain_L[c] : e->L;
            e->Nc[c] = tune->chain_N ? tune->chain_N[c] : e->N;
            if (e->Lc[c] < 1 || e->Lc[c] > e->L || e->Nc[c] < 1 || e->Nc[c] > e->N) {
                const int lc = e->Lc[c], ncs = e->Nc[c];
                delete e;
                return set_err(DQMC_ERR_ARG, "chain %d: own size L=%d N=%d outside 1..%d, 1..%d", c,
                               lc, ncs, p0->L, p0->N);
            }
        }
    }
    e->pe.resize(n_chains);
    e->meas_at.assign(e->L, 0);
    for (int c = 0; c < n_chains; c++) {
        e->pe[c] = P[c]->period_eqlt;
        const int lc = e->ragged ? e->Lc[c] : e->L;
        for (int l = 0; l < lc; l++)
            if (e->pe[c] > 0 && (l + 1) % e->pe[c] == 0) e->meas_at[l] = 1;
    }
    e->sweep.assign(n_chains, 0);
    e->n_sweep.assign(n_chains, 0);
    e->n_sweep_warm.assign(n_chains, 0);
    if (hipStreamCreateWithFlags(&e->st, hipStreamNonBlocking) != hipSuccess) {
        delete e;
        return set_err(DQMC_ERR_HIP, "hipStreamCreate failed");
    }
    for (auto& ev : e->done) {
        if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
            dqmc_engine_destroy(e);
            return set_err(DQMC_ERR_HIP, "hipEventCreate failed");
        }
    }
    int rc = engine_setup(e, P);
    if (!rc) {
        const int t = tune ? tune->gemm_tile : 0;
        if (t < 0 || t > GT_W) rc = set_err(DQMC_ERR_ARG, "bad gemm_tile %d", t);
        else if (t > 0) e->gemm_tile = t;
        else rc = e->cplx ? autotune_gemm<zd>(e) : autotune_gemm<double>(e);
    }
    if (rc) {
        dqmc_engine_destroy(e);
        return rc;
    }
    *out = e;
    return 0;
}

extern "C" void dqmc_engine_destroy(dqmc_engine* e)
{
    if (!e) return;
    (void)hipSetDevice(e->device);
    if (e->st) (void)hipStreamSynchronize(e->st);
    for (auto& ev : e->gemm_events) {
        (void)hipEventDestroy(ev.first);
        (void)hipEventDestroy(ev.second);
    }
    for (int k = 0; k < 4; k++) {
        if (e->gexec[k]) (void)hipGraphExecDestroy(e->gexec[k]);
        if (e->graph[k]) (void)hipGraphDestroy(e->graph[k]);
    }
    for (auto& ev : e->done)
        if (ev) (void)hipEventDestroy(ev);
    prof_collect(e);
    for (auto ev : e->ev_pool) (void)hipEventDestroy(ev);
    for (void* p : e->allocs) (void)hipFree(p);
    if (e->st) (void)hipStreamDestroy(e->st);
    delete e;
}

extern "C" int dqmc_engine_set_state(dqmc_engine* e, int c, const uint64_t* rng, int sweep,
                                     const int32_t* hs)
{
    if (!e || c < 0 || c >= e->nc) return set_err(DQMC_ERR_ARG, "bad chain");
    HIPCHK(hipSetDevice(e->device));
    const long LN = (long)e->L * e->N;
    HIPCHK(hipMemcpy(e->rng + c * 17, rng, 17 * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(e->hs + c * LN, hs, LN * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(e->d_sweep + c, &sweep, 4, hipMemcpyHostToDevice));
    e->sweep[c] = sweep;
    e->stale = true;
    return 0;
}

extern "C" int dqmc_engine_get_state(dqmc_engine* e, int c, uint64_t* rng, int* sweep, int32_t* hs)
{
    if (!e || c < 0 || c >= e->nc) return set_err(DQMC_ERR_ARG, "bad chain");
    HIPCHK(hipSetDevice(e->device));
    HIPCHK(hipStreamSynchronize(e->st));
    const long LN = (long)e->L * e->N;
    if (rng) HIPCHK(hipMemcpy(rng, e->rng + c * 17, 17 * 8, hipMemcpyDeviceToHost));
    if (hs) HIPCHK(hipMemcpy(hs, e->hs + c * LN, LN * 4, hipMemcpyDeviceToHost));
    if (sweep) *sweep = e->sweep[c];
    return 0;
}

static int meas_copy(dqmc_engine* e, int c, dqmc_meas* m, int ue, int to_dev)
{
    const long* off = ue ? e->ue_off : e->eq_off;
    const long* len = ue ? e->ue_len : e->eq_len;
    const long total = ue ? e->ue_total : e->eq_total;
    char* acc = (char*)(ue ? e->acc_ue : e->acc_eq) + (size_t)c * total * e->esz;
    const int nk = ue ? DQ_UE_COUNT : DQ_EQ_COUNT;
    for (int k = 0; k < nk; k++) {
        if (len[k] <= 0 || !m->acc[k]) continue;
        if (to_dev)
            HIPCHK(hipMemcpy(acc + off[k] * e->esz, m->acc[k], len[k] * e->esz, hipMemcpyHostToDevice));
        else
            HIPCHK(hipMemcpy(m->acc[k], acc + off[k] * e->esz, len[k] * e->esz, hipMemcpyDeviceToHost));
    }
    int* ns = (ue ? e->ns_ue : e->ns_eq) + c;
    char* sg = (char*)(ue ? e->sign_ue : e->sign_eq) + c * e->esz;
    if (to_dev) {
        HIPCHK(hipMemcpy(ns, &m->n_sample, 4, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(sg, m->sign, e->esz, hipMemcpyHostToDevice));
    } else {
        HIPCHK(hipMemcpy(&m->n_sample, ns, 4, hipMemcpyDeviceToHost));
        double sgv[2] = {0.0, 0.0};
        HIPCHK(hipMemcpy(sgv, sg, e->esz, hipMemcpyDeviceToHost));
        m->sign[0] = sgv[0];
        m->sign[1] = sgv[1];
    }
    return 0;
}

extern "C" int dqmc_engine_set_meas(dqmc_engine* e, int c, const dqmc_meas* eq, const dqmc_meas* ue)
{
    if (!e || c < 0 || c >= e->nc) return set_err(DQMC_ERR_ARG, "bad chain");
    HIPCHK(hipSetDevice(e->device));
    HIPCHK(hipStreamSynchronize(e->st));
    if (eq) RC(meas_copy(e, c, (dqmc_meas*)eq, 0, 1));
    if (ue && e->period_uneqlt > 0) RC(meas_copy(e, c, (dqmc_meas*)ue, 1, 1));
    return 0;
}

extern "C" int dqmc_engine_get_meas(dqmc_engine* e, int c, dqmc_meas* eq, dqmc_meas* ue)
{
    if (!e || c < 0 || c >= e->nc) return set_err(DQMC_ERR_ARG, "bad chain");
    HIPCHK(hipSetDevice(e->device));
    HIPCHK(hipStreamSynchronize(e->st));
    if (eq) RC(meas_copy(e, c, eq, 0, 0));
    if (ue && e->period_uneqlt > 0) RC(meas_copy(e, c, ue, 1, 0));
    return 0;
}

extern "C" int dqmc_engine_run(dqmc_engine* e, int n_sweeps, const volatile int* stop_flag)
{
    if (!e || n_sweeps < 0) return set_err(DQMC_ERR_ARG, "bad arguments");
    HIPCHK(hipSetDevice(e->device));
    return e->cplx ? run_prof<zd>(e, n_sweeps, stop_flag) : run_prof<double>(e, n_sweeps, stop_flag);
}

extern "C" int dqmc_engine_green(dqmc_engine* e, int c, int spin, double* G, double* phase)
{
    if (!e || c < 0 || c >= e->nc || spin < 0 || spin > 1) return set_err(DQMC_ERR_ARG, "bad args");
    HIPCHK(hipSetDevice(e->device));
    if (e->stale) RC(e->cplx ? init_green<zd>(e) : init_green<double>(e));
    HIPCHK(hipStreamSynchronize(e->st));
    const size_t mb = (size_t)e->N * e->N * e->esz;
    if (G) HIPCHK(hipMemcpy(G, (char*)e->G + (2 * c + spin) * mb, mb, hipMemcpyDeviceToHost));
    if (phase) {
        double v[2] = {0.0, 0.0};
        HIPCHK(hipMemcpy(v, (char*)e->phase + c * e->esz, e->esz, hipMemcpyDeviceToHost));
        phase[0] = v[0];
        phase[1] = v[1];
    }
    return 0;
}

template <typename T>
static int calc_eq_g_hook(dqmc_engine* e, int c, int spin, int f, double* G, double* phase)
{
    const int N = e->N;
    const long NN = (long)N * N;
    k_scales<<<nblk((long)e->nc * e->L * N), 256, 0, e->st>>>(N, e->L, e->nc, e->hs, e->exp_lambda,
                                                               e->escale, e->iscale);
    LCHK();
    RC(build_C<T>(e, 0, e->F));
    T* out = dp<T>(e->W);  // free after build_C
    RC(calc_eq_g<T>(e, f, out, dp<T>(e->phb2)));
    HIPCHK(hipStreamSynchronize(e->st));
    const int b = 2 * c + spin;
    HIPCHK(hipMemcpy(G, out + b * NN, NN * sizeof(T), hipMemcpyDeviceToHost));
    T ph;
    HIPCHK(hipMemcpy(&ph, dp<T>(e->phb2) + b, sizeof(T), hipMemcpyDeviceToHost));
    phase[0] = re(ph);
    phase[1] = im(ph);
    e->stale = true;  // C_f was rebuilt for the current field; G must follow
    return 0;
}

extern "C" int dqmc_engine_calc_eq_g(dqmc_engine* e, int c, int spin, int f, double* G,
                                     double* phase)
{
    if (!e || c < 0 || c >= e->nc || spin < 0 || spin > 1 || f < 0 || f >= e->F)
        return set_err(DQMC_ERR_ARG, "bad args");
    HIPCHK(hipSetDevice(e->device));
    return e->cplx ? calc_eq_g_hook<zd>(e, c, spin, f, G, phase)
                   : calc_eq_g_hook<double>(e, c, spin, f, G, phase);
}

// --------------------------------------------------------------------------
// calc_eq_g without allocation: the work space of calc_eq_g lives in one
// file-scope device array per GPU.  In host code `scratch` is only a stand-in
// for the symbol; its device address comes from hipGetSymbolAddress, once per
// device, and only that address is passed to kernels.
// --------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[DQMC_NOALLOC_SCRATCH_BYTES];

#define NOALLOC_MAX_DEVICES 64
static std::mutex g_scratch_mu;
static void* g_scratch_dev[NOALLOC_MAX_DEVICES];

// device address of this GPU's `scratch` (the current device), resolved on
// the first call on that device
static int scratch_address(void** out)
{
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES)
        return set_err(DQMC_ERR_UNSUPPORTED, "device %d beyond %d", dev, NOALLOC_MAX_DEVICES);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        const hipError_t err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (err != hipSuccess || !p)
            return set_err(DQMC_ERR_HIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(err));
        g_scratch_dev[dev] = p;
    }
    *out = g_scratch_dev[dev];
    return 0;
}

// Carves calc_eq_g's work space for `batch` matrices of order N out of the
// region at `base` (nullptr: only count); returns the bytes used.
static size_t eqg_carve(dqmc_engine* e, char* base, int N, size_t es, int batch)
{
    const long NN = (long)N * N, B = batch;
    size_t off = 0;
    auto take = [&](size_t bytes) -> void* {
        void* p = base ? (void*)(base + off) : nullptr;
        off += (bytes + 255) & ~(size_t)255;
        return p;
    };
    e->sVb = NN;
    e->sTb = 2L * N * 64 + 64 * 64;
    e->sWq = 64L * 2 * N;
    void** mats[] = {&e->X, &e->Y, &e->Y2, &e->Z, &e->Q, &e->Tm, &e->Rs, &e->Tp, &e->Mm};
    for (void** m : mats) *m = take(B * NN * es);
    e->tau = take(B * 2 * N * es);
    e->Vb = take(B * e->sVb * es);
    e->Tb = take(B * e->sTb * es);
    e->Wq1 = take(B * e->sWq * es);
    e->Wq2 = take(B * e->sWq * es);
    e->detq = take(B * es);
    e->detm = take(B * es);
    e->jpvt = (int*)take(B * N * 4);
    e->perm = (int*)take(B * N * 4);
    e->dvec = (double*)take(B * N * 8);
    e->vvec = (double*)take(B * N * 8);
    return off;
}

__global__ __launch_bounds__(256) void k_zero_u64(unsigned long long* p, long n)
{
    for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256) p[i] = 0ull;
}

template <typename T>
static int calc_eq_g_noalloc_t(int N, int F, int n_chains, int f, const double* C, double* G,
                               double* phase, hipStream_t st)
{
    dqmc_engine tmp;  // no allocation: every buffer points into the caller's or `scratch`
    dqmc_engine* e = &tmp;
    e->nc = n_chains; e->N = N; e->L = F; e->F = F; e->n_matmul = 1;
    e->cplx = is_cplx<T>::value; e->esz = sizeof(T);
    e->st = st;
    e->gemm_tile = GT_64;
    const size_t need = eqg_carve(e, nullptr, N, sizeof(T), 2 * n_chains);
    if (need > (size_t)DQMC_NOALLOC_SCRATCH_BYTES)
        return set_err(DQMC_ERR_ARG,
                       "calc_eq_g of %d matrices of order %d needs %zu bytes of work space; "
                       "the static scratch holds %zu", 2 * n_chains, N, need,
                       (size_t)DQMC_NOALLOC_SCRATCH_BYTES);
    void* base = nullptr;
    RC(scratch_address(&base));
    eqg_carve(e, (char*)base, N, sizeof(T), 2 * n_chains);
    // every call starts from a cleared work space, as a freshly created engine
    // does: the previous call may have left other contents at other strides
    const long n64 = (long)(need / 8);
    k_zero_u64<<<nblk(n64, 256), 256, 0, st>>>((unsigned long long*)base, n64);
    LCHK();
    e->C = (void*)C;
    return calc_eq_g<T>(e, f, (T*)G, (T*)phase);
}

extern "C" int dqmc_calc_eq_g_noalloc(int is_complex, int N, int F, int n_chains, int f,
                                      const double* C, double* G, double* phase, void* stream)
{
    if (N <= 0 || N > QR_MAXM || F <= 0 || n_chains <= 0 || f < 0 || f >= F || !C || !G || !phase)
        return set_err(DQMC_ERR_ARG, "bad arguments (N=%d F=%d n_chains=%d f=%d)", N, F, n_chains, f);
    hipStream_t st = (hipStream_t)stream;
    return is_complex ? calc_eq_g_noalloc_t<zd>(N, F, n_chains, f, C, G, phase, st)
                      : calc_eq_g_noalloc_t<double>(N, F, n_chains, f, C, G, phase, st);
}

extern "C" int dqmc_engine_calc_ue_g(dqmc_engine* e, int c, int spin, double* G0t, double* Gtt,
                                     double* Gt0)
{
    if (!e || c < 0 || c >= e->nc || spin < 0 || spin > 1) return set_err(DQMC_ERR_ARG, "bad args");
    if (e->period_uneqlt <= 0)
        return set_err(DQMC_ERR_ARG, "engine created with period_uneqlt == 0");
    if (e->ragged) return set_err(DQMC_ERR_UNSUPPORTED, "padded chains: measured blocks only");
    HIPCHK(hipSetDevice(e->device));
    return e->cplx ? ue_hook<zd>(e, c, spin, G0t, Gtt, Gt0)
                   : ue_hook<double>(e, c, spin, G0t, Gtt, Gt0);
}

extern "C" int64_t dqmc_engine_accepts(dqmc_engine* e)
{
    if (!e) return -1;
    (void)hipSetDevice(e->device);
    (void)hipStreamSynchronize(e->st);
    std::vector<unsigned long long> v(e->nc);
    if (hipMemcpy(v.data(), e->nacc, e->nc * 8, hipMemcpyDeviceToHost) != hipSuccess) return -1;
    int64_t s = 0;
    for (auto x : v) s += (int64_t)x;
    return s;
}

extern "C" int dqmc_engine_slice_accepts(dqmc_engine* e, int c, int64_t* counts)
{
    if (!e || c < 0 || c >= e->nc || !counts) return set_err(DQMC_ERR_ARG, "bad arguments");
    HIPCHK(hipSetDevice(e->device));
    HIPCHK(hipStreamSynchronize(e->st));
    std::vector<unsigned long long> v(e->L);
    HIPCHK(hipMemcpy(v.data(), e->nacc_sl + (long)c * e->L, e->L * 8, hipMemcpyDeviceToHost));
    for (int l = 0; l < e->L; l++) counts[l] = (int64_t)v[l];
    return 0;
}

extern "C" int dqmc_engine_drift(dqmc_engine* e, double* max_abs)
{
    if (!e || !max_abs) return set_err(DQMC_ERR_ARG, "bad arguments");
    if (!e->check_drift) return set_err(DQMC_ERR_ARG, "engine created without the drift check");
    HIPCHK(hipSetDevice(e->device));
    HIPCHK(hipStreamSynchronize(e->st));
    std::vector<unsigned long long> v(e->nc, 0ull);
    HIPCHK(hipMemcpy(v.data(), e->d_drift, e->nc * 8, hipMemcpyDeviceToHost));
    for (int c = 0; c < e->nc; c++) memcpy(&max_abs[c], &v[c], 8);
    std::fill(v.begin(), v.end(), 0ull);
    HIPCHK(hipMemcpy(e->d_drift, v.data(), e->nc * 8, hipMemcpyHostToDevice));
    return 0;
}

extern "C" int dqmc_engine_profile(dqmc_engine* e, const char** names, double* ms,
                                   int64_t* calls, int max)
{
    if (!e) return set_err(DQMC_ERR_ARG, "null engine");
    if (!e->prof) return set_err(DQMC_ERR_ARG, "engine created without profiling (flags bit 3)");
    prof_collect(e);
    for (int k = 0; k < PS_COUNT && k < max; k++) {
        if (names) names[k] = g_prof_names[k];
        if (ms) ms[k] = e->prof_ms[k];
        if (calls) calls[k] = e->prof_calls[k];
    }
    return PS_COUNT;
}

extern "C" int dqmc_engine_gemm_stats(dqmc_engine* e, double* avg_ms, int64_t* launches,
                                      double* flops_per_launch)
{
    if (!e) return set_err(DQMC_ERR_ARG, "null engine");
    HIPCHK(hipStreamSynchronize(e->st));
    double tot = 0.0;
    for (auto& ev : e->gemm_events) {
        float t = 0.f;
        (void)hipEventElapsedTime(&t, ev.first, ev.second);
        tot += t;
        (void)hipEventDestroy(ev.first);
        (void)hipEventDestroy(ev.second);
    }
    const long long n = (long long)e->gemm_events.size();
    e->gemm_events.clear();
    if (avg_ms) *avg_ms = n ? tot / n : 0.0;
    if (launches) *launches = n;
    if (flops_per_launch) *flops_per_launch = e->gemm_flops;
    return 0;
}

// Kernel-level test hook: one GEMM through the engine's MFMA kernel on host
// buffers, C = alpha diag(r) op(A) diag(u) op(B) diag(w) + beta C.
template <typename T>
static int test_gemm_t(int opa, int opb, int M, int N, int K, const double* alpha,
                       const double* A, int lda, const double* B, int ldb, const double* beta,
                       double* Cm, int ldc, const double* r, const double* u, const double* w,
                       int tile)
{
    dqmc_engine tmp;
    dqmc_engine* e = &tmp;
    e->gemm_tile = tile;
    HIPCHK(hipStreamCreate(&e->st));
    const int ka = (opa == OP_N) ? K : M, kb = (opb == OP_N) ? N : K;
    void *dA, *dB, *dC, *dr = 0, *du = 0, *dw = 0;
    const size_t es = sizeof(T);
    DALLOC(dA, (size_t)lda * ka * es);
    DALLOC(dB, (size_t)ldb * kb * es);
    DALLOC(dC, (size_t)ldc * N * es);
    HIPCHK(hipMemcpy(dA, A, (size_t)lda * ka * es, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dB, B, (size_t)ldb * kb * es, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dC, Cm, (size_t)ldc * N * es, hipMemcpyHostToDevice));
    if (r) { DALLOC(dr, M * 8); HIPCHK(hipMemcpy(dr, r, M * 8, hipMemcpyHostToDevice)); }
    if (u) { DALLOC(du, K * 8); HIPCHK(hipMemcpy(du, u, K * 8, hipMemcpyHostToDevice)); }
    if (w) { DALLOC(dw, N * 8); HIPCHK(hipMemcpy(dw, w, N * 8, hipMemcpyHostToDevice)); }
    T al, be;
    memcpy(&al, alpha, es);
    memcpy(&be, beta, es);
    RC(gemm<T>(e, opa, opb, M, N, K, al, (T*)dA, lda, 0, (T*)dB, ldb, 0, be, (T*)dC, ldc, 0, 1,
               (double*)dr, 0, (double*)du, 0, (double*)dw, 0));
    HIPCHK(hipStreamSynchronize(e->st));
    HIPCHK(hipMemcpy(Cm, dC, (size_t)ldc * N * es, hipMemcpyDeviceToHost));
    for (void* p : e->allocs) (void)hipFree(p);
    e->allocs.clear();
    (void)hipStreamDestroy(e->st);
    return 0;
}

extern "C" int dqmc_test_gemm(int is_complex, int opa, int opb, int M, int N, int K,
                              const double* alpha, const double* A, int lda, const double* B,
                              int ldb, const double* beta, double* Cm, int ldc, const double* r,
                              const double* u, const double* w, int tile)
{
    if (M <= 0 || N <= 0 || K <= 0) return set_err(DQMC_ERR_ARG, "bad GEMM shape");
    if (tile < 0 || tile > GT_W) return set_err(DQMC_ERR_ARG, "bad tile %d", tile);
    return is_complex
               ? test_gemm_t<zd>(opa, opb, M, N, K, alpha, A, lda, B, ldb, beta, Cm, ldc, r, u, w, tile)
               : test_gemm_t<double>(opa, opb, M, N, K, alpha, A, lda, B, ldb, beta, Cm, ldc, r, u, w,
                                     tile);
}

// Sustained FP64 matrix-core rate (no counterpart in the reference): every
// wave runs `iters` rounds of 8 independent MFMA accumulation chains out of
// registers (no memory traffic), 4 workgroups of 4 waves per CU.  form 0:
// v_mfma_f64_16x16x4f64, form 1: v_mfma_f64_4x4x4f64.  The rate is the
// ceiling a compute-bound FP64 kernel of that form can reach on this device
// and is read beside the published peak.
template <int FORM>
__global__ __launch_bounds__(256) void k_mfma_loop(double* out, int iters, double a0)
{
    const double a = a0 + threadIdx.x * 1e-9, b = a0 - threadIdx.x * 1e-9;
    double s = 0.0;
    if constexpr (FORM == 0) {
        dbl4 acc[8];
#pragma unroll
        for (int c = 0; c < 8; c++) acc[c] = (dbl4){0.0, 0.0, 0.0, 0.0};
        for (int it = 0; it < iters; it++) {
#pragma unroll
            for (int c = 0; c < 8; c++) acc[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[c], 0, 0, 0);
        }
#pragma unroll
        for (int c = 0; c < 8; c++) s += acc[c][0] + acc[c][1] + acc[c][2] + acc[c][3];
    } else {
        double acc[8];
#pragma unroll
        for (int c = 0; c < 8; c++) acc[c] = 0.0;
        for (int it = 0; it < iters; it++) {
#pragma unroll
            for (int c = 0; c < 8; c++) acc[c] = __builtin_amdgcn_mfma_f64_4x4x4f64(a, b, acc[c], 0, 0, 0);
        }
#pragma unroll
        for (int c = 0; c < 8; c++) s += acc[c];
    }
    out[blockIdx.x * 256 + threadIdx.x] = s;
}

extern "C" int dqmc_probe_mfma(int form, double* tflops)
{
    if ((form != 0 && form != 1) || !tflops) return set_err(DQMC_ERR_ARG, "bad probe form %d", form);
    int dev = 0, ncu = 0;
    HIPCHK(hipGetDevice(&dev));
    HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    const int blocks = 4 * ncu, iters = form == 0 ? 4000 : 16000;
    double* out = nullptr;
    HIPCHK(hipMalloc(&out, sizeof(double) * 256 * blocks));
    hipStream_t st;
    hipEvent_t e0, e1;
    int rc = 0;
    if (hipStreamCreate(&st) != hipSuccess) {
        (void)hipFree(out);
        return set_err(DQMC_ERR_HIP, "probe stream");
    }
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    float ms = 0.f;
    for (int r = 0; r < 3 && !rc; r++) {  // the last of three launches is timed
        (void)hipEventRecord(e0, st);
        if (form == 0) k_mfma_loop<0><<<blocks, 256, 0, st>>>(out, iters, 1e-3);
        else k_mfma_loop<1><<<blocks, 256, 0, st>>>(out, iters, 1e-3);
        (void)hipEventRecord(e1, st);
        if (hipEventSynchronize(e1) != hipSuccess || hipGetLastError() != hipSuccess)
            rc = set_err(DQMC_ERR_HIP, "probe launch");
        (void)hipEventElapsedTime(&ms, e0, e1);
    }
    const double flop_per_inst = form == 0 ? 2.0 * 16 * 16 * 4 : 4 * 2.0 * 4 * 4 * 4;
    if (!rc) *tflops = (double)blocks * 4 * iters * 8 * flop_per_inst / (ms * 1e-3) * 1e-12;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipStreamDestroy(st);
    (void)hipFree(out);
    return rc;
}
