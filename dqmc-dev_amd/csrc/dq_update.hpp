// dq_update.hpp -- the Metropolis sweep over one time slice with delayed
// rank-k Green's-function updates (src/updates.c:7-84), the RNG
// (src/rand.h:6-32) and the HS-field scalings of B_l (src/dqmc.c:29-59).
//
// General-N fallback (N > 512; dq_update2.hpp covers N <= 512).  One workgroup
// per chain handles both spins, because the acceptance test couples them
// (u < |r_up * r_dn|).  Lane 0 scans the precomputed site order until the next
// acceptance without any workgroup barrier; the whole workgroup then builds
// the new delay columns a_k, b_k for both spins (thread j owns row j), and
// flushes G += A B^T every `kd` acceptances and at the end of the slice.
// The delay buffers and G stay in HBM/L2 (coherent inside the workgroup).
#pragma once
#include "dq_common.hpp"

#define UPD_THREADS 256
#define UPD_MAXN 1024

__device__ inline uint64_t xs_next(uint64_t* s)
{
    const uint64_t s0 = s[s[16]];
    const int p = (int)((s[16] + 1) & 15);
    s[16] = (uint64_t)p;
    uint64_t s1 = s[p];
    s1 ^= s1 << 31;
    s[p] = s1 ^ s0 ^ (s1 >> 11) ^ (s0 >> 30);
    return s[p] * 1181783497276652981ULL;
}

__device__ inline double xs_doub(uint64_t* s)
{
    const uint64_t bits = (0x3FFULL << 52) | (xs_next(s) >> 12);
    return __longlong_as_double((long long)bits) - 1.0;
}

template <typename T>
struct UpdArgs {
    int N, L, l, kd;
    T* G; long sG;                  // G[b], b = 2c + s
    T* Ab; T* Bb; long sAB;         // delay buffers [b][N*kd]
    const double* del;              // [c][2N]
    const double* exp_lambda;       // [c][2N]
    int* hs;                        // [c][L*N]
    const int* order;               // [c][L][N] site order (dq_rng.hpp)
    const double* unif;             // [c][L][N] uniform per visited site
    T* phase;                       // [c]
    double* escale; double* iscale; // [b][L*N]
    unsigned long long* n_acc;      // [c]
    unsigned long long* n_acc_sl;   // [c][L] accepted flips per slice (or null)
    const int* active;              // [c]
    const int* Lc;                  // [c] own slice counts of padded chains (or null)
};

template <typename T>
__device__ inline void upd_flush(int N, int k, T* G, const T* A, const T* B, T* d, int tid)
{
    // G += A(:, 0:k) B(:, 0:k)^T ;  d = diag(G)
    for (int j = tid; j < N; j += blockDim.x) {
        for (int c = 0; c < N; c++) {
            T s = G[j + (long)N * c];
            for (int m = 0; m < k; m++) s += A[j + (long)N * m] * B[c + (long)N * m];
            G[j + (long)N * c] = s;
        }
        d[j] = G[j + (long)N * j];
    }
}

template <typename T>
__global__ __launch_bounds__(UPD_THREADS) void k_update(UpdArgs<T> a)
{
    __shared__ T du[UPD_MAXN], dd[UPD_MAXN];
    __shared__ int order[UPD_MAXN];
    __shared__ int s_site;
    __shared__ T s_fu, s_fd;
    const int c = blockIdx.x;
    if (!a.active[c]) return;
    if (a.Lc && a.l >= a.Lc[c]) return;  // identity slice of a padded chain
    const int N = a.N, tid = threadIdx.x;
    T* Gu = a.G + (2 * c) * a.sG;
    T* Gd = a.G + (2 * c + 1) * a.sG;
    T* Au = a.Ab + (2 * c) * a.sAB;
    T* Bu = a.Bb + (2 * c) * a.sAB;
    T* Ad = a.Ab + (2 * c + 1) * a.sAB;
    T* Bd = a.Bb + (2 * c + 1) * a.sAB;
    const double* del = a.del + (long)c * 2 * N;
    int* hs = a.hs + (long)c * a.L * N + (long)N * a.l;
    const long ol = ((long)c * a.L + a.l) * N;
    const double* unif = a.unif + ol;
    for (int j = tid; j < N; j += blockDim.x) {
        du[j] = Gu[j + (long)N * j];
        dd[j] = Gd[j + (long)N * j];
        order[j] = a.order[ol + j];
    }
    __syncthreads();
    T phase = from_re<T>(0.0);
    unsigned long long nacc = 0;
    if (tid == 0) phase = a.phase[c];
    int ii = 0;  // scan position (meaningful in thread 0)
    int k = 0;
    for (;;) {
        if (tid == 0) {
            int site = -1;
            for (; ii < N; ii++) {
                const int i = order[ii];
                const int sv = hs[i];
                const double delu = del[i + N * sv];
                const double deld = del[i + N * (1 - sv)];
                if (delu == 0.0 && deld == 0.0) continue;
                const T ru = from_re<T>(1.0) + (from_re<T>(1.0) - du[i]) * delu;
                const T rd = from_re<T>(1.0) + (from_re<T>(1.0) - dd[i]) * deld;
                const T prob = ru * rd;
                const double absprob = absv(prob);
                if (unif[ii] < absprob) {
                    site = i;
                    s_fu = rdiv(from_re<T>(delu), ru);
                    s_fd = rdiv(from_re<T>(deld), rd);
                    hs[i] = 1 - sv;
                    phase = phase * (prob * (1.0 / absprob));
                    nacc++;
                    ii++;
                    break;
                }
            }
            s_site = site;
        }
        __syncthreads();
        const int i = s_site;
        if (i < 0) break;
        const T fu = s_fu, fd = s_fd;
        for (int j = tid; j < N; j += blockDim.x) {
            T au = Gu[j + (long)N * i], bu = Gu[i + (long)N * j];
            T ad = Gd[j + (long)N * i], bd = Gd[i + (long)N * j];
            for (int m = 0; m < k; m++) {
                au += Au[j + (long)N * m] * Bu[i + (long)N * m];
                bu += Bu[j + (long)N * m] * Au[i + (long)N * m];
                ad += Ad[j + (long)N * m] * Bd[i + (long)N * m];
                bd += Bd[j + (long)N * m] * Ad[i + (long)N * m];
            }
            if (j == i) { au -= from_re<T>(1.0); ad -= from_re<T>(1.0); }
            au = au * fu;
            ad = ad * fd;
            du[j] += au * bu;
            dd[j] += ad * bd;
            Au[j + (long)N * k] = au;
            Bu[j + (long)N * k] = bu;
            Ad[j + (long)N * k] = ad;
            Bd[j + (long)N * k] = bd;
        }
        k++;
        __syncthreads();
        if (k == a.kd) {
            upd_flush(N, k, Gu, Au, Bu, du, tid);
            upd_flush(N, k, Gd, Ad, Bd, dd, tid);
            k = 0;
            __syncthreads();
        }
    }
    if (k > 0) {
        upd_flush(N, k, Gu, Au, Bu, du, tid);
        upd_flush(N, k, Gd, Ad, Bd, dd, tid);
    }
    // scalings of B_l and B_l^-1 for the final field of this slice
    const double* el = a.exp_lambda + (long)c * 2 * N;
    double* esu = a.escale + (2 * c) * (long)a.L * N + (long)N * a.l;
    double* esd = a.escale + (2 * c + 1) * (long)a.L * N + (long)N * a.l;
    double* isu = a.iscale + (2 * c) * (long)a.L * N + (long)N * a.l;
    double* isd = a.iscale + (2 * c + 1) * (long)a.L * N + (long)N * a.l;
    for (int j = tid; j < N; j += blockDim.x) {
        const int sv = hs[j];
        esu[j] = el[j + N * sv];
        esd[j] = el[j + N * (1 - sv)];
        isu[j] = el[j + N * (1 - sv)];
        isd[j] = el[j + N * sv];
    }
    if (tid == 0) {
        a.phase[c] = phase;
        a.n_acc[c] += nacc;
        if (a.n_acc_sl) a.n_acc_sl[(long)c * a.L + a.l] += nacc;
    }
}

// Scalings of B_l / B_l^-1 for every slice of every chain from hs
// (src/dqmc.c:29-59): escale[b][l][j] multiplies column j of expK,
// iscale[b][l][j] multiplies row j of inv_expK.  Spin up couples to hs,
// spin down to !hs.
__global__ void k_scales(int N, int L, int nc, const int* hs, const double* exp_lambda,
                         double* escale, double* iscale)
{
    const long total = (long)nc * L * N;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total;
         e += (long)gridDim.x * blockDim.x) {
        const int c = e / ((long)L * N);
        const long r = e % ((long)L * N);
        const int j = r % N;
        const int sv = hs[e];
        const double* el = exp_lambda + (long)c * 2 * N;
        escale[(2 * c) * (long)L * N + r] = el[j + N * sv];
        escale[(2 * c + 1) * (long)L * N + r] = el[j + N * (1 - sv)];
        iscale[(2 * c) * (long)L * N + r] = el[j + N * (1 - sv)];
        iscale[(2 * c + 1) * (long)L * N + r] = el[j + N * sv];
    }
}

// X[b] = A[b / adiv] * diag(s[b]) (column scaling); used for the first B_l of
// a product chain.  The identity instead for an identity slice (l >=
// Lc[b / 2], chains padded to a common L; Lc null: none).
template <typename T>
__global__ void k_colscale(int N, const T* A, long sA, int adiv, const double* s, long ss, T* X,
                           long sX, int batch, const int* Lc = nullptr, int l = 0)
{
    EW_IB(N);
    EW_FOR_J(N) {
        const long r = i + (long)N * j;
        if (Lc && l >= Lc[b >> 1])
            X[b * sX + r] = from_re<T>(i == j ? 1.0 : 0.0);
        else
            X[b * sX + r] = A[(b / adiv) * sA + r] * s[b * ss + j];
    }
}

template <typename T>
__global__ void k_copy(long n, const T* src, long ssrc, T* dst, long sdst, int batch, T alpha)
{
    const long total = n * batch;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total;
         e += (long)gridDim.x * blockDim.x) {
        const int b = e / n;
        const long r = e % n;
        dst[b * sdst + r] = alpha * src[b * ssrc + r];
    }
}

// phase[c] = phase_b[2c] * phase_b[2c+1] for active chains (src/dqmc.c:219, 358)
template <typename T>
__global__ void k_phase_combine(int nc, const T* phb, T* phase, const int* active)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < nc && active[c]) phase[c] = phb[2 * c] * phb[2 * c + 1];
}
