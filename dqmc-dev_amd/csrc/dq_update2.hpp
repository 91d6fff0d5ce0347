// dq_update2.hpp -- Metropolis sweep over one time slice with delayed rank-KD
// updates (src/updates.c:7-84), register-resident delay buffers.
//
// One 256-thread workgroup per chain, both spins.  Thread t owns the rows
// j = t + 256 r (r < R) of the delay buffers A_up, B_up, A_dn, B_dn and keeps
// them in registers; the delay slots are unrolled, so every register index is
// a compile-time constant.  The site order and uniforms come precomputed
// (dq_rng.hpp); the HS row, Delta, and the running G diagonal sit in LDS, so
// lane 0's scan for the next acceptance touches no global memory.  On an
// acceptance at site i: all threads fetch G[:, i] and G[i, :] (both spins),
// the owner of row i broadcasts its A/B rows through LDS, and every thread
// forms its entries of the new columns a_k, b_k (the gemv of updates.c:33-36)
// and updates the diagonal.  Every KD acceptances and at the end of the slice
// G += A B^T is applied (B staged through LDS in chunks of UPD2_CH rows).
#pragma once
#include "dq_common.hpp"

#define UPD2_THREADS 256
#define UPD2_MAXN 512
#define UPD2_CH 64

template <typename T>
struct Upd2Args {
    int N, L, l;
    T* G; long sG;                  // G[b], b = 2c + s
    const double* del;              // [c][2N]
    const double* exp_lambda;       // [c][2N]
    int* hs;                        // [c][L*N]
    const int* order;               // [c][L][N]
    const double* unif;             // [c][L][N]
    T* phase;                       // [c]
    double* escale; double* iscale; // [b][L*N]
    unsigned long long* n_acc;      // [c]
    unsigned long long* n_acc_sl;   // [c][L] accepted flips per slice (or null)
    const int* active;              // [c]
    const int* Lc;                  // [c] own slice counts of padded chains (or null)
};

// G += A(:, 0:KD) B(:, 0:KD)^T for the rows this thread owns, d = diag(G).
// B is staged through LDS in chunks of UPD2_CH rows; within a chunk each
// thread keeps FB columns of its G row in flight at once so the global-memory
// latency overlaps instead of being paid once per column.
#define UPD2_FB 8
template <typename T, int R, int KD>
__device__ inline void upd2_flush(int N, int tid, T* G, const T (&A)[R][KD], const T (&Bm)[R][KD],
                                  T* d, T* Bst)
{
    for (int c0 = 0; c0 < N; c0 += UPD2_CH) {
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int j = tid + UPD2_THREADS * r;
            if (j >= c0 && j < c0 + UPD2_CH && j < N) {
#pragma unroll
                for (int m = 0; m < KD; m++) Bst[(j - c0) * KD + m] = Bm[r][m];
            }
        }
        __syncthreads();
        const int cend = (N - c0 < UPD2_CH) ? N - c0 : UPD2_CH;
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int j = tid + UPD2_THREADS * r;
            if (j < N) {
                T* grow = G + j + (long)N * c0;
                for (int cb = 0; cb < cend; cb += UPD2_FB) {
                    T g[UPD2_FB];
#pragma unroll
                    for (int t = 0; t < UPD2_FB; t++)
                        if (cb + t < cend) g[t] = grow[(long)N * (cb + t)];
#pragma unroll
                    for (int t = 0; t < UPD2_FB; t++) {
                        T s = g[t];
#pragma unroll
                        for (int m = 0; m < KD; m++) s += A[r][m] * Bst[(cb + t) * KD + m];
                        g[t] = s;
                    }
#pragma unroll
                    for (int t = 0; t < UPD2_FB; t++)
                        if (cb + t < cend) grow[(long)N * (cb + t)] = g[t];
                }
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int j = tid + UPD2_THREADS * r;
        if (j < N) d[j] = G[j + (long)N * j];
    }
}

// state of one workgroup while it fills the delay slots of a round
template <typename T, int R, int KD>
struct Upd2State {
    T Au[R][KD], Bu[R][KD], Ad[R][KD], Bd[R][KD];
};

struct Upd2Shared {
    int* hsl; const int* ordl; const double* ul; const double* dl0; const double* dl1;
    int* s_site;
};

// Slot K of a round: lane 0 scans for the next acceptance, then every thread
// forms its entries of column K of A and B for both spins.  Returns false when
// the slice has no further acceptance (slots K.. stay zero).  K is a
// compile-time constant, so all register indices below are static.
template <typename T, int R, int KD, int K>
__device__ inline bool upd2_slot(Upd2State<T, R, KD>& S, int N, int tid, T* Gu, T* Gd, T* du,
                                 T* dd, T* rAu, T* rBu, T* rAd, T* rBd, T* s_fu, T* s_fd,
                                 const Upd2Shared& sh, int& ii, T& phase,
                                 unsigned long long& nacc)
{
    if (tid == 0) {
        int site = -1;
        for (; ii < N; ii++) {
            const int i = sh.ordl[ii];
            const int sv = sh.hsl[i];
            const double delu = sv ? sh.dl1[i] : sh.dl0[i];
            const double deld = sv ? sh.dl0[i] : sh.dl1[i];
            if (delu == 0.0 && deld == 0.0) continue;
            const T ru = from_re<T>(1.0) + (from_re<T>(1.0) - du[i]) * delu;
            const T rd = from_re<T>(1.0) + (from_re<T>(1.0) - dd[i]) * deld;
            const T prob = ru * rd;
            const double absprob = absv(prob);
            if (sh.ul[ii] < absprob) {
                site = i;
                *s_fu = rdiv(from_re<T>(delu), ru);
                *s_fd = rdiv(from_re<T>(deld), rd);
                sh.hsl[i] = 1 - sv;
                phase = phase * (prob * (1.0 / absprob));
                nacc++;
                ii++;
                break;
            }
        }
        *sh.s_site = site;
    }
    __syncthreads();
    const int i = *sh.s_site;
    if (i < 0) return false;
    T gcu[R], gru[R], gcd[R], grd[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int j = tid + UPD2_THREADS * r;
        const int jj = j < N ? j : 0;
        gcu[r] = Gu[jj + (long)N * i];
        gru[r] = Gu[i + (long)N * jj];
        gcd[r] = Gd[jj + (long)N * i];
        grd[r] = Gd[i + (long)N * jj];
    }
    if (K > 0) {
        if (tid == (i & (UPD2_THREADS - 1))) {
            const int oq = i / UPD2_THREADS;
#pragma unroll
            for (int r = 0; r < R; r++)
                if (r == oq) {
#pragma unroll
                    for (int m = 0; m < K; m++) {
                        rAu[m] = S.Au[r][m]; rBu[m] = S.Bu[r][m];
                        rAd[m] = S.Ad[r][m]; rBd[m] = S.Bd[r][m];
                    }
                }
        }
        __syncthreads();
    }
    const T fu = *s_fu, fd = *s_fd;
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int j = tid + UPD2_THREADS * r;
        T au = gcu[r], bu = gru[r], ad = gcd[r], bd = grd[r];
#pragma unroll
        for (int m = 0; m < K; m++) {
            au += S.Au[r][m] * rBu[m];
            bu += S.Bu[r][m] * rAu[m];
            ad += S.Ad[r][m] * rBd[m];
            bd += S.Bd[r][m] * rAd[m];
        }
        if (j == i) { au -= from_re<T>(1.0); ad -= from_re<T>(1.0); }
        au = au * fu;
        ad = ad * fd;
        if (j < N) {
            du[j] += au * bu;
            dd[j] += ad * bd;
            S.Au[r][K] = au; S.Bu[r][K] = bu; S.Ad[r][K] = ad; S.Bd[r][K] = bd;
        }
    }
    __syncthreads();
    return true;
}

template <typename T, int R, int KD, int K>
struct Upd2Round {
    // returns the number of slots filled (K.. KD)
    __device__ static inline int run(Upd2State<T, R, KD>& S, int N, int tid, T* Gu, T* Gd, T* du,
                                     T* dd, T* rAu, T* rBu, T* rAd, T* rBd, T* s_fu, T* s_fd,
                                     const Upd2Shared& sh, int& ii, T& phase,
                                     unsigned long long& nacc)
    {
        if (!upd2_slot<T, R, KD, K>(S, N, tid, Gu, Gd, du, dd, rAu, rBu, rAd, rBd, s_fu, s_fd, sh,
                                    ii, phase, nacc))
            return K;
        return Upd2Round<T, R, KD, K + 1>::run(S, N, tid, Gu, Gd, du, dd, rAu, rBu, rAd, rBd,
                                                s_fu, s_fd, sh, ii, phase, nacc);
    }
};
template <typename T, int R, int KD>
struct Upd2Round<T, R, KD, KD> {
    __device__ static inline int run(Upd2State<T, R, KD>&, int, int, T*, T*, T*, T*, T*, T*, T*,
                                     T*, T*, T*, const Upd2Shared&, int&, T&, unsigned long long&)
    {
        return KD;
    }
};

template <typename T, int R, int KD>
__global__ __launch_bounds__(UPD2_THREADS) void k_update_reg(Upd2Args<T> a)
{
    __shared__ T du[UPD2_MAXN], dd[UPD2_MAXN];
    __shared__ int hsl[UPD2_MAXN], ordl[UPD2_MAXN];
    __shared__ double ul[UPD2_MAXN], dl0[UPD2_MAXN], dl1[UPD2_MAXN];
    __shared__ T rAu[KD], rBu[KD], rAd[KD], rBd[KD];
    __shared__ T Bst[UPD2_CH * KD];
    __shared__ int s_site;
    __shared__ T s_fu, s_fd;
    const int c = blockIdx.x;
    if (!a.active[c]) return;
    if (a.Lc && a.l >= a.Lc[c]) return;  // identity slice of a padded chain
    const int N = a.N, tid = threadIdx.x;
    T* Gu = a.G + (2 * c) * a.sG;
    T* Gd = a.G + (2 * c + 1) * a.sG;
    int* hs = a.hs + (long)c * a.L * N + (long)N * a.l;
    const long ol = ((long)c * a.L + a.l) * N;
    for (int j = tid; j < N; j += UPD2_THREADS) {
        du[j] = Gu[j + (long)N * j];
        dd[j] = Gd[j + (long)N * j];
        hsl[j] = hs[j];
        ordl[j] = a.order[ol + j];
        ul[j] = a.unif[ol + j];
        dl0[j] = a.del[2L * N * c + j];
        dl1[j] = a.del[2L * N * c + N + j];
    }
    T phase = from_re<T>(0.0);
    unsigned long long nacc = 0;
    if (tid == 0) phase = a.phase[c];
    int ii = 0;
    Upd2Shared sh;
    sh.hsl = hsl; sh.ordl = ordl; sh.ul = ul; sh.dl0 = dl0; sh.dl1 = dl1; sh.s_site = &s_site;
    __syncthreads();

    Upd2State<T, R, KD> S;
    for (;;) {
#pragma unroll
        for (int r = 0; r < R; r++)
#pragma unroll
            for (int m = 0; m < KD; m++) {
                S.Au[r][m] = from_re<T>(0.0); S.Bu[r][m] = from_re<T>(0.0);
                S.Ad[r][m] = from_re<T>(0.0); S.Bd[r][m] = from_re<T>(0.0);
            }
        const int kused = Upd2Round<T, R, KD, 0>::run(S, N, tid, Gu, Gd, du, dd, rAu, rBu, rAd,
                                                       rBd, &s_fu, &s_fd, sh, ii, phase, nacc);
        if (kused > 0) {
            upd2_flush<T, R, KD>(N, tid, Gu, S.Au, S.Bu, du, Bst);
            upd2_flush<T, R, KD>(N, tid, Gd, S.Ad, S.Bd, dd, Bst);
            __syncthreads();
        }
        if (kused < KD) break;
    }
    // final field of this slice: HS row and the scalings of B_l / B_l^-1
    const double* el = a.exp_lambda + (long)c * 2 * N;
    double* esu = a.escale + (2 * c) * (long)a.L * N + (long)N * a.l;
    double* esd = a.escale + (2 * c + 1) * (long)a.L * N + (long)N * a.l;
    double* isu = a.iscale + (2 * c) * (long)a.L * N + (long)N * a.l;
    double* isd = a.iscale + (2 * c + 1) * (long)a.L * N + (long)N * a.l;
    for (int j = tid; j < N; j += UPD2_THREADS) {
        const int sv = hsl[j];
        hs[j] = sv;
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
