// dq_qrp.hpp -- column-pivoted Householder QR (xgeqp3 semantics) with the
// whole matrix resident on one CU, for the first factorization of calc_eq_g
// (src/greens.c:91-101: dgeqp3 + dorgqr on the first product).
//
// One 512-thread workgroup (8 waves) per matrix.  Column q belongs to wave
// q % 8, slot q / 8; the first SL slots of a wave live in LDS, the rest in
// registers (lane = row, RPL rows per lane).  Columns never move (implicit
// pivoting): pivot j retires its slot, and is written out as column j.
// Step j:
//   (1) the pivot is the largest of the 8 per-wave candidates (the best
//       squared norm of each wave's live columns), first column on ties;
//   (2) every wave reads the pivot column from the candidate's LDS buffer
//       and builds the reflector itself (xlarfg), so v never crosses waves;
//       the owner writes column j out (R above/at the diagonal, v below) and
//       the explicit unit-lower V;
//   (3) every wave applies H^H to its live columns, recomputes their norms
//       exactly (rows > j), and copies its best column to its candidate
//       buffer; reductions are batched over columns and stay in registers.
// One workgroup barrier per step (candidate arrays alternate by step parity;
// 9 candidate buffers rotate so that the one read as the pivot is never
// written in the same step); no HBM traffic inside the loop.
// Ties between equal norms go to the lowest original column, where xgeqp3
// takes the lowest current position; only exactly equal norms can differ.
// The explicit Q is formed afterwards from V with compact-WY T factors and
// FP64-MFMA GEMMs.
#pragma once
#include "dq_common.hpp"
#include "dq_factor.hpp"

#define QRP_WAVES 8
#define QRP_THREADS (64 * QRP_WAVES)

__device__ inline double readlane_d(double v, int l)
{
    const long long u = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)u, l);
    const int hi = __builtin_amdgcn_readlane((int)(u >> 32), l);
    return __longlong_as_double((long long)(unsigned)lo | ((long long)hi << 32));
}

// --- wave64 reductions in registers (no LDS round trips) -------------------
// gfx950 lane exchanges:  permlane32_swap(x, y) -> ([x lanes 0-31, y lanes
// 0-31], [x lanes 32-63, y lanes 32-63]);  permlane16_swap(x, y) ->
// ([x row 0, y row 0, x row 2, y row 2], [x row 1, y row 1, x row 3, y row 3]);
// DPP row_ror:n reads lane (i - n) mod 16 of the same 16-lane row.
__device__ inline void swap32_d(double& x, double& y)
{
    const long long ux = __double_as_longlong(x), uy = __double_as_longlong(y);
    const auto lo = __builtin_amdgcn_permlane32_swap((unsigned)ux, (unsigned)uy, false, false);
    const auto hi = __builtin_amdgcn_permlane32_swap((unsigned)(ux >> 32), (unsigned)(uy >> 32),
                                                     false, false);
    x = __longlong_as_double(((long long)hi[0] << 32) | lo[0]);
    y = __longlong_as_double(((long long)hi[1] << 32) | lo[1]);
}
__device__ inline void swap16_d(double& x, double& y)
{
    const long long ux = __double_as_longlong(x), uy = __double_as_longlong(y);
    const auto lo = __builtin_amdgcn_permlane16_swap((unsigned)ux, (unsigned)uy, false, false);
    const auto hi = __builtin_amdgcn_permlane16_swap((unsigned)(ux >> 32), (unsigned)(uy >> 32),
                                                     false, false);
    x = __longlong_as_double(((long long)hi[0] << 32) | lo[0]);
    y = __longlong_as_double(((long long)hi[1] << 32) | lo[1]);
}
template <int CTRL>
__device__ inline double dpp_d(double v)
{
    const long long u = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)u, CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(u >> 32), CTRL, 0xf, 0xf, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
#define DPP_ROR(n) (0x120 + (n))
#define DPP_MIRROR 0x140
#define DPP_HALF_MIRROR 0x141
#define DPP_QUAD_REV 0x1B   // quad_perm [3,2,1,0]
#define DPP_QUAD_XOR1 0xB1  // quad_perm [1,0,3,2]

// Sums K in {1, 2, 4, 8} per-lane values over the wave at once, transposed:
// each exchange halves the values a lane carries (lanes 32-63 take the upper
// half, then odd rows, then the upper 8 lanes of a row), the rest is summed
// within the lane groups.  Value i ends in lanes (64 / K) i ... ; on return
// every a[i] holds its full sum, uniform over the wave.
template <int K>
__device__ inline void wave_sum_many(double (&a)[K], int lane)
{
    static_assert(K == 1 || K == 2 || K == 4 || K == 8, "K in {1, 2, 4, 8}");
    constexpr int KA = K > 1 ? K / 2 : 1;
    constexpr int KB = KA > 1 ? KA / 2 : 1;
    double c1[KA], c2[KB];
#pragma unroll
    for (int i = 0; i < KA; i++) {
        double x = a[i], y = a[K > 1 ? i + KA : i];
        swap32_d(x, y);
        c1[i] = x + y;
    }
#pragma unroll
    for (int i = 0; i < KB; i++) {
        double x = c1[i], y = c1[KA > 1 ? i + KB : i];
        swap16_d(x, y);
        c2[i] = x + y;
    }
    double cur;
    if constexpr (K == 8) {
        const bool up = (lane & 8) != 0;
        const double keep = up ? c2[1] : c2[0], send = up ? c2[0] : c2[1];
        cur = keep + dpp_d<DPP_MIRROR>(send);
        cur += dpp_d<DPP_HALF_MIRROR>(cur);
        cur += dpp_d<DPP_QUAD_REV>(cur);
        cur += dpp_d<DPP_QUAD_XOR1>(cur);
    } else {
        cur = c2[0];
        cur += dpp_d<DPP_ROR(8)>(cur);
        cur += dpp_d<DPP_ROR(4)>(cur);
        cur += dpp_d<DPP_ROR(2)>(cur);
        cur += dpp_d<DPP_ROR(1)>(cur);
    }
#pragma unroll
    for (int i = 0; i < K; i++) a[i] = readlane_d(cur, (64 / K) * i);
}
__device__ inline double wave_sum_r(double v, int lane)
{
    double a[1] = {v};
    wave_sum_many<1>(a, lane);
    return a[0];
}
__device__ inline zd wave_sum_r(zd v, int lane)
{
    double a[2] = {v.x, v.y};
    wave_sum_many<2>(a, lane);
    return mkz(a[0], a[1]);
}

// component-wise select (a conditional over whole complex values would be
// lowered through memory)
__device__ inline double selv(bool c, double a, double b) { return c ? a : b; }
__device__ inline zd selv(bool c, zd a, zd b) { return mkz(c ? a.x : b.x, c ? a.y : b.y); }

// xlarfg (as larfg in dq_factor.hpp) written branch-free, both outcomes
// computed and selected, so that complex values stay in registers
__device__ inline void larfg_sel(double alpha, double xnorm2, double& beta, double& tau,
                                 double& scal)
{
    const bool triv = xnorm2 == 0.0;
    const double b = -copysign(sqrt(alpha * alpha + xnorm2), alpha);
    tau = triv ? 0.0 : (b - alpha) / b;
    scal = triv ? 0.0 : 1.0 / (alpha - b);
    beta = triv ? alpha : b;
}
__device__ inline void larfg_sel(zd alpha, double xnorm2, zd& beta, zd& tau, zd& scal)
{
    const bool triv = xnorm2 == 0.0 && alpha.y == 0.0;
    const double b = -copysign(sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xnorm2), alpha.x);
    const zd t = mkz((b - alpha.x) / b, -alpha.y / b);
    const zd sc = zdiv(mkz(1.0, 0.0), mkz(alpha.x - b, alpha.y));
    tau = selv(triv, mkz(0.0, 0.0), t);
    scal = selv(triv, mkz(0.0, 0.0), sc);
    beta = selv(triv, alpha, mkz(b, 0.0));
}

template <typename T> struct qrp_parts { static const int n = 1; };
template <> struct qrp_parts<zd> { static const int n = 2; };
__device__ inline void put_part(double* a, double v) { a[0] = v; }
__device__ inline void put_part(double* a, zd v) { a[0] = v.x; a[1] = v.y; }
__device__ inline void get_part(const double* a, double& v) { v = a[0]; }
__device__ inline void get_part(const double* a, zd& v) { v = mkz(a[0], a[1]); }

// Apply H^H = I - conj(tau) v v^H of step j to a batch of CB columns xb
// (columns q0 + 8 u, live where bit u of lv is set) and recompute their
// squared norms over rows > j exactly; the dots v^H x and the norms each take
// one batched register reduction.  The best live column (largest norm, first
// on ties) is copied to the candidate buffer cw.
template <typename T, int RPL, int CB>
__device__ __forceinline__ void qrp_apply_batch(T (&xb)[CB][RPL], unsigned lv, int q0, int j,
                                                int lane, const T (&v)[RPL], T ct, int t0,
                                                double rm0, double& bestv, int& bestq, T* cw)
{
    constexpr int NP = qrp_parts<T>::n;
    const T zero = from_re<T>(0.0);
    double wr[CB * NP];
#pragma unroll
    for (int u = 0; u < CB * NP; u++) wr[u] = 0.0;
#pragma unroll
    for (int t = 0; t < RPL; t++) {
        if (64 * t + 63 >= j) {
#pragma unroll
            for (int u = 0; u < CB; u++) {
                T d;
                get_part(&wr[u * NP], d);
                d += cj(v[t]) * xb[u][t];
                put_part(&wr[u * NP], d);
            }
        }
    }
    wave_sum_many<CB * NP>(wr, lane);
    // a retired column gets f = 0 (and rows above j see v = 0): unchanged
    T f[CB];
#pragma unroll
    for (int u = 0; u < CB; u++) {
        T w;
        get_part(&wr[u * NP], w);
        f[u] = selv((lv >> u) & 1, ct * w, zero);
    }
    double sn[CB];
#pragma unroll
    for (int u = 0; u < CB; u++) sn[u] = 0.0;
#pragma unroll
    for (int t = 0; t < RPL; t++) {
        if (64 * t + 63 >= j) {
            const double mk = (t == t0) ? rm0 : 1.0;
#pragma unroll
            for (int u = 0; u < CB; u++) {
                xb[u][t] -= v[t] * f[u];
                sn[u] += mk * abs2(xb[u][t]);
            }
        }
    }
    wave_sum_many<CB>(sn, lane);
#pragma unroll
    for (int u = 0; u < CB; u++) {
        if (((lv >> u) & 1) && sn[u] > bestv) {
            bestv = sn[u];
            bestq = q0 + QRP_WAVES * u;
#pragma unroll
            for (int t = 0; t < RPL; t++) cw[lane + 64 * t] = xb[u][t];
        }
    }
}

// m x n matrix A[b] (lda), n <= m <= 64 RPL, n <= 8 (SR + SL), SR a multiple
// of CB.  Outputs: A holds R (upper) and the reflectors (below the diagonal)
// in pivoted column order, tau[b][0..n), jpvt[b][j] = 1-based original column
// of pivot j, and the explicit unit-lower V[b] (ldv, m x n).  CB: register
// columns per batched reduction (LDS columns go in pairs).
template <typename T, int RPL, int SR, int SL, int CB>
__global__ __launch_bounds__(QRP_THREADS) void k_qrp_res(int m, int n, T* A_, int lda, long sA,
                                                         T* tau_, long stau, int* jpvt_, long spiv,
                                                         T* V_, int ldv, long sV)
{
    static_assert(SR % CB == 0, "register columns come in whole batches");
    constexpr int MR = 64 * RPL;
    __shared__ T Ls[QRP_WAVES][SL > 0 ? SL : 1][MR];
    __shared__ T Cw[QRP_WAVES + 1][MR];     // candidate columns (one spare)
    __shared__ double candv[2][QRP_WAVES];  // per wave: best norm, column, buffer
    __shared__ int candq[2][QRP_WAVES];
    __shared__ int candb[2][QRP_WAVES];
    const int b = blockIdx.x;
    T* A = A_ + b * sA;
    T* V = V_ + b * sV;
    T* tau = tau_ + b * stau;
    int* jp = jpvt_ + b * spiv;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // uniform: scalar branches
    const T zero = from_re<T>(0.0);

    // load: column q = wave + 8 s; LDS slots s < SL, register slots after
    T x[SR][RPL];
#pragma unroll
    for (int s = 0; s < SR; s++) {
        const int q = wave + QRP_WAVES * (SL + s);
#pragma unroll
        for (int t = 0; t < RPL; t++) {
            const int r = lane + 64 * t;
            x[s][t] = (q < n && r < m) ? A[r + (long)lda * q] : zero;
        }
    }
    for (int s = 0; s < SL; s++) {
        const int q = wave + QRP_WAVES * s;
#pragma unroll
        for (int t = 0; t < RPL; t++) {
            const int r = lane + 64 * t;
            Ls[wave][s][r] = (q < n && r < m) ? A[r + (long)lda * q] : zero;
        }
    }
    // live columns of this wave (bit s), uniform
    unsigned long long live = 0;
    for (int s = 0; s < SR + SL; s++)
        if (wave + QRP_WAVES * s < n) live |= 1ull << s;
    int mybuf = wave, spare = QRP_WAVES;
    // initial norms; the running best column goes to the wave's buffer
    {
        double bestv = -1.0;
        int bestq = -1;
        for (int s = 0; s < SL; s++) {
            if (!((live >> s) & 1)) continue;
            double sn = 0.0;
#pragma unroll
            for (int t = 0; t < RPL; t++) sn += abs2(Ls[wave][s][lane + 64 * t]);
            sn = wave_sum_r(sn, lane);
            const int q = wave + QRP_WAVES * s;
            if (sn > bestv) {
                bestv = sn;
                bestq = q;
#pragma unroll
                for (int t = 0; t < RPL; t++) Cw[mybuf][lane + 64 * t] = Ls[wave][s][lane + 64 * t];
            }
        }
#pragma unroll
        for (int s = 0; s < SR; s++) {
            if (!((live >> (SL + s)) & 1)) continue;
            double sn = 0.0;
#pragma unroll
            for (int t = 0; t < RPL; t++) sn += abs2(x[s][t]);
            sn = wave_sum_r(sn, lane);
            const int q = wave + QRP_WAVES * (SL + s);
            if (sn > bestv) {
                bestv = sn;
                bestq = q;
#pragma unroll
                for (int t = 0; t < RPL; t++) Cw[mybuf][lane + 64 * t] = x[s][t];
            }
        }
        if (lane == 0) { candv[0][wave] = bestv; candq[0][wave] = bestq; candb[0][wave] = mybuf; }
    }
    __syncthreads();

    for (int j = 0; j < n; j++) {
        const int cb = j & 1;
        // (1) pivot: the largest candidate norm, first column on ties
        double pv = candv[cb][0];
        int p = candq[cb][0], wp = 0;
#pragma unroll
        for (int w = 1; w < QRP_WAVES; w++) {
            const double cv = candv[cb][w];
            const int cq = candq[cb][w];
            if (cv > pv || (cv == pv && cq >= 0 && (p < 0 || cq < p))) { pv = cv; p = cq; wp = w; }
        }
        p = __builtin_amdgcn_readfirstlane(p);
        wp = __builtin_amdgcn_readfirstlane(wp);
        const int bp = __builtin_amdgcn_readfirstlane(candb[cb][wp]);
        // (2) reflector of the pivot column, built by every wave
        T xv[RPL];
        double xn2 = 0.0;
#pragma unroll
        for (int t = 0; t < RPL; t++) {
            const int r = lane + 64 * t;
            xv[t] = Cw[bp][r];
            if (r > j) xn2 += abs2(xv[t]);
        }
        xn2 = wave_sum_r(xn2, lane);
        const T alpha = Cw[bp][j];
        T beta, tj, scal;
        larfg_sel(alpha, xn2, beta, tj, scal);
        T v[RPL];
#pragma unroll
        for (int t = 0; t < RPL; t++) {
            const int r = lane + 64 * t;
            v[t] = selv(r > j, xv[t] * scal, from_re<T>(r == j ? 1.0 : 0.0));
        }
        if (wave == wp) {
            // retire the pivot column as column j of the factorization
#pragma unroll
            for (int t = 0; t < RPL; t++) {
                const int r = lane + 64 * t;
                if (r < m) {
                    A[r + (long)lda * j] = selv(r < j, xv[t], selv(r == j, beta, v[t]));
                    V[r + (long)ldv * j] = v[t];
                }
            }
            if (lane == 0) { tau[j] = tj; jp[j] = p + 1; }
            live &= ~(1ull << ((p - wave) / QRP_WAVES));
            // the pivot's buffer is being read this step: switch to the spare
            mybuf = spare;
        }
        spare = bp;  // free again once every wave has read it (next barrier)
        // (3) apply H^H to the live columns, recompute their norms, and keep
        //     the best one as this wave's next candidate
        const T ct = cj(tj);
        const int t0 = j >> 6;
        const double rm0 = (lane > (j & 63)) ? 1.0 : 0.0;   // rows > j in row slot t0
        double bestv = -1.0;
        int bestq = -1;
        for (int s0 = 0; s0 < SL; s0 += 2) {
            const unsigned lv = (unsigned)(live >> s0) & (s0 + 1 < SL ? 3u : 1u);
            if (lv == 0) continue;
            T xb[2][RPL];
#pragma unroll
            for (int u = 0; u < 2; u++)
#pragma unroll
                for (int t = 0; t < RPL; t++)
                    xb[u][t] = (s0 + u < SL) ? Ls[wave][s0 + u][lane + 64 * t] : zero;
            qrp_apply_batch<T, RPL, 2>(xb, lv, wave + QRP_WAVES * s0, j, lane, v, ct, t0, rm0,
                                       bestv, bestq, Cw[mybuf]);
#pragma unroll
            for (int u = 0; u < 2; u++)
                if (s0 + u < SL) {
#pragma unroll
                    for (int t = 0; t < RPL; t++) Ls[wave][s0 + u][lane + 64 * t] = xb[u][t];
                }
        }
#pragma unroll
        for (int c0 = 0; c0 < SR; c0 += CB) {
            const unsigned lv = (unsigned)(live >> (SL + c0)) & ((1u << CB) - 1);
            if (lv == 0) continue;  // whole batch retired
            qrp_apply_batch<T, RPL, CB>(*reinterpret_cast<T(*)[CB][RPL]>(&x[c0][0]), lv,
                                        wave + QRP_WAVES * (SL + c0), j, lane, v, ct, t0, rm0,
                                        bestv, bestq, Cw[mybuf]);
        }
        if (lane == 0) {
            candv[cb ^ 1][wave] = bestv;
            candq[cb ^ 1][wave] = bestq;
            candb[cb ^ 1][wave] = mybuf;
        }
        __syncthreads();
    }
}

// Compact-WY factor of each NB-column panel of reflectors from its Gram
// matrix (xlarft): on entry Tk[p] holds V_p^H V_p (NB x NB, ld NB), on exit
// the upper-triangular T with H_{k0} ... H_{k0+nb-1} = I - V_p T V_p^H,
//   T(j, j) = tau_j,  T(0:j, j) = -tau_j T(0:j, 0:j) (V_p^H v_j)(0:j).
// One wave per (panel, matrix); NB <= 64.
template <typename T>
__global__ __launch_bounds__(64) void k_larft_gram(int NB, int kmax, const T* tau_, long stau,
                                                   T* Tb_, long sT)
{
    const int p = blockIdx.x, b = blockIdx.y, lane = threadIdx.x;
    const int k0 = p * NB;
    const int nbp = (kmax - k0 < NB) ? kmax - k0 : NB;
    const T* tau = tau_ + b * stau + k0;
    T* Tk = Tb_ + b * sT + (long)p * NB * NB;
    for (int j = 0; j < nbp; j++) {
        const T z = (lane < j) ? Tk[lane + NB * j] : from_re<T>(0.0);
        const T tj = tau[j];
        T s = from_re<T>(0.0);
        for (int c = 0; c < j; c++) {
            const T zc = shfl(z, c);
            if (lane <= c) s += Tk[lane + NB * c] * zc;
        }
        if (lane < nbp) {
            T out = from_re<T>(0.0);
            if (lane < j) out = -(tj * s);
            else if (lane == j) out = tj;
            Tk[lane + NB * j] = out;
        }
    }
}
