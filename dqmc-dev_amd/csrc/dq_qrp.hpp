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

// Apply H^H = I - conj(tau) v v^H of step j to a batch of CB columns xb
// (columns q0 + 8 u, live where bit u of lv is set) and recompute their
// squared norms over rows > j exactly; the dots v^H x and the norms each take
// one batched register reduction.  The best live column (largest norm, first
// on ties) is copied to the candidate buffer cw.
// Every row slot takes part: v is zero above row j, so those rows add exact
// zeros to the dots and stay unchanged by the update, and the row mask mk
// (1 below row j, else 0) keeps them out of the norms.  (Skipping the slots
// above j by a test per slot was compiled to selects on every value.)
template <typename T, int RPL, int CB>
__device__ __forceinline__ void qrp_apply_batch(T (&xb)[CB][RPL], unsigned lv, int q0, int lane,
                                                const T (&v)[RPL], T ct, const double (&mk)[RPL],
                                                double& bestv, int& bestq, T* cw)
{
    constexpr int NP = val_parts<T>::n;
    const T zero = from_re<T>(0.0);
    double wr[CB * NP];
#pragma unroll
    for (int u = 0; u < CB * NP; u++) wr[u] = 0.0;
#pragma unroll
    for (int t = 0; t < RPL; t++) {
#pragma unroll
        for (int u = 0; u < CB; u++) {
            T d;
            get_part(&wr[u * NP], d);
            d += cj(v[t]) * xb[u][t];
            put_part(&wr[u * NP], d);
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
#pragma unroll
        for (int u = 0; u < CB; u++) {
            xb[u][t] -= v[t] * f[u];
            sn[u] += mk[t] * abs2(xb[u][t]);
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
// of pivot j, the explicit unit-lower V[b] (ldv, m x n) and, if det_ is
// non-null, det_[b] = prod_j det(H_j) = prod_j (1 - tau_j v_j^H v_j).  CB:
// register columns per batched reduction (LDS columns go in pairs).
template <typename T, int RPL, int SR, int SL, int CB>
__global__ __launch_bounds__(QRP_THREADS) void k_qrp_res(int m, int n, T* A_, int lda, long sA,
                                                         T* tau_, long stau, int* jpvt_, long spiv,
                                                         T* V_, int ldv, long sV, T* det_)
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

    T det = from_re<T>(1.0);  // product of the reflectors' determinants (uniform)
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
        // det(H_j) = 1 - tau v^H v, v^H v = 1 + |scal|^2 |x1|^2
        det = det * (from_re<T>(1.0) - tj * (1.0 + abs2(scal) * xn2));
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
        double mk[RPL];  // rows > j
#pragma unroll
        for (int t = 0; t < RPL; t++) mk[t] = (lane + 64 * t > j) ? 1.0 : 0.0;
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
            qrp_apply_batch<T, RPL, 2>(xb, lv, wave + QRP_WAVES * s0, lane, v, ct, mk, bestv,
                                       bestq, Cw[mybuf]);
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
                                        wave + QRP_WAVES * (SL + c0), lane, v, ct, mk, bestv,
                                        bestq, Cw[mybuf]);
        }
        if (lane == 0) {
            candv[cb ^ 1][wave] = bestv;
            candq[cb ^ 1][wave] = bestq;
            candb[cb ^ 1][wave] = mybuf;
        }
        __syncthreads();
    }
    if (det_ && tid == 0) det_[b] = det;
}

// Compact-WY factor of each NB-column panel of reflectors from its Gram
// matrix (xlarft): on entry Tk[p] holds V_p^H V_p (NB x NB, ld NB), on exit
// the upper-triangular T with H_{k0} ... H_{k0+nb-1} = I - V_p T V_p^H,
//   T(j, j) = tau_j,  T(0:j, j) = -tau_j T(0:j, 0:j) (V_p^H v_j)(0:j).
// One 512-thread workgroup per (panel, matrix), NB <= 64, the panel in LDS,
// in blocks of 8 columns: the diagonal blocks by the recurrence (one wave
// each, in parallel), then block column B from the finished T(0:8B, 0:8B)
// as T12 = -T1 (V1^H V2) T2 (the xlarft merge).  (One wave walking the 64
// columns in order took 222 us per call at 512 matrices of order 256.)
template <typename T>
__global__ __launch_bounds__(512) void k_larft_gram(int NB, int kmax, const T* tau_, long stau,
                                                    T* Tb_, long sT)
{
    __shared__ T S[64 * 64];  // strict upper: the Gram entries, then T
    __shared__ T Wb[64 * 8];
    const int p = blockIdx.x, b = blockIdx.y, tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const int k0 = p * NB;
    const int nbp = (kmax - k0 < NB) ? kmax - k0 : NB;
    const T* tau = tau_ + b * stau + k0;
    T* Tk = Tb_ + b * sT + (long)p * NB * NB;
    const T zero = from_re<T>(0.0);
    // columns past nbp: tau = 0 and no Gram entries, so their T columns are 0
    for (int e = tid; e < 64 * 64; e += 512) {
        const int r = e & 63, c = e >> 6;
        T v = zero;
        if (r < nbp && c < nbp) v = (r == c) ? tau[c] : (r < c ? Tk[r + NB * c] : zero);
        S[e] = v;
    }
    __syncthreads();
    const int nbk = (nbp + 7) / 8;  // blocks past the panel's columns stay zero
    if (wave < nbk && lane < 8) {  // diagonal block `wave`
        const int o = 8 * wave, i = o + lane;
        for (int jj = 1; jj < 8; jj++) {
            const int jc = o + jj;
            T sum = zero;
            for (int a = i; a < jc; a++) sum += S[i + 64 * a] * S[a + 64 * jc];
            if (i < jc) S[i + 64 * jc] = -(S[jc + 64 * jc] * sum);
        }
    }
    __syncthreads();
    for (int B = 1; B < nbk; B++) {
        const int rows = 8 * B, c0 = 8 * B;
        const int i = tid % rows, cc = tid / rows;
        const bool act = tid < 8 * rows;
        if (act) {
            T w = zero;
            for (int a = i; a < rows; a++) w += S[i + 64 * a] * S[a + 64 * (c0 + cc)];
            Wb[i + rows * cc] = w;
        }
        __syncthreads();
        if (act) {
            T t = zero;
            for (int c2 = 0; c2 <= cc; c2++) t += Wb[i + rows * c2] * S[(c0 + c2) + 64 * (c0 + cc)];
            S[i + 64 * (c0 + cc)] = -t;
        }
        __syncthreads();
    }
    for (int e = tid; e < 64 * 64; e += 512) {
        const int r = e & 63, c = e >> 6;
        if (r < nbp && c < nbp) Tk[r + NB * c] = S[e];
    }
}
