// dq_qr_blocked.hpp -- blocked Householder QR (geqrf / orgqr semantics) for a
// batch of matrices: a register-resident panel kernel plus FP64-MFMA GEMMs.
//
// Per panel of nb columns (starting at column k0):
//   k_qr_panel: the (m-k0) x nb panel lives in registers (lane = row slot,
//     wave = column slot), each Householder step is a wave64 reduction for the
//     norm, an LDS broadcast of v, and one wave reduction per remaining panel
//     column.  The same reductions also give V(:,0:j)^H v_j, from which the
//     compact-WY factor T of the panel is built (the xlarft recurrence
//     T(0:j,j) = -tau_j T(0:j,0:j) V(:,0:j)^H v_j).  The determinant of the
//     panel's reflectors (1 - tau v^H v each) is multiplied into det[b].
//   trailing update: A(k0:, k0+nb:) -= V T^H (V^H A(k0:, k0+nb:)), three GEMMs.
// The explicit Q = H_0 H_1 ... is accumulated backwards with the same
// three-GEMM step.  Matrices are column-major with leading dimension lda.
#pragma once
#include "dq_common.hpp"
#include "dq_factor.hpp"

#define QRB_THREADS 1024
#define QRB_WAVES 16

// QM: row slots per lane (panel height <= 64*QM); NCW: columns per wave
// (panel width nb = 16*NCW).
template <typename T, int QM, int NCW>
__global__ __launch_bounds__(QRB_THREADS) void k_qr_panel(int m, int k0, int nbp, T* A_, int lda,
                                                          long sA, T* tau_, long stau, T* V_,
                                                          int ldv, long sV, T* Tw_, long sT,
                                                          T* det_, int det_init)
{
    constexpr int NB = 16 * NCW;
    __shared__ T v[64 * QM];
    __shared__ T Ts[NB * NB];
    __shared__ T z[NB];
    __shared__ T s_tau;
    __shared__ T s_det;
    const int b = blockIdx.x;
    T* A = A_ + b * sA + k0 + (long)lda * k0;  // panel origin
    const int mp = m - k0;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    T x[QM][NCW];
#pragma unroll
    for (int cs = 0; cs < NCW; cs++) {
        const int jc = wave + 16 * cs;
#pragma unroll
        for (int q = 0; q < QM; q++) {
            const int r = lane + 64 * q;
            x[q][cs] = (jc < nbp && r < mp) ? A[r + (long)lda * jc] : from_re<T>(0.0);
        }
    }
    for (int e = tid; e < NB * NB; e += QRB_THREADS) Ts[e] = from_re<T>(0.0);
    if (tid == 0) s_det = from_re<T>(1.0);
    __syncthreads();

#pragma unroll
    for (int cs = 0; cs < NCW; cs++) {
        for (int w0 = 0; w0 < 16; w0++) {
            const int j = w0 + 16 * cs;
            if (j >= nbp) break;
            // (a) reflector of panel column j (owner wave w0, register column cs)
            if (wave == w0) {
                T alpha = from_re<T>(0.0);
                double xn2 = 0.0;
#pragma unroll
                for (int q = 0; q < QM; q++) {
                    const int r = lane + 64 * q;
                    if (r == j) alpha = x[q][cs];
                    if (r > j && r < mp) xn2 += abs2(x[q][cs]);
                }
                xn2 = wave_sum_r(xn2, lane);
                alpha = shfl(alpha, j & 63);
                T beta, t, scal;
                larfg(alpha, xn2, &beta, &t, &scal);
#pragma unroll
                for (int q = 0; q < QM; q++) {
                    const int r = lane + 64 * q;
                    if (r > j && r < mp) x[q][cs] = x[q][cs] * scal;
                    if (r == j) x[q][cs] = beta;
                    if (r < mp) v[r] = (r > j) ? x[q][cs] : from_re<T>(r == j ? 1.0 : 0.0);
                }
                if (lane == 0) {
                    s_tau = t;
                    Ts[j + NB * j] = t;
                    // det(H_j) = 1 - tau v^H v, v^H v = 1 + |scal|^2 |x1|^2
                    const double vv = 1.0 + abs2(scal) * xn2;
                    s_det = s_det * (from_re<T>(1.0) - t * vv);
                }
            }
            __syncthreads();
            // (b) apply H_j^H to the panel columns right of j; dots with the
            //     reflectors left of j for the T recurrence.  The NCW dots of
            //     a wave are summed over the wave in one register reduction.
            const T tj = s_tau;
            const T ct = cj(tj);
            constexpr int NP = val_parts<T>::n;
            double dr[NCW * NP];
#pragma unroll
            for (int c2 = 0; c2 < NCW; c2++) {
                const int jc = wave + 16 * c2;
                T d = from_re<T>(0.0);
                if (jc < nbp && jc != j) {
#pragma unroll
                    for (int q = 0; q < QM; q++) {
                        const int r = lane + 64 * q;
                        if (r >= j && r < mp) {
                            if (jc > j) d += cj(v[r]) * x[q][c2];
                            else d += cj(x[q][c2]) * v[r];
                        }
                    }
                }
                put_part(&dr[c2 * NP], d);
            }
            wave_sum_many<NCW * NP>(dr, lane);
#pragma unroll
            for (int c2 = 0; c2 < NCW; c2++) {
                const int jc = wave + 16 * c2;
                if (jc < nbp && jc != j) {
                    T d;
                    get_part(&dr[c2 * NP], d);
                    if (jc > j) {
                        const T f = ct * d;
#pragma unroll
                        for (int q = 0; q < QM; q++) {
                            const int r = lane + 64 * q;
                            if (r >= j && r < mp) x[q][c2] -= v[r] * f;
                        }
                    } else if (lane == 0) {
                        z[jc] = d;
                    }
                }
            }
            __syncthreads();
            // (c) T(0:j, j) = -tau_j T(0:j, 0:j) z(0:j)   (j <= 63), by the wave
            //     that built reflector j: the owner of reflector j + 1 is
            //     another wave, so this overlaps the next step's (a)
            if (wave == w0 && lane < j) {
                // T is upper triangular (zeros below the diagonal), so the sum
                // may start at 0; the bound j is uniform, and the unrolled
                // loop issues its independent LDS reads together
                T s = from_re<T>(0.0);
#pragma unroll
                for (int a = 0; a < NB; a++)
                    if (a < j) s += Ts[lane + NB * a] * z[a];
                Ts[lane + NB * j] = -(tj * s);
            }
        }
    }
    __syncthreads();
    // write back: R (upper) and the reflectors (below) into A, explicit V, T, tau
    T* V = V_ + b * sV + k0 + (long)ldv * k0;
#pragma unroll
    for (int cs = 0; cs < NCW; cs++) {
        const int jc = wave + 16 * cs;
        if (jc < nbp) {
#pragma unroll
            for (int q = 0; q < QM; q++) {
                const int r = lane + 64 * q;
                if (r < mp) {
                    A[r + (long)lda * jc] = x[q][cs];
                    V[r + (long)ldv * jc] =
                        (r > jc) ? x[q][cs] : from_re<T>(r == jc ? 1.0 : 0.0);
                }
            }
        }
    }
    T* Tw = Tw_ + b * sT;
    for (int e = tid; e < NB * NB; e += QRB_THREADS) Tw[e] = Ts[e];
    T* tau = tau_ + b * stau + k0;
    for (int jc = tid; jc < nbp; jc += QRB_THREADS) tau[jc] = Ts[jc + NB * jc];
    if (tid == 0 && det_) det_[b] = det_init ? s_det : det_[b] * s_det;
}

// Solve U_kk X_k = X_k for one diagonal block of order nk <= 64 (U staged in
// LDS, LU_NC right-hand sides per wave, one row per lane).  Used by the blocked
// upper-triangular solve below.
template <typename T>
__global__ __launch_bounds__(256) void k_trsm_diag64(int nk, int ncols, const T* U_, long sU,
                                                     int ldu, T* X_, long sX, int ldx)
{
    __shared__ T Us[64 * 64];
    __shared__ T Ui[64];  // 1 / U_kk, formed once instead of once per column
    const int b = blockIdx.y;
    const T* U = U_ + b * sU;
    T* X = X_ + b * sX;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int e = tid; e < nk * nk; e += blockDim.x) {
        const int i = e % nk, k = e / nk;
        Us[i + 64 * k] = (i <= k) ? U[i + (long)ldu * k] : from_re<T>(0.0);
    }
    if (tid < nk) Ui[tid] = rdiv(from_re<T>(1.0), U[tid + (long)ldu * tid]);
    __syncthreads();
    const int c0 = (blockIdx.x * (blockDim.x >> 6) + wave) * LU_NC;
    if (c0 >= ncols) return;
    T x[LU_NC];
#pragma unroll
    for (int q = 0; q < LU_NC; q++)
        x[q] = (lane < nk && c0 + q < ncols) ? X[lane + (long)ldx * (c0 + q)] : from_re<T>(0.0);
    for (int k = nk - 1; k >= 0; k--) {
        const T rkk = Ui[k];
        const T uik = Us[lane + 64 * k];
#pragma unroll
        for (int q = 0; q < LU_NC; q++) {
            T xk = shfl(x[q], k);
            xk = xk * rkk;
            if (lane == k) x[q] = xk;
            if (lane < k) x[q] -= uik * xk;
        }
    }
#pragma unroll
    for (int q = 0; q < LU_NC; q++)
        if (lane < nk && c0 + q < ncols) X[lane + (long)ldx * (c0 + q)] = x[q];
}

// phase of the determinant of a product of Householder reflectors stored
// LAPACK-style (V below the diagonal of A, tau): prod_i (1 - tau_i v_i^H v_i)
template <typename T>
__global__ __launch_bounds__(64) void k_reflector_det(int m, int k, const T* A_, int lda, long sA,
                                                      const T* tau_, long stau, T* det)
{
    const int b = blockIdx.x, lane = threadIdx.x;
    const T* A = A_ + b * sA;
    const T* tau = tau_ + b * stau;
    T d = from_re<T>(1.0);
    for (int i = 0; i < k; i++) {
        double s = 0.0;
        for (int r = i + 1 + lane; r < m; r += 64) s += abs2(A[r + (long)lda * i]);
        s = wave_sum(s);
        d = d * (from_re<T>(1.0) - tau[i] * (1.0 + s));
    }
    if (lane == 0) det[b] = d;
}

// phase of det(I + B_L..B_1) from the stabilized factorization used by
// calc_eq_g on the GPU: G^-1 = Q Db M' with M' = Q_M R_M, so
//   phase = phase(det Q) * sign(det Db) * phase(det Q_M) * prod phase(R_M,ii)
// (the quantity src/greens.c:194-207 returns; Db = diag(1/v) is real).
template <typename T>
__global__ __launch_bounds__(64) void k_eqg_phase(int n, const T* detq, const T* detm,
                                                  const double* v_, long sv, const T* R_, long sR,
                                                  T* phase_out)
{
    const int b = blockIdx.x, lane = threadIdx.x;
    const double* v = v_ + b * sv;
    const T* R = R_ + b * sR;
    T ph = from_re<T>(1.0);
    int neg = 0;
    for (int i = lane; i < n; i += 64) {
        const T r = R[i + (long)n * i];
        ph = ph * (r * (1.0 / absv(r)));
        neg += (v[i] < 0.0);
    }
    // product over lanes
    for (int o = 32; o > 0; o >>= 1) {
        const T other = shfl(ph, lane ^ o);
        ph = ph * other;
        neg += __shfl_xor(neg, o, 64);
    }
    if (lane == 0) {
        T p = ph;
        const T dq = detq[b], dm = detm[b];
        p = p * (dq * (1.0 / absv(dq)));
        p = p * (dm * (1.0 / absv(dm)));
        if (neg & 1) p = p * (-1.0);
        phase_out[b] = p;
    }
}
