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


// QM: row slots per lane (panel height <= 64*QM); NW waves of NCW columns
// (panel width nb = NW*NCW).  Panel column jc belongs to wave jc % NW,
// register slot jc / NW.  8 waves where the registers allow it: two
// workgroups then share a CU (at most 128 VGPRs each), and the latency of one
// workgroup's steps hides behind the other's.
//
// One workgroup barrier per Householder step j.  Between two barriers:
//   (b) every wave applies H_j^H to its columns right of j and takes the dots
//       V(:,jc)^H v_j of its columns left of j (the compact-WY recurrence);
//       v_j is zero outside rows j .. mp-1, so the row loops run unmasked;
//   (a) the owner of column j + 1, right after its own update of that column,
//       builds reflector j + 1 into the other v buffer;
//   the dots V(:,jc)^H v_j (jc < j) go into the strict upper triangle of the
//   T buffer, tau_j onto its diagonal.
// After the last step T comes from these in blocks of 8 columns: the diagonal
// blocks by the recurrence (one wave per block, in parallel), then block
// column B from the finished T(0:8B, 0:8B) as T12 = -T1 (V1^H V2) T2 (the
// xlarft merge).  Building T(0:j-1, j-1) inside the steps put one wave's
// serial sum of up to 64 LDS reads on every step's critical path (176 against
// 124 us per panel without it).
// f(integral_constant<int, i>) for i = 0 .. K-1, unrolled at compile time
template <int K, int I = 0, typename F>
__device__ __forceinline__ void static_for(F&& f)
{
    if constexpr (I < K) {
        f(std::integral_constant<int, I>());
        static_for<K, I + 1>(f);
    }
}

template <typename T, int QM, int NCW, int NW>
__global__ __launch_bounds__(64 * NW, 4) void k_qr_panel(int m, int k0, int nbp, T* A_, int lda,
                                                          long sA, T* tau_, long stau, T* V_,
                                                          int ldv, long sV, T* Tw_, long sT,
                                                          T* det_, int det_init)
{
    constexpr int NB = NW * NCW, NT = 64 * NW;
    constexpr int NP = val_parts<T>::n;
    __shared__ T v[2][64 * QM];
    __shared__ T Ts[NB * NB];
    __shared__ T s_det;
    const int b = blockIdx.x;
    T* A = A_ + b * sA + k0 + (long)lda * k0;  // panel origin
    const int mp = m - k0;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const T zero = from_re<T>(0.0), one = from_re<T>(1.0);
    T x[QM][NCW];
#pragma unroll
    for (int cs = 0; cs < NCW; cs++) {
        const int jc = wave + NW * cs;
#pragma unroll
        for (int q = 0; q < QM; q++) {
            const int r = lane + 64 * q;
            x[q][cs] = (jc < nbp && r < mp) ? A[r + (long)lda * jc] : zero;
        }
    }
    for (int e = tid; e < NB * NB; e += NT) Ts[e] = zero;
    if (tid == 0) s_det = one;
    __syncthreads();

    // (a) reflector of panel column jn (register slot cs of its owner wave):
    // v into buffer jn & 1 (zero outside rows jn .. mp-1), tau on the
    // diagonal of T, det(H) = 1 - tau v^H v with v^H v = 1 + |scal|^2 |x1|^2
    auto reflector = [&](auto csc, int jn) __attribute__((always_inline)) {
        constexpr int cs = decltype(csc)::value;
        T alpha = zero;
        double xn2 = 0.0;
#pragma unroll
        for (int q = 0; q < QM; q++) {
            const int r = lane + 64 * q;
            if (r == jn) alpha = x[q][cs];
            if (r > jn && r < mp) xn2 += abs2(x[q][cs]);
        }
        xn2 = wave_sum_r(xn2, lane);
        alpha = bcast(alpha, jn & 63);
        T beta, t, scal;
        larfg(alpha, xn2, &beta, &t, &scal);
        T* vb = v[jn & 1];
#pragma unroll
        for (int q = 0; q < QM; q++) {
            const int r = lane + 64 * q;
            const bool below = r > jn && r < mp;
            // component-wise selects: a conditional over whole complex
            // values would be lowered through scratch memory
            x[q][cs] = selv(below, x[q][cs] * scal, selv(r == jn, beta, x[q][cs]));
            vb[r] = selv(below, x[q][cs], selv(r == jn, one, zero));
        }
        if (lane == 0) {
            Ts[jn + NB * jn] = t;
            const double vv = 1.0 + abs2(scal) * xn2;
            s_det = s_det * (one - t * vv);
        }
    };
    if (nbp > 0 && wave == 0) reflector(std::integral_constant<int, 0>(), 0);
    __syncthreads();
    // the register slot of the step's columns is a compile-time constant
    // (a run-time index into x would move the panel to scratch memory)
    static_for<NCW>([&](auto csc) __attribute__((always_inline)) {
        constexpr int cs = decltype(csc)::value;
        for (int w0 = 0; w0 < NW; w0++) {
            const int j = w0 + NW * cs;
            if (j >= nbp) break;
            // the owner of column j + 1 is the step's critical path: its wave
            // issues ahead of the others until its reflector is out
            const int nxt = (j + 1) % NW;
            const bool crit = j + 1 < nbp && wave == nxt;
            if (crit) __builtin_amdgcn_s_setprio(2);
            // (b) d_jc = v_j^H x_jc for every column of the wave, summed over
            //     the wave in one register reduction
            const T* vb = v[j & 1];
            const T ct = cj(Ts[j + NB * j]);
            double dr[NCW * NP];
#pragma unroll
            for (int c2 = 0; c2 < NCW; c2++) {
                T d = zero;
#pragma unroll
                for (int q = 0; q < QM; q++) d += cj(vb[lane + 64 * q]) * x[q][c2];
                put_part(&dr[c2 * NP], d);
            }
            wave_sum_many<NCW * NP>(dr, lane);
#pragma unroll
            for (int c2 = 0; c2 < NCW; c2++) {
                const int jc = wave + NW * c2;
                T d;
                get_part(&dr[c2 * NP], d);
                if (jc > j) {
                    const T f = ct * d;
#pragma unroll
                    for (int q = 0; q < QM; q++) x[q][c2] -= vb[lane + 64 * q] * f;
                } else if (jc < j && lane == 0) {
                    Ts[jc + NB * j] = cj(d);  // V(:,jc)^H v_j
                }
            }
            // (a) of step j + 1 by the owner of column j + 1
            if (j + 1 < nbp) {
                if (w0 < NW - 1) {
                    if (wave == w0 + 1) reflector(csc, j + 1);
                } else if constexpr (cs + 1 < NCW) {
                    if (wave == 0) reflector(std::integral_constant<int, cs + 1>(), j + 1);
                }
            }
            if (crit) __builtin_amdgcn_s_setprio(0);
            __syncthreads();
        }
    });
    __syncthreads();
    // T from S = strict upper V^H V (in Ts) and tau (its diagonal).  Columns
    // past nbp have tau = 0 and S = 0, so their T columns come out zero.
    constexpr int NBK = NB / 8;
    static_assert(NB % 8 == 0 && NBK <= NW, "one wave per diagonal block");
    const int nbk = (nbp + 7) / 8;  // blocks past the panel's columns stay zero
    // diagonal blocks: T(i, jj) = -tau_jj sum_{a = i}^{jj-1} T(i, a) S(a, jj)
    if (wave < nbk && lane < 8) {
        const int o = 8 * wave, i = o + lane;
        for (int jj = 1; jj < 8; jj++) {
            const int jc = o + jj;
            T sum = zero;
            for (int a = i; a < jc; a++) sum += Ts[i + NB * a] * Ts[a + NB * jc];
            if (i < jc) Ts[i + NB * jc] = -(Ts[jc + NB * jc] * sum);
        }
    }
    __syncthreads();
    // block column B: W = T(0:8B, 0:8B) S(0:8B, cols), then T(0:8B, cols) =
    // -W T_BB, W passing through the (now free) v buffer
    T* Wb = &v[0][0];
    for (int B = 1; B < nbk; B++) {
        const int rows = 8 * B, c0 = 8 * B;
        const int i = tid % rows, cc = tid / rows;
        const bool act = tid < 8 * rows;
        if (act) {
            T w = zero;
            for (int a = i; a < rows; a++) w += Ts[i + NB * a] * Ts[a + NB * (c0 + cc)];
            Wb[i + rows * cc] = w;
        }
        __syncthreads();
        if (act) {
            T t = zero;
            for (int c2 = 0; c2 <= cc; c2++)
                t += Wb[i + rows * c2] * Ts[(c0 + c2) + NB * (c0 + cc)];
            Ts[i + NB * (c0 + cc)] = -t;
        }
        __syncthreads();
    }
    // write back: R (upper) and the reflectors (below) into A, explicit V, T, tau
    T* V = V_ + b * sV + k0 + (long)ldv * k0;
    // the row index passes through an empty asm statement, so the write
    // addresses are formed here rather than kept live from the loads
    int lw = lane;
    asm volatile("" : "+v"(lw));
#pragma unroll
    for (int cs = 0; cs < NCW; cs++) {
        const int jc = wave + NW * cs;
        if (jc < nbp) {
#pragma unroll
            for (int q = 0; q < QM; q++) {
                const int r = lw + 64 * q;
                if (r < mp) {
                    A[r + (long)lda * jc] = x[q][cs];
                    V[r + (long)ldv * jc] = selv(r > jc, x[q][cs], from_re<T>(r == jc ? 1.0 : 0.0));
                }
            }
        }
    }
    T* Tw = Tw_ + b * sT;
    for (int e = tid; e < NB * NB; e += NT) Tw[e] = Ts[e];
    T* tau = tau_ + b * stau + k0;
    for (int jc = tid; jc < nbp; jc += NT) tau[jc] = Ts[jc + NB * jc];
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
            T xk = bcast(x[q], k);
            xk = xk * rkk;
            if (lane == k) x[q] = xk;
            if (lane < k) x[q] -= uik * xk;
        }
    }
#pragma unroll
    for (int q = 0; q < LU_NC; q++)
        if (lane < nk && c0 + q < ncols) X[lane + (long)ldx * (c0 + q)] = x[q];
}

// The same solve for a full diagonal block of order 64 of a real U, one thread
// per right-hand side: the thread's column of X_k in registers, U_kk staged
// once per block in LDS with coalesced loads and read back with uniform
// addresses (broadcasts), the reciprocals of its diagonal computed once, back
// substitution as straight-line FMAs.  The lane-per-row form above spends a
// broadcast, a select and a masked update per column and step.
constexpr int TRSM_COLS = 256;  // right-hand sides per block
__global__ __launch_bounds__(TRSM_COLS) void k_trsm_diag64_cols(int ncols, const double* U_,
                                                                long sU, int ldu, double* X_,
                                                                long sX, int ldx)
{
    __shared__ double Us[64 * 64];  // column-major, leading dimension 64
    __shared__ double dinv[64];
    const int b = blockIdx.y;
    const double* U = U_ + b * sU;
    double* X = X_ + b * sX;
    const int t = threadIdx.x;
    const int c = blockIdx.x * TRSM_COLS + t;
    const long cl = c < ncols ? c : ncols - 1;  // in-bounds column for the load
    double x[64];
#pragma unroll
    for (int i = 0; i < 64; i++) x[i] = X[i + ldx * cl];
    for (int q = t; q < 64 * 64; q += TRSM_COLS) Us[q] = U[(q & 63) + (long)ldu * (q >> 6)];
    __syncthreads();
    if (t < 64) dinv[t] = 1.0 / Us[t + 64 * t];
    __syncthreads();
#pragma unroll
    for (int k = 63; k >= 0; k--) {
        const double xk = x[k] * dinv[k];
        x[k] = xk;
#pragma unroll
        for (int i = 0; i < k; i++) x[i] -= Us[i + 64 * k] * xk;
    }
    if (c < ncols) {
#pragma unroll
        for (int i = 0; i < 64; i++) X[i + ldx * (long)c] = x[i];
    }
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
