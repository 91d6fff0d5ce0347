// dq_factor.hpp -- per-matrix factorization kernels for the stabilized
// Green's function (one workgroup per (chain, spin) matrix, batched over the
// grid).  Restates the LAPACK steps src/greens.c:69-208 relies on:
//   geqp3 / geqrf (Householder QR, optional column pivoting) + explicit Q,
//   the pre-pivoting of Algorithm 3 (Tomas et al., IPDPS 2012),
//   Eq. 2.12 of Bai et al. 2011 solved by LU with partial pivoting,
//   and the phase of det(I + B_L..B_1) (src/greens.c:194-207).
// Matrices live in HBM/L2; the Householder vector of the current step is
// staged in LDS; dot products are wave64 reductions; each wave owns a strided
// set of columns, so a reflector is applied with no workgroup barrier between
// columns.
#pragma once
#include "dq_common.hpp"

#define QR_THREADS 1024
#define QR_WAVES (QR_THREADS / 64)
#define QR_MAXM 1024

// Householder vector for x = [alpha; x1], LAPACK xlarfg convention:
// H = I - tau v v^H, v[0] = 1, H^H [alpha; x1] = [beta; 0], beta real.
__device__ inline void larfg(double alpha, double xnorm2, double* beta, double* tau, double* scal)
{
    if (xnorm2 == 0.0) {
        *tau = 0.0; *beta = alpha; *scal = 0.0;
        return;
    }
    const double b = -copysign(sqrt(alpha * alpha + xnorm2), alpha);
    *tau = (b - alpha) / b;
    *scal = 1.0 / (alpha - b);
    *beta = b;
}
__device__ inline void larfg(zd alpha, double xnorm2, zd* beta, zd* tau, zd* scal)
{
    if (xnorm2 == 0.0 && alpha.y == 0.0) {
        *tau = mkz(0.0, 0.0); *beta = alpha; *scal = mkz(0.0, 0.0);
        return;
    }
    const double b = -copysign(sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xnorm2), alpha.x);
    *tau = mkz((b - alpha.x) / b, -alpha.y / b);
    *scal = zdiv(mkz(1.0, 0.0), mkz(alpha.x - b, alpha.y));
    *beta = mkz(b, 0.0);
}

// QR of the m x n matrix A[b] (lda) in place: R in the upper triangle, the
// Householder vectors below it, tau[b][0..k).  PIVOT: column pivoting on the
// largest remaining column norm (geqp3, norms recomputed exactly after every
// step); jpvt[b][j] = 1-based original column of column j.  Q[b] (ldq, m x m),
// when non-null, receives Q = H_0 ... H_{k-1}.
// Row i lives in lane i%64, slot i/64 (RPL slots, m <= 64*RPL); each wave
// updates CB of its columns at once so their loads are in flight together.
template <typename T, bool PIVOT, int RPL>
__global__ __launch_bounds__(QR_THREADS) void k_qr(int m, int n, T* A_, int lda, long sA, T* tau_,
                                                    long stau, int* jpvt_, long spiv, T* Q_, int ldq,
                                                    long sQ)
{
    constexpr int CB = (is_cplx<T>::value ? 8 : 16) / RPL > 0 ? (is_cplx<T>::value ? 8 : 16) / RPL : 1;
    __shared__ T v[QR_MAXM];
    __shared__ double nrm[QR_MAXM];
    __shared__ int piv[QR_MAXM];
    __shared__ T s_tau;
    __shared__ int s_p;
    const int b = blockIdx.x;
    T* A = A_ + b * sA;
    T* tau = tau_ + b * stau;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int k = m < n ? m : n;

    if (PIVOT) {
        for (int c = wave; c < n; c += QR_WAVES) {
            double s = 0.0;
#pragma unroll
            for (int r = 0; r < RPL; r++) {
                const int i = lane + 64 * r;
                if (i < m) s += abs2(A[i + (long)lda * c]);
            }
            s = wave_sum_r(s, lane);
            if (lane == 0) { nrm[c] = s; piv[c] = c; }
        }
        __syncthreads();
    }
    for (int j = 0; j < k; j++) {
        if (PIVOT) {
            if (wave == 0) {
                double best = -1.0;
                int bi = j;
                for (int c = j + lane; c < n; c += 64)
                    if (nrm[c] > best) { best = nrm[c]; bi = c; }
                for (int o = 32; o > 0; o >>= 1) {
                    const double ob = __shfl_xor(best, o, 64);
                    const int oi = __shfl_xor(bi, o, 64);
                    if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
                }
                if (lane == 0) s_p = bi;
            }
            __syncthreads();
            const int p = s_p;
            if (p != j) {
                for (int i = tid; i < m; i += QR_THREADS) {
                    const T t = A[i + (long)lda * j];
                    A[i + (long)lda * j] = A[i + (long)lda * p];
                    A[i + (long)lda * p] = t;
                }
                if (tid == 0) {
                    const double tn = nrm[j]; nrm[j] = nrm[p]; nrm[p] = tn;
                    const int tp = piv[j]; piv[j] = piv[p]; piv[p] = tp;
                }
            }
            __syncthreads();
        }
        // reflector for column j (wave 0)
        if (wave == 0) {
            T xv[RPL];
            double s = 0.0;
            T alpha = from_re<T>(0.0);
#pragma unroll
            for (int r = 0; r < RPL; r++) {
                const int i = lane + 64 * r;
                xv[r] = (i >= j && i < m) ? A[i + (long)lda * j] : from_re<T>(0.0);
                if (i > j && i < m) s += abs2(xv[r]);
                if (i == j) alpha = xv[r];
            }
            s = wave_sum_r(s, lane);
            alpha = shfl(alpha, j & 63);
            T beta, t, scal;
            larfg(alpha, s, &beta, &t, &scal);
#pragma unroll
            for (int r = 0; r < RPL; r++) {
                const int i = lane + 64 * r;
                if (i > j && i < m) {
                    const T x = xv[r] * scal;
                    A[i + (long)lda * j] = x;
                    v[i] = x;
                }
            }
            if (lane == 0) {
                A[j + (long)lda * j] = beta;
                v[j] = from_re<T>(1.0);
                tau[j] = t;
                s_tau = t;
            }
        }
        __syncthreads();
        const T ct = cj(s_tau);
        for (int c0 = j + 1 + wave; c0 < n; c0 += QR_WAVES * CB) {
            T x[CB][RPL];
            T w[CB];
#pragma unroll
            for (int t = 0; t < CB; t++) {
                const int c = c0 + QR_WAVES * t;
                w[t] = from_re<T>(0.0);
#pragma unroll
                for (int r = 0; r < RPL; r++) {
                    const int i = lane + 64 * r;
                    x[t][r] = (c < n && i >= j && i < m) ? A[i + (long)lda * c] : from_re<T>(0.0);
                }
            }
            constexpr int NP = val_parts<T>::n;
            double wr[CB * NP];
#pragma unroll
            for (int t = 0; t < CB; t++) {
#pragma unroll
                for (int r = 0; r < RPL; r++) {
                    const int i = lane + 64 * r;
                    if (i >= j && i < m) w[t] += cj(v[i]) * x[t][r];
                }
                put_part(&wr[t * NP], w[t]);
            }
            wave_sum_many<CB * NP>(wr, lane);  // one in-register reduction for the batch
#pragma unroll
            for (int t = 0; t < CB; t++) get_part(&wr[t * NP], w[t]);
#pragma unroll
            for (int t = 0; t < CB; t++) {
                const int c = c0 + QR_WAVES * t;
                if (c >= n) continue;
                const T f = ct * w[t];
                double sn = 0.0;
#pragma unroll
                for (int r = 0; r < RPL; r++) {
                    const int i = lane + 64 * r;
                    if (i >= j && i < m) {
                        const T y = x[t][r] - v[i] * f;
                        A[i + (long)lda * c] = y;
                        if (PIVOT && i > j) sn += abs2(y);
                    }
                }
                if (PIVOT) {
                    sn = wave_sum_r(sn, lane);
                    if (lane == 0) nrm[c] = sn;
                }
            }
        }
        __syncthreads();
    }
    if (PIVOT) {
        int* jp = jpvt_ + b * spiv;
        for (int c = tid; c < n; c += QR_THREADS) jp[c] = piv[c] + 1;
    }
    if (Q_) {
        T* Q = Q_ + b * sQ;
        for (long e = tid; e < (long)m * m; e += QR_THREADS) {
            const int i = e % m, c = e / m;
            Q[i + (long)ldq * c] = from_re<T>(i == c ? 1.0 : 0.0);
        }
        __syncthreads();
        for (int j = k - 1; j >= 0; j--) {
            for (int i = j + 1 + tid; i < m; i += QR_THREADS) v[i] = A[i + (long)lda * j];
            if (tid == 0) v[j] = from_re<T>(1.0);
            __syncthreads();
            const T t = tau[j];
            if (!is_zero(t)) {
                for (int c0 = j + wave; c0 < m; c0 += QR_WAVES * CB) {
                    T x[CB][RPL];
                    T w[CB];
#pragma unroll
                    for (int u = 0; u < CB; u++) {
                        const int c = c0 + QR_WAVES * u;
                        w[u] = from_re<T>(0.0);
#pragma unroll
                        for (int r = 0; r < RPL; r++) {
                            const int i = lane + 64 * r;
                            x[u][r] = (c < m && i >= j && i < m) ? Q[i + (long)ldq * c]
                                                                  : from_re<T>(0.0);
                        }
                    }
#pragma unroll
                    for (int u = 0; u < CB; u++) {
#pragma unroll
                        for (int r = 0; r < RPL; r++) {
                            const int i = lane + 64 * r;
                            if (i >= j && i < m) w[u] += cj(v[i]) * x[u][r];
                        }
                        w[u] = wave_sum_r(w[u], lane);
                    }
#pragma unroll
                    for (int u = 0; u < CB; u++) {
                        const int c = c0 + QR_WAVES * u;
                        if (c >= m) continue;
                        const T f = t * w[u];
#pragma unroll
                        for (int r = 0; r < RPL; r++) {
                            const int i = lane + 64 * r;
                            if (i >= j && i < m) Q[i + (long)ldq * c] = x[u][r] - v[i] * f;
                        }
                    }
                }
            }
            __syncthreads();
        }
    }
}

// Explicit unit-lower V[b] (ld n, n x n) from reflectors stored below the
// diagonal of A[b] (ld n), for the compact-WY formation of Q.
template <typename T>
__global__ void k_extract_v(int n, const T* A_, long sA, T* V_, long sV)
{
    const int b = blockIdx.y, c = blockIdx.x;
    const T* A = A_ + b * sA;
    T* V = V_ + b * sV;
    for (int i = threadIdx.x; i < n; i += blockDim.x)
        V[i + (long)n * c] = (i > c) ? A[i + (long)n * c] : from_re<T>(i == c ? 1.0 : 0.0);
}

// d_i = R_ii of an upper-triangular R (1 if 0; R's diagonal is real, the
// Householder betas), read once per matrix: the kernels below take d from
// here instead of each gathering the whole diagonal again.
template <typename T>
__global__ __launch_bounds__(256) void k_eqg_diag(int n, const T* R_, long sR, double* d_, long sd)
{
    const int b = blockIdx.y;
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    double di = re(R_[b * sR + i + (long)n * i]);
    d_[b * sd + i] = (di == 0.0) ? 1.0 : di;
}

// After the pivoted QR of the first product (src/greens.c:91-101), with d
// from k_eqg_diag: T[:, jpvt_j - 1] = diag(1/d) R[:, j] (upper part).
// Grid (row blocks, column groups of EQG_COLS, batch).
#define EQG_COLS 16
template <typename T>
__global__ void k_eqg_init_T(int n, const T* R_, long sR, const int* jpvt_, long spiv,
                             const double* d_, long sd, T* Tm_, long sT)
{
    const int b = blockIdx.z;
    const T* R = R_ + b * sR;
    const int* jp = jpvt_ + b * spiv;
    const double* d = d_ + b * sd;
    T* Tm = Tm_ + b * sT;
    // thread: row i of EQG_COLS columns j (1 / d_i formed once per row)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double vi = 1.0 / d[i];
    const int j0 = blockIdx.y * EQG_COLS, j1 = min(j0 + EQG_COLS, n);
    for (int j = j0; j < j1; j++)
        Tm[i + (long)n * (jp[j] - 1)] = (i <= j) ? vi * R[i + (long)n * j] : from_re<T>(0.0);
}

// Pre-pivoting (src/greens.c:116-132): squared column norms of Y, stable
// descending order (the insertion sort's order: ties keep index order),
// Z[:, k] = Y[:, perm_k].  One workgroup per matrix.
template <typename T>
__global__ __launch_bounds__(1024) void k_prepivot(int n, const T* Y_, long sY, T* Z_, long sZ,
                                                   int* perm_, long sp)
{
    __shared__ double nrm[QR_MAXM];
    __shared__ int perm[QR_MAXM];
    const int b = blockIdx.x;
    const T* Y = Y_ + b * sY;
    T* Z = Z_ + b * sZ;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int nw = blockDim.x >> 6;
    for (int c = wave; c < n; c += nw) {
        double s = 0.0;
        for (int i = lane; i < n; i += 64) s += abs2(Y[i + (long)n * c]);
        s = wave_sum(s);
        if (lane == 0) nrm[c] = s;
    }
    __syncthreads();
    for (int j = tid; j < n; j += blockDim.x) {
        const double dj = nrm[j];
        int r = 0;
        for (int i = 0; i < n; i++) {
            const double di = nrm[i];
            r += (di > dj) || (di == dj && i < j);
        }
        perm[r] = j;
    }
    __syncthreads();
    int* pg = perm_ + b * sp;
    for (int j = tid; j < n; j += blockDim.x) pg[j] = perm[j];
    // column by column: no 64-bit division per element
    for (int c = wave; c < n; c += nw) {
        const T* y = Y + (long)n * perm[c];
        for (int i = lane; i < n; i += 64) Z[i + (long)n * c] = y[i];
    }
}

// After the QR of a pre-pivoted product (src/greens.c:137-154), with d from
// k_eqg_diag: Rs = diag(1/d) * triu(R); Tp[i,:] = T[perm_i, :].
template <typename T>
__global__ void k_eqg_post_qr(int n, const T* R_, long sR, const int* perm_, long sp, const T* Tm_,
                              long sT, const double* d_, long sd, T* Rs_, T* Tp_, long sW)
{
    const int b = blockIdx.z;
    const T* R = R_ + b * sR;
    const int* perm = perm_ + b * sp;
    const T* Tm = Tm_ + b * sT;
    const double* d = d_ + b * sd;
    T* Rs = Rs_ + b * sW;
    T* Tp = Tp_ + b * sW;
    // thread: row i of EQG_COLS columns c (1 / d_i and perm_i once per row)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double vi = 1.0 / d[i];
    const int pi = perm[i];
    const int c0 = blockIdx.y * EQG_COLS, c1 = min(c0 + EQG_COLS, n);
    for (int c = c0; c < c1; c++) {
        Rs[i + (long)n * c] = (i <= c) ? vi * R[i + (long)n * c] : from_re<T>(0.0);
        Tp[i + (long)n * c] = Tm[pi + (long)n * c];
    }
}

// Eq. 2.12 set-up (src/greens.c:158-175): split d into Db/Ds, then
//   RHS = diag(v) Q^H,  M = diag(d) T + RHS.
// 32 x 32 tiles (blockDim 32 x 8): Q^H is read along Q's columns and written
// along RHS's columns through an LDS tile.
template <typename T>
__global__ __launch_bounds__(256) void k_eqg_final_prep(int n, const T* Q_, long sQ,
                                                        const double* d_, long sd, const T* Tm_,
                                                        long sT, T* M_, T* RHS_, long sW, double* v_,
                                                        long sv)
{
    __shared__ T tile[32][33];  // tile[k][x] = Q[c0 + x, i0 + k]
    const int b = blockIdx.z;
    const int i0 = 32 * blockIdx.x, c0 = 32 * blockIdx.y;
    const int tx = threadIdx.x, ty = threadIdx.y;
    const T* Q = Q_ + b * sQ;
    for (int k = ty; k < 32; k += 8) {
        const int i = i0 + k, c = c0 + tx;
        if (i < n && c < n) tile[k][tx] = Q[c + (long)n * i];
    }
    __syncthreads();
    const double* d = d_ + b * sd;
    const T* Tm = Tm_ + b * sT;
    T* M = M_ + b * sW;
    T* RHS = RHS_ + b * sW;
    const int i = i0 + tx;
    if (i >= n) return;
    const double di = d[i];
    double vi, ds;
    if (fabs(di) > 1.0) { vi = 1.0 / di; ds = 1.0; }
    else { vi = 1.0; ds = di; }
    if (c0 == 0 && ty == 0) v_[b * sv + i] = vi;
    for (int k = ty; k < 32; k += 8) {
        const int c = c0 + k;
        if (c < n) {
            const T r = vi * cj(tile[tx][k]);
            RHS[i + (long)n * c] = r;
            M[i + (long)n * c] = ds * Tm[i + (long)n * c] + r;
        }
    }
}

// LU with partial pivoting of M (getrf, src/greens.c:177), the row
// interchanges applied to RHS as they happen (getrs's laswp).  ipiv[b][j] is
// the 0-based row swapped with row j.
template <typename T>
__global__ __launch_bounds__(1024) void k_getrf(int n, T* M_, T* RHS_, long sW, int* ipiv_, long sp)
{
    __shared__ T colj[QR_MAXM];
    __shared__ int s_p;
    const int b = blockIdx.x;
    T* M = M_ + b * sW;
    T* X = RHS_ + b * sW;
    int* ipiv = ipiv_ + b * sp;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int nw = blockDim.x >> 6;
    for (int j = 0; j < n; j++) {
        if (wave == 0) {
            double best = -1.0;
            int bi = j;
            for (int i = j + lane; i < n; i += 64) {
                const double a = abs1(M[i + (long)n * j]);
                if (a > best) { best = a; bi = i; }
            }
            for (int o = 32; o > 0; o >>= 1) {
                const double ob = __shfl_xor(best, o, 64);
                const int oi = __shfl_xor(bi, o, 64);
                if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
            }
            if (lane == 0) { s_p = bi; ipiv[j] = bi; }
        }
        __syncthreads();
        const int p = s_p;
        if (p != j) {
            for (int c = tid; c < n; c += blockDim.x) {
                T t = M[j + (long)n * c]; M[j + (long)n * c] = M[p + (long)n * c]; M[p + (long)n * c] = t;
                t = X[j + (long)n * c]; X[j + (long)n * c] = X[p + (long)n * c]; X[p + (long)n * c] = t;
            }
        }
        __syncthreads();
        const T rp = rdiv(from_re<T>(1.0), M[j + (long)n * j]);
        for (int i = j + 1 + tid; i < n; i += blockDim.x) {
            const T l = M[i + (long)n * j] * rp;
            M[i + (long)n * j] = l;
            colj[i] = l;
        }
        __syncthreads();
        for (int c = j + 1 + wave; c < n; c += nw) {
            T* col = M + (long)n * c;
            const T mjc = col[j];
            for (int i = j + 1 + lane; i < n; i += 64) col[i] -= colj[i] * mjc;
        }
        __syncthreads();
    }
}

// Triangular solves with the LU factors (getrs, src/greens.c:178): X <- U^-1 L^-1 X,
// LU_NC right-hand sides per wave held in registers (RPL rows per lane).
// Block 0 of each matrix also computes the phase of det(I + B..B) from the LU
// diagonal, the pivots, v = Db^-1 and the Householder reflectors (V below the
// diagonal of Vq, tau) of the last QR (src/greens.c:194-207), writing 1/phase.
#define LU_NC 4
template <typename T, int RPL>
__global__ __launch_bounds__(256) void k_getrs_phase(int n, const T* M_, T* RHS_, long sW,
                                                     const int* ipiv_, long sp, const double* v_, long sv,
                                                     const T* Vq_, long sVq, const T* tau_,
                                                     long stau, T* phase_out)
{
    const int b = blockIdx.y;
    const T* M = M_ + b * sW;
    T* X = RHS_ + b * sW;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int gw = blockIdx.x * (blockDim.x >> 6) + wave;  // global wave index
    const int c0 = gw * LU_NC;
    if (c0 < n) {
        T x[LU_NC][RPL];
#pragma unroll
        for (int q = 0; q < LU_NC; q++)
#pragma unroll
            for (int r = 0; r < RPL; r++) {
                const int i = lane + 64 * r;
                x[q][r] = (i < n && c0 + q < n) ? X[i + (long)n * (c0 + q)] : from_re<T>(0.0);
            }
        for (int kk = 0; kk < n; kk++) {
            const int owner = kk & 63, rk = kk >> 6;
            T xk[LU_NC];
#pragma unroll
            for (int q = 0; q < LU_NC; q++) {
                T val = from_re<T>(0.0);
#pragma unroll
                for (int r = 0; r < RPL; r++)
                    if (r == rk) val = x[q][r];
                xk[q] = bcast(val, owner);
            }
#pragma unroll
            for (int r = 0; r < RPL; r++) {
                const int i = lane + 64 * r;
                if (i > kk && i < n) {
                    const T l = M[i + (long)n * kk];
#pragma unroll
                    for (int q = 0; q < LU_NC; q++) x[q][r] -= l * xk[q];
                }
            }
        }
        for (int kk = n - 1; kk >= 0; kk--) {
            const int owner = kk & 63, rk = kk >> 6;
            const T ukk = M[kk + (long)n * kk];
            T xk[LU_NC];
#pragma unroll
            for (int q = 0; q < LU_NC; q++) {
                T val = from_re<T>(0.0);
#pragma unroll
                for (int r = 0; r < RPL; r++)
                    if (r == rk) val = x[q][r];
                val = rdiv(val, ukk);
                xk[q] = bcast(val, owner);
#pragma unroll
                for (int r = 0; r < RPL; r++)
                    if (r == rk && lane == owner) x[q][r] = xk[q];
            }
#pragma unroll
            for (int r = 0; r < RPL; r++) {
                const int i = lane + 64 * r;
                if (i < kk) {
                    const T u = M[i + (long)n * kk];
#pragma unroll
                    for (int q = 0; q < LU_NC; q++) x[q][r] -= u * xk[q];
                }
            }
        }
#pragma unroll
        for (int q = 0; q < LU_NC; q++)
#pragma unroll
            for (int r = 0; r < RPL; r++) {
                const int i = lane + 64 * r;
                if (i < n && c0 + q < n) X[i + (long)n * (c0 + q)] = x[q][r];
            }
    }
    if (blockIdx.x == 0 && wave == 0 && phase_out) {
        const int* ipiv = ipiv_ + b * sp;
        const double* v = v_ + b * sv;
        const T* Vq = Vq_ + b * sVq;
        const T* tau = tau_ + b * stau;
        T ph = from_re<T>(1.0);
        for (int i = 0; i < n; i++) {
            double s = 0.0;
            for (int j = i + 1 + lane; j < n; j += 64) s += abs2(Vq[j + (long)n * i]);
            s = wave_sum(s);
            const T c = rdiv(from_re<T>(v[i]), M[i + (long)n * i]);
            ph = ph * (c * (1.0 / absv(c)));
            const T ref = from_re<T>(1.0) - tau[i] * (1.0 + s);
            ph = ph * rdiv(from_re<T>(absv(ref)), ref);
            if (ipiv[i] != i) ph = ph * (-1.0);
        }
        if (lane == 0) phase_out[b] = rdiv(from_re<T>(1.0), ph);
    }
}
