// dq_adjoint.hpp -- adjoint of log|det M_s| through the stabilized B-chain.
//
// No reference counterpart: the reference computes only the phase of
// det M (its determinant code is commented out, src/greens.c:180-191).
//
// With M = I + B_{L-1} ... B_0, B_l = E e^{V_l} (E = e^{-dt K}), and G_l the
// Green's function that has B_l at the right end of the product (the G on
// which the reference's updates of slice l act),
//   d log det M = sum_l tr[(I - G_l) B_l^-1 dB_l]
// (push-through identity).  Hence, for fixed HS fields:
//   dE:      d log det M = tr[W E^-1 dE],  W = sum_l e^{V_l} (I - G_l) e^{-V_l}
//   d mu:    E = e^{dt mu} E_0      =>  d/d mu     = dt sum_l tr(I - G_l)
//   d lambda V_l = sigma lambda s_l =>  d/d lambda = sum_l sum_i sigma s_li (1 - G_l,ii)
// The engine walks the slices as a sweep does (wraps, and a fresh
// stabilized G every n_matmul slices -- the checkpoints of the UDT chain),
// accumulating W, sum tr G_l and the lambda term; log|det M| comes from the
// factorisation of the first calc_eq_g:
//   I + Q D T = Q diag(v)^-1 M',  M' = Q_M R_M  =>  log|det M| = sum log|R_ii| - sum log|v_i|.
//
// The same walk gives derivatives of a G-based observable O = tr(X G), G the
// equal-time Green's function at slice 0 and X fixed: dG = -G dM G and
// dM = sum_l B_{L-1}..B_{l+1} dB_l B_{l-1}..B_0 give
//   dO = sum_l tr(Z_l B_l^-1 dB_l),   Z_l = G(t_l,0) X G^(0,t_l),
// with the unequal-time Green's functions G(t_l,0) = B_{l-1}..B_0 G and
// G^(0,t_l) = -(I - G) (B_{l-1}..B_0)^-1 (at l = 0: G - I, the limit from
// above, where calc_ue_g stores G itself).  B_l^-1 dB_l is e^{-V_l} E^-1 dE
// e^{V_l} for the hopping and dV_l for the HS coupling, so O has the
// log-determinant's chain rule with (I - G_l) replaced by Z_l:
//   W_X = sum_l e^{V_l} Z_l e^{-V_l},  dO/dmu = dt sum_l tr Z_l,
//   dO/dlambda = sum_l sum_i sigma s_li (Z_l)_ii.
#pragma once
#include "dq_common.hpp"

// W[b] += e^{V_l} (a I + c G) e^{-V_l} (elementwise with the slice's scales);
// a = 1, c = -1 for the log-determinant, a = 0, c = 1 for Z_l
template <typename T>
__global__ void k_adj_accum_W(int N, const T* G, long sG, const double* esc, const double* isc,
                              long ss, int l, T* W, long sW, int B, double a, double c)
{
    const long NN = (long)N * N;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < NN * B;
         e += (long)gridDim.x * blockDim.x) {
        const int b = (int)(e / NN);
        const long r = e % NN;
        const int i = (int)(r % N), j = (int)(r / N);
        const T g = G[b * sG + r];
        const T ig = from_re<T>(i == j ? a : 0.0) + c * g;
        W[b * sW + r] += (esc[b * ss + (long)N * l + i] * isc[b * ss + (long)N * l + j]) * ig;
    }
}

// trG[b] += tr G;  sG[b] += sigma_b sum_i s_li (a + c G_ii), s = 2 hs - 1,
// sigma = +1 (up, b even) / -1 (down)
template <typename T>
__global__ __launch_bounds__(64) void k_adj_accum_diag(int N, int L, const T* G, long sG,
                                                       const int* hs, int l, T* trG, T* sGl,
                                                       double a, double c)
{
    const int b = blockIdx.x, lane = threadIdx.x;
    const int cc = b >> 1;
    const double sigma = (b & 1) ? -1.0 : 1.0;
    const int* h = hs + (long)cc * L * N + (long)N * l;
    T t = from_re<T>(0.0), s = t;
    for (int i = lane; i < N; i += 64) {
        const T g = G[b * sG + i + (long)N * i];
        t += g;
        s += (sigma * (2 * h[i] - 1)) * (a + c * g);
    }
    t = wave_sum(t);
    s = wave_sum(s);
    if (lane == 0) {
        trG[b] += t;
        sGl[b] += s;
    }
}

// log|det M| = sum log|R_ii| - sum log|v_i| of the last calc_eq_g
template <typename T>
__global__ __launch_bounds__(64) void k_eqg_logdet(int n, const double* v_, long sv, const T* R_,
                                                   long sR, double* out)
{
    const int b = blockIdx.x, lane = threadIdx.x;
    const double* v = v_ + b * sv;
    const T* R = R_ + b * sR;
    double s = 0.0;
    for (int i = lane; i < n; i += 64) s += log(absv(R[i + (long)n * i])) - log(fabs(v[i]));
    s = wave_sum(s);
    if (lane == 0) out[b] = s;
}

// out[b] = tr(X_b G_b) = sum_ij X_ij G_ji (column-major N x N)
template <typename T>
__global__ __launch_bounds__(64) void k_adj_trace_prod(int N, const T* X, long sX, const T* G,
                                                       long sG, T* out)
{
    const int b = blockIdx.x, lane = threadIdx.x;
    const long NN = (long)N * N;
    T s = from_re<T>(0.0);
    for (long r = lane; r < NN; r += 64) {
        const int i = (int)(r % N), j = (int)(r / N);
        s += X[b * sX + r] * G[b * sG + j + (long)N * i];
    }
    s = wave_sum(s);
    if (lane == 0) out[b] = s;
}

// X[b] += v I
template <typename T>
__global__ void k_adj_add_diag(int N, T* X, long sX, double v, int B)
{
    const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (e >= (long)N * B) return;
    const int b = (int)(e / N), i = (int)(e % N);
    X[b * sX + i + (long)N * i] += from_re<T>(v);
}
