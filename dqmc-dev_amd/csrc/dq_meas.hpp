// dq_meas.hpp -- equal- and unequal-time measurement contractions
// (src/meas.c:66-120 measure_eqlt, src/meas.c:728-801 measure_uneqlt).
//
// Deterministic scatter-free reductions: the host builds, per chain, the
// inverse of map_ij / map_i (bin -> list of linear indices i + N*j in
// increasing order).  One thread owns one accumulator bin and adds the terms
// of its list in that order, which is the order in which the reference's
// (j outer, i inner) loops add them, so no atomics are needed and results do
// not depend on scheduling.  Accumulators stay resident in HBM for the whole
// run and are copied to the host only when the caller asks for them.
#pragma once
#include "dq_common.hpp"

// extra bin maps of the optional groups (CSR per chain, like map_bb / map_bs)
enum { PM_bb2 = 0, PM_b2b, PM_b2b2, PM_plaq, PM_b, PM_b2, PM_COUNT };

template <typename T>
struct MeasArgs {
    int N, L, num_i, num_ij, num_bb, num_b;
    int energy, bond;
    const T* G; long sG;              // half-wrapped equal-time G[b] (eq) / G(t,t) etc (ue)
    const T* G0t; const T* Gtt; const T* Gt0; long sU; // ue arrays [b][L*N*N]
    const T* G0tT;                    // G(0,t)^T per slice, same layout (bond-pair kernels)
    const T* btab;                    // [b][t][f, b, n][num_b]: bond hop terms, densities
    const T* bph;                     // [c][pf_u, pb_u, pf_d, pb_d][num_b]: bond phases
    const T* b2tab;                   // [b][t][f, b, n][num_b2]: the same for two-hop bonds
    const T* phase;                   // [c]
    const int* flag;                  // [c]: measure this chain now
    const int* rp_i; const int* ix_i; long srp_i, six_i;     // CSR of map_i  [c]
    const int* rp_ij; const int* ix_ij; long srp_ij, six_ij; // CSR of map_ij [c]
    const int* degen;                 // [c][4]: degen_i, degen_ij, degen_bb, degen_bs
    const T* peierlsu; const T* peierlsd; long sP; // [c][N*N] (complex only)
    T* acc; long sacc;                // accumulator block of chain c
    const long* off;                  // offsets of each accumulator in the block (-1: off)
    int* n_sample; T* sign;           // [c]
    const int* bonds; long sbonds;    // [c][2*num_b]
    const int* rp_bb; const int* ix_bb; long srp_bb, six_bb;  // CSR of map_bb
    const int* rp_bs; const int* ix_bs; long srp_bs, six_bs;  // CSR of map_bs
    int num_bs;
    // optional groups (dq_meas2.hpp)
    int thermal, twobond, pbbonly, chiral, localjq;
    int Nx, Ny, num_b2, num_bb2, num_b2b, num_b2b2, num_plaq, num_plaq_acc, num_b_acc, num_b2_acc;
    const int* rp2[PM_COUNT]; const int* ix2[PM_COUNT]; long srp2[PM_COUNT], six2[PM_COUNT];
    const int* degen2;                 // [c][PM_COUNT]
    const int* bond2s; long sbond2s;   // [c][2 num_b2]
    const T* pp; long spp;             // [c][4][num_b2]: pp_u, ppr_u, pp_d, ppr_d
    const int* plaqs; long splaqs;     // [c][3 num_plaq]
    T* gs; long sgs;                   // gen_suscept partial sums [c][L][4][num_ij]
    const int* invij; long sinv;       // [c][j][num_ij]: the row i that column j adds to bin r
                                       // (when every column feeds every bin once)
};

// accumulator indices (match enum dqmc_eq_acc / dqmc_ue_acc)
enum { EQ_density = 0, EQ_density_u, EQ_density_d, EQ_double_occ, EQ_g00, EQ_g00_u, EQ_g00_d,
       EQ_nn, EQ_xx, EQ_zz, EQ_pair_sw, EQ_kk = 16, EQ_kv = 17, EQ_kn = 18, EQ_vv = 19, EQ_vn = 20 };
enum { UE_gt0 = 0, UE_gt0_u, UE_gt0_d, UE_nn, UE_xx, UE_zz, UE_pair_sw, UE_pair_bb = 11, UE_jj,
       UE_jsjs, UE_kk, UE_ksks, UE_vv = 30, UE_vn = 31 };

template <typename T>
__device__ inline T peierls_mul(const T* P, long idx, T g);
template <>
__device__ inline double peierls_mul<double>(const double*, long, double g) { return g; }
template <>
__device__ inline zd peierls_mul<zd>(const zd* P, long idx, zd g) { return g * P[idx]; }

template <typename T>
__global__ __launch_bounds__(256) void k_meas_eq(MeasArgs<T> a)
{
    const int c = blockIdx.y;
    if (!a.flag[c]) return;
    const int N = a.N;
    const T* gu = a.G + (2 * c) * a.sG;
    const T* gd = a.G + (2 * c + 1) * a.sG;
    const T phase = a.phase[c];
    T* acc = a.acc + c * a.sacc;
    const long* off = a.off;
    const int degen_i = a.degen[4 * c], degen_ij = a.degen[4 * c + 1];
    const T* pu = a.peierlsu + c * a.sP;
    const T* pd = a.peierlsd + c * a.sP;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t == 0) {
        a.n_sample[c] += 1;
        a.sign[c] += phase;
    }
    // one-site bins
    if (t < a.num_i) {
        const int* rp = a.rp_i + c * a.srp_i;
        const int* ix = a.ix_i + c * a.six_i;
        const T pre = phase * (1.0 / degen_i);
        T s0 = acc[off[EQ_density] + t], s1 = acc[off[EQ_density_u] + t];
        T s2 = acc[off[EQ_density_d] + t], s3 = acc[off[EQ_double_occ] + t];
        for (int q = rp[t]; q < rp[t + 1]; q++) {
            const int i = ix[q];
            const T guii = gu[i + (long)N * i], gdii = gd[i + (long)N * i];
            const T one = from_re<T>(1.0);
            s0 += pre * (from_re<T>(2.0) - guii - gdii);
            s1 += pre * (one - guii);
            s2 += pre * (one - gdii);
            s3 += pre * (one - guii) * (one - gdii);
        }
        acc[off[EQ_density] + t] = s0;
        acc[off[EQ_density_u] + t] = s1;
        acc[off[EQ_density_d] + t] = s2;
        acc[off[EQ_double_occ] + t] = s3;
    }
    if (t < a.num_ij) {
        const int* rp = a.rp_ij + c * a.srp_ij;
        const int* ix = a.ix_ij + c * a.six_ij;
        const T pre = phase * (1.0 / degen_ij);
        T g00 = acc[off[EQ_g00] + t], g00u = acc[off[EQ_g00_u] + t], g00d = acc[off[EQ_g00_d] + t];
        T nn = acc[off[EQ_nn] + t], xx = acc[off[EQ_xx] + t], zz = acc[off[EQ_zz] + t];
        T psw = acc[off[EQ_pair_sw] + t];
        T vv = from_re<T>(0.0), vn = from_re<T>(0.0);
        if (a.energy) { vv = acc[off[EQ_vv] + t]; vn = acc[off[EQ_vn] + t]; }
        const T one = from_re<T>(1.0), two = from_re<T>(2.0);
        for (int q = rp[t]; q < rp[t + 1]; q++) {
            const int lin = ix[q];
            const int i = lin % N, j = lin / N;
            const double delta = (i == j) ? 1.0 : 0.0;
            const T guii = gu[i + (long)N * i], gdii = gd[i + (long)N * i];
            const T guij = gu[i + (long)N * j], gdij = gd[i + (long)N * j];
            const T guji = gu[j + (long)N * i], gdji = gd[j + (long)N * i];
            const T gujj = gu[j + (long)N * j], gdjj = gd[j + (long)N * j];
            const T pgu = peierls_mul(pu, j + (long)N * i, guij);
            const T pgd = peierls_mul(pd, j + (long)N * i, gdij);
            g00 += 0.5 * (pre * (pgu + pgd));
            g00u += pre * pgu;
            g00d += pre * pgd;
            const T x = delta * (guii + gdii) - (guji * guij + gdji * gdij);
            nn += pre * ((two - guii - gdii) * (two - gujj - gdjj) + x);
            xx += 0.25 * (pre * (delta * (guii + gdii) - (guji * gdij + gdji * guij)));
            zz += 0.25 * (pre * ((gdii - guii) * (gdjj - gujj) + x));
            psw += pre * guij * gdij;
            if (a.energy) {
                const T nuinuj = (one - guii) * (one - gujj) + (delta - guji) * guij;
                const T ndindj = (one - gdii) * (one - gdjj) + (delta - gdji) * gdij;
                vv += pre * nuinuj * ndindj;
                vn += pre * (nuinuj * (one - gdii) + (one - guii) * ndindj);
            }
        }
        acc[off[EQ_g00] + t] = g00; acc[off[EQ_g00_u] + t] = g00u; acc[off[EQ_g00_d] + t] = g00d;
        acc[off[EQ_nn] + t] = nn; acc[off[EQ_xx] + t] = xx; acc[off[EQ_zz] + t] = zz;
        acc[off[EQ_pair_sw] + t] = psw;
        if (a.energy) { acc[off[EQ_vv] + t] = vv; acc[off[EQ_vn] + t] = vn; }
    }
}

// unequal time, default set + vv/vn (src/meas.c:749-801): one thread per (bin, t)
template <typename T>
__global__ __launch_bounds__(256) void k_meas_ue(MeasArgs<T> a)
{
    const int c = blockIdx.z;
    if (!a.flag[c]) return;
    const int N = a.N;
    const long NN = (long)N * N;
    const int tt = blockIdx.y;
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    const T phase = a.phase[c];
    T* acc = a.acc + c * a.sacc;
    const long* off = a.off;
    if (r == 0 && tt == 0) {
        a.n_sample[c] += 1;
        a.sign[c] += phase;
    }
    if (r >= a.num_ij) return;
    const int degen_ij = a.degen[4 * c + 1];
    const T* pu = a.peierlsu + c * a.sP;
    const T* pd = a.peierlsd + c * a.sP;
    const T* Gu0t = a.G0t + (2 * c) * a.sU + NN * tt;
    const T* Gutt = a.Gtt + (2 * c) * a.sU + NN * tt;
    const T* Gut0 = a.Gt0 + (2 * c) * a.sU + NN * tt;
    const T* Gd0t = a.G0t + (2 * c + 1) * a.sU + NN * tt;
    const T* Gdtt = a.Gtt + (2 * c + 1) * a.sU + NN * tt;
    const T* Gdt0 = a.Gt0 + (2 * c + 1) * a.sU + NN * tt;
    const T* Gu00 = a.Gtt + (2 * c) * a.sU;
    const T* Gd00 = a.Gtt + (2 * c + 1) * a.sU;
    const int* rp = a.rp_ij + c * a.srp_ij;
    const int* ix = a.ix_ij + c * a.six_ij;
    const T pre = phase * (1.0 / degen_ij);
    const long o = r + (long)a.num_ij * tt;
    T gt0 = acc[off[UE_gt0] + o], gt0u = acc[off[UE_gt0_u] + o], gt0d = acc[off[UE_gt0_d] + o];
    T nn = acc[off[UE_nn] + o], xx = acc[off[UE_xx] + o], zz = acc[off[UE_zz] + o];
    T psw = acc[off[UE_pair_sw] + o];
    T vv = from_re<T>(0.0), vn = from_re<T>(0.0);
    if (a.energy) { vv = acc[off[UE_vv] + o]; vn = acc[off[UE_vn] + o]; }
    const T one = from_re<T>(1.0), two = from_re<T>(2.0);
    const double delta_t = (tt == 0) ? 1.0 : 0.0;
    for (int q = rp[r]; q < rp[r + 1]; q++) {
        const int lin = ix[q];
        const int i = lin % N, j = lin / N;
        const double dtij = delta_t * ((i == j) ? 1.0 : 0.0);
        const T guii = Gutt[i + (long)N * i], guij = Gut0[i + (long)N * j];
        const T guji = Gu0t[j + (long)N * i], gujj = Gu00[j + (long)N * j];
        const T gdii = Gdtt[i + (long)N * i], gdij = Gdt0[i + (long)N * j];
        const T gdji = Gd0t[j + (long)N * i], gdjj = Gd00[j + (long)N * j];
        const T pgu = peierls_mul(pu, j + (long)N * i, guij);
        const T pgd = peierls_mul(pd, j + (long)N * i, gdij);
        gt0 += 0.5 * (pre * (pgu + pgd));
        gt0u += pre * pgu;
        gt0d += pre * pgd;
        const T x = dtij * (guii + gdii) - (guji * guij + gdji * gdij);
        nn += pre * ((two - guii - gdii) * (two - gujj - gdjj) + x);
        xx += 0.25 * (pre * (dtij * (guii + gdii) - (guji * gdij + gdji * guij)));
        zz += 0.25 * (pre * ((gdii - guii) * (gdjj - gujj) + x));
        psw += pre * guij * gdij;
        if (a.energy) {
            const T nuinuj = (one - guii) * (one - gujj) + (dtij - guji) * guij;
            const T ndindj = (one - gdii) * (one - gdjj) + (dtij - gdji) * gdij;
            vv += pre * nuinuj * ndindj;
            vn += pre * (nuinuj * (one - gdii) + (one - guii) * ndindj);
        }
    }
    acc[off[UE_gt0] + o] = gt0; acc[off[UE_gt0_u] + o] = gt0u; acc[off[UE_gt0_d] + o] = gt0d;
    acc[off[UE_nn] + o] = nn; acc[off[UE_xx] + o] = xx; acc[off[UE_zz] + o] = zz;
    acc[off[UE_pair_sw] + o] = psw;
    if (a.energy) { acc[off[UE_vv] + o] = vv; acc[off[UE_vn] + o] = vn; }
}

// The same contraction for maps in which every column j adds exactly one term
// to every bin (translation-symmetric maps, num_ij = N): one workgroup per
// (t, chain), one thread per bin.  Row strips of G(0,t) (and of the Peierls
// factors) are staged through LDS with line-coalesced reads, and the diagonals
// of G(t,t) and G(0,0) are kept in LDS; each bin still adds its terms in
// increasing j, the reference's order.
#define UES_NMAX 256
template <typename T, int CH>
__global__ __launch_bounds__(256) void k_meas_ue_strip(MeasArgs<T> a)
{
    constexpr bool Z = is_cplx<T>::value;
    __shared__ T strip[2][CH][UES_NMAX];               // G0t_s[j0 + jj, i]
    __shared__ T pstrip[Z ? 2 : 1][Z ? CH : 1][Z ? UES_NMAX : 1];  // P_s[j0 + jj + N i]
    __shared__ T dtt[2][UES_NMAX], doo[2][UES_NMAX];
    const int c = blockIdx.y;
    if (!a.flag[c]) return;
    const int tt = blockIdx.x;
    const int N = a.N, nij = a.num_ij;
    const long NN = (long)N * N;
    const int tid = threadIdx.x, nt = blockDim.x;
    const T phase = a.phase[c];
    if (tid == 0 && tt == 0) {
        a.n_sample[c] += 1;
        a.sign[c] += phase;
    }
    const T* G0t[2] = {a.G0t + (2 * c) * a.sU + NN * tt, a.G0t + (2 * c + 1) * a.sU + NN * tt};
    const T* Gt0[2] = {a.Gt0 + (2 * c) * a.sU + NN * tt, a.Gt0 + (2 * c + 1) * a.sU + NN * tt};
    const T* P[2] = {a.peierlsu + c * a.sP, a.peierlsd + c * a.sP};
    for (int i = tid; i < N; i += nt) {
#pragma unroll
        for (int s = 0; s < 2; s++) {
            dtt[s][i] = a.Gtt[(2 * c + s) * a.sU + NN * tt + i + (long)N * i];
            doo[s][i] = a.Gtt[(2 * c + s) * a.sU + i + (long)N * i];
        }
    }
    const int* inv = a.invij + c * a.sinv;
    const int r = tid;
    const T pre = phase * (1.0 / a.degen[4 * c + 1]);
    const T zero = from_re<T>(0.0), one = from_re<T>(1.0), two = from_re<T>(2.0);
    T* acc = a.acc + c * a.sacc;
    const long* off = a.off;
    const long o = r + (long)nij * tt;
    T gt0 = zero, gt0u = zero, gt0d = zero, nn = zero, xx = zero, zz = zero, psw = zero;
    T vv = zero, vn = zero;
    if (r < nij) {   // terms add straight into the running sums, as the reference's do
        gt0 = acc[off[UE_gt0] + o]; gt0u = acc[off[UE_gt0_u] + o]; gt0d = acc[off[UE_gt0_d] + o];
        nn = acc[off[UE_nn] + o]; xx = acc[off[UE_xx] + o]; zz = acc[off[UE_zz] + o];
        psw = acc[off[UE_pair_sw] + o];
        if (a.energy) { vv = acc[off[UE_vv] + o]; vn = acc[off[UE_vn] + o]; }
    }
    const double delta_t = (tt == 0) ? 1.0 : 0.0;
    // this thread's rows i of the strip's columns j and their G(t,0)[i, j],
    // fetched one strip ahead into registers so that the loads of a strip
    // are all in flight at once instead of one dependent chain per column
    int iv[2][CH];
    T cu[2][CH], cd[2][CH];
    auto fetch = [&](int j0, auto hc) {
        constexpr int h = decltype(hc)::value;
#pragma unroll
        for (int jj = 0; jj < CH; jj++) {
            const int j = j0 + jj;
            iv[h][jj] = (r < nij && j < N) ? inv[(long)j * nij + r] : 0;
        }
#pragma unroll
        for (int jj = 0; jj < CH; jj++) {
            const int j = j0 + jj;
            const bool ok = r < nij && j < N;
            cu[h][jj] = ok ? Gt0[0][iv[h][jj] + (long)N * j] : zero;
            cd[h][jj] = ok ? Gt0[1][iv[h][jj] + (long)N * j] : zero;
        }
    };
    fetch(0, std::integral_constant<int, 0>());
    // strips in pairs, so that the register set of a strip is a constant
    auto strip_step = [&](int j0, auto hc) {
        constexpr int h = decltype(hc)::value;
        const int ch = (N - j0 < CH) ? N - j0 : CH;
        __syncthreads();
        for (int e = tid; e < CH * N; e += nt) {
            const int jj = e % CH, i = e / CH;
            if (jj < ch) {
#pragma unroll
                for (int s = 0; s < 2; s++) {
                    strip[s][jj][i] = G0t[s][(j0 + jj) + (long)N * i];
                    if constexpr (Z) pstrip[s][jj][i] = P[s][(j0 + jj) + (long)N * i];
                }
            }
        }
        if (j0 + CH < N) fetch(j0 + CH, std::integral_constant<int, 1 - h>());
        __syncthreads();
        if (r < nij) {
#pragma unroll
            for (int jj = 0; jj < CH; jj++) {
                if (jj >= ch) continue;
                const int j = j0 + jj;
                const int i = iv[h][jj];
                const double dtij = delta_t * ((i == j) ? 1.0 : 0.0);
                const T guii = dtt[0][i], gdii = dtt[1][i], gujj = doo[0][j], gdjj = doo[1][j];
                const T guij = cu[h][jj], gdij = cd[h][jj];
                const T guji = strip[0][jj][i], gdji = strip[1][jj][i];
                T pgu = guij, pgd = gdij;
                if constexpr (Z) {   // g * p, the product order of peierls_mul
                    pgu = guij * pstrip[0][jj][i];
                    pgd = gdij * pstrip[1][jj][i];
                }
                gt0 += 0.5 * (pre * (pgu + pgd));
                gt0u += pre * pgu;
                gt0d += pre * pgd;
                const T x = dtij * (guii + gdii) - (guji * guij + gdji * gdij);
                nn += pre * ((two - guii - gdii) * (two - gujj - gdjj) + x);
                xx += 0.25 * (pre * (dtij * (guii + gdii) - (guji * gdij + gdji * guij)));
                zz += 0.25 * (pre * ((gdii - guii) * (gdjj - gujj) + x));
                psw += pre * guij * gdij;
                if (a.energy) {
                    const T nuinuj = (one - guii) * (one - gujj) + (dtij - guji) * guij;
                    const T ndindj = (one - gdii) * (one - gdjj) + (dtij - gdji) * gdij;
                    vv += pre * nuinuj * ndindj;
                    vn += pre * (nuinuj * (one - gdii) + (one - guii) * ndindj);
                }
            }
        }
    };
    for (int j0 = 0; j0 < N; j0 += 2 * CH) {
        strip_step(j0, std::integral_constant<int, 0>());
        if (j0 + CH < N) strip_step(j0 + CH, std::integral_constant<int, 1>());
    }
    if (r < nij) {
        acc[off[UE_gt0] + o] = gt0; acc[off[UE_gt0_u] + o] = gt0u; acc[off[UE_gt0_d] + o] = gt0d;
        acc[off[UE_nn] + o] = nn; acc[off[UE_xx] + o] = xx; acc[off[UE_zz] + o] = zz;
        acc[off[UE_pair_sw] + o] = psw;
        if (a.energy) { acc[off[UE_vv] + o] = vv; acc[off[UE_vn] + o] = vn; }
    }
}

// Equal time (src/meas.c:66-120) for maps in which every column j adds exactly
// one term to every bin (translation-symmetric maps, num_ij = N): the
// unequal-time strip contraction at t = 0 (G(0,0) = G(t,t) = G(t,0) = G(0,t),
// delta_t = 1).  One workgroup per chain, one thread per bin; rows of G are
// staged through LDS in strips of CH with line-coalesced reads, where the CSR
// form reads row j of the column-major G one cache line per element.  Each
// bin adds its terms in increasing j, as the CSR form (and the reference)
// does, so the sums are the same to the bit.
template <typename T, int CH>
__global__ __launch_bounds__(256) void k_meas_eq_strip(MeasArgs<T> a)
{
    constexpr bool Z = is_cplx<T>::value;
    __shared__ T strip[2][CH][UES_NMAX];                            // G_s[j0 + jj, i]
    __shared__ T pstrip[Z ? 2 : 1][Z ? CH : 1][Z ? UES_NMAX : 1];  // P_s[j0 + jj + N i]
    __shared__ T dg[2][UES_NMAX];
    const int c = blockIdx.x;
    if (!a.flag[c]) return;
    const int N = a.N, nij = a.num_ij;
    const int tid = threadIdx.x, nt = blockDim.x;
    const T phase = a.phase[c];
    if (tid == 0) {
        a.n_sample[c] += 1;
        a.sign[c] += phase;
    }
    const T* G[2] = {a.G + (2 * c) * a.sG, a.G + (2 * c + 1) * a.sG};
    const T* P[2] = {a.peierlsu + c * a.sP, a.peierlsd + c * a.sP};
    for (int i = tid; i < N; i += nt) {
#pragma unroll
        for (int s = 0; s < 2; s++) dg[s][i] = G[s][i + (long)N * i];
    }
    __syncthreads();
    T* acc = a.acc + c * a.sacc;
    const long* off = a.off;
    const T zero = from_re<T>(0.0), one = from_re<T>(1.0), two = from_re<T>(2.0);
    const int r = tid;
    // one-site bins (their CSR lists, diagonals from LDS)
    if (r < a.num_i) {
        const int* rp = a.rp_i + c * a.srp_i;
        const int* ix = a.ix_i + c * a.six_i;
        const T pre = phase * (1.0 / a.degen[4 * c]);
        T s0 = acc[off[EQ_density] + r], s1 = acc[off[EQ_density_u] + r];
        T s2 = acc[off[EQ_density_d] + r], s3 = acc[off[EQ_double_occ] + r];
        for (int q = rp[r]; q < rp[r + 1]; q++) {
            const int i = ix[q];
            const T guii = dg[0][i], gdii = dg[1][i];
            s0 += pre * (two - guii - gdii);
            s1 += pre * (one - guii);
            s2 += pre * (one - gdii);
            s3 += pre * (one - guii) * (one - gdii);
        }
        acc[off[EQ_density] + r] = s0;
        acc[off[EQ_density_u] + r] = s1;
        acc[off[EQ_density_d] + r] = s2;
        acc[off[EQ_double_occ] + r] = s3;
    }
    const int* inv = a.invij + c * a.sinv;
    const T pre = phase * (1.0 / a.degen[4 * c + 1]);
    T g00 = zero, g00u = zero, g00d = zero, nn = zero, xx = zero, zz = zero, psw = zero;
    T vv = zero, vn = zero;
    if (r < nij) {
        g00 = acc[off[EQ_g00] + r]; g00u = acc[off[EQ_g00_u] + r]; g00d = acc[off[EQ_g00_d] + r];
        nn = acc[off[EQ_nn] + r]; xx = acc[off[EQ_xx] + r]; zz = acc[off[EQ_zz] + r];
        psw = acc[off[EQ_pair_sw] + r];
        if (a.energy) { vv = acc[off[EQ_vv] + r]; vn = acc[off[EQ_vn] + r]; }
    }
    // as in k_meas_ue_strip: the strip's columns fetched one strip ahead
    int iv[2][CH];
    T cu[2][CH], cd[2][CH];
    auto fetch = [&](int j0, auto hc) {
        constexpr int h = decltype(hc)::value;
#pragma unroll
        for (int jj = 0; jj < CH; jj++) {
            const int j = j0 + jj;
            iv[h][jj] = (r < nij && j < N) ? inv[(long)j * nij + r] : 0;
        }
#pragma unroll
        for (int jj = 0; jj < CH; jj++) {
            const int j = j0 + jj;
            const bool ok = r < nij && j < N;
            cu[h][jj] = ok ? G[0][iv[h][jj] + (long)N * j] : zero;
            cd[h][jj] = ok ? G[1][iv[h][jj] + (long)N * j] : zero;
        }
    };
    fetch(0, std::integral_constant<int, 0>());
    // strips in pairs, so that the register set of a strip is a constant
    auto strip_step = [&](int j0, auto hc) {
        constexpr int h = decltype(hc)::value;
        const int ch = (N - j0 < CH) ? N - j0 : CH;
        __syncthreads();
        for (int e = tid; e < CH * N; e += nt) {
            const int jj = e % CH, i = e / CH;
            if (jj < ch) {
#pragma unroll
                for (int s = 0; s < 2; s++) {
                    strip[s][jj][i] = G[s][(j0 + jj) + (long)N * i];
                    if constexpr (Z) pstrip[s][jj][i] = P[s][(j0 + jj) + (long)N * i];
                }
            }
        }
        if (j0 + CH < N) fetch(j0 + CH, std::integral_constant<int, 1 - h>());
        __syncthreads();
        if (r < nij) {
#pragma unroll
            for (int jj = 0; jj < CH; jj++) {
                if (jj >= ch) continue;
                const int j = j0 + jj;
                const int i = iv[h][jj];
                const double delta = (i == j) ? 1.0 : 0.0;
                const T guii = dg[0][i], gdii = dg[1][i], gujj = dg[0][j], gdjj = dg[1][j];
                const T guij = cu[h][jj], gdij = cd[h][jj];
                const T guji = strip[0][jj][i], gdji = strip[1][jj][i];
                T pgu = guij, pgd = gdij;
                if constexpr (Z) {   // g * p, the product order of peierls_mul
                    pgu = guij * pstrip[0][jj][i];
                    pgd = gdij * pstrip[1][jj][i];
                }
                g00 += 0.5 * (pre * (pgu + pgd));
                g00u += pre * pgu;
                g00d += pre * pgd;
                const T x = delta * (guii + gdii) - (guji * guij + gdji * gdij);
                nn += pre * ((two - guii - gdii) * (two - gujj - gdjj) + x);
                xx += 0.25 * (pre * (delta * (guii + gdii) - (guji * gdij + gdji * guij)));
                zz += 0.25 * (pre * ((gdii - guii) * (gdjj - gujj) + x));
                psw += pre * guij * gdij;
                if (a.energy) {
                    const T nuinuj = (one - guii) * (one - gujj) + (delta - guji) * guij;
                    const T ndindj = (one - gdii) * (one - gdjj) + (delta - gdji) * gdij;
                    vv += pre * nuinuj * ndindj;
                    vn += pre * (nuinuj * (one - gdii) + (one - guii) * ndindj);
                }
            }
        }
    };
    for (int j0 = 0; j0 < N; j0 += 2 * CH) {
        strip_step(j0, std::integral_constant<int, 0>());
        if (j0 + CH < N) strip_step(j0 + CH, std::integral_constant<int, 1>());
    }
    if (r < nij) {
        acc[off[EQ_g00] + r] = g00; acc[off[EQ_g00_u] + r] = g00u; acc[off[EQ_g00_d] + r] = g00d;
        acc[off[EQ_nn] + r] = nn; acc[off[EQ_xx] + r] = xx; acc[off[EQ_zz] + r] = zz;
        acc[off[EQ_pair_sw] + r] = psw;
        if (a.energy) { acc[off[EQ_vv] + r] = vv; acc[off[EQ_vn] + r] = vn; }
    }
}

// Peierls factor p[i + N j] (complex builds; 1 in real builds)
template <typename T>
__device__ inline T pf(const T* P, long idx);
template <>
__device__ inline double pf<double>(const double*, long) { return 1.0; }
template <>
__device__ inline zd pf<zd>(const zd* P, long idx) { return P[idx]; }

// equal time, meas_energy_corr bond terms: kv, kn (1 bond - 1 site,
// src/meas.c:290-330) and kk (1 bond - 1 bond, src/meas.c:333-406)
template <typename T>
__global__ __launch_bounds__(256) void k_meas_eq_energy(MeasArgs<T> a)
{
    const int c = blockIdx.y;
    if (!a.flag[c]) return;
    const int N = a.N, nb = a.num_b;
    const T* gu = a.G + (2 * c) * a.sG;
    const T* gd = a.G + (2 * c + 1) * a.sG;
    const T phase = a.phase[c];
    T* acc = a.acc + c * a.sacc;
    const long* off = a.off;
    const T* pu = a.peierlsu + c * a.sP;
    const T* pd = a.peierlsd + c * a.sP;
    const int* bonds = a.bonds + c * a.sbonds;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const T one = from_re<T>(1.0), two = from_re<T>(2.0);
    if (t < a.num_bs) {
        const int* rp = a.rp_bs + c * a.srp_bs;
        const int* ix = a.ix_bs + c * a.six_bs;
        const T pre = phase * (1.0 / a.degen[4 * c + 3]);
        T kv = acc[off[EQ_kv] + t], kn = acc[off[EQ_kn] + t];
        for (int q = rp[t]; q < rp[t + 1]; q++) {
            const int lin = ix[q];
            const int b = lin % nb, j = lin / nb;
            const int i0 = bonds[b], i1 = bonds[b + nb];
            const T pui0i1 = pf(pu, i0 + (long)N * i1), pui1i0 = pf(pu, i1 + (long)N * i0);
            const T pdi0i1 = pf(pd, i0 + (long)N * i1), pdi1i0 = pf(pd, i1 + (long)N * i0);
            const double d0j = (i0 == j) ? 1.0 : 0.0, d1j = (i1 == j) ? 1.0 : 0.0;
            const T gui0j = gu[i0 + (long)N * j], guji0 = gu[j + (long)N * i0];
            const T gdi0j = gd[i0 + (long)N * j], gdji0 = gd[j + (long)N * i0];
            const T gui1j = gu[i1 + (long)N * j], guji1 = gu[j + (long)N * i1];
            const T gdi1j = gd[i1 + (long)N * j], gdji1 = gd[j + (long)N * i1];
            const T gui0i1 = gu[i0 + (long)N * i1], gui1i0 = gu[i1 + (long)N * i0];
            const T gdi0i1 = gd[i0 + (long)N * i1], gdi1i0 = gd[i1 + (long)N * i0];
            const T gujj = gu[j + (long)N * j], gdjj = gd[j + (long)N * j];
            const T ku = pui1i0 * (0.0 - gui0i1) + pui0i1 * (0.0 - gui1i0);
            const T kd = pdi1i0 * (0.0 - gdi0i1) + pdi0i1 * (0.0 - gdi1i0);
            const T xu = pui0i1 * (d0j - guji0) * gui1j + pui1i0 * (d1j - guji1) * gui0j;
            const T xd = pdi0i1 * (d0j - gdji0) * gdi1j + pdi1i0 * (d1j - gdji1) * gdi0j;
            kv += pre * ((ku * (one - gujj) + xu) * (one - gdjj) + (kd * (one - gdjj) + xd) * (one - gujj));
            kn += pre * ((ku + kd) * (two - gujj - gdjj) + xu + xd);
        }
        acc[off[EQ_kv] + t] = kv;
        acc[off[EQ_kn] + t] = kn;
    }
    if (t < a.num_bb) {
        const int* rp = a.rp_bb + c * a.srp_bb;
        const int* ix = a.ix_bb + c * a.six_bb;
        const T pre = phase * (1.0 / a.degen[4 * c + 2]);
        T kk = acc[off[EQ_kk] + t];
        for (int q = rp[t]; q < rp[t + 1]; q++) {
            const int lin = ix[q];
            const int b = lin % nb, cc = lin / nb;
            const int j0 = bonds[cc], j1 = bonds[cc + nb];
            const int i0 = bonds[b], i1 = bonds[b + nb];
            const T puj0j1 = pf(pu, j0 + (long)N * j1), puj1j0 = pf(pu, j1 + (long)N * j0);
            const T pdj0j1 = pf(pd, j0 + (long)N * j1), pdj1j0 = pf(pd, j1 + (long)N * j0);
            const T pui0i1 = pf(pu, i0 + (long)N * i1), pui1i0 = pf(pu, i1 + (long)N * i0);
            const T pdi0i1 = pf(pd, i0 + (long)N * i1), pdi1i0 = pf(pd, i1 + (long)N * i0);
            const double d00 = (i0 == j0), d10 = (i1 == j0), d01 = (i0 == j1), d11 = (i1 == j1);
#define GU(x, y) gu[(x) + (long)N * (y)]
#define GD(x, y) gd[(x) + (long)N * (y)]
            const T x = pui0i1 * puj0j1 * (d01 - GU(j1, i0)) * GU(i1, j0) +
                        pui1i0 * puj1j0 * (d10 - GU(j0, i1)) * GU(i0, j1) +
                        pdi0i1 * pdj0j1 * (d01 - GD(j1, i0)) * GD(i1, j0) +
                        pdi1i0 * pdj1j0 * (d10 - GD(j0, i1)) * GD(i0, j1);
            const T y = pui0i1 * puj1j0 * (d00 - GU(j0, i0)) * GU(i1, j1) +
                        pui1i0 * puj0j1 * (d11 - GU(j1, i1)) * GU(i0, j0) +
                        pdi0i1 * pdj1j0 * (d00 - GD(j0, i0)) * GD(i1, j1) +
                        pdi1i0 * pdj0j1 * (d11 - GD(j1, i1)) * GD(i0, j0);
            kk += pre * ((pui1i0 * GU(i0, i1) + pui0i1 * GU(i1, i0) + pdi1i0 * GD(i0, i1) +
                          pdi0i1 * GD(i1, i0)) *
                             (puj1j0 * GU(j0, j1) + puj0j1 * GU(j1, j0) + pdj1j0 * GD(j0, j1) +
                              pdj0j1 * GD(j1, j0)) +
                         x + y);
#undef GU
#undef GD
        }
        acc[off[EQ_kk] + t] = kk;
    }
}

