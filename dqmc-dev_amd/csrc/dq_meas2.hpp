// dq_meas2.hpp -- the optional measurement groups of src/meas.c on the GPU:
//   unequal time: bond (pair_bb, jj, jsjs, kk, ksks), thermal (jnj, jjn,
//   jnjn, j2jn, jnj2), 2-bond (pair_b2b2, j2j2, js2js2, k2k2, ks2ks2, j2j,
//   jj2), energy (kv, kn), gen_suscept, nematic;
//   equal time: gen_suscept, chiral, local JQ (j, jn, j2).
//
// The reference writes each correlator as a hand-expanded Wick sum.  Here
// every correlator of two bond operators is assembled from per-spin pieces,
// because G is spin-diagonal and Wick's theorem factorises by spin:
//   hop   H = p c+_x c_y:   <H>   = -p G[y,x]           (no delta: the
//                                   reference's convention, also for a
//                                   two-hop bond that returns to its site)
//   density n_x:           <n_x> = 1 - G[x,x]
//   connected <c+_a c_b (t) c+_c c_d (0)> = (dt d_ad - G0t[d,a]) Gt0[b,c]
// A bond operator has the forward hop (x0 -> x1, phase pf) and the backward
// hop (x1 -> x0, phase pb); J = F - B, K = F + B, N = n_x0 + n_x1, and the
// heat-current piece jn = sum_s J_s N_{-s}.  Then, per (bin, t), e.g.
//   jj   = (J_u + J_d)^a (J_u + J_d)^b + C_u(J,J) + C_d(J,J)
//   jnj  = sum_s [ (J_s^a J_s^b + C_s(J,J)) N_-s^a + J_s^a (N_-s^a J_-s^b + C_-s(N,J)) ]
// and likewise for the others (see the kernels).  The per-spin form was
// checked against a term-by-term restatement of the reference formulas (the
// oracle) on random G, real and complex, to 1e-16.
//
// Every kernel owns one accumulator bin and sums the inputs of that bin in
// increasing linear index (host-built CSR of the reference's map), so results
// are deterministic and need no atomics.
#pragma once
#include "dq_common.hpp"
#include "dq_meas.hpp"

enum { UE_uuuu = 7, UE_dddd = 8, UE_uudd = 9, UE_dduu = 10, UE_pair_b2b2 = 16, UE_j2jn = 17,
       UE_jnj2 = 18, UE_jnjn = 19, UE_j2j = 20, UE_jj2 = 21, UE_j2j2 = 22, UE_js2js2 = 23,
       UE_k2k2 = 24, UE_ks2ks2 = 25, UE_jnj = 26, UE_jjn = 27, UE_kv = 28, UE_kn = 29,
       UE_nem_nnnn = 32, UE_nem_ssss = 33 };
enum { EQ_uuuu = 11, EQ_dddd = 12, EQ_uudd = 13, EQ_dduu = 14, EQ_chi = 15, EQ_j = 21,
       EQ_j2 = 22, EQ_jn = 23 };

// Green's functions of one spin at one slice (the unequal-time blocks).  G(0,t)
// is read from its transpose, otT[a + N b] = G0t[b, a]: a wave's lanes share
// the bond at 0 and differ in the bond at t, so the reads go down one column.
template <typename T>
struct SpinSlice {
    const T *tt, *oo, *otT, *t0;  // G(t,t), G(0,0), G(0,t)^T, G(t,0)
    int N;
    double dt;                   // 1 at t = 0 (the equal-time delta terms)
};

template <typename T>
struct BondOp {
    int x0, x1;
    T pf, pb;   // phases of c+_x0 c_x1 and c+_x1 c_x0
    T f, bk, n; // (tables) <F>, <B> and the bond density at the operator's slice
};

template <typename T>
struct PairVals {
    T ja, jb, ka, kb, cjj, ckk;  // <J>, <K> at t (a) and 0 (b); connected J-J, K-K
    T na, nb, cnj, cjn, cnn;     // bond densities and their connected terms
};

// per-spin pieces of bond operator a at t and bond operator b at 0; DA / DB:
// also the density pieces of a / b; TAB: the hop terms and densities come
// with the operators (k_bond_tab, the same expressions) instead of from G
template <typename T, bool DA, bool DB, bool TAB = false>
__device__ inline void pair_vals(const SpinSlice<T>& g, const BondOp<T>& a, const BondOp<T>& b,
                                 PairVals<T>& v)
{
    const long N = g.N;
    const int sa[2] = {a.x0, a.x1}, sb[2] = {b.x0, b.x1};
    T ot[2][2], t0[2][2];  // ot[q][p] = G0t[sb[q], sa[p]], t0[p][q] = Gt0[sa[p], sb[q]]
#pragma unroll
    for (int p = 0; p < 2; p++)
#pragma unroll
        for (int q = 0; q < 2; q++) {
            ot[q][p] = g.otT[sa[p] + N * sb[q]];
            t0[p][q] = g.t0[sa[p] + N * sb[q]];
        }
    // connected <c+_{sa[pa]} c_{sa[qa]} (t)  c+_{sb[pb]} c_{sb[qb]} (0)>
    auto cn = [&](int pa, int qa, int pb, int qb) -> T {
        return (from_re<T>(sa[pa] == sb[qb] ? g.dt : 0.0) - ot[qb][pa]) * t0[qa][pb];
    };
    T fa, ba, fb, bb;
    if constexpr (TAB) {
        fa = a.f; ba = a.bk; fb = b.f; bb = b.bk;
    } else {
        fa = -(a.pf * g.tt[a.x1 + N * a.x0]);
        ba = -(a.pb * g.tt[a.x0 + N * a.x1]);
        fb = -(b.pf * g.oo[b.x1 + N * b.x0]);
        bb = -(b.pb * g.oo[b.x0 + N * b.x1]);
    }
    const T cff = a.pf * b.pf * cn(0, 1, 0, 1), cfb = a.pf * b.pb * cn(0, 1, 1, 0);
    const T cbf = a.pb * b.pf * cn(1, 0, 0, 1), cbb = a.pb * b.pb * cn(1, 0, 1, 0);
    v.ja = fa - ba;
    v.ka = fa + ba;
    v.jb = fb - bb;
    v.kb = fb + bb;
    v.cjj = cff - cfb - cbf + cbb;
    v.ckk = cff + cfb + cbf + cbb;
    if (DA) {
        if constexpr (TAB) v.na = a.n;
        else v.na = from_re<T>(2.0) - g.tt[a.x0 + N * a.x0] - g.tt[a.x1 + N * a.x1];
        v.cnj = b.pf * (cn(0, 0, 0, 1) + cn(1, 1, 0, 1)) - b.pb * (cn(0, 0, 1, 0) + cn(1, 1, 1, 0));
    }
    if (DB) {
        if constexpr (TAB) v.nb = b.n;
        else v.nb = from_re<T>(2.0) - g.oo[b.x0 + N * b.x0] - g.oo[b.x1 + N * b.x1];
        v.cjn = a.pf * (cn(0, 1, 0, 0) + cn(0, 1, 1, 1)) - a.pb * (cn(1, 0, 0, 0) + cn(1, 0, 1, 1));
    }
    if (DA && DB) v.cnn = cn(0, 0, 0, 0) + cn(0, 0, 1, 1) + cn(1, 1, 0, 0) + cn(1, 1, 1, 1);
}

// spin-resolved combinations (s = 0 up, 1 down)
template <typename T>
__device__ inline T corr_jj(const PairVals<T>* v, bool spin_odd)
{
    const T sa = spin_odd ? v[0].ja - v[1].ja : v[0].ja + v[1].ja;
    const T sb = spin_odd ? v[0].jb - v[1].jb : v[0].jb + v[1].jb;
    return sa * sb + v[0].cjj + v[1].cjj;
}
template <typename T>
__device__ inline T corr_kk(const PairVals<T>* v, bool spin_odd)
{
    const T sa = spin_odd ? v[0].ka - v[1].ka : v[0].ka + v[1].ka;
    const T sb = spin_odd ? v[0].kb - v[1].kb : v[0].kb + v[1].kb;
    return sa * sb + v[0].ckk + v[1].ckk;
}
// <jn^a J^b>
template <typename T>
__device__ inline T corr_jnj(const PairVals<T>* v)
{
    T r = from_re<T>(0.0);
#pragma unroll
    for (int s = 0; s < 2; s++) {
        const PairVals<T>&x = v[s], &y = v[1 - s];
        r += (x.ja * x.jb + x.cjj) * y.na + x.ja * (y.na * y.jb + y.cnj);
    }
    return r;
}
// <J^a jn^b>
template <typename T>
__device__ inline T corr_jjn(const PairVals<T>* v)
{
    T r = from_re<T>(0.0);
#pragma unroll
    for (int s = 0; s < 2; s++) {
        const PairVals<T>&x = v[s], &y = v[1 - s];
        r += (x.ja * x.jb + x.cjj) * y.nb + x.jb * (y.ja * y.nb + y.cjn);
    }
    return r;
}
// <jn^a jn^b>
template <typename T>
__device__ inline T corr_jnjn(const PairVals<T>* v)
{
    T r = from_re<T>(0.0);
#pragma unroll
    for (int s = 0; s < 2; s++) {
        const PairVals<T>&x = v[s], &y = v[1 - s];
        r += (x.ja * x.jb + x.cjj) * (y.na * y.nb + y.cnn) + (x.ja * x.nb + x.cjn) * (y.na * y.jb + y.cnj);
    }
    return r;
}

// phases of bond b, spin s
template <typename T>
__device__ inline BondOp<T> bond1(const int* bonds, int nb, int b, const T* P, long N)
{
    BondOp<T> o;
    o.x0 = bonds[b];
    o.x1 = bonds[b + nb];
    o.pf = pf<T>(P, o.x0 + N * o.x1);
    o.pb = pf<T>(P, o.x1 + N * o.x0);
    return o;
}
// bond b of spin s from the tables of k_bond_tab: phases bph (the chain's)
// and hop terms ht = [f, b, n][b] of the operator's slice and spin
template <typename T>
__device__ inline BondOp<T> bond1_tab(const int* bonds, int nb, int b, const T* bph, int s,
                                      const T* ht)
{
    BondOp<T> o;
    o.x0 = bonds[b];
    o.x1 = bonds[b + nb];
    o.pf = bph[(2 * s) * nb + b];
    o.pb = bph[(2 * s + 1) * nb + b];
    o.f = ht[b];
    o.bk = ht[nb + b];
    o.n = ht[2 * nb + b];
    return o;
}

// The tables read by bond1_tab, per (chain, slice t, bond b): for each spin
// <F> = -pf G(t,t)[x1, x0], <B> = -pb G(t,t)[x0, x1], the bond density
// 2 - G(t,t)[x0, x0] - G(t,t)[x1, x1] (as pair_vals forms them from G), and
// at t = 0 the phases.  A wave of the pair kernel differs in the bond at t,
// so reading these by bond index replaces gathers over G and the phases.
template <typename T>
__global__ __launch_bounds__(256) void k_bond_tab(MeasArgs<T> a, T* btab, T* bph)
{
    const int c = blockIdx.z;
    if (!a.flag[c]) return;
    const int tt = blockIdx.y, nb = a.num_b;
    const int b = blockIdx.x * 256 + threadIdx.x;
    if (b >= nb) return;
    const long N = a.N, NN = N * N;
    const int* bonds = a.bonds + c * a.sbonds;
    const int x0 = bonds[b], x1 = bonds[b + nb];
#pragma unroll
    for (int s = 0; s < 2; s++) {
        const T* P = (s ? a.peierlsd : a.peierlsu) + c * a.sP;
        const T pfv = pf<T>(P, x0 + N * x1), pbv = pf<T>(P, x1 + N * x0);
        const T* g = a.Gtt + (2 * c + s) * a.sU + NN * tt;
        T* h = btab + ((long)(2 * c + s) * a.L + tt) * 3 * nb;
        h[b] = -(pfv * g[x1 + N * x0]);
        h[nb + b] = -(pbv * g[x0 + N * x1]);
        h[2 * nb + b] = from_re<T>(2.0) - g[x0 + N * x0] - g[x1 + N * x1];
        if (tt == 0) {
            bph[((long)c * 4 + 2 * s) * nb + b] = pfv;
            bph[((long)c * 4 + 2 * s + 1) * nb + b] = pbv;
        }
    }
}

// two-hop bond c of spin s: phases from pp = [pp_u, ppr_u, pp_d, ppr_d][c]
// and hop terms ht = [f, b, n][c] of the operator's slice (k_bond2_tab)
template <typename T>
__device__ inline BondOp<T> bond2_tab(const int* b2s, int nb2, int c, const T* pp, int s,
                                      const T* ht)
{
    BondOp<T> o;
    o.x0 = b2s[c];
    o.x1 = b2s[c + nb2];
    o.pf = pp[(2 * s) * nb2 + c];
    o.pb = pp[(2 * s + 1) * nb2 + c];
    o.f = ht[c];
    o.bk = ht[nb2 + c];
    o.n = ht[2 * nb2 + c];
    return o;
}

// k_bond_tab for the two-hop bonds (phases from pp, no phase table)
template <typename T>
__global__ __launch_bounds__(256) void k_bond2_tab(MeasArgs<T> a, T* b2tab)
{
    const int c = blockIdx.z;
    if (!a.flag[c]) return;
    const int tt = blockIdx.y, nb2 = a.num_b2;
    const int b = blockIdx.x * 256 + threadIdx.x;
    if (b >= nb2) return;
    const long N = a.N, NN = N * N;
    const int* b2s = a.bond2s + c * a.sbond2s;
    const T* pp = a.pp + c * a.spp;
    const int x0 = b2s[b], x1 = b2s[b + nb2];
#pragma unroll
    for (int s = 0; s < 2; s++) {
        const T pfv = pp[(2 * s) * nb2 + b], pbv = pp[(2 * s + 1) * nb2 + b];
        const T* g = a.Gtt + (2 * c + s) * a.sU + NN * tt;
        T* h = b2tab + ((long)(2 * c + s) * a.L + tt) * 3 * nb2;
        h[b] = -(pfv * g[x1 + N * x0]);
        h[nb2 + b] = -(pbv * g[x0 + N * x1]);
        h[2 * nb2 + b] = from_re<T>(2.0) - g[x0 + N * x0] - g[x1 + N * x1];
    }
}

// two-hop bond c: pp = [pp_u, ppr_u, pp_d, ppr_d][c]
template <typename T>
__device__ inline BondOp<T> bond2(const int* b2s, int nb2, int c, const T* pp, int s)
{
    BondOp<T> o;
    o.x0 = b2s[c];
    o.x1 = b2s[c + nb2];
    o.pf = pp[(2 * s) * nb2 + c];
    o.pb = pp[(2 * s + 1) * nb2 + c];
    return o;
}

enum { PK_BB = 0, PK_B2B = 1, PK_BB2 = 2, PK_B2B2 = 3 };

// unequal-time bond-type correlators over one of the reference's pair maps:
//   PK_BB   map_bb  [b + c nb ]: bond b (t)  x bond c (0)   pair_bb jj jsjs kk ksks jnj jjn jnjn
//   PK_B2B  map_bb2 [b + c nb2]: 2-bond b (t) x bond c (0)  j2jn j2j
//   PK_BB2  map_b2b [b + c nb ]: bond b (t)  x 2-bond c (0) jnj2 jj2
//   PK_B2B2 map_b2b2[b + c nb2]: 2-bond x 2-bond            pair_b2b2 j2j2 js2js2 k2k2 ks2ks2
// (src/meas.c:860-1365).  One thread per (bin, t).
template <typename T, int PK>
__global__ __launch_bounds__(256) void k_meas_ue_pairs(MeasArgs<T> a)
{
    const int c = blockIdx.z;
    if (!a.flag[c]) return;
    const int tt = blockIdx.y;
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    const int nbins = PK == PK_BB ? a.num_bb : PK == PK_B2B ? a.num_bb2 : PK == PK_BB2 ? a.num_b2b : a.num_b2b2;
    if (r >= nbins) return;
    const int N = a.N, nb = a.num_b, nb2 = a.num_b2;
    const long NN = (long)N * N;
    const T phase = a.phase[c];
    T* acc = a.acc + c * a.sacc;
    const long* off = a.off;
    const int* bonds = a.bonds + c * a.sbonds;
    const int* b2s = a.bond2s + c * a.sbond2s;
    const T* pp = a.pp + c * a.spp;
    // per-bond tables (k_bond_tab, k_bond2_tab) at slice t and at 0, per spin
    const T* bph = a.bph ? a.bph + (long)c * 4 * nb : nullptr;
    const T *T1t[2], *T1z[2], *T2t[2], *T2z[2];
#pragma unroll
    for (int s = 0; s < 2; s++) {
        const long bt = (long)(2 * c + s) * a.L;
        T1t[s] = a.btab ? a.btab + (bt + tt) * 3 * nb : nullptr;
        T1z[s] = a.btab ? a.btab + bt * 3 * nb : nullptr;
        T2t[s] = a.b2tab ? a.b2tab + (bt + tt) * 3 * nb2 : nullptr;
        T2z[s] = a.b2tab ? a.b2tab + bt * 3 * nb2 : nullptr;
    }
    SpinSlice<T> g[2];
#pragma unroll
    for (int s = 0; s < 2; s++) {
        const long bo = (2 * c + s) * a.sU;
        g[s].tt = a.Gtt + bo + NN * tt;
        g[s].oo = a.Gtt + bo;
        g[s].otT = a.G0tT + bo + NN * tt;
        g[s].t0 = a.Gt0 + bo + NN * tt;
        g[s].N = N;
        g[s].dt = (tt == 0) ? 1.0 : 0.0;
    }
    const int mi = PK == PK_BB ? 0 : PK == PK_B2B ? PM_bb2 : PK == PK_BB2 ? PM_b2b : PM_b2b2;
    const int* rp = PK == PK_BB ? a.rp_bb + c * a.srp_bb : a.rp2[mi] + c * a.srp2[mi];
    const int* ix = PK == PK_BB ? a.ix_bb + c * a.six_bb : a.ix2[mi] + c * a.six2[mi];
    const int deg = PK == PK_BB ? a.degen[4 * c + 2] : a.degen2[PM_COUNT * c + mi];
    const T pre = phase * (1.0 / deg);
    const long o = r + (long)nbins * tt;
    const bool bond = a.bond, thermal = a.thermal, two = a.twobond, pbb = a.pbbonly;
    T s0 = from_re<T>(0.0), s1 = s0, s2 = s0, s3 = s0, s4 = s0, s5 = s0, s6 = s0, s7 = s0;
    for (int q = rp[r]; q < rp[r + 1]; q++) {
        const int lin = ix[q];
        PairVals<T> v[2];
        if (PK == PK_BB) {
            const int b = lin % nb, cc = lin / nb;
#pragma unroll
            for (int s = 0; s < 2; s++) {
                const BondOp<T> A = bond1_tab<T>(bonds, nb, b, bph, s, T1t[s]),
                                B = bond1_tab<T>(bonds, nb, cc, bph, s, T1z[s]);
                if (thermal) pair_vals<T, true, true, true>(g[s], A, B, v[s]);
                else pair_vals<T, false, false, true>(g[s], A, B, v[s]);
            }
            if (bond || pbb) {
                const int i0 = bonds[b], i1 = bonds[b + nb], j0 = bonds[cc], j1 = bonds[cc + nb];
                const T* gu = g[0].t0;
                const T* gd = g[1].t0;
                s0 += 0.5 * (pre * (gu[i0 + N * (long)j0] * gd[i1 + N * (long)j1] +
                                    gu[i1 + N * (long)j0] * gd[i0 + N * (long)j1] +
                                    gu[i0 + N * (long)j1] * gd[i1 + N * (long)j0] +
                                    gu[i1 + N * (long)j1] * gd[i0 + N * (long)j0]));
            }
            if (bond) {
                s1 += pre * corr_jj(v, false);
                s2 += pre * corr_jj(v, true);
                s3 += pre * corr_kk(v, false);
                s4 += pre * corr_kk(v, true);
            }
            if (thermal) {
                s5 += pre * corr_jnj(v);
                s6 += pre * corr_jjn(v);
                s7 += pre * corr_jnjn(v);
            }
        } else if (PK == PK_B2B) {
            const int b = lin % nb2, cc = lin / nb2;
#pragma unroll
            for (int s = 0; s < 2; s++) {
                const BondOp<T> A = bond2_tab<T>(b2s, nb2, b, pp, s, T2t[s]),
                                B = bond1_tab<T>(bonds, nb, cc, bph, s, T1z[s]);
                if (thermal) pair_vals<T, false, true, true>(g[s], A, B, v[s]);
                else pair_vals<T, false, false, true>(g[s], A, B, v[s]);
            }
            if (thermal) s0 += pre * corr_jjn(v);
            if (two) s1 += pre * corr_jj(v, false);
        } else if (PK == PK_BB2) {
            const int b = lin % nb, cc = lin / nb;
#pragma unroll
            for (int s = 0; s < 2; s++) {
                const BondOp<T> A = bond1_tab<T>(bonds, nb, b, bph, s, T1t[s]),
                                B = bond2_tab<T>(b2s, nb2, cc, pp, s, T2z[s]);
                if (thermal) pair_vals<T, true, false, true>(g[s], A, B, v[s]);
                else pair_vals<T, false, false, true>(g[s], A, B, v[s]);
            }
            if (thermal) s0 += pre * corr_jnj(v);
            if (two) s1 += pre * corr_jj(v, false);
        } else {
            const int b = lin % nb2, cc = lin / nb2;
#pragma unroll
            for (int s = 0; s < 2; s++) {
                const BondOp<T> A = bond2_tab<T>(b2s, nb2, b, pp, s, T2t[s]),
                                B = bond2_tab<T>(b2s, nb2, cc, pp, s, T2z[s]);
                pair_vals<T, false, false, true>(g[s], A, B, v[s]);
            }
            const int i0 = b2s[b], i2 = b2s[b + nb2], j0 = b2s[cc], j2 = b2s[cc + nb2];
            const T* gu = g[0].t0;
            const T* gd = g[1].t0;
            s0 += 0.5 * (pre * (gu[i0 + N * (long)j0] * gd[i2 + N * (long)j2] +
                                gu[i2 + N * (long)j0] * gd[i0 + N * (long)j2] +
                                gu[i0 + N * (long)j2] * gd[i2 + N * (long)j0] +
                                gu[i2 + N * (long)j2] * gd[i0 + N * (long)j0]));
            s1 += pre * corr_jj(v, false);
            s2 += pre * corr_jj(v, true);
            s3 += pre * corr_kk(v, false);
            s4 += pre * corr_kk(v, true);
        }
    }
    if (PK == PK_BB) {
        if (bond || pbb) acc[off[UE_pair_bb] + o] += s0;
        if (bond) {
            acc[off[UE_jj] + o] += s1;
            acc[off[UE_jsjs] + o] += s2;
            acc[off[UE_kk] + o] += s3;
            acc[off[UE_ksks] + o] += s4;
        }
        if (thermal) {
            acc[off[UE_jnj] + o] += s5;
            acc[off[UE_jjn] + o] += s6;
            acc[off[UE_jnjn] + o] += s7;
        }
    } else if (PK == PK_B2B) {
        if (thermal) acc[off[UE_j2jn] + o] += s0;
        if (two) acc[off[UE_j2j] + o] += s1;
    } else if (PK == PK_BB2) {
        if (thermal) acc[off[UE_jnj2] + o] += s0;
        if (two) acc[off[UE_jj2] + o] += s1;
    } else {
        acc[off[UE_pair_b2b2] + o] += s0;
        acc[off[UE_j2j2] + o] += s1;
        acc[off[UE_js2js2] + o] += s2;
        acc[off[UE_k2k2] + o] += s3;
        acc[off[UE_ks2ks2] + o] += s4;
    }
}

// unequal-time kv, kn: bond b (t) x site j (0), map_bs[b + nb j]
// (src/meas.c:409-471).  kv = <K (t) n_ju n_jd (0)>, kn = <K (t) (n_ju + n_jd)(0)>.
template <typename T>
__global__ __launch_bounds__(256) void k_meas_ue_kv(MeasArgs<T> a)
{
    const int c = blockIdx.z;
    if (!a.flag[c]) return;
    const int tt = blockIdx.y;
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= a.num_bs) return;
    const int N = a.N, nb = a.num_b;
    const long NN = (long)N * N;
    const T phase = a.phase[c];
    T* acc = a.acc + c * a.sacc;
    const int* bonds = a.bonds + c * a.sbonds;
    const T* bph = a.bph + (long)c * 4 * nb;
    const int* rp = a.rp_bs + c * a.srp_bs;
    const int* ix = a.ix_bs + c * a.six_bs;
    const T pre = phase * (1.0 / a.degen[4 * c + 3]);
    const double dt = (tt == 0) ? 1.0 : 0.0;
    T kv = from_re<T>(0.0), kn = kv;
    for (int q = rp[r]; q < rp[r + 1]; q++) {
        const int lin = ix[q];
        const int b = lin % nb, j = lin / nb;
        T k[2], cn[2], n[2];
#pragma unroll
        for (int s = 0; s < 2; s++) {
            const long bo = (2 * c + s) * a.sU;
            const T* goo = a.Gtt + bo;
            const T* gotT = a.G0tT + bo + NN * tt;  // G(0,t)^T
            const T* gt0 = a.Gt0 + bo + NN * tt;
            // the bond's phases and hop terms at t from k_bond_tab's tables
            const BondOp<T> A =
                bond1_tab<T>(bonds, nb, b, bph, s, a.btab + ((long)(2 * c + s) * a.L + tt) * 3 * nb);
            k[s] = A.f + A.bk;
            n[s] = from_re<T>(1.0) - goo[j + N * (long)j];
            // connected <c+_x c_y (t) n_j (0)> = (dt d_xj - G0t[j,x]) Gt0[y,j]
            cn[s] = A.pf * ((from_re<T>(A.x0 == j ? dt : 0.0) - gotT[A.x0 + N * (long)j]) * gt0[A.x1 + N * (long)j]) +
                    A.pb * ((from_re<T>(A.x1 == j ? dt : 0.0) - gotT[A.x1 + N * (long)j]) * gt0[A.x0 + N * (long)j]);
        }
        kv += pre * ((k[0] * n[0] + cn[0]) * n[1] + (k[1] * n[1] + cn[1]) * n[0]);
        kn += pre * ((k[0] + k[1]) * (n[0] + n[1]) + cn[0] + cn[1]);
    }
    const long o = r + (long)a.num_bs * tt;
    acc[a.off[UE_kv] + o] += kv;
    acc[a.off[UE_kn] + o] += kn;
}

// gen_suscept (src/meas.c:122-157 eq, 473-516 uneqlt).  The reference's
// quadruple loop adds (dt d - G0t[j2,i1]) * Gt0[j1,i2] into bin
// map_ij[i1 + N j2] + num_ij map_ij[i2 + N j1]: the sum factorises into
//   A_s[k] = sum_{map_ij[i + N j] = k} (dt d_ij - G0t_s[j, i])
//   B_s[k] = sum_{map_ij[i + N j] = k} Gt0_s[j, i]
// and uuuu[k1 + nij k2] += pre A_u[k1] B_u[k2] (dddd: d d, uudd: A_u B_d,
// dduu: A_d B_u), O(N^2) instead of O(N^4).  Step 1: A, B per (chain, t).
template <typename T>
__global__ __launch_bounds__(256) void k_gensus_ab(MeasArgs<T> a, int ue)
{
    const int c = blockIdx.z;
    if (!a.flag[c]) return;
    const int tt = blockIdx.y;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int nij = a.num_ij;
    if (k >= nij) return;
    const int N = a.N;
    const long NN = (long)N * N;
    const int* rp = a.rp_ij + c * a.srp_ij;
    const int* ix = a.ix_ij + c * a.six_ij;
    const double dt = (tt == 0) ? 1.0 : 0.0;
    T* out = a.gs + c * a.sgs + (long)tt * 4 * nij;
#pragma unroll
    for (int s = 0; s < 2; s++) {
        const T *G0, *G1;
        if (ue) {
            G0 = a.G0t + (2 * c + s) * a.sU + NN * tt;
            G1 = a.Gt0 + (2 * c + s) * a.sU + NN * tt;
        } else {
            G0 = G1 = a.G + (2 * c + s) * a.sG;
        }
        T A = from_re<T>(0.0), B = A;
        for (int q = rp[k]; q < rp[k + 1]; q++) {
            const int lin = ix[q];
            const int i = lin % N, j = lin / N;
            A += from_re<T>(i == j ? dt : 0.0) - G0[j + (long)N * i];
            B += G1[j + (long)N * i];
        }
        out[s * nij + k] = A;
        out[(2 + s) * nij + k] = B;
    }
}

// step 2: the outer products into the num_ij^2 bins (per t for uneqlt)
template <typename T>
__global__ __launch_bounds__(256) void k_gensus_outer(MeasArgs<T> a, int ue)
{
    const int c = blockIdx.z;
    if (!a.flag[c]) return;
    const int tt = blockIdx.y;
    const int nij = a.num_ij;
    const long r = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (r >= (long)nij * nij) return;
    const int k1 = (int)(r % nij), k2 = (int)(r / nij);
    const T* ab = a.gs + c * a.sgs + (long)tt * 4 * nij;
    const T pre = a.phase[c] * (1.0 / ((double)a.Nx * a.Ny * a.Nx * a.Ny));
    const T Au = ab[k1], Ad = ab[nij + k1], Bu = ab[2 * nij + k2], Bd = ab[3 * nij + k2];
    T* acc = a.acc + c * a.sacc;
    const long o = r + (long)nij * nij * tt;
    const int ku = ue ? UE_uuuu : EQ_uuuu;
    acc[a.off[ku] + o] += pre * (Au * Bu);
    acc[a.off[ku + 1] + o] += pre * (Ad * Bd);
    acc[a.off[ku + 2] + o] += pre * (Au * Bd);
    acc[a.off[ku + 3] + o] += pre * (Ad * Bu);
}

// All 16 principal minors of a 4x4 matrix, mn[mask] = det of the rows and
// columns in mask (mn[0] = 1), by cofactor expansion with every index fixed at
// compile time, so that the matrix and the minors stay in registers.
__device__ inline void principal_minors4(const double (&M)[4][4], double (&mn)[16])
{
    mn[0] = 1.0;
#pragma unroll
    for (int i = 0; i < 4; i++) mn[1 << i] = M[i][i];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = i + 1; j < 4; j++) mn[(1 << i) | (1 << j)] = M[i][i] * M[j][j] - M[i][j] * M[j][i];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = i + 1; j < 4; j++)
#pragma unroll
            for (int k = j + 1; k < 4; k++)
                mn[(1 << i) | (1 << j) | (1 << k)] =
                    M[i][i] * (M[j][j] * M[k][k] - M[j][k] * M[k][j]) -
                    M[i][j] * (M[j][i] * M[k][k] - M[j][k] * M[k][i]) +
                    M[i][k] * (M[j][i] * M[k][j] - M[j][j] * M[k][i]);
    // Laplace expansion along rows 0 and 1
    auto s2 = [&](int a_, int b_) { return M[0][a_] * M[1][b_] - M[0][b_] * M[1][a_]; };
    auto c2 = [&](int a_, int b_) { return M[2][a_] * M[3][b_] - M[2][b_] * M[3][a_]; };
    mn[15] = s2(0, 1) * c2(2, 3) - s2(0, 2) * c2(1, 3) + s2(0, 3) * c2(1, 2) +
             s2(1, 2) * c2(0, 3) - s2(1, 3) * c2(0, 2) + s2(2, 3) * c2(0, 1);
}

// nematic (real build only, src/meas.c:518-715): over the first 2N bonds,
// <n_i0 n_i1 (t) n_j0 n_j1 (0)> for every spin assignment, summed (nnnn) and
// with the sign (-1)^(number of down spins) (ssss).  For one assignment the
// expectation is the product over spins of the determinant of the
// contractions of the densities carrying that spin; the reference takes every
// delta as 0, so the contraction matrix is
//   diagonal 1 - G[x,x];  same time -G[y,x];  t->0 -G0t[y,x];  0->t -Gt0[y,x]
// (checked against the reference's 16 expanded terms).
__global__ __launch_bounds__(256) void k_meas_ue_nematic(MeasArgs<double> a)
{
    const int c = blockIdx.z;
    if (!a.flag[c]) return;
    const int tt = blockIdx.y;
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= a.num_bb) return;
    const int N = a.N, nb = a.num_b;
    const long NN = (long)N * N;
    const int* bonds = a.bonds + c * a.sbonds;
    const int* rp = a.rp_bb + c * a.srp_bb;
    const int* ix = a.ix_bb + c * a.six_bb;
    const double pre = a.phase[c] / a.degen[4 * c + 2];
    double sn = 0.0, ss = 0.0;
    for (int q = rp[r]; q < rp[r + 1]; q++) {
        const int lin = ix[q];
        const int b = lin % nb, cc = lin / nb;
        if (b >= 2 * N || cc >= 2 * N) continue;
        const int site[4] = {bonds[b], bonds[b + nb], bonds[cc], bonds[cc + nb]};
        double minor[2][16];
#pragma unroll
        for (int s = 0; s < 2; s++) {
            const long bo = (2 * c + s) * a.sU;
            const double* gtt = a.Gtt + bo + NN * tt;
            const double* goo = a.Gtt + bo;
            const double* got = a.G0t + bo + NN * tt;
            const double* gt0 = a.Gt0 + bo + NN * tt;
            double M[4][4];
#pragma unroll
            for (int u = 0; u < 4; u++)
#pragma unroll
                for (int w = 0; w < 4; w++) {
                    const long xu = site[u], xw = site[w];
                    double e;
                    if (u == w) e = 1.0 - (u < 2 ? gtt[xu + N * xu] : goo[xu + N * xu]);
                    else if (u < 2 && w < 2) e = -gtt[xw + N * xu];
                    else if (u >= 2 && w >= 2) e = -goo[xw + N * xu];
                    else if (u < 2) e = -got[xw + N * xu];
                    else e = -gt0[xw + N * xu];
                    M[u][w] = e;
                }
            principal_minors4(M, minor[s]);
        }
        // bit a of mask = 1: operator a has spin up
#pragma unroll
        for (int m = 0; m < 16; m++) {
            const double v = minor[0][m] * minor[1][15 - m];
            sn += v;
            ss += (__popc(15 - m) & 1) ? -v : v;
        }
    }
    const long o = r + (long)a.num_bb * tt;
    a.acc[c * a.sacc + a.off[UE_nem_nnnn] + o] += pre * sn;
    a.acc[c * a.sacc + a.off[UE_nem_ssss] + o] += pre * ss;
}

// equal-time chiral (src/meas.c:159-208) and local JQ (:211-299) on the
// half-wrapped G; one thread per bin of map_plaq / map_b / map_b2
template <typename T>
__global__ __launch_bounds__(256) void k_meas_eq_extra(MeasArgs<T> a)
{
    const int c = blockIdx.y;
    if (!a.flag[c]) return;
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    const int N = a.N;
    const long Nl = N;
    const T* gu = a.G + (2 * c) * a.sG;
    const T* gd = a.G + (2 * c + 1) * a.sG;
    const T phase = a.phase[c];
    T* acc = a.acc + c * a.sacc;
    const long* off = a.off;
    if (a.chiral && r < a.num_plaq_acc) {
        const int* rp = a.rp2[PM_plaq] + c * a.srp2[PM_plaq];
        const int* ix = a.ix2[PM_plaq] + c * a.six2[PM_plaq];
        const int* pl = a.plaqs + c * a.splaqs;
        const int np = a.num_plaq;
        const T pre = phase * (1.0 / a.degen2[PM_COUNT * c + PM_plaq]);
        T s = from_re<T>(0.0);
        for (int q = rp[r]; q < rp[r + 1]; q++) {
            const int p = ix[q];
            const long x[3] = {pl[p], pl[p + np], pl[p + 2 * np]};
            // u[a][b] = G_up(x_a, x_b), d[a][b] = G_dn(x_a, x_b)
            T u[3][3], d[3][3];
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) {
                    u[i][j] = gu[x[i] + Nl * x[j]];
                    d[i][j] = gd[x[i] + Nl * x[j]];
                }
            s += pre * (d[1][1] * d[2][0] * u[0][2] - d[0][2] * d[2][1] * u[1][0] + d[0][1] * d[2][2] * u[1][0] -
                        d[2][1] * u[0][2] * u[1][0] - d[2][0] * u[0][2] * u[1][1] + d[0][1] * d[2][0] * u[1][2] -
                        d[0][0] * d[2][1] * u[1][2] + d[2][1] * u[0][0] * u[1][2] + d[2][0] * u[0][1] * u[1][2] -
                        d[0][2] * d[1][1] * u[2][0] + d[0][2] * u[1][1] * u[2][0] + d[0][1] * u[1][2] * u[2][0] -
                        d[0][2] * u[1][0] * u[2][1] +
                        d[1][2] * (d[2][0] * u[0][1] + (d[0][1] + u[0][1]) * u[2][0] + (d[0][0] - u[0][0]) * u[2][1]) -
                        d[1][0] * (d[2][1] * u[0][2] + (d[0][2] + u[0][2]) * u[2][1] + u[0][1] * (d[2][2] - u[2][2])) -
                        d[0][1] * u[1][0] * u[2][2]);
        }
        acc[off[EQ_chi] + r] += s;
    }
    if (a.localjq && r < a.num_b_acc) {
        const int* rp = a.rp2[PM_b] + c * a.srp2[PM_b];
        const int* ix = a.ix2[PM_b] + c * a.six2[PM_b];
        const int* bonds = a.bonds + c * a.sbonds;
        const int nb = a.num_b;
        const T* pu = a.peierlsu + c * a.sP;
        const T* pd = a.peierlsd + c * a.sP;
        const T pre = phase * (1.0 / a.degen2[PM_COUNT * c + PM_b]);
        T sj = from_re<T>(0.0), sjn = sj;
        for (int q = rp[r]; q < rp[r + 1]; q++) {
            const int b = ix[q];
            const long i0 = bonds[b], i1 = bonds[b + nb];
            const double dl = (i0 == i1) ? 1.0 : 0.0;
            const T u01 = pf<T>(pu, i0 + Nl * i1), u10 = pf<T>(pu, i1 + Nl * i0);
            const T d01 = pf<T>(pd, i0 + Nl * i1), d10 = pf<T>(pd, i1 + Nl * i0);
            const T hu10 = from_re<T>(dl) - gu[i1 + Nl * i0], hu01 = from_re<T>(dl) - gu[i0 + Nl * i1];
            const T hd10 = from_re<T>(dl) - gd[i1 + Nl * i0], hd01 = from_re<T>(dl) - gd[i0 + Nl * i1];
            const T nu = from_re<T>(2.0) - gu[i0 + Nl * i0] - gu[i1 + Nl * i1];
            const T nd = from_re<T>(2.0) - gd[i0 + Nl * i0] - gd[i1 + Nl * i1];
            // the reference pairs the up phase and density with the down hop
            // (and vice versa) in jn (src/meas.c:252-259); kept as is
            sjn += pre * (u01 * nu * hd10 + d01 * nd * hu10 - u10 * nu * hd01 - d10 * nd * hu01);
            sj += pre * (u01 * hu10 + d01 * hd10 - u10 * hu01 - d10 * hd01);
        }
        acc[off[EQ_jn] + r] += sjn;
        acc[off[EQ_j] + r] += sj;
    }
    if (a.localjq && r < a.num_b2_acc) {
        const int* rp = a.rp2[PM_b2] + c * a.srp2[PM_b2];
        const int* ix = a.ix2[PM_b2] + c * a.six2[PM_b2];
        const int* b2s = a.bond2s + c * a.sbond2s;
        const int nb2 = a.num_b2;
        const T* pp = a.pp + c * a.spp;
        const T pre = phase * (1.0 / a.degen2[PM_COUNT * c + PM_b2]);
        T s = from_re<T>(0.0);
        for (int q = rp[r]; q < rp[r + 1]; q++) {
            const int cc = ix[q];
            const long i0 = b2s[cc], i2 = b2s[cc + nb2];
            const double dl = (i0 == i2) ? 1.0 : 0.0;
            s += pre * (pp[cc] * (from_re<T>(dl) - gu[i2 + Nl * i0]) + pp[2 * nb2 + cc] * (from_re<T>(dl) - gd[i2 + Nl * i0]) -
                        pp[nb2 + cc] * (from_re<T>(dl) - gu[i0 + Nl * i2]) - pp[3 * nb2 + cc] * (from_re<T>(dl) - gd[i0 + Nl * i2]));
        }
        acc[off[EQ_j2] + r] += s;
    }
}
