// dq_rng.hpp -- the random stream of one sweep, generated ahead of the
// Metropolis kernels.
//
// The reference draws, per slice, N numbers for the inside-out Fisher-Yates
// site order and then one number per visited site whose Delta's are not both
// zero (src/rand.h:6-32, src/updates.c:21-26).  None of this depends on the
// acceptance outcomes, so the whole sweep's stream is known in advance:
// k_rng_sweep produces, per chain and slice, the site order and the uniform
// for every visited site (+inf where the reference draws nothing, so the site
// is never accepted), and advances the chain's xorshift1024* state exactly as
// the reference does.
//
// The state lives in wave-uniform registers (the generator runs on the scalar
// path); it is rotated once so that, in groups of 16 draws, every index is a
// compile-time constant.
#pragma once
#include "dq_common.hpp"

#define RNG_MAXN 1024

struct XsState {
    uint64_t r[16];  // r[q] = s[(p0 + q) & 15]
};

// one xorshift1024* step for the static position q (0..15) of a 16-group
template <int Q>
__device__ inline uint64_t xs_step(XsState& st)
{
    const uint64_t s0 = st.r[Q];
    uint64_t s1 = st.r[(Q + 1) & 15];
    s1 ^= s1 << 31;
    const uint64_t v = s1 ^ s0 ^ (s1 >> 11) ^ (s0 >> 30);
    st.r[(Q + 1) & 15] = v;
    return v * 1181783497276652981ULL;
}

template <int Q>
__device__ inline void xs_group(XsState& st, uint64_t* out, int lane)
{
    const uint64_t x = xs_step<Q>(st);
    if (lane == 0) out[Q] = x;
    if (Q < 15) xs_group<(Q < 15 ? Q + 1 : 15)>(st, out, lane);
}
template <>
__device__ inline void xs_group<15>(XsState& st, uint64_t* out, int lane)
{
    const uint64_t x = xs_step<15>(st);
    if (lane == 0) out[15] = x;
}

// raw[c][0..T) <- the next T = L*(N + draws[c]) draws of chain c (one sweep);
// state advanced.  One wave per chain; all lanes run the (uniform) generator,
// lane 0 stores.
__global__ __launch_bounds__(64) void k_rng_raw(int N, int L, const int* draws, uint64_t* rng,
                                                uint64_t* raw, long sraw, const int* active,
                                                const int* Lc, const int* Nc)
{
    const int c = blockIdx.x;
    if (!active[c]) return;
    // a chain padded to N, L draws for its own sites and slices only
    const int Lr = Lc ? Lc[c] : L, Nr = Nc ? Nc[c] : N;
    const int T = Lr * (Nr + draws[c]);
    const int lane = threadIdx.x;
    uint64_t* s = rng + 17L * c;
    uint64_t* out = raw + c * sraw;
    const int p0 = (int)(s[16] & 15);
    XsState st;
#pragma unroll
    for (int q = 0; q < 16; q++) st.r[q] = s[(p0 + q) & 15];
    int g = 0;
    for (; g + 16 <= T; g += 16) xs_group<0>(st, out + g, lane);
    // tail (< 16 draws): rotate through LDS with a dynamic index
    __shared__ uint64_t tail[16];
    const int rem = T - g;
    if (lane == 0) {
#pragma unroll
        for (int q = 0; q < 16; q++) tail[q] = st.r[q];
        int p = 0;
        for (int t = 0; t < rem; t++) {
            const uint64_t s0 = tail[p];
            const int pn = (p + 1) & 15;
            uint64_t s1 = tail[pn];
            s1 ^= s1 << 31;
            tail[pn] = s1 ^ s0 ^ (s1 >> 11) ^ (s0 >> 30);
            out[g + t] = tail[pn] * 1181783497276652981ULL;
            p = pn;
        }
        // write back: s[(p0 + q) & 15] = tail[q]; index advanced by T
        for (int q = 0; q < 16; q++) s[(p0 + q) & 15] = tail[q];
        s[16] = (uint64_t)((p0 + T) & 15);
    }
}

// From the raw draws: per slice, the inside-out Fisher-Yates order and the
// uniform of each visited site.  draws[c] = number of sites that draw.
// One 256-thread block per (chain); slices processed in sequence.
__global__ __launch_bounds__(256) void k_rng_orders(int N, int L, const uint64_t* raw, long sraw,
                                                    const double* del, const int* draws,
                                                    int* order, double* unif, const int* active,
                                                    const int* Lc, const int* Nc)
{
    const int c = blockIdx.x;
    if (!active[c]) return;
    // padded chains: orders over the chain's own Nr sites for its own Lr
    // slices; the padding sites follow in place and are never accepted
    const int Lr = Lc ? Lc[c] : L, Nr = Nc ? Nc[c] : N;
    __shared__ int jj[RNG_MAXN];
    __shared__ int ord[RNG_MAXN];
    const int tid = threadIdx.x;
    const uint64_t* x = raw + c * sraw;
    const double* dl = del + 2L * N * c;
    const int nd = draws[c];
    long pos = 0;
    for (int l = 0; l < Lr; l++) {
        for (int i = tid; i < Nr; i += blockDim.x)
            jj[i] = (int)((x[pos + i] >> 3) % (uint64_t)(i + 1));
        __syncthreads();
        if (tid == 0) {
            for (int i = 0; i < Nr; i++) {
                const int j = jj[i];
                if (i != j) ord[i] = ord[j];
                ord[j] = i;
            }
        }
        __syncthreads();
        int* o = order + ((long)c * L + l) * N;
        double* u = unif + ((long)c * L + l) * N;
        for (int i = tid; i < N; i += blockDim.x) o[i] = (i < Nr) ? ord[i] : i;
        for (int i = Nr + tid; i < N; i += blockDim.x)
            u[i] = __longlong_as_double(0x7FF0000000000000LL);  // +inf: never accepted
        if (nd == Nr) {
            for (int i = tid; i < Nr; i += blockDim.x) {
                const uint64_t bits = (0x3FFULL << 52) | (x[pos + Nr + i] >> 12);
                u[i] = __longlong_as_double((long long)bits) - 1.0;
            }
        } else if (tid == 0) {
            long q = pos + Nr;
            for (int ii = 0; ii < Nr; ii++) {
                const int i = ord[ii];
                if (dl[i] == 0.0 && dl[i + N] == 0.0) {
                    u[ii] = __longlong_as_double(0x7FF0000000000000LL);  // +inf: never accepted
                } else {
                    const uint64_t bits = (0x3FFULL << 52) | (x[q++] >> 12);
                    u[ii] = __longlong_as_double((long long)bits) - 1.0;
                }
            }
        }
        pos += Nr + nd;
        __syncthreads();
    }
}
