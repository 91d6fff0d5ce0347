// gemm_bench -- the engine's batched N x N x N product (the wrap and B-chain
// shape: C = A diag(u) B, batch 2 x chains) for every tile of dq_gemm.hpp,
// beside rocBLAS's strided-batched dgemm on the same operands (C = A B).
// Measurement tool, not part of the library.
//
//   gemm_bench [N=256] [batch=512] [reps=20]
//
// Prints per variant the mean time per launch over `reps` back-to-back
// launches (HIP events, after 3 warm-up launches) and FP64 TFLOP/s, and for
// every tile the largest difference from the 64x64 tile relative to max |C|.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "dq_gemm.hpp"

#define CK(x)                                                                   \
    do {                                                                        \
        hipError_t e_ = (x);                                                    \
        if (e_ != hipSuccess) {                                                 \
            fprintf(stderr, "HIP %s at line %d\n", hipGetErrorString(e_), __LINE__); \
            exit(1);                                                            \
        }                                                                       \
    } while (0)

int main(int argc, char** argv)
{
    const int N = argc > 1 ? atoi(argv[1]) : 256;
    const int B = argc > 2 ? atoi(argv[2]) : 512;
    const int reps = argc > 3 ? atoi(argv[3]) : 20;
    if (N <= 0 || B <= 0 || reps <= 0) return 2;
    const long NN = (long)N * N;
    double *A, *Bm, *C, *C0, *u;
    CK(hipMalloc(&A, B * NN * 8));
    CK(hipMalloc(&Bm, B * NN * 8));
    CK(hipMalloc(&C, B * NN * 8));
    CK(hipMalloc(&C0, B * NN * 8));
    CK(hipMalloc(&u, (long)B * N * 8));
    {
        std::vector<double> h(B * NN);
        srand(1);
        for (auto& x : h) x = rand() / (double)RAND_MAX - 0.5;
        CK(hipMemcpy(A, h.data(), B * NN * 8, hipMemcpyHostToDevice));
        for (auto& x : h) x = rand() / (double)RAND_MAX - 0.5;
        CK(hipMemcpy(Bm, h.data(), B * NN * 8, hipMemcpyHostToDevice));
        std::vector<double> hu((long)B * N);
        for (auto& x : hu) x = 0.5 + rand() / (double)RAND_MAX;
        CK(hipMemcpy(u, hu.data(), (long)B * N * 8, hipMemcpyHostToDevice));
    }
    hipStream_t st;
    CK(hipStreamCreate(&st));
    hipEvent_t e0, e1;
    CK(hipEventCreate(&e0));
    CK(hipEventCreate(&e1));
    GemmArgs<double> g{};
    g.M = g.N = g.K = N;
    g.A = A; g.lda = N; g.sA = NN;
    g.B = Bm; g.ldb = N; g.sB = NN;
    g.C = C; g.ldc = N; g.sC = NN;
    g.u = u; g.su = N;
    g.alpha = 1.0; g.beta = 0.0; g.beta_zero = 1;
    g.adiv = g.bdiv = 1;
    const double flop = 2.0 * N * N * (double)N * B;
    auto timeit = [&](const char* name, auto launch) {
        for (int w = 0; w < 3; w++) launch();
        CK(hipStreamSynchronize(st));
        CK(hipEventRecord(e0, st));
        for (int r = 0; r < reps; r++) launch();
        CK(hipEventRecord(e1, st));
        CK(hipEventSynchronize(e1));
        float ms = 0.f;
        CK(hipEventElapsedTime(&ms, e0, e1));
        const double t = ms / reps;
        printf("%-44s %8.1f us  %6.2f TFLOP/s\n", name, t * 1e3, flop / (t * 1e-3) / 1e12);
        fflush(stdout);
    };
    auto diff = [&]() {
        std::vector<double> a(B * NN), b(B * NN);
        CK(hipMemcpy(a.data(), C, B * NN * 8, hipMemcpyDeviceToHost));
        CK(hipMemcpy(b.data(), C0, B * NN * 8, hipMemcpyDeviceToHost));
        double m = 0, mx = 0;
        for (long i = 0; i < B * NN; i++) {
            m = fmax(m, fabs(a[i] - b[i]));
            mx = fmax(mx, fabs(b[i]));
        }
        printf("    max |C - C(64x64)| / max |C| = %.2e\n", m / mx);
    };
    const struct {
        int tile;
        const char* name;
    } tiles[] = {{GT_64, "k_gemm 64x64, 4x4x4 form (tile 1)"},
                 {GT_128, "k_gemm 128x64 K16, 4x4x4 form (tile 2)"},
                 {GT_128x64, "k_gemm 128x64 K8, 4x4x4 form (tile 3)"},
                 {GT_W, "k_gemm_w 128x64 K16, 16x16x4 form (tile 4)"}};
    for (const auto& t : tiles) {
        timeit(t.name, [&] { CK(gemm_launch<double>(0, 0, g, B, st, t.tile)); });
        if (t.tile == GT_64)
            CK(hipMemcpy(C0, C, B * NN * 8, hipMemcpyDeviceToDevice));
        else
            diff();
    }
    rocblas_handle hdl;
    if (rocblas_create_handle(&hdl) != rocblas_status_success) return 1;
    rocblas_set_stream(hdl, st);
    const double one = 1.0, zero = 0.0;
    timeit("rocblas_dgemm_strided_batched (C = A B)", [&] {
        rocblas_dgemm_strided_batched(hdl, rocblas_operation_none, rocblas_operation_none, N, N, N,
                                      &one, A, N, NN, Bm, N, NN, &zero, C, N, NN, B);
    });
    rocblas_destroy_handle(hdl);
    return 0;
}
