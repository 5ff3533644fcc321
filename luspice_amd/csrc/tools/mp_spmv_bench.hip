// mp_spmv_bench -- fp64 vs fp32-value SpMV on one s0 matrix, with the
// reference CLI contract (src/mp/mp_spmv_bench.cu:211-341):
//
//   mp_spmv_bench <s0.mtx> [iters=200] [groups=3]     (iters >= 10, groups >= 1)
//
// Same stages and JSON keys as the reference: x = ones; the fp32-vs-fp64
// check is max|y32 - y64| / max|y64| on the CSR path; 10 warm-ups of each
// kernel, then `groups` groups of `iters` back-to-back launches per kernel,
// each group timed with hipEvents, minimum over groups.  The reference's
// bandwidth figure 2*nnz*bytes/t (values + x, :300) is kept under its keys.
//
// The kernels are this library's:
//   csr_fp64 / csr_fp32  one thread per row, 256-thread blocks, fp64 x and
//                        fp64 sum (fp32 only narrows the stored values; the
//                        reference's k_csr32 also narrows x and y);
//   sell_fp64 / sell_fp32  SELL-C-64 (one wave per 64-row slice, column-major
//                        slots, whole-matrix length sort), fp64 x and sum.
// Added keys: the algorithmic bytes of one product,
//   B = nnz*(v + 4) + 4*(n+1) + 8n (x) + 8n (y)   (CSR; SELL reads the row
//   lengths, n*4, in place of the row pointers), and that figure over the
//   time as achieved GB/s against the 8000 GB/s HBM peak of MI355X.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "host_internal.h"
#include "launch.hpp"
#include "spmv.h"

#define CK(x)                                                                                \
    do {                                                                                     \
        hipError_t e_ = (x);                                                                 \
        if (e_ != hipSuccess) {                                                              \
            fprintf(stderr, "HIP FAIL %s:%d %s: %s\n", __FILE__, __LINE__, #x,               \
                    hipGetErrorString(e_));                                                  \
            exit(1);                                                                         \
        }                                                                                    \
    } while (0)

static const double kPeakGBs = 8000.0;

template <typename X>
static X *up(const std::vector<X> &v)
{
    X *d = nullptr;
    CK(hipMalloc((void **)&d, std::max<size_t>(1, v.size()) * sizeof(X)));
    if (!v.empty()) CK(hipMemcpy(d, v.data(), v.size() * sizeof(X), hipMemcpyHostToDevice));
    return d;
}

template <typename F>
static double group_ms(F &&f, int iters)
{
    hipEvent_t e0, e1;
    CK(hipEventCreate(&e0));
    CK(hipEventCreate(&e1));
    CK(hipEventRecord(e0, 0));
    for (int i = 0; i < iters; i++) f();
    CK(hipEventRecord(e1, 0));
    CK(hipEventSynchronize(e1));
    float ms = 0.0f;
    CK(hipEventElapsedTime(&ms, e0, e1));
    CK(hipEventDestroy(e0));
    CK(hipEventDestroy(e1));
    return (double)ms / iters;
}

int main(int argc, char **argv)
{
    if (argc < 2) {
        fprintf(stderr, "usage: %s <s0.mtx> [iters] [groups]\n", argv[0]);
        return 2;
    }
    CSRMatrix A{};
    if (read_s0_to_csr(argv[1], &A)) {
        fprintf(stderr, "read fail %s\n", argv[1]);
        return 1;
    }
    int iters = argc >= 3 ? atoi(argv[2]) : 200;
    int groups = argc >= 4 ? atoi(argv[3]) : 3;
    if (iters < 10) iters = 10;
    if (groups < 1) groups = 1;
    const int n = A.n;
    const long long nnz = A.nnz;
    if (n <= 0) {
        fprintf(stderr, "empty matrix %s\n", argv[1]);
        return 1;
    }

    // the reader skips rows outside 1..n and columns below 1 only: every
    // column must index x before any kernel runs
    for (long long e = 0; e < nnz; e++)
        if (A.colidx[e] < 0 || A.colidx[e] >= n) {
            fprintf(stderr, "column %d outside 1..%d in %s\n", A.colidx[e] + 1, n, argv[1]);
            return 1;
        }
    // CSR, fp64 and fp32 values
    std::vector<int> rp(A.rowptr, A.rowptr + n + 1), ci(A.colidx, A.colidx + nnz);
    std::vector<double> v64(A.values, A.values + nnz);
    std::vector<float> v32(nnz);
    for (long long e = 0; e < nnz; e++) v32[e] = (float)v64[e];
    int *d_rp = up(rp), *d_ci = up(ci);
    double *d_v64 = up(v64);
    float *d_v32 = up(v32);

    // SELL-C-64, rows sorted by length over the whole matrix
    LspSell S;
    if (lsp_build_sell(n, rp.data(), ci.data(), nullptr, 64, 0, S)) {
        fprintf(stderr, "sell build failed\n");
        return 1;
    }
    std::vector<int> rlen(n);
    for (int k = 0; k < n; k++) rlen[k] = rp[S.perm[k] + 1] - rp[S.perm[k]];
    std::vector<double> sv64(S.slots, 0.0);
    std::vector<float> sv32(S.slots, 0.0f);
    for (int64_t t = 0; t < S.slots; t++)
        if (S.src[t] >= 0) { sv64[t] = v64[S.src[t]]; sv32[t] = (float)v64[S.src[t]]; }
    int64_t *d_off = up(S.off);
    int *d_rlen = up(rlen), *d_scol = up(S.col), *d_perm = up(S.perm);
    double *d_sv64 = up(sv64);
    float *d_sv32 = up(sv32);
    lsp::SellDesc<double> D64{d_off, d_rlen, d_scol, d_sv64, d_perm, S.nslices, n};
    lsp::SellDesc<float> D32{d_off, d_rlen, d_scol, d_sv32, d_perm, S.nslices, n};
    std::vector<lsp::SellDesc<double>> h64{D64};
    std::vector<lsp::SellDesc<float>> h32{D32};
    lsp::SellDesc<double> *d_D64 = up(h64);
    lsp::SellDesc<float> *d_D32 = up(h32);
    std::vector<int2> tasks(S.nslices);
    for (int s = 0; s < S.nslices; s++) tasks[s] = make_int2(0, s);
    int2 *d_tasks = up(tasks);
    std::vector<int64_t> voff{0};
    int64_t *d_voff = up(voff);

    std::vector<double> x(n, 1.0);
    double *d_x = up(x), *d_y64 = nullptr, *d_y32 = nullptr;
    CK(hipMalloc((void **)&d_y64, n * sizeof(double)));
    CK(hipMalloc((void **)&d_y32, n * sizeof(double)));

    auto csr64 = [&]() { CK(lsp::launch_csr_spmv(n, d_rp, d_ci, d_v64, d_x, d_y64, 0)); };
    auto csr32 = [&]() { CK(lsp::launch_csr_spmv_f32(n, d_rp, d_ci, d_v32, d_x, d_y32, 0)); };
    auto sel64 = [&]() {
        CK(lsp::launch_sell_spmv<double>(d_D64, d_tasks, S.nslices, d_voff, d_x, d_y64, nullptr, nullptr, 0));
    };
    auto sel32 = [&]() {
        CK(lsp::launch_sell_spmv<float>(d_D32, d_tasks, S.nslices, d_voff, d_x, d_y32, nullptr, nullptr, 0));
    };

    // fp32 vs fp64 on the CSR path, and SELL against CSR (fp64: same per-row order)
    csr64();
    csr32();
    CK(hipDeviceSynchronize());
    std::vector<double> y64(n), y32(n), ys(n);
    CK(hipMemcpy(y64.data(), d_y64, n * sizeof(double), hipMemcpyDeviceToHost));
    CK(hipMemcpy(y32.data(), d_y32, n * sizeof(double), hipMemcpyDeviceToHost));
    sel64();
    CK(hipDeviceSynchronize());
    CK(hipMemcpy(ys.data(), d_y64, n * sizeof(double), hipMemcpyDeviceToHost));
    double m2 = 0.0, me = 0.0;
    long sell_diff = 0;
    for (int i = 0; i < n; i++) {
        m2 = std::max(m2, std::fabs(y64[i]));
        me = std::max(me, std::fabs(y32[i] - y64[i]));
        if (ys[i] != y64[i]) sell_diff++;
    }
    const double rel_err = m2 > 0 ? me / m2 : 0.0;

    for (int i = 0; i < 10; i++) { csr64(); csr32(); sel64(); sel32(); }
    CK(hipDeviceSynchronize());
    double t_c64 = 1e30, t_c32 = 1e30, t_s64 = 1e30, t_s32 = 1e30;
    for (int g = 0; g < groups; g++) {
        t_c64 = std::min(t_c64, group_ms(csr64, iters));
        t_c32 = std::min(t_c32, group_ms(csr32, iters));
        t_s64 = std::min(t_s64, group_ms(sel64, iters));
        t_s32 = std::min(t_s32, group_ms(sel32, iters));
    }
    auto bw_ref = [&](double t, int bytes) { return 2.0 * (double)nnz * bytes / (t * 1e-3) / 1e9; };
    auto alg = [&](int vbytes, bool sell) {
        return (double)nnz * (vbytes + 4) + (sell ? 4.0 * n : 4.0 * (n + 1)) + 16.0 * n;
    };
    auto gbs = [&](double bytes, double t) { return bytes / (t * 1e-3) / 1e9; };
    const double a_c64 = gbs(alg(8, false), t_c64), a_c32 = gbs(alg(4, false), t_c32);
    const double a_s64 = gbs(alg(8, true), t_s64), a_s32 = gbs(alg(4, true), t_s32);

    char name[256] = "matrix";
    {
        const char *b = strrchr(argv[1], '/');
        snprintf(name, sizeof name, "%s", b ? b + 1 : argv[1]);
        char *dot = strrchr(name, '.');
        if (dot) *dot = '\0';
    }
    const double pad = 100.0 * ((double)S.slots - (double)nnz) / std::max(1.0, (double)nnz);
    printf("== mp_spmv_bench %s n=%d nnz=%lld (iters=%d groups=%d, sell stored=%lld pad=%.1f%%) ==\n",
           name, n, nnz, iters, groups, (long long)S.slots, pad);
    printf("%-10s %10s %9s %9s %9s %10s\n", "format", "t_ms", "bw_gbs", "alg_gbs", "peak%", "vs_fp64");
    printf("%-10s %10.6f %9.1f %9.1f %9.2f %10s\n", "csr_fp64", t_c64, bw_ref(t_c64, 8), a_c64,
           100 * a_c64 / kPeakGBs, "1.00x");
    printf("%-10s %10.6f %9.1f %9.1f %9.2f %9.2fx\n", "csr_fp32", t_c32, bw_ref(t_c32, 4), a_c32,
           100 * a_c32 / kPeakGBs, t_c64 / t_c32);
    printf("%-10s %10.6f %9.1f %9.1f %9.2f %10s\n", "sell_fp64", t_s64, bw_ref(t_s64, 8), a_s64,
           100 * a_s64 / kPeakGBs, "1.00x");
    printf("%-10s %10.6f %9.1f %9.1f %9.2f %9.2fx\n", "sell_fp32", t_s32, bw_ref(t_s32, 4), a_s32,
           100 * a_s32 / kPeakGBs, t_s64 / t_s32);
    printf("rel_err(fp32 vs fp64) = %.3e, sell vs csr fp64 rows differing: %ld\n", rel_err, sell_diff);
    printf("JSON:{\"matrix\":\"%s\",\"n\":%d,\"nnz\":%lld,\"iters\":%d,\"groups\":%d,"
           "\"t_csr_fp64_ms\":%.6f,\"t_csr_fp32_ms\":%.6f,\"t_sell_fp64_ms\":%.6f,\"t_sell_fp32_ms\":%.6f,"
           "\"bw_csr_fp64_gbs\":%.2f,\"bw_csr_fp32_gbs\":%.2f,\"bw_sell_fp64_gbs\":%.2f,\"bw_sell_fp32_gbs\":%.2f,"
           "\"speedup_csr_fp32\":%.3f,\"speedup_sell_fp32\":%.3f,\"rel_err_fp32\":%.3e,"
           "\"sell_stored\":%lld,\"sell_pad_pct\":%.1f,"
           "\"alg_gbs_csr_fp64\":%.2f,\"alg_gbs_csr_fp32\":%.2f,\"alg_gbs_sell_fp64\":%.2f,"
           "\"alg_gbs_sell_fp32\":%.2f,\"peak_gbs\":%.0f,\"sell_csr_rows_differing\":%ld}\n",
           name, n, nnz, iters, groups, t_c64, t_c32, t_s64, t_s32, bw_ref(t_c64, 8), bw_ref(t_c32, 4),
           bw_ref(t_s64, 8), bw_ref(t_s32, 4), t_c64 / t_c32, t_s64 / t_s32, rel_err, (long long)S.slots, pad,
           a_c64, a_c32, a_s64, a_s32, kPeakGBs, sell_diff);
    fflush(stdout);

    hipFree(d_rp); hipFree(d_ci); hipFree(d_v64); hipFree(d_v32);
    hipFree(d_off); hipFree(d_rlen); hipFree(d_scol); hipFree(d_perm); hipFree(d_sv64); hipFree(d_sv32);
    hipFree(d_D64); hipFree(d_D32); hipFree(d_tasks); hipFree(d_voff);
    hipFree(d_x); hipFree(d_y64); hipFree(d_y32);
    csr_free(&A);
    return sell_diff ? 3 : 0;
}
