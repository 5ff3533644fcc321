// mp_lu_bench -- mixed-precision sparse LU stage benchmark with the reference
// CLI contract (src/mp/mp_lu_bench.cu:239-495):
//
//   mp_lu_bench <s0.mtx> <s0.rhs> [pin] [flt] [max_n_st] [reps]
//   env DSH_ST_REPS (default 2) repeats of the single-thread kernels
//
// Same pipeline and JSON keys as the reference: s0 read (zero RHS -> 1.0),
// .flt diagonal pin, COLAMD (all knobs zero) -> q, q-order CSC, cpu_lu_ref
// (fp64 oracle x and lnnz/unnz), capacities Lcap = Ucap = max(2 nnz + 64,
// 1.5 lnnz_cpu + 64), then the reference's kernels launched with their own
// shapes through lu_gpu.h / lu_gpu_par.h (<<<1,1>>> single-thread fp64 and
// fp32, <<<1,256,12n+256>>> single-workgroup fp64, <<<1,256>>> trisolves),
// timed as warm-up + min over reps with hipEvents, each checked by rel L2
// against the CPU x and the residual ||Ax-b||/||b||.
//
// Added after the reference's stages (extra JSON keys): the device-resident
// path of luspice.h -- persistent level-ordered refactor with fp32 and fp64
// factor storage, the level-scheduled trisolve, and the fp64 GMRES-IR solve.
// The single-workgroup kernel is skipped (n/a, -1) when 12n + 256 bytes of
// LDS exceed the device limit; the reference fails there.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <utility>
#include <vector>

#include "cpu_ref_lu.h"
#include "lu_gpu.h"
#include "lu_gpu_par.h"
#include "luspice.h"
#include "s0io.h"

#define CK(x)                                                                                \
    do {                                                                                     \
        hipError_t e_ = (x);                                                                 \
        if (e_ != hipSuccess) {                                                              \
            fprintf(stderr, "HIP FAIL %s:%d %s: %s\n", __FILE__, __LINE__, #x,               \
                    hipGetErrorString(e_));                                                  \
            exit(1);                                                                         \
        }                                                                                    \
    } while (0)

static void build_csc_q(const lsp_s0 &M, const int *q, std::vector<int> &Ap, std::vector<int> &Ai,
                        std::vector<double> &Ax)
{
    const int n = M.n;
    std::vector<int> invq(n);
    for (int j = 0; j < n; j++) invq[q ? q[j] : j] = j;
    Ap.assign(n + 1, 0);
    for (long i = 0; i < M.nnz; i++) Ap[invq[M.col[i] - 1] + 1]++;
    for (int i = 0; i < n; i++) Ap[i + 1] += Ap[i];
    Ai.assign(M.nnz, 0);
    Ax.assign(M.nnz, 0.0);
    std::vector<int> cur(Ap.begin(), Ap.end() - 1);
    for (long i = 0; i < M.nnz; i++) {
        const int p = cur[invq[M.col[i] - 1]]++;
        Ai[p] = M.row[i] - 1;
        Ax[p] = M.val[i];
    }
}

static double residual(const lsp_s0 &M, const double *x)
{
    std::vector<double> bh(M.n, 0.0);
    for (long i = 0; i < M.nnz; i++) bh[M.row[i] - 1] += M.val[i] * x[M.col[i] - 1];
    double num = 0, den = 0;
    for (int i = 0; i < M.n; i++) {
        const double d = bh[i] - M.b[i + 1];
        num += d * d;
        den += M.b[i + 1] * M.b[i + 1];
    }
    return std::sqrt(num / den);
}

static double rel_l2(int n, const double *a, const double *b)
{
    double num = 0, den = 0;
    for (int i = 0; i < n; i++) {
        const double d = a[i] - b[i];
        num += d * d;
        den += b[i] * b[i];
    }
    return std::sqrt(num / den);
}

template <typename F>
static double time_best(F &&f, int reps)
{
    hipEvent_t e0, e1;
    CK(hipEventCreate(&e0));
    CK(hipEventCreate(&e1));
    f();
    CK(hipDeviceSynchronize());
    double best = 1e300;
    for (int r = 0; r < reps; r++) {
        CK(hipEventRecord(e0, nullptr));
        f();
        CK(hipEventRecord(e1, nullptr));
        CK(hipDeviceSynchronize());
        float ms = 0;
        CK(hipEventElapsedTime(&ms, e0, e1));
        best = std::min(best, (double)ms);
    }
    CK(hipEventDestroy(e0));
    CK(hipEventDestroy(e1));
    return best;
}

template <typename T>
static T *dmalloc(size_t count)
{
    T *p = nullptr;
    CK(hipMalloc((void **)&p, std::max<size_t>(count, 1) * sizeof(T)));
    return p;
}

int main(int argc, char **argv)
{
    if (argc < 3) {
        fprintf(stderr, "usage: %s <s0.mtx> <s0.rhs> [pin] [flt] [max_n_st] [reps]\n", argv[0]);
        return 2;
    }
    lsp_s0 M;
    if (lsp_s0_read(argv[1], argv[2], &M)) return 1;
    const double pin_val = argc >= 4 ? atof(argv[3]) : 0.0;
    const char *flt = argc >= 5 ? argv[4] : nullptr;
    const int max_n_st = argc >= 6 ? atoi(argv[5]) : 15000;
    int reps = argc >= 7 ? atoi(argv[6]) : 0;
    if (reps <= 0) reps = 5;
    int st_reps = 2;
    if (getenv("DSH_ST_REPS")) st_reps = std::max(1, atoi(getenv("DSH_ST_REPS")));
    printf("== mp_lu_bench %s ==\nn=%d nnz=%ld pin=%.1e flt=%s max_n_st=%d reps=%d\n", argv[1],
           M.n, M.nnz, pin_val, flt ? flt : "-", max_n_st, reps);
    lsp_flt_pin(&M, flt, pin_val);
    const int n = M.n;
    // every entry must index the n x n matrix before it reaches a kernel
    for (long i = 0; i < M.nnz; i++)
        if (M.row[i] < 1 || M.row[i] > n || M.col[i] < 1 || M.col[i] > n) {
            fprintf(stderr, "entry %ld (%d, %d) outside 1..%d\n", i + 1, M.row[i], M.col[i], n);
            return 1;
        }

    // COLAMD on the 0-based pattern
    std::vector<int> r0(M.nnz), c0(M.nnz), q(n);
    for (long i = 0; i < M.nnz; i++) { r0[i] = M.row[i] - 1; c0[i] = M.col[i] - 1; }
    if (lsp_colamd(n, M.nnz, r0.data(), c0.data(), q.data()) != 0) {
        fprintf(stderr, "colamd failed\n");
        return 1;
    }
    printf("colamd ok\n");
    std::vector<int> Ap, Ai, Ap0, Ai0;
    std::vector<double> Ax, Ax0;
    build_csc_q(M, q.data(), Ap, Ai, Ax);
    build_csc_q(M, nullptr, Ap0, Ai0, Ax0);

    std::vector<double> x_cpu(n);
    long lnnz_cpu = 0, unnz_cpu = 0;
    int rc = cpu_lu_ref(n, Ap0.data(), Ai0.data(), Ax0.data(), q.data(), M.b + 1, 1.0, x_cpu.data(),
                        &lnnz_cpu, &unnz_cpu);
    if (rc >= 0) {
        fprintf(stderr, "CPU reference singular at column %d\n", rc);
        return 1;
    }
    printf("CPU ref: lnnz=%ld unnz=%ld (L+U=%ld)\n", lnnz_cpu, unnz_cpu, lnnz_cpu + unnz_cpu);
    long Lcap = 2L * M.nnz + 64, Ucap = 2L * M.nnz + 64;
    if (lnnz_cpu > Lcap) Lcap = (long)(lnnz_cpu * 1.5) + 64;
    if (unnz_cpu > Ucap) Ucap = (long)(unnz_cpu * 1.5) + 64;

    int *d_Ap = dmalloc<int>(n + 1), *d_Ai = dmalloc<int>(M.nnz), *d_diag = dmalloc<int>(n),
        *d_q = dmalloc<int>(n);
    double *d_Ax = dmalloc<double>(M.nnz);
    float *d_Ax32 = dmalloc<float>(M.nnz);
    int *d_Lp = dmalloc<int>(n + 1), *d_Li = dmalloc<int>(Lcap), *d_Up = dmalloc<int>(n + 1),
        *d_Ui = dmalloc<int>(Ucap);
    double *d_Lx = dmalloc<double>(Lcap), *d_Ux = dmalloc<double>(Ucap);
    int *d_Lp32 = dmalloc<int>(n + 1), *d_Up32 = dmalloc<int>(n + 1), *d_Li32 = dmalloc<int>(Lcap),
        *d_Ui32 = dmalloc<int>(Ucap);
    float *d_Lx32 = dmalloc<float>(Lcap), *d_Ux32 = dmalloc<float>(Ucap);
    int *d_pinv = dmalloc<int>(n), *d_pinv32 = dmalloc<int>(n), *d_xi = dmalloc<int>(2 * n + 1),
        *d_xi32 = dmalloc<int>(2 * n + 1), *d_xip = dmalloc<int>(n + 4 * LU_PAR_CAP);
    double *d_x = dmalloc<double>(n), *d_b = dmalloc<double>(n), *d_xo = dmalloc<double>(n),
           *d_wd = dmalloc<double>(2 * n);
    float *d_x32 = dmalloc<float>(n);
    int *d_wi = dmalloc<int>(n), *d_lnnz = dmalloc<int>(1), *d_unnz = dmalloc<int>(1),
        *d_sing = dmalloc<int>(1);
    std::vector<float> Ax32(M.nnz);
    for (long i = 0; i < M.nnz; i++) Ax32[i] = (float)Ax[i];
    CK(hipMemcpy(d_Ap, Ap.data(), (n + 1) * 4, hipMemcpyHostToDevice));
    CK(hipMemcpy(d_Ai, Ai.data(), M.nnz * 4, hipMemcpyHostToDevice));
    CK(hipMemcpy(d_Ax, Ax.data(), M.nnz * 8, hipMemcpyHostToDevice));
    CK(hipMemcpy(d_Ax32, Ax32.data(), M.nnz * 4, hipMemcpyHostToDevice));
    CK(hipMemcpy(d_diag, q.data(), n * 4, hipMemcpyHostToDevice));
    CK(hipMemcpy(d_q, q.data(), n * 4, hipMemcpyHostToDevice));
    CK(hipMemcpy(d_b, M.b + 1, n * 8, hipMemcpyHostToDevice));

    auto read_status = [&](const char *tag) {
        int sing = 0, ln = 0, un = 0;
        CK(hipMemcpy(&sing, d_sing, 4, hipMemcpyDeviceToHost));
        CK(hipMemcpy(&ln, d_lnnz, 4, hipMemcpyDeviceToHost));
        CK(hipMemcpy(&un, d_unnz, 4, hipMemcpyDeviceToHost));
        if (sing == -2) { fprintf(stderr, "[%s] capacity overflow\n", tag); exit(1); }
        if (sing == -3) { fprintf(stderr, "[%s] structure overflow\n", tag); exit(1); }
        if (sing >= 0) { fprintf(stderr, "[%s] singular at column %d\n", tag, sing); exit(1); }
        return std::pair<int, int>(ln, un);
    };
    auto download_xo = [&]() {
        std::vector<double> h(n);
        CK(hipMemcpy(h.data(), d_xo, n * 8, hipMemcpyDeviceToHost));
        return h;
    };
    const bool do_st = n <= max_n_st;
    double t_f_st64 = -1, t_f_st32 = -1, t_f_par64 = -1;
    double t_t_st64 = -1, t_t_st32 = -1, t_t_par64 = -1, t_t_par32 = -1;
    long lnnz64 = -1, unnz64 = -1, lnnz32 = -1, unnz32 = -1;

    if (do_st) {  // 1) fp64 single thread
        auto fac = [&]() {
            lu_factor_kernel<<<1, 1>>>(n, d_Ap, d_Ai, d_Ax, d_diag, d_Lp, d_Li, d_Lx, Lcap, d_Up,
                                       d_Ui, d_Ux, Ucap, d_pinv, d_xi, d_x, d_lnnz, d_unnz, d_sing);
            CK(hipGetLastError());
        };
        auto sol = [&]() {
            trisolve_kernel<<<1, 1>>>(n, d_Lp, d_Li, d_Lx, d_Up, d_Ui, d_Ux, d_pinv, d_q, d_b, d_xo,
                                      d_wi, d_wd);
            CK(hipGetLastError());
        };
        fac();
        auto p = read_status("fp64 ST factor");
        lnnz64 = p.first;
        unnz64 = p.second;
        t_f_st64 = time_best(fac, st_reps);
        t_t_st64 = time_best(sol, reps);
        auto h = download_xo();
        printf("fp64 ST: factor=%.3f ms trisolve=%.3f ms  L2(vs CPU)=%.3e res=%.3e\n", t_f_st64,
               t_t_st64, rel_l2(n, h.data(), x_cpu.data()), residual(M, h.data()));
    }
    if (do_st) {  // 2) fp32 single thread
        auto fac = [&]() {
            lu_factor_f32_kernel<<<1, 1>>>(n, d_Ap, d_Ai, d_Ax32, d_diag, d_Lp32, d_Li32, d_Lx32,
                                           Lcap, d_Up32, d_Ui32, d_Ux32, Ucap, d_pinv32, d_xi32,
                                           d_x32, d_lnnz, d_unnz, d_sing);
            CK(hipGetLastError());
        };
        auto sol = [&]() {
            trisolve_f32_kernel<<<1, 1>>>(n, d_Lp32, d_Li32, d_Lx32, d_Up32, d_Ui32, d_Ux32,
                                          d_pinv32, d_q, d_b, d_xo, d_wi, d_wd);
            CK(hipGetLastError());
        };
        fac();
        auto p = read_status("fp32 ST factor");
        lnnz32 = p.first;
        unnz32 = p.second;
        t_f_st32 = time_best(fac, st_reps);
        t_t_st32 = time_best(sol, reps);
        auto h = download_xo();
        printf("fp32 ST: factor=%.3f ms trisolve=%.3f ms  L2(vs CPU)=%.3e res=%.3e\n", t_f_st32,
               t_t_st32, rel_l2(n, h.data(), x_cpu.data()), residual(M, h.data()));
    }
    {  // 3) fp64 single workgroup
        int dev = 0, maxlds = 0;
        CK(hipGetDevice(&dev));
        CK(hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
        const size_t shb = LU_PAR_SHARED_BYTES(n);
        if (shb <= (size_t)maxlds || shb <= 160 * 1024) {
            if (shb > 65536)
                CK(hipFuncSetAttribute((const void *)lu_factor_par_kernel,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)shb));
            auto fac = [&]() {
                lu_factor_par_kernel<<<1, 256, shb>>>(n, d_Ap, d_Ai, d_Ax, d_diag, d_Lp, d_Li, d_Lx,
                                                      Lcap, d_Up, d_Ui, d_Ux, Ucap, d_pinv, d_xip,
                                                      nullptr, d_lnnz, d_unnz, d_sing);
                CK(hipGetLastError());
            };
            auto sol = [&]() {
                trisolve_par_kernel<<<1, 256>>>(n, d_Lp, d_Li, d_Lx, d_Up, d_Ui, d_Ux, d_pinv, d_q,
                                                d_b, d_xo, d_wi, d_wd);
                CK(hipGetLastError());
            };
            fac();
            auto p = read_status("fp64 PAR factor");
            lnnz64 = p.first;
            unnz64 = p.second;
            t_f_par64 = time_best(fac, reps);
            t_t_par64 = time_best(sol, reps);
            auto h = download_xo();
            printf("fp64 PAR: factor=%.3f ms trisolve=%.3f ms  L2(vs CPU)=%.3e res=%.3e\n",
                   t_f_par64, t_t_par64, rel_l2(n, h.data(), x_cpu.data()), residual(M, h.data()));
        } else {
            printf("fp64 PAR: n/a (needs %zu B of LDS, device limit %d B)\n", shb, maxlds);
        }
    }
    if (t_f_st32 >= 0) {  // 4) fp32 parallel trisolve on the fp32 ST factors
        auto sol = [&]() {
            trisolve_par_f32_kernel<<<1, 256>>>(n, d_Lp32, d_Li32, d_Lx32, d_Up32, d_Ui32, d_Ux32,
                                                d_pinv32, d_q, d_b, d_xo, d_wi, d_wd);
            CK(hipGetLastError());
        };
        t_t_par32 = time_best(sol, reps);
        auto h = download_xo();
        printf("fp32 PAR trisolve: %.3f ms  L2(vs CPU)=%.3e res=%.3e\n", t_t_par32,
               rel_l2(n, h.data(), x_cpu.data()), residual(M, h.data()));
    } else {
        printf("fp32 PAR trisolve: n/a (fp32 ST factor skipped at n>%d)\n", max_n_st);
    }

    // 5) device-resident path (luspice.h)
    double t_rf32 = -1, t_rf64 = -1, t_tri32 = -1, t_ir = -1, ir_err = -1, ir_res = -1;
    int ir_inner = -1;
    {
        std::vector<double> v0(M.nnz);
        for (long i = 0; i < M.nnz; i++) v0[i] = M.val[i];
        lsp_circuit_in cin{n, M.nnz, r0.data(), c0.data(), v0.data(), q.data()};
        lsp_batch *B = nullptr;
        if (lsp_batch_create(&cin, 1, nullptr, &B) == LSP_OK) {
            lsp_refactor_report rr{};
            lsp_solve_report sr{};
            std::vector<double> xs(n);
            double rf = 0, fw = 0, bw = 0;
            t_rf64 = 1e300;
            for (int r = 0; r <= reps; r++)
                if (lsp_batch_refactor(B, LSP_PREC_FP64, &rr) == LSP_OK && r > 0)
                    t_rf64 = std::min(t_rf64, rr.ms);
            t_rf32 = 1e300;
            for (int r = 0; r <= reps; r++)
                if (lsp_batch_refactor(B, LSP_PREC_FP32, &rr) == LSP_OK && r > 0)
                    t_rf32 = std::min(t_rf32, rr.ms);
            t_tri32 = 1e300;
            for (int r = 0; r <= reps; r++) {
                if (lsp_batch_trisolve(B, M.b + 1, xs.data(), 0) == LSP_OK && r > 0) {
                    lsp_batch_last_times(B, &rf, &fw, &bw);
                    t_tri32 = std::min(t_tri32, fw + bw);
                }
            }
            hipEvent_t e0, e1;
            CK(hipEventCreate(&e0));
            CK(hipEventCreate(&e1));
            CK(hipEventRecord(e0, nullptr));
            lsp_batch_solve(B, M.b + 1, xs.data(), 0, &sr);
            CK(hipEventRecord(e1, nullptr));
            CK(hipDeviceSynchronize());
            float ms = 0;
            CK(hipEventElapsedTime(&ms, e0, e1));
            t_ir = ms;
            ir_inner = sr.inner_total;
            ir_res = residual(M, xs.data());
            ir_err = rel_l2(n, xs.data(), x_cpu.data());
            printf("device path: refactor fp64=%.3f ms fp32=%.3f ms, trisolve fp32=%.3f ms, "
                   "GMRES-IR %.3f ms (%d inner) L2(vs CPU)=%.3e res=%.3e\n",
                   t_rf64, t_rf32, t_tri32, t_ir, ir_inner, ir_err, ir_res);
            lsp_batch_destroy(B);
        } else {
            printf("device path: batch creation failed\n");
        }
    }

    const double l64 = lnnz64 > 0 ? lnnz64 : lnnz_cpu, u64 = unnz64 > 0 ? unnz64 : unnz_cpu;
    const double l32 = lnnz32 > 0 ? lnnz32 : lnnz_cpu, u32 = unnz32 > 0 ? unnz32 : unnz_cpu;
    const double mem64 = (l64 + u64) * 12.0, mem32 = (l32 + u32) * 8.0;
    const double ratio_st = (t_f_st64 > 0 && t_f_st32 > 0) ? t_f_st64 / t_f_st32 : -1;
    const double ratio_ts_st = (t_t_st64 > 0 && t_t_st32 > 0) ? t_t_st64 / t_t_st32 : -1;
    const double ratio_ts_par = (t_t_par64 > 0 && t_t_par32 > 0) ? t_t_par64 / t_t_par32 : -1;
    printf("\n== SUMMARY %s n=%d nnz=%ld ==\n", argv[1], n, M.nnz);
    printf("JSON:{\"matrix\":\"%s\",\"n\":%d,\"nnz\":%ld,\"lnnz_cpu\":%ld,\"unnz_cpu\":%ld,"
           "\"lnnz64\":%ld,\"unnz64\":%ld,\"lnnz32\":%ld,\"unnz32\":%ld,"
           "\"mem_fp64_mb\":%.3f,\"mem_fp32_mb\":%.3f,"
           "\"t_factor_st64_ms\":%.3f,\"t_factor_st32_ms\":%.3f,"
           "\"t_factor_par64_ms\":%.3f,"
           "\"t_trisolve_st64_ms\":%.3f,\"t_trisolve_st32_ms\":%.3f,"
           "\"t_trisolve_par64_ms\":%.3f,\"t_trisolve_par32_ms\":%.3f,"
           "\"speedup_factor_st\":%.3f,\"speedup_trisolve_st\":%.3f,\"speedup_trisolve_par\":%.3f,"
           "\"t_refactor_dev64_ms\":%.3f,\"t_refactor_dev32_ms\":%.3f,\"t_trisolve_dev32_ms\":%.3f,"
           "\"t_gmres_ir_ms\":%.3f,\"ir_inner\":%d,\"ir_rel_err\":%.3e,\"ir_res\":%.3e}\n",
           argv[1], n, M.nnz, lnnz_cpu, unnz_cpu, lnnz64, unnz64, lnnz32, unnz32, mem64 / 1e6,
           mem32 / 1e6, t_f_st64, t_f_st32, t_f_par64, t_t_st64, t_t_st32, t_t_par64, t_t_par32,
           ratio_st, ratio_ts_st, ratio_ts_par, t_rf64, t_rf32, t_tri32, t_ir, ir_inner, ir_err,
           ir_res);
    lsp_s0_free(&M);
    return 0;
}
