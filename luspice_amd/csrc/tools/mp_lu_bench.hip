// mp_lu_bench -- stage benchmark of the sparse LU path, with the reference
// CLI's contract (src/mp/mp_lu_bench.cu:239-495): its argv, its stdout lines
// and its JSON keys.
//
//   mp_lu_bench <s0.mtx> <s0.rhs> [pin] [flt] [max_n_st] [reps]
//   env DSH_ST_REPS (default 2): timed repeats of the single-thread factor
//
// Preparation: the s0 files (a zero or missing RHS entry reads as 1.0), the
// .flt diagonal pin, COLAMD (all knobs zero) -> q, the q-order CSC with rows
// in file order, and the CPU reference cpu_lu_ref (fp64 x, lnnz / unnz).
// Factor capacities are max(2 nnz + 64, 1.5 x the CPU count + 64).
//
// Stages: a table.  Each entry names its stdout tag, the factor set it
// writes or reads (fp64 or fp32 storage), the shape of its factor kernel
// (none, <<<1,1>>> or one 256-thread workgroup) and of its solve kernel.  One
// runner gates a stage (single-thread kernels only up to max_n_st, the
// workgroup factor only when its 12n + 256 bytes of LDS fit), launches it
// through the lu_gpu.h / lu_gpu_par.h launchers, times it with HIP events
// (one warm-up call, then the fastest of the repeats) and checks the solve
// against the CPU x (relative L2) and the residual ||A x - b|| / ||b||.
//
// After the table, the device-resident path of luspice.h (extra JSON keys):
// persistent refactor with fp64 and fp32 factor storage, the level-ordered
// trisolve and the fp64 GMRES-IR solve.  Where the reference stops with an
// error (the workgroup kernel's LDS need above the device limit), this tool
// reports the stage as n/a (-1) and goes on.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "cpu_ref_lu.h"
#include "csc.h"
#include "lu_gpu.h"
#include "lu_gpu_par.h"
#include "luspice.h"
#include "s0io.h"

namespace {

[[noreturn]] void die(const char *what)
{
    fprintf(stderr, "mp_lu_bench: %s\n", what);
    exit(1);
}

void hip_ok(hipError_t e, const char *what)
{
    if (e != hipSuccess) {
        fprintf(stderr, "mp_lu_bench: %s: %s\n", what, hipGetErrorString(e));
        exit(1);
    }
}

// ---- the problem and its checks -------------------------------------------
struct Problem {
    lsp_s0 M{};
    int n = 0;
    std::vector<int> row, col, q;          // 0-based COO, column order
    std::vector<int> Ap, Ai;               // q-order CSC (rows: original ids, file order)
    std::vector<double> Ax;
    std::vector<float> Ax32;
    std::vector<double> b, x_cpu;          // b[0..n)
    long lnnz_cpu = 0, unnz_cpu = 0;

    // ||x - x_cpu|| / ||x_cpu||
    double err(const std::vector<double> &x) const
    {
        double d2 = 0, r2 = 0;
        for (int i = 0; i < n; i++) {
            d2 += (x[i] - x_cpu[i]) * (x[i] - x_cpu[i]);
            r2 += x_cpu[i] * x_cpu[i];
        }
        return std::sqrt(d2 / r2);
    }
    // ||A x - b|| / ||b||
    double res(const std::vector<double> &x) const
    {
        std::vector<double> r(b);
        for (size_t e = 0; e < row.size(); e++) r[row[e]] -= M.val[e] * x[col[e]];
        double a = 0, c = 0;
        for (int i = 0; i < n; i++) {
            a += r[i] * r[i];
            c += b[i] * b[i];
        }
        return std::sqrt(a / c);
    }
};

// Device arrays, freed together at exit.
struct Arena {
    std::vector<void *> blocks;
    template <typename T> T *get(size_t count, const void *init = nullptr)
    {
        void *p = nullptr;
        hip_ok(hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T)), "device allocation");
        if (init) hip_ok(hipMemcpy(p, init, count * sizeof(T), hipMemcpyHostToDevice), "upload");
        blocks.push_back(p);
        return (T *)p;
    }
    ~Arena()
    {
        for (void *p : blocks) (void)hipFree(p);
    }
};

// The fastest of `reps` timed calls after one warm-up call.
struct Clock {
    hipEvent_t a{}, b{};
    Clock()
    {
        hip_ok(hipEventCreate(&a), "event");
        hip_ok(hipEventCreate(&b), "event");
    }
    ~Clock()
    {
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
    }
    template <typename F> double fastest(F &&call, int reps)
    {
        call();
        hip_ok(hipDeviceSynchronize(), "warm-up");
        double best = INFINITY;
        for (int r = 0; r < reps; r++) {
            hip_ok(hipEventRecord(a, nullptr), "event");
            call();
            hip_ok(hipEventRecord(b, nullptr), "event");
            hip_ok(hipEventSynchronize(b), "event sync");
            float ms = 0;
            hip_ok(hipEventElapsedTime(&ms, a, b), "event time");
            best = std::min(best, (double)ms);
        }
        return best;
    }
};

// ---- stage table ------------------------------------------------------------
enum Store { F64 = 0, F32 = 1 };
enum Shape { NONE, SINGLE_THREAD, WORKGROUP };

struct FactorSet {
    int *Lp, *Li, *Up, *Ui, *pinv;
    void *Lx, *Ux, *work;                  // values and the factor's T work vector
    long lnnz = -1, unnz = -1;
    bool ready = false;
};

struct Stage {
    const char *tag;
    Store store;
    Shape factor, solve;
    double t_factor = -1, t_solve = -1;
};

struct Run {
    const Problem &P;
    Arena &dev;
    long Lcap, Ucap;
    int max_n_st, reps, st_reps;
    FactorSet set[2];
    int *Ap, *Ai, *diag, *q, *xi, *xi_wg, *wi, *lnnz, *unnz, *sing;
    double *Ax, *b, *xo, *wd;
    float *Ax32;

    void factor(Stage &s)
    {
        FactorSet &F = set[s.store];
        const int n = P.n;
        auto launch = [&]() {
            int rc;
            if (s.factor == WORKGROUP)
                rc = lu_factor_par_launch(n, Ap, Ai, Ax, diag, F.Lp, F.Li, (double *)F.Lx, Lcap, F.Up, F.Ui,
                                          (double *)F.Ux, Ucap, F.pinv, xi_wg, lnnz, unnz, sing, nullptr);
            else if (s.store == F64)
                rc = lu_factor_launch(n, Ap, Ai, Ax, diag, F.Lp, F.Li, (double *)F.Lx, Lcap, F.Up, F.Ui,
                                      (double *)F.Ux, Ucap, F.pinv, xi, (double *)F.work, lnnz, unnz, sing,
                                      nullptr);
            else
                rc = lu_factor_f32_launch(n, Ap, Ai, Ax32, diag, F.Lp, F.Li, (float *)F.Lx, Lcap, F.Up, F.Ui,
                                          (float *)F.Ux, Ucap, F.pinv, xi, (float *)F.work, lnnz, unnz, sing,
                                          nullptr);
            if (rc) die("factor launch failed");
        };
        launch();
        int st[3];
        hip_ok(hipMemcpy(&st[0], sing, 4, hipMemcpyDeviceToHost), "status");
        hip_ok(hipMemcpy(&st[1], lnnz, 4, hipMemcpyDeviceToHost), "status");
        hip_ok(hipMemcpy(&st[2], unnz, 4, hipMemcpyDeviceToHost), "status");
        if (st[0] != -1) {
            fprintf(stderr, "[%s factor] %s %d\n", s.tag,
                    st[0] == -2 ? "capacity overflow" : st[0] == -3 ? "structure overflow" : "singular at column",
                    st[0]);
            exit(1);
        }
        F.lnnz = st[1];
        F.unnz = st[2];
        F.ready = true;
        Clock clk;
        s.t_factor = clk.fastest(launch, s.factor == SINGLE_THREAD ? st_reps : reps);
    }

    std::vector<double> solve(Stage &s)
    {
        const FactorSet &F = set[s.store];
        const int n = P.n;
        const bool wg = s.solve == WORKGROUP;
        auto launch = [&]() {
            const int rc =
                s.store == F64
                    ? (wg ? trisolve_par_launch : trisolve_launch)(n, F.Lp, F.Li, (double *)F.Lx, F.Up, F.Ui,
                                                                   (double *)F.Ux, F.pinv, q, b, xo, wi, wd,
                                                                   nullptr)
                    : trisolve_f32_launch(n, F.Lp, F.Li, (float *)F.Lx, F.Up, F.Ui, (float *)F.Ux, F.pinv, q,
                                          b, xo, wi, wd, wg ? 1 : 0, nullptr);
            if (rc) die("trisolve launch failed");
        };
        Clock clk;
        s.t_solve = clk.fastest(launch, reps);
        std::vector<double> x(n);
        hip_ok(hipMemcpy(x.data(), xo, (size_t)n * 8, hipMemcpyDeviceToHost), "download x");
        return x;
    }

    // Why the stage cannot run here, or nullptr.
    const char *gate(const Stage &s, std::string &why) const
    {
        char buf[160];
        if (s.factor == SINGLE_THREAD && P.n > max_n_st) return "";
        if (s.factor == WORKGROUP) {
            int dev = 0, lim = 0;
            hip_ok(hipGetDevice(&dev), "device");
            hip_ok(hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, dev), "attribute");
            const size_t need = LU_PAR_SHARED_BYTES(P.n);
            if (need > (size_t)lim && need > 160u * 1024u) {
                snprintf(buf, sizeof buf, "n/a (needs %zu B of LDS, device limit %d B)", need, lim);
                why = buf;
                return why.c_str();
            }
        }
        if (s.factor == NONE && !set[s.store].ready) {
            snprintf(buf, sizeof buf, "n/a (fp32 ST factor skipped at n>%d)", max_n_st);
            why = buf;
            return why.c_str();
        }
        return nullptr;
    }

    void run(Stage &s)
    {
        std::string why;
        if (const char *skip = gate(s, why)) {
            if (*skip) printf("%s: %s\n", s.tag, skip);
            return;
        }
        if (s.factor != NONE) factor(s);
        const std::vector<double> x = solve(s);
        if (s.factor != NONE)
            printf("%s: factor=%.3f ms trisolve=%.3f ms  L2(vs CPU)=%.3e res=%.3e\n", s.tag, s.t_factor,
                   s.t_solve, P.err(x), P.res(x));
        else
            printf("%s: %.3f ms  L2(vs CPU)=%.3e res=%.3e\n", s.tag, s.t_solve, P.err(x), P.res(x));
    }
};

// ---- the device-resident path (luspice.h) -----------------------------------
struct DevicePath {
    double rf64 = -1, rf32 = -1, tri32 = -1, ir_ms = -1, err = -1, res = -1;
    int inner = -1;
};

DevicePath device_path(const Problem &P, int reps)
{
    DevicePath d;
    std::vector<double> vals(P.M.val, P.M.val + P.M.nnz);
    lsp_circuit_in cin{P.n, (long)P.row.size(), P.row.data(), P.col.data(), vals.data(), P.q.data()};
    lsp_batch *B = nullptr;
    if (lsp_batch_create(&cin, 1, nullptr, &B) != LSP_OK) {
        printf("device path: batch creation failed\n");
        return d;
    }
    auto best_refactor = [&](int prec) {
        double best = INFINITY;
        lsp_refactor_report rr{};
        for (int r = 0; r <= reps; r++)
            if (lsp_batch_refactor(B, prec, &rr) == LSP_OK && r > 0) best = std::min(best, rr.ms);
        return best;
    };
    d.rf64 = best_refactor(LSP_PREC_FP64);
    d.rf32 = best_refactor(LSP_PREC_FP32);
    std::vector<double> x(P.n);
    d.tri32 = INFINITY;
    for (int r = 0; r <= reps; r++) {
        double rf = 0, fw = 0, bw = 0;
        if (lsp_batch_trisolve(B, P.b.data(), x.data(), 0) == LSP_OK && r > 0) {
            lsp_batch_last_times(B, &rf, &fw, &bw);
            d.tri32 = std::min(d.tri32, fw + bw);
        }
    }
    Clock clk;
    lsp_solve_report sr{};
    hip_ok(hipEventRecord(clk.a, nullptr), "event");
    lsp_batch_solve(B, P.b.data(), x.data(), 0, &sr);
    hip_ok(hipEventRecord(clk.b, nullptr), "event");
    hip_ok(hipEventSynchronize(clk.b), "event sync");
    float ms = 0;
    hip_ok(hipEventElapsedTime(&ms, clk.a, clk.b), "event time");
    d.ir_ms = ms;
    d.inner = sr.inner_total;
    d.err = P.err(x);
    d.res = P.res(x);
    printf("device path: refactor fp64=%.3f ms fp32=%.3f ms, trisolve fp32=%.3f ms, "
           "GMRES-IR %.3f ms (%d inner) L2(vs CPU)=%.3e res=%.3e\n",
           d.rf64, d.rf32, d.tri32, d.ir_ms, d.inner, d.err, d.res);
    lsp_batch_destroy(B);
    return d;
}

// ---- JSON summary -----------------------------------------------------------
struct Json {
    std::string body;
    void add(const char *key, const char *fmt, ...) __attribute__((format(printf, 3, 4)))
    {
        char val[256];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(val, sizeof val, fmt, ap);
        va_end(ap);
        body += body.empty() ? "{" : ",";
        body += "\"";
        body += key;
        body += "\":";
        body += val;
    }
};

double ratio(double a, double b) { return a > 0 && b > 0 ? a / b : -1; }

}  // namespace

int main(int argc, char **argv)
{
    if (argc < 3) {
        fprintf(stderr, "usage: %s <s0.mtx> <s0.rhs> [pin] [flt] [max_n_st] [reps]\n", argv[0]);
        return 2;
    }
    Problem P;
    if (lsp_s0_read(argv[1], argv[2], &P.M)) return 1;
    const double pin = argc > 3 ? atof(argv[3]) : 0.0;
    const char *flt = argc > 4 ? argv[4] : nullptr;
    const int max_n_st = argc > 5 ? atoi(argv[5]) : 15000;
    const int reps = argc > 6 && atoi(argv[6]) > 0 ? atoi(argv[6]) : 5;
    const char *env = getenv("DSH_ST_REPS");
    const int st_reps = env ? std::max(1, atoi(env)) : 2;
    printf("== mp_lu_bench %s ==\nn=%d nnz=%ld pin=%.1e flt=%s max_n_st=%d reps=%d\n", argv[1], P.M.n,
           P.M.nnz, pin, flt ? flt : "-", max_n_st, reps);
    lsp_flt_pin(&P.M, flt, pin);

    // 0-based COO; every entry must index the n x n matrix before a kernel sees it
    P.n = P.M.n;
    const long nnz = P.M.nnz;
    P.row.resize(nnz);
    P.col.resize(nnz);
    for (long e = 0; e < nnz; e++) {
        P.row[e] = P.M.row[e] - 1;
        P.col[e] = P.M.col[e] - 1;
        if (P.row[e] < 0 || P.row[e] >= P.n || P.col[e] < 0 || P.col[e] >= P.n) {
            fprintf(stderr, "mp_lu_bench: entry %ld (%d, %d) lies outside 1..%d\n", e + 1, P.M.row[e],
                    P.M.col[e], P.n);
            return 1;
        }
    }
    P.b.assign(P.M.b + 1, P.M.b + 1 + P.n);
    P.q.resize(P.n);
    if (lsp_colamd(P.n, nnz, P.row.data(), P.col.data(), P.q.data()) != 0) die("colamd failed");
    printf("colamd ok\n");
    std::vector<int> src;
    lsp_coo_to_csc(P.n, nnz, P.row.data(), P.col.data(), P.q.data(), P.Ap, P.Ai, src);
    P.Ax.resize(nnz);
    P.Ax32.resize(nnz);
    for (long p = 0; p < nnz; p++) {
        P.Ax[p] = P.M.val[src[p]];
        P.Ax32[p] = (float)P.Ax[p];
    }

    // the CPU reference takes the natural-order CSC and q
    {
        std::vector<int> Ap0, Ai0, src0;
        lsp_coo_to_csc(P.n, nnz, P.row.data(), P.col.data(), nullptr, Ap0, Ai0, src0);
        std::vector<double> Ax0(nnz);
        for (long p = 0; p < nnz; p++) Ax0[p] = P.M.val[src0[p]];
        P.x_cpu.resize(P.n);
        const int rc = cpu_lu_ref(P.n, Ap0.data(), Ai0.data(), Ax0.data(), P.q.data(), P.b.data(), 1.0,
                                  P.x_cpu.data(), &P.lnnz_cpu, &P.unnz_cpu);
        if (rc >= 0) {
            fprintf(stderr, "mp_lu_bench: CPU reference singular at column %d\n", rc);
            return 1;
        }
    }
    printf("CPU ref: lnnz=%ld unnz=%ld (L+U=%ld)\n", P.lnnz_cpu, P.unnz_cpu, P.lnnz_cpu + P.unnz_cpu);

    Arena dev;
    const int n = P.n;
    const long base = 2L * nnz + 64;
    Run R{P, dev, std::max(base, P.lnnz_cpu > base ? (long)(P.lnnz_cpu * 1.5) + 64 : base),
          std::max(base, P.unnz_cpu > base ? (long)(P.unnz_cpu * 1.5) + 64 : base), max_n_st, reps, st_reps};
    for (int s = 0; s < 2; s++) {
        FactorSet &F = R.set[s];
        const size_t v = s == F64 ? sizeof(double) : sizeof(float);
        F.Lp = dev.get<int>(n + 1);
        F.Up = dev.get<int>(n + 1);
        F.Li = dev.get<int>(R.Lcap);
        F.Ui = dev.get<int>(R.Ucap);
        F.pinv = dev.get<int>(n);
        F.Lx = dev.get<char>(R.Lcap * v);
        F.Ux = dev.get<char>(R.Ucap * v);
        F.work = dev.get<char>((size_t)n * v);
    }
    R.Ap = dev.get<int>(n + 1, P.Ap.data());
    R.Ai = dev.get<int>(nnz, P.Ai.data());
    R.Ax = dev.get<double>(nnz, P.Ax.data());
    R.Ax32 = dev.get<float>(nnz, P.Ax32.data());
    R.diag = dev.get<int>(n, P.q.data());           // diagrow[k] = q[k]
    R.q = dev.get<int>(n, P.q.data());
    R.b = dev.get<double>(n, P.b.data());
    R.xo = dev.get<double>(n);
    R.xi = dev.get<int>(2 * (size_t)n + 1);
    R.xi_wg = dev.get<int>((size_t)n + 4 * LU_PAR_CAP);
    R.wi = dev.get<int>(n);
    R.wd = dev.get<double>(2 * (size_t)n);
    R.lnnz = dev.get<int>(1);
    R.unnz = dev.get<int>(1);
    R.sing = dev.get<int>(1);

    Stage stages[] = {
        {"fp64 ST", F64, SINGLE_THREAD, SINGLE_THREAD},
        {"fp32 ST", F32, SINGLE_THREAD, SINGLE_THREAD},
        {"fp64 PAR", F64, WORKGROUP, WORKGROUP},
        {"fp32 PAR trisolve", F32, NONE, WORKGROUP},
    };
    for (Stage &s : stages) R.run(s);
    const Stage &st64 = stages[0], &st32 = stages[1], &wg64 = stages[2], &wg32 = stages[3];

    const DevicePath d = device_path(P, reps);

    const FactorSet &f64 = R.set[F64], &f32 = R.set[F32];
    const double l64 = f64.lnnz > 0 ? f64.lnnz : P.lnnz_cpu, u64 = f64.unnz > 0 ? f64.unnz : P.unnz_cpu;
    const double l32 = f32.lnnz > 0 ? f32.lnnz : P.lnnz_cpu, u32 = f32.unnz > 0 ? f32.unnz : P.unnz_cpu;
    Json J;
    J.add("matrix", "\"%s\"", argv[1]);
    J.add("n", "%d", n);
    J.add("nnz", "%ld", nnz);
    J.add("lnnz_cpu", "%ld", P.lnnz_cpu);
    J.add("unnz_cpu", "%ld", P.unnz_cpu);
    J.add("lnnz64", "%ld", f64.lnnz);
    J.add("unnz64", "%ld", f64.unnz);
    J.add("lnnz32", "%ld", f32.lnnz);
    J.add("unnz32", "%ld", f32.unnz);
    J.add("mem_fp64_mb", "%.3f", (l64 + u64) * 12.0 / 1e6);
    J.add("mem_fp32_mb", "%.3f", (l32 + u32) * 8.0 / 1e6);
    J.add("t_factor_st64_ms", "%.3f", st64.t_factor);
    J.add("t_factor_st32_ms", "%.3f", st32.t_factor);
    J.add("t_factor_par64_ms", "%.3f", wg64.t_factor);
    J.add("t_trisolve_st64_ms", "%.3f", st64.t_solve);
    J.add("t_trisolve_st32_ms", "%.3f", st32.t_solve);
    J.add("t_trisolve_par64_ms", "%.3f", wg64.t_solve);
    J.add("t_trisolve_par32_ms", "%.3f", wg32.t_solve);
    J.add("speedup_factor_st", "%.3f", ratio(st64.t_factor, st32.t_factor));
    J.add("speedup_trisolve_st", "%.3f", ratio(st64.t_solve, st32.t_solve));
    J.add("speedup_trisolve_par", "%.3f", ratio(wg64.t_solve, wg32.t_solve));
    J.add("t_refactor_dev64_ms", "%.3f", d.rf64);
    J.add("t_refactor_dev32_ms", "%.3f", d.rf32);
    J.add("t_trisolve_dev32_ms", "%.3f", d.tri32);
    J.add("t_gmres_ir_ms", "%.3f", d.ir_ms);
    J.add("ir_inner", "%d", d.inner);
    J.add("ir_rel_err", "%.3e", d.err);
    J.add("ir_res", "%.3e", d.res);
    printf("\n== SUMMARY %s n=%d nnz=%ld ==\nJSON:%s}\n", argv[1], n, nnz, J.body.c_str());
    lsp_s0_free(&P.M);
    return 0;
}
