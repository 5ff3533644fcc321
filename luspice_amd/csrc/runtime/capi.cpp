// capi.cpp -- extern "C" entry points of include/luspice.h and cpu_ref_lu.h.
#include <algorithm>
#include <cstring>
#include <new>
#include <vector>

#include "../host/csc.h"
#include "../host/host_internal.h"
#include "../host/mna.h"
#include "batch.hpp"
#include "cpu_ref_lu.h"
#include "luspice.h"

extern "C" {

const char *lsp_strerror(int code)
{
    switch (code) {
    case LSP_OK: return "ok";
    case LSP_E_ARG: return "invalid argument";
    case LSP_E_ALLOC: return "device allocation failed";
    case LSP_E_HIP: return "HIP runtime error";
    case LSP_E_SINGULAR: return "singular matrix";
    case LSP_E_STRUCTURE: return "inconsistent symbolic structure";
    case LSP_E_DEVICE: return "device dependency wait gave up";
    case LSP_E_CAPACITY: return "column exceeds scratch capacity";
    case LSP_E_NOT_CONVERGED: return "iterative refinement did not converge";
    default: return "unknown error";
    }
}

int lsp_options_default(lsp_options *o)
{
    if (!o) return LSP_E_ARG;
    std::memset(o, 0, sizeof(*o));
    o->pivot_tol = 1.0;
    o->refactor_ptol = 1e-3;
    o->ir_outer = 7;
    o->ir_inner = 50;
    o->ir_inner_tol = 1e-10;
    o->ir_outer_tol = 1e-10;
    o->lds_cap = 1024;
    o->blocks_per_cu = 0;
    o->chunk_rows = 0;
    o->sell_sigma = 0;
    o->fallback_fp64 = 1;
    o->repivot = 1;
    return LSP_OK;
}

int lsp_batch_create(const lsp_circuit_in *circuits, int ncircuits, const lsp_options *opt,
                     lsp_batch **out)
{
    if (!out) return LSP_E_ARG;
    *out = nullptr;
    lsp_batch *b = new (std::nothrow) lsp_batch();
    if (!b) return LSP_E_ALLOC;
    int rc = b->impl.create(circuits, ncircuits, opt);
    if (rc) {
        delete b;
        return rc;
    }
    *out = b;
    return LSP_OK;
}

void lsp_batch_destroy(lsp_batch *b) { delete b; }

int lsp_batch_adopt(lsp_batch *const *parts, int nparts, const int *part_of, const int *circuit_of,
                    int ncircuits, lsp_batch **out)
{
    int rc = LSP_OK;
    if (!out || !parts || nparts <= 0 || !part_of || !circuit_of || ncircuits <= 0) rc = LSP_E_ARG;
    if (out) *out = nullptr;
    std::vector<std::pair<lsp::Batch *, int>> picks;
    for (int i = 0; !rc && i < ncircuits; i++) {
        if (part_of[i] < 0 || part_of[i] >= nparts || !parts[part_of[i]]) rc = LSP_E_ARG;
        else picks.emplace_back(&parts[part_of[i]]->impl, circuit_of[i]);
    }
    lsp_batch *b = nullptr;
    if (!rc) {
        b = new (std::nothrow) lsp_batch();
        if (!b) rc = LSP_E_ALLOC;
    }
    if (!rc) rc = b->impl.adopt(picks);
    // the sources go whatever the result (each handle once)
    if (parts)
        for (int p = 0; p < nparts; p++) {
            bool seen = false;
            for (int r = 0; r < p && !seen; r++) seen = parts[r] == parts[p];
            if (!seen) delete parts[p];
        }
    if (rc) {
        delete b;
        return rc;
    }
    *out = b;
    return LSP_OK;
}

int lsp_batch_set_active(lsp_batch *b, const int *active)
{
    return b ? b->impl.set_active(active) : LSP_E_ARG;
}

int lsp_batch_set_stream(lsp_batch *b, void *s)
{
    return b ? b->impl.set_stream((hipStream_t)s) : LSP_E_ARG;
}

int lsp_batch_info(const lsp_batch *b, int c, lsp_circuit_info *inf)
{
    return b ? b->impl.info(c, inf) : LSP_E_ARG;
}

int lsp_batch_sizes(const lsp_batch *b, long *n, long *nnz)
{
    if (!b) return LSP_E_ARG;
    if (n) *n = b->impl.total_n();
    if (nnz) *nnz = b->impl.total_nnz();
    return LSP_OK;
}

int lsp_batch_set_values(lsp_batch *b, const double *vals, int on_device)
{
    return b ? b->impl.set_values(vals, on_device != 0) : LSP_E_ARG;
}

int lsp_batch_refactor(lsp_batch *b, int prec, lsp_refactor_report *rep)
{
    return b ? b->impl.refactor(prec, rep) : LSP_E_ARG;
}

int lsp_batch_trisolve(lsp_batch *b, const double *rhs, double *x, int on_device)
{
    return b ? b->impl.trisolve(rhs, x, on_device != 0) : LSP_E_ARG;
}

int lsp_batch_pivot_range(lsp_batch *b, int circuit, double *min_abs, double *max_abs)
{
    return b ? b->impl.diag_stats(circuit, min_abs, max_abs) : LSP_E_ARG;
}

int lsp_ruiz(int n, long nnz, const int *row, const int *col, const double *val, int sweeps,
             double *R, double *C)
{
    if (n <= 0 || nnz < 0 || sweeps < 1 || !row || !col || !val || !R || !C) return LSP_E_ARG;
    for (long e = 0; e < nnz; e++)
        if (row[e] < 0 || row[e] >= n || col[e] < 0 || col[e] >= n) return LSP_E_ARG;
    lsp_ruiz_scales(n, nnz, row, col, val, sweeps, R, C);
    return LSP_OK;
}

int lsp_batch_ir_step(lsp_batch *b, const double *rhs, double *x, int on_device, double *rel_res)
{
    return b && rhs && x ? b->impl.ir_step(rhs, x, on_device != 0, rel_res) : LSP_E_ARG;
}

int lsp_batch_solve(lsp_batch *b, const double *rhs, double *x, int on_device,
                    lsp_solve_report *rep)
{
    return b ? b->impl.solve(rhs, x, on_device != 0, rep) : LSP_E_ARG;
}

int lsp_batch_spmv(lsp_batch *b, const double *x, double *y, int on_device, int prec)
{
    return b ? b->impl.spmv(x, y, on_device != 0, prec) : LSP_E_ARG;
}

int lsp_batch_export_factors(lsp_batch *b, int c, int *pinv, int *Lp, int *Li, double *Lx,
                             int *Up, int *Ui, double *Ux)
{
    return b ? b->impl.export_factors(c, pinv, Lp, Li, Lx, Up, Ui, Ux) : LSP_E_ARG;
}

int lsp_batch_graph_step(lsp_batch *b, int prec, lsp_refactor_report *rrep, lsp_solve_report *srep)
{
    return b ? b->impl.graph_step(prec, rrep, srep) : LSP_E_ARG;
}

int lsp_batch_device_vectors(lsp_batch *b, double **rhs, double **x)
{
    return b ? b->impl.device_vectors(rhs, x) : LSP_E_ARG;
}

int lsp_batch_set_rhs(lsp_batch *b, const double *rhs)
{
    return b && rhs ? b->impl.copy_vector(rhs, true) : LSP_E_ARG;
}

int lsp_batch_get_x(lsp_batch *b, double *x)
{
    return b && x ? b->impl.copy_vector(x, false) : LSP_E_ARG;
}

int lsp_batch_prepare(const lsp_circuit_in *circuits, int ncircuits, const lsp_options *opt,
                      lsp_batch **out)
{
    if (!out) return LSP_E_ARG;
    *out = nullptr;
    lsp_batch *b = new (std::nothrow) lsp_batch();
    if (!b) return LSP_E_ALLOC;
    int rc = b->impl.prepare(circuits, ncircuits, opt);
    if (rc) {
        delete b;
        return rc;
    }
    *out = b;
    return LSP_OK;
}

int lsp_batch_estimate_bytes(const lsp_batch *b, size_t *bytes)
{
    if (!b || !bytes) return LSP_E_ARG;
    *bytes = b->impl.estimate_device_bytes();
    return LSP_OK;
}

int lsp_batch_upload(lsp_batch *b)
{
    return b ? b->impl.upload_all() : LSP_E_ARG;
}

int lsp_batch_device_bytes(const lsp_batch *b, size_t *bytes)
{
    if (!b || !bytes) return LSP_E_ARG;
    *bytes = b->impl.device_bytes();
    return LSP_OK;
}

int lsp_dev_alloc(void **ptr, size_t bytes)
{
    if (!ptr) return LSP_E_ARG;
    *ptr = nullptr;
    return hipMalloc(ptr, bytes ? bytes : 16) == hipSuccess ? LSP_OK : LSP_E_ALLOC;
}

int lsp_dev_free(void *ptr)
{
    return hipFree(ptr) == hipSuccess ? LSP_OK : LSP_E_HIP;
}

int lsp_dev_copy(void *dst, const void *src, size_t bytes)
{
    if (!dst || !src) return LSP_E_ARG;
    return hipMemcpy(dst, src, bytes, hipMemcpyDefault) == hipSuccess ? LSP_OK : LSP_E_HIP;
}

int lsp_dev_sync(void)
{
    return hipDeviceSynchronize() == hipSuccess ? LSP_OK : LSP_E_HIP;
}

int lsp_dev_mem_info(size_t *free_bytes, size_t *total_bytes)
{
    size_t f = 0, t = 0;
    if (hipMemGetInfo(&f, &t) != hipSuccess) return LSP_E_HIP;
    if (free_bytes) *free_bytes = f;
    if (total_bytes) *total_bytes = t;
    return LSP_OK;
}

int lsp_batch_last_times(const lsp_batch *b, double *rf, double *fw, double *bw)
{
    return b ? b->impl.last_times(rf, fw, bw) : LSP_E_ARG;
}

const char *lsp_batch_last_error(const lsp_batch *b)
{
    return b ? b->impl.last_error.c_str() : "";
}

int lsp_batch_kernel_stats(lsp_batch *b, lsp_kernel_stats *st, int reset)
{
    return b ? b->impl.kernel_stats(st, reset) : LSP_E_ARG;
}

// ---------------------------------------------------------------------------
int lsp_colamd(int n, long nnz, const int *row, const int *col, int *q)
{
    if (n < 0 || nnz < 0 || (nnz && (!row || !col)) || !q) return LSP_E_ARG;
    for (long e = 0; e < nnz; e++)
        if (row[e] < 0 || row[e] >= n || col[e] < 0 || col[e] >= n) return LSP_E_ARG;
    return lsp_colamd_coo(n, nnz, row, col, q) == 0 ? LSP_OK : LSP_E_STRUCTURE;
}

int lsp_symbolic_factor(int n, long nnz, const int *row, const int *col, const double *val,
                        const int *q, double tol, int *pinv, int *Lp, int *Li, double *Lx,
                        long lcap, int *Up, int *Ui, double *Ux, long ucap, long *lnnz,
                        long *unnz)
{
    if (n < 1 || !row || !col || !val) return -2;
    std::vector<int> qq(n);
    for (int k = 0; k < n; k++) qq[k] = q ? q[k] : k;
    std::vector<int> Ap, Ai, src;
    lsp_coo_to_csc(n, nnz, row, col, qq.data(), Ap, Ai, src);
    std::vector<double> Ax(nnz);
    for (long p = 0; p < nnz; p++) Ax[p] = val[src[p]];
    LspGpLu F;
    int rc = lsp_gp_factor(n, Ap.data(), Ai.data(), Ax.data(), qq.data(), tol, F);
    if (rc != -1) return rc;
    if (lnnz) *lnnz = F.lnnz;
    if (unnz) *unnz = F.unnz;
    if (F.lnnz > lcap || F.unnz > ucap) return -3;
    if (pinv) std::copy(F.pinv.begin(), F.pinv.end(), pinv);
    if (Lp) std::copy(F.Lp.begin(), F.Lp.end(), Lp);
    if (Li) std::copy(F.Li.begin(), F.Li.end(), Li);
    if (Lx) std::copy(F.Lx.begin(), F.Lx.end(), Lx);
    if (Up) std::copy(F.Up.begin(), F.Up.end(), Up);
    if (Ui) std::copy(F.Ui.begin(), F.Ui.end(), Ui);
    if (Ux) std::copy(F.Ux.begin(), F.Ux.end(), Ux);
    return -1;
}

int lsp_mna_generate(int kind, int n_req, double lr, uint64_t seed, int *n, long *nnz, int *row,
                     int *col, double *val, double *rhs)
{
    LspMnaSpec sp;
    if (kind == 0) lsp_spec_c0(&sp, seed);
    else if (kind == 1) lsp_spec_c1(&sp, seed);
    else if (kind == 2) {
        if (n_req < 3) return LSP_E_ARG;
        lsp_spec_sized(&sp, n_req, lr, seed);
    } else if (kind == 3) {
        if (n_req < 3) return LSP_E_ARG;
        lsp_spec_c2(&sp, n_req, seed);
    } else {
        return LSP_E_ARG;
    }
    LspCoo A;
    if (lsp_mna_build(&sp, A) != 0) return LSP_E_ARG;
    if (n) *n = A.n;
    if (nnz) *nnz = (long)A.val.size();
    if (row) std::copy(A.row.begin(), A.row.end(), row);
    if (col) std::copy(A.col.begin(), A.col.end(), col);
    if (val) std::copy(A.val.begin(), A.val.end(), val);
    if (rhs) std::copy(A.b.begin(), A.b.end(), rhs);
    return LSP_OK;
}

void lsp_mna_perturb_values(uint64_t seed, int step, double sigma, long nnz, const double *val0,
                            double *val, int n, double *rhs)
{
    lsp_mna_perturb(seed, step, sigma, (int)nnz, val0, val, n, rhs);
}

// cpu_ref_lu.h ---------------------------------------------------------------
int cpu_lu_ref(int n, const int *Ap, const int *Ai, const double *Ax, const int *q,
               const double *b, double tol, double *x, long *lnnz, long *unnz)
{
    if (n < 1) return -2;
    // q-order CSC view of the natural-order input (columns permuted, rows kept
    // in their stored order, as lu_factor reads A(:,q[k]))
    std::vector<int> Aq(n + 1, 0), Aiq((size_t)Ap[n]);
    std::vector<double> Axq((size_t)Ap[n]);
    std::vector<int> qq(n);
    for (int k = 0; k < n; k++) qq[k] = q ? q[k] : k;
    for (int k = 0; k < n; k++) {
        const int c = qq[k];
        Aq[k + 1] = Aq[k] + (Ap[c + 1] - Ap[c]);
        std::copy(Ai + Ap[c], Ai + Ap[c + 1], Aiq.begin() + Aq[k]);
        std::copy(Ax + Ap[c], Ax + Ap[c + 1], Axq.begin() + Aq[k]);
    }
    LspGpLu F;
    int rc = lsp_gp_factor(n, Aq.data(), Aiq.data(), Axq.data(), qq.data(), tol, F);
    if (rc != -1) return rc;
    lsp_gp_solve(F, qq.data(), b, x);
    if (lnnz) *lnnz = F.lnnz;
    if (unnz) *unnz = F.unnz;
    return -1;
}

}  // extern "C"
