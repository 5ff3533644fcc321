// lu_static.cpp -- the reference's static-LU lifecycle (include/lu_static.h)
// on the device-resident batch runtime.
//
// Reference behaviour kept (src/lu_static.c):
//   * dsh_symbolic reads the COO pattern with 1-based row/col (coo_to_csc,
//     :65-97), keeps rows sorted inside each column, stores the column order
//     (NULL = natural) and counts calls process-wide;
//   * dsh_numeric rebuilds the values from COO (a later duplicate of an entry
//     overwrites an earlier one, as splu's scatter does), adds pin to the
//     diagonal entries present in the pattern (:199-209), factors in fp64 with
//     splu's partial pivoting (tol = 1.0) -- P equals splu's on every call --
//     and fills min_diag_U (:230-240), lnnz, unnz and cap;
//   * F->factor_cache points at the factors in splu's lu_factors layout
//     (splu.h:29-41): the reference's own caller reads U.Up / U.Ui / U.Ux and
//     L.Lp / L.Li / L.Lx through it (bench/test_lu.c:150-171);
//   * dsh_trisolve is x = Q U^-1 L^-1 P b with the fp64 factors
//     (lu_solve_perm, :275-281), so ||b - A x|| / ||b|| is at fp64 level;
//   * return codes 0 OK, 1 SINGULAR, 2 ALLOC, 3 ARG / NOT_IMPL and their texts.
// What changes: the factorization and the solves run on the GPU; dsh_numeric
// after the first call is a refactorization over the kept structure, and the
// pivot sequence is re-derived on the host only when splu would pick another
// pivot on the new values; dsh_gmres_ir is implemented (fp32 factors as the
// right preconditioner of an fp64 GMRES-IR; the reference's returns
// NOT_IMPL, :283-290).
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <vector>

#include "batch.hpp"
#include "lu_static.h"

namespace {

std::atomic<long> g_symbolic_count{0};

// splu's lu_csc / lu_factors, field for field (ints, pointers and longs in
// the same order), so a caller compiled against splu.h reads the factors
// through factor_cache.  L uses the lp/li/lx fields, U the up/ui/ux fields.
struct SpluCsc {
    int n;
    int *lp, *li;
    double *lx;
    int *up, *ui;
    double *ux;
    long lcap, ucap, lnnz, unnz;
};
struct SpluFactors {
    int n;
    int *pinv;
    SpluCsc L, U;
    int singular_col;
    void *owner;                           // past splu's layout: the StaticState
};

struct StaticState {
    std::vector<int> Ap, Ai;               // pattern, 0-based rows sorted per column
    std::vector<int> q;
    std::vector<int> row0, col0;           // pattern in CSC order (0-based COO)
    std::vector<double> vals;              // values in CSC order
    std::unique_ptr<lsp_batch> batch;
    double max_abs_a = 0;
    SpluFactors *fac = nullptr;            // what factor_cache points at
};

StaticState *state(const dsh_lu_static *F)
{
    const SpluFactors *f = (const SpluFactors *)F->factor_cache;
    return f ? (StaticState *)f->owner : nullptr;
}

void free_factors(SpluFactors *f)
{
    if (!f) return;
    std::free(f->pinv);
    std::free(f->L.lp); std::free(f->L.li); std::free(f->L.lx);
    std::free(f->U.up); std::free(f->U.ui); std::free(f->U.ux);
    std::memset(&f->L, 0, sizeof(f->L));
    std::memset(&f->U, 0, sizeof(f->U));
    f->pinv = nullptr;
}

int csc_slot(const StaticState &S, int r, int c)
{
    const int *b = S.Ai.data() + S.Ap[c], *e = S.Ai.data() + S.Ap[c + 1];
    const int *it = std::lower_bound(b, e, r);
    return (it != e && *it == r) ? (int)(it - S.Ai.data()) : -1;
}

// The first numeric call: host symbolic on these values (splu's pivots) and
// upload.  Later calls refactor; repivot mode 3 re-derives the pivots on the
// host whenever splu would pick another one (lsp_options::repivot).
int build_batch(StaticState &S, int n)
{
    lsp_circuit_in cin{};
    cin.n = n;
    cin.nnz = (long)S.vals.size();
    cin.row = S.row0.data();
    cin.col = S.col0.data();
    cin.val = S.vals.data();
    cin.q = S.q.data();
    lsp_options opt;
    lsp_options_default(&opt);
    opt.repivot = 3;
    S.batch.reset(new lsp_batch());
    int rc = S.batch->impl.create(&cin, 1, &opt);
    if (rc) S.batch.reset();
    return rc;
}

// Copies the batch's fp64 factors into the splu-layout host arrays.
int export_host(StaticState &S, int n)
{
    lsp::Batch &B = S.batch->impl;
    lsp_circuit_info inf{};
    B.info(0, &inf);
    SpluFactors *f = S.fac;
    if (f->L.lnnz != inf.lnnz || f->U.unnz != inf.unnz || f->n != n || !f->pinv) {
        free_factors(f);
        f->n = n;
        f->pinv = (int *)std::malloc(n * sizeof(int));
        f->L.n = f->U.n = n;
        f->L.lp = (int *)std::malloc((n + 1) * sizeof(int));
        f->L.li = (int *)std::malloc(inf.lnnz * sizeof(int));
        f->L.lx = (double *)std::malloc(inf.lnnz * sizeof(double));
        f->U.up = (int *)std::malloc((n + 1) * sizeof(int));
        f->U.ui = (int *)std::malloc(inf.unnz * sizeof(int));
        f->U.ux = (double *)std::malloc(inf.unnz * sizeof(double));
        if (!f->pinv || !f->L.lp || !f->L.li || !f->L.lx || !f->U.up || !f->U.ui || !f->U.ux) return 2;
        f->L.lcap = f->L.lnnz = inf.lnnz;
        f->U.ucap = f->U.unnz = inf.unnz;
    }
    f->singular_col = -1;
    return B.export_factors(0, f->pinv, f->L.lp, f->L.li, f->L.lx, f->U.up, f->U.ui, f->U.ux) ? 3 : 0;
}

// The batch's current factors in `prec` (refactors when another precision was
// loaded since; the pivot sequence is kept).
int load_prec(StaticState &S, int prec)
{
    lsp::Batch &B = S.batch->impl;
    if (B.cur_prec() == prec) return LSP_OK;
    int rc = B.set_repivot(0);
    if (rc) return rc;
    lsp_refactor_report rep{};
    rc = B.refactor(prec, &rep);
    const int rc2 = B.set_repivot(3);
    return rc ? rc : rc2;
}

}  // namespace

extern "C" {

long dsh_symbolic_count(void) { return g_symbolic_count.load(); }

int dsh_lu_static_init(dsh_lu_static *F)
{
    if (!F) return 3;
    std::memset(F, 0, sizeof(*F));
    return 0;
}

int dsh_lu_static_alloc(dsh_lu_static *F)
{
    // device memory is owned by the resident state created by dsh_numeric
    if (!F || F->n <= 0) return 3;
    F->dev_allocated = 1;
    return 0;
}

void dsh_lu_static_free(dsh_lu_static *F)
{
    if (!F) return;
    StaticState *S = state(F);
    if (S) {
        free_factors(S->fac);
        std::free(S->fac);
        delete S;
    }
    F->factor_cache = nullptr;
    std::free(F->Ap);
    std::free(F->Ai);
    std::free(F->Ax);
    std::free(F->P);
    std::free(F->Q);
    F->Ap = F->Ai = nullptr;
    F->Ax = nullptr;
    F->P = F->Q = nullptr;
    F->dev_allocated = 0;
}

const char *dsh_lu_status_str(int rc)
{
    switch (rc) {
    case 0: return "OK";
    case 1: return "SINGULAR";
    case 2: return "ALLOC";
    case 3: return "ARG/NOT_IMPL";
    default: return "UNKNOWN";
    }
}

int dsh_symbolic(int n, long cnnz, const int *coo_row, const int *coo_col, const int *colamd_q,
                 dsh_lu_static *F)
{
    if (!F || n <= 0 || cnnz <= 0 || !coo_row || !coo_col) return 3;
    for (long i = 0; i < cnnz; i++)
        if (coo_row[i] < 1 || coo_row[i] > n || coo_col[i] < 1 || coo_col[i] > n) return 3;
    if (colamd_q) {
        std::vector<char> seen(n, 0);
        for (int k = 0; k < n; k++) {
            if (colamd_q[k] < 0 || colamd_q[k] >= n || seen[colamd_q[k]]) return 3;
            seen[colamd_q[k]] = 1;
        }
    }
    StaticState *S = state(F);
    if (!S) {
        S = new StaticState();
        S->fac = (SpluFactors *)std::calloc(1, sizeof(SpluFactors));
        if (!S->fac) {
            delete S;
            return 2;
        }
        S->fac->owner = S;
        F->factor_cache = S->fac;
    }
    S->batch.reset();
    // column counts, then rows bucketed per column, sorted, repeats dropped
    std::vector<int> cnt(n + 1, 0);
    for (long i = 0; i < cnnz; i++) cnt[coo_col[i]]++;
    for (int c = 0; c < n; c++) cnt[c + 1] += cnt[c];
    std::vector<int> rows(cnnz);
    {
        std::vector<int> fill(cnt.begin(), cnt.end() - 1);
        for (long i = 0; i < cnnz; i++) rows[fill[coo_col[i] - 1]++] = coo_row[i] - 1;
    }
    S->Ap.assign(n + 1, 0);
    S->Ai.clear();
    S->Ai.reserve(cnnz);
    for (int c = 0; c < n; c++) {
        auto b = rows.begin() + cnt[c], e = rows.begin() + cnt[c + 1];
        std::sort(b, e);
        S->Ai.insert(S->Ai.end(), b, std::unique(b, e));
        S->Ap[c + 1] = (int)S->Ai.size();
    }
    S->q.resize(n);
    for (int k = 0; k < n; k++) S->q[k] = colamd_q ? colamd_q[k] : k;
    S->row0.assign(S->Ai.begin(), S->Ai.end());
    S->col0.resize(S->Ai.size());
    for (int c = 0; c < n; c++) std::fill(S->col0.begin() + S->Ap[c], S->col0.begin() + S->Ap[c + 1], c);
    F->n = n;
    F->cnnz = (long)S->Ai.size();
    std::free(F->Ap);
    std::free(F->Ai);
    std::free(F->Q);
    F->Ap = (int *)std::malloc((n + 1) * sizeof(int));
    F->Ai = (int *)std::malloc(S->Ai.size() * sizeof(int));
    F->Q = (int *)std::malloc(n * sizeof(int));
    if (!F->Ap || !F->Ai || !F->Q) return 2;
    std::copy(S->Ap.begin(), S->Ap.end(), F->Ap);
    std::copy(S->Ai.begin(), S->Ai.end(), F->Ai);
    std::copy(S->q.begin(), S->q.end(), F->Q);
    g_symbolic_count++;
    return 0;
}

int dsh_numeric(dsh_lu_static *F, const int *coo_row, const int *coo_col, const double *coo_val,
                int nnz)
{
    StaticState *S = F ? state(F) : nullptr;
    if (!S || F->n <= 0 || nnz <= 0 || !coo_row || !coo_col || !coo_val) return 3;
    const int n = F->n;
    S->vals.assign(S->Ai.size(), 0.0);
    for (int i = 0; i < nnz; i++) {
        const int r = coo_row[i] - 1, c = coo_col[i] - 1;
        if (r < 0 || r >= n || c < 0 || c >= n) return 3;
        const int p = csc_slot(*S, r, c);
        if (p < 0) return 3;                      // entry outside the symbolic pattern
        S->vals[p] = coo_val[i];
    }
    if (F->pin > 0)
        for (int c = 0; c < n; c++) {
            const int p = csc_slot(*S, c, c);
            if (p >= 0) S->vals[p] += F->pin;
        }
    S->max_abs_a = 0;
    for (double v : S->vals) S->max_abs_a = std::max(S->max_abs_a, std::fabs(v));
    int rc;
    lsp_refactor_report rep{};
    if (!S->batch) {
        rc = build_batch(*S, n);                  // host pivots = splu's on these values
        if (rc == LSP_E_SINGULAR) {
            F->min_diag_U = 0.0;
            return 1;
        }
        if (rc) return rc == LSP_E_ALLOC ? 2 : 3;
    } else {
        rc = S->batch->impl.set_values(S->vals.data(), false);
        if (rc) return 3;
    }
    rc = S->batch->impl.refactor(LSP_PREC_FP64, &rep);
    if (rc == LSP_E_SINGULAR) {
        F->min_diag_U = 0.0;
        return 1;
    }
    if (rc) return rc == LSP_E_ALLOC ? 2 : 3;
    rc = export_host(*S, n);
    if (rc) return rc;
    const SpluFactors *f = S->fac;
    F->lnnz = f->L.lnnz;
    F->unnz = f->U.unnz;
    F->cap = F->lnnz + F->unnz;
    double dmin = INFINITY, dmax = 0.0;
    for (int k = 0; k < n; k++) {
        const double d = std::fabs(f->U.ux[f->U.up[k + 1] - 1]);   // U's diagonal is last
        if (d != 0.0) dmin = std::min(dmin, d);
        dmax = std::max(dmax, d);
    }
    F->min_diag_U = std::isinf(dmin) ? 0.0 : dmin;
    F->pivot_growth = S->max_abs_a > 0 ? dmax / S->max_abs_a : 0.0;
    std::free(F->P);
    F->P = (int *)std::malloc(n * sizeof(int));
    if (!F->P) return 2;
    std::copy(f->pinv, f->pinv + n, F->P);
    F->dev_allocated = 1;
    return 0;
}

int dsh_trisolve(const dsh_lu_static *F, const double *b, double *x)
{
    StaticState *S = F ? state(F) : nullptr;
    if (!S || !S->batch || !b || !x) return 3;
    if (load_prec(*S, LSP_PREC_FP64)) return 3;
    return S->batch->impl.trisolve(b, x, false) == LSP_OK ? 0 : 3;
}

int dsh_gmres_ir(dsh_lu_static *F, const double *b, double *x, int max_outer, int max_inner,
                 double rtol, double *residual_seq, int *iters)
{
    StaticState *S = F ? state(F) : nullptr;
    if (!S || !S->batch || !b || !x || max_outer < 1 || max_inner < 1 || !(rtol > 0)) return 3;
    lsp::Batch &B = S->batch->impl;
    if (B.set_ir_params(max_outer, max_inner, rtol, rtol) != LSP_OK) return 3;
    if (load_prec(*S, LSP_PREC_FP32)) return 3;
    lsp_solve_report rep{};
    const int rc = B.solve(b, x, false, &rep);
    if (iters) *iters = rep.inner_total;
    if (residual_seq) {
        const auto &h = B.residual_history();
        for (int i = 0; i < max_outer + 1; i++)
            residual_seq[i] = i < (int)h.size() ? h[i] : rep.max_rel_res;
    }
    return (rc == LSP_OK && rep.max_rel_res <= std::max(rtol * 10, 1e-9)) ? 0 : 3;
}

}  // extern "C"
