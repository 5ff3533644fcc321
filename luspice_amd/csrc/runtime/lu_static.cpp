// lu_static.cpp -- the reference's static-LU lifecycle (include/lu_static.h)
// on the device-resident batch runtime.
//
// Reference behaviour kept (src/lu_static.c): dsh_symbolic reads the COO
// pattern with 1-based row/col (coo_to_csc, :78-97), sorts rows inside each
// column, stores the column order (NULL = natural) and counts calls;
// dsh_numeric rebuilds values from COO (a later duplicate of an entry
// overwrites an earlier one, as splu's scatter does), adds pin to diagonal
// entries present in the pattern (:288-298), factors, and reports SINGULAR
// (1) / ARG (3); dsh_lu_status_str gives the same texts.  What changes: the
// factorization runs on the GPU with the pivot sequence reused across calls,
// and dsh_gmres_ir is implemented.
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

struct StaticState {
    std::vector<int> Ap, Ai;              // pattern, 0-based rows sorted per column
    std::vector<int> q;
    std::vector<int> row0, col0;          // pattern in CSC order (0-based COO)
    std::vector<double> vals;             // values in CSC order
    std::unique_ptr<lsp_batch> batch;
    double max_abs_a = 0;
};

StaticState *state(const dsh_lu_static *F) { return (StaticState *)F->factor_cache; }

int csc_slot(const StaticState &S, int r, int c)
{
    const int *b = S.Ai.data() + S.Ap[c], *e = S.Ai.data() + S.Ap[c + 1];
    const int *it = std::lower_bound(b, e, r);
    return (it != e && *it == r) ? (int)(it - S.Ai.data()) : -1;
}

int build_batch(StaticState &S, int n, double pin)
{
    (void)pin;
    lsp_circuit_in cin{};
    cin.n = n;
    cin.nnz = (long)S.vals.size();
    cin.row = S.row0.data();
    cin.col = S.col0.data();
    cin.val = S.vals.data();
    cin.q = S.q.data();
    lsp_options opt;
    lsp_options_default(&opt);
    S.batch.reset(new lsp_batch());
    int rc = S.batch->impl.create(&cin, 1, &opt);
    if (rc) S.batch.reset();
    return rc;
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
    delete state(F);
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
    StaticState *S = state(F);
    if (!S) {
        S = new StaticState();
        F->factor_cache = S;
    }
    S->batch.reset();
    S->Ap.assign(n + 1, 0);
    for (long i = 0; i < cnnz; i++) S->Ap[coo_col[i]]++;
    for (int c = 0; c < n; c++) S->Ap[c + 1] += S->Ap[c];
    std::vector<int> rows(cnnz);
    std::vector<int> cur(S->Ap.begin(), S->Ap.end() - 1);
    for (long i = 0; i < cnnz; i++) rows[cur[coo_col[i] - 1]++] = coo_row[i] - 1;
    // sort and drop repeated (row, col) pairs of the pattern
    std::vector<int> Ap(n + 1, 0), Ai;
    Ai.reserve(cnnz);
    for (int c = 0; c < n; c++) {
        std::sort(rows.begin() + S->Ap[c], rows.begin() + S->Ap[c + 1]);
        for (int p = S->Ap[c]; p < S->Ap[c + 1]; p++)
            if (p == S->Ap[c] || rows[p] != rows[p - 1]) Ai.push_back(rows[p]);
        Ap[c + 1] = (int)Ai.size();
    }
    S->Ap.swap(Ap);
    S->Ai.swap(Ai);
    S->q.resize(n);
    if (colamd_q) {
        std::vector<char> seen(n, 0);
        for (int k = 0; k < n; k++) {
            if (colamd_q[k] < 0 || colamd_q[k] >= n || seen[colamd_q[k]]) return 3;
            seen[colamd_q[k]] = 1;
            S->q[k] = colamd_q[k];
        }
    } else {
        for (int k = 0; k < n; k++) S->q[k] = k;
    }
    S->row0.resize(S->Ai.size());
    S->col0.resize(S->Ai.size());
    for (int c = 0; c < n; c++)
        for (int p = S->Ap[c]; p < S->Ap[c + 1]; p++) {
            S->row0[p] = S->Ai[p];
            S->col0[p] = c;
        }
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
        rc = build_batch(*S, n, F->pin);
        if (rc == LSP_E_SINGULAR) return 1;
        if (rc) return rc == LSP_E_ALLOC ? 2 : 3;
        rc = S->batch->impl.refactor(LSP_PREC_FP32, &rep);
    } else {
        rc = S->batch->impl.set_values(S->vals.data(), false);
        if (rc) return 3;
        rc = S->batch->impl.refactor(LSP_PREC_FP32, &rep);
        if (rc == LSP_OK && rep.bad_pivots > 0) rc = LSP_E_SINGULAR;  // pivot no longer acceptable
        if (rc == LSP_E_SINGULAR) {
            // new values need a new pivot sequence: re-derive it on the host
            rc = build_batch(*S, n, F->pin);
            if (rc == LSP_E_SINGULAR) return 1;
            if (rc) return rc == LSP_E_ALLOC ? 2 : 3;
            rc = S->batch->impl.refactor(LSP_PREC_FP32, &rep);
        }
    }
    if (rc == LSP_E_SINGULAR) return 1;
    if (rc) return 3;
    lsp_circuit_info inf{};
    S->batch->impl.info(0, &inf);
    F->lnnz = inf.lnnz;
    F->unnz = inf.unnz;
    F->cap = inf.lnnz + inf.unnz;
    double dmin = 0, dmax = 0;
    S->batch->impl.diag_stats(0, &dmin, &dmax);
    F->min_diag_U = dmin;
    F->pivot_growth = S->max_abs_a > 0 ? dmax / S->max_abs_a : 0.0;
    std::free(F->P);
    F->P = (int *)std::malloc(n * sizeof(int));
    if (!F->P) return 2;
    const auto &pinv = S->batch->impl.circuit(0).plan.pinv;
    std::copy(pinv.begin(), pinv.end(), F->P);
    F->dev_allocated = 1;
    return 0;
}

int dsh_trisolve(const dsh_lu_static *F, const double *b, double *x)
{
    StaticState *S = F ? state(F) : nullptr;
    if (!S || !S->batch || !b || !x) return 3;
    return S->batch->impl.trisolve(b, x, false) == LSP_OK ? 0 : 3;
}

int dsh_gmres_ir(dsh_lu_static *F, const double *b, double *x, int max_outer, int max_inner,
                 double rtol, double *residual_seq, int *iters)
{
    StaticState *S = F ? state(F) : nullptr;
    if (!S || !S->batch || !b || !x || max_outer < 1 || max_inner < 1 || !(rtol > 0)) return 3;
    lsp::Batch &B = S->batch->impl;
    if (B.set_ir_params(max_outer, max_inner, rtol, rtol) != LSP_OK) return 3;
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
