// csc.cpp -- COO -> CSC / CSR conversions with the reference's ordering rules.
//
// lsp_coo_to_csc mirrors build_csc_q (src/mp/mp_lu_bench.cu:150-186): column
// k of the result is original column q[k] (q = nullptr: natural order), row
// ids stay original, and rows inside a column keep the COO (file) order.
// src[p] records which COO entry landed in slot p so that later value arrays
// in COO order can be gathered without rebuilding the structure.
// lsp_colamd_input sorts rows inside each column, as run_colamd does before
// calling colamd (src/mp/mp_lu_bench.cu:131-138).
#include <algorithm>
#include <vector>

#include "host_internal.h"
#include "csc.h"

void lsp_coo_to_csc(int n, long nnz, const int *row, const int *col, const int *q,
                    std::vector<int> &Ap, std::vector<int> &Ai, std::vector<int> &src)
{
    std::vector<int> invq(n);
    for (int k = 0; k < n; k++) invq[q ? q[k] : k] = k;
    Ap.assign(n + 1, 0);
    for (long e = 0; e < nnz; e++) Ap[invq[col[e]] + 1]++;
    for (int k = 0; k < n; k++) Ap[k + 1] += Ap[k];
    Ai.assign(nnz, 0);
    src.assign(nnz, 0);
    std::vector<int> cur(Ap.begin(), Ap.end() - 1);
    for (long e = 0; e < nnz; e++) {
        int p = cur[invq[col[e]]]++;
        Ai[p] = row[e];
        src[p] = (int)e;
    }
}

void lsp_coo_to_csr(int n, long nnz, const int *row, const int *col,
                    std::vector<int> &rp, std::vector<int> &ci, std::vector<int> &src)
{
    rp.assign(n + 1, 0);
    for (long e = 0; e < nnz; e++) rp[row[e] + 1]++;
    for (int i = 0; i < n; i++) rp[i + 1] += rp[i];
    ci.assign(nnz, 0);
    src.assign(nnz, 0);
    std::vector<int> cur(rp.begin(), rp.end() - 1);
    for (long e = 0; e < nnz; e++) {
        int p = cur[row[e]]++;
        ci[p] = col[e];
        src[p] = (int)e;
    }
}

int lsp_colamd_coo(int n, long nnz, const int *row, const int *col, int *q)
{
    std::vector<int> Ap, Ai, src;
    lsp_coo_to_csc(n, nnz, row, col, nullptr, Ap, Ai, src);
    for (int c = 0; c < n; c++) std::sort(Ai.begin() + Ap[c], Ai.begin() + Ap[c + 1]);
    return lsp_colamd_order(n, n, Ap.data(), Ai.data(), q);
}
