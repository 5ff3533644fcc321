/* s0io.c -- s0 matrix / RHS / .flt file handling.
 *
 * read_s0_to_csr, csr_free, read_rhs: the reference's spmv.h contract
 * (src/spmv.cu:42-145): header "n <n>", two passes over "A i j v" lines,
 * entries with i outside [1, n] or j < 1 skipped, rows kept in file order.
 *
 * lsp_s0_read: the CLI reader of src/mp/mp_lu_bench.cu:46-85 -- COO kept in
 * file order (1-based), RHS from "RHS i v" lines with 1 <= i <= n, and any
 * RHS entry that is exactly zero or missing becomes 1.0 (:83).
 * lsp_flt_pin: src/mp/mp_lu_bench.cu:87-111 -- for each "F <row>" line of the
 * .flt file add pin to the diagonal entry (first match in file order), or
 * append a new diagonal entry; only when pin > 0 and the file opens.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "spmv.h"
#include "s0io.h"

int read_s0_to_csr(const char *path, CSRMatrix *A)
{
    FILE *f = fopen(path, "r");
    if (!f) { fprintf(stderr, "[spmv] error: cannot open %s\n", path); return 1; }
    char line[512];
    int n = 0;
    if (!fgets(line, sizeof line, f) || sscanf(line, "n %d", &n) != 1 || n <= 0) {
        fclose(f);
        return 1;
    }
    long long cnt = 0;
    while (fgets(line, sizeof line, f)) if (line[0] == 'A') cnt++;
    if (cnt == 0) { fclose(f); return 1; }
    int *ri = (int *)malloc((size_t)cnt * sizeof(int));
    int *rj = (int *)malloc((size_t)cnt * sizeof(int));
    double *rv = (double *)malloc((size_t)cnt * sizeof(double));
    rewind(f);
    if (!fgets(line, sizeof line, f)) { fclose(f); free(ri); free(rj); free(rv); return 1; }
    long long k = 0;
    while (k < cnt && fgets(line, sizeof line, f)) {
        int i, j;
        double v;
        if (line[0] != 'A' || sscanf(line, "A %d %d %lf", &i, &j, &v) != 3) continue;
        if (i < 1 || i > n || j < 1) continue;
        ri[k] = i - 1; rj[k] = j - 1; rv[k] = v;
        k++;
    }
    fclose(f);
    int *rowptr = (int *)calloc((size_t)n + 1, sizeof(int));
    for (long long t = 0; t < k; t++) rowptr[ri[t] + 1]++;
    for (int i = 0; i < n; i++) rowptr[i + 1] += rowptr[i];
    int *colidx = (int *)malloc((size_t)(k > 0 ? k : 1) * sizeof(int));
    double *values = (double *)malloc((size_t)(k > 0 ? k : 1) * sizeof(double));
    int *cur = (int *)malloc((size_t)n * sizeof(int));
    memcpy(cur, rowptr, (size_t)n * sizeof(int));
    for (long long t = 0; t < k; t++) {
        int p = cur[ri[t]]++;
        colidx[p] = rj[t];
        values[p] = rv[t];
    }
    free(ri); free(rj); free(rv); free(cur);
    A->n = n;
    A->nnz = k;
    A->rowptr = rowptr;
    A->colidx = colidx;
    A->values = values;
    return 0;
}

void csr_free(CSRMatrix *A)
{
    if (!A) return;
    free(A->rowptr);
    free(A->colidx);
    free(A->values);
    A->rowptr = A->colidx = NULL;
    A->values = NULL;
    A->n = 0;
    A->nnz = 0;
}

int read_rhs(const char *path, int n, double *x)
{
    FILE *f = fopen(path, "r");
    if (!f) { fprintf(stderr, "[spmv] error: cannot open %s\n", path); return 1; }
    char line[512];
    if (!fgets(line, sizeof line, f)) { fclose(f); return 1; }
    int cnt = 0;
    while (fgets(line, sizeof line, f)) {
        int i;
        double v;
        if (line[0] != 'R' || sscanf(line, "RHS %d %lf", &i, &v) != 2) continue;
        if (i >= 1 && i <= n) { x[i - 1] = v; cnt++; }
    }
    fclose(f);
    return cnt > 0 ? 0 : 1;
}

/* ------------------------------------------------------------------------ */
int lsp_s0_read(const char *mtx, const char *rhs, lsp_s0 *M)
{
    memset(M, 0, sizeof(*M));
    FILE *f = fopen(mtx, "r");
    if (!f) { fprintf(stderr, "open fail %s\n", mtx); return 1; }
    char line[512];
    if (!fgets(line, sizeof line, f) || sscanf(line, "n %d", &M->n) != 1 || M->n <= 0) {
        fclose(f);
        return 1;
    }
    long cap = 1024;
    M->row = (int *)malloc((size_t)cap * sizeof(int));
    M->col = (int *)malloc((size_t)cap * sizeof(int));
    M->val = (double *)malloc((size_t)cap * sizeof(double));
    while (fgets(line, sizeof line, f)) {
        if (line[0] != 'A') continue;
        if (M->nnz >= cap) {
            cap *= 2;
            M->row = (int *)realloc(M->row, (size_t)cap * sizeof(int));
            M->col = (int *)realloc(M->col, (size_t)cap * sizeof(int));
            M->val = (double *)realloc(M->val, (size_t)cap * sizeof(double));
        }
        if (sscanf(line, "A %d %d %lf", &M->row[M->nnz], &M->col[M->nnz], &M->val[M->nnz]) == 3)
            M->nnz++;
    }
    fclose(f);
    M->b = (double *)calloc((size_t)M->n + 1, sizeof(double));
    if (rhs && (f = fopen(rhs, "r"))) {
        while (fgets(line, sizeof line, f)) {
            int idx;
            double v;
            if (sscanf(line, "RHS %d %lf", &idx, &v) == 2 && idx >= 1 && idx <= M->n) M->b[idx] = v;
        }
        fclose(f);
    }
    for (int i = 1; i <= M->n; i++) if (M->b[i] == 0) M->b[i] = 1.0;
    return 0;
}

void lsp_s0_free(lsp_s0 *M)
{
    free(M->row); free(M->col); free(M->val); free(M->b);
    memset(M, 0, sizeof(*M));
}

int lsp_flt_pin(lsp_s0 *M, const char *flt, double pin)
{
    if (!flt || pin <= 0) return 0;
    FILE *f = fopen(flt, "r");
    if (!f) return 0;
    char line[256];
    int added = 0;
    while (fgets(line, sizeof line, f)) {
        int r;
        if (sscanf(line, "F %d", &r) != 1) continue;
        int found = 0;
        for (long k = 0; k < M->nnz; k++) {
            if (M->row[k] == r && M->col[k] == r) { M->val[k] += pin; found = 1; break; }
        }
        if (!found) {
            M->row = (int *)realloc(M->row, (size_t)(M->nnz + 1) * sizeof(int));
            M->col = (int *)realloc(M->col, (size_t)(M->nnz + 1) * sizeof(int));
            M->val = (double *)realloc(M->val, (size_t)(M->nnz + 1) * sizeof(double));
            M->row[M->nnz] = r; M->col[M->nnz] = r; M->val[M->nnz] = pin;
            M->nnz++;
            added++;
        }
    }
    fclose(f);
    return added;
}

int lsp_s0_write(const char *path, int n, long nnz, const int *row0, const int *col0,
                 const double *val)
{
    FILE *f = fopen(path, "w");
    if (!f) return 1;
    fprintf(f, "n %d\n", n);
    for (long e = 0; e < nnz; e++) fprintf(f, "A %d %d %.17g\n", row0[e] + 1, col0[e] + 1, val[e]);
    return fclose(f) ? 1 : 0;
}

int lsp_rhs_write(const char *path, int n, const double *b)
{
    FILE *f = fopen(path, "w");
    if (!f) return 1;
    fprintf(f, "n %d\n", n);
    for (int i = 0; i < n; i++) fprintf(f, "RHS %d %.17g\n", i + 1, b[i]);
    return fclose(f) ? 1 : 0;
}
