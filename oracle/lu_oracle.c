/* oracle/lu_oracle.c -- CPU restatement of the reference's fp64 sparse LU path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library (luspice_amd/)
 * links, loads or calls this file; only tests/, __graft_entry__.smoke() and
 * the cpu_baseline leg of bench.py use it, and only as the checker.
 *
 * What it restates (reference paths relative to the reference repository):
 *   orc_lu_factor     src/third_party/splu.h:151-206  (lu_factor: left-looking
 *                     Gilbert-Peierls LU, P*A(:,q) = L*U, DFS reach per column
 *                     src/third_party/splu.h:77-123, sparse triangular solve
 *                     :126-146, first strict maximum in reach order, diagonal
 *                     row q[k] kept when |x| >= tol*max, L unit diagonal first,
 *                     U diagonal last, L rows remapped to pivot positions :202)
 *   orc_lu_solve_perm src/third_party/splu.h:210-239
 *   orc_cpu_lu_ref    src/cpu_ref_lu.c:15-58 (capacity grows instead of the
 *                     n*n+4096 preallocation at :22, which is infeasible for
 *                     n >= ~20k; growth does not change any result)
 *   orc_level_compute src/level.c:20-40
 *   orc_level_hist    src/level.c:43-48
 *
 * Pinning: tests/test_oracle_golden.py compares every output (pinv, the L/U
 * arrays in storage order, lnnz/unnz, x, levels) with fixtures written by the
 * reference's own C compiled in oracle/_ref (tests/golden/make_golden.py),
 * against a dense partial-pivoting solve on tiny matrices, and against the
 * closed form x = Q U^-1 L^-1 P b.
 *
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off, no -march).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

/* ------------------------------------------------------------------------ */
/* growable factor storage                                                   */
/* ------------------------------------------------------------------------ */
static int orc_reserve(int **idx, double **val, long *cap, long need)
{
    if (need <= *cap) return 0;
    long nc = (*cap) * 2 + 16;
    if (nc < need) nc = need;
    int *ni = (int *)realloc(*idx, (size_t)nc * sizeof(int));
    if (!ni) return -1;
    *idx = ni;
    double *nv = (double *)realloc(*val, (size_t)nc * sizeof(double));
    if (!nv) return -1;
    *val = nv;
    *cap = nc;
    return 0;
}

void orc_lu_free(orc_lu *F)
{
    if (!F) return;
    free(F->pinv); free(F->Lp); free(F->Li); free(F->Lx);
    free(F->Up); free(F->Ui); free(F->Ux);
    memset(F, 0, sizeof(*F));
}

/* ------------------------------------------------------------------------ */
/* reach: nodes are original row ids.  Row i that is already pivotal (pinv[i]
 * = J >= 0) has out-edges to the rows stored in L column J.  Returns top;
 * xi[top..n) lists the reach of column `col` of A in topological order, the
 * same order the reference's recursive-stack DFS produces.                 */
/* ------------------------------------------------------------------------ */
static int orc_reach(int n, const int *Lp, const int *Li, const int *pinv,
                     const int *Bp, const int *Bi, int col,
                     int *xi, int *stk, int *nxt, int *mark, int stamp)
{
    int top = n;
    for (int p = Bp[col]; p < Bp[col + 1]; p++) {
        int start = Bi[p];
        if (mark[start] == stamp) continue;
        int depth = 0;
        stk[0] = start;
        while (depth >= 0) {
            int node = stk[depth];
            int J = pinv[node];
            if (mark[node] != stamp) {          /* first visit */
                mark[node] = stamp;
                nxt[depth] = (J < 0) ? 0 : Lp[J];
            }
            int end = (J < 0) ? 0 : Lp[J + 1];
            int pushed = 0;
            for (int e = nxt[depth]; e < end; e++) {
                int child = Li[e];
                if (mark[child] == stamp) continue;
                nxt[depth] = e;                 /* resume here */
                stk[++depth] = child;
                pushed = 1;
                break;
            }
            if (!pushed) {                      /* all children done */
                depth--;
                xi[--top] = node;
            }
        }
    }
    return top;
}

int orc_lu_factor(int n, const int *Ap, const int *Ai, const double *Ax,
                  const int *q, double tol, orc_lu *F)
{
    memset(F, 0, sizeof(*F));
    if (n < 1) return -2;
    F->n = n;
    F->pinv = (int *)malloc((size_t)n * sizeof(int));
    F->Lp = (int *)calloc((size_t)n + 1, sizeof(int));
    F->Up = (int *)calloc((size_t)n + 1, sizeof(int));
    int *xi = (int *)malloc((size_t)n * sizeof(int));
    int *stk = (int *)malloc((size_t)n * sizeof(int));
    int *nxt = (int *)malloc((size_t)n * sizeof(int));
    int *mark = (int *)calloc((size_t)n, sizeof(int));
    double *x = (double *)calloc((size_t)n, sizeof(double));
    if (!F->pinv || !F->Lp || !F->Up || !xi || !stk || !nxt || !mark || !x) {
        free(xi); free(stk); free(nxt); free(mark); free(x);
        return -2;
    }
    for (int i = 0; i < n; i++) F->pinv[i] = -1;
    long lnz = 0, unz = 0;
    int rc = -1;
    for (int k = 0; k < n; k++) {
        F->Lp[k] = (int)lnz;
        F->Up[k] = (int)unz;
        if (orc_reserve(&F->Li, &F->Lx, &F->lcap, lnz + n + 16) ||
            orc_reserve(&F->Ui, &F->Ux, &F->ucap, unz + n + 16)) { rc = -2; break; }
        int col = q ? q[k] : k;
        /* x = L \ A(:,col) over the reach */
        int top = orc_reach(n, F->Lp, F->Li, F->pinv, Ap, Ai, col,
                            xi, stk, nxt, mark, k + 1);
        for (int p = top; p < n; p++) x[xi[p]] = 0.0;
        for (int p = Ap[col]; p < Ap[col + 1]; p++) x[Ai[p]] = Ax[p];
        for (int p = top; p < n; p++) {
            int r = xi[p];
            int J = F->pinv[r];
            if (J < 0) continue;
            x[r] /= F->Lx[F->Lp[J]];              /* unit diagonal, stored first */
            double xr = x[r];
            for (int e = F->Lp[J] + 1; e < F->Lp[J + 1]; e++)
                x[F->Li[e]] -= F->Lx[e] * xr;
        }
        /* pivot: first strict maximum over unpivoted rows in reach order;
         * pivotal rows go to U in reach order */
        int ipiv = -1;
        double amax = -1.0;
        for (int p = top; p < n; p++) {
            int r = xi[p];
            if (F->pinv[r] < 0) {
                double t = fabs(x[r]);
                if (t > amax) { amax = t; ipiv = r; }
            } else {
                F->Ui[unz] = F->pinv[r];
                F->Ux[unz++] = x[r];
            }
        }
        if (ipiv == -1 || amax <= 0) { rc = k; break; }
        if (F->pinv[col] < 0 && fabs(x[col]) >= amax * tol) ipiv = col;
        double piv = x[ipiv];
        F->Ui[unz] = k;
        F->Ux[unz++] = piv;
        F->pinv[ipiv] = k;
        F->Li[lnz] = ipiv;
        F->Lx[lnz++] = 1.0;
        for (int p = top; p < n; p++) {
            int r = xi[p];
            if (F->pinv[r] < 0) {
                F->Li[lnz] = r;
                F->Lx[lnz++] = x[r] / piv;
            }
            x[r] = 0.0;
        }
    }
    free(xi); free(stk); free(nxt); free(mark); free(x);
    F->singular_col = rc;
    if (rc != -1) return rc;
    F->Lp[n] = (int)lnz;
    F->Up[n] = (int)unz;
    for (long p = 0; p < lnz; p++) F->Li[p] = F->pinv[F->Li[p]];
    F->lnnz = lnz;
    F->unnz = unz;
    return -1;
}

void orc_lu_solve_perm(int n, const orc_lu *F, const int *q,
                       const double *b, double *x)
{
    double *w = (double *)malloc((size_t)n * sizeof(double));
    double *y = (double *)malloc((size_t)n * sizeof(double));
    int *row_of = (int *)malloc((size_t)n * sizeof(int));
    for (int k = 0; k < n; k++) row_of[k] = -1;
    for (int i = 0; i < n; i++) if (F->pinv[i] >= 0) row_of[F->pinv[i]] = i;
    for (int k = 0; k < n; k++) w[k] = row_of[k] >= 0 ? b[row_of[k]] : 0.0;
    for (int k = 0; k < n; k++) {                   /* L y = P b */
        y[k] = w[k] / F->Lx[F->Lp[k]];
        for (int e = F->Lp[k] + 1; e < F->Lp[k + 1]; e++)
            w[F->Li[e]] -= F->Lx[e] * y[k];
    }
    for (int k = n - 1; k >= 0; k--) {              /* U z = y */
        y[k] /= F->Ux[F->Up[k + 1] - 1];
        for (int e = F->Up[k]; e < F->Up[k + 1] - 1; e++)
            y[F->Ui[e]] -= F->Ux[e] * y[k];
    }
    for (int k = 0; k < n; k++) x[q ? q[k] : k] = y[k];
    free(w); free(y); free(row_of);
}

int orc_cpu_lu_ref(int n, const int *Ap, const int *Ai, const double *Ax,
                   const int *q, const double *b, double tol,
                   double *x, long *lnnz, long *unnz)
{
    if (n < 1) return -2;
    orc_lu F;
    int rc = orc_lu_factor(n, Ap, Ai, Ax, q, tol, &F);
    if (rc != -1) { orc_lu_free(&F); return rc; }
    orc_lu_solve_perm(n, &F, q, b, x);
    if (lnnz) *lnnz = F.lnnz;
    if (unnz) *unnz = F.unnz;
    orc_lu_free(&F);
    return -1;
}

/* Copy the factor arrays out (caller sizes them from lnnz/unnz). */
void orc_lu_export(const orc_lu *F, int *pinv, int *Lp, int *Li, double *Lx,
                   int *Up, int *Ui, double *Ux)
{
    int n = F->n;
    if (pinv) memcpy(pinv, F->pinv, (size_t)n * sizeof(int));
    if (Lp) memcpy(Lp, F->Lp, (size_t)(n + 1) * sizeof(int));
    if (Up) memcpy(Up, F->Up, (size_t)(n + 1) * sizeof(int));
    if (Li) memcpy(Li, F->Li, (size_t)F->lnnz * sizeof(int));
    if (Lx) memcpy(Lx, F->Lx, (size_t)F->lnnz * sizeof(double));
    if (Ui) memcpy(Ui, F->Ui, (size_t)F->unnz * sizeof(int));
    if (Ux) memcpy(Ux, F->Ux, (size_t)F->unnz * sizeof(double));
}

/* One-shot convenience for ctypes: factor and copy out.  The caller passes
 * capacities; returns -1 ok, k singular column, -2 bad argument, -3 when the
 * caller's arrays are too small (lnnz/unnz are still reported). */
int orc_lu_factor_export(int n, const int *Ap, const int *Ai, const double *Ax,
                         const int *q, double tol,
                         int *pinv, int *Lp, int *Li, double *Lx, long lcap,
                         int *Up, int *Ui, double *Ux, long ucap,
                         long *lnnz, long *unnz)
{
    orc_lu F;
    int rc = orc_lu_factor(n, Ap, Ai, Ax, q, tol, &F);
    if (rc != -1) { orc_lu_free(&F); return rc; }
    if (lnnz) *lnnz = F.lnnz;
    if (unnz) *unnz = F.unnz;
    if (F.lnnz > lcap || F.unnz > ucap) { orc_lu_free(&F); return -3; }
    orc_lu_export(&F, pinv, Lp, Li, Lx, Up, Ui, Ux);
    orc_lu_free(&F);
    return -1;
}

/* ------------------------------------------------------------------------ */
/* level sets over the U pattern (position space, diagonal last)             */
/* ------------------------------------------------------------------------ */
int orc_level_compute(int n, const int *Up, const int *Ui, int *level,
                      int *nlevels)
{
    int deepest = 0;
    for (int k = 0; k < n; k++) {
        int lv = 0;
        for (int e = Up[k]; e < Up[k + 1] - 1; e++) {
            int j = Ui[e];
            if (j >= 0 && j < k && level[j] + 1 > lv) lv = level[j] + 1;
        }
        level[k] = lv;
        if (lv > deepest) deepest = lv;
    }
    *nlevels = deepest + 1;
    return 0;
}

int orc_level_hist(int n, const int *level, int nlevels, int *hist)
{
    for (int l = 0; l < nlevels; l++) hist[l] = 0;
    for (int k = 0; k < n; k++)
        if (level[k] < nlevels) hist[level[k]]++;
    return 0;
}
