/* oracle/ir_oracle.c -- CPU restatement of the reference's mixed solve.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 *   orc_coo_spmv      bench/test_lu_gpu_f32.cu:172-178 (host_spmv: COO loop in
 *                     file order, y zeroed first)
 *   orc_residual      src/mp/mp_lu_bench.cu:188-203 (||A x - b||_2 / ||b||_2)
 *   orc_rel_l2        src/mp/mp_lu_bench.cu:205-214
 *   orc_csr_spmv      src/spmv.cu:154-171 (one row, entries in CSR order)
 *   orc_trisolve_f32  src/lu_gpu_f32.cu:94-124 (fp32 factors read as double,
 *                     fp64 accumulation, x[q[k]] = z[k])
 *   orc_gmres_ir      bench/test_lu_gpu_f32.cu:540-715 (right-preconditioned
 *                     restarted GMRES(m), x0 = 0, two-pass modified
 *                     Gram-Schmidt, Givens, inner stop |g_{j+1}|/||b|| <= tol,
 *                     outer stop on the true residual).  One deliberate change:
 *                     H is zeroed at every restart; the reference accumulates
 *                     into H with += and never clears it across restarts
 *                     (:550, :624-625), which corrupts every restart after the
 *                     first.
 *   orc_ruiz          Ruiz equilibration (D. Ruiz, "A scaling algorithm to
 *                     equilibrate both rows and columns norms in matrices",
 *                     RAL-TR-2001-034): each sweep divides rows by sqrt(row max)
 *                     then columns by sqrt(col max) of the current scaled
 *                     matrix.  The reference's loop (bench/test_lu_gpu_f32.cu:
 *                     247-269) multiplies the values by the cumulative scale
 *                     every sweep, so its returned R, C do not describe the
 *                     matrix it factors; this restatement keeps
 *                     A_s = diag(R) A diag(C) exact.
 *   orc_diag_shift    bench/test_lu_gpu_f32.cu:354-364 and
 *                     bench/test_lu_gpu_ir64.cu:206-212 (factor-only shift: for
 *                     each row i, the first COO entry in file order with
 *                     row == col == i becomes a + shift*|a|; later duplicates and
 *                     every other entry are copied unchanged)
 *   orc_sell_order    src/sellc_spmv.cu:199-257 (rows sorted by length,
 *                     descending, by counting sort -- equal lengths come out in
 *                     reverse index order; slice length = max row length in the
 *                     slice, slice height C)
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

void orc_coo_spmv(int n, long nnz, const int *row1, const int *col1,
                  const double *val, const double *x, double *y)
{
    for (int i = 0; i < n; i++) y[i] = 0.0;
    for (long e = 0; e < nnz; e++)
        y[row1[e] - 1] += val[e] * x[col1[e] - 1];
}

double orc_residual(int n, long nnz, const int *row1, const int *col1,
                    const double *val, const double *b, const double *x)
{
    double *ax = (double *)malloc((size_t)n * sizeof(double));
    orc_coo_spmv(n, nnz, row1, col1, val, x, ax);
    double num = 0, den = 0;
    for (int i = 0; i < n; i++) {
        double d = ax[i] - b[i];
        num += d * d;
        den += b[i] * b[i];
    }
    free(ax);
    return sqrt(num / den);
}

double orc_rel_l2(int n, const double *a, const double *b)
{
    double num = 0, den = 0;
    for (int i = 0; i < n; i++) {
        double d = a[i] - b[i];
        num += d * d;
        den += b[i] * b[i];
    }
    return sqrt(num / den);
}

void orc_csr_spmv(int n, const int *rowptr, const int *colidx,
                  const double *val, const double *x, double *y)
{
    for (int i = 0; i < n; i++) {
        double s = 0.0;
        for (int e = rowptr[i]; e < rowptr[i + 1]; e++) s += val[e] * x[colidx[e]];
        y[i] = s;
    }
}

void orc_trisolve_f32(int n, const int *Lp, const int *Li, const float *Lx,
                      const int *Up, const int *Ui, const float *Ux,
                      const int *pinv, const int *q, const double *b,
                      double *x)
{
    int *row_of = (int *)malloc((size_t)n * sizeof(int));
    double *w = (double *)malloc((size_t)n * sizeof(double));
    double *y = (double *)malloc((size_t)n * sizeof(double));
    for (int k = 0; k < n; k++) row_of[k] = -1;
    for (int i = 0; i < n; i++) if (pinv[i] >= 0) row_of[pinv[i]] = i;
    for (int k = 0; k < n; k++) w[k] = row_of[k] >= 0 ? b[row_of[k]] : 0.0;
    for (int k = 0; k < n; k++) {
        y[k] = w[k] / (double)Lx[Lp[k]];
        for (int e = Lp[k] + 1; e < Lp[k + 1]; e++)
            w[Li[e]] -= (double)Lx[e] * y[k];
    }
    for (int k = n - 1; k >= 0; k--) {
        y[k] /= (double)Ux[Up[k + 1] - 1];
        for (int e = Up[k]; e < Up[k + 1] - 1; e++)
            y[Ui[e]] -= (double)Ux[e] * y[k];
    }
    for (int k = 0; k < n; k++) x[q ? q[k] : k] = y[k];
    free(row_of); free(w); free(y);
}

int orc_gmres_ir(int n, long nnz, const int *row1, const int *col1,
                 const double *val, const double *b,
                 const int *Lp, const int *Li, const float *Lx,
                 const int *Up, const int *Ui, const float *Ux,
                 const int *pinv, const int *q,
                 int max_outer, int m, double inner_tol, double outer_tol,
                 double *x, int *total_inner, int *outer_done,
                 double *final_res)
{
    double *V = (double *)malloc((size_t)(m + 1) * n * sizeof(double));
    double *H = (double *)malloc((size_t)(m + 1) * m * sizeof(double));
    double *cs = (double *)calloc((size_t)m, sizeof(double));
    double *sn = (double *)calloc((size_t)m, sizeof(double));
    double *g = (double *)calloc((size_t)m + 1, sizeof(double));
    double *yv = (double *)calloc((size_t)m + 1, sizeof(double));
    double *w = (double *)malloc((size_t)n * sizeof(double));
    double *z = (double *)malloc((size_t)n * sizeof(double));
    double *r = (double *)malloc((size_t)n * sizeof(double));
    double nb = 0.0;
    for (int i = 0; i < n; i++) nb += b[i] * b[i];
    nb = sqrt(nb);
    for (int i = 0; i < n; i++) x[i] = 0.0;
    int inner_sum = 0, outer = 0;
    double rel = 0.0;
    for (outer = 0; outer < max_outer; outer++) {
        orc_coo_spmv(n, nnz, row1, col1, val, x, w);
        for (int i = 0; i < n; i++) r[i] = b[i] - w[i];
        double beta = 0.0;
        for (int i = 0; i < n; i++) beta += r[i] * r[i];
        beta = sqrt(beta);
        if (beta / nb <= outer_tol) { rel = beta / nb; break; }
        memset(H, 0, (size_t)(m + 1) * m * sizeof(double));
        for (int i = 0; i < n; i++) V[i] = r[i] / beta;
        g[0] = beta;
        int j;
        for (j = 0; j < m; j++) {
            orc_trisolve_f32(n, Lp, Li, Lx, Up, Ui, Ux, pinv, q, V + (size_t)j * n, z);
            orc_coo_spmv(n, nnz, row1, col1, val, z, w);
            for (int pass = 0; pass < 2; pass++) {
                for (int i = 0; i <= j; i++) {
                    const double *vi = V + (size_t)i * n;
                    double h = 0.0;
                    for (int t = 0; t < n; t++) h += vi[t] * w[t];
                    H[i * m + j] += h;
                    for (int t = 0; t < n; t++) w[t] -= h * vi[t];
                }
            }
            double hn = 0.0;
            for (int t = 0; t < n; t++) hn += w[t] * w[t];
            hn = sqrt(hn);
            if (hn < 1e-14) { j++; break; }
            H[(j + 1) * m + j] = hn;
            for (int t = 0; t < n; t++) V[(size_t)(j + 1) * n + t] = w[t] / hn;
            for (int i = 0; i < j; i++) {
                double t0 = H[i * m + j];
                H[i * m + j] = cs[i] * t0 + sn[i] * H[(i + 1) * m + j];
                H[(i + 1) * m + j] = -sn[i] * t0 + cs[i] * H[(i + 1) * m + j];
            }
            double a = H[j * m + j], c = H[(j + 1) * m + j];
            double rr = sqrt(a * a + c * c);
            cs[j] = a / rr;
            sn[j] = c / rr;
            H[j * m + j] = rr;
            H[(j + 1) * m + j] = 0.0;
            g[j + 1] = -sn[j] * g[j];
            g[j] = cs[j] * g[j];
            if (fabs(g[j + 1]) / nb <= inner_tol) { j++; break; }
        }
        inner_sum += j;
        for (int i = j - 1; i >= 0; i--) {
            double s = g[i];
            for (int t = i + 1; t < j; t++) s -= H[i * m + t] * yv[t];
            yv[i] = s / H[i * m + i];
        }
        for (int t = 0; t < n; t++) w[t] = 0.0;
        for (int i = 0; i < j; i++)
            for (int t = 0; t < n; t++) w[t] += yv[i] * V[(size_t)i * n + t];
        orc_trisolve_f32(n, Lp, Li, Lx, Up, Ui, Ux, pinv, q, w, z);
        for (int t = 0; t < n; t++) x[t] += z[t];
        orc_coo_spmv(n, nnz, row1, col1, val, x, w);
        double nr = 0.0;
        for (int i = 0; i < n; i++) { double d = b[i] - w[i]; nr += d * d; }
        rel = sqrt(nr) / nb;
        if (rel <= outer_tol) { outer++; break; }
    }
    if (total_inner) *total_inner = inner_sum;
    if (outer_done) *outer_done = outer;
    if (final_res) *final_res = orc_residual(n, nnz, row1, col1, val, b, x);
    free(V); free(H); free(cs); free(sn); free(g); free(yv);
    free(w); free(z); free(r);
    return rel <= outer_tol ? 0 : 1;
}

void orc_diag_shift(int n, long nnz, const int *row1, const int *col1,
                    const double *val, double shift, double *out)
{
    /* the reference scans the entries once per row i and stops at the first
     * (i, i) entry; one pass with a per-row "seen" mark visits the same entry */
    char *seen = (char *)calloc((size_t)n + 1, 1);
    for (long e = 0; e < nnz; e++) {
        out[e] = val[e];
        const int i = row1[e];
        if (i == col1[e] && i >= 1 && i <= n && !seen[i]) {
            seen[i] = 1;
            out[e] = val[e] + shift * fabs(val[e]);
        }
    }
    free(seen);
}

void orc_ruiz(int n, long nnz, const int *row1, const int *col1, double *val,
              int sweeps, double *R, double *C)
{
    double *mx = (double *)malloc((size_t)n * sizeof(double));
    for (int i = 0; i < n; i++) { R[i] = 1.0; C[i] = 1.0; }
    for (int it = 0; it < sweeps; it++) {
        for (int i = 0; i < n; i++) mx[i] = 0.0;
        for (long e = 0; e < nnz; e++) {
            double a = fabs(val[e]);
            if (a > mx[row1[e] - 1]) mx[row1[e] - 1] = a;
        }
        for (int i = 0; i < n; i++) mx[i] = mx[i] > 0 ? 1.0 / sqrt(mx[i]) : 1.0;
        for (int i = 0; i < n; i++) R[i] *= mx[i];
        for (long e = 0; e < nnz; e++) val[e] *= mx[row1[e] - 1];
        for (int i = 0; i < n; i++) mx[i] = 0.0;
        for (long e = 0; e < nnz; e++) {
            double a = fabs(val[e]);
            if (a > mx[col1[e] - 1]) mx[col1[e] - 1] = a;
        }
        for (int i = 0; i < n; i++) mx[i] = mx[i] > 0 ? 1.0 / sqrt(mx[i]) : 1.0;
        for (int i = 0; i < n; i++) C[i] *= mx[i];
        for (long e = 0; e < nnz; e++) val[e] *= mx[col1[e] - 1];
    }
    free(mx);
}

long orc_sell_order(int n, const int *rowptr, int C, int *order, int *lens,
                    int *offs)
{
    int maxlen = 0;
    for (int i = 0; i < n; i++) {
        int l = rowptr[i + 1] - rowptr[i];
        if (l > maxlen) maxlen = l;
    }
    int *cnt = (int *)calloc((size_t)maxlen + 2, sizeof(int));
    for (int i = 0; i < n; i++) cnt[rowptr[i + 1] - rowptr[i]]++;
    for (int l = maxlen - 1; l >= 0; l--) cnt[l] += cnt[l + 1];
    for (int i = 0; i < n; i++) order[--cnt[rowptr[i + 1] - rowptr[i]]] = i;
    free(cnt);
    int ns = (n + C - 1) / C;
    long total = 0;
    offs[0] = 0;
    for (int s = 0; s < ns; s++) {
        int sl = 0;
        for (int r = 0; r < C; r++) {
            int k = s * C + r;
            if (k < n) {
                int l = rowptr[order[k] + 1] - rowptr[order[k]];
                if (l > sl) sl = l;
            }
        }
        lens[s] = sl;
        total += (long)C * sl;
        offs[s + 1] = (int)total;
    }
    return total;
}
