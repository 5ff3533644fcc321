/* oracle/oracle.h -- CPU restatement of the reference's sparse-LU hot path.
 * TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker; never by luspice_amd/.
 */
#ifndef LUSPICE_ORACLE_H
#define LUSPICE_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int n;
    int *pinv;
    int *Lp, *Li; double *Lx; long lcap, lnnz;
    int *Up, *Ui; double *Ux; long ucap, unnz;
    int singular_col;
} orc_lu;

/* lu_oracle.c */
int  orc_lu_factor(int n, const int *Ap, const int *Ai, const double *Ax,
                   const int *q, double tol, orc_lu *F);
void orc_lu_free(orc_lu *F);
void orc_lu_solve_perm(int n, const orc_lu *F, const int *q,
                       const double *b, double *x);
int  orc_cpu_lu_ref(int n, const int *Ap, const int *Ai, const double *Ax,
                    const int *q, const double *b, double tol,
                    double *x, long *lnnz, long *unnz);
void orc_lu_export(const orc_lu *F, int *pinv, int *Lp, int *Li, double *Lx,
                   int *Up, int *Ui, double *Ux);
int  orc_lu_factor_export(int n, const int *Ap, const int *Ai, const double *Ax,
                          const int *q, double tol,
                          int *pinv, int *Lp, int *Li, double *Lx, long lcap,
                          int *Up, int *Ui, double *Ux, long ucap,
                          long *lnnz, long *unnz);
int  orc_level_compute(int n, const int *Up, const int *Ui, int *level,
                       int *nlevels);
int  orc_level_hist(int n, const int *level, int nlevels, int *hist);

/* ir_oracle.c */
void orc_coo_spmv(int n, long nnz, const int *row1, const int *col1,
                  const double *val, const double *x, double *y);
double orc_residual(int n, long nnz, const int *row1, const int *col1,
                    const double *val, const double *b, const double *x);
double orc_rel_l2(int n, const double *a, const double *b);
void orc_csr_spmv(int n, const int *rowptr, const int *colidx,
                  const double *val, const double *x, double *y);
void orc_trisolve_f32(int n, const int *Lp, const int *Li, const float *Lx,
                      const int *Up, const int *Ui, const float *Ux,
                      const int *pinv, const int *q, const double *b,
                      double *x);
int  orc_gmres_ir(int n, long nnz, const int *row1, const int *col1,
                  const double *val, const double *b,
                  const int *Lp, const int *Li, const float *Lx,
                  const int *Up, const int *Ui, const float *Ux,
                  const int *pinv, const int *q,
                  int max_outer, int m, double inner_tol, double outer_tol,
                  double *x, int *total_inner, int *outer_done,
                  double *final_res);
void orc_diag_shift(int n, long nnz, const int *row1, const int *col1,
                    const double *val, double shift, double *out);
void orc_ruiz(int n, long nnz, const int *row1, const int *col1, double *val,
              int sweeps, double *R, double *C);
long orc_sell_order(int n, const int *rowptr, int C, int *order, int *lens,
                    int *offs);

#ifdef __cplusplus
}
#endif
#endif
