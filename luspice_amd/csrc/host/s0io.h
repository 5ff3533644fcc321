/* s0io.h -- s0 matrix files as read by the reference's CLIs. */
#ifndef LSP_S0IO_H
#define LSP_S0IO_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int n;
    long nnz;
    int *row, *col;     /* 1-based, file order */
    double *val;
    double *b;          /* n+1 entries, b[1..n]; zero / missing -> 1.0 */
} lsp_s0;

int  lsp_s0_read(const char *mtx, const char *rhs, lsp_s0 *M);
void lsp_s0_free(lsp_s0 *M);
int  lsp_flt_pin(lsp_s0 *M, const char *flt, double pin);
int  lsp_s0_write(const char *path, int n, long nnz, const int *row0, const int *col0,
                  const double *val);
int  lsp_rhs_write(const char *path, int n, const double *b);

#ifdef __cplusplus
}
#endif
#endif
