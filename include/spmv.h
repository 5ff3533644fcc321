/* spmv.h -- s0 file I/O and CSR SpMV with the reference's C API
 * (src/spmv.h:18-60 of the reference; same struct, names and return codes).
 *
 * s0 matrix format: first line "n <n>", then "A <i> <j> <value>" lines,
 * 1-based; other lines are ignored.  RHS format: "n <n>" then "RHS <i> <v>".
 *
 * Differences from the reference that a caller can observe: none in results;
 * `stream` is a hipStream_t passed as void*, and spmv_launch runs on gfx950.
 * Implementation: luspice_amd/csrc/host/s0io.c (file I/O) and
 * luspice_amd/csrc/runtime/spmv_api.cpp (device buffers, launches).
 */
#ifndef RACE7_SOLVER_SPMV_H
#define RACE7_SOLVER_SPMV_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int n;            /* order */
    long long nnz;    /* stored entries */
    int *rowptr;      /* n+1 */
    int *colidx;      /* nnz, 0-based */
    double *values;   /* nnz */
} CSRMatrix;

/* Read an s0 .mtx into CSR (0-based, rows in file order).  0 on success. */
int read_s0_to_csr(const char *path, CSRMatrix *A);
/* Free what read_s0_to_csr allocated. */
void csr_free(CSRMatrix *A);
/* Read an s0 RHS into x[0..n).  0 on success (at least one entry read). */
int read_rhs(const char *path, int n, double *x);
/* One y = A x round trip through the device; time_ms (optional) is the
 * kernel time after one warm-up.  0 on success. */
int spmv_host(const CSRMatrix *A, const double *x, double *y, double *time_ms);
/* Allocate device CSR buffers and copy A in.  0 on success. */
int csr_alloc_device(const CSRMatrix *A, int **d_rowptr, int **d_colidx, double **d_values);
/* y = A x on the device, one thread per row (stream = hipStream_t or 0). */
void spmv_launch(int n, const int *d_rowptr, const int *d_colidx, const double *d_values,
                 const double *d_x, double *d_y, void *stream);

#ifdef __cplusplus
}
#endif

#endif
