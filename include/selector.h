/* selector.h -- format x precision selector.
 * Same types, enums and functions as the reference's src/selector.h:5-60;
 * adds an extern "C" guard so C++ and HIP hosts can include it (the ABI is
 * unchanged).  Implementation: luspice_amd/csrc/host/selector.c.
 */
#ifndef DSH_SELECTOR_H
#define DSH_SELECTOR_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int n;
    long nnz;
    double avg_rowlen, max_rowlen, p50_rowlen;
    double row_skew;        /* max / mean row length */
    double kappa_ruiz;      /* condition estimate after Ruiz (<= 0 unknown) */
    double kappa_est;       /* raw condition estimate (<= 0 unknown) */
    double dyn_range_log10; /* log10(max|A| / min|A_ij != 0|) (<= 0 unknown) */
    int    singular_flag;   /* 1 = known singular */
} dsh_matrix_info;

enum { DSH_FMT_CSR = 0, DSH_FMT_SELL_C_SIGMA = 1, DSH_FMT_ELL = 2 };
enum { DSH_PREC_FP32 = 0, DSH_PREC_FP64 = 1 };
enum { DSH_MP_PFAST = 0, DSH_MP_PSAFE = 1, DSH_MP_PHARD = 2, DSH_MP_PSING = 3 };

typedef struct {
    int spmv_format;        /* DSH_FMT_* */
    int sigma;              /* SELL sort window */
    int use_ruiz;
    int use_shift;
    double shift;           /* 1e-7 */
    int factor_prec;        /* DSH_PREC_* */
    int ir_outer, ir_inner;
    double ir_rtol;
    int mp_path;            /* DSH_MP_* */
    int pre_ruiz_hint;
    int fp16_candidate;
} dsh_config;

void dsh_matrix_analyze(int n, long nnz, const int *rowptr, dsh_matrix_info *info);
void dsh_format_select(const dsh_matrix_info *info, dsh_config *cfg);
const char *dsh_fmt_name(int fmt);
void dsh_mp_initial(const dsh_matrix_info *info, dsh_config *cfg);
int  dsh_mp_escalate(int cur);
const char *dsh_mp_name(int p);

#ifdef __cplusplus
}
#endif
#endif
