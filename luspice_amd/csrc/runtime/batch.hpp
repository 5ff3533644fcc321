// batch.hpp -- device-resident batch of circuits (one circuit = batch of 1).
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../host/host_internal.h"
#include "../kernels/launch.hpp"
#include "luspice.h"

namespace lsp {

// Typed device allocation owned by a Batch.
struct DevMem {
    void *p = nullptr;
    size_t bytes = 0;
};

struct CircuitHost {
    int n = 0;
    long nnz = 0;
    std::vector<int> row, col;           // COO pattern (0-based)
    std::vector<double> val0;            // values used for the symbolic phase
    std::vector<int> q;
    LspGpLu sym;                         // CPU symbolic factors (fp64)
    LspPlan plan;
    LspSell sell;
    // fused IR step: rows below the dense block grouped by forward / backward
    // level (row lists padded to n entries; nlv + 1 level pointers each)
    std::vector<int> frow, flvp, brow, blvp;
    long voff = 0, aoff = 0;             // vector / value arena offsets
    // offsets into the concatenated device arrays
    long o_np1 = 0;                      // n+1 sized arrays
    long o_n = 0;                        // n sized arrays
    long o_slot = 0;                     // slot arrays
    long o_a = 0;                    // A-entry arrays
    long o_rel = 0;                      // update map
    long o_l = 0, o_u = 0;               // strict L / U entry arrays
    long o_sl = 0, o_slp1 = 0, o_ss = 0; // SELL slices, slices+1, slots
    long o_t = 0, o_t2 = 0;              // dense block: T, T^2
    long o_fpk = 0, o_bpk = 0;           // trisolve packets forward / backward
    long o_tw = 0;                       // wide-supernode block pairs
    long o_lf = 0, o_lb = 0;             // fused IR step: level pointers forward / backward
};

template <typename T>
struct FactorSet {                       // per storage precision
    T *val = nullptr, *lrow = nullptr, *urow = nullptr, *diag = nullptr;
    T *tailF = nullptr;                  // dense trailing blocks, storage precision
    FactorDesc<T> *fdesc = nullptr;      // device
    TriDesc<T> *tdesc = nullptr;         // device
    IrDesc<T> *idesc = nullptr;          // device (fused IR step)
    bool ready = false;                  // arrays allocated
};

// Everything upload_all() places on the device, with the sizes it derives.
// Batch inherits it, so a rebuild (new pivot sequences) frees the memory and
// resets every pointer and count with one assignment.
struct DeviceArrays {
    long ntot_ = 0, nnz_tot_ = 0;
    int cu_count_ = 0;
    int mm_alloc_ = 0;
    // structure (device)
    int64_t *d_cp_ = nullptr, *d_relp_ = nullptr, *d_soff_ = nullptr;
    int *d_dp_ = nullptr, *d_pos_ = nullptr, *d_aptr_ = nullptr, *d_asrc_ = nullptr,
        *d_aslot_ = nullptr, *d_dst_ = nullptr, *d_lrp_ = nullptr, *d_lci_ = nullptr,
        *d_urp_ = nullptr, *d_uci_ = nullptr, *d_rowat_ = nullptr, *d_q_ = nullptr,
        *d_done_ = nullptr, *d_srlen_ = nullptr,
        *d_scol_ = nullptr, *d_sperm_ = nullptr, *d_ssrc_ = nullptr;
    uint16_t *d_rel_ = nullptr;
    int *d_snend_ = nullptr, *d_ru_ = nullptr, *d_fcut_ = nullptr, *d_bcut_ = nullptr;
    int *d_dpos_ = nullptr;              // per column: position of the diagonal row q[k]
    int *d_fpk_ = nullptr, *d_bpk_ = nullptr, *d_twide_ = nullptr;
    double *d_tailW_ = nullptr;
    // Ruiz equilibration (FP16 factor storage)
    int *d_arow_ = nullptr, *d_acol_ = nullptr;
    long *d_raoff_ = nullptr, *d_rvoff_ = nullptr;
    double *d_ains_ = nullptr, *d_rscale_ = nullptr, *d_cscale_ = nullptr;
    long *d_sidx_ = nullptr;             // factor-only shift: first diagonal entry per row
    long nsidx_ = 0;
    unsigned long long *d_rmx_ = nullptr;
    double *d_ain_ = nullptr, *d_sval64_ = nullptr;
    float *d_sval32_ = nullptr;
    __half *d_sval16_ = nullptr;         // FP16 SELL values (krylov_spmv == 2 only)
    SellDesc<__half> *d_sdesc16_ = nullptr;
    SellDesc<double> *d_sdesc64_ = nullptr;
    SellDesc<float> *d_sdesc32_ = nullptr;
    int2 *d_ftask_ = nullptr, *d_fwtask_ = nullptr, *d_bwtask_ = nullptr, *d_stask_ = nullptr;
    long nslot_tot_ = 0, nl_tot_ = 0, nu_tot_ = 0, nsell_tot_ = 0;
    int nstask_ = 0;
    int *d_ctr_ = nullptr;               // [0..1] factor, [2..3] fwd, [4..5] bwd
    int *d_epoch_ = nullptr;             // epoch counters: [0] factor, [1] fwd, [2] bwd
    int *d_abort_ = nullptr;             // [0] factor, [1] trisolve
    int *d_mstat_ = nullptr;             // kMstat per circuit
    int *d_dbg_ = nullptr;               // first stuck wait of the refactor
    long long *d_timeline_ = nullptr;    // optional refactor timeline (LSP_REFACTOR_TIMELINE)
    long long *d_tri_tl_ = nullptr;      // optional trisolve timeline (LSP_TRISOLVE_TIMELINE)
    double *d_scratch_ = nullptr;
    long scratch_stride_ = 0;
    int fblocks_ = 0, tblocks_f_ = 0, tblocks_b_ = 0;
    int nq_f_ = 1, nq_t_ = 1;                          // queue counters: refactor, solves
    long width_f_ = 0, width_tf_ = 0, width_tb_ = 0;   // widest merged level
    // dense trailing blocks
    int *d_tail_ds_ = nullptr, *d_tail_lcut_ = nullptr, *d_tail_ctr_ = nullptr, *d_tails_ = nullptr;
    double *d_tailS_ = nullptr;
    unsigned long long *d_colmax_ = nullptr;
    int *d_dflag_ = nullptr;                   // per dense job: the last diagonal block factored
    DenseJob *d_djobs_ = nullptr;
    int njobs_ = 0, max_tail_T_ = 0;
    long nt_tot_ = 0, nt2_tot_ = 0;
    long nfw_tot_ = 0, nbw_tot_ = 0;     // row tasks of the sparse solves
    FactorSet<float> f32_;
    FactorSet<double> f64_;
    FactorSet<__half> f16_;
    // vectors and Krylov state
    double *d_b_ = nullptr, *d_x_ = nullptr, *d_r_ = nullptr, *d_w_ = nullptr,
           *d_z_ = nullptr, *d_t_ = nullptr, *d_ty_ = nullptr, *d_tz_ = nullptr, *d_typ_ = nullptr,
           *d_tzp_ = nullptr, *d_V_ = nullptr,
           *d_Z_ = nullptr;
    long *d_voff_ = nullptr;
    int4 *d_chunks_ = nullptr;
    int *d_cbeg_ = nullptr;
    int nchunks_ = 0;
    double *d_partial_ = nullptr, *d_H_ = nullptr, *d_cs_ = nullptr, *d_sn_ = nullptr,
           *d_g_ = nullptr, *d_yv_ = nullptr, *d_beta_ = nullptr, *d_nb_ = nullptr,
           *d_hn_ = nullptr, *d_hcol_ = nullptr;
    int *d_act_ = nullptr, *d_jfin_ = nullptr, *d_conv_ = nullptr, *d_cycled_ = nullptr,
        *d_nact_ = nullptr, *d_subset_ = nullptr;
    // fused IR step: rows below the dense block by level, and per-circuit state
    int *d_frow_ = nullptr, *d_flvp_ = nullptr, *d_brow_ = nullptr, *d_blvp_ = nullptr;
    int *d_irconv_ = nullptr;
    double *d_rn2_ = nullptr, *d_bn2_ = nullptr;
    int ir_lds_n_ = 0;                   // doubles of LDS for the work vector (0: not set up)
};

class Batch : private DeviceArrays {
public:
    ~Batch();
    int create(const lsp_circuit_in *c, int nc, const lsp_options *opt);
    int prepare(const lsp_circuit_in *c, int nc, const lsp_options *opt);   // host symbolic only
    // A prepared batch made of circuits moved out of other batches (their host
    // structures, no symbolic recomputation); the sources are left empty.
    int adopt(const std::vector<std::pair<Batch *, int>> &picks);
    // Circuits that refactor() and solve() work on (nullptr: all).
    int set_active(const int *mask);
    int upload_all();                                                         // device phase
    size_t estimate_device_bytes() const;
    size_t device_bytes() const;
    int set_stream(hipStream_t s);
    int set_values(const double *vals, bool on_device);
    int check_prec(int prec);
    bool repivot_wanted(const int *mstat_words) const;
    int refactor(int prec, lsp_refactor_report *rep, const std::vector<int> *subset = nullptr);
    int launch_refactor_prec(int prec, const int *d_active);
    int record_step(int prec);
    int trisolve(const double *b, double *x, bool on_device);
    int solve(const double *b, double *x, bool on_device, lsp_solve_report *rep);
    int spmv(const double *x, double *y, bool on_device, int prec);
    int info(int c, lsp_circuit_info *inf) const;
    int export_factors(int c, int *pinv, int *Lp, int *Li, double *Lx, int *Up, int *Ui,
                       double *Ux);
    int device_vectors(double **b, double **x);
    int graph_step(int prec, lsp_refactor_report *rrep, lsp_solve_report *srep);
    int copy_vector(const double *p, bool to_rhs);
    int last_times(double *rf, double *fw, double *bw) const;
    int kernel_stats(lsp_kernel_stats *st, int reset);
    int set_ir_params(int outer, int inner, double inner_tol, double outer_tol);
    // repivot policy of later refactors (lsp_options::repivot)
    int set_repivot(int mode)
    {
        if (mode < 0 || mode > 3) return LSP_E_ARG;
        opt_.repivot = mode;
        return LSP_OK;
    }
    int diag_stats(int c, double *min_abs, double *max_abs);
    const std::vector<double> &residual_history() const { return res_hist0_; }
    const CircuitHost &circuit(int c) const { return circ_[c]; }
    long total_n() const { return ntot_; }
    long total_nnz() const { return nnz_tot_; }
    int ncircuits() const { return (int)circ_.size(); }

    // device-resident pieces used by the serving loop and the bench
    int gmres_ir(const double *d_b, double *d_x, int prec, const std::vector<char> *subset,
                 lsp_solve_report *rep, bool keep_x = false);
    // Fused Richardson steps x += M^-1 (b - A x) from x = 0, one launch per step
    // for every unconverged circuit, up to max_steps; conv (host) gets 1 for
    // each circuit that met ir_outer_tol.
    int fused_ir(const double *d_b, double *d_x, int prec, const std::vector<char> *subset, int max_steps,
                 std::vector<int> &conv, lsp_solve_report *rep);
    // One fused step on the given vectors for every circuit (no convergence
    // skip); rel_res gets ||b - A x|| / ||b|| before the update, per circuit.
    int ir_step(const double *b, double *x, bool on_device, double *rel_res);
    int launch_trisolve_prec(int prec, const double *d_b, double *d_x, const int *active,
                             const int *d_nact = nullptr);
    int launch_spmv_krylov(const double *d_xv, double *d_y, const int *active);
    int launch_spmv64(const double *d_x, double *d_y, const double *d_b, const int *active);
    hipStream_t stream() const { return stream_; }
    int cur_prec() const { return cur_prec_; }
    std::string last_error;

private:
    template <typename T> int ensure_factor_set(FactorSet<T> &fs);
    template <typename T> int ir_launch(FactorSet<T> &fs, const double *d_b, double *d_x, double tol2, bool first);
    int ir_launch_prec(int prec, const double *d_b, double *d_x, double tol2, bool first);
    int ir_prepare();
    bool use_fused(int prec) const;      // ir_method 1 and factors the fused step reads
    int fused_steps() const;
    int refactor_pass(int prec, const std::vector<int> *subset, std::vector<int> &ms);
    int repivot(const std::vector<int> &cands, int *changed);
    void release_device();
    template <typename T> int refactor_t(FactorSet<T> &fs, const int *d_active);
    template <typename T> int trisolve_t(FactorSet<T> &fs, const double *d_b, double *d_x,
                                         const int *d_active, const int *d_nact = nullptr);
    void *dalloc(size_t bytes);
    void dfree(void *p);
    int alloc_krylov(int mm);
    template <typename X> X *upload(const std::vector<X> &v);
    int check(hipError_t e, const char *what);
    int sync_read_int(const int *d, int *h, int count);
    KrylovArgs kargs() const;
    void account_trisolve();

    lsp_options opt_{};
    hipStream_t stream_ = nullptr;
    bool own_stream_ = false;
    std::vector<CircuitHost> circ_;
    std::vector<int> active_;             // set_active mask (empty: all circuits)
    std::vector<DevMem> mem_;
    int cur_prec_ = -1;
    bool uploaded_ = false;              // upload_all done (cleared by release_device)

    // HIP graph of one Newton step (graph_step)
    hipGraph_t graph_ = nullptr;
    hipGraphExec_t graph_exec_ = nullptr;
    int graph_prec_ = -1;
    bool capturing_ = false;
    int *h_pin_ = nullptr;               // pinned host words for flag reads
    const char *refactor_tl_path_ = nullptr;   // LSP_REFACTOR_TIMELINE, read at creation
    const char *trisolve_tl_path_ = nullptr;   // LSP_TRISOLVE_TIMELINE, read at creation
    hipEvent_t ev_[6] = {};
    float t_refactor_ = 0, t_fwd_ = 0, t_bwd_ = 0;
    bool tri_pending_ = false;
    std::vector<double> res_hist0_;      // ||r||/||b|| of circuit 0 at each restart
    lsp_kernel_stats stats_{};
};

}  // namespace lsp

struct lsp_batch {
    lsp::Batch impl;
};
