// launch.hpp -- launch argument blocks and host launchers of the kernels.
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "desc.hpp"

namespace lsp {

struct FactorLaunch {
    const void *mats;       // FactorDesc<T>[nmat]
    const int2 *task;       // (matrix, column) in level order
    const int *active;      // per-matrix mask (nullptr = all)
    int ntask;
    int *epoch;             // device epoch counter: read at start, bumped by the last wave
    int *ctr;               // kQueueInts: queue counters, waves finished
    int *abort_word;
    double *scratch;        // per-wave global fallback
    long scratch_stride;    // doubles per wave slice
    int cap;                // LDS doubles per wave
    int total_waves;
    double ptol;
    double stol;            // > 0: also classify each pivot against splu's choice at this tol
    int *dbg;               // 8 ints: first stuck wait (dep, seen, epoch, spins, column)
    long long *timeline;    // optional, 4 per task (100 MHz wall clock): start, flag
                            // published (end for a dense-block column), end of the
                            // last dependency wait, column
    int poll_backoff;       // dependency polls back off to at most poll_backoff + 1 sleeps
    int nq;                 // queue counters (1, 2 or 4) in the ctr block
};

struct TriLaunch {
    const void *mats;       // TriDesc<T>[nmat]
    const int2 *task;       // (matrix, row) in level order
    const long *voff;       // per-matrix vector offsets
    const int *active;      // per-matrix mask (nullptr = all)
    const double *b;        // right-hand sides (original row order)
    double *x;              // solutions (original column order)
    double *y, *z;          // work vectors (position order), unset before the solve
    double *yp, *zp;        // partial sums of supernode rows (position order), unset before the solve
    const double *rs, *cs;  // optional row / column scales (Ruiz): b_s = R b, x = C z
    const int *n_active;    // optional: when *n_active == 0 no circuit is active, tasks are skipped
    long long *timeline;    // optional diagnostic, 2 per task: ticket granted, task finished
    int ntask;
    int *ctr;               // kQueueInts: queue counters, waves finished
    int *abort_word;
    int total_waves;
    int nq;                 // queue counters (1, 2 or 4) in the ctr block
};

struct KrylovArgs {
    int nmat;
    int mmax;               // restart length m
    long ntot;              // arena length
    const long *voff;       // per matrix
    const int4 *chunks;     // (matrix, r0, r1, 0)
    int nchunks;
    const int *cbeg;        // nmat+1 chunk ranges per matrix
    double *partial;        // nchunks * (mmax+1)
    double *H;              // nmat * (mmax+1) * mmax, H[i*mmax + j]
    double *cs, *sn;        // nmat * mmax
    double *g;              // nmat * (mmax+1)
    double *yv;             // nmat * mmax
    double *beta;           // nmat: ||r|| at cycle start
    double *nb;             // nmat: ||b||
    double *hn;             // nmat: ||w|| of the current step
    double *hcol;           // nmat * (mmax+1): scratch for one Arnoldi column
    int *act;               // nmat: 1 while the inner cycle runs
    int *jfin;              // nmat: inner steps taken in the cycle
    int *conv;              // nmat: 1 once the outer test passed
    int *cycled;            // nmat: 1 if the cycle produced an update
    double inner_tol, outer_tol;
};

// One fused IR step over the circuits of a batch (irstep.hip).
struct IrStepArgs {
    const void *descs;      // IrDesc<T>[nmat]
    const double *b;        // right-hand sides (arena, original row order)
    double *x;              // solutions (arena, original column order), updated in place
    double *r;              // residual arena (original row order)
    double *yw;             // work arena (position order) for circuits larger than lds_n
    int *conv;              // per circuit: 1 = converged (the launch leaves it alone)
    double *rn2, *bn2;      // per circuit: ||b - A x||^2 and ||b||^2 at its last step
    double tol2;            // squared relative residual tolerance (< 0: every circuit steps)
    int lds_n;              // doubles of LDS per workgroup for the work vector
    int first;              // 1: x starts at 0 (r = b, x is overwritten rather than updated)
};
template <typename T> hipError_t launch_ir_step(const IrStepArgs &A, int nmat, hipStream_t st);
template <typename T> hipError_t ir_step_set_lds(int lds_n);

// factor.hip
template <typename T> hipError_t launch_refactor(const FactorLaunch &L, int blocks, hipStream_t st);
template <typename T> hipError_t refactor_set_lds(int cap);
hipError_t launch_reset_mstat(int *mstat, int nmat, int *abort_word, int *dbg, hipStream_t st);
template <typename T>
hipError_t launch_fill_unset(const void *mats, const int *active, int nmat, long max_slots, hipStream_t st);

// dense trailing blocks (dense.hip)
hipError_t dense_set_lds();
hipError_t launch_dense_lu(const DenseJob *jobs, int njobs, int maxT, const int *active, int *abort_word,
                           hipStream_t st);
template <typename T>
hipError_t launch_dense_finish(const DenseJob *jobs, int njobs, int maxT, const void *mats,
                               const int *active, double ptol, double stol, hipStream_t st);
template <typename T>
hipError_t launch_tail_solve(const void *mats, const int *tails, int ntails, int maxT, const long *voff,
                             const int *active, const double *b, double *y, double *yp, double *z,
                             double *x, bool forward, int *abort_word, const double *rs, const double *cs,
                             hipStream_t st);

// trisolve.hip
template <typename T>
hipError_t launch_trisolve_half(const TriLaunch &L, bool forward, int blocks, hipStream_t st);
hipError_t launch_solve_unset(const TriLaunch &L, long n, hipStream_t st);

// spmv.hip
template <typename V>
hipError_t launch_sell_spmv(const SellDesc<V> *mats, const int2 *tasks, int ntask,
                            const int64_t *voff, const double *x, double *y, const double *b,
                            const int *active, hipStream_t st);
hipError_t launch_csr_spmv(int n, const int *rowptr, const int *colidx, const double *values,
                           const double *x, double *y, hipStream_t st);
hipError_t launch_csr_spmv_f32(int n, const int *rowptr, const int *colidx, const float *values,
                               const double *x, double *y, hipStream_t st);
template <typename Vout>
hipError_t launch_gather_cast(long count, const int *idx, const double *in, Vout *out,
                              hipStream_t st);

// gmres.hip
hipError_t k_chunk_dots(const KrylovArgs &K, const double *V, const double *w, int nv, hipStream_t s);
hipError_t k_reduce(const KrylovArgs &K, int nv, int mode, int j, double *out, hipStream_t s);
hipError_t k_chunk_update(const KrylovArgs &K, const double *V, double *w, int nv, hipStream_t s);
hipError_t k_chunk_scale(const KrylovArgs &K, const double *src, double *dst, const double *den,
                         hipStream_t s);
hipError_t k_start_cycle(const KrylovArgs &K, int *n_active, hipStream_t s);
hipError_t k_givens(const KrylovArgs &K, int j, int *n_active, hipStream_t s);
hipError_t k_backsolve(const KrylovArgs &K, hipStream_t s);
hipError_t k_chunk_combine(const KrylovArgs &K, const double *V, double *t, hipStream_t s);
hipError_t k_chunk_axpy(const KrylovArgs &K, double *x, const double *z, hipStream_t s);
hipError_t k_chunk_fill(const KrylovArgs &K, double *x, double v, hipStream_t s);
hipError_t k_activate(const KrylovArgs &K, hipStream_t s);
hipError_t k_set_all_active(const KrylovArgs &K, int v, hipStream_t s);
hipError_t k_zero(double *p, long n, hipStream_t s);
hipError_t k_zero_int(int *p, long n, hipStream_t s);
hipError_t k_copy(const double *src, double *dst, long n, hipStream_t s);

// Ruiz equilibration of every circuit's values (one workgroup per circuit).
struct RuizArgs {
    const double *ain;      // values, caller order, all circuits
    double *aout;           // scaled values R A C, same layout
    const int *row, *col;   // per entry: circuit-local original row / column
    const long *aoff;       // ncirc+1 entry offsets
    const long *voff;       // ncirc+1 vector offsets
    double *rs, *cs;        // row / column scales (vector layout)
    unsigned long long *mx; // scratch, vector layout
    const int *active;      // per-circuit mask (nullptr = all)
    int sweeps;
    int pad_;
};
hipError_t launch_ruiz(const RuizArgs &R, int ncirc, hipStream_t st);
hipError_t launch_round_f32(const double *in, float *out, long count, hipStream_t st);
// Factor-only diagonal shift: dst[e] = src[e] + shift*|src[e]| for e in idx[0..nidx).
hipError_t launch_shift(const double *src, double *dst, const long *idx, long nidx, double shift,
                        hipStream_t st);

}  // namespace lsp
