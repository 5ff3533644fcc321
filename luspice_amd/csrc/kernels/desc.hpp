// desc.hpp -- per-matrix device descriptors shared by kernels and runtime.
//
// Every kernel works on a table of matrices so that one launch can serve one
// circuit or a whole batch of circuits of uneven size: a task is (matrix,
// index) and the matrix's arrays come from its descriptor.
#pragma once

#include <stdint.h>

namespace lsp {

// Abort codes raised by the persistent kernels.
enum : int {
    kErrNone = 0,
    kErrSpin = 1,        // a dependency wait exceeded its spin limit
    kErrCapacity = 2,    // a column/row exceeded its scratch capacity
};

// Ints in one persistent kernel's queue block: up to 4 queue counters, a
// cache line each, and the count of waves finished.
constexpr int kQueueInts = 128;

// Status words per matrix written by the refactor (FactorDesc::mstat).
constexpr int kMstat = 8;

// Trisolve task codes at or above kTriPacket name a packet of short rows
// (each lane of a packet row gathers up to kTriPacketLane of its entries);
// codes in [kTriWide, kTriPacket) name one 64-row block of a wide supernode.
constexpr int kTriPacket = 1 << 30;
constexpr int kTriWide = 1 << 29;
constexpr int kTriPacketLane = 4;

// Numeric refactorization over the static column patterns.
template <typename T>
struct FactorDesc {
    const int64_t *cp;      // n+1 slot pointers
    const int *dp;          // n diagonal slots (absolute)
    const int *pos;         // slot -> row position
    const int *aptr;        // n+1 q-order CSC pointers of A
    const int *asrc;        // A entry -> input value index
    const int *aslot;       // A entry -> local slot (-1 skipped duplicate)
    const double *ain;      // input values (fp64, caller order)
    const int64_t *relp;    // n+1 segment-map pointers
    const uint16_t *rel;    // segment maps: local slots of each run's last L rows
    const int *snend;       // n: one past the last column of j's supernode
    const int *dst;         // slot -> index in lrow / urow / diag
    const int *dpos;        // n: position of the diagonal row q[k] (pinv[q[k]])
    T *val;                 // column storage of L and U
    T *lrow, *urow, *diag;  // row-wise copies for the trisolve
    int *done;              // per-column epoch flags
    int *mstat;             // kMstat ints: bad pivots, first bad column,
                            // first zero-pivot column, non-finite count,
                            // columns whose pivot splu would not pick, columns
                            // too close to call, first of either, pad
    // dense trailing block (tail_T = 0: none, tail_k0 = n)
    const int *tail_ds;     // tail_T: sparse dependencies of tail column c
    double *tail_S;         // tail_T^2 fp64 work block, column-major
    unsigned long long *tail_colmax;   // tail_T: max |L(i,c)| as bit patterns
    T *tail_F;              // tail_T^2 factor copy in storage precision
    int n;
    int tail_k0;
    int tail_T;
    int pad_;
};

// Row-gather triangular solves.  Vectors are passed per launch as arena
// bases (original order for b and x, position order for the work vectors),
// matrix m's slice starting at voff[m].
template <typename T>
struct TriDesc {
    const int *lrp, *lci;   // strict L by rows (positions)
    const int *urp, *uci;   // strict U by rows
    const T *lrow, *urow, *diag;
    const int *rowat;       // position -> original row (P)
    const int *q;           // position -> original column (Q)
    const int *twide;       // (supernode start, block) pairs of the wide-supernode block tasks
    // supernodes (rows below tail_k0)
    const T *val;           // column storage of L and U (FactorDesc::val)
    const int *dp;          // n diagonal slots (absolute)
    const int *snend;       // n: one past the last row of i's supernode
    const int *ru;          // n: U(j - ru[j] .. j-1, j) is column j's run in its supernode
    const int *fcut;        // n: end of row i's L entries left of its supernode
    const int *bcut;        // n: start of row i's U entries right of its supernode
    const int *fpk, *bpk;   // row packets forward / backward: [g, count, row codes]
    // dense trailing block: rows [tail_k0, n) are solved with tail_F
    const T *tail_F;        // tail_T^2, column-major, unit-L below / U on and above
    const int *tail_lcut;   // tail_T: end of the sparse part of L row tail_k0 + r
    int *tail_ctr;          // 4 ints: block ticket counter, exit counter, pad
    int n;
    int tail_k0;
    int tail_T;
    int pad_;
};

// One dense trailing block to factor (fp64 work, fixed pivot order).
struct DenseJob {
    double *S;              // T x T, column-major, leading dimension T
    double *W;              // 2 x 64 x 64 workspace: inverses of the current diagonal block
    unsigned long long *colmax;
    int *flag;              // panel whose diagonal block (and W) is final, during an update launch
    int T;
    int mat;                // matrix index (descriptor, mstat, active mask)
};

// SELL-C-sigma SpMV with C = 64 (one wave per slice, one lane per row).
template <typename V>
struct SellDesc {
    const int64_t *off;     // nslices+1
    const int *rlen;        // per sorted row: stored length
    const int *col;         // slots
    const V *val;           // slots
    const int *perm;        // sorted row -> original row
    int nslices;
    int n;
};

// Fused IR step (irstep.hip): one circuit's arrays.
template <typename T>
struct IrDesc {
    SellDesc<double> A;     // A with fp64 values, for the residual
    const int *lrp, *lci;   // strict L by rows (positions)
    const int *urp, *uci;   // strict U by rows
    const T *lrow, *urow, *diag;
    const T *tail_F;        // dense trailing block, column-major (tail_T = 0: none)
    const int *rowat;       // position -> original row (P)
    const int *q;           // position -> original column (Q)
    const int *tail_lcut;   // tail_T: end of the sparse part of L row tail_k0 + r
    const int *frow, *flvp; // rows below tail_k0 by forward level; nlv_f + 1 level pointers
    const int *brow, *blvp; // the same for the backward solve
    long voff;              // first entry of the circuit in the vector arenas
    int n, tail_k0, tail_T, nlv_f, nlv_b, pad_;
};

// Status words written by the persistent kernels.
struct KStatus {
    int abort_code;         // kErr* (0 = none)
    int bad_pivot_count;    // columns whose pivot failed the tolerance test
    int first_bad_pivot;    // smallest failing column (INT_MAX if none)
    int first_singular;     // smallest column with a zero pivot (INT_MAX)
    int bad_matrix;         // matrix index of the first failure (-1)
    int pad_[3];
};

}  // namespace lsp
