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
    const int64_t *relp;    // n+1 update-map pointers
    const uint16_t *rel;    // update map (local slots)
    const int *dst;         // slot -> index in lrow / urow / diag
    T *val;                 // column storage of L and U
    T *lrow, *urow, *diag;  // row-wise copies for the trisolve
    int *done;              // per-column epoch flags
    int *mstat;             // 4 ints: bad pivots, first bad column,
                            // first zero-pivot column, non-finite count
    int n;
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
    int *fdone, *bdone;     // per-row epoch flags
    int n;
    int pad_;
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
