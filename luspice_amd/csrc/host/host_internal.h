// host_internal.h -- host-side structures shared by the C++ runtime.
//
// Index spaces used throughout:
//   original   row/column ids of A as read from the s0 file (0-based here)
//   q-order    column k of A(:,q) is original column q[k]
//   position   row position after partial pivoting: pinv[orig_row] = position
// The factorization is P*A(:,q) = L*U (src/third_party/splu.h:3-8 of the
// reference): L unit lower triangular, U upper triangular, both indexed by
// position.
#pragma once

#include <cstdint>
#include <vector>

// ---------------------------------------------------------------------------
// COO with 0-based indices; b has n entries.
struct LspCoo {
    int n = 0;
    std::vector<int> row, col;
    std::vector<double> val;
    std::vector<double> b;
};

// ---------------------------------------------------------------------------
// colamd.cpp
int lsp_colamd_order(int n_row, int n_col, const int *Ap, const int *Ai, int *q);

// ---------------------------------------------------------------------------
// Gilbert-Peierls LU with the reference's pivoting rule (symbolic.cpp).
struct LspGpLu {
    int n = 0;
    std::vector<int> pinv;                // orig row -> position
    std::vector<int> Lp, Li;              // L CSC, Li = positions, unit diag first
    std::vector<double> Lx;
    std::vector<int> Up, Ui;              // U CSC, Ui = positions, diag last
    std::vector<double> Ux;
    long lnnz = 0, unnz = 0;
    int singular = -1;                    // -1 ok, else column index
};

// A in q-order CSC with original row ids (Ap/Ai/Ax), diag_row[k] = q[k]
// (the row preferred as pivot for column k), tol as in splu.h:185.
int lsp_gp_factor(int n, const int *Ap, const int *Ai, const double *Ax,
                  const int *diag_row, double tol, LspGpLu &F);
// x = Q U^-1 L^-1 P b  (src/third_party/splu.h:210-239)
void lsp_gp_solve(const LspGpLu &F, const int *q, const double *b, double *x);

// ---------------------------------------------------------------------------
// Static structure for the GPU numeric refactorization and trisolve.
constexpr int kTriPacketBase = 1 << 30;   // trisolve task codes at or above: packets
constexpr int kTriWideBase = 1 << 29;     // codes in [kTriWideBase, kTriPacketBase): wide supernode blocks
constexpr int kTriPacketPerLane = 4;      // entries per lane of a packet row
constexpr int kTriBlockRows = 64;         // rows of one triangle task (one wave, one row per lane)

struct LspPlan {
    int n = 0;
    long nnz_a = 0;
    // column patterns in position space, sorted ascending:
    //   slots [cp[k], dp[k])     U(:,k) off-diagonal (positions < k)
    //   slot  dp[k]              diagonal (position k)
    //   slots (dp[k], cp[k+1])   L(:,k) off-diagonal (positions > k)
    std::vector<int64_t> cp;              // n+1
    std::vector<int> dp;                  // n
    std::vector<int> pos;                 // nslots
    // A scatter: q-order CSC entries of A -> local slot in their column
    std::vector<int> aptr;                // n+1 (q-order CSC column pointers)
    std::vector<int> asrc;                // nnz_a: index into the caller's value array
    std::vector<int> aslot;               // nnz_a: local slot, -1 = overwritten duplicate
    // Supernodes of L among positions [0, tail_k0): j and j+1 share one when
    // L(:, j+1) = L(:, j) minus row j+1.  snend[j] = one past the last column
    // of j's supernode (j+1 for a single column).
    std::vector<int> snend;               // n
    // Segment update maps.  The dependencies of column k that lie in one
    // supernode D form a run of consecutive U slots (positions j_s .. e-1 with
    // e = min(snend, k)); every dependency's L rows are the run's later
    // positions followed by L(:, e-1).  For each run, in slot order, the map
    // holds the local slot in k of every row of L(:, e-1).
    std::vector<int64_t> relp;            // n+1
    std::vector<uint16_t> rel;            // local slots (max_col <= 65535)
    int64_t n_segments = 0;               // runs over all columns
    int max_col = 0;                      // max column pattern length
    // dependency levels of the factorization (U pattern, level.c semantics)
    std::vector<int> level;               // n
    int nlevels = 0;
    std::vector<int> ftask;               // columns sorted by (level, k)
    // row-wise copies for the row-gather trisolve
    std::vector<int> lrp, lci;            // L strictly lower, CSR by position
    std::vector<int> urp, uci;            // U strictly upper, CSR by position
    std::vector<int> dst;                 // nslots: index into lrow/urow/diag
    std::vector<int> lvf, lvb;            // row levels forward / backward
    int nlv_f = 0, nlv_b = 0;             // dependency levels of the solve tasks
    // Trisolve tasks over rows [0, tail_k0), in dependency order.  A row of a
    // single-column supernode is one task (code 2i); a supernode D = [s, e)
    // of two or more rows becomes one gather task per row (code 2i+1: the
    // row's entries outside D, left of s forward / right of e backward, into
    // a partial sum) and one triangle task (code -(s+1)) that solves D's
    // dense diagonal block for all its rows.
    // Row tasks of one level with at most 32 kTriPacketPerLane entries are
    // packed: code kTriPacketBase + o names a packet at offset o of fpk /
    // bpk, laid out as [g, count, row code x count]: g lanes per row (2, 4,
    // 8, 16 or 32, with g kTriPacketPerLane >= the row's entry count), up to
    // 64 / g rows.  Triangle tasks of supernodes of at most 16 rows are
    // packed as [-g, count, supernode start x count], g >= the width.
    // Supernodes wider than kTriBlockRows are solved by one task per 64-row
    // block (code kTriWideBase + x): twide[2x], twide[2x + 1] = (s, b), the
    // b-th block of the supernode at s in solve order (from the top forward,
    // from the bottom backward); block b waits for blocks < b.
    std::vector<int> fwd_task, bwd_task;
    std::vector<int> twide;
    std::vector<int> fwd_key, bwd_key;    // ordering keys (merge across circuits)
    std::vector<int> fpk, bpk;            // packets (see above)
    std::vector<int> fcut;                // n: end of row i's L entries left of its supernode
    std::vector<int> bcut;                // n: start of row i's U entries right of its supernode
    std::vector<int> ru;                  // n: U(j - ru[j] .. j-1, j) is column j's run in its supernode
    // permutations
    std::vector<int> q;                   // n: q[k] = original column
    std::vector<int> rowat;               // n: position -> original row
    std::vector<int> pinv;                // n
    // scatter count (update FMAs) and per-column statistics
    int64_t n_updates = 0;
    int64_t n_l = 0, n_u = 0;             // off-diagonal counts
    // Dense trailing block: positions [tail_k0, n) form a T x T block
    // (T = n - tail_k0, 0 = none) factored densely after the sparse columns
    // have been applied to it.  Set by lsp_plan_set_tail.
    int tail_k0 = 0;
    int tail_T = 0;
    std::vector<int> tail_ds;             // T: U slots of tail column k0+c with pos < k0
    std::vector<int> tail_lcut;           // T: first lci index of tail row k0+r with col >= k0
    std::vector<int> fkey;                // n: ordering key of column k's refactor task
    int64_t tail_updates = 0;             // updates moved into the dense block
};

// Chooses the dense trailing block: the largest T in [min_T, max_T] whose
// pattern density nnz(block) / T^2 is at least min_density (no tail when none
// qualifies or max_T <= 0).  Rebuilds ftask (sparse columns by level, tail
// columns keyed after their sparse dependencies), fwd_task / bwd_task (rows
// below k0 only, backward levels recomputed with the tail rows solved first)
// and the tail maps.  Returns T.
int lsp_plan_set_tail(LspPlan &P, int min_T, int max_T, double min_density);

// Supernodes over positions [0, tail_k0) and the per-segment update maps
// (called by lsp_build_plan and lsp_plan_set_tail).  Returns 0, or <0 when a
// run does not end at its supernode boundary or a row falls outside the
// target pattern (both mean the symbolic pattern is not closed under fill).
int lsp_plan_build_segments(LspPlan &P);

// Ruiz equilibration scales: after `sweeps` sweeps, R A C has every row and
// column maximum close to 1 (used to pivot on the scaled matrix).
void lsp_ruiz_scales(int n, long nnz, const int *row, const int *col, const double *val, int sweeps,
                     double *R, double *C);

// Build the plan from A (q-order CSC, values only used for nothing here) and
// the symbolic factors.  src_index maps each q-CSC entry to the caller's value
// order (nullptr = identity).  Returns 0 ok, <0 on an inconsistent structure.
int lsp_build_plan(int n, const int *Ap, const int *Ai, const int *src_index,
                   const LspGpLu &F, const int *q, LspPlan &P);

// ---------------------------------------------------------------------------
// SELL-C-sigma layout of A in original row/column order (sell.cpp)
struct LspSell {
    int n = 0, C = 64, sigma = 0;
    int nslices = 0;
    std::vector<int64_t> off;             // nslices+1 (slot offsets)
    std::vector<int> len;                 // nslices
    std::vector<int> col;                 // slots (padding col = row's own 0)
    std::vector<int> src;                 // slots -> caller value index, -1 pad
    std::vector<int> perm;                // slice-order row k -> original row
    int64_t slots = 0;
};
int lsp_build_sell(int n, const int *rowptr, const int *colidx,
                   const int *src_index, int C, int sigma, LspSell &S);

// ---------------------------------------------------------------------------
// level.c / selector.c are plain C with the reference's C API.
