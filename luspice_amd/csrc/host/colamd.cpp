// colamd.cpp -- column approximate minimum degree ordering.
//
// Written from the published algorithm: T. A. Davis, J. R. Gilbert,
// S. I. Larimore, E. G. Ng, "A column approximate minimum degree ordering
// algorithm", ACM TOMS 30(3), 2004, 353-376.  The reference links an external
// SuiteSparse/ngspice colamd() that is not vendored
// (src/mp/mp_lu_bench.cu:31, :113-148) and calls it with every knob zero
// (:139).  With zero knobs the published semantics are:
//   * dense-row threshold   max(16, 0*10*sqrt(n_col)) = 16: longer rows are
//     ignored;
//   * dense-column threshold max(16, 0*sqrt(min(n_row,n_col))) = 16: longer
//     columns are ordered last;
//   * aggressive row absorption off.
// Rows act as elements of the column-intersection graph of A^T A.  Each step
// eliminates the column of least approximate degree (head of the lowest
// degree bucket, LIFO buckets as in the paper), merges every row that holds
// it into one new pivot row, and re-scores the pivot row's columns with the
// COLAMD bound  score(c) = |pivot row| - 1 + sum_{r in Col(c), r live}
// |Row(r) \ pivot row|, clipped to the number of remaining columns.
// Indistinguishable columns are detected by hashing their row lists and
// merged into supercolumns (mass elimination), as in the paper's Section 3.4;
// a supercolumn's members are numbered consecutively when its principal
// column is eliminated.
//
// Output q[k] = original column eliminated k-th (the colamd p[] array).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "host_internal.h"

namespace {

struct Colamd {
    int nr, nc;
    std::vector<std::vector<int>> row;   // columns of each live row
    std::vector<std::vector<int>> col;   // rows of each live column
    std::vector<char> row_alive, col_alive;
    std::vector<int> row_deg;            // live thickness-weighted degree
    std::vector<int> thick;              // supercolumn thickness (0 = absorbed)
    std::vector<int> score;
    std::vector<int> head, nxt, prv;     // degree buckets
    std::vector<int> super_next;         // members chained after principal
    std::vector<int> mark;               // per-row scratch (set differences)
    std::vector<int> cmark;              // per-column scratch
    int stamp = 0;

    void bucket_insert(int c) {
        int s = score[c];
        prv[c] = -1;
        nxt[c] = head[s];
        if (head[s] >= 0) prv[head[s]] = c;
        head[s] = c;
    }
    void bucket_remove(int c) {
        int s = score[c];
        if (prv[c] >= 0) nxt[prv[c]] = nxt[c]; else head[s] = nxt[c];
        if (nxt[c] >= 0) prv[nxt[c]] = prv[c];
        prv[c] = nxt[c] = -1;
    }
};

}  // namespace

int lsp_colamd_order(int n_row, int n_col, const int *Ap, const int *Ai, int *q)
{
    if (n_row < 0 || n_col < 0) return -1;
    if (n_col == 0) return 0;
    Colamd S;
    S.nr = n_row;
    S.nc = n_col;
    const int dense_row = 16, dense_col = 16;
    S.row.assign(n_row, {});
    S.col.assign(n_col, {});
    S.row_alive.assign(n_row, 1);
    S.col_alive.assign(n_col, 1);
    S.thick.assign(n_col, 1);
    S.super_next.assign(n_col, -1);
    S.score.assign(n_col, 0);
    S.head.assign(n_col + 1, -1);
    S.nxt.assign(n_col, -1);
    S.prv.assign(n_col, -1);
    S.mark.assign(n_row, -1);
    S.cmark.assign(n_col, 0);

    // Column lists with duplicate rows removed; rows built from them.
    std::vector<int> seen(n_row, -1);
    for (int c = 0; c < n_col; c++) {
        for (int p = Ap[c]; p < Ap[c + 1]; p++) {
            int r = Ai[p];
            if (r < 0 || r >= n_row) return -2;
            if (seen[r] == c) continue;
            seen[r] = c;
            S.col[c].push_back(r);
        }
    }
    std::vector<int> order_tail;   // dense / null columns, numbered last
    // dense columns first (c descending, as the paper's reference code does)
    for (int c = n_col - 1; c >= 0; c--) {
        if ((int)S.col[c].size() > dense_col) {
            S.col_alive[c] = 0;
            order_tail.push_back(c);
        }
    }
    for (int c = 0; c < n_col; c++) {
        if (!S.col_alive[c]) continue;
        for (int r : S.col[c]) S.row[r].push_back(c);
    }
    for (int r = 0; r < n_row; r++) {
        int d = (int)S.row[r].size();
        if (d > dense_row || d == 0) S.row_alive[r] = 0;
    }
    // drop dead rows from column lists; columns left empty are null columns
    for (int c = n_col - 1; c >= 0; c--) {
        if (!S.col_alive[c]) continue;
        auto &L = S.col[c];
        L.erase(std::remove_if(L.begin(), L.end(),
                               [&](int r) { return !S.row_alive[r]; }), L.end());
        if (L.empty()) {
            S.col_alive[c] = 0;
            order_tail.push_back(c);
        }
    }
    S.row_deg.assign(n_row, 0);
    for (int r = 0; r < n_row; r++) {
        if (!S.row_alive[r]) { S.row[r].clear(); continue; }
        auto &R = S.row[r];
        R.erase(std::remove_if(R.begin(), R.end(),
                               [&](int c) { return !S.col_alive[c]; }), R.end());
        S.row_deg[r] = (int)R.size();
    }
    int n_live = 0;
    for (int c = 0; c < n_col; c++) if (S.col_alive[c]) n_live++;
    // initial scores: sum over rows of (|row| - 1), clipped
    for (int c = n_col - 1; c >= 0; c--) {
        if (!S.col_alive[c]) continue;
        long sc = 0;
        for (int r : S.col[c]) sc += S.row_deg[r] - 1;
        if (sc > n_live - 1) sc = n_live - 1;
        if (sc < 0) sc = 0;
        S.score[c] = (int)sc;
        S.bucket_insert(c);
    }

    int k = 0;            // columns ordered so far (thickness-weighted)
    int min_score = 0;
    std::vector<int> pivot_row_cols;
    std::vector<int> touched_rows;
    std::vector<uint64_t> hashv(n_col, 0);
    while (k < n_live) {
        while (min_score <= n_col && (min_score >= (int)S.head.size() || S.head[min_score] < 0))
            min_score++;
        if (min_score > n_col) return -3;
        int pc = S.head[min_score];
        S.bucket_remove(pc);
        // number the supercolumn
        for (int c = pc; c >= 0; c = S.super_next[c]) q[k++] = c;
        S.col_alive[pc] = 0;

        // pivot row = union of live rows of pc, minus pc
        if (++S.stamp == 0x7fffffff) { std::fill(S.cmark.begin(), S.cmark.end(), 0); S.stamp = 1; }
        int st = S.stamp;
        pivot_row_cols.clear();
        int pr = -1;
        for (int r : S.col[pc]) {
            if (!S.row_alive[r]) continue;
            if (pr < 0) pr = r;           // reuse the first row's slot
            for (int c : S.row[r]) {
                if (!S.col_alive[c] || S.cmark[c] == st) continue;
                S.cmark[c] = st;
                pivot_row_cols.push_back(c);
            }
            S.row_alive[r] = 0;
        }
        S.col[pc].clear();
        if (pivot_row_cols.empty()) continue;
        S.row_alive[pr] = 1;
        S.row[pr] = pivot_row_cols;
        int prdeg = 0;
        for (int c : pivot_row_cols) prdeg += S.thick[c];
        S.row_deg[pr] = prdeg;

        // remove pivot-row columns from their buckets, prune dead rows, and
        // compute set differences |Row(r) \ pivot row| for neighbouring rows
        touched_rows.clear();
        for (int c : pivot_row_cols) {
            S.bucket_remove(c);
            auto &L = S.col[c];
            size_t w = 0;
            for (size_t t = 0; t < L.size(); t++) {
                int r = L[t];
                if (!S.row_alive[r] || r == pr) continue;
                L[w++] = r;
                if (S.mark[r] < 0) {
                    S.mark[r] = S.row_deg[r];
                    touched_rows.push_back(r);
                }
                S.mark[r] -= S.thick[c];
            }
            L.resize(w);
        }
        // new approximate degrees + hash of row lists for supercolumn search
        int remaining = n_live - k;
        for (int c : pivot_row_cols) {
            long sc = prdeg - S.thick[c];
            uint64_t h = 0;
            for (int r : S.col[c]) {
                sc += S.mark[r] > 0 ? S.mark[r] : 0;
                h += (uint64_t)r * 0x9E3779B97F4A7C15ULL;
            }
            int lim = remaining - S.thick[c];
            if (sc > lim) sc = lim;
            if (sc < 0) sc = 0;
            S.score[c] = (int)sc;
            hashv[c] = h ^ (uint64_t)S.col[c].size();
        }
        for (int r : touched_rows) S.mark[r] = -1;
        // live rows of each column get the pivot row
        for (int c : pivot_row_cols) S.col[c].push_back(pr);

        // supercolumn detection among the pivot row's columns: identical row
        // lists (same hash, same length, same set) merge into the first one
        if (pivot_row_cols.size() > 1) {
            std::vector<int> idx(pivot_row_cols.size());
            for (size_t t = 0; t < idx.size(); t++) idx[t] = (int)t;
            std::sort(idx.begin(), idx.end(), [&](int a, int b) {
                int ca = pivot_row_cols[a], cb = pivot_row_cols[b];
                if (hashv[ca] != hashv[cb]) return hashv[ca] < hashv[cb];
                return a < b;
            });
            std::vector<int> tmpa, tmpb;
            for (size_t t = 0; t < idx.size();) {
                size_t u = t + 1;
                int ct = pivot_row_cols[idx[t]];
                while (u < idx.size() && hashv[pivot_row_cols[idx[u]]] == hashv[ct]) u++;
                // group [t,u): pairwise compare against the group's members
                for (size_t a = t; a < u; a++) {
                    int ca = pivot_row_cols[idx[a]];
                    if (S.thick[ca] == 0) continue;
                    tmpa = S.col[ca];
                    std::sort(tmpa.begin(), tmpa.end());
                    for (size_t b = a + 1; b < u; b++) {
                        int cb = pivot_row_cols[idx[b]];
                        if (S.thick[cb] == 0) continue;
                        if (S.col[cb].size() != tmpa.size()) continue;
                        tmpb = S.col[cb];
                        std::sort(tmpb.begin(), tmpb.end());
                        if (tmpa != tmpb) continue;
                        // absorb cb into ca; ca's degree no longer counts cb
                        S.score[ca] -= S.thick[cb];
                        if (S.score[ca] < 0) S.score[ca] = 0;
                        S.thick[ca] += S.thick[cb];
                        S.thick[cb] = 0;
                        int tail = ca;
                        while (S.super_next[tail] >= 0) tail = S.super_next[tail];
                        S.super_next[tail] = cb;
                        S.col_alive[cb] = 0;
                        for (int r : S.col[cb]) {
                            auto &R = S.row[r];
                            R.erase(std::remove(R.begin(), R.end(), cb), R.end());
                        }
                        S.col[cb].clear();
                    }
                }
                t = u;
            }
        }
        // re-insert surviving columns; scores of merged supercolumns lose the
        // absorbed thickness counted inside the pivot row
        for (int c : pivot_row_cols) {
            if (!S.col_alive[c] || S.thick[c] == 0) continue;
            int lim = n_live - k - S.thick[c];
            if (lim < 0) lim = 0;
            if (S.score[c] > lim) S.score[c] = lim;
            S.bucket_insert(c);
            if (S.score[c] < min_score) min_score = S.score[c];
        }
        // the pivot row only keeps live principal columns
        auto &P = S.row[pr];
        P.erase(std::remove_if(P.begin(), P.end(),
                               [&](int c) { return !S.col_alive[c] || S.thick[c] == 0; }),
                P.end());
        int d = 0;
        for (int c : P) d += S.thick[c];
        S.row_deg[pr] = d;
        if (P.empty()) S.row_alive[pr] = 0;
    }
    // dense and null columns last, in the order they were set aside
    for (int c : order_tail) q[k++] = c;
    return k == n_col ? 0 : -4;
}
