// sell.cpp -- SELL-C-sigma layout (C = 64 by default: one wave64 per slice).
//
// Rows are sorted by stored length, descending, inside windows of sigma rows
// (sigma <= 0: the whole matrix, which is what the reference's "SELL-C-s32"
// does, src/sellc_spmv.cu:199-257).  Ties keep ascending row order (the
// reference's counting sort emits them in reverse; the order of rows inside a
// slice does not change any row's sum).  Slice s holds rows perm[s*C ..
// s*C+C); slot j of the r-th row of the slice is at off[s] + j*C + r
// (column-major, so lane r of a wave reads consecutive addresses), and
// len[s] is the longest row of the slice.  Padding slots get col 0 and
// src -1; the kernel stops at each row's own length and never reads them.
#include <algorithm>
#include <numeric>

#include "host_internal.h"

int lsp_build_sell(int n, const int *rowptr, const int *colidx, const int *src_index, int C,
                   int sigma, LspSell &S)
{
    if (n < 0 || C <= 0) return -1;
    S = LspSell();
    S.n = n;
    S.C = C;
    S.sigma = sigma;
    S.perm.resize(n);
    std::iota(S.perm.begin(), S.perm.end(), 0);
    const int win = sigma > 0 ? sigma : std::max(n, 1);
    for (int w0 = 0; w0 < n; w0 += win) {
        const int w1 = std::min(n, w0 + win);
        std::stable_sort(S.perm.begin() + w0, S.perm.begin() + w1, [&](int a, int b) {
            return rowptr[a + 1] - rowptr[a] > rowptr[b + 1] - rowptr[b];
        });
    }
    S.nslices = (n + C - 1) / C;
    S.off.assign(S.nslices + 1, 0);
    S.len.assign(S.nslices, 0);
    for (int s = 0; s < S.nslices; s++) {
        int mx = 0;
        for (int r = 0; r < C && s * C + r < n; r++) {
            const int row = S.perm[s * C + r];
            mx = std::max(mx, rowptr[row + 1] - rowptr[row]);
        }
        S.len[s] = mx;
        S.off[s + 1] = S.off[s] + (int64_t)mx * C;
    }
    S.slots = S.off[S.nslices];
    S.col.assign(S.slots, 0);
    S.src.assign(S.slots, -1);
    for (int s = 0; s < S.nslices; s++) {
        for (int r = 0; r < C && s * C + r < n; r++) {
            const int row = S.perm[s * C + r];
            for (int e = rowptr[row]; e < rowptr[row + 1]; e++) {
                const int64_t slot = S.off[s] + (int64_t)(e - rowptr[row]) * C + r;
                S.col[slot] = colidx[e];
                S.src[slot] = src_index ? src_index[e] : e;
            }
        }
    }
    return 0;
}
