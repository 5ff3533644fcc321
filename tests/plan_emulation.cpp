// CPU emulation of the device kernels' use of the static plan, for the CPU
// test suite (tests/test_plan_emulation.py builds and runs it).
//
// 1. Refactor segment updates: every column is updated the way
//    factor_column does it, lane by lane, with the dependencies of each
//    64-slot chunk applied in parts cut at random ready prefixes (the
//    streaming waits cut them wherever the flags happen to stand).  Every
//    value index is checked against its column, every row slot against the
//    column, and the map cursor must end at the column's map end.  The
//    result must match a plain per-dependency left-looking update.
// 2. Trisolve task lists: the forward and backward task lists, packets of
//    rows and of small triangles included, are run in ticket order; every
//    wait must already be satisfied (the order is a valid schedule) and each
//    packet's rows must fit their lane budget.  The solution must match plain
//    substitution with the same factors.
//
// usage: plan_emulation <kind: 0 = c0, 1 = c1, else n of a sized circuit> <tail max T (0 = none)> <seed>
//                       [plant: 1 = forward tasks in reverse order, 2 = one run's map advance lost,
//                               3 = the first two blocks of a wide supernode swapped (forward and backward)]
// prints one line "ok ..." and exits 0, or a message and a nonzero code.  A
// planted mistake must make it fail (the checks are not vacuous).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "host_internal.h"
#include "mna.h"
#include "csc.h"

static int fail(const char *what, long a = 0, long b = 0)
{
    std::printf("FAIL %s (%ld, %ld)\n", what, a, b);
    return 2;
}

int main(int argc, char **argv)
{
    const int kind = argc > 1 ? std::atoi(argv[1]) : 0;
    const int tail_max = argc > 2 ? std::atoi(argv[2]) : 0;
    const unsigned seed = argc > 3 ? (unsigned)std::atoi(argv[3]) : 1u;
    const int plant = argc > 4 ? std::atoi(argv[4]) : 0;
    std::mt19937_64 rng(seed);

    LspMnaSpec sp;
    if (kind == 0) lsp_spec_c0(&sp, 0);
    else if (kind == 1) lsp_spec_c1(&sp, 1);
    else lsp_spec_sized(&sp, kind, 0.02, 3);
    LspCoo A;
    lsp_mna_build(&sp, A);
    const int n = A.n;
    const long nnz = (long)A.val.size();
    std::vector<int> q(n);
    lsp_colamd_coo(n, nnz, A.row.data(), A.col.data(), q.data());
    std::vector<int> Ap, Ai, src;
    lsp_coo_to_csc(n, nnz, A.row.data(), A.col.data(), q.data(), Ap, Ai, src);
    std::vector<double> Ax(nnz);
    for (long p = 0; p < nnz; p++) Ax[p] = A.val[src[p]];
    LspGpLu F;
    if (lsp_gp_factor(n, Ap.data(), Ai.data(), Ax.data(), q.data(), 1.0, F) != -1) return fail("factor");
    LspPlan P;
    if (lsp_build_plan(n, Ap.data(), Ai.data(), src.data(), F, q.data(), P) != 0) return fail("plan");
    if (tail_max > 0 && lsp_plan_set_tail(P, 64, tail_max, 0.5) < 0) return fail("tail");
    const int k0 = P.tail_k0;

    // ---- 1. refactor segment updates ----
    std::vector<double> vref(P.cp[n], 0.0), vemu(P.cp[n], 0.0);
    std::vector<int> where(n, -1);
    std::vector<double> x(P.max_col + 1), xr(P.max_col + 1);
    double maxdiff = 0.0;
    long parts = 0;
    for (int k = 0; k < n; k++) {
        const long c0 = P.cp[k];
        const int m = (int)(P.cp[k + 1] - c0);
        const int dl = (int)(P.dp[k] - c0);
        for (long s = c0; s < c0 + m; s++) where[P.pos[s]] = (int)(s - c0);
        for (int i = 0; i < m; i++) x[i] = xr[i] = 0.0;
        for (int p = P.aptr[k]; p < P.aptr[k + 1]; p++) {
            const int s = P.aslot[p];
            if (s >= 0) x[s] = xr[s] = A.val[P.asrc[p]];
        }
        const int ndep = k < k0 ? dl : P.tail_ds[k - k0];
        for (int s = 0; s < ndep; s++) {               // reference: one dependency at a time
            const int j = P.pos[c0 + s];
            for (long t = P.dp[j] + 1; t < P.cp[j + 1]; t++) xr[where[P.pos[t]]] -= vref[t] * xr[s];
        }
        long u = P.relp[k];
        for (int s0 = 0; s0 < ndep; s0 += 64) {
            const int nd = std::min(64, ndep - s0);
            int J[64], E[64], LB[64], NL[64];
            double XC[64];
            unsigned long long heads = 0;
            for (int l = 0; l < 64; l++) {
                J[l] = E[l] = -2; LB[l] = NL[l] = 0;
                if (l < nd) {
                    J[l] = P.pos[c0 + s0 + l];
                    E[l] = std::min(P.snend[J[l]], k);
                    LB[l] = P.dp[J[l]] + 1;
                    NL[l] = (int)(P.cp[E[l]] - P.dp[E[l] - 1]) - 1;
                }
            }
            for (int l = 0; l < nd; l++)
                if (l == 0 || J[l] != J[l - 1] + 1 || E[l] != E[l - 1]) heads |= 1ull << l;
            for (int l = 0; l < 64; l++) XC[l] = l < nd ? x[s0 + l] : 0.0;
            for (int have = 0; have < nd;) {
                int p = have + 1 + (int)(rng() % (unsigned)(nd - have));
                if (rng() % 3 == 0) p = nd;
                unsigned long long hs = (heads & ~((1ull << have) - 1)) | (1ull << have);
                if (p < 64) hs &= (1ull << p) - 1;
                while (hs) {
                    const int ta = __builtin_ctzll(hs);
                    hs &= hs - 1;
                    const int tb = hs ? __builtin_ctzll(hs) : p;
                    const int bw = tb - ta;
                    const int jl = E[ta] - 1;
                    const int tri = jl - J[tb - 1];
                    const int nlast = NL[ta];
                    parts++;
                    for (int qq = 0; qq < bw - 1; qq++) {       // triangle
                        const int tp = ta + qq;
                        for (int l = tp + 1; l < tb; l++) {
                            const long idx = LB[tp] + (l - tp - 1);
                            if (idx >= P.cp[J[tp] + 1]) return fail("triangle index", k, idx);
                            XC[l] -= vemu[idx] * XC[tp];
                        }
                    }
                    for (int l = ta; l < tb; l++) x[s0 + l] = XC[l];
                    const int R = tri + nlast;
                    for (int i = 0; i < R; i++) {                // rows below the part
                        double acc = 0.0;
                        for (int tp = ta; tp < tb; tp++) {
                            const long idx = LB[tp] + (tb - tp - 1) + i;
                            if (idx >= P.cp[J[tp] + 1]) return fail("row index", k, idx);
                            acc += vemu[idx] * XC[tp];
                        }
                        const int slot = i < tri ? s0 + tb + i : P.rel[u + (i - tri)];
                        if (slot < 0 || slot >= m) return fail("row slot", k, slot);
                        x[slot] -= acc;
                    }
                    if (tri == 0 && !(plant == 2 && k == n / 2 && nlast > 0)) u += nlast;
                    for (int l = tb; l < nd; l++) XC[l] = x[s0 + l];
                }
                have = p;
            }
        }
        if (u != P.relp[k + 1]) return fail("map cursor", k, u);
        if (k >= k0) {
            double cm = 0.0;
            for (int i = 0; i < m; i++) cm = std::max(cm, std::fabs(xr[i]));
            for (int i = 0; i < m; i++) maxdiff = std::max(maxdiff, std::fabs(x[i] - xr[i]) / (cm + 1e-300));
            continue;
        }
        const double pr = xr[dl], pe = x[dl];
        double cm = 0.0;
        for (int i = 0; i < m; i++) cm = std::max(cm, std::fabs(i < dl ? xr[i] : i == dl ? pr : xr[i] / pr));
        for (int i = 0; i < m; i++) {
            const double a = i < dl ? xr[i] : i == dl ? pr : xr[i] / pr;
            const double b = i < dl ? x[i] : i == dl ? pe : x[i] * (1.0 / pe);
            vref[c0 + i] = a;
            vemu[c0 + i] = b;
            maxdiff = std::max(maxdiff, std::fabs(a - b) / (cm + 1e-300));
        }
    }
    if (!(maxdiff < 1e-9)) return fail("segment update differs", (long)(maxdiff * 1e15));

    // ---- 2. trisolve task lists (factors: the plain left-looking ones) ----
    if (plant == 1) std::reverse(P.fwd_task.begin(), P.fwd_task.end());
    if (plant == 3) {
        for (auto *tl : {&P.fwd_task, &P.bwd_task})
            for (size_t t = 0; t + 1 < tl->size(); t++)
                if ((*tl)[t] >= kTriWideBase && (*tl)[t] < kTriPacketBase && (*tl)[t + 1] == (*tl)[t] + 1) {
                    std::swap((*tl)[t], (*tl)[t + 1]);
                    break;
                }
    }
    std::vector<double> &val = vref;
    // tail columns: factor them plainly so that every row has factors
    for (int k = k0; k < n; k++) {
        const long c0 = P.cp[k];
        const int m = (int)(P.cp[k + 1] - c0);
        const int dl = (int)(P.dp[k] - c0);
        for (long s = c0; s < c0 + m; s++) where[P.pos[s]] = (int)(s - c0);
        for (int i = 0; i < m; i++) xr[i] = 0.0;
        for (int p = P.aptr[k]; p < P.aptr[k + 1]; p++) {
            const int s = P.aslot[p];
            if (s >= 0) xr[s] = A.val[P.asrc[p]];
        }
        for (int s = 0; s < dl; s++) {
            const int j = P.pos[c0 + s];
            for (long t = P.dp[j] + 1; t < P.cp[j + 1]; t++) xr[where[P.pos[t]]] -= val[t] * xr[s];
        }
        const double piv = xr[dl];
        for (int i = 0; i < m; i++) val[c0 + i] = i <= dl ? xr[i] : xr[i] / piv;
    }
    std::vector<double> lrow(P.lrp[n]), urow(P.urp[n]), diag(n);
    for (int k = 0; k < n; k++) {
        for (long s = P.cp[k]; s < P.dp[k]; s++) urow[P.dst[s]] = val[s];
        diag[k] = val[P.dp[k]];
        for (long s = P.dp[k] + 1; s < P.cp[k + 1]; s++) lrow[P.dst[s]] = val[s];
    }
    std::vector<double> b(n), yr(n), zr(n), y(n, 0.0), z(n, 0.0);
    for (int i = 0; i < n; i++) b[i] = std::sin(1.0 + i);
    for (int i = 0; i < n; i++) {
        double a = b[P.rowat[i]];
        for (int e = P.lrp[i]; e < P.lrp[i + 1]; e++) a -= lrow[e] * yr[P.lci[e]];
        yr[i] = a;
    }
    for (int i = n - 1; i >= 0; i--) {
        double a = yr[i];
        for (int e = P.urp[i]; e < P.urp[i + 1]; e++) a -= urow[e] * zr[P.uci[e]];
        zr[i] = a / diag[i];
    }
    const int EP = 5;
    std::vector<int> ff(n, 0), fb(n, 0);
    long packets = 0;
    // forward over rows below k0
    for (int code : P.fwd_task) {
        if (code >= kTriPacketBase) {
            const int o = code - kTriPacketBase;
            const int g = P.fpk[o], cnt = P.fpk[o + 1];
            packets++;
            if (g < 0) {
                const int gg = -g;
                if (!(gg == 4 || gg == 8 || gg == 16) || cnt < 2 || cnt > 64 / gg) return fail("triangle packet", o);
                for (int r = 0; r < cnt; r++) {
                    const int s = P.fpk[o + 2 + r], e = P.snend[s];
                    if (e - s > gg || e - s < 2) return fail("triangle packet width", s);
                    for (int i = s; i < e; i++) if (ff[i] != -EP) return fail("fwd wait", i);
                }
                for (int r = 0; r < cnt; r++) {
                    const int s = P.fpk[o + 2 + r], e = P.snend[s];
                    for (int i = s; i < e; i++) {
                        double p = y[i];
                        for (int j = s; j < i; j++) p -= val[P.dp[j] + (i - j)] * y[j];
                        y[i] = p;
                        ff[i] = EP;
                    }
                }
                continue;
            }
            if (!(g == 2 || g == 4 || g == 8 || g == 16 || g == 32) || cnt < 2 || cnt > 64 / g)
                return fail("row packet", o);
            std::vector<double> acc(cnt, 0.0);
            for (int r = 0; r < cnt; r++) {
                const int rc = P.fpk[o + 2 + r], i = rc >> 1, part = rc & 1;
                const int e1 = part ? P.fcut[i] : P.lrp[i + 1];
                if (e1 - P.lrp[i] > g * kTriPacketPerLane) return fail("row packet budget", i);
                for (int e = P.lrp[i]; e < e1; e++) {
                    if (ff[P.lci[e]] != EP) return fail("fwd wait", P.lci[e]);
                    acc[r] += lrow[e] * y[P.lci[e]];
                }
            }
            for (int r = 0; r < cnt; r++) {
                const int rc = P.fpk[o + 2 + r], i = rc >> 1, part = rc & 1;
                y[i] = b[P.rowat[i]] - acc[r];
                ff[i] = part ? -EP : EP;
            }
            continue;
        }
        if (code < 0 || code >= kTriWideBase) {
            // a whole supernode (at most 64 rows), or block b of a wide one
            // (its earlier blocks must be final)
            int s = -code - 1, b = 0, whole = 1;
            if (code >= 0) {
                const long x = code - kTriWideBase;
                if (2 * x + 1 >= (long)P.twide.size()) return fail("wide block code", x);
                s = P.twide[2 * x], b = P.twide[2 * x + 1], whole = 0;
            }
            const int e = P.snend[s];
            if (whole && e - s > kTriBlockRows) return fail("wide supernode in one task", s);
            const int r0 = whole ? s : s + kTriBlockRows * b;
            const int r1 = whole ? e : std::min(e, r0 + kTriBlockRows);
            if (r0 >= r1) return fail("empty block", s, b);
            for (int i = s; i < r0; i++) if (ff[i] != EP) return fail("fwd block wait", i);
            for (int i = r0; i < r1; i++) if (ff[i] != -EP) return fail("fwd wait", i);
            for (int i = r0; i < r1; i++) {
                double p = y[i];
                for (int j = s; j < i; j++) p -= val[P.dp[j] + (i - j)] * y[j];
                y[i] = p;
                ff[i] = EP;
            }
            continue;
        }
        const int i = code >> 1, part = code & 1;
        const int e1 = part ? P.fcut[i] : P.lrp[i + 1];
        double a = 0.0;
        for (int e = P.lrp[i]; e < e1; e++) {
            if (ff[P.lci[e]] != EP) return fail("fwd wait", P.lci[e]);
            a += lrow[e] * y[P.lci[e]];
        }
        y[i] = b[P.rowat[i]] - a;
        ff[i] = part ? -EP : EP;
    }
    for (int i = k0; i < n; i++) {                      // the dense block's rows
        double a = b[P.rowat[i]];
        for (int e = P.lrp[i]; e < P.lrp[i + 1]; e++) a -= lrow[e] * y[P.lci[e]];
        y[i] = a;
    }
    for (int i = n - 1; i >= k0; i--) {
        double a = y[i];
        for (int e = P.urp[i]; e < P.urp[i + 1]; e++) a -= urow[e] * z[P.uci[e]];
        z[i] = a / diag[i];
        fb[i] = EP;
    }
    for (int code : P.bwd_task) {
        if (code >= kTriPacketBase) {
            const int o = code - kTriPacketBase;
            const int g = P.bpk[o], cnt = P.bpk[o + 1];
            packets++;
            if (g < 0) {
                const int gg = -g;
                if (!(gg == 4 || gg == 8 || gg == 16) || cnt < 2 || cnt > 64 / gg) return fail("triangle packet", o);
                for (int r = 0; r < cnt; r++) {
                    const int s = P.bpk[o + 2 + r], e = P.snend[s];
                    if (e - s > gg || e - s < 2) return fail("triangle packet width", s);
                    for (int i = s; i < e; i++) if (fb[i] != -EP) return fail("bwd wait", i);
                }
                for (int r = 0; r < cnt; r++) {
                    const int s = P.bpk[o + 2 + r], e = P.snend[s];
                    for (int i = e - 1; i >= s; i--) {
                        double num = z[i];
                        for (int j = i + 1; j < e; j++) {
                            const int d = j - i;
                            if (d <= P.ru[j]) num -= val[P.dp[j] - d] * z[j];
                        }
                        z[i] = num / diag[i];
                        fb[i] = EP;
                    }
                }
                continue;
            }
            if (!(g == 2 || g == 4 || g == 8 || g == 16 || g == 32) || cnt < 2 || cnt > 64 / g)
                return fail("row packet", o);
            std::vector<double> acc(cnt, 0.0);
            for (int r = 0; r < cnt; r++) {
                const int rc = P.bpk[o + 2 + r], i = rc >> 1, part = rc & 1;
                const int e0 = part ? P.bcut[i] : P.urp[i];
                if (P.urp[i + 1] - e0 > g * kTriPacketPerLane) return fail("row packet budget", i);
                for (int e = e0; e < P.urp[i + 1]; e++) {
                    if (fb[P.uci[e]] != EP) return fail("bwd wait", P.uci[e]);
                    acc[r] += urow[e] * z[P.uci[e]];
                }
            }
            for (int r = 0; r < cnt; r++) {
                const int rc = P.bpk[o + 2 + r], i = rc >> 1, part = rc & 1;
                z[i] = part ? y[i] - acc[r] : (y[i] - acc[r]) / diag[i];
                fb[i] = part ? -EP : EP;
            }
            continue;
        }
        if (code < 0 || code >= kTriWideBase) {
            // a whole supernode, or block b from the bottom of a wide one
            // (its later blocks must be final)
            int s = -code - 1, b = 0, whole = 1;
            if (code >= 0) {
                const long x = code - kTriWideBase;
                if (2 * x + 1 >= (long)P.twide.size()) return fail("wide block code", x);
                s = P.twide[2 * x], b = P.twide[2 * x + 1], whole = 0;
            }
            const int e = P.snend[s];
            if (whole && e - s > kTriBlockRows) return fail("wide supernode in one task", s);
            const int r1 = whole ? e : e - kTriBlockRows * b;
            const int r0 = whole ? s : std::max(s, r1 - kTriBlockRows);
            if (r0 >= r1) return fail("empty block", s, b);
            for (int i = r1; i < e; i++) if (fb[i] != EP) return fail("bwd block wait", i);
            for (int i = r0; i < r1; i++) if (fb[i] != -EP) return fail("bwd wait", i);
            for (int i = r1 - 1; i >= r0; i--) {
                double num = z[i];
                for (int j = i + 1; j < e; j++) {
                    const int d = j - i;
                    if (d <= P.ru[j]) num -= val[P.dp[j] - d] * z[j];
                }
                z[i] = num / diag[i];
                fb[i] = EP;
            }
            continue;
        }
        const int i = code >> 1, part = code & 1;
        const int e0 = part ? P.bcut[i] : P.urp[i];
        double a = 0.0;
        for (int e = e0; e < P.urp[i + 1]; e++) {
            if (fb[P.uci[e]] != EP) return fail("bwd wait", P.uci[e]);
            a += urow[e] * z[P.uci[e]];
        }
        z[i] = part ? y[i] - a : (y[i] - a) / diag[i];
        fb[i] = part ? -EP : EP;
    }
    for (int i = 0; i < k0; i++)
        if (ff[i] != EP || fb[i] != EP) return fail("row not final", i);
    double ef = 0.0, eb = 0.0;
    for (int i = 0; i < n; i++) {
        ef = std::max(ef, std::fabs(y[i] - yr[i]) / (1.0 + std::fabs(yr[i])));
        eb = std::max(eb, std::fabs(z[i] - zr[i]) / (1.0 + std::fabs(zr[i])));
    }
    if (!(ef < 1e-10 && eb < 1e-10)) return fail("trisolve differs", (long)(ef * 1e15), (long)(eb * 1e15));
    std::printf("ok n %d k0 %d parts %ld segment-diff %.2e tasks %zu/%zu packets %ld wide-blocks %zu "
                "trisolve-diff %.2e/%.2e\n", n, k0, parts, maxdiff, P.fwd_task.size(), P.bwd_task.size(), packets,
                P.twide.size() / 2, ef, eb);
    return 0;
}
