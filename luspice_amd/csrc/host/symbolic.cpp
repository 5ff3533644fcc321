// symbolic.cpp -- CPU symbolic analysis: pivot sequence, L/U pattern and the
// static plan the GPU refactorization and trisolve consume.
//
// lsp_gp_factor follows the reference's fp64 LU (src/third_party/splu.h:151-206,
// wrapped by src/cpu_ref_lu.c:15-58): left-looking Gilbert-Peierls LU with a
// depth-first reach per column, threshold partial pivoting (first strict
// maximum |x| among unpivoted rows in reach order; the preferred row
// diag_row[k] wins when |x| >= tol * max), L unit diagonal stored first, U
// diagonal stored last, L rows remapped to pivot positions at the end.  The
// pattern is structural: every reached row is stored, including exact zeros,
// so P, nnz(L), nnz(U) and the per-column entry order equal the reference's.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "host_internal.h"

int lsp_gp_factor(int n, const int *Ap, const int *Ai, const double *Ax,
                  const int *diag_row, double tol, LspGpLu &F)
{
    F = LspGpLu();
    if (n < 1) return -2;
    F.n = n;
    F.pinv.assign(n, -1);
    F.Lp.assign(n + 1, 0);
    F.Up.assign(n + 1, 0);
    F.Li.reserve((size_t)Ap[n] * 4 + 16);
    F.Lx.reserve((size_t)Ap[n] * 4 + 16);
    F.Ui.reserve((size_t)Ap[n] * 4 + 16);
    F.Ux.reserve((size_t)Ap[n] * 4 + 16);
    std::vector<int> reach(n), stack(n), resume(n), visited(n, 0);
    std::vector<double> x(n, 0.0);
    int *pinv = F.pinv.data();
    for (int k = 0; k < n; k++) {
        F.Lp[k] = (int)F.Li.size();
        F.Up[k] = (int)F.Ui.size();
        const int stamp = k + 1;
        // depth-first reach of A(:,k) through the L columns built so far
        int top = n;
        for (int p = Ap[k]; p < Ap[k + 1]; p++) {
            int root = Ai[p];
            if (visited[root] == stamp) continue;
            int d = 0;
            stack[0] = root;
            while (d >= 0) {
                int v = stack[d];
                int J = pinv[v];
                if (visited[v] != stamp) {
                    visited[v] = stamp;
                    resume[d] = J < 0 ? 0 : F.Lp[J];
                }
                int stop = J < 0 ? 0 : F.Lp[J + 1];
                bool descended = false;
                for (int e = resume[d]; e < stop; e++) {
                    int w = F.Li[e];
                    if (visited[w] == stamp) continue;
                    resume[d] = e;
                    stack[++d] = w;
                    descended = true;
                    break;
                }
                if (!descended) {
                    reach[--top] = v;
                    d--;
                }
            }
        }
        // sparse triangular solve over the reach (topological order)
        for (int t = top; t < n; t++) x[reach[t]] = 0.0;
        for (int p = Ap[k]; p < Ap[k + 1]; p++) x[Ai[p]] = Ax[p];
        for (int t = top; t < n; t++) {
            int v = reach[t];
            int J = pinv[v];
            if (J < 0) continue;
            x[v] /= F.Lx[F.Lp[J]];
            const double xv = x[v];
            for (int e = F.Lp[J] + 1; e < F.Lp[J + 1]; e++) x[F.Li[e]] -= F.Lx[e] * xv;
        }
        int piv = -1;
        double best = -1.0;
        for (int t = top; t < n; t++) {
            int v = reach[t];
            if (pinv[v] < 0) {
                double a = std::fabs(x[v]);
                if (a > best) { best = a; piv = v; }
            } else {
                F.Ui.push_back(pinv[v]);
                F.Ux.push_back(x[v]);
            }
        }
        if (piv == -1 || best <= 0) { F.singular = k; return k; }
        const int dr = diag_row ? diag_row[k] : k;
        if (pinv[dr] < 0 && std::fabs(x[dr]) >= best * tol) piv = dr;
        const double pv = x[piv];
        F.Ui.push_back(k);
        F.Ux.push_back(pv);
        pinv[piv] = k;
        F.Li.push_back(piv);
        F.Lx.push_back(1.0);
        for (int t = top; t < n; t++) {
            int v = reach[t];
            if (pinv[v] < 0) {
                F.Li.push_back(v);
                F.Lx.push_back(x[v] / pv);
            }
            x[v] = 0.0;
        }
    }
    F.Lp[n] = (int)F.Li.size();
    F.Up[n] = (int)F.Ui.size();
    for (auto &r : F.Li) r = pinv[r];
    F.lnnz = (long)F.Li.size();
    F.unnz = (long)F.Ui.size();
    return -1;
}

void lsp_gp_solve(const LspGpLu &F, const int *q, const double *b, double *x)
{
    const int n = F.n;
    std::vector<int> rowat(n, -1);
    for (int i = 0; i < n; i++) if (F.pinv[i] >= 0) rowat[F.pinv[i]] = i;
    std::vector<double> w(n), y(n);
    for (int k = 0; k < n; k++) w[k] = rowat[k] >= 0 ? b[rowat[k]] : 0.0;
    for (int k = 0; k < n; k++) {
        y[k] = w[k] / F.Lx[F.Lp[k]];
        for (int e = F.Lp[k] + 1; e < F.Lp[k + 1]; e++) w[F.Li[e]] -= F.Lx[e] * y[k];
    }
    for (int k = n - 1; k >= 0; k--) {
        y[k] /= F.Ux[F.Up[k + 1] - 1];
        for (int e = F.Up[k]; e < F.Up[k + 1] - 1; e++) y[F.Ui[e]] -= F.Ux[e] * y[k];
    }
    for (int k = 0; k < n; k++) x[q ? q[k] : k] = y[k];
}

// ---------------------------------------------------------------------------
// plan
// ---------------------------------------------------------------------------
static void sort_by_level(const std::vector<int> &lv, int nlv, std::vector<int> &out)
{
    const int n = (int)lv.size();
    std::vector<int> cnt(nlv + 1, 0);
    for (int i = 0; i < n; i++) cnt[lv[i] + 1]++;
    for (int l = 0; l < nlv; l++) cnt[l + 1] += cnt[l];
    out.assign(n, 0);
    for (int i = 0; i < n; i++) out[cnt[lv[i]]++] = i;
}

int lsp_build_plan(int n, const int *Ap, const int *Ai, const int *src_index,
                   const LspGpLu &F, const int *q, LspPlan &P)
{
    P = LspPlan();
    if (n >= kTriWideBase / 2) return -1;        // row task codes 2i + 1 stay below the block codes
    P.n = n;
    P.nnz_a = Ap[n];
    P.pinv = F.pinv;
    P.q.assign(q, q + n);
    P.rowat.assign(n, -1);
    for (int i = 0; i < n; i++) {
        if (F.pinv[i] < 0 || F.pinv[i] >= n) return -1;
        P.rowat[F.pinv[i]] = i;
    }
    // column patterns
    P.cp.assign(n + 1, 0);
    P.dp.assign(n, 0);
    for (int k = 0; k < n; k++) {
        int nu = F.Up[k + 1] - F.Up[k] - 1;
        int nl = F.Lp[k + 1] - F.Lp[k] - 1;
        if (nu < 0 || nl < 0) return -2;
        if (F.Ui[F.Up[k + 1] - 1] != k || F.Li[F.Lp[k]] != k) return -3;
        P.cp[k + 1] = P.cp[k] + nu + 1 + nl;
        P.max_col = std::max(P.max_col, nu + 1 + nl);
    }
    if (P.max_col > 65535) return -9;            // local slots are 16-bit
    const int64_t nslots = P.cp[n];
    P.pos.assign(nslots, 0);
    for (int k = 0; k < n; k++) {
        int64_t s = P.cp[k];
        for (int e = F.Up[k]; e < F.Up[k + 1] - 1; e++) P.pos[s++] = F.Ui[e];
        std::sort(P.pos.begin() + P.cp[k], P.pos.begin() + s);
        P.dp[k] = (int)(s - P.cp[k]);            // local offset of the diagonal
        P.pos[s++] = k;
        int64_t l0 = s;
        for (int e = F.Lp[k] + 1; e < F.Lp[k + 1]; e++) P.pos[s++] = F.Li[e];
        std::sort(P.pos.begin() + l0, P.pos.begin() + s);
        for (int64_t t = P.cp[k]; t < P.dp[k] + P.cp[k]; t++)
            if (P.pos[t] >= k) return -4;
        for (int64_t t = l0; t < s; t++)
            if (P.pos[t] <= k) return -4;
        P.n_u += P.dp[k];
        P.n_l += s - l0;
    }
    // dp as absolute slot index
    for (int k = 0; k < n; k++) P.dp[k] = (int)(P.cp[k] + P.dp[k]);
    if (nslots > 0x7fffffffLL) return -5;
    // position -> local slot lookup, one column at a time
    std::vector<int> where(n, -1), owner(n, -1);
    P.aptr.assign(Ap, Ap + n + 1);
    P.asrc.assign(P.nnz_a, 0);
    P.aslot.assign(P.nnz_a, -1);
    // left-looking update count (statistic; the map itself is built per
    // supernode segment by lsp_plan_build_segments)
    P.n_updates = 0;
    for (int k = 0; k < n; k++)
        for (int64_t s = P.cp[k]; s < P.dp[k]; s++) {
            const int j = P.pos[s];
            P.n_updates += P.cp[j + 1] - P.dp[j] - 1;
        }
    std::vector<int> slot_writer;
    for (int k = 0; k < n; k++) {
        for (int64_t s = P.cp[k]; s < P.cp[k + 1]; s++) {
            where[P.pos[s]] = (int)(s - P.cp[k]);
            owner[P.pos[s]] = k;
        }
        // A entries; a later duplicate of the same row overwrites (splu assigns)
        slot_writer.assign(P.cp[k + 1] - P.cp[k], -1);
        for (int p = Ap[k]; p < Ap[k + 1]; p++) {
            int r = Ai[p];
            if (r < 0 || r >= n) return -6;
            int ps = F.pinv[r];
            if (owner[ps] != k) return -7;         // A entry outside the pattern
            int loc = where[ps];
            if (slot_writer[loc] >= 0) P.aslot[slot_writer[loc]] = -1;
            slot_writer[loc] = p;
            P.aslot[p] = loc;
            P.asrc[p] = src_index ? src_index[p] : p;
        }
    }
    // factor levels (src/level.c:20-40 on the U pattern)
    P.level.assign(n, 0);
    int deepest = 0;
    for (int k = 0; k < n; k++) {
        int lv = 0;
        for (int64_t s = P.cp[k]; s < P.dp[k]; s++) lv = std::max(lv, P.level[P.pos[s]] + 1);
        P.level[k] = lv;
        deepest = std::max(deepest, lv);
    }
    P.nlevels = deepest + 1;
    sort_by_level(P.level, P.nlevels, P.ftask);
    // row-wise L (strict lower) and U (strict upper)
    P.lrp.assign(n + 1, 0);
    P.urp.assign(n + 1, 0);
    for (int k = 0; k < n; k++) {
        for (int64_t s = P.cp[k]; s < P.dp[k]; s++) P.urp[P.pos[s] + 1]++;
        for (int64_t s = P.dp[k] + 1; s < P.cp[k + 1]; s++) P.lrp[P.pos[s] + 1]++;
    }
    for (int i = 0; i < n; i++) { P.lrp[i + 1] += P.lrp[i]; P.urp[i + 1] += P.urp[i]; }
    P.lci.assign(P.lrp[n], 0);
    P.uci.assign(P.urp[n], 0);
    P.dst.assign(nslots, 0);
    {
        std::vector<int> lcur(P.lrp.begin(), P.lrp.end() - 1);
        std::vector<int> ucur(P.urp.begin(), P.urp.end() - 1);
        for (int k = 0; k < n; k++) {               // k ascending -> rows sorted by column
            for (int64_t s = P.cp[k]; s < P.dp[k]; s++) {
                int i = P.pos[s];
                int d = ucur[i]++;
                P.uci[d] = k;
                P.dst[s] = d;
            }
            P.dst[P.dp[k]] = k;
            for (int64_t s = P.dp[k] + 1; s < P.cp[k + 1]; s++) {
                int i = P.pos[s];
                int d = lcur[i]++;
                P.lci[d] = k;
                P.dst[s] = d;
            }
        }
    }
    // trisolve row levels
    P.lvf.assign(n, 0);
    P.lvb.assign(n, 0);
    int df = 0, db = 0;
    for (int i = 0; i < n; i++) {
        int lv = 0;
        for (int e = P.lrp[i]; e < P.lrp[i + 1]; e++) lv = std::max(lv, P.lvf[P.lci[e]] + 1);
        P.lvf[i] = lv;
        df = std::max(df, lv);
    }
    for (int i = n - 1; i >= 0; i--) {
        int lv = 0;
        for (int e = P.urp[i]; e < P.urp[i + 1]; e++) lv = std::max(lv, P.lvb[P.uci[e]] + 1);
        P.lvb[i] = lv;
        db = std::max(db, lv);
    }
    P.nlv_f = df + 1;
    P.nlv_b = db + 1;
    sort_by_level(P.lvf, P.nlv_f, P.fwd_task);
    sort_by_level(P.lvb, P.nlv_b, P.bwd_task);
    P.tail_k0 = n;
    P.tail_T = 0;
    P.fkey = P.level;
    return lsp_plan_build_segments(P);
}

static void plan_tri_supernodes(LspPlan &P);

// ---------------------------------------------------------------------------
// supernode segments
// ---------------------------------------------------------------------------
int lsp_plan_build_segments(LspPlan &P)
{
    const int n = P.n;
    const int k0 = P.tail_k0;
    auto llen = [&](int j) { return (int)(P.cp[j + 1] - P.dp[j] - 1); };
    // j joins j-1's supernode when L(:, j-1) = {j} + L(:, j)
    P.snend.assign(n, 0);
    for (int j = n - 1; j >= 0; j--) {
        bool next_joins = false;
        const int c = j + 1;
        if (c < k0 && llen(j) == llen(c) + 1 && llen(j) > 0 && P.pos[P.dp[j] + 1] == c) {
            next_joins = true;
            const int64_t a = P.dp[j] + 2, b = P.dp[c] + 1;
            for (int t = 0; t < llen(c); t++)
                if (P.pos[a + t] != P.pos[b + t]) { next_joins = false; break; }
        }
        P.snend[j] = next_joins ? P.snend[c] : j + 1;
    }
    // runs and their maps
    std::vector<int> where(n, -1), owner(n, -1);
    P.relp.assign(n + 1, 0);
    P.rel.clear();
    P.n_segments = 0;
    for (int k = 0; k < n; k++) {
        const int64_t c0 = P.cp[k];
        for (int64_t s = c0; s < P.cp[k + 1]; s++) {
            where[P.pos[s]] = (int)(s - c0);
            owner[P.pos[s]] = k;
        }
        const int ndep = k < k0 ? P.dp[k] - (int)c0 : P.tail_ds[k - k0];
        for (int t = 0; t < ndep; t++) {
            const int j = P.pos[c0 + t];
            const int e = std::min(P.snend[j], k);
            const bool last = t + 1 == ndep || P.pos[c0 + t + 1] != j + 1 ||
                              std::min(P.snend[P.pos[c0 + t + 1]], k) != e;
            if (!last) continue;
            if (j != e - 1) return -10;            // run stops inside its supernode
            P.n_segments++;
            for (int64_t s = P.dp[j] + 1; s < P.cp[j + 1]; s++) {
                const int ps = P.pos[s];
                if (owner[ps] != k) return -8;     // fill outside the pattern
                P.rel.push_back((uint16_t)where[ps]);
            }
        }
        P.relp[k + 1] = (int64_t)P.rel.size();
    }
    plan_tri_supernodes(P);
    return 0;
}

// ---------------------------------------------------------------------------
// supernodal trisolve tasks
// ---------------------------------------------------------------------------
static void plan_tri_supernodes(LspPlan &P)
{
    constexpr int kWave64 = 64;             // lanes per wave of the solve kernel
    const int n = P.n;
    const int k0 = P.tail_k0;
    std::vector<int> sst(n);
    for (int j = 0; j < n; j++) sst[j] = (j > 0 && j < k0 && P.snend[j - 1] == P.snend[j]) ? sst[j - 1] : j;
    P.fcut.assign(n, 0);
    P.bcut.assign(n, 0);
    P.ru.assign(n, 0);
    for (int i = 0; i < n; i++) {
        int e = P.lrp[i];
        while (e < P.lrp[i + 1] && P.lci[e] < sst[i]) e++;
        P.fcut[i] = e;
        const int end = i < k0 ? P.snend[i] : i + 1;
        int f = P.urp[i];
        while (f < P.urp[i + 1] && P.uci[f] < end) f++;
        P.bcut[i] = f;
        int r = 0;
        for (int64_t t = P.dp[i] - 1; t >= P.cp[i] && P.pos[t] >= sst[i] && P.pos[t] == i - 1 - r; t--) r++;
        P.ru[i] = r;
    }
    // blocks of the supernodes wider than one triangle task
    std::vector<int> wide0(k0, -1);
    P.twide.clear();
    for (int s = 0; s < k0; s = P.snend[s]) {
        const int w = P.snend[s] - s;
        if (w <= kTriBlockRows) continue;
        wide0[s] = (int)P.twide.size() / 2;
        for (int b = 0; b * kTriBlockRows < w; b++) {
            P.twide.push_back(s);
            P.twide.push_back(b);
        }
    }
    // F(D) = 1 + max F over the supernodes D's rows read outside D
    auto build = [&](bool fwd, std::vector<int> &task, std::vector<int> &key, std::vector<int> &pk, int &nlv) {
        std::vector<int> F(k0, 0);
        int top = 0;
        auto visit = [&](int i) {
            const int s = sst[i];
            int l = F[s];
            if (fwd) {
                for (int e = P.lrp[i]; e < P.fcut[i]; e++) l = std::max(l, F[sst[P.lci[e]]] + 1);
            } else {
                for (int e = P.bcut[i]; e < P.urp[i + 1]; e++) {
                    const int j = P.uci[e];
                    if (j < k0) l = std::max(l, F[sst[j]] + 1);
                }
            }
            F[s] = l;
        };
        // a supernode's level is final once all its rows are visited; rows
        // of earlier supernodes (forward) / later ones (backward) come first
        if (fwd) for (int i = 0; i < k0; i++) visit(i);
        else for (int i = k0 - 1; i >= 0; i--) visit(i);
        std::vector<std::pair<int, int>> items;
        items.reserve(k0 + k0 / 4);
        for (int i = 0; i < k0; i++) {
            const int s = sst[i];
            const int w = P.snend[s] - s;
            top = std::max(top, F[s]);
            if (w == 1) {
                items.push_back({2 * F[s] + 1, 2 * i});
            } else {
                items.push_back({2 * F[s], 2 * i + 1});
                if (i == s && w <= kTriBlockRows) items.push_back({2 * F[s] + 1, -(s + 1)});
                if (i == s && w > kTriBlockRows)
                    for (int b = 0; b * kTriBlockRows < w; b++)
                        items.push_back({2 * F[s] + 1, kTriWideBase + wide0[s] + b});
            }
        }
        std::stable_sort(items.begin(), items.end(),
                         [](const std::pair<int, int> &a, const std::pair<int, int> &b) { return a.first < b.first; });
        // Pack the short row tasks of each key: rows with at most
        // g kTriPacketPerLane entries (g = 2, 4, 8, 16, 32 lanes) go 64 / g to
        // a packet, so one ticket serves up to 32 rows.  (One lane per row,
        // 64 rows to a packet, made c1's solves faster and a 512-circuit
        // batch slower: a packet completes with its slowest row.)  Rows and triangles of
        // one key are independent, so any grouping of them keeps the order a
        // valid schedule; the blocks of a wide supernode stay in their order.
        auto entries = [&](int code) {
            const int i = code >> 1;
            const bool part = code & 1;
            return fwd ? (part ? P.fcut[i] : P.lrp[i + 1]) - P.lrp[i]
                       : P.urp[i + 1] - (part ? P.bcut[i] : P.urp[i]);
        };
        // Triangle tasks of supernodes of at most g rows (g = 4, 8, 16) are
        // packed the same way, 64 / g to a packet laid out as
        // [-g, count, supernode start x count].
        std::vector<std::pair<int, int>> out;
        out.reserve(items.size());
        pk.clear();
        std::vector<int> cls[5], tcls[3];
        auto emit = [&](int key_v, std::vector<int> &v, int g, bool tri) {
            const int per = kWave64 / g;
            for (size_t x = 0; x < v.size(); x += per) {
                const int cnt = (int)std::min<size_t>(per, v.size() - x);
                if (cnt == 1) {
                    out.push_back({key_v, tri ? -(v[x] + 1) : v[x]});
                    continue;
                }
                out.push_back({key_v, kTriPacketBase + (int)pk.size()});
                pk.push_back(tri ? -g : g);
                pk.push_back(cnt);
                for (int r = 0; r < cnt; r++) pk.push_back(v[x + r]);
            }
        };
        for (size_t a = 0; a < items.size();) {
            size_t b = a;
            while (b < items.size() && items[b].first == items[a].first) b++;
            for (auto &c : cls) c.clear();
            for (auto &c : tcls) c.clear();
            for (size_t t = a; t < b; t++) {
                const int code = items[t].second;
                if (code >= kTriWideBase) {
                    out.push_back(items[t]);
                    continue;
                }
                if (code >= 0) {
                    const int ne = entries(code);
                    const int k = kTriPacketPerLane;
                    const int c = ne <= 2 * k ? 4 : ne <= 4 * k ? 0 : ne <= 8 * k ? 1 : ne <= 16 * k ? 2
                                : ne <= 32 * k ? 3 : -1;
                    if (c >= 0) { cls[c].push_back(code); continue; }
                } else {
                    const int s0 = -code - 1;
                    const int w = P.snend[s0] - s0;
                    const int c = w <= 4 ? 0 : w <= 8 ? 1 : w <= 16 ? 2 : -1;
                    if (c >= 0) { tcls[c].push_back(s0); continue; }
                }
                out.push_back(items[t]);
            }
            for (int c = 0; c < 4; c++) emit(items[a].first, cls[c], 4 << c, false);
            emit(items[a].first, cls[4], 2, false);   // rows of at most 8 entries, 32 to a packet
            for (int c = 0; c < 3; c++) emit(items[a].first, tcls[c], 4 << c, true);
            a = b;
        }
        task.resize(out.size());
        key.resize(out.size());
        for (size_t t = 0; t < out.size(); t++) { key[t] = out[t].first; task[t] = out[t].second; }
        nlv = k0 > 0 ? top + 1 : 0;
    };
    build(true, P.fwd_task, P.fwd_key, P.fpk, P.nlv_f);
    build(false, P.bwd_task, P.bwd_key, P.bpk, P.nlv_b);
}

// ---------------------------------------------------------------------------
// dense trailing block
// ---------------------------------------------------------------------------
int lsp_plan_set_tail(LspPlan &P, int min_T, int max_T, double min_density)
{
    const int n = P.n;
    // nz(k0) = pattern entries inside [k0, n) x [k0, n), grown one position at a
    // time: column k0-1 adds its diagonal and L part (all below), row k0-1 adds
    // its U entries (all to the right).
    int best = 0;
    int64_t nz = 0;
    for (int T = 1; T <= std::min(n, max_T); T++) {
        const int k = n - T;
        nz += (P.cp[k + 1] - P.dp[k]) + (P.urp[k + 1] - P.urp[k]);
        if (T >= min_T && (double)nz >= min_density * (double)T * (double)T) best = T;
    }
    const int k0 = n - best;
    P.tail_T = best;
    P.tail_k0 = k0;
    P.tail_ds.assign(best, 0);
    P.tail_lcut.assign(best, 0);
    P.tail_updates = 0;
    P.fkey = P.level;
    if (best == 0) return lsp_plan_build_segments(P);
    for (int c = 0; c < best; c++) {
        const int k = k0 + c;
        int ds = 0;
        int key = 0;
        for (int64_t s = P.cp[k]; s < P.dp[k] && P.pos[s] < k0; s++) {
            ds++;
            key = std::max(key, P.level[P.pos[s]] + 1);
        }
        P.tail_ds[c] = ds;
        P.fkey[k] = key;
        for (int64_t s = P.cp[k] + ds; s < P.dp[k]; s++) {
            const int j = P.pos[s];
            P.tail_updates += P.cp[j + 1] - P.dp[j] - 1;
        }
        int e = P.lrp[k];
        while (e < P.lrp[k + 1] && P.lci[e] < k0) e++;
        P.tail_lcut[c] = e;
    }
    int nkey = 0;
    for (int k = 0; k < n; k++) nkey = std::max(nkey, P.fkey[k] + 1);
    sort_by_level(P.fkey, nkey, P.ftask);
    // forward rows below k0 keep their levels (their L entries are all below k0)
    std::vector<int> fw;
    fw.reserve(k0);
    for (int i : P.fwd_task)
        if (i < k0) fw.push_back(i);
    P.fwd_task.swap(fw);
    int df = 0;
    for (int i = 0; i < k0; i++) df = std::max(df, P.lvf[i]);
    P.nlv_f = k0 > 0 ? df + 1 : 0;
    // backward: tail rows are solved densely first; rows below k0 wait only
    // on rows below k0
    int db = 0;
    for (int i = n - 1; i >= 0; i--) {
        if (i >= k0) { P.lvb[i] = -1; continue; }
        int lv = 0;
        for (int e = P.urp[i]; e < P.urp[i + 1]; e++) {
            const int c = P.uci[e];
            if (c < k0) lv = std::max(lv, P.lvb[c] + 1);
        }
        P.lvb[i] = lv;
        db = std::max(db, lv);
    }
    P.nlv_b = k0 > 0 ? db + 1 : 0;
    std::vector<int> bw;
    bw.reserve(k0);
    {
        std::vector<int> cnt(P.nlv_b + 1, 0);
        for (int i = 0; i < k0; i++) cnt[P.lvb[i] + 1]++;
        for (int l = 0; l < P.nlv_b; l++) cnt[l + 1] += cnt[l];
        bw.assign(k0, 0);
        for (int i = 0; i < k0; i++) bw[cnt[P.lvb[i]]++] = i;
    }
    P.bwd_task.swap(bw);
    const int rc = lsp_plan_build_segments(P);
    return rc != 0 ? rc : best;
}

// ---------------------------------------------------------------------------
// Ruiz equilibration scales (D. Ruiz, "A scaling algorithm to equilibrate both
// rows and columns norms in matrices", RAL-TR-2001-034): each sweep divides
// every row by the square root of its largest |entry| of R A C, then every
// column likewise.  Reference behaviour: test_lu_gpu_f32.cu:247-269, with the
// per-sweep factor applied (so that the scaled matrix is exactly R A C).
void lsp_ruiz_scales(int n, long nnz, const int *row, const int *col, const double *val, int sweeps,
                     double *R, double *C)
{
    std::vector<double> mx(n);
    for (int i = 0; i < n; i++) { R[i] = 1.0; C[i] = 1.0; }
    for (int it = 0; it < sweeps; it++) {
        for (int side = 0; side < 2; side++) {
            const int *idx = side == 0 ? row : col;
            double *sc = side == 0 ? R : C;
            std::fill(mx.begin(), mx.end(), 0.0);
            for (long e = 0; e < nnz; e++) {
                const double v = std::fabs(val[e] * R[row[e]] * C[col[e]]);
                if (v > mx[idx[e]]) mx[idx[e]] = v;
            }
            for (int i = 0; i < n; i++)
                if (mx[i] > 0) sc[i] *= 1.0 / std::sqrt(mx[i]);
        }
    }
}

