// mna.cpp -- seeded synthetic MNA circuit matrices.
//
// The reference benchmarks real SuiteSparse / ngspice matrices that are not
// in its repository (docs/06_reproduce.md:22-30); these generators make the
// workloads self-contained.  One builder covers every workload:
//
//   grid of rows x cols nodes (the last row may be partial), node v = r*cols+c
//   each grid edge      conductance g ~ logU[g_lo, g_hi], stamped +g on both
//                       diagonals and -g on both off-diagonals
//   diagonal edges      (r,c)-(r+1,c+1) with probability diag_frac, drawn
//                       right after the grid edges of node v (none when 0)
//   long-range branches round(lr_frac * n_grid_edges) distinct non-adjacent
//                       node pairs, same conductance law
//   node capacitors     C ~ logU[c_lo, c_hi] stamped as C/h on the diagonal
//                       (skipped when c_hi <= 0)
//   ideal sources       n_src distinct random nodes v, tied to ground through
//                       a branch current unknown s: A[v][N+s] = A[N+s][v] = 1,
//                       zero diagonal on row N+s
//   right-hand side     b_i ~ U[-1, 1]
//
// Random draws come from xoshiro256** (Blackman & Vigna, 2018) seeded with
// splitmix64, in a fixed order (edges, long-range pairs, capacitors, sources,
// b), so a seed always yields the same matrix on any host.  Entries are
// accumulated per (row, col) and emitted sorted by (row, col).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <map>
#include <unordered_set>
#include <vector>

#include "host_internal.h"
#include "mna.h"

namespace {

struct Rng {
    uint64_t s[4];
    explicit Rng(uint64_t seed) {
        uint64_t z = seed;
        for (int i = 0; i < 4; i++) {
            z += 0x9E3779B97F4A7C15ULL;
            uint64_t t = z;
            t = (t ^ (t >> 30)) * 0xBF58476D1CE4E5B9ULL;
            t = (t ^ (t >> 27)) * 0x94D049BB133111EBULL;
            s[i] = t ^ (t >> 31);
        }
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        uint64_t r = rotl(s[1] * 5, 7) * 9;
        uint64_t t = s[1] << 17;
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    double u01() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    double uniform(double a, double b) { return a + (b - a) * u01(); }
    double loguniform(double a, double b) {
        return std::exp(std::log(a) + (std::log(b) - std::log(a)) * u01());
    }
    int below(int m) { return (int)(u01() * m); }
    double normal() {   // Box-Muller, one value per call
        double u1 = u01(), u2 = u01();
        if (u1 < 1e-300) u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

}  // namespace

int lsp_mna_build(const LspMnaSpec *sp, LspCoo &out)
{
    if (sp->rows <= 0 || sp->cols <= 0 || sp->n_nodes <= 0 ||
        sp->n_nodes > sp->rows * sp->cols || sp->n_src < 0 || sp->n_src > sp->n_nodes)
        return -1;
    Rng rng(sp->seed);
    const int N = sp->n_nodes;
    const int n = N + sp->n_src;
    std::map<std::pair<int, int>, double> acc;
    auto stamp_g = [&](int a, int b, double g) {
        acc[{a, a}] += g;
        acc[{b, b}] += g;
        acc[{a, b}] -= g;
        acc[{b, a}] -= g;
    };
    std::unordered_set<uint64_t> adj;
    auto key = [](int a, int b) {
        if (a > b) std::swap(a, b);
        return ((uint64_t)a << 32) | (uint32_t)b;
    };
    int n_edges = 0;
    for (int v = 0; v < N; v++) {
        int r = v / sp->cols, c = v % sp->cols;
        if (c + 1 < sp->cols && v + 1 < N) {
            stamp_g(v, v + 1, rng.loguniform(sp->g_lo, sp->g_hi));
            adj.insert(key(v, v + 1));
            n_edges++;
        }
        if (v + sp->cols < N) {
            stamp_g(v, v + sp->cols, rng.loguniform(sp->g_lo, sp->g_hi));
            adj.insert(key(v, v + sp->cols));
            n_edges++;
        }
        if (sp->diag_frac > 0 && c + 1 < sp->cols && v + sp->cols + 1 < N && rng.u01() < sp->diag_frac) {
            stamp_g(v, v + sp->cols + 1, rng.loguniform(sp->g_lo, sp->g_hi));
            adj.insert(key(v, v + sp->cols + 1));
            n_edges++;
        }
        (void)r;
    }
    int n_lr = (int)std::lround(sp->lr_frac * n_edges);
    for (int t = 0, guard = 0; t < n_lr && guard < 100 * (n_lr + 10); guard++) {
        int a = rng.below(N), b = rng.below(N);
        if (a == b || adj.count(key(a, b))) continue;
        adj.insert(key(a, b));
        stamp_g(a, b, rng.loguniform(sp->g_lo, sp->g_hi));
        t++;
    }
    if (sp->c_hi > 0) {
        for (int v = 0; v < N; v++) acc[{v, v}] += rng.loguniform(sp->c_lo, sp->c_hi) / sp->h;
    }
    // distinct source nodes: partial Fisher-Yates over the node ids
    std::vector<int> nodes(N);
    for (int v = 0; v < N; v++) nodes[v] = v;
    for (int s = 0; s < sp->n_src; s++) {
        int j = s + rng.below(N - s);
        std::swap(nodes[s], nodes[j]);
        int v = nodes[s];
        acc[{v, N + s}] = 1.0;
        acc[{N + s, v}] = 1.0;
    }
    out = LspCoo();
    out.n = n;
    out.row.reserve(acc.size());
    out.col.reserve(acc.size());
    out.val.reserve(acc.size());
    for (auto &kv : acc) {           // std::map order = (row, col) ascending
        out.row.push_back(kv.first.first);
        out.col.push_back(kv.first.second);
        out.val.push_back(kv.second);
    }
    out.b.resize(n);
    for (int i = 0; i < n; i++) out.b[i] = rng.uniform(-1.0, 1.0);
    return 0;
}

// Newton-step perturbation: every stored value times lognormal(0, sigma),
// fresh b ~ U[-1,1]; seeded per (circuit seed, step).
void lsp_mna_perturb(uint64_t seed, int step, double sigma, int nnz,
                     const double *val0, double *val, int n, double *b)
{
    Rng rng(seed * 0x100000001B3ULL + (uint64_t)step * 0x9E3779B97F4A7C15ULL + 17);
    for (int e = 0; e < nnz; e++) val[e] = val0[e] * std::exp(sigma * rng.normal());
    for (int i = 0; i < n; i++) b[i] = rng.uniform(-1.0, 1.0);
}

// Workload presets --------------------------------------------------------
// Config 0: 25x40 grid, g logU[1e-6,1e2], 2% long-range, caps logU[1e-12,1e-9]
// at h = 1e-9, 20 sources (n = 1020), seed 0.
void lsp_spec_c0(LspMnaSpec *sp, uint64_t seed)
{
    *sp = LspMnaSpec();
    sp->rows = 25; sp->cols = 40; sp->n_nodes = 1000;
    sp->g_lo = 1e-6; sp->g_hi = 1e2; sp->lr_frac = 0.02;
    sp->c_lo = 1e-12; sp->c_hi = 1e-9; sp->h = 1e-9;
    sp->n_src = 20; sp->seed = seed;
}

// Config 1: 316x316 5-point grid, g logU[1e-3,1e1], 50 pad sources, no
// capacitors (n = 99906), seed 1.
void lsp_spec_c1(LspMnaSpec *sp, uint64_t seed)
{
    *sp = LspMnaSpec();
    sp->rows = 316; sp->cols = 316; sp->n_nodes = 316 * 316;
    sp->g_lo = 1e-3; sp->g_hi = 1e1; sp->lr_frac = 0.0;
    sp->c_lo = 0; sp->c_hi = 0; sp->h = 1e-9;
    sp->n_src = 50; sp->seed = seed;
}

// A circuit "like config 0" with n unknowns: nodes = round(n / 1.02), the rest
// sources (2% as in config 0), near-square grid, long-range fraction chosen
// for the requested nonzeros per row (lr_frac 0.02 gives ~4.9 per row).
void lsp_spec_sized(LspMnaSpec *sp, int n, double lr_frac, uint64_t seed)
{
    *sp = LspMnaSpec();
    int N = (int)std::lround(n / 1.02);
    if (N < 2) N = 2;
    if (N > n) N = n;
    int rows = (int)std::floor(std::sqrt((double)N));
    if (rows < 1) rows = 1;
    int cols = (N + rows - 1) / rows;
    sp->rows = rows; sp->cols = cols; sp->n_nodes = N;
    sp->g_lo = 1e-6; sp->g_hi = 1e2; sp->lr_frac = lr_frac;
    sp->c_lo = 1e-12; sp->c_hi = 1e-9; sp->h = 1e-9;
    sp->n_src = n - N; sp->seed = seed;
}

// Config 2 (precision sweep): like config 0 with n unknowns and ~6 nonzeros
// per row (2% long-range branches as in config 0, plus diagonal grid edges
// with probability 0.55, which keeps the fill of a 2-D mesh); conductances
// logU[g_lo, 1e3] with
// g_lo = 10^-(6 + 4u), u ~ U[0,1] drawn from the seed (about 10-35% of the entries
// below 6e-5), and node capacitors logU[1e-18, 1e-15] F at h = 1e-9, small
// enough not to make the diagonal dominant.
void lsp_spec_c2(LspMnaSpec *sp, int n, uint64_t seed)
{
    lsp_spec_sized(sp, n, 0.02, seed);
    sp->diag_frac = 0.55;
    Rng rng(seed ^ 0xC2C2C2C2ULL);
    sp->g_lo = std::pow(10.0, -(6.0 + 4.0 * rng.u01()));
    sp->g_hi = 1e3;
    sp->c_lo = 1e-18;
    sp->c_hi = 1e-15;
}

