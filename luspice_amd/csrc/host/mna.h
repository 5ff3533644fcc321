// mna.h -- synthetic MNA generator (host C++).
#pragma once
#include <cstdint>

#include "host_internal.h"

struct LspMnaSpec {
    int rows = 0, cols = 0, n_nodes = 0;
    double g_lo = 1e-6, g_hi = 1e2;
    double lr_frac = 0.0;
    double diag_frac = 0.0;   // probability of a diagonal grid edge (v, v+cols+1)
    double c_lo = 0.0, c_hi = 0.0, h = 1e-9;
    int n_src = 0;
    uint64_t seed = 0;
};

int lsp_mna_build(const LspMnaSpec *sp, LspCoo &out);
void lsp_mna_perturb(uint64_t seed, int step, double sigma, int nnz,
                     const double *val0, double *val, int n, double *b);
void lsp_spec_c0(LspMnaSpec *sp, uint64_t seed);
void lsp_spec_c1(LspMnaSpec *sp, uint64_t seed);
void lsp_spec_sized(LspMnaSpec *sp, int n, double lr_frac, uint64_t seed);
void lsp_spec_c2(LspMnaSpec *sp, int n, uint64_t seed);
