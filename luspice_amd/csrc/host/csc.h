// csc.h -- sparse format conversions (host C++).
#pragma once
#include <vector>

void lsp_coo_to_csc(int n, long nnz, const int *row, const int *col, const int *q,
                    std::vector<int> &Ap, std::vector<int> &Ai, std::vector<int> &src);
void lsp_coo_to_csr(int n, long nnz, const int *row, const int *col,
                    std::vector<int> &rp, std::vector<int> &ci, std::vector<int> &src);
int lsp_colamd_coo(int n, long nnz, const int *row, const int *col, int *q);
