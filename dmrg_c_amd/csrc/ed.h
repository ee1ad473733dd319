// Device-resident exact-diagonalization problem (ed.hip, ed_abi.cpp).
#pragma once

#include "entry.h"

struct dmrgx_ed {
  int32_t L = 0, d = 0;
  int64_t n = 0;
  dmrgx::DevBuf<int64_t> basis, pw;
  std::vector<std::vector<int64_t>> hrow;
  std::vector<std::vector<int32_t>> hcol;
  std::vector<std::vector<double>> hval;
  std::vector<dmrgx::DevBuf<int64_t>> op_row;
  std::vector<dmrgx::DevBuf<int32_t>> op_col;
  std::vector<dmrgx::DevBuf<double>> op_val;
  std::vector<dmrgx_ed_term> hterms;
  std::vector<int32_t> hn_ele;
  std::vector<int64_t> hpw;     // host sources of the uploads, kept until the problem is freed
  std::vector<char> hops;
  dmrgx::DevBuf<dmrgx_ed_term> terms;
  dmrgx::DevBuf<int32_t> n_ele;
  struct DOp {
    const int64_t *row;
    const int32_t *col;
    const double *val;
  };
  dmrgx::DevBuf<DOp> ops;
};

namespace dmrgx {
std::unique_ptr<dmrgx_ed> ed_create(int32_t L, int32_t d, int32_t n_ops, const CRS1 *ops, const int32_t *n_ele,
                                    int32_t n_terms, const dmrgx_ed_term *terms);
void ed_basis(dmrgx_ed &E, int64_t n, const int64_t *basis);
void ed_sector(dmrgx_ed &E, const int32_t *qn1, int32_t t1, const int32_t *qn2, int32_t t2);
std::unique_ptr<dmrgx_matrix> ed_matrix(const dmrgx_ed &E);
EdRows ed_rows(const dmrgx_ed &out, const dmrgx_ed &in, int32_t op, int32_t site, int32_t sign_flag);
}  // namespace dmrgx
