// Exact-diagonalization entry points: the reference's EXACT_* symbols (include/exact.h:113-125)
// and the model builders of dmrgx_ed_model_build.
//
// EXACT_DIAGONALIZE_HAMILTONIAN, EXACT_V_M_Q0..Q3 and EXACT_EXPECTATION_* run on the device.
// The row helpers EXACT_MAKE_ELEM_ON/INTER, EXACT_GET_A_BASIS, EXACT_FREE_A_BASIS,
// EXACT_FIND_SITE_STATE and EXACT_SSD_COEFF stay on the host with the reference's results,
// so a program's own GET_HAM still links; dmrgx_ed_assemble is the device replacement for
// that loop.
#include <cmath>
#include <cstdlib>
#include <cstring>

#include "ed.h"

#pragma clang fp contract(off)

namespace dmrgx {
void launch_start_vector(int64_t n, uint64_t seed, double *out, hipStream_t s);
}  // namespace dmrgx

using namespace dmrgx;

namespace {

// ------------------------------------------------------------ device V_M helpers
struct DevOnsite {
  DevBuf<int64_t> row;
  DevBuf<int32_t> col;
  DevBuf<double> val;
};

void upload_onsite(const char *fn, const CRS1 *M, int d, DevOnsite &O) {
  if (!M || M->row_dim < d || !M->Row) fatal(fn, "M_On->row_dim=%d is smaller than dim_onsite=%d\n", M ? M->row_dim : -1, d);
  const int64_t nnz = M->Row[d];
  O.row.resize(d + 1);
  O.col.resize(std::max<int64_t>(nnz, 1));
  O.val.resize(std::max<int64_t>(nnz, 1));
  ctx().copy_in(O.row.p, M->Row, sizeof(int64_t) * (d + 1));
  if (nnz > 0) {
    ctx().copy_in(O.col.p, M->Col, sizeof(int32_t) * nnz);
    ctx().copy_in(O.val.p, M->Val, sizeof(double) * nnz);
  }
}

void upload_codes(const long *B, int n, DevBuf<int64_t> &out) {
  out.resize(std::max(n, 1));
  if (n > 0)
    ctx().copy_in(out.p, B, sizeof(int64_t) * n);
}

int64_t site_power(int d, int site) {
  int64_t p = 1;
  for (int s = 0; s < site; s++) p *= d;
  return p;
}

// Out = O(site) Vec from basis_in (n_in) to basis_out (n_out), EXACT_V_M_Q0.c:20-35 and siblings
void v_m(const char *fn, const CRS1 *M, int d, int site, const long *basis_out, int n_out, const long *basis_in,
         int n_in, const double *Vec, double *Out, bool sign, const int *N_Ele) {
  if (d <= 0 || site < 0) fatal(fn, "dim_onsite=%d, site=%d\n", d, site);
  run_or_die(fn, [&] {
    if (n_out <= 0) return;
    hipStream_t st = ctx().stream;
    DevOnsite O;
    upload_onsite(fn, M, d, O);
    DevBuf<int64_t> bo, bi;
    upload_codes(basis_out, n_out, bo);
    const int64_t *bin = bo.p;
    if (basis_in != basis_out || n_in != n_out) {
      upload_codes(basis_in, n_in, bi);
      bin = bi.p;
    }
    DevBuf<int32_t> ne(d);
    if (sign) ctx().copy_in(ne.p, N_Ele, sizeof(int32_t) * d);
    DevVec x(std::max(n_in, 1)), y(n_out);
    if (n_in > 0) x.up(Vec);
    const EdRows rows{d, site, sign ? 1 : 0, site_power(d, site), n_in, bo.p, bin, O.row.p, O.col.p, O.val.p, ne.p};
    Reduce R;
    launch_ed_apply(n_out, rows, x.d.p, y.d.p, nullptr, nullptr, R, st);
    y.down(Out);
  });
}

// ------------------------------------------------------------ on-site operators
struct HostOp {
  int d = 0;
  std::vector<int64_t> row;
  std::vector<int32_t> col;
  std::vector<double> val;
  CRS1 view() {
    CRS1 c{};
    c.max_val = (long)val.size();
    c.max_row = d + 1;
    c.row_dim = c.col_dim = d;
    c.Row = reinterpret_cast<long *>(row.data());
    c.Col = col.data();
    c.Val = val.data();
    return c;
  }
};

constexpr double kOpZero = 1.0e-15;  // the |val| > 10^-15 filter of onsite_op/*.c

// element-wise construction with the filter of onsite_op/ONSITE_*.c (row-major, columns ascending)
template <class F>
HostOp make_op(int d, F &&elem) {
  HostOp o;
  o.d = d;
  o.row.assign(d + 1, 0);
  for (int r = 0; r < d; r++) {
    for (int c = 0; c < d; c++) {
      const double v = elem(r, c);
      if (std::fabs(v) > kOpZero) {
        o.col.push_back(c);
        o.val.push_back(v);
      }
    }
    o.row[r + 1] = (int64_t)o.col.size();
  }
  return o;
}

// sml/MATRIX_SUM_CRS1.c: row-wise merge; coinciding entries are summed and dropped when
// |sum| <= 1e-15, the others copied as they are
HostOp op_sum(const HostOp &a, const HostOp &b) {
  HostOp o;
  o.d = a.d;
  o.row.assign(a.d + 1, 0);
  for (int r = 0; r < a.d; r++) {
    int64_t i = a.row[r], j = b.row[r];
    while (i < a.row[r + 1] || j < b.row[r + 1]) {
      if (j >= b.row[r + 1] || (i < a.row[r + 1] && a.col[i] < b.col[j])) {
        o.col.push_back(a.col[i]);
        o.val.push_back(a.val[i++]);
      } else if (i >= a.row[r + 1] || b.col[j] < a.col[i]) {
        o.col.push_back(b.col[j]);
        o.val.push_back(b.val[j++]);
      } else {
        const double s = a.val[i] + b.val[j];
        if (std::fabs(s) > kOpZero) {
          o.col.push_back(a.col[i]);
          o.val.push_back(s);
        }
        i++;
        j++;
      }
    }
    o.row[r + 1] = (int64_t)o.col.size();
  }
  return o;
}

HostOp op_scaled(HostOp o, double c) {
  for (double &v : o.val) v = v * c;
  return o;
}

HostOp zero_diag(int d) {
  HostOp o;
  o.d = d;
  for (int r = 0; r < d; r++) {
    o.col.push_back(r);
    o.val.push_back(0.0);
    o.row.push_back(r);
  }
  o.row.push_back(d);
  return o;
}

// onsite_op/ONSITE_{SP,SM,SZ,SZSZ}_SZBASIS_HB.c (spin = 2S, 1-based a = row+1, b = col+1)
HostOp hb_sp(int spin) {
  return make_op(spin + 1, [&](int r, int c) {
    const int a = r + 1, b = c + 1;
    return (a + 1 == b) ? std::sqrt((spin * 0.5 + 1) * (a + b - 1.0) - a * b) : 0.0;
  });
}
HostOp hb_sm(int spin) {
  return make_op(spin + 1, [&](int r, int c) {
    const int a = r + 1, b = c + 1;
    return (a == b + 1) ? std::sqrt((spin * 0.5 + 1) * (a + b - 1.0) - a * b) : 0.0;
  });
}
HostOp hb_sz(int spin) {
  return make_op(spin + 1, [&](int r, int c) { return (spin * 0.5 + 1.0 - (c + 1)) * (r == c ? 1 : 0); });
}
HostOp hb_szsz(int spin) {
  return make_op(spin + 1, [&](int r, int c) {
    const double v = (spin * 0.5 + 1.0 - (c + 1)) * (r == c ? 1 : 0);
    return v * v;
  });
}

// onsite_op/ONSITE_*_SZBASIS_HUBBARD.c, local states 0 empty, 1 up, 2 down, 3 up+down
HostOp hub_op(double (*f)(int, int)) { return make_op(4, f); }
double h_cupd(int r, int c) { return ((c == 0 && r == 1) || (c == 2 && r == 3)) ? 1.0 : 0.0; }
double h_cup(int r, int c) { return ((c == 1 && r == 0) || (c == 3 && r == 2)) ? 1.0 : 0.0; }
double h_cdnd(int r, int c) { return (c == 0 && r == 2) ? 1.0 : (c == 1 && r == 3) ? -1.0 : 0.0; }
double h_cdn(int r, int c) { return (c == 2 && r == 0) ? 1.0 : (c == 3 && r == 1) ? -1.0 : 0.0; }
double h_nup(int r, int c) { return ((c == 1 && r == 1) || (c == 3 && r == 3)) ? 1.0 : 0.0; }
double h_ndn(int r, int c) { return ((c == 2 && r == 2) || (c == 3 && r == 3)) ? 1.0 : 0.0; }
double h_nc(int r, int c) { return ((c == 1 && r == 1) || (c == 2 && r == 2)) ? 1.0 : (c == 3 && r == 3) ? 2.0 : 0.0; }
double h_szc(int r, int c) { return (c == 1 && r == 1) ? 0.5 : (c == 2 && r == 2) ? -0.5 : 0.0; }
double h_ncnc(int r, int c) { return (c == 3 && r == 3) ? 1.0 : 0.0; }

HostOp hub_scaled(double (*f)(int, int), double coeff) {
  return make_op(4, [&](int r, int c) { return f(r, c) * coeff; });
}

double ssd_coeff(int site, int tot_site, const char *bc, bool onsite) {
  // exact/EXACT_SSD_COEFF.c:8-30
  site = site + 1;
  double temp;
  if (strcmp(bc, "SSD") == 0) temp = onsite ? sin(M_PI * (site - 0.5) / tot_site) : sin(M_PI * site / tot_site);
  else temp = 1.0;
  return temp * temp;
}

dmrgx_ed_term on_term(int site, int op, double c) { return dmrgx_ed_term{0, site, 0, op, 0, 0, 0, 1, c}; }
dmrgx_ed_term inter(int s1, int s2, int o1, int o2, double c, int sign, int lo = 0, int hi = 0) {
  return dmrgx_ed_term{1, s1, s2, o1, o2, lo, hi, sign, c};
}

struct ModelDesc {
  int d = 0;
  std::vector<HostOp> ops;
  std::vector<dmrgx_ed_term> terms;
  std::vector<int32_t> qn1, qn2, n_ele;
  int32_t t1 = 0, t2 = 0;
};

// Host-only description of a model: operators, term list, sector tables.
ModelDesc describe_model(const dmrgx_ed_model &m) {
  const int L = m.L;
  if (L < 2 || L > 63) throw Failure(DMRGX_ERR_ARG, "ED model needs 2 <= L <= 63");
  if (m.bc < 0 || m.bc > 2) throw Failure(DMRGX_ERR_ARG, "bc must be 0 (OBC), 1 (PBC) or 2 (SSD)");
  const char *bc = m.bc == 1 ? "PBC" : m.bc == 2 ? "SSD" : "OBC";
  ModelDesc D;
  std::vector<HostOp> &ops = D.ops;
  std::vector<dmrgx_ed_term> &terms = D.terms;
  std::vector<int32_t> &qn1 = D.qn1, &qn2 = D.qn2, &n_ele = D.n_ele;
  int &d = D.d;
  for (int s = 0; s < L; s++) terms.push_back(on_term(s, 0, ssd_coeff(s, L, bc, true)));
  if (m.model == 0) {
    if (m.spin <= 0 || m.spin > 15) throw Failure(DMRGX_ERR_ARG, "spin (2S) must be in 1..15");
    d = m.spin + 1;
    // onsite D_z SzSz + h_z Sz (main/dmrg/xxz_vf/src/main/MAKE_ONSITE_HAM.c)
    ops.push_back(op_sum(op_scaled(hb_szsz(m.spin), m.D_z), op_scaled(hb_sz(m.spin), m.h_z)));
    ops.push_back(hb_sp(m.spin));
    ops.push_back(hb_sm(m.spin));
    ops.push_back(hb_sz(m.spin));
    ops.push_back(zero_diag(d));
    for (int s = 0; s + 1 < L; s++) {
      const double c = ssd_coeff(s, L, bc, false);
      terms.push_back(inter(s, s + 1, 1, 2, 0.5 * m.J_xy * c, 1));
      terms.push_back(inter(s, s + 1, 2, 1, 0.5 * m.J_xy * c, 1));
      terms.push_back(inter(s, s + 1, 3, 3, m.J_z * c, 1));
    }
    if (m.bc == 1 && L > 2) {
      terms.push_back(inter(0, L - 1, 1, 2, 0.5 * m.J_xy, 1));
      terms.push_back(inter(0, L - 1, 2, 1, 0.5 * m.J_xy, 1));
      terms.push_back(inter(0, L - 1, 3, 3, m.J_z, 1));
    }
    terms.push_back(on_term(0, 4, 0.0));
    for (int a = 0; a < d; a++) qn1.push_back(m.spin - 2 * a);
  } else if (m.model == 1) {
    d = 4;
    // onsite h_z Sz + U n_up n_dn + mu n (onsite_op/ONSITE_MAKE_HAM_SZBASIS_HUBBARD.c:14-22)
    ops.push_back(op_sum(op_sum(hub_scaled(h_szc, m.h_z), hub_scaled(h_ncnc, m.U)), hub_scaled(h_nc, m.mu)));
    ops.push_back(hub_op(h_cupd));
    ops.push_back(hub_op(h_cup));
    ops.push_back(hub_op(h_cdnd));
    ops.push_back(hub_op(h_cdn));
    ops.push_back(hub_op(h_nup));
    ops.push_back(hub_op(h_ndn));
    ops.push_back(zero_diag(4));
    ops.push_back(hub_op(h_nc));
    ops.push_back(hub_op(h_szc));
    for (int s = 0; s + 1 < L; s++) {
      const double c = ssd_coeff(s, L, bc, false);
      // hopping sign from the electron parity of site s of the row state
      terms.push_back(inter(s, s + 1, 1, 2, m.t * c, 1, s, s + 1));
      terms.push_back(inter(s, s + 1, 2, 1, m.t * c, -1, s, s + 1));
      terms.push_back(inter(s, s + 1, 3, 4, m.t * c, 1, s, s + 1));
      terms.push_back(inter(s, s + 1, 4, 3, m.t * c, -1, s, s + 1));
      terms.push_back(inter(s, s + 1, 5, 5, m.V * c, 1));
      terms.push_back(inter(s, s + 1, 6, 6, m.V * c, 1));
    }
    if (m.bc == 1 && L > 2) {
      // wrap: parity of sites 0..L-2
      terms.push_back(inter(0, L - 1, 1, 2, m.t, 1, 0, L - 1));
      terms.push_back(inter(0, L - 1, 2, 1, m.t, -1, 0, L - 1));
      terms.push_back(inter(0, L - 1, 3, 4, m.t, 1, 0, L - 1));
      terms.push_back(inter(0, L - 1, 4, 3, m.t, -1, 0, L - 1));
      terms.push_back(inter(0, L - 1, 5, 5, m.V, 1));
      terms.push_back(inter(0, L - 1, 6, 6, m.V, 1));
    }
    terms.push_back(on_term(0, 7, 0.0));
    n_ele = {0, 1, 1, 2};   // onsite_op/ONSITE_FIND_SITE_ELE_SZBASIS_HUBBARD.c
    qn1 = n_ele;
    qn2 = {0, 1, -1, 0};    // onsite_op/ONSITE_FIND_SITE_SZ_SZBASIS_HUBBARD.c
  } else {
    throw Failure(DMRGX_ERR_ARG, "model must be 0 (XXZ) or 1 (Hubbard)");
  }
  if (m.model == 0) {
    D.t1 = m.sz;
  } else {
    D.t1 = m.N;
    D.t2 = m.sz;
  }
  return D;
}

std::unique_ptr<dmrgx_ed> build_model(const dmrgx_ed_model &m) {
  ModelDesc D = describe_model(m);
  std::vector<CRS1> views;
  for (HostOp &o : D.ops) views.push_back(o.view());
  auto E = ed_create(m.L, D.d, (int32_t)views.size(), views.data(), D.n_ele.empty() ? nullptr : D.n_ele.data(),
                     (int32_t)D.terms.size(), D.terms.data());
  ed_sector(*E, D.qn1.data(), D.t1, D.qn2.empty() ? nullptr : D.qn2.data(), D.t2);
  return E;
}

// ------------------------------------------------------------ host row accumulation
// Insert-or-add into the sorted accumulator (Check ascending, Basis == Check):
// EXACT_MAKE_ELEM_ON.c:18-29; an ordered insertion gives the quicksort's order.
void acc_put(const char *fn, EXACT_A_BASIS *A, long *elem_num, long code, double v) {
  long lo = 0, hi = *elem_num;
  while (lo < hi) {
    const long mid = lo + (hi - lo) / 2;
    if (A->Check[mid] < code) lo = mid + 1;
    else hi = mid;
  }
  if (lo < *elem_num && A->Check[lo] == code) {
    A->Val[lo] = A->Val[lo] + v;
    return;
  }
  if (*elem_num >= A->max_row) fatal(fn, "A_Basis->max_row=%d is too small\n", A->max_row);
  const long tail = *elem_num - lo;
  memmove(A->Check + lo + 1, A->Check + lo, sizeof(long) * tail);
  memmove(A->Basis + lo + 1, A->Basis + lo, sizeof(long) * tail);
  memmove(A->Val + lo + 1, A->Val + lo, sizeof(double) * tail);
  A->Check[lo] = code;
  A->Basis[lo] = code;
  A->Val[lo] = v;
  (*elem_num)++;
}

long ipow_l(int d, int site) {
  long p = 1;
  for (int s = 0; s < site; s++) p *= d;
  return p;
}

}  // namespace

extern "C" {

// ------------------------------------------------------------ reference symbols
int EXACT_FIND_SITE_STATE(long basis, int site, int dim_onsite) {
  for (int i = 0; i < site; i++) basis = basis / (long)dim_onsite;
  return (int)(basis % dim_onsite);
}

double EXACT_SSD_COEFF(int site, int tot_site, char BC[], char Inter_Name[]) {
  if (strcmp(BC, "SSD") == 0 && strcmp(Inter_Name, "Onsite") != 0 && strcmp(Inter_Name, "Intersite") != 0)
    fatal("EXACT_SSD_COEFF", "%s", "");
  return ssd_coeff(site, tot_site, BC, strcmp(Inter_Name, "Onsite") == 0);
}

EXACT_A_BASIS **EXACT_GET_A_BASIS(int p_threads, int max_row) {
  EXACT_A_BASIS **A = (EXACT_A_BASIS **)malloc(sizeof(EXACT_A_BASIS *) * p_threads);
  if (!A) fatal("EXACT_GET_A_BASIS", "Need more memory\n");
  for (int t = 0; t < p_threads; t++) {
    A[t] = (EXACT_A_BASIS *)calloc(1, sizeof(EXACT_A_BASIS));
    A[t]->Basis = (long *)calloc(max_row, sizeof(long));
    A[t]->Check = (long *)malloc(sizeof(long) * max_row);
    A[t]->Val = (double *)calloc(max_row, sizeof(double));
    if (!A[t] || !A[t]->Basis || !A[t]->Check || !A[t]->Val) fatal("EXACT_GET_A_BASIS", "Need more memory\n");
    for (int i = 0; i < max_row; i++) A[t]->Check[i] = -1;
    A[t]->max_row = max_row;
    A[t]->elem_num = 0;
  }
  return A;
}

void EXACT_FREE_A_BASIS(EXACT_A_BASIS **A, int p_threads) {
  for (int t = 0; t < p_threads; t++) {
    free(A[t]->Basis);
    free(A[t]->Check);
    free(A[t]->Val);
    free(A[t]);
  }
  free(A);
}

void EXACT_MAKE_ELEM_ON(long basis, int site, int dim_onsite, CRS1 *M_On, long *elem_num, double coeef,
                        EXACT_A_BASIS *A_Basis) {
  const int local = EXACT_FIND_SITE_STATE(basis, site, dim_onsite);
  const long p = ipow_l(dim_onsite, site);
  for (long it = M_On->Row[local]; it < M_On->Row[local + 1]; it++)
    acc_put("EXACT_MAKE_ELEM_ON", A_Basis, elem_num, basis + (M_On->Col[it] - local) * p, coeef * M_On->Val[it]);
}

void EXACT_MAKE_ELEM_INTER(long basis, int site1, int site2, int dim_onsite, CRS1 *M1, CRS1 *M2, long *elem_num,
                           double coeef, int sign, EXACT_A_BASIS *A_Basis) {
  const int l1 = EXACT_FIND_SITE_STATE(basis, site1, dim_onsite);
  const int l2 = EXACT_FIND_SITE_STATE(basis, site2, dim_onsite);
  const long p1 = ipow_l(dim_onsite, site1), p2 = ipow_l(dim_onsite, site2);
  for (long i = M1->Row[l1]; i < M1->Row[l1 + 1]; i++)
    for (long j = M2->Row[l2]; j < M2->Row[l2 + 1]; j++)
      acc_put("EXACT_MAKE_ELEM_INTER", A_Basis, elem_num, basis + (M1->Col[i] - l1) * p1 + (M2->Col[j] - l2) * p2,
              sign * coeef * M1->Val[i] * M2->Val[j]);
}

void EXACT_V_M_Q0(CRS1 *M_On, double *Vec, double *Out_Vec, int dim_onsite, int site, int p_threads,
                  EXACT_BASIS_INFO *Basis_Info) {
  (void)p_threads;
  v_m("EXACT_V_M_Q0", M_On, dim_onsite, site, Basis_Info->Basis, Basis_Info->dim, Basis_Info->Basis, Basis_Info->dim,
      Vec, Out_Vec, false, nullptr);
}

void EXACT_V_M_Q1(CRS1 *M_On, int qn_out, double *Vec, int qn_in, double *Out_Vec, int dim_onsite, int site,
                  int p_threads, EXACT_WHOLE_BASIS_Q1 *W) {
  (void)p_threads;
  if (qn_out < 0 || qn_in < 0) fatal("EXACT_V_M_Q1", "qn_out=%d,qn_in=%d\n", qn_out, qn_in);
  v_m("EXACT_V_M_Q1", M_On, dim_onsite, site, W->Basis[qn_out], W->Dim[qn_out], W->Basis[qn_in], W->Dim[qn_in], Vec,
      Out_Vec, false, nullptr);
}

void EXACT_V_M_Q2(CRS1 *M_On, int qn1_out, int qn2_out, double *Vec, int qn1_in, int qn2_in, double *Out_Vec,
                  char Sign_Flag[], int *N_Ele, int dim_onsite, int op_site, int p_threads, EXACT_WHOLE_BASIS_Q2 *W) {
  (void)p_threads;
  if (qn1_out < 0 || qn2_out < 0 || qn1_in < 0 || qn2_in < 0)
    fatal("EXACT_V_M_Q2", "qn1_out=%d,qn2_out=%d,qn1_in=%d,qn2_in=%d\n", qn1_out, qn2_out, qn1_in, qn2_in);
  const bool sign = strcmp(Sign_Flag, "Yes") == 0;
  v_m("EXACT_V_M_Q2", M_On, dim_onsite, op_site, W->Basis[qn1_out][qn2_out], W->Dim[qn1_out][qn2_out],
      W->Basis[qn1_in][qn2_in], W->Dim[qn1_in][qn2_in], Vec, Out_Vec, sign, N_Ele);
}

void EXACT_V_M_Q3(CRS1 *M_On, int qn1_out, int qn2_out, int qn3_out, double *Vec, int qn1_in, int qn2_in, int qn3_in,
                  double *Out_Vec, char Sign_Flag[], int *N_Ele, int dim_onsite, int op_site, int p_threads,
                  EXACT_WHOLE_BASIS_Q3 *W) {
  (void)p_threads;
  if (qn1_out < 0 || qn2_out < 0 || qn3_out < 0 || qn1_in < 0 || qn2_in < 0 || qn3_in < 0)
    fatal("EXACT_V_M_Q3", "qn1_out=%d,qn2_out=%d,qn3_out=%d,qn1_in=%d,qn2_in=%d,qn3_in=%d\n", qn1_out, qn2_out,
          qn3_out, qn1_in, qn2_in, qn3_in);
  const bool sign = strcmp(Sign_Flag, "Yes") == 0;
  v_m("EXACT_V_M_Q3", M_On, dim_onsite, op_site, W->Basis[qn1_out][qn2_out][qn3_out], W->Dim[qn1_out][qn2_out][qn3_out],
      W->Basis[qn1_in][qn2_in][qn3_in], W->Dim[qn1_in][qn2_in][qn3_in], Vec, Out_Vec, sign, N_Ele);
}

// exact/EXACT_EXPECTATION_ONSITE.c:6-11: Out[site] = <Vec | O_site Vec>, T_Vec = the last product
void EXACT_EXPECTATION_ONSITE(CRS1 *M_On, double *Out, double *Vec, double *T_Vec, int dim_onsite, int tot_site,
                              int p_threads, EXACT_BASIS_INFO *BI) {
  (void)p_threads;
  static const char *fn = "EXACT_EXPECTATION_ONSITE";
  run_or_die(fn, [&] {
    const int n = BI->dim;
    if (n <= 0 || tot_site <= 0) return;
    hipStream_t st = ctx().stream;
    DevOnsite O;
    upload_onsite(fn, M_On, dim_onsite, O);
    DevBuf<int64_t> b;
    upload_codes(BI->Basis, n, b);
    DevVec v(n), t(n);
    v.up(Vec);
    DevBuf<double> res(tot_site);
    DevBuf<int32_t> ne(1);
    Reduce R;
    for (int site = 0; site < tot_site; site++) {
      const EdRows rows{dim_onsite, site, 0, site_power(dim_onsite, site), n, b.p, b.p, O.row.p, O.col.p, O.val.p, ne.p};
      launch_ed_apply(n, rows, v.d.p, t.d.p, v.d.p, res.p + site, R, st);
    }
    ctx().copy_out(Out, res.p, sizeof(double) * tot_site);
    if (T_Vec) t.down(T_Vec);
    HIPCHK(hipStreamSynchronize(st));
  });
}

// exact/EXACT_EXPECTATION_INTERSITE_Q0.c:6-14: T1 = O_start Vec; Out[r] = <T1 | R_{start+r} Vec>
void EXACT_EXPECTATION_INTERSITE_Q0(CRS1 *M_O, CRS1 *M_R, int start, int end, double *Out, double *Vec,
                                    double *T_Vec1, double *T_Vec2, int dim_onsite, int p_threads,
                                    EXACT_BASIS_INFO *BI) {
  (void)p_threads;
  static const char *fn = "EXACT_EXPECTATION_INTERSITE_Q0";
  run_or_die(fn, [&] {
    const int n = BI->dim;
    if (n <= 0 || start < 0) return;
    hipStream_t st = ctx().stream;
    DevOnsite O, Rop;
    upload_onsite(fn, M_O, dim_onsite, O);
    upload_onsite(fn, M_R, dim_onsite, Rop);
    DevBuf<int64_t> b;
    upload_codes(BI->Basis, n, b);
    DevVec v(n), t1(n), t2(n);
    v.up(Vec);
    const int len = end - start;
    DevBuf<double> res(std::max(len, 1));
    DevBuf<int32_t> ne(1);
    Reduce R;
    const EdRows r0{dim_onsite, start, 0, site_power(dim_onsite, start), n, b.p, b.p, O.row.p, O.col.p, O.val.p, ne.p};
    launch_ed_apply(n, r0, v.d.p, t1.d.p, nullptr, nullptr, R, st);
    for (int site = start; site < end; site++) {
      const EdRows rr{dim_onsite, site, 0, site_power(dim_onsite, site), n, b.p, b.p, Rop.row.p, Rop.col.p, Rop.val.p, ne.p};
      launch_ed_apply(n, rr, v.d.p, t2.d.p, t1.d.p, res.p + (site - start), R, st);
    }
    if (len > 0) ctx().copy_out(Out, res.p, sizeof(double) * len);
    if (T_Vec1) t1.down(T_Vec1);
    if (T_Vec2 && len > 0) t2.down(T_Vec2);
    HIPCHK(hipStreamSynchronize(st));
  });
}

// exact/EXACT_DIAGONALIZE_HAMILTONIAN.c:8-70
void EXACT_DIAGONALIZE_HAMILTONIAN(EXACT_HAM_INFO *Ham_Info, EXACT_PARAMETER *Param, EXACT_TIME *Time,
                                   int p_threads) {
  (void)p_threads;
  const bool lan = strcmp(Param->Diag_Method, "Lanczos") == 0;
  const bool slow = strcmp(Param->Diag_Method, "Lanczos_Slow") == 0;
  if (!lan && !slow) fatal("DMRG_DIAGONALIZE_HAMILTONIAN", "%s", "");
  CRS1 *H = Ham_Info->Ham;
  const char *ln = slow ? "LANCZOS_SLOW" : "LANCZOS";
  check_square(ln, H);
  if (Param->diag_max_step <= Param->diag_min_step)
    fatal(ln, "Box->max_step=%d,Box->min_step=%d\n", Param->diag_max_step, Param->diag_min_step);
  Ham_Info->diag_num = Param->diag_num;
  const int n = H->row_dim;
  run_or_die("EXACT_DIAGONALIZE_HAMILTONIAN", [&] {
    const double t0 = now_s();
    auto A = upload_csr(n, H->Row, H->Col, H->Val, false);
    DevVec v(n);
    SolveOut out;
    double eig;
    if (n == 1) {
      eig = H->Val[0];
      const double one = 1.0;
      v.up(&one);
    } else {
      launch_start_vector(n, ctx().seed, v.d.p, ctx().stream);  // Guess "No" (:22)
      LanczosOpts o;
      o.acc = Param->diag_acc;
      o.min_step = Param->diag_min_step;
      o.max_step = Param->diag_max_step;
      o.max_block_step = slow ? 0 : Param->diag_max_step;
      o.type = 0;  // "Normal" (:21)
      lanczos(op_of(*A), o, v.d.p, out);
      if (!out.lan_converged) {
        sml_append("error.txt", "Error in %s\nDoes not converge (max step is %d)\n", ln, Param->diag_max_step);
        fatal(ln, "Does not converge (max step is %d)\n", Param->diag_max_step);
      }
      eig = out.eig;
      sml_append(slow ? "LANCZOS_SLOW_Step.txt" : "LANCZOS_Step.txt", "%d\n", out.lan_steps);
    }
    const double t1 = now_s();
    inverse_iteration(op_of(*A), eig, v.d.p, Param->inv_iter_acc, Param->inv_iter_diag_add, Param->inv_iter_max_step,
                      Param->cg_acc, Param->cg_max_step, out);
    if (out.ii_converged) {
      sml_append("Inv_iter_step.txt", "%d\n", out.ii_steps);
    } else {
      printf("Error in INVERSE_ITERATION\nDoes not converge(%e > %e) (max step is %d)\ncontinue...\n", out.err,
             Param->inv_iter_acc, Param->inv_iter_max_step);
      sml_append("error.txt", "Error in INVERSE_ITERATION\nDoes not converge(%e > %e) (max step is %d)\ncontinue...\n",
                 out.err, Param->inv_iter_acc, Param->inv_iter_max_step);
    }
    if (Ham_Info->diag_num >= 2) {
      printf("Underconstruction\n");
      exit(1);
    }
    v.down(Ham_Info->Vector[0]);
    Ham_Info->Value[0] = eig;
    Ham_Info->Error[0] = converge_check(op_of(*A), v.d.p, eig);  // CONVERGE_CHECK (:65)
    Ham_Info->mem_ham = (double)H->Row[n] * 12 * pow(10, -9) + (double)n * 8 * pow(10, -9);
    Ham_Info->dim = n;
    if (Time) {
      Time->diag = t1 - t0;
      Time->inv_iter = now_s() - t1;
    }
  });
}

// ------------------------------------------------------------ dmrgx_ed models
int dmrgx_ed_model_build(const dmrgx_ed_model *m, dmrgx_ed **out) {
  if (!m || !out) return DMRGX_ERR_ARG;
  return run_status([&] { *out = build_model(*m).release(); });
}

int dmrgx_ed_model_describe(const dmrgx_ed_model *m, int32_t *d, int32_t *n_ops, int64_t *op_nnz, int64_t *op_row,
                            int32_t *op_col, double *op_val, int32_t *n_terms, dmrgx_ed_term *terms, int32_t *n_ele,
                            int32_t *qn1, int32_t *qn2, int32_t *targets) {
  if (!m || !d || !n_ops || !n_terms) return DMRGX_ERR_ARG;
  return run_status([&] {
    const ModelDesc D = describe_model(*m);
    *d = D.d;
    *n_ops = (int32_t)D.ops.size();
    *n_terms = (int32_t)D.terms.size();
    int64_t at = 0;
    for (size_t k = 0; k < D.ops.size(); k++) {
      const HostOp &o = D.ops[k];
      if (op_nnz) op_nnz[k] = (int64_t)o.col.size();
      if (op_row) std::copy(o.row.begin(), o.row.end(), op_row + k * (D.d + 1));
      if (op_col) std::copy(o.col.begin(), o.col.end(), op_col + at);
      if (op_val) std::copy(o.val.begin(), o.val.end(), op_val + at);
      at += (int64_t)o.col.size();
    }
    if (terms) std::copy(D.terms.begin(), D.terms.end(), terms);
    for (int a = 0; a < D.d; a++) {
      if (n_ele) n_ele[a] = D.n_ele.empty() ? 0 : D.n_ele[a];
      if (qn1) qn1[a] = D.qn1[a];
      if (qn2) qn2[a] = D.qn2.empty() ? 0 : D.qn2[a];
    }
    if (targets) {
      targets[0] = D.t1;
      targets[1] = D.qn2.empty() ? 0 : D.t2;
      targets[2] = D.qn2.empty() ? 0 : 1;
    }
  });
}

}  // extern "C"
