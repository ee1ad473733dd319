// Lowest eigenpair of a symmetric tridiagonal matrix, the part of dstev_ the Lanczos
// drivers use (sml/LANCZOS_SLOW_SYM.c:127-137 take Lap_D[0] and the first eigenvector).
//
// Eigenvalue: bisection on the Sturm count of T - x I (number of negative pivots of its
// LDL^T factorisation), started from the Gershgorin interval, run until the interval
// stops shrinking (full double precision).  Eigenvector: inverse iteration with the
// tridiagonal LU factorisation of T - lam I using partial pivoting (row interchanges as
// in LAPACK's dgttrf), three sweeps from a constant vector.
#include <algorithm>
#include <cmath>
#include <vector>

#include "internal.h"

namespace dmrgx {

namespace {
int count_below(int k, const double *a, const double *b, double x) {
  int c = 0;
  double d = a[0] - x;
  if (d < 0) c++;
  for (int i = 1; i < k; i++) {
    if (d == 0.0) d = -1e-300;  // perturb an exact zero pivot
    d = (a[i] - x) - (b[i - 1] / d) * b[i - 1];
    if (d < 0) c++;
  }
  return c;
}
}  // namespace

double tridiag_lowest_value(int k, const double *a, const double *b, double upper_hint) {
  if (k == 1) return a[0];
  double lo = a[0], hi = a[0];
  for (int i = 0; i < k; i++) {
    const double rad = (i > 0 ? std::fabs(b[i - 1]) : 0.0) + (i + 1 < k ? std::fabs(b[i]) : 0.0);
    lo = std::min(lo, a[i] - rad);
    hi = std::max(hi, a[i] + rad);
  }
  const double pad = 1e-13 * (hi - lo) + 1e-300;
  lo -= pad;
  hi += pad;
  if (std::isfinite(upper_hint) && upper_hint > lo && upper_hint < hi &&
      count_below(k, a, b, upper_hint) >= 1)
    hi = upper_hint;
  for (int it = 0; it < 256; it++) {
    const double mid = lo + 0.5 * (hi - lo);
    if (!(mid > lo && mid < hi)) break;
    if (count_below(k, a, b, mid) >= 1) hi = mid;
    else lo = mid;
  }
  return lo + 0.5 * (hi - lo);
}

void tridiag_vector(int k, const double *a, const double *b, double lam, double *v) {
  if (k == 1) {
    v[0] = 1.0;
    return;
  }
  std::vector<double> d(k), lo(k), up(k), up2(k);
  std::vector<char> swapped(k, 0);
  double scale = 0.0;
  for (int i = 0; i < k; i++) scale = std::max(scale, std::fabs(a[i]) + (i + 1 < k ? std::fabs(b[i]) : 0.0));
  const double tiny = std::max(scale, 1e-300) * 1e-16;
  for (int i = 0; i < k; i++) {
    d[i] = a[i] - lam;
    lo[i] = (i + 1 < k) ? b[i] : 0.0;
    up[i] = lo[i];
    up2[i] = 0.0;
  }
  // LU with partial pivoting: row i+1 is swapped up when its sub-diagonal is larger
  for (int i = 0; i + 1 < k; i++) {
    if (std::fabs(d[i]) >= std::fabs(lo[i])) {
      if (d[i] == 0.0) d[i] = tiny;
      const double m = lo[i] / d[i];
      lo[i] = m;
      d[i + 1] -= m * up[i];
    } else {
      swapped[i] = 1;
      const double m = d[i] / lo[i];
      d[i] = lo[i];
      lo[i] = m;
      const double t = up[i];
      up[i] = d[i + 1];
      d[i + 1] = t - m * d[i + 1];
      if (i + 2 < k) {
        up2[i] = up[i + 1];
        up[i + 1] = -m * up[i + 1];
      }
    }
  }
  if (d[k - 1] == 0.0) d[k - 1] = tiny;
  for (int i = 0; i < k; i++) v[i] = 1.0;
  for (int sweep = 0; sweep < 3; sweep++) {
    for (int i = 0; i + 1 < k; i++) {
      if (!swapped[i]) {
        v[i + 1] -= lo[i] * v[i];
      } else {
        const double t = v[i];
        v[i] = v[i + 1];
        v[i + 1] = t - lo[i] * v[i];
      }
    }
    v[k - 1] /= d[k - 1];
    v[k - 2] = (v[k - 2] - up[k - 2] * v[k - 1]) / d[k - 2];
    for (int i = k - 3; i >= 0; i--) v[i] = (v[i] - up[i] * v[i + 1] - up2[i] * v[i + 2]) / d[i];
    double nrm = 0.0;
    for (int i = 0; i < k; i++) nrm += v[i] * v[i];
    nrm = 1.0 / std::sqrt(nrm);
    for (int i = 0; i < k; i++) v[i] *= nrm;
  }
}

}  // namespace dmrgx
