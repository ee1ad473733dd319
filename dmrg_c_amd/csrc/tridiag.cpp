// Lowest eigenpair of a symmetric tridiagonal matrix, the part of dstev_ the Lanczos
// drivers use (sml/LANCZOS_SLOW_SYM.c:127-137 take Lap_D[0] and the first eigenvector).
//
// Eigenvalue: Newton's method on det(T - x I) from below, bracketed by the Sturm count of
// T - x I (number of negative pivots of its LDL^T factorisation) and safeguarded by
// bisection, run until it stops moving (full double precision).  Eigenvector: inverse iteration with the
// tridiagonal LU factorisation of T - lam I using partial pivoting (row interchanges as
// in LAPACK's dgttrf), three sweeps from a constant vector.
#include <algorithm>
#include <cmath>
#include <limits>
#include <vector>

#include "internal.h"

namespace dmrgx {


namespace {
// Sturm count and Newton ratio of T - x I.  Returns the number of pivots d_i of its LDL^T
// factorisation that are not positive (an exact zero pivot counts and is replaced by a tiny
// negative one to go on): zero means x is below every eigenvalue.  Then *ratio = p'(x)/p(x)
// for p(x) = det(T - x I).
// Computed from the leading principal minors p_i = (a_i - x) p_{i-1} - b_{i-1}^2 p_{i-2}
// and their x-derivatives (d_i = p_i / p_{i-1}, so d_i <= 0 when p_i is zero or differs in
// sign from p_{i-1}): one multiply-add per step on the dependent chain instead of a
// division.  The minors are rescaled by powers of two (exact) to stay in range.
int pivots(int k, const double *a, const double *b, double x, double *ratio) {
  double pm = 1.0, qm = 0.0;        // p_{i-1}, p'_{i-1}   (p_{-1} = 1)
  double p = a[0] - x, q = -1.0;    // p_0, p'_0
  int neg = 0;
  for (int i = 0;;) {
    if (p == 0.0) {
      p = -1e-300 * pm;
      neg++;
    } else if (std::signbit(p) != std::signbit(pm)) {
      neg++;
    }
    if (++i == k) break;
    const double c = a[i] - x, b2 = b[i - 1] * b[i - 1];
    const double pn = c * p - b2 * pm;
    const double qn = c * q - b2 * qm - p;
    pm = p;
    qm = q;
    p = pn;
    q = qn;
    const double mag = std::fabs(p);
    if (mag > 0x1p+400 || (mag < 0x1p-400 && mag > 0.0)) {
      const int e = std::ilogb(mag);
      p = std::scalbn(p, -e);
      q = std::scalbn(q, -e);
      pm = std::scalbn(pm, -e);
      qm = std::scalbn(qm, -e);
    }
  }
  *ratio = q / p;
  return neg;
}
}  // namespace

// Lowest eigenvalue by Newton's method on det(T - x I) from below, safeguarded by
// bisection.  Left of the lowest eigenvalue every pivot is positive and the Newton step
// x - p/p' stays left of it and approaches it monotonically (quadratically once close, the
// lowest eigenvalue being simple); a bisection midpoint that tests negative-free becomes the
// new Newton point, one that does not becomes the new upper end.  upper_hint (an upper
// bound, e.g. the previous Lanczos step's value) and lower_guess (a guess below the value)
// shorten the search; pass +inf / NaN without them.
double tridiag_lowest_value(int k, const double *a, const double *b, double upper_hint, double lower_guess) {
  if (k == 1) return a[0];
  double glo = a[0], ghi = a[0];
  for (int i = 0; i < k; i++) {
    const double rad = (i > 0 ? std::fabs(b[i - 1]) : 0.0) + (i + 1 < k ? std::fabs(b[i]) : 0.0);
    glo = std::min(glo, a[i] - rad);
    ghi = std::max(ghi, a[i] + rad);
  }
  const double pad = 1e-13 * (ghi - glo) + 1e-300;
  glo -= pad;
  ghi += pad;
  double hi = ghi, lo = glo, r = 0.0;
  bool have_lo = false;
  if (std::isfinite(upper_hint) && upper_hint > glo && upper_hint < ghi) {
    if (pivots(k, a, b, upper_hint, &r) == 0) {
      // the hint is below the value (equal to it to rounding): it is a lower end
      lo = upper_hint;
      have_lo = true;
    } else {
      hi = upper_hint;
    }
  }
  // lower end: the guess if it is below the value, else walk down from hi in growing steps
  if (!have_lo && std::isfinite(lower_guess) && lower_guess > glo && lower_guess < hi &&
      pivots(k, a, b, lower_guess, &r) == 0) {
    lo = lower_guess;
    have_lo = true;
  }
  if (!have_lo) {
    double step = std::max(1e-12 * std::max(std::fabs(hi), 1.0), (hi - glo) * 1e-6);
    while (true) {
      const double x = hi - step;
      if (!(x > glo)) {
        lo = glo;
        pivots(k, a, b, lo, &r);
        break;
      }
      if (pivots(k, a, b, x, &r) == 0) {
        lo = x;
        break;
      }
      hi = x;
      step *= 16.0;
    }
  }
  // r = p'/p at lo (< 0); Newton moves lo up, bisection shrinks hi when Newton is slow
  double prev_gain = std::numeric_limits<double>::infinity();
  for (int it = 0; it < 300; it++) {
    double x = lo - 1.0 / r;
    if (!(x > lo)) break;                      // no progress: lo is the eigenvalue to rounding
    if (!(x < hi)) x = lo + 0.5 * (hi - lo);   // left the bracket (rounding): bisect
    if (!(x > lo && x < hi)) break;            // bracket at rounding width
    double rx;
    if (pivots(k, a, b, x, &rx) == 0) {
      const double gain = x - lo, width = hi - lo;
      lo = x;
      r = rx;
      const bool slow = gain > 0.25 * prev_gain && gain < 0.125 * width;
      prev_gain = gain;
      if (slow) {
        // linear, not quadratic, progress (far left, or a close second eigenvalue): also
        // test the midpoint
        const double m = lo + 0.5 * (hi - lo);
        double rm;
        if (m > lo && m < hi) {
          if (pivots(k, a, b, m, &rm) == 0) {
            lo = m;
            r = rm;
          } else {
            hi = m;
          }
        }
      }
    } else {
      hi = x;
    }
  }
  return lo;
}

double RitzTrack::next(int k, const double *a, const double *b) {
  double guess = std::numeric_limits<double>::quiet_NaN();
  if (std::isfinite(last) && std::isfinite(drop))
    guess = last - 2.0 * std::max(drop, 0.0) - 1e-12 * std::max(1.0, std::fabs(last));
  const double lam = tridiag_lowest_value(k, a, b, last, guess);
  if (std::isfinite(last)) drop = last - lam;
  last = lam;
  return lam;
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
