// The fitted target's evaluator and the target cache's fill, restated for the
// tests (test infrastructure, never the product):
//   _solid_harmonic_fold    src/fit.jl:214-275  the real solid harmonics
//                                                S_lm(xi) = rho^l y~_lm(xi/rho)
//   _radial_sum             src/fit.jl:699-711  sum_k C[slot, k, v] rho^2k
//   fit_variables_at        src/fit.jl:721-733  the twenty fit variables at x, t
//   state_from_fit          src/fit.jl:182-212  (log a, b^i, g_ij - d_ij, Pi) -> (h, Pi)
//   fit_state               src/fit.jl:749-755  + bounds_project into the fit's bounds
//   fit_target_kernel!      src/fit.jl:1070-1116 A and S of the latest fit (and the
//   fill_target!            src/fit.jl:1131-1151 previous one) inside r_fill
// Plain scalar arithmetic in the reference's evaluation order, compiled
// without contraction (oracle/Makefile), so the device copy, which turns
// contraction off too, differs from it only where the two exp's round apart.
#pragma once
#include <cmath>
#include <vector>

#include "gh_oracle_bounds.hpp"

namespace orc {

// FitParams (src/fit.jl:492-502) with the coefficient array of InteriorFit
// (src/fit.jl:504-539) beside it: coeffs has shape ((L+1)^2, cont+1, 20),
// column-major, slot fastest; t is the time the fit was built at.
struct FitParams {
  int lmax = 0, cont = 0;
  bool tilde = false;
  double rbar = 1.0;
  double c0[3] = {0, 0, 0}, v[3] = {0, 0, 0};  // HoleCenter: c(t) = c0 + v t
  StateBounds bounds;                           // on: the projection (always, no gate)
  const double* coeffs = nullptr;
  double t = 0.0;
};

// One (l, m) of the fold (src/fit.jl:214-222): the m = 0 slot, or the cosine
// and the sine slot of |m| (0-based slots l^2 + l +- m), each adding
// s * _radial_sum to the twenty accumulators.
inline void fit_fold_slot(const FitParams& p, int slot, double s, double P2, double P4, double acc[20]) {
  const int NS = (p.lmax + 1) * (p.lmax + 1), nb = p.cont + 1;
  for (int v = 0; v < 20; ++v) {
    const double* C = p.coeffs + slot + (long)NS * nb * v;
    double a = C[0];
    if (p.cont >= 1) a += P2 * C[NS];
    if (p.cont >= 2) a += P4 * C[2 * NS];
    acc[v] = acc[v] + s * a;
  }
}
inline void fit_fold_term(const FitParams& p, int l, int m, double q, double C, double S, double s2, double P2,
                          double P4, double acc[20]) {
  if (m == 0) {
    fit_fold_slot(p, l * l + l, q, P2, P4, acc);
  } else {
    fit_fold_slot(p, l * l + l + m, s2 * q * C, P2, P4, acc);
    fit_fold_slot(p, l * l + l - m, -(s2 * q * S), P2, P4, acc);
  }
}

// fit_variables_at (src/fit.jl:721-733) with _solid_harmonic_fold's
// recurrence (src/fit.jl:243-275): Q_mm by the sectoral step, Q_{m+1,m} =
// sqrt(2m+3) z Q_mm, Q_lm = alpha (z Q_{l-1,m} - beta rho^2 Q_{l-2,m}),
// (x + i y)^m for the azimuth.
inline void fit_variables_at(const FitParams& p, const double x[3], double t, double out[20]) {
  double c[3];
  for (int d = 0; d < 3; ++d) c[d] = p.c0[d] + t * p.v[d];
  const double ir = 1.0 / p.rbar;
  const double xi1 = (x[0] - c[0]) * ir, xi2 = (x[1] - c[1]) * ir, xi3 = (x[2] - c[2]) * ir;
  const double r2 = xi1 * xi1 + xi2 * xi2 + xi3 * xi3;
  const double P4 = r2 * r2;
  for (int v = 0; v < 20; ++v) out[v] = 0.0;
  const double s2 = std::sqrt(2.0);
  double qmm = 1.0 / std::sqrt(4.0 * M_PI);
  double C = 1.0, S = 0.0;
  for (int m = 0; m <= p.lmax; ++m) {
    if (m > 0) {
      qmm = -std::sqrt((double)(2 * m + 1) / (double)(2 * m)) * qmm;
      const double Cn = C * xi1 - S * xi2, Sn = C * xi2 + S * xi1;
      C = Cn;
      S = Sn;
    }
    fit_fold_term(p, m, m, qmm, C, S, s2, r2, P4, out);
    if (m + 1 <= p.lmax) {
      double qa = qmm;
      double qb = std::sqrt((double)(2 * m + 3)) * xi3 * qmm;
      fit_fold_term(p, m + 1, m, qb, C, S, s2, r2, P4, out);
      for (int l = m + 2; l <= p.lmax; ++l) {
        const double al = std::sqrt((double)(4 * l * l - 1) / (double)(l * l - m * m));
        const double be = std::sqrt((double)((l - 1) * (l - 1) - m * m) / (double)(4 * (l - 1) * (l - 1) - 1));
        const double qc = al * (xi3 * qb - be * r2 * qa);
        fit_fold_term(p, l, m, qc, C, S, s2, r2, P4, out);
        qa = qb;
        qb = qc;
      }
    }
  }
}

// state_from_fit (src/fit.jl:182-212): alpha^2 = exp(2 log alpha), gamma_ij =
// delta_ij + v_ij, beta_i = gamma_ij beta^j, h_tt = (1 - alpha^2) + beta_i beta^i,
// h_ti = beta_i, h_ij = v_ij; Pi as it is, or sqrt(det gamma / alpha^2) Pi~.
inline void state_from_fit(const double v[20], bool tilde, double h[10], double P[10]) {
  const double o = 1.0;
  const double a2 = std::exp(2.0 * v[0]);
  const double be1 = v[1], be2 = v[2], be3 = v[3];
  const double g11 = o + v[4], g12 = v[5], g13 = v[6], g22 = o + v[7], g23 = v[8], g33 = o + v[9];
  const double b1 = g11 * be1 + g12 * be2 + g13 * be3;
  const double b2 = g12 * be1 + g22 * be2 + g23 * be3;
  const double b3 = g13 * be1 + g23 * be2 + g33 * be3;
  const double bb = b1 * be1 + b2 * be2 + b3 * be3;
  h[0] = (o - a2) + bb;
  h[1] = b1;
  h[2] = b2;
  h[3] = b3;
  for (int n = 4; n < 10; ++n) h[n] = v[n];
  const double detg =
      g11 * (g22 * g33 - g23 * g23) - g12 * (g12 * g33 - g23 * g13) + g13 * (g12 * g23 - g22 * g13);
  const double f = !tilde ? o : detg > 0.0 ? std::sqrt(detg / a2) : NAN;
  for (int n = 0; n < 10; ++n) P[n] = f * v[10 + n];
}

// fit_state (src/fit.jl:749-755): the reassembled state projected into the
// fit's bounds (everywhere: the fit's projection has no gate).
inline void fit_state(const FitParams& p, const double x[3], double t, double h[10], double P[10]) {
  double v[20], h0[10], P0[10];
  fit_variables_at(p, x, t, v);
  state_from_fit(v, p.tilde, h0, P0);
  StateBounds bd = p.bounds;
  bd.on = true;
  bool nf;
  bounds_project(h0, P0, bd, h, P, nf);
}

// fill_target! + fit_target_kernel! (src/fit.jl:1070-1151) over a state
// layout of P^3 points per block: at every owned point with |x - c(t)| <
// r_fill (c the interior's center at t) the 40 slots A_h, A_Pi, S_h, S_Pi of
// the latest fit fa and, when fb is given and older, the previous one;
// zero elsewhere when clear. pos(b, i, j, k, x) gives a point's position.
template <class Pos>
void fill_target(const FitParams& fa, const FitParams* fb, double t, double r_fill, double h_int,
                 const double icenter[3], const double ivel[3], bool rate, bool clear, long nblocks, int P,
                 Pos pos, double* out) {
  const bool hasb = fb && fa.t > fb->t;
  const double kappa = hasb ? 1.0 / (fa.t - fb->t) : 0.0;
  const double speed = std::sqrt(fa.v[0] * fa.v[0] + fa.v[1] * fa.v[1] + fa.v[2] * fa.v[2]);
  const double dt = speed > 0.0 ? h_int / (8.0 * speed) : 1.0;
  const double t_a = fa.t, t_f = t;
  const long P3 = (long)P * P * P;
  for (long b = 0; b < nblocks; ++b)
    for (int k = 0; k < P; ++k)
      for (int j = 0; j < P; ++j)
        for (int i = 0; i < P; ++i) {
          const long pt = i + (long)P * (j + (long)P * k);
          double* o = out + b * 40 * P3 + pt;
          double x[3];
          pos(b, i, j, k, x);
          const double c[3] = {icenter[0] + t_f * ivel[0], icenter[1] + t_f * ivel[1], icenter[2] + t_f * ivel[2]};
          const double d1 = x[0] - c[0], d2 = x[1] - c[1], d3 = x[2] - c[2];
          const double r = std::sqrt(d1 * d1 + d2 * d2 + d3 * d3);
          if (r < r_fill) {
            double ha[10], Pa[10], hp[10], Pp[10], hm[10], Pm[10];
            fit_state(fa, x, t_f, ha, Pa);
            if (rate) {
              fit_state(fa, x, t_f + dt, hp, Pp);
              fit_state(fa, x, t_f - dt, hm, Pm);
            }
            const double id = rate ? 1.0 / (2.0 * dt) : 0.0;
            if (hasb) {
              double hb[10], Pb[10];
              fit_state(*fb, x, t_f, hb, Pb);
              for (int v = 0; v < 10; ++v) {
                const double sh = kappa * (ha[v] - hb[v]);
                const double sP = kappa * (Pa[v] - Pb[v]);
                o[v * P3] = ha[v] + (t_f - t_a) * sh;
                o[(10 + v) * P3] = Pa[v] + (t_f - t_a) * sP;
                o[(20 + v) * P3] = rate ? sh + id * (hp[v] - hm[v]) : sh;
                o[(30 + v) * P3] = rate ? sP + id * (Pp[v] - Pm[v]) : sP;
              }
            } else {
              for (int v = 0; v < 10; ++v) {
                o[v * P3] = ha[v];
                o[(10 + v) * P3] = Pa[v];
                o[(20 + v) * P3] = rate ? id * (hp[v] - hm[v]) : 0.0;
                o[(30 + v) * P3] = rate ? id * (Pp[v] - Pm[v]) : 0.0;
              }
            }
          } else if (clear) {
            for (int v = 0; v < 40; ++v) o[v * P3] = 0.0;
          }
        }
}

}  // namespace orc
