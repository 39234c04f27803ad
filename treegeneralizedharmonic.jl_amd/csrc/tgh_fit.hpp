// The fitted target's evaluator and the target cache's fill on the device
// (fit_target_kernel!, src/fit.jl:1070-1116; fill_target!, :1131-1151):
// at the refill cadence, one thread per owned point evaluates the latest
// fit -- and the previous one, and the latest at t +- dt for its translation
// rate -- and writes the 40 slots A_h, A_Pi, S_h, S_Pi that the :fitted
// kernels read as A + (t - t_f) S.
//
// The fit (InteriorFit, src/fit.jl:504-539) is a sum over the real solid
// harmonics S_lm(xi) = |xi|^l y~_lm(xi/|xi|) of radial polynomials in
// |xi|^2, xi = (x - c(t)) / rbar, for twenty variables (log alpha, beta^i,
// gamma_ij - delta_ij, Pi_ab or (alpha/sqrt(det gamma)) Pi_ab). Its
// coefficients are read at the same address by every lane of a wave, so they
// come through the scalar cache; the harmonics' recurrence runs in registers.
// Contraction is off so that the evaluation rounds as the reference's
// left-to-right scalar arithmetic does.
#pragma once
#include "tgh_bounds.hpp"

namespace tgh {

struct FitDev {
  int lmax, cont, tilde;
  double rbar, c0[3], v[3];
  Bounds bd;                     // the fit's bounds (the projection has no gate)
  const double* __restrict__ coeffs;  // ((L+1)^2, cont+1, 20), slot fastest
};

// acc_v += s * sum_k C[slot, k, v] rho^2k (_radial_sum, src/fit.jl:699-711).
TGH_HD void fit_slot(const FitDev& f, int slot, double s, double P2, double P4, double acc[20]) {
#pragma clang fp contract(off)
  const int NS = (f.lmax + 1) * (f.lmax + 1);
  const long vs = (long)NS * (f.cont + 1);
#pragma unroll
  for (int v = 0; v < 20; ++v) {
    const double* C = f.coeffs + slot + vs * v;
    double a = C[0];
    if (f.cont >= 1) a += P2 * C[NS];
    if (f.cont >= 2) a += P4 * C[2 * NS];
    acc[v] = acc[v] + s * a;
  }
}
// One (l, m) of _solid_harmonic_fold (src/fit.jl:214-222): the m = 0 slot,
// or the cosine and the sine slot of |m|, 0-based at l^2 + l +- m.
TGH_HD void fit_term(const FitDev& f, int l, int m, double q, double C, double S, double s2, double P2, double P4,
                     double acc[20]) {
#pragma clang fp contract(off)
  if (m == 0) {
    fit_slot(f, l * l + l, q, P2, P4, acc);
  } else {
    fit_slot(f, l * l + l + m, s2 * q * C, P2, P4, acc);
    fit_slot(f, l * l + l - m, -(s2 * q * S), P2, P4, acc);
  }
}
// fit_variables_at (src/fit.jl:721-733): the normalised Legendre recurrences
// in xi_z with rho^2 in the three-term step (src/fit.jl:243-275), so that
// every term is a polynomial in xi and the center needs no special case.
TGH_HD void fit_variables_at(const FitDev& f, const double x[3], double t, double out[20]) {
#pragma clang fp contract(off)
  const double c0 = f.c0[0] + t * f.v[0], c1 = f.c0[1] + t * f.v[1], c2 = f.c0[2] + t * f.v[2];
  const double ir = 1.0 / f.rbar;
  const double x1 = (x[0] - c0) * ir, x2 = (x[1] - c1) * ir, x3 = (x[2] - c2) * ir;
  const double r2 = x1 * x1 + x2 * x2 + x3 * x3;
  const double P4 = r2 * r2;
#pragma unroll
  for (int v = 0; v < 20; ++v) out[v] = 0.0;
  const double s2 = sqrt(2.0);
  double qmm = 1.0 / sqrt(4.0 * M_PI);
  double C = 1.0, S = 0.0;
#pragma unroll 1
  for (int m = 0; m <= f.lmax; ++m) {
    if (m > 0) {
      qmm = -sqrt((double)(2 * m + 1) / (double)(2 * m)) * qmm;
      const double Cn = C * x1 - S * x2, Sn = C * x2 + S * x1;
      C = Cn;
      S = Sn;
    }
    fit_term(f, m, m, qmm, C, S, s2, r2, P4, out);
    if (m + 1 <= f.lmax) {
      double qa = qmm;
      double qb = sqrt((double)(2 * m + 3)) * x3 * qmm;
      fit_term(f, m + 1, m, qb, C, S, s2, r2, P4, out);
#pragma unroll 1
      for (int l = m + 2; l <= f.lmax; ++l) {
        const double al = sqrt((double)(4 * l * l - 1) / (double)(l * l - m * m));
        const double be = sqrt((double)((l - 1) * (l - 1) - m * m) / (double)(4 * (l - 1) * (l - 1) - 1));
        const double qc = al * (x3 * qb - be * r2 * qa);
        fit_term(f, l, m, qc, C, S, s2, r2, P4, out);
        qa = qb;
        qb = qc;
      }
    }
  }
}
// state_from_fit (src/fit.jl:182-212): alpha^2 = exp(2 log alpha), beta_i =
// gamma_ij beta^j, h_tt = (1 - alpha^2) + beta_i beta^i, h_ti = beta_i,
// h_ij = gamma_ij - delta_ij; Pi, or sqrt(det gamma / alpha^2) Pi~ (NaN for a
// fit that is not a metric, which the projection then repairs).
TGH_HD void state_from_fit(const double v[20], bool tilde, double h[NC], double P[NC]) {
#pragma clang fp contract(off)
  const double a2 = exp(2.0 * v[0]);
  const double be1 = v[1], be2 = v[2], be3 = v[3];
  const double g11 = 1.0 + v[4], g12 = v[5], g13 = v[6], g22 = 1.0 + v[7], g23 = v[8], g33 = 1.0 + v[9];
  const double b1 = g11 * be1 + g12 * be2 + g13 * be3;
  const double b2 = g12 * be1 + g22 * be2 + g23 * be3;
  const double b3 = g13 * be1 + g23 * be2 + g33 * be3;
  const double bb = b1 * be1 + b2 * be2 + b3 * be3;
  h[0] = (1.0 - a2) + bb;
  h[1] = b1;
  h[2] = b2;
  h[3] = b3;
#pragma unroll
  for (int n = 4; n < NC; ++n) h[n] = v[n];
  const double detg =
      g11 * (g22 * g33 - g23 * g23) - g12 * (g12 * g33 - g23 * g13) + g13 * (g12 * g23 - g22 * g13);
  const double fct = !tilde ? 1.0 : detg > 0.0 ? sqrt(detg / a2) : __builtin_nan("");
#pragma unroll
  for (int n = 0; n < NC; ++n) P[n] = fct * v[NC + n];
}
// fit_state (src/fit.jl:749-755): the reassembled state projected into the
// fit's bounds.
TGH_HD void fit_state(const FitDev& f, const double x[3], double t, double h[NC], double P[NC]) {
  double v[20];
  fit_variables_at(f, x, t, v);
  state_from_fit(v, f.tilde != 0, h, P);
  bool nf;
  bounds_project(h, P, f.bd, nf);
}

struct FitTargetArgs {
  Geom g;
  double ic[3], iv[3];  // the interior's center c(t) = ic + iv t
  FitDev fa, fb;
  int hasb, rate, clear;
  double t_a, kappa, t_f, r_fill, dt;
  double* __restrict__ out;  // (N+1)^3 x 40 x nblocks
};
// fit_target_kernel! (src/fit.jl:1070-1116): one thread per owned point.
__global__ __launch_bounds__(256) void k_fit_target(FitTargetArgs a) {
#pragma clang fp contract(off)
  const Geom& g = a.g;
  const int P = g.P;
  const long P3 = (long)P * P * P;
  int b, wx;
  if (!wg_map((int)((P3 + 255) / 256), g.nblocks, b, wx)) return;
  const long pt = (long)wx * blockDim.x + threadIdx.x;
  if (pt >= P3) return;
  const int i = (int)(pt % P), j = (int)((pt / P) % P), k = (int)(pt / ((long)P * P));
  double x[3];
  position(g, b, i, j, k, x);
  double* o = a.out + (long)b * 40 * P3 + pt;
  const double t_f = a.t_f;
  const double c0 = a.ic[0] + t_f * a.iv[0], c1 = a.ic[1] + t_f * a.iv[1], c2 = a.ic[2] + t_f * a.iv[2];
  const double d1 = x[0] - c0, d2 = x[1] - c1, d3 = x[2] - c2;
  const double r = sqrt(d1 * d1 + d2 * d2 + d3 * d3);
  if (!(r < a.r_fill)) {
    if (a.clear)
      for (int v = 0; v < 40; ++v) o[v * P3] = 0.0;
    return;
  }
  double ha[NC], Pa[NC];
  fit_state(a.fa, x, t_f, ha, Pa);
  // S = kappa (F_a - F_b) of the two fits; A = F_a + (t_f - t_a) S
  double sh[NC], sP[NC];
#pragma unroll
  for (int v = 0; v < NC; ++v) sh[v] = sP[v] = 0.0;
  if (a.hasb) {
    double hb[NC], Pb[NC];
    fit_state(a.fb, x, t_f, hb, Pb);
#pragma unroll
    for (int v = 0; v < NC; ++v) {
      sh[v] = a.kappa * (ha[v] - hb[v]);
      sP[v] = a.kappa * (Pa[v] - Pb[v]);
      o[v * P3] = ha[v] + (t_f - a.t_a) * sh[v];
      o[(NC + v) * P3] = Pa[v] + (t_f - a.t_a) * sP[v];
    }
  } else {
#pragma unroll
    for (int v = 0; v < NC; ++v) {
      o[v * P3] = ha[v];
      o[(NC + v) * P3] = Pa[v];
    }
  }
  // the cache's slope: S, plus the latest fit's translation at this point
  if (a.rate) {
    const double id = 1.0 / (2.0 * a.dt);
    double hp[NC], Pp[NC], hm[NC], Pm[NC];
    fit_state(a.fa, x, t_f + a.dt, hp, Pp);
    fit_state(a.fa, x, t_f - a.dt, hm, Pm);
#pragma unroll
    for (int v = 0; v < NC; ++v) {
      const double rh = id * (hp[v] - hm[v]), rP = id * (Pp[v] - Pm[v]);
      sh[v] = a.hasb ? sh[v] + rh : rh;
      sP[v] = a.hasb ? sP[v] + rP : rP;
    }
  }
#pragma unroll
  for (int v = 0; v < NC; ++v) {
    o[(2 * NC + v) * P3] = sh[v];
    o[(3 * NC + v) * P3] = sP[v];
  }
}

}  // namespace tgh
