// The range projection on the device (bounds_project, src/bounds.jl:260-589):
// a pointwise map holding a point's state inside ranges stated in ADM
// variables -- the spectrum of gamma_ij, the shift's norm |beta|, the lapse alpha
// and the scale max |(alpha/sqrt(gamma)) Pi_ab| -- written back only where it
// fired. Healthy quantities keep their bits, every test carries a slack of
// 8 eps of the terms it is made of, and a fired result is tested again by the
// same arithmetic until it passes, so the map is idempotent on the state and
// on its flag. The arithmetic is plain multiplies and adds evaluated left to
// right with floating-point contraction off, so that every call site -- the
// first pass, the re-test, the next stage's call -- computes the same bits.
#pragma once
#include "tgh_device.hpp"

namespace tgh {

struct Bounds {
  double a_min, a_max, l_min, l_max, b_max, K_max, r_gate;  // r_gate = 0: no projection
};

constexpr double kDblEps = 2.220446049250313e-16;

struct Adm {
  double g[6];  // gamma xx, xy, xz, yy, yz, zz
  double b[3];  // the lowered shift beta_i = h_ti
  double detg, bb, a2;
};

// The 3+1 split of the packed offset (src/bounds.jl:269-300): the adjugate
// inverse of gamma, beta_i beta^i, alpha^2 = (1 - h_tt) + beta_i beta^i.
TGH_HD Adm adm_split(const double h[NC]) {
#pragma clang fp contract(off)
  Adm s;
  const double g11 = 1.0 + h[4], g12 = h[5], g13 = h[6], g22 = 1.0 + h[7], g23 = h[8], g33 = 1.0 + h[9];
  const double b1 = h[1], b2 = h[2], b3 = h[3];
  const double c11 = g22 * g33 - g23 * g23;
  const double c12 = g13 * g23 - g12 * g33;
  const double c13 = g12 * g23 - g13 * g22;
  const double c22 = g11 * g33 - g13 * g13;
  const double c23 = g12 * g13 - g11 * g23;
  const double c33 = g11 * g22 - g12 * g12;
  s.detg = g11 * c11 + g12 * c12 + g13 * c13;
  const double qf = b1 * (c11 * b1 + c12 * b2 + c13 * b3) + b2 * (c12 * b1 + c22 * b2 + c23 * b3) +
                    b3 * (c13 * b1 + c23 * b2 + c33 * b3);
  s.bb = qf / s.detg;
  s.a2 = (1.0 - h[0]) + s.bb;
  s.g[0] = g11;
  s.g[1] = g12;
  s.g[2] = g13;
  s.g[3] = g22;
  s.g[4] = g23;
  s.g[5] = g33;
  s.b[0] = b1;
  s.b[1] = b2;
  s.b[2] = b3;
  return s;
}

// Sylvester's criterion with each leading minor above 16 eps of the sum of
// its terms' magnitudes: true is certain positive definiteness.
TGH_HD bool surely_posdef(double a11, double a12, double a13, double a22, double a23, double a33) {
#pragma clang fp contract(off)
  const double tol = 16 * kDblEps;
  const double m2 = a11 * a22 - a12 * a12;
  const double e2 = fabs(a11 * a22) + a12 * a12;
  const double c1 = a22 * a33 - a23 * a23;
  const double c2 = a12 * a33 - a23 * a13;
  const double c3 = a12 * a23 - a22 * a13;
  const double m3 = a11 * c1 - a12 * c2 + a13 * c3;
  const double e3 = fabs(a11) * (fabs(a22 * a33) + a23 * a23) + fabs(a12) * (fabs(a12 * a33) + fabs(a23 * a13)) +
                    fabs(a13) * (fabs(a12 * a23) + fabs(a22 * a13));
  return (a11 > 0.0) & (m2 > tol * e2) & (m3 > tol * e3);
}

// The Jacobi rotation that annihilates A[p][q] (Golub & Van Loan's sym.schur2
// with the accurate diagonal update), accumulated into the columns of V.
TGH_HD void jacobi_rotate(double A[3][3], double V[3][3], int p, int q) {
#pragma clang fp contract(off)
  const double apq = A[p][q];
  if (apq == 0.0) return;
  const double app = A[p][p], aqq = A[q][q];
  const double tau = (aqq - app) / (2 * apq);
  const double t = fabs(tau) > 1.0 / sqrt(kDblEps) ? 1.0 / (2 * tau)
                                                   : copysign(1.0, tau) / (fabs(tau) + sqrt(1.0 + tau * tau));
  const double c = 1.0 / sqrt(1.0 + t * t);
  const double s = t * c;
  const int r = 3 - p - q;
  const double arp = A[r][p], arq = A[r][q];
  A[p][p] = app - t * apq;
  A[q][q] = aqq + t * apq;
  A[p][q] = 0.0;
  A[q][p] = 0.0;
  const double nrp = c * arp - s * arq, nrq = s * arp + c * arq;
  A[r][p] = nrp;
  A[p][r] = nrp;
  A[r][q] = nrq;
  A[q][r] = nrq;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const double vp = V[i][p], vq = V[i][q];
    V[i][p] = c * vp - s * vq;
    V[i][q] = s * vp + c * vq;
  }
}

// Eigenvalues (diagonal of A on return) and eigenvectors (V's columns) of a
// symmetric 3x3 by cyclic Jacobi: a few eps of |A| at any spectrum, double
// eigenvalues included.
TGH_HD void sym_eigen3(double A[3][3], double V[3][3]) {
#pragma clang fp contract(off)
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) V[i][j] = i == j ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 16; ++sweep) {
    const double off = A[0][1] * A[0][1] + A[0][2] * A[0][2] + A[1][2] * A[1][2];
    const double on = A[0][0] * A[0][0] + A[1][1] * A[1][1] + A[2][2] * A[2][2];
    if (!(off > kDblEps * kDblEps * on)) break;
    jacobi_rotate(A, V, 0, 1);
    jacobi_rotate(A, V, 0, 2);
    jacobi_rotate(A, V, 1, 2);
  }
}

// max with NaN propagation (a NaN argument is the result)
TGH_HD double max_nan(double x, double y) { return (x != x || y != y) ? x + y : (x > y ? x : y); }

// gamma's spectrum clamped into [l_min, l_max]: gamma as it came unless an
// eigenvalue is outside by more than 8 eps of the spectrum's scale, then the
// whole gamma recomposed from the eigenvectors and the clamped eigenvalues.
TGH_HD bool spectrum_clamp(const double g[6], double l_min, double l_max, double out[6]) {
#pragma clang fp contract(off)
#pragma unroll
  for (int n = 0; n < 6; ++n) out[n] = g[n];
  const double a11 = g[0], a12 = g[1], a13 = g[2], a22 = g[3], a23 = g[4], a33 = g[5];
  const bool lo = surely_posdef(a11 - l_min, a12, a13, a22 - l_min, a23, a33 - l_min);
  const bool hi = surely_posdef(l_max - a11, -a12, -a13, l_max - a22, -a23, l_max - a33);
  if (lo & hi) return false;
  double A[3][3] = {{a11, a12, a13}, {a12, a22, a23}, {a13, a23, a33}}, V[3][3];
  sym_eigen3(A, V);
  const double lam[3] = {A[0][0], A[1][1], A[2][2]};
  const double S = max_nan(max_nan(fabs(lam[0]), fabs(lam[1])), fabs(lam[2]));
  const double s = 8 * kDblEps * S;
  bool inside = true;
#pragma unroll
  for (int k = 0; k < 3; ++k) inside = inside & ((l_min - s <= lam[k]) & (lam[k] <= l_max + s));
  if (inside) return false;
  double lc[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) lc[k] = lam[k] > l_max ? l_max : (lam[k] < l_min ? l_min : lam[k]);
  const int I[6] = {0, 0, 0, 1, 1, 2}, J[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
  for (int n = 0; n < 6; ++n)
    out[n] = V[I[n]][0] * lc[0] * V[J[n]][0] + V[I[n]][1] * lc[1] * V[J[n]][1] + V[I[n]][2] * lc[2] * V[J[n]][2];
  return true;
}

// One pass of the five steps: (1) non-finite components to Minkowski's 0;
// (2) the spectrum of gamma; (3) the shift's norm with the projected gamma;
// (4) the lapse by moving h_tt alone; (5) Pi rescaled to keep (alpha/sqrt(gamma)) Pi
// where the lapse was raised from a positive value, then its scale capped.
TGH_HD bool project_once(const double h[NC], const double P[NC], const Bounds& bd, double ho[NC], double Po[NC],
                         bool& nonfinite) {
#pragma clang fp contract(off)
  const double slack = 8 * kDblEps;
  bool fin = true;
#pragma unroll
  for (int n = 0; n < NC; ++n) fin = fin & isfinite(h[n]) & isfinite(P[n]);
  nonfinite = !fin;
  double h1[NC], P1[NC];
#pragma unroll
  for (int n = 0; n < NC; ++n) {
    h1[n] = isfinite(h[n]) ? h[n] : 0.0;
    P1[n] = isfinite(P[n]) ? P[n] : 0.0;
  }
  const Adm s0 = adm_split(h1);
  double gc[6];
  const bool gfire = spectrum_clamp(s0.g, bd.l_min, bd.l_max, gc);
  double hp[NC];
#pragma unroll
  for (int n = 0; n < NC; ++n) hp[n] = h1[n];
  if (gfire) {
    hp[4] = gc[0] - 1.0;
    hp[5] = gc[1];
    hp[6] = gc[2];
    hp[7] = gc[3] - 1.0;
    hp[8] = gc[4];
    hp[9] = gc[5] - 1.0;
  }
  const Adm s1 = gfire ? adm_split(hp) : s0;
  const double bmax2 = bd.b_max * bd.b_max;
  const bool bfire = !(s1.bb <= bmax2 + slack * bmax2);
  if (bfire) {
    const double f = bd.b_max / sqrt(s1.bb);
    hp[1] = f * s1.b[0];
    hp[2] = f * s1.b[1];
    hp[3] = f * s1.b[2];
  }
  const Adm s2 = bfire ? adm_split(hp) : s1;
  const double lo2 = bd.a_min * bd.a_min, hi2 = bd.a_max * bd.a_max;
  const double sa = slack * (1.0 + fabs(hp[0]) + fabs(s2.bb));
  const bool alow = !(s2.a2 >= lo2 - sa);
  const bool ahigh = !(s2.a2 <= hi2 + sa);
  const bool afire = alow | ahigh;
  if (afire) hp[0] = (1.0 - (alow ? lo2 : hi2)) + s2.bb;
  const Adm s3 = afire ? adm_split(hp) : s2;
  const bool hfire = gfire | bfire | afire;
  const bool raise = alow & (s0.a2 > 0.0) & (s0.detg > 0.0) & (s3.a2 > s0.a2);
  double Pr[NC];
  const double fr = raise ? (sqrt(s0.a2) / sqrt(s3.a2)) * (sqrt(s3.detg) / sqrt(s0.detg)) : 1.0;
#pragma unroll
  for (int n = 0; n < NC; ++n) Pr[n] = raise ? P1[n] * fr : P1[n];
  const double a = sqrt(s3.a2) / sqrt(s3.detg);
  double m = fabs(a * Pr[0]);
#pragma unroll
  for (int n = 1; n < NC; ++n) m = max_nan(m, fabs(a * Pr[n]));
  const bool Kfire = !(m <= bd.K_max + slack * bd.K_max);
  const double fk = bd.K_max / m;
#pragma unroll
  for (int n = 0; n < NC; ++n) {
    Po[n] = Kfire ? Pr[n] * fk : Pr[n];
    ho[n] = hfire ? hp[n] : h1[n];
  }
  return nonfinite | hfire | raise | Kfire;
}

// bounds_project: one pass, then, where it fired, the result tested again by
// the next call's own arithmetic and projected until it passes (4 at most).
TGH_HD bool bounds_project(double h[NC], double P[NC], const Bounds& bd, bool& nonfinite) {
  double ho[NC], Po[NC];
  const bool hit = project_once(h, P, bd, ho, Po, nonfinite);
  if (!hit) return false;
#pragma unroll 1
  for (int pass = 0; pass < 4; ++pass) {
    double h2[NC], P2[NC];
    bool nf2;
    if (!project_once(ho, Po, bd, h2, P2, nf2)) break;
#pragma unroll
    for (int n = 0; n < NC; ++n) {
      ho[n] = h2[n];
      Po[n] = P2[n];
    }
  }
#pragma unroll
  for (int n = 0; n < NC; ++n) {
    h[n] = ho[n];
    P[n] = Po[n];
  }
  return true;
}

}  // namespace tgh
