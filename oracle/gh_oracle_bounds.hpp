// The range projection (src/bounds.jl:260-589): a pointwise map that holds a
// point's state inside ranges stated in ADM variables -- the spectrum of
// gamma_ij, the shift's norm, the lapse and the scale of (alpha/sqrt(gamma)) Pi --
// written back only where it fired, idempotent on the state and on its flag.
// Restated from the reference's description and code, in plain scalar
// arithmetic evaluated left to right (the reference's own evaluation order),
// with no fused multiply-add (this file is compiled without one; the device
// copy turns contraction off), so every call site computes the same bits.
#pragma once
#include <cfloat>
#include <cmath>

namespace orc {

// StateBounds (bounds.jl:81-136); on = false is `bounds = nothing`.
struct StateBounds {
  bool on = false;
  double a_min = 0, a_max = 0, l_min = 0, l_max = 0, b_max = 0, K_max = 0, r_gate = 0;
};

constexpr double kEps = DBL_EPSILON;
constexpr int kBoundsSlack = 8;   // BOUNDS_SLACK (bounds.jl:406)
constexpr int kJacobiSweeps = 16; // JACOBI_SWEEPS (bounds.jl:365)
constexpr int kBoundsPasses = 4;  // BOUNDS_PASSES (bounds.jl:530)

// _adm_split (bounds.jl:269-300) of the packed offset h (tt,tx,ty,tz,xx,xy,xz,yy,yz,zz):
// gamma_ij = delta_ij + h_ij, the lowered shift beta_i = h_ti, det gamma by the
// adjugate, beta_i beta^i and alpha^2 = (1 - h_tt) + beta_i beta^i.
struct AdmSplit {
  double g[6];  // gamma xx, xy, xz, yy, yz, zz
  double b[3];
  double detg, bb, a2;
};
inline AdmSplit adm_split(const double h[10]) {
  AdmSplit s;
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
  s.g[0] = g11; s.g[1] = g12; s.g[2] = g13; s.g[3] = g22; s.g[4] = g23; s.g[5] = g33;
  s.b[0] = b1; s.b[1] = b2; s.b[2] = b3;
  return s;
}

// _surely_posdef (bounds.jl:305-318): Sylvester's criterion, each leading minor
// above 16 eps of the magnitudes it is made of; true is certain, false is not.
inline bool surely_posdef(double a11, double a12, double a13, double a22, double a23, double a33) {
  const double tol = 16 * kEps;
  const double m2 = a11 * a22 - a12 * a12;
  const double e2 = std::fabs(a11 * a22) + a12 * a12;
  const double c1 = a22 * a33 - a23 * a23;
  const double c2 = a12 * a33 - a23 * a13;
  const double c3 = a12 * a23 - a22 * a13;
  const double m3 = a11 * c1 - a12 * c2 + a13 * c3;
  const double e3 = std::fabs(a11) * (std::fabs(a22 * a33) + a23 * a23) +
                    std::fabs(a12) * (std::fabs(a12 * a33) + std::fabs(a23 * a13)) +
                    std::fabs(a13) * (std::fabs(a12 * a23) + std::fabs(a22 * a13));
  return (a11 > 0.0) & (m2 > tol * e2) & (m3 > tol * e3);
}

// One Jacobi rotation annihilating A[p][q], accumulated into V's columns
// (_jacobi_rotation, bounds.jl:325-363; Golub & Van Loan sym.schur2 with the
// Numerical Recipes diagonal update).
inline void jacobi_rotation(double A[3][3], double V[3][3], int p, int q) {
  const double apq = A[p][q];
  if (apq == 0.0) return;
  const double app = A[p][p], aqq = A[q][q];
  const double tau = (aqq - app) / (2 * apq);
  const double t = std::fabs(tau) > 1.0 / std::sqrt(kEps) ? 1.0 / (2 * tau)
                                                           : std::copysign(1.0, tau) / (std::fabs(tau) + std::sqrt(1.0 + tau * tau));
  const double c = 1.0 / std::sqrt(1.0 + t * t);
  const double s = t * c;
  const int r = 3 - p - q;
  const double arp = A[r][p], arq = A[r][q];
  A[p][p] = app - t * apq;
  A[q][q] = aqq + t * apq;
  A[p][q] = A[q][p] = 0.0;
  A[r][p] = A[p][r] = c * arp - s * arq;
  A[r][q] = A[q][r] = s * arp + c * arq;
  for (int i = 0; i < 3; ++i) {
    const double vp = V[i][p], vq = V[i][q];
    V[i][p] = c * vp - s * vq;
    V[i][q] = s * vp + c * vq;
  }
}

// sym_eigen3 (bounds.jl:382-396): cyclic Jacobi until the off-diagonal is below
// eps^2 of the diagonal, at most 16 sweeps; eigenvalues lam, eigenvectors V's columns.
inline void sym_eigen3(double A[3][3], double lam[3], double V[3][3]) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) V[i][j] = i == j ? 1.0 : 0.0;
  for (int sweep = 0; sweep < kJacobiSweeps; ++sweep) {
    const double off = A[0][1] * A[0][1] + A[0][2] * A[0][2] + A[1][2] * A[1][2];
    const double on = A[0][0] * A[0][0] + A[1][1] * A[1][1] + A[2][2] * A[2][2];
    if (!(off > kEps * kEps * on)) break;
    jacobi_rotation(A, V, 0, 1);
    jacobi_rotation(A, V, 0, 2);
    jacobi_rotation(A, V, 1, 2);
  }
  lam[0] = A[0][0];
  lam[1] = A[1][1];
  lam[2] = A[2][2];
}

inline double recompose(const double V[3][3], const double lam[3], int i, int j) {
  return V[i][0] * lam[0] * V[j][0] + V[i][1] * lam[1] * V[j][1] + V[i][2] * lam[2] * V[j][2];
}

inline double clamp_(double x, double lo, double hi) { return x > hi ? hi : (x < lo ? lo : x); }
// Julia's max: a NaN in either argument is the result.
inline double max_nan(double x, double y) { return (x != x || y != y) ? x + y : (x > y ? x : y); }

// _spectrum_clamp (bounds.jl:412-432): gamma as it came unless an eigenvalue is
// outside [l_min, l_max] by more than 8 eps of the spectrum's scale; then the
// whole gamma recomposed from the eigenvectors and the clamped eigenvalues.
inline bool spectrum_clamp(const double g[6], double l_min, double l_max, double out[6]) {
  for (int n = 0; n < 6; ++n) out[n] = g[n];
  const double a11 = g[0], a12 = g[1], a13 = g[2], a22 = g[3], a23 = g[4], a33 = g[5];
  const bool lo = surely_posdef(a11 - l_min, a12, a13, a22 - l_min, a23, a33 - l_min);
  const bool hi = surely_posdef(l_max - a11, -a12, -a13, l_max - a22, -a23, l_max - a33);
  if (lo & hi) return false;
  double A[3][3] = {{a11, a12, a13}, {a12, a22, a23}, {a13, a23, a33}}, lam[3], V[3][3];
  sym_eigen3(A, lam, V);
  const double S = max_nan(max_nan(std::fabs(lam[0]), std::fabs(lam[1])), std::fabs(lam[2]));
  const double s = kBoundsSlack * kEps * S;
  bool inside = true;
  for (int k = 0; k < 3; ++k) inside = inside & ((l_min - s <= lam[k]) & (lam[k] <= l_max + s));
  if (inside) return false;
  double lc[3];
  for (int k = 0; k < 3; ++k) lc[k] = clamp_(lam[k], l_min, l_max);
  out[0] = recompose(V, lc, 0, 0);
  out[1] = recompose(V, lc, 0, 1);
  out[2] = recompose(V, lc, 0, 2);
  out[3] = recompose(V, lc, 1, 1);
  out[4] = recompose(V, lc, 1, 2);
  out[5] = recompose(V, lc, 2, 2);
  return true;
}

// One application of the five steps (_project_once, bounds.jl:532-588).
inline bool project_once(const double h[10], const double P[10], const StateBounds& bd, double ho[10],
                         double Po[10], bool& nonfinite) {
  const double slack = kBoundsSlack * kEps;
  // (1) non-finite components take Minkowski's value
  bool fin = true;
  for (int n = 0; n < 10; ++n) fin = fin & std::isfinite(h[n]) & std::isfinite(P[n]);
  nonfinite = !fin;
  double h1[10], P1[10];
  for (int n = 0; n < 10; ++n) {
    h1[n] = nonfinite && !std::isfinite(h[n]) ? 0.0 : h[n];
    P1[n] = nonfinite && !std::isfinite(P[n]) ? 0.0 : P[n];
  }
  const AdmSplit s0 = adm_split(h1);
  // (2) the spectrum of gamma
  double gc[6];
  const bool gfire = spectrum_clamp(s0.g, bd.l_min, bd.l_max, gc);
  double hg[10];
  for (int n = 0; n < 10; ++n) hg[n] = h1[n];
  if (gfire) {
    hg[4] = gc[0] - 1.0; hg[5] = gc[1]; hg[6] = gc[2];
    hg[7] = gc[3] - 1.0; hg[8] = gc[4]; hg[9] = gc[5] - 1.0;
  }
  const AdmSplit s1 = gfire ? adm_split(hg) : s0;
  // (3) the shift's norm, with the projected gamma (a NaN takes the clamping branch)
  const double bmax2 = bd.b_max * bd.b_max;
  const bool bfire = !(s1.bb <= bmax2 + slack * bmax2);
  const double f = bfire ? bd.b_max / std::sqrt(s1.bb) : 1.0;
  double hb[10];
  for (int n = 0; n < 10; ++n) hb[n] = hg[n];
  if (bfire) {
    hb[1] = f * s1.b[0]; hb[2] = f * s1.b[1]; hb[3] = f * s1.b[2];
  }
  const AdmSplit s2 = bfire ? adm_split(hb) : s1;
  // (4) the lapse, by moving g_tt alone; the slack is of the terms alpha^2 is made of
  const double lo2 = bd.a_min * bd.a_min, hi2 = bd.a_max * bd.a_max;
  const double sa = slack * (1.0 + std::fabs(hb[0]) + std::fabs(s2.bb));
  const bool alow = !(s2.a2 >= lo2 - sa);
  const bool ahigh = !(s2.a2 <= hi2 + sa);
  const bool afire = alow | ahigh;
  const double at2 = alow ? lo2 : hi2;
  double hp[10];
  for (int n = 0; n < 10; ++n) hp[n] = hb[n];
  if (afire) hp[0] = (1.0 - at2) + s2.bb;
  const AdmSplit s3 = afire ? adm_split(hp) : s2;
  const bool hfire = gfire | bfire | afire;
  // (5) the momentum: keep (alpha/sqrt(gamma)) Pi where the lapse was raised, then cap its scale
  const bool raise = alow & (s0.a2 > 0.0) & (s0.detg > 0.0) & (s3.a2 > s0.a2);
  double Pr[10];
  const double fr = raise ? (std::sqrt(s0.a2) / std::sqrt(s3.a2)) * (std::sqrt(s3.detg) / std::sqrt(s0.detg)) : 1.0;
  for (int n = 0; n < 10; ++n) Pr[n] = raise ? P1[n] * fr : P1[n];
  const double a = std::sqrt(s3.a2) / std::sqrt(s3.detg);
  double m = std::fabs(a * Pr[0]);
  for (int n = 1; n < 10; ++n) m = max_nan(m, std::fabs(a * Pr[n]));
  const bool Kfire = !(m <= bd.K_max + slack * bd.K_max);
  const double fk = bd.K_max / m;
  for (int n = 0; n < 10; ++n) Po[n] = Kfire ? Pr[n] * fk : Pr[n];
  for (int n = 0; n < 10; ++n) ho[n] = hfire ? hp[n] : h1[n];
  return nonfinite | hfire | raise | Kfire;
}

// bounds_project (bounds.jl:510-527): once, then, where it fired, again on its
// own result until the next call's test passes (at most 4 more times).
inline bool bounds_project(const double h[10], const double P[10], const StateBounds& bd, double ho[10],
                           double Po[10], bool& nonfinite) {
  const bool hit = project_once(h, P, bd, ho, Po, nonfinite);
  if (hit) {
    for (int pass = 0; pass < kBoundsPasses; ++pass) {
      double h2[10], P2[10];
      bool nf2;
      if (!project_once(ho, Po, bd, h2, P2, nf2)) break;
      for (int n = 0; n < 10; ++n) {
        ho[n] = h2[n];
        Po[n] = P2[n];
      }
    }
  }
  return hit;
}

}  // namespace orc
