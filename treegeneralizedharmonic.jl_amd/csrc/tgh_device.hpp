// Device-side pointwise algebra of the generalized-harmonic system, written for
// gfx950: straight-line FP64 code over packed symmetric tensors, every loop
// unrolled at compile time so that every index is a constant and every
// intermediate a register.
//
// Reference (eschnett/TreeGeneralizedHarmonic.jl): src/pointwise.jl (packing
// :83-112, metric_quantities :145-184, gh_node_source_lean :612-741,
// metric_divergences :754-778, matter_source :667-698), src/gauge.jl
// (KerrSchildSource :293-326, damping profiles :1094-1156), src/interior.jl
// (smoothstep :117-120, profiles :567-590, core_position :615-625),
// src/initialdata.jl (background_state :1674-1694, point_draw :2004-2013),
// and SpacetimeMetrics' closed-form backgrounds.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <type_traits>

#define TGH_HD __host__ __device__ __forceinline__

namespace tgh {

constexpr int NC = 10;

// Packed slot of the symmetric pair (a, b): order tt,tx,ty,tz,xx,xy,xz,yy,yz,zz.
TGH_HD constexpr int pidx(int a, int b) {
  return a >= b ? b * 4 - (b * (b - 1)) / 2 + (a - b) : a * 4 - (a * (a - 1)) / 2 + (b - a);
}
// Row and column of packed slot n (row >= column).
TGH_HD constexpr int prow(int n) { return n <= 3 ? n : n <= 6 ? n - 3 : n <= 8 ? n - 5 : 3; }
TGH_HD constexpr int pcol(int n) { return n <= 3 ? 0 : n <= 6 ? 1 : n <= 8 ? 2 : 3; }
// Packed slot of a spatial pair (i, j), i, j in 0..2: xx,xy,xz,yy,yz,zz.
TGH_HD constexpr int sidx3(int i, int j) {
  return i >= j ? j * 3 - (j * (j - 1)) / 2 + (i - j) : i * 3 - (i * (i - 1)) / 2 + (j - i);
}
TGH_HD constexpr double eta(int a, int b) { return a != b ? 0.0 : (a == 0 ? -1.0 : 1.0); }

// ---------------------------------------------------------------------------
// The coefficient set of a point, from the offset h = g - eta, by the 3+1
// split: gamma_ij = delta_ij + h_ij, beta_i = h_ti, beta^i = gamma^ij beta_j,
// alpha^2 = (1 - h_tt) + beta^i beta_i, and
//   g^tt = -1/alpha^2, g^ti = beta^i/alpha^2, g^ij = gamma^ij - beta^i beta^j/alpha^2.
// Every quantity is formed from the offsets (det(gamma) - 1 from the
// elementary symmetric polynomials of h_ij), so alpha, beta and sqrt(gamma)
// keep full relative precision for small h, as metric_quantities
// (src/pointwise.jl:145-184) does by its 4x4 offset identities.
struct Coef {
  double gu[NC];   // packed g^{ab}
  double alpha, sqrtg;
  double beta[3];  // beta^i
  double gamu[6];  // gamma^{ij}, packed xx,xy,xz,yy,yz,zz
};

TGH_HD void metric_quantities(const double h[NC], Coef& c) {
  const double hxx = h[4], hxy = h[5], hxz = h[6], hyy = h[7], hyz = h[8], hzz = h[9];
  // det(I + H) - 1 = e1 + e2 + e3 of the spatial offset H.
  const double e1 = hxx + hyy + hzz;
  const double e2 = (hxx * hyy - hxy * hxy) + (hxx * hzz - hxz * hxz) + (hyy * hzz - hyz * hyz);
  const double e3 = hxx * (hyy * hzz - hyz * hyz) - hxy * (hxy * hzz - hyz * hxz) +
                    hxz * (hxy * hyz - hyy * hxz);
  const double detg = 1.0 + (e1 + (e2 + e3));
  c.sqrtg = sqrt(detg);
  const double gxx = 1.0 + hxx, gyy = 1.0 + hyy, gzz = 1.0 + hzz;
  const double idet = 1.0 / detg;
  double gi[6];
  gi[0] = (gyy * gzz - hyz * hyz) * idet;
  gi[1] = (hxz * hyz - hxy * gzz) * idet;
  gi[2] = (hxy * hyz - hxz * gyy) * idet;
  gi[3] = (gxx * gzz - hxz * hxz) * idet;
  gi[4] = (hxy * hxz - gxx * hyz) * idet;
  gi[5] = (gxx * gyy - hxy * hxy) * idet;
#pragma unroll
  for (int n = 0; n < 6; ++n) c.gamu[n] = gi[n];
  const double bl0 = h[1], bl1 = h[2], bl2 = h[3];
  const double b0 = gi[0] * bl0 + gi[1] * bl1 + gi[2] * bl2;
  const double b1 = gi[1] * bl0 + gi[3] * bl1 + gi[4] * bl2;
  const double b2 = gi[2] * bl0 + gi[4] * bl1 + gi[5] * bl2;
  c.beta[0] = b0;
  c.beta[1] = b1;
  c.beta[2] = b2;
  const double a2 = (1.0 - h[0]) + (b0 * bl0 + b1 * bl1 + b2 * bl2);
  c.alpha = sqrt(a2);
  const double ia2 = 1.0 / a2;
  c.gu[0] = -ia2;
  c.gu[1] = b0 * ia2;
  c.gu[2] = b1 * ia2;
  c.gu[3] = b2 * ia2;
  c.gu[4] = gi[0] - b0 * b0 * ia2;
  c.gu[5] = gi[1] - b1 * b0 * ia2;
  c.gu[6] = gi[2] - b2 * b0 * ia2;
  c.gu[7] = gi[3] - b1 * b1 * ia2;
  c.gu[8] = gi[4] - b2 * b1 * ia2;
  c.gu[9] = gi[5] - b2 * b2 * ia2;
}

// ---------------------------------------------------------------------------
// The source -alpha sqrt(gamma) (S0 + Z) (gh_node_source_lean,
// src/pointwise.jl:612-741): D[a] is d_a g packed (a = t, x, y, z), Hl the
// lowered gauge source and dH[c][a] = d_c H_a.
template <bool HAS_H>
TGH_HD void gh_source(const Coef& c, const double h[NC], const double D[4][NC], const double Hl[4],
                      const double dH[4][4], double g0, double g2, double out[NC]) {
  const double* Gp = c.gu;
  // The terms are summed in an order that lets the inputs die early: the
  // gauge source's derivative first, then C2 (D and G), then g^{ab} Gamma^c_ab
  // and its contraction with D (from D directly), then the forty
  // Christoffels (D's last use) and their products.
  double S[NC];
#pragma unroll
  for (int n = 0; n < NC; ++n) S[n] = HAS_H ? -(dH[prow(n)][pcol(n)] + dH[pcol(n)][prow(n)]) : 0.0;
  // (1) C2 + C2^T, one first index a at a time: E = G D_a, Cuu = E G,
  //     C2[a][b] = sum_{mu,nu} Cuu[mu nu] D[mu](nu, b).
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    double E[4][4];
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int col = 0; col < 4; ++col) {
        double s = Gp[pidx(r, 0)] * D[a][pidx(0, col)];
#pragma unroll
        for (int x = 1; x < 4; ++x) s += Gp[pidx(r, x)] * D[a][pidx(x, col)];
        E[r][col] = s;
      }
    double Cuu[NC];
#pragma unroll
    for (int n = 0; n < NC; ++n) {
      const int i = prow(n), j = pcol(n);
      double s = E[i][0] * Gp[pidx(0, j)];
#pragma unroll
      for (int y = 1; y < 4; ++y) s += E[i][y] * Gp[pidx(y, j)];
      Cuu[n] = s;
    }
    double C2[4];
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      double s = 0.0;
#pragma unroll
      for (int m = 0; m < 16; ++m) s += Cuu[pidx(m % 4, m / 4)] * D[m % 4][pidx(m / 4, b)];
      C2[b] = s;
    }
#pragma unroll
    for (int n = 0; n < NC; ++n) {
      const int i = prow(n), j = pcol(n);
      if (i == a) S[n] += C2[j];
      if (j == a) S[n] += C2[i];
    }
  }
  // (2) Gup^c = g^{ab} Gamma^c_ab = G^{cx} L_x with L_x = g^{ab} Gamma_{x ab}, the
  //     lowered Christoffels Gamma_{x bc} = (D_b[x,c] + D_c[x,b] - D_x[b,c]) / 2;
  //     then -Gup^c D_c[n]
  double Gup[4];
  {
    double L[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int n = 0; n < NC; ++n) {
      const int i = prow(n), j = pcol(n);
      const double gn = (i == j ? 1.0 : 2.0) * Gp[n];
#pragma unroll
      for (int x = 0; x < 4; ++x) L[x] += gn * (0.5 * (D[i][pidx(x, j)] + D[j][pidx(x, i)] - D[x][n]));
    }
#pragma unroll
    for (int cc = 0; cc < 4; ++cc)
      Gup[cc] = Gp[pidx(cc, 0)] * L[0] + Gp[pidx(cc, 1)] * L[1] + Gp[pidx(cc, 2)] * L[2] + Gp[pidx(cc, 3)] * L[3];
  }
#pragma unroll
  for (int n = 0; n < NC; ++n) S[n] -= Gup[0] * D[0][n] + Gup[1] * D[1][n] + Gup[2] * D[2][n] + Gup[3] * D[3][n];
  //     the forty Christoffels Gamma^a_bc = G^{ax} Gamma_{x bc}, and
  //     -2 Gamma^c_{d a} Gamma^d_{c b} + 2 Gamma^c_ab H_c
  double Gam[4][NC];
#pragma unroll
  for (int n = 0; n < NC; ++n) {
    const int i = prow(n), j = pcol(n);
    double low[4];
#pragma unroll
    for (int x = 0; x < 4; ++x) low[x] = 0.5 * (D[i][pidx(x, j)] + D[j][pidx(x, i)] - D[x][n]);
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      double s = Gp[pidx(a, 0)] * low[0];
#pragma unroll
      for (int x = 1; x < 4; ++x) s += Gp[pidx(a, x)] * low[x];
      Gam[a][n] = s;
    }
  }
#pragma unroll
  for (int n = 0; n < NC; ++n) {
    const int a = prow(n), b = pcol(n);
    double t2 = 0.0;
#pragma unroll
    for (int m = 0; m < 16; ++m) t2 += Gam[m % 4][pidx(m / 4, a)] * Gam[m / 4][pidx(m % 4, b)];
    double s = -2.0 * t2;
    if (HAS_H) s += 2.0 * (Gam[0][n] * Hl[0] + Gam[1][n] * Hl[1] + Gam[2][n] * Hl[2] + Gam[3][n] * Hl[3]);
    S[n] += s;
  }
  // (3) Gundlach–Pretorius damping, t_a = -alpha delta_a^t.
  if (g0 != 0.0) {
    double Cup[4];
#pragma unroll
    for (int cc = 0; cc < 4; ++cc) {
      double gh = 0.0;
      if (HAS_H) {
        gh = Gp[pidx(cc, 0)] * Hl[0] + Gp[pidx(cc, 1)] * Hl[1] + Gp[pidx(cc, 2)] * Hl[2] +
             Gp[pidx(cc, 3)] * Hl[3];
      }
      Cup[cc] = Gup[cc] + gh;
    }
    double Cl[4];
#pragma unroll
    for (int cc = 0; cc < 4; ++cc) {
      double s = 0.0;
#pragma unroll
      for (int x = 0; x < 4; ++x) s += (eta(cc, x) + h[pidx(cc, x)]) * Cup[x];
      Cl[cc] = s;
    }
    const double tC = -c.alpha * (Gp[0] * Cl[0] + Gp[1] * Cl[1] + Gp[2] * Cl[2] + Gp[3] * Cl[3]);
    const double tl = -c.alpha;
    const double k1 = 1.0 + g2;
#pragma unroll
    for (int n = 0; n < NC; ++n) {
      const int a = prow(n), b = pcol(n);
      double d = 0.0;
      if (a == 0) d += tl * Cl[b];
      if (b == 0) d += tl * Cl[a];
      S[n] += g0 * (d - k1 * (eta(a, b) + h[n]) * tC);
    }
  }
  const double s = -(c.alpha * c.sqrtg);
#pragma unroll
  for (int n = 0; n < NC; ++n) out[n] = s * S[n];
}

// d_i beta^i and d_i (alpha sqrt(gamma) gamma^{ij}) (metric_divergences,
// src/pointwise.jl:754-778), from rows t and i of d_i g^{ab} = -g^{ac}(d_i g_cd)g^{db}.
TGH_HD void metric_divergences(const Coef& c, const double dh[3][NC], double& divb, double divA[3]) {
  const double* Gp = c.gu;
  double r1[3][4], ri[3][4];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    double v0[4], vi[4];
#pragma unroll
    for (int d = 0; d < 4; ++d) {
      double s0 = 0.0, si = 0.0;
#pragma unroll
      for (int cc = 0; cc < 4; ++cc) {
        s0 += Gp[pidx(0, cc)] * dh[i][pidx(cc, d)];
        si += Gp[pidx(i + 1, cc)] * dh[i][pidx(cc, d)];
      }
      v0[d] = s0;
      vi[d] = si;
    }
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      double s0 = 0.0, si = 0.0;
#pragma unroll
      for (int d = 0; d < 4; ++d) {
        s0 += v0[d] * Gp[pidx(d, b)];
        si += vi[d] * Gp[pidx(d, b)];
      }
      r1[i][b] = -s0;
      ri[i][b] = -si;
    }
  }
  const double gutt = Gp[0];
  divb = -((r1[0][1] + c.beta[0] * r1[0][0]) + (r1[1][2] + c.beta[1] * r1[1][0]) +
           (r1[2][3] + c.beta[2] * r1[2][0])) / gutt;
  const double a3 = c.alpha * c.alpha * c.alpha;
  double da[3], ds[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    da[i] = 0.5 * a3 * r1[i][0];
    double s = 0.0;
#pragma unroll
    for (int m = 0; m < 9; ++m) s += c.gamu[sidx3(m % 3, m / 3)] * dh[i][pidx(m % 3 + 1, m / 3 + 1)];
    ds[i] = 0.5 * c.sqrtg * s;
  }
  const double as = c.alpha * c.sqrtg;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const double gik = c.gamu[sidx3(i, k)];
      s += (da[i] * c.sqrtg + c.alpha * ds[i]) * gik +
           as * (ri[i][k + 1] + c.beta[k] * r1[i][i + 1] + c.beta[i] * r1[i][k + 1] +
                 c.beta[i] * c.beta[k] * r1[i][0]);
    }
    divA[k] = s;
  }
}

// The algebraic Kerr-Schild gauge source of a hole at rest with spin a along z
// (src/gauge.jl:293-326 with u = (1,0,0,0), S = (0,0,0,a)): H_a = -K w_a / D,
// w_a = h_at, K = h_tt, s = a h_tz, D = M + sqrt(M^2 - s^2); dH[c][a] = d_c H_a.
TGH_HD void ks_source(double M, double a, const double h[NC], const double D[4][NC], double Hl[4],
                      double dH[4][4]) {
  const double w[4] = {h[0], h[1], h[2], h[3]};
  const double K = w[0];
  const double sw = a * w[3];
  const double rad = M * M - sw * sw;
  const double root = rad > 0.0 ? sqrt(rad) : 0.0;
  const double Dn = M + root;
  const double iD = 1.0 / Dn;
#pragma unroll
  for (int q = 0; q < 4; ++q) Hl[q] = (-K * iD) * w[q];
#pragma unroll
  for (int cc = 0; cc < 4; ++cc) {
    const double dw0 = D[cc][0], dw3 = D[cc][3];
    const double dK = dw0;
    const double dD = root > 0.0 ? -sw * (a * dw3) / root : 0.0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const double dwq = D[cc][q];
      dH[cc][q] = -(dK * w[q] + K * dwq) * iD + K * w[q] * dD * (iD * iD);
    }
  }
}

// KerrSchildSource with a velocity (src/gauge.jl:257-359): with w_a = k_ab u^b,
// K = u^a w_a, s = S^a w_a and D = M + sqrt(M^2 - s^2), H_a = -K w_a / D and
// its gradient, u = gamma (1, v) and S the boosted spin.
TGH_HD void ks_source_moving(double M, const double S[4], const double u[4], const double h[NC],
                             const double D[4][NC], double Hl[4], double dH[4][4]) {
  double w[4];
#pragma unroll
  for (int q = 0; q < 4; ++q)
    w[q] = h[pidx(q, 0)] * u[0] + h[pidx(q, 1)] * u[1] + h[pidx(q, 2)] * u[2] + h[pidx(q, 3)] * u[3];
  const double K = u[0] * w[0] + u[1] * w[1] + u[2] * w[2] + u[3] * w[3];
  const double sw = S[0] * w[0] + S[1] * w[1] + S[2] * w[2] + S[3] * w[3];
  const double rad = M * M - sw * sw;
  const double root = rad > 0.0 ? sqrt(rad) : 0.0;
  const double Dn = M + root;
#pragma unroll
  for (int q = 0; q < 4; ++q) Hl[q] = (-K / Dn) * w[q];
#pragma unroll
  for (int cc = 0; cc < 4; ++cc) {
    double dw[4];
#pragma unroll
    for (int q = 0; q < 4; ++q)
      dw[q] = D[cc][pidx(q, 0)] * u[0] + D[cc][pidx(q, 1)] * u[1] + D[cc][pidx(q, 2)] * u[2] +
              D[cc][pidx(q, 3)] * u[3];
    const double dK = u[0] * dw[0] + u[1] * dw[1] + u[2] * dw[2] + u[3] * dw[3];
    const double Sdw = S[0] * dw[0] + S[1] * dw[1] + S[2] * dw[2] + S[3] * dw[3];
    const double dD = root > 0.0 ? -sw * Sdw / root : 0.0;
#pragma unroll
    for (int q = 0; q < 4; ++q) dH[cc][q] = -(dK * w[q] + K * dw[q]) / Dn + K * w[q] * dD / (Dn * Dn);
  }
}

// matter_source (src/pointwise.jl:667-698): 16 pi alpha sqrt(gamma) (T_ab - g_ab T/2).
TGH_HD void matter_source(const Coef& c, const double h[NC], double rho, const double S[3],
                          const double Sij[6], double out[NC]) {
  const double pi = 3.14159265358979323846;
  const double c8 = 8.0 * pi, c16 = 16.0 * pi;
  const double* b = c.beta;
  const double Sb0 = Sij[0] * b[0] + Sij[1] * b[1] + Sij[2] * b[2];
  const double Sb1 = Sij[1] * b[0] + Sij[3] * b[1] + Sij[4] * b[2];
  const double Sb2 = Sij[2] * b[0] + Sij[4] * b[1] + Sij[5] * b[2];
  const double bS = b[0] * S[0] + b[1] * S[1] + b[2] * S[2];
  const double Ttt = c.alpha * c.alpha * rho - 2.0 * (c.alpha * bS) + (b[0] * Sb0 + b[1] * Sb1 + b[2] * Sb2);
  const double Tt[3] = {Sb0 - c.alpha * S[0], Sb1 - c.alpha * S[1], Sb2 - c.alpha * S[2]};
  const double* gi = c.gamu;
  const double trS = gi[0] * Sij[0] + gi[3] * Sij[3] + gi[5] * Sij[5] +
                     2.0 * (gi[1] * Sij[1] + gi[2] * Sij[2] + gi[4] * Sij[4]);
  const double tr8 = c8 * (trS - rho);
  const double am = c.alpha * c.sqrtg;
  out[0] = am * (c16 * Ttt - (h[0] - 1.0) * tr8);
#pragma unroll
  for (int i = 0; i < 3; ++i) out[1 + i] = am * (c16 * Tt[i] - h[1 + i] * tr8);
#pragma unroll
  for (int n = 4; n < NC; ++n) {
    const int i = prow(n) - 1, j = pcol(n) - 1;
    out[n] = am * (c16 * Sij[sidx3(i, j)] - (eta(i + 1, j + 1) + h[n]) * tr8);
  }
}

// ---------------------------------------------------------------------------
// Profiles.
TGH_HD double smoothstep(double s) {  // src/interior.jl:117-120
  double u = fmin(fmax(s, 0.0), 1.0);
  double v = u * u * u * (10.0 + u * (6.0 * u - 15.0));
  return fmin(fmax(v, 0.0), 1.0);
}

// ---------------------------------------------------------------------------
// Forward-mode dual numbers for the analytic backgrounds' first derivatives.
template <int K>
struct Dual {
  double v;
  double d[K];
};
template <int K>
TGH_HD Dual<K> dconst(double x) {
  Dual<K> r;
  r.v = x;
#pragma unroll
  for (int k = 0; k < K; ++k) r.d[k] = 0.0;
  return r;
}
template <int K> TGH_HD Dual<K> operator+(const Dual<K>& a, const Dual<K>& b) {
  Dual<K> r; r.v = a.v + b.v;
#pragma unroll
  for (int k = 0; k < K; ++k) r.d[k] = a.d[k] + b.d[k];
  return r;
}
template <int K> TGH_HD Dual<K> operator-(const Dual<K>& a, const Dual<K>& b) {
  Dual<K> r; r.v = a.v - b.v;
#pragma unroll
  for (int k = 0; k < K; ++k) r.d[k] = a.d[k] - b.d[k];
  return r;
}
template <int K> TGH_HD Dual<K> operator-(const Dual<K>& a) {
  Dual<K> r; r.v = -a.v;
#pragma unroll
  for (int k = 0; k < K; ++k) r.d[k] = -a.d[k];
  return r;
}
template <int K> TGH_HD Dual<K> operator*(const Dual<K>& a, const Dual<K>& b) {
  Dual<K> r; r.v = a.v * b.v;
#pragma unroll
  for (int k = 0; k < K; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
  return r;
}
template <int K> TGH_HD Dual<K> operator*(double a, const Dual<K>& b) {
  Dual<K> r; r.v = a * b.v;
#pragma unroll
  for (int k = 0; k < K; ++k) r.d[k] = a * b.d[k];
  return r;
}
template <int K> TGH_HD Dual<K> operator+(double a, const Dual<K>& b) {
  Dual<K> r = b; r.v = a + b.v; return r;
}
template <int K> TGH_HD Dual<K> operator/(const Dual<K>& a, const Dual<K>& b) {
  Dual<K> r; const double ib = 1.0 / b.v; r.v = a.v * ib;
#pragma unroll
  for (int k = 0; k < K; ++k) r.d[k] = (a.d[k] - r.v * b.d[k]) * ib;
  return r;
}
template <int K> TGH_HD Dual<K> dsqrt(const Dual<K>& a) {
  Dual<K> r; r.v = sqrt(a.v); const double f = 0.5 / r.v;
#pragma unroll
  for (int k = 0; k < K; ++k) r.d[k] = a.d[k] * f;
  return r;
}
template <int K> TGH_HD Dual<K> dsin(const Dual<K>& a) {
  Dual<K> r; r.v = sin(a.v); const double cc = cos(a.v);
#pragma unroll
  for (int k = 0; k < K; ++k) r.d[k] = a.d[k] * cc;
  return r;
}
template <int K> TGH_HD Dual<K> dtanh(const Dual<K>& a) {
  Dual<K> r; r.v = tanh(a.v); const double s = 1.0 - r.v * r.v;
#pragma unroll
  for (int k = 0; k < K; ++k) r.d[k] = a.d[k] * s;
  return r;
}

enum { BG_MINKOWSKI = 0, BG_GAUGE_WAVE = 1, BG_KERR_SCHILD = 2, BG_SHIFTED = 3, BG_HARMONIC = 4 };

struct Bg {
  int kind;
  double M, a;
  double c[3];
  double A, len;  // gauge wave A, d; shifted Minkowski A, w
  double u[3];    // boost velocity of a Kerr-Schild or harmonic hole (it moves at -u)
};

// Second-order jets (value, gradient and packed Hessian in the four
// coordinates t, x, y, z) for the sampled gauge source, which needs the
// metric's second derivatives.
struct Jet2 {
  double v;
  double d[4];
  double dd[NC];  // d_a d_b, packed as the metric's components
};
TGH_HD Jet2 jconst(double x) {
  Jet2 r;
  r.v = x;
#pragma unroll
  for (int k = 0; k < 4; ++k) r.d[k] = 0.0;
#pragma unroll
  for (int n = 0; n < NC; ++n) r.dd[n] = 0.0;
  return r;
}
TGH_HD Jet2 operator+(const Jet2& a, const Jet2& b) {
  Jet2 r; r.v = a.v + b.v;
#pragma unroll
  for (int k = 0; k < 4; ++k) r.d[k] = a.d[k] + b.d[k];
#pragma unroll
  for (int n = 0; n < NC; ++n) r.dd[n] = a.dd[n] + b.dd[n];
  return r;
}
TGH_HD Jet2 operator-(const Jet2& a, const Jet2& b) {
  Jet2 r; r.v = a.v - b.v;
#pragma unroll
  for (int k = 0; k < 4; ++k) r.d[k] = a.d[k] - b.d[k];
#pragma unroll
  for (int n = 0; n < NC; ++n) r.dd[n] = a.dd[n] - b.dd[n];
  return r;
}
TGH_HD Jet2 operator-(const Jet2& a) {
  Jet2 r; r.v = -a.v;
#pragma unroll
  for (int k = 0; k < 4; ++k) r.d[k] = -a.d[k];
#pragma unroll
  for (int n = 0; n < NC; ++n) r.dd[n] = -a.dd[n];
  return r;
}
TGH_HD Jet2 operator*(const Jet2& a, const Jet2& b) {
  Jet2 r; r.v = a.v * b.v;
#pragma unroll
  for (int k = 0; k < 4; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
#pragma unroll
  for (int n = 0; n < NC; ++n) {
    const int i = prow(n), j = pcol(n);
    r.dd[n] = a.dd[n] * b.v + a.v * b.dd[n] + a.d[i] * b.d[j] + a.d[j] * b.d[i];
  }
  return r;
}
TGH_HD Jet2 operator*(double a, const Jet2& b) {
  Jet2 r; r.v = a * b.v;
#pragma unroll
  for (int k = 0; k < 4; ++k) r.d[k] = a * b.d[k];
#pragma unroll
  for (int n = 0; n < NC; ++n) r.dd[n] = a * b.dd[n];
  return r;
}
TGH_HD Jet2 operator+(double a, const Jet2& b) { Jet2 r = b; r.v = a + b.v; return r; }
// A unary function f applied to a jet: f(a), f'(a), f''(a) given.
TGH_HD Jet2 jchain(const Jet2& a, double f0, double f1, double f2) {
  Jet2 r; r.v = f0;
#pragma unroll
  for (int k = 0; k < 4; ++k) r.d[k] = f1 * a.d[k];
#pragma unroll
  for (int n = 0; n < NC; ++n) {
    const int i = prow(n), j = pcol(n);
    r.dd[n] = f1 * a.dd[n] + f2 * a.d[i] * a.d[j];
  }
  return r;
}
TGH_HD Jet2 operator/(const Jet2& a, const Jet2& b) {
  const double ib = 1.0 / b.v;
  return a * jchain(b, ib, -ib * ib, 2.0 * ib * ib * ib);
}
TGH_HD Jet2 dsqrt(const Jet2& a) {
  const double s = sqrt(a.v);
  return jchain(a, s, 0.5 / s, -0.25 / (s * a.v));
}
TGH_HD Jet2 dsin(const Jet2& a) {
  const double sn = sin(a.v), cs = cos(a.v);
  return jchain(a, sn, cs, -sn);
}
TGH_HD Jet2 dtanh(const Jet2& a) {
  const double th = tanh(a.v), s = 1.0 - th * th;
  return jchain(a, th, s, -2.0 * th * s);
}
template <class J> TGH_HD J jet_const(double x);
template <> TGH_HD Dual<4> jet_const<Dual<4>>(double x) { return dconst<4>(x); }
template <> TGH_HD Jet2 jet_const<Jet2>(double x) { return jconst(x); }
TGH_HD void jet_seed(Dual<4>& x, int k) { x.d[k] = 1.0; }
TGH_HD void jet_seed(Jet2& x, int k) { x.d[k] = 1.0; }

// Nested forward-mode duals for the harmonic chart, whose Jacobian needs its
// own derivatives. Every operation rounds in a fixed order with contraction
// off, the order of the oracle's restatement (a dual times a constant is the
// product with the constant's zero-derivative dual, a quotient's derivative
// (a' b - a b') / b^2, a root's a' / (2 sqrt a)), so that the chart's values
// -- the Dirichlet data of a box around the hole, where the right-hand side's
// stencils multiply them by 1/h^2 -- agree with it to the last bits.
template <class V, int K>
struct NDual {
  V v;
  V d[K];
};
TGH_HD double nd_cst(double x, const double*) { return x; }
template <class V, int K>
TGH_HD NDual<V, K> nd_cst(double x, const NDual<V, K>*) {
  NDual<V, K> r;
  r.v = nd_cst(x, (const V*)nullptr);
#pragma unroll
  for (int k = 0; k < K; ++k) r.d[k] = nd_cst(0.0, (const V*)nullptr);
  return r;
}
TGH_HD double nd_add(double a, double b) {
#pragma clang fp contract(off)
  return a + b;
}
TGH_HD double nd_sub(double a, double b) {
#pragma clang fp contract(off)
  return a - b;
}
TGH_HD double nd_mul(double a, double b) {
#pragma clang fp contract(off)
  return a * b;
}
TGH_HD double nd_div(double a, double b) {
#pragma clang fp contract(off)
  return a / b;
}
TGH_HD double nd_neg(double a) { return -a; }
TGH_HD double nd_sqrt(double a) { return sqrt(a); }
template <class V, int K>
TGH_HD NDual<V, K> nd_add(const NDual<V, K>& a, const NDual<V, K>& b) {
  NDual<V, K> r;
  r.v = nd_add(a.v, b.v);
#pragma unroll
  for (int k = 0; k < K; ++k) r.d[k] = nd_add(a.d[k], b.d[k]);
  return r;
}
template <class V, int K>
TGH_HD NDual<V, K> nd_sub(const NDual<V, K>& a, const NDual<V, K>& b) {
  NDual<V, K> r;
  r.v = nd_sub(a.v, b.v);
#pragma unroll
  for (int k = 0; k < K; ++k) r.d[k] = nd_sub(a.d[k], b.d[k]);
  return r;
}
template <class V, int K>
TGH_HD NDual<V, K> nd_neg(const NDual<V, K>& a) {
  NDual<V, K> r;
  r.v = nd_neg(a.v);
#pragma unroll
  for (int k = 0; k < K; ++k) r.d[k] = nd_neg(a.d[k]);
  return r;
}
template <class V, int K>
TGH_HD NDual<V, K> nd_mul(const NDual<V, K>& a, const NDual<V, K>& b) {
  NDual<V, K> r;
  r.v = nd_mul(a.v, b.v);
#pragma unroll
  for (int k = 0; k < K; ++k) r.d[k] = nd_add(nd_mul(a.d[k], b.v), nd_mul(a.v, b.d[k]));
  return r;
}
template <class V, int K>
TGH_HD NDual<V, K> nd_div(const NDual<V, K>& a, const NDual<V, K>& b) {
  NDual<V, K> r;
  r.v = nd_div(a.v, b.v);
  const V bb = nd_mul(b.v, b.v);
#pragma unroll
  for (int k = 0; k < K; ++k) r.d[k] = nd_div(nd_sub(nd_mul(a.d[k], b.v), nd_mul(a.v, b.d[k])), bb);
  return r;
}
template <class V, int K>
TGH_HD NDual<V, K> nd_sqrt(const NDual<V, K>& a) {
  NDual<V, K> r;
  r.v = nd_sqrt(a.v);
  const V two_r = nd_mul(nd_cst(2.0, (const V*)nullptr), r.v);
#pragma unroll
  for (int k = 0; k < K; ++k) r.d[k] = nd_div(a.d[k], two_r);
  return r;
}
template <class V, int K>
TGH_HD NDual<V, K> operator+(const NDual<V, K>& a, const NDual<V, K>& b) { return nd_add(a, b); }
template <class V, int K>
TGH_HD NDual<V, K> operator-(const NDual<V, K>& a, const NDual<V, K>& b) { return nd_sub(a, b); }
template <class V, int K>
TGH_HD NDual<V, K> operator-(const NDual<V, K>& a) { return nd_neg(a); }
template <class V, int K>
TGH_HD NDual<V, K> operator*(const NDual<V, K>& a, const NDual<V, K>& b) { return nd_mul(a, b); }
template <class V, int K>
TGH_HD NDual<V, K> operator/(const NDual<V, K>& a, const NDual<V, K>& b) { return nd_div(a, b); }
template <class V, int K>
TGH_HD NDual<V, K> operator+(const NDual<V, K>& a, double b) { return nd_add(a, nd_cst(b, &a)); }
template <class V, int K>
TGH_HD NDual<V, K> operator-(const NDual<V, K>& a, double b) { return nd_sub(a, nd_cst(b, &a)); }
template <class V, int K>
TGH_HD NDual<V, K> operator*(double a, const NDual<V, K>& b) { return nd_mul(nd_cst(a, &b), b); }
template <class V, int K>
TGH_HD NDual<V, K> operator/(const NDual<V, K>& a, double b) { return nd_div(a, nd_cst(b, &a)); }
template <class V, int K>
TGH_HD NDual<V, K> operator/(double a, const NDual<V, K>& b) { return nd_div(nd_cst(a, &b), b); }

// The Kerr metric in harmonic coordinates (SpacetimeMetrics' Harmonic(M, a),
// restated from Cook & Scheel, Phys. Rev. D 56, 4775 (1997)) as the packed
// offset h = g - eta with first derivatives in (t, x, y, z), spin along +z
// about c. The harmonic chart is ingoing Kerr-Schild with the radius lowered
// by M:
//   X + i Y = (R + i a) e^{i phi} sin(theta), Z = R cos(theta), R = r - M,
//   t_H = t_KS - 2 M ln(r - r_-),
// so a harmonic point maps to the Kerr-Schild point
//   x + i y = (X + i Y)(r + i a)/(R + i a), z = r Z / R, t = t_H + 2 M ln(r - r_-),
// R solving the Kerr-Schild quartic in (X, Y, Z), and g_H = J^T g_KS J with
// J the map's Jacobian, J^t_i = 2 M / (r - r_-) d_i R. The map is evaluated
// in duals over (X, Y, Z) whose components are duals over (t, x, y, z), so
// each entry of J carries its own first derivatives.
using NS = NDual<double, 4>;  // a value with its derivatives in (t, x, y, z)
TGH_HD NS ns_cst(double x) { return nd_cst(x, (const NS*)nullptr); }

// The harmonic chart's full metric g at the hole-relative harmonic position
// (X, Y, Z), each carried with its derivatives in (t, x, y, z).
TGH_HD void harmonic_metric_exact(const Bg& bg, const NS& Xs, const NS& Ys, const NS& Zs, NS g[4][4]) {
  using S = NS;
  using D3 = NDual<S, 3>;
  const double a = bg.a, M = bg.M;
  const double rm = M - sqrt(M * M - a * a);
  const double aa = a * a;
  D3 P[3];
  const S* in[3] = {&Xs, &Ys, &Zs};
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    P[i] = nd_cst(0.0, (const D3*)nullptr);
    P[i].v = *in[i];
    P[i].d[i] = ns_cst(1.0);
  }
  const D3 &X = P[0], &Y = P[1], &Z = P[2];
  D3 rho2 = X * X + Y * Y + Z * Z;
  D3 hb = (rho2 - aa) / 2.0;
  D3 R2 = hb + nd_sqrt(hb * hb + aa * Z * Z);
  D3 R = nd_sqrt(R2);
  D3 r = R + M;
  D3 den = R2 + aa;
  D3 c1 = (r * R + aa) / den, c2 = (-a * M) / den;
  D3 m[3];
  m[0] = c1 * X - c2 * Y;
  m[1] = c2 * X + c1 * Y;
  m[2] = r * Z / R;
  // J[c][b] = d x_KS^c / d x_H^b; the time column is the identity's.
  S J[4][4];
#pragma unroll
  for (int c = 0; c < 4; ++c)
#pragma unroll
    for (int b = 0; b < 4; ++b) J[c][b] = ns_cst(c == 0 && b == 0 ? 1.0 : 0.0);
  const S tfac = (2.0 * M) / (r.v - rm);
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    J[0][i + 1] = tfac * R.d[i];
#pragma unroll
    for (int c = 0; c < 3; ++c) J[c + 1][i + 1] = m[c].d[i];
  }
  // g_KS at the mapped point, whose Kerr-Schild radius is r = R + M.
  const S rv = r.v, r2v = rv * rv, xv = m[0].v, yv = m[1].v, zv = m[2].v;
  const S f = (2.0 * M) * rv * r2v / (r2v * r2v + aa * zv * zv);
  const S dl = r2v + aa;
  S l[4];
  l[0] = ns_cst(1.0);
  l[1] = (rv * xv + a * yv) / dl;
  l[2] = (rv * yv - a * xv) / dl;
  l[3] = zv / rv;
  S gk[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) gk[i][j] = ns_cst(eta(i, j)) + f * l[i] * l[j];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      S sum = ns_cst(0.0);
#pragma unroll
      for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int e = 0; e < 4; ++e) sum = sum + J[c][i] * J[e][j] * gk[c][e];
      g[i][j] = sum;
    }
}

// The Kerr-Schild metric g = eta + f l l at the hole-relative position, in the
// same fixed order (the moving hole's chart).
TGH_HD void ks_metric_exact(const Bg& bg, const NS& X, const NS& Y, const NS& Z, NS g[4][4]) {
  const double a = bg.a, M = bg.M;
  NS rho2 = X * X + Y * Y + Z * Z;
  NS hb = (rho2 - a * a) / 2.0;
  NS r2 = hb + nd_sqrt(hb * hb + (a * a) * Z * Z);
  NS r = nd_sqrt(r2);
  NS f = (2.0 * M) * r * r2 / (r2 * r2 + (a * a) * Z * Z);
  NS den = r2 + a * a;
  NS l[4] = {ns_cst(1.0), (r * X + a * Y) / den, (r * Y - a * X) / den, Z / r};
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) g[i][j] = ns_cst(eta(i, j)) + f * l[i] * l[j];
}

TGH_HD void bg_harmonic_offset(const Bg& bg, const double x[3], Dual<4> hd[NC]) {
  NS P[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    P[i] = ns_cst(x[i]);
    P[i].d[i + 1] = 1.0;
    P[i] = P[i] - bg.c[i];
  }
  NS g[4][4];
  harmonic_metric_exact(bg, P[0], P[1], P[2], g);
#pragma unroll
  for (int n = 0; n < NC; ++n) {
    const int i = prow(n), j = pcol(n);
    hd[n].v = nd_sub(g[i][j].v, eta(i, j));
#pragma unroll
    for (int k = 0; k < 4; ++k) hd[n].d[k] = g[i][j].d[k];
  }
}

// boost(m, u) and translate (SpacetimeMetrics, as the reference states it at
// src/interior.jl:66-72, src/initialdata.jl:93-96): m evaluated at
// x' = Lambda (t, x - c), Lambda the symmetric Lorentz boost of velocity u,
//   Lambda^0_0 = gamma, Lambda^0_i = Lambda^i_0 = gamma u_i,
//   Lambda^i_j = delta_ij + (gamma - 1) u_i u_j / u^2,
// and pulled back, g_ab = Lambda^c_a Lambda^d_b m_cd(x'): the hole moves at -u.
TGH_HD bool bg_boosted(const Bg& bg) {
  return (bg.kind == BG_KERR_SCHILD || bg.kind == BG_HARMONIC) && (bg.u[0] != 0.0 || bg.u[1] != 0.0 || bg.u[2] != 0.0);
}
TGH_HD void lorentz_boost(const double u[3], double L[4][4]) {
#pragma clang fp contract(off)
  const double u2 = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
  const double gam = 1.0 / sqrt(1.0 - u2);
  const double f = (gam - 1.0) / u2;
  L[0][0] = gam;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    L[0][i + 1] = gam * u[i];
    L[i + 1][0] = gam * u[i];
#pragma unroll
    for (int j = 0; j < 3; ++j) L[i + 1][j + 1] = (i == j ? 1.0 : 0.0) + f * u[i] * u[j];
  }
}

// The oracle-order 4x4 helpers behind metric_quantities' alpha, beta and
// sqrt(gamma) for the moving hole's Pi (see bg_state_exact).
TGH_HD double ex_det3(double a, double b, double c, double d, double e, double f, double g, double h, double i) {
#pragma clang fp contract(off)
  return a * (e * i - f * h) - b * (d * i - f * g) + c * (d * h - e * g);
}
TGH_HD double ex_det4(const double m[4][4]) {
#pragma clang fp contract(off)
  const double c0 = ex_det3(m[1][1], m[1][2], m[1][3], m[2][1], m[2][2], m[2][3], m[3][1], m[3][2], m[3][3]);
  const double c1 = ex_det3(m[1][0], m[1][2], m[1][3], m[2][0], m[2][2], m[2][3], m[3][0], m[3][2], m[3][3]);
  const double c2 = ex_det3(m[1][0], m[1][1], m[1][3], m[2][0], m[2][1], m[2][3], m[3][0], m[3][1], m[3][3]);
  const double c3 = ex_det3(m[1][0], m[1][1], m[1][2], m[2][0], m[2][1], m[2][2], m[3][0], m[3][1], m[3][2]);
  return m[0][0] * c0 - m[0][1] * c1 + m[0][2] * c2 - m[0][3] * c3;
}
TGH_HD void ex_matmul(const double A[4][4], const double B[4][4], double C[4][4]) {
#pragma clang fp contract(off)
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      double s = A[i][0] * B[0][j];
#pragma unroll
      for (int k = 1; k < 4; ++k) s = s + A[i][k] * B[k][j];
      C[i][j] = s;
    }
}
TGH_HD void ex_inv4(const double m[4][4], double out[4][4]) {
#pragma clang fp contract(off)
  double adj[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      int r[3], c[3], nr = 0, nc = 0;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        if (k != j) r[nr++] = k;
        if (k != i) c[nc++] = k;
      }
      const double mn = ex_det3(m[r[0]][c[0]], m[r[0]][c[1]], m[r[0]][c[2]], m[r[1]][c[0]], m[r[1]][c[1]],
                                m[r[1]][c[2]], m[r[2]][c[0]], m[r[2]][c[1]], m[r[2]][c[2]]);
      adj[i][j] = ((i + j) % 2 == 0) ? mn : -mn;
    }
  const double d = ex_det4(m);
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) out[i][j] = adj[i][j] / d;
}
// alpha, beta^i and sqrt(gamma) of metric_quantities (src/pointwise.jl:145-184)
// in the oracle's order: g^{ab} by the offset -(g^-1 h) eta, sqrt(gamma) from
// the elementary symmetric polynomials of eta h.
TGH_HD void ex_lapse_shift(const double hv[NC], double& alpha, double beta[3], double& sqrtg) {
#pragma clang fp contract(off)
  double h[4][4], e4[4][4], g4[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      h[a][b] = hv[pidx(a, b)];
      e4[a][b] = eta(a, b);
      g4[a][b] = eta(a, b) + h[a][b];
    }
  double gi[4][4], gih[4][4], tt[4][4];
  ex_inv4(g4, gi);
  ex_matmul(gi, h, gih);
  ex_matmul(gih, e4, tt);
  const double q1 = -tt[0][0];
  const double gutt = eta(0, 0) + q1;
  alpha = 1.0 / sqrt(1.0 - q1);
#pragma unroll
  for (int i = 0; i < 3; ++i) beta[i] = -(eta(0, i + 1) + -tt[0][i + 1]) / gutt;
  double B[4][4], B2[4][4], B3[4][4];
  ex_matmul(e4, h, B);
  ex_matmul(B, B, B2);
  ex_matmul(B2, B, B3);
  const double trB = B[0][0] + B[1][1] + B[2][2] + B[3][3];
  const double trB2 = B2[0][0] + B2[1][1] + B2[2][2] + B2[3][3];
  const double trB3 = B3[0][0] + B3[1][1] + B3[2][2] + B3[3][3];
  const double e1 = trB;
  const double e2 = (trB * trB - trB2) / 2.0;
  const double e3 = (trB * trB * trB - 3.0 * trB * trB2 + 2.0 * trB3) / 6.0;
  const double ee4 = ex_det4(B);
  const double d1 = -(e1 + e2 + e3 + ee4);
  const double detgm1 = -d1 - q1 + d1 * q1;
  sqrtg = sqrt(1.0 + detgm1);
}

// background_state's one dual pass (src/initialdata.jl:1683-1694) for a moving
// hole: h = g - eta, Pi = (sqrt(gamma)/alpha)(d_t g - beta^i d_i g) and d_t h,
// every operation in the oracle's order with contraction off. The layer's and
// the core's target rates are centered differences of these values over
// 2 delta = 2 cbrt(eps) r_0, which multiplies any difference in their last
// bits by 1/(2 delta) ~ 1e5; evaluated this way they have none.
// Kept out of line: a moving hole's core and layer call it up to three times
// a point, and inlined copies of it would multiply the kernels' size.
__host__ __device__ __noinline__ inline void bg_state_exact(const Bg& bg, double t, const double x[3], double h[NC],
                                                     double P[NC], double dtg[NC]) {
  NS X[4];
  const double p[4] = {t, x[0], x[1], x[2]};
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    X[k] = ns_cst(p[k]);
    X[k].d[k] = 1.0;
  }
  NS g[4][4];
  if (bg_boosted(bg)) {
    double L[4][4];
    lorentz_boost(bg.u, L);
    const NS xr[4] = {X[0], X[1] - bg.c[0], X[2] - bg.c[1], X[3] - bg.c[2]};
    NS xp[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) xp[c] = L[c][0] * xr[0] + L[c][1] * xr[1] + L[c][2] * xr[2] + L[c][3] * xr[3];
    NS gp[4][4];
    if (bg.kind == BG_HARMONIC)
      harmonic_metric_exact(bg, xp[1] - 0.0, xp[2] - 0.0, xp[3] - 0.0, gp);
    else
      ks_metric_exact(bg, xp[1] - 0.0, xp[2] - 0.0, xp[3] - 0.0, gp);
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        NS sum = ns_cst(0.0);
#pragma unroll
        for (int c = 0; c < 4; ++c)
#pragma unroll
          for (int d = 0; d < 4; ++d) sum = sum + nd_mul(L[c][a], L[d][b]) * gp[c][d];
        g[a][b] = sum;
      }
  } else if (bg.kind == BG_HARMONIC) {
    harmonic_metric_exact(bg, X[1] - bg.c[0], X[2] - bg.c[1], X[3] - bg.c[2], g);
  } else {
    ks_metric_exact(bg, X[1] - bg.c[0], X[2] - bg.c[1], X[3] - bg.c[2], g);
  }
  double dh[3][NC];
#pragma unroll
  for (int n = 0; n < NC; ++n) {
    const int i = prow(n), j = pcol(n);
    h[n] = nd_sub(g[i][j].v, eta(i, j));
    dtg[n] = g[i][j].d[0];
#pragma unroll
    for (int k = 0; k < 3; ++k) dh[k][n] = g[i][j].d[k + 1];
  }
  double alpha, beta[3], sqrtg;
  ex_lapse_shift(h, alpha, beta, sqrtg);
  const double s = nd_div(sqrtg, alpha);
#pragma unroll
  for (int n = 0; n < NC; ++n)
    P[n] = nd_mul(s, nd_sub(nd_sub(nd_sub(dtg[n], nd_mul(beta[0], dh[0][n])), nd_mul(beta[1], dh[1][n])),
                            nd_mul(beta[2], dh[2][n])));
}

// Sets of backgrounds, as bit masks over the kinds: a kernel compiled for one
// set keeps only those backgrounds' code (and registers). The gauge-source
// kind fixes the set: H_a = 0 admits only harmonic charts, and every chart
// that is not harmonic needs a source.
constexpr int bg_bit(int kind) { return 1 << kind; }
constexpr int BGS_MOVING = 1 << 5;  // the boosted (moving) Kerr-Schild and harmonic charts
constexpr int BGS_ALL = 0x1f | BGS_MOVING;
constexpr int BGS_STATIC = 0x1f;  // every chart at rest
constexpr int BGS_HARMONIC = bg_bit(BG_MINKOWSKI) | bg_bit(BG_GAUGE_WAVE) | bg_bit(BG_HARMONIC);
constexpr int BGS_SOURCED = bg_bit(BG_MINKOWSKI) | bg_bit(BG_KERR_SCHILD) | bg_bit(BG_SHIFTED);

// The packed offset h = g - eta of a background at (t, x), carried as jets in
// (t, x, y, z): Dual<4> for first derivatives, Jet2 for second. A kind
// outside SET gives NaN.
template <class J, int SET = BGS_ALL>
TGH_HD void bg_offset(const Bg& bg, double t, const double x[3], J hd[NC]) {
#pragma unroll
  for (int n = 0; n < NC; ++n) hd[n] = jet_const<J>(0.0);
  if (bg.kind == BG_MINKOWSKI) return;
  if (bg.kind < 0 || bg.kind > BG_HARMONIC || !(SET & bg_bit(bg.kind)) || bg_boosted(bg)) {
#pragma unroll
    for (int n = 0; n < NC; ++n) hd[n] = jet_const<J>(__builtin_nan(""));
    return;
  }
  if constexpr ((SET & bg_bit(BG_HARMONIC)) != 0) {
    if (bg.kind == BG_HARMONIC) {
      // First derivatives only: the harmonic chart has H_a = 0, so no gauge
      // source (which needs second derivatives) is ever taken from it.
      if constexpr (std::is_same<J, Dual<4>>::value) {
        bg_harmonic_offset(bg, x, hd);
      } else {
#pragma unroll
        for (int n = 0; n < NC; ++n) hd[n] = jet_const<J>(__builtin_nan(""));
      }
      return;
    }
  }
  J X[4];
  const double p[4] = {t, x[0], x[1], x[2]};
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    X[k] = jet_const<J>(p[k]);
    jet_seed(X[k], k);
  }
  if constexpr ((SET & bg_bit(BG_GAUGE_WAVE)) != 0) {
    if (bg.kind == BG_GAUGE_WAVE) {
      const double pi = 3.14159265358979323846;
      J s = dsin((2.0 * pi / bg.len) * (X[1] - X[0]));
      J Hm1 = -(bg.A * s);  // H - 1
      hd[pidx(0, 0)] = -Hm1;
      hd[pidx(1, 1)] = Hm1;
      return;
    }
  }
  if constexpr ((SET & bg_bit(BG_SHIFTED)) != 0) {
    if (bg.kind == BG_SHIFTED) {
      J th = dtanh((1.0 / bg.len) * X[1]);
      J pp = bg.A * (1.0 + (-(th * th)));
      hd[pidx(0, 1)] = -pp;
      hd[pidx(1, 1)] = -(pp * pp);
      return;
    }
  }
  if constexpr ((SET & bg_bit(BG_KERR_SCHILD)) != 0) {
    // Kerr-Schild with spin a along +z about c:
    //   r^4 - r^2 (rho^2 - a^2) - a^2 z^2 = 0, f = 2 M r^3 / (r^4 + a^2 z^2),
    //   l = (1, (r x + a y)/(r^2 + a^2), (r y - a x)/(r^2 + a^2), z / r), h = f l l.
    const double a = bg.a, M = bg.M;
    J Xr = X[1] + jet_const<J>(-bg.c[0]), Yr = X[2] + jet_const<J>(-bg.c[1]),
      Zr = X[3] + jet_const<J>(-bg.c[2]);
    J rho2 = Xr * Xr + Yr * Yr + Zr * Zr;
    J hb = 0.5 * (rho2 + jet_const<J>(-a * a));
    J r2 = hb + dsqrt(hb * hb + (a * a) * (Zr * Zr));
    J r = dsqrt(r2);
    J f = (2.0 * M) * (r * r2) / (r2 * r2 + (a * a) * (Zr * Zr));
    J den = r2 + jet_const<J>(a * a);
    J l[4];
    l[0] = jet_const<J>(1.0);
    l[1] = (r * Xr + a * Yr) / den;
    l[2] = (r * Yr + (-a) * Xr) / den;
    l[3] = Zr / r;
#pragma unroll
    for (int n = 0; n < NC; ++n) hd[n] = f * (l[prow(n)] * l[pcol(n)]);
  }
}

// The sampled gauge source of a background (gauge.jl:156-212): the lowered
// H_b = -Gamma_b = -g^{de} Gamma_{b,de} and dH[a][b] = d_a H_b, with
// Gamma_{b,de} = (d_d g_be + d_e g_bd - d_b g_de)/2 and
// d_a g^{de} = -g^{dp} (d_a g_pq) g^{qe}.
TGH_HD void gauge_source_grad(const Bg& bg, double t, const double x[3], double Hl[4],
                              double dH[4][4]) {
  Jet2 hj[NC];
  bg_offset<Jet2, BGS_SOURCED>(bg, t, x, hj);
  // g, dg[c][n] = d_c g_n, ddg[c][e][n] = d_c d_e g_n
  double g[4][4], gi[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) g[a][b] = eta(a, b) + hj[pidx(a, b)].v;
  // general 4x4 inverse by cofactors (values only, once per point)
  {
    double m[16], inv[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) m[i] = g[i / 4][i % 4];
    inv[0] = m[5] * m[10] * m[15] - m[5] * m[11] * m[14] - m[9] * m[6] * m[15] + m[9] * m[7] * m[14] +
             m[13] * m[6] * m[11] - m[13] * m[7] * m[10];
    inv[4] = -m[4] * m[10] * m[15] + m[4] * m[11] * m[14] + m[8] * m[6] * m[15] - m[8] * m[7] * m[14] -
             m[12] * m[6] * m[11] + m[12] * m[7] * m[10];
    inv[8] = m[4] * m[9] * m[15] - m[4] * m[11] * m[13] - m[8] * m[5] * m[15] + m[8] * m[7] * m[13] +
             m[12] * m[5] * m[11] - m[12] * m[7] * m[9];
    inv[12] = -m[4] * m[9] * m[14] + m[4] * m[10] * m[13] + m[8] * m[5] * m[14] - m[8] * m[6] * m[13] -
              m[12] * m[5] * m[10] + m[12] * m[6] * m[9];
    inv[1] = -m[1] * m[10] * m[15] + m[1] * m[11] * m[14] + m[9] * m[2] * m[15] - m[9] * m[3] * m[14] -
             m[13] * m[2] * m[11] + m[13] * m[3] * m[10];
    inv[5] = m[0] * m[10] * m[15] - m[0] * m[11] * m[14] - m[8] * m[2] * m[15] + m[8] * m[3] * m[14] +
             m[12] * m[2] * m[11] - m[12] * m[3] * m[10];
    inv[9] = -m[0] * m[9] * m[15] + m[0] * m[11] * m[13] + m[8] * m[1] * m[15] - m[8] * m[3] * m[13] -
             m[12] * m[1] * m[11] + m[12] * m[3] * m[9];
    inv[13] = m[0] * m[9] * m[14] - m[0] * m[10] * m[13] - m[8] * m[1] * m[14] + m[8] * m[2] * m[13] +
              m[12] * m[1] * m[10] - m[12] * m[2] * m[9];
    inv[2] = m[1] * m[6] * m[15] - m[1] * m[7] * m[14] - m[5] * m[2] * m[15] + m[5] * m[3] * m[14] +
             m[13] * m[2] * m[7] - m[13] * m[3] * m[6];
    inv[6] = -m[0] * m[6] * m[15] + m[0] * m[7] * m[14] + m[4] * m[2] * m[15] - m[4] * m[3] * m[14] -
             m[12] * m[2] * m[7] + m[12] * m[3] * m[6];
    inv[10] = m[0] * m[5] * m[15] - m[0] * m[7] * m[13] - m[4] * m[1] * m[15] + m[4] * m[3] * m[13] +
              m[12] * m[1] * m[7] - m[12] * m[3] * m[5];
    inv[14] = -m[0] * m[5] * m[14] + m[0] * m[6] * m[13] + m[4] * m[1] * m[14] - m[4] * m[2] * m[13] -
              m[12] * m[1] * m[6] + m[12] * m[2] * m[5];
    inv[3] = -m[1] * m[6] * m[11] + m[1] * m[7] * m[10] + m[5] * m[2] * m[11] - m[5] * m[3] * m[10] -
             m[9] * m[2] * m[7] + m[9] * m[3] * m[6];
    inv[7] = m[0] * m[6] * m[11] - m[0] * m[7] * m[10] - m[4] * m[2] * m[11] + m[4] * m[3] * m[10] +
             m[8] * m[2] * m[7] - m[8] * m[3] * m[6];
    inv[11] = -m[0] * m[5] * m[11] + m[0] * m[7] * m[9] + m[4] * m[1] * m[11] - m[4] * m[3] * m[9] -
              m[8] * m[1] * m[7] + m[8] * m[3] * m[5];
    inv[15] = m[0] * m[5] * m[10] - m[0] * m[6] * m[9] - m[4] * m[1] * m[10] + m[4] * m[2] * m[9] +
              m[8] * m[1] * m[6] - m[8] * m[2] * m[5];
    const double det = m[0] * inv[0] + m[1] * inv[4] + m[2] * inv[8] + m[3] * inv[12];
    const double idet = 1.0 / det;
#pragma unroll
    for (int i = 0; i < 16; ++i) gi[i / 4][i % 4] = inv[i] * idet;
  }
  // Gamma_{b,de} and its derivative d_a Gamma_{b,de}
  double Gam[4][4][4];
#pragma unroll
  for (int b = 0; b < 4; ++b)
#pragma unroll
    for (int d = 0; d < 4; ++d)
#pragma unroll
      for (int e = 0; e < 4; ++e)
        Gam[b][d][e] = 0.5 * (hj[pidx(b, e)].d[d] + hj[pidx(b, d)].d[e] - hj[pidx(d, e)].d[b]);
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    double s = 0.0;
#pragma unroll
    for (int d = 0; d < 4; ++d)
#pragma unroll
      for (int e = 0; e < 4; ++e) s += gi[d][e] * Gam[b][d][e];
    Hl[b] = -s;
  }
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    // d_a g^{de} = -g^{dp} (d_a g_pq) g^{qe}
    double t1[4][4], dgi[4][4];
#pragma unroll
    for (int d = 0; d < 4; ++d)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        double s = 0.0;
#pragma unroll
        for (int pp = 0; pp < 4; ++pp) s += gi[d][pp] * hj[pidx(pp, q)].d[a];
        t1[d][q] = s;
      }
#pragma unroll
    for (int d = 0; d < 4; ++d)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        double s = 0.0;
#pragma unroll
        for (int q = 0; q < 4; ++q) s += t1[d][q] * gi[q][e];
        dgi[d][e] = -s;
      }
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      double s = 0.0;
#pragma unroll
      for (int d = 0; d < 4; ++d)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const double dG = 0.5 * (hj[pidx(b, e)].dd[pidx(a, d)] + hj[pidx(b, d)].dd[pidx(a, e)] -
                                   hj[pidx(d, e)].dd[pidx(a, b)]);
          s += dgi[d][e] * Gam[b][d][e] + gi[d][e] * dG;
        }
      dH[a][b] = -s;
    }
  }
}

// background_state (src/initialdata.jl:1674-1694): h and
// Pi = (sqrt(gamma)/alpha)(d_t g - beta^i d_i g) at (t, x).
template <int SET = BGS_ALL>
TGH_HD void background_state(const Bg& bg, double t, const double x[3], double h[NC], double P[NC]) {
  if constexpr ((SET & BGS_MOVING) != 0) {
    if (bg_boosted(bg)) {  // a moving hole: every value in the oracle's order
      double dtg[NC];
      bg_state_exact(bg, t, x, h, P, dtg);
      return;
    }
  }
  Dual<4> hd[NC];
  bg_offset<Dual<4>, SET>(bg, t, x, hd);
#pragma unroll
  for (int n = 0; n < NC; ++n) h[n] = hd[n].v;
  Coef c;
  metric_quantities(h, c);
  const double s = c.sqrtg / c.alpha;
#pragma unroll
  for (int n = 0; n < NC; ++n)
    P[n] = s * (hd[n].d[0] - c.beta[0] * hd[n].d[1] - c.beta[1] * hd[n].d[2] - c.beta[2] * hd[n].d[3]);
}

// SplitMix64 point draw in [0, 1) (src/initialdata.jl:2004-2013).
TGH_HD double point_draw(uint64_t base, const double x[3], int v) {
  uint64_t z = base;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    uint64_t cc = i < 3 ? (uint64_t)__double_as_longlong(x[i]) : (uint64_t)v;
    z = (z ^ cc) + 0x9e3779b97f4a7c15ULL;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    z ^= z >> 31;
  }
  return (double)(z >> 11) * 0x1.0p-53;
}

}  // namespace tgh
