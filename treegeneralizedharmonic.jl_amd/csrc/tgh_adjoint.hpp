// The adjoint of the fused right-hand side: the vector-Jacobian products of
// the pointwise algebra (the metric quantities, the lean source, the metric
// divergences, the closed-form Kerr-Schild gauge source, the layer), written
// by hand in the reverse order of the forward code (tgh_device.hpp,
// tgh_kernels.hpp), and the kernels that apply them on the brick layout:
//
//   k_adj_point   one evolved brick point per thread: the stage input's
//                 stencils and the cotangent of F at the point give the
//                 point's local cotangents -- of the point values (h, Pi), of
//                 the first derivatives of h and Pi, of the second derivatives
//                 of h and of the Kreiss-Oliger sums -- written to scratch
//                 fields, and the parameters' cotangent, folded per workgroup;
//   k_adj_gather  one evolved brick point per thread: the transposed stencils
//                 (the forward offsets reversed) gather the scratch fields into
//                 the cotangent of the stage input, and the RK4 combine's
//                 transpose forms the next stage's cotangent.
//
// The parameters differentiated are the constraint damping's Gaussian rates
// (gamma0_near, gamma0_far), gamma2, and the layer's rho_max and rho ramp
// fraction (the interior layer, src/interior.jl:178-238; the damping,
// src/gauge.jl:1094-1156). The reference has no adjoint: it differentiates
// nothing; the oracle's dual-number tangent of the same right-hand side is
// what pins these transposes (<lambda, J v> = <J^T lambda, v>).
#pragma once
#include "tgh_kernels.hpp"

namespace tgh {

enum { ADJ_NPAR = 5 };  // gamma0_near, gamma0_far, gamma2, rho_max, rho_ramp

// The cotangent of the coefficient set.
struct CoefBar {
  double gu[NC], alpha, sqrtg, beta[3], gamu[6];
};
TGH_HD void coefbar_zero(CoefBar& c) {
#pragma unroll
  for (int n = 0; n < NC; ++n) c.gu[n] = 0.0;
  c.alpha = c.sqrtg = 0.0;
#pragma unroll
  for (int i = 0; i < 3; ++i) c.beta[i] = 0.0;
#pragma unroll
  for (int n = 0; n < 6; ++n) c.gamu[n] = 0.0;
}

// h-bar += (d Coef / d h)^T cbar: metric_quantities (tgh_device.hpp) run
// forward for its intermediate values, then reversed statement by statement
// (gu, beta and gamma^{ij} from gi, b and 1/alpha^2; alpha^2 from h_tt, b and
// the shift; b from gi; gi = cof / det; det(I + H), whose derivative in h_ii
// is the cofactor cof_ii and in h_ij (i != j) twice cof_ij, since h_ij and
// h_ji are one packed entry).
TGH_HD void metric_quantities_vjp(const double h[NC], const CoefBar& cb, double hbar[NC]) {
  const double hxx = h[4], hxy = h[5], hxz = h[6], hyy = h[7], hyz = h[8], hzz = h[9];
  const double gxx = 1.0 + hxx, gyy = 1.0 + hyy, gzz = 1.0 + hzz;
  const double e1 = hxx + hyy + hzz;
  const double e2 = (hxx * hyy - hxy * hxy) + (hxx * hzz - hxz * hxz) + (hyy * hzz - hyz * hyz);
  const double e3 = hxx * (hyy * hzz - hyz * hyz) - hxy * (hxy * hzz - hyz * hxz) +
                    hxz * (hxy * hyz - hyy * hxz);
  const double detg = 1.0 + (e1 + (e2 + e3));
  const double sqrtg = sqrt(detg);
  const double idet = 1.0 / detg;
  const double cof[6] = {gyy * gzz - hyz * hyz, hxz * hyz - hxy * gzz, hxy * hyz - hxz * gyy,
                         gxx * gzz - hxz * hxz, hxy * hxz - gxx * hyz, gxx * gyy - hxy * hxy};
  double gi[6];
#pragma unroll
  for (int n = 0; n < 6; ++n) gi[n] = cof[n] * idet;
  const double bl[3] = {h[1], h[2], h[3]};
  // gi's packed index of the spatial pair (i, j), and the pair of packed n
  const int gidx[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
  const int pi_[6] = {0, 0, 0, 1, 1, 2}, pj_[6] = {0, 1, 2, 1, 2, 2};
  double b[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) b[i] = gi[gidx[i][0]] * bl[0] + gi[gidx[i][1]] * bl[1] + gi[gidx[i][2]] * bl[2];
  const double a2 = (1.0 - h[0]) + (b[0] * bl[0] + b[1] * bl[1] + b[2] * bl[2]);
  const double alpha = sqrt(a2);
  const double ia2 = 1.0 / a2;
  // gu[0] = -1/a2, gu[1+i] = b_i / a2, gu[4+n] = gi[n] - b_i b_j / a2; beta = b; gamu = gi
  double gib[6], bb[3];
  double ia2b = -cb.gu[0];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    bb[i] = cb.beta[i] + cb.gu[1 + i] * ia2;
    ia2b += cb.gu[1 + i] * b[i];
  }
#pragma unroll
  for (int n = 0; n < 6; ++n) {
    const int i = pi_[n], j = pj_[n];
    const double g = cb.gu[4 + n];
    gib[n] = cb.gamu[n] + g;
    ia2b -= g * b[i] * b[j];
    bb[i] -= g * b[j] * ia2;
    bb[j] -= g * b[i] * ia2;
  }
  // 1/a2 and alpha = sqrt(a2); a2 = (1 - h_tt) + b . bl
  const double a2b = cb.alpha * 0.5 / alpha - ia2b * ia2 * ia2;
  hbar[0] -= a2b;
  double blb[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    bb[i] += a2b * bl[i];
    blb[i] = a2b * b[i];
  }
  // b_i = sum_j gi(ij) bl_j
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      gib[gidx[i][j]] += bb[i] * bl[j];
      blb[j] += bb[i] * gi[gidx[i][j]];
    }
#pragma unroll
  for (int i = 0; i < 3; ++i) hbar[1 + i] += blb[i];
  // gi = cof / det; sqrt(det)
  double cofb[6], idetb = 0.0;
#pragma unroll
  for (int n = 0; n < 6; ++n) {
    cofb[n] = gib[n] * idet;
    idetb += gib[n] * cof[n];
  }
  const double detgb = cb.sqrtg * 0.5 / sqrtg - idetb * idet * idet;
  double hxxb = detgb * cof[0], hxyb = 2.0 * detgb * cof[1], hxzb = 2.0 * detgb * cof[2];
  double hyyb = detgb * cof[3], hyzb = 2.0 * detgb * cof[4], hzzb = detgb * cof[5];
  // the cofactors
  hyyb += cofb[0] * gzz;
  hzzb += cofb[0] * gyy;
  hyzb -= 2.0 * cofb[0] * hyz;
  hxzb += cofb[1] * hyz;
  hyzb += cofb[1] * hxz;
  hxyb -= cofb[1] * gzz;
  hzzb -= cofb[1] * hxy;
  hxyb += cofb[2] * hyz;
  hyzb += cofb[2] * hxy;
  hxzb -= cofb[2] * gyy;
  hyyb -= cofb[2] * hxz;
  hxxb += cofb[3] * gzz;
  hzzb += cofb[3] * gxx;
  hxzb -= 2.0 * cofb[3] * hxz;
  hxyb += cofb[4] * hxz;
  hxzb += cofb[4] * hxy;
  hxxb -= cofb[4] * hyz;
  hyzb -= cofb[4] * gxx;
  hxxb += cofb[5] * gyy;
  hyyb += cofb[5] * gxx;
  hxyb -= 2.0 * cofb[5] * hxy;
  hbar[4] += hxxb;
  hbar[5] += hxyb;
  hbar[6] += hxzb;
  hbar[7] += hyyb;
  hbar[8] += hyzb;
  hbar[9] += hzzb;
}

// The reverse of gh_source<HAS_H> (the lean source, tgh_device.hpp): given
// ob = the cotangent of out, accumulate the cotangents of the coefficient
// set (cb), of h (hb: through eta + h in the damping term), of D (Db), of
// the lowered gauge source (Hb) and of g0 and g2, and return Sb, the
// cotangent of the sum S (out = -alpha sqrt(gamma) S), from which that of
// d_c H_a follows: dH_ab enters S_ab and S_ba with a minus sign, so its
// cotangent is -Sb_ab off the diagonal and -2 Sb_aa on it. The forward's
// intermediate values are recomputed here first; S itself is never held:
// only its contraction with ob (for the cotangent of alpha sqrt(gamma)) and
// with Sb the damping term's (for g0's) are accumulated as it is formed.
template <bool HAS_H>
TGH_HD void gh_source_vjp(const Coef& c, const double h[NC], const double D[4][NC], const double Hl[4],
                          const double dH[4][4], double g0, double g2, const double ob[NC], CoefBar& cb,
                          double hb[NC], double Db[4][NC], double Hb[4], double Sb[NC], double& g0b,
                          double& g2b) {
  const double* Gp = c.gu;
  const double as = c.alpha * c.sqrtg;
#pragma unroll
  for (int n = 0; n < NC; ++n) Sb[n] = -as * ob[n];
  double sob = 0.0;  // sum_n ob[n] S[n]
  // ---- forward: S's contraction, the Christoffels, Gup, and the damping's Cup, Cl, tC ----
#pragma unroll 1
  for (int a = 0; a < 4; ++a) {
    double E[4][4];
    for (int r = 0; r < 4; ++r)
      for (int col = 0; col < 4; ++col) {
        double s = Gp[pidx(r, 0)] * D[a][pidx(0, col)];
        for (int x = 1; x < 4; ++x) s += Gp[pidx(r, x)] * D[a][pidx(x, col)];
        E[r][col] = s;
      }
    double Cuu[NC];
    for (int n = 0; n < NC; ++n) {
      const int i = prow(n), j = pcol(n);
      double s = E[i][0] * Gp[pidx(0, j)];
      for (int y = 1; y < 4; ++y) s += E[i][y] * Gp[pidx(y, j)];
      Cuu[n] = s;
    }
    for (int b = 0; b < 4; ++b) {
      double s = 0.0;
      for (int m = 0; m < 16; ++m) s += Cuu[pidx(m % 4, m / 4)] * D[m % 4][pidx(m / 4, b)];
      // C2 + C2^T: S_ab and S_ba each take s, the diagonal twice
      sob += s * ob[pidx(a, b)] * (a == b ? 2.0 : 1.0);
    }
  }
  double Gam[4][NC];
  for (int n = 0; n < NC; ++n) {
    const int i = prow(n), j = pcol(n);
    double low[4];
    for (int x = 0; x < 4; ++x) low[x] = 0.5 * (D[i][pidx(x, j)] + D[j][pidx(x, i)] - D[x][n]);
    for (int a = 0; a < 4; ++a) {
      double s = Gp[pidx(a, 0)] * low[0];
      for (int x = 1; x < 4; ++x) s += Gp[pidx(a, x)] * low[x];
      Gam[a][n] = s;
    }
  }
  double Gup[4];
  for (int cc = 0; cc < 4; ++cc) {
    double s = 0.0;
    for (int n = 0; n < NC; ++n) s += (prow(n) == pcol(n) ? 1.0 : 2.0) * Gp[n] * Gam[cc][n];
    Gup[cc] = s;
  }
  for (int n = 0; n < NC; ++n) {
    const int a = prow(n), b = pcol(n);
    double t2 = 0.0;
    for (int m = 0; m < 16; ++m) t2 += Gam[m % 4][pidx(m / 4, a)] * Gam[m / 4][pidx(m % 4, b)];
    const double t5 = Gup[0] * D[0][n] + Gup[1] * D[1][n] + Gup[2] * D[2][n] + Gup[3] * D[3][n];
    double s = -2.0 * t2 - t5;
    if (HAS_H) {
      const double t4 = Gam[0][n] * Hl[0] + Gam[1][n] * Hl[1] + Gam[2][n] * Hl[2] + Gam[3][n] * Hl[3];
      s += 2.0 * t4 - (dH[a][b] + dH[b][a]);
    }
    sob += ob[n] * s;
  }
  double Cup[4], Cl[4], tC = 0.0;
  const double tl = -c.alpha, k1 = 1.0 + g2;
  for (int cc = 0; cc < 4; ++cc) {
    double gh = 0.0;
    if (HAS_H) gh = Gp[pidx(cc, 0)] * Hl[0] + Gp[pidx(cc, 1)] * Hl[1] + Gp[pidx(cc, 2)] * Hl[2] + Gp[pidx(cc, 3)] * Hl[3];
    Cup[cc] = Gup[cc] + gh;
  }
  for (int cc = 0; cc < 4; ++cc) {
    double s = 0.0;
    for (int x = 0; x < 4; ++x) s += (eta(cc, x) + h[pidx(cc, x)]) * Cup[x];
    Cl[cc] = s;
  }
  tC = -c.alpha * (Gp[0] * Cl[0] + Gp[1] * Cl[1] + Gp[2] * Cl[2] + Gp[3] * Cl[3]);
  for (int n = 0; n < NC; ++n) {
    const int a = prow(n), b = pcol(n);
    double d = 0.0;
    if (a == 0) d += tl * Cl[b];
    if (b == 0) d += tl * Cl[a];
    const double dmp = d - k1 * (eta(a, b) + h[n]) * tC;
    sob += ob[n] * (g0 * dmp);
    g0b += Sb[n] * dmp;
  }
  // ---- reverse ----
  // (4) out = -alpha sqrt(gamma) S
  const double asb = -sob;
  cb.alpha += asb * c.sqrtg;
  cb.sqrtg += asb * c.alpha;
  // (3) the damping term, g0 dmp
  double Gupb[4] = {0.0, 0.0, 0.0, 0.0};
  {
    double Clb[4] = {0.0, 0.0, 0.0, 0.0}, tlb = 0.0, tCb = 0.0;
    for (int n = 0; n < NC; ++n) {
      const int a = prow(n), b = pcol(n);
      g2b += Sb[n] * g0 * (-(eta(a, b) + h[n]) * tC);
      const double sg = g0 * Sb[n];
      if (a == 0) {
        Clb[b] += sg * tl;
        tlb += sg * Cl[b];
      }
      if (b == 0) {
        Clb[a] += sg * tl;
        tlb += sg * Cl[a];
      }
      tCb -= sg * k1 * (eta(a, b) + h[n]);
      hb[n] -= sg * k1 * tC;
    }
    // tC = -alpha sum_x Gp[0x] Cl[x]; tl = -alpha
    const double gcl = Gp[0] * Cl[0] + Gp[1] * Cl[1] + Gp[2] * Cl[2] + Gp[3] * Cl[3];
    cb.alpha -= tCb * gcl + tlb;
    for (int x = 0; x < 4; ++x) {
      cb.gu[pidx(0, x)] -= c.alpha * tCb * Cl[x];
      Clb[x] -= c.alpha * tCb * Gp[pidx(0, x)];
    }
    // Cl[c] = sum_x (eta + h)[cx] Cup[x]
    double Cupb[4] = {0.0, 0.0, 0.0, 0.0};
    for (int cc = 0; cc < 4; ++cc)
      for (int x = 0; x < 4; ++x) {
        hb[pidx(cc, x)] += Clb[cc] * Cup[x];
        Cupb[x] += Clb[cc] * (eta(cc, x) + h[pidx(cc, x)]);
      }
    // Cup[c] = Gup[c] + sum_x Gp[cx] Hl[x]
    for (int cc = 0; cc < 4; ++cc) {
      Gupb[cc] += Cupb[cc];
      if (HAS_H)
        for (int x = 0; x < 4; ++x) {
          cb.gu[pidx(cc, x)] += Cupb[cc] * Hl[x];
          Hb[x] += Cupb[cc] * Gp[pidx(cc, x)];
        }
    }
  }
  // (2) -2 t2 - t5 + 2 t4 - (dH_ab + dH_ba)
  double Gamb[4][NC];
  for (int a = 0; a < 4; ++a)
    for (int n = 0; n < NC; ++n) Gamb[a][n] = 0.0;
  for (int n = 0; n < NC; ++n) {
    const int a = prow(n), b = pcol(n);
    const double sb = Sb[n];
    const double t2b = -2.0 * sb, t5b = -sb;
    for (int cc = 0; cc < 4; ++cc) {
      Gupb[cc] += t5b * D[cc][n];
      Db[cc][n] += t5b * Gup[cc];
    }
    if (HAS_H) {
      const double t4b = 2.0 * sb;
      for (int cc = 0; cc < 4; ++cc) {
        Gamb[cc][n] += t4b * Hl[cc];
        Hb[cc] += t4b * Gam[cc][n];
      }
    }
    for (int m = 0; m < 16; ++m) {
      const int p1 = pidx(m / 4, a), p2 = pidx(m % 4, b);
      Gamb[m % 4][p1] += t2b * Gam[m / 4][p2];
      Gamb[m / 4][p2] += t2b * Gam[m % 4][p1];
    }
  }
  // Gup[c] = sum_n mult(n) Gp[n] Gam[c][n]
  for (int cc = 0; cc < 4; ++cc)
    for (int n = 0; n < NC; ++n) {
      const double mult = prow(n) == pcol(n) ? 1.0 : 2.0;
      cb.gu[n] += Gupb[cc] * mult * Gam[cc][n];
      Gamb[cc][n] += Gupb[cc] * mult * Gp[n];
    }
  // Gam[a][n] = sum_x Gp[ax] low_x(n), low_x(n) = (D_i[xj] + D_j[xi] - D_x[n]) / 2
  for (int n = 0; n < NC; ++n) {
    const int i = prow(n), j = pcol(n);
    double low[4];
    for (int x = 0; x < 4; ++x) low[x] = 0.5 * (D[i][pidx(x, j)] + D[j][pidx(x, i)] - D[x][n]);
    for (int x = 0; x < 4; ++x) {
      double lb = 0.0;
      for (int a = 0; a < 4; ++a) {
        cb.gu[pidx(a, x)] += Gamb[a][n] * low[x];
        lb += Gamb[a][n] * Gp[pidx(a, x)];
      }
      Db[i][pidx(x, j)] += 0.5 * lb;
      Db[j][pidx(x, i)] += 0.5 * lb;
      Db[x][n] -= 0.5 * lb;
    }
  }
  // (1) C2 + C2^T, first index a at a time: E = G D_a, Cuu = E G, C2 = Cuu : D
#pragma unroll 1
  for (int a = 0; a < 4; ++a) {
    double C2b[4] = {0.0, 0.0, 0.0, 0.0};
    for (int n = 0; n < NC; ++n) {
      if (prow(n) == a) C2b[pcol(n)] += Sb[n];
      if (pcol(n) == a) C2b[prow(n)] += Sb[n];
    }
    double E[4][4];
    for (int r = 0; r < 4; ++r)
      for (int col = 0; col < 4; ++col) {
        double s = Gp[pidx(r, 0)] * D[a][pidx(0, col)];
        for (int x = 1; x < 4; ++x) s += Gp[pidx(r, x)] * D[a][pidx(x, col)];
        E[r][col] = s;
      }
    double Cuu[NC], Cuub[NC];
    for (int n = 0; n < NC; ++n) {
      const int i = prow(n), j = pcol(n);
      double s = E[i][0] * Gp[pidx(0, j)];
      for (int y = 1; y < 4; ++y) s += E[i][y] * Gp[pidx(y, j)];
      Cuu[n] = s;
      Cuub[n] = 0.0;
    }
    for (int b = 0; b < 4; ++b)
      for (int m = 0; m < 16; ++m) {
        const int mu = m % 4, nu = m / 4;
        Cuub[pidx(mu, nu)] += C2b[b] * D[mu][pidx(nu, b)];
        Db[mu][pidx(nu, b)] += C2b[b] * Cuu[pidx(mu, nu)];
      }
    double Eb[4][4];
    for (int r = 0; r < 4; ++r)
      for (int col = 0; col < 4; ++col) Eb[r][col] = 0.0;
    for (int n = 0; n < NC; ++n) {
      const int i = prow(n), j = pcol(n);
      for (int y = 0; y < 4; ++y) {
        Eb[i][y] += Cuub[n] * Gp[pidx(y, j)];
        cb.gu[pidx(y, j)] += Cuub[n] * E[i][y];
      }
    }
    for (int r = 0; r < 4; ++r)
      for (int col = 0; col < 4; ++col)
        for (int x = 0; x < 4; ++x) {
          cb.gu[pidx(r, x)] += Eb[r][col] * D[a][pidx(x, col)];
          Db[a][pidx(x, col)] += Eb[r][col] * Gp[pidx(r, x)];
        }
  }
}

// The reverse of metric_divergences (tgh_device.hpp): given the cotangents
// of d_i beta^i (divbb) and of d_i A^{ij} (divAb), accumulate those of the
// coefficient set and of the first derivatives dh[i] = d_{i+1} h.
TGH_HD void metric_divergences_vjp(const Coef& c, const double dh[3][NC], double divbb, const double divAb[3],
                                   CoefBar& cb, double dhb[3][NC]) {
  const double* Gp = c.gu;
  double r1[3][4], ri[3][4], v0s[3][4], vis[3][4];
  for (int i = 0; i < 3; ++i) {
    for (int d = 0; d < 4; ++d) {
      double s0 = 0.0, si = 0.0;
      for (int cc = 0; cc < 4; ++cc) {
        s0 += Gp[pidx(0, cc)] * dh[i][pidx(cc, d)];
        si += Gp[pidx(i + 1, cc)] * dh[i][pidx(cc, d)];
      }
      v0s[i][d] = s0;
      vis[i][d] = si;
    }
    for (int b = 0; b < 4; ++b) {
      double s0 = 0.0, si = 0.0;
      for (int d = 0; d < 4; ++d) {
        s0 += v0s[i][d] * Gp[pidx(d, b)];
        si += vis[i][d] * Gp[pidx(d, b)];
      }
      r1[i][b] = -s0;
      ri[i][b] = -si;
    }
  }
  const double gutt = Gp[0];
  const double a3 = c.alpha * c.alpha * c.alpha;
  double da[3], ds[3], dss[3];
  for (int i = 0; i < 3; ++i) {
    da[i] = 0.5 * a3 * r1[i][0];
    double s = 0.0;
    for (int m = 0; m < 9; ++m) s += c.gamu[sidx3(m % 3, m / 3)] * dh[i][pidx(m % 3 + 1, m / 3 + 1)];
    dss[i] = s;
    ds[i] = 0.5 * c.sqrtg * s;
  }
  const double as = c.alpha * c.sqrtg;
  double r1b[3][4], rib[3][4], dab[3] = {0.0, 0.0, 0.0}, dsb[3] = {0.0, 0.0, 0.0}, asb = 0.0;
  for (int i = 0; i < 3; ++i)
    for (int b = 0; b < 4; ++b) r1b[i][b] = rib[i][b] = 0.0;
  // divA[k] = sum_i (da_i sqrt(g) + alpha ds_i) gamu_ik + as (ri[i][k+1] + beta_k r1[i][i+1]
  //           + beta_i r1[i][k+1] + beta_i beta_k r1[i][0])
  for (int k = 0; k < 3; ++k)
    for (int i = 0; i < 3; ++i) {
      const double gik = c.gamu[sidx3(i, k)];
      const double e = da[i] * c.sqrtg + c.alpha * ds[i];
      const double T = ri[i][k + 1] + c.beta[k] * r1[i][i + 1] + c.beta[i] * r1[i][k + 1] +
                       c.beta[i] * c.beta[k] * r1[i][0];
      const double eb = divAb[k] * gik;
      dab[i] += eb * c.sqrtg;
      cb.sqrtg += eb * da[i];
      cb.alpha += eb * ds[i];
      dsb[i] += eb * c.alpha;
      cb.gamu[sidx3(i, k)] += divAb[k] * e;
      asb += divAb[k] * T;
      const double Tb = divAb[k] * as;
      rib[i][k + 1] += Tb;
      cb.beta[k] += Tb * r1[i][i + 1];
      r1b[i][i + 1] += Tb * c.beta[k];
      cb.beta[i] += Tb * r1[i][k + 1];
      r1b[i][k + 1] += Tb * c.beta[i];
      cb.beta[i] += Tb * c.beta[k] * r1[i][0];
      cb.beta[k] += Tb * c.beta[i] * r1[i][0];
      r1b[i][0] += Tb * c.beta[i] * c.beta[k];
    }
  cb.alpha += asb * c.sqrtg;
  cb.sqrtg += asb * c.alpha;
  double a3b = 0.0;
  for (int i = 0; i < 3; ++i) {
    // ds_i = sqrt(g)/2 sum_m gamu dh
    cb.sqrtg += dsb[i] * 0.5 * dss[i];
    for (int m = 0; m < 9; ++m) {
      cb.gamu[sidx3(m % 3, m / 3)] += dsb[i] * 0.5 * c.sqrtg * dh[i][pidx(m % 3 + 1, m / 3 + 1)];
      dhb[i][pidx(m % 3 + 1, m / 3 + 1)] += dsb[i] * 0.5 * c.sqrtg * c.gamu[sidx3(m % 3, m / 3)];
    }
    // da_i = a3/2 r1[i][0]
    a3b += dab[i] * 0.5 * r1[i][0];
    r1b[i][0] += dab[i] * 0.5 * a3;
  }
  cb.alpha += a3b * 3.0 * c.alpha * c.alpha;
  // divb = -Q / gutt, Q = sum_i (r1[i][i+1] + beta_i r1[i][0])
  {
    double Q = 0.0;
    for (int i = 0; i < 3; ++i) Q += r1[i][i + 1] + c.beta[i] * r1[i][0];
    const double Qb = -divbb / gutt;
    cb.gu[0] += divbb * Q / (gutt * gutt);
    for (int i = 0; i < 3; ++i) {
      r1b[i][i + 1] += Qb;
      cb.beta[i] += Qb * r1[i][0];
      r1b[i][0] += Qb * c.beta[i];
    }
  }
  // r1[i][b] = -sum_d v0[d] Gp[db], v0[d] = sum_c Gp[0c] dh[i][cd]; ri with row i+1
  for (int i = 0; i < 3; ++i) {
    double v0b[4] = {0.0, 0.0, 0.0, 0.0}, vib[4] = {0.0, 0.0, 0.0, 0.0};
    for (int b = 0; b < 4; ++b)
      for (int d = 0; d < 4; ++d) {
        v0b[d] -= r1b[i][b] * Gp[pidx(d, b)];
        vib[d] -= rib[i][b] * Gp[pidx(d, b)];
        cb.gu[pidx(d, b)] -= r1b[i][b] * v0s[i][d] + rib[i][b] * vis[i][d];
      }
    for (int d = 0; d < 4; ++d)
      for (int cc = 0; cc < 4; ++cc) {
        cb.gu[pidx(0, cc)] += v0b[d] * dh[i][pidx(cc, d)];
        cb.gu[pidx(i + 1, cc)] += vib[d] * dh[i][pidx(cc, d)];
        dhb[i][pidx(cc, d)] += v0b[d] * Gp[pidx(0, cc)] + vib[d] * Gp[pidx(i + 1, cc)];
      }
  }
}

// The reverse of ks_source (the closed-form Kerr-Schild gauge source of a
// hole at rest, spin along z): the cotangents of h (its tt, tx, ty, tz
// components) and of D[c] (the same components) from those of H_a and d_c H_a.
TGH_HD void ks_source_vjp(double M, double a, const double h[NC], const double D[4][NC], const double Hb[4],
                          const double Sb[NC], double hb[NC], double Db[4][NC]) {
  const double w[4] = {h[0], h[1], h[2], h[3]};
  const double K = w[0];
  const double sw = a * w[3];
  const double rad = M * M - sw * sw;
  const double root = rad > 0.0 ? sqrt(rad) : 0.0;
  const double Dn = M + root;
  const double iD = 1.0 / Dn;
  double wb[4] = {0.0, 0.0, 0.0, 0.0}, Kb = 0.0, iDb = 0.0, swb = 0.0, rootb = 0.0;
  // Hl[q] = -K iD w_q
  for (int q = 0; q < 4; ++q) {
    wb[q] -= Hb[q] * K * iD;
    Kb -= Hb[q] * iD * w[q];
    iDb -= Hb[q] * K * w[q];
  }
  // dH[c][q] = -(dK w_q + K dw_q) iD + K w_q dD iD^2, dK = D[c][0], dD = -sw a D[c][3] / root
  for (int cc = 0; cc < 4; ++cc) {
    const double dK = D[cc][0];
    const double dD = root > 0.0 ? -sw * (a * D[cc][3]) / root : 0.0;
    double dKb = 0.0, dDb = 0.0;
    for (int q = 0; q < 4; ++q) {
      const double g = -(cc == q ? 2.0 : 1.0) * Sb[pidx(cc, q)];  // the cotangent of d_cc H_q
      const double dwq = D[cc][q];
      dKb -= g * w[q] * iD;
      wb[q] += -g * dK * iD + g * K * dD * iD * iD;
      Kb += -g * dwq * iD + g * w[q] * dD * iD * iD;
      Db[cc][q] -= g * K * iD;
      dDb += g * K * w[q] * iD * iD;
      iDb += -g * (dK * w[q] + K * dwq) + 2.0 * g * K * w[q] * dD * iD;
    }
    Db[cc][0] += dKb;
    if (root > 0.0) {
      Db[cc][3] += dDb * (-sw * a / root);
      swb += dDb * (-a * D[cc][3] / root);
      rootb += dDb * (sw * a * D[cc][3] / (root * root));
    }
  }
  rootb += -iDb * iD * iD;
  if (root > 0.0) swb += rootb * (-sw / root);
  wb[3] += a * swb;
  wb[0] += Kb;
  for (int q = 0; q < 4; ++q) hb[q] += wb[q];
}

// ---------------------------------------------------------------------------
// The scratch fields of one VJP (brick layout, zero on the frame): per
// variable v the cotangents of the point value h_v (AH) and Pi_v (AP), of the
// first derivatives d_i h_v (BH) and d_i Pi_v (BP) -- the scaled derivatives
// (1/h) st_d1 --, of the second derivatives of h_v in A's order xx, xy, xz,
// yy, yz, zz (CH, the mixed ones carrying the forward's factor 2), and of the
// Kreiss-Oliger sums of h_v (DH) and Pi_v (DP) with the amplitude over h
// folded in.
enum {
  AF_AH = 0,
  AF_AP = 10,
  AF_BH = 20,
  AF_BP = 50,
  AF_CH = 80,
  AF_DH = 140,
  AF_DP = 150,
  AF_N = 160
};

struct AdjPointArgs {
  Brick k;
  CaseP c;
  double t;
  const double* __restrict__ W;     // the stage input (brick, frame filled)
  const double* __restrict__ Hsrc;  // the sampled gauge source (brick) or null
  const double* __restrict__ kb;    // the cotangent of this stage's F (brick)
  double* __restrict__ F;           // AF_N scratch fields (brick)
  double* __restrict__ thp;         // ADJ_NPAR accumulators per logical workgroup
};

struct AdjGatherArgs {
  Brick k;
  const double* __restrict__ F;
  const double* __restrict__ lam;
  double* __restrict__ lacc;
  double* __restrict__ kbn;
  double* __restrict__ ybar;  // the VJP alone (tgh_vjp_rhs), or null
  double cl, cy;
  int first, last;
};
struct AdjInitArgs {
  Brick k;
  CaseP c;
  double t, J, nw, cl;
  const double* __restrict__ u;
  double* __restrict__ lam;
  double* __restrict__ kbn;
};
struct AdjCopiesArgs {
  Geom g;
  Brick k;
  const double* __restrict__ src;  // block layout
  double* __restrict__ dst;        // brick layout
};
// One evolved brick point: F(u) recomputed in the forward's order, then its
// transpose applied to the point's cotangent (lh for d_t h, lP for d_t Pi).
template <int Q, int GAUGE, int HOLE>
__device__ __forceinline__ void adj_point(const AdjPointArgs& a, int x, int y, int z, double th[ADJ_NPAR]) {
  const Brick& k = a.k;
  const CaseP& c = a.c;
  const long off = brick_index(k, x, y, z), fs = k.fs;
  const int sy = (int)k.py, sz = (int)k.pz;
  const double xp[3] = {k.lo[0] + (double)x * k.h, k.lo[1] + (double)y * k.h, k.lo[2] + (double)z * k.h};
  double r = 0.0;
  if (frozen_core<HOLE>(c, a.t, xp, r)) {
    // a hole at rest: du = 0 in the core, whatever u holds
#pragma unroll 1
    for (int f = 0; f < AF_N; ++f) a.F[off + f * fs] = 0.0;
    return;
  }
  const double* w0 = a.W + off;
  const double ih = k.inv_h, ih2 = ih * ih;
  // ---- forward ----
  double hv[NC], Pv[NC], D[4][NC];
#pragma unroll
  for (int v = 0; v < NC; ++v) {
    const double* pv = w0 + v * fs;
    hv[v] = pv[0];
    Pv[v] = w0[(NC + v) * fs];
    D[1][v] = ih * st_d1<Q>(pv, 1);
    D[2][v] = ih * st_d1<Q>(pv, sy);
    D[3][v] = ih * st_d1<Q>(pv, sz);
  }
  Coef cf;
  metric_quantities(hv, cf);
  const double a_div = cf.alpha / cf.sqrtg;
  const double epsh = dissipation_rate(c, a.t, xp) * ih;
  for (int v = 0; v < NC; ++v) {
    const double* pv = w0 + v * fs;
    double s = cf.beta[0] * D[1][v] + cf.beta[1] * D[2][v] + cf.beta[2] * D[3][v] + a_div * Pv[v];
    if (c.has_diss) s += epsh * (st_ko<Q>(pv, 1) + st_ko<Q>(pv, sy) + st_ko<Q>(pv, sz));
    D[0][v] = s;
  }
  double Hl[4] = {0.0, 0.0, 0.0, 0.0}, dH[4][4];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) dH[i][j] = 0.0;
  if (GAUGE == GAUGE_SAMPLED) {
    const double* Hp = a.Hsrc + off;
    for (int q = 0; q < 4; ++q) Hl[q] = Hp[q * fs];
    for (int n = 0; n < 16; ++n) dH[n % 4][n / 4] = Hp[(4 + n) * fs];
  } else if (GAUGE == GAUGE_KS) {
    ks_source(c.bg.M, c.bg.a, hv, D, Hl, dH);
  }
  double e_g = 0.0;  // the Gaussian's exp(-r^2 / 2 w^2)
  double g0 = c.g0;
  if (c.damping != 0) {
    double cc[3];
    center_at(c, a.t, cc);
    const double d1 = xp[0] - cc[0], d2 = xp[1] - cc[1], d3 = xp[2] - cc[2];
    e_g = exp(-(d1 * d1 + d2 * d2 + d3 * d3) / (2.0 * c.g0_width * c.g0_width));
    g0 = damping_rate(c, a.t, xp);
  }
  PiCoef pc;
  metric_divergences(cf, &D[1], pc.divb, pc.divA);
  const double as = cf.alpha * cf.sqrtg;
  double A[6];
  for (int n = 0; n < 6; ++n) A[n] = as * cf.gamu[n];
  // the layer
  double wgt = 1.0, rho = 0.0, sl = 0.0, ss = 0.0, arg = 0.0;
  const bool layer = HOLE && r < c.r1;
  double ue[20];
  if (layer) {
    sl = (r - c.r0) / (c.r1 - c.r0);
    wgt = smoothstep(sl / c.w_ramp);
    arg = (1.0 - sl) / c.rho_ramp;
    ss = smoothstep(arg);
    rho = c.interior == INT_FROZEN ? 0.0 : c.rho_max * ss;
    case_state<bg_set_for(GAUGE)>(c, a.t, xp, ue);
  }
  // ---- reverse ----
  double lh[NC], pb[NC], msb[NC];
  double aH[NC], aP[NC], Db[4][NC];
  double bb[3] = {0.0, 0.0, 0.0}, divbb = 0.0, divAb[3] = {0.0, 0.0, 0.0}, Ab[6];
  double rhob = 0.0;
  for (int n = 0; n < 6; ++n) Ab[n] = 0.0;
  for (int v = 0; v < NC; ++v) {
    lh[v] = a.kb[off + v * fs];
    const double lP = a.kb[off + (NC + v) * fs];
    pb[v] = wgt * lP;  // the principal part's cotangent
    msb[v] = layer ? wgt * lP : lP;
    aH[v] = 0.0;
    aP[v] = 0.0;
    if (layer) {
      aP[v] -= rho * lP;
      rhob -= lP * (Pv[v] - ue[NC + v]);
      aH[v] -= rho * lh[v];
      rhob -= lh[v] * (hv[v] - ue[v]);
    }
    for (int i = 0; i < 4; ++i) Db[i][v] = 0.0;
  }
  // the principal part of the Pi components: beta^i d_i Pi + divb Pi +
  // divA^j d_j h + A^{ij} d_i d_j h + KO(Pi)
#pragma unroll 1
  for (int v = 0; v < NC; ++v) {
    const double* ph = w0 + v * fs;
    const double* pP = w0 + (NC + v) * fs;
    const double p_ = pb[v];
    const double dP[3] = {ih * st_d1<Q>(pP, 1), ih * st_d1<Q>(pP, sy), ih * st_d1<Q>(pP, sz)};
    for (int i = 0; i < 3; ++i) {
      bb[i] += p_ * dP[i];
      a.F[off + (AF_BP + 10 * i + v) * fs] = p_ * cf.beta[i];
      divAb[i] += p_ * D[i + 1][v];
      Db[i + 1][v] += p_ * pc.divA[i];
    }
    divbb += p_ * Pv[v];
    aP[v] += p_ * pc.divb;
    const double d2[6] = {ih2 * st_d2<Q>(ph, 1), 2.0 * ih2 * st_mixed<Q>(ph, 1, sy), 2.0 * ih2 * st_mixed<Q>(ph, 1, sz),
                          ih2 * st_d2<Q>(ph, sy), 2.0 * ih2 * st_mixed<Q>(ph, sy, sz), ih2 * st_d2<Q>(ph, sz)};
    for (int n = 0; n < 6; ++n) {
      Ab[n] += p_ * d2[n];
      const double mult = (n == 1 || n == 2 || n == 4) ? 2.0 : 1.0;
      a.F[off + (AF_CH + 10 * n + v) * fs] = p_ * A[n] * mult;
    }
    a.F[off + (AF_DP + v) * fs] = c.has_diss ? p_ * epsh : 0.0;
  }
  CoefBar cb;
  coefbar_zero(cb);
  // the source: msrc = gh_source(cf, h, D, H, dH, g0, g2)
  double Hb[4] = {0.0, 0.0, 0.0, 0.0}, Sb[NC], g0b = 0.0, g2b = 0.0;
  gh_source_vjp<gauge_has_source(GAUGE)>(cf, hv, D, Hl, dH, g0, c.g2, msb, cb, aH, Db, Hb, Sb, g0b, g2b);
  if (GAUGE == GAUGE_KS) ks_source_vjp(c.bg.M, c.bg.a, hv, D, Hb, Sb, aH, Db);
  // d_t h = D[0] = beta . d h + (alpha / sqrt g) Pi + KO(h); the output
  // for h is w D[0] - rho (h - h_e)
#pragma unroll 1
  for (int v = 0; v < NC; ++v) {
    const double d0b = Db[0][v] + wgt * lh[v];
    double adb = d0b * Pv[v];
    for (int i = 0; i < 3; ++i) {
      bb[i] += d0b * D[i + 1][v];
      Db[i + 1][v] += d0b * cf.beta[i];
    }
    aP[v] += d0b * a_div;
    a.F[off + (AF_DH + v) * fs] = c.has_diss ? d0b * epsh : 0.0;
    cb.alpha += adb / cf.sqrtg;
    cb.sqrtg -= adb * cf.alpha / (cf.sqrtg * cf.sqrtg);
  }
  // A = alpha sqrt(g) gamma^{ij}; the divergences
  {
    double asb = 0.0;
    for (int n = 0; n < 6; ++n) {
      asb += Ab[n] * cf.gamu[n];
      cb.gamu[n] += Ab[n] * as;
    }
    cb.alpha += asb * cf.sqrtg;
    cb.sqrtg += asb * cf.alpha;
  }
  for (int i = 0; i < 3; ++i) cb.beta[i] += bb[i];
  metric_divergences_vjp(cf, &D[1], divbb, divAb, cb, &Db[1]);
  metric_quantities_vjp(hv, cb, aH);
  for (int v = 0; v < NC; ++v) {
    a.F[off + (AF_AH + v) * fs] = aH[v];
    a.F[off + (AF_AP + v) * fs] = aP[v];
    for (int i = 0; i < 3; ++i) a.F[off + (AF_BH + 10 * i + v) * fs] = Db[i + 1][v];
  }
  // the parameters: gamma0 = far + (near - far) e; rho = rho_max s(arg)
  if (c.damping != 0) {
    th[0] += g0b * e_g;
    th[1] += g0b * (1.0 - e_g);
  }
  th[2] += g2b;
  if (layer && c.interior != INT_FROZEN) {
    th[3] += rhob * ss;
    const double u0 = fmin(fmax(arg, 0.0), 1.0);
    const double dss = (arg > 0.0 && arg < 1.0) ? 30.0 * u0 * u0 * (1.0 - u0) * (1.0 - u0) : 0.0;
    th[4] += rhob * c.rho_max * dss * (-(1.0 - sl) / (c.rho_ramp * c.rho_ramp));
  }
}

#ifdef TGH_ADJOINT_KERNELS  // the kernels: built in the adjoint's translation unit
template <int Q, int GAUGE, int HOLE>
__global__ __launch_bounds__(256) void k_adj_point(AdjPointArgs a) {
  const int total = brick_runs(a.k);
  const int L = wg_logical(total);
  if (L >= total) return;
  double th[ADJ_NPAR] = {0.0, 0.0, 0.0, 0.0, 0.0};
  int x, y, z;
  if (brick_run(a.k, L, x, y, z)) adj_point<Q, GAUGE, HOLE>(a, x, y, z, th);
  __shared__ double sh[ADJ_NPAR][256];
  const int tid = threadIdx.x;
#pragma unroll
  for (int q = 0; q < ADJ_NPAR; ++q) sh[q][tid] = th[q];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (tid < w)
#pragma unroll
      for (int q = 0; q < ADJ_NPAR; ++q) sh[q][tid] += sh[q][tid + w];
    __syncthreads();
  }
  if (tid < ADJ_NPAR) a.thp[(long)L * ADJ_NPAR + tid] += sh[tid][0];
}

// The transposed stencils at one evolved brick point q and the RK4 combine's
// transpose: ybar = the cotangent of the stage input at q; then
//   lacc = (first ? lam : lacc) + ybar,  kbn = cl * lam + cy * ybar
// or, with last (the step's first stage, reversed last), kbn = cl * lacc.
// The brick index of (x, y, z) shifted by -o along axis d, wrapped on a
// periodic axis (its frame holds copies of the points the shift lands on);
// on a Dirichlet axis the frame's scratch fields are zero.
__device__ __forceinline__ long adj_shift(const Brick& k, int x, int y, int z, int d, int o) {
  int c[3] = {x, y, z};
  c[d] -= o;
  if (k.periodic[d]) c[d] = (c[d] % k.n[d] + k.n[d]) % k.n[d];
  return brick_index(k, c[0], c[1], c[2]);
}
__device__ __forceinline__ long adj_shift2(const Brick& k, int x, int y, int z, int d1, int o1, int d2, int o2) {
  int c[3] = {x, y, z};
  c[d1] -= o1;
  c[d2] -= o2;
  if (k.periodic[d1]) c[d1] = (c[d1] % k.n[d1] + k.n[d1]) % k.n[d1];
  if (k.periodic[d2]) c[d2] = (c[d2] % k.n[d2] + k.n[d2]) % k.n[d2];
  return brick_index(k, c[0], c[1], c[2]);
}
template <int Q>
__global__ __launch_bounds__(256) void k_adj_gather(AdjGatherArgs a) {
  const Brick& k = a.k;
  int x, y, z;
  if (!brick_map(k, x, y, z)) return;
  constexpr int R = W<Q>::R;
  const long off = brick_index(k, x, y, z), fs = k.fs;
  const double ih = k.inv_h, ih2 = ih * ih;
  const double* F = a.F;
#pragma unroll 1
  for (int v = 0; v < NC; ++v) {
    double uh = F[off + (AF_AH + v) * fs], uP = F[off + (AF_AP + v) * fs];
    double s1h = 0.0, s1P = 0.0, s2 = 0.0, sk = 0.0, skP = 0.0;
    for (int d = 0; d < 3; ++d) {
      for (int m = 0; m < 2 * R + 1; ++m) {
        const long i = adj_shift(k, x, y, z, d, m - R);
        if (m != R) {
          s1h += W<Q>::d1[m] * F[i + (AF_BH + 10 * d + v) * fs];
          s1P += W<Q>::d1[m] * F[i + (AF_BP + 10 * d + v) * fs];
        }
        s2 += W<Q>::d2[m] * F[i + (AF_CH + 10 * sidx3(d, d) + v) * fs];
      }
      for (int m = 0; m < 2 * R + 3; ++m) {
        const long i = adj_shift(k, x, y, z, d, m - R - 1);
        sk += W<Q>::ko[m] * F[i + (AF_DH + v) * fs];
        skP += W<Q>::ko[m] * F[i + (AF_DP + v) * fs];
      }
    }
    // the mixed derivatives, outer axis d1 < inner axis d2 as the forward
    const int pairs[3][2] = {{0, 1}, {0, 2}, {1, 2}};
    for (int pq = 0; pq < 3; ++pq) {
      const int d1 = pairs[pq][0], d2 = pairs[pq][1];
      const int f = AF_CH + 10 * sidx3(d1, d2) + v;
      for (int m1 = 0; m1 < 2 * R + 1; ++m1) {
        if (m1 == R) continue;
        double inner = 0.0;
        for (int m2 = 0; m2 < 2 * R + 1; ++m2) {
          if (m2 == R) continue;
          inner += W<Q>::d1[m2] * F[adj_shift2(k, x, y, z, d1, m1 - R, d2, m2 - R) + f * fs];
        }
        s2 += W<Q>::d1[m1] * inner;
      }
    }
    uh += ih * s1h + ih2 * s2 + sk;
    uP += ih * s1P + skP;
    const double yb[2] = {uh, uP};
    for (int q = 0; q < 2; ++q) {
      const long idx = off + (v + NC * q) * fs;
      if (a.ybar) {
        a.ybar[idx] = yb[q];
        continue;
      }
      const double la = (a.first ? a.lam[idx] : a.lacc[idx]) + yb[q];
      a.lacc[idx] = la;
      a.kbn[idx] = a.last ? a.cl * la : a.cl * a.lam[idx] + a.cy * yb[q];
    }
  }
}

// lambda_T = dJ/du_T for J = the masked L2 error sqrt(sum e^2 / nw) over the
// evolved brick points (r >= r_1; each brick point one primary, weight 1),
// and kbn = c lambda_T (the last stage's cotangent).
template <bool MOVE>
__global__ __launch_bounds__(256) void k_adj_init(AdjInitArgs a) {
  const Brick& k = a.k;
  int x, y, z;
  if (!brick_map(k, x, y, z)) return;
  const long off = brick_index(k, x, y, z), fs = k.fs;
  const double xp[3] = {k.lo[0] + (double)x * k.h, k.lo[1] + (double)y * k.h, k.lo[2] + (double)z * k.h};
  const bool keep = is_evolved(a.c, a.t, xp);
  double s[20];
  if (keep) case_state_any<MOVE ? BGS_ALL : BGS_STATIC>(a.c, a.t, xp, s);
  const double f = (a.J > 0.0 && a.nw > 0.0) ? 1.0 / (a.J * a.nw) : 0.0;
  for (int v = 0; v < 20; ++v) {
    const double l = keep ? (a.u[off + v * fs] - s[v]) * f : 0.0;
    a.lam[off + v * fs] = l;
    a.kbn[off + v * fs] = a.cl * l;
  }
}

// A block-layout cotangent summed over the copies of each brick point (a
// point on a block face is an own point of every block holding it, and each
// copy's F is the brick point's), into the brick.
__global__ __launch_bounds__(256) void k_adj_fold_copies(AdjCopiesArgs a) {
  const Brick& k = a.k;
  const Geom& g = a.g;
  int x, y, z;
  if (!brick_map(k, x, y, z)) return;
  const int gp[3] = {x, y, z};
  int nc[3], ob[3][2], lc[3][2];
  for (int d = 0; d < 3; ++d) {
    nc[d] = 0;
    const int o = gp[d] / g.N, l = gp[d] - o * g.N;
    if (o < g.nb[d]) {
      ob[d][nc[d]] = o;
      lc[d][nc[d]++] = l;
    }
    if (l == 0) {
      if (o > 0) {
        ob[d][nc[d]] = o - 1;
        lc[d][nc[d]++] = g.N;
      } else if (g.periodic[d]) {
        ob[d][nc[d]] = g.nb[d] - 1;
        lc[d][nc[d]++] = g.N;
      }
    }
  }
  const long psz = (long)g.P * g.P * g.P, off = brick_index(k, x, y, z);
  for (int v = 0; v < 20; ++v) {
    double s = 0.0;
    for (int i = 0; i < nc[0]; ++i)
      for (int j = 0; j < nc[1]; ++j)
        for (int m = 0; m < nc[2]; ++m) {
          const int b = g.lookup[ob[0][i] + g.nb[0] * (ob[1][j] + g.nb[1] * ob[2][m])];
          s += a.src[(long)b * 20 * psz + v * psz + lc[0][i] + (long)g.P * (lc[1][j] + (long)g.P * lc[2][m])];
        }
    a.dst[off + v * k.fs] = s;
  }
}
// A brick cotangent onto the block layout's primaries (the copies, which the
// fill overwrites before every evaluation, take zero); dst is zeroed first.
__global__ __launch_bounds__(256) void k_adj_to_primaries(AdjCopiesArgs a) {
  const Brick& k = a.k;
  const Geom& g = a.g;
  int x, y, z;
  if (!brick_map(k, x, y, z)) return;
  const int gp[3] = {x, y, z};
  int loc[3];
  const int b = brick_primary(g, gp, loc);
  const long psz = (long)g.P * g.P * g.P, off = brick_index(k, x, y, z);
  double* d = a.dst + (long)b * 20 * psz + loc[0] + (long)g.P * (loc[1] + (long)g.P * loc[2]);
  for (int v = 0; v < 20; ++v) d[v * psz] = a.src[off + v * k.fs];
}

#endif  // TGH_ADJOINT_KERNELS

}  // namespace tgh
