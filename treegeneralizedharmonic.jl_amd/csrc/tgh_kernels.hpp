// The HIP kernels of the library: ghost fill, the fused right-hand side with
// its RK4 stage epilogue, and the pointwise kernels around them.
//
// Layouts (device, FP64, column-major as the reference's TreeAMR arrays):
//   state  u[i, j, k, v, b], (N+1)^3 owned points per block (closed vertex
//          blocks), v = 0..19 (h_ab then Pi_ab), blocks in Morton order;
//   work   W[i, j, k, v, b], (N+1+2G)^3 per block, owned point (i,j,k) at
//          (i+G, j+G, k+G); G = q/2 + 1 (src/evolution.jl:1047-1053).
#pragma once
#include "tgh_bounds.hpp"
#include "tgh_device.hpp"

namespace tgh {

enum { INT_NONE = 0, INT_DAMPED = 1, INT_FROZEN = 2, INT_PASTED = 3, INT_FITTED = 4 };
// Gauge-source kinds as the kernels are specialised: the case's kind, and for
// a moving hole its own (its kernels carry the moving core and layer).
enum { GAUGE_NONE = 0, GAUGE_SAMPLED = 1, GAUGE_KS = 2, GAUGE_KS_MOVING = 3, GAUGE_NONE_MOVING = 4 };
// The kernels' interior kinds (template HOLE): 0 none, 1 the sphere, 2 the
// general interior (tracked geometry, :fitted), 3 the sphere of a moving
// :damped hole or of a harmonic hole at rest, whose core and layer targets
// come from the box of k_brick_layer (per stage, or once at rest) instead of
// the dual passes in the head.
enum { HOLE_BOXED = 3 };
constexpr bool gauge_has_source(int gauge) { return gauge == GAUGE_SAMPLED || gauge == GAUGE_KS || gauge == GAUGE_KS_MOVING; }
enum { EPI_STORE = 0, EPI_RK4 = 1 };

// Everything a kernel needs to know about the mesh and the case: passed by
// value (kernel argument space), so nothing here is a device allocation.
struct Geom {
  int N, G, S, P;          // cells per block, ghost width, work and state extents
  int nb[3];               // brick of blocks
  int periodic[3];
  int refl[3][2];          // reflecting low / high face per axis
  int rot;                 // the rotating seam: the low x and y faces glued by a quarter turn about z
  int nblocks;
  double lo[3];
  double h, inv_h;         // level 0's spacing
  const int* bcoord;       // 3 ints per block (in the brick of its level)
  const int* lookup;       // brick index -> block (a uniform forest)
  // A frozen hierarchy: per-block levels (spacing h / 2^level) and per owned
  // point a primary flag (nblocks P^3 bytes); both null on a uniform forest.
  const int* blevel;
  const unsigned char* prim;
};

TGH_HD int block_level(const Geom& g, int b) { return g.blevel ? g.blevel[b] : 0; }
// h / 2^l and 2^l / h: scalings by powers of two, exact.
TGH_HD double block_h(const Geom& g, int b) { return g.h / (double)(1 << block_level(g, b)); }
TGH_HD double block_inv_h(const Geom& g, int b) { return g.inv_h * (double)(1 << block_level(g, b)); }

struct CaseP {
  Bg bg;
  int gauge, damping, interior, has_diss;
  double g0, g0_near, g0_far, g0_width, g2;
  double eps_ko;  // the number, or the horizon profile's eps_out
  int diss;       // 0 the number, 1 the horizon profile
  double eps_in, eps_r1, eps_rh;
  double center[3];  // the hole's center at t = 0
  double r0, r1, rho_max, w_ramp, rho_ramp;
  double vel[3];     // its velocity: c(t) = center + vel t (HoleCenter)
  int moving;        // vel != 0
  double delta;      // layer_rate_step: cbrt(eps) r_0, the targets' rate difference step
  double ks_u[4], ks_S[4];  // the moving KerrSchildSource: u = gamma (1, v), S the boosted spin
  // FittedInterior's tracked geometry: the layer's radii along the ray from
  // the surface series (device array of (lmax + 1)^2 coefficients) clamped
  // into [r_in, r_out], less offset and thickness.
  int shaped, lmax;
  const double* shape;
  double r_in, r_out, offset, thickness;
  // The :fitted target cache (40 variables per owned point, A then S, in
  // the state layout), its fill time, whether its rate is fed forward, and
  // the trailing-side narrowing of the rho ramp.
  const double* tw;
  double t_f, trail;
  int rate;
  int general;  // shaped or :fitted: the kernels' general interior (HOLE = 2)
};

// shape_series (src/interior.jl:1294-1329): sum_lm a_lm y~_lm(n), the fully
// normalised Legendre recurrences in z and (n_x + i n_y)^m for the azimuth,
// in the oracle's order (no contraction).
TGH_HD double shape_term(const double* sh, int l, int m, double q, double C, double S, double s2) {
#pragma clang fp contract(off)
  if (m == 0) return sh[l * l + l] * q;
  const double ac = sh[l * l + l + m], as = sh[l * l + l - m];
  return s2 * q * (ac * C - as * S);
}
TGH_HD double shape_series(const double* sh, int lmax, const double n[3]) {
#pragma clang fp contract(off)
  const double x = n[0], y = n[1], z = n[2];
  const double s2 = sqrt(2.0);
  double acc = 0.0;
  double qmm = 1.0 / sqrt(4.0 * M_PI);
  double C = 1.0, S = 0.0;
  for (int m = 0; m <= lmax; ++m) {
    if (m > 0) {
      qmm = -sqrt((double)(2 * m + 1) / (double)(2 * m)) * qmm;
      const double Cn = C * x - S * y, Sn = C * y + S * x;
      C = Cn;
      S = Sn;
    }
    acc += shape_term(sh, m, m, qmm, C, S, s2);
    if (m + 1 <= lmax) {
      double qa = qmm;
      double qb = sqrt((double)(2 * m + 3)) * z * qmm;
      acc += shape_term(sh, m + 1, m, qb, C, S, s2);
      for (int l = m + 2; l <= lmax; ++l) {
        const double al = sqrt((double)(4 * l * l - 1) / (double)(l * l - m * m));
        const double be = sqrt((double)((l - 1) * (l - 1) - m * m) / (double)(4 * (l - 1) * (l - 1) - 1));
        const double qc = al * (z * qb - be * qa);
        acc += shape_term(sh, l, m, qc, C, S, s2);
        qa = qb;
        qb = qc;
      }
    }
  }
  return acc;
}

TGH_HD void position(const Geom& g, int b, int i, int j, int k, double x[3]) {
  const int* B = g.bcoord + 3 * b;
  const double h = block_h(g, b);
  x[0] = g.lo[0] + (double)(B[0] * g.N + i) * h;
  x[1] = g.lo[1] + (double)(B[1] * g.N + j) * h;
  x[2] = g.lo[2] + (double)(B[2] * g.N + k) * h;
}

// center_at (src/interior.jl:86-93): c(t) = c0 + v t. It and the radius and
// the core rule below round as the oracle does (no contraction): the moving
// core's target rate is a difference over 2 delta of values taken at them.
TGH_HD void center_at(const CaseP& c, double t, double cc[3]) {
#pragma clang fp contract(off)
  cc[0] = c.center[0] + t * c.vel[0];
  cc[1] = c.center[1] + t * c.vel[1];
  cc[2] = c.center[2] + t * c.vel[2];
}

TGH_HD double damping_rate(const CaseP& c, double t, const double x[3]) {  // src/gauge.jl:1100,1140-1147
  if (c.damping == 0) return c.g0;
  double cc[3];
  center_at(c, t, cc);
  const double d1 = x[0] - cc[0], d2 = x[1] - cc[1], d3 = x[2] - cc[2];
  const double r2 = d1 * d1 + d2 * d2 + d3 * d3;
  return c.g0_far + (c.g0_near - c.g0_far) * exp(-r2 / (2.0 * c.g0_width * c.g0_width));
}

// dissipation_rate (src/gauge.jl:1180,1254-1264): eps_KO itself, or the
// HorizonDissipation profile about the hole's center -- eps_out at and
// outside r_h, eps_in at and inside r_1, each through a branch, and the
// quintic smoothstep of (r_h - r)/(r_h - r_1) between.
TGH_HD double dissipation_rate(const CaseP& c, double t, const double x[3]) {
  if (c.diss == 0) return c.eps_ko;
  double cc[3];
  center_at(c, t, cc);
  const double d1 = x[0] - cc[0], d2 = x[1] - cc[1], d3 = x[2] - cc[2];
  const double r = sqrt(d1 * d1 + d2 * d2 + d3 * d3);
  if (r >= c.eps_rh) return c.eps_ko;
  if (r <= c.eps_r1) return c.eps_in;
  const double s = (c.eps_rh - r) / (c.eps_rh - c.eps_r1);
  return c.eps_ko + (c.eps_in - c.eps_ko) * smoothstep(s);
}

// interior_radius (src/interior.jl:303-310): |x - c(t)|.
TGH_HD double hole_radius(const CaseP& c, double t, const double x[3]) {
#pragma clang fp contract(off)
  double cc[3];
  center_at(c, t, cc);
  const double d1 = x[0] - cc[0], d2 = x[1] - cc[1], d3 = x[2] - cc[2];
  return sqrt(d1 * d1 + d2 * d2 + d3 * d3);
}

// interior_point (src/interior.jl:341,1868-1886): the distance from c(t) and
// the layer's radii along the ray -- the sphere's r_0, r_1, or the tracked
// surface's r_1 = r_h(n) - offset and r_0 = r_1 - thickness with the two
// exact fast paths (outside r_out - offset, inside r_in - offset - thickness).
struct IGeo {
  double r, r0, r1;
};
TGH_HD IGeo interior_geo(const CaseP& c, double t, const double x[3]) {
#pragma clang fp contract(off)
  double cc[3];
  center_at(c, t, cc);
  const double d1 = x[0] - cc[0], d2 = x[1] - cc[1], d3 = x[2] - cc[2];
  const double r = sqrt(d1 * d1 + d2 * d2 + d3 * d3);
  if (!c.shaped) return {r, c.r0, c.r1};
  const double r1o = c.r_out - c.offset;
  if (r >= r1o) return {r, r1o - c.thickness, r1o};
  const double r1i = c.r_in - c.offset, r0i = r1i - c.thickness;
  if (r < r0i) return {r, r0i, r1i};
  double n[3] = {0.0, 0.0, 1.0};
  if (r != 0.0) {
    n[0] = d1 / r;
    n[1] = d2 / r;
    n[2] = d3 / r;
  }
  const double r1 = fmin(fmax(shape_series(c.shape, c.lmax, n), c.r_in), c.r_out) - c.offset;
  return {r, r1 - c.thickness, r1};
}
// core_position on the tracked geometry (src/interior.jl:1950-1965): the core
// surface's point on the ray, r_0(n) = (r_h(n) - offset) - thickness, where
// r < r_0(n); x itself elsewhere.
TGH_HD void shaped_core_point(const CaseP& c, double t, const double x[3], double xc[3]) {
#pragma clang fp contract(off)
  xc[0] = x[0];
  xc[1] = x[1];
  xc[2] = x[2];
  double cc[3];
  center_at(c, t, cc);
  const double d[3] = {x[0] - cc[0], x[1] - cc[1], x[2] - cc[2]};
  const double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  if (r >= (c.r_out - c.offset) - c.thickness) return;
  double n[3] = {0.0, 0.0, 1.0};
  if (r != 0.0) {
    n[0] = d[0] / r;
    n[1] = d[1] / r;
    n[2] = d[2] / r;
  }
  const double r0 = (fmin(fmax(shape_series(c.shape, c.lmax, n), c.r_in), c.r_out) - c.offset) - c.thickness;
  if (r >= r0) return;
  for (int q = 0; q < 3; ++q) xc[q] = cc[q] + r0 * n[q];
}

// core_position (src/interior.jl:615-625): inside the core, the point on the
// sphere r_0 about c(t) along the ray (+z at the center).
TGH_HD void core_point(const CaseP& c, double t, const double x[3], double xc[3]) {
#pragma clang fp contract(off)
  xc[0] = x[0];
  xc[1] = x[1];
  xc[2] = x[2];
  if (c.interior != INT_NONE) {
    double cc[3];
    center_at(c, t, cc);
    const double d[3] = {x[0] - cc[0], x[1] - cc[1], x[2] - cc[2]};
    const double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    if (r < c.r0) {
      double n[3] = {0.0, 0.0, 1.0};
      if (r != 0.0) {
        n[0] = d[0] / r;
        n[1] = d[1] / r;
        n[2] = d[2] / r;
      }
      for (int q = 0; q < 3; ++q) xc[q] = cc[q] + c.r0 * n[q];
    }
  }
}
template <int SET = BGS_ALL>
TGH_HD void case_state(const CaseP& c, double t, const double x[3], double out[20]) {
  double xc[3];
  core_point(c, t, x, xc);
  background_state<SET>(c.bg, t, xc, out, out + NC);
}
// case_state with the tracked geometry's core rule where the case has one
// (the fill, the initial state, the error norms and the sampled source).
template <int SET>
TGH_HD void case_state_any(const CaseP& c, double t, const double x[3], double out[20]) {
  if (c.shaped && c.interior != INT_NONE) {
    double xc[3];
    shaped_core_point(c, t, x, xc);
    background_state<SET>(c.bg, t, xc, out, out + NC);
    return;
  }
  case_state<SET>(c, t, x, out);
}
// The backgrounds a right-hand side compiled for a gauge-source kind can meet:
// H_a = 0 the harmonic charts (at rest or moving), a sampled or the static
// closed-form source the charts at rest that need one, the moving closed form
// the moving Kerr-Schild hole.
constexpr int bg_set_for(int gauge) {
  return gauge == GAUGE_NONE          ? BGS_HARMONIC
         : gauge == GAUGE_NONE_MOVING ? (bg_bit(BG_HARMONIC) | BGS_MOVING)
         : gauge == GAUGE_KS_MOVING   ? (bg_bit(BG_KERR_SCHILD) | BGS_MOVING)
                                      : BGS_SOURCED;
}

// ---------------------------------------------------------------------------
// The moving hole's core and layer (src/interior.jl:390-549,
// src/initialdata.jl:1697-1764), in the oracle's order without contraction.

// background_state_rate: the state at (t, x), its d_t h from the same pass,
// and d_t Pi as the centered difference of Pi at t +- delta over the
// difference of the two rounded times.
TGH_HD void background_state_rate(const Bg& bg, double t, const double x[3], double delta, double h[NC],
                                  double P[NC], double dth[NC], double dtP[NC]) {
#pragma clang fp contract(off)
  bg_state_exact(bg, t, x, h, P, dth);
  const double tp = t + delta, tm = t - delta;
  double hx[NC], Pp[NC], Pm[NC], dx[NC];
  bg_state_exact(bg, tp, x, hx, Pp, dx);
  bg_state_exact(bg, tm, x, hx, Pm, dx);
#pragma unroll
  for (int n = 0; n < NC; ++n) dtP[n] = (Pp[n] - Pm[n]) / (tp - tm);
}
// core_target_rate: the rate of t -> u_exact(t, c(t) + r_0 n(t)) at x by the
// centered difference, weighted by core_rate_weight = smoothstep(4 r / r_0 - 1);
// no pass where that is zero (r <= r_0 / 4).
TGH_HD void core_target_rate(const CaseP& c, double t, const double x[3], double dth[NC], double dtP[NC]) {
#pragma clang fp contract(off)
  const double s = smoothstep(4.0 * hole_radius(c, t, x) / c.r0 - 1.0);
#pragma unroll
  for (int n = 0; n < NC; ++n) dth[n] = dtP[n] = 0.0;
  if (s == 0.0) return;
  const double tp = t + c.delta, tm = t - c.delta;
  double xp[3], xm[3];
  core_point(c, tp, x, xp);
  core_point(c, tm, x, xm);
  double hp[NC], Pp[NC], hm[NC], Pm[NC], dx[NC];
  bg_state_exact(c.bg, tp, xp, hp, Pp, dx);
  bg_state_exact(c.bg, tm, xm, hm, Pm, dx);
  const double f = s / (tp - tm);
#pragma unroll
  for (int n = 0; n < NC; ++n) {
    dth[n] = f * (hp[n] - hm[n]);
    dtP[n] = f * (Pp[n] - Pm[n]);
  }
}
TGH_HD void inv3_ex(const double m[3][3], double o[3][3]) {
#pragma clang fp contract(off)
  const double c00 = m[1][1] * m[2][2] - m[1][2] * m[2][1], c01 = m[1][2] * m[2][0] - m[1][0] * m[2][2],
               c02 = m[1][0] * m[2][1] - m[1][1] * m[2][0];
  const double det = m[0][0] * c00 + m[0][1] * c01 + m[0][2] * c02;
  o[0][0] = c00 / det;
  o[1][0] = c01 / det;
  o[2][0] = c02 / det;
  o[0][1] = (m[0][2] * m[2][1] - m[0][1] * m[2][2]) / det;
  o[1][1] = (m[0][0] * m[2][2] - m[0][2] * m[2][0]) / det;
  o[2][1] = (m[0][1] * m[2][0] - m[0][0] * m[2][1]) / det;
  o[0][2] = (m[0][1] * m[1][2] - m[0][2] * m[1][1]) / det;
  o[1][2] = (m[0][2] * m[1][0] - m[0][0] * m[1][2]) / det;
  o[2][2] = (m[0][0] * m[1][1] - m[0][1] * m[1][0]) / det;
}
TGH_HD void mv3_ex(const double m[3][3], const double v[3], double o[3]) {
#pragma clang fp contract(off)
#pragma unroll
  for (int i = 0; i < 3; ++i) o[i] = m[i][0] * v[0] + m[i][1] * v[1] + m[i][2] * v[2];
}
// _adm_split (src/interior.jl:505-515) without the square root.
TGH_HD void adm_split_ex(const double h[NC], double& A, double bu[3], double bl[3], double g[3][3]) {
#pragma clang fp contract(off)
  g[0][0] = 1.0 + h[4]; g[0][1] = h[5]; g[0][2] = h[6];
  g[1][0] = h[5]; g[1][1] = 1.0 + h[7]; g[1][2] = h[8];
  g[2][0] = h[6]; g[2][1] = h[8]; g[2][2] = 1.0 + h[9];
  bl[0] = h[1];
  bl[1] = h[2];
  bl[2] = h[3];
  double gi[3][3];
  inv3_ex(g, gi);
  mv3_ex(gi, bl, bu);
  A = (1.0 - h[0]) + (bu[0] * bl[0] + bu[1] * bl[1] + bu[2] * bl[2]);
}
// core_relaxation_rate (src/interior.jl:450-503): the metric relaxed toward
// h_t at rate rho in (alpha^2, beta^i, gamma_ij), plus the target's ADM rate
// read off d_t h_t at h_t, mapped back to h by the chain rule of
// g_tt = -alpha^2 + gamma_ij beta^i beta^j, g_ti = gamma_ij beta^j.
TGH_HD void core_relaxation_rate(const double h[NC], const double ht[NC], double rho, const double dht[NC],
                                 double o[NC]) {
#pragma clang fp contract(off)
  double A, bu[3], bl[3], g[3][3], At, but[3], blt[3], gt[3][3];
  adm_split_ex(h, A, bu, bl, g);
  adm_split_ex(ht, At, but, blt, gt);
  const double dgt[3][3] = {{dht[4], dht[5], dht[6]}, {dht[5], dht[7], dht[8]}, {dht[6], dht[8], dht[9]}};
  double gti[3][3], dgbt[3], rh[3], dbt[3];
  inv3_ex(gt, gti);
  mv3_ex(dgt, but, dgbt);
#pragma unroll
  for (int i = 0; i < 3; ++i) rh[i] = dht[1 + i] - dgbt[i];
  mv3_ex(gti, rh, dbt);
  const double dAt = -dht[0] + (but[0] * dgbt[0] + but[1] * dgbt[1] + but[2] * dgbt[2]) +
                     2.0 * (blt[0] * dbt[0] + blt[1] * dbt[1] + blt[2] * dbt[2]);
  const double dA = dAt - rho * (A - At);
  double db[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) db[i] = dbt[i] - rho * (bu[i] - but[i]);
  const double dd[3][3] = {{h[4] - ht[4], h[5] - ht[5], h[6] - ht[6]},
                           {h[5] - ht[5], h[7] - ht[7], h[8] - ht[8]},
                           {h[6] - ht[6], h[8] - ht[8], h[9] - ht[9]}};
  double dg[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) dg[i][j] = dgt[i][j] - rho * dd[i][j];
  double dgb[3], gdb[3];
  mv3_ex(dg, bu, dgb);
  mv3_ex(g, db, gdb);
  const double dtt = -dA + (bu[0] * dgb[0] + bu[1] * dgb[1] + bu[2] * dgb[2]) +
                     2.0 * (bl[0] * db[0] + bl[1] * db[1] + bl[2] * db[2]);
  o[0] = dtt;
#pragma unroll
  for (int i = 0; i < 3; ++i) o[1 + i] = dgb[i] + gdb[i];
  o[4] = dg[0][0];
  o[5] = dg[1][0];
  o[6] = dg[2][0];
  o[7] = dg[1][1];
  o[8] = dg[2][1];
  o[9] = dg[2][2];
}
// The moving :damped core's rates at x (gh_rhs_point!, src/evolution.jl:
// 526-583): the ADM relaxation of h and Pi's linear one toward the target at
// the core rule's point, at core_rate = max(rho_max, 12 |v| / r_0), plus the
// target's rates. hu and Pu are the point's state.
// Its two halves: the target (the core rule's point's exact state hd, Pd and
// the target's rates dthd, dtPd, which need the dual passes) and the
// relaxation of the point's own state toward it.
TGH_HD void moving_core_target(const CaseP& c, double t, const double x[3], double hd[NC], double Pd[NC],
                               double dthd[NC], double dtPd[NC]) {
#pragma clang fp contract(off)
  double xc[3], dx[NC];
  core_point(c, t, x, xc);
  bg_state_exact(c.bg, t, xc, hd, Pd, dx);
  core_target_rate(c, t, x, dthd, dtPd);
}
TGH_HD void moving_core_relax(const CaseP& c, const double hu[NC], const double Pu[NC], const double hd[NC],
                              const double Pd[NC], const double dthd[NC], const double dtPd[NC], double dh[NC],
                              double dP[NC]) {
#pragma clang fp contract(off)
  const double s = sqrt(c.vel[0] * c.vel[0] + c.vel[1] * c.vel[1] + c.vel[2] * c.vel[2]);
  const double rc = fmax(c.rho_max, 12.0 * s / c.r0);
  core_relaxation_rate(hu, hd, rc, dthd, dh);
#pragma unroll
  for (int v = 0; v < NC; ++v) dP[v] = dtPd[v] - rc * (Pu[v] - Pd[v]);
}
TGH_HD void moving_core_rates(const CaseP& c, double t, const double x[3], const double hu[NC],
                              const double Pu[NC], double dh[NC], double dP[NC]) {
  double hd[NC], Pd[NC], dthd[NC], dtPd[NC];
  moving_core_target(c, t, x, hd, Pd, dthd, dtPd);
  moving_core_relax(c, hu, Pu, hd, Pd, dthd, dtPd, dh, dP);
}

// Where a point of block b's work array takes its value from (TreeAMR 0.3's
// closed-block fill on a uniform forest): -2 the Dirichlet hook, -1 an own
// primary (left as scattered), 0 a copy of (owner, loc). A face point is a
// primary in the block where its local index is below N, and in the last
// block of a non-periodic axis also at N (the evolved boundary plane).
// The parity of variable v (h_ab then Pi_ab, packed tt, tx, ty, tz, xx, xy, xz,
// yy, yz, zz) under the mirror x^d -> -x^d (state_parity, src/boundaries.jl:
// 84-120): bit d is set when the index d occurs an odd number of times in ab.
TGH_HD int parity_odd_mask(int v) {
  const int c = v % 10;
  return c == 1 ? 1 : c == 2 ? 2 : c == 3 ? 4 : c == 5 ? 3 : c == 6 ? 5 : c == 8 ? 6 : 0;
}
// -1 when variable v changes sign under the mirrors of the axes in mask m.
TGH_HD double parity_sign(int v, int m) {
  const int o = parity_odd_mask(v) & m;
  return ((o ^ (o >> 1) ^ (o >> 2)) & 1) ? -1.0 : 1.0;
}
// The quarter turn about z (state_rotation, src/boundaries.jl:145-205): with
// R (x, y, z) -> (-y, x, z), u(R p) = Q u(p), (Q u)[v] = rot_sign(v) u[rot_src(v)]
// -- the covariant tensor turned index by index: t and z stay, the turned
// tensor's index x is minus the original's y, its y the original's x.
TGH_HD int rot_src(int v) {
  const int c = v % 10;
  const int s = c == 1 ? 2 : c == 2 ? 1 : c == 4 ? 7 : c == 6 ? 8 : c == 7 ? 4 : c == 8 ? 6 : c;
  return v - c + s;
}
TGH_HD double rot_sign(int v) {
  const int c = v % 10;
  return (c == 1 || c == 5 || c == 6) ? -1.0 : 1.0;
}
// Q^tn applied to variable v: the source variable and the sign.
TGH_HD void rot_power(int v, int tn, int& sv, double& sg) {
  sv = v;
  sg = 1.0;
  for (int n = 0; n < tn; ++n) {
    sg *= rot_sign(sv);
    sv = rot_src(sv);
  }
}
// The seam's map of a global point (x, y) beyond the low x and/or y face: the
// point inside it is the image of, and the number of quarter turns tn, with
// u(point) = Q^tn u(inside point). Unchanged (tn = 0) inside the quadrant.
TGH_HD int rot_unfold(int& x, int& y) {
  const int x0 = x, y0 = y;
  if (x0 < 0 && y0 < 0) {
    x = -x0;
    y = -y0;
    return 2;
  }
  if (x0 < 0) {
    x = y0;
    y = -x0;
    return 1;
  }
  if (y0 < 0) {
    x = -y0;
    y = x0;
    return 3;
  }
  return 0;
}

// Where work point l (owned-relative, -G .. N + G) of block b takes its value:
// -2 the Dirichlet hook, -1 its own primary, 0 a copy of the primary (owner,
// loc); *mirror gets the axes across whose reflecting face the point lies
// (the copy then takes each variable's parity). Beyond a face that neither
// wraps nor reflects the point is the hook's, also where it lies beyond a
// reflecting face on another axis.
TGH_HD int classify(const Geom& g, int b, const int l[3], int& owner, int loc[3], int* mirror = nullptr,
                    int* turn = nullptr) {
  int ob[3];
  int mm = 0;
  const int* B = g.bcoord + 3 * b;
  int gs[3] = {B[0] * g.N + l[0], B[1] * g.N + l[1], B[2] * g.N + l[2]};
  const int tn = g.rot ? rot_unfold(gs[0], gs[1]) : 0;
  if (turn) *turn = tn;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    int gi = gs[d];
    const int tot = g.nb[d] * g.N;
    if (g.periodic[d]) {
      const int gw = ((gi % tot) + tot) % tot;
      ob[d] = gw / g.N;
      loc[d] = gw - ob[d] * g.N;
    } else {
      if (gi < 0) {
        if (!g.refl[d][0]) return -2;
        gi = -gi;
        mm |= 1 << d;
      } else if (gi > tot) {
        if (!g.refl[d][1]) return -2;
        gi = 2 * tot - gi;
        mm |= 1 << d;
      }
      const int o = gi / g.N;
      ob[d] = o < g.nb[d] - 1 ? o : g.nb[d] - 1;
      loc[d] = gi - ob[d] * g.N;
    }
  }
  if (mirror) *mirror = mm;
  owner = g.lookup[ob[0] + g.nb[0] * (ob[1] + g.nb[1] * ob[2])];
  if (mm == 0 && tn == 0 && owner == b && loc[0] == l[0] && loc[1] == l[1] && loc[2] == l[2]) return -1;
  return 0;
}
// The axes whose reflecting wall plane holds owned point l of block b.
TGH_HD int wall_mask(const Geom& g, int b, const int l[3]) {
  int m = 0;
  const int* B = g.bcoord + 3 * b;
  const int L = g.blevel ? g.blevel[b] : 0;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const int gi = B[d] * g.N + l[d], tot = (g.nb[d] << L) * g.N;
    if ((g.refl[d][0] && gi == 0) || (g.refl[d][1] && gi == tot)) m |= 1 << d;
  }
  return m;
}
TGH_HD bool reflects(const Geom& g) {
  return g.refl[0][0] | g.refl[0][1] | g.refl[1][0] | g.refl[1][1] | g.refl[2][0] | g.refl[2][1] | g.rot;
}

// XCD-aware workgroup order. MI355X dispatches consecutive workgroups of a
// launch round-robin over its 8 XCDs, each with its own L2. A launch here is
// 1-D with 8 * ceil(total / 8) workgroups, and hardware workgroup w takes
// logical workgroup (w mod 8) * per + w / 8, so each XCD runs one contiguous
// range of logical workgroups: a run of whole blocks, consecutive in Morton
// order, whose work arrays and neighbours stay in that XCD's L2.
constexpr int NUM_XCD = 8;
__device__ __forceinline__ int wg_logical(int total) {
  const int per = (total + NUM_XCD - 1) / NUM_XCD;
  const int flat = blockIdx.x;
  return (flat % NUM_XCD) * per + flat / NUM_XCD;
}
__device__ __forceinline__ bool wg_map(int wgpb, int nblocks, int& b, int& wx) {
  const int total = wgpb * nblocks;
  const int L = wg_logical(total);
  if (L >= total) return false;
  b = L / wgpb;
  wx = L - b * wgpb;
  return true;
}

// ---------------------------------------------------------------------------
// Stencils on the work array. Weights are exact rationals rounded once
// (src/stencils.jl:124-157); the zero center weight of the first derivative is
// skipped (bitwise-neutral). Sums run from the lowest offset to the highest.
template <int Q> struct W;
template <> struct W<2> {
  static constexpr int R = 1;
  static constexpr double d1[3] = {-1.0 / 2, 0.0, 1.0 / 2};
  static constexpr double d2[3] = {1.0, -2.0, 1.0};
  static constexpr double ko[5] = {-1.0 / 16, 4.0 / 16, -6.0 / 16, 4.0 / 16, -1.0 / 16};
};
template <> struct W<4> {
  static constexpr int R = 2;
  static constexpr double d1[5] = {1.0 / 12, -2.0 / 3, 0.0, 2.0 / 3, -1.0 / 12};
  static constexpr double d2[5] = {-1.0 / 12, 4.0 / 3, -5.0 / 2, 4.0 / 3, -1.0 / 12};
  static constexpr double ko[7] = {1.0 / 64, -6.0 / 64, 15.0 / 64, -20.0 / 64, 15.0 / 64, -6.0 / 64, 1.0 / 64};
};
template <> struct W<6> {
  static constexpr int R = 3;
  static constexpr double d1[7] = {-1.0 / 60, 3.0 / 20, -3.0 / 4, 0.0, 3.0 / 4, -3.0 / 20, 1.0 / 60};
  static constexpr double d2[7] = {1.0 / 90, -3.0 / 20, 3.0 / 2, -49.0 / 18, 3.0 / 2, -3.0 / 20, 1.0 / 90};
  static constexpr double ko[9] = {-1.0 / 256, 8.0 / 256, -28.0 / 256, 56.0 / 256, -70.0 / 256,
                                   56.0 / 256, -28.0 / 256, 8.0 / 256, -1.0 / 256};
};
template <> struct W<8> {
  static constexpr int R = 4;
  static constexpr double d1[9] = {1.0 / 280, -4.0 / 105, 1.0 / 5, -4.0 / 5, 0.0,
                                   4.0 / 5, -1.0 / 5, 4.0 / 105, -1.0 / 280};
  static constexpr double d2[9] = {-1.0 / 560, 8.0 / 315, -1.0 / 5, 8.0 / 5, -205.0 / 72,
                                   8.0 / 5, -1.0 / 5, 8.0 / 315, -1.0 / 560};
  static constexpr double ko[11] = {1.0 / 1024, -10.0 / 1024, 45.0 / 1024, -120.0 / 1024,
                                    210.0 / 1024, -252.0 / 1024, 210.0 / 1024, -120.0 / 1024,
                                    45.0 / 1024, -10.0 / 1024, 1.0 / 1024};
};

template <int Q>
__device__ __forceinline__ double st_d1(const double* __restrict__ p, int s) {
  constexpr int R = W<Q>::R;
  double acc = 0.0;
#pragma unroll
  for (int k = 0; k < 2 * R + 1; ++k)
    if (k != R) acc += W<Q>::d1[k] * p[(k - R) * s];
  return acc;
}
template <int Q>
__device__ __forceinline__ double st_d2(const double* __restrict__ p, int s) {
  constexpr int R = W<Q>::R;
  double acc = 0.0;
#pragma unroll
  for (int k = 0; k < 2 * R + 1; ++k) acc += W<Q>::d2[k] * p[(k - R) * s];
  return acc;
}
template <int Q>
__device__ __forceinline__ double st_ko(const double* __restrict__ p, int s) {
  constexpr int R = W<Q>::R + 1;
  double acc = 0.0;
#pragma unroll
  for (int k = 0; k < 2 * R + 1; ++k) acc += W<Q>::ko[k] * p[(k - R) * s];
  return acc;
}
// The mixed derivative: outer sum along s1, inner along s2 (src/stencils.jl:363-374).
template <int Q>
__device__ __forceinline__ double st_mixed(const double* __restrict__ p, int s1, int s2) {
  constexpr int R = W<Q>::R;
  double outer = 0.0;
#pragma unroll
  for (int a = 0; a < 2 * R + 1; ++a) {
    if (a == R) continue;
    double inner = 0.0;
#pragma unroll
    for (int e = 0; e < 2 * R + 1; ++e)
      if (e != R) inner += W<Q>::d1[e] * p[(a - R) * s1 + (e - R) * s2];
    outer += W<Q>::d1[a] * inner;
  }
  return outer;
}

// ---------------------------------------------------------------------------
// The stage time. A kernel captured in a graph that is replayed step after
// step reads the step's time t_n from a device scalar (clock) that the host
// sets before each replay, and adds its stage's offset (toff: 0, dt/2 or dt,
// the same sum t_n + dt/2 the host forms); without a clock it is the
// argument t itself.
TGH_HD double stage_time(double t, const double* clock, double toff) { return clock ? *clock + toff : t; }

struct RhsArgs {
  Geom g;
  CaseP c;
  double t;
  const double* __restrict__ W;  // stage input, work layout, ghosts filled
  const double* __restrict__ Hsrc;
  const double* __restrict__ m_rho;
  const double* __restrict__ m_S;
  const double* __restrict__ m_Sij;
  // EPI_STORE
  double* du;
  // EPI_RK4: stage s = 1..4 of the acc/Y recurrences; us is u_n, read in
  // stages 1-3 and overwritten by u_{n+1} in stage 4
  double* us;
  double* acc;
  double* Wn;
  double cacc, cnext;
  int stage;
  // the split form: per-point coefficients the head leaves for the Pi kernel
  double* __restrict__ coef;
  const double* clock;  // stage_time's clock (NULL: t as given) and offset
  double toff;
  // a moving hole's layer targets and rates on the brick, evaluated per stage by
  // k_brick_layer (LAYER_FIELDS fields of lB^3 slots, slot = brick coordinates mod lB);
  // NULL: the head evaluates them itself
  const double* lbox;
  long lB3;
  int lB;
  int lLmax;  // the box's grid: the finest level's (a block point of level L is
              // at index (B N + i) 2^(lLmax - L) on it)
};
// The layer box's slot of brick point (x, y, z).
TGH_HD long layer_slot(int x, int y, int z, int B) {
  return (long)(x % B) + (long)B * ((long)(y % B) + (long)B * (long)(z % B));
}


// The epilogue: store F (the rhs! verb) or apply RK4 stage s to it: with k = F,
//   s = 1: acc = u + c_acc k,  Y = u + c_next k
//   s = 2, 3: acc += c_acc k,  Y = u + c_next k
//   s = 4: u = acc + c_acc k (and Y = u, the next step's first stage input)
// where Y is the next stage's input, written into the other work array's
// owned point. The pointers are kernel arguments marked __restrict__ so that
// the stores do not order the stencil loads around them.
#define TGH_OUT_PARAMS                                                                          \
  double *__restrict__ du, double *__restrict__ us, double *__restrict__ acc,                   \
      double *__restrict__ Wn
#define TGH_OUT_ARGS du, us, acc, Wn
template <int EPI>
__device__ __forceinline__ void emit(const RhsArgs& a, TGH_OUT_PARAMS, long si, long wi, double k) {
  if (EPI == EPI_STORE) {
    du[si] = k;
  } else {
    if (a.stage == 1) {
      const double u = us[si];
      acc[si] = u + a.cacc * k;
      Wn[wi] = u + a.cnext * k;
    } else if (a.stage < 4) {
      const double u = us[si];
      acc[si] = acc[si] + a.cacc * k;
      Wn[wi] = u + a.cnext * k;
    } else {
      const double un = acc[si] + a.cacc * k;
      us[si] = un;
      Wn[wi] = un;
    }
  }
}

// The RK4 epilogue in two halves, so that a kernel can issue the loads of u_n
// and of the accumulator early: epi_load reads what stage s needs (u_n in
// stages 1-3, the accumulator in stages 2-4), epi_store applies the stage.
__device__ __forceinline__ void epi_load(const RhsArgs& a, const double* __restrict__ us,
                                         const double* __restrict__ acc, long si, double& u, double& ac) {
  u = a.stage < 4 ? us[si] : 0.0;
  ac = a.stage > 1 ? acc[si] : 0.0;
}
__device__ __forceinline__ void epi_store(const RhsArgs& a, double* __restrict__ us, double* __restrict__ acc,
                                          double* __restrict__ Wn, long si, long wi, double k, double u,
                                          double ac) {
  if (a.stage == 1) {
    acc[si] = u + a.cacc * k;
    Wn[wi] = u + a.cnext * k;
  } else if (a.stage < 4) {
    acc[si] = ac + a.cacc * k;
    Wn[wi] = u + a.cnext * k;
  } else {
    const double un = ac + a.cacc * k;
    us[si] = un;
    Wn[wi] = un;
  }
}

// Where one point's inputs and outputs live: the stage input's center w0 with
// its field stride sv and row and plane strides s2, s3 (the x stride is 1);
// the epilogue's state index si0 with field stride ssv (us, acc, du) and work
// index wi0 with field stride sv (Wn); the sampled gauge source Hp with field
// stride hs; the matter fields (rho, S_i, S_ij) with field stride ms.
struct PointRef {
  const double* __restrict__ w0;
  long sv;
  int s2, s3;
  double x[3];
  double ih;  // 1/h of the point's block (the block layout)
  long ti0;   // the :fitted target cache's index of variable 0 (the block layout)
  long si0, ssv, wi0;
  const double* __restrict__ Hp;
  long hs;
  const double* __restrict__ m_rho;
  const double* __restrict__ m_S;
  const double* __restrict__ m_Sij;
  long ms;
  long lslot;  // the layer box's slot of the point (read only where RhsArgs::lbox is set)
};

// What the Pi components need from the head at a point: beta^i, d_i beta^i,
// d_i A^{ij}, A^{ij} = alpha sqrt(gamma) gamma^{ij} (xx,xy,xz,yy,yz,zz), the
// layer weight w and c1_v = w msrc_v - rho (Pi_v - Pi_exact,v), so that
//   dPi_v = w (principal part of component v) + c1_v.
struct PiCoef {
  double b[3], divb, divA[3], A[6], w;
  double c1[NC];
  double epsh;  // the Kreiss-Oliger amplitude over h at the point
};

// The head at one point (gh_rhs_head, src/evolution.jl:258-297, inside
// gh_rhs_point!, src/evolution.jl:448-695,846-863): the state, the 30 first
// derivatives and the coefficients, d_t h complete (through the epilogue), the
// source, the interior layer's weights. Returns false in the frozen core,
// where F is not evaluated (a hole at rest: d_t h = 0 is emitted).
// Where the head reads a point's stencils: along x from X(v) (stride 1), along
// y from Y(v) (stride sy), along z from Z(v) (stride sz), the centre h_v at
// X(v)[0] and Pi_v from P(v). GlobalSrc reads the stage input in place (the
// three are the same pointer); the LDS head stages them per tile.
struct GlobalSrc {
  const double* __restrict__ w0;
  long sv;
  int sy, sz;
  __device__ __forceinline__ const double* X(int v) const { return w0 + v * sv; }
  __device__ __forceinline__ const double* Y(int v) const { return w0 + v * sv; }
  __device__ __forceinline__ const double* Z(int v) const { return w0 + v * sv; }
  __device__ __forceinline__ double P(int v) const { return w0[(NC + v) * sv]; }
};

// The frozen core (F is not evaluated there: a hole at rest, du = 0) and the
// point's radius from the hole's centre.
template <int HOLE>
__device__ __forceinline__ bool frozen_core(const CaseP& c, double t, const double x[3], double& r) {
  r = 0.0;
  if (!HOLE) return false;
  r = hole_radius(c, t, x);
  return c.interior == INT_PASTED ? r < c.r1 : r < c.r0;
}

// The head's first half (gh_rhs_head, src/evolution.jl:258-307): the state,
// the 30 first derivatives, the coefficient set and d_t h with its
// dissipation (D[0]); D[1..3] = d_i h.
template <int Q, class SRC>
__device__ __forceinline__ void head_derivs(const CaseP& c, const SRC& src, double inv_h, double epsh,
                                            double hv[NC], double D[4][NC], Coef& cf) {
  const double ih = inv_h;
#pragma unroll
  for (int v = 0; v < NC; ++v) {
    hv[v] = src.X(v)[0];
    D[1][v] = ih * st_d1<Q>(src.X(v), 1);
    D[2][v] = ih * st_d1<Q>(src.Y(v), src.sy);
    D[3][v] = ih * st_d1<Q>(src.Z(v), src.sz);
  }
  metric_quantities(hv, cf);
  const double a_div = cf.alpha / cf.sqrtg;
#pragma unroll
  for (int v = 0; v < NC; ++v) {
    double s = cf.beta[0] * D[1][v] + cf.beta[1] * D[2][v] + cf.beta[2] * D[3][v] + a_div * src.P(v);
    if (c.has_diss)
      s += epsh * (st_ko<Q>(src.X(v), 1) + st_ko<Q>(src.Y(v), src.sy) + st_ko<Q>(src.Z(v), src.sz));
    D[0][v] = s;
  }
}

// The Pi components' geometric coefficients: beta^i, d_i beta^i, d_i A^{ij} and
// A^{ij} = alpha sqrt(gamma) gamma^{ij} (metric_divergences, src/pointwise.jl:754-778).
__device__ __forceinline__ void head_geometry(const Coef& cf, const double D[4][NC], PiCoef& pc) {
  metric_divergences(cf, &D[1], pc.divb, pc.divA);
  const double as = cf.alpha * cf.sqrtg;
#pragma unroll
  for (int n = 0; n < 6; ++n) pc.A[n] = as * cf.gamu[n];
  pc.b[0] = cf.beta[0];
  pc.b[1] = cf.beta[1];
  pc.b[2] = cf.beta[2];
}

// The head's second half: the gauge source, the source msrc (gh_node_source_lean),
// matter, and the interior layer -- d_t h through the epilogue, and the
// weight w and c1_v = w msrc_v - rho (Pi_v - Pi_exact,v) into pc. Pi_v is
// read through src.P.
template <int GAUGE, int HOLE, int EPI, class SRC>
__device__ __forceinline__ void head_source(const RhsArgs& a, const PointRef& pr, const SRC& src, double r,
                                            const Coef& cf, const double hv[NC], const double D[4][NC],
                                            PiCoef& pc, TGH_OUT_PARAMS) {
  const long sv = pr.sv, psz = pr.ssv, si0 = pr.si0, wi0 = pr.wi0;
  const CaseP& c = a.c;
  double wgt = 1.0, rho = 0.0;
  const bool layer = HOLE && r < c.r1;
  if (layer) {
    const double sl = (r - c.r0) / (c.r1 - c.r0);
    wgt = smoothstep(sl / c.w_ramp);
    rho = c.interior == INT_FROZEN ? 0.0 : c.rho_max * smoothstep((1.0 - sl) / c.rho_ramp);
  }
  // The layer's target u_e (r >= r0 here: the core rule is the identity) and,
  // for a moving :damped hole, its rates, which (1 - w) d_t u_e adds so that
  // the moving solution is a steady state of the layer (src/evolution.jl:652-680).
  // At rest it is evaluated twice in the layer, for d_t h before the source
  // and for d_t Pi after it, so that no part of it is held through the
  // source. A moving background's exact state and rates cost three dual
  // passes each time: its variants evaluate them once and hold the Pi half
  // through the source (their register budget is spent on those passes).
  constexpr bool MOVE = (bg_set_for(GAUGE) & BGS_MOVING) != 0;
  const bool follows = MOVE && c.moving && c.interior == INT_DAMPED;
  const double om = 1.0 - wgt;
  double uP[MOVE ? NC : 1], rP[MOVE ? NC : 1];
  {
    double ue[20], re[20];
    if (layer) {
      if constexpr (HOLE == HOLE_BOXED && MOVE) {
        // a moving :damped hole's box (k_brick_layer): background_state_rate's
        // values, its d_t Pi formed here from Pi at t +- delta as it forms it
#pragma clang fp contract(off)
        const double* lb = a.lbox + pr.lslot;
        const double tp = a.t + c.delta, tm = a.t - c.delta;
#pragma unroll
        for (int v = 0; v < NC; ++v) {
          ue[v] = lb[v * a.lB3];
          ue[NC + v] = lb[(NC + v) * a.lB3];
          re[v] = lb[(2 * NC + v) * a.lB3];
          re[NC + v] = (lb[(3 * NC + v) * a.lB3] - lb[(4 * NC + v) * a.lB3]) / (tp - tm);
        }
      } else if constexpr (HOLE == HOLE_BOXED) {
        // a harmonic hole at rest: its target from the box, filled once
        const double* lb = a.lbox + pr.lslot;
#pragma unroll
        for (int v = 0; v < 20; ++v) ue[v] = lb[v * a.lB3];
      } else if (follows) {
        background_state_rate(c.bg, a.t, pr.x, c.delta, ue, ue + NC, re, re + NC);
      } else {
        case_state<bg_set_for(GAUGE)>(c, a.t, pr.x, ue);
      }
    }
    if constexpr (MOVE) {
#pragma unroll
      for (int v = 0; v < NC; ++v) {
        uP[v] = ue[NC + v];
        rP[v] = re[NC + v];
      }
    }
#pragma unroll
    for (int v = 0; v < NC; ++v) {
      double f = D[0][v];
      if (layer) f = follows ? wgt * f + om * re[v] - rho * (hv[v] - ue[v]) : wgt * f - rho * (hv[v] - ue[v]);
      emit<EPI>(a, TGH_OUT_ARGS, si0 + v * psz, wi0 + v * sv, f);
    }
  }
  const double g0 = damping_rate(c, a.t, pr.x);
  double msrc[NC];
  {
    double Hl[4] = {0.0, 0.0, 0.0, 0.0}, dH[4][4];
    if (GAUGE == GAUGE_SAMPLED) {
      const double* Hp = pr.Hp;
#pragma unroll
      for (int q = 0; q < 4; ++q) Hl[q] = Hp[q * pr.hs];
#pragma unroll
      for (int n = 0; n < 16; ++n) dH[n % 4][n / 4] = Hp[(4 + n) * pr.hs];
    } else if (GAUGE == GAUGE_KS) {
      ks_source(c.bg.M, c.bg.a, hv, D, Hl, dH);
    } else if (GAUGE == GAUGE_KS_MOVING) {
      ks_source_moving(c.bg.M, c.ks_S, c.ks_u, hv, D, Hl, dH);
    }
    gh_source<gauge_has_source(GAUGE)>(cf, hv, D, Hl, dH, g0, c.g2, msrc);
  }
  if (pr.m_rho != nullptr) {
    const double rr = pr.m_rho[0];
    const double Sm[3] = {pr.m_S[0], pr.m_S[pr.ms], pr.m_S[2 * pr.ms]};
    double Sij[6];
#pragma unroll
    for (int n = 0; n < 6; ++n) Sij[n] = pr.m_Sij[n * pr.ms];
    double ms[NC];
    matter_source(cf, hv, rr, Sm, Sij, ms);
#pragma unroll
    for (int n = 0; n < NC; ++n) msrc[n] += ms[n];
  }
  pc.w = layer ? wgt : 1.0;
  if constexpr (MOVE) {
#pragma unroll
    for (int v = 0; v < NC; ++v) {
      double c1 = msrc[v];
      if (layer)
        c1 = follows ? wgt * c1 + om * rP[v] - rho * (src.P(v) - uP[v]) : wgt * c1 - rho * (src.P(v) - uP[v]);
      pc.c1[v] = c1;
    }
    return;
  }
  double ue[20], re[20];
  if (layer) {
    if constexpr (HOLE == HOLE_BOXED) {
      const double* lb = a.lbox + pr.lslot;
#pragma unroll
      for (int v = 0; v < 20; ++v) ue[v] = lb[v * a.lB3];
    } else if (follows) {
      background_state_rate(c.bg, a.t, pr.x, c.delta, ue, ue + NC, re, re + NC);
    } else {
      case_state<bg_set_for(GAUGE)>(c, a.t, pr.x, ue);
    }
  }
#pragma unroll
  for (int v = 0; v < NC; ++v) {
    double c1 = msrc[v];
    if (layer)
      c1 = follows ? wgt * c1 + om * re[NC + v] - rho * (src.P(v) - ue[NC + v])
                   : wgt * c1 - rho * (src.P(v) - ue[NC + v]);
    pc.c1[v] = c1;
  }
}

// The frozen core's rates (gh_rhs_point!, src/evolution.jl:500-583): zero for
// a hole at rest; a moving :damped core relaxes toward its target.
template <int GAUGE, int HOLE, class SRC>
__device__ __forceinline__ void core_rates(const RhsArgs& a, const double x[3], long lslot, const SRC& src,
                                           double dh[NC], double dP[NC]) {
  constexpr bool MOVE = (bg_set_for(GAUGE) & BGS_MOVING) != 0;
  if constexpr (MOVE) {
    if (a.c.moving && a.c.interior == INT_DAMPED) {
      double hu[NC], Pu[NC];
#pragma unroll
      for (int v = 0; v < NC; ++v) {
        hu[v] = src.X(v)[0];
        Pu[v] = src.P(v);
      }
      if constexpr (HOLE == HOLE_BOXED) {
        // the target from the box (k_brick_layer) and its rates formed here as
        // core_target_rate forms them
#pragma clang fp contract(off)
        const CaseP& c = a.c;
        const double* lb = a.lbox + lslot;
        double hd[NC], Pd[NC], dthd[NC], dtPd[NC];
        const double w = smoothstep(4.0 * hole_radius(c, a.t, x) / c.r0 - 1.0);
        const double tp = a.t + c.delta, tm = a.t - c.delta;
        const double f = w / (tp - tm);
#pragma unroll
        for (int v = 0; v < NC; ++v) {
          hd[v] = lb[v * a.lB3];
          Pd[v] = lb[(NC + v) * a.lB3];
          dthd[v] = w == 0.0 ? 0.0 : f * (lb[(2 * NC + v) * a.lB3] - lb[(4 * NC + v) * a.lB3]);
          dtPd[v] = w == 0.0 ? 0.0 : f * (lb[(3 * NC + v) * a.lB3] - lb[(5 * NC + v) * a.lB3]);
        }
        moving_core_relax(a.c, hu, Pu, hd, Pd, dthd, dtPd, dh, dP);
      } else {
        moving_core_rates(a.c, a.t, x, hu, Pu, dh, dP);
      }
      return;
    }
  }
#pragma unroll
  for (int v = 0; v < NC; ++v) dh[v] = dP[v] = 0.0;
}

// The general interior (HOLE = 2, the block layout): the tracked geometry's
// radii along the ray or the sphere's, and the :fitted variant, whose core and
// layer relax toward the target cache, A + (t - t_f) S, with its rate S fed
// forward if asked and rho's ramp narrowed on the trailing side by trail
// (gh_rhs_point!, src/evolution.jl:500-700; _trail_profiles :1256-1272;
// _cached_target and _cached_rate, src/fit.jl:1050-1055).
__device__ __forceinline__ double fitted_target(const CaseP& c, const PointRef& pr, double t, int v, double& S) {
#pragma clang fp contract(off)
  const double A = c.tw[pr.ti0 + (long)v * pr.ssv];
  S = c.tw[pr.ti0 + (long)(20 + v) * pr.ssv];
  return A + (t - c.t_f) * S;
}
template <int GAUGE, int EPI, class SRC>
__device__ __forceinline__ void head_source_general(const RhsArgs& a, const PointRef& pr, const SRC& src,
                                                    const IGeo& g, const Coef& cf, const double hv[NC],
                                                    const double D[4][NC], PiCoef& pc, TGH_OUT_PARAMS) {
  const long sv = pr.sv, psz = pr.ssv, si0 = pr.si0, wi0 = pr.wi0;
  const CaseP& c = a.c;
  const double g0 = damping_rate(c, a.t, pr.x);
  double msrc[NC];
  {
    double Hl[4] = {0.0, 0.0, 0.0, 0.0}, dH[4][4];
    if (GAUGE == GAUGE_SAMPLED) {
      const double* Hp = pr.Hp;
#pragma unroll
      for (int q = 0; q < 4; ++q) Hl[q] = Hp[q * pr.hs];
#pragma unroll
      for (int n = 0; n < 16; ++n) dH[n % 4][n / 4] = Hp[(4 + n) * pr.hs];
    } else if (GAUGE == GAUGE_KS) {
      ks_source(c.bg.M, c.bg.a, hv, D, Hl, dH);
    } else if (GAUGE == GAUGE_KS_MOVING) {
      ks_source_moving(c.bg.M, c.ks_S, c.ks_u, hv, D, Hl, dH);
    }
    gh_source<gauge_has_source(GAUGE)>(cf, hv, D, Hl, dH, g0, c.g2, msrc);
  }
  if (pr.m_rho != nullptr) {
    const double rr = pr.m_rho[0];
    const double Sm[3] = {pr.m_S[0], pr.m_S[pr.ms], pr.m_S[2 * pr.ms]};
    double Sij[6];
#pragma unroll
    for (int n = 0; n < 6; ++n) Sij[n] = pr.m_Sij[n * pr.ms];
    double ms[NC];
    matter_source(cf, hv, rr, Sm, Sij, ms);
#pragma unroll
    for (int n = 0; n < NC; ++n) msrc[n] += ms[n];
  }
  const bool layer = g.r < g.r1;
  const bool fitted = c.interior == INT_FITTED;
  double wgt = 1.0, rho = 0.0;
  if (layer) {
#pragma clang fp contract(off)
    double rramp = c.rho_ramp;
    if (fitted && c.trail != 0.0) {
      double cc[3];
      center_at(c, a.t, cc);
      const double sp = sqrt(c.vel[0] * c.vel[0] + c.vel[1] * c.vel[1] + c.vel[2] * c.vel[2]);
      double zeta = 0.0;
      if (!(sp == 0.0 || g.r == 0.0))
        zeta = fmax(0.0, -((pr.x[0] - cc[0]) * c.vel[0] + (pr.x[1] - cc[1]) * c.vel[1] +
                           (pr.x[2] - cc[2]) * c.vel[2]) / (g.r * sp));
      rramp = c.rho_ramp * (1.0 - c.trail * zeta);
    }
    const double sl = (g.r - g.r0) / (g.r1 - g.r0);
    wgt = smoothstep(sl / c.w_ramp);
    rho = c.interior == INT_FROZEN ? 0.0 : c.rho_max * smoothstep((1.0 - sl) / rramp);
  }
  // the :damped / :frozen layer's target: the background at x (r >= r_0 here)
  double ue[20];
  if (layer && !fitted) background_state<bg_set_for(GAUGE)>(c.bg, a.t, pr.x, ue, ue + NC);
  const double om = 1.0 - wgt;
#pragma unroll 1
  for (int v = 0; v < NC; ++v) {
    double f = D[0][v];
    if (layer) {
      if (fitted) {
        double S;
        const double tg = fitted_target(c, pr, a.t, v, S);
        f = wgt * f + (c.rate ? om * S : 0.0) - rho * (hv[v] - tg);
      } else {
        f = wgt * f - rho * (hv[v] - ue[v]);
      }
    }
    emit<EPI>(a, TGH_OUT_ARGS, si0 + v * psz, wi0 + v * sv, f);
  }
  pc.w = layer ? wgt : 1.0;
#pragma unroll 1
  for (int v = 0; v < NC; ++v) {
    double c1 = msrc[v];
    if (layer) {
      if (fitted) {
        double S;
        const double tg = fitted_target(c, pr, a.t, NC + v, S);
        c1 = wgt * c1 + (c.rate ? om * S : 0.0) - rho * (src.P(v) - tg);
      } else {
        c1 = wgt * c1 - rho * (src.P(v) - ue[NC + v]);
      }
    }
    pc.c1[v] = c1;
  }
}

// The whole head at one point (gh_rhs_head, src/evolution.jl:258-297, inside
// gh_rhs_point!, :448-695,846-863). Returns false in the frozen core, where
// the core's d_t h is emitted and its d_t Pi left in pc.c1 (w = -1).
template <int Q, int GAUGE, int HOLE, int EPI, class SRC>
__device__ __forceinline__ bool head_point(const RhsArgs& a, const PointRef& pr, const SRC& src, double inv_h,
                                           PiCoef& pc, TGH_OUT_PARAMS) {
  if constexpr (HOLE == 2) {
    const CaseP& c = a.c;
    const IGeo g = interior_geo(c, a.t, pr.x);
    if (c.interior == INT_PASTED ? g.r < g.r1 : g.r < g.r0) {
      // the core: zero for a hole at rest, the relaxation toward the target
      // at rho_max for :fitted, du = [S] - rho_max (u - (A + (t - t_f) S))
      double dh[NC];
#pragma unroll 1
      for (int v = 0; v < NC; ++v) {
        dh[v] = 0.0;
        pc.c1[v] = 0.0;
        if (c.interior == INT_FITTED) {
#pragma clang fp contract(off)
          double S;
          double tg = fitted_target(c, pr, a.t, v, S);
          dh[v] = (c.rate ? S : 0.0) - c.rho_max * (src.X(v)[0] - tg);
          tg = fitted_target(c, pr, a.t, NC + v, S);
          pc.c1[v] = (c.rate ? S : 0.0) - c.rho_max * (src.P(v) - tg);
        }
        emit<EPI>(a, TGH_OUT_ARGS, pr.si0 + v * pr.ssv, pr.wi0 + v * pr.sv, dh[v]);
      }
      pc.w = -1.0;
      return false;
    }
    double hv[NC], D[4][NC];
    Coef cf;
    pc.epsh = dissipation_rate(c, a.t, pr.x) * inv_h;
    head_derivs<Q>(c, src, inv_h, pc.epsh, hv, D, cf);
    head_source_general<GAUGE, EPI>(a, pr, src, g, cf, hv, D, pc, TGH_OUT_ARGS);
    head_geometry(cf, D, pc);
    return true;
  }
  double r;
  if (frozen_core<HOLE>(a.c, a.t, pr.x, r)) {
    double dh[NC];
    core_rates<GAUGE, HOLE>(a, pr.x, pr.lslot, src, dh, pc.c1);
#pragma unroll 1
    for (int v = 0; v < NC; ++v) emit<EPI>(a, TGH_OUT_ARGS, pr.si0 + v * pr.ssv, pr.wi0 + v * pr.sv, dh[v]);
    pc.w = -1.0;
    return false;
  }
  double hv[NC], D[4][NC];
  Coef cf;
  pc.epsh = dissipation_rate(a.c, a.t, pr.x) * inv_h;
  head_derivs<Q>(a.c, src, inv_h, pc.epsh, hv, D, cf);
  head_source<GAUGE, HOLE, EPI>(a, pr, src, r, cf, hv, D, pc, TGH_OUT_ARGS);
  head_geometry(cf, D, pc);
  return true;
}

// The principal part of Pi component v (gh_rhs_pi, src/evolution.jl:324-354)
// from the stencils of h_v (ph) and Pi_v (pP), with strides 1, s2, s3 --
// of a work array or of a staged copy in LDS:
//   beta^i d_i Pi + d_i beta^i Pi + d_i A^{ij} d_j h + A^{ij} d_i d_j h + KO(Pi).
template <int Q>
__device__ __forceinline__ double pi_principal(const double* ph, const double* pP, int s2, int s3,
                                               const double b[3], double divb, const double dA[3],
                                               const double A[6], double ih, double ih2, double epsh,
                                               bool diss) {
  const int s1 = 1;
  const double Pi = pP[0];
  const double dh1 = ih * st_d1<Q>(ph, s1), dh2 = ih * st_d1<Q>(ph, s2), dh3 = ih * st_d1<Q>(ph, s3);
  const double dP1 = ih * st_d1<Q>(pP, s1), dP2 = ih * st_d1<Q>(pP, s2), dP3 = ih * st_d1<Q>(pP, s3);
  double s = b[0] * dP1 + b[1] * dP2 + b[2] * dP3 + divb * Pi + dA[0] * dh1 + dA[1] * dh2 + dA[2] * dh3;
  s += A[0] * (ih2 * st_d2<Q>(ph, s1)) + A[3] * (ih2 * st_d2<Q>(ph, s2)) + A[5] * (ih2 * st_d2<Q>(ph, s3));
  const double dxy = ih2 * st_mixed<Q>(ph, s1, s2);
  const double dxz = ih2 * st_mixed<Q>(ph, s1, s3);
  const double dyz = ih2 * st_mixed<Q>(ph, s2, s3);
  s += 2.0 * (A[1] * dxy + A[2] * dxz + A[4] * dyz);
  if (diss) s += epsh * (st_ko<Q>(pP, s1) + st_ko<Q>(pP, s2) + st_ko<Q>(pP, s3));
  return s;
}

// The fused right-hand side at one point in the streaming order of
// src/evolution.jl:258-388: the head, then the ten Pi components one at a time.
template <int Q, int GAUGE, int HOLE, int EPI>
__device__ __forceinline__ void rhs_point(const RhsArgs& a, const PointRef& pr, double inv_h, TGH_OUT_PARAMS) {
  PiCoef pc;
  const long sv = pr.sv, psz = pr.ssv;
  const GlobalSrc src{pr.w0, pr.sv, pr.s2, pr.s3};
  if (!head_point<Q, GAUGE, HOLE, EPI>(a, pr, src, inv_h, pc, TGH_OUT_ARGS)) {
#pragma unroll 1
    for (int v = 0; v < NC; ++v)
      emit<EPI>(a, TGH_OUT_ARGS, pr.si0 + (NC + v) * psz, pr.wi0 + (NC + v) * sv, pc.c1[v]);
    return;
  }
  const double ih = inv_h, ih2 = inv_h * inv_h, epsh = pc.epsh;
#pragma unroll 1
  for (int v = 0; v < NC; ++v) {
    const double* ph = pr.w0 + v * sv;
    const double s = pi_principal<Q>(ph, ph + NC * sv, pr.s2, pr.s3, pc.b, pc.divb, pc.divA, pc.A, ih, ih2,
                                     epsh, a.c.has_diss);
    emit<EPI>(a, TGH_OUT_ARGS, pr.si0 + (NC + v) * psz, pr.wi0 + (NC + v) * sv, pc.w * s + pc.c1[v]);
  }
}

// Coefficient slots of the split form, each an array of cs doubles indexed
// like the points: the PiCoef fields in order (w = -1 marks the frozen core).
enum { CF_B0 = 0, CF_DIVB = 3, CF_DIVA = 4, CF_A = 7, CF_W = 13, CF_C1 = 14, CF_EPS = 24, CF_N = 25 };
__device__ __forceinline__ void store_coef(double* __restrict__ cf0, long cs, const PiCoef& pc) {
#pragma unroll
  for (int n = 0; n < 3; ++n) cf0[(CF_B0 + n) * cs] = pc.b[n];
  cf0[CF_DIVB * cs] = pc.divb;
#pragma unroll
  for (int n = 0; n < 3; ++n) cf0[(CF_DIVA + n) * cs] = pc.divA[n];
#pragma unroll
  for (int n = 0; n < 6; ++n) cf0[(CF_A + n) * cs] = pc.A[n];
  cf0[CF_W * cs] = pc.w;
#pragma unroll
  for (int n = 0; n < NC; ++n) cf0[(CF_C1 + n) * cs] = pc.c1[n];
  cf0[CF_EPS * cs] = pc.epsh;
}
// The geometric and the source halves of the coefficients, stored as each is
// known.
__device__ __forceinline__ void store_coef_geometry(double* __restrict__ cf0, long cs, const PiCoef& pc) {
#pragma unroll
  for (int n = 0; n < 3; ++n) cf0[(CF_B0 + n) * cs] = pc.b[n];
  cf0[CF_DIVB * cs] = pc.divb;
#pragma unroll
  for (int n = 0; n < 3; ++n) cf0[(CF_DIVA + n) * cs] = pc.divA[n];
#pragma unroll
  for (int n = 0; n < 6; ++n) cf0[(CF_A + n) * cs] = pc.A[n];
  cf0[CF_EPS * cs] = pc.epsh;
}
__device__ __forceinline__ void store_coef_source(double* __restrict__ cf0, long cs, const PiCoef& pc) {
  cf0[CF_W * cs] = pc.w;
#pragma unroll
  for (int n = 0; n < NC; ++n) cf0[(CF_C1 + n) * cs] = pc.c1[n];
}
__device__ __forceinline__ void load_coef(const double* __restrict__ cf0, long cs, PiCoef& pc) {
#pragma unroll
  for (int n = 0; n < 3; ++n) pc.b[n] = cf0[(CF_B0 + n) * cs];
  pc.divb = cf0[CF_DIVB * cs];
#pragma unroll
  for (int n = 0; n < 3; ++n) pc.divA[n] = cf0[(CF_DIVA + n) * cs];
#pragma unroll
  for (int n = 0; n < 6; ++n) pc.A[n] = cf0[(CF_A + n) * cs];
  pc.w = cf0[CF_W * cs];
  pc.epsh = cf0[CF_EPS * cs];
}

// The block layout: one owned point of one block per thread (workgroup wx of
// block b).
__device__ __forceinline__ bool block_point_at(const RhsArgs& a, const double* __restrict__ W, int b, int wx,
                                               PointRef& pr, long& pidx) {
  const Geom& g = a.g;
  const int P = g.P, S = g.S, G = g.G;
  const int p = wx * blockDim.x + threadIdx.x;
  if (p >= P * P * P) return false;
  const int i = p % P, j = (p / P) % P, k = p / (P * P);
  const long sv = (long)S * S * S, psz = (long)P * P * P;
  const long off = (long)(i + G) + (long)S * ((j + G) + (long)S * (k + G));
  pr.w0 = W + (long)b * 20 * sv + off;
  pr.sv = sv;
  pr.s2 = S;
  pr.s3 = S * S;
  position(g, b, i, j, k, pr.x);
  pr.ih = block_inv_h(g, b);
  pr.ti0 = (long)b * 40 * psz + p;
  pr.si0 = (long)b * 20 * psz + p;
  pr.ssv = psz;
  pr.wi0 = (long)b * 20 * sv + off;
  pr.Hp = a.Hsrc ? a.Hsrc + pr.si0 : nullptr;
  pr.hs = psz;
  pr.m_rho = a.m_rho ? a.m_rho + (long)b * psz + p : nullptr;
  pr.m_S = a.m_S ? a.m_S + (long)b * 3 * psz + p : nullptr;
  pr.m_Sij = a.m_Sij ? a.m_Sij + (long)b * 6 * psz + p : nullptr;
  pr.ms = psz;
  pr.lslot = 0;
  if (a.lbox) {
    const int* Bc = g.bcoord + 3 * b;
    const int sh = a.lLmax - block_level(g, b);
    pr.lslot = layer_slot((Bc[0] * g.N + i) << sh, (Bc[1] * g.N + j) << sh, (Bc[2] * g.N + k) << sh, a.lB);
  }
  pidx = (long)b * psz + p;
  return true;
}
__device__ __forceinline__ bool block_point(const RhsArgs& a, const double* __restrict__ W, PointRef& pr,
                                            long& pidx) {
  int b, wx;
  if (!wg_map((a.g.P * a.g.P * a.g.P + 255) / 256, a.g.nblocks, b, wx)) return false;
  return block_point_at(a, W, b, wx, pr, pidx);
}

template <int Q, int GAUGE, int HOLE, int EPI>
__global__ __launch_bounds__(256) void k_rhs(RhsArgs a, const double* __restrict__ W, TGH_OUT_PARAMS) {
  a.t = stage_time(a.t, a.clock, a.toff);
  PointRef pr;
  long pi;
  if (!block_point(a, W, pr, pi)) return;
  rhs_point<Q, GAUGE, HOLE, EPI>(a, pr, pr.ih, TGH_OUT_ARGS);
}

// The split form, first half: the head at one owned point per thread, leaving
// the Pi components' coefficients in coef (CF_N arrays of nblocks (N+1)^3).
template <int Q, int GAUGE, int HOLE, int EPI>
__global__ __launch_bounds__(256) void k_head(RhsArgs a, const double* __restrict__ W,
                                              double* __restrict__ coef, TGH_OUT_PARAMS) {
  a.t = stage_time(a.t, a.clock, a.toff);
  PointRef pr;
  long pi;
  if (!block_point(a, W, pr, pi)) return;
  const long cs = (long)a.g.nblocks * a.g.P * a.g.P * a.g.P;
  PiCoef pc;
  const GlobalSrc src{pr.w0, pr.sv, pr.s2, pr.s3};
  if (head_point<Q, GAUGE, HOLE, EPI>(a, pr, src, pr.ih, pc, TGH_OUT_ARGS))
    store_coef(coef + pi, cs, pc);
  else
    store_coef_source(coef + pi, cs, pc);  // w = -1 and the core's d_t Pi
}

// The split form, second half: the ten Pi components at one owned point per
// thread, from the work array and the coefficients k_head left.
#ifndef TGH_PI_WAVES_PER_EU
#define TGH_PI_WAVES_PER_EU 2
#endif
template <int Q, int EPI>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(TGH_PI_WAVES_PER_EU))) void k_pi(
    RhsArgs a, const double* __restrict__ W, const double* __restrict__ coef, TGH_OUT_PARAMS) {
  a.t = stage_time(a.t, a.clock, a.toff);
  PointRef pr;
  long pi;
  if (!block_point(a, W, pr, pi)) return;
  const long cs = (long)a.g.nblocks * a.g.P * a.g.P * a.g.P;
  const long sv = pr.sv, psz = pr.ssv;
  const double* __restrict__ cf0 = coef + pi;
  PiCoef pc;
  load_coef(cf0, cs, pc);
  if (pc.w < 0.0) {  // the frozen core: its d_t Pi is the coefficient c1
#pragma unroll 1
    for (int v = 0; v < NC; ++v)
      emit<EPI>(a, TGH_OUT_ARGS, pr.si0 + (NC + v) * psz, pr.wi0 + (NC + v) * sv, cf0[(CF_C1 + v) * cs]);
    return;
  }
  const double ih = pr.ih, ih2 = ih * ih, epsh = pc.epsh;
#pragma unroll 1
  for (int v = 0; v < NC; ++v) {
    const double* ph = pr.w0 + v * sv;
    const double s = pi_principal<Q>(ph, ph + NC * sv, pr.s2, pr.s3, pc.b, pc.divb, pc.divA, pc.A, ih, ih2,
                                     epsh, a.c.has_diss);
    emit<EPI>(a, TGH_OUT_ARGS, pr.si0 + (NC + v) * psz, pr.wi0 + (NC + v) * sv,
              pc.w * s + cf0[(CF_C1 + v) * cs]);
  }
}

// Whether x is evolved (outside the layer's outer surface): the points the
// masked norms count.
__device__ __forceinline__ bool is_evolved(const CaseP& c, double t, const double x[3]) {
  if (c.interior == INT_NONE) return true;
  if (c.shaped) {
    const IGeo g = interior_geo(c, t, x);
    return g.r >= g.r1;
  }
  return hole_radius(c, t, x) >= c.r1;
}

#ifndef TGH_RHS_ONLY  // the pointwise kernels: built once, in the API's translation unit
// ---------------------------------------------------------------------------
// The ghost fill of a work array (prepare!, src/evolution.jl:1349-1357).
struct FillArgs {
  Geom g;
  CaseP c;
  double t;
  double* __restrict__ W;
  const double* clock;  // stage_time's clock and offset
  double toff;
};
template <bool MOVE>
__global__ __launch_bounds__(256) void k_fill(FillArgs a) {
  a.t = stage_time(a.t, a.clock, a.toff);
  const Geom& g = a.g;
  const int S = g.S, G = g.G;
  int b, wx;
  if (!wg_map((g.S * g.S * g.S + 255) / 256, g.nblocks, b, wx)) return;
  const int p = wx * blockDim.x + threadIdx.x;
  if (p >= S * S * S) return;
  const int l[3] = {p % S - G, (p / S) % S - G, p / (S * S) - G};
  int owner, loc[3], mm = 0, tn = 0;
  const int cl = classify(g, b, l, owner, loc, &mm, &tn);
  const long sv = (long)S * S * S;
  double* dst = a.W + (long)b * 20 * sv + p;
  if (cl == -1) {
    // a primary on a reflecting wall takes the parity: its odd variables are
    // zero (TreeAMR's projection of the wall at every fill of closed blocks)
    const int wm = wall_mask(g, b, l);
    if (wm) {
#pragma unroll
      for (int v = 0; v < 20; ++v)
        if (parity_odd_mask(v) & wm) dst[v * sv] = 0.0;
    }
    return;
  }
  if (cl == -2) {
    double x[3], s[20];
    position(g, b, l[0], l[1], l[2], x);
    case_state_any<MOVE ? BGS_ALL : BGS_STATIC>(a.c, a.t, x, s);
#pragma unroll
    for (int v = 0; v < 20; ++v) dst[v * sv] = s[v];
  } else {
    // the primary's value, projected as its own wall would be (the thread
    // that owns it may not have written the projection yet), times the
    // parities of the mirrors crossed
    const double* src = a.W + (long)owner * 20 * sv + (loc[0] + G) + (long)S * ((loc[1] + G) + (long)S * (loc[2] + G));
    const int sw = wall_mask(g, owner, loc);
    if (tn == 0) {
#pragma unroll
      for (int v = 0; v < 20; ++v)
        dst[v * sv] = (parity_odd_mask(v) & sw) ? 0.0 : parity_sign(v, mm) * src[v * sv];
    } else {
      // across the rotating seam: Q^tn of the (projected) source, then the mirrors
      for (int v = 0; v < 20; ++v) {
        int q;
        double sg;
        rot_power(v, tn, q, sg);
        dst[v * sv] = (parity_odd_mask(q) & sw) ? 0.0 : parity_sign(v, mm) * sg * src[q * sv];
      }
    }
  }
}

// The ghost fill of a forest with levels, from a plan the host builds once
// (one entry per work point that is not its block's primary, grouped in
// passes that the stream orders): pass 0 holds every copy (same-level,
// injection from a finer leaf, coincident coarse vertex: each reads a
// primary, which no entry writes) and every Dirichlet point; pass l >= 1 the
// prolongations into the ghosts of level-l blocks, each reading one level-
// (l-1) block's points, filled by the passes before. An entry is
//   dst = b S^3 + p  (block and work point), src = b' S^3 + p' (the source
//   point, or the prolongation stencil's first point), mode: bits 0-2 the
//   axes on which the point is a coarse midpoint (p points with the
//   midpoint Lagrange weights; none set: a copy), 8 the Dirichlet hook,
//   bits 4-6 the axes whose reflecting face the point lies beyond and bits
//   7-8 the quarter turns across the rotating seam (the value is then the
//   turned and mirrored one of the point inside).
enum { PLAN_DIRICHLET = 8 };
struct PlanOp {
  int dst, src, mode, pad;
};
struct PlanArgs {
  Geom g;
  CaseP c;
  double t;
  double* __restrict__ W;
  const PlanOp* __restrict__ ops;
  int n;
  const double* clock;  // stage_time's clock and offset
  double toff;
};
// The midpoint Lagrange weights of NP nodes at -NP/2+1 .. NP/2 (exact
// dyadic rationals, equal to the host's product formula).
template <int NP> struct MidW;
template <> struct MidW<2> { static constexpr double w[2] = {1.0 / 2, 1.0 / 2}; };
template <> struct MidW<4> { static constexpr double w[4] = {-1.0 / 16, 9.0 / 16, 9.0 / 16, -1.0 / 16}; };
template <> struct MidW<6> {
  static constexpr double w[6] = {3.0 / 256, -25.0 / 256, 150.0 / 256, 150.0 / 256, -25.0 / 256, 3.0 / 256};
};
template <> struct MidW<8> {
  static constexpr double w[8] = {-5.0 / 2048,    49.0 / 2048,  -245.0 / 2048, 1225.0 / 2048,
                                  1225.0 / 2048, -245.0 / 2048, 49.0 / 2048,   -5.0 / 2048};
};
template <> struct MidW<10> {
  static constexpr double w[10] = {35.0 / 65536,    -405.0 / 65536, 2268.0 / 65536,  -8820.0 / 65536,
                                   39690.0 / 65536, 39690.0 / 65536, -8820.0 / 65536, 2268.0 / 65536,
                                   -405.0 / 65536,  35.0 / 65536};
};
// One variable's tensor-product interpolation from the stencil's first point f
// (strides 1, S, S^2): NX, NY, NZ points per axis (1 or NP), sum-factorised
// with x innermost, each sum from 0 in stencil order, no contraction.
template <int NX, int NY, int NZ, int NP>
__device__ __forceinline__ double prolong_sum(const double* __restrict__ f, int S) {
#pragma clang fp contract(off)
  double sz = 0.0;
#pragma unroll
  for (int c2 = 0; c2 < NZ; ++c2) {
    double sy = 0.0;
#pragma unroll
    for (int c1 = 0; c1 < NY; ++c1) {
      double sx = 0.0;
#pragma unroll
      for (int c0 = 0; c0 < NX; ++c0) sx += (NX > 1 ? MidW<NP>::w[c0] : 1.0) * f[c0 + S * (c1 + S * c2)];
      sy += (NY > 1 ? MidW<NP>::w[c1] : 1.0) * sx;
    }
    sz += (NZ > 1 ? MidW<NP>::w[c2] : 1.0) * sy;
  }
  return sz;
}
// A plan entry's value: a copy (mode 0) or the prolongation over the
// midpoint axes that the mode's bits name (x 1, y 2, z 4).
template <int NP>
__device__ __forceinline__ double prolong_mode(int mode, const double* __restrict__ f, int S) {
  switch (mode) {
    case 0: return prolong_sum<1, 1, 1, NP>(f, S);
    case 1: return prolong_sum<NP, 1, 1, NP>(f, S);
    case 2: return prolong_sum<1, NP, 1, NP>(f, S);
    case 3: return prolong_sum<NP, NP, 1, NP>(f, S);
    case 4: return prolong_sum<1, 1, NP, NP>(f, S);
    case 5: return prolong_sum<NP, 1, NP, NP>(f, S);
    case 6: return prolong_sum<1, NP, NP, NP>(f, S);
    default: return prolong_sum<NP, NP, NP, NP>(f, S);
  }
}
// The copies and prolongations: one thread per entry and variable
// (blockIdx.y = v); a pass's entries are sorted by mode on the host, so a
// wave takes one branch.
template <int NP>
__global__ __launch_bounds__(256) void k_fill_plan(PlanArgs a) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= a.n) return;
  const int v = blockIdx.y;
  const PlanOp o = a.ops[e];
  if (o.mode == PLAN_DIRICHLET) return;  // k_fill_plan_dirichlet's
  const Geom& g = a.g;
  const int S = g.S;
  const long S3 = (long)S * S * S;
  const int b = o.dst / (int)S3, p = o.dst - b * (int)S3;
  double* dst = a.W + (long)b * 20 * S3 + p;
  const int sb = o.src / (int)S3, sp = o.src - sb * (int)S3;
  const int mm = (o.mode >> 4) & 7, tn = (o.mode >> 7) & 3;
  int sv = v;
  double sg = 1.0;
  if (mm | tn) {
    rot_power(v, tn, sv, sg);
    sg *= parity_sign(v, mm);
  }
  const double* f = a.W + (long)sb * 20 * S3 + sp + (long)sv * S3;
  const double val = prolong_mode<NP>(o.mode & 7, f, S);
  dst[v * S3] = (mm | tn) ? sg * val : val;
}
// The regrid's transfer: one thread per (entry, variable) of a transfer
// plan, the new forest's state point from the old forest's prepared work
// array -- a copy or a prolongation, as the old forest's fill computes it.
struct TransferArgs {
  const double* __restrict__ W;
  double* __restrict__ U;
  const PlanOp* __restrict__ ops;
  int n, S, P3;
};
template <int NP>
__global__ __launch_bounds__(256) void k_transfer(TransferArgs a) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= a.n) return;
  const int v = blockIdx.y;
  const PlanOp o = a.ops[e];
  const long S3 = (long)a.S * a.S * a.S;
  const int b = o.dst / a.P3, p = o.dst - b * a.P3;
  const int sb = o.src / (int)S3, sp = o.src - sb * (int)S3;
  const double* f = a.W + (long)sb * 20 * S3 + sp + (long)v * S3;
  a.U[((long)b * 20 + v) * a.P3 + p] = prolong_mode<NP>(o.mode, f, a.S);
}
// The Dirichlet entries: one thread per entry, the exact state at the
// ghost's own position.
template <bool MOVE>
__global__ __launch_bounds__(256) void k_fill_plan_dirichlet(PlanArgs a) {
  a.t = stage_time(a.t, a.clock, a.toff);
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= a.n) return;
  const PlanOp o = a.ops[e];
  if (o.mode != PLAN_DIRICHLET) return;
  const Geom& g = a.g;
  const int S = g.S, G = g.G;
  const long S3 = (long)S * S * S;
  const int b = o.dst / (int)S3, p = o.dst - b * (int)S3;
  double* dst = a.W + (long)b * 20 * S3 + p;
  double x[3], s[20];
  position(g, b, p % S - G, (p / S) % S - G, p / (S * S) - G, x);
  case_state_any<MOVE ? BGS_ALL : BGS_STATIC>(a.c, a.t, x, s);
#pragma unroll
  for (int q = 0; q < 20; ++q) dst[q * S3] = s[q];
}

// The reflecting walls' odd variables in a forest's work array, zeroed at the
// start of every fill (before the plan's copies and prolongations read them).
struct WallArgs {
  Geom g;
  double* __restrict__ W;
};
__global__ __launch_bounds__(256) void k_wall_work(WallArgs a) {
  const Geom& g = a.g;
  const int P = g.P, S = g.S, G = g.G;
  int b, wx;
  if (!wg_map((g.P * g.P * g.P + 255) / 256, g.nblocks, b, wx)) return;
  const int p = wx * blockDim.x + threadIdx.x;
  if (p >= P * P * P) return;
  const int l[3] = {p % P, (p / P) % P, p / (P * P)};
  const int wm = wall_mask(g, b, l);
  if (!wm) return;
  const long sv = (long)S * S * S;
  double* w = a.W + (long)b * 20 * sv + (l[0] + G) + (long)S * ((l[1] + G) + (long)S * (l[2] + G));
  for (int v = 0; v < 20; ++v)
    if (parity_odd_mask(v) & wm) w[v * sv] = 0.0;
}

// scatter! / gather! between the state and a work array's owned points.
struct CopyArgs {
  Geom g;
  const double* __restrict__ src;
  double* __restrict__ dst;
  int to_work;
};
__global__ __launch_bounds__(256) void k_scatter(CopyArgs a) {
  const Geom& g = a.g;
  const int P = g.P, S = g.S, G = g.G;
  int b, wx;
  if (!wg_map((g.P * g.P * g.P + 255) / 256, g.nblocks, b, wx)) return;
  const int p = wx * blockDim.x + threadIdx.x;
  if (p >= P * P * P) return;
  const int i = p % P, j = (p / P) % P, k = p / (P * P);
  const long sv = (long)S * S * S, psz = (long)P * P * P;
  const long wi = (long)b * 20 * sv + (i + G) + (long)S * ((j + G) + (long)S * (k + G));
  const long si = (long)b * 20 * psz + p;
#pragma unroll 4
  for (int v = 0; v < 20; ++v) {
    if (a.to_work)
      a.dst[wi + v * sv] = a.src[si + v * psz];
    else
      a.dst[si + v * psz] = a.src[wi + v * sv];
  }
}

// sample_gauge_source! (src/gauge.jl:156-184): H_b and d_a H_b of the
// background at every owned point, through the core rule, into Hsrc (state
// layout, 20 variables: H_b then d_a H_b in column-major (a, b) order).
struct SampleArgs {
  Geom g;
  CaseP c;
  double t;
  double* __restrict__ Hsrc;
};
__global__ __launch_bounds__(256) void k_sample_gauge(SampleArgs a) {
  const Geom& g = a.g;
  const int P = g.P;
  int b, wx;
  if (!wg_map((g.P * g.P * g.P + 255) / 256, g.nblocks, b, wx)) return;
  const int p = wx * blockDim.x + threadIdx.x;
  if (p >= P * P * P) return;
  const int i = p % P, j = (p / P) % P, k = p / (P * P);
  const long psz = (long)P * P * P;
  double x[3], xc[3];
  position(g, b, i, j, k, x);
  if (a.c.shaped && a.c.interior != INT_NONE) shaped_core_point(a.c, a.t, x, xc);
  else core_point(a.c, a.t, x, xc);
  double Hl[4], dH[4][4];
  gauge_source_grad(a.c.bg, a.t, xc, Hl, dH);
  double* out = a.Hsrc + (long)b * 20 * psz + p;
  for (int q = 0; q < 4; ++q) out[q * psz] = Hl[q];
  for (int n = 0; n < 16; ++n) out[(4 + n) * psz] = dH[n % 4][n / 4];
}

// Pointwise kernels over owned points of a state vector.
struct PointArgs {
  Geom g;
  CaseP c;
  double t;
  double* __restrict__ u;
  uint64_t base;
  double amp;
  int mode;  // 0 exact state, 1 add noise, 2 paste inside r1, 3 project the reflecting walls
};
template <bool MOVE>
__global__ __launch_bounds__(256) void k_pointwise(PointArgs a) {
  const Geom& g = a.g;
  const int P = g.P;
  int b, wx;
  if (!wg_map((g.P * g.P * g.P + 255) / 256, g.nblocks, b, wx)) return;
  const int p = wx * blockDim.x + threadIdx.x;
  if (p >= P * P * P) return;
  const int i = p % P, j = (p / P) % P, k = p / (P * P);
  const long psz = (long)P * P * P;
  double* ub = a.u + (long)b * 20 * psz + p;
  double x[3];
  position(g, b, i, j, k, x);
  if (a.mode == 1) {
    // on a low reflecting wall the noise is projected onto the parity: the
    // odd variables there are zero (add_noise!, src/initialdata.jl:1938-1990,
    // which projects the low walls only)
    const int l[3] = {i, j, k};
    int low = 0;
    const int* B = g.bcoord + 3 * b;
#pragma unroll
    for (int d = 0; d < 3; ++d)
      if (g.refl[d][0] && B[d] * g.N + l[d] == 0) low |= 1 << d;
    // the two seam planes are the same points a quarter turn apart: they keep their data
    const bool seam = g.rot && ((B[0] * g.N + l[0]) == 0 || (B[1] * g.N + l[1]) == 0);
    for (int v = 0; v < 20; ++v) {
      if (!seam) ub[v * psz] += a.amp * (2.0 * point_draw(a.base, x, v + 1) - 1.0);
      if (parity_odd_mask(v) & low) ub[v * psz] = 0.0;
    }
    return;
  }
  if (a.mode == 3) {
    // settle!'s wall: the odd variables of every reflecting wall plane zero
    const int l[3] = {i, j, k};
    const int wm = wall_mask(g, b, l);
    if (wm)
      for (int v = 0; v < 20; ++v)
        if (parity_odd_mask(v) & wm) ub[v * psz] = 0.0;
    return;
  }
  if (a.mode == 2 && !(hole_radius(a.c, a.t, x) < a.c.r1)) return;
  double s[20];
  case_state_any<MOVE ? BGS_ALL : BGS_STATIC>(a.c, a.t, x, s);
  for (int v = 0; v < 20; ++v) ub[v * psz] = s[v];
}

// Reductions over owned points: each workgroup of 256 points of one block
// folds its points in shared memory and writes RED_STRIDE partials; the host
// folds the partials in (block, workgroup) order, so the answer does not
// depend on the launch.
//  mode 0 (max_speed): [0] the largest speed over evolved points, [1] the
//    number of evolved points with a non-finite value;
//  mode 1 (gh_error_kernel! + error_norms, src/constraints.jl:311-396,
//    629-671): at every primary the error e = |u - u_exact(t)|, the 2-norm
//    over the 20 variables summed in variable order; [0] sum w e^2 and [1]
//    sum w over evolved primaries (w the volume relative to level 0's on a
//    forest, 1 on a uniform one), [2] the largest e there, [3] the layer's
//    residual (the largest e at r_0 <= r < r_1; for :fitted the distance
//    from the target cache), [4] the drift (the largest |h_tt - h_tt,exact|
//    at evolved primaries with shell_lo <= r <= shell_hi).
enum { RED_STRIDE = 5 };
struct ReduceArgs {
  Geom g;
  CaseP c;
  double t;
  const double* __restrict__ u;
  double* __restrict__ out;  // RED_STRIDE partials per workgroup
  int mode;                  // 0 speed, 1 error rows
  double shell_lo, shell_hi;
};
// the point's radius and whether it lies in the layer r_0 <= r < r_1
__device__ __forceinline__ bool in_layer(const CaseP& c, double t, const double x[3], double& r) {
  if (c.interior == INT_NONE) {
    r = 0.0;
    return false;
  }
  if (c.shaped) {
    const IGeo g = interior_geo(c, t, x);
    r = g.r;
    return g.r0 <= g.r && g.r < g.r1;
  }
  r = hole_radius(c, t, x);
  return c.r0 <= r && r < c.r1;
}
template <bool MOVE>
__global__ __launch_bounds__(256) void k_reduce(ReduceArgs a) {
  const Geom& g = a.g;
  const int P = g.P;
  int b, wx;
  if (!wg_map((g.P * g.P * g.P + 255) / 256, g.nblocks, b, wx)) return;
  const int p = wx * blockDim.x + threadIdx.x;
  const long psz = (long)P * P * P;
  double r[RED_STRIDE] = {0.0, 0.0, 0.0, 0.0, 0.0};
  if (p < P * P * P) {
    const int i = p % P, j = (p / P) % P, k = p / (P * P);
    const double* ub = a.u + (long)b * 20 * psz + p;
    double x[3];
    position(g, b, i, j, k, x);
    const bool keep = is_evolved(a.c, a.t, x);
    if (a.mode == 0) {
      if (keep) {
        double hv[NC];
        bool fin = true;
#pragma unroll
        for (int v = 0; v < 20; ++v) {
          const double val = ub[v * psz];
          fin = fin && isfinite(val);
          if (v < NC) hv[v] = val;
        }
        r[1] = fin ? 0.0 : 1.0;
        Coef cf;
        metric_quantities(hv, cf);
        r[0] = cf.alpha * sqrt(cf.gamu[0] + cf.gamu[3] + cf.gamu[5]) +
               sqrt(cf.beta[0] * cf.beta[0] + cf.beta[1] * cf.beta[1] + cf.beta[2] * cf.beta[2]);
      }
    } else {
      bool prim;
      if (g.prim) {
        prim = g.prim[(long)b * psz + p] != 0;
      } else {
        int l[3] = {i, j, k}, owner, loc[3];
        prim = classify(g, b, l, owner, loc) == -1;
      }
      double rad;
      const bool lay = in_layer(a.c, a.t, x, rad);
      if (prim && (keep || lay)) {
        double s[20];
        case_state_any<MOVE ? BGS_ALL : BGS_STATIC>(a.c, a.t, x, s);
        double e2 = 0.0;
#pragma unroll
        for (int v = 0; v < 20; ++v) {
          const double e = ub[v * psz] - s[v];
          e2 += e * e;
        }
        const double e = sqrt(e2);
        if (keep) {
          const double wv = g.prim ? ldexp(1.0, -3 * block_level(g, b)) : 1.0;
          r[0] = g.prim ? e2 * wv : e2;
          r[1] = wv;
          r[2] = e;
          if (a.shell_lo <= rad && rad <= a.shell_hi) r[4] = fabs(ub[0] - s[0]);
        }
        if (lay) {
          double ef = e;
          if (a.c.interior == INT_FITTED && a.c.tw) {
            const long ti = (long)b * 40 * psz + p;
            double f2 = 0.0;
            for (int v = 0; v < 20; ++v) {
#pragma clang fp contract(off)
              const double tg = a.c.tw[ti + v * psz] + (a.t - a.c.t_f) * a.c.tw[ti + (20 + v) * psz];
              const double d = ub[v * psz] - tg;
              f2 += d * d;
            }
            ef = sqrt(f2);
          }
          r[3] = ef;
        }
      }
    }
  }
  __shared__ double sh[RED_STRIDE][256];
  const int tid = threadIdx.x;
#pragma unroll
  for (int q = 0; q < RED_STRIDE; ++q) sh[q][tid] = r[q];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (tid < w) {
      if (a.mode == 0) {
        const double o0 = sh[0][tid + w], m0 = sh[0][tid];
        sh[0][tid] = (o0 > m0 || isnan(o0)) ? o0 : m0;
        sh[1][tid] = sh[1][tid] + sh[1][tid + w];
      } else {
        sh[0][tid] = sh[0][tid] + sh[0][tid + w];
        sh[1][tid] = sh[1][tid] + sh[1][tid + w];
#pragma unroll
        for (int q = 2; q < RED_STRIDE; ++q) {
          const double o = sh[q][tid + w], m = sh[q][tid];
          sh[q][tid] = (o > m || isnan(o)) ? o : m;
        }
      }
    }
    __syncthreads();
  }
  if (tid == 0) {
    double* o = a.out + RED_STRIDE * ((long)b * ((g.P * g.P * g.P + 255) / 256) + wx);
#pragma unroll
    for (int q = 0; q < RED_STRIDE; ++q) o[q] = sh[q][0];
  }
}

// The gauge constraint monitor (gh_constraint_kernel!, src/constraints.jl:
// 211-309; gauge_constraint_at_node, src/pointwise.jl:947-960) at every owned
// point of a prepared work array: C_a = g_ab (Gamma^b + g^{bc} H_c), Gamma^b =
// g^{cd} g^{be} Gamma_{ecd}, with d_t h = beta^i d_i h + (alpha / sqrt gamma) Pi
// (no dissipation) and the case's source; each workgroup folds its points'
// sums of w C_a^2 (w the volume weight on a forest, 1 on a uniform one), the
// maxima of |C_a| and the weights into 9 partials, over the points the
// error norms count.
struct GCArgs {
  Geom g;
  CaseP c;
  double t;
  const double* __restrict__ W;
  const double* __restrict__ Hsrc;
  double* __restrict__ out;
};
template <int Q>
__global__ __launch_bounds__(256) void k_gauge_constraint(GCArgs a) {
  const Geom& g = a.g;
  const int P = g.P, S = g.S, G = g.G;
  int b, wx;
  if (!wg_map((P * P * P + 255) / 256, g.nblocks, b, wx)) return;
  const int p = wx * blockDim.x + threadIdx.x;
  const long psz = (long)P * P * P, sv = (long)S * S * S;
  double r[9];
#pragma unroll
  for (int n = 0; n < 9; ++n) r[n] = 0.0;
  if (p < P * P * P) {
    const int i = p % P, j = (p / P) % P, k = p / (P * P);
    double x[3];
    position(g, b, i, j, k, x);
    const CaseP& c = a.c;
    if (is_evolved(c, a.t, x)) {
      int l[3] = {i, j, k}, owner, loc[3];
      const bool prim = g.prim ? g.prim[(long)b * psz + p] != 0 : classify(g, b, l, owner, loc) == -1;
      if (!(g.prim && !prim)) {
        const double* w0 = a.W + (long)b * 20 * sv + (i + G) + (long)S * ((j + G) + (long)S * (k + G));
        const double ih = block_inv_h(g, b);
        double hv[NC], D[4][NC];
        for (int v = 0; v < NC; ++v) {
          const double* wv = w0 + v * sv;
          hv[v] = wv[0];
          D[1][v] = ih * st_d1<Q>(wv, 1);
          D[2][v] = ih * st_d1<Q>(wv, S);
          D[3][v] = ih * st_d1<Q>(wv, S * S);
        }
        Coef cf;
        metric_quantities(hv, cf);
        const double ad = cf.alpha / cf.sqrtg;
        for (int v = 0; v < NC; ++v)
          D[0][v] = cf.beta[0] * D[1][v] + cf.beta[1] * D[2][v] + cf.beta[2] * D[3][v] + ad * w0[(NC + v) * sv];
        double Hl[4] = {0.0, 0.0, 0.0, 0.0}, dH[4][4];
        if (c.gauge == GAUGE_SAMPLED) {
          for (int q = 0; q < 4; ++q) Hl[q] = a.Hsrc[(long)b * 20 * psz + p + q * psz];
        } else if (c.gauge == GAUGE_KS) {
          ks_source(c.bg.M, c.bg.a, hv, D, Hl, dH);
        } else if (c.gauge == GAUGE_KS_MOVING) {
          ks_source_moving(c.bg.M, c.ks_S, c.ks_u, hv, D, Hl, dH);
        }
        // T_e = g^{cd} Gamma_{ecd} = g^{cd} (d_d g_ec - d_e g_cd / 2)
        double Te[4];
        for (int e = 0; e < 4; ++e) {
          double s = 0.0;
          for (int cc = 0; cc < 4; ++cc)
            for (int d = 0; d < 4; ++d)
              s += cf.gu[pidx(cc, d)] * (0.5 * (D[d][pidx(e, cc)] + D[cc][pidx(e, d)]) - 0.5 * D[e][pidx(cc, d)]);
          Te[e] = s;
        }
        double Cu[4];
        for (int m = 0; m < 4; ++m) {
          double gam = 0.0, hu = 0.0;
          for (int e = 0; e < 4; ++e) {
            gam += cf.gu[pidx(m, e)] * Te[e];
            hu += cf.gu[pidx(m, e)] * Hl[e];
          }
          Cu[m] = gam + hu;
        }
        const double wgt = g.blevel ? ldexp(1.0, -3 * block_level(g, b)) : 1.0;
        for (int q = 0; q < 4; ++q) {
          double C = 0.0;
          for (int m = 0; m < 4; ++m) C += ((q == m ? (q == 0 ? -1.0 : 1.0) : 0.0) + hv[pidx(q, m)]) * Cu[m];
          r[4 + q] = fabs(C);
          if (prim) r[q] = C * C * wgt;
        }
        if (prim) r[8] = wgt;
      }
    }
  }
  __shared__ double sh[9][256];
  const int tid = threadIdx.x;
#pragma unroll
  for (int n = 0; n < 9; ++n) sh[n][tid] = r[n];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (tid < w) {
#pragma unroll
      for (int n = 0; n < 9; ++n)
        sh[n][tid] = (n >= 4 && n < 8) ? fmax(sh[n][tid], sh[n][tid + w]) : sh[n][tid] + sh[n][tid + w];
    }
    __syncthreads();
  }
  if (tid < 9) a.out[9 * ((long)b * ((P * P * P + 255) / 256) + wx) + tid] = sh[tid][0];
}

// The ADM monitor (adm_constraint_kernel!, src/constraints.jl:432-534, and
// adm_constraints_at_node, :100-210) at every owned point: d_t g and d_t Pi
// from the evolution equations without dissipation, d_i d_t g and d_t d_t g
// by differentiating the first one (metric_derivatives_along,
// src/pointwise.jl:530-545), the textbook Ricci tensor of the 4-metric, and
// H = 2 R_nn + R, M_i = -n^a (R_ai - g_ai R / 2). Partials as the gauge
// monitor's, 9 per workgroup: sums of w H^2 and w M_i^2, maxima of |H| and
// |M_i|, the weights.
__device__ __forceinline__ double g4_at(const double h[NC], int a, int b) {
  return (a == b ? (a == 0 ? -1.0 : 1.0) : 0.0) + h[pidx(a, b)];
}
// d alpha, d beta^j, d sqrt(gamma) along a derivative dg of g (packed).
__device__ __forceinline__ void adm_along(const Coef& cf, const double dg[NC], double& da, double db[3],
                                          double& dsq) {
  double t00 = 0.0, t0[3] = {0.0, 0.0, 0.0};
  for (int c = 0; c < 4; ++c)
    for (int d = 0; d < 4; ++d) {
      const double x = cf.gu[pidx(0, c)] * dg[pidx(c, d)];
      t00 += x * cf.gu[pidx(d, 0)];
      for (int j = 0; j < 3; ++j) t0[j] += x * cf.gu[pidx(d, j + 1)];
    }
  const double dgu00 = -t00;
  da = cf.alpha * cf.alpha * cf.alpha * dgu00 / 2.0;
  for (int j = 0; j < 3; ++j) db[j] = -(-t0[j] + cf.beta[j] * dgu00) / cf.gu[0];
  double sq = 0.0;
  for (int j = 0; j < 3; ++j)
    for (int k = 0; k < 3; ++k) sq += cf.gamu[pidx(j + 1, k + 1) - 4] * dg[pidx(j + 1, k + 1)];
  dsq = cf.sqrtg * sq / 2.0;
}
template <int Q>
__global__ __launch_bounds__(256) void k_adm_constraint(GCArgs a) {
  const Geom& g = a.g;
  const int P = g.P, S = g.S, G = g.G;
  int b, wx;
  if (!wg_map((P * P * P + 255) / 256, g.nblocks, b, wx)) return;
  const int p = wx * blockDim.x + threadIdx.x;
  const long psz = (long)P * P * P, sv = (long)S * S * S;
  double r[9];
  for (int n = 0; n < 9; ++n) r[n] = 0.0;
  if (p < P * P * P) {
    const int i = p % P, j = (p / P) % P, k = p / (P * P);
    double x[3];
    position(g, b, i, j, k, x);
    const CaseP& c = a.c;
    if (is_evolved(c, a.t, x)) {
      int l[3] = {i, j, k}, owner, loc[3];
      const bool prim = g.prim ? g.prim[(long)b * psz + p] != 0 : classify(g, b, l, owner, loc) == -1;
      if (!(g.prim && !prim)) {
        const double* w0 = a.W + (long)b * 20 * sv + (i + G) + (long)S * ((j + G) + (long)S * (k + G));
        const double ih = block_inv_h(g, b), ih2 = ih * ih;
        const int st[3] = {1, S, S * S};
        double hv[NC], Pv[NC], D[4][NC], dP[3][NC], dd[10][NC];  // dd: tt, tx, ty, tz, xx, xy, xz, yy, yz, zz
        for (int v = 0; v < NC; ++v) {
          const double* wh = w0 + v * sv;
          const double* wp = w0 + (NC + v) * sv;
          hv[v] = wh[0];
          Pv[v] = wp[0];
          for (int d = 0; d < 3; ++d) {
            D[1 + d][v] = ih * st_d1<Q>(wh, st[d]);
            dP[d][v] = ih * st_d1<Q>(wp, st[d]);
          }
          for (int d1 = 0; d1 < 3; ++d1)
            for (int d2 = d1; d2 < 3; ++d2)
              dd[pidx(d1 + 1, d2 + 1)][v] =
                  ih2 * (d1 == d2 ? st_d2<Q>(wh, st[d1]) : st_mixed<Q>(wh, st[d1], st[d2]));
        }
        Coef cf;
        metric_quantities(hv, cf);
        const double phi = cf.alpha / cf.sqrtg;
        for (int v = 0; v < NC; ++v)
          D[0][v] = cf.beta[0] * D[1][v] + cf.beta[1] * D[2][v] + cf.beta[2] * D[3][v] + phi * Pv[v];
        double Hl[4] = {0.0, 0.0, 0.0, 0.0}, dH[4][4];
        for (int q = 0; q < 16; ++q) dH[q % 4][q / 4] = 0.0;
        if (c.gauge == GAUGE_SAMPLED) {
          const double* Hp = a.Hsrc + (long)b * 20 * psz + p;
          for (int q = 0; q < 4; ++q) Hl[q] = Hp[q * psz];
          for (int n = 0; n < 16; ++n) dH[n % 4][n / 4] = Hp[(4 + n) * psz];
        } else if (c.gauge == GAUGE_KS) {
          ks_source(c.bg.M, c.bg.a, hv, D, Hl, dH);
        } else if (c.gauge == GAUGE_KS_MOVING) {
          ks_source_moving(c.bg.M, c.ks_S, c.ks_u, hv, D, Hl, dH);
        }
        double msrc[NC];
        gh_source<true>(cf, hv, D, Hl, dH, damping_rate(c, a.t, x), c.g2, msrc);
        PiCoef pc;
        head_geometry(cf, D, pc);
        double dtP[NC];
        for (int v = 0; v < NC; ++v)
          dtP[v] = pc.b[0] * dP[0][v] + pc.b[1] * dP[1][v] + pc.b[2] * dP[2][v] + pc.divb * Pv[v] +
                   pc.divA[0] * D[1][v] + pc.divA[1] * D[2][v] + pc.divA[2] * D[3][v] + pc.A[0] * dd[4][v] +
                   pc.A[3] * dd[7][v] + pc.A[5] * dd[9][v] +
                   2.0 * (pc.A[1] * dd[5][v] + pc.A[2] * dd[6][v] + pc.A[4] * dd[8][v]) + msrc[v];
        // d_i d_t g, then d_t d_t g
        for (int ii = 0; ii < 3; ++ii) {
          double da, db[3], dsq;
          adm_along(cf, D[1 + ii], da, db, dsq);
          const double dphi = (da - phi * dsq) / cf.sqrtg;
          for (int v = 0; v < NC; ++v)
            dd[1 + ii][v] = (dphi * Pv[v] + phi * dP[ii][v]) +
                            (db[0] * D[1][v] + db[1] * D[2][v] + db[2] * D[3][v]) +
                            (cf.beta[0] * dd[pidx(ii + 1, 1)][v] + cf.beta[1] * dd[pidx(ii + 1, 2)][v] +
                             cf.beta[2] * dd[pidx(ii + 1, 3)][v]);
        }
        {
          double da, db[3], dsq;
          adm_along(cf, D[0], da, db, dsq);
          const double dphi = (da - phi * dsq) / cf.sqrtg;
          for (int v = 0; v < NC; ++v)
            dd[0][v] = dphi * Pv[v] + phi * dtP[v] + (db[0] * D[1][v] + db[1] * D[2][v] + db[2] * D[3][v]) +
                       (cf.beta[0] * dd[1][v] + cf.beta[1] * dd[2][v] + cf.beta[2] * dd[3][v]);
        }
        // dgu[e] = -g^-1 (d_e g) g^-1, Gamma_{a,bc}, Gamma^a_bc, Gamma^c_cd
        double dgu[4][NC], Gl[4][NC], G4[4][NC], Gtr[4];
        for (int e = 0; e < 4; ++e)
          for (int a2 = 0; a2 < 4; ++a2)
            for (int b2 = a2; b2 < 4; ++b2) {
              double s2 = 0.0;
              for (int cc = 0; cc < 4; ++cc)
                for (int d = 0; d < 4; ++d) s2 += cf.gu[pidx(a2, cc)] * D[e][pidx(cc, d)] * cf.gu[pidx(d, b2)];
              dgu[e][pidx(a2, b2)] = -s2;
            }
        for (int a2 = 0; a2 < 4; ++a2)
          for (int b2 = 0; b2 < 4; ++b2)
            for (int c2 = b2; c2 < 4; ++c2)
              Gl[a2][pidx(b2, c2)] = (D[b2][pidx(a2, c2)] + D[c2][pidx(a2, b2)] - D[a2][pidx(b2, c2)]) / 2;
        for (int a2 = 0; a2 < 4; ++a2)
          for (int n = 0; n < NC; ++n) {
            double s2 = 0.0;
            for (int xx = 0; xx < 4; ++xx) s2 += cf.gu[pidx(a2, xx)] * Gl[xx][n];
            G4[a2][n] = s2;
          }
        for (int d = 0; d < 4; ++d) {
          double s2 = 0.0;
          for (int cc = 0; cc < 4; ++cc) s2 += G4[cc][pidx(cc, d)];
          Gtr[d] = s2;
        }
        // d_m d_n g_ab
        auto DD = [&](int m, int n, int a2, int b2) { return dd[pidx(m, n)][pidx(a2, b2)]; };
        // R_ab is symmetric: its ten components a <= b, mirrored
        double R[4][4];
        for (int a2 = 0; a2 < 4; ++a2)
          for (int b2 = a2; b2 < 4; ++b2) {
            double A = 0.0, B = 0.0, Cc = 0.0, Dd = 0.0;
#pragma unroll
            for (int cc = 0; cc < 4; ++cc)
#pragma unroll
              for (int d = 0; d < 4; ++d) {
                const double gcd = cf.gu[pidx(cc, d)];
                A += dgu[cc][pidx(cc, d)] * Gl[d][pidx(a2, b2)] +
                     gcd * (DD(cc, a2, d, b2) + DD(cc, b2, d, a2) - DD(cc, d, a2, b2)) / 2;
                B += dgu[b2][pidx(cc, d)] * Gl[d][pidx(cc, a2)] +
                     gcd * (DD(b2, cc, d, a2) + DD(b2, a2, d, cc) - DD(b2, d, cc, a2)) / 2;
                Dd += G4[cc][pidx(b2, d)] * G4[d][pidx(cc, a2)];
              }
#pragma unroll
            for (int d = 0; d < 4; ++d) Cc += Gtr[d] * G4[d][pidx(a2, b2)];
            R[a2][b2] = A - B + Cc - Dd;
            R[b2][a2] = R[a2][b2];
          }
        double Rs = 0.0;
        for (int a2 = 0; a2 < 4; ++a2)
          for (int b2 = 0; b2 < 4; ++b2) Rs += cf.gu[pidx(a2, b2)] * R[a2][b2];
        const double nv[4] = {1.0 / cf.alpha, -cf.beta[0] / cf.alpha, -cf.beta[1] / cf.alpha, -cf.beta[2] / cf.alpha};
        double Rnn = 0.0;
        for (int a2 = 0; a2 < 4; ++a2)
          for (int b2 = 0; b2 < 4; ++b2) Rnn += R[a2][b2] * nv[a2] * nv[b2];
        double vals[4];
        vals[0] = 2 * Rnn + Rs;
        for (int q = 0; q < 3; ++q) {
          double s2 = 0.0;
          for (int a2 = 0; a2 < 4; ++a2) s2 += nv[a2] * (R[a2][q + 1] - g4_at(hv, a2, q + 1) * Rs / 2);
          vals[1 + q] = -s2;
        }
        const double wgt = g.blevel ? ldexp(1.0, -3 * block_level(g, b)) : 1.0;
        for (int q = 0; q < 4; ++q) {
          r[4 + q] = fabs(vals[q]);
          if (prim) r[q] = vals[q] * vals[q] * wgt;
        }
        if (prim) r[8] = wgt;
      }
    }
  }
  __shared__ double sh[9][256];
  const int tid = threadIdx.x;
  for (int n = 0; n < 9; ++n) sh[n][tid] = r[n];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (tid < w) {
      for (int n = 0; n < 9; ++n)
        sh[n][tid] = (n >= 4 && n < 8) ? fmax(sh[n][tid], sh[n][tid + w]) : sh[n][tid] + sh[n][tid + w];
    }
    __syncthreads();
  }
  if (tid < 9) a.out[9 * ((long)b * ((P * P * P + 255) / 256) + wx) + tid] = sh[tid][0];
}

// The refinement indicator (lohner, cell_tau, gh_tau_kernel!, tau_max,
// src/refinement.jl:152-219,312-378) on a prepared work array: the largest
// Loehner ratio over the ten h components and three axes at every evolved
// owned point, folded per workgroup into its maximum (the host takes each
// block's over its workgroups); with scale < 0 the kernel folds the largest
// |h_v| over the evolved owned points instead (field_scale, :221-300).
struct TauArgs {
  Geom g;
  CaseP c;
  double t, scale, eps;
  const double* __restrict__ W;
  double* __restrict__ out;
};
__global__ __launch_bounds__(256) void k_tau(TauArgs a) {
  const Geom& g = a.g;
  const int P = g.P, S = g.S, G = g.G;
  int b, wx;
  if (!wg_map((P * P * P + 255) / 256, g.nblocks, b, wx)) return;
  const int p = wx * blockDim.x + threadIdx.x;
  const long sv = (long)S * S * S;
  double tau = 0.0;
  if (p < P * P * P) {
    const int i = p % P, j = (p / P) % P, k = p / (P * P);
    double x[3];
    position(g, b, i, j, k, x);
    if (is_evolved(a.c, a.t, x)) {
      const double* w0 = a.W + (long)b * 20 * sv + (i + G) + (long)S * ((j + G) + (long)S * (k + G));
      const long st[3] = {1, S, (long)S * S};
      for (int v = 0; v < NC; ++v) {
        const double* wv = w0 + v * sv;
        const double u0 = wv[0];
        if (a.scale < 0.0) {
          tau = fmax(tau, fabs(u0));
          continue;
        }
        for (int d = 0; d < 3; ++d) {
          const double up = wv[st[d]], um = wv[-st[d]];
          const double num = fabs(up - 2 * u0 + um);
          const double den = fabs(up - u0) + fabs(u0 - um) + 4 * a.eps * fabs(a.scale);
          tau = fmax(tau, den == 0.0 ? 0.0 : num / den);
        }
      }
    }
  }
  __shared__ double sh[256];
  const int tid = threadIdx.x;
  sh[tid] = tau;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (tid < w) sh[tid] = fmax(sh[tid], sh[tid + w]);
    __syncthreads();
  }
  if (tid == 0) a.out[(long)b * ((P * P * P + 255) / 256) + wx] = sh[0];
}

#endif  // TGH_RHS_ONLY

// ---------------------------------------------------------------------------
// The brick layout of the RK4 stepper. On a uniform forest of closed vertex
// blocks every mesh point is the primary of exactly one block, and the
// stepper holds the primaries of all blocks in one array per variable over
// the whole brick of points, with a frame of G ghost points on every side:
// point (x, y, z) of variable v, x in [-G, n_0 + G) and so on, sits at
//   (x + XO) + py (y + G) + pz (z + G) + fs v,
// with n_d = nb_d N + 1 points on a non-periodic axis (the boundary planes
// are evolved) and nb_d N on a periodic one. A stencil of an interior point
// then reads its neighbours' values where they are stored, no per-block ghost
// copies exist, and only the frame is filled: from the Dirichlet hook or by
// periodic wrap, with the same rule as the block fill (a point outside a
// non-periodic axis takes the exact state at its own unwrapped position).
// The primary (block, local index) of brick point gp.
__device__ __forceinline__ int brick_primary(const Geom& g, const int gp[3], int loc[3]) {
  int ob[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const int o = gp[d] / g.N;
    ob[d] = o < g.nb[d] - 1 ? o : g.nb[d] - 1;
    loc[d] = gp[d] - ob[d] * g.N;
  }
  return g.lookup[ob[0] + g.nb[0] * (ob[1] + g.nb[1] * ob[2])];
}

struct Brick {
  int n[3];      // evolved points per axis
  int tot[3];    // nb_d N
  int periodic[3];
  int refl[3][2];  // reflecting low / high face per axis
  int rot;         // the rotating seam (low x and y faces, a quarter turn about z)
  int G, XO;     // ghost width; x offset of x = 0 in a row (XO >= G)
  long py, pz, fs;
  double lo[3], h, inv_h;
  // point (x, y, z) is at lo + (x + gx0) h: 0 for the brick; a block's first
  // index (B N) when a block's work array is run as a brick, so that positions
  // round as the block layout's do; lsh: its layer-box slot's shift to the
  // finest level's grid
  int gx0[3];
  int lsh;
};
TGH_HD long brick_index(const Brick& k, int x, int y, int z) {
  return (long)(x + k.XO) + k.py * (long)(y + k.G) + k.pz * (long)(z + k.G);
}
// Interior points, one per thread: each z-plane is cut into runs of 256
// consecutive points in x-fastest order, so a wave reads 64 consecutive x
// values of a row (or the end of one row and the start of the next), and the
// runs are dealt to the XCDs in contiguous ranges (wg_map's order): each XCD
// sweeps one slab of planes, whose stencil neighbours it shares in its L2.
// The runs of the brick: wpp per plane, brick_runs in all.
__device__ __forceinline__ int brick_runs(const Brick& k) {
  const long plane = (long)k.n[0] * k.n[1];
  return (int)((plane + 255) / 256) * k.n[2];
}
// This thread's point of logical run L (0 <= L < brick_runs(k)).
__device__ __forceinline__ bool brick_run(const Brick& k, int L, int& x, int& y, int& z) {
  const long plane = (long)k.n[0] * k.n[1];
  const int wpp = (int)((plane + 255) / 256);
  z = L / wpp;
  const int r = L - z * wpp;
  const long lin = (long)r * 256 + threadIdx.x;
  if (lin >= plane) return false;
  y = (int)(lin / k.n[0]);
  x = (int)(lin - (long)y * k.n[0]);
  return true;
}
__device__ __forceinline__ bool brick_map(const Brick& k, int& x, int& y, int& z) {
  const int total = brick_runs(k);
  const int L = wg_logical(total);
  if (L >= total) return false;
  return brick_run(k, L, x, y, z);
}

// Tiles of TX x TY x TZ points over the brick, dealt to the XCDs in
// contiguous ranges of an x-fastest order.
__device__ __forceinline__ int brick_tiles(const Brick& k, int TX, int TY, int TZ) {
  return ((k.n[0] + TX - 1) / TX) * ((k.n[1] + TY - 1) / TY) * ((k.n[2] + TZ - 1) / TZ);
}
// The corner of logical tile L (0 <= L < brick_tiles).
__device__ __forceinline__ void tile_at(const Brick& k, int TX, int TY, int TZ, int tile, int& X0, int& Y0,
                                        int& Z0) {
  const int ntx = (k.n[0] + TX - 1) / TX, nty = (k.n[1] + TY - 1) / TY;
  X0 = (tile % ntx) * TX;
  Y0 = ((tile / ntx) % nty) * TY;
  Z0 = (tile / (ntx * nty)) * TZ;
}
__device__ __forceinline__ bool tile_map(const Brick& k, int TX, int TY, int TZ, int& X0, int& Y0, int& Z0) {
  const int total = brick_tiles(k, TX, TY, TZ);
  const int L = wg_logical(total);
  if (L >= total) return false;
  tile_at(k, TX, TY, TZ, L, X0, Y0, Z0);
  return true;
}

// The fused right-hand side with the RK4 epilogue on the brick layout: the
// stage input W, its successor Wn, u_n and the accumulator all in the brick
// layout, so one index addresses a point in every one of them.
template <int Q, int GAUGE, int HOLE>
__global__ __launch_bounds__(256) void k_brick_rhs(RhsArgs a, Brick k, const double* __restrict__ W,
                                                   TGH_OUT_PARAMS) {
  a.t = stage_time(a.t, a.clock, a.toff);
  int x, y, z;
  if (!brick_map(k, x, y, z)) return;
  const long off = brick_index(k, x, y, z);
  PointRef pr;
  pr.w0 = W + off;
  pr.sv = k.fs;
  pr.s2 = (int)k.py;
  pr.s3 = (int)k.pz;
  pr.x[0] = k.lo[0] + (double)(x + k.gx0[0]) * k.h;
  pr.x[1] = k.lo[1] + (double)(y + k.gx0[1]) * k.h;
  pr.x[2] = k.lo[2] + (double)(z + k.gx0[2]) * k.h;
  pr.si0 = off;
  pr.ssv = k.fs;
  pr.wi0 = off;
  pr.Hp = a.Hsrc ? a.Hsrc + off : nullptr;
  pr.hs = k.fs;
  pr.m_rho = nullptr;
  pr.m_S = nullptr;
  pr.m_Sij = nullptr;
  pr.ms = 0;
  pr.lslot = a.lbox ? layer_slot((x + k.gx0[0]) << k.lsh, (y + k.gx0[1]) << k.lsh, (z + k.gx0[2]) << k.lsh, a.lB)
                    : 0;
  rhs_point<Q, GAUGE, HOLE, EPI_RK4>(a, pr, k.inv_h, TGH_OUT_ARGS);
}

// The split form on the brick layout, first half: the head at one point per
// thread (brick_map's order), leaving the Pi coefficients in coef (CF_N
// arrays of the brick's field size, indexed like the brick).
template <int Q, int GAUGE, int HOLE>
__device__ __forceinline__ void brick_head_at(const RhsArgs& a, const Brick& k, const double* __restrict__ W,
                                              const double* __restrict__ Hsrc, double* __restrict__ coef,
                                              TGH_OUT_PARAMS, int x, int y, int z);
template <int Q, int GAUGE, int HOLE>
__global__ __launch_bounds__(256) void k_brick_head(RhsArgs a, Brick k, const double* __restrict__ W,
                                                    double* __restrict__ coef, TGH_OUT_PARAMS) {
  a.t = stage_time(a.t, a.clock, a.toff);
  int x, y, z;
  if (!brick_map(k, x, y, z)) return;
  brick_head_at<Q, GAUGE, HOLE>(a, k, W, a.Hsrc, coef, TGH_OUT_ARGS, x, y, z);
}
template <int Q, int GAUGE, int HOLE>
__device__ __forceinline__ void brick_head_at(const RhsArgs& a, const Brick& k, const double* __restrict__ W,
                                              const double* __restrict__ Hsrc, double* __restrict__ coef,
                                              TGH_OUT_PARAMS, int x, int y, int z) {
  const long off = brick_index(k, x, y, z);
  PointRef pr;
  pr.w0 = W + off;
  pr.sv = k.fs;
  pr.s2 = (int)k.py;
  pr.s3 = (int)k.pz;
  pr.x[0] = k.lo[0] + (double)(x + k.gx0[0]) * k.h;
  pr.x[1] = k.lo[1] + (double)(y + k.gx0[1]) * k.h;
  pr.x[2] = k.lo[2] + (double)(z + k.gx0[2]) * k.h;
  pr.si0 = off;
  pr.ssv = k.fs;
  pr.wi0 = off;
  pr.Hp = Hsrc ? Hsrc + off : nullptr;
  pr.hs = k.fs;
  pr.m_rho = nullptr;
  pr.m_S = nullptr;
  pr.m_Sij = nullptr;
  pr.ms = 0;
  pr.lslot = a.lbox ? layer_slot((x + k.gx0[0]) << k.lsh, (y + k.gx0[1]) << k.lsh, (z + k.gx0[2]) << k.lsh, a.lB)
                    : 0;
  const GlobalSrc src{pr.w0, pr.sv, pr.s2, pr.s3};
  double r;
  if (frozen_core<HOLE>(a.c, a.t, pr.x, r)) {
    double dh[NC], dP[NC];
    core_rates<GAUGE, HOLE>(a, pr.x, pr.lslot, src, dh, dP);
#pragma unroll 1
    for (int v = 0; v < NC; ++v) {
      emit<EPI_RK4>(a, TGH_OUT_ARGS, off + v * k.fs, off + v * k.fs, dh[v]);
      coef[(CF_C1 + v) * k.fs + off] = dP[v];  // the core's d_t Pi, for the Pi kernel
    }
    coef[CF_W * k.fs + off] = -1.0;
    return;
  }
  // the geometric coefficients are stored as soon as they are known, so that
  // they are not held through the source
  double hv[NC], D[4][NC];
  Coef cf;
  PiCoef pc;
  pc.epsh = dissipation_rate(a.c, a.t, pr.x) * k.inv_h;
  head_derivs<Q>(a.c, src, k.inv_h, pc.epsh, hv, D, cf);
  head_geometry(cf, D, pc);
  store_coef_geometry(coef + off, k.fs, pc);
  head_source<GAUGE, HOLE, EPI_RK4>(a, pr, src, r, cf, hv, D, pc, TGH_OUT_ARGS);
  store_coef_source(coef + off, k.fs, pc);
}

// The split form on the brick layout, second half: a workgroup takes a tile
// of points (PiTile) and, one component at a time,
// stages the tile's box of h_v and Pi_v (the tile and a halo of G = q/2 + 1
// points, the reach of the dissipation stencil) in LDS with coalesced loads;
// every stencil of the component is then read from LDS. Tiles go to the XCDs
// in contiguous ranges of x-fastest tile order.
#ifndef TGH_PI_TILE_X
#define TGH_PI_TILE_X 8
#endif
#ifndef TGH_PI_TILE_Y
#define TGH_PI_TILE_Y 8
#endif
#ifndef TGH_PI_TILE_Z
#define TGH_PI_TILE_Z 4
#endif
#ifndef TGH_PI_POINTS_Z
#define TGH_PI_POINTS_Z 1
#endif
// The tile: X x Y x Z threads, each taking PZ points along z (at z, z + Z, ...);
// at q = 8 (halo 5) two planes of threads, one point each.
// The box's LDS row pitch: 24 doubles (at least the row's BX), so that the
// eight rows a wave reads fall on distinct banks of the 64 four-byte LDS
// banks (24 = -8 mod 32: row r starts at bank 48 r mod 64).
// The Pi kernel's tile shapes the autotuner chooses among (tgh_autotune):
// TC 0 the measured default, TC 1 a taller tile, TC 2 a wider square one,
// TC 3 a flat one whose box no longer fits one column per thread, TC 4 and 5
// the default tile with an LDS row pitch of 25 and 26 doubles instead of 24
// (the banks a wave's rows land on). Every shape computes each point with the
// same arithmetic in the same order: only the staging differs, so the
// results are the same bits.
constexpr int PI_TILES = 6;
// The tile of a block run as a brick (17 x 5 x 3 threads, one point each): a
// block's 17 points along x in one row, 4 x 6 tiles a block, 80 % of the
// threads busy, where the brick's 8 x 8 x 4 tile would keep 43 %.
constexpr int PI_TILE_BLOCK = 6;
template <int Q, int TC = 0> struct PiTile {
  static constexpr int X = TC == PI_TILE_BLOCK ? 17 : TC == 0 || TC >= 4 ? TGH_PI_TILE_X : TC == 1 ? 8 : TC == 2 ? 10 : 16;
  static constexpr int Y = TC == PI_TILE_BLOCK ? 5 : TC == 0 || TC >= 4 ? TGH_PI_TILE_Y : TC == 1 ? 12 : TC == 2 ? 10 : 8;
  static constexpr int Z = TC == PI_TILE_BLOCK ? 3 : Q == 8 ? 2 : TC == 0 || TC >= 4 ? TGH_PI_TILE_Z : TC == 3 ? 2 : 4;
  static constexpr int PZ = TC == PI_TILE_BLOCK || Q == 8 || !(TC == 0 || TC >= 4) ? 1 : TGH_PI_POINTS_Z;
  static constexpr int pitch(int bx) {
    return TC == 4 ? (bx <= 25 ? 25 : bx) : TC == 5 ? (bx <= 26 ? 26 : bx) : (bx <= 24 ? 24 : bx);
  }
};
template <int Q, int TC = 0>
__device__ __forceinline__ void brick_pi_tile(const RhsArgs& a, const Brick& k, const double* __restrict__ W,
                                              const double* __restrict__ coef, TGH_OUT_PARAMS, int X0, int Y0,
                                              int Z0);
template <int Q, int TC>
__global__ __launch_bounds__((PiTile<Q, TC>::X * PiTile<Q, TC>::Y * PiTile<Q, TC>::Z)) void k_brick_pi(
    RhsArgs a, Brick k, const double* __restrict__ W, const double* __restrict__ coef, TGH_OUT_PARAMS) {
  a.t = stage_time(a.t, a.clock, a.toff);
  using T = PiTile<Q, TC>;
  int X0, Y0, Z0;
  if (!tile_map(k, T::X, T::Y, T::Z * T::PZ, X0, Y0, Z0)) return;  // uniform
  brick_pi_tile<Q, TC>(a, k, W, coef, TGH_OUT_ARGS, X0, Y0, Z0);
}
// One tile of the Pi kernel, corner (X0, Y0, Z0), one workgroup.
template <int Q, int TC>
__device__ __forceinline__ void brick_pi_tile(const RhsArgs& a, const Brick& k, const double* __restrict__ W,
                                              const double* __restrict__ coef, TGH_OUT_PARAMS, int X0, int Y0,
                                              int Z0) {
  constexpr int TX = PiTile<Q, TC>::X, TY = PiTile<Q, TC>::Y, TZ = PiTile<Q, TC>::Z, PZ = PiTile<Q, TC>::PZ;
  constexpr int NT = TX * TY * TZ, DZ = TZ * PZ;  // threads; the tile's depth in points
  constexpr int H = ::tgh::W<Q>::R + 1;
  // the box: BX x BY x BZ points, stored with row pitch RX >= BX (a pitch
  // that spreads a wave's rows over the LDS banks) and plane pitch RX * BY
  constexpr int BX = TX + 2 * H, BY = TY + 2 * H, BZ = DZ + 2 * H, BV = BX * BY * BZ;
  constexpr int RX = PiTile<Q, TC>::pitch(BX), RV = RX * BY * BZ;
  __shared__ double sh[2][RV];
  const int t = threadIdx.x;
  const int lx = t % TX, ly = (t / TX) % TY, lz = t / (TX * TY);
  const int x = X0 + lx, y = Y0 + ly;
  const double ih = k.inv_h, ih2 = ih * ih;
  bool valid[PZ];
  long off[PZ];
  int c0[PZ];
  PiCoef pc[PZ];
#pragma unroll
  for (int q = 0; q < PZ; ++q) {
    const int z = Z0 + lz + q * TZ;
    valid[q] = x < k.n[0] && y < k.n[1] && z < k.n[2];
    off[q] = valid[q] ? brick_index(k, x, y, z) : 0;
    pc[q].w = -1.0;
    pc[q].epsh = 0.0;
    if (valid[q]) load_coef(coef + off[q], k.fs, pc[q]);
    c0[q] = (lx + H) + RX * ((ly + H) + BY * (lz + q * TZ + H));
  }
  // box point (bx, by, bz) is brick point (X0 - H + bx, ...): never below -G,
  // and skipped at or beyond n + G, where only points outside the brick
  // (no thread computes them) would read it
  const int ex = k.n[0] + k.G - (X0 - H), ey = k.n[1] + k.G - (Y0 - H), ez = k.n[2] + k.G - (Z0 - H);
  const long box0 = brick_index(k, X0 - H, Y0 - H, Z0 - H);
  // Staging: each thread owns at most one box column (bx, by) (BX * BY <= NT
  // for q <= 6; q = 8 takes the loop below) and copies it plane by plane. The
  // next component's column is loaded into registers while the current one
  // is computed, so the loads' latency hides behind the arithmetic.
  // Only what the stencils read is staged: h_v in the box of halo R = H - 1
  // (first, second and mixed derivatives), Pi_v on the three axis lines
  // through the tile (first derivatives and the dissipation, reach H): a
  // column inside the tile's footprint takes every plane, a column beside it
  // along x or y the tile's own planes, a corner column none -- 2176 values
  // of the 3920 in the two boxes at q = 4.
  const int nbz = ez < BZ ? ez : BZ;  // planes inside the stored extent
  constexpr bool ONE_COL = BX * BY <= NT;
  const int sbx = t % BX, sby = t / BX;
  const bool stager = ONE_COL && t < BX * BY;
  const bool col_in = stager && sbx < ex && sby < ey;
  const bool hcol = col_in && sbx >= 1 && sbx < BX - 1 && sby >= 1 && sby < BY - 1;
  const bool ix = sbx >= H && sbx < BX - H, iy = sby >= H && sby < BY - H;
  const int pz0 = (ix && iy) ? 0 : ((ix || iy) ? H : BZ);
  const int pz1 = (ix && iy) ? BZ : ((ix || iy) ? H + DZ : BZ);
  const double* __restrict__ Wc = W + box0 + (sbx + k.py * sby);
  const int lcol = sbx + RX * sby;
  double ph[ONE_COL ? BZ : 1], pp[ONE_COL ? BZ : 1];
  auto fetch = [&](int v) {
    if constexpr (ONE_COL) {
#pragma unroll
      for (int bz = 0; bz < BZ; ++bz) {
        const bool inh = hcol && bz >= 1 && bz < BZ - 1 && bz < nbz;
        const bool inp = col_in && bz >= pz0 && bz < pz1 && bz < nbz;
        ph[bz] = inh ? Wc[v * k.fs + (long)bz * k.pz] : 0.0;
        pp[bz] = inp ? Wc[(NC + v) * k.fs + (long)bz * k.pz] : 0.0;
      }
    }
  };
  fetch(0);
#pragma unroll 1
  for (int v = 0; v < NC; ++v) {
    __syncthreads();  // the previous component's reads are done
    if constexpr (ONE_COL) {
      if (stager) {
#pragma unroll
        for (int bz = 0; bz < BZ; ++bz) {
          if (bz >= 1 && bz < BZ - 1) sh[0][lcol + RX * BY * bz] = ph[bz];
          if (bz >= pz0 && bz < pz1) sh[1][lcol + RX * BY * bz] = pp[bz];
        }
      }
    } else {
      for (int col = t; col < BX * BY; col += NT) {
        const int cx = col % BX, cy = col / BX;
        const bool cin = cx < ex && cy < ey;
        const double* __restrict__ Wh = W + box0 + v * k.fs + (cx + k.py * cy);
        const int lc = cx + RX * cy;
        for (int bz = 0; bz < BZ; ++bz) {
          const bool in = cin && bz < nbz;
          sh[0][lc + RX * BY * bz] = in ? Wh[(long)bz * k.pz] : 0.0;
          sh[1][lc + RX * BY * bz] = in ? Wh[NC * k.fs + (long)bz * k.pz] : 0.0;
        }
      }
    }
    __syncthreads();
    // this component's epilogue inputs first, then the next component's column
    double u0[PZ], a0[PZ], c1[PZ];
#pragma unroll
    for (int q = 0; q < PZ; ++q) {
      const long si = off[q] + (NC + v) * k.fs;
      if (valid[q]) {
        epi_load(a, us, acc, si, u0[q], a0[q]);
        c1[q] = coef[(CF_C1 + v) * k.fs + off[q]];
      }
    }
    if (v + 1 < NC) fetch(v + 1);
#pragma unroll
    for (int q = 0; q < PZ; ++q) {
      if (valid[q]) {
        // c1 is added after both branches: the wait for it (which, counted in
        // order, also waits for the next component's prefetched columns) then
        // comes after the principal part. In the frozen core (w < 0) the rate
        // is c1 itself.
        double ws = 0.0;
        if (pc[q].w >= 0.0) {
          const double s = pi_principal<Q>(&sh[0][c0[q]], &sh[1][c0[q]], RX, RX * BY, pc[q].b, pc[q].divb,
                                           pc[q].divA, pc[q].A, ih, ih2, pc[q].epsh, a.c.has_diss);
          ws = pc[q].w * s;
        }
        const double f = ws + c1[q];
        const long si = off[q] + (NC + v) * k.fs;
        epi_store(a, us, acc, Wn, si, si, f, u0[q], a0[q]);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Several problems in one launch (independent holes of uneven block counts):
// a device-side table holds each problem's stage arguments, its brick and the
// first logical workgroup of its range in the head and the Pi launches. A
// workgroup takes its logical index in wg_logical's XCD-contiguous order over
// the whole launch, finds the problem whose range holds it, and runs that
// problem's run (head) or tile (Pi) -- the same code as the one-problem
// kernels, so a problem in a batch computes the bits it computes alone.
// The table's arrays, typed as global memory on the device: pointers loaded
// from memory are otherwise generic, and their loads and stores flat.
#if defined(__HIP_DEVICE_COMPILE__)
#define TGH_GLOBAL __attribute__((address_space(1)))
#else
#define TGH_GLOBAL
#endif
struct BatchEntry {
  RhsArgs a;  // the stage's t, stage and stage coefficients (and the case)
  Brick k;
  TGH_GLOBAL const double* W;  // the stage input
  TGH_GLOBAL const double* Hsrc;
  TGH_GLOBAL double* coef;
  TGH_GLOBAL double* us;
  TGH_GLOBAL double* acc;
  TGH_GLOBAL double* Wn;  // the next stage's input
  int head0, pi0;  // first logical workgroup in the head and the Pi launch
  int lLmax;       // the block layout run as bricks: the problem's finest level
};
__device__ __forceinline__ int batch_find(const BatchEntry* __restrict__ tab, int n, int L, bool pi) {
  int p = 0;
  while (p + 1 < n && L >= (pi ? tab[p + 1].pi0 : tab[p + 1].head0)) ++p;
  return p;
}
template <int Q, int GAUGE, int HOLE>
__global__ __launch_bounds__(256) void k_brick_head_batch(const BatchEntry* __restrict__ tab, int n,
                                                                        int total) {
  const int L = wg_logical(total);
  if (L >= total) return;
  const BatchEntry& e = tab[batch_find(tab, n, L, false)];
  int x, y, z;
  if (!brick_run(e.k, L - e.head0, x, y, z)) return;
  brick_head_at<Q, GAUGE, HOLE>(e.a, e.k, (const double*)e.W, (const double*)e.Hsrc, (double*)e.coef, nullptr,
                                (double*)e.us, (double*)e.acc, (double*)e.Wn, x, y, z);
}
template <int Q>
__global__ __launch_bounds__(PiTile<Q>::X * PiTile<Q>::Y * PiTile<Q>::Z) void k_brick_pi_batch(
    const BatchEntry* __restrict__ tab, int n, int total) {
  const int L = wg_logical(total);
  if (L >= total) return;  // uniform over the workgroup
  const BatchEntry& e = tab[batch_find(tab, n, L, true)];
  int X0, Y0, Z0;
  tile_at(e.k, PiTile<Q>::X, PiTile<Q>::Y, PiTile<Q>::Z * PiTile<Q>::PZ, L - e.pi0, X0, Y0, Z0);
  brick_pi_tile<Q>(e.a, e.k, (const double*)e.W, (const double*)e.coef, nullptr, (double*)e.us, (double*)e.acc,
                   (double*)e.Wn, X0, Y0, Z0);
}

// The block layout's split form with the brick's code: block b's work array
// is a brick of P^3 points (strides S, S^2, S^3 per variable, G ghosts) at
// index offset B N, run by the brick head (one point a thread) and the
// LDS-tiled Pi kernel (PI_TILE_BLOCK tiles). During a step u_n, the
// accumulator and the Pi coefficients sit in the work layout beside W, so
// that one index addresses a point in all of them, as on the brick.
struct BlockBricks {
  const double* __restrict__ W;
  double* __restrict__ coef;  // CF_N fields of S^3 points a block
  double* __restrict__ us;
  double* __restrict__ acc;
  double* __restrict__ Wn;
  int lLmax;  // the finest level (the layer box's grid)
};
__device__ __forceinline__ Brick block_brick(const Geom& g, int b, int lLmax) {
  Brick k;
  const int* B = g.bcoord + 3 * b;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    k.n[d] = g.P;
    k.tot[d] = g.N;
    k.periodic[d] = 0;
    k.refl[d][0] = k.refl[d][1] = 0;
    k.lo[d] = g.lo[d];
    k.gx0[d] = B[d] * g.N;
  }
  k.rot = 0;
  k.G = g.G;
  k.XO = g.G;
  k.py = g.S;
  k.pz = (long)g.S * g.S;
  k.fs = (long)g.S * g.S * g.S;
  k.h = block_h(g, b);
  k.inv_h = block_inv_h(g, b);
  k.lsh = lLmax - block_level(g, b);
  return k;
}
template <int Q, int GAUGE, int HOLE>
__global__ __launch_bounds__(256) void k_block_head(RhsArgs a, BlockBricks bb) {
  a.t = stage_time(a.t, a.clock, a.toff);
  const Geom& g = a.g;
  const int P = g.P, P3 = g.P * g.P * g.P;
  int b, wx;
  if (!wg_map((P3 + 255) / 256, g.nblocks, b, wx)) return;
  const int p = wx * 256 + (int)threadIdx.x;
  if (p >= P3) return;
  const Brick k = block_brick(g, b, bb.lLmax);
  const long wb = (long)b * 20 * k.fs;
  brick_head_at<Q, GAUGE, HOLE>(a, k, bb.W + wb, nullptr, bb.coef + (long)b * CF_N * k.fs, nullptr, bb.us + wb,
                                bb.acc + wb, bb.Wn + wb, p % P, (p / P) % P, p / (P * P));
}
template <int Q>
__global__ __launch_bounds__((PiTile<Q, PI_TILE_BLOCK>::X * PiTile<Q, PI_TILE_BLOCK>::Y *
                              PiTile<Q, PI_TILE_BLOCK>::Z)) void k_block_pi(RhsArgs a, BlockBricks bb) {
  a.t = stage_time(a.t, a.clock, a.toff);
  using T = PiTile<Q, PI_TILE_BLOCK>;
  const Geom& g = a.g;
  const int P = g.P;
  const int tx = (P + T::X - 1) / T::X, ty = (P + T::Y - 1) / T::Y, tz = (P + T::Z - 1) / T::Z;
  int b, wx;
  if (!wg_map(tx * ty * tz, g.nblocks, b, wx)) return;  // uniform over the workgroup
  const int X0 = (wx % tx) * T::X, Y0 = ((wx / tx) % ty) * T::Y, Z0 = (wx / (tx * ty)) * T::Z;
  const Brick k = block_brick(g, b, bb.lLmax);
  const long wb = (long)b * 20 * k.fs;
  brick_pi_tile<Q, PI_TILE_BLOCK>(a, k, bb.W + wb, bb.coef + (long)b * CF_N * k.fs, nullptr, bb.us + wb,
                                  bb.acc + wb, bb.Wn + wb, X0, Y0, Z0);
}

// A batch on the block layout with its blocks run as bricks: one head and one
// Pi launch over every block of every problem, each entry's blocks a range of
// logical workgroups from head0 (head) and pi0 (Pi tiles); W, Wn, us, acc and
// coef the problem's work-layout arrays.
template <int Q, int GAUGE, int HOLE>
__global__ __launch_bounds__(256) void k_block_head_batch(const BatchEntry* __restrict__ tab, int n, int total) {
  const int L = wg_logical(total);
  if (L >= total) return;
  const BatchEntry& e = tab[batch_find(tab, n, L, false)];
  const Geom& g = e.a.g;
  const int P = g.P, P3 = g.P * g.P * g.P, wgpb = (P3 + 255) / 256;
  const int loc = L - e.head0, b = loc / wgpb;
  const int p = (loc - b * wgpb) * 256 + (int)threadIdx.x;
  if (p >= P3) return;
  const Brick k = block_brick(g, b, e.lLmax);
  const long wb = (long)b * 20 * k.fs;
  brick_head_at<Q, GAUGE, HOLE>(e.a, k, (const double*)e.W + wb, nullptr, (double*)e.coef + (long)b * CF_N * k.fs,
                                nullptr, (double*)e.us + wb, (double*)e.acc + wb, (double*)e.Wn + wb, p % P,
                                (p / P) % P, p / (P * P));
}
template <int Q>
__global__ __launch_bounds__((PiTile<Q, PI_TILE_BLOCK>::X * PiTile<Q, PI_TILE_BLOCK>::Y *
                              PiTile<Q, PI_TILE_BLOCK>::Z)) void k_block_pi_batch(const BatchEntry* __restrict__ tab,
                                                                                   int n, int total) {
  using T = PiTile<Q, PI_TILE_BLOCK>;
  const int L = wg_logical(total);
  if (L >= total) return;  // uniform over the workgroup
  const BatchEntry& e = tab[batch_find(tab, n, L, true)];
  const Geom& g = e.a.g;
  const int P = g.P;
  const int tx = (P + T::X - 1) / T::X, ty = (P + T::Y - 1) / T::Y, tz = (P + T::Z - 1) / T::Z;
  const int loc = L - e.pi0, b = loc / (tx * ty * tz), wx = loc - b * (tx * ty * tz);
  const int X0 = (wx % tx) * T::X, Y0 = ((wx / tx) % ty) * T::Y, Z0 = (wx / (tx * ty)) * T::Z;
  const Brick k = block_brick(g, b, e.lLmax);
  const long wb = (long)b * 20 * k.fs;
  brick_pi_tile<Q, PI_TILE_BLOCK>(e.a, k, (const double*)e.W + wb, (double*)e.coef + (long)b * CF_N * k.fs, nullptr,
                                  (double*)e.us + wb, (double*)e.acc + wb, (double*)e.Wn + wb, X0, Y0, Z0);
}

// One RK4 stage of a batch on the block layout: the fused kernel over every
// block of every problem in the table (forests of any depth and problems with
// the general interior among them), each entry's blocks a contiguous range of
// logical workgroups from head0.
template <int Q, int GAUGE, int HOLE>
__global__ __launch_bounds__(256) void k_rhs_batch(const BatchEntry* __restrict__ tab, int n, int total) {
  const int L = wg_logical(total);
  if (L >= total) return;
  const BatchEntry& e = tab[batch_find(tab, n, L, false)];
  const RhsArgs& a = e.a;
  const int wgpb = (a.g.P * a.g.P * a.g.P + 255) / 256;
  const int loc = L - e.head0;
  const int b = loc / wgpb;
  PointRef pr;
  long pi;
  if (!block_point_at(a, (const double*)e.W, b, loc - b * wgpb, pr, pi)) return;
  rhs_point<Q, GAUGE, HOLE, EPI_RK4>(a, pr, pr.ih, nullptr, (double*)e.us, (double*)e.acc, (double*)e.Wn);
}

#ifndef TGH_RHS_ONLY
// The stage clock's kernels: one thread sets the step's time before a graph
// replay; k_batch_clock gives a batch's stage tables (4 x n entries, stage
// major) the stage times clock + (0, dt/2, dt/2, dt), the first node of the
// batch's captured step.
__global__ void k_set_clock(double* clock, double t) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *clock = t;
}
__global__ __launch_bounds__(256) void k_batch_clock(BatchEntry* tab, int n, const double* clock, double half,
                                                     double dt) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= 4 * n) return;
  const int s = e / n;
  tab[e].a.t = *clock + (s == 0 ? 0.0 : s == 3 ? dt : half);
}

// The frame of a brick array at time t, one point of the whole extent per
// thread (interior points return at once).
struct FrameArgs {
  Brick k;
  CaseP c;
  double t;
  double* __restrict__ W;
  const double* clock;  // stage_time's clock and offset
  double toff;
  int shell;  // 1: one thread per frame point (brick_shell_point); 0: per extent point
};
// A moving :damped hole's core and layer on the brick: per stage, the exact
// background's dual passes (bg_state_exact) that the head's core and layer
// targets are made of, for every core and layer point of the B^3 box about the
// hole (its corner one point below the layer's bounding box about c(t)), into
// slot layer_slot(x, y, z, B) of LAYER_FIELDS fields; grid.y is the pass, so the
// three passes of a point run on three threads:
//   a layer point (background_state_rate): pass 0 the state h, Pi and d_t h
//     at t (fields 0-29), passes 1 and 2 Pi at t +- delta (30-39, 40-49);
//   a core point (moving_core_target): pass 0 the state at the core rule's
//     point at t (0-19), passes 1 and 2 the state at t +- delta at that
//     time's core point (20-39, 40-59), only where core_target_rate's weight
//     is not zero.
// The head (HOLE_BOXED) combines them as those functions do.
enum { LAYER_FIELDS = 60 };
struct LayerArgs {
  Brick k;
  CaseP c;
  double t;
  const double* clock;  // stage_time's clock and offset
  double toff;
  double* __restrict__ box;
  int B;
};
template <bool MOVING>
__global__ __launch_bounds__(256) void k_brick_layer(LayerArgs a) {
  a.t = stage_time(a.t, a.clock, a.toff);
  const Brick& k = a.k;
  const CaseP& c = a.c;
  const int B = a.B, pass = blockIdx.y;
  const long B3 = (long)B * B * B;
  const long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= B3) return;
  const int l[3] = {(int)(s % B), (int)((s / B) % B), (int)(s / ((long)B * B))};
  double cc[3];
  center_at(c, a.t, cc);
  int p[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    p[d] = (int)floor((cc[d] - c.r1 - k.lo[d]) * k.inv_h) - 1 + l[d];
    if (p[d] < 0 || p[d] >= k.n[d]) return;
  }
  const double x[3] = {k.lo[0] + (double)p[0] * k.h, k.lo[1] + (double)p[1] * k.h, k.lo[2] + (double)p[2] * k.h};
  double r;
  const bool core = frozen_core<1>(c, a.t, x, r);
  if (!core && !(r < c.r1)) return;  // the head reads core and layer points only
  double* dst = a.box + layer_slot(p[0], p[1], p[2], B);
  if constexpr (!MOVING) {
    // a harmonic hole at rest: the layer's target, once (the core's rates are zero)
    if (core) return;
    double ue[20];
    case_state<bg_set_for(GAUGE_NONE)>(c, a.t, x, ue);
#pragma unroll
    for (int v = 0; v < 20; ++v) dst[v * B3] = ue[v];
    return;
  }
  double hh[NC], PP[NC], dd[NC];
  if (core) {
    double xq[3];
    double tq = a.t;
    if (pass > 0) {
#pragma clang fp contract(off)
      const double w = smoothstep(4.0 * hole_radius(c, a.t, x) / c.r0 - 1.0);
      if (w == 0.0) return;  // no rate there
      tq = pass == 1 ? a.t + c.delta : a.t - c.delta;
    }
    core_point(c, tq, x, xq);
    bg_state_exact(c.bg, tq, xq, hh, PP, dd);
    const int f0 = 20 * pass;
#pragma unroll
    for (int v = 0; v < NC; ++v) {
      dst[(f0 + v) * B3] = hh[v];
      dst[(f0 + NC + v) * B3] = PP[v];
    }
    return;
  }
  const double tq = pass == 0 ? a.t : pass == 1 ? a.t + c.delta : a.t - c.delta;
  bg_state_exact(c.bg, tq, x, hh, PP, dd);
  if (pass == 0) {
#pragma unroll
    for (int v = 0; v < NC; ++v) {
      dst[v * B3] = hh[v];
      dst[(NC + v) * B3] = PP[v];
      dst[(2 * NC + v) * B3] = dd[v];
    }
  } else {
    const int f0 = pass == 1 ? 3 * NC : 4 * NC;
#pragma unroll
    for (int v = 0; v < NC; ++v) dst[(f0 + v) * B3] = PP[v];
  }
}

// The frame's points (the brick's extent less its n0 x n1 x n2 inside) in x-fastest
// order: the G planes below and above, then per inner plane the G rows below
// and above, then per inner row the G points on each side. Point s of
// shell_size(k) as extent coordinates (0 .. n + 2G - 1).
__device__ __forceinline__ long shell_size(const Brick& k) {
  const long E0 = k.n[0] + 2 * k.G, E1 = k.n[1] + 2 * k.G, E2 = k.n[2] + 2 * k.G;
  return E0 * E1 * E2 - (long)k.n[0] * k.n[1] * k.n[2];
}
__device__ __forceinline__ void brick_shell_point(const Brick& k, long s, int e[3]) {
  const int G = k.G;
  const long E0 = k.n[0] + 2 * G, E1 = k.n[1] + 2 * G;
  const long A = 2L * G * E0 * E1, B = 2L * G * E0 * k.n[2];
  if (s < A) {
    const long pl = s / (E0 * E1), r = s - pl * E0 * E1;
    e[2] = (int)(pl < G ? pl : pl + k.n[2]);
    e[1] = (int)(r / E0);
    e[0] = (int)(r - (r / E0) * E0);
  } else if (s < A + B) {
    const long s2 = s - A, row = s2 / E0;  // row of the 2G per inner plane
    const long pz = row / (2 * G), q = row - pz * 2 * G;
    e[2] = (int)(G + pz);
    e[1] = (int)(q < G ? q : q + k.n[1]);
    e[0] = (int)(s2 - row * E0);
  } else {
    const long s3 = s - A - B, per = 2L * G * k.n[1];
    const long pz = s3 / per, r = s3 - pz * per, py = r / (2 * G), q = r - py * 2 * G;
    e[2] = (int)(G + pz);
    e[1] = (int)(G + py);
    e[0] = (int)(q < G ? q : q + k.n[0]);
  }
}
template <bool MOVE>
__global__ __launch_bounds__(256) void k_brick_frame(FrameArgs a) {
  a.t = stage_time(a.t, a.clock, a.toff);
  const Brick& k = a.k;
  const long E0 = k.n[0] + 2 * k.G, E1 = k.n[1] + 2 * k.G, E2 = k.n[2] + 2 * k.G;
  const long t = (long)blockIdx.x * 256 + threadIdx.x;
  int gp[3];
  if (a.shell) {
    // the frame alone: a wave's lanes are frame points, none idle beside an inner row
    if (t >= shell_size(k)) return;
    int e[3];
    brick_shell_point(k, t, e);
#pragma unroll
    for (int d = 0; d < 3; ++d) gp[d] = e[d] - k.G;
  } else {
    if (t >= E0 * E1 * E2) return;
    gp[0] = (int)(t % E0) - k.G;
    gp[1] = (int)((t / E0) % E1) - k.G;
    gp[2] = (int)(t / (E0 * E1)) - k.G;
  }
  bool inside = true, dirichlet = false;
  int w[3], mm = 0, wall = 0;
  // beyond the rotating seam the point inside is the image of, tn quarter turns
  int gs[3] = {gp[0], gp[1], gp[2]};
  const int tn = k.rot ? rot_unfold(gs[0], gs[1]) : 0;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    if (gp[d] < 0 || gp[d] >= k.n[d]) inside = false;
    if (gs[d] >= 0 && gs[d] < k.n[d]) {
      w[d] = gs[d];
      if (!k.periodic[d] && ((k.refl[d][0] && gp[d] == 0) || (k.refl[d][1] && gp[d] == k.n[d] - 1)))
        wall |= 1 << d;
    } else {
      if (k.periodic[d]) {
        w[d] = ((gs[d] % k.tot[d]) + k.tot[d]) % k.tot[d];
      } else if (gs[d] < 0 && k.refl[d][0]) {
        w[d] = -gs[d];
        mm |= 1 << d;
      } else if (gs[d] >= k.n[d] && k.refl[d][1]) {
        w[d] = 2 * (k.n[d] - 1) - gs[d];
        mm |= 1 << d;
      } else {
        dirichlet = true;
      }
    }
  }
  double* dst = a.W + brick_index(k, gp[0], gp[1], gp[2]);
  if (inside) {
    // a point of a reflecting wall takes the parity: its odd variables are zero
    if (wall) {
#pragma unroll
      for (int v = 0; v < 20; ++v)
        if (parity_odd_mask(v) & wall) dst[v * k.fs] = 0.0;
    }
    return;
  }
  if (dirichlet) {
    double x[3], s[20];
#pragma unroll
    for (int d = 0; d < 3; ++d) x[d] = k.lo[d] + (double)gp[d] * k.h;
    case_state<MOVE ? BGS_ALL : BGS_STATIC>(a.c, a.t, x, s);
#pragma unroll
    for (int v = 0; v < 20; ++v) dst[v * k.fs] = s[v];
  } else {
    // the source brick point, projected as its own walls would be, times the
    // parities of the mirrors crossed
    int sw = 0;
#pragma unroll
    for (int d = 0; d < 3; ++d)
      if (!k.periodic[d] && ((k.refl[d][0] && w[d] == 0) || (k.refl[d][1] && w[d] == k.n[d] - 1))) sw |= 1 << d;
    const double* src = a.W + brick_index(k, w[0], w[1], w[2]);
    if (tn == 0) {
#pragma unroll
      for (int v = 0; v < 20; ++v)
        dst[v * k.fs] = (parity_odd_mask(v) & sw) ? 0.0 : parity_sign(v, mm) * src[v * k.fs];
    } else {
      for (int v = 0; v < 20; ++v) {
        int q;
        double sg;
        rot_power(v, tn, q, sg);
        dst[v * k.fs] = (parity_odd_mask(q) & sw) ? 0.0 : parity_sign(v, mm) * sg * src[q * k.fs];
      }
    }
  }
}

// Between the state (block layout) and the brick. In: every brick point reads
// the owned point of the block whose primary it is (into two brick arrays:
// the stage input and u_n). Out, after the steps, in two passes so that no
// value is read after it is overwritten: first every copy (an owned point
// that is not its brick point's primary) advances by its primary's increment,
// u_c += u_brick - u_primary, which is the reference's stepping of the whole
// state vector, where a copy keeps its own value plus the primary's stage
// sums (periodic copies sit at the unwrapped position and can differ); then
// every primary takes its brick point's value.
struct BrickCopyArgs {
  Geom g;
  Brick k;
  const double* __restrict__ src;
  double* __restrict__ dst0;
  double* __restrict__ dst1;
  int pass;  // out: 0 copies, 1 primaries
};
__global__ __launch_bounds__(256) void k_brick_in(BrickCopyArgs a) {
  const Brick& k = a.k;
  const Geom& g = a.g;
  int x, y, z;
  if (!brick_map(k, x, y, z)) return;
  const int gp[3] = {x, y, z};
  int loc[3];
  const int b = brick_primary(g, gp, loc);
  const long psz = (long)g.P * g.P * g.P;
  const double* s = a.src + (long)b * 20 * psz + loc[0] + (long)g.P * (loc[1] + (long)g.P * loc[2]);
  const long off = brick_index(k, x, y, z);
#pragma unroll 4
  for (int v = 0; v < 20; ++v) {
    const double val = s[v * psz];
    a.dst0[off + v * k.fs] = val;
    a.dst1[off + v * k.fs] = val;
  }
}
__global__ __launch_bounds__(256) void k_brick_out(BrickCopyArgs a) {
  const Geom& g = a.g;
  const Brick& k = a.k;
  const int P = g.P;
  int b, wx;
  if (!wg_map((P * P * P + 255) / 256, g.nblocks, b, wx)) return;
  const int p = wx * blockDim.x + threadIdx.x;
  if (p >= P * P * P) return;
  const int l[3] = {p % P, (p / P) % P, p / (P * P)};
  const int* B = g.bcoord + 3 * b;
  int gp[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    gp[d] = B[d] * g.N + l[d];
    if (k.periodic[d] && gp[d] == k.tot[d]) gp[d] = 0;
  }
  int loc[3];
  const int ob = brick_primary(g, gp, loc);
  const bool primary = ob == b && loc[0] == l[0] && loc[1] == l[1] && loc[2] == l[2];
  if (primary != (a.pass == 1)) return;
  const long off = brick_index(k, gp[0], gp[1], gp[2]);
  const long psz = (long)P * P * P;
  double* d0 = a.dst0 + (long)b * 20 * psz + p;
  if (primary) {
#pragma unroll 4
    for (int v = 0; v < 20; ++v) d0[v * psz] = a.src[off + v * k.fs];
  } else {
    const double* up = a.dst0 + (long)ob * 20 * psz + loc[0] + (long)P * (loc[1] + (long)P * loc[2]);
#pragma unroll 4
    for (int v = 0; v < 20; ++v) d0[v * psz] = d0[v * psz] + (a.src[off + v * k.fs] - up[v * psz]);
  }
}

// The sampled gauge source on the brick layout (as k_sample_gauge).
struct BrickSampleArgs {
  Brick k;
  CaseP c;
  double t;
  double* __restrict__ Hsrc;
};
__global__ __launch_bounds__(256) void k_brick_sample(BrickSampleArgs a) {
  const Brick& k = a.k;
  int x, y, z;
  if (!brick_map(k, x, y, z)) return;
  const double xp[3] = {k.lo[0] + (double)x * k.h, k.lo[1] + (double)y * k.h, k.lo[2] + (double)z * k.h};
  double xc[3];
  core_point(a.c, a.t, xp, xc);
  double Hl[4], dH[4][4];
  gauge_source_grad(a.c.bg, a.t, xc, Hl, dH);
  double* out = a.Hsrc + brick_index(k, x, y, z);
  for (int q = 0; q < 4; ++q) out[q * k.fs] = Hl[q];
  for (int n = 0; n < 16; ++n) out[(4 + n) * k.fs] = dH[n % 4][n / 4];
}


// The range projection over points within the gate r < r_gate (apply_bounds!,
// src/bounds.jl:613-660,733-770), written back only where it fired.
//   mode 0: the state u (block layout), one partial (hits, non-finite, r_max)
//           per workgroup into out, folded on the host in order;
//   mode 1: a block work array's owned points (src = dst0 = W), and the
//           state at the same point too when dst1 (the RK4 result) is given;
//   mode 2: brick points of the box [lo, lo + ext) (src = dst0), and dst1 at
//           the same index when given.
struct BoundsArgs {
  Geom g;
  Brick k;
  CaseP c;
  Bounds bd;
  double* __restrict__ dst0;
  double* __restrict__ dst1;
  double* __restrict__ out;
  int mode;
  int lo[3], ext[3];
};
__global__ __launch_bounds__(256) void k_bounds(BoundsArgs a) {
  const Geom& g = a.g;
  double x[3];
  long i0 = 0, s0 = 0, i1 = 0, s1 = 0;  // index and field stride in dst0 / dst1
  bool valid = false;
  int b = 0, wx = 0;
  if (a.mode == 2) {
    const long n = (long)a.ext[0] * a.ext[1] * a.ext[2];
    const long t = (long)blockIdx.x * 256 + threadIdx.x;
    if (t < n) {
      const int gp[3] = {a.lo[0] + (int)(t % a.ext[0]), a.lo[1] + (int)((t / a.ext[0]) % a.ext[1]),
                         a.lo[2] + (int)(t / ((long)a.ext[0] * a.ext[1]))};
#pragma unroll
      for (int d = 0; d < 3; ++d) x[d] = a.k.lo[d] + (double)gp[d] * a.k.h;
      i0 = i1 = brick_index(a.k, gp[0], gp[1], gp[2]);
      s0 = s1 = a.k.fs;
      valid = true;
    }
  } else {
    const int P = g.P, S = g.S, G = g.G;
    if (!wg_map((P * P * P + 255) / 256, g.nblocks, b, wx)) return;
    const int p = wx * blockDim.x + threadIdx.x;
    if (p < P * P * P) {
      const int i = p % P, j = (p / P) % P, kk = p / (P * P);
      position(g, b, i, j, kk, x);
      const long psz = (long)P * P * P, sv = (long)S * S * S;
      const long si = (long)b * 20 * psz + p;
      if (a.mode == 0) {
        i0 = si;
        s0 = psz;
      } else {
        i0 = (long)b * 20 * sv + (i + G) + (long)S * ((j + G) + (long)S * (kk + G));
        s0 = sv;
        i1 = si;
        s1 = psz;
      }
      valid = true;
    }
  }
  double hit = 0.0, nf = 0.0, rm = 0.0;
  if (valid) {
    const double r = hole_radius(a.c, 0.0, x);  // a hole at rest (the library refuses bounds on a moving one)
    if (r < a.bd.r_gate) {
      double h[NC], Pi[NC];
#pragma unroll
      for (int v = 0; v < NC; ++v) {
        h[v] = a.dst0[i0 + v * s0];
        Pi[v] = a.dst0[i0 + (NC + v) * s0];
      }
      bool nonfin;
      if (bounds_project(h, Pi, a.bd, nonfin)) {
        hit = 1.0;
        nf = nonfin ? 1.0 : 0.0;
        rm = r;
#pragma unroll
        for (int v = 0; v < NC; ++v) {
          a.dst0[i0 + v * s0] = h[v];
          a.dst0[i0 + (NC + v) * s0] = Pi[v];
        }
        if (a.dst1 != nullptr) {
#pragma unroll
          for (int v = 0; v < NC; ++v) {
            a.dst1[i1 + v * s1] = h[v];
            a.dst1[i1 + (NC + v) * s1] = Pi[v];
          }
        }
      }
    }
  }
  if (a.mode != 0) return;
  __shared__ double sh[3][256];
  const int tid = threadIdx.x;
  sh[0][tid] = hit;
  sh[1][tid] = nf;
  sh[2][tid] = rm;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (tid < w) {
      sh[0][tid] += sh[0][tid + w];
      sh[1][tid] += sh[1][tid + w];
      sh[2][tid] = fmax(sh[2][tid], sh[2][tid + w]);
    }
    __syncthreads();
  }
  if (tid == 0) {
    double* o = a.out + RED_STRIDE * ((long)b * ((g.P * g.P * g.P + 255) / 256) + wx);
    o[0] = sh[0][0];
    o[1] = sh[1][0];
    o[2] = sh[2][0];
  }
}

// The :pasted limiter (gh_paste_kernel!, src/evolution.jl:880-905) on a
// stage value: the exact state, with the core rule, at every point inside the
// layer's outer surface (r < r_1).
//   mode 0: block layout: the owned points of the work array dst0 and, when
//           given, of the state dst1;
//   mode 1: brick layout: every brick point of dst0 and, when given, dst1.
struct PasteArgs {
  Geom g;
  Brick k;
  CaseP c;
  double t;
  double* __restrict__ dst0;
  double* __restrict__ dst1;
  int mode;
  const double* clock;  // stage_time's clock and offset
  double toff;
};
template <bool MOVE>
__global__ __launch_bounds__(256) void k_paste(PasteArgs a) {
  a.t = stage_time(a.t, a.clock, a.toff);
  double x[3];
  long i0, s0, i1 = 0, s1 = 0;
  if (a.mode == 1) {
    const Brick& k = a.k;
    int bx, by, bz;
    if (!brick_map(k, bx, by, bz)) return;
    x[0] = k.lo[0] + (double)bx * k.h;
    x[1] = k.lo[1] + (double)by * k.h;
    x[2] = k.lo[2] + (double)bz * k.h;
    i0 = i1 = brick_index(k, bx, by, bz);
    s0 = s1 = k.fs;
  } else {
    const Geom& g = a.g;
    const int P = g.P, S = g.S, G = g.G;
    int b, wx;
    if (!wg_map((P * P * P + 255) / 256, g.nblocks, b, wx)) return;
    const int p = wx * blockDim.x + threadIdx.x;
    if (p >= P * P * P) return;
    const int i = p % P, j = (p / P) % P, kk = p / (P * P);
    position(g, b, i, j, kk, x);
    const long sv = (long)S * S * S, psz = (long)P * P * P;
    i0 = (long)b * 20 * sv + (i + G) + (long)S * ((j + G) + (long)S * (kk + G));
    s0 = sv;
    i1 = (long)b * 20 * psz + p;
    s1 = psz;
  }
  if (!(hole_radius(a.c, a.t, x) < a.c.r1)) return;
  double st[20];
  case_state_any<MOVE ? BGS_ALL : BGS_STATIC>(a.c, a.t, x, st);
#pragma unroll 4
  for (int v = 0; v < 20; ++v) {
    a.dst0[i0 + v * s0] = st[v];
    if (a.dst1) a.dst1[i1 + v * s1] = st[v];
  }
}

#endif  // TGH_RHS_ONLY

}  // namespace tgh
