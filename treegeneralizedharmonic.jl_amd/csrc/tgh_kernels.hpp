// The HIP kernels of the library: ghost fill, the fused right-hand side with
// its RK4 stage epilogue, and the pointwise kernels around them.
//
// Layouts (device, FP64, column-major as the reference's TreeAMR arrays):
//   state  u[i, j, k, v, b], (N+1)^3 owned points per block (closed vertex
//          blocks), v = 0..19 (h_ab then Pi_ab), blocks in Morton order;
//   work   W[i, j, k, v, b], (N+1+2G)^3 per block, owned point (i,j,k) at
//          (i+G, j+G, k+G); G = q/2 + 1 (src/evolution.jl:1047-1053).
#pragma once
#include "tgh_device.hpp"

namespace tgh {

enum { INT_NONE = 0, INT_DAMPED = 1, INT_FROZEN = 2, INT_PASTED = 3 };
enum { GAUGE_NONE = 0, GAUGE_SAMPLED = 1, GAUGE_KS = 2 };
enum { EPI_STORE = 0, EPI_RK4 = 1 };

// Everything a kernel needs to know about the mesh and the case: passed by
// value (kernel argument space), so nothing here is a device allocation.
struct Geom {
  int N, G, S, P;          // cells per block, ghost width, work and state extents
  int nb[3];               // brick of blocks
  int periodic[3];
  int nblocks;
  double lo[3];
  double h, inv_h;
  const int* bcoord;       // 3 ints per block
  const int* lookup;       // brick index -> block
};

struct CaseP {
  Bg bg;
  int gauge, damping, interior, has_diss;
  double g0, g0_near, g0_far, g0_width, g2;
  double eps_ko;
  double center[3];
  double r0, r1, rho_max, w_ramp, rho_ramp;
};

TGH_HD void position(const Geom& g, int b, int i, int j, int k, double x[3]) {
  const int* B = g.bcoord + 3 * b;
  x[0] = g.lo[0] + (double)(B[0] * g.N + i) * g.h;
  x[1] = g.lo[1] + (double)(B[1] * g.N + j) * g.h;
  x[2] = g.lo[2] + (double)(B[2] * g.N + k) * g.h;
}

TGH_HD double damping_rate(const CaseP& c, const double x[3]) {  // src/gauge.jl:1100,1140-1147
  if (c.damping == 0) return c.g0;
  const double d1 = x[0] - c.center[0], d2 = x[1] - c.center[1], d3 = x[2] - c.center[2];
  const double r2 = d1 * d1 + d2 * d2 + d3 * d3;
  return c.g0_far + (c.g0_near - c.g0_far) * exp(-r2 / (2.0 * c.g0_width * c.g0_width));
}

TGH_HD double hole_radius(const CaseP& c, const double x[3]) {
  const double d1 = x[0] - c.center[0], d2 = x[1] - c.center[1], d3 = x[2] - c.center[2];
  return sqrt(d1 * d1 + d2 * d2 + d3 * d3);
}

// case_state_tuple with the core rule (src/initialdata.jl:1795-1796,
// src/interior.jl:615-625).
TGH_HD void core_point(const CaseP& c, const double x[3], double xc[3]) {
  xc[0] = x[0];
  xc[1] = x[1];
  xc[2] = x[2];
  if (c.interior != INT_NONE) {
    const double d[3] = {x[0] - c.center[0], x[1] - c.center[1], x[2] - c.center[2]};
    const double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    if (r < c.r0) {
      double n[3] = {0.0, 0.0, 1.0};
      if (r != 0.0) {
        n[0] = d[0] / r;
        n[1] = d[1] / r;
        n[2] = d[2] / r;
      }
      for (int q = 0; q < 3; ++q) xc[q] = c.center[q] + c.r0 * n[q];
    }
  }
}
TGH_HD void case_state(const CaseP& c, double t, const double x[3], double out[20]) {
  double xc[3];
  core_point(c, x, xc);
  background_state(c.bg, t, xc, out, out + NC);
}

// Where a point of block b's work array takes its value from (TreeAMR 0.3's
// closed-block fill on a uniform forest): -2 the Dirichlet hook, -1 an own
// primary (left as scattered), 0 a copy of (owner, loc). A face point is a
// primary in the block where its local index is below N, and in the last
// block of a non-periodic axis also at N (the evolved boundary plane).
TGH_HD int classify(const Geom& g, int b, const int l[3], int& owner, int loc[3]) {
  int ob[3];
  const int* B = g.bcoord + 3 * b;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const int gi = B[d] * g.N + l[d];
    const int tot = g.nb[d] * g.N;
    if (g.periodic[d]) {
      const int gw = ((gi % tot) + tot) % tot;
      ob[d] = gw / g.N;
      loc[d] = gw - ob[d] * g.N;
    } else {
      if (gi < 0 || gi > tot) return -2;
      const int o = gi / g.N;
      ob[d] = o < g.nb[d] - 1 ? o : g.nb[d] - 1;
      loc[d] = gi - ob[d] * g.N;
    }
  }
  owner = g.lookup[ob[0] + g.nb[0] * (ob[1] + g.nb[1] * ob[2])];
  if (owner == b && loc[0] == l[0] && loc[1] == l[1] && loc[2] == l[2]) return -1;
  return 0;
}

// XCD-aware workgroup order. MI355X dispatches consecutive workgroups of a
// launch round-robin over its 8 XCDs, each with its own L2. A launch here is
// 1-D with 8 * ceil(total / 8) workgroups, and hardware workgroup w takes
// logical workgroup (w mod 8) * per + w / 8, so each XCD runs one contiguous
// range of logical workgroups: a run of whole blocks, consecutive in Morton
// order, whose work arrays and neighbours stay in that XCD's L2.
constexpr int NUM_XCD = 8;
__device__ __forceinline__ bool wg_map(int wgpb, int nblocks, int& b, int& wx) {
  const int total = wgpb * nblocks;
  const int per = (total + NUM_XCD - 1) / NUM_XCD;
  const int flat = blockIdx.x;
  const int L = (flat % NUM_XCD) * per + flat / NUM_XCD;
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
};

// Coefficient slots of the split form, each an array over the (N+1)^3 x nb
// owned points: beta^i, d_i beta^i, d_i A^{ij}, A^{ij} (xx,xy,xz,yy,yz,zz),
// the layer weight w (-1 marks the frozen core, where F is not evaluated),
// then per Pi component c1_v = w msrc_v - rho (Pi_v - Pi_exact,v), so that
// dPi_v = w (principal part) + c1_v.
enum { CF_B0 = 0, CF_DIVB = 3, CF_DIVA = 4, CF_A = 7, CF_W = 13, CF_C1 = 14, CF_N = 24 };

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

// The fused right-hand side at one owned point per thread
// (gh_rhs_kernel!/gh_rhs_point!, src/evolution.jl:448-695,846-863) in the
// streaming order of src/evolution.jl:258-388: the state, the 30 first
// derivatives and the coefficients; d_t h complete; the source; then the ten
// Pi components one at a time.
template <int Q, int GAUGE, bool HOLE, int EPI>
__global__ __launch_bounds__(256) void k_rhs(RhsArgs a, const double* __restrict__ W, TGH_OUT_PARAMS) {
  const Geom& g = a.g;
  const int P = g.P, S = g.S, G = g.G;
  int b, wx;
  if (!wg_map((g.P * g.P * g.P + 255) / 256, g.nblocks, b, wx)) return;
  const int p = wx * blockDim.x + threadIdx.x;
  if (p >= P * P * P) return;
  const int i = p % P, j = (p / P) % P, k = p / (P * P);
  const long sv = (long)S * S * S;
  const int s1 = 1, s2 = S, s3 = S * S;
  const double* __restrict__ Wb = W + (long)b * 20 * sv;
  const long off = (long)(i + G) + (long)S * ((j + G) + (long)S * (k + G));
  const double* __restrict__ w0 = Wb + off;
  const long psz = (long)P * P * P;
  const long si0 = (long)b * 20 * psz + p;
  const long wi0 = (long)b * 20 * sv + off;

  double x[3];
  position(g, b, i, j, k, x);
  const CaseP& c = a.c;
  double wgt = 1.0, rho = 0.0, r = 0.0;
  if (HOLE) {
    r = hole_radius(c, x);
    const bool frozen = c.interior == INT_PASTED ? r < c.r1 : r < c.r0;
    if (frozen) {  // F is not evaluated in the core (a hole at rest: du = 0)
#pragma unroll 1
      for (int v = 0; v < 20; ++v) emit<EPI>(a, TGH_OUT_ARGS, si0 + v * psz, wi0 + v * sv, 0.0);
      return;
    }
  }
  const double g0 = damping_rate(c, x);
  const double epsh = c.eps_ko * g.inv_h;
  const double ih = g.inv_h, ih2 = g.inv_h * g.inv_h;

  double hv[NC], D[4][NC];
#pragma unroll
  for (int v = 0; v < NC; ++v) {
    const double* pv = w0 + v * sv;
    hv[v] = pv[0];
    D[1][v] = ih * st_d1<Q>(pv, s1);
    D[2][v] = ih * st_d1<Q>(pv, s2);
    D[3][v] = ih * st_d1<Q>(pv, s3);
  }
  Coef cf;
  metric_quantities(hv, cf);
  const double a_div = cf.alpha / cf.sqrtg;
#pragma unroll
  for (int v = 0; v < NC; ++v) {
    const double* pv = w0 + v * sv;
    double s = cf.beta[0] * D[1][v] + cf.beta[1] * D[2][v] + cf.beta[2] * D[3][v] +
               a_div * pv[NC * sv];
    if (c.has_diss) s += epsh * (st_ko<Q>(pv, s1) + st_ko<Q>(pv, s2) + st_ko<Q>(pv, s3));
    D[0][v] = s;
  }
  double msrc[NC];
  {
    double Hl[4] = {0.0, 0.0, 0.0, 0.0}, dH[4][4];
    if (GAUGE == GAUGE_SAMPLED) {
      const double* Hp = a.Hsrc + si0;
#pragma unroll
      for (int q = 0; q < 4; ++q) Hl[q] = Hp[q * psz];
#pragma unroll
      for (int n = 0; n < 16; ++n) dH[n % 4][n / 4] = Hp[(4 + n) * psz];
    } else if (GAUGE == GAUGE_KS) {
      ks_source(c.bg.M, c.bg.a, hv, D, Hl, dH);
    }
    gh_source<GAUGE != GAUGE_NONE>(cf, hv, D, Hl, dH, g0, c.g2, msrc);
  }
  if (a.m_rho != nullptr) {
    const long mi = (long)b * psz + p;
    const double rr = a.m_rho[mi];
    const double Sm[3] = {a.m_S[(long)b * 3 * psz + p], a.m_S[(long)b * 3 * psz + psz + p],
                          a.m_S[(long)b * 3 * psz + 2 * psz + p]};
    double Sij[6];
#pragma unroll
    for (int n = 0; n < 6; ++n) Sij[n] = a.m_Sij[(long)b * 6 * psz + n * psz + p];
    double ms[NC];
    matter_source(cf, hv, rr, Sm, Sij, ms);
#pragma unroll
    for (int n = 0; n < NC; ++n) msrc[n] += ms[n];
  }
  double divb, divA[3];
  metric_divergences(cf, &D[1], divb, divA);
  const double as = cf.alpha * cf.sqrtg;
  const double A00 = as * cf.gamu[0], A01 = as * cf.gamu[1], A02 = as * cf.gamu[2];
  const double A11 = as * cf.gamu[3], A12 = as * cf.gamu[4], A22 = as * cf.gamu[5];
  const double b0 = cf.beta[0], b1 = cf.beta[1], b2 = cf.beta[2];

  if (HOLE) {
    if (r < c.r1) {
      double w_, rho_;
      const double sl = (r - c.r0) / (c.r1 - c.r0);
      w_ = smoothstep(sl / c.w_ramp);
      rho_ = c.interior == INT_FROZEN ? 0.0 : c.rho_max * smoothstep((1.0 - sl) / c.rho_ramp);
      wgt = w_;
      rho = rho_;
    }
  }
  double ue[20];
  const bool layer = HOLE && r < c.r1;
  if (layer) case_state(c, a.t, x, ue);  // r >= r0 here: the core rule is the identity

#pragma unroll
  for (int v = 0; v < NC; ++v) {
    double f = D[0][v];
    if (layer) f = wgt * f - rho * (hv[v] - ue[v]);
    emit<EPI>(a, TGH_OUT_ARGS, si0 + v * psz, wi0 + v * sv, f);
  }
#pragma unroll 1
  for (int v = 0; v < NC; ++v) {
    const double* ph = w0 + v * sv;
    const double* pP = ph + NC * sv;
    const double Pi = pP[0];
    const double dh1 = ih * st_d1<Q>(ph, s1), dh2 = ih * st_d1<Q>(ph, s2), dh3 = ih * st_d1<Q>(ph, s3);
    const double dP1 = ih * st_d1<Q>(pP, s1), dP2 = ih * st_d1<Q>(pP, s2), dP3 = ih * st_d1<Q>(pP, s3);
    double s = b0 * dP1 + b1 * dP2 + b2 * dP3 + divb * Pi + divA[0] * dh1 + divA[1] * dh2 + divA[2] * dh3;
    s += A00 * (ih2 * st_d2<Q>(ph, s1)) + A11 * (ih2 * st_d2<Q>(ph, s2)) + A22 * (ih2 * st_d2<Q>(ph, s3));
    const double dxy = ih2 * st_mixed<Q>(ph, s1, s2);
    const double dxz = ih2 * st_mixed<Q>(ph, s1, s3);
    const double dyz = ih2 * st_mixed<Q>(ph, s2, s3);
    s += 2.0 * (A01 * dxy + A02 * dxz + A12 * dyz);
    if (c.has_diss) s += epsh * (st_ko<Q>(pP, s1) + st_ko<Q>(pP, s2) + st_ko<Q>(pP, s3));
    double f = s + msrc[v];
    if (layer) f = wgt * f - rho * (Pi - ue[NC + v]);
    emit<EPI>(a, TGH_OUT_ARGS, si0 + (NC + v) * psz, wi0 + (NC + v) * sv, f);
  }
}

// The split form, first half: gh_rhs_head (src/evolution.jl:258-297) at one
// owned point per thread -- the state, the 30 first derivatives, the
// coefficient set, d_t h complete (stored through the epilogue), the source --
// leaving the Pi components' coefficients in a.coef for k_pi. The frozen core
// is decided here, before any stencil is read.
template <int Q, int GAUGE, bool HOLE, int EPI>
__global__ __launch_bounds__(256) void k_head(RhsArgs a, const double* __restrict__ W,
                                              double* __restrict__ coef, TGH_OUT_PARAMS) {
  const Geom& g = a.g;
  const int P = g.P, S = g.S, G = g.G;
  int b, wx;
  if (!wg_map((g.P * g.P * g.P + 255) / 256, g.nblocks, b, wx)) return;
  const int p = wx * blockDim.x + threadIdx.x;
  if (p >= P * P * P) return;
  const int i = p % P, j = (p / P) % P, k = p / (P * P);
  const long sv = (long)S * S * S;
  const int s1 = 1, s2 = S, s3 = S * S;
  const double* __restrict__ Wb = W + (long)b * 20 * sv;
  const long off = (long)(i + G) + (long)S * ((j + G) + (long)S * (k + G));
  const double* __restrict__ w0 = Wb + off;
  const long psz = (long)P * P * P;
  const long si0 = (long)b * 20 * psz + p;
  const long wi0 = (long)b * 20 * sv + off;

  double x[3];
  position(g, b, i, j, k, x);
  const CaseP& c = a.c;
  double wgt = 1.0, rho = 0.0, r = 0.0;
  if (HOLE) {
    r = hole_radius(c, x);
    const bool frozen = c.interior == INT_PASTED ? r < c.r1 : r < c.r0;
    if (frozen) {  // F is not evaluated in the core (a hole at rest: du = 0)
#pragma unroll 1
      for (int v = 0; v < NC; ++v) emit<EPI>(a, TGH_OUT_ARGS, si0 + v * psz, wi0 + v * sv, 0.0);
      coef[(long)CF_W * g.nblocks * psz + (long)b * psz + p] = -1.0;
      return;
    }
  }
  const double g0 = damping_rate(c, x);
  const double epsh = c.eps_ko * g.inv_h;
  const double ih = g.inv_h, ih2 = g.inv_h * g.inv_h;

  double hv[NC], D[4][NC];
#pragma unroll
  for (int v = 0; v < NC; ++v) {
    const double* pv = w0 + v * sv;
    hv[v] = pv[0];
    D[1][v] = ih * st_d1<Q>(pv, s1);
    D[2][v] = ih * st_d1<Q>(pv, s2);
    D[3][v] = ih * st_d1<Q>(pv, s3);
  }
  Coef cf;
  metric_quantities(hv, cf);
  const double a_div = cf.alpha / cf.sqrtg;
#pragma unroll
  for (int v = 0; v < NC; ++v) {
    const double* pv = w0 + v * sv;
    double s = cf.beta[0] * D[1][v] + cf.beta[1] * D[2][v] + cf.beta[2] * D[3][v] +
               a_div * pv[NC * sv];
    if (c.has_diss) s += epsh * (st_ko<Q>(pv, s1) + st_ko<Q>(pv, s2) + st_ko<Q>(pv, s3));
    D[0][v] = s;
  }
  double msrc[NC];
  {
    double Hl[4] = {0.0, 0.0, 0.0, 0.0}, dH[4][4];
    if (GAUGE == GAUGE_SAMPLED) {
      const double* Hp = a.Hsrc + si0;
#pragma unroll
      for (int q = 0; q < 4; ++q) Hl[q] = Hp[q * psz];
#pragma unroll
      for (int n = 0; n < 16; ++n) dH[n % 4][n / 4] = Hp[(4 + n) * psz];
    } else if (GAUGE == GAUGE_KS) {
      ks_source(c.bg.M, c.bg.a, hv, D, Hl, dH);
    }
    gh_source<GAUGE != GAUGE_NONE>(cf, hv, D, Hl, dH, g0, c.g2, msrc);
  }
  if (a.m_rho != nullptr) {
    const long mi = (long)b * psz + p;
    const double rr = a.m_rho[mi];
    const double Sm[3] = {a.m_S[(long)b * 3 * psz + p], a.m_S[(long)b * 3 * psz + psz + p],
                          a.m_S[(long)b * 3 * psz + 2 * psz + p]};
    double Sij[6];
#pragma unroll
    for (int n = 0; n < 6; ++n) Sij[n] = a.m_Sij[(long)b * 6 * psz + n * psz + p];
    double ms[NC];
    matter_source(cf, hv, rr, Sm, Sij, ms);
#pragma unroll
    for (int n = 0; n < NC; ++n) msrc[n] += ms[n];
  }
  double divb, divA[3];
  metric_divergences(cf, &D[1], divb, divA);
  const double as = cf.alpha * cf.sqrtg;
  const double A00 = as * cf.gamu[0], A01 = as * cf.gamu[1], A02 = as * cf.gamu[2];
  const double A11 = as * cf.gamu[3], A12 = as * cf.gamu[4], A22 = as * cf.gamu[5];
  const double b0 = cf.beta[0], b1 = cf.beta[1], b2 = cf.beta[2];

  if (HOLE) {
    if (r < c.r1) {
      double w_, rho_;
      const double sl = (r - c.r0) / (c.r1 - c.r0);
      w_ = smoothstep(sl / c.w_ramp);
      rho_ = c.interior == INT_FROZEN ? 0.0 : c.rho_max * smoothstep((1.0 - sl) / c.rho_ramp);
      wgt = w_;
      rho = rho_;
    }
  }
  double ue[20];
  const bool layer = HOLE && r < c.r1;
  if (layer) case_state(c, a.t, x, ue);  // r >= r0 here: the core rule is the identity

#pragma unroll
  for (int v = 0; v < NC; ++v) {
    double f = D[0][v];
    if (layer) f = wgt * f - rho * (hv[v] - ue[v]);
    emit<EPI>(a, TGH_OUT_ARGS, si0 + v * psz, wi0 + v * sv, f);
  }
  {
    const long np = (long)g.nblocks * psz;
    double* cf0 = coef + (long)b * psz + p;
    cf0[CF_B0 * np] = b0;
    cf0[(CF_B0 + 1) * np] = b1;
    cf0[(CF_B0 + 2) * np] = b2;
    cf0[CF_DIVB * np] = divb;
    cf0[CF_DIVA * np] = divA[0];
    cf0[(CF_DIVA + 1) * np] = divA[1];
    cf0[(CF_DIVA + 2) * np] = divA[2];
    cf0[CF_A * np] = A00;
    cf0[(CF_A + 1) * np] = A01;
    cf0[(CF_A + 2) * np] = A02;
    cf0[(CF_A + 3) * np] = A11;
    cf0[(CF_A + 4) * np] = A12;
    cf0[(CF_A + 5) * np] = A22;
    cf0[CF_W * np] = layer ? wgt : 1.0;
#pragma unroll
    for (int v = 0; v < NC; ++v) {
      double c1 = msrc[v];
      if (layer) c1 = wgt * c1 - rho * (w0[(NC + v) * sv] - ue[NC + v]);
      cf0[(CF_C1 + v) * np] = c1;
    }
  }
}

// The split form, second half: the ten Pi components (gh_rhs_pi,
// src/evolution.jl:324-354) at one owned point per thread, from the work
// array and the coefficients k_head left: nine stencils of h_v and seven of
// Pi_v per component, contracted and stored through the epilogue. Light on
// registers, so many waves hide the loads.
#ifndef TGH_PI_WAVES_PER_EU
#define TGH_PI_WAVES_PER_EU 2
#endif
template <int Q, int EPI>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(TGH_PI_WAVES_PER_EU))) void k_pi(RhsArgs a, const double* __restrict__ W,
                                            const double* __restrict__ coef, TGH_OUT_PARAMS) {
  const Geom& g = a.g;
  const int P = g.P, S = g.S, G = g.G;
  int b, wx;
  if (!wg_map((g.P * g.P * g.P + 255) / 256, g.nblocks, b, wx)) return;
  const int p = wx * blockDim.x + threadIdx.x;
  if (p >= P * P * P) return;
  const int i = p % P, j = (p / P) % P, k = p / (P * P);
  const long sv = (long)S * S * S;
  const int s1 = 1, s2 = S, s3 = S * S;
  const long off = (long)(i + G) + (long)S * ((j + G) + (long)S * (k + G));
  const double* __restrict__ w0 = W + (long)b * 20 * sv + off;
  const long psz = (long)P * P * P;
  const long si0 = (long)b * 20 * psz + p;
  const long wi0 = (long)b * 20 * sv + off;
  const long np = (long)g.nblocks * psz;
  const double* __restrict__ cf0 = coef + (long)b * psz + p;
  const double wl = cf0[CF_W * np];
  if (wl < 0.0) {
#pragma unroll 1
    for (int v = 0; v < NC; ++v) emit<EPI>(a, TGH_OUT_ARGS, si0 + (NC + v) * psz, wi0 + (NC + v) * sv, 0.0);
    return;
  }
  const double b0 = cf0[CF_B0 * np], b1 = cf0[(CF_B0 + 1) * np], b2 = cf0[(CF_B0 + 2) * np];
  const double divb = cf0[CF_DIVB * np];
  const double dA0 = cf0[CF_DIVA * np], dA1 = cf0[(CF_DIVA + 1) * np], dA2 = cf0[(CF_DIVA + 2) * np];
  const double A00 = cf0[CF_A * np], A01 = cf0[(CF_A + 1) * np], A02 = cf0[(CF_A + 2) * np];
  const double A11 = cf0[(CF_A + 3) * np], A12 = cf0[(CF_A + 4) * np], A22 = cf0[(CF_A + 5) * np];
  const double epsh = a.c.eps_ko * g.inv_h;
  const double ih = g.inv_h, ih2 = g.inv_h * g.inv_h;
  const bool diss = a.c.has_diss;
#pragma unroll 1
  for (int v = 0; v < NC; ++v) {
    const double* ph = w0 + v * sv;
    const double* pP = ph + NC * sv;
    const double Pi = pP[0];
    const double dh1 = ih * st_d1<Q>(ph, s1), dh2 = ih * st_d1<Q>(ph, s2), dh3 = ih * st_d1<Q>(ph, s3);
    const double dP1 = ih * st_d1<Q>(pP, s1), dP2 = ih * st_d1<Q>(pP, s2), dP3 = ih * st_d1<Q>(pP, s3);
    double s = b0 * dP1 + b1 * dP2 + b2 * dP3 + divb * Pi + dA0 * dh1 + dA1 * dh2 + dA2 * dh3;
    s += A00 * (ih2 * st_d2<Q>(ph, s1)) + A11 * (ih2 * st_d2<Q>(ph, s2)) + A22 * (ih2 * st_d2<Q>(ph, s3));
    const double dxy = ih2 * st_mixed<Q>(ph, s1, s2);
    const double dxz = ih2 * st_mixed<Q>(ph, s1, s3);
    const double dyz = ih2 * st_mixed<Q>(ph, s2, s3);
    s += 2.0 * (A01 * dxy + A02 * dxz + A12 * dyz);
    if (diss) s += epsh * (st_ko<Q>(pP, s1) + st_ko<Q>(pP, s2) + st_ko<Q>(pP, s3));
    const double f = wl * s + cf0[(CF_C1 + v) * np];
    emit<EPI>(a, TGH_OUT_ARGS, si0 + (NC + v) * psz, wi0 + (NC + v) * sv, f);
  }
}

// ---------------------------------------------------------------------------
// The ghost fill of a work array (prepare!, src/evolution.jl:1349-1357).
struct FillArgs {
  Geom g;
  CaseP c;
  double t;
  double* __restrict__ W;
};
__global__ __launch_bounds__(256) void k_fill(FillArgs a) {
  const Geom& g = a.g;
  const int S = g.S, G = g.G;
  int b, wx;
  if (!wg_map((g.S * g.S * g.S + 255) / 256, g.nblocks, b, wx)) return;
  const int p = wx * blockDim.x + threadIdx.x;
  if (p >= S * S * S) return;
  const int l[3] = {p % S - G, (p / S) % S - G, p / (S * S) - G};
  int owner, loc[3];
  const int cl = classify(g, b, l, owner, loc);
  if (cl == -1) return;
  const long sv = (long)S * S * S;
  double* dst = a.W + (long)b * 20 * sv + p;
  if (cl == -2) {
    double x[3], s[20];
    position(g, b, l[0], l[1], l[2], x);
    case_state(a.c, a.t, x, s);
#pragma unroll
    for (int v = 0; v < 20; ++v) dst[v * sv] = s[v];
  } else {
    const double* src = a.W + (long)owner * 20 * sv + (loc[0] + G) + (long)S * ((loc[1] + G) + (long)S * (loc[2] + G));
#pragma unroll
    for (int v = 0; v < 20; ++v) dst[v * sv] = src[v * sv];
  }
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
  core_point(a.c, x, xc);
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
  int mode;  // 0 exact state, 1 add noise, 2 paste inside r1
};
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
    for (int v = 0; v < 20; ++v) ub[v * psz] += a.amp * (2.0 * point_draw(a.base, x, v + 1) - 1.0);
    return;
  }
  if (a.mode == 2 && !(hole_radius(a.c, x) < a.c.r1)) return;
  double s[20];
  case_state(a.c, a.t, x, s);
  for (int v = 0; v < 20; ++v) ub[v * psz] = s[v];
}

// Reductions over owned points: each workgroup of 256 points of one block
// folds its points in shared memory and writes one partial; the host folds
// the partials in (block, workgroup) order, so the answer does not depend on
// the launch.
struct ReduceArgs {
  Geom g;
  CaseP c;
  double t;
  const double* __restrict__ u;
  double* __restrict__ out;  // per workgroup: [0] max speed | sum e^2, [1] nonfinite | count, [2] linf
  int mode;                  // 0 speed, 1 error
};
__device__ __forceinline__ bool is_evolved(const CaseP& c, const double x[3]) {
  return c.interior == INT_NONE || hole_radius(c, x) >= c.r1;
}
__global__ __launch_bounds__(256) void k_reduce(ReduceArgs a) {
  const Geom& g = a.g;
  const int P = g.P;
  int b, wx;
  if (!wg_map((g.P * g.P * g.P + 255) / 256, g.nblocks, b, wx)) return;
  const int p = wx * blockDim.x + threadIdx.x;
  const long psz = (long)P * P * P;
  double r0 = 0.0, r1 = 0.0, r2 = 0.0;
  if (p < P * P * P) {
    const int i = p % P, j = (p / P) % P, k = p / (P * P);
    const double* ub = a.u + (long)b * 20 * psz + p;
    double x[3];
    position(g, b, i, j, k, x);
    if (is_evolved(a.c, x)) {
      if (a.mode == 0) {
        double hv[NC];
        bool fin = true;
#pragma unroll
        for (int v = 0; v < 20; ++v) {
          const double val = ub[v * psz];
          fin = fin && isfinite(val);
          if (v < NC) hv[v] = val;
        }
        r1 = fin ? 0.0 : 1.0;
        Coef cf;
        metric_quantities(hv, cf);
        r0 = cf.alpha * sqrt(cf.gamu[0] + cf.gamu[3] + cf.gamu[5]) +
             sqrt(cf.beta[0] * cf.beta[0] + cf.beta[1] * cf.beta[1] + cf.beta[2] * cf.beta[2]);
      } else {
        double s[20];
        case_state(a.c, a.t, x, s);
        double e2 = 0.0;
#pragma unroll
        for (int v = 0; v < 20; ++v) {
          const double e = ub[v * psz] - s[v];
          e2 += e * e;
          r2 = fmax(r2, fabs(e));
        }
        int l[3] = {i, j, k}, owner, loc[3];
        if (classify(g, b, l, owner, loc) == -1) {
          r0 = e2;
          r1 = 1.0;
        }
      }
    }
  }
  __shared__ double sh[3][256];
  const int tid = threadIdx.x;
  sh[0][tid] = r0;
  sh[1][tid] = r1;
  sh[2][tid] = r2;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (tid < w) {
      const double o0 = sh[0][tid + w], m0 = sh[0][tid];
      if (a.mode == 0)
        sh[0][tid] = (o0 > m0 || isnan(o0)) ? o0 : m0;
      else
        sh[0][tid] = m0 + o0;
      sh[1][tid] = sh[1][tid] + sh[1][tid + w];
      sh[2][tid] = fmax(sh[2][tid], sh[2][tid + w]);
    }
    __syncthreads();
  }
  if (tid == 0) {
    double* o = a.out + 3 * ((long)b * ((g.P * g.P * g.P + 255) / 256) + wx);
    o[0] = sh[0][0];
    o[1] = sh[1][0];
    o[2] = sh[2][0];
  }
}

}  // namespace tgh
