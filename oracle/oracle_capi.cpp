// C entry points of the CPU oracle, for the tests (ctypes) and bench.py's
// cpu_baseline leg. TEST INFRASTRUCTURE ONLY: the product library never links
// this. It reads the case descriptor of include/tgh.h so that a test hands
// the same case to the oracle and to the library.
#include <omp.h>

#include <cfloat>
#include <cmath>

#include "../include/tgh.h"
#include "gh_oracle_fit.hpp"
#include "gh_oracle_mesh.hpp"

using namespace orc;

namespace {
BgParams bg_from(const tgh_desc* d) {
  BgParams p;
  p.kind = d->background;
  p.M = d->M;
  p.a = d->a;
  for (int i = 0; i < 3; ++i) p.c[i] = d->center[i];
  for (int i = 0; i < 3; ++i) p.u[i] = -d->velocity[i];  // boost(bg, -v): the hole moves at v
  if (d->background == TGH_BG_GAUGE_WAVE) {
    p.A = d->wave_amp;
    p.d = d->wave_len;
  } else if (d->background == TGH_BG_SHIFTED_MINKOWSKI) {
    p.A = d->wave_amp;
    p.w = d->wave_len;
  }
  return p;
}
Case case_from(const tgh_desc* d) {
  Case c;
  c.bg = bg_from(d);
  c.gauge = d->gauge;
  {
    // KerrSchildSource(T; M, spin = (0, 0, a), velocity = v) (gauge.jl:263-278):
    // u = gamma (1, v), S = (gamma v.a, a + (gamma - 1)(v.a)/v^2 v)
    const double* v = d->velocity;
    const double v2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    const double gam = 1.0 / std::sqrt(1.0 - v2);
    const double av[3] = {0.0, 0.0, d->a};
    const double va = v[0] * av[0] + v[1] * av[1] + v[2] * av[2];
    c.ks_S[0] = gam * va;
    for (int i = 0; i < 3; ++i) c.ks_S[1 + i] = v2 == 0.0 ? av[i] : av[i] + ((gam - 1.0) * va / v2) * v[i];
    c.ks_u[0] = gam;
    for (int i = 0; i < 3; ++i) c.ks_u[1 + i] = gam * v[i];
  }
  c.damping = d->damping;
  c.g0 = d->gamma0;
  c.g0_near = d->gamma0_near;
  c.g0_far = d->gamma0_far;
  c.g0_width = d->gamma0_width;
  c.g2 = d->gamma2;
  c.eps_ko = d->eps_ko;
  c.diss = d->dissipation;
  c.eps_in = d->eps_in;
  c.eps_r1 = d->eps_r1;
  c.eps_rh = d->eps_rh;
  c.interior = d->interior;
  for (int i = 0; i < 3; ++i) c.center[i] = d->center[i];
  for (int i = 0; i < 3; ++i) c.vel[i] = d->velocity[i];
  c.r0 = d->r0;
  c.delta = std::cbrt(DBL_EPSILON) * d->r0;  // layer_rate_step (interior.jl:524-537)
  c.r1 = d->r1;
  c.rho_max = d->rho_max;
  c.w_ramp = d->w_ramp;
  c.rho_ramp = d->rho_ramp;
  c.shaped = d->shaped != 0;
  c.lmax = d->shape_lmax;
  for (int i = 0; i < 81; ++i) c.shape[i] = d->shape[i];
  c.r_in = d->r_in;
  c.r_out = d->r_out;
  c.offset = d->offset;
  c.thickness = d->thickness;
  return c;
}
V10<double> v10(const double* p) {
  V10<double> v;
  for (int i = 0; i < NC; ++i) v[i] = p[i];
  return v;
}
void mq_out(const MQ<double>& m, double* out) {
  // alpha, beta[3], sqrtg, gamu[9] (row-major), gu4[16], guo[16], g4[16]
  out[0] = m.alpha;
  for (int i = 0; i < 3; ++i) out[1 + i] = m.beta[i];
  out[4] = m.sqrtg;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) out[5 + 3 * i + j] = m.gamu[i][j];
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b) {
      out[14 + 4 * a + b] = m.gu4[a][b];
      out[30 + 4 * a + b] = m.guo[a][b];
      out[46 + 4 * a + b] = m.g4[a][b];
    }
}
}  // namespace

extern "C" {

// ---- stencils ----
int orc_derivative_weights(int q, int m, long long* num, long long* den, double* w) {
  auto ws = rational_derivative_weights(q, m);
  for (size_t i = 0; i < ws.size(); ++i) {
    num[i] = (long long)ws[i].n;
    den[i] = (long long)ws[i].d;
    w[i] = (double)num[i] / (double)den[i];
  }
  return (int)ws.size();
}
int orc_dissipation_weights(int r, long long* num, long long* den, double* w) {
  auto ws = rational_dissipation_weights(r);
  for (size_t i = 0; i < ws.size(); ++i) {
    num[i] = (long long)ws[i].n;
    den[i] = (long long)ws[i].d;
    w[i] = (double)num[i] / (double)den[i];
  }
  return (int)ws.size();
}

// ---- pointwise algebra ----
void orc_metric_quantities(const double* h, double* out) { mq_out(metric_quantities(v10(h)), out); }

// Forward-mode derivative of metric_quantities along direction dh (for the
// chain-rule checks): out = d(alpha), d(beta)[3], d(sqrtg), d(gamu)[9].
void orc_metric_quantities_jvp(const double* h, const double* dh, double* out) {
  using D = Dual<double, 1>;
  V10<D> hv;
  for (int i = 0; i < NC; ++i) {
    hv[i] = D(h[i]);
    hv[i].d[0] = dh[i];
  }
  MQ<D> m = metric_quantities(hv);
  out[0] = m.alpha.d[0];
  for (int i = 0; i < 3; ++i) out[1 + i] = m.beta[i].d[0];
  out[4] = m.sqrtg.d[0];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) out[5 + 3 * i + j] = m.gamu[i][j].d[0];
}

void orc_gh_node_source(const double* h, const double* dg, const double* Hl, const double* dHl,
                        double g0, double g2, double* out) {
  MQ<double> m = metric_quantities(v10(h));
  double d[4][4][4], H2[4][4];
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b) {
      H2[a][b] = dHl[4 * a + b];
      for (int c = 0; c < 4; ++c) d[a][b][c] = dg[16 * a + 4 * b + c];
    }
  auto s = gh_node_source(m, d, Hl, H2, g0, g2);
  for (int n = 0; n < NC; ++n) out[n] = s[n];
}
void orc_gh_node_source_lean(const double* h, const double* dth, const double* dh, const double* Hl,
                             const double* dHl, double g0, double g2, double* out) {
  MQ<double> m = metric_quantities(v10(h));
  V10<double> d[3] = {v10(dh), v10(dh + 10), v10(dh + 20)};
  double H2[4][4];
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b) H2[a][b] = dHl[4 * a + b];
  auto s = gh_node_source_lean(m, v10(dth), d, Hl, H2, g0, g2);
  for (int n = 0; n < NC; ++n) out[n] = s[n];
}
void orc_metric_divergences(const double* h, const double* dh, double* out) {
  MQ<double> m = metric_quantities(v10(h));
  V10<double> d[3] = {v10(dh), v10(dh + 10), v10(dh + 20)};
  double divb, divA[3];
  metric_divergences(m, d, divb, divA);
  out[0] = divb;
  for (int i = 0; i < 3; ++i) out[1 + i] = divA[i];
}
void orc_metric_derivatives(const double* h, const double* dh, double* out) {
  MQ<double> m = metric_quantities(v10(h));
  V10<double> d[3] = {v10(dh), v10(dh + 10), v10(dh + 20)};
  MDer<double> md = metric_derivatives(m, d);
  for (int i = 0; i < 3; ++i) out[i] = md.dalpha[i];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) out[3 + 3 * i + j] = md.dbeta[i][j];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      for (int k = 0; k < 3; ++k) out[12 + 9 * i + 3 * j + k] = md.dA[i][j][k];
}
void orc_gh_node_rhs_expanded(const double* h, const double* P, const double* dh, const double* dP,
                              const double* ddh, const double* Hl, const double* dHl, double g0,
                              double g2, double* dth, double* dtP) {
  V10<double> d[3] = {v10(dh), v10(dh + 10), v10(dh + 20)};
  V10<double> e[3] = {v10(dP), v10(dP + 10), v10(dP + 20)};
  V10<double> dd[6];
  for (int n = 0; n < 6; ++n) dd[n] = v10(ddh + 10 * n);
  double H2[4][4];
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b) H2[a][b] = dHl[4 * a + b];
  V10<double> o1, o2;
  gh_node_rhs_expanded(v10(h), v10(P), d, e, dd, Hl, H2, g0, g2, o1, o2);
  for (int n = 0; n < NC; ++n) {
    dth[n] = o1[n];
    dtP[n] = o2[n];
  }
}
void orc_gh_node_rhs_flux(const double* h, const double* P, const double* dh, const double* Hl,
                          const double* dHl, double g0, double g2, double* dtg, double* F,
                          double* msrc) {
  V10<double> d[3] = {v10(dh), v10(dh + 10), v10(dh + 20)};
  double H2[4][4];
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b) H2[a][b] = dHl[4 * a + b];
  V10<double> o, f[3], m;
  gh_node_rhs_flux(v10(h), v10(P), d, Hl, H2, g0, g2, o, f, m);
  for (int n = 0; n < NC; ++n) {
    dtg[n] = o[n];
    msrc[n] = m[n];
    for (int i = 0; i < 3; ++i) F[10 * i + n] = f[i][n];
  }
}
void orc_matter_source(const double* h, double rho, const double* S, const double* Sij6, double* out) {
  MQ<double> m = metric_quantities(v10(h));
  double Sij[3][3] = {{Sij6[0], Sij6[1], Sij6[2]}, {Sij6[1], Sij6[3], Sij6[4]}, {Sij6[2], Sij6[4], Sij6[5]}};
  auto o = matter_source(m, rho, S, Sij);
  for (int n = 0; n < NC; ++n) out[n] = o[n];
}
void orc_ks_algebraic_source(double M, const double* S, const double* u, const double* h,
                             const double* dth, const double* dh, double* Hl, double* dHl) {
  V10<double> d[3] = {v10(dh), v10(dh + 10), v10(dh + 20)};
  double H2[4][4];
  ks_algebraic_source<double>(M, S, u, v10(h), v10(dth), d, Hl, H2);
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b) dHl[4 * a + b] = H2[a][b];
}

// ---- backgrounds ----
void orc_dmetric(const tgh_desc* desc, double t, const double* x, double* g, double* dg) {
  double G[4][4], DG[4][4][4];
  dmetric(bg_from(desc), t, x[0], x[1], x[2], G, DG);
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b) {
      g[4 * a + b] = G[a][b];
      for (int c = 0; c < 4; ++c) dg[16 * a + 4 * b + c] = DG[a][b][c];
    }
}
void orc_background_state(const tgh_desc* desc, double t, const double* x, double* h, double* P,
                          double* dh) {
  V10<double> hh, pp, d[3];
  background_state(bg_from(desc), t, x, hh, pp, d);
  for (int n = 0; n < NC; ++n) {
    h[n] = hh[n];
    P[n] = pp[n];
    for (int i = 0; i < 3; ++i) dh[10 * i + n] = d[i][n];
  }
}
void orc_gauge_source_grad(const tgh_desc* desc, double t, const double* x, double* Hl, double* dHl) {
  double H2[4][4];
  gauge_source_grad(bg_from(desc), t, x[0], x[1], x[2], Hl, H2);
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b) dHl[4 * a + b] = H2[a][b];
}
void orc_case_state(const tgh_desc* desc, double t, const double* x, double* out) {
  Case c = case_from(desc);
  case_state(c, t, x, out);
}
void orc_background_state_rate(const tgh_desc* desc, double t, const double* x, double* h, double* P,
                               double* dth, double* dtP) {
  Case c = case_from(desc);
  V10<double> a, b, e, f;
  background_state_rate(c.bg, t, x, c.delta, a, b, e, f);
  for (int n = 0; n < NC; ++n) {
    h[n] = a[n];
    P[n] = b[n];
    dth[n] = e[n];
    dtP[n] = f[n];
  }
}
void orc_core_target_rate(const tgh_desc* desc, double t, const double* x, double* dth, double* dtP) {
  Case c = case_from(desc);
  V10<double> e, f;
  core_target_rate(c, t, x, e, f);
  for (int n = 0; n < NC; ++n) {
    dth[n] = e[n];
    dtP[n] = f[n];
  }
}
void orc_core_relaxation_rate(const double* h, const double* ht, double rho, const double* dht, double* out) {
  V10<double> o = core_relaxation_rate(v10(h), v10(ht), rho, v10(dht));
  for (int n = 0; n < NC; ++n) out[n] = o[n];
}
double orc_layer_rate_step(const tgh_desc* desc) { return case_from(desc).delta; }
double orc_point_draw(uint64_t base, const double* x, int v) { return point_draw(base, x, v); }
double orc_smoothstep(double s) { return smoothstep(s); }

// ---- the mesh problem ----
static Problem* problem_from(const tgh_desc* d) {
  auto* p = new Problem();
  p->mesh.q = d->q;
  p->mesh.N = d->N;
  for (int i = 0; i < 3; ++i) {
    p->mesh.nb[i] = d->nblocks[i];
    p->mesh.lo[i] = d->lo[i];
    p->mesh.periodic[i] = d->periodic[i] != 0;
    p->mesh.reflect[i][0] = d->reflecting[i][0] != 0;
    p->mesh.reflect[i][1] = d->reflecting[i][1] != 0;
  }
  p->mesh.rotating = d->rotating != 0;
  p->mesh.h = (d->hi[0] - d->lo[0]) / ((double)d->nblocks[0] * d->N);
  p->cs = case_from(d);
  if (d->r_gate > 0) {
    p->bounds.on = true;
    p->bounds.a_min = d->alpha_min;
    p->bounds.a_max = d->alpha_max;
    p->bounds.l_min = d->lambda_min;
    p->bounds.l_max = d->lambda_max;
    p->bounds.b_max = d->beta_max;
    p->bounds.K_max = d->K_max;
    p->bounds.r_gate = d->r_gate;
  }
  return p;
}
void* orc_problem_create(const tgh_desc* d) {
  Problem* p = problem_from(d);
  p->setup();
  return p;
}
// A frozen hierarchy: leaves as (level, i, j, k) quadruples, i, j, k the
// leaf's coordinates in the brick of its level; p the prolongation order.
void* orc_problem_create_forest(const tgh_desc* d, const int* leaves, int nleaves, int prolong) {
  std::vector<std::array<int, 4>> lv;
  for (int n = 0; n < nleaves; ++n) lv.push_back({leaves[4 * n], leaves[4 * n + 1], leaves[4 * n + 2], leaves[4 * n + 3]});
  Problem* p = problem_from(d);
  p->mesh.half_open = d->ownership == TGH_OWNERSHIP_HALF_OPEN;
  p->mesh.init_forest(lv, prolong);
  p->setup();
  return p;
}
// The :fitted target: tw host array of 40 (N+1)^3 nblocks doubles (copied).
void orc_set_target(void* pp, const double* tw, double t_f, int rate, double trail) {
  Problem* p = (Problem*)pp;
  const long n = 40 * p->mesh.P * p->mesh.P * p->mesh.P * p->mesh.nblocks();
  p->tw.assign(tw, tw + n);
  p->t_f = t_f;
  p->rate = rate != 0;
  p->trail = trail;
}
double orc_shape_series(const double* shape, int lmax, const double* n) { return shape_series(shape, lmax, n); }
namespace {
// tgh_fit (coeffs a HOST pointer here) as the oracle's FitParams.
FitParams fit_from(const tgh_fit* f) {
  FitParams p;
  p.lmax = f->lmax;
  p.cont = f->cont;
  p.tilde = f->tilde != 0;
  p.rbar = f->rbar;
  for (int d = 0; d < 3; ++d) {
    p.c0[d] = f->center[d];
    p.v[d] = f->velocity[d];
  }
  p.bounds.on = true;
  p.bounds.a_min = f->alpha_min;
  p.bounds.a_max = f->alpha_max;
  p.bounds.l_min = f->lambda_min;
  p.bounds.l_max = f->lambda_max;
  p.bounds.b_max = f->beta_max;
  p.bounds.K_max = f->K_max;
  p.coeffs = f->coeffs;
  p.t = f->t;
  return p;
}
}  // namespace
// fit_variables_at / state_from_fit / fit_state at one point (src/fit.jl:182-212,
// 721-755), for the tests.
void orc_fit_variables_at(const tgh_fit* f, const double* x, double t, double* out) {
  fit_variables_at(fit_from(f), x, t, out);
}
void orc_state_from_fit(const double* v, int tilde, double* h, double* P) { state_from_fit(v, tilde != 0, h, P); }
void orc_fit_state(const tgh_fit* f, const double* x, double t, double* h, double* P) {
  fit_state(fit_from(f), x, t, h, P);
}
// fill_target! on the problem's owned points (src/fit.jl:1070-1151): r_fill =
// (r_out - offset) + h about the interior's center; out has 40 P^3 nblocks doubles.
void orc_fill_target(void* pp, double* out, const tgh_fit* fa, const tgh_fit* fb, double t, double h, int rate,
                     int clear) {
  Problem* p = (Problem*)pp;
  const Mesh& m = p->mesh;
  const FitParams a = fit_from(fa);
  FitParams b;
  if (fb) b = fit_from(fb);
  const double r_fill = (p->cs.r_out - p->cs.offset) + h;
  fill_target(a, fb ? &b : nullptr, t, r_fill, h, p->cs.center, p->cs.vel, rate != 0, clear != 0, m.nblocks(),
              (int)m.P, [&](long bb, int i, int j, int k, double x[3]) { m.position(bb, i, j, k, x); }, out);
}
void orc_interior_geo(const tgh_desc* d, double t, const double* x, double* out) {
  Case c = case_from(d);
  IGeo g = interior_geo(c, t, x);
  out[0] = g.r;
  out[1] = g.r0;
  out[2] = g.r1;
}
int orc_block_level(void* p, long b) {
  auto& m = ((Problem*)p)->mesh;
  return m.forest() ? m.level[b] : 0;
}
// Where work point l of block b takes its value from: 0 primary, 1 copy,
// 2 Dirichlet, 3 prolongation; src gets (block, loc or stencil start, half mask).
int orc_forest_classify(void* pp, long b, long i, long j, long k, long* src) {
  auto& m = ((Problem*)pp)->mesh;
  long l[3] = {i, j, k};
  Mesh::Src s;
  int c = m.forest_classify(b, l, s);
  src[0] = s.b;
  for (int d = 0; d < 3; ++d) src[1 + d] = s.loc[d];
  src[4] = s.half[0] | (s.half[1] << 1) | (s.half[2] << 2);
  src[5] = s.mirror | (s.turn << 3);  // the mirrored axes and the quarter turns
  return c;
}
void orc_prolong_weights(void* pp, double* out) {
  auto& m = ((Problem*)pp)->mesh;
  for (size_t k = 0; k < m.wmid.size(); ++k) out[k] = m.wmid[k];
}
void orc_problem_destroy(void* p) { delete (Problem*)p; }
long orc_state_length(void* p) { return ((Problem*)p)->mesh.state_size(); }
long orc_work_length(void* p) { return ((Problem*)p)->mesh.work_size(); }
const double* orc_work(void* p) { return ((Problem*)p)->work.data(); }
void orc_block_coords(void* p, int* out) {
  auto& bl = ((Problem*)p)->mesh.blocks;
  for (size_t b = 0; b < bl.size(); ++b)
    for (int d = 0; d < 3; ++d) out[3 * b + d] = bl[b][d];
}
void orc_fill_exact(void* p, double* u, double t) { ((Problem*)p)->fill_exact(u, t); }
void orc_add_noise(void* p, double* u, uint64_t base, double amp) { ((Problem*)p)->add_noise(u, base, amp); }
void orc_prepare(void* p, const double* u, double t) { ((Problem*)p)->prepare(u, t); }
void orc_rhs(void* p, double* du, double t, int nthreads) { ((Problem*)p)->rhs(du, t, nthreads); }
void orc_gh_rhs(void* p, double* du, const double* u, double t, int nthreads) {
  ((Problem*)p)->gh_rhs(du, u, t, nthreads);
}
void orc_host_rhs_at(void* p, long b, long i, long j, long k, double t, double* out) {
  ((Problem*)p)->host_rhs_at(b, i, j, k, t, out);
}
double orc_rhs_scale(void* p, long b, long i, long j, long k) { return ((Problem*)p)->rhs_scale(b, i, j, k); }
void orc_rk4_step(void* p, double* u, double t, double dt, int nthreads) {
  ((Problem*)p)->rk4_step(u, t, dt, nthreads);
}
void orc_settle(void* p, double* u, double t) { ((Problem*)p)->settle(u, t); }
void orc_stage_limit(void* p, double* u, double t) { ((Problem*)p)->stage_limit(u, t); }
long orc_apply_bounds(void* p, double* u, double t) { return ((Problem*)p)->apply_bounds(u, t); }

// bounds_project at one point (bounds.jl:510-527), for the tests: h, P in,
// ho, Po out; returns hit | nonfinite << 1.
int orc_bounds_project(const double* h, const double* P, const double* bnd, double* ho, double* Po) {
  StateBounds bd;
  bd.on = true;
  bd.a_min = bnd[0];
  bd.a_max = bnd[1];
  bd.l_min = bnd[2];
  bd.l_max = bnd[3];
  bd.b_max = bnd[4];
  bd.K_max = bnd[5];
  bd.r_gate = bnd[6];
  bool nf;
  const bool hit = bounds_project(h, P, bd, ho, Po, nf);
  return (hit ? 1 : 0) | (nf ? 2 : 0);
}
// sym_eigen3 (bounds.jl:382-396) of a symmetric 3x3 (row-major in A), for the tests.
void orc_sym_eigen3(const double* A, double* lam, double* V) {
  double a[3][3], v[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) a[i][j] = A[3 * i + j];
  sym_eigen3(a, lam, v);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) V[3 * i + j] = v[i][j];
}
double orc_max_speed(void* p, const double* u, double t) { return ((Problem*)p)->max_speed(u, t); }
// gh_rhs! with its dual tangent along (udot, pd[5] = d(gamma0_near,
// gamma0_far, gamma2, rho_max, rho_ramp)); 0, or 1 for a case the tangent
// pass does not cover.
int orc_gh_rhs_jvp(void* p, const double* u, const double* udot, double t, const double* pd, double* du,
                   double* du_dot, int nthreads) {
  try {
    ((Problem*)p)->gh_rhs_jvp(u, udot, t, pd, du, du_dot, nthreads);
  } catch (const std::exception&) {
    return 1;
  }
  return 0;
}
int orc_rk4_step_jvp(void* p, double* u, double* udot, double t, double dt, const double* pd, int nthreads) {
  try {
    ((Problem*)p)->rk4_step_jvp(u, udot, t, dt, pd, nthreads);
  } catch (const std::exception&) {
    return 1;
  }
  return 0;
}
double orc_error_tangent(void* p, const double* u, const double* udot, double t) {
  return ((Problem*)p)->error_tangent(u, udot, t);
}
// The mirror parity of variable v: bit d set when it is odd across x^d = 0.
int orc_parity_odd_mask(int v) { return parity_odd_mask(v); }
// The quarter turn's map of variable v: (Q u)[v] = sign * u[src], returned as
// sign * (src + 1), the reference's state_rotation entry.
int orc_rotation(int v) { return (int)rot_sign(v) * (rot_src(v) + 1); }
// The matter fields every right-hand side adds (NULL rho: vacuum again).
void orc_set_matter(void* pp, const double* rho, const double* S, const double* Sij) {
  Problem* p = (Problem*)pp;
  const long n = p->mesh.P * p->mesh.P * p->mesh.P * p->mesh.nblocks();
  if (!rho) {
    p->mat_rho.clear();
    p->mat_S.clear();
    p->mat_Sij.clear();
    return;
  }
  p->mat_rho.assign(rho, rho + n);
  p->mat_S.assign(S, S + 3 * n);
  p->mat_Sij.assign(Sij, Sij + 6 * n);
}
// with_interior(p, with_rho_max(p.interior, rho)): the layer's rate.
void orc_set_rho_max(void* p, double rho) { ((Problem*)p)->cs.rho_max = rho; }
void orc_error_rows(void* p, const double* u, double t, double shell_lo, double shell_hi, double* out) {
  ((Problem*)p)->error_rows(u, t, shell_lo, shell_hi, out);
}
void orc_error_norms(void* p, const double* u, double t, double* l2, double* linf) {
  ((Problem*)p)->error_norms(u, t, *l2, *linf);
}
void orc_gauge_constraint(void* p, const double* u, double t, double* l2, double* linf) {
  ((Problem*)p)->gauge_constraint(u, t, l2, linf);
}
void orc_adm_constraint(void* p, const double* u, double t, double* ham, double* mom_l2, double* mom_linf) {
  ((Problem*)p)->adm_constraint(u, t, ham, mom_l2, mom_linf);
}
void orc_tau_blocks(void* p, const double* u, double t, double scale, double eps, double* out) {
  ((Problem*)p)->tau_blocks(u, t, scale, eps, out);
}
double orc_field_scale(void* p, const double* u, double t) { return ((Problem*)p)->field_scale(u, t); }
// 0, or 1 when the forests admit no transfer (a jump of two levels).
int orc_transfer(void* to, void* from, const double* u_from, double t, double* u_new) {
  try {
    ((Problem*)to)->transfer_from(*(Problem*)from, u_from, t, u_new);
  } catch (const std::exception&) {
    return 1;
  }
  return 0;
}
int orc_is_primary(void* p, long b, long i, long j, long k) {
  return ((Problem*)p)->mesh.is_primary(b, i, j, k) ? 1 : 0;
}

// The FP64 operation count of F at one interior point of the :none branch
// (q, dissipation on iff eps_ko != 0, no gauge source): the streaming kernel
// instantiated with the counting type on a prepared work array.
double orc_count_rhs_flops(void* pp, long b, long i, long j, long k, int head_only) {
  Problem* p = (Problem*)pp;
  double x[3];
  p->mesh.position(b, i, j, k, x);
  Cnt::ops = 0;
  V10<Cnt> dth, dtP;
  Cnt g0(damping_rate(p->cs, 0.0, x)), g2(p->cs.g2), epsh(p->cs.eps_ko / p->mesh.h);
  p->rhs_point_F<Cnt>(b, i, j, k, 0.0, g0, g2, epsh, p->cs.eps_ko != 0.0, dth, dtP, head_only != 0);
  return (double)Cnt::ops;
}

int orc_max_threads(void) { return omp_get_max_threads(); }
}
