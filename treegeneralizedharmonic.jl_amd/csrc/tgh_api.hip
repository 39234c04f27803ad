// The C ABI (include/tgh.h): problem set-up, the verbs and the fused RK4
// stepper, over the kernels of tgh_kernels.hpp.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <exception>
#include <mutex>
#include <string>
#include <vector>

#include "tgh_forest.hpp"
#include "tgh_problem.hpp"
#include "tgh_adjoint.hpp"
#include "tgh_fit.hpp"

using namespace tgh;

namespace {
thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}
#define HIP_TRY(expr)                                                                      \
  do {                                                                                     \
    hipError_t e_ = (expr);                                                                \
    if (e_ != hipSuccess)                                                                  \
      return fail(TGH_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(e_));         \
  } while (0)

uint64_t morton3(uint32_t x, uint32_t y, uint32_t z) {
  uint64_t k = 0;
  for (int bit = 0; bit < 21; ++bit) {
    k |= (uint64_t)((x >> bit) & 1) << (3 * bit);
    k |= (uint64_t)((y >> bit) & 1) << (3 * bit + 1);
    k |= (uint64_t)((z >> bit) & 1) << (3 * bit + 2);
  }
  return k;
}
}  // namespace


namespace tgh {

// 1-D grids of 8 * ceil(total / 8) workgroups of 256 threads, which the
// kernels map to (block, workgroup-in-block) XCD by XCD (wg_map).
dim3 xcd_grid(long total) {
  const long per = (total + NUM_XCD - 1) / NUM_XCD;
  return dim3((unsigned)(NUM_XCD * per));
}
dim3 grid_points(const tgh_problem* p) {
  const long P3 = (long)p->geom.P * p->geom.P * p->geom.P;
  return xcd_grid((P3 + 255) / 256 * p->geom.nblocks);
}
dim3 grid_work(const tgh_problem* p) {
  const long S3 = (long)p->geom.S * p->geom.S * p->geom.S;
  return xcd_grid((S3 + 255) / 256 * p->geom.nblocks);
}

}  // namespace tgh

namespace {

// lmax: the finest level of a forest (0 on a uniform one); the radius checks
// take its spacing, the one at a hole that hole_forest builds around.
int check_desc(const tgh_desc* d, int lmax = 0) {
  if (!d) return fail(TGH_EINVAL, "desc is NULL");
  if (d->float_bits != 64)
    return fail(TGH_EINVAL, "float_bits = " + std::to_string(d->float_bits) + ": the kernels compute in FP64 "
                            "only (64)");
  if (d->q < 2 || d->q > 8 || d->q % 2)
    return fail(TGH_EINVAL, "q = " + std::to_string(d->q) + ": the centred stencils are built for the "
                            "even orders 2, 4, 6 and 8");
  const int G = d->q / 2 + 1;
  if (d->N < 2 * G)
    return fail(TGH_EINVAL, "a closed vertex block needs N >= 2G = " + std::to_string(2 * G) +
                                " at q=" + std::to_string(d->q) + ", got N=" + std::to_string(d->N));
  for (int i = 0; i < 3; ++i)
    if (d->nblocks[i] < 1) return fail(TGH_EINVAL, "nblocks: each axis holds one block or more");
  double h0 = (d->hi[0] - d->lo[0]) / ((double)d->nblocks[0] * d->N);
  for (int i = 0; i < 3; ++i) {
    double hi_ = (d->hi[i] - d->lo[i]) / ((double)d->nblocks[i] * d->N);
    if (!(hi_ > 0) || std::fabs(hi_ - h0) > 1e-12 * h0)
      return fail(TGH_EINVAL, "the box widths over nblocks * N give a different spacing on each "
                              "axis; the blocks' cells are cubes");
  }
  h0 /= (double)(1 << lmax);
  if (d->ownership != TGH_OWNERSHIP_CLOSED && d->ownership != TGH_OWNERSHIP_HALF_OPEN)
    return fail(TGH_EINVAL, "unknown ownership of block faces");
  for (int i = 0; i < 3; ++i)
    if (d->periodic[i] && (d->reflecting[i][0] || d->reflecting[i][1]))
      return fail(TGH_EINVAL, "axis " + std::to_string(i) + " is marked periodic and reflecting: a wrap "
                              "joins its two faces, a mirror ends the box at one, so pick one");
  if (d->rotating) {
    // the seam glues the low x face to the low y face (src/initialdata.jl:240-290's checks)
    for (int i = 0; i < 2; ++i) {
      if (d->periodic[i]) return fail(TGH_EINVAL, "the rotating seam's axis " + std::to_string(i) + " is periodic");
      if (d->reflecting[i][0] || d->reflecting[i][1])
        return fail(TGH_EINVAL, "the rotating seam's axis " + std::to_string(i) + " also reflects");
    }
    if (d->lo[0] != d->lo[1] || d->nblocks[0] != d->nblocks[1])
      return fail(TGH_EINVAL, "the rotating seam turns x into y about the axis x = y = lo: it needs "
                              "lo[0] = lo[1] and the same number of blocks along x and y");
  }
  if (d->background < 0 || d->background > 4) return fail(TGH_EINVAL, "unknown background");
  if (d->gauge < 0 || d->gauge > 2) return fail(TGH_EINVAL, "unknown gauge source kind");
  if (d->gauge == TGH_GAUGE_SAMPLED && d->background == TGH_BG_GAUGE_WAVE)
    return fail(TGH_EINVAL, "the gauge wave is a harmonic background (H_a = 0); a sampled source "
                            "is for static non-harmonic backgrounds only");
  if (d->gauge == TGH_GAUGE_KERR_SCHILD && d->background != TGH_BG_KERR_SCHILD)
    return fail(TGH_EINVAL, "TGH_GAUGE_KERR_SCHILD is the Kerr-Schild hole's own H_a; this "
                            "background is not Kerr-Schild");
  if (d->gauge == TGH_GAUGE_SAMPLED && d->background == TGH_BG_HARMONIC)
    return fail(TGH_EINVAL, "TGH_BG_HARMONIC needs TGH_GAUGE_NONE: its chart is harmonic, H_a "
                            "vanishes identically");
  if (d->gauge == TGH_GAUGE_NONE &&
      (d->background == TGH_BG_KERR_SCHILD || d->background == TGH_BG_SHIFTED_MINKOWSKI))
    return fail(TGH_EINVAL, "TGH_GAUGE_NONE on a non-harmonic background: its coordinates have "
                            "H_a = -Gamma_a != 0, so give it a sampled or closed-form source");
  if (d->interior < 0 || d->interior > 4) return fail(TGH_EINVAL, "unknown interior variant");
  if (d->shaped) {
    // FittedInterior's checks (src/interior.jl:1770-1816)
    if (d->interior == TGH_INTERIOR_NONE)
      return fail(TGH_EINVAL, "a tracked geometry (shaped = 1) belongs to an interior; this case has none");
    if (d->interior == TGH_INTERIOR_PASTED)
      return fail(TGH_EUNSUPPORTED, "the :pasted limiter on a tracked geometry is not built");
    if (d->shape_lmax < 0 || d->shape_lmax > TGH_SHAPE_MAX_LMAX)
      return fail(TGH_EINVAL, "shape_lmax = " + std::to_string(d->shape_lmax) + " is outside 0 .. 8");
    for (int i = 0; i < (d->shape_lmax + 1) * (d->shape_lmax + 1); ++i)
      if (!std::isfinite(d->shape[i])) return fail(TGH_EINVAL, "the shape's coefficients must be finite");
    if (!(0 < d->r_in && d->r_in <= d->r_out))
      return fail(TGH_EINVAL, "r_in = " + std::to_string(d->r_in) + ", r_out = " + std::to_string(d->r_out) +
                                  ": the clamp interval of the tracked surface needs 0 < r_in <= r_out");
    if (!(d->offset > 0 && d->thickness > 0))
      return fail(TGH_EINVAL, "offset and thickness are distances: both have to be positive");
    if (!((d->r_in - d->offset) - d->thickness > 0))
      return fail(TGH_EINVAL, "r_in - offset - thickness <= 0: the tracked core would reach the center");
    if (!(d->w_ramp > 0 && d->w_ramp <= 1 && d->rho_ramp > 0 && d->rho_ramp <= 1))
      return fail(TGH_EINVAL, "w_ramp and rho_ramp are fractions of the layer, in (0, 1]");
    if (!(d->rho_max >= 0)) return fail(TGH_EINVAL, "the relaxation rate must satisfy rho_max >= 0");
    if (d->thickness < 2.0 * (G + 1) * h0)
      return fail(TGH_EINVAL, "thickness < 2(G+1)h: an evolved point's stencil (G points) plus the "
                              "ramps would reach into the tracked core");
    if ((d->background == TGH_BG_KERR_SCHILD || d->background == TGH_BG_HARMONIC) &&
        !((d->r_in - d->offset) - d->thickness > std::fabs(d->a)))
      return fail(TGH_EINVAL, "r_in - offset - thickness <= |a|: the Kerr ring singularity (the disk "
                              "|x| <= |a| in z = 0) would lie outside the core");
    const double* v = d->velocity;
    if ((v[0] != 0.0 || v[1] != 0.0 || v[2] != 0.0) && d->interior != TGH_INTERIOR_FITTED)
      return fail(TGH_EUNSUPPORTED, "a moving tracked geometry is built for the :fitted variant only");
  } else if (d->interior != TGH_INTERIOR_NONE) {
    if (!(d->r0 > 0))
      return fail(TGH_EINVAL, "r0 = " + std::to_string(d->r0) + ": the core radius has to be positive");
    if (!(d->r1 > d->r0)) return fail(TGH_EINVAL, "the layer needs r_1 > r_0");
    if (!(d->w_ramp > 0 && d->w_ramp <= 1 && d->rho_ramp > 0 && d->rho_ramp <= 1))
      return fail(TGH_EINVAL, "w_ramp and rho_ramp are fractions of the layer, in (0, 1]");
    if (!(d->rho_max >= 0)) return fail(TGH_EINVAL, "the relaxation rate must satisfy rho_max >= 0");
    // check_interior_radii (src/interior.jl:1104-1150): no stencil of an
    // evolved point reaches the core.
    if (d->r1 - d->r0 < 2.0 * (G + 1) * h0)
      return fail(TGH_EINVAL, "r1 - r0 < 2(G+1)h: an evolved point's stencil plus the ramps would "
                              "reach into the core");
    // singular_radius (src/interior.jl:997-1010, checked at :1120-1126): both
    // Kerr charts are singular on the disk |x| <= |a|, z = 0.
    if ((d->background == TGH_BG_KERR_SCHILD || d->background == TGH_BG_HARMONIC) &&
        !(d->r0 > std::fabs(d->a)))
      return fail(TGH_EINVAL, "r0 <= |a|: the Kerr ring singularity (the disk |x| <= |a| in z = 0) "
                              "would lie outside the core");
  }
  {
    const double* v = d->velocity;
    if (!(std::isfinite(v[0]) && std::isfinite(v[1]) && std::isfinite(v[2])))
      return fail(TGH_EINVAL, "the hole's velocity must be finite");
    const double v2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    if (v2 != 0.0) {
      if (!(v2 < 1.0))
        return fail(TGH_EINVAL, "|velocity| = " + std::to_string(std::sqrt(v2)) + " >= 1");
      if (d->background != TGH_BG_KERR_SCHILD && d->background != TGH_BG_HARMONIC)
        return fail(TGH_EINVAL, "a velocity is defined for the Kerr-Schild and harmonic holes only");
      if (d->interior == TGH_INTERIOR_NONE)
        return fail(TGH_EINVAL, "a moving hole needs its interior (hole_case)");
      if (d->interior == TGH_INTERIOR_PASTED)
        return fail(TGH_EUNSUPPORTED, "the :pasted interior of a moving hole is not built");
      if (d->gauge == TGH_GAUGE_SAMPLED)
        return fail(TGH_EINVAL, "a moving hole's H_a changes with t and a source sampled at creation "
                                "would freeze it: use TGH_GAUGE_KERR_SCHILD");
      if (d->r_gate != 0)
        return fail(TGH_EUNSUPPORTED, "the range projection about a moving hole is not built");
    }
  }
  if (d->dissipation != TGH_DISSIPATION_CONSTANT && d->dissipation != TGH_DISSIPATION_HORIZON)
    return fail(TGH_EINVAL, "unknown dissipation kind");
  if (!(d->eps_ko >= 0))
    return fail(TGH_EINVAL, "eps_ko < 0 would amplify the grid-scale modes (the weights already "
                            "carry the damping sign)");
  if (d->dissipation == TGH_DISSIPATION_HORIZON) {
    // HorizonDissipation's checks (src/gauge.jl:1229-1252)
    if (!(d->eps_in >= 0))
      return fail(TGH_EINVAL, "eps_in = " + std::to_string(d->eps_in) + " is negative");
    if (!(d->eps_ko <= 4 && d->eps_in <= 4))
      return fail(TGH_EINVAL, "an amplitude above 4 puts the corner mode's rate 3 eps/h beyond RK4's "
                              "real-axis stability limit 2.79/dt at the stepper's CFL");
    if (!(0 < d->eps_r1 && d->eps_r1 < d->eps_rh))
      return fail(TGH_EINVAL, "eps_r1 = " + std::to_string(d->eps_r1) + ", eps_rh = " + std::to_string(d->eps_rh) +
                                  ": the dissipation ramp needs 0 < eps_r1 < eps_rh");
  }
  if (d->r_gate != 0) {
    // StateBounds (src/bounds.jl:94-135) and check_bounds_gate (:214-236)
    const double v[7] = {d->alpha_min, d->alpha_max, d->lambda_min, d->lambda_max, d->beta_max, d->K_max, d->r_gate};
    for (double x : v)
      if (!std::isfinite(x))
        return fail(TGH_EINVAL, "a range bound is not finite (no bounds at all is r_gate = 0)");
    if (!(0 < d->alpha_min && d->alpha_min < 1 && 1 < d->alpha_max))
      return fail(TGH_EINVAL, "alpha_min/alpha_max must bracket 1 and stay positive (the repair of "
                              "a non-finite point sets alpha = 1)");
    if (!(0 < d->lambda_min && d->lambda_min < 1 && 1 < d->lambda_max))
      return fail(TGH_EINVAL, "lambda_min/lambda_max must bracket 1 and stay positive");
    if (!(d->beta_max > 0)) return fail(TGH_EINVAL, "beta_max is not positive");
    if (!(d->K_max > 0)) return fail(TGH_EINVAL, "K_max is not positive");
    if (!(d->r_gate > 0))
      return fail(TGH_EINVAL, "r_gate < 0 (0 turns the bounds off)");
    if (d->interior == TGH_INTERIOR_NONE)
      return fail(TGH_EINVAL, "range bounds act inside r_gate about a hole's center, and this case "
                              "has no interior");
    const double reach = std::max((double)G, std::sqrt(2.0) * (d->q / 2));
    if (!(d->r_gate <= d->r1 - reach * h0))
      return fail(TGH_EINVAL, "r_gate = " + std::to_string(d->r_gate) +
                                  " but an evolved point (r >= r_1) reads points down to r_1 - R h = " +
                                  std::to_string(d->r1 - reach * h0) + " (check_bounds_gate)");
  }
  if (d->damping == TGH_DAMPING_GAUSSIAN && !(d->gamma0_width > 0))
    return fail(TGH_EINVAL, "gamma0_width is not positive");
  return TGH_OK;
}

template <int NP>
void launch_plan(dim3 grid, hipStream_t st, const PlanArgs& a) {
  hipLaunchKernelGGL((k_fill_plan<NP>), grid, dim3(256), 0, st, a);
}
template <int NP>
void launch_transfer(dim3 grid, hipStream_t st, const TransferArgs& a) {
  hipLaunchKernelGGL((k_transfer<NP>), grid, dim3(256), 0, st, a);
}
void launch_fill(tgh_problem* p, double* W, double t, double toff = 0.0) {
  if (p->forest) {
    if (reflects(p->geom)) {
      WallArgs wa{p->geom, W};
      hipLaunchKernelGGL(k_wall_work, grid_points(p), dim3(256), 0, p->stream, wa);
    }
    // the plan's passes, in order on the stream
    PlanArgs a;
    std::memset(&a, 0, sizeof(a));
    a.g = p->geom;
    a.c = p->cs;
    a.t = t;
    a.W = W;
    a.clock = p->clock_use;
    a.toff = toff;
    // pass 0's Dirichlet tail on its own kernel (it evaluates the background)
    if (p->plan_dir[1] > p->plan_dir[0]) {
      a.n = p->plan_dir[1] - p->plan_dir[0];
      a.ops = (const PlanOp*)p->d_plan + p->plan_dir[0];
      hipLaunchKernelGGL(p->cs.moving ? k_fill_plan_dirichlet<true> : k_fill_plan_dirichlet<false>,
                         dim3((a.n + 255) / 256), dim3(256), 0, p->stream, a);
    }
    for (size_t s = 0; s + 1 < p->plan_off.size(); ++s) {
      const int end = s == 0 ? p->plan_dir[0] : p->plan_off[s + 1];
      a.n = end - p->plan_off[s];
      if (a.n == 0) continue;
      a.ops = (const PlanOp*)p->d_plan + p->plan_off[s];
      const dim3 grid((a.n + 255) / 256, 20);
      switch (p->prolong) {
        case 2: launch_plan<2>(grid, p->stream, a); break;
        case 4: launch_plan<4>(grid, p->stream, a); break;
        case 6: launch_plan<6>(grid, p->stream, a); break;
        case 8: launch_plan<8>(grid, p->stream, a); break;
        default: launch_plan<10>(grid, p->stream, a); break;
      }
    }
    return;
  }
  FillArgs a{p->geom, p->cs, t, W, p->clock_use, toff};
  hipLaunchKernelGGL(p->cs.moving ? k_fill<true> : k_fill<false>, grid_work(p), dim3(256), 0, p->stream, a);
}
void launch_scatter(tgh_problem* p, const double* u, double* W, int to_work) {
  CopyArgs a{p->geom, to_work ? u : W, to_work ? W : (double*)u, to_work};
  hipLaunchKernelGGL(k_scatter, grid_points(p), dim3(256), 0, p->stream, a);
}

// One right-hand side on the block layout (the fused kernel, or the head and
// the Pi components as two launches).
void launch_rhs(tgh_problem* p, const RhsArgs& a, int epi) {
  switch (p->q) {
    case 2: launch_rhs_blocks<2>(p, a, epi); break;
    case 4: launch_rhs_blocks<4>(p, a, epi); break;
    case 6: launch_rhs_blocks<6>(p, a, epi); break;
    default: launch_rhs_blocks<8>(p, a, epi); break;
  }
}

RhsArgs rhs_args(tgh_problem* p, const double* W, double t, const tgh_matter* m) {
  RhsArgs a;
  std::memset(&a, 0, sizeof(a));
  a.g = p->geom;
  a.c = p->cs;
  a.t = t;
  a.W = W;
  a.Hsrc = p->Hsrc;
  a.coef = p->coef;
  if (m && m->rho) {
    a.m_rho = m->rho;
    a.m_S = m->S;
    a.m_Sij = m->Sij;
  }
  return a;
}

// The error rows from the reduction's partials, folded in workgroup order:
// (err_l2, err_linf, residual, drift).
void fold_error_rows(const double* red, size_t nwg, double out[4]) {
  double num = 0, den = 0, mx[3] = {0, 0, 0};
  for (size_t w = 0; w < nwg; ++w) {
    num += red[RED_STRIDE * w];
    den += red[RED_STRIDE * w + 1];
    for (int q = 0; q < 3; ++q) {
      const double v = red[RED_STRIDE * w + 2 + q];
      mx[q] = (v > mx[q] || std::isnan(v)) ? v : mx[q];
    }
  }
  out[0] = den > 0 ? std::sqrt(num / den) : 0.0;
  out[1] = mx[0];
  out[2] = mx[1];
  out[3] = mx[2];
}

int sync(tgh_problem* p) {
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(p->stream));
  return TGH_OK;
}

int reduce(tgh_problem* p, const double* u, double t, int mode, double shell_lo = 0.0, double shell_hi = -1.0) {
  ReduceArgs a{p->geom, p->cs, t, u, p->red, mode, shell_lo, shell_hi};
  hipLaunchKernelGGL(p->cs.moving ? k_reduce<true> : k_reduce<false>, grid_points(p), dim3(256), 0, p->stream, a);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(p->red_host.data(), p->red, sizeof(double) * p->red_host.size(),
                         hipMemcpyDeviceToHost, p->stream));
  HIP_TRY(hipStreamSynchronize(p->stream));
  return TGH_OK;
}

int next_event(tgh_problem* p) {
  if (p->ev_used == p->ev_pool.size()) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return -1;
    p->ev_pool.push_back(e);
  }
  return (int)p->ev_used++;
}
}  // namespace

int tgh::mark(tgh_problem* p) {
  int i = next_event(p);
  if (i >= 0) (void)hipEventRecord(p->ev_pool[i], p->stream);
  return i;
}

namespace {
// Fold the recorded event pairs into the running totals (after a sync).
void collect_profile(tgh_problem* p) {
  for (auto& pr : p->ev_rhs) {
    float ms = 0;
    if (pr.first >= 0 && pr.second >= 0 &&
        hipEventElapsedTime(&ms, p->ev_pool[pr.first], p->ev_pool[pr.second]) == hipSuccess) {
      p->prof[0] += ms;
      p->prof[1] += 1;
    }
  }
  for (auto& pr : p->ev_fill) {
    float ms = 0;
    if (pr.first >= 0 && pr.second >= 0 &&
        hipEventElapsedTime(&ms, p->ev_pool[pr.first], p->ev_pool[pr.second]) == hipSuccess) {
      p->prof[2] += ms;
      p->prof[3] += 1;
    }
  }
  for (auto& pr : p->ev_pi) {
    float ms = 0;
    if (pr.first >= 0 && pr.second >= 0 &&
        hipEventElapsedTime(&ms, p->ev_pool[pr.first], p->ev_pool[pr.second]) == hipSuccess) {
      p->prof[4] += ms;
      p->prof[5] += 1;
    }
  }
  p->ev_rhs.clear();
  p->ev_fill.clear();
  p->ev_pi.clear();
  p->ev_used = 0;
}

}  // namespace

dim3 tgh::grid_brick(const tgh_problem* p) {
  const Brick& k = p->brick;
  const long wpp = ((long)k.n[0] * k.n[1] + 255) / 256;
  return xcd_grid(wpp * k.n[2]);
}
dim3 tgh::grid_brick_tiles(const tgh_problem* p, int tx, int ty, int tz) {
  const Brick& k = p->brick;
  const long nt = (long)((k.n[0] + tx - 1) / tx) * ((k.n[1] + ty - 1) / ty) * ((k.n[2] + tz - 1) / tz);
  return xcd_grid(nt);
}

namespace {

// ---------------------------------------------------------------------------
// The brick stepper. Its default right-hand side form (measured on MI355X):
constexpr bool kBrickSplitDefault = true;
void launch_frame(tgh_problem* p, double* W, double t, double toff = 0.0) {
  const Brick& k = p->brick;
  const long E = (long)(k.n[0] + 2 * k.G) * (k.n[1] + 2 * k.G) * (k.n[2] + 2 * k.G);
  // without mirrors or the seam only the frame's points have work: one thread
  // each; with them the walls' inside points are projected too
  const int shell = reflects(p->geom) ? 0 : 1;
  const long n = shell ? E - (long)k.n[0] * k.n[1] * k.n[2] : E;
  FrameArgs a{k, p->cs, t, W, p->clock_use, toff, shell};
  hipLaunchKernelGGL(p->cs.moving ? k_brick_frame<true> : k_brick_frame<false>, dim3((unsigned)((n + 255) / 256)),
                     dim3(256), 0, p->stream, a);
}

// A background that does not depend on t (the gauge wave is the one that does).
bool static_background(const CaseP& c) {
  const int kind = c.bg.kind;
  return !c.moving && (kind == BG_MINKOWSKI || kind == BG_KERR_SCHILD || kind == BG_SHIFTED || kind == BG_HARMONIC);
}

// The layer box (k_brick_layer) of a moving :damped hole or of a harmonic hole
// at rest, on the finest level's grid: the layer's bounding box about c(t) with
// a point of margin each side, whatever t (slots wrap modulo B). Nothing for
// other cases.
Brick layer_grid(const tgh_problem* p);
int ensure_layer_box(tgh_problem* p) {
  if (p->Blayer || p->cs.general || p->cs.interior == INT_NONE) return TGH_OK;
  const bool moving = p->cs.moving && p->cs.interior == INT_DAMPED;
  // a harmonic hole at rest (no gauge source): its layer's target is a dual
  // pass of the harmonic chart, the same at every t, so the box is filled once
  const bool harmonic = !p->cs.moving && p->cs.gauge == GAUGE_NONE;
  if (!moving && !harmonic) return TGH_OK;
  int lmax = 0;
  for (int l : p->level_host) lmax = std::max(lmax, l);
  p->layer_lmax = lmax;
  p->layer_B = 2 * (int)std::ceil(p->cs.r1 * p->geom.inv_h * (double)(1 << lmax)) + 4;
  const size_t lb = sizeof(double) * LAYER_FIELDS * (size_t)p->layer_B * p->layer_B * p->layer_B;
  hipError_t e = hipMalloc(&p->Blayer, lb);
  if (e != hipSuccess) {
    p->Blayer = nullptr;
    return fail(TGH_EDEVICE, std::string("allocating the layer box: ") + hipGetErrorString(e));
  }
  p->layer_static = harmonic;
  if (harmonic) {
    const long B3 = (long)p->layer_B * p->layer_B * p->layer_B;
    LayerArgs la{layer_grid(p), p->cs, 0.0, nullptr, 0.0, p->Blayer, p->layer_B};
    hipLaunchKernelGGL(k_brick_layer<false>, dim3((unsigned)((B3 + 255) / 256), 1), dim3(256), 0, p->stream, la);
    HIP_TRY(hipGetLastError());
  }
  return TGH_OK;
}
// The box kernel's view of the finest level's grid (its corner, spacing and
// closed point counts; the other fields are unused there).
Brick layer_grid(const tgh_problem* p) {
  Brick k;
  std::memset(&k, 0, sizeof(k));
  const double f = (double)(1 << p->layer_lmax);
  for (int d = 0; d < 3; ++d) {
    k.lo[d] = p->geom.lo[d];
    k.n[d] = (p->geom.nb[d] << p->layer_lmax) * p->geom.N + 1;
  }
  k.h = p->geom.h / f;
  k.inv_h = p->geom.inv_h * f;
  return k;
}
void launch_layer_box(tgh_problem* p, double t, double toff, RhsArgs& a) {
  const long B3 = (long)p->layer_B * p->layer_B * p->layer_B;
  if (!p->layer_static) {
    LayerArgs la{layer_grid(p), p->cs, t, p->clock_use, toff, p->Blayer, p->layer_B};
    hipLaunchKernelGGL(k_brick_layer<true>, dim3((unsigned)((B3 + 255) / 256), 3), dim3(256), 0, p->stream, la);
  }
  a.lbox = p->Blayer;
  a.lB3 = B3;
  a.lB = p->layer_B;
  a.lLmax = p->layer_lmax;
}

int ensure_brick(tgh_problem* p) {
  if (p->brick_ready) return TGH_OK;
  const Geom& g = p->geom;
  Brick& k = p->brick;
  std::memset(&k, 0, sizeof(k));
  k.G = g.G;
  k.XO = 16;  // x = 0 starts a 128-byte line; G <= 5 ghosts sit before it
  for (int d = 0; d < 3; ++d) {
    k.tot[d] = g.nb[d] * g.N;
    k.periodic[d] = g.periodic[d];
    k.refl[d][0] = g.refl[d][0];
    k.refl[d][1] = g.refl[d][1];
    k.rot = g.rot;
    k.n[d] = k.tot[d] + (g.periodic[d] ? 0 : 1);
    k.lo[d] = g.lo[d];
  }
  k.h = g.h;
  k.inv_h = g.inv_h;
  k.py = ((long)k.XO + k.n[0] + k.G + 15) / 16 * 16;
  k.pz = k.py * (k.n[1] + 2 * k.G);
  k.fs = k.pz * (k.n[2] + 2 * k.G);
  const size_t bytes = sizeof(double) * 20 * (size_t)k.fs;
  double** arrs[4] = {&p->BW[0], &p->BW[1], &p->Bus, &p->Bacc};
  for (double** a : arrs) {
    hipError_t e = hipMalloc(a, bytes);
    if (e != hipSuccess)
      return fail(TGH_EDEVICE, std::string("allocating the stepper's brick arrays (") + std::to_string(bytes >> 20) +
                                   " MiB each): " + hipGetErrorString(e));
    HIP_TRY(hipMemsetAsync(*a, 0, bytes, p->stream));
  }
  const bool static_bg = static_background(p->cs);
  // a wrapped or mirrored frame reads the stage's own values: refilled every stage
  p->frame_static = static_bg && !g.periodic[0] && !g.periodic[1] && !g.periodic[2] && !reflects(g);
  if (p->frame_static) {
    launch_frame(p, p->BW[0], 0.0);
    launch_frame(p, p->BW[1], 0.0);
  }
  {
    const int rc = ensure_layer_box(p);
    if (rc) return rc;
  }
  if (p->cs.gauge == GAUGE_SAMPLED) {
    HIP_TRY(hipMalloc(&p->BH, bytes));
    HIP_TRY(hipMemsetAsync(p->BH, 0, bytes, p->stream));
    BrickSampleArgs sa{k, p->cs, 0.0, p->BH};
    hipLaunchKernelGGL(k_brick_sample, grid_brick(p), dim3(256), 0, p->stream, sa);
  }
  int rc = sync(p);
  if (rc) return rc;
  p->brick_ready = true;
  return TGH_OK;
}

void launch_bounds_brick(tgh_problem* p, double* A, double* B);
void launch_paste(tgh_problem* p, int mode, double* A, double* B, double t, double toff = 0.0);

// One RK4 step on the brick arrays (u_n in Bus, the stage input in BW[0]).
void rk4_brick_launches(tgh_problem* p, double t, double dt) {
  const double c_acc[4] = {dt / 6, dt / 3, dt / 3, dt / 6};
  const double c_next[4] = {dt / 2, dt / 2, dt, 0.0};
  const double t_off[4] = {0.0, dt / 2, dt / 2, dt};
  const double t_stage[4] = {t, t + dt / 2, t + dt / 2, t + dt};
  for (int s = 0; s < 4; ++s) {
    double* Wc = p->BW[s % 2];
    double* Wn = p->BW[(s + 1) % 2];
    if (!p->frame_static) {
      const bool prof = p->profiling;
      int e0 = prof ? mark(p) : -1;
      launch_frame(p, Wc, t_stage[s], t_off[s]);
      if (prof) p->ev_fill.push_back({e0, mark(p)});
    }
    RhsArgs a = rhs_args(p, Wc, t_stage[s], nullptr);
    a.clock = p->clock_use;
    a.toff = t_off[s];
    if (p->Blayer) launch_layer_box(p, t_stage[s], t_off[s], a);
    a.Hsrc = p->BH;
    a.us = p->Bus;
    a.acc = p->Bacc;
    a.Wn = Wn;
    a.cacc = c_acc[s];
    a.cnext = c_next[s];
    a.stage = s + 1;
    switch (p->q) {
      case 2: launch_rhs_brick<2>(p, a); break;
      case 4: launch_rhs_brick<4>(p, a); break;
      case 6: launch_rhs_brick<6>(p, a); break;
      default: launch_rhs_brick<8>(p, a); break;
    }
    // the stage limiter on the stage value (and on the step's result): the
    // range projection, then the :pasted overwrite
    if (p->has_bounds) launch_bounds_brick(p, s == 3 ? p->Bus : Wn, s == 3 ? Wn : nullptr);
    if (p->cs.interior == INT_PASTED)
      launch_paste(p, 1, s == 3 ? p->Bus : Wn, s == 3 ? Wn : nullptr, t_stage[s < 3 ? s + 1 : 3],
                   t_off[s < 3 ? s + 1 : 3]);
  }
}

// The range projection on the state (mode 0, with accounting partials), on
// a block work array's owned points (+ the state), or on the brick points
// inside the gate's bounding box (+ a second brick array).
void launch_bounds_state(tgh_problem* p, double* u) {
  BoundsArgs a{};
  a.g = p->geom;
  a.c = p->cs;
  a.bd = p->bd;
  a.dst0 = u;
  a.out = p->red;
  a.mode = 0;
  hipLaunchKernelGGL(k_bounds, grid_points(p), dim3(256), 0, p->stream, a);
}
void launch_bounds_work(tgh_problem* p, double* W, double* us) {
  BoundsArgs a{};
  a.g = p->geom;
  a.c = p->cs;
  a.bd = p->bd;
  a.dst0 = W;
  a.dst1 = us;
  a.mode = 1;
  hipLaunchKernelGGL(k_bounds, grid_points(p), dim3(256), 0, p->stream, a);
}
void launch_bounds_brick(tgh_problem* p, double* A, double* B) {
  const Brick& k = p->brick;
  BoundsArgs a{};
  a.g = p->geom;
  a.k = k;
  a.c = p->cs;
  a.bd = p->bd;
  a.dst0 = A;
  a.dst1 = B;
  a.mode = 2;
  long n = 1;
  for (int d = 0; d < 3; ++d) {  // the gate's bounding box, one point of margin each side
    const int i0 = (int)std::floor((p->cs.center[d] - p->bd.r_gate - k.lo[d]) * k.inv_h) - 1;
    const int i1 = (int)std::ceil((p->cs.center[d] + p->bd.r_gate - k.lo[d]) * k.inv_h) + 1;
    a.lo[d] = std::max(0, i0);
    a.ext[d] = std::max(0, std::min(k.n[d] - 1, i1) - a.lo[d] + 1);
    n *= a.ext[d];
  }
  if (n > 0) hipLaunchKernelGGL(k_bounds, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, p->stream, a);
}

// The :pasted limiter on a stage value at time t (mode 0: a work array's
// owned points and optionally the state; mode 1: brick arrays).
void launch_paste(tgh_problem* p, int mode, double* A, double* B, double t, double toff) {
  PasteArgs a{};
  a.g = p->geom;
  a.k = p->brick;
  a.c = p->cs;
  a.t = t;
  a.clock = p->clock_use;
  a.toff = toff;
  a.dst0 = A;
  a.dst1 = B;
  a.mode = mode;
  const dim3 grid = mode == 1 ? grid_brick(p) : grid_points(p);
  hipLaunchKernelGGL(p->cs.moving ? k_paste<true> : k_paste<false>, grid, dim3(256), 0, p->stream, a);
}

// One RK4 step on the block layout run as bricks (form 5): u_n in Wus, the
// accumulator in Wacc, the coefficients in Wcoef (all in the work layout), the
// stage input in W[0] holding u_n with its ghosts unfilled.
void rk4_block_lds_launches(tgh_problem* p, double t, double dt) {
  const double c_acc[4] = {dt / 6, dt / 3, dt / 3, dt / 6};
  const double c_next[4] = {dt / 2, dt / 2, dt, 0.0};
  const double t_off[4] = {0.0, dt / 2, dt / 2, dt};
  const double t_stage[4] = {t, t + dt / 2, t + dt / 2, t + dt};
  for (int s = 0; s < 4; ++s) {
    double* Wc = p->W[s % 2];
    double* Wn = p->W[(s + 1) % 2];
    const bool prof = p->profiling;
    int e0 = prof ? mark(p) : -1;
    launch_fill(p, Wc, t_stage[s], t_off[s]);
    if (prof) p->ev_fill.push_back({e0, mark(p)});
    RhsArgs a = rhs_args(p, Wc, t_stage[s], nullptr);
    a.clock = p->clock_use;
    a.toff = t_off[s];
    if (p->Blayer) launch_layer_box(p, t_stage[s], t_off[s], a);
    a.cacc = c_acc[s];
    a.cnext = c_next[s];
    a.stage = s + 1;
    const BlockBricks bb{Wc, p->Wcoef, p->Wus, p->Wacc, Wn, p->layer_lmax};
    switch (p->q) {
      case 2: launch_rhs_blocks_lds<2>(p, a, bb); break;
      case 4: launch_rhs_blocks_lds<4>(p, a, bb); break;
      case 6: launch_rhs_blocks_lds<6>(p, a, bb); break;
      default: launch_rhs_blocks_lds<8>(p, a, bb); break;
    }
  }
}

// One RK4 step on device arrays: u (state, in place), stage work arrays
// W[0] (holding u_n with ghosts unfilled) and W[1].
void rk4_step_launches(tgh_problem* p, double* u, double t, double dt) {
  const double c_acc[4] = {dt / 6, dt / 3, dt / 3, dt / 6};
  const double c_next[4] = {dt / 2, dt / 2, dt, 0.0};
  const double t_off[4] = {0.0, dt / 2, dt / 2, dt};
  const double t_stage[4] = {t, t + dt / 2, t + dt / 2, t + dt};
  for (int s = 0; s < 4; ++s) {
    double* Wc = p->W[s % 2];
    double* Wn = p->W[(s + 1) % 2];
    const bool prof = p->profiling;
    int e0 = prof ? mark(p) : -1;
    launch_fill(p, Wc, t_stage[s], t_off[s]);
    if (prof) p->ev_fill.push_back({e0, mark(p)});
    RhsArgs a = rhs_args(p, Wc, t_stage[s], p->mat.rho ? &p->mat : nullptr);
    a.clock = p->clock_use;
    a.toff = t_off[s];
    if (p->Blayer) launch_layer_box(p, t_stage[s], t_off[s], a);
    a.us = u;
    a.acc = p->acc;
    a.Wn = Wn;
    a.cacc = c_acc[s];
    a.cnext = c_next[s];
    a.stage = s + 1;
    launch_rhs(p, a, EPI_RK4);
    // the stage limiter on the stage value (and on the step's result): the
    // range projection, then the :pasted overwrite
    if (p->has_bounds) launch_bounds_work(p, Wn, s == 3 ? u : nullptr);
    if (p->cs.interior == INT_PASTED)
      launch_paste(p, 0, Wn, s == 3 ? u : nullptr, t_stage[s < 3 ? s + 1 : 3], t_off[s < 3 ? s + 1 : 3]);
  }
}

}  // namespace

extern "C" {

const char* tgh_last_error(void) { return g_last_error.c_str(); }

// The problem on a uniform forest (leaves NULL) or on a forest with levels.
static int create_problem(const tgh_desc* d, const int32_t* leaves, int32_t nleaves, int32_t prolong,
                          tgh_problem** out) {
  if (!out) return fail(TGH_EINVAL, "out is NULL");
  *out = nullptr;
  ForestPlan plan;
  std::vector<int> lev, fbc;
  int lmax = 0;
  if (leaves) {
    for (int32_t n = 0; n < nleaves; ++n) {
      lev.push_back(leaves[4 * n]);
      for (int c = 0; c < 3; ++c) fbc.push_back(leaves[4 * n + 1 + c]);
      lmax = std::max(lmax, std::min(std::max(leaves[4 * n], 0), 10));
    }
  }
  int rc = check_desc(d, lmax);
  if (rc) return rc;
  const bool mirrors = d->reflecting[0][0] || d->reflecting[0][1] || d->reflecting[1][0] ||
                       d->reflecting[1][1] || d->reflecting[2][0] || d->reflecting[2][1] || d->rotating;
  if (mirrors && d->ownership == TGH_OWNERSHIP_HALF_OPEN)
    return fail(TGH_EUNSUPPORTED, "reflecting faces are built for closed blocks, whose wall plane is owned");
  if (!leaves && d->ownership == TGH_OWNERSHIP_HALF_OPEN)
    return fail(TGH_EUNSUPPORTED, "half-open ownership of block faces is built for forests: create the "
                                  "problem with tgh_problem_create_forest (a uniform brick is the forest "
                                  "of its level-0 leaves)");
  if (leaves) {
    const int per[3] = {d->periodic[0] ? 1 : 0, d->periodic[1] ? 1 : 0, d->periodic[2] ? 1 : 0};
    std::string msg;
    try {
      msg = forest_plan(d->N, d->q / 2 + 1, d->nblocks, per, prolong, lev, fbc, plan,
                        d->ownership == TGH_OWNERSHIP_HALF_OPEN, d->reflecting, d->rotating);
    } catch (const std::exception& ex) {  // host threads or memory
      return fail(TGH_EDEVICE, std::string("building the fill plan: ") + ex.what());
    }
    if (!msg.empty()) return fail(TGH_EINVAL, msg);
  }
  int device = 0;
  HIP_TRY(hipGetDevice(&device));
  auto* p = new tgh_problem();
  p->device = device;
  p->desc = *d;
  p->q = d->q;
  Geom& g = p->geom;
  g.N = d->N;
  g.G = d->q / 2 + 1;
  g.S = d->N + 1 + 2 * g.G;
  g.P = d->N + 1;
  for (int i = 0; i < 3; ++i) {
    g.nb[i] = d->nblocks[i];
    g.periodic[i] = d->periodic[i] ? 1 : 0;
    g.refl[i][0] = d->reflecting[i][0] ? 1 : 0;
    g.refl[i][1] = d->reflecting[i][1] ? 1 : 0;
    g.rot = d->rotating ? 1 : 0;
    g.lo[i] = d->lo[i];
  }
  g.nblocks = leaves ? nleaves : d->nblocks[0] * d->nblocks[1] * d->nblocks[2];
  g.h = (d->hi[0] - d->lo[0]) / ((double)d->nblocks[0] * d->N);
  g.inv_h = 1.0 / g.h;
  std::vector<int> lookup;
  if (leaves) {
    // a forest: the leaves in the caller's order
    p->forest = true;
    p->bcoord_host = fbc;
    p->level_host = lev;
    p->prolong = prolong;
    p->plan_off = plan.pass_off;
    p->plan_dir[0] = plan.dir_begin;
    p->plan_dir[1] = plan.dir_end;
    lookup.assign(1, 0);
  } else {
    // Blocks in Morton order of the brick.
    std::vector<std::array<int, 3>> bl;
    for (int k = 0; k < g.nb[2]; ++k)
      for (int j = 0; j < g.nb[1]; ++j)
        for (int i = 0; i < g.nb[0]; ++i) bl.push_back({i, j, k});
    std::stable_sort(bl.begin(), bl.end(), [](const auto& a, const auto& b) {
      return morton3(a[0], a[1], a[2]) < morton3(b[0], b[1], b[2]);
    });
    p->bcoord_host.resize(3 * bl.size());
    lookup.assign((size_t)g.nblocks, 0);
    for (size_t b = 0; b < bl.size(); ++b) {
      for (int d2 = 0; d2 < 3; ++d2) p->bcoord_host[3 * b + d2] = bl[b][d2];
      lookup[bl[b][0] + g.nb[0] * (bl[b][1] + g.nb[1] * bl[b][2])] = (int)b;
    }
  }
  CaseP& c = p->cs;
  std::memset(&c, 0, sizeof(c));
  c.bg.kind = d->background;
  c.bg.M = d->M;
  c.bg.a = d->a;
  for (int i = 0; i < 3; ++i) c.bg.c[i] = d->center[i];
  c.bg.A = d->wave_amp;
  c.bg.len = d->wave_len;
  c.gauge = d->gauge;
  c.damping = d->damping;
  c.interior = d->interior;
  c.diss = d->dissipation;
  c.eps_in = d->eps_in;
  c.eps_r1 = d->eps_r1;
  c.eps_rh = d->eps_rh;
  if (d->r_gate > 0) {
    p->has_bounds = true;
    p->bd = Bounds{d->alpha_min, d->alpha_max, d->lambda_min, d->lambda_max, d->beta_max, d->K_max, d->r_gate};
  }
  c.has_diss = d->dissipation == TGH_DISSIPATION_CONSTANT ? d->eps_ko != 0.0
                                                           : !(d->eps_ko == 0.0 && d->eps_in == 0.0);
  c.g0 = d->gamma0;
  c.g0_near = d->gamma0_near;
  c.g0_far = d->gamma0_far;
  c.g0_width = d->gamma0_width;
  c.g2 = d->gamma2;
  c.eps_ko = d->eps_ko;
  for (int i = 0; i < 3; ++i) c.center[i] = d->center[i];
  c.r0 = d->r0;
  c.r1 = d->r1;
  c.rho_max = d->rho_max;
  c.w_ramp = d->w_ramp;
  c.rho_ramp = d->rho_ramp;
  c.shaped = d->shaped ? 1 : 0;
  c.lmax = d->shape_lmax;
  c.shape = nullptr;
  c.r_in = d->r_in;
  c.r_out = d->r_out;
  c.offset = d->offset;
  c.thickness = d->thickness;
  c.tw = nullptr;
  c.t_f = 0.0;
  c.trail = 0.0;
  c.rate = 0;
  c.general = d->interior != TGH_INTERIOR_NONE && (d->shaped || d->interior == TGH_INTERIOR_FITTED);
  {
    // the moving hole: translate(boost(bg, -v), center), distances from
    // center + v t, the core's and the layer's rates (src/interior.jl:390-549)
    const double* v = d->velocity;
    c.moving = v[0] != 0.0 || v[1] != 0.0 || v[2] != 0.0;
    for (int i = 0; i < 3; ++i) {
      c.vel[i] = v[i];
      c.bg.u[i] = c.moving ? -v[i] : 0.0;
    }
    c.delta = std::cbrt(DBL_EPSILON) * d->r0;  // layer_rate_step (src/interior.jl:524-537)
    // KerrSchildSource(T; M, spin = (0, 0, a), velocity = v) (src/gauge.jl:263-278)
    const double v2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    const double gam = 1.0 / std::sqrt(1.0 - v2);
    const double av[3] = {0.0, 0.0, d->a};
    const double va = v[0] * av[0] + v[1] * av[1] + v[2] * av[2];
    c.ks_S[0] = gam * va;
    for (int i = 0; i < 3; ++i) c.ks_S[1 + i] = v2 == 0.0 ? av[i] : av[i] + ((gam - 1.0) * va / v2) * v[i];
    c.ks_u[0] = gam;
    for (int i = 0; i < 3; ++i) c.ks_u[1 + i] = gam * v[i];
    if (c.moving) c.gauge = c.gauge == GAUGE_KS ? GAUGE_KS_MOVING : GAUGE_NONE_MOVING;
  }

  const long P3 = (long)g.P * g.P * g.P, S3 = (long)g.S * g.S * g.S;
  p->state_len = P3 * 20 * g.nblocks;
  p->work_len = S3 * 20 * g.nblocks;
  const long wg_per_block = (P3 + 255) / 256;
  p->red_host.assign(RED_STRIDE * (size_t)g.nblocks * wg_per_block, 0.0);
  auto cleanup = [&](int code, const std::string& m) {
    tgh_problem_destroy(p);
    return fail(code, m);
  };
  hipError_t e;
  if ((e = hipStreamCreate(&p->stream)) != hipSuccess) return cleanup(TGH_EDEVICE, hipGetErrorString(e));
  p->own_stream = true;
  if ((e = hipMalloc(&p->d_bcoord, sizeof(int) * p->bcoord_host.size())) != hipSuccess ||
      (e = hipMalloc(&p->d_lookup, sizeof(int) * lookup.size())) != hipSuccess ||
      (e = hipMalloc(&p->W[0], sizeof(double) * p->work_len)) != hipSuccess ||
      (e = hipMalloc(&p->W[1], sizeof(double) * p->work_len)) != hipSuccess ||
      (e = hipMalloc(&p->acc, sizeof(double) * p->state_len)) != hipSuccess ||
      (e = hipMalloc(&p->coef, sizeof(double) * (p->state_len / 20) * CF_N)) != hipSuccess ||
      (e = hipMalloc(&p->red, sizeof(double) * p->red_host.size())) != hipSuccess)
    return cleanup(TGH_EDEVICE, std::string("allocating the problem's device arrays: ") + hipGetErrorString(e));
  if ((e = hipMemcpy(p->d_bcoord, p->bcoord_host.data(), sizeof(int) * p->bcoord_host.size(),
                     hipMemcpyHostToDevice)) != hipSuccess ||
      (e = hipMemcpy(p->d_lookup, lookup.data(), sizeof(int) * lookup.size(), hipMemcpyHostToDevice)) !=
          hipSuccess)
    return cleanup(TGH_EDEVICE, hipGetErrorString(e));
  // The ghosts of both work arrays start defined (a fill writes every one of them
  // before any kernel reads it; this only keeps inspection deterministic).
  if ((e = hipMemset(p->W[0], 0, sizeof(double) * p->work_len)) != hipSuccess ||
      (e = hipMemset(p->W[1], 0, sizeof(double) * p->work_len)) != hipSuccess)
    return cleanup(TGH_EDEVICE, hipGetErrorString(e));
  if (c.shaped) {
    if ((e = hipMalloc(&p->d_shape, sizeof(d->shape))) != hipSuccess ||
        (e = hipMemcpy(p->d_shape, d->shape, sizeof(d->shape), hipMemcpyHostToDevice)) != hipSuccess)
      return cleanup(TGH_EDEVICE, std::string("uploading the shape: ") + hipGetErrorString(e));
    c.shape = p->d_shape;
  }
  g.bcoord = p->d_bcoord;
  g.lookup = p->d_lookup;
  g.blevel = nullptr;
  g.prim = nullptr;
  if (p->forest) {
    if ((e = hipMalloc(&p->d_level, sizeof(int) * p->level_host.size())) != hipSuccess ||
        (e = hipMalloc(&p->d_prim, plan.prim.size())) != hipSuccess ||
        (e = hipMalloc(&p->d_plan, sizeof(int) * std::max<size_t>(plan.ops.size(), 4))) != hipSuccess)
      return cleanup(TGH_EDEVICE, std::string("allocating the forest's tables: ") + hipGetErrorString(e));
    if ((e = hipMemcpy(p->d_level, p->level_host.data(), sizeof(int) * p->level_host.size(),
                       hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(p->d_prim, plan.prim.data(), plan.prim.size(), hipMemcpyHostToDevice)) != hipSuccess ||
        (!plan.ops.empty() && (e = hipMemcpy(p->d_plan, plan.ops.data(), sizeof(int) * plan.ops.size(),
                                             hipMemcpyHostToDevice)) != hipSuccess))
      return cleanup(TGH_EDEVICE, hipGetErrorString(e));
    g.blevel = p->d_level;
    g.prim = p->d_prim;
  }
  if (c.gauge == GAUGE_SAMPLED) {
    // The sample is taken once, here, at t = 0: the background is static (a
    // moving non-harmonic background is refused, src/initialdata.jl GHCase).
    if ((e = hipMalloc(&p->Hsrc, sizeof(double) * p->state_len)) != hipSuccess)
      return cleanup(TGH_EDEVICE, hipGetErrorString(e));
    SampleArgs sa{g, c, 0.0, p->Hsrc};
    hipLaunchKernelGGL(k_sample_gauge, grid_points(p), dim3(256), 0, p->stream, sa);
    if ((e = hipGetLastError()) != hipSuccess || (e = hipStreamSynchronize(p->stream)) != hipSuccess)
      return cleanup(TGH_EDEVICE, std::string("sampling the gauge source: ") + hipGetErrorString(e));
  }
  *out = p;
  return TGH_OK;
}

int tgh_problem_create(const tgh_desc* d, tgh_problem** out) { return create_problem(d, nullptr, 0, 0, out); }

int tgh_problem_create_forest(const tgh_desc* d, const int32_t* leaves, int32_t nleaves, int32_t prolongation,
                              tgh_problem** out) {
  if (!out) return fail(TGH_EINVAL, "out is NULL");
  *out = nullptr;
  if (!leaves || nleaves < 1) return fail(TGH_EINVAL, "a forest needs a leaf list of at least one leaf");
  return create_problem(d, leaves, nleaves, prolongation, out);
}

int tgh_set_target(tgh_problem* p, const double* target, double t_target, int32_t rate, double trail) {
  if (!p) return fail(TGH_EINVAL, "NULL problem");
  if (!std::isfinite(t_target) || !std::isfinite(trail))
    return fail(TGH_EINVAL, "the target's time and the trail factor must be finite");
  if (!(trail >= 0 && trail <= 1))
    return fail(TGH_EINVAL, "the trail factor narrows the ramp by 1 - trail zeta and must lie in [0, 1]");
  HIP_TRY(hipStreamSynchronize(p->stream));
  p->cs.tw = target;
  p->cs.t_f = t_target;
  p->cs.rate = rate != 0;
  p->cs.trail = trail;
  ++p->version;
  if (p->graph_exec) {  // a captured graph holds the old arguments
    HIP_TRY(hipGraphExecDestroy(p->graph_exec));
    p->graph_exec = nullptr;
  }
  return TGH_OK;
}

namespace {
// A tgh_fit checked and turned into the kernel's view, or an error message.
std::string fit_dev(const tgh_fit* f, const char* which, FitDev& o) {
  const std::string w(which);
  if (!f->coeffs) return w + " fit has no coefficients";
  if (f->lmax < 0 || f->lmax > 32) return w + " fit's degree lmax = " + std::to_string(f->lmax) + " is outside 0..32";
  if (f->cont < 0 || f->cont > 2)
    return w + " fit's radial order cont = " + std::to_string(f->cont) + " is not 0, 1 or 2";
  const double nums[] = {f->rbar,      f->center[0],   f->center[1],   f->center[2], f->velocity[0],
                         f->velocity[1], f->velocity[2], f->alpha_min, f->alpha_max,   f->lambda_min,
                         f->lambda_max, f->beta_max,    f->K_max,      f->t};
  for (double x : nums)
    if (!std::isfinite(x)) return w + " fit holds a non-finite number";
  if (!(f->rbar > 0)) return w + " fit's scale radius rbar must be positive";
  if (!(0 < f->alpha_min && f->alpha_min < 1 && 1 < f->alpha_max) ||
      !(0 < f->lambda_min && f->lambda_min < 1 && 1 < f->lambda_max) || !(f->beta_max > 0) || !(f->K_max > 0))
    return w + " fit's bounds need 0 < alpha_min < 1 < alpha_max, 0 < lambda_min < 1 < lambda_max and "
               "positive beta_max and K_max";
  o.lmax = f->lmax;
  o.cont = f->cont;
  o.tilde = f->tilde != 0;
  o.rbar = f->rbar;
  for (int d = 0; d < 3; ++d) {
    o.c0[d] = f->center[d];
    o.v[d] = f->velocity[d];
  }
  o.bd = Bounds{f->alpha_min, f->alpha_max, f->lambda_min, f->lambda_max, f->beta_max, f->K_max, 0.0};
  o.coeffs = f->coeffs;
  return "";
}
}  // namespace

int tgh_fill_target(tgh_problem* p, double* target, const tgh_fit* latest, const tgh_fit* previous, double t,
                    double h, int32_t rate, int32_t clear) {
  if (!p || !target || !latest) return fail(TGH_EINVAL, "NULL problem, target or latest fit");
  if (!p->desc.shaped)
    return fail(TGH_EINVAL, "fill_target reads the tracked geometry (FittedInterior): this problem's interior "
                            "is not shaped");
  if (!std::isfinite(t) || !(h >= 0) || !std::isfinite(h))
    return fail(TGH_EINVAL, "the fill time must be finite and the geometry's spacing h finite and >= 0");
  FitTargetArgs a;
  std::memset(&a, 0, sizeof(a));
  std::string msg = fit_dev(latest, "the latest", a.fa);
  if (msg.empty() && previous) msg = fit_dev(previous, "the previous", a.fb);
  if (!msg.empty()) return fail(TGH_EINVAL, msg);
  // fill_target! (src/fit.jl:1131-1151): the previous fit counts when it is older
  a.hasb = previous && latest->t > previous->t;
  if (!a.hasb) a.fb = a.fa;
  a.kappa = a.hasb ? 1.0 / (latest->t - previous->t) : 0.0;
  const double* v = latest->velocity;
  const double speed = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  if (rate && speed > 0 && !(h > 0))
    return fail(TGH_EINVAL, "a moving fit's rate is a difference over h / (8 |v|): it needs the geometry's "
                            "spacing h > 0");
  a.dt = speed > 0 ? h / (8.0 * speed) : 1.0;
  a.g = p->geom;
  for (int d = 0; d < 3; ++d) {
    a.ic[d] = p->desc.center[d];
    a.iv[d] = p->desc.velocity[d];
  }
  a.rate = rate != 0;
  a.clear = clear != 0;
  a.t_a = latest->t;
  a.t_f = t;
  a.r_fill = (p->desc.r_out - p->desc.offset) + h;
  a.out = target;
  hipLaunchKernelGGL(k_fit_target, grid_points(p), dim3(256), 0, p->stream, a);
  HIP_TRY(hipGetLastError());
  return sync(p);
}

int tgh_set_matter(tgh_problem* p, const tgh_matter* m) {
  if (!p) return fail(TGH_EINVAL, "NULL problem");
  if (m && m->rho && (!m->S || !m->Sij)) return fail(TGH_EINVAL, "matter needs rho, S and Sij together");
  HIP_TRY(hipStreamSynchronize(p->stream));
  p->mat = (m && m->rho) ? *m : tgh_matter{nullptr, nullptr, nullptr};
  ++p->version;
  if (p->graph_exec) {  // a captured graph holds the old arguments
    HIP_TRY(hipGraphExecDestroy(p->graph_exec));
    p->graph_exec = nullptr;
  }
  return TGH_OK;
}

int tgh_set_rho_max(tgh_problem* p, double rho_max) {
  if (!p) return fail(TGH_EINVAL, "NULL problem");
  if (p->desc.interior == TGH_INTERIOR_NONE) return fail(TGH_EINVAL, "this problem's case has no interior");
  if (!(rho_max >= 0) || !std::isfinite(rho_max))
    return fail(TGH_EINVAL, "the relaxation rate must be finite and satisfy rho_max >= 0");
  HIP_TRY(hipStreamSynchronize(p->stream));
  p->desc.rho_max = rho_max;
  p->cs.rho_max = rho_max;
  ++p->version;
  if (p->graph_exec) {  // a captured graph holds the old arguments
    HIP_TRY(hipGraphExecDestroy(p->graph_exec));
    p->graph_exec = nullptr;
  }
  return TGH_OK;
}

int tgh_transfer(tgh_problem* from, const double* u_from, double t, tgh_problem* to, double* u_to) {
  if (!from || !to || !u_from || !u_to) return fail(TGH_EINVAL, "NULL argument");
  if (!std::isfinite(t)) return fail(TGH_EINVAL, "the transfer's time must be finite");
  const tgh_desc &a = from->desc, &b = to->desc;
  bool same = a.N == b.N;
  for (int d = 0; d < 3; ++d)
    same = same && a.nblocks[d] == b.nblocks[d] && (a.periodic[d] != 0) == (b.periodic[d] != 0) &&
           a.lo[d] == b.lo[d] && a.hi[d] == b.hi[d];
  if (!same)
    return fail(TGH_EINVAL, "the transfer needs two forests of the same root brick: N, nblocks, periodic, lo and hi");
  if (reflects(from->geom) || reflects(to->geom))
    return fail(TGH_EUNSUPPORTED, "the transfer's plan does not mirror: regrid is built without reflecting faces");
  // a uniform problem is the forest of its level-0 blocks; the interpolation
  // order is from's, or to's where from has no levels
  const std::vector<int> flev = from->forest ? from->level_host : std::vector<int>(from->geom.nblocks, 0);
  const std::vector<int> tlev = to->forest ? to->level_host : std::vector<int>(to->geom.nblocks, 0);
  const int np = from->forest ? from->prolong : to->forest ? to->prolong : 2;
  if (a.ownership != b.ownership)
    return fail(TGH_EINVAL, "the transfer needs two forests with the same ownership of block faces");
  ForestIndex ix;
  ix.half_open = a.ownership == TGH_OWNERSHIP_HALF_OPEN;
  std::string msg;
  std::vector<int> ops;
  try {
    msg = ix.build(a.N, from->geom.nb, from->geom.periodic, flev, from->bcoord_host);
    if (msg.empty()) msg = transfer_plan(ix, from->geom.G, np, tlev, to->bcoord_host, ops);
  } catch (const std::exception& ex) {  // host threads or memory
    return fail(TGH_EDEVICE, std::string("building the transfer plan: ") + ex.what());
  }
  if (!msg.empty()) return fail(TGH_EINVAL, msg);
  HIP_TRY(hipStreamSynchronize(to->stream));  // the caller's work on u_to
  void* d_ops = nullptr;
  HIP_TRY(hipMalloc(&d_ops, sizeof(int) * ops.size()));
  hipError_t e = hipMemcpy(d_ops, ops.data(), sizeof(int) * ops.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    launch_scatter(from, u_from, from->W[0], 1);
    launch_fill(from, from->W[0], t);
    TransferArgs ta{from->W[0], u_to, (const PlanOp*)d_ops, (int)(ops.size() / 4), from->geom.S,
                    from->geom.P * from->geom.P * from->geom.P};
    const dim3 grid((ta.n + 255) / 256, 20);
    switch (np) {
      case 2: launch_transfer<2>(grid, from->stream, ta); break;
      case 4: launch_transfer<4>(grid, from->stream, ta); break;
      case 6: launch_transfer<6>(grid, from->stream, ta); break;
      case 8: launch_transfer<8>(grid, from->stream, ta); break;
      default: launch_transfer<10>(grid, from->stream, ta); break;
    }
    e = hipGetLastError();
    const hipError_t e2 = hipStreamSynchronize(from->stream);
    if (e == hipSuccess) e = e2;
  }
  (void)hipFree(d_ops);
  if (e != hipSuccess) return fail(TGH_EDEVICE, std::string("tgh_transfer: ") + hipGetErrorString(e));
  return TGH_OK;
}

int tgh_block_levels(const tgh_problem* p, int32_t* out) {
  if (!p || !out) return fail(TGH_EINVAL, "NULL argument");
  for (int b = 0; b < p->geom.nblocks; ++b) out[b] = p->forest ? p->level_host[b] : 0;
  return TGH_OK;
}

int tgh_problem_destroy(tgh_problem* p) {
  if (!p) return TGH_OK;
  if (p->stream) (void)hipStreamSynchronize(p->stream);
  if (p->graph_exec) (void)hipGraphExecDestroy(p->graph_exec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  for (hipEvent_t e : p->ev_pool) (void)hipEventDestroy(e);
  (void)hipFree(p->d_clock);
  (void)hipFree(p->Wus);
  (void)hipFree(p->Wacc);
  (void)hipFree(p->Wcoef);
  (void)hipFree(p->Blayer);
  (void)hipFree(p->d_bcoord);
  (void)hipFree(p->d_lookup);
  (void)hipFree(p->d_level);
  (void)hipFree(p->d_shape);
  (void)hipFree(p->gc);
  (void)hipFree(p->d_prim);
  (void)hipFree(p->d_plan);
  (void)hipFree(p->W[0]);
  (void)hipFree(p->W[1]);
  (void)hipFree(p->acc);
  (void)hipFree(p->coef);
  (void)hipFree(p->Hsrc);
  (void)hipFree(p->red);
  (void)hipFree(p->BW[0]);
  (void)hipFree(p->BW[1]);
  (void)hipFree(p->Bus);
  (void)hipFree(p->Bacc);
  (void)hipFree(p->BH);
  (void)hipFree(p->Bcoef);
  (void)hipFree(p->adjF);
  for (int i = 0; i < 2; ++i) {
    (void)hipFree(p->adjL[i]);
    (void)hipFree(p->adjK[i]);
  }
  (void)hipFree(p->adjTh);
  (void)hipFree(p->adjSlots);
  (void)hipFree(p->adjCkpt);
  if (p->own_stream && p->stream) (void)hipStreamDestroy(p->stream);
  if (p->cap_stream) (void)hipStreamDestroy(p->cap_stream);
  delete p;
  return TGH_OK;
}

int64_t tgh_state_length(const tgh_problem* p) { return p ? p->state_len : -1; }

int tgh_block_coords(const tgh_problem* p, int32_t* out) {
  if (!p || !out) return fail(TGH_EINVAL, "NULL argument");
  std::memcpy(out, p->bcoord_host.data(), sizeof(int) * p->bcoord_host.size());
  return TGH_OK;
}

int tgh_set_stream(tgh_problem* p, void* stream) {
  if (!p) return fail(TGH_EINVAL, "NULL problem");
  HIP_TRY(hipStreamSynchronize(p->stream));
  if (p->own_stream) HIP_TRY(hipStreamDestroy(p->stream));
  p->stream = (hipStream_t)stream;  // NULL: the device's null stream
  p->own_stream = false;
  if (p->graph_exec) {  // a captured graph belongs to the old stream's work order
    HIP_TRY(hipGraphExecDestroy(p->graph_exec));
    p->graph_exec = nullptr;
  }
  return TGH_OK;
}

int tgh_prepare(tgh_problem* p, const double* u, double t) {
  if (!p || !u) return fail(TGH_EINVAL, "NULL argument");
  launch_scatter(p, u, p->W[0], 1);
  launch_fill(p, p->W[0], t);
  return sync(p);
}

int tgh_rhs(tgh_problem* p, double* du, double t, const tgh_matter* matter) {
  if (!p || !du) return fail(TGH_EINVAL, "NULL argument");
  if (p->cs.interior == INT_FITTED && !p->cs.tw)
    return fail(TGH_EINVAL, "the :fitted variant reads its target cache, and none is set: call "
                            "tgh_set_target first");
  if (matter && matter->rho && (!matter->S || !matter->Sij))
    return fail(TGH_EINVAL, "a matter adapter needs rho, S and Sij");
  RhsArgs a = rhs_args(p, p->W[0], t, matter);
  a.du = du;
  launch_rhs(p, a, EPI_STORE);
  return sync(p);
}

int tgh_gh_rhs(tgh_problem* p, double* du, const double* u, double t, const tgh_matter* matter) {
  int rc = tgh_prepare(p, u, t);
  if (rc) return rc;
  return tgh_rhs(p, du, t, matter);
}

int tgh_settle(tgh_problem* p, double* u, double t) {
  int rc = tgh_prepare(p, u, t);
  if (rc) return rc;
  launch_scatter(p, u, p->W[0], 0);
  return sync(p);
}

int tgh_apply_bounds(tgh_problem* p, double* u, double t, int64_t* moved, int64_t* nonfinite, double* r_max) {
  if (!p || !u) return fail(TGH_EINVAL, "NULL argument");
  (void)t;  // the gate's center is static
  int64_t n = 0, nf = 0;
  double rm = -1.0;
  if (p->has_bounds) {
    launch_bounds_state(p, u);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(p->red_host.data(), p->red, sizeof(double) * p->red_host.size(),
                           hipMemcpyDeviceToHost, p->stream));
    HIP_TRY(hipStreamSynchronize(p->stream));
    double hits = 0, nfs = 0, r = 0;
    for (size_t w = 0; w < p->red_host.size() / RED_STRIDE; ++w) {
      hits += p->red_host[RED_STRIDE * w];
      nfs += p->red_host[RED_STRIDE * w + 1];
      r = std::max(r, p->red_host[RED_STRIDE * w + 2]);
    }
    n = (int64_t)hits;
    nf = (int64_t)nfs;
    if (n > 0) rm = r;
  }
  if (moved) *moved = n;
  if (nonfinite) *nonfinite = nf;
  if (r_max) *r_max = rm;
  return TGH_OK;
}

int tgh_stage_limit(tgh_problem* p, double* u, double t) {
  if (!p || !u) return fail(TGH_EINVAL, "NULL argument");
  if (p->has_bounds) {
    launch_bounds_state(p, u);
    int rc = sync(p);
    if (rc) return rc;
  }
  if (p->cs.interior != INT_PASTED) return TGH_OK;
  PointArgs a{p->geom, p->cs, t, u, 0, 0.0, 2};
  hipLaunchKernelGGL(p->cs.moving ? k_pointwise<true> : k_pointwise<false>, grid_points(p), dim3(256), 0, p->stream, a);
  return sync(p);
}
int tgh_step_limit(tgh_problem* p, double* u, double t) { return tgh_stage_limit(p, u, t); }

int tgh_initial_state(tgh_problem* p, double* u, double t) {
  if (!p || !u) return fail(TGH_EINVAL, "NULL argument");
  PointArgs a{p->geom, p->cs, t, u, 0, 0.0, 0};
  hipLaunchKernelGGL(p->cs.moving ? k_pointwise<true> : k_pointwise<false>, grid_points(p), dim3(256), 0, p->stream, a);
  int rc = sync(p);
  if (rc) return rc;
  return tgh_stage_limit(p, u, t);
}

int tgh_add_noise(tgh_problem* p, double* u, uint64_t base, double amplitude) {
  if (!p || !u) return fail(TGH_EINVAL, "NULL argument");
  if (!(amplitude >= 0)) return fail(TGH_EINVAL, "the noise amplitude must be >= 0");
  PointArgs a{p->geom, p->cs, 0.0, u, base, amplitude, 1};
  hipLaunchKernelGGL(p->cs.moving ? k_pointwise<true> : k_pointwise<false>, grid_points(p), dim3(256), 0, p->stream, a);
  return sync(p);
}

int tgh_max_speed(tgh_problem* p, const double* u, double t, double* out) {
  if (!p || !u || !out) return fail(TGH_EINVAL, "NULL argument");
  int rc = reduce(p, u, t, 0);
  if (rc) return rc;
  double lam = 0.0, nbad = 0.0;
  for (size_t w = 0; w < p->red_host.size() / RED_STRIDE; ++w) {
    const double l = p->red_host[RED_STRIDE * w];
    lam = (l > lam || std::isnan(l)) ? l : lam;
    nbad += p->red_host[RED_STRIDE * w + 1];
  }
  if (nbad > 0)
    return fail(TGH_ENOTFINITE, std::to_string((long)nbad) + " evolved points hold NaN or Inf at t = " +
                std::to_string(t));
  if (!(std::isfinite(lam) && lam > 0))
    return fail(TGH_ENOTFINITE, "max over evolved points of alpha sqrt(tr gamma^ij) + |beta| = " +
                std::to_string(lam) + ": no time step follows from it");
  *out = lam;
  return TGH_OK;
}

int tgh_error_rows(tgh_problem* p, const double* u, double t, double shell_lo, double shell_hi, double* out) {
  if (!p || !u || !out) return fail(TGH_EINVAL, "NULL argument");
  int rc = reduce(p, u, t, 1, shell_lo, shell_hi);
  if (rc) return rc;
  fold_error_rows(p->red_host.data(), p->red_host.size() / RED_STRIDE, out);
  return TGH_OK;
}

int tgh_error_norms(tgh_problem* p, const double* u, double t, double* l2, double* linf) {
  if (!p || !l2 || !linf) return fail(TGH_EINVAL, "NULL argument");
  double r[4];
  int rc = tgh_error_rows(p, u, t, 0.0, -1.0, r);
  if (rc) return rc;
  *l2 = r[0];
  *linf = r[1];
  return TGH_OK;
}

// The two constraint monitors: prepare u at t, launch the monitor (adm
// selects the ADM one), and fold its 9 partials per workgroup into the masked
// norms of its four components.
// The monitor's launches: the state into the work array, the ghost fill, and
// the constraint kernel, which writes 9 partials per workgroup into gc
// ([0..3] the weighted squares of the four components, [4..7] their largest
// magnitudes, [8] the weight).
static int launch_monitor(tgh_problem* p, const double* u, double t, bool adm, double* gc) {
  const Geom& g = p->geom;
  launch_scatter(p, u, p->W[0], 1);
  launch_fill(p, p->W[0], t);
  GCArgs a{g, p->cs, t, p->W[0], p->Hsrc, gc};
  if (adm) {
    switch (p->q) {
      case 2: hipLaunchKernelGGL(k_adm_constraint<2>, grid_points(p), dim3(256), 0, p->stream, a); break;
      case 4: hipLaunchKernelGGL(k_adm_constraint<4>, grid_points(p), dim3(256), 0, p->stream, a); break;
      case 6: hipLaunchKernelGGL(k_adm_constraint<6>, grid_points(p), dim3(256), 0, p->stream, a); break;
      default: hipLaunchKernelGGL(k_adm_constraint<8>, grid_points(p), dim3(256), 0, p->stream, a); break;
    }
  } else {
    switch (p->q) {
      case 2: hipLaunchKernelGGL(k_gauge_constraint<2>, grid_points(p), dim3(256), 0, p->stream, a); break;
      case 4: hipLaunchKernelGGL(k_gauge_constraint<4>, grid_points(p), dim3(256), 0, p->stream, a); break;
      case 6: hipLaunchKernelGGL(k_gauge_constraint<6>, grid_points(p), dim3(256), 0, p->stream, a); break;
      default: hipLaunchKernelGGL(k_gauge_constraint<8>, grid_points(p), dim3(256), 0, p->stream, a); break;
    }
  }
  HIP_TRY(hipGetLastError());
  return TGH_OK;
}

// The norms from the folded partials f: [0..3] the sums of the weighted
// squares, [4..7] the largest magnitudes, [8] the sum of the weights.
static void monitor_norms(const double f[9], double* l2, double* linf) {
  for (int q = 0; q < 4; ++q) {
    l2[q] = f[8] > 0 ? std::sqrt(f[q] / f[8]) : 0.0;
    linf[q] = f[4 + q];
  }
}

static int monitor(tgh_problem* p, const double* u, double t, bool adm, double* l2, double* linf) {
  const Geom& g = p->geom;
  const long nwg = (long)g.nblocks * ((g.P * g.P * g.P + 255) / 256);
  if (!p->gc) HIP_TRY(hipMalloc(&p->gc, sizeof(double) * 9 * nwg));
  int rc = launch_monitor(p, u, t, adm, p->gc);
  if (rc) return rc;
  std::vector<double> h(9 * nwg);
  HIP_TRY(hipMemcpyAsync(h.data(), p->gc, sizeof(double) * h.size(), hipMemcpyDeviceToHost, p->stream));
  HIP_TRY(hipStreamSynchronize(p->stream));
  // in workgroup order, so the answer does not depend on the launch
  double f[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (long w = 0; w < nwg; ++w) {
    for (int q = 0; q < 4; ++q) {
      f[q] += h[9 * w + q];
      f[4 + q] = std::max(f[4 + q], h[9 * w + 4 + q]);
    }
    f[8] += h[9 * w + 8];
  }
  monitor_norms(f, l2, linf);
  return TGH_OK;
}

// ---- the gauge monitor without allocation ----------------------------------
// A file-scope device array holds tgh_gauge_constraint_noalloc's partials
// (9 per workgroup from offset 0) and their fold (9 doubles at
// kScratchFold). Each device the library runs on holds its own copy until the
// process ends; host code reaches it only through the address
// hipGetSymbolAddress returns on that device (scratch_on). Every call on one
// device shares the array: calls that overlap there race.
constexpr long kScratchDoubles = 1L << 20;  // 8 MiB
constexpr long kScratchFold = kScratchDoubles - 16;
constexpr long kScratchMaxWorkgroups = kScratchFold / 9;
static __device__ __align__(256) double scratch[kScratchDoubles];
constexpr int kScratchDevices = 64;
static std::mutex scratch_mutex;
static double* scratch_addr[kScratchDevices] = {};

// The array's address on `device` (current), looked up once per device.
static int scratch_on(int device, double** out) {
  if (device < 0 || device >= kScratchDevices)
    return fail(TGH_EUNSUPPORTED, "device " + std::to_string(device) + ": the no-allocation entry points "
                                  "keep the scratch addresses of devices 0.." +
                                      std::to_string(kScratchDevices - 1) + " only");
  std::lock_guard<std::mutex> lock(scratch_mutex);
  if (!scratch_addr[device]) {
    void* ptr = nullptr;
    HIP_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    scratch_addr[device] = (double*)ptr;
  }
  *out = scratch_addr[device];
  return TGH_OK;
}

// The fold of monitor() on the device: thread q folds column q of the nwg
// partials in workgroup order (sums for 0..3 and 8, the larger-of for 4..7,
// the same comparisons as the host's), so the norms are the allocating
// entry point's to the bit.
__global__ __launch_bounds__(64) void k_monitor_fold(double* s, long nwg) {
  const int q = threadIdx.x;
  if (q >= 9) return;
  double acc = 0.0;
  if (q >= 4 && q < 8) {
    for (long w = 0; w < nwg; ++w) {
      const double v = s[9 * w + q];
      acc = acc < v ? v : acc;
    }
  } else {
    for (long w = 0; w < nwg; ++w) acc += s[9 * w + q];
  }
  s[kScratchFold + q] = acc;
}

int tgh_gauge_constraint_noalloc(tgh_problem* p, const double* u, double t, double* l2, double* linf) {
  if (!p || !u || !l2 || !linf) return fail(TGH_EINVAL, "NULL argument");
  const Geom& g = p->geom;
  const long wgpb = (g.P * g.P * g.P + 255) / 256;
  const long nwg = (long)g.nblocks * wgpb;
  if (nwg > kScratchMaxWorkgroups)
    return fail(TGH_EINVAL, "tgh_gauge_constraint_noalloc: " + std::to_string(g.nblocks) + " blocks of " +
                                std::to_string(g.P) + "^3 points need " + std::to_string(nwg) +
                                " workgroup partials, above the static scratch's " +
                                std::to_string(kScratchMaxWorkgroups) + " (" +
                                std::to_string(kScratchMaxWorkgroups / wgpb) +
                                " such blocks): use tgh_gauge_constraint");
  int device = 0;
  HIP_TRY(hipGetDevice(&device));
  if (device != p->device)
    return fail(TGH_EINVAL, "the problem's arrays live on device " + std::to_string(p->device) +
                                " but device " + std::to_string(device) + " is current");
  double* s = nullptr;
  int rc = scratch_on(device, &s);
  if (rc) return rc;
  rc = launch_monitor(p, u, t, false, s);
  if (rc) return rc;
  hipLaunchKernelGGL(k_monitor_fold, dim3(1), dim3(64), 0, p->stream, s, nwg);
  HIP_TRY(hipGetLastError());
  double f[9];
  HIP_TRY(hipMemcpyAsync(f, s + kScratchFold, sizeof(f), hipMemcpyDeviceToHost, p->stream));
  HIP_TRY(hipStreamSynchronize(p->stream));
  monitor_norms(f, l2, linf);
  return TGH_OK;
}

// The indicator's (scale >= 0) or the field scale's (scale < 0) partials,
// folded per block into out (nblocks doubles).
static int tau_pass(tgh_problem* p, const double* u, double t, double scale, double eps, double* out) {
  const Geom& g = p->geom;
  const int wgpb = (g.P * g.P * g.P + 255) / 256;
  const long nwg = (long)g.nblocks * wgpb;
  if (!p->gc) HIP_TRY(hipMalloc(&p->gc, sizeof(double) * 9 * nwg));
  if (scale < 0.0) {
    launch_scatter(p, u, p->W[0], 1);  // owned points only: no fill needed
  } else {
    launch_scatter(p, u, p->W[0], 1);
    launch_fill(p, p->W[0], t);
  }
  TauArgs a{g, p->cs, t, scale, eps, p->W[0], p->gc};
  hipLaunchKernelGGL(k_tau, grid_points(p), dim3(256), 0, p->stream, a);
  HIP_TRY(hipGetLastError());
  std::vector<double> h(nwg);
  HIP_TRY(hipMemcpyAsync(h.data(), p->gc, sizeof(double) * nwg, hipMemcpyDeviceToHost, p->stream));
  HIP_TRY(hipStreamSynchronize(p->stream));
  for (int b = 0; b < g.nblocks; ++b) {
    double m = 0.0;
    for (int w = 0; w < wgpb; ++w) m = std::max(m, h[(long)b * wgpb + w]);
    out[b] = m;
  }
  return TGH_OK;
}

int tgh_refinement_indicator(tgh_problem* p, const double* u, double t, double scale, double eps,
                             double* tau_blocks) {
  if (!p || !u || !tau_blocks) return fail(TGH_EINVAL, "NULL argument");
  if (!(scale >= 0 && std::isfinite(scale)) || !(eps >= 0 && std::isfinite(eps)))
    return fail(TGH_EINVAL, "the indicator's reference amplitude and noise fraction must be finite and >= 0");
  return tau_pass(p, u, t, scale, eps, tau_blocks);
}

int tgh_field_scale(tgh_problem* p, const double* u, double t, double* scale) {
  if (!p || !u || !scale) return fail(TGH_EINVAL, "NULL argument");
  std::vector<double> per(p->geom.nblocks);
  int rc = tau_pass(p, u, t, -1.0, 0.0, per.data());
  if (rc) return rc;
  double s = 0.0;
  for (double v : per) s = std::max(s, v);
  *scale = s;
  return TGH_OK;
}

int tgh_gauge_constraint(tgh_problem* p, const double* u, double t, double* l2, double* linf) {
  if (!p || !u || !l2 || !linf) return fail(TGH_EINVAL, "NULL argument");
  return monitor(p, u, t, false, l2, linf);
}

int tgh_adm_constraint(tgh_problem* p, const double* u, double t, double* l2, double* linf) {
  if (!p || !u || !l2 || !linf) return fail(TGH_EINVAL, "NULL argument");
  return monitor(p, u, t, true, l2, linf);
}

int tgh_rk4(tgh_problem* p, double* u, double t0, double dt, int32_t nsteps, int32_t flags) {
  if (!p || !u) return fail(TGH_EINVAL, "NULL argument");
  if (nsteps < 0 || !(dt > 0)) return fail(TGH_EINVAL, "tgh_rk4 needs dt > 0 and nsteps >= 0");
  if (p->cs.interior == INT_FITTED && !p->cs.tw)
    return fail(TGH_EINVAL, "the :fitted variant reads its target cache, and none is set: call "
                            "tgh_set_target first");
  if (nsteps == 0) return TGH_OK;
  const bool graph = (flags & TGH_RK4_GRAPH) != 0;
  // a forest with levels, the general interior and matter step on the block layout
  const bool blocks = (flags & TGH_RK4_BLOCKS) != 0 || p->forest || p->cs.general || p->mat.rho != nullptr;
  if (flags & ~(TGH_RK4_GRAPH | TGH_RK4_FUSED | TGH_RK4_BLOCKS | TGH_RK4_SPLIT | TGH_RK4_PER_BLOCK))
    return fail(TGH_EINVAL, "unknown tgh_rk4 flags " + std::to_string(flags));
  if ((flags & TGH_RK4_PER_BLOCK) && (flags & TGH_RK4_SPLIT))
    return fail(TGH_EINVAL, "TGH_RK4_PER_BLOCK launches the fused kernel: it excludes TGH_RK4_SPLIT");
  if (flags & TGH_RK4_PER_BLOCK) flags |= TGH_RK4_BLOCKS | TGH_RK4_FUSED;
  if ((flags & TGH_RK4_FUSED) && (flags & TGH_RK4_SPLIT))
    return fail(TGH_EINVAL, "TGH_RK4_FUSED and TGH_RK4_SPLIT exclude each other");
  // forms: 0 block layout, head + Pi; 1 block layout, fused; 2 brick layout,
  // fused; 3 brick layout, head + LDS-tiled Pi; 4 block layout, fused, one
  // launch per block; 5 block layout, each block run as a brick by the
  // brick's head and LDS-tiled Pi kernel (not with bounds, :pasted, matter,
  // the sampled source or the general interior: form 0 there). Without FUSED
  // or SPLIT the layout's faster form on MI355X (measured: the brick's split
  // form; on the block layout form 5 where it runs, else the fused kernel).
  const bool lds_ok = !p->has_bounds && p->cs.interior != INT_PASTED && !p->mat.rho &&
                      p->cs.gauge != GAUGE_SAMPLED && !p->cs.general;
  const bool split = (flags & TGH_RK4_SPLIT) != 0 ||
                     (!(flags & TGH_RK4_FUSED) && (blocks ? lds_ok : kBrickSplitDefault));
  const int form = (flags & TGH_RK4_PER_BLOCK) ? 4 : blocks ? (split ? (lds_ok ? 5 : 0) : 1) : (split ? 3 : 2);
  const bool in_brick = form == 2 || form == 3;
  if (in_brick) {
    int rc = ensure_brick(p);
    if (rc) return rc;
    if (form == 3 && !p->Bcoef) {
      HIP_TRY(hipMalloc(&p->Bcoef, sizeof(double) * CF_N * (size_t)p->brick.fs));
      HIP_TRY(hipMemsetAsync(p->Bcoef, 0, sizeof(double) * CF_N * (size_t)p->brick.fs, p->stream));
    }
    // u_n into the stepper's u_n and first stage input
    BrickCopyArgs ca{p->geom, p->brick, u, p->BW[0], p->Bus, 0};
    hipLaunchKernelGGL(k_brick_in, grid_brick(p), dim3(256), 0, p->stream, ca);
  } else {
    int rc = ensure_layer_box(p);
    if (rc) return rc;
    // the first stage input: u_n scattered into W[0]
    launch_scatter(p, u, p->W[0], 1);
    if (form == 5) {
      const size_t wbytes = sizeof(double) * (size_t)p->work_len;
      if (!p->Wus) {
        HIP_TRY(hipMalloc(&p->Wus, wbytes));
        HIP_TRY(hipMalloc(&p->Wacc, wbytes));
        HIP_TRY(hipMalloc(&p->Wcoef, wbytes / 20 * CF_N));
        HIP_TRY(hipMemsetAsync(p->Wcoef, 0, wbytes / 20 * CF_N, p->stream));
      }
      launch_scatter(p, u, p->Wus, 1);  // and u_n, in the work layout too
    }
  }
  if (form != p->graph_form && p->graph_exec) {  // the captured graph holds another form
    HIP_TRY(hipGraphExecDestroy(p->graph_exec));
    p->graph_exec = nullptr;
  }
  p->fused = form == 1 || form == 2 || form == 4;
  p->per_block = form == 4;
  auto step = [&](double t) {
    if (in_brick)
      rk4_brick_launches(p, t, dt);
    else if (form == 5)
      rk4_block_lds_launches(p, t, dt);
    else
      rk4_step_launches(p, u, t, dt);
  };
  int rc = TGH_OK;
  if (!graph) {
    for (int n = 0; n < nsteps && rc == TGH_OK; ++n) {
      step(t0 + n * dt);
      if (p->profiling && p->ev_used > 4096) {
        rc = sync(p);
        if (rc == TGH_OK) collect_profile(p);
      }
    }
  } else {
    // A hipGraph of one step (4 x (fill + fused RHS)), replayed nsteps times.
    // The captured kernels read the step's time from the stage clock, which a
    // one-thread kernel sets before each replay, so one graph serves every
    // step of every case (moving holes, :fitted targets, the gauge wave
    // included); it is re-captured only when dt, u or the form change.
    if (!p->d_clock) HIP_TRY(hipMalloc(&p->d_clock, sizeof(double)));
    for (int n = 0; n < nsteps; ++n) {
      const double t = t0 + n * dt;
      const bool reuse =
          p->graph_exec && p->graph_key[0] == dt && p->graph_key[1] == (double)(uintptr_t)u;
      if (!reuse) {
        if (p->graph_exec) {
          HIP_TRY(hipGraphExecDestroy(p->graph_exec));
          p->graph_exec = nullptr;
        }
        if (p->graph) {
          HIP_TRY(hipGraphDestroy(p->graph));
          p->graph = nullptr;
        }
        // The null stream cannot be captured: record on a private stream then,
        // and launch the graph on the problem's stream below.
        hipStream_t saved = p->stream;
        if (saved == nullptr) {
          if (!p->cap_stream) HIP_TRY(hipStreamCreateWithFlags(&p->cap_stream, hipStreamNonBlocking));
          p->stream = p->cap_stream;
        }
        hipError_t ec = hipStreamBeginCapture(p->stream, hipStreamCaptureModeThreadLocal);
        if (ec == hipSuccess) {
          p->clock_use = p->d_clock;
          step(t);
          p->clock_use = nullptr;
          ec = hipStreamEndCapture(p->stream, &p->graph);
        }
        p->stream = saved;
        if (ec != hipSuccess) {
          p->fused = false;
          return fail(TGH_EDEVICE, std::string("capturing the RK4 step: ") + hipGetErrorString(ec));
        }
        HIP_TRY(hipGraphInstantiate(&p->graph_exec, p->graph, nullptr, nullptr, 0));
        p->graph_form = form;
        p->graph_key[0] = dt;
        p->graph_key[1] = (double)(uintptr_t)u;
      }
      hipLaunchKernelGGL(k_set_clock, dim3(1), dim3(1), 0, p->stream, p->d_clock, t);
      HIP_TRY(hipGraphLaunch(p->graph_exec, p->stream));
    }
  }
  if (rc == TGH_OK && form == 5) launch_scatter(p, u, p->Wus, 0);  // u_{n+1} back into the state
  if (rc == TGH_OK && in_brick) {
    // u_{n+1} back into the state: the copies, then the primaries
    for (int pass = 0; pass < 2; ++pass) {
      BrickCopyArgs ca{p->geom, p->brick, p->Bus, u, nullptr, pass};
      hipLaunchKernelGGL(k_brick_out, grid_points(p), dim3(256), 0, p->stream, ca);
    }
  }
  if (rc == TGH_OK && reflects(p->geom)) {
    // gh_solve returns a settled state (src/stepping.jl:235-239): the walls'
    // odd variables, which the stages carried at roundoff, are zero
    PointArgs wa{p->geom, p->cs, t0 + nsteps * dt, u, 0, 0.0, 3};
    hipLaunchKernelGGL(p->cs.moving ? k_pointwise<true> : k_pointwise<false>, grid_points(p), dim3(256), 0,
                       p->stream, wa);
  }
  if (rc == TGH_OK) rc = sync(p);
  p->fused = false;
  p->per_block = false;
  if (rc == TGH_OK && p->profiling && !graph) collect_profile(p);
  return rc;
}

// ---------------------------------------------------------------------------
// The adjoint (the reference has none; the parameters are the constraint
// damping's Gaussian rates, gamma2, and the layer's rho_max and rho ramp).
namespace {

long adj_runs(const Brick& k) {
  const long plane = (long)k.n[0] * k.n[1];
  return (plane + 255) / 256 * k.n[2];
}

// The cases the adjoint covers, and its scratch (allocated once).
int adj_prepare(tgh_problem* p) {
  if (p->forest || p->cs.general)
    return fail(TGH_EUNSUPPORTED, "the adjoint runs on the brick layout: a uniform forest and a spherical "
                                  "interior (no levels, no tracked or :fitted geometry)");
  if (p->cs.moving) return fail(TGH_EUNSUPPORTED, "the adjoint is built for holes at rest");
  if (p->has_bounds) return fail(TGH_EUNSUPPORTED, "the adjoint is built for cases without range bounds");
  if (p->mat.rho) return fail(TGH_EUNSUPPORTED, "the adjoint is built for vacuum (no matter fields)");
  if (reflects(p->geom)) return fail(TGH_EUNSUPPORTED, "the adjoint's transposed fill does not mirror");
  if (p->cs.interior == INT_PASTED) return fail(TGH_EUNSUPPORTED, "the adjoint of the :pasted limiter is not built");
  if (p->cs.gauge != GAUGE_NONE && p->cs.gauge != GAUGE_SAMPLED && p->cs.gauge != GAUGE_KS)
    return fail(TGH_EUNSUPPORTED, "the adjoint is built for the static gauge-source kinds");
  int rc = ensure_brick(p);
  if (rc) return rc;
  const size_t fb = sizeof(double) * (size_t)p->brick.fs;
  if (!p->adjF) {
    HIP_TRY(hipMalloc(&p->adjF, AF_N * fb));
    HIP_TRY(hipMemsetAsync(p->adjF, 0, AF_N * fb, p->stream));  // the frame stays zero
    for (int i = 0; i < 2; ++i) {
      HIP_TRY(hipMalloc(&p->adjL[i], 20 * fb));
      HIP_TRY(hipMemsetAsync(p->adjL[i], 0, 20 * fb, p->stream));
      HIP_TRY(hipMalloc(&p->adjK[i], 20 * fb));
      HIP_TRY(hipMemsetAsync(p->adjK[i], 0, 20 * fb, p->stream));
    }
    HIP_TRY(hipMalloc(&p->adjTh, sizeof(double) * ADJ_NPAR * adj_runs(p->brick)));
  }
  return TGH_OK;
}

// The parameters' cotangent: the per-workgroup accumulators folded in order.
int adj_theta(tgh_problem* p, double* out) {
  const long n = adj_runs(p->brick);
  std::vector<double> h((size_t)(ADJ_NPAR * n));
  HIP_TRY(hipMemcpyAsync(h.data(), p->adjTh, sizeof(double) * h.size(), hipMemcpyDeviceToHost, p->stream));
  HIP_TRY(hipStreamSynchronize(p->stream));
  for (int q = 0; q < ADJ_NPAR; ++q) out[q] = 0.0;
  for (long w = 0; w < n; ++w)
    for (int q = 0; q < ADJ_NPAR; ++q) out[q] += h[(size_t)(ADJ_NPAR * w + q)];
  return TGH_OK;
}

// One RK4 stage on explicit brick arrays (rk4_brick_launches' stage s with
// the stage input Wc and the next stage's Wn).
void brick_stage(tgh_problem* p, double* Wc, double* Wn, double t, double dt, int s) {
  const double c_acc[4] = {dt / 6, dt / 3, dt / 3, dt / 6};
  const double c_next[4] = {dt / 2, dt / 2, dt, 0.0};
  const double t_stage[4] = {t, t + dt / 2, t + dt / 2, t + dt};
  if (!p->frame_static) launch_frame(p, Wc, t_stage[s]);
  RhsArgs a = rhs_args(p, Wc, t_stage[s], nullptr);
  a.Hsrc = p->BH;
  a.us = p->Bus;
  a.acc = p->Bacc;
  a.Wn = Wn;
  a.cacc = c_acc[s];
  a.cnext = c_next[s];
  a.stage = s + 1;
  switch (p->q) {
    case 2: launch_rhs_brick<2>(p, a); break;
    case 4: launch_rhs_brick<4>(p, a); break;
    case 6: launch_rhs_brick<6>(p, a); break;
    default: launch_rhs_brick<8>(p, a); break;
  }
}

}  // namespace

int tgh_vjp_rhs(tgh_problem* p, const double* u, double t, const double* du_bar, double* u_bar, double* theta_bar) {
  if (!p || !u || !du_bar || !u_bar || !theta_bar) return fail(TGH_EINVAL, "NULL argument");
  int rc = adj_prepare(p);
  if (rc) return rc;
  const Brick& k = p->brick;
  BrickCopyArgs ca{p->geom, k, u, p->BW[0], p->Bus, 0};
  hipLaunchKernelGGL(k_brick_in, grid_brick(p), dim3(256), 0, p->stream, ca);
  if (!p->frame_static) launch_frame(p, p->BW[0], t);
  launch_adj_fold_copies(p, AdjCopiesArgs{p->geom, k, du_bar, p->adjK[0]});
  HIP_TRY(hipMemsetAsync(p->adjTh, 0, sizeof(double) * ADJ_NPAR * adj_runs(k), p->stream));
  launch_adj_point(p, AdjPointArgs{k, p->cs, t, p->BW[0], p->BH, p->adjK[0], p->adjF, p->adjTh});
  launch_adj_gather(p, AdjGatherArgs{k, p->adjF, nullptr, nullptr, nullptr, p->adjL[0], 0.0, 0.0, 0, 0});
  HIP_TRY(hipMemsetAsync(u_bar, 0, sizeof(double) * (size_t)p->state_len, p->stream));
  launch_adj_to_primaries(p, AdjCopiesArgs{p->geom, k, p->adjL[0], u_bar});
  HIP_TRY(hipGetLastError());
  return adj_theta(p, theta_bar);
}

int tgh_adjoint(tgh_problem* p, double* u, double t0, double dt, int32_t nsteps, int32_t every, double* J,
                double* grad, double* ms) {
  if (!p || !u || !J || !grad) return fail(TGH_EINVAL, "NULL argument");
  if (nsteps < 1 || every < 1 || !(dt > 0) || !std::isfinite(t0))
    return fail(TGH_EINVAL, "tgh_adjoint needs nsteps >= 1, every >= 1, dt > 0 and a finite t0");
  int rc = adj_prepare(p);
  if (rc) return rc;
  const Brick& k = p->brick;
  const size_t fb20 = sizeof(double) * 20 * (size_t)k.fs;
  const size_t nslot = 4 * (size_t)every + 1, nseg = ((size_t)nsteps + every - 1) / every;
  if (p->adjSlotsN < nslot) {
    HIP_TRY(hipFree(p->adjSlots));
    p->adjSlots = nullptr;
    p->adjSlotsN = 0;
    HIP_TRY(hipMalloc(&p->adjSlots, nslot * fb20));
    HIP_TRY(hipMemsetAsync(p->adjSlots, 0, nslot * fb20, p->stream));
    p->adjSlotsN = nslot;
    if (p->frame_static)
      for (size_t i = 0; i < nslot; ++i) launch_frame(p, p->adjSlots + i * 20 * (size_t)k.fs, 0.0);
  }
  if (p->adjCkptN < nseg) {
    HIP_TRY(hipFree(p->adjCkpt));
    p->adjCkpt = nullptr;
    p->adjCkptN = 0;
    HIP_TRY(hipMalloc(&p->adjCkpt, nseg * fb20));
    p->adjCkptN = nseg;
  }
  if (!p->Bcoef) {
    HIP_TRY(hipMalloc(&p->Bcoef, sizeof(double) * CF_N * (size_t)k.fs));
    HIP_TRY(hipMemsetAsync(p->Bcoef, 0, sizeof(double) * CF_N * (size_t)k.fs, p->stream));
  }
  auto slot = [&](size_t i) { return p->adjSlots + i * 20 * (size_t)k.fs; };
  auto ckpt = [&](size_t i) { return p->adjCkpt + i * 20 * (size_t)k.fs; };
  hipEvent_t ev[4];
  for (auto& e : ev) HIP_TRY(hipEventCreate(&e));
  p->fused = false;  // the split form (head + LDS-tiled Pi), the stepper's default
  // ---- forward: u_0 into the brick, a checkpoint at each segment's start ----
  HIP_TRY(hipEventRecord(ev[0], p->stream));
  BrickCopyArgs ca{p->geom, k, u, p->BW[0], p->Bus, 0};
  hipLaunchKernelGGL(k_brick_in, grid_brick(p), dim3(256), 0, p->stream, ca);
  for (size_t sg = 0; sg < nseg; ++sg) {
    HIP_TRY(hipMemcpyAsync(ckpt(sg), p->BW[0], fb20, hipMemcpyDeviceToDevice, p->stream));
    const int n0 = (int)(sg * every), n1 = std::min(nsteps, (int)((sg + 1) * every));
    for (int n = n0; n < n1; ++n) rk4_brick_launches(p, t0 + n * dt, dt);
  }
  HIP_TRY(hipEventRecord(ev[1], p->stream));
  for (int pass = 0; pass < 2; ++pass) {
    BrickCopyArgs co{p->geom, k, p->Bus, u, nullptr, pass};
    hipLaunchKernelGGL(k_brick_out, grid_points(p), dim3(256), 0, p->stream, co);
  }
  // J = the masked L2 error at T, and lambda_T = dJ/du_T
  const double T = t0 + nsteps * dt;
  rc = reduce(p, u, T, 1);
  if (rc) return rc;
  double num = 0.0, den = 0.0;
  for (size_t w = 0; w < p->red_host.size() / RED_STRIDE; ++w) {
    num += p->red_host[RED_STRIDE * w];
    den += p->red_host[RED_STRIDE * w + 1];
  }
  *J = den > 0 ? std::sqrt(num / den) : 0.0;
  double *lam = p->adjL[0], *lacc = p->adjL[1], *kb = p->adjK[0], *kbn = p->adjK[1];
  launch_adj_init(p, AdjInitArgs{k, p->cs, T, *J, den, dt / 6, p->Bus, lam, kb});
  HIP_TRY(hipMemsetAsync(p->adjTh, 0, sizeof(double) * ADJ_NPAR * adj_runs(k), p->stream));
  // ---- backward: per segment, its stage inputs recomputed from the
  // checkpoint, then its steps reversed ----
  HIP_TRY(hipEventRecord(ev[2], p->stream));
  const double cl[4] = {dt / 6, dt / 6, dt / 3, dt / 3}, cy[4] = {0.0, dt / 2, dt / 2, dt};
  for (size_t sg = nseg; sg-- > 0;) {
    const int n0 = (int)(sg * every), n1 = std::min(nsteps, (int)((sg + 1) * every)), ns = n1 - n0;
    HIP_TRY(hipMemcpyAsync(slot(0), ckpt(sg), fb20, hipMemcpyDeviceToDevice, p->stream));
    HIP_TRY(hipMemcpyAsync(p->Bus, ckpt(sg), fb20, hipMemcpyDeviceToDevice, p->stream));
    for (int j = 0; j < ns; ++j)
      for (int s = 0; s < 4; ++s) brick_stage(p, slot(4 * j + s), slot(4 * j + s + 1), t0 + (n0 + j) * dt, dt, s);
    for (int j = ns; j-- > 0;) {
      const double t = t0 + (n0 + j) * dt;
      const double t_stage[4] = {t, t + dt / 2, t + dt / 2, t + dt};
      for (int s = 4; s-- > 0;) {
        double* Wc = slot(4 * j + s);
        if (!p->frame_static) launch_frame(p, Wc, t_stage[s]);
        launch_adj_point(p, AdjPointArgs{k, p->cs, t_stage[s], Wc, p->BH, kb, p->adjF, p->adjTh});
        launch_adj_gather(p, AdjGatherArgs{k, p->adjF, lam, lacc, kbn, nullptr, cl[s], cy[s], s == 3, s == 0});
        std::swap(kb, kbn);
      }
      std::swap(lam, lacc);
    }
  }
  HIP_TRY(hipEventRecord(ev[3], p->stream));
  HIP_TRY(hipGetLastError());
  rc = adj_theta(p, grad);
  if (rc) return rc;
  if (ms) {
    float f0 = 0.0f, f1 = 0.0f;
    HIP_TRY(hipEventElapsedTime(&f0, ev[0], ev[1]));
    HIP_TRY(hipEventElapsedTime(&f1, ev[2], ev[3]));
    ms[0] = f0;
    ms[1] = f1;
  }
  for (auto& e : ev) HIP_TRY(hipEventDestroy(e));
  return TGH_OK;
}

int tgh_set_tile(tgh_problem* p, int32_t tile) {
  if (!p) return fail(TGH_EINVAL, "NULL problem");
  if (tile < 0 || tile >= PI_TILES)
    return fail(TGH_EINVAL, "the Pi kernel's tile shape is one of 0 .. " + std::to_string(PI_TILES - 1) + ", got " +
                                std::to_string(tile));
  HIP_TRY(hipStreamSynchronize(p->stream));
  p->pi_tile = tile;
  if (p->graph_exec) {  // a captured graph launches the old shape
    HIP_TRY(hipGraphExecDestroy(p->graph_exec));
    p->graph_exec = nullptr;
  }
  return TGH_OK;
}

int tgh_autotune(tgh_problem* p, double* u, double t0, double dt, double* ms) {
  if (!p || !u) return fail(TGH_EINVAL, "NULL argument");
  if (p->forest || p->cs.general)
    return fail(TGH_EUNSUPPORTED, "the tile shapes are the brick stepper's; this problem steps on the block layout");
  const bool was_profiling = p->profiling;
  int best = 0;
  double best_ms = 0.0;
  for (int tc = 0; tc < PI_TILES; ++tc) {
    int rc = tgh_set_tile(p, tc);
    if (rc) return rc;
    // one step to warm the shape, then two timed (hipEvents around each Pi launch)
    rc = tgh_rk4(p, u, t0, dt, 1, TGH_RK4_SPLIT);
    if (rc) return rc;
    rc = tgh_profile(p, 1);
    if (rc) return rc;
    rc = tgh_rk4(p, u, t0 + dt, dt, 2, TGH_RK4_SPLIT);
    if (rc) return rc;
    const double pi_ms = p->prof[5] > 0 ? p->prof[4] / p->prof[5] : 0.0;
    if (ms) ms[tc] = pi_ms;
    if (tc == 0 || pi_ms < best_ms) {
      best = tc;
      best_ms = pi_ms;
    }
    t0 += 3 * dt;
  }
  int rc = tgh_profile(p, was_profiling ? 1 : 0);
  if (rc) return rc;
  return tgh_set_tile(p, best);
}

int tgh_profile(tgh_problem* p, int32_t enable) {
  if (!p) return fail(TGH_EINVAL, "NULL problem");
  HIP_TRY(hipStreamSynchronize(p->stream));
  p->profiling = enable != 0;
  p->ev_rhs.clear();
  p->ev_fill.clear();
  p->ev_pi.clear();
  p->ev_used = 0;
  for (double& x : p->prof) x = 0;
  return TGH_OK;
}

int tgh_profile_read(tgh_problem* p, double* out) {
  if (!p || !out) return fail(TGH_EINVAL, "NULL argument");
  for (int i = 0; i < 6; ++i) out[i] = p->prof[i];
  return TGH_OK;
}

int64_t tgh_work_length(const tgh_problem* p) { return p ? p->work_len : -1; }

int tgh_copy_work(tgh_problem* p, double* dst) {
  if (!p || !dst) return fail(TGH_EINVAL, "NULL argument");
  HIP_TRY(hipMemcpyAsync(dst, p->W[0], sizeof(double) * p->work_len, hipMemcpyDeviceToDevice, p->stream));
  return sync(p);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Batches: several problems stepped by one head and one Pi launch per stage.
struct tgh_batch {
  std::vector<tgh_problem*> probs;
  int q = 4, gauge = 0;
  int hole = 0;  // the head's interior kind: 0, 1 (the sphere) or HOLE_BOXED
  // The block layout (a problem on a forest with levels or with the general
  // interior among them): one fused launch over every block of every problem
  // per stage; the RK4 epilogue writes the states in place, so the tables and
  // the graph hold the states' pointers.
  bool blocks = false;
  bool lds = false;                  // the block layout with its blocks run as bricks
  int hole_kind = 0;                 // 0 none, 1 the sphere, 2 the general interior
  std::vector<double*> tab_u, graph_u;
  BatchEntry* d_tab = nullptr;  // 4 stage tables of n entries
  BatchEntry* h_tab = nullptr;  // pinned staging of the tables
  int head_total = 0, pi_total = 0;
  double tab_key[2] = {0, 0};   // (dt, t) the uploaded tables hold
  std::vector<unsigned> tab_ver;  // the problems' versions the tables were built from
  bool tab_valid = false;
  hipStream_t stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  double graph_dt = 0;  // the step of the captured graph
  // The stage clock (the step's time on the device, set before each step):
  // the step's first launch stamps the stage times into the tables, and the
  // problems' frames and fills read it, so tables and graph outlive a step.
  double* d_clock = nullptr;
};

namespace {
void pi_tile_dims(int q, int& tx, int& ty, int& tz) {
  switch (q) {
    case 2: tx = PiTile<2>::X; ty = PiTile<2>::Y; tz = PiTile<2>::Z * PiTile<2>::PZ; break;
    case 4: tx = PiTile<4>::X; ty = PiTile<4>::Y; tz = PiTile<4>::Z * PiTile<4>::PZ; break;
    case 6: tx = PiTile<6>::X; ty = PiTile<6>::Y; tz = PiTile<6>::Z * PiTile<6>::PZ; break;
    default: tx = PiTile<8>::X; ty = PiTile<8>::Y; tz = PiTile<8>::Z * PiTile<8>::PZ; break;
  }
}
// The stage tables at step time t: problem i's stage-s arguments, as
// rk4_brick_launches sets them for the problem alone.
int batch_tables(tgh_batch* b, double t, double dt) {
  const int n = (int)b->probs.size();
  const double c_acc[4] = {dt / 6, dt / 3, dt / 3, dt / 6};
  const double c_next[4] = {dt / 2, dt / 2, dt, 0.0};
  const double t_stage[4] = {t, t + dt / 2, t + dt / 2, t + dt};
  int h0 = 0, p0 = 0, tx, ty, tz;
  pi_tile_dims(b->q, tx, ty, tz);
  for (int i = 0; i < n; ++i) {
    tgh_problem* p = b->probs[i];
    const Brick& k = p->brick;
    const long plane = (long)k.n[0] * k.n[1];
    const int runs = (int)((plane + 255) / 256) * k.n[2];
    const int tiles = ((k.n[0] + tx - 1) / tx) * ((k.n[1] + ty - 1) / ty) * ((k.n[2] + tz - 1) / tz);
    for (int s = 0; s < 4; ++s) {
      BatchEntry& e = b->h_tab[s * n + i];
      e.a = rhs_args(p, p->BW[s % 2], t_stage[s], nullptr);
      e.a.Hsrc = p->BH;
      e.a.us = p->Bus;
      e.a.acc = p->Bacc;
      e.a.Wn = p->BW[(s + 1) % 2];
      e.a.du = nullptr;
      e.a.coef = p->Bcoef;
      e.a.cacc = c_acc[s];
      e.a.cnext = c_next[s];
      e.a.stage = s + 1;
      if (b->hole == HOLE_BOXED) {
        e.a.lbox = p->Blayer;
        e.a.lB3 = (long)p->layer_B * p->layer_B * p->layer_B;
        e.a.lB = p->layer_B;
      }
      e.k = k;
      e.W = (decltype(e.W))e.a.W;
      e.Hsrc = (decltype(e.Hsrc))e.a.Hsrc;
      e.coef = (decltype(e.coef))e.a.coef;
      e.us = (decltype(e.us))e.a.us;
      e.acc = (decltype(e.acc))e.a.acc;
      e.Wn = (decltype(e.Wn))e.a.Wn;
      e.head0 = h0;
      e.pi0 = p0;
    }
    h0 += runs;
    p0 += tiles;
  }
  b->head_total = h0;
  b->pi_total = p0;
  HIP_TRY(hipMemcpyAsync(b->d_tab, b->h_tab, sizeof(BatchEntry) * 4 * n, hipMemcpyHostToDevice, b->stream));
  // the pinned tables are rewritten only after this copy has run
  HIP_TRY(hipStreamSynchronize(b->stream));
  b->tab_key[0] = dt;
  b->tab_key[1] = t;
  b->tab_valid = true;
  return TGH_OK;
}
// The block layout's stage tables: problem i's stage-s arguments as
// rk4_step_launches sets them for the problem alone, u_i its state.
int batch_tables_blocks(tgh_batch* b, double* const* u, double t, double dt) {
  const int n = (int)b->probs.size();
  const double c_acc[4] = {dt / 6, dt / 3, dt / 3, dt / 6};
  const double c_next[4] = {dt / 2, dt / 2, dt, 0.0};
  const double t_stage[4] = {t, t + dt / 2, t + dt / 2, t + dt};
  int w0 = 0, pw0 = 0;
  for (int i = 0; i < n; ++i) {
    tgh_problem* p = b->probs[i];
    const int wgpb = (p->geom.P * p->geom.P * p->geom.P + 255) / 256;
    for (int s = 0; s < 4; ++s) {
      BatchEntry& e = b->h_tab[s * n + i];
      std::memset(&e, 0, sizeof(e));
      e.a = rhs_args(p, p->W[s % 2], t_stage[s], nullptr);
      e.a.us = b->lds ? p->Wus : u[i];
      e.a.acc = b->lds ? p->Wacc : p->acc;
      e.a.Wn = p->W[(s + 1) % 2];
      e.a.du = nullptr;
      e.a.cacc = c_acc[s];
      e.a.cnext = c_next[s];
      e.a.stage = s + 1;
      if (b->lds) e.a.coef = p->Wcoef;
      e.W = (decltype(e.W))e.a.W;
      e.Hsrc = (decltype(e.Hsrc))e.a.Hsrc;
      e.us = (decltype(e.us))e.a.us;
      e.acc = (decltype(e.acc))e.a.acc;
      e.Wn = (decltype(e.Wn))e.a.Wn;
      e.coef = (decltype(e.coef))e.a.coef;
      e.head0 = w0;
      e.pi0 = pw0;
      e.lLmax = p->layer_lmax;
    }
    w0 += wgpb * p->geom.nblocks;
    if (b->lds) {
      const int P = p->geom.P;
      using T = PiTile<4, PI_TILE_BLOCK>;  // the block tile's shape is the same at every order
      pw0 += ((P + T::X - 1) / T::X) * ((P + T::Y - 1) / T::Y) * ((P + T::Z - 1) / T::Z) * p->geom.nblocks;
    }
  }
  b->head_total = w0;
  b->pi_total = pw0;
  HIP_TRY(hipMemcpyAsync(b->d_tab, b->h_tab, sizeof(BatchEntry) * 4 * n, hipMemcpyHostToDevice, b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  b->tab_key[0] = dt;
  b->tab_key[1] = t;
  b->tab_u.assign(u, u + n);
  b->tab_valid = true;
  return TGH_OK;
}
void batch_step_launches_blocks(tgh_batch* b, double t, double dt) {
  const int n = (int)b->probs.size();
  const double t_off[4] = {0.0, dt / 2, dt / 2, dt};
  const double t_stage[4] = {t, t + dt / 2, t + dt / 2, t + dt};
  for (int s = 0; s < 4; ++s) {
    for (tgh_problem* p : b->probs) launch_fill(p, p->W[s % 2], t_stage[s], t_off[s]);
    const BatchEntry* tab = b->d_tab + s * n;
    if (b->lds) {
      const int ht = b->head_total, pt = b->pi_total, hk = b->hole_kind ? 1 : 0;
      switch (b->q) {
        case 2: launch_batch_blocks_lds<2>(b->stream, tab, n, ht, pt, b->gauge, hk); break;
        case 4: launch_batch_blocks_lds<4>(b->stream, tab, n, ht, pt, b->gauge, hk); break;
        case 6: launch_batch_blocks_lds<6>(b->stream, tab, n, ht, pt, b->gauge, hk); break;
        default: launch_batch_blocks_lds<8>(b->stream, tab, n, ht, pt, b->gauge, hk); break;
      }
    } else {
      switch (b->q) {
        case 2: launch_batch_blocks<2>(b->stream, tab, n, b->head_total, b->gauge, b->hole_kind); break;
        case 4: launch_batch_blocks<4>(b->stream, tab, n, b->head_total, b->gauge, b->hole_kind); break;
        case 6: launch_batch_blocks<6>(b->stream, tab, n, b->head_total, b->gauge, b->hole_kind); break;
        default: launch_batch_blocks<8>(b->stream, tab, n, b->head_total, b->gauge, b->hole_kind); break;
      }
    }
    for (int i = 0; i < n; ++i) {
      tgh_problem* p = b->probs[i];
      if (p->has_bounds) launch_bounds_work(p, p->W[(s + 1) % 2], s == 3 ? b->tab_u[i] : nullptr);
    }
  }
}
void batch_step_launches(tgh_batch* b, double t, double dt) {
  // the stage times from the clock, into the tables and the problems' launches
  const int n = (int)b->probs.size();
  hipLaunchKernelGGL(k_batch_clock, dim3((4 * n + 255) / 256), dim3(256), 0, b->stream, b->d_tab, n, b->d_clock,
                     dt / 2, dt);
  for (tgh_problem* p : b->probs) p->clock_use = b->d_clock;
  if (b->blocks) {
    batch_step_launches_blocks(b, t, dt);
    for (tgh_problem* p : b->probs) p->clock_use = nullptr;
    return;
  }
  const double t_off[4] = {0.0, dt / 2, dt / 2, dt};
  const double t_stage[4] = {t, t + dt / 2, t + dt / 2, t + dt};
  for (int s = 0; s < 4; ++s) {
    for (tgh_problem* p : b->probs) {
      if (!p->frame_static) launch_frame(p, p->BW[s % 2], t_stage[s], t_off[s]);
      if (b->hole == HOLE_BOXED) {
        RhsArgs unused = rhs_args(p, p->BW[s % 2], t_stage[s], nullptr);
        launch_layer_box(p, t_stage[s], t_off[s], unused);
      }
    }
    const BatchEntry* tab = b->d_tab + s * n;
    switch (b->q) {
      case 2: launch_batch<2>(b->stream, tab, n, b->head_total, b->pi_total, b->gauge, b->hole); break;
      case 4: launch_batch<4>(b->stream, tab, n, b->head_total, b->pi_total, b->gauge, b->hole); break;
      case 6: launch_batch<6>(b->stream, tab, n, b->head_total, b->pi_total, b->gauge, b->hole); break;
      default: launch_batch<8>(b->stream, tab, n, b->head_total, b->pi_total, b->gauge, b->hole); break;
    }
    for (tgh_problem* p : b->probs)
      if (p->has_bounds) launch_bounds_brick(p, s == 3 ? p->Bus : p->BW[(s + 1) % 2], s == 3 ? p->BW[(s + 1) % 2] : nullptr);
  }
  for (tgh_problem* p : b->probs) p->clock_use = nullptr;
}
void batch_drop_graph(tgh_batch* b) {
  if (b->exec) hipGraphExecDestroy(b->exec);
  if (b->graph) hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}
}  // namespace

extern "C" {

int tgh_batch_create(tgh_problem* const* probs, int32_t n, tgh_batch** out) {
  if (!probs || !out) return fail(TGH_EINVAL, "NULL argument");
  *out = nullptr;
  if (n < 1 || n > 1024) return fail(TGH_EINVAL, "a batch holds 1 to 1024 problems, got " + std::to_string(n));
  for (int i = 0; i < n; ++i) {
    if (!probs[i]) return fail(TGH_EINVAL, "problem " + std::to_string(i) + " of the batch is NULL");
    for (int j = 0; j < i; ++j)
      if (probs[j] == probs[i])
        return fail(TGH_EINVAL, "problem " + std::to_string(i) + " appears twice in the batch: its brick arrays "
                                "would be stepped twice per stage");
  }
  const tgh_problem* p0 = probs[0];
  auto hole_kind = [](const tgh_problem* p) { return p->cs.interior == INT_NONE ? 0 : p->cs.general ? 2 : 1; };
  bool blocks = false;
  for (int i = 0; i < n; ++i) blocks = blocks || probs[i]->forest || probs[i]->cs.general;
  for (int i = 0; i < n; ++i) {
    const tgh_problem* p = probs[i];
    if (p->q != p0->q || p->cs.gauge != p0->cs.gauge || hole_kind(p) != hole_kind(p0))
      return fail(TGH_EINVAL, "the problems of a batch share the order, the gauge-source kind and the kind of "
                              "interior (one kernel specialisation runs them all); problem " +
                                  std::to_string(i) + " differs from problem 0");
    if (p->cs.interior == INT_PASTED)
      return fail(TGH_EUNSUPPORTED, "a batch does not run the :pasted limiter; step problem " + std::to_string(i) +
                                        " alone (tgh_rk4)");
    if (reflects(p->geom))
      return fail(TGH_EUNSUPPORTED, "a batch steps problems without reflecting faces; step problem " +
                                        std::to_string(i) + " alone (tgh_rk4)");
    if (p->mat.rho)
      return fail(TGH_EUNSUPPORTED, "a batch steps vacuum problems; problem " + std::to_string(i) + " has matter");
    if (blocks && (p->cs.gauge == GAUGE_KS_MOVING || p->cs.gauge == GAUGE_NONE_MOVING))
      return fail(TGH_EUNSUPPORTED, "a batch on the block layout (forests with levels, tracked or :fitted "
                                    "interiors) of moving holes is not built; problem " + std::to_string(i) + " moves");
  }
  tgh_batch* b = new tgh_batch;
  b->probs.assign(probs, probs + n);
  b->q = p0->q;
  b->gauge = p0->cs.gauge;
  b->hole = p0->cs.interior != INT_NONE ? 1 : 0;
  b->blocks = blocks;
  b->hole_kind = hole_kind(p0);
  auto bail = [&](int rc) {
    tgh_batch_destroy(b);
    return rc;
  };
  if (blocks && b->hole_kind != 2) {
    // the blocks run as bricks where every problem allows it (as tgh_rk4's form 5)
    bool ok = true;
    for (tgh_problem* p : b->probs)
      ok = ok && !p->has_bounds && p->cs.interior != INT_PASTED && !p->mat.rho && p->cs.gauge != GAUGE_SAMPLED &&
           !p->cs.general;
    b->lds = ok;
    for (tgh_problem* p : b->probs) {
      if (!b->lds || p->Wus) continue;
      const size_t wbytes = sizeof(double) * (size_t)p->work_len;
      if (hipMalloc(&p->Wus, wbytes) != hipSuccess || hipMalloc(&p->Wacc, wbytes) != hipSuccess ||
          hipMalloc(&p->Wcoef, wbytes / 20 * CF_N) != hipSuccess ||
          hipMemset(p->Wcoef, 0, wbytes / 20 * CF_N) != hipSuccess)
        return bail(fail(TGH_EDEVICE, "allocating a batch's work-layout arrays"));
    }
  }
  for (tgh_problem* p : b->probs) {
    if (blocks) continue;
    int rc = ensure_brick(p);
    if (rc) return bail(rc);
    if (!p->Bcoef) {
      const size_t bytes = sizeof(double) * CF_N * (size_t)p->brick.fs;
      if (hipMalloc(&p->Bcoef, bytes) != hipSuccess) return bail(fail(TGH_EDEVICE, "allocating a batch's coefficients"));
      if (hipMemsetAsync(p->Bcoef, 0, bytes, p->stream) != hipSuccess) return bail(fail(TGH_EDEVICE, "clearing them"));
    }
    rc = sync(p);
    if (rc) return bail(rc);
  }
  // moving :damped holes with their layer boxes: the head reads the boxes, as
  // each problem's own steps do, so that every problem computes its own bits
  if (!blocks && b->hole) {
    bool boxed = true;
    for (tgh_problem* p : b->probs) boxed = boxed && p->Blayer;
    if (boxed) b->hole = HOLE_BOXED;
  }
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess)
    return bail(fail(TGH_EDEVICE, "creating the batch's stream"));
  if (hipMalloc(&b->d_tab, sizeof(BatchEntry) * 4 * n) != hipSuccess ||
      hipHostMalloc(&b->h_tab, sizeof(BatchEntry) * 4 * n) != hipSuccess ||
      hipMalloc(&b->d_clock, sizeof(double)) != hipSuccess)
    return bail(fail(TGH_EDEVICE, "allocating the batch's tables"));
  *out = b;
  return TGH_OK;
}

int tgh_batch_rk4(tgh_batch* b, double* const* u, double t0, double dt, int32_t nsteps, int32_t flags) {
  if (!b || !u) return fail(TGH_EINVAL, "NULL argument");
  if (nsteps < 0 || !(dt > 0)) return fail(TGH_EINVAL, "tgh_batch_rk4 needs dt > 0 and nsteps >= 0");
  if (flags & ~TGH_RK4_GRAPH) return fail(TGH_EINVAL, "tgh_batch_rk4 takes 0 or TGH_RK4_GRAPH, got " + std::to_string(flags));
  const int n = (int)b->probs.size();
  for (int i = 0; i < n; ++i)
    if (!u[i]) return fail(TGH_EINVAL, "state " + std::to_string(i) + " of the batch is NULL");
  if (nsteps == 0) return TGH_OK;
  const bool graph = (flags & TGH_RK4_GRAPH) != 0;
  // every launch of the batch (the problems' own frame, bounds and copies
  // included) goes on the batch's stream
  std::vector<hipStream_t> saved(n);
  for (int i = 0; i < n; ++i) {
    saved[i] = b->probs[i]->stream;
    HIP_TRY(hipStreamSynchronize(saved[i]));
    b->probs[i]->stream = b->stream;
  }
  int rc = TGH_OK;
  for (int i = 0; i < n; ++i) {
    tgh_problem* p = b->probs[i];
    if (b->blocks) {
      if (p->cs.interior == INT_FITTED && !p->cs.tw) {
        rc = fail(TGH_EINVAL, "problem " + std::to_string(i) + " has a :fitted interior and no target "
                              "(tgh_set_target)");
        break;
      }
      launch_scatter(p, u[i], p->W[0], 1);  // u_n into the first stage's work array
      if (b->lds) launch_scatter(p, u[i], p->Wus, 1);  // and the work-layout u_n
      continue;
    }
    BrickCopyArgs ca{p->geom, p->brick, u[i], p->BW[0], p->Bus, 0};
    hipLaunchKernelGGL(k_brick_in, grid_brick(p), dim3(256), 0, b->stream, ca);
  }
  const std::vector<double*> uv(u, u + n);
  // on the block layout the tables carry each problem's case, whose :fitted
  // target (tgh_set_target) may have changed since the last call
  if (b->blocks) b->tab_valid = false;
  // a setter (tgh_set_rho_max, tgh_set_target) changed a problem's case
  std::vector<unsigned> ver;
  for (tgh_problem* q : b->probs) ver.push_back(q->version);
  if (ver != b->tab_ver) {
    b->tab_valid = false;
    b->tab_ver = ver;
  }
  for (int s = 0; s < nsteps && rc == TGH_OK; ++s) {
    const double t = t0 + s * dt;
    // the stage times come from the clock: the tables (and the graph) are kept
    // across steps and rebuilt only when dt, the states or a problem's case change
    const bool stale = !b->tab_valid || b->tab_key[0] != dt || (b->blocks && b->tab_u != uv);
    if (stale) {
      rc = b->blocks ? batch_tables_blocks(b, u, t, dt) : batch_tables(b, t, dt);
      if (rc) break;
    }
    hipLaunchKernelGGL(k_set_clock, dim3(1), dim3(1), 0, b->stream, b->d_clock, t);
    if (!graph) {
      batch_step_launches(b, t, dt);
      continue;
    }
    const bool reuse = b->exec && b->graph_dt == dt && (!b->blocks || b->graph_u == uv);
    if (!reuse) {
      batch_drop_graph(b);
      hipError_t ec = hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal);
      if (ec == hipSuccess) {
        batch_step_launches(b, t, dt);
        ec = hipStreamEndCapture(b->stream, &b->graph);
      }
      if (ec == hipSuccess) ec = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
      if (ec != hipSuccess) {
        batch_drop_graph(b);
        rc = fail(TGH_EDEVICE, std::string("capturing the batch's RK4 step: ") + hipGetErrorString(ec));
        break;
      }
      b->graph_dt = dt;
      b->graph_u = uv;
    }
    if (hipGraphLaunch(b->exec, b->stream) != hipSuccess) {
      rc = fail(TGH_EDEVICE, "launching the batch's RK4 graph");
      break;
    }
  }
  if (rc == TGH_OK) {
    for (int i = 0; i < n && b->lds; ++i) launch_scatter(b->probs[i], u[i], b->probs[i]->Wus, 0);
    for (int i = 0; i < n && !b->blocks; ++i) {  // (the block layout wrote u in place)
      tgh_problem* p = b->probs[i];
      for (int pass = 0; pass < 2; ++pass) {
        BrickCopyArgs ca{p->geom, p->brick, p->Bus, u[i], nullptr, pass};
        hipLaunchKernelGGL(k_brick_out, grid_points(p), dim3(256), 0, b->stream, ca);
      }
    }
    hipError_t e = hipStreamSynchronize(b->stream);
    if (e == hipSuccess) e = hipGetLastError();
    if (e != hipSuccess) rc = fail(TGH_EDEVICE, std::string("the batch's RK4 steps: ") + hipGetErrorString(e));
  } else {
    hipStreamSynchronize(b->stream);
  }
  for (int i = 0; i < n; ++i) b->probs[i]->stream = saved[i];
  return rc;
}

int tgh_batch_destroy(tgh_batch* b) {
  if (!b) return TGH_OK;
  if (b->stream) hipStreamSynchronize(b->stream);
  batch_drop_graph(b);
  if (b->d_tab) hipFree(b->d_tab);
  if (b->h_tab) hipHostFree(b->h_tab);
  if (b->d_clock) hipFree(b->d_clock);
  if (b->stream) hipStreamDestroy(b->stream);
  delete b;
  return TGH_OK;
}

}  // extern "C"
