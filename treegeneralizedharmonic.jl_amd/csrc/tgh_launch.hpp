// The right-hand side launches, templated on the order: one translation unit
// per order (tgh_q<Q>.hip) includes this and instantiates them.
#pragma once
#define TGH_RHS_ONLY
#include "tgh_problem.hpp"

namespace tgh {


// The brick Pi kernel with tile shape TC.
template <int Q, int TC>
void launch_brick_pi(tgh_problem* p, const RhsArgs& a) {
  using T = PiTile<Q, TC>;
  hipLaunchKernelGGL((k_brick_pi<Q, TC>), grid_brick_tiles(p, T::X, T::Y, T::Z * T::PZ), dim3(T::X * T::Y * T::Z),
                     0, p->stream, a, p->brick, a.W, (const double*)p->Bcoef, a.du, a.us, a.acc, a.Wn);
}

// One right-hand side: the fused kernel, or the head and the Pi components as
// two launches. While profiling, the head is timed as the "rhs" pair and the
// Pi kernel as its own pair.
template <int Q, int GAUGE, int HOLE, int EPI>
void launch_rhs_q(tgh_problem* p, const RhsArgs& a) {
  const bool prof = p->profiling && EPI == EPI_RK4;
  if (p->per_block) {
    // One launch per block (the fused kernel over that block alone): the
    // plain per-block form, kept as the baseline of the batched forms. Block
    // b sees a one-block mesh whose coordinates and arrays start at its own.
    const long P3 = (long)a.g.P * a.g.P * a.g.P, S3 = (long)a.g.S * a.g.S * a.g.S;
    const dim3 grid = xcd_grid((P3 + 255) / 256);
    for (int b = 0; b < a.g.nblocks; ++b) {
      RhsArgs ab = a;
      ab.g.nblocks = 1;
      ab.g.bcoord = a.g.bcoord + 3 * b;
      if (a.g.blevel) ab.g.blevel = a.g.blevel + b;
      const long ws = (long)b * 20 * S3, ss = (long)b * 20 * P3;
      ab.W = a.W + ws;
      if (a.Wn) ab.Wn = a.Wn + ws;
      if (a.us) ab.us = a.us + ss;
      if (a.acc) ab.acc = a.acc + ss;
      if (a.du) ab.du = a.du + ss;
      if (a.Hsrc) ab.Hsrc = a.Hsrc + ss;
      if (a.c.tw) ab.c.tw = a.c.tw + 2 * ss;  // the target cache: 40 variables a point
      hipLaunchKernelGGL((k_rhs<Q, GAUGE, HOLE, EPI>), grid, dim3(256), 0, p->stream, ab, ab.W, ab.du, ab.us,
                         ab.acc, ab.Wn);
    }
    return;
  }
  if (p->fused) {
    int e0 = prof ? mark(p) : -1;
    hipLaunchKernelGGL((k_rhs<Q, GAUGE, HOLE, EPI>), grid_points(p), dim3(256), 0, p->stream, a, a.W,
                       a.du, a.us, a.acc, a.Wn);
    if (prof) p->ev_rhs.push_back({e0, mark(p)});
  } else {
    int e0 = prof ? mark(p) : -1;
    hipLaunchKernelGGL((k_head<Q, GAUGE, HOLE, EPI>), grid_points(p), dim3(256), 0, p->stream, a, a.W,
                       a.coef, a.du, a.us, a.acc, a.Wn);
    int e1 = prof ? mark(p) : -1;
    hipLaunchKernelGGL((k_pi<Q, EPI>), grid_points(p), dim3(256), 0, p->stream, a, a.W,
                       (const double*)a.coef, a.du, a.us, a.acc, a.Wn);
    if (prof) {
      p->ev_rhs.push_back({e0, e1});
      p->ev_pi.push_back({e1, mark(p)});
    }
  }
}
template <int Q, int EPI>
void launch_rhs_g(tgh_problem* p, const RhsArgs& a) {
  const bool hole = p->cs.interior != INT_NONE;
  if (hole && p->cs.general) {
    // the general interior (tracked geometry, :fitted): its own specialisation
    switch (p->cs.gauge) {
      case GAUGE_KS: launch_rhs_q<Q, GAUGE_KS, 2, EPI>(p, a); break;
      case GAUGE_SAMPLED: launch_rhs_q<Q, GAUGE_SAMPLED, 2, EPI>(p, a); break;
      case GAUGE_KS_MOVING: launch_rhs_q<Q, GAUGE_KS_MOVING, 2, EPI>(p, a); break;
      case GAUGE_NONE_MOVING: launch_rhs_q<Q, GAUGE_NONE_MOVING, 2, EPI>(p, a); break;
      default: launch_rhs_q<Q, GAUGE_NONE, 2, EPI>(p, a); break;
    }
    return;
  }
  if constexpr (EPI == EPI_RK4) {
    if (a.lbox) {  // a moving :damped hole or a harmonic hole at rest, with its layer box
      if (p->cs.gauge == GAUGE_KS_MOVING)
        launch_rhs_q<Q, GAUGE_KS_MOVING, HOLE_BOXED, EPI>(p, a);
      else if (p->cs.gauge == GAUGE_NONE_MOVING)
        launch_rhs_q<Q, GAUGE_NONE_MOVING, HOLE_BOXED, EPI>(p, a);
      else
        launch_rhs_q<Q, GAUGE_NONE, HOLE_BOXED, EPI>(p, a);
      return;
    }
  }
  switch (p->cs.gauge) {
    case GAUGE_KS:
      if (hole) launch_rhs_q<Q, GAUGE_KS, true, EPI>(p, a);
      else launch_rhs_q<Q, GAUGE_KS, false, EPI>(p, a);
      break;
    case GAUGE_SAMPLED:
      if (hole) launch_rhs_q<Q, GAUGE_SAMPLED, true, EPI>(p, a);
      else launch_rhs_q<Q, GAUGE_SAMPLED, false, EPI>(p, a);
      break;
    case GAUGE_KS_MOVING:  // a moving hole always has an interior
      launch_rhs_q<Q, GAUGE_KS_MOVING, true, EPI>(p, a);
      break;
    case GAUGE_NONE_MOVING:
      launch_rhs_q<Q, GAUGE_NONE_MOVING, true, EPI>(p, a);
      break;
    default:
      if (hole) launch_rhs_q<Q, GAUGE_NONE, true, EPI>(p, a);
      else launch_rhs_q<Q, GAUGE_NONE, false, EPI>(p, a);
      break;
  }
}
template <int Q>
void launch_rhs_blocks(tgh_problem* p, const RhsArgs& a, int epi) {
  if (epi == EPI_RK4)
    launch_rhs_g<Q, EPI_RK4>(p, a);
  else
    launch_rhs_g<Q, EPI_STORE>(p, a);
}

template <int Q, int GAUGE, int HOLE>
void launch_brick_rhs_q(tgh_problem* p, const RhsArgs& a) {
  const bool prof = p->profiling;
  int e0 = prof ? mark(p) : -1;
  if (p->fused) {
    hipLaunchKernelGGL((k_brick_rhs<Q, GAUGE, HOLE>), grid_brick(p), dim3(256), 0, p->stream, a, p->brick,
                       a.W, a.du, a.us, a.acc, a.Wn);
    if (prof) p->ev_rhs.push_back({e0, mark(p)});
  } else {
    hipLaunchKernelGGL((k_brick_head<Q, GAUGE, HOLE>), grid_brick(p), dim3(256), 0, p->stream, a, p->brick,
                       a.W, p->Bcoef, a.du, a.us, a.acc, a.Wn);
    int e1 = prof ? mark(p) : -1;
    switch (p->pi_tile) {
      case 1: launch_brick_pi<Q, 1>(p, a); break;
      case 2: launch_brick_pi<Q, 2>(p, a); break;
      case 3: launch_brick_pi<Q, 3>(p, a); break;
      case 4: launch_brick_pi<Q, 4>(p, a); break;
      case 5: launch_brick_pi<Q, 5>(p, a); break;
      default: launch_brick_pi<Q, 0>(p, a); break;
    }
    if (prof) {
      p->ev_rhs.push_back({e0, e1});
      p->ev_pi.push_back({e1, mark(p)});
    }
  }
}
template <int Q>
void launch_brick_rhs_g(tgh_problem* p, const RhsArgs& a) {
  const bool hole = p->cs.interior != INT_NONE;
  if (a.lbox) {  // a moving :damped hole or a harmonic hole at rest, with its layer box
    if (p->cs.gauge == GAUGE_KS_MOVING)
      launch_brick_rhs_q<Q, GAUGE_KS_MOVING, HOLE_BOXED>(p, a);
    else if (p->cs.gauge == GAUGE_NONE_MOVING)
      launch_brick_rhs_q<Q, GAUGE_NONE_MOVING, HOLE_BOXED>(p, a);
    else
      launch_brick_rhs_q<Q, GAUGE_NONE, HOLE_BOXED>(p, a);
    return;
  }
  switch (p->cs.gauge) {
    case GAUGE_KS:
      if (hole) launch_brick_rhs_q<Q, GAUGE_KS, true>(p, a);
      else launch_brick_rhs_q<Q, GAUGE_KS, false>(p, a);
      break;
    case GAUGE_SAMPLED:
      if (hole) launch_brick_rhs_q<Q, GAUGE_SAMPLED, true>(p, a);
      else launch_brick_rhs_q<Q, GAUGE_SAMPLED, false>(p, a);
      break;
    case GAUGE_KS_MOVING:
      launch_brick_rhs_q<Q, GAUGE_KS_MOVING, true>(p, a);
      break;
    case GAUGE_NONE_MOVING:
      launch_brick_rhs_q<Q, GAUGE_NONE_MOVING, true>(p, a);
      break;
    default:
      if (hole) launch_brick_rhs_q<Q, GAUGE_NONE, true>(p, a);
      else launch_brick_rhs_q<Q, GAUGE_NONE, false>(p, a);
      break;
  }
}

template <int Q>
void launch_rhs_brick(tgh_problem* p, const RhsArgs& a) {
  launch_brick_rhs_g<Q>(p, a);
}

// The block layout's split form run with the brick's code (k_block_head, then
// k_block_pi over every block's tiles); timed while profiling as the brick's is.
template <int Q, int GAUGE, int HOLE>
void launch_block_lds_q(tgh_problem* p, const RhsArgs& a, const BlockBricks& bb) {
  using T = PiTile<Q, PI_TILE_BLOCK>;
  const int P = a.g.P;
  const long P3 = (long)P * P * P;
  const long tiles = (long)((P + T::X - 1) / T::X) * ((P + T::Y - 1) / T::Y) * ((P + T::Z - 1) / T::Z);
  const bool prof = p->profiling;
  int e0 = prof ? mark(p) : -1;
  hipLaunchKernelGGL((k_block_head<Q, GAUGE, HOLE>), xcd_grid((P3 + 255) / 256 * a.g.nblocks), dim3(256), 0,
                     p->stream, a, bb);
  int e1 = prof ? mark(p) : -1;
  hipLaunchKernelGGL((k_block_pi<Q>), xcd_grid(tiles * a.g.nblocks), dim3(T::X * T::Y * T::Z), 0, p->stream, a,
                     bb);
  if (prof) {
    p->ev_rhs.push_back({e0, e1});
    p->ev_pi.push_back({e1, mark(p)});
  }
}
template <int Q>
void launch_rhs_blocks_lds(tgh_problem* p, const RhsArgs& a, const BlockBricks& bb) {
  const bool hole = p->cs.interior != INT_NONE;
  if (a.lbox) {  // a moving :damped hole or a harmonic hole at rest, with its layer box
    if (p->cs.gauge == GAUGE_KS_MOVING)
      launch_block_lds_q<Q, GAUGE_KS_MOVING, HOLE_BOXED>(p, a, bb);
    else if (p->cs.gauge == GAUGE_NONE_MOVING)
      launch_block_lds_q<Q, GAUGE_NONE_MOVING, HOLE_BOXED>(p, a, bb);
    else
      launch_block_lds_q<Q, GAUGE_NONE, HOLE_BOXED>(p, a, bb);
    return;
  }
  switch (p->cs.gauge) {
    case GAUGE_KS:
      if (hole) launch_block_lds_q<Q, GAUGE_KS, true>(p, a, bb);
      else launch_block_lds_q<Q, GAUGE_KS, false>(p, a, bb);
      break;
    case GAUGE_KS_MOVING:
      launch_block_lds_q<Q, GAUGE_KS_MOVING, true>(p, a, bb);
      break;
    case GAUGE_NONE_MOVING:
      launch_block_lds_q<Q, GAUGE_NONE_MOVING, true>(p, a, bb);
      break;
    default:  // GAUGE_NONE (the sampled source is read in the state layout: not this form)
      if (hole) launch_block_lds_q<Q, GAUGE_NONE, true>(p, a, bb);
      else launch_block_lds_q<Q, GAUGE_NONE, false>(p, a, bb);
      break;
  }
}

// One RK4 stage of a batch of problems: the head and the Pi kernel, each one
// launch over every problem's runs or tiles (the table's ranges).
template <int Q, int GAUGE, int HOLE>
void launch_batch_q(hipStream_t st, const BatchEntry* tab, int n, int head_total, int pi_total) {
  using T = PiTile<Q>;
  hipLaunchKernelGGL((k_brick_head_batch<Q, GAUGE, HOLE>), xcd_grid(head_total), dim3(256), 0, st, tab, n,
                     head_total);
  hipLaunchKernelGGL((k_brick_pi_batch<Q>), xcd_grid(pi_total), dim3(T::X * T::Y * T::Z), 0, st, tab, n, pi_total);
}
template <int Q>
void launch_batch(hipStream_t st, const BatchEntry* tab, int n, int head_total, int pi_total, int gauge, int hole) {
  if (hole == HOLE_BOXED) {  // holes with their layer boxes
    if (gauge == GAUGE_KS_MOVING)
      launch_batch_q<Q, GAUGE_KS_MOVING, HOLE_BOXED>(st, tab, n, head_total, pi_total);
    else if (gauge == GAUGE_NONE_MOVING)
      launch_batch_q<Q, GAUGE_NONE_MOVING, HOLE_BOXED>(st, tab, n, head_total, pi_total);
    else
      launch_batch_q<Q, GAUGE_NONE, HOLE_BOXED>(st, tab, n, head_total, pi_total);
    return;
  }
  switch (gauge) {
    case GAUGE_KS:
      if (hole) launch_batch_q<Q, GAUGE_KS, true>(st, tab, n, head_total, pi_total);
      else launch_batch_q<Q, GAUGE_KS, false>(st, tab, n, head_total, pi_total);
      break;
    case GAUGE_SAMPLED:
      if (hole) launch_batch_q<Q, GAUGE_SAMPLED, true>(st, tab, n, head_total, pi_total);
      else launch_batch_q<Q, GAUGE_SAMPLED, false>(st, tab, n, head_total, pi_total);
      break;
    case GAUGE_KS_MOVING:
      launch_batch_q<Q, GAUGE_KS_MOVING, true>(st, tab, n, head_total, pi_total);
      break;
    case GAUGE_NONE_MOVING:
      launch_batch_q<Q, GAUGE_NONE_MOVING, true>(st, tab, n, head_total, pi_total);
      break;
    default:
      if (hole) launch_batch_q<Q, GAUGE_NONE, true>(st, tab, n, head_total, pi_total);
      else launch_batch_q<Q, GAUGE_NONE, false>(st, tab, n, head_total, pi_total);
      break;
  }
}

// One RK4 stage of a batch on the block layout (static gauge-source kinds;
// hole 0 none, 1 the sphere, 2 the general interior).
template <int Q, int GAUGE>
void launch_batch_blocks_g(hipStream_t st, const BatchEntry* tab, int n, int total, int hole) {
  const dim3 grid = xcd_grid(total);
  if (hole == 2)
    hipLaunchKernelGGL((k_rhs_batch<Q, GAUGE, 2>), grid, dim3(256), 0, st, tab, n, total);
  else if (hole == 1)
    hipLaunchKernelGGL((k_rhs_batch<Q, GAUGE, 1>), grid, dim3(256), 0, st, tab, n, total);
  else
    hipLaunchKernelGGL((k_rhs_batch<Q, GAUGE, 0>), grid, dim3(256), 0, st, tab, n, total);
}
// The same with the blocks run as bricks (static gauge-source kinds; hole 0 or 1).
template <int Q, int GAUGE, int HOLE>
void launch_batch_blocks_lds_q(hipStream_t st, const BatchEntry* tab, int n, int head_total, int pi_total) {
  using T = PiTile<Q, PI_TILE_BLOCK>;
  hipLaunchKernelGGL((k_block_head_batch<Q, GAUGE, HOLE>), xcd_grid(head_total), dim3(256), 0, st, tab, n,
                     head_total);
  hipLaunchKernelGGL((k_block_pi_batch<Q>), xcd_grid(pi_total), dim3(T::X * T::Y * T::Z), 0, st, tab, n, pi_total);
}
template <int Q>
void launch_batch_blocks_lds(hipStream_t st, const BatchEntry* tab, int n, int head_total, int pi_total, int gauge,
                             int hole) {
  if (gauge == GAUGE_KS) {
    if (hole) launch_batch_blocks_lds_q<Q, GAUGE_KS, 1>(st, tab, n, head_total, pi_total);
    else launch_batch_blocks_lds_q<Q, GAUGE_KS, 0>(st, tab, n, head_total, pi_total);
  } else {
    if (hole) launch_batch_blocks_lds_q<Q, GAUGE_NONE, 1>(st, tab, n, head_total, pi_total);
    else launch_batch_blocks_lds_q<Q, GAUGE_NONE, 0>(st, tab, n, head_total, pi_total);
  }
}
template <int Q>
void launch_batch_blocks(hipStream_t st, const BatchEntry* tab, int n, int total, int gauge, int hole) {
  switch (gauge) {
    case GAUGE_KS: launch_batch_blocks_g<Q, GAUGE_KS>(st, tab, n, total, hole); break;
    case GAUGE_SAMPLED: launch_batch_blocks_g<Q, GAUGE_SAMPLED>(st, tab, n, total, hole); break;
    default: launch_batch_blocks_g<Q, GAUGE_NONE>(st, tab, n, total, hole); break;
  }
}

}  // namespace tgh

#define TGH_INSTANTIATE_ORDER(Q)                                                            \
  template void tgh::launch_rhs_blocks<Q>(tgh_problem*, const tgh::RhsArgs&, int);          \
  template void tgh::launch_rhs_brick<Q>(tgh_problem*, const tgh::RhsArgs&);                \
  template void tgh::launch_rhs_blocks_lds<Q>(tgh_problem*, const tgh::RhsArgs&, const tgh::BlockBricks&); \
  template void tgh::launch_batch_blocks<Q>(hipStream_t, const tgh::BatchEntry*, int, int, int, int); \
  template void tgh::launch_batch_blocks_lds<Q>(hipStream_t, const tgh::BatchEntry*, int, int, int, int, int); \
  template void tgh::launch_batch<Q>(hipStream_t, const tgh::BatchEntry*, int, int, int, int, int);
