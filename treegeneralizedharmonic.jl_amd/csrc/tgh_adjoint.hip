// The adjoint's kernels in a translation unit of their own (they compile in
// parallel with the right-hand side's): the point VJP at each order, gauge
// kind and interior flag a hole at rest takes, the transposed stencils at
// each order, and the launchers tgh_api.hip calls.
#define TGH_RHS_ONLY
#define TGH_ADJOINT_KERNELS
#include "tgh_problem.hpp"
#include "tgh_adjoint.hpp"

namespace tgh {

template <int Q, int GAUGE>
static void adj_point_hole(const tgh_problem* p, const AdjPointArgs& a, bool hole) {
  if (hole)
    hipLaunchKernelGGL((k_adj_point<Q, GAUGE, 1>), grid_brick(p), dim3(256), 0, p->stream, a);
  else
    hipLaunchKernelGGL((k_adj_point<Q, GAUGE, 0>), grid_brick(p), dim3(256), 0, p->stream, a);
}
template <int Q>
static void adj_point_q(const tgh_problem* p, const AdjPointArgs& a, bool hole) {
  switch (a.c.gauge) {
    case GAUGE_SAMPLED: adj_point_hole<Q, GAUGE_SAMPLED>(p, a, hole); break;
    case GAUGE_KS: adj_point_hole<Q, GAUGE_KS>(p, a, hole); break;
    default: adj_point_hole<Q, GAUGE_NONE>(p, a, hole); break;
  }
}
void launch_adj_point(const tgh_problem* p, const AdjPointArgs& a) {
  const bool hole = a.c.interior != INT_NONE;
  switch (p->q) {
    case 2: adj_point_q<2>(p, a, hole); break;
    case 4: adj_point_q<4>(p, a, hole); break;
    case 6: adj_point_q<6>(p, a, hole); break;
    default: adj_point_q<8>(p, a, hole); break;
  }
}
void launch_adj_gather(const tgh_problem* p, const AdjGatherArgs& a) {
  switch (p->q) {
    case 2: hipLaunchKernelGGL(k_adj_gather<2>, grid_brick(p), dim3(256), 0, p->stream, a); break;
    case 4: hipLaunchKernelGGL(k_adj_gather<4>, grid_brick(p), dim3(256), 0, p->stream, a); break;
    case 6: hipLaunchKernelGGL(k_adj_gather<6>, grid_brick(p), dim3(256), 0, p->stream, a); break;
    default: hipLaunchKernelGGL(k_adj_gather<8>, grid_brick(p), dim3(256), 0, p->stream, a); break;
  }
}
void launch_adj_init(const tgh_problem* p, const AdjInitArgs& a) {
  hipLaunchKernelGGL(k_adj_init<false>, grid_brick(p), dim3(256), 0, p->stream, a);
}
void launch_adj_fold_copies(const tgh_problem* p, const AdjCopiesArgs& a) {
  hipLaunchKernelGGL(k_adj_fold_copies, grid_brick(p), dim3(256), 0, p->stream, a);
}
void launch_adj_to_primaries(const tgh_problem* p, const AdjCopiesArgs& a) {
  hipLaunchKernelGGL(k_adj_to_primaries, grid_brick(p), dim3(256), 0, p->stream, a);
}

}  // namespace tgh
