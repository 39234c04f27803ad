// The library's problem object and the pieces of host code shared between
// the API's translation unit (tgh_api.hip) and the right-hand side's, one per
// finite-difference order (tgh_q2.hip ... tgh_q8.hip), which compile in
// parallel.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <utility>
#include <vector>

#include "../../include/tgh.h"
#include "tgh_kernels.hpp"

struct tgh_problem {
  tgh_desc desc;
  tgh::Geom geom;
  tgh::CaseP cs;
  tgh::Bounds bd{};                   // the range projection's ranges
  bool has_bounds = false;
  int q = 4;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  std::vector<int> bcoord_host;
  int* d_bcoord = nullptr;
  int* d_lookup = nullptr;
  // A forest with levels (tgh_problem_create_forest): per-block levels, the
  // primary flags and the fill plan, in passes.
  bool forest = false;
  std::vector<int> level_host;
  int* d_level = nullptr;
  unsigned char* d_prim = nullptr;
  int* d_plan = nullptr;
  std::vector<int> plan_off;
  int plan_dir[2] = {0, 0};           // pass 0's Dirichlet entries
  int prolong = 0;
  double* d_shape = nullptr;          // the tracked surface's coefficients
  double* gc = nullptr;               // the gauge constraint's partials (9 per workgroup)
  double* W[2] = {nullptr, nullptr};  // work arrays (prepare uses W[0])
  double* acc = nullptr;              // RK4 accumulator, state layout
  double* coef = nullptr;             // the split form's per-point coefficients (CF_N x points)
  double* Hsrc = nullptr;             // sampled gauge source, state layout x 20
  double* red = nullptr;              // reduction partials (RED_STRIDE per workgroup)
  std::vector<double> red_host;
  long state_len = 0, work_len = 0;
  hipGraphExec_t graph_exec = nullptr;
  hipGraph_t graph = nullptr;
  double graph_key[2] = {0, 0};       // (dt, u pointer) the graph was captured for
  // The stage clock: a device scalar holding the step's time, set before each
  // replay of the graph, which the captured kernels read (stage_time); clock_use
  // points at it while launches are recorded for a graph, else NULL.
  double* d_clock = nullptr;
  const double* clock_use = nullptr;
  unsigned version = 0;               // bumped when a setter changes the case (batches re-read it)
  tgh_matter mat{nullptr, nullptr, nullptr};  // tgh_set_matter's fields for tgh_rk4 (rho NULL: vacuum)
  int device = 0;                     // the HIP device current at creation, which holds the arrays
  // Benchmark timing: event pairs around each launch while profiling.
  bool profiling = false;
  bool fused = false;                 // one kernel per RHS instead of head + Pi
  bool per_block = false;             // the fused block kernel launched once per block
  int pi_tile = 0;                    // the brick Pi kernel's tile shape (tgh_autotune, tgh_set_tile)
  int graph_form = -1;                // the form the captured graph launches
  // The stepper's brick layout (allocated at the first tgh_rk4 that uses it):
  // two stage arrays, u_n and the accumulator, and the sampled gauge source.
  tgh::Brick brick;
  double* BW[2] = {nullptr, nullptr};
  double* Bus = nullptr;
  double* Bacc = nullptr;
  double* BH = nullptr;
  double* Bcoef = nullptr;            // the split form's coefficients (CF_N brick fields)
  double* Blayer = nullptr;           // a moving :damped hole's layer box (LAYER_FIELDS x layer_B^3)
  int layer_B = 0;
  int layer_lmax = 0;                 // the finest level, whose grid the box is on
  bool layer_static = false;          // filled once (a hole at rest)
  // the block layout's LDS form: u_n, the accumulator and the Pi coefficients
  // in the work layout during a step (allocated at its first use)
  double* Wus = nullptr;
  double* Wacc = nullptr;
  double* Wcoef = nullptr;
  bool brick_ready = false;
  bool frame_static = false;          // the frame never changes: filled once
  hipStream_t cap_stream = nullptr;   // capture stream when the problem runs on the null stream
  std::vector<hipEvent_t> ev_pool;
  std::vector<std::pair<int, int>> ev_rhs, ev_fill, ev_pi;  // indices into ev_pool
  size_t ev_used = 0;
  double prof[6] = {0, 0, 0, 0, 0, 0};
  // The adjoint's device memory (tgh_vjp_rhs, tgh_adjoint; allocated at
  // their first call): the VJP's scratch fields (AF_N brick fields), the
  // cotangents lambda, its accumulator and the stage cotangents (20 brick
  // fields each), the parameters' per-workgroup accumulators; the stage
  // inputs of one checkpoint segment and the checkpoints (grown on demand).
  double* adjF = nullptr;
  double* adjL[2] = {nullptr, nullptr};
  double* adjK[2] = {nullptr, nullptr};
  double* adjTh = nullptr;
  double* adjSlots = nullptr;
  size_t adjSlotsN = 0;
  double* adjCkpt = nullptr;
  size_t adjCkptN = 0;
};

namespace tgh {
struct AdjPointArgs;
struct AdjGatherArgs;
struct AdjInitArgs;
struct AdjCopiesArgs;
// The adjoint's launchers (tgh_adjoint.hip).
void launch_adj_point(const tgh_problem* p, const AdjPointArgs& a);
void launch_adj_gather(const tgh_problem* p, const AdjGatherArgs& a);
void launch_adj_init(const tgh_problem* p, const AdjInitArgs& a);
void launch_adj_fold_copies(const tgh_problem* p, const AdjCopiesArgs& a);
void launch_adj_to_primaries(const tgh_problem* p, const AdjCopiesArgs& a);
dim3 xcd_grid(long total);
dim3 grid_points(const tgh_problem* p);
dim3 grid_brick(const tgh_problem* p);
dim3 grid_brick_tiles(const tgh_problem* p, int tx, int ty, int tz);
int mark(tgh_problem* p);  // a profiling event on the problem's stream

// One right-hand side on the block layout (EPI_STORE or EPI_RK4) and one RK4
// stage's right-hand side on the brick layout, at order Q; defined in
// tgh_launch.hpp, instantiated in tgh_q<Q>.hip.
template <int Q> void launch_rhs_blocks(tgh_problem* p, const RhsArgs& a, int epi);
template <int Q> void launch_rhs_brick(tgh_problem* p, const RhsArgs& a);
// the block layout's split form with the brick's head and LDS-tiled Pi kernel
template <int Q> void launch_rhs_blocks_lds(tgh_problem* p, const RhsArgs& a, const BlockBricks& bb);
extern template void launch_rhs_blocks_lds<2>(tgh_problem*, const RhsArgs&, const BlockBricks&);
extern template void launch_rhs_blocks_lds<4>(tgh_problem*, const RhsArgs&, const BlockBricks&);
extern template void launch_rhs_blocks_lds<6>(tgh_problem*, const RhsArgs&, const BlockBricks&);
extern template void launch_rhs_blocks_lds<8>(tgh_problem*, const RhsArgs&, const BlockBricks&);
extern template void launch_rhs_blocks<2>(tgh_problem*, const RhsArgs&, int);
extern template void launch_rhs_blocks<4>(tgh_problem*, const RhsArgs&, int);
extern template void launch_rhs_blocks<6>(tgh_problem*, const RhsArgs&, int);
extern template void launch_rhs_blocks<8>(tgh_problem*, const RhsArgs&, int);
extern template void launch_rhs_brick<2>(tgh_problem*, const RhsArgs&);
extern template void launch_rhs_brick<4>(tgh_problem*, const RhsArgs&);
extern template void launch_rhs_brick<6>(tgh_problem*, const RhsArgs&);
extern template void launch_rhs_brick<8>(tgh_problem*, const RhsArgs&);
// One RK4 stage of a batch of problems on the brick layout (the head and the
// Pi launch over the table's ranges), at order Q.
template <int Q>
void launch_batch(hipStream_t st, const BatchEntry* tab, int n, int head_total, int pi_total, int gauge, int hole);
// One RK4 stage of a batch on the block layout (the fused kernel over every
// block of every problem), at order Q.
template <int Q>
void launch_batch_blocks(hipStream_t st, const BatchEntry* tab, int n, int total, int gauge, int hole);
extern template void launch_batch_blocks<2>(hipStream_t, const BatchEntry*, int, int, int, int);
extern template void launch_batch_blocks<4>(hipStream_t, const BatchEntry*, int, int, int, int);
extern template void launch_batch_blocks<6>(hipStream_t, const BatchEntry*, int, int, int, int);
extern template void launch_batch_blocks<8>(hipStream_t, const BatchEntry*, int, int, int, int);
template <int Q>
void launch_batch_blocks_lds(hipStream_t st, const BatchEntry* tab, int n, int head_total, int pi_total, int gauge,
                             int hole);
extern template void launch_batch_blocks_lds<2>(hipStream_t, const BatchEntry*, int, int, int, int, int);
extern template void launch_batch_blocks_lds<4>(hipStream_t, const BatchEntry*, int, int, int, int, int);
extern template void launch_batch_blocks_lds<6>(hipStream_t, const BatchEntry*, int, int, int, int, int);
extern template void launch_batch_blocks_lds<8>(hipStream_t, const BatchEntry*, int, int, int, int, int);
extern template void launch_batch<2>(hipStream_t, const BatchEntry*, int, int, int, int, int);
extern template void launch_batch<4>(hipStream_t, const BatchEntry*, int, int, int, int, int);
extern template void launch_batch<6>(hipStream_t, const BatchEntry*, int, int, int, int, int);
extern template void launch_batch<8>(hipStream_t, const BatchEntry*, int, int, int, int, int);
}  // namespace tgh
