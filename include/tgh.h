/*
 * tgh.h — the C ABI of the MI355X generalized-harmonic right-hand side, its
 * RK4 step and its Kreiss–Oliger term (libtgh_mi355x.so).
 *
 * The library replaces the device paths of eschnett/TreeGeneralizedHarmonic.jl
 * behind the package's own verbs. Every entry point below names the reference
 * interface it stands in for (file:line in that repository). Arrays passed in
 * are DEVICE pointers (hipMalloc'd or a torch CUDA tensor's data_ptr()), FP64,
 * in the reference's state layout: column-major (i, j, k, v, b) with
 * (N+1)^3 owned points per block (closed vertex blocks), 20 variables
 * (h_ab then Pi_ab in the packed order tt,tx,ty,tz,xx,xy,xz,yy,yz,zz) and the
 * blocks in the order tgh_block_coords reports (Morton order of the brick).
 *
 * Every function returns TGH_OK or an error code; tgh_last_error() gives the
 * message of the last failure on the calling thread. All work is ordered on
 * the problem's HIP stream and every verb synchronises before returning.
 * Device memory is allocated by tgh_problem_create and tgh_batch_create,
 * by the first brick-layout step and the first constraint monitor, and by
 * tgh_transfer for its plan (freed before it returns).
 */
#ifndef TGH_H
#define TGH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  TGH_OK = 0,
  TGH_EINVAL = 1,      /* an argument the reference refuses (ArgumentError) */
  TGH_EDEVICE = 2,     /* a HIP runtime failure */
  TGH_ENOTFINITE = 3,  /* max_speed: non-finite evolved state or speed */
  TGH_EUNSUPPORTED = 4 /* a reference feature this library does not build */
};

enum { /* SpacetimeMetrics backgrounds, restated in closed form */
  TGH_BG_MINKOWSKI = 0,
  TGH_BG_GAUGE_WAVE = 1,        /* AwA gauge wave, H = 1 - A sin(2 pi (x - t)/d) */
  TGH_BG_KERR_SCHILD = 2,       /* KerrSchild(M, a), spin along +z, at center */
  TGH_BG_SHIFTED_MINKOWSKI = 3, /* t^ = t + A w tanh(x / w) */
  TGH_BG_HARMONIC = 4           /* Harmonic(M, a): Kerr in harmonic coordinates, spin along +z */
};
enum { /* gauge source kinds, src/gauge.jl:342-359 */
  TGH_GAUGE_NONE = 0,        /* harmonic background: H_a = 0 */
  TGH_GAUGE_SAMPLED = 1,     /* Hsrc sampled once at problem creation */
  TGH_GAUGE_KERR_SCHILD = 2  /* closed-form KerrSchildSource, in-kernel */
};
enum { TGH_DAMPING_CONSTANT = 0, TGH_DAMPING_GAUSSIAN = 1 }; /* src/gauge.jl:1094-1156 */
enum { /* interior variants, src/interior.jl:178-238 */
  TGH_INTERIOR_NONE = 0,
  TGH_INTERIOR_DAMPED = 1,
  TGH_INTERIOR_FROZEN = 2,
  TGH_INTERIOR_PASTED = 3,
  TGH_INTERIOR_FITTED = 4  /* :fitted: core and layer relax toward the target cache (tgh_set_target) */
};
enum { TGH_SHAPE_MAX_LMAX = 8 }; /* the tracked surface's largest series degree here */
enum { /* the Kreiss–Oliger amplitude (src/gauge.jl:1172-1264) */
  TGH_DISSIPATION_CONSTANT = 0,
  TGH_DISSIPATION_HORIZON = 1
};
enum { /* which block holds a point on a block face as its own (tgh_desc.ownership) */
  TGH_OWNERSHIP_CLOSED = 0,   /* TreeAMR 0.3, the reference's mesh today: a block owns both face
                                 planes; a shared point's primary is the finest leaf holding it as
                                 a vertex, and both outer planes of a Dirichlet box are evolved */
  TGH_OWNERSHIP_HALF_OPEN = 1 /* TreeAMR 0.1: a block owns [0, N) per axis; a point belongs to the
                                 leaf whose half-open extent holds it, and the upper outer plane of
                                 a Dirichlet axis is a ghost the hook fills. Forests only
                                 (tgh_problem_create_forest); the mesh the reference's published
                                 frozen-hierarchy tables and driver records were measured on */
};
enum { /* tgh_rk4 flags */
  TGH_RK4_GRAPH = 1,  /* capture the step in a hipGraph and replay it */
  TGH_RK4_FUSED = 2,  /* one kernel per right-hand side */
  TGH_RK4_BLOCKS = 4, /* run the stages on the per-block ghosted work arrays
                         (a ghost fill per stage) instead of the default brick
                         layout (every point once, a frame of ghosts around
                         the whole brick) */
  TGH_RK4_SPLIT = 8,  /* the right-hand side as two launches: the head (with
                         d_t h) and the ten Pi components (on the block layout
                         each block run as a brick, LDS-tiled, where the case
                         has no bounds, :pasted core, matter, sampled source
                         or general interior); without FUSED or SPLIT the
                         layout's faster form is used */
  TGH_RK4_PER_BLOCK = 16 /* the plain per-block form: the per-block work arrays
                         and the fused kernel launched once per block and
                         stage (implies BLOCKS and FUSED; the baseline the
                         batched forms are measured against) */
};

/* The case and its uniform forest: GHCase (src/initialdata.jl:158-311) plus
 * gh_forest (src/initialdata.jl:1370-1400) and GHProblem's q
 * (src/evolution.jl:1034-1149). */
typedef struct tgh_desc {
  int32_t float_bits;   /* 64: FP64 only; anything else is refused */
  int32_t q;            /* finite-difference order: 2, 4, 6 or 8; G = q/2 + 1 */
  int32_t N;            /* cells per block and axis */
  int32_t nblocks[3];   /* the brick of root blocks */
  int32_t periodic[3];  /* 1 = periodic, 0 = Dirichlet from the background */
  int32_t background;   /* TGH_BG_* */
  int32_t gauge;        /* TGH_GAUGE_* */
  int32_t damping;      /* TGH_DAMPING_* */
  int32_t interior;     /* TGH_INTERIOR_* */
  int32_t dissipation;  /* TGH_DISSIPATION_*: eps_ko everywhere, or the horizon profile */
  double lo[3], hi[3];  /* the box; cells must be cubic */
  double M, a;          /* hole mass and spin parameter */
  double center[3];     /* the hole's (static) center */
  double wave_amp, wave_len; /* gauge wave A, d; shifted Minkowski A, w */
  double gamma0;        /* constant damping rate */
  double gamma0_near, gamma0_far, gamma0_width; /* Gaussian profile */
  double gamma2;
  double eps_ko;        /* Kreiss–Oliger amplitude (eps_out of the horizon profile);
                           0 (and eps_in 0) compiles the term out */
  double r0, r1, rho_max, w_ramp, rho_ramp; /* the interior's layer */
  double eps_in, eps_r1, eps_rh; /* HorizonDissipation (src/gauge.jl:1192-1264): eps_in at
                                    and inside eps_r1 < eps_rh, eps_ko at and outside
                                    eps_rh, about center */
  /* StateBounds (src/bounds.jl:51-136): the range projection's ranges, applied at
   * r < r_gate from center; r_gate = 0 is a case without bounds (`nothing`). */
  double alpha_min, alpha_max, lambda_min, lambda_max, beta_max, K_max, r_gate;
  /* The hole's coordinate velocity v (HoleCenter, src/interior.jl:60-93):
   * with v != 0 a Kerr-Schild or harmonic background is
   * translate(boost(bg, -v), center) (kerr_schild_case / harmonic_kerr_case's
   * moving hole, src/initialdata.jl:1283-1298), the hole at center + v t;
   * every distance to the hole is measured from center + v t, and a :damped
   * core follows the hole (src/interior.jl:390-549). |v| < 1. */
  double velocity[3];
  /* FittedInterior's tracked geometry (src/interior.jl:1752-1990): with
   * shaped = 1 the layer's radii along the ray through x are r_1 = r_h(n) -
   * offset and r_0 = r_1 - thickness instead of r1 and r0, where r_h(n) is
   * the real spherical-harmonic series of shape (shape_lmax <= 8, the
   * (lmax + 1)^2 coefficients in real_harmonic_index order, l^2 + l + m)
   * clamped into [r_in, r_out], its bounding radii. */
  int32_t shaped;
  int32_t shape_lmax;
  double shape[(TGH_SHAPE_MAX_LMAX + 1) * (TGH_SHAPE_MAX_LMAX + 1)];
  double r_in, r_out, offset, thickness;
  int32_t ownership;    /* TGH_OWNERSHIP_*; 0 (closed blocks) when zero-initialised */
  /* Reflecting faces (GHCase's reflecting, src/initialdata.jl:30-60; the
   * mirrors of src/boundaries.jl:84-142): reflecting[d][0] / [d][1] nonzero
   * makes the low / high face of axis d a symmetry plane. Its ghosts are the
   * mirror image of the interior, each variable times its parity (component
   * ab of h and Pi is odd across x^d = 0 when the index d occurs an odd
   * number of times in ab), the wall plane itself is owned and evolved, and
   * at every fill its odd variables are set to zero; the Dirichlet hook
   * fills only the faces that neither reflect nor wrap. Not with periodic on
   * the same axis, nor with half-open ownership (the wall plane must be
   * owned); on uniform forests and on forests with levels, not in batches,
   * the adjoint or tgh_transfer. Zero-initialised: no mirror. */
  int32_t reflecting[3][2];
  /* The rotating seam (GHCase's rotating = (1, 2), src/boundaries.jl:145-205):
   * nonzero glues the low x face to the low y face by a quarter turn about
   * the z axis through (lo[0], lo[1]): the ghosts beyond x = lo[0] hold the
   * state a quarter turn away, u(R p) = Q u(p) with R (x, y) -> (-y, x) and Q
   * the tensor's signed permutation of the 20 variables; both seam planes
   * are owned and evolved. Needs x and y non-periodic and non-reflecting,
   * lo[0] = lo[1] and equal block counts on both. Zero: no seam. */
  int32_t rotating;
} tgh_desc;

/* Matter's 3+1 projections at every owned point (src/evolution.jl:1295-1311,
 * src/pointwise.jl:667-698): rho (N+1)^3 x nb, S (N+1)^3 x 3 x nb,
 * Sij (N+1)^3 x 6 x nb in (xx, xy, xz, yy, yz, zz) order; device pointers. */
typedef struct tgh_matter {
  const double* rho;
  const double* S;
  const double* Sij;
} tgh_matter;

typedef struct tgh_problem tgh_problem;
typedef struct tgh_batch tgh_batch;

/* GHProblem(U, schedule, case; q) — src/evolution.jl:1034-1149. Checks what
 * the reference's constructor checks (order, ghost width, interior radii
 * against the mesh) and allocates the work arrays on the current device. */
int tgh_problem_create(const tgh_desc* desc, tgh_problem** out);
/* GHProblem(fs, GhostSchedule(fs, Operators(prolongation = p, ...)), case; q)
 * on a frozen hierarchy — gh_forest(; refined = true) and hole_forest
 * (src/initialdata.jl:1370-1515) with TreeAMR's ghost schedule
 * (test/evolution_cases.jl:34-41). desc->nblocks is the brick of root
 * blocks; leaves holds nleaves quadruples (level, i, j, k), (i, j, k) the
 * leaf's coordinates in the brick of its level (nblocks_d 2^level blocks of
 * N cells of spacing h / 2^level); the state's blocks are the leaves in this
 * order. The leaves must tile the box, be 2:1 balanced across faces, edges,
 * corners and periodic seams, and N must be even and at least 2G.
 * prolongation is the coarse-to-fine interpolation order p (even, 2 <= p <=
 * q + 2; the interface-order rule of test/interface_tests.jl:64-96 asks for
 * q + 2). Restriction on a vertex-centred mesh is injection, whatever its
 * order, and takes no argument. Such a problem steps on the block layout,
 * and a batch it joins runs on the block layout. */
int tgh_problem_create_forest(const tgh_desc* desc, const int32_t* leaves, int32_t nleaves, int32_t prolongation,
                              tgh_problem** out);
/* The level of every block, 1 per block (0 on a uniform forest;
 * level(k) of TreeAMR's leaves). */
int tgh_block_levels(const tgh_problem* p, int32_t* out);
/* The regrid's state transfer (src/refinement.jl:973-1021: refine!/coarsen!
 * then the state carried over onto the new leaves): u_to (to's state, DEVICE)
 * takes at every owned point of to's forest the value that from's ghost
 * fill at time t gives there -- a copy of from's owner of the point (an
 * unchanged leaf, injection from a finer one) or from's order-prolongation
 * interpolation out of the leaf one level coarser that holds it (to's
 * order where from is uniform). Both cover the same root brick (N,
 * nblocks, periodic, lo, hi); a uniform problem is the forest of its level-0
 * blocks; to's leaves may be at most one level finer than from's where they
 * lie. from's work array is overwritten. Synchronous. */
int tgh_transfer(tgh_problem* from, const double* u_from, double t, tgh_problem* to, double* u_to);
/* GHProblem(...; target, t_target, target_rate, trail) (src/evolution.jl:
 * 1007-1021,1034-1037) for a :fitted interior: target is a DEVICE array of
 * 40 variables per owned point in the state layout ((N+1)^3 x 40 x nblocks:
 * A_v then S_v, v = 1..20), read as A + (t - t_target) S
 * (_cached_target, src/fit.jl:1050-1055) and, with rate != 0, S as the
 * target's rate (_cached_rate); trail narrows the rho ramp on the trailing
 * side by 1 - trail max(0, -n.v/|v|) (_trail_profiles,
 * src/evolution.jl:1256-1272). The library keeps the pointer (the caller owns
 * the array and may refill it between calls); NULL detaches it. */
int tgh_set_target(tgh_problem* p, const double* target, double t_target, int32_t rate, double trail);
/* A fitted target as the cache's fill reads it: FitParams (src/fit.jl:492-502)
 * and InteriorFit's coefficients and time (src/fit.jl:504-539). The fit is
 *   f_v(x) = sum_lm S_lm(xi) sum_{k=0..cont} C[lm, k, v] |xi|^{2k},
 *   xi = (x - c(t)) / rbar,  c(t) = center + velocity t,
 * S_lm the real solid harmonics (shape_series' normalisation, slots l^2 + l + m),
 * for the twenty fit variables (log alpha, beta^i, gamma_ij - delta_ij, Pi_ab
 * or, with tilde, (alpha / sqrt(det gamma)) Pi_ab); fit_state reassembles the
 * packed state from them and projects it into the bounds (src/fit.jl:182-212,
 * 749-755). */
typedef struct tgh_fit {
  int32_t lmax;       /* degree L of the harmonics, 0..32 */
  int32_t cont;       /* radial order, 0, 1 or 2 */
  int32_t tilde;      /* nonzero: the momentum slots hold (alpha / sqrt(det gamma)) Pi */
  int32_t pad;
  double rbar;        /* the radius the ansatz is scaled by (> 0) */
  double center[3];   /* HoleCenter: c(t) = center + velocity t */
  double velocity[3];
  double alpha_min, alpha_max, lambda_min, lambda_max, beta_max, K_max; /* StateBounds */
  double t;           /* the time the fit was built at */
  const double* coeffs; /* DEVICE array of (L+1)^2 x (cont+1) x 20 doubles, column-major
                           (slot fastest, then k, then the variable) */
} tgh_fit;
/* fill_target!(target, origins, spacings, interior, (latest, previous), t;
 * rate, clear) with fit_target_kernel! (src/fit.jl:1070-1151) on a problem
 * whose interior is the tracked geometry (tgh_desc.shaped = 1): at every
 * owned point with |x - c(t)| < r_fill = (r_out - offset) + h (c the
 * interior's center at t, h the geometry's spacing, FittedInterior's h), the
 * 40 slots of target ((N+1)^3 x 40 x nblocks, the layout tgh_set_target
 * reads) take A = F_a + (t - t_a) S and S = (F_a - F_b) / (t_a - t_b), F the
 * fits' states at t (S = 0 without an older previous fit: previous NULL or
 * previous->t >= latest->t); with rate != 0 S also carries the latest fit's
 * translation, the centred difference of F_a over t +- h / (8 |v|). With
 * clear != 0 every other point is zeroed, else left as it is. target is a
 * DEVICE array. Synchronous. */
int tgh_fill_target(tgh_problem* p, double* target, const tgh_fit* latest, const tgh_fit* previous, double t,
                    double h, int32_t rate, int32_t clear);
/* gh_integrator(p, u, tspan; matter) — src/stepping.jl:150-154,208-225
 * (MatterRHS) with the in-kernel adapter of src/evolution.jl:1295-1311: the
 * matter fields every right-hand side tgh_rk4 evaluates adds to the source,
 * held fixed over the steps (device arrays owned by the caller, laid out as
 * tgh_matter says; NULL or rho NULL: vacuum again). A problem with matter
 * steps on the block layout. */
int tgh_set_matter(tgh_problem* p, const tgh_matter* m);
/* with_interior(p, with_rho_max(p.interior, rho_max)) — src/evolution.jl
 * with_interior, src/interior.jl with_ρ_max; what evolve!'s chunk loop does
 * per chunk (chunk_interior, src/driver.jl:2391-2411): the layer's
 * relaxation rate, the rest of the problem unchanged. */
int tgh_set_rho_max(tgh_problem* p, double rho_max);
/* No reference counterpart: releases the problem's device memory. */
int tgh_problem_destroy(tgh_problem* p);

/* Doubles in one state vector: (N+1)^3 * 20 * nblocks (statevector(fs)). */
int64_t tgh_state_length(const tgh_problem* p);
/* Block b's integer coordinates in the brick, 3 per block (block_corners,
 * src/evolution.jl:120-130). */
int tgh_block_coords(const tgh_problem* p, int32_t* out);
/* Order the problem's work on a caller's hipStream_t from now on (NULL = the
 * device's null stream), e.g. the stream a framework's tensors are produced
 * on; until then the problem uses a stream of its own. */
int tgh_set_stream(tgh_problem* p, void* stream);

/* prepare!(p, u, t) — src/evolution.jl:1349-1357: scatter u into the work
 * array and fill the ghosts (periodic copies, closed-block face copies and the
 * Dirichlet hook of src/boundaries.jl:49-63 at this t). u is not written. */
int tgh_prepare(tgh_problem* p, const double* u, double t);
/* rhs!(du, p, t; matter) — src/evolution.jl:1386-1390: the fused kernel
 * gh_rhs_kernel! (src/evolution.jl:846-863) at every owned point of the work
 * array prepare left. matter may be NULL (vacuum). */
int tgh_rhs(tgh_problem* p, double* du, double t, const tgh_matter* matter);
/* gh_rhs!(du, u, p, t; matter) — src/evolution.jl:1327-1331. */
int tgh_gh_rhs(tgh_problem* p, double* du, const double* u, double t,
               const tgh_matter* matter);
/* settle!(p, u, t) — src/evolution.jl:1370-1374: prepare, then gather back. */
int tgh_settle(tgh_problem* p, double* u, double t);
/* stage_limit!(u, p, t) / step_limit!(u, p, t) — src/stepping.jl:91-97: the
 * range projection where the case has bounds (src/bounds.jl:780-800), then
 * the :pasted overwrite (src/evolution.jl:888-905); a no-op for a case with
 * neither. */
int tgh_stage_limit(tgh_problem* p, double* u, double t);
int tgh_step_limit(tgh_problem* p, double* u, double t);
/* apply_bounds!(p, u, t) — src/bounds.jl:733-770: the range projection
 * (bounds_project, :510-589) at every owned point with r < r_gate, written back
 * only where it fired; the number of points moved, of those with a non-finite
 * component, and the outermost radius it fired at (-1: it did not). A no-op
 * returning 0 for a case without bounds. tgh_stage_limit and tgh_step_limit
 * apply it first, then the :pasted overwrite; tgh_rk4 applies it to every
 * stage value and to the result. */
int tgh_apply_bounds(tgh_problem* p, double* u, double t, int64_t* moved, int64_t* nonfinite, double* r_max);
/* max_speed(p, u, t) — src/evolution.jl:1539-1560: the largest
 * alpha sqrt(tr gamma^ij) + |beta| over evolved points, checked:
 * TGH_ENOTFINITE if an evolved value or the speed is not finite and positive. */
int tgh_max_speed(tgh_problem* p, const double* u, double t, double* out);
/* initial_state!(u, p, t) — src/stepping.jl:110-115: the case's exact state
 * with the core rule, then the limiter. */
int tgh_initial_state(tgh_problem* p, double* u, double t);
/* add_noise!(U, rng; amplitude) — src/initialdata.jl:1961-2013 with the
 * SplitMix64 point draw; base is the draw's seed word. */
int tgh_add_noise(tgh_problem* p, double* u, uint64_t base, double amplitude);
/* gh_solve / IRK.step! x nsteps — src/stepping.jl:208-239: classical RK4 with
 * the limiter on every stage value and on the result, fixed dt, in place.
 * The stage update is fused into the right-hand side's epilogue. The stages
 * run on a copy of u in the brick layout (each mesh point stored once) unless
 * TGH_RK4_BLOCKS is set; either way every owned point of u, face copies
 * included, advances as the reference steps its state vector (a copy by its
 * primary's increment). The first call in the brick layout allocates its
 * arrays (4 x 20 doubles per point of the ghosted brick, and 24 more with
 * TGH_RK4_SPLIT). */
int tgh_rk4(tgh_problem* p, double* u, double t0, double dt, int32_t nsteps, int32_t flags);
/* gh_constraint!(p, u, t) then constraint_norms(p)'s gauge rows
 * (src/constraints.jl:211-309,559-566,848-876): prepares u at t and evaluates
 * the gauge constraint C_a = g_ab (Gamma^b + g^{bc} H_c) at every owned
 * point, d_t h without the dissipation, H the case's source; l2 and linf get
 * the masked norms of the four components over the evolved points
 * (4 doubles each, host memory), counted as tgh_error_norms counts them. */
int tgh_gauge_constraint(tgh_problem* p, const double* u, double t, double* l2, double* linf);
/* No counterpart in the reference: tgh_gauge_constraint for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc). The same kernels and the same fold
 * order, so the same norms to the bit; the per-workgroup partials and their
 * fold live in the library's static device array (8 MiB on each device it
 * runs on, held until the process ends) instead of memory the problem
 * allocates on its first monitor call. Calls no hipMalloc or hipFree. Limit:
 * 116 506 workgroup partials, i.e. 5825 blocks of 17^3 points (N = 16); a
 * larger problem gets TGH_EINVAL. The problem's device must be current.
 * Calls that overlap on one device share the array and race. */
int tgh_gauge_constraint_noalloc(tgh_problem* p, const double* u, double t, double* l2, double* linf);
/* adm_constraint!(p, u, t) then constraint_norms(p)'s Hamiltonian and
 * momentum rows (src/constraints.jl:100-210,432-534,584-605,848-876): the
 * ADM constraints H = 2 G_nn and M_i = -G_n i at every owned point, from the
 * state's first and second derivatives and the time derivatives the
 * evolution equations give (no dissipation); l2 and linf get the masked
 * norms of (H, M_x, M_y, M_z) (4 doubles each, host memory), counted as
 * tgh_error_norms counts them. Vacuum only (no matter adapter). */
int tgh_adm_constraint(tgh_problem* p, const double* u, double t, double* l2, double* linf);
/* gh_tau!(...; scale, eps) then each block's maximum (src/refinement.jl:
 * 152-219,312-378): prepares u at t and writes into tau_blocks (nblocks
 * doubles, host memory) each block's largest Loehner indicator over its
 * evolved owned points, the ten h components and the three axes, with the
 * noise floor 4 eps |scale|; tau_max is their maximum. The regrid that reads
 * it is not built. */
int tgh_refinement_indicator(tgh_problem* p, const double* u, double t, double scale, double eps,
                             double* tau_blocks);
/* field_scale(U, mask, ...) (src/refinement.jl:221-300): the largest |h_v|
 * over the evolved owned points, the indicator's reference amplitude. */
int tgh_field_scale(tgh_problem* p, const double* u, double t, double* scale);
/* No reference counterpart (the reference's kernels take no launch shape):
 * the autotuner of the brick stepper's tile shapes. Steps u three times per
 * candidate shape of the Pi kernel's LDS tile (TGH_PI_TILES of them; every
 * shape computes the same bits, so u is advanced 3 TGH_PI_TILES steps from
 * t0 as the default form would), times each shape's Pi launches with
 * hipEvents, writes the average ms per launch into ms (TGH_PI_TILES doubles,
 * or NULL) and keeps the fastest for the problem's later tgh_rk4 calls.
 * Pass a scratch copy of the state. TGH_EUNSUPPORTED on the block layout. */
enum { TGH_PI_TILES = 6 };
int tgh_autotune(tgh_problem* p, double* u, double t0, double dt, double* ms);
/* No reference counterpart: pins the Pi kernel's tile shape (0 .. TGH_PI_TILES - 1;
 * 0 is the default). */
int tgh_set_tile(tgh_problem* p, int32_t tile);
/* Several independent problems stepped together (the batched form: holes of
 * uneven block counts in one launch per kernel, driven by a device-side table
 * of each problem's ranges and arguments). The reference steps each problem
 * by its own gh_solve (src/stepping.jl:208-239); a problem in a batch computes
 * the bits tgh_rk4 computes for it alone in the brick layout's split form.
 * With a forest with levels or a tracked or :fitted interior among them the
 * batch runs on the block layout instead (one head and one Pi launch over
 * every block of every problem per stage, each problem computing the bits of
 * tgh_rk4 with TGH_RK4_BLOCKS | TGH_RK4_SPLIT; with a range projection or a
 * general interior among them one fused launch, the bits of TGH_RK4_BLOCKS |
 * TGH_RK4_FUSED), and moving holes are refused there
 * (TGH_EUNSUPPORTED). The problems share the order q, the gauge-source kind
 * and the kind of interior (one kernel specialisation), and none is :pasted;
 * each keeps its own arrays (brick arrays allocated here if missing). The
 * batch holds the problems, which must outlive it. TGH_EINVAL on a mismatch. */
int tgh_batch_create(tgh_problem* const* probs, int32_t n, tgh_batch** out);
/* gh_solve for every problem of the batch over the same (t0, dt, nsteps):
 * u[i] is problem i's state (in place). flags: 0 or TGH_RK4_GRAPH. */
int tgh_batch_rk4(tgh_batch* b, double* const* u, double t0, double dt, int32_t nsteps, int32_t flags);
int tgh_batch_destroy(tgh_batch* b);
/* gh_error! + error_norms — src/constraints.jl:311-396,629-671: at every
 * primary the error e = |u - u_exact(t)|, the 2-norm over the 20 variables;
 * the masked volume-weighted L2 of e and its L-inf over evolved primaries. */
int tgh_error_norms(tgh_problem* p, const double* u, double t, double* l2, double* linf);
/* gh_error!(p, u, t; shell = (shell_lo, shell_hi)) + error_norms —
 * src/constraints.jl:311-396,629-671: out[4] (host) = err_l2, err_linf (as
 * tgh_error_norms), residual (the layer's L-inf of e at r_0 <= r < r_1; for
 * :fitted the distance from the target cache A + (t - t_f) S) and drift (the
 * L-inf of |h_tt - h_tt,exact| at evolved primaries with shell_lo <= r <=
 * shell_hi; 0 for the default empty band shell_lo = 0 > shell_hi = -1). */
int tgh_error_rows(tgh_problem* p, const double* u, double t, double shell_lo, double shell_hi, double* out);
/* The adjoint. The reference differentiates nothing; these entry points give
 * the vector-Jacobian product of its right-hand side (gh_rhs!,
 * src/evolution.jl:1327) and the gradient of its masked L2 error
 * (gh_error! + error_norms, src/constraints.jl:311-396,629-671) after RK4
 * (src/stepping.jl:208-239) with respect to the parameters
 *   theta = (gamma0_near, gamma0_far, gamma2, rho_max, rho_ramp)
 * of GaussianDamping (src/gauge.jl:1094-1156; 0 for ConstantDamping) and of
 * the interior layer (src/interior.jl:178-238). Built for a uniform forest
 * with a hole at rest (or none), the :damped or :frozen sphere, no range
 * bounds, a static gauge-source kind; TGH_EUNSUPPORTED otherwise. */
/* u_bar = (dF/du)^T du_bar and theta_bar = (dF/dtheta)^T du_bar at (u, t),
 * F = gh_rhs!(., u, p, t): u, du_bar and u_bar are state vectors (device);
 * du_bar's face copies each count (the copies of a point share its F),
 * u_bar's copies are zero (the fill overwrites them); theta_bar has 5
 * doubles (host). */
int tgh_vjp_rhs(tgh_problem* p, const double* u, double t, const double* du_bar, double* u_bar, double* theta_bar);
/* From u (device state, in: u_0 at t0; out: u_T) nsteps RK4 steps of dt
 * (gh_solve's stepping, checkpointed every `every` steps), J = the masked L2
 * error of u_T against the exact state at T = t0 + nsteps dt, and grad[5] =
 * dJ/dtheta (host), the stages of each segment recomputed from its
 * checkpoint in the reverse sweep. ms[2] (host, or NULL): the forward sweep's
 * and the reverse sweep's device time. */
int tgh_adjoint(tgh_problem* p, double* u, double t0, double dt, int32_t nsteps, int32_t every, double* J,
                double* grad, double* ms);
/* Doubles in the problem's work array: (N+1+2G)^3 * 20 * nblocks. */
int64_t tgh_work_length(const tgh_problem* p);
/* Copy the work array prepare filled (ghosts included) into a caller's
 * device buffer of tgh_work_length doubles, for inspection; the layout is the
 * reference's ghosted FieldSet work array. */
int tgh_copy_work(tgh_problem* p, double* dst);
/* No reference counterpart: per-kernel timing for the benchmark. While
 * enabled, tgh_rk4 without TGH_RK4_GRAPH brackets every kernel launch with
 * hipEvents on the problem's stream; tgh_profile_read gives out[0..5] =
 * summed milliseconds and count of the right-hand side's first launch (the
 * head, or the whole fused kernel), of the ghost fills, and of the
 * Pi-component launches, since the last tgh_profile(p, 1). */
int tgh_profile(tgh_problem* p, int32_t enable);
int tgh_profile_read(tgh_problem* p, double* out);
/* Message of the last failure on this thread ("" if none). */
const char* tgh_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* TGH_H */
