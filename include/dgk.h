/* libdgk: the DG time-stepping hot path of povanberg/DGFEM-Acoustic on gfx950.
 *
 * Drop-in boundary for the reference's in-process solver API:
 *   dgk_create       Mesh::Mesh connectivity/geometry/BC setup (src/Mesh.cpp:20-435)
 *                    + Mesh::precomputeMassMatrix (src/Mesh.cpp:440-466)
 *   dgk_rk4          the time-step body of solver::rungeKutta (src/solver.cpp:247-285):
 *                    hard source overwrite, then classic RK4 with numStep(beta=0) stages
 *   dgk_euler        the time-step body of solver::forwardEuler (src/solver.cpp:139-153)
 *   dgk_rhs          one Mesh::updateFlux + solver::numStep(beta=0)
 *                    (src/Mesh.cpp:569-674, src/solver.cpp:35-52): out = dt*Minv*(S-F)
 *   dgk_set_state /
 *   dgk_get_state    the caller-owned state u[4][K*Np] (src/dgalerkin.cpp:36), same layout:
 *                    equation-major (p, vx, vy, vz), then element, then Gmsh element node
 * Conventions: plain C types only; every call returns a dgk_status; the library owns
 * device memory, the caller owns host buffers; calls on one context are ordered on the
 * context's HIP stream and dgk_get_state / dgk_sync block until prior work is done.
 * One context must not be used by two host threads at once.
 */
#ifndef DGK_H
#define DGK_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  DGK_OK = 0,
  DGK_ERR_INVALID = 1,      /* bad argument */
  DGK_ERR_UNSUPPORTED = 2,  /* element type/order or geometry the kernels do not handle */
  DGK_ERR_MESH = 3,         /* inconsistent mesh (non-conforming, non-manifold, inverted) */
  DGK_ERR_DEVICE = 4,       /* HIP runtime failure or missing GPU */
  DGK_ERR_NONFINITE = 5     /* state check found NaN/Inf */
} dgk_status;

enum { DGK_FP64 = 0, DGK_FP32 = 1 };
/* Interior penalty sign. REFERENCE reproduces the reference's fc * orientation rule
 * (src/Mesh.cpp:210-211, 278-294, 519-527) with the normal sign taken at the first point of
 * the face quadrature rule (Gauss-Legendre, order+2 points, collapsed on triangles).
 * STABLE uses the dissipative (upwind-consistent) sign on every interior face. */
enum { DGK_SIGN_REFERENCE = 0, DGK_SIGN_STABLE = 1 };

typedef struct {
  int dim;                   /* element dimension 1..3 */
  int order;                 /* Lagrange order 1..5 */
  int64_t num_elements;      /* K */
  const double* node_xyz;    /* [K*Np*3] element node coordinates, Gmsh node order */
  const int64_t* node_tags;  /* [K*Np] node tags; conforming neighbours share tags */
  const uint8_t* node_bc;    /* [K*Np] 1 reflecting / 0 absorbing; a boundary face takes
                                the code of its first node (src/Mesh.cpp:364-384). NULL = all absorbing */
} dgk_mesh;

typedef struct {
  double c0, rho0;
  double v0[3];
} dgk_phys;

typedef struct {
  int precision;   /* DGK_FP64 (drop-in) or DGK_FP32 (FP32 arithmetic, FP64 geometry rounded once) */
  int flux_sign;   /* DGK_SIGN_REFERENCE or DGK_SIGN_STABLE */
  int device;      /* HIP device ordinal */
  int variant;     /* stage-kernel variant, 0 = autotuned/default (see dgk_variant_name) */
  int fp32_flags;  /* DGK_FP32 only: DGK_FP32_* bits, 0 = the default FP32 form */
  int addr64;      /* 1: the stage kernel addresses every array through 64-bit per-tile bases even
                      when all fit 32-bit offsets (0: only where an array reaches 4 GiB) */
  int reserved[4];
} dgk_opts;

/* fp32_flags bits. By default the FP32 form keeps the solution u as an FP32 value plus a bf16
 * low word (compensated: the RK4 update u += dt/6 (k1+2k2+2k3+k4) is summed exactly), which
 * removes the step-by-step accumulation of FP32 rounding in u. PLAIN_STATE turns that off
 * (plain FP32 u). OPERATOR_CORRECTION adds a bf16 residual of the FP32-rounded operator
 * matrices on MFMA (about 6% slower, small further gain). */
enum { DGK_FP32_PLAIN_STATE = 1, DGK_FP32_OPERATOR_CORRECTION = 2 };

typedef struct {
  int dim, order, Np, Nf, Nfp;
  int64_t num_elements, num_faces, num_boundary_faces;
  int precision, variant;
  double fc;              /* the reference's flux coefficient for this element type */
  int64_t device_bytes;   /* device memory held by the context */
} dgk_info;

typedef struct dgk_ctx dgk_ctx;

/* Context lifecycle. On failure *out is NULL and dgk_last_error(NULL) describes it. */
dgk_status dgk_create(const dgk_mesh* mesh, const dgk_phys* phys, const dgk_opts* opts, dgk_ctx** out);
dgk_status dgk_destroy(dgk_ctx* ctx);
const char* dgk_last_error(const dgk_ctx* ctx);
dgk_status dgk_info_get(const dgk_ctx* ctx, dgk_info* info);

/* State transfer, u[4][K*Np] FP64 host arrays. */
dgk_status dgk_set_state(dgk_ctx* ctx, const double* u);
dgk_status dgk_get_state(dgk_ctx* ctx, double* u);

/* Point sources (src/configParser.cpp:65-102): n rows of 9 doubles
 * {pole, x, y, z, radius, amp, freq, phase, duration}. At the start of every step with
 * t < duration, p is overwritten with amp*sin(2*pi*freq*t + phase) at the element nodes
 * strictly inside the radius (src/solver.cpp:197-210, 248-256); later rows win where
 * sources overlap. node_xyz [K*Np*3] are the element node coordinates (as in dgk_mesh). */
dgk_status dgk_set_sources(dgk_ctx* ctx, int n, const double* src, const double* node_xyz);

/* Advance nsteps RK4 (or forward-Euler) steps of size dt. t_start[i] is the time at the
 * start of step i (used by the sources; may be NULL when none are set). Asynchronous. */
dgk_status dgk_rk4(dgk_ctx* ctx, int nsteps, double dt, const double* t_start);
dgk_status dgk_euler(dgk_ctx* ctx, int nsteps, double dt, const double* t_start);

/* out = dt * L(u): one updateFlux + numStep(beta = 0), host arrays [4][K*Np]. */
dgk_status dgk_rhs(dgk_ctx* ctx, const double* u, double* out, double dt);

dgk_status dgk_sync(dgk_ctx* ctx);
/* Returns DGK_ERR_NONFINITE if the device state holds NaN/Inf. */
dgk_status dgk_check_finite(dgk_ctx* ctx);

/* Face connectivity built on the device, per element face e*Nf+j:
 *   nbr[e*Nf+j]   neighbour element, or -1 (absorbing boundary) / -2 (reflecting boundary)
 *   nface[e*Nf+j] neighbour's local face (-1 on the boundary)
 *   perm[e*Nf+j]  index of the vertex permutation taking this face's vertex order to the
 *                 neighbour's (-1 on the boundary)
 *   sign[e*Nf+j]  interior penalty sign (+1/-1; 0 on the boundary) */
/* Receivers: pressure at n points (xyz[3n]), interpolated with the Lagrange basis of the element
 * in which the point is most interior (DG fields jump across faces; on a face either side may be
 * taken) and recorded after every time step (rk4 and euler), up to `capacity` steps. node_xyz is the mesh's
 * element-node coordinate array (as in dgk_set_sources). Setting receivers clears the record.
 * No reference counterpart in the solver (the reference writes whole-field views only). */
dgk_status dgk_set_receivers(dgk_ctx* ctx, int n, const double* xyz, const double* node_xyz, int capacity);
/* rows = steps recorded; out (may be NULL to query rows) receives rows x n values, row-major */
dgk_status dgk_get_receivers(dgk_ctx* ctx, double* out, int capacity_rows, int* rows);
/* Launch shape of the stage kernel (no reference counterpart; tuning): the persistent grid as a
 * percentage of the occupancy-derived grid (100 = one full wave of resident workgroups) and the
 * tile order (0 = XCD-aware slabs, 1 = round-robin). Results do not depend on either. */
dgk_status dgk_set_launch(dgk_ctx* ctx, int grid_pct, int tile_order);
/* Tile shape of the MFMA stage kernel (no reference counterpart; tuning): 0 the built-in choice
 * per (precision, dim, order); 1 4-wave workgroups over 16-element tiles; 2 8-wave workgroups over
 * 32-element tiles; 3 the row split (4 waves over a 4-element tile, each wave a quarter of the
 * operator's 16-row blocks, the operator in LDS); 4 4-wave workgroups over 16-element tiles with
 * the operator read from global memory; 5 the row split over 8-wave workgroups and 8-element tiles
 * (two element groups, two waves per SIMD); 6, 7, 8 4-wave workgroups over 16-element tiles compiled
 * for 3, 4 or 5 waves per SIMD (at most 512/n registers per lane; where that many workgroups' LDS
 * fits a CU); 9, 10, 11 as 7, 8, 6 with the RK epilogue operands loaded after the volume k-steps.
 * Shapes 1-11 exist for FP64 and the default FP32 form in
 * 2D and 3D at orders 2-5, for buffers below 4 GiB; a shape that has no kernel for the context, or
 * does not fit its element type, returns DGK_ERR_UNSUPPORTED and leaves the setting unchanged. If
 * the context's form later changes to one without that shape (a 2D state with nonzero vz), the
 * built-in shape runs. Results agree with the built-in shape to rounding: bitwise between
 * element-split shapes compiled for the same waves per SIMD; the row split sums the k-steps in
 * another order, and another register budget may fuse a product into a multiply-add differently. */
dgk_status dgk_set_tile_shape(dgk_ctx* ctx, int shape);
/* Capacity planning (no reference counterpart; no device use): *addr64 = 1 when a context of
 * num_elements elements of this dim/order/precision addresses its arrays through 64-bit per-tile
 * bases (some array of the stage kernel -- state, J^-1 or face geometry -- reaches 4 GiB; the
 * same arithmetic, bitwise equal results), 0 when 32-bit offsets suffice. Computed for the state's
 * initial live fields (dim+1 in 1D/2D, 4 in 3D). DGK_ERR_UNSUPPORTED beyond 2^31 element faces. */
dgk_status dgk_stage_addressing(int dim, int order, int precision, int64_t num_elements, int* addr64);
/* Load a tuning table (text lines "dim order precision variant grid_pct tile_order [tile_shape
 * stage_us | stage_us]", '#' comments), applied to contexts created afterwards: launch shape
 * always, tile shape where the context has a kernel of that shape, variant when
 * dgk_opts.variant is 0. Replaces any table loaded before. */
dgk_status dgk_load_tuning(const char* path);
dgk_status dgk_connectivity(dgk_ctx* ctx, int32_t* nbr, int32_t* nface, int32_t* perm, int32_t* sign);

/* Timed run for measurement: nsteps RK4 steps with a HIP event pair around every stage
 * launch on the context's stream. Writes the summed stage-kernel time and the wall time of
 * the whole sequence (both ms, device clock). */
dgk_status dgk_rk4_timed(dgk_ctx* ctx, int nsteps, double dt, double* stage_ms, double* total_ms);

const char* dgk_variant_name(int variant);

/* ---- Batched ensembles -------------------------------------------------------------------
 * Many independent simulations (any mix of dimensions, orders and sizes) advanced together.
 * No reference counterpart: the reference runs one simulation per process. Simulations with the
 * same (dim, order, physics, flux sign) share one device context built on their disjoint union;
 * each simulation keeps its own time step (folded into its elements' geometry), so one stage
 * launch advances all of them. dgk_batch_rk4 captures the groups' stage launches as parallel
 * branches of one hipGraph. States use the single-simulation layout u[4][K_i*Np_i]. */
typedef struct dgk_batch dgk_batch;
dgk_status dgk_batch_create(int n, const dgk_mesh* meshes, const dgk_phys* phys, const double* dt,
                            const dgk_opts* opts, dgk_batch** out);
dgk_status dgk_batch_destroy(dgk_batch* b);
const char* dgk_batch_last_error(const dgk_batch* b);
dgk_status dgk_batch_set_state(dgk_batch* b, int sim, const double* u);
dgk_status dgk_batch_get_state(dgk_batch* b, int sim, double* u);
/* schedule: 0 = groups as parallel graph branches, 1 = groups in sequence on one stream */
dgk_status dgk_batch_rk4(dgk_batch* b, int nsteps, int schedule);
/* receivers of member sim (node_xyz: that member's mesh), recorded after every following step
 * for at least `capacity` steps. Members of one group share one record: setting a member's
 * receivers restarts that member's series (dgk_batch_get_receivers returns the steps since) and
 * keeps the rows already recorded for the other members. */
dgk_status dgk_batch_set_receivers(dgk_batch* b, int sim, int n, const double* xyz, const double* node_xyz,
                                   int capacity);
dgk_status dgk_batch_get_receivers(dgk_batch* b, int sim, double* out, int capacity_rows, int* rows);
/* launch shape of one group's stage kernel (see dgk_set_launch) */
dgk_status dgk_batch_set_launch(dgk_batch* b, int group, int grid_pct, int tile_order);
/* tile shape of one group's stage kernel (see dgk_set_tile_shape) */
dgk_status dgk_batch_set_tile_shape(dgk_batch* b, int group, int shape);
/* capture (and keep) the graphs that dgk_batch_rk4(b, nsteps, schedule) will replay (up to 100 steps
 * each, the remainder in a graph of its own length), without running them */
dgk_status dgk_batch_prepare(dgk_batch* b, int nsteps, int schedule);
/* nsteps RK4 steps group by group without a graph, with a HIP event pair around every stage
 * launch; group_stage_ms[g] receives the summed stage-kernel time of group g */
dgk_status dgk_batch_rk4_timed(dgk_batch* b, int nsteps, double* group_stage_ms);
dgk_status dgk_batch_sync(dgk_batch* b);
/* number of groups; per group g (arrays of length >= groups, any may be NULL): dim[g], order[g],
 * elements[g] (sum of K_i), sims[g], faces[g] (unique faces), boundary_faces[g] */
dgk_status dgk_batch_groups(const dgk_batch* b, int* groups, int* dim, int* order, int64_t* elements, int* sims,
                            int64_t* faces, int64_t* boundary_faces);

#ifdef __cplusplus
}
#endif
#endif
