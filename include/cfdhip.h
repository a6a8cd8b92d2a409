/* C ABI of libcfdhip.so: the gfx950 implementation of tackandr/CFD-demo's
 * degree-4 nodal DG Euler hot path (RHS, limiting, CFL step, SSP-RK3 step and
 * the run loop).
 *
 * Every entry point returns an int status (CFD_OK on success); the text of
 * the last error on the calling thread is cfd_last_error(). No C++ exception
 * crosses this boundary and no function frees memory it did not allocate.
 *
 * State arrays use the reference layout: float64, shape (Np=15, K, 4),
 * C order, conserved variables (rho, rho*u, rho*v, E) on the last axis.
 * Functions named *_device take HBM pointers (e.g. torch data_ptr()); all
 * others take host pointers and copy at the call edge.
 *
 * A context owns its HBM tables and buffers and one HIP stream. Contexts are
 * not thread-safe: use one context per host thread.
 *
 * The reference binds four jitted callables at src/cfd_demo/solver.py:118-125
 * (_rhs, _limit, _time_step, _step) and drives them from EulerSolver.run
 * (solver.py:167-216); the functions below replace exactly those.
 */
#ifndef CFDHIP_H
#define CFDHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  CFD_OK = 0,
  CFD_EINVAL = 1,        /* bad argument -> ValueError */
  CFD_EHIP = 2,          /* HIP runtime error */
  CFD_ENOMEM = 3,        /* device allocation failed */
  CFD_ENOTINIT = 4,      /* cfd_init was not called */
  CFD_EINITIAL = 10,     /* run: initial state not admissible (solver.py:171-174) */
  CFD_EINVALID_DT = 11,  /* run: non-finite or non-positive dt (solver.py:182-183) */
  CFD_EREJECTED = 12     /* run: max_rejections+1 rejections (solver.py:195-197) */
};

enum { CFD_FLUX_RUSANOV = 0, CFD_FLUX_ROE = 1, CFD_FLUX_HLLC = 2 };
enum { CFD_LIM_MINMOD = 0, CFD_LIM_MC = 1, CFD_LIM_VAN_LEER = 2, CFD_LIM_SUPERBEE = 3, CFD_LIM_UNLIMITED = 4 };

/* Degree-4 reference-element tables (reference_element.py:264-316 and
 * limiting.py:97-121); all matrices row-major. */
typedef struct {
  const double* dr;             /* 15x15 */
  const double* ds;             /* 15x15 */
  const double* lift;           /* 15x15, columns = face slots 5*f+k */
  const double* cell_weights;   /* 15 */
  const double* average_weights;/* 15: cell_weights / sum(cell_weights) (limiting.py:107) */
  const double* face_weights;   /* 15, GLL weight of each face slot */
  const double* p1_midpoint;    /* 3x15 */
  const double* linear_basis;   /* 15x3 */
  const int32_t* face_nodes;    /* 15 */
  const int32_t* midpoint_nodes;/* 3 */
} CfdReferenceElement;

/* Compact per-element mesh tables replacing DGOperators (dg.py:30-49) and
 * LimiterOperators (limiting.py:34-48).
 *  geo  (K x 16): rx ry sx sy | nx0 nx1 nx2 | ny0 ny1 ny2 | fscale0..2 | 0.5|e|0..2
 *  conn (K x 4 int32): field f<3 is the neighbour element across face f, or the
 *       fixed-state index on a FIXED face; field 3 packs one byte per face:
 *       bits 0-1 neighbour's local face, bit 2 reversed node order,
 *       bits 3-5 kind (0 interior, 1 slip wall, 2 fixed state, 3 transmissive,
 *       4 exact state), bits 6-7 Cockburn-Shu partner face j of face i.
 *       On an exact-state face field f is the first of six bc_states rows:
 *       the exterior states at own face nodes 0..4, then the ghost-cell
 *       average the limiter uses (the far-field condition proposed in the
 *       reference's docs/adr/0012-sheared-rectangle-mesh-and-far-field-
 *       boundaries.md:32-44).
 *  lim  (K x 8): alpha0..2 | beta0..2 | dt_geometry | jacobian
 *  bc_states (num_bc_states x 4): conserved fixed exterior states. */
typedef struct {
  int64_t num_elements;
  const double* geo;
  const int32_t* conn;
  const double* lim;
  const double* bc_states;
  int32_t num_bc_states;
} CfdMeshTables;

/* StepReport (time_integration.py:22-29). */
typedef struct {
  double boundary_flux[4];
  int32_t admissible;
  double min_density;
  double min_pressure;
  int64_t troubled_cells;
} CfdStepReport;

/* RunResult (solver.py:62-88) plus the values an error message needs. */
typedef struct {
  double time;
  int64_t steps;
  int64_t rejected_steps;
  double boundary_flux_integral[4];
  double min_density;
  double min_pressure;
  int64_t max_troubled_cells;
  double fail_time; /* t at CFD_EINVALID_DT / CFD_EREJECTED */
  double fail_dt;   /* dt at CFD_EINVALID_DT */
} CfdRunResult;

typedef struct CfdCtx CfdCtx;

const char* cfd_last_error(void);
int cfd_device_count(int* count);

/* Upload the reference-element tables to the current device. Replaces the
 * jnp.asarray copies of build_dg_operators (dg.py:103-114). */
int cfd_init(const CfdReferenceElement* ref);

/* Replaces EulerSolver.__init__'s table upload and jit binding
 * (solver.py:94-125). flux/limiter ids follow INTERFACE_FLUXES
 * (fluxes.py:146-150) and LIMITER_FUNCTIONS (limiters.py:42-48). */
int cfd_ctx_create(const CfdMeshTables* mesh, int flux_id, int limiter_id, double gamma, double cfl,
                   int max_rejections, CfdCtx** out);
int cfd_ctx_destroy(CfdCtx* ctx);
/* Tuning: stage block 64 (default, or the CFDHIP_STAGE_BLOCK environment
 * variable) runs the four-lane stage kernel (32 elements per workgroup);
 * 128 or 256 run the one-lane kernel with that many elements per workgroup.
 * No reference counterpart. */
int cfd_ctx_set_stage_block(CfdCtx* ctx, int elements_per_workgroup);

/* _rhs = jit(dg_rhs) (solver.py:118, dg.py:140-165): rhs (15,K,4), bflux (4). */
int cfd_rhs(CfdCtx* ctx, const double* q, double* rhs, double* boundary_flux);
int cfd_rhs_device(CfdCtx* ctx, const double* q_dev, double* rhs_dev, double* boundary_flux);

/* _limit = jit(apply_limiters) (solver.py:119-121, limiting.py:203-216).
 * avg (K,4) and flags (K, 1 = troubled) may be NULL. */
int cfd_limit(CfdCtx* ctx, const double* q, double* q_out, double* averages, int64_t* troubled_count,
              uint8_t* flags);

/* _time_step = jit(stable_time_step) (solver.py:122, time_integration.py:37-41). */
int cfd_time_step(CfdCtx* ctx, const double* q, double* dt);

/* _step = jit(ssp_rk3_step) (solver.py:123-125, time_integration.py:44-84).
 * flags (3, K) receives each stage's troubled-cell flags; may be NULL. */
int cfd_step(CfdCtx* ctx, const double* q, double dt, double* q_out, CfdStepReport* report, uint8_t* flags);
int cfd_step_device(CfdCtx* ctx, const double* q_dev, double dt, double* q_out_dev, CfdStepReport* report);

/* Allocation-free SSP-RK3 step. No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). The same step,
 * kernels and bits as cfd_step_device, without a context: the tables are the
 * caller's device arrays, and every other buffer (the intermediate stage
 * state, limiter summaries, worklist, partial rows, counters, report row)
 * lives in one static device array of the library, about 40 MB per GPU,
 * sized for at most CFD_STATIC_MAX_ELEMENTS elements; larger meshes are
 * rejected with CFD_EINVAL. cfd_init must have been called on the current
 * device. Runs on the null stream and synchronises before returning. All
 * calls on one GPU share the array: calls that overlap on one GPU (several
 * host threads) race and must be serialised by the caller.
 *  geo (K x 16), conn (K x 4 int32): as in CfdMeshTables (conn is not
 *       checked here; use tables cfd_ctx_create accepts);
 *  alpha_beta (K x 6): alpha0..2 | beta0..2 (the first six lim columns);
 *  dt_geometry (K): lim column 6;  bc_states (num_bc_states x 4). */
#define CFD_STATIC_MAX_ELEMENTS 65536
typedef struct {
  int64_t num_elements;
  const double* geo;
  const int32_t* conn;
  const double* alpha_beta;
  const double* dt_geometry;
  const double* bc_states;
  int32_t num_bc_states;
} CfdDeviceMesh;
int cfd_step_device_static(const CfdDeviceMesh* mesh, int flux_id, int limiter_id, double gamma,
                           const double* q_dev, double dt, double* q_out_dev, CfdStepReport* report);

/* cell_averages (dg.py:168-171) -> (K,4); domain_integral (dg.py:174-176) -> (4). */
int cfd_cell_averages(CfdCtx* ctx, const double* q, double* averages);
int cfd_integrate(CfdCtx* ctx, const double* q, double* integral);

/* EulerSolver.run (solver.py:167-216) with the state resident in HBM.
 * max_steps < 0 means no limit (the reference has none). */
int cfd_run(CfdCtx* ctx, const double* q0, double final_time, int64_t max_steps, double* q_out,
            CfdRunResult* result);

/* Resident-state stepping for benchmarks: upload once, then advance. */
int cfd_state_upload(CfdCtx* ctx, const double* q);
int cfd_state_download(CfdCtx* ctx, double* q);
int cfd_state_advance(CfdCtx* ctx, int64_t nsteps, double* dt_used, CfdStepReport* last);
int cfd_synchronize(CfdCtx* ctx);

/* Kernel timing with HIP events on the context's stream (benchmarks):
 * cfd_timing_read gives [stage-kernel ms total, launches, limiter ms total,
 * launches] accumulated by cfd_state_advance since cfd_timing_enable;
 * cfd_time_rhs gives the mean duration of the RHS-only kernel on the
 * resident state over reps launches. */
int cfd_timing_enable(CfdCtx* ctx, int on);
int cfd_timing_read(CfdCtx* ctx, double* out4);
int cfd_time_rhs(CfdCtx* ctx, int reps, double* avg_ms);

/* Batched solves (no reference counterpart: the reference runs one case per
 * EulerSolver). Many meshes are concatenated into one CfdMeshTables whose
 * element count is a multiple of 256; case i owns elements
 * [first, first + padded), padded a multiple of 256, of which the first
 * `count` are real. Each case runs EulerSolver.run's loop (solver.py:167-216)
 * with its own flux, limiter, CFL and final time; one batched SSP-RK3 step
 * advances every active case by one of its own steps. */
typedef struct {
  int64_t first;
  int64_t count;
  int64_t padded;
  int32_t flux_id;
  int32_t limiter_id;
  double cfl;
  double final_time;
  int32_t max_rejections;
} CfdBatchCase;

typedef struct {
  double time;
  int64_t steps;
  int64_t rejected_steps;
  double boundary_flux_integral[4];
  double min_density;
  double min_pressure;
  int64_t max_troubled_cells;
  int32_t status; /* 0 running, 1 done, 2 bad initial state, 3 invalid dt, 4 rejected too often */
  double fail_time;
  double fail_dt;
} CfdBatchCaseResult;

typedef struct CfdBatch CfdBatch;

int cfd_batch_create(const CfdMeshTables* mesh, int ncase, const CfdBatchCase* cases, double gamma, CfdBatch** out);
int cfd_batch_destroy(CfdBatch* batch);
/* Upload the concatenated (15, K, 4) state (padding elements hold an
 * admissible state of their case), check it and compute each first dt. */
int cfd_batch_begin(CfdBatch* batch, const double* q);
/* Up to nsteps batched steps (use_graph: replay a captured hipGraph). */
int cfd_batch_advance(CfdBatch* batch, int64_t nsteps, int use_graph, int64_t* active_left);
int cfd_batch_download(CfdBatch* batch, double* q);
int cfd_batch_results(CfdBatch* batch, CfdBatchCaseResult* results);
/* Autotuner: time the stage kernel on the current state in its five layouts
 * and keep the fastest: one lane per element with 64, 128 or 256 elements
 * per workgroup (codes 64, 128, 256), four lanes per element with the flux
 * chosen at run time (code 4) or one four-lane launch per flux present
 * (code 5). chosen receives the code; ms_out[5] (may be NULL) the mean
 * times. */
int cfd_batch_autotune(CfdBatch* batch, int reps, int* chosen, double* ms_out);

/* Discrete adjoint (no reference counterpart: the reference differentiates
 * only its pointwise fluxes and limiters with jax.grad,
 * tests/unit/test_physics.py:164-182,224-229). Derivatives are those of the
 * SSP-RK3 map (time_integration.py:44-84) with each limiter pass's troubled
 * flags and positivity scale held at their forward values and dt fixed.
 *
 * cfd_rhs_vjp:   out = (dL/dq)^T mu for the DG RHS L of dg.py:140-165.
 * cfd_limit_vjp: out = (dLim/dq)^T lam for apply_limiters
 *                (limiting.py:203-216) at q; host arrays (15, K, 4). */
int cfd_rhs_vjp(CfdCtx* ctx, const double* q, const double* mu, double* out);
int cfd_limit_vjp(CfdCtx* ctx, const double* q, const double* lam, double* out);

/* A discrete-adjoint run: num_steps SSP-RK3 steps with the time steps dts
 * (NULL: the first forward records the CFL steps, cfl of the context, and
 * later calls reuse them). The forward keeps a checkpoint every
 * checkpoint_every steps; the backward recomputes each segment from its
 * checkpoint, storing the stage states and a tape of the limiter's repaired
 * elements, then runs the adjoint steps in reverse. State pointers are
 * device pointers to (15, K, 4) float64; both calls synchronise. */
typedef struct CfdAdjoint CfdAdjoint;
typedef struct {
  int64_t num_steps;
  int64_t checkpoints;
  int64_t tape_records;      /* records the tape holds */
  int64_t tape_peak;         /* most records one segment used */
  int64_t tape_regrows;      /* segments recomputed after growing the tape */
  int64_t positivity_scaled; /* tape records whose positivity scale was < 1 (last backward) */
  double forward_ms;         /* last forward, host wall clock */
  double backward_ms;        /* last backward, host wall clock */
  double recompute_ms;       /* part of backward_ms spent recomputing segments */
  int64_t device_bytes;      /* device memory held by the run */
  int64_t skipped_passes;    /* limiter passes the last backward's recomputation skipped (unlimited runs:
                                passes the forward found to repair nothing) */
} CfdAdjointStats;

int cfd_adjoint_create(CfdCtx* ctx, int64_t num_steps, const double* dts, int checkpoint_every, CfdAdjoint** out);
int cfd_adjoint_destroy(CfdAdjoint* adj);
int cfd_adjoint_forward(CfdAdjoint* adj, const double* q0_dev, double* qn_dev);
int cfd_adjoint_backward(CfdAdjoint* adj, const double* gn_dev, double* g0_dev);
/* The same march with every limiter pass's troubled decisions taken from the
 * last cfd_adjoint_forward (positivity is applied as usual): the map whose
 * derivative cfd_adjoint_backward returns, for finite-difference checks.
 * Leaves the checkpoints untouched. */
int cfd_adjoint_forward_replay(CfdAdjoint* adj, const double* q0_dev, double* qn_dev);
int cfd_adjoint_time_steps(CfdAdjoint* adj, double* dts);
int cfd_adjoint_stats(CfdAdjoint* adj, CfdAdjointStats* stats);

/* Host-side table assembly (no GPU needed): connectivity and boundary tags,
 * the periodic patch, face-node orientation, affine geometry, Cockburn-Shu
 * coefficients and CFL geometry of a straight-sided triangle mesh, written as
 * the compact tables CfdMeshTables describes. Replaces the reference's NumPy
 * setup in build_connectivity (mesh.py:135-185), element_geometry
 * (mesh.py:216-243), pair_face_nodes (mesh.py:254-291), build_dg_operators
 * (dg.py:76-116) and neighbor_coefficients (limiting.py:51-94), with the same
 * error texts (CFD_EINVAL). Boundary tags are given in sorted name order,
 * tag_names newline-separated; tag kinds 1 slip wall, 2 fixed state (tag_state
 * = its bc_states row), 3 transmissive, 4 periodic (tag_partner, tag_shift:
 * x_here = x_partner + shift), 5 exact state (the field of such a face is
 * left 0 for the caller to point at its bc_states rows). The output arrays
 * geo (K x 16), conn (K x 4), lim (K x 8) are required; the others may be
 * NULL: neighbor_element, neighbor_face, face_tag (K x 3 int64, periodic
 * faces patched to interior faces), centroid (K x 2), normal (K x 3 x 2),
 * edge_length (K x 3), face_midpoint (K x 3 x 2), periodic_shift (K x 3 x 2). */
typedef struct {
  int64_t num_vertices;
  const double* vertices;     /* Nv x 2 */
  int64_t num_elements;
  const int64_t* elements;    /* K x 3, counterclockwise */
  int32_t num_tags;
  const char* tag_names;
  const int64_t* tag_offsets; /* num_tags + 1: tag t owns edges tag_offsets[t] .. tag_offsets[t + 1] - 1 */
  const int64_t* tag_edges;   /* vertex pairs */
  const int32_t* tag_kind;
  const int32_t* tag_state;
  const int32_t* tag_partner;
  const double* tag_shift;    /* num_tags x 2 */
} CfdMeshInput;
typedef struct {
  double* geo;
  int32_t* conn;
  double* lim;
  int64_t* neighbor_element;
  int64_t* neighbor_face;
  int64_t* face_tag;
  double* centroid;
  double* normal;
  double* edge_length;
  double* face_midpoint;
  double* periodic_shift;
} CfdMeshOutput;
int cfd_build_tables(const CfdMeshInput* in, CfdMeshOutput* out);

/* Pointwise interface flux (fluxes.py:28-143) on n face states: qm, qp (n,4),
 * nx, ny (n), out (n,4); host arrays. */
int cfd_flux_pointwise(int flux_id, int64_t n, const double* qm, const double* qp, const double* nx,
                       const double* ny, double gamma, double* out);
/* Pointwise limiter function phi(r) (limiters.py:16-39). */
int cfd_limiter_pointwise(int limiter_id, int64_t n, const double* r, double* out);

#ifdef __cplusplus
}
#endif

#endif /* CFDHIP_H */
