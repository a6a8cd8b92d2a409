// Device-side tables, argument blocks and helpers shared by the kernels.
#pragma once

#include "cfd_physics.hpp"

namespace cfd {

constexpr int LBLOCK = 256;      // threads per block of the per-element passes
constexpr int FBLOCK = 64;       // threads per block of the limiter repair pass
constexpr int FGRID = 4096;      // blocks of the limiter repair pass (4 elements in flight each)
constexpr int SUMM_STRIDE = 14;  // doubles of limiter summary per element (7 double2 pieces)
constexpr int PART_STRIDE = 16;  // doubles per partial row (PART_COLS used)
constexpr int PART_COLS = 9;
constexpr int GEO_STRIDE = 16;   // doubles of geometry per element
constexpr int LIM_STRIDE = 8;    // doubles of limiter geometry per element in the host tables
constexpr int LIM_DEV = 6;       // ... on the device: alpha[3], beta[3] (dt geometry and J go to their own arrays)
constexpr double INF = __builtin_huge_val();

// Summary layout (per element, written by the stage kernels, read by the
// limiter): seven double2 pieces stored piece-major, piece p of element e at
// ((double2*)summ)[p * K + e], so a wave's 16-byte loads of one piece for
// consecutive elements are contiguous. Offsets below are in doubles within
// the element's gathered row (piece p = doubles 2p, 2p + 1), except that the
// cell average is stored as the pieces (rho, E) and (rho u, rho v), so the
// troubled-cell test, which compares densities and energies only, gathers one
// piece per neighbour; load_summary puts the average back in order.
constexpr int SUMM_AVG = 0;    // 4: cell average of the unlimited stage state
constexpr int SUMM_MID = 4;    // 6: (rho, E) at the three face-midpoint nodes
constexpr int SUMM_RMIN = 10;  // min nodal density
constexpr int SUMM_PMIN = 11;  // min nodal pressure
constexpr int SUMM_FLAGS = 12; // SF_* bits
constexpr int SUMM_LAMBDA = 13;  // max nodal |u| + c (last stage only)

// Partial-row columns: boundary flux (sum), nodal minima, troubled count,
// admissibility (min of 0/1), CFL candidate dtg / lambda (min).
constexpr int COL_RMIN = 4;
constexpr int COL_PMIN = 5;
constexpr int COL_TROUBLED = 6;
constexpr int COL_ADM = 7;
constexpr int COL_DT = 8;
constexpr int SF_NEED_RHO = 1;  // some nodal density < epsilon
constexpr int SF_NEED_P = 2;    // some nodal pressure < epsilon
constexpr int SF_NONFINITE = 4; // some value is not finite

enum StageMode : int { MODE_RHS = 0, MODE_STAGE1 = 1, MODE_STAGE2 = 2, MODE_STAGE3 = 3, MODE_SUMMARY = 4 };

// Face slot s = 5*f + k -> own node index (face_nodes of the reference
// element, reference_element.py:274-279; checked on the host at init).
constexpr int FACE_NODE[NSLOT] = {0, 1, 2, 3, 4, 4, 8, 11, 13, 14, 0, 5, 9, 12, 14};
constexpr int MID_NODE[3] = {2, 11, 9};  // face_nodes[:, 2] (reference_element.py:295)

constexpr uint64_t pack_face_nodes() {
  uint64_t v = 0;
  for (int s = 0; s < NSLOT; ++s) v |= (uint64_t)FACE_NODE[s] << (4 * s);
  return v;
}
constexpr uint64_t FACE_NODE_PACKED = pack_face_nodes();

// For each node, the (up to two) face slots it owns, one byte each, 0xff = none.
constexpr int node_slots(int j) {
  int v = 0xffff, t = 0;
  for (int s = 0; s < NSLOT; ++s)
    if (FACE_NODE[s] == j) {
      v = (v & ~(0xff << (8 * t))) | (s << (8 * t));
      ++t;
    }
  return v;
}
constexpr int NODE_SLOTS[NP] = {node_slots(0), node_slots(1), node_slots(2),  node_slots(3),  node_slots(4),
                                node_slots(5), node_slots(6), node_slots(7),  node_slots(8),  node_slots(9),
                                node_slots(10), node_slots(11), node_slots(12), node_slots(13), node_slots(14)};

__device__ __forceinline__ int face_node_dyn(int s) { return (int)((FACE_NODE_PACKED >> (4 * s)) & 15u); }

extern __constant__ double c_drT[NP * NP];
extern __constant__ double c_dsT[NP * NP];
extern __constant__ double c_liftT[NSLOT * NP];
extern __constant__ double c_avgw[NP];
extern __constant__ double c_cellw[NP];
extern __constant__ double c_facew[NSLOT];
extern __constant__ double c_p1mid[3 * NP];
extern __constant__ double c_linb[NP * 3];

// Batched solves: many meshes concatenated, each padded to a multiple of
// 256 elements, so every workgroup (64, 128 or 256 elements) and every
// 256-element limiter block lies inside one case.
struct BatchView {
  const int* lcase;        // case of each 256-element block
  const int64_t* begin;    // first element of each case
  const int* nreal;        // real (unpadded) elements of each case
  const int* flux;         // flux id per case
  const int* limiter;      // limiter id per case
  const double* dt;        // dt of the current attempt per case (<= 0: case inactive)
  int* wcount;             // per 256-element block: worklist length
  // The 256-element blocks a launch covers (nullptr: all of them): workgroup
  // w of a launch works on the w-th listed block (or the part of it its
  // workgroup size covers), so launches shrink as cases finish.
  const int* sblk = nullptr;
};

// Element block of workgroup w when each 256-element block holds `per`
// workgroups: through the launch's block list, or w itself.
__device__ __forceinline__ int listed_block(const BatchView& bv, int w, int per) {
  return bv.sblk ? bv.sblk[w / per] * per + w % per : w;
}

struct StageArgs {
  const double* q;
  const double* q0;
  double* out;
  const double* geo;
  const int* conn;
  const double* bcs;
  double* summ;
  double* part;
  int64_t K;
  double gamma;
  double dt;
  BatchView bv;  // batched solves only
  int blk0 = 0;  // batched one-flux launches: first element block of the launch
  // Row stride of geo in doubles: the stage kernels read only rx, ry, sx, sy,
  // from the context's compact K x 4 copy (stride 4) or from the full
  // geometry rows (GEO_STRIDE) of caller-owned tables.
  int geo_stride = GEO_STRIDE;
  // Device-controlled runs (nullptr: off): the step's dt, read from device
  // memory; a value that is not > 0 marks a step enqueued after the run
  // stopped, and the kernel returns at once.
  const double* dtp = nullptr;
};

struct LimitArgs {
  double* q;
  const double* summ;
  const double* geo;
  const int* conn;
  const double* lim;
  const double* bcs;
  double* part;
  uint8_t* flags;
  int* work;        // worklist of element ids (bit 31 = troubled)
  int* work_count;  // worklist length, zero on entry to k_limit_detect
  const double* dtg;
  int64_t K;
  double gamma;
  int limiter;
  int limit_slopes;
  int want_dt;
  BatchView bv;  // bv.lcase != nullptr: batched mode, per-block worklists
  // Adjoint tape (nullptr: off). Worklist slot s of this pass is recorded at
  // record *tape_head + s when that is below tape_cap.
  double* tape = nullptr;
  const int64_t* tape_head = nullptr;
  int64_t tape_cap = 0;
  // Replay (nullptr: off): the troubled decision of element e is force[e]
  // instead of the detector's (frozen-decision forward for the adjoint).
  const uint8_t* force = nullptr;
  // Stage-kernel partial rows (nullptr: none): detection block b folds rows
  // [b * stage_per_blk, (b + 1) * stage_per_blk) of the n_stage_rows into its
  // own row, so the report reduction reads one row per LBLOCK elements.
  const double* stage_rows = nullptr;
  int stage_per_blk = 0;
  int n_stage_rows = 0;
  // Repair pass (nullptr: off): the worklist length of this pass is stored
  // here (the adjoint's recomputation skips passes that repaired nothing).
  int* count_out = nullptr;
  const double* dtp = nullptr;  // as StageArgs::dtp
};

// A step of a device-controlled run that the control kernel has stopped.
__device__ __forceinline__ bool step_stopped(const double* dtp) { return dtp && !(*dtp > 0.0); }

// EulerSolver.run's loop state (solver.py:167-216) for runs whose steps are
// enqueued ahead of the host: run_control applies the loop body after each
// step, so the host reads this block once per batch of steps instead of a
// report per step.
enum RunStatus : int { RUN_GO = 0, RUN_DONE = 1, RUN_RETRY = 2, RUN_INVALID_DT = 3, RUN_REJECTED = 4 };
struct RunCtl {
  double dt;          // dt of the next step (the step kernels read it); 0 while stopped
  double retry_dt;    // RUN_RETRY: the halved dt of the next attempt
  double next_dt;     // RUN_DONE: cfl x the last step's CFL candidate
  double t, final_time, cfl;
  double fint[4];     // boundary-flux integral
  double mind, minp;  // running minima (Python min semantics)
  double fail_time, fail_dt;
  double last_dt, last_bflux[4], last_rmin, last_pmin;  // the last accepted step and its report
  int64_t steps, rejected, max_troubled, max_steps, last_troubled;
  int status, attempt, max_rejections, reaches_end;
};

// Tape record of one limiter worklist element (TAPE_STRIDE doubles):
// element id, troubled (0/1), positivity scale theta, the scaled neighbour
// deviations sn[3][4] and the pre-limit nodal values [NP][4].
constexpr int TAPE_STRIDE = 80;
constexpr int TAPE_E = 0;
constexpr int TAPE_TROUBLED = 1;
constexpr int TAPE_THETA = 2;
constexpr int TAPE_SN = 4;
constexpr int TAPE_Q = 16;

struct DtArgs {
  const double* q;
  const double* dtg;  // per-element factor (dt kernel) or Jacobian (integral kernel)
  double* part;
  int64_t K;
  double gamma;
};

struct ElemGeo {
  double rx, ry, sx, sy;
  double nx[3], ny[3];
  double fs[3];  // face_scale = 0.5 |e| / J (dg.py:99)
  double he[3];  // 0.5 |e|, the edge-weight factor (dg.py:101)
};

__device__ __forceinline__ void load4(const double* p, double v[4]) {
  const double2 a = reinterpret_cast<const double2*>(p)[0];
  const double2 b = reinterpret_cast<const double2*>(p)[1];
  v[0] = a.x;
  v[1] = a.y;
  v[2] = b.x;
  v[3] = b.y;
}

__device__ __forceinline__ void store4(double* p, const double v[4]) {
  reinterpret_cast<double2*>(p)[0] = make_double2(v[0], v[1]);
  reinterpret_cast<double2*>(p)[1] = make_double2(v[2], v[3]);
}

// Non-temporal store (the nt bit: evict-first in the caches) for streams the
// same kernel does not read back and that do not fit in the caches anyway:
// they then stop evicting data the next kernel reads (the stage kernel's
// state output, written beside the limiter summary that detection reads
// next, took 12 % off the limiter that way).
__device__ __forceinline__ void store4_nt(double* p, const double v[4]) {
#pragma unroll
  for (int k = 0; k < 4; ++k) __builtin_nontemporal_store(v[k], p + k);
}

// Face frame of an affine triangle from its inverse-Jacobian entries. The
// reference faces s = -1, r + s = 0 and r = -1 map to the physical edge
// vectors 2J (sy, -sx), 2J (-ry - sy, rx + sx) and 2J (ry, -rx), J the volume
// Jacobian 1 / (rx sy - ry sx). With g_f = -(sx, sy), (rx + sx, ry + sy),
// -(rx, ry) and sigma the sign of J: outward unit normal sigma g_f / |g_f|,
// face scale 0.5 |e_f| / J = sigma |g_f|, half edge length 0.5 |e_f| = |g_f| |J|.
// The geometry row holds the same quantities computed from the vertices
// (mesh.py:216-243, dg.py:99-101); the two agree to a few ulps. The stage
// and adjoint kernels compute them instead of loading them: nine of the
// row's sixteen doubles, loaded by every lane of an element, were a third of
// the data the four-lane kernel's loads return. The limiter keeps reading
// the row's normals (slip-wall ghost averages).
__device__ __forceinline__ void face_frame(double rx, double ry, double sx, double sy, int f, double& nx, double& ny,
                                           double& fs) {
  const double gx = f == 0 ? -sx : (f == 1 ? rx + sx : -rx);
  const double gy = f == 0 ? -sy : (f == 1 ? ry + sy : -ry);
  const bool ccw = rx * sy - ry * sx > 0.0;
  const double len = fsqrt(gx * gx + gy * gy);
  const double il = frcp(len);
  nx = gx * (ccw ? il : -il);
  ny = gy * (ccw ? il : -il);
  fs = ccw ? len : -len;
}

// The affine part of a geometry row (rx, ry, sx, sy; rows of `stride`
// doubles) and the face frames derived from it (face_frame; half edge length
// = face scale x J).
__device__ __forceinline__ void load_geo_affine(const double* geo, int stride, int64_t e, ElemGeo& g) {
  const double2* p = reinterpret_cast<const double2*>(geo + e * stride);
  const double2 a = p[0], b = p[1];
  g.rx = a.x;
  g.ry = a.y;
  g.sx = b.x;
  g.sy = b.y;
  const double jac = frcp(g.rx * g.sy - g.ry * g.sx);
#pragma unroll
  for (int f = 0; f < 3; ++f) {
    face_frame(g.rx, g.ry, g.sx, g.sy, f, g.nx[f], g.ny[f], g.fs[f]);
    g.he[f] = g.fs[f] * jac;
  }
}

// Workgroups are dispatched round-robin over the 8 XCDs; give each XCD one
// contiguous range of element blocks so neighbour traces hit its own L2.
__device__ __forceinline__ int xcd_block(int b, int nb) {
  const int x = b & 7;
  const int q = nb >> 3, r = nb & 7;
  return x * q + (x < r ? x : r) + (b >> 3);
}

// The same ranges, each walked from its end: the limiter's detection pass
// reads the summaries the stage kernel wrote just before in the same per-XCD
// order, so starting from the most recent ones finds more of them still in
// the Infinity Cache (detection + repair 0.107 -> 0.100 ms per stage at 2M
// elements). Partial rows stay indexed by block, so results are unchanged.
__device__ __forceinline__ int xcd_block_rev(int b, int nb) {
  const int x = b & 7;
  const int q = nb >> 3, r = nb & 7;
  const int len = q + (x < r ? 1 : 0);
  return x * q + (x < r ? x : r) + (len - 1 - (b >> 3));
}

// Correctly rounded x / 3 without a division (a division is eleven
// dependent instructions, and the last stage takes fifteen per lane; this is
// seven). q = RN(x RN(1/3)) is within one ulp of x / 3, so the remainder
// x - 3q is exact as an FMA (also for subnormal x, where x and 3q are
// multiples of the same quantum, and near overflow, since the FMA does not
// round 3q), and Markstein's correction RN(q + r RN(1/3)) is RN(x / 3): x / 3
// is never a rounding midpoint, which would take a 55-bit x. copysign keeps
// the sign of a zero x; a NaN remainder (x infinite or NaN) returns q.
__device__ __forceinline__ double div3(double x) {
#pragma clang fp contract(off)
  constexpr double third = 0x1.5555555555555p-2;  // RN(1/3)
  const double q = x * third;
  const double r = fma(-q, 3.0, x);
  const double res = copysign(fma(r, third, q), x);
  return r == r ? res : q;
}

// SSP-RK3 stage combinations with the reference's association
// (time_integration.py:70,72,74); contraction is off so the bits match.
template <int MODE>
__device__ __forceinline__ double rk3_combine(double q0, double qi, double dt, double rhs) {
#pragma clang fp contract(off)
  if constexpr (MODE == MODE_STAGE1) {
    return qi + dt * rhs;
  } else if constexpr (MODE == MODE_STAGE2) {
    return 0.75 * q0 + 0.25 * (qi + dt * rhs);
  } else {
    return div3(q0) + 2.0 / 3.0 * (qi + dt * rhs);
  }
}

__device__ __forceinline__ void part_neutral(double row[PART_COLS]) {
#pragma unroll
  for (int c = 0; c < PART_COLS; ++c) row[c] = 0.0;
  row[COL_RMIN] = INF;
  row[COL_PMIN] = INF;
  row[COL_ADM] = 1.0;
  row[COL_DT] = INF;
}

__device__ __forceinline__ double part_combine(int c, double a, double b) {
  if (c == COL_RMIN || c == COL_PMIN || c == COL_DT) return nan_min(a, b);
  if (c == COL_ADM) return a < b ? a : b;
  return a + b;
}

// A double moved across lanes by a DPP control: quad permutations (0x00-0xFF)
// and rotations within rows of 16 lanes (0x121-0x12F) are VALU moves, with no
// trip through the LDS crossbar that a lane shuffle takes.
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}

// Fixed-order reduction of one column over a wave; lane 0 ends with
// ((q0 q1)(q2 q3))((q4 q5)(q6 q7)) ... in a balanced tree over the sixteen
// quads. Within a row of 16 lanes: partner lane ^1 and ^2 (quad
// permutations), then the lane 4 and 8 places on (row rotations, which meet
// the butterfly's partners on the lanes lane 0 depends on); across rows:
// lane shuffles ^16 and ^32.
template <int C>
__device__ __forceinline__ double wave_reduce_col(double v) {
  v = part_combine(C, v, dpp_f64<0xB1>(v));   // quad_perm [1,0,3,2]
  v = part_combine(C, v, dpp_f64<0x4E>(v));   // quad_perm [2,3,0,1]
  v = part_combine(C, v, dpp_f64<0x124>(v));  // row_ror 4
  v = part_combine(C, v, dpp_f64<0x128>(v));  // row_ror 8
  v = part_combine(C, v, __shfl_xor(v, 16, 64));
  return part_combine(C, v, __shfl_xor(v, 32, 64));
}

template <int C = 0>
__device__ __forceinline__ void wave_reduce_row(double v[PART_COLS]) {
  if constexpr (C < PART_COLS) {
    v[C] = wave_reduce_col<C>(v[C]);
    wave_reduce_row<C + 1>(v);
  }
}

// Block-wide deterministic reduction of one partial row (NT threads): a fixed
// tree inside each wave (wave_reduce_col), then wave results in order through
// LDS.
template <int NT>
__device__ __forceinline__ void block_reduce_row(const double row[PART_COLS], double* out) {
  constexpr int NW = NT / 64;
  __shared__ double sh[NW][PART_COLS];
  double v[PART_COLS];
#pragma unroll
  for (int c = 0; c < PART_COLS; ++c) v[c] = row[c];
  wave_reduce_row(v);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int c = 0; c < PART_COLS; ++c) sh[w][c] = v[c];
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int c = 0; c < PART_COLS; ++c) {
      double t = sh[0][c];
      for (int k = 1; k < NW; ++k) t = part_combine(c, t, sh[k][c]);
      out[c] = t;
    }
  }
}

// The stage kernels' partial row: only the four boundary-flux sums vary, the
// other columns are neutral. Same tree and wave order as block_reduce_row on
// those columns (identical bits), without the work on the constant ones.
template <int NT>
__device__ __forceinline__ void block_reduce_bflux(const double bf[4], double* out) {
  constexpr int NW = NT / 64;
  __shared__ double sh[NW][4];
  double v[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) v[c] = bf[c];
  // Most waves hold no boundary face; their sums are +0 (every lane starts at
  // +0 and only boundary slots add), so they skip the butterfly's 48 lane
  // shuffles. A NaN counts as nonzero and still goes through.
  if (__ballot(v[0] != 0.0 || v[1] != 0.0 || v[2] != 0.0 || v[3] != 0.0) != 0) {
    v[0] = wave_reduce_col<0>(v[0]);
    v[1] = wave_reduce_col<1>(v[1]);
    v[2] = wave_reduce_col<2>(v[2]);
    v[3] = wave_reduce_col<3>(v[3]);
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int c = 0; c < 4; ++c) sh[w][c] = v[c];
  __syncthreads();
  if (threadIdx.x == 0) {
    double row[PART_COLS];
    part_neutral(row);
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      double t = sh[0][c];
      for (int k = 1; k < NW; ++k) t = t + sh[k][c];
      row[c] = t;
    }
#pragma unroll
    for (int c = 0; c < PART_COLS; ++c) out[c] = row[c];
  }
}

// sqrt(u^2 + v^2) + c with the reference's operation order
// (time_integration.py:39-40); contraction off so k_dt and the fused stage
// summary give the same bits as each other.
__device__ __forceinline__ double wave_speed(const double q[4], double gamma) {
#pragma clang fp contract(off)
  const double u = q[1] / q[0];
  const double v = q[2] / q[0];
  const double p = (gamma - 1.0) * (q[3] - 0.5 * (q[1] * u + q[2] * v));
  return sqrt(u * u + v * v) + sqrt(gamma * p / q[0]);
}

// pressure (euler.py:30-32) and sqrt(u^2 + v^2) + c from one pair of
// divisions; bit-identical to pressure_of and wave_speed.
__device__ __forceinline__ void pressure_and_speed(const double q[4], double gamma, double& p, double& ws) {
#pragma clang fp contract(off)
  const double u = q[1] / q[0];
  const double v = q[2] / q[0];
  p = (gamma - 1.0) * (q[3] - 0.5 * (q[1] * u + q[2] * v));
  ws = sqrt(u * u + v * v) + sqrt(gamma * p / q[0]);
}

// A lower bound of the correctly rounded pressure (pressure_of) from the
// pressure p = gm1 * (E - ke) computed with the refined reciprocal and
// possibly contracted into FMAs. The two differ by a few ulps of the terms
// that cancel, |E| and ke (the reciprocal, the products, the sum and the
// difference each add at most a few half-ulps); 2^-47 of (|E| + |ke|) is 64
// half-ulps, so p minus that never exceeds the correctly rounded value.
__device__ __forceinline__ double pressure_floor(double p, double E, double ke, double gm1) {
  return p - gm1 * (fabs(E) + fabs(ke)) * 0x1p-47;
}

__device__ __forceinline__ double2* summ_piece(double* summ, int64_t K, int p, int64_t e) {
  return reinterpret_cast<double2*>(summ) + (int64_t)p * K + e;
}

// Where the two doubles of stored piece p go in the gathered row: the cell
// average's pieces hold (avg 0, avg 3) and (avg 1, avg 2).
__device__ __forceinline__ constexpr int summ_lo(int p) { return p == 0 ? 0 : (p == 1 ? 1 : 2 * p); }
__device__ __forceinline__ constexpr int summ_hi(int p) { return p == 0 ? 3 : (p == 1 ? 2 : 2 * p + 1); }

// Gathers pieces [0, npieces) of element e's summary into sm (the row order
// of the SUMM_ offsets).
template <int NPIECES = SUMM_STRIDE / 2>
__device__ __forceinline__ void load_summary(const double* summ, int64_t K, int64_t e, double* sm) {
  const double2* s = reinterpret_cast<const double2*>(summ);
#pragma unroll
  for (int p = 0; p < NPIECES; ++p) {
    const double2 v = s[(int64_t)p * K + e];
    sm[summ_lo(p)] = v.x;
    sm[summ_hi(p)] = v.y;
  }
}

// Per-element limiter summary of a stage state held in registers.
template <bool WITH_LAMBDA>
__device__ __forceinline__ void write_summary(const double q[NP][NV], double gamma, double* summ, int64_t K,
                                              int64_t e) {
  const double gm1 = gamma - 1.0;
  double avg[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int i = 0; i < NP; ++i)
#pragma unroll
    for (int c = 0; c < 4; ++c) avg[c] = fma(c_avgw[i], q[i][c], avg[c]);
  const double pbar = pressure_of(avg, gm1);
  const double eps = nan_min(1e-13, nan_min(avg[0], pbar));
  // Nodal pressures and wave speeds with the refined hardware reciprocal and
  // square root. need_p only sends an element to the repair pass, which
  // decides with correctly rounded pressures; it tests a lower bound of the
  // correctly rounded pressure, so no node below epsilon is missed.
  double rmin = INF, pmin = INF, lam = -INF;
  bool finite = true, need_p = false;
#pragma unroll
  for (int i = 0; i < NP; ++i) {
    const double inv = frcp(q[i][0]);
    const double u = q[i][1] * inv, v = q[i][2] * inv;
    const double ke = 0.5 * (q[i][1] * u + q[i][2] * v);
    const double p = gm1 * (q[i][3] - ke);
    if constexpr (WITH_LAMBDA) {
      const double ws = fsqrt(u * u + v * v) + fsqrt(gamma * p * inv);
      lam = i == 0 ? ws : nan_max(lam, ws);
    }
    rmin = nan_min(rmin, q[i][0]);
    pmin = nan_min(pmin, p);
    need_p = need_p || (pressure_floor(p, q[i][3], ke, gm1) < eps);
#pragma unroll
    for (int c = 0; c < 4; ++c) finite = finite && isfinite(q[i][c]);
  }
  const int flags = ((rmin < eps) ? SF_NEED_RHO : 0) | (need_p ? SF_NEED_P : 0) | (finite ? 0 : SF_NONFINITE);
  *summ_piece(summ, K, 0, e) = make_double2(avg[0], avg[3]);
  *summ_piece(summ, K, 1, e) = make_double2(avg[1], avg[2]);
  *summ_piece(summ, K, 2, e) = make_double2(q[MID_NODE[0]][0], q[MID_NODE[0]][3]);
  *summ_piece(summ, K, 3, e) = make_double2(q[MID_NODE[1]][0], q[MID_NODE[1]][3]);
  *summ_piece(summ, K, 4, e) = make_double2(q[MID_NODE[2]][0], q[MID_NODE[2]][3]);
  *summ_piece(summ, K, 5, e) = make_double2(rmin, pmin);
  *summ_piece(summ, K, 6, e) = make_double2((double)flags, lam);
}

// _minmod (limiting.py:124-125).
__device__ __forceinline__ double minmod2(double a, double b) {
#pragma clang fp contract(off)
  if (a * b > 0.0) {
    const double m = fmin(fabs(a), fabs(b));
    return a > 0.0 ? m : -m;
  }
  return 0.0;
}

// Troubled-cell detection (limiting.py:128-151). With SLOPES it also returns
// the scaled neighbour deviations nu * (alpha d_i + beta d_j) of all four
// components for the reconstruction; without, only the density and energy
// entries of sn are meaningful (the test reads no others), and a neighbour's
// average costs one 16-byte piece instead of two.
template <bool SLOPES>
__device__ __forceinline__ bool detect_troubled(const LimitArgs& a, int64_t e, const double avg[4],
                                                const double* sm, double sn[3][4]) {
#pragma clang fp contract(off)
  const int4 cn = reinterpret_cast<const int4*>(a.conn)[e];
  const int nbr[3] = {cn.x, cn.y, cn.z};
  const uint32_t code = (uint32_t)cn.w;
  const double* lim = a.lim + e * LIM_DEV;
  const double* geo = a.geo + e * GEO_STRIDE;
  // Neighbour-average differences per face, kept as three separate arrays so
  // the partner-face selection below stays a register select.
  double d0[4], d1[4], d2[4];
  int pf[3];
#pragma unroll
  for (int f = 0; f < 3; ++f) {
    const uint32_t fc = (code >> (8 * f)) & 0xffu;
    const int kind = (fc >> 3) & 7;
    pf[f] = (fc >> 6) & 3;
    double nb[4];
    if (kind == BC_INTERIOR) {
      if constexpr (SLOPES) {
        load_summary<2>(a.summ, a.K, nbr[f], nb);
      } else {
        const double2 v = reinterpret_cast<const double2*>(a.summ)[nbr[f]];  // piece 0: (rho, E)
        nb[0] = v.x;
        nb[1] = avg[1];
        nb[2] = avg[2];
        nb[3] = v.y;
      }
    } else if (kind == BC_SLIP) {
      const double nx = geo[4 + f], ny = geo[7 + f];
      const double mn = avg[1] * nx + avg[2] * ny;
      nb[0] = avg[0];
      nb[1] = avg[1] - 2.0 * mn * nx;
      nb[2] = avg[2] - 2.0 * mn * ny;
      nb[3] = avg[3];
    } else if (kind == BC_FIXED || kind == BC_EXACT) {
      const int64_t row = (int64_t)nbr[f] + (kind == BC_EXACT ? NFP : 0);  // an exact face's ghost average
#pragma unroll
      for (int c = 0; c < 4; ++c) nb[c] = a.bcs[row * 4 + c];
    } else {
#pragma unroll
      for (int c = 0; c < 4; ++c) nb[c] = avg[c];
    }
    double* d = f == 0 ? d0 : (f == 1 ? d1 : d2);
#pragma unroll
    for (int c = 0; c < 4; ++c) d[c] = nb[c] - avg[c];
  }
  bool troubled = false;
#pragma unroll
  for (int f = 0; f < 3; ++f) {
    const double al = lim[f], be = lim[3 + f];
    const double* dself = f == 0 ? d0 : (f == 1 ? d1 : d2);
    // The partner face is one of the two other faces.
    const int fa = (f + 1) % 3, fb = (f + 2) % 3;
    const double* da = fa == 0 ? d0 : (fa == 1 ? d1 : d2);
    const double* db = fb == 0 ? d0 : (fb == 1 ? d1 : d2);
    const bool use_a = pf[f] == fa;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const double paired = use_a ? da[c] : db[c];
      sn[f][c] = 1.5 * (al * dself[c] + be * paired);
    }
#pragma unroll
    for (int cc = 0; cc < 2; ++cc) {
      const int c = cc == 0 ? 0 : 3;
      const double t = sm[SUMM_MID + 2 * f + cc] - avg[c];
      const double excess = fabs(minmod2(t, sn[f][c]) - t);
      troubled = troubled || (excess > 1e-10 * fabs(avg[c]));
    }
  }
  return troubled;
}

}  // namespace cfd
