// gfx950 kernels for the degree-4 nodal DG Euler hot path:
//   k_stage        strong-form RHS (volume + face fluxes + lift), fused with
//                  the SSP-RK3 stage combination and a per-element limiter
//                  summary (cell average, midpoint traces, minima, flags, and
//                  in the last stage the CFL wave speed)
//   k_limit_detect troubled-cell detection from the summaries; elements that
//                  need repair go to a worklist, the rest report directly
//                  (k_limit_detect_direct: the same without slope limiting)
//   k_limit_fix    P1 reconstruction and Zhang-Shu positivity, in place, for
//                  worklist elements only (16 lanes per element)
//   k_limit_block  both passes per 256-element block, for small meshes
//   k_dt           CFL time step of a state (first step / time_step API)
//   k_finalize     deterministic two-level reduction of per-block partials
//
// State layout in HBM is the reference's (Np, K, 4) float64, node-major:
// node j of element e starts at q + (j*K + e)*4. The 15x15 operators live in
// __constant__ memory and are read with wave-uniform indices, i.e. as scalar
// FMA operands.
//
// Two stage-kernel layouts: k_stage4 (cfd_stage4.inc; four lanes per
// element, 15 accumulators per lane, three waves per SIMD) is the default for
// a fixed flux; k_stage below (one lane per element, 60 accumulators, two
// waves per SIMD) serves the run-time flux of batched solves, the
// summary-only mode and other workgroup sizes. Contractions on FP64 MFMA
// (in two lane layouts; MI355X's FP64 matrix and vector rates are equal), an
// LDS-staged one-lane-per-element variant and a sixteen-lanes-per-element
// layout with DPP row-broadcast FMAs were all measured slower and are not
// kept.
#include "cfd_device.hpp"

namespace cfd {

__constant__ double c_drT[NP * NP];       // Dr transposed: column j at [j*NP]
__constant__ double c_dsT[NP * NP];
__constant__ double c_liftT[NSLOT * NP];  // LIFT transposed: column s at [s*NP]
__constant__ double c_avgw[NP];           // cell weights / sum(cell weights)
__constant__ double c_cellw[NP];          // cell weights
__constant__ double c_facew[NSLOT];       // GLL edge weight of each face slot
__constant__ double c_p1mid[3 * NP];
__constant__ double c_linb[NP * 3];

// ---------------------------------------------------------------------------
// Stage kernel: one lane per element, SB elements per workgroup.
// ---------------------------------------------------------------------------
template <int FLUX, int MODE, int SB>
__global__ void __launch_bounds__(SB) __attribute__((amdgpu_waves_per_eu(2))) k_stage(StageArgs a) {
  int blk = xcd_block(blockIdx.x, gridDim.x);
  if constexpr (FLUX == FLUX_DYNAMIC) blk = listed_block(a.bv, blk, LBLOCK / SB);
  const int64_t K = a.K;
  const int64_t e0 = (int64_t)blk * SB;
  const int ne = (int)((K - e0) < SB ? (K - e0) : SB);
  const int t = threadIdx.x;
  const int64_t e = e0 + t;
  bool live = t < ne;
  const double gamma = a.gamma;
  const double gm1 = gamma - 1.0;
  double dt = a.dt;
  int flux_rt = FLUX;
  if constexpr (FLUX != FLUX_DYNAMIC) {
    if (a.dtp) {
      dt = *a.dtp;
      if (!(dt > 0.0)) return;
    }
  }
  if constexpr (FLUX == FLUX_DYNAMIC) {
    const int cs = a.bv.lcase[e0 >> 8];
    dt = a.bv.dt[cs];
    flux_rt = a.bv.flux[cs];
    live = live && (e - a.bv.begin[cs]) < a.bv.nreal[cs];
    if (!(dt > 0.0)) {
      // Inactive case: the last stage carries the step input over unchanged.
      if constexpr (MODE == MODE_STAGE3) {
        if (live)
#pragma unroll
          for (int i = 0; i < NP; ++i) {
            double v[4];
            load4(a.q0 + ((int64_t)i * K + e) * 4, v);
            store4(a.out + ((int64_t)i * K + e) * 4, v);
          }
      }
      return;
    }
  }

  double acc[NP][NV];
#pragma unroll
  for (int i = 0; i < NP; ++i)
#pragma unroll
    for (int c = 0; c < NV; ++c) acc[i][c] = 0.0;
  double bfl[NV] = {0.0, 0.0, 0.0, 0.0};

  if (live) {
    if constexpr (MODE == MODE_SUMMARY) {
#pragma unroll
      for (int i = 0; i < NP; ++i) load4(a.q + ((int64_t)i * K + e) * 4, acc[i]);
    } else {
      ElemGeo g;
      load_geo_affine(a.geo, a.geo_stride, e, g);
      const int4 cn = reinterpret_cast<const int4*>(a.conn)[e];
      const int nbr[3] = {cn.x, cn.y, cn.z};
      const uint32_t code = (uint32_t)cn.w;

      // Volume term, node by node; the j loop stays rolled so only one node's
      // fluxes are live next to the 60 accumulators. c_drT/c_dsT/c_liftT are
      // column-major so each column is one run of scalar loads.
#pragma unroll 1
      for (int j = 0; j < NP; ++j) {
        double qj[4];
        load4(a.q + ((int64_t)j * K + e) * 4, qj);
        const Prim w = to_prim(qj, gm1);
        const double F[4] = {qj[1], qj[1] * w.u + w.p, qj[2] * w.u, (qj[3] + w.p) * w.u};
        const double G[4] = {qj[2], qj[1] * w.v, qj[2] * w.v + w.p, (qj[3] + w.p) * w.v};
        double Ft[4], Gt[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          Ft[c] = g.rx * F[c] + g.ry * G[c];
          Gt[c] = g.sx * F[c] + g.sy * G[c];
        }
#pragma unroll
        for (int i = 0; i < NP; ++i) {
          const double dr = c_drT[j * NP + i];
          const double ds = c_dsT[j * NP + i];
#pragma unroll
          for (int c = 0; c < 4; ++c) acc[i][c] = fma(-dr, Ft[c], fma(-ds, Gt[c], acc[i][c]));
        }
        // Face slots whose own node is j (at most two: the corners).
        const int slots = NODE_SLOTS[j];
#pragma unroll 1
        for (int tt = 0; tt < 2; ++tt) {
          const int s = (slots >> (8 * tt)) & 0xff;
          if (s == 0xff) break;
          const int f = s / NFP;
          const int kk = s - f * NFP;
          const double nx = f == 0 ? g.nx[0] : (f == 1 ? g.nx[1] : g.nx[2]);
          const double ny = f == 0 ? g.ny[0] : (f == 1 ? g.ny[1] : g.ny[2]);
          const double fs = f == 0 ? g.fs[0] : (f == 1 ? g.fs[1] : g.fs[2]);
          const int nb = f == 0 ? nbr[0] : (f == 1 ? nbr[1] : nbr[2]);
          Side m;
#pragma unroll
          for (int c = 0; c < 4; ++c) {
            m.q[c] = qj[c];
            m.f[c] = F[c] * nx + G[c] * ny;  // f_minus as in dg.py:161
          }
          m.w = w;
          m.c = sqrt(gamma * w.p / w.rho);
          m.un = w.u * nx + w.v * ny;
          const uint32_t fc = (code >> (8 * f)) & 0xffu;
          const int kind = (fc >> 3) & 7;
          double qp[4];
          if (kind == BC_INTERIOR) {
            const int nf = fc & 3;
            const int pk = (fc & 4) ? (NFP - 1 - kk) : kk;
            const int pn = face_node_dyn(nf * NFP + pk);
            load4(a.q + ((int64_t)pn * K + nb) * 4, qp);
          } else if (kind == BC_SLIP) {
            const double mn = qj[1] * nx + qj[2] * ny;
            qp[0] = qj[0];
            qp[1] = qj[1] - 2.0 * mn * nx;
            qp[2] = qj[2] - 2.0 * mn * ny;
            qp[3] = qj[3];
          } else if (kind == BC_FIXED || kind == BC_EXACT) {
            load4(a.bcs + ((int64_t)nb + (kind == BC_EXACT ? kk : 0)) * 4, qp);
          } else {
#pragma unroll
            for (int c = 0; c < 4; ++c) qp[c] = qj[c];
          }
          Side p;
          make_side(p, qp, nx, ny, gamma, gm1);
          double fh[4];
          if constexpr (FLUX == FLUX_DYNAMIC) {
            interface_flux_rt(flux_rt, m, p, nx, ny, gm1, fh);
          } else {
            interface_flux<FLUX>(m, p, nx, ny, gm1, fh);
          }
          double jump[4];
#pragma unroll
          for (int c = 0; c < 4; ++c) jump[c] = fs * (m.f[c] - fh[c]);
#pragma unroll
          for (int i = 0; i < NP; ++i) {
            const double l = c_liftT[s * NP + i];
#pragma unroll
            for (int c = 0; c < 4; ++c) acc[i][c] = fma(l, jump[c], acc[i][c]);
          }
          if (kind != BC_INTERIOR) {
            const double he = f == 0 ? g.he[0] : (f == 1 ? g.he[1] : g.he[2]);
            const double bw = c_facew[s] * he;
#pragma unroll
            for (int c = 0; c < 4; ++c) bfl[c] = fma(bw, fh[c], bfl[c]);
          }
        }
      }

      // Epilogue: write the RHS, or the SSP-RK3 combination of the stage.
#pragma unroll
      for (int i = 0; i < NP; ++i) {
        double* o = a.out + ((int64_t)i * K + e) * 4;
        if constexpr (MODE == MODE_RHS) {
          store4(o, acc[i]);
        } else {
          double qi[4];
          load4(a.q + ((int64_t)i * K + e) * 4, qi);
          double q0[4] = {0.0, 0.0, 0.0, 0.0};
          if constexpr (MODE != MODE_STAGE1) load4(a.q0 + ((int64_t)i * K + e) * 4, q0);
#pragma unroll
          for (int c = 0; c < 4; ++c) acc[i][c] = rk3_combine<MODE>(q0[c], qi[c], dt, acc[i][c]);
          store4(o, acc[i]);
        }
      }
    }
  }

  if constexpr (MODE != MODE_RHS) {
    if (live && a.summ) write_summary<MODE == MODE_STAGE3>(acc, gamma, a.summ, K, e);
  }
  if constexpr (MODE != MODE_SUMMARY) block_reduce_bflux<SB>(bfl, a.part + (int64_t)blk * PART_STRIDE);
}

// ---------------------------------------------------------------------------
// Limiter, pass 1: detection for every element from the stage summaries.
// ---------------------------------------------------------------------------
// Detection of element e from its summary row sm (limiting.py:136-151, and the
// positivity trigger of limiting.py:166-200): troubled flag, report values
// folded into row, and whether the element goes to the repair pass (entry:
// e, bit 31 set when troubled).
__device__ __forceinline__ bool detect_element(const LimitArgs& a, int64_t e, const double* sm,
                                               double row[PART_COLS], int& entry) {
  double avg[4];
  load4(sm, avg);
  const int fl = (int)sm[SUMM_FLAGS];
  bool troubled = false;
  if (a.limit_slopes) {
    double sn[3][4];
    troubled = detect_troubled<false>(a, e, avg, sm, sn);
    if (a.force) troubled = a.force[e] != 0;
  }
  const bool work = troubled || (fl & (SF_NEED_RHO | SF_NEED_P));
  const double pbar = pressure_of(avg, a.gamma - 1.0);
  const bool avg_ok = (avg[0] > 0.0) && (pbar > 0.0);
  if (troubled) row[COL_TROUBLED] += 1.0;
  if (work) {
    entry = (int)e | (troubled ? (int)0x80000000u : 0);
    if (!avg_ok) row[COL_ADM] = 0.0;  // finiteness is checked by the repair pass
  } else {
    row[COL_RMIN] = nan_min(row[COL_RMIN], sm[SUMM_RMIN]);
    row[COL_PMIN] = nan_min(row[COL_PMIN], sm[SUMM_PMIN]);
    if (!(avg_ok && !(fl & SF_NONFINITE))) row[COL_ADM] = 0.0;
    if (a.want_dt) row[COL_DT] = nan_min(row[COL_DT], a.dtg[e] / sm[SUMM_LAMBDA]);
  }
  if (a.flags) a.flags[e] = troubled ? 1 : 0;
  return work;
}

// The stage rows of detection block blk (lanes 0 .. stage_per_blk - 1 take
// one each).
__device__ __forceinline__ void fold_stage_rows(const LimitArgs& a, int blk, double row[PART_COLS]) {
  if (a.stage_rows && (int)threadIdx.x < a.stage_per_blk) {
    const int r = blk * a.stage_per_blk + (int)threadIdx.x;
    if (r < a.n_stage_rows) {
      const double* p = a.stage_rows + (int64_t)r * PART_STRIDE;
#pragma unroll
      for (int c = 0; c < PART_COLS; ++c) row[c] = part_combine(c, row[c], p[c]);
    }
  }
}

// Worklist append aggregated per block: one atomic on the shared counter per
// block that has entries, each lane's slot from its rank among the block's
// entries. Late in a run with strong discontinuities a quarter of the
// elements go to the worklist, and atomics on one address serialise (one per
// wave still cost 0.25 ms per pass at 2M elements). Called by every thread.
template <int NT>
__device__ __forceinline__ void append_work(int* count, int* list, bool want, int entry) {
  constexpr int NW = NT / 64;
  __shared__ int wbase[NW];
  __shared__ int bbase;
  const uint64_t m = __ballot(want);
  const int lane = (int)__lane_id(), w = (int)threadIdx.x >> 6;
  if (lane == 0) wbase[w] = __popcll(m);
  __syncthreads();
  if (threadIdx.x == 0) {
    int total = 0;
#pragma unroll
    for (int i = 0; i < NW; ++i) {
      const int n = wbase[i];
      wbase[i] = total;
      total += n;
    }
    bbase = total > 0 ? atomicAdd(count, total) : 0;
  }
  __syncthreads();
  if (want) list[bbase + wbase[w] + __popcll(m & ((1ull << lane) - 1ull))] = entry;
}

__global__ void __launch_bounds__(LBLOCK) k_limit_detect(LimitArgs a) {
  if (step_stopped(a.dtp)) return;
  const int blk = a.bv.lcase ? listed_block(a.bv, xcd_block(blockIdx.x, gridDim.x), 1)
                             : xcd_block_rev(blockIdx.x, gridDim.x);
  const int64_t e = (int64_t)blk * LBLOCK + threadIdx.x;
  double row[PART_COLS];
  part_neutral(row);
  bool live = e < a.K;
  const bool batched = a.bv.lcase != nullptr;
  if (batched && live) {
    const int cs = a.bv.lcase[blk];
    live = (a.bv.dt[cs] > 0.0) && (e - a.bv.begin[cs]) < a.bv.nreal[cs];
    a.limiter = a.bv.limiter[cs];
    a.limit_slopes = a.limiter != LIM_UNLIMITED;
  }
  // The summary pieces are stored piece-major: each of the seven 16-byte
  // loads is contiguous across the block's lanes.
  int entry = 0;
  bool want = false;
  if (live) {
    double sm[SUMM_STRIDE];
    load_summary(a.summ, a.K, e, sm);
    want = detect_element(a, e, sm, row, entry);
  }
  if (batched) {
    append_work<LBLOCK>(&a.bv.wcount[blk], a.work + (int64_t)blk * LBLOCK, want, entry);
  } else {
    append_work<LBLOCK>(a.work_count, a.work, want, entry);
  }
  fold_stage_rows(a, blk, row);
  block_reduce_row<LBLOCK>(row, a.part + (int64_t)blk * PART_STRIDE);
}

// Unlimited runs (no slopes, so no neighbour rows and no batched cases).
__global__ void __launch_bounds__(LBLOCK) k_limit_detect_direct(LimitArgs a) {
  if (step_stopped(a.dtp)) return;
  const int blk = xcd_block_rev(blockIdx.x, gridDim.x);
  const int64_t e = (int64_t)blk * LBLOCK + threadIdx.x;
  double row[PART_COLS];
  part_neutral(row);
  int entry = 0;
  bool want = false;
  if (e < a.K) {
    double sm[SUMM_STRIDE];
    load_summary(a.summ, a.K, e, sm);
    want = detect_element(a, e, sm, row, entry);
  }
  append_work<LBLOCK>(a.work_count, a.work, want, entry);
  fold_stage_rows(a, blk, row);
  block_reduce_row<LBLOCK>(row, a.part + (int64_t)blk * PART_STRIDE);
}

// ---------------------------------------------------------------------------
// Limiter, pass 2: reconstruction and positivity for worklist elements. A
// fixed grid strides over the worklist; every lane reaches the exit.
// ---------------------------------------------------------------------------
// Repair of worklist elements by groups of RG = 16 lanes: lane i of a group
// owns node i (lane 15 repeats node 14), so an element's repair takes the
// latency of one node's arithmetic instead of fifteen. Element-wide values
// (the detection, the P1 projections at the edge midpoints, the Cockburn-Shu
// and positivity scales) are evaluated with the operations and order of a
// one-lane loop over the nodes (projections summed in node order from the
// group's node buffer), so the repaired state does not depend on the grouping;
// minima, maxima and "all finite" over the nodes go through 16-lane shuffles.
// With a tape, worklist slot `slot` also records the pre-limit values, the
// neighbour deviations and the positivity scale for the adjoint.
constexpr int RG = 16;  // lanes per repaired element

// Group-local LDS exchange: a group lies inside one wave.
__device__ __forceinline__ void grp_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
// Group reductions: a group is one 16-lane DPP row, so rotations of the row
// by 8, 4, 2 and 1 lanes leave every lane with the whole group's result, as
// VALU moves (a lane shuffle goes through the LDS crossbar, a round trip per
// step on the repair's dependent chain). Min, max and or do not depend on
// the order of combination.
template <int R>
__device__ __forceinline__ int row_ror(int v) {
  return __builtin_amdgcn_mov_dpp(v, 0x120 + R, 0xF, 0xF, false);
}
template <int R>
__device__ __forceinline__ double row_ror(double v) {
  return __hiloint2double(row_ror<R>(__double2hiint(v)), row_ror<R>(__double2loint(v)));
}
static_assert(RG == 16, "group reductions rotate one 16-lane DPP row");
__device__ __forceinline__ double grp_min(double v) {
  v = nan_min(v, row_ror<8>(v));
  v = nan_min(v, row_ror<4>(v));
  v = nan_min(v, row_ror<2>(v));
  return nan_min(v, row_ror<1>(v));
}
__device__ __forceinline__ double grp_max(double v) {
  v = nan_max(v, row_ror<8>(v));
  v = nan_max(v, row_ror<4>(v));
  v = nan_max(v, row_ror<2>(v));
  return nan_max(v, row_ror<1>(v));
}
__device__ __forceinline__ bool grp_any(bool b) {
  int v = b ? 1 : 0;
  v |= row_ror<8>(v);
  v |= row_ror<4>(v);
  v |= row_ror<2>(v);
  v |= row_ror<1>(v);
  return v != 0;
}

// qs: the group's node buffer [NP][4]; ds: its 12 midpoint slopes.
__device__ __forceinline__ void repair_group(const LimitArgs& a, int entry, int64_t slot, int lane, double (*qs)[4],
                                             double* ds, double row[PART_COLS]) {
#pragma clang fp contract(off)
  const double gm1 = a.gamma - 1.0;
  const int64_t e = entry & 0x7fffffff;
  const bool troubled = entry < 0;
  const int64_t K = a.K;
  const int j = min(lane, NP - 1);
  double sm[10];  // pieces 0-4: cell average, midpoint (rho, E)
  load_summary<5>(a.summ, K, e, sm);
  double avg[4];
  load4(sm, avg);
  double q[4];
  load4(a.q + ((int64_t)j * K + e) * 4, q);
  double sn[3][4];
  if (troubled) detect_troubled<true>(a, e, avg, sm, sn);
  double* rec = nullptr;
  if (a.tape) {
    const int64_t pos = *a.tape_head + slot;
    if (pos < a.tape_cap) rec = a.tape + pos * TAPE_STRIDE;
  }
  if (rec) {
    if (lane == 0) {
      rec[TAPE_E] = (double)e;
      rec[TAPE_TROUBLED] = troubled ? 1.0 : 0.0;
#pragma unroll
      for (int f = 0; f < 3; ++f)
#pragma unroll
        for (int c = 0; c < 4; ++c) rec[TAPE_SN + 4 * f + c] = troubled ? sn[f][c] : 0.0;
    }
    if (lane < NP) store4(rec + TAPE_Q + 4 * lane, q);
  }
  if (troubled) {
    // Cockburn-Shu P1 reconstruction (limiting.py:153-163): lane 4f + c
    // projects component c at midpoint f, every lane rebuilds its node.
    store4(qs[j], q);
    grp_sync();
    if (lane < 12) {
      const int f = lane >> 2, c = lane & 3;
      double s = 0.0;
#pragma unroll
      for (int n = 0; n < NP; ++n) s = s + c_p1mid[f * NP + n] * qs[n][c];
      double snv = 0.0;
#pragma unroll
      for (int k = 0; k < 12; ++k) snv = k == lane ? sn[k >> 2][k & 3] : snv;
      const double proj = s - sm[c];
      double dv = 0.0;
      if (proj != 0.0) {
        const double ratio = snv / proj;
        dv = proj * fmin(1.0, limiter_phi(a.limiter, ratio));
      }
      ds[lane] = dv;
    }
    grp_sync();
    double d[3][4];
#pragma unroll
    for (int f = 0; f < 3; ++f)
#pragma unroll
      for (int c = 0; c < 4; ++c) d[f][c] = ds[4 * f + c];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const double pos = (fmax(d[0][c], 0.0) + fmax(d[1][c], 0.0)) + fmax(d[2][c], 0.0);
      const double neg = (fmax(-d[0][c], 0.0) + fmax(-d[1][c], 0.0)) + fmax(-d[2][c], 0.0);
      const double tp = pos > 0.0 ? fmin(1.0, neg / pos) : 0.0;
      const double tn = neg > 0.0 ? fmin(1.0, pos / neg) : 0.0;
#pragma unroll
      for (int f = 0; f < 3; ++f) d[f][c] = tp * fmax(d[f][c], 0.0) - tn * fmax(-d[f][c], 0.0);
      const double lin = (c_linb[j * 3 + 0] * d[0][c] + c_linb[j * 3 + 1] * d[1][c]) + c_linb[j * 3 + 2] * d[2][c];
      q[c] = avg[c] + lin;
    }
    grp_sync();  // the buffers are free for the group's next element
  }
  // Zhang-Shu positivity scaling (limiting.py:166-200).
  const double pbar = pressure_of(avg, gm1);
  const double eps = nan_min(1e-13, nan_min(avg[0], pbar));
  double scale = 1.0;
  const double rlow = grp_min(q[0]);
  if (rlow < eps) {
    const double span = avg[0] - rlow;
    double theta = (avg[0] - eps) / span;
    theta = fmin(fmax(theta, 0.0), 1.0);
    scale = theta;
#pragma unroll
    for (int c = 0; c < 4; ++c) q[c] = avg[c] + theta * (q[c] - avg[c]);
  }
  const bool need = pressure_of(q, gm1) < eps;
  if (grp_any(need)) {
    double lo = 1.0;
    if (need) {
      double dev[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) dev[c] = q[c] - avg[c];
      double l = 0.0, h = 1.0;
      for (int it = 0; it < 60; ++it) {
        const double mid = 0.5 * (l + h);
        double sv[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) sv[c] = avg[c] + mid * dev[c];
        if (pressure_of(sv, gm1) >= eps) {
          l = mid;
        } else {
          h = mid;
        }
      }
      lo = l;
    }
    const double theta = grp_min(lo);
    if (theta < 1.0) {
#pragma unroll
      for (int c = 0; c < 4; ++c) q[c] = avg[c] + theta * (q[c] - avg[c]);
      scale = scale * theta;
    }
  }
  if (rec && lane == 0) rec[TAPE_THETA] = scale;
  if (lane < NP) store4(a.q + ((int64_t)j * K + e) * 4, q);
  bool fin = true;
#pragma unroll
  for (int c = 0; c < 4; ++c) fin = fin && isfinite(q[c]);
  const bool finite = !grp_any(!fin);
  const double rmin = grp_min(q[0]);
  const double pmin = grp_min(pressure_of(q, gm1));
  row[COL_RMIN] = nan_min(row[COL_RMIN], rmin);
  row[COL_PMIN] = nan_min(row[COL_PMIN], pmin);
  if (!finite) row[COL_ADM] = 0.0;
  if (a.want_dt) {
    const double lam = grp_max(wave_speed(q, a.gamma));
    row[COL_DT] = nan_min(row[COL_DT], a.dtg[e] / lam);
  }
}

// Held to four waves per SIMD (128 VGPRs, from 164): the repair is a chain of
// dependent loads per element, so more groups in flight pay; late in the
// four-quadrant run (a quarter of the elements repaired per pass) the limiter
// went from 0.495 to 0.415 ms per stage. Five waves spilled and were slower.
__global__ void __launch_bounds__(FBLOCK) __attribute__((amdgpu_waves_per_eu(4))) k_limit_fix(LimitArgs a) {
  if (step_stopped(a.dtp)) return;
  __shared__ double qs[FBLOCK / RG][NP][4];
  __shared__ double ds[FBLOCK / RG][12];
  double row[PART_COLS];
  part_neutral(row);
  const int n = *a.work_count;
  if (a.count_out && blockIdx.x == 0 && threadIdx.x == 0) *a.count_out = n;
  if ((int)blockIdx.x * (FBLOCK / RG) >= n) {
    // No worklist entry reaches this block: its partial row is the neutral one.
    if (threadIdx.x == 0)
#pragma unroll
      for (int c = 0; c < PART_COLS; ++c) a.part[(int64_t)blockIdx.x * PART_STRIDE + c] = row[c];
    return;
  }
  const int g = threadIdx.x / RG, lane = threadIdx.x % RG;
  // A group repairs its entries one after another; the next entry is read
  // while the current one is repaired, so the list read leaves the chain of
  // dependent loads each repair starts with.
  const int stride = gridDim.x * (FBLOCK / RG);
  int idx = blockIdx.x * (FBLOCK / RG) + g;
  int entry = idx < n ? a.work[idx] : 0;
  for (; idx < n; idx += stride) {
    const int cur = entry;
    if (idx + stride < n) entry = a.work[idx + stride];
    repair_group(a, cur, idx, lane, qs[g], ds[g], row);
  }
  block_reduce_row<FBLOCK>(row, a.part + (int64_t)blockIdx.x * PART_STRIDE);
}

// Batched variant: block b repairs the worklist of 256-element block b, so
// each case's report rows stay contiguous; it re-arms its list counter.
__global__ void __launch_bounds__(FBLOCK) __attribute__((amdgpu_waves_per_eu(4))) k_limit_fix_blocks(LimitArgs a) {
  __shared__ double qs[FBLOCK / RG][NP][4];
  __shared__ double ds[FBLOCK / RG][12];
  double row[PART_COLS];
  part_neutral(row);
  const int b = listed_block(a.bv, blockIdx.x, 1);
  const int n = a.bv.wcount[b];
  if (n > 0) {
    const int cs = a.bv.lcase[b];
    a.limiter = a.bv.limiter[cs];
    const int g = threadIdx.x / RG, lane = threadIdx.x % RG;
    for (int idx = g; idx < n; idx += FBLOCK / RG)
      repair_group(a, a.work[(int64_t)b * LBLOCK + idx], -1, lane, qs[g], ds[g], row);
  }
  block_reduce_row<FBLOCK>(row, a.part + (int64_t)b * PART_STRIDE);
  if (threadIdx.x == 0) a.bv.wcount[b] = 0;
}

// Single-case limiter without a tape in one pass: block b detects its
// LBLOCK elements (as k_limit_detect) and its 16-lane groups then repair the
// block's own worklist, kept in LDS (as k_limit_fix). A repair writes only its
// element's nodes and detection reads only summaries, so no grid-wide barrier
// is needed between the two; one partial row per block. The block runs
// OBLOCK threads for its LBLOCK elements: the first LBLOCK detect, then all
// OBLOCK / RG groups repair, so the block's worklist (troubled elements
// cluster at the discontinuities) drains four times sooner than with one
// group per 16 elements; the summary rows are staged by all threads.
constexpr int OBLOCK = 1024;
__global__ void __launch_bounds__(OBLOCK) k_limit_block(LimitArgs a) {
  if (step_stopped(a.dtp)) return;
  const int blk = xcd_block(blockIdx.x, gridDim.x);
  const int64_t e = (int64_t)blk * LBLOCK + threadIdx.x;
  double row[PART_COLS];
  part_neutral(row);
  __shared__ double srow[LBLOCK][18];
  __shared__ int wl[LBLOCK];
  __shared__ int wn;
  __shared__ double qs[OBLOCK / RG][NP][4];
  __shared__ double ds[OBLOCK / RG][12];
  if (threadIdx.x == 0) wn = 0;
  {
    const int64_t r0 = (int64_t)blk * LBLOCK;
    const int nrow = (int)min((int64_t)LBLOCK, a.K - r0);
    const double2* src = reinterpret_cast<const double2*>(a.summ);
    for (int i = threadIdx.x; i < nrow * (SUMM_STRIDE / 2); i += OBLOCK) {
      const int p = i / nrow, r = i - p * nrow;
      const double2 v = src[(int64_t)p * a.K + r0 + r];
      srow[r][summ_lo(p)] = v.x;
      srow[r][summ_hi(p)] = v.y;
    }
    __syncthreads();
  }
  if (threadIdx.x < LBLOCK && e < a.K) {
    int entry = 0;
    if (detect_element(a, e, &srow[threadIdx.x][0], row, entry)) wl[atomicAdd(&wn, 1)] = entry;
  }
  __syncthreads();
  const int n = wn;
  const int g = threadIdx.x / RG, lane = threadIdx.x % RG;
  for (int idx = g; idx < n; idx += OBLOCK / RG) repair_group(a, wl[idx], -1, lane, qs[g], ds[g], row);
  fold_stage_rows(a, blk, row);
  block_reduce_row<OBLOCK>(row, a.part + (int64_t)blk * PART_STRIDE);
}

// ---------------------------------------------------------------------------
// CFL time step: min over elements of dtg / max_i(|u| + c) (time_integration.py:37-41).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(LBLOCK) k_dt(DtArgs a) {
  const int blk = xcd_block(blockIdx.x, gridDim.x);
  const int64_t e = (int64_t)blk * LBLOCK + threadIdx.x;
  double row[PART_COLS];
  part_neutral(row);
  if (e < a.K) {
    double lam = -INF;
#pragma unroll
    for (int i = 0; i < NP; ++i) {
      double q[4];
      load4(a.q + ((int64_t)i * a.K + e) * 4, q);
      const double ws = wave_speed(q, a.gamma);
      lam = i == 0 ? ws : nan_max(lam, ws);
    }
    row[COL_DT] = a.dtg[e] / lam;
  }
  block_reduce_row<LBLOCK>(row, a.part + (int64_t)blk * PART_STRIDE);
}

// Initial-state check (solver.py:171-176): rho_min, p_min and "all admissible".
__global__ void __launch_bounds__(LBLOCK) k_check(DtArgs a) {
  const int blk = xcd_block(blockIdx.x, gridDim.x);
  const int64_t e = (int64_t)blk * LBLOCK + threadIdx.x;
  double row[PART_COLS];
  part_neutral(row);
  if (e < a.K) {
    const double gm1 = a.gamma - 1.0;
#pragma unroll
    for (int i = 0; i < NP; ++i) {
      double q[4];
      load4(a.q + ((int64_t)i * a.K + e) * 4, q);
      const double p = pressure_of(q, gm1);
      bool fin = true;
#pragma unroll
      for (int c = 0; c < 4; ++c) fin = fin && isfinite(q[c]);
      if (!(fin && q[0] > 0.0 && p > 0.0)) row[COL_ADM] = 0.0;
      row[COL_RMIN] = nan_min(row[COL_RMIN], q[0]);
      row[COL_PMIN] = nan_min(row[COL_PMIN], p);
    }
  }
  block_reduce_row<LBLOCK>(row, a.part + (int64_t)blk * PART_STRIDE);
}

// Domain integral sum_k J_k sum_i w_i q[i,k,:] (dg.py:174-176), per block.
__global__ void __launch_bounds__(LBLOCK) k_integrate(DtArgs a) {
  const int blk = xcd_block(blockIdx.x, gridDim.x);
  const int64_t e = (int64_t)blk * LBLOCK + threadIdx.x;
  double row[PART_COLS];
  part_neutral(row);
  if (e < a.K) {
    double s[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int i = 0; i < NP; ++i) {
      double q[4];
      load4(a.q + ((int64_t)i * a.K + e) * 4, q);
#pragma unroll
      for (int c = 0; c < 4; ++c) s[c] += c_cellw[i] * q[c];
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) row[c] = s[c] * a.dtg[e];
  }
  block_reduce_row<LBLOCK>(row, a.part + (int64_t)blk * PART_STRIDE);
}

// ---------------------------------------------------------------------------
// Device-controlled runs: EulerSolver.run's loop body (solver.py:167-216)
// after each step, from the step's report rows (the host loop's fill_report
// and acceptance logic, operation for operation). Run by the thread of the
// last stage's report reduction that writes the report row.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void run_control(RunCtl* __restrict__ r, const double* __restrict__ rep) {
#pragma clang fp contract(off)
  if (r->status != RUN_GO) return;
  const double dt = r->dt;
  const double* b0 = rep;
  const double* b1 = rep + PART_STRIDE;
  const double* b2 = rep + 2 * PART_STRIDE;
  double bf[4];
  for (int k = 0; k < 4; ++k) bf[k] = dt * (b0[k] / 6.0 + b1[k] / 6.0 + 2.0 * b2[k] / 3.0);
  bool adm = true;
  double rmin = b0[COL_RMIN], pmin = b0[COL_PMIN], tmax = b0[COL_TROUBLED];
  for (int s = 0; s < 3; ++s) {
    const double* row = rep + s * PART_STRIDE;
    adm = adm && (row[COL_ADM] >= 1.0);
    if (row[COL_RMIN] != row[COL_RMIN] || row[COL_RMIN] < rmin) rmin = row[COL_RMIN];
    if (row[COL_PMIN] != row[COL_PMIN] || row[COL_PMIN] < pmin) pmin = row[COL_PMIN];
    if (row[COL_TROUBLED] > tmax) tmax = row[COL_TROUBLED];
  }
  if (!adm) {
    // Rejected: retry from the same input with half the step, up to
    // max_rejections times. The host re-enqueues from that input.
    ++r->rejected;
    if (r->attempt >= r->max_rejections) {
      r->status = RUN_REJECTED;
      r->fail_time = r->t;
      r->dt = 0.0;
      return;
    }
    ++r->attempt;
    r->reaches_end = 0;
    r->retry_dt = dt * 0.5;
    r->status = RUN_RETRY;
    r->dt = 0.0;
    return;
  }
  r->t = r->reaches_end ? r->final_time : r->t + dt;
  ++r->steps;
  for (int k = 0; k < 4; ++k) r->fint[k] += bf[k];
  r->mind = rmin < r->mind ? rmin : r->mind;
  r->minp = pmin < r->minp ? pmin : r->minp;
  const int64_t tr = (int64_t)tmax;
  if (tr > r->max_troubled) r->max_troubled = tr;
  r->attempt = 0;
  r->last_dt = dt;
  for (int k = 0; k < 4; ++k) r->last_bflux[k] = bf[k];
  r->last_rmin = rmin;
  r->last_pmin = pmin;
  r->last_troubled = tr;
  // Head of the next iteration.
  const double dt_next = r->cfl * b2[COL_DT];
  if (!(r->t < r->final_time) || (r->max_steps >= 0 && r->steps >= r->max_steps)) {
    r->status = RUN_DONE;
    r->next_dt = dt_next;
    r->dt = 0.0;
    return;
  }
  double d = dt_next;
  if (!isfinite(d) || d <= 0.0) {
    r->status = RUN_INVALID_DT;
    r->fail_time = r->t;
    r->fail_dt = d;
    r->dt = 0.0;
    return;
  }
  const double remaining = r->final_time - r->t;
  r->reaches_end = d >= remaining;
  if (r->reaches_end) d = remaining;
  r->dt = d;
}

// Resume after a rejected attempt: the next attempt runs with the halved dt.
__global__ void k_run_resume(RunCtl* __restrict__ r) {
  if (threadIdx.x != 0 || r->status != RUN_RETRY) return;
  r->status = RUN_GO;
  r->dt = r->retry_dt;
}

// Deterministic reduction of nrows partial rows into out[PART_COLS]. Each
// block reduces a fixed contiguous chunk; the last block to finish reduces
// the chunk results in order, writes out, and re-arms the counters (the
// block-completion counter and, if given, the limiter worklist counter).
// Device-controlled runs: a stopped step's reduction returns at once (dtp),
// and for the last stage (ctl) the thread that wrote out applies the loop
// body to the report rows at rep (out is one of them, hence not restrict).
__global__ void __launch_bounds__(LBLOCK) k_finalize(const double* __restrict__ part, int nrows,
                                                    double* __restrict__ chunk, int* done, int* work_count,
                                                    double* out, const double* dtp, RunCtl* ctl,
                                                    const double* rep) {
  if (step_stopped(dtp)) return;
  __shared__ int last;
  const int per = (nrows + gridDim.x - 1) / gridDim.x;
  const int r0 = blockIdx.x * per;
  const int r1 = min(nrows, r0 + per);
  double row[PART_COLS];
  part_neutral(row);
  for (int r = r0 + threadIdx.x; r < r1; r += LBLOCK) {
    const double* p = part + (int64_t)r * PART_STRIDE;
#pragma unroll
    for (int c = 0; c < PART_COLS; ++c) row[c] = part_combine(c, row[c], p[c]);
  }
  if (gridDim.x == 1) {
    // Small meshes reduce in one block: its result is the report row.
    block_reduce_row<LBLOCK>(row, out);
    if (threadIdx.x == 0 && work_count) *work_count = 0;
    if (threadIdx.x == 0 && ctl) run_control(ctl, rep);
    return;
  }
  block_reduce_row<LBLOCK>(row, chunk + (int64_t)blockIdx.x * PART_STRIDE);
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) last = (atomicAdd(done, 1) == (int)gridDim.x - 1);
  __syncthreads();
  if (!last) return;
  __threadfence();
  part_neutral(row);
  for (int b = threadIdx.x; b < (int)gridDim.x; b += LBLOCK) {
    const volatile double* p = chunk + (int64_t)b * PART_STRIDE;
#pragma unroll
    for (int c = 0; c < PART_COLS; ++c) row[c] = part_combine(c, row[c], p[c]);
  }
  block_reduce_row<LBLOCK>(row, out);
  if (threadIdx.x == 0) {
    *done = 0;
    if (work_count) *work_count = 0;
    if (ctl) run_control(ctl, rep);
  }
}

// Pointwise evaluation of an interface flux on n face states (API tests).
template <int FLUX>
__global__ void k_flux_pointwise(int64_t n, const double* __restrict__ qm, const double* __restrict__ qp,
                                 const double* __restrict__ nx, const double* __restrict__ ny, double gamma,
                                 double* __restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const double gm1 = gamma - 1.0;
  double a[4], b[4];
  load4(qm + t * 4, a);
  load4(qp + t * 4, b);
  Side m, p;
  make_side(m, a, nx[t], ny[t], gamma, gm1);
  make_side(p, b, nx[t], ny[t], gamma, gm1);
  double f[4];
  interface_flux<FLUX>(m, p, nx[t], ny[t], gm1, f);
  store4(out + t * 4, f);
}

__global__ void k_limiter_pointwise(int64_t n, int id, const double* __restrict__ r, double* __restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) out[t] = limiter_phi(id, r[t]);
}

}  // namespace cfd

#include "cfd_stage4.inc"
#include "cfd_adjoint_kernels.inc"
#include "cfd_capi.inc"
#include "cfd_batched.inc"
#include "cfd_adjoint.inc"
#include "cfd_tables.inc"
