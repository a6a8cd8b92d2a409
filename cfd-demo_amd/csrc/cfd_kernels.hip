// gfx950 kernels for the degree-4 nodal DG Euler hot path:
//   k_stage   : strong-form RHS (volume + face fluxes + lift), fused with the
//               SSP-RK3 stage combination and a per-element limiter summary
//   k_limit   : troubled-cell detection / P1 reconstruction and Zhang-Shu
//               positivity, in place, touching only elements that need it
//   k_dt      : CFL time step reduction
//   k_finalize: deterministic reduction of per-block partials
//
// State layout in HBM is the reference's (Np, K, 4) float64, node-major:
// node j of element e starts at q + (j*K + e)*4. Consecutive lanes own
// consecutive elements, so every per-node load of a wave is one contiguous
// 2 KiB stretch. The 15x15 operators live in __constant__ memory and are
// read with wave-uniform indices, i.e. as scalar operands of the FMAs.
#include "cfd_device.hpp"

namespace cfd {

__constant__ double c_drT[NP * NP];    // Dr transposed: column j at [j*NP]
__constant__ double c_dsT[NP * NP];
__constant__ double c_liftT[NSLOT * NP];  // LIFT transposed: column s at [s*NP]
__constant__ double c_avgw[NP];    // cell weights / sum(cell weights)
__constant__ double c_cellw[NP];   // cell weights
__constant__ double c_facew[NSLOT];  // GLL edge weight of each face slot
__constant__ double c_p1mid[3 * NP];
__constant__ double c_linb[NP * 3];

// ---------------------------------------------------------------------------
// Stage kernel: one lane per element.
// ---------------------------------------------------------------------------
template <int FLUX, int MODE>
__global__ void __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(2)))
k_stage(StageArgs a) {
  const int blk = xcd_block(blockIdx.x, gridDim.x);
  const int64_t K = a.K;
  const int64_t e = (int64_t)blk * BLOCK + threadIdx.x;
  const bool live = e < K;
  const double gamma = a.gamma;
  const double gm1 = gamma - 1.0;

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
      load_geo(a.geo, e, g);
      const int4 cn = reinterpret_cast<const int4*>(a.conn)[e];
      const int nbr[3] = {cn.x, cn.y, cn.z};
      const uint32_t code = (uint32_t)cn.w;

      // Volume term, node by node; the j loop stays rolled so only one node's
      // fluxes are live next to the 60 accumulators. c_drT/c_dsT/c_liftT are
      // stored column-major so each column is one run of scalar loads.
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
        for (int t = 0; t < 2; ++t) {
          const int s = (slots >> (8 * t)) & 0xff;
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
          } else if (kind == BC_FIXED) {
            load4(a.bcs + (int64_t)nb * 4, qp);
          } else {
#pragma unroll
            for (int c = 0; c < 4; ++c) qp[c] = qj[c];
          }
          Side p;
          make_side(p, qp, nx, ny, gamma, gm1);
          double fh[4];
          interface_flux<FLUX>(m, p, nx, ny, gm1, fh);
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
      const double dt = a.dt;
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

  if constexpr (MODE == MODE_RHS) {
    block_sum4(bfl, a.bpart + (int64_t)blk * PART_STRIDE);
  } else {
    if (live) write_summary(acc, gamma, a.summ + e * SUMM_STRIDE);
    if constexpr (MODE != MODE_SUMMARY) block_sum4(bfl, a.bpart + (int64_t)blk * PART_STRIDE);
  }
}

// ---------------------------------------------------------------------------
// Limiter kernel: one lane per element, in place on the stage state.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(BLOCK) k_limit(LimitArgs a) {
  const int blk = xcd_block(blockIdx.x, gridDim.x);
  const int64_t K = a.K;
  const int64_t e = (int64_t)blk * BLOCK + threadIdx.x;
  double rmin = INF, pmin = INF;
  double troubled_count = 0.0, adm = 1.0;
  if (e < K) {
    const double* sm = a.summ + e * SUMM_STRIDE;
    double avg[4];
    load4(sm, avg);
    const double flags = sm[SUMM_FLAGS];
    const int fl = (int)flags;
    const double gm1 = a.gamma - 1.0;
    bool troubled = false;
    double sn[3][4];
    if (a.limit_slopes) troubled = detect_troubled(a, e, avg, sm, sn);
    const bool work = troubled || (fl & (SF_NEED_RHO | SF_NEED_P));
    if (work) {
      double q[NP][4];
      double* base = a.q;
#pragma unroll
      for (int i = 0; i < NP; ++i) load4(base + ((int64_t)i * K + e) * 4, q[i]);
      if (troubled) reconstruct_p1(q, avg, sn, a.limiter);
      positivity(q, avg, a.gamma);
      bool finite = true;
      rmin = INF;
      pmin = INF;
#pragma unroll
      for (int i = 0; i < NP; ++i) {
        store4(base + ((int64_t)i * K + e) * 4, q[i]);
#pragma unroll
        for (int c = 0; c < 4; ++c) finite = finite && isfinite(q[i][c]);
        rmin = nan_min(rmin, q[i][0]);
        pmin = nan_min(pmin, pressure_of(q[i], gm1));
      }
      adm = finite ? 1.0 : 0.0;
    } else {
      rmin = sm[SUMM_RMIN];
      pmin = sm[SUMM_PMIN];
      adm = (fl & SF_NONFINITE) ? 0.0 : 1.0;
    }
    const double pbar = pressure_of(avg, gm1);
    if (!(avg[0] > 0.0) || !(pbar > 0.0)) adm = 0.0;
    troubled_count = troubled ? 1.0 : 0.0;
    if (a.flags) a.flags[e] = troubled ? 1 : 0;
  }
  block_limit_reduce(rmin, pmin, troubled_count, adm, a.part + (int64_t)blk * PART_STRIDE + 4);
}

// ---------------------------------------------------------------------------
// CFL time step: min over elements of dtg / max_i(|u| + c) (time_integration.py:37-41).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(BLOCK) k_dt(DtArgs a) {
  const int blk = xcd_block(blockIdx.x, gridDim.x);
  const int64_t e = (int64_t)blk * BLOCK + threadIdx.x;
  double val = INF;
  if (e < a.K) {
#pragma clang fp contract(off)
    const double gm1 = a.gamma - 1.0;
    double lam = -INF;
    bool first = true;
#pragma unroll
    for (int i = 0; i < NP; ++i) {
      double q[4];
      load4(a.q + ((int64_t)i * a.K + e) * 4, q);
      const double u = q[1] / q[0];
      const double v = q[2] / q[0];
      const double p = gm1 * (q[3] - 0.5 * (q[1] * u + q[2] * v));
      const double ws = sqrt(u * u + v * v) + sqrt(a.gamma * p / q[0]);
      lam = first ? ws : nan_max(lam, ws);
      first = false;
    }
    val = a.dtg[e] / lam;
  }
  block_min1(val, a.part + (int64_t)blk * PART_STRIDE);
}

// Initial-state check (solver.py:171-176): per block [rho_min, p_min, all_ok].
__global__ void __launch_bounds__(BLOCK) k_check(DtArgs a) {
  const int blk = xcd_block(blockIdx.x, gridDim.x);
  const int64_t e = (int64_t)blk * BLOCK + threadIdx.x;
  double rmin = INF, pmin = INF, ok = 1.0;
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
      if (!(fin && q[0] > 0.0 && p > 0.0)) ok = 0.0;
      rmin = nan_min(rmin, q[0]);
      pmin = nan_min(pmin, p);
    }
  }
  block_limit_reduce(rmin, pmin, 0.0, ok, a.part + (int64_t)blk * PART_STRIDE);
}

// Domain integral sum_k J_k sum_i w_i q[i,k,:] (dg.py:174-176), per block.
__global__ void __launch_bounds__(BLOCK) k_integrate(DtArgs a) {
  const int blk = xcd_block(blockIdx.x, gridDim.x);
  const int64_t e = (int64_t)blk * BLOCK + threadIdx.x;
  double s[4] = {0.0, 0.0, 0.0, 0.0};
  if (e < a.K) {
#pragma unroll
    for (int i = 0; i < NP; ++i) {
      double q[4];
      load4(a.q + ((int64_t)i * a.K + e) * 4, q);
#pragma unroll
      for (int c = 0; c < 4; ++c) s[c] += c_cellw[i] * q[c];
    }
    const double J = a.dtg[e];
#pragma unroll
    for (int c = 0; c < 4; ++c) s[c] *= J;
  }
  block_sum4(s, a.part + (int64_t)blk * PART_STRIDE);
}

// Deterministic reduction of nb partial rows (PART_STRIDE doubles each).
// ops[c]: 0 = sum, 1 = nan-min, 2 = and (min of 0/1), 3 = max.
__global__ void __launch_bounds__(BLOCK) k_finalize(const double* __restrict__ part, int nb, int ncols,
                                                   int4 ops_lo, int4 ops_hi, double* __restrict__ out) {
  __shared__ double sh[BLOCK];
  const int ops[8] = {ops_lo.x, ops_lo.y, ops_lo.z, ops_lo.w, ops_hi.x, ops_hi.y, ops_hi.z, ops_hi.w};
  for (int c = 0; c < ncols; ++c) {
    const int op = ops[c];
    double v = (op == 0) ? 0.0 : (op == 1 || op == 2) ? INF : -INF;
    for (int b = threadIdx.x; b < nb; b += BLOCK) v = reduce_op(op, v, part[(int64_t)b * PART_STRIDE + c]);
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int w = BLOCK / 2; w > 0; w >>= 1) {
      if (threadIdx.x < w) sh[threadIdx.x] = reduce_op(op, sh[threadIdx.x], sh[threadIdx.x + w]);
      __syncthreads();
    }
    if (threadIdx.x == 0) out[c] = sh[0];
    __syncthreads();
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

#include "cfd_capi.inc"
