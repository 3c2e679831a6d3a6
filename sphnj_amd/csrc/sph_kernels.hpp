// SPH kernels of libsphnj for gfx950.  One thread per particle; sorted SoA arrays so that a
// wave's lanes walk the same or adjacent cells; every kernel is batch-aware (the case of a
// particle is found in the per-species case offset table) and templated on the precision
// policy (P64 / M32), so there is one code path per kernel.
//
// Neighbour order everywhere: the 9 cells of Sim::createCell (src/sim.cu:175-193) -- outer
// loop over the x offset, inner over the y offset -- and inside a cell fluid, then solid,
// then wall particles (the layout of Cell::particle_list, src/sort.cu:59-92), each in slot
// (= (cell, original index)) order.
#pragma once

#include "device.hpp"

namespace sphnj {

constexpr unsigned kStOutOfGrid = 1u, kStCellOver64 = 2u, kStNaN = 4u, kStRefListLeft = 8u;

__device__ __forceinline__ int cxy_pack(int cx, int cy) { return (cx << 16) | cy; }
__device__ __forceinline__ int cxy_x(int p) { return p >> 16; }
__device__ __forceinline__ int cxy_y(int p) { return p & 0xffff; }

// Position view of a particle: M32 stores float offsets from the origin of the particle's own
// cell; P64 stores the absolute double position.
template <class Pol, class CP>
__device__ __forceinline__ typename Pol::R view_x(double x, int cx, const CP &cp) {
  if constexpr (Pol::MIXED)
    return (float)(x - (cp.cxmin + (double)cx * cp.cdx));
  else
    return x;
}
template <class Pol, class CP>
__device__ __forceinline__ typename Pol::R view_y(double y, int cy, const CP &cp) {
  if constexpr (Pol::MIXED)
    return (float)(y - (cp.cymin + (double)cy * cp.cdy));
  else
    return y;
}
// Offset to subtract from a particle view so that (neighbour view - base) is the separation
// when the neighbour sits in cell (cx + a, cy + b).
template <class Pol, class CP>
__device__ __forceinline__ typename Pol::R cell_shift_x(int a, const CP &cp) {
  if constexpr (Pol::MIXED)
    return (float)a * cp.cdx_r;
  else
    return 0.0;
}
template <class Pol, class CP>
__device__ __forceinline__ typename Pol::R cell_shift_y(int b, const CP &cp) {
  if constexpr (Pol::MIXED)
    return (float)b * cp.cdy_r;
  else
    return 0.0;
}

// ---- small 2x2 helpers ------------------------------------------------------------------
template <class R>
struct Mat2 {
  R a, b, c, d;
};
template <class R>
__device__ __forceinline__ Mat2<R> mmul(const Mat2<R> &x, const Mat2<R> &y) {
  return {x.a * y.a + x.b * y.c, x.a * y.b + x.b * y.d, x.c * y.a + x.d * y.c, x.c * y.b + x.d * y.d};
}

// ======================================================================================
// CELLS stage: hash, key sort (radix_sort.hpp), permute, cell ranges
// ======================================================================================

// findPositionCell + fillFluidCellID (src/pair.cu:8-27): key of every fluid slot, scattered to
// original-index order so that a stable sort over original order yields (cell, orig) order.
template <class Pol>
__global__ void __launch_bounds__(256) k_hash_fluid(Dev<Pol> d) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= d.nf) return;
  const int c = case_of(s, d.f_case_off, d.ncase);
  const auto &cp = d.cases[c];
  unsigned fl = 0;
  const int cx = axis_cell(d.fa.x[s], cp.cxmin, cp.inv_cdx, cp.nx, &fl);
  const int cy = axis_cell(d.fa.y[s], cp.cymin, cp.inv_cdy, cp.ny, &fl);
  const int o = d.f_orig[s];
  d.f_key_by_orig[o] = cp.cell_off + cy + cx * cp.ny;
  d.f_slot_of_orig[o] = s;
  if (fl) atomicOr(d.status, kStOutOfGrid);
}

// fillSolidCellID (src/pair.cu:29-36); solids stay in original order, the key array is the
// sort input directly.  Also refreshes the M32 cell-relative position view.
template <class Pol>
__global__ void __launch_bounds__(256) k_hash_solid(Dev<Pol> d, const double *x, const double *y, int *keys,
                                                    int *cxy_out, int update_view) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= d.ns) return;
  const int c = case_of(i, d.s_case_off, d.ncase);
  const auto &cp = d.cases[c];
  unsigned fl = 0;
  const int cx = axis_cell(x[i], cp.cxmin, cp.inv_cdx, cp.nx, &fl);
  const int cy = axis_cell(y[i], cp.cymin, cp.inv_cdy, cp.ny, &fl);
  keys[i] = cp.cell_off + cy + cx * cp.ny;
  if (cxy_out) cxy_out[i] = cxy_pack(cx, cy);
  if (update_view) {
    if constexpr (Pol::MIXED) {
      d.s_vx[i] = view_x<Pol>(x[i], cx, cp);
      d.s_vy[i] = view_y<Pol>(y[i], cy, cp);
    }
  }
  if (fl) atomicOr(d.status, kStOutOfGrid);
}

// Wall key (once, at init).
template <class Pol>
__global__ void __launch_bounds__(256) k_hash_wall(Dev<Pol> d, const double *x, const double *y, int *keys) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= d.nw) return;
  const int c = case_of(i, d.w_case_off, d.ncase);
  const auto &cp = d.cases[c];
  unsigned fl = 0;
  const int cx = axis_cell(x[i], cp.cxmin, cp.inv_cdx, cp.nx, &fl);
  const int cy = axis_cell(y[i], cp.cymin, cp.inv_cdy, cp.ny, &fl);
  keys[i] = cp.cell_off + cy + cx * cp.ny;
  if (fl) atomicOr(d.status, kStOutOfGrid);
}

// Gather the fluid state into sorted order (the zip permute of Fluid_List::sort,
// src/sort.cu:10-20 / include/particle.cuh:93-96), rebuild the views, reset the per-case
// maximum speed of this step.
template <class Pol>
__global__ void __launch_bounds__(256) k_permute_fluid(Dev<Pol> d, const int *sorted_vals) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < d.ncase) d.umax2[t] = 0ull;
  if (t >= d.nf) return;
  const int o = sorted_vals[t];
  const int s = d.f_slot_of_orig[o];
  const double x = d.fa.x[s], y = d.fa.y[s], u = d.fa.u[s], v = d.fa.v[s], rho = d.fa.rho[s];
  d.fb.x[t] = x;
  d.fb.y[t] = y;
  d.fb.u[t] = u;
  d.fb.v[t] = v;
  d.fb.rho[t] = rho;
  d.fb.p[t] = d.fa.p[s];
  d.f_phase2[t] = d.f_phase[s];
  d.f_cls2[t] = d.f_cls[s];
  d.f_orig[t] = o;
  const int c = case_of(t, d.f_case_off, d.ncase);
  const auto &cp = d.cases[c];
  unsigned fl = 0;
  const int cx = axis_cell(x, cp.cxmin, cp.inv_cdx, cp.nx, &fl);
  const int cy = axis_cell(y, cp.cymin, cp.inv_cdy, cp.ny, &fl);
  d.f_cxy[t] = cxy_pack(cx, cy);
  if constexpr (Pol::MIXED) {
    d.fb.vx[t] = view_x<Pol>(x, cx, cp);
    d.fb.vy[t] = view_y<Pol>(y, cy, cp);
    d.fb.vu[t] = (float)u;
    d.fb.vv[t] = (float)v;
    d.fb.vrho[t] = (float)rho;
  }
}

// Gather the static wall arrays into sorted order (init only).
template <class Pol>
__global__ void __launch_bounds__(256) k_permute_wall(Dev<Pol> d, const int *sorted_vals, const double *x0,
                                                      const double *y0, const int8_t *bc0) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= d.nw) return;
  const int o = sorted_vals[t];
  const double x = x0[o], y = y0[o];
  d.wx[t] = x;
  d.wy[t] = y;
  d.w_bc[t] = bc0[o];
  d.w_orig[t] = o;
  const int c = case_of(t, d.w_case_off, d.ncase);
  const auto &cp = d.cases[c];
  unsigned fl = 0;
  const int cx = axis_cell(x, cp.cxmin, cp.inv_cdx, cp.nx, &fl);
  const int cy = axis_cell(y, cp.cymin, cp.inv_cdy, cp.ny, &fl);
  d.w_cxy[t] = cxy_pack(cx, cy);
  d.w_vx[t] = view_x<Pol>(x, cx, cp);
  d.w_vy[t] = view_y<Pol>(y, cy, cp);
}

__global__ void __launch_bounds__(256) k_clear_ranges(int2 *r0, int2 *r1, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  r0[i] = make_int2(0, 0);
  if (r1) r1[i] = make_int2(0, 0);
}

// Cell [start, end) from the sorted keys (replaces reduce_by_key + exclusive_scan +
// updateCell*, src/sort.cu:16-19, 59-137).  Empty cells keep [0, 0).
__global__ void __launch_bounds__(256) k_cell_bounds(const int *keys, int n, int2 *range) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int k = keys[i];
  if (i == 0 || keys[i - 1] != k) range[k].x = i;
  if (i == n - 1 || keys[i + 1] != k) range[k].y = i + 1;
}

// The reference keeps at most CELL_SIZE = 64 particles per cell (include/global.h:9).
__global__ void __launch_bounds__(256) k_cell_audit(const int2 *fr, const int2 *sr, const int2 *wr, int n,
                                                    unsigned *status) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  bool over = false;
  if (i < n) {
    const int2 a = fr[i], b = sr[i], w = wr[i];
    over = (a.y - a.x) + (b.y - b.x) + (w.y - w.x) > 64;
  }
  if (__ballot(over ? 1 : 0) != 0ull && (threadIdx.x & 63) == 0) atomicOr(status, kStCellOver64);
}

// ======================================================================================
// WALL: normals (init) and Shepard update (every step)
// ======================================================================================

// getVirtNormal (src/virt.cu:8-48) over the wall cells.
template <class Pol>
__global__ void __launch_bounds__(256) k_wall_normal(Dev<Pol> d) {
  using R = typename Pol::R;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= d.nw) return;
  const int c = case_of(i, d.w_case_off, d.ncase);
  const auto &cp = d.cases[c];
  const int pk = d.w_cxy[i], cx = cxy_x(pk), cy = cxy_y(pk);
  const R xi = d.w_vx[i], yi = d.w_vy[i];
  R nx = 0, ny = 0;
  for (int a = -1; a <= 1; ++a) {
    const int jx = cx + a;
    if (jx < 0 || jx >= cp.nx) continue;
    const R bx = xi - cell_shift_x<Pol>(a, cp);
    for (int b = -1; b <= 1; ++b) {
      const int jy = cy + b;
      if (jy < 0 || jy >= cp.ny) continue;
      const R by = yi - cell_shift_y<Pol>(b, cp);
      const int2 r = d.w_range[cp.cell_off + jx * cp.ny + jy];
      for (int s = r.x; s < r.y; ++s) {
        const R dx = d.w_vx[s] - bx, dy = d.w_vy[s] - by;
        const R r2 = dx * dx + dy * dy;
        if (r2 < cp.rc2) {
          const R w = kexp<R>(-(r2 * cp.hinv2)) * cp.kdc;
          nx -= dx * w;
          ny -= dy * w;
        }
      }
    }
  }
  const R nn = sqrt(nx * nx + ny * ny);
  if (nn < R(0.1)) {  // EPS_FOR_SUM (include/global.h:6)
    d.w_n[2 * i] = 0;
    d.w_n[2 * i + 1] = 0;
  } else {
    d.w_n[2 * i] = nx / nn;
    d.w_n[2 * i + 1] = ny / nn;
  }
}

// getVirtData + virtDataCalculate (src/virt.cu:61-149): Shepard-interpolated velocity and
// hydrostatically corrected pressure from the fluid, slip / no-slip mirroring, phase of the
// nearest fluid particle (first strict minimum in iteration order), density from the EOS.
template <class Pol>
__global__ void __launch_bounds__(256) k_wall_update(Dev<Pol> d) {
  using R = typename Pol::R;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= d.nw) return;
  const int c = case_of(i, d.w_case_off, d.ncase);
  const auto &cp = d.cases[c];
  const int pk = d.w_cxy[i], cx = cxy_x(pk), cy = cxy_y(pk);
  const R xi = d.w_vx[i], yi = d.w_vy[i];
  const auto &F = d.fa;
  R S = 0, U = 0, V = 0, P = 0, nearest = 0;
  int phase = 0;
  bool found = false;
  for (int a = -1; a <= 1; ++a) {
    const int jx = cx + a;
    if (jx < 0 || jx >= cp.nx) continue;
    const R bx = xi - cell_shift_x<Pol>(a, cp);
    for (int b = -1; b <= 1; ++b) {
      const int jy = cy + b;
      if (jy < 0 || jy >= cp.ny) continue;
      const R by = yi - cell_shift_y<Pol>(b, cp);
      const int2 r = d.f_range[cp.cell_off + jx * cp.ny + jy];
      for (int s = r.x; s < r.y; ++s) {
        const R dx = F.vx[s] - bx, dy = F.vy[s] - by;
        const R r2 = dx * dx + dy * dy;
        if (!(r2 < cp.rc2)) continue;
        const R w = (kexp<R>(-(r2 * cp.hinv2)) - cp.eR) * cp.k2c;
        S += w;
        U += F.vu[s] * w;
        V += F.vv[s] * w;
        P += (F.p[s] - F.vrho[s] * (cp.gx * dx + cp.gy * dy)) * w;
        if (!found || r2 < nearest) {
          found = true;
          nearest = r2;
          phase = d.f_phase[s];
        }
      }
    }
  }
  R u = 0, v = 0, p = 0;
  if (S != R(0)) {
    u = U / S;
    v = V / S;
    p = P / S;
  }
  const R nx = d.w_n[2 * i], ny = d.w_n[2 * i + 1];
  const int bc = d.w_bc[i];
  if (bc == 0) {  // SLIP
    const R un = u * nx + v * ny;
    u = u - R(2) * un * nx;
    v = v - R(2) * un * ny;
  } else if (bc == 1) {  // NO_SLIP
    u = -u;
    v = -v;
  }
  d.w_u[i] = u;
  d.w_v[i] = v;
  d.w_p[i] = p;
  d.w_phase[i] = (int8_t)phase;
  // pressureToDensity (src/fluid.cu:6-13)
  d.w_rho[i] = (R)(cp.rho0_d[phase] * pow(((double)p - cp.p_bg_d) * cp.p2rho_d[phase] + 1.0, cp.ginv_d[phase]));
}

// ======================================================================================
// LDS neighbour tiles
// ======================================================================================
// A block owns blockDim.x consecutive sorted fluid slots, whose global keys lie in [k0, k1].
// The 3x3 neighbour cells of any of them that sit in neighbour column a (-1, 0, +1) have keys
// in [k0 + a ny - 1, k1 + a ny + 1]: a contiguous key range, hence a contiguous slot segment.
// The block stages the three segments in LDS once (coalesced) and every lane reads its
// candidates from there.  A tile that straddles two cases, spans more than kTileMaxCells
// cells per segment or holds more than the kernel's capacity falls back to global loads.
constexpr int kTileMaxCells = 96;

struct TileHdr {
  int ok;
  int c0, c1;  // cases of the first and last slot of the block
  int k0[3], nc[3], lo[3], hi[3], base[3];
};

template <class R>
struct V2;
template <>
struct V2<float> {
  using T = float2;
};
template <>
struct V2<double> {
  using T = double2;
};

// Every thread of the block must call this (it synchronises).  On return T.ok says whether
// the tile is staged; cr[a][k] holds the slot range of cell T.k0[a] + k.
template <class Pol>
__device__ __forceinline__ void tile_setup(const Dev<Pol> &d, int i0, int cap, TileHdr &T,
                                           int2 (*cr)[kTileMaxCells]) {
  const int t = threadIdx.x;
  if (t == 0) {
    const int il = min(i0 + (int)blockDim.x, d.nf) - 1;
    const int c0 = case_of(i0, d.f_case_off, d.ncase), c1 = case_of(il, d.f_case_off, d.ncase);
    const auto &cp = d.cases[c0];
    const int k0 = d.f_key[i0], k1 = d.f_key[il];
    const int cmin = cp.cell_off, cmax = cp.cell_off + cp.ncell - 1;
    for (int a = 0; a < 3; ++a) {
      const int ka = max(k0 + (a - 1) * cp.ny - 1, cmin), kb = min(k1 + (a - 1) * cp.ny + 1, cmax);
      T.k0[a] = ka;
      T.nc[a] = kb >= ka ? kb - ka + 1 : 0;
      T.lo[a] = 0x7fffffff;
      T.hi[a] = -1;
    }
    T.ok = (c0 == c1) && (k1 - k0 + 3 <= kTileMaxCells);
    T.c0 = c0;
    T.c1 = c1;
  }
  __syncthreads();
  if (T.ok) {
    for (int e = t; e < 3 * kTileMaxCells; e += blockDim.x) {
      const int a = e / kTileMaxCells, k = e - a * kTileMaxCells;
      if (k < T.nc[a]) {
        const int2 r = d.f_range[T.k0[a] + k];
        cr[a][k] = r;
        if (r.y > r.x) {
          atomicMin(&T.lo[a], r.x);
          atomicMax(&T.hi[a], r.y);
        }
      }
    }
  }
  __syncthreads();
  if (t == 0 && T.ok) {
    int base = 0;
    for (int a = 0; a < 3; ++a) {
      if (T.hi[a] < 0) {
        T.lo[a] = 0;
        T.hi[a] = 0;
      }
      T.base[a] = base;
      base += T.hi[a] - T.lo[a];
    }
    T.ok = base <= cap;
  }
  __syncthreads();
}

// Copy field f of the three segments into LDS (dst[base[a] + k] = f(lo[a] + k)).
template <class F>
__device__ __forceinline__ void tile_stage(const TileHdr &T, F &&f) {
  for (int a = 0; a < 3; ++a) {
    const int n = T.hi[a] - T.lo[a];
    for (int k = threadIdx.x; k < n; k += blockDim.x) f(T.base[a] + k, T.lo[a] + k);
  }
}

// Gradient weight w'(r^2) = exp(-r^2 / h^2) kd c32 (Kernel::initiate, src/kernel.cpp:37-40).
template <class R, class CP>
__device__ __forceinline__ R grad_w(R r2, const CP &cp) {
  if constexpr (sizeof(R) == 4)
    return __builtin_amdgcn_exp2f(r2 * cp.nhl2) * cp.kdc;
  else
    return exp(-(r2 * cp.hinv2)) * cp.kdc;
}

// Calls f(s) for s in [lo, hi), U at a time so the loads of U candidates are in flight together.
template <int U, class F>
__device__ __forceinline__ void for_range(int lo, int hi, F &&f) {
  int s = lo;
  for (; s + U <= hi; s += U) {
#pragma unroll
    for (int k = 0; k < U; ++k) f(s + k);
  }
  for (; s < hi; ++s) f(s);
}

// Per-case constants a neighbour loop needs, held in registers.
template <class R>
struct LoopConsts {
  R rc2, nhl2, hinv2, kdc;
  template <class CP>
  __device__ __forceinline__ explicit LoopConsts(const CP &cp)
      : rc2(cp.rc2), nhl2(cp.nhl2), hinv2(cp.hinv2), kdc(cp.kdc) {}
  // Gradient weight w'(r^2) = exp(-r^2/h^2) kd c32 inside the cut-off, exactly 0 outside, so a
  // candidate outside the cut-off adds exact zeros and the loop needs no branch.
  __device__ __forceinline__ R w(R r2) const {
    R e;
    if constexpr (sizeof(R) == 4)
      e = __builtin_amdgcn_exp2f(r2 * nhl2);
    else
      e = exp(-(r2 * hinv2));
    return r2 < rc2 ? e * kdc : R(0);
  }
};

// ======================================================================================
// FLUID: Riemann pair loop fused with the explicit update
// ======================================================================================

template <class R>
struct PairSums {
  R ax = 0, ay = 0, drho = 0, L00 = 0, L01 = 0, L10 = 0, L11 = 0, col = 0;
  int near = 0;
};

// Riemann terms of fluidFluidCalculate / fluidSolidCalculate / fluidVirtCalculate
// (src/fluid.cu:25-191), with the algebra rearranged around q = (u_i - u_j).dx:
//   du = q / r,  (u_i - v*).dx = rho_j q / (rho_i + rho_j) - (p_i - p_j) r / ((rho_i + rho_j) c).
template <class R, bool VISCID>
__device__ __forceinline__ void riemann_pair(R dx, R dy, R r2, R w, R ui, R vi, R uj, R vj, R rhoi, R rhoi_inv,
                                             R rhoj, R pi, R pj, R cs, R csinv, R mu8, R eps_h2, PairSums<R> &A) {
  const R inv_r = r2 > R(0) ? krsqrt<R>(r2) : R(0);
  const R rn = r2 * inv_r;  // |dx|
  const R du_ = ui - uj, dv_ = vi - vj;
  const R q = du_ * dx + dv_ * dy;
  const R isum = krcp<R>(rhoi + rhoj);
  R ps = (rhoi * pj + rhoj * pi) * isum;
  const R du = q * inv_r;
  if (du >= R(0)) ps += rhoi * rhoj * du * isum * fmin(R(3) * du, cs);
  const R t = (pi - pj) * isum * csinv;
  const R s = R(-2) * ps * rhoi_inv * w;
  A.ax += s * dx;
  A.ay += s * dy;
  A.drho += R(2) * rhoi * w * (rhoj * isum * q - t * rn);
  const R wdx = w * dx, wdy = w * dy;
  A.L00 -= du_ * wdx;
  A.L01 -= du_ * wdy;
  A.L10 -= dv_ * wdx;
  A.L11 -= dv_ * wdy;
  if constexpr (VISCID) {
    const R vis = -q / (r2 + eps_h2);
    const R tv = mu8 * vis * rhoi_inv * w;
    A.ax += tv * dx;
    A.ay += tv * dy;
  }
}

// LDS capacity of a tile (particles) for a block of BLOCK slots: about 3.4 x BLOCK particles
// are staged in a dense region; the rest is headroom for sparse or compressed tiles.
template <class Pol, int BLOCK>
struct PairCap {
  static constexpr int value = (Pol::MIXED ? 1536 : 1024) * BLOCK / 256;
};
template <int BLOCK>
struct ShiftCap {
  static constexpr int value = 2048 * BLOCK / 256;
};

// findFluidPair (src/fluid.cu:193-254) + fluidUpdate (src/fluid.cu:256-287).  Reads the
// current buffers (fa) of every neighbour and writes particle i's updated state into the
// alternate buffers (fb), so neighbours never see a half-updated state.  Also reduces the
// per-case max |u|^2 of the updated velocities (getMaxSpeed, src/fluid.cu:429-440).
template <class Pol, bool VISCID, int BLOCK>
__global__ void __launch_bounds__(BLOCK) k_fluid_pair_update(Dev<Pol> d) {
  using R = typename Pol::R;
  using R2 = typename V2<R>::T;
  constexpr int CAP = PairCap<Pol, BLOCK>::value;
  __shared__ R2 s_pos[CAP], s_vel[CAP], s_rp[CAP];
  __shared__ int s_fl[CAP];
  __shared__ int2 s_cr[3][kTileMaxCells];
  __shared__ TileHdr T;
  const int i0 = blockIdx.x * blockDim.x;
  const int i = i0 + threadIdx.x;
  const auto &A = d.fa;
  tile_setup(d, i0, CAP, T, s_cr);
  const bool staged = T.ok;
  if (staged) {
    tile_stage(T, [&](int l, int g) {
      s_pos[l] = R2{A.vx[g], A.vy[g]};
      s_vel[l] = R2{A.vu[g], A.vv[g]};
      s_rp[l] = R2{A.vrho[g], A.p[g]};
      s_fl[l] = (int)d.f_phase[g] | ((int)d.f_cls[g] << 8);
    });
  }
  __syncthreads();
  unsigned long long my_u2 = 0ull;
  const int c = T.c0 == T.c1 ? T.c0 : (i < d.nf ? case_of(i, d.f_case_off, d.ncase) : T.c1);
  if (i < d.nf) {
    const auto &cp = d.cases[c];
    const int pk = d.f_cxy[i], cx = cxy_x(pk), cy = cxy_y(pk);
    const R xi = A.vx[i], yi = A.vy[i], ui = A.vu[i], vi = A.vv[i], rhoi = A.vrho[i], pi = A.p[i];
    const int phi = d.f_phase[i];
    const bool liq = phi == 0;
    const R cs = cp.c[phi], csinv = cp.cinv[phi], mu8 = cp.mu8[phi];
    const R rhoi_inv = krcp<R>(rhoi);
    const LoopConsts<R> K(cp);
    const R eps_h2 = cp.eps_h2, qrc2 = cp.quarter_rc2;
    PairSums<R> S;
    // fluidFluidCalculate on one candidate (w = 0 outside the cut-off: exact zero terms)
    auto fluid_j = [&](int s, R dx, R dy, R r2, R uj, R vj, R rhoj, R pj, int fl) {
      const R w = K.w(r2);
      riemann_pair<R, VISCID>(dx, dy, r2, w, ui, vi, uj, vj, rhoi, rhoi_inv, rhoj, pi, pj, cs, csinv, mu8, eps_h2, S);
      if (!(liq && (fl & 0xff) == 1)) S.col += r2 * w;
      // FREE_SURFACE neighbour: HAS_NEAR (2) within half the cut-off, else HAS (1); the
      // reference's sticky update (src/fluid.cu:67-73) ends at the maximum over neighbours.
      const bool fs = r2 < K.rc2 && liq && s != i && (fl >> 8) == 1;
      S.near = max(S.near, fs ? (r2 < qrc2 ? 2 : 1) : 0);
    };
    for (int a = -1; a <= 1; ++a) {
      const int jx = cx + a;
      if (jx < 0 || jx >= cp.nx) continue;
      const R bx = xi - cell_shift_x<Pol>(a, cp);
      for (int b = -1; b <= 1; ++b) {
        const int jy = cy + b;
        if (jy < 0 || jy >= cp.ny) continue;
        const R by = yi - cell_shift_y<Pol>(b, cp);
        const int q = cp.cell_off + jx * cp.ny + jy;
        if (staged) {
          const int2 rf = s_cr[a + 1][q - T.k0[a + 1]];
          const int off = T.base[a + 1] - T.lo[a + 1];
          for_range<2>(rf.x, rf.y, [&](int s) {
            const R2 pj = s_pos[s + off], vj = s_vel[s + off], rp = s_rp[s + off];
            const R dx = pj.x - bx, dy = pj.y - by;
            fluid_j(s, dx, dy, dx * dx + dy * dy, vj.x, vj.y, rp.x, rp.y, s_fl[s + off]);
          });
        } else {
          const int2 rf = d.f_range[q];
          for (int s = rf.x; s < rf.y; ++s) {
            const R dx = A.vx[s] - bx, dy = A.vy[s] - by;
            const R r2 = dx * dx + dy * dy;
            if (!(r2 < K.rc2)) continue;
            fluid_j(s, dx, dy, r2, A.vu[s], A.vv[s], A.vrho[s], A.p[s], (int)d.f_phase[s] | ((int)d.f_cls[s] << 8));
          }
        }
        const int2 rs = d.s_range[q];
        for (int k = rs.x; k < rs.y; ++k) {  // fluidSolidCalculate: rho_j = rho_i
          const int j = d.s_sorted[k];
          const R dx = d.s_vx[j] - bx, dy = d.s_vy[j] - by;
          const R r2 = dx * dx + dy * dy;
          if (!(r2 < K.rc2)) continue;
          const R w = K.w(r2);
          riemann_pair<R, VISCID>(dx, dy, r2, w, ui, vi, (R)d.su[j], (R)d.sv[j], rhoi, rhoi_inv, rhoi, pi, d.s_p[j],
                                  cs, csinv, mu8, eps_h2, S);
        }
        const int2 rw = d.w_range[q];
        for (int s = rw.x; s < rw.y; ++s) {  // fluidVirtCalculate: rho_j = rho0[wall phase]
          const R dx = d.w_vx[s] - bx, dy = d.w_vy[s] - by;
          const R r2 = dx * dx + dy * dy;
          if (!(r2 < K.rc2)) continue;
          const R w = K.w(r2);
          const int phj = d.w_phase[s];
          riemann_pair<R, VISCID>(dx, dy, r2, w, ui, vi, d.w_u[s], d.w_v[s], rhoi, rhoi_inv, cp.rho0[phj], pi,
                                  d.w_p[s], cs, csinv, mu8, eps_h2, S);
          if (!(liq && phj == 1)) S.col += r2 * w;
        }
      }
    }
    d.f_col[i] = S.col;
    d.f_near[i] = (int8_t)S.near;
    d.f_L[4 * i + 0] = S.L00;
    d.f_L[4 * i + 1] = S.L01;
    d.f_L[4 * i + 2] = S.L10;
    d.f_L[4 * i + 3] = S.L11;
    // fluidUpdate, integrated in double
    const double x0 = A.x[i], y0 = A.y[i], u0 = A.u[i], v0 = A.v[i], r0 = A.rho[i];
    double ax = (double)S.ax, ay = (double)S.ay;
    if (cp.steady) {
      ax += -0.1 * u0;
      ay += -0.1 * v0;
    }
    double x = x0 + u0 * cp.dt, y = y0 + v0 * cp.dt;
    const double u = u0 + (ax + cp.gx_d) * cp.dt, v = v0 + (ay + cp.gy_d) * cp.dt;
    double rho = r0 + (double)S.drho * cp.dt;
    rho = fmax(rho, cp.rho_min_d[phi]);
    // densityToPressure (src/fluid.cu:15-23)
    const double p = fmax((pow(rho * cp.rho0_inv_d[phi], cp.gam_d[phi]) - 1.0) * cp.rho2p_d[phi] + cp.p_bg_d,
                          cp.p_min_d);
    x = fmin(fmax(x, cp.dxmin), cp.dxmax);
    y = fmin(fmax(y, cp.dymin), cp.dymax);
    const auto &B = d.fb;
    B.x[i] = x;
    B.y[i] = y;
    B.u[i] = u;
    B.v[i] = v;
    B.rho[i] = rho;
    B.p[i] = (R)p;
    if constexpr (Pol::MIXED) {
      B.vx[i] = view_x<Pol>(x, cx, cp);
      B.vy[i] = view_y<Pol>(y, cy, cp);
      B.vu[i] = (float)u;
      B.vv[i] = (float)v;
      B.vrho[i] = (float)rho;
    }
    const double u2 = u * u + v * v;
    my_u2 = (unsigned long long)__double_as_longlong(u2);
    if (isnan(rho) || isnan(x) || isnan(y) || isnan(u2)) atomicOr(d.status, kStNaN);
  }
  // per-case max |u|^2 (non-negative doubles order like their bit patterns): one atomic per
  // wave when the wave lies in one case, else one per lane
  const int c_lane0 = __shfl(c, 0, 64);
  if (__ballot(c != c_lane0 ? 1 : 0) == 0ull) {
    const unsigned long long m = wave_max_u64(my_u2);
    if ((threadIdx.x & 63) == 0 && m != 0ull) atomicMax(&d.umax2[c], m);
  } else if (i < d.nf && my_u2 != 0ull) {
    atomicMax(&d.umax2[c], my_u2);
  }
}

// ======================================================================================
// SHIFT: gradient pass fused with the shift update (ping-pong positions)
// ======================================================================================

// correctionCalculate (src/fluid.cu:316-367) + correctionUpdate / shiftDisplacementCalculate
// (src/fluid.cu:369-427).  The gradient sums over all species with no phase filter.
template <class Pol, int BLOCK>
__global__ void __launch_bounds__(BLOCK) k_fluid_shift(Dev<Pol> d, int pass) {
  using R = typename Pol::R;
  using R2 = typename V2<R>::T;
  constexpr int CAP = ShiftCap<BLOCK>::value;
  __shared__ R2 s_pos[CAP];
  __shared__ int2 s_cr[3][kTileMaxCells];
  __shared__ TileHdr T;
  const int i0 = blockIdx.x * blockDim.x;
  const int i = i0 + threadIdx.x;
  const auto &A = d.fa;
  tile_setup(d, i0, CAP, T, s_cr);
  const bool staged = T.ok;
  if (staged) tile_stage(T, [&](int l, int g) { s_pos[l] = R2{A.vx[g], A.vy[g]}; });
  __syncthreads();
  if (i >= d.nf) return;
  const int c = T.c0 == T.c1 ? T.c0 : case_of(i, d.f_case_off, d.ncase);
  const auto &cp = d.cases[c];
  if (pass >= cp.K) {  // this case has fewer passes than the batch: carry positions over
    d.fb.x[i] = A.x[i];
    d.fb.y[i] = A.y[i];
    if constexpr (Pol::MIXED) {
      d.fb.vx[i] = A.vx[i];
      d.fb.vy[i] = A.vy[i];
    }
    return;
  }
  const int pk = d.f_cxy[i], cx = cxy_x(pk), cy = cxy_y(pk);
  const R xi = A.vx[i], yi = A.vy[i];
  const LoopConsts<R> K(cp);
  R gx = 0, gy = 0;
  auto add = [&](R dx, R dy) {
    const R w = K.w(dx * dx + dy * dy);
    gx += dx * w;
    gy += dy * w;
  };
  for (int a = -1; a <= 1; ++a) {
    const int jx = cx + a;
    if (jx < 0 || jx >= cp.nx) continue;
    const R bx = xi - cell_shift_x<Pol>(a, cp);
    for (int b = -1; b <= 1; ++b) {
      const int jy = cy + b;
      if (jy < 0 || jy >= cp.ny) continue;
      const R by = yi - cell_shift_y<Pol>(b, cp);
      const int q = cp.cell_off + jx * cp.ny + jy;
      if (staged) {
        const int2 rf = s_cr[a + 1][q - T.k0[a + 1]];
        const int off = T.base[a + 1] - T.lo[a + 1];
        for_range<4>(rf.x + off, rf.y + off, [&](int s) {
          const R2 pj = s_pos[s];
          add(pj.x - bx, pj.y - by);
        });
      } else {
        const int2 rf = d.f_range[q];
        for (int s = rf.x; s < rf.y; ++s) add(A.vx[s] - bx, A.vy[s] - by);
      }
      const int2 rs = d.s_range[q];
      for (int k = rs.x; k < rs.y; ++k) {
        const int j = d.s_sorted[k];
        add(d.s_vx[j] - bx, d.s_vy[j] - by);
      }
      const int2 rw = d.w_range[q];
      for (int s = rw.x; s < rw.y; ++s) add(d.w_vx[s] - bx, d.w_vy[s] - by);
    }
  }
  const R umax = fmin((R)sqrt(__longlong_as_double((long long)d.umax2[c])), cp.umax_cap);
  const R dxs = -cp.shift_k * gx, dys = -cp.shift_k * gy;
  R sx = dxs, sy = dys;
  int cls;
  bool project = false;
  const R col = d.f_col[i];
  const int near = d.f_near[i];
  if (d.f_phase[i] == 1) {
    cls = 3;  // gas: INNER
  } else if (col < R(1.5)) {
    if (near == 0) {
      cls = 0;  // SPLASH
      sx = 0;
      sy = 0;
    } else {
      cls = 1;  // FREE_SURFACE
      project = true;
    }
  } else if (col < R(2.0) && near == 2) {
    cls = 2;  // VICINITY
    project = true;
  } else {
    cls = 3;  // INNER
  }
  if (project) {
    R nx = 0, ny = 0;
    const R g2 = gx * gx + gy * gy;
    if (g2 > R(0)) {
      const R ig = krsqrt<R>(g2);
      nx = gx * ig;
      ny = gy * ig;
    }
    if (sqrt(nx * nx + ny * ny) < R(0.1)) {
      nx = 0;
      ny = 0;
    }
    sx = (R(1) - nx * nx) * dxs - nx * ny * dys;
    sy = -ny * nx * dxs + (R(1) - ny * ny) * dys;
  }
  d.f_cls[i] = (int8_t)cls;
  sx *= umax;
  sy *= umax;
  const double x = A.x[i] + (double)sx, y = A.y[i] + (double)sy;
  const auto &B = d.fb;
  B.x[i] = x;
  B.y[i] = y;
  if constexpr (Pol::MIXED) {
    B.vx[i] = view_x<Pol>(x, cx, cp);
    B.vy[i] = view_y<Pol>(y, cy, cp);
  }
  const R *L = d.f_L + 4 * i;
  A.u[i] += (double)(L[0] * sx + L[1] * sy);
  A.v[i] += (double)(L[2] * sx + L[3] * sy);
}

// ======================================================================================
// SOLID
// ======================================================================================
// Solid kernels give each particle a group of kSolidLanes lanes: the lanes split the
// particle's candidates (cell ranges, reference list) and combine their partial sums with
// shuffles, so a few thousand solids still fill the chip.
constexpr int kSolidLanes = 16;

template <class T>
__device__ __forceinline__ T group_sum(T v) {
#pragma unroll
  for (int o = kSolidLanes / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kSolidLanes);
  return v;
}
__device__ __forceinline__ int group_or(int v) {
#pragma unroll
  for (int o = kSolidLanes / 2; o > 0; o >>= 1) v |= __shfl_xor(v, o, kSolidLanes);
  return v;
}

// solidCoupleUpdate (src/solid.cu:355-444): Shepard pressure from the fluid, normal from the
// solid neighbours in the current cells; and the check that every reference-list neighbour
// still lies within the 3x3 current cells.
template <class Pol>
__global__ void __launch_bounds__(256) k_solid_couple(Dev<Pol> d) {
  using R = typename Pol::R;
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = gid / kSolidLanes, l = gid % kSolidLanes;
  const bool active = i < d.ns;
  R nx = 0, ny = 0, P = 0, S = 0;
  int left = 0;
  int c = 0;
  if (active) {
    c = case_of(i, d.s_case_off, d.ncase);
    const auto &cp = d.cases[c];
    const int pk = d.s_cxy[i], cx = cxy_x(pk), cy = cxy_y(pk);
    const R xi = d.s_vx[i], yi = d.s_vy[i];
    const LoopConsts<R> K(cp);
    const R eR = cp.eR, k2c = cp.k2c, hinv2 = cp.hinv2;
    const auto &F = d.fa;
    for (int a = -1; a <= 1; ++a) {
      const int jx = cx + a;
      if (jx < 0 || jx >= cp.nx) continue;
      const R bx = xi - cell_shift_x<Pol>(a, cp);
      for (int b = -1; b <= 1; ++b) {
        const int jy = cy + b;
        if (jy < 0 || jy >= cp.ny) continue;
        const R by = yi - cell_shift_y<Pol>(b, cp);
        const int q = cp.cell_off + jx * cp.ny + jy;
        const int2 rf = d.f_range[q];
        for (int s = rf.x + l; s < rf.y; s += kSolidLanes) {
          const R dx = F.vx[s] - bx, dy = F.vy[s] - by;
          const R r2 = dx * dx + dy * dy;
          if (r2 < K.rc2) {
            const R w = (kexp<R>(-(r2 * hinv2)) - eR) * k2c;
            S += w;
            P += F.p[s] * w;
          }
        }
        const int2 rs = d.s_range[q];
        for (int k = rs.x + l; k < rs.y; k += kSolidLanes) {
          const int j = d.s_sorted[k];
          const R dx = d.s_vx[j] - bx, dy = d.s_vy[j] - by;
          const R w = K.w(dx * dx + dy * dy);
          nx -= dx * w;
          ny -= dy * w;
        }
      }
    }
    for (int e = d.s_ref_off[i] + l; e < d.s_ref_off[i + 1]; e += kSolidLanes) {
      const int pj = d.s_cxy[d.s_ref_j[e]];
      left |= abs(cxy_x(pj) - cx) > 1 || abs(cxy_y(pj) - cy) > 1;
    }
  }
  S = group_sum(S);
  P = group_sum(P);
  nx = group_sum(nx);
  ny = group_sum(ny);
  left = group_or(left);
  if (!active || l != 0) return;
  const auto &cp = d.cases[c];
  const R nn = sqrt(nx * nx + ny * ny);
  if (nn < R(0.1)) {
    d.s_n[2 * i] = 0;
    d.s_n[2 * i + 1] = 0;
  } else {
    d.s_n[2 * i] = nx / nn;
    d.s_n[2 * i + 1] = ny / nn;
  }
  d.s_p[i] = S != R(0) ? fmax(P / S, cp.p_min) : R(0);
  if (left) atomicOr(d.status, kStRefListLeft);
}

// plasticElasticFunction (src/solid.cu:80-108)
template <class R, class CP>
__device__ __forceinline__ R plastic_elastic(const CP &cp, R strain, R &max_strain) {
  const R ts = strain > R(-0.5) ? sqrt(R(1) + R(2) * strain) - R(1) : R(-1);
  if (ts > max_strain) return cp.aux_mod * (ts - max_strain) - cp.yield_stress;
  if (ts > -cp.harden_strain) {
    if (max_strain > ts && cp.plastic) max_strain = ts;
    return -cp.yield_stress;
  }
  if (max_strain > ts && cp.plastic) max_strain = ts;
  const R e = fabs(-cp.harden_strain - ts);
  return -(cp.harden_coef * e * e + cp.yield_stress);
}

// solidDeformationCalculate (src/solid.cu:169-209) in displacement form.  With H~ = sum
// (d_j - d_i) (w' dX)^T over the reference list (d = x - X) and B = sum dX (w' dX)^T, the
// reference's F = (B + H~) B^-1 = I + H~ B^-1, so F - I is formed without cancellation.
// B^-1 depends only on reference positions and is precomputed at init.  Stress as
// solidStressCalculation (src/solid.cu:131-167) plus the neo-Hookean extension.
template <class Pol>
__global__ void __launch_bounds__(256) k_solid_deform(Dev<Pol> d, int sub) {
  using R = typename Pol::R;
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = gid / kSolidLanes, l = gid % kSolidLanes;
  const bool active = i < d.ns;
  int c = 0;
  R h00 = 0, h01 = 0, h10 = 0, h11 = 0;
  if (active) {
    c = case_of(i, d.s_case_off, d.ncase);
    if (sub < d.cases[c].S) {
      const double dix = d.sx[i] - d.sX[i], diy = d.sy[i] - d.sY[i];
      for (int e = d.s_ref_off[i] + l; e < d.s_ref_off[i + 1]; e += kSolidLanes) {
        const int j = d.s_ref_j[e];
        const R gX = d.s_ref_g[2 * e], gY = d.s_ref_g[2 * e + 1];
        const R ux = (R)((d.sx[j] - d.sX[j]) - dix), uy = (R)((d.sy[j] - d.sY[j]) - diy);
        h00 += ux * gX;
        h01 += ux * gY;
        h10 += uy * gX;
        h11 += uy * gY;
      }
    }
  }
  h00 = group_sum(h00);
  h01 = group_sum(h01);
  h10 = group_sum(h10);
  h11 = group_sum(h11);
  if (!active || l != 0) return;
  const auto &cp = d.cases[c];
  if (sub >= cp.S) return;
  const R *Lr = d.s_Lref + 4 * i;
  const Mat2<R> L = {Lr[0], Lr[1], Lr[2], Lr[3]};
  Mat2<R> H;
  if (d.s_degen[i]) {  // det(B) == 0: L = I, F = B + H~
    const R *Bt = d.s_Bt + 4 * i;
    H = {Bt[0] + h00 - R(1), Bt[1] + h01, Bt[2] + h10, Bt[3] + h11 - R(1)};
  } else {
    H = mmul(Mat2<R>{h00, h01, h10, h11}, L);
  }
  const Mat2<R> F = {R(1) + H.a, H.b, H.c, R(1) + H.d};
  // Green strain E = (H + H^T + H^T H) / 2
  const R e00 = H.a + R(0.5) * (H.a * H.a + H.c * H.c);
  const R e11 = H.d + R(0.5) * (H.b * H.b + H.d * H.d);
  const R e01 = R(0.5) * (H.b + H.c + H.a * H.b + H.c * H.d);
  Mat2<R> P = {0, 0, 0, 0};
  const int mat = d.s_mat[i];
  R *Mp = d.s_M + 4 * i;
  if (mat == 0) {  // LINEAR_ELASTIC = St Venant-Kirchhoff (linearStressCalculation)
    const R trE = e00 + e11;
    const Mat2<R> S = {R(2) * cp.mu_s * e00 + cp.lambda * trE, R(2) * cp.mu_s * e01, R(2) * cp.mu_s * e01,
                       R(2) * cp.mu_s * e11 + cp.lambda * trE};
    P = mmul(F, S);
  } else if (mat == 1) {  // AUXETIC (auxeticStressCalculation, getMatrixEigenValues/Vectors)
    const R m = R(0.5) * (e00 + e11);
    const R rad = sqrt(R(0.25) * (e00 - e11) * (e00 - e11) + e01 * e01);
    const R al = R(0.5) * atan2(e01 + e01, e00 - e11);
    const R ca = cos(al), sa = sin(al);
    const Mat2<R> Rm = {ca, -sa, sa, ca};
    const R rdet = ca * ca + sa * sa;
    const Mat2<R> Ri = {ca / rdet, sa / rdet, -sa / rdet, ca / rdet};
    Mat2<R> M = {Mp[0], Mp[1], Mp[2], Mp[3]};
    Mat2<R> Mq = mmul(mmul(Ri, M), Rm);
    const R s0 = plastic_elastic<R>(cp, m + rad, Mq.a);
    const R s1 = plastic_elastic<R>(cp, m - rad, Mq.d);
    M = mmul(mmul(Rm, Mq), Ri);
    Mp[0] = M.a;
    Mp[1] = M.b;
    Mp[2] = M.c;
    Mp[3] = M.d;
    const Mat2<R> Sg = {s0, 0, 0, s1};
    P = mmul(mmul(mmul(F, Rm), Sg), Ri);
  } else if (mat == 2) {  // NEO_HOOKEAN extension: P = mu (F - F^-T) + lambda ln(J) F^-T
    const R jm1 = H.a + H.d + (H.a * H.d - H.b * H.c);  // J - 1
    const R J = R(1) + jm1;
    const R iJ = R(1) / J;
    const R lnJ = log1p(jm1);
    // F^-T = cof(F) / J = [[1 + H11, -H10], [-H01, 1 + H00]] / J
    const Mat2<R> FiT = {(R(1) + H.d) * iJ, -H.c * iJ, -H.b * iJ, (R(1) + H.a) * iJ};
    const Mat2<R> D = {(jm1 + H.a * J - H.d) * iJ, H.b + H.c * iJ, H.c + H.b * iJ, (jm1 + H.d * J - H.a) * iJ};
    P = {cp.mu_s * D.a + cp.lambda * lnJ * FiT.a, cp.mu_s * D.b + cp.lambda * lnJ * FiT.b,
         cp.mu_s * D.c + cp.lambda * lnJ * FiT.c, cp.mu_s * D.d + cp.lambda * lnJ * FiT.d};
  }
  const Mat2<R> PL = mmul(P, L);
  R *o = d.s_F + 4 * i;
  o[0] = F.a;
  o[1] = F.b;
  o[2] = F.c;
  o[3] = F.d;
  o = d.s_L + 4 * i;
  o[0] = L.a;
  o[1] = L.b;
  o[2] = L.c;
  o[3] = L.d;
  o = d.s_P + 4 * i;
  o[0] = P.a;
  o[1] = P.b;
  o[2] = P.c;
  o[3] = P.d;
  o = d.s_PL + 4 * i;
  o[0] = PL.a;
  o[1] = PL.b;
  o[2] = PL.c;
  o[3] = PL.d;
}

// solidForceCalculate (src/solid.cu:307-353: solidFluidCouple :211-255, solidSolidForce
// :257-305) fused with solidUpdate (:446-471); positions and velocities ping-pong.
template <class Pol>
__global__ void __launch_bounds__(256) k_solid_force_update(Dev<Pol> d, int sub) {
  using R = typename Pol::R;
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = gid / kSolidLanes, l = gid % kSolidLanes;
  const bool active = i < d.ns;
  int c = 0;
  bool run = false;
  R cpx = 0, cpy = 0, ax = 0, ay = 0;
  if (active) {
    c = case_of(i, d.s_case_off, d.ncase);
    const auto &cp = d.cases[c];
    run = sub < cp.S;
    if (run) {
      const int pk = d.s_cxy[i], cx = cxy_x(pk), cy = cxy_y(pk);
      const double xi = d.sx[i], yi = d.sy[i], ui = d.su[i], vi = d.sv[i];
      const int mat = d.s_mat[i];
      const R rr_inv = mat == 1 ? cp.rho_aux_inv : cp.rho_s_inv;
      const LoopConsts<R> K(cp);
      // fluid -> solid Riemann force (rho_i = rho_j, u_i = u_s, p_i = p_s, c = c[phase_j])
      const R xiv = view_x<Pol>(xi, cx, cp), yiv = view_y<Pol>(yi, cy, cp);
      const R uiv = (R)ui, viv = (R)vi, psi = d.s_p[i];
      const R c0 = cp.c[0], c1 = cp.c[1];
      const auto &F = d.fa;
      for (int a = -1; a <= 1; ++a) {
        const int jx = cx + a;
        if (jx < 0 || jx >= cp.nx) continue;
        const R bx = xiv - cell_shift_x<Pol>(a, cp);
        for (int b = -1; b <= 1; ++b) {
          const int jy = cy + b;
          if (jy < 0 || jy >= cp.ny) continue;
          const R by = yiv - cell_shift_y<Pol>(b, cp);
          const int2 rf = d.f_range[cp.cell_off + jx * cp.ny + jy];
          for (int s = rf.x + l; s < rf.y; s += kSolidLanes) {
            const R dx = F.vx[s] - bx, dy = F.vy[s] - by;
            const R r2 = dx * dx + dy * dy;
            if (!(r2 < K.rc2)) continue;
            const R w = K.w(r2);
            const R inv_r = r2 > R(0) ? krsqrt<R>(r2) : R(0);
            const R rhoj = F.vrho[s];
            R ps = R(0.5) * (F.p[s] + psi);
            const R du = ((uiv - F.vu[s]) * dx + (viv - F.vv[s]) * dy) * inv_r;
            if (du >= R(0)) ps += R(0.5) * rhoj * du * fmin(R(3) * du, d.f_phase[s] ? c1 : c0);
            const R t = R(-2) * ps * rr_inv * w;
            cpx += t * dx;
            cpy += t * dy;
          }
        }
      }
      // solid -> solid over the reference-configuration list
      const R *Fi = d.s_F + 4 * i, *Li = d.s_L + 4 * i, *PLi = d.s_PL + 4 * i;
      const Mat2<R> FL = mmul(Mat2<R>{Fi[0], Fi[1], Fi[2], Fi[3]}, Mat2<R>{Li[0], Li[1], Li[2], Li[3]});
      const R pl0 = PLi[0], pl1 = PLi[1], pl2 = PLi[2], pl3 = PLi[3];
      const R h = cp.h, eps_h2 = cp.eps_h2, alpha = cp.alpha, beta = cp.beta, cs = cp.c_s;
      for (int e = d.s_ref_off[i] + l; e < d.s_ref_off[i + 1]; e += kSolidLanes) {
        const int j = d.s_ref_j[e];
        const R gX = d.s_ref_g[2 * e], gY = d.s_ref_g[2 * e + 1];
        const R *PLj = d.s_PL + 4 * j;
        ax += ((PLj[0] + pl0) * gX + (PLj[1] + pl1) * gY) * rr_inv;
        ay += ((PLj[2] + pl2) * gX + (PLj[3] + pl3) * gY) * rr_inv;
        const R dx = (R)(d.sx[j] - xi), dy = (R)(d.sy[j] - yi);
        const R du = (R)(d.su[j] - ui), dv = (R)(d.sv[j] - vi);
        const R nu = h * (du * dx + dv * dy) / (dx * dx + dy * dy + eps_h2);
        const R vis = alpha * nu * cs - beta * nu * nu;
        ax += vis * (FL.a * gX + FL.b * gY);
        ay += vis * (FL.c * gX + FL.d * gY);
      }
    }
  }
  cpx = group_sum(cpx);
  cpy = group_sum(cpy);
  ax = group_sum(ax);
  ay = group_sum(ay);
  if (!active || l != 0) return;
  const auto &cp = d.cases[c];
  if (!run) {  // this case has fewer sub-steps than the batch: carry the state over
    d.sx2[i] = d.sx[i];
    d.sy2[i] = d.sy[i];
    d.su2[i] = d.su[i];
    d.sv2[i] = d.sv[i];
    if constexpr (Pol::MIXED) {
      d.s_vx2[i] = d.s_vx[i];
      d.s_vy2[i] = d.s_vy[i];
    }
    return;
  }
  const int pk = d.s_cxy[i], cx = cxy_x(pk), cy = cxy_y(pk);
  const double xi = d.sx[i], yi = d.sy[i], ui = d.su[i], vi = d.sv[i];
  d.s_cp[2 * i] = cpx;
  d.s_cp[2 * i + 1] = cpy;
  double tax = (double)ax + (double)cpx, tay = (double)ay + (double)cpy;
  double x = xi, y = yi, u = ui, v = vi;
  if (d.s_bc[i] == 1) {  // FIXED_PARTICLE
    u = cp.fixed_u;
    v = cp.fixed_v;
    x += u * cp.dt_s;
    y += v * cp.dt_s;
    tax = 0.0;
    tay = 0.0;
  } else {
    if (cp.steady) {
      tax += -0.1 * ui;
      tay += -0.1 * vi;
    }
    x += u * cp.dt_s;
    y += v * cp.dt_s;
    u += (tax + cp.gx_d) * cp.dt_s;
    v += (tay + cp.gy_d) * cp.dt_s;
  }
  d.s_a[2 * i] = tax;
  d.s_a[2 * i + 1] = tay;
  d.sx2[i] = x;
  d.sy2[i] = y;
  d.su2[i] = u;
  d.sv2[i] = v;
  if constexpr (Pol::MIXED) {
    d.s_vx2[i] = view_x<Pol>(x, cx, cp);
    d.s_vy2[i] = view_y<Pol>(y, cy, cp);
  }
  if (isnan(x) || isnan(y)) atomicOr(d.status, kStNaN);
}

// ---- init: reference-configuration neighbour lists --------------------------------------

// Count of reference neighbours (|X_j - X_i|^2 < rc^2, 9 cells of the reference hash).
template <class Pol>
__global__ void __launch_bounds__(256) k_ref_count(Dev<Pol> d, const int *cxy, const int2 *range,
                                                   const int *sorted, const double *rc2_d, int *count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= d.ns) return;
  const int c = case_of(i, d.s_case_off, d.ncase);
  const auto &cp = d.cases[c];
  const int cx = cxy_x(cxy[i]), cy = cxy_y(cxy[i]);
  const double Xi = d.sX[i], Yi = d.sY[i], rc2 = rc2_d[c];
  int n = 0;
  for (int a = -1; a <= 1; ++a) {
    const int jx = cx + a;
    if (jx < 0 || jx >= cp.nx) continue;
    for (int b = -1; b <= 1; ++b) {
      const int jy = cy + b;
      if (jy < 0 || jy >= cp.ny) continue;
      const int2 r = range[cp.cell_off + jx * cp.ny + jy];
      for (int k = r.x; k < r.y; ++k) {
        const int j = sorted[k];
        const double dX = d.sX[j] - Xi, dY = d.sY[j] - Yi;
        if (dX * dX + dY * dY < rc2) ++n;
      }
    }
  }
  count[i] = n;
}

// Fill the lists (j, w'(dX) dX) and precompute B, B^-1 of the reference configuration.
template <class Pol>
__global__ void __launch_bounds__(256) k_ref_fill(Dev<Pol> d, const int *cxy, const int2 *range, const int *sorted,
                                                  const double *kdc_d, const double *hinv2_d, const double *rc2_d) {
  using R = typename Pol::R;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= d.ns) return;
  const int c = case_of(i, d.s_case_off, d.ncase);
  const auto &cp = d.cases[c];
  const double kdc = kdc_d[c], hinv2 = hinv2_d[c], rc2 = rc2_d[c];
  const int cx = cxy_x(cxy[i]), cy = cxy_y(cxy[i]);
  const double Xi = d.sX[i], Yi = d.sY[i];
  int e = d.s_ref_off[i];
  const int e_end = d.s_ref_off[i + 1];  // from k_ref_count; never write past it
  double b00 = 0, b01 = 0, b10 = 0, b11 = 0;
  for (int a = -1; a <= 1; ++a) {
    const int jx = cx + a;
    if (jx < 0 || jx >= cp.nx) continue;
    for (int b = -1; b <= 1; ++b) {
      const int jy = cy + b;
      if (jy < 0 || jy >= cp.ny) continue;
      const int2 r = range[cp.cell_off + jx * cp.ny + jy];
      for (int k = r.x; k < r.y; ++k) {
        const int j = sorted[k];
        const double dX = d.sX[j] - Xi, dY = d.sY[j] - Yi, r2 = dX * dX + dY * dY;
        if (r2 < rc2 && e < e_end) {
          const double w = exp(-(r2 * hinv2)) * kdc;
          d.s_ref_j[e] = j;
          d.s_ref_g[2 * e] = (R)(dX * w);
          d.s_ref_g[2 * e + 1] = (R)(dY * w);
          b00 += dX * dX * w;
          b01 += dX * dY * w;
          b10 += dY * dX * w;
          b11 += dY * dY * w;
          ++e;
        }
      }
    }
  }
  const double det = b00 * b11 - b01 * b10;
  R *Bt = d.s_Bt + 4 * i, *Lr = d.s_Lref + 4 * i;
  Bt[0] = (R)b00;
  Bt[1] = (R)b01;
  Bt[2] = (R)b10;
  Bt[3] = (R)b11;
  if (det != 0.0) {
    const double id = 1.0 / det;
    Lr[0] = (R)(b11 * id);
    Lr[1] = (R)(-b01 * id);
    Lr[2] = (R)(-b10 * id);
    Lr[3] = (R)(b00 * id);
    d.s_degen[i] = 0;
  } else {
    Lr[0] = 1;
    Lr[1] = 0;
    Lr[2] = 0;
    Lr[3] = 1;
    d.s_degen[i] = 1;
  }
}

// ---- inspection: neighbour lists of the fluid particles ---------------------------------
template <class Pol>
__global__ void __launch_bounds__(256) k_neighbors(Dev<Pol> d, int lo, int hi, const int *offset, int *count,
                                                   int *list, const int *limit) {
  using R = typename Pol::R;
  const int i = lo + blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= hi) return;
  const int c = case_of(i, d.f_case_off, d.ncase);
  const auto &cp = d.cases[c];
  const auto &A = d.fa;
  const int pk = d.f_cxy[i], cx = cxy_x(pk), cy = cxy_y(pk);
  const R xi = A.vx[i], yi = A.vy[i];
  int n = 0;
  int *out = list ? list + offset[i - lo] : nullptr;
  const int lim = limit ? limit[i - lo] : 0x7fffffff;  // count of the first pass
  for (int a = -1; a <= 1; ++a) {
    const int jx = cx + a;
    if (jx < 0 || jx >= cp.nx) continue;
    const R bx = xi - cell_shift_x<Pol>(a, cp);
    for (int b = -1; b <= 1; ++b) {
      const int jy = cy + b;
      if (jy < 0 || jy >= cp.ny) continue;
      const R by = yi - cell_shift_y<Pol>(b, cp);
      const int q = cp.cell_off + jx * cp.ny + jy;
      const int2 rf = d.f_range[q];
      for (int s = rf.x; s < rf.y; ++s) {
        const R dx = A.vx[s] - bx, dy = A.vy[s] - by;
        if (dx * dx + dy * dy < cp.rc2) {
          if (out && n < lim) out[n] = (0 << 28) | (d.f_orig[s] - cp.f_off);
          ++n;
        }
      }
      const int2 rs = d.s_range[q];
      for (int k = rs.x; k < rs.y; ++k) {
        const int j = d.s_sorted[k];
        const R dx = d.s_vx[j] - bx, dy = d.s_vy[j] - by;
        if (dx * dx + dy * dy < cp.rc2) {
          if (out && n < lim) out[n] = (1 << 28) | (j - cp.s_off);
          ++n;
        }
      }
      const int2 rw = d.w_range[q];
      for (int s = rw.x; s < rw.y; ++s) {
        const R dx = d.w_vx[s] - bx, dy = d.w_vy[s] - by;
        if (dx * dx + dy * dy < cp.rc2) {
          if (out && n < lim) out[n] = (2 << 28) | (d.w_orig[s] - cp.w_off);
          ++n;
        }
      }
    }
  }
  if (count) count[i - lo] = n;
}

// ---- FP64 reductions (fixed order: per-block partials, then one block) ------------------
template <class Pol>
__global__ void __launch_bounds__(256) k_energy_partial(Dev<Pol> d, int flo, int fhi, int slo, int shi,
                                                        const double *consts, double *part) {
  // consts: d^2, rho0_L, rho0_G, rho_s, rho_aux, lambda, mu_s
  __shared__ double sm[5][256];
  const int t = threadIdx.x;
  double ke = 0, mass = 0, m2 = 0, kes = 0, ep = 0;
  const double d2 = consts[0];
  for (int i = flo + blockIdx.x * 256 + t; i < fhi; i += gridDim.x * 256) {
    const double u = d.fa.u[i], v = d.fa.v[i], u2 = u * u + v * v;
    ke += 0.5 * consts[1 + d.f_phase[i]] * d2 * u2;
    mass += d.fa.rho[i] * d2;
    m2 = fmax(m2, u2);
  }
  for (int i = slo + blockIdx.x * 256 + t; i < shi; i += gridDim.x * 256) {
    const double u = d.su[i], v = d.sv[i];
    kes += 0.5 * (d.s_mat[i] == 1 ? consts[4] : consts[3]) * d2 * (u * u + v * v);
    const double f0 = d.s_F[4 * i], f1 = d.s_F[4 * i + 1], f2 = d.s_F[4 * i + 2], f3 = d.s_F[4 * i + 3];
    const double e00 = 0.5 * (f0 * f0 + f2 * f2 - 1.0), e11 = 0.5 * (f1 * f1 + f3 * f3 - 1.0);
    const double e01 = 0.5 * (f0 * f1 + f2 * f3);
    const double tr = e00 + e11;
    ep += 0.5 * consts[5] * tr * tr * d2 + consts[6] * (e00 * e00 + 2.0 * e01 * e01 + e11 * e11) * d2;
  }
  sm[0][t] = ke;
  sm[1][t] = mass;
  sm[2][t] = m2;
  sm[3][t] = kes;
  sm[4][t] = ep;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) {
      sm[0][t] += sm[0][t + w];
      sm[1][t] += sm[1][t + w];
      sm[2][t] = fmax(sm[2][t], sm[2][t + w]);
      sm[3][t] += sm[3][t + w];
      sm[4][t] += sm[4][t + w];
    }
    __syncthreads();
  }
  if (t < 5) part[blockIdx.x * 5 + t] = sm[t][0];
}

}  // namespace sphnj
