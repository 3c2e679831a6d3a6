// SPH kernels of libsphnj for gfx950.  One thread per particle; sorted SoA arrays so that a
// wave's lanes walk the same or adjacent cells; every kernel is batch-aware (the case of a
// particle is found in the per-species case offset table) and templated on the precision
// policy (P64 / M32), so there is one code path per kernel.
//
// Neighbour order of the wall, solid and inspection kernels: the 9 cells of Sim::createCell
// (src/sim.cu:175-193) -- outer loop over the x offset, inner over the y offset -- and inside
// a cell fluid (quadrant, original index), then solid, then wall particles (the layout of
// Cell::particle_list, src/sort.cu:59-92).  The fluid pair and shift kernels visit the same
// candidate set (the 3x3 cells of the particle's hashed cell) by half-cell sub-columns and
// skip sub-cells that provably lie beyond the cut-off; their sums differ from the cell order
// only in rounding.
#pragma once

#include "device.hpp"
#include "radix_sort.hpp"

namespace sphnj {

constexpr unsigned kStOutOfGrid = 1u, kStCellOver64 = 2u, kStNaN = 4u, kStRefListLeft = 8u, kStSyncTimeout = 16u;

// Packed as unsigned halves: a sub-column reaches 2 nx - 1 <= 65533, past the sign bit.
__device__ __forceinline__ int cxy_pack(int cx, int cy) { return (int)(((unsigned)cx << 16) | (unsigned)cy); }
__device__ __forceinline__ int cxy_x(int p) { return (int)((unsigned)p >> 16); }
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

// ---- the half-cell grid of the fluid order ----------------------------------------------
// Fluid slots are ordered on the half-cell ("sub-cell") grid, x-major: (case, sub-column
// sx, sub-row sy, original index), with sx = 2 cx + hx and sy = 2 cy + hy.  A run of
// consecutive sub-rows of one sub-column is one contiguous slot range.  A cell's fluids are
// two such runs, sub-column 2cx rows 2cy..2cy+1 then sub-column 2cx+1, which together list
// them in (quadrant, original index) order.

__device__ __forceinline__ int2 run_join(int2 a, int2 b) {
  if (a.y <= a.x) return b;
  if (b.y <= b.x) return a;
  return make_int2(a.x, b.y);
}

template <class Pol, class CP>
__device__ __forceinline__ void fluid_cell_runs(const Dev<Pol> &d, const CP &cp, int jx, int jy, int2 &r0,
                                                int2 &r1) {
  const int sny = 2 * cp.ny;
  const int k = 4 * cp.cell_off + 2 * jx * sny + 2 * jy;
  r0 = run_join(d.f_srange[k], d.f_srange[k + 1]);
  r1 = run_join(d.f_srange[k + sny], d.f_srange[k + sny + 1]);
}

// Run r (0 .. kCellRuns - 1) of the fluid in the 3x3 cells around cell (cx, cy): cell
// (cx + r / 6 - 1, cy + (r / 2) % 3 - 1), sub-column r % 2, both of its sub-rows.  False when the
// cell lies off the grid; else the run's slots and the lane's position relative to the run's
// cell origin (M32 views are relative to their own cell).
constexpr int kCellRuns = 18;
template <class Pol, class CP, class R>
__device__ __forceinline__ bool cell_run(const Dev<Pol> &d, const CP &cp, int cx, int cy, int r, R xi, R yi,
                                         int2 &rr, R &bx, R &by) {
  const int a = r / 6 - 1, b = (r >> 1) % 3 - 1, jx = cx + a, jy = cy + b;
  if (jx < 0 || jx >= cp.nx || jy < 0 || jy >= cp.ny) return false;
  const int sny = 2 * cp.ny;
  const int k = 4 * cp.cell_off + (2 * jx + (r & 1)) * sny + 2 * jy;
  rr = run_join(d.f_srange[k], d.f_srange[k + 1]);
  bx = xi - cell_shift_x<Pol>(a, cp);
  by = yi - cell_shift_y<Pol>(b, cp);
  return true;
}

// Sum over the aligned groups of W lanes (W a power of two dividing 64).
template <int W, class T>
__device__ __forceinline__ T lanes_sum(T v) {
#pragma unroll
  for (int o = W / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, W);
  return v;
}

// A fluid neighbour's fields as the solid loops read them.
template <class R>
struct FluidNb {
  R x, y, u, v, rho, p;
  int phase;
  template <class FA>
  __device__ __forceinline__ static FluidNb load(const FA &F, const int8_t *phase, int s) {
    return FluidNb{F.vx[s], F.vy[s], F.vu[s], F.vv[s], F.vrho[s], F.p[s], (int)phase[s]};
  }
};

// Position and pressure of a fluid neighbour.
template <class R>
struct R2c {
  R x, y, p;
};

// Calls use(s, v) for s in [lo, hi) in order, with v = load(s) read U slots at a time: the
// loads of U slots are issued together (indices clamped into the range), so a lane waits once
// per U slots instead of once per slot.
template <int U, class Ld, class F>
__device__ __forceinline__ void for_run(int lo, int hi, Ld &&load, F &&use) {
  for (int b = lo; b < hi; b += U) {
    decltype(load(0)) v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) v[u] = load(min(b + u, hi - 1));
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (b + u < hi) use(b + u, v[u]);
  }
}

// Case owning global cell i (cases hold consecutive cell ranges).
template <class Pol>
__device__ __forceinline__ int case_of_cell(int i, const Dev<Pol> &d) {
  int lo = 0, hi = d.ncase;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (d.cases[mid].cell_off <= i)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
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
// CELLS stage: hash, key sort (a counting sort for the fluid, radix_sort.hpp for the solids),
// permute, cell ranges
// ======================================================================================

// The step size of each case for the step that begins (the first thing of its cell stage),
// and the simulated time at its end: dt, or with adaptive_cfl C > 0
// dt_n = min(dt, C h / (c_max + U), S C h / c_s), U = sqrt(max |u|^2 after the last update).
// Threads t, t + n, ... of one block.
template <class Pol>
__device__ __forceinline__ void step_dt_body(const Dev<Pol> &d, int t, int n) {
  for (int c = t; c < d.ncase; c += n) {
    const auto &cp = d.cases[c];
    double dt = cp.dt;
    if (cp.adaptive) {
      dt = fmin(dt, cp.cfl_h / (cp.c_max + sqrt(d.u2_last[c])));
      if (cp.dt_solid_cap > 0.0) dt = fmin(dt, cp.dt_solid_cap);
    }
    d.dtv[c] = dt;
    d.tsim[c] += dt;
  }
}

// findPositionCell + fillFluidCellID (src/pair.cu:8-27) on the half-cell grid, and the count
// of a counting sort (replacing thrust::sort_by_key, src/sort.cu:10-20): every fluid slot's
// sub-cell key, (sy + sx * 2 ny) after the case's 4 ncell sub-cells offset, and its rank among
// its key's slots.  Slots sorted last step hold equal keys in runs, so each run of equal keys
// among a wave's consecutive lanes takes one atomic; the permute later fixes the order
// inside each key, which the atomics leave open.  A wave takes kCellRounds rounds of 64
// consecutive slots: the loads of all rounds, then the atomics of all rounds, then the
// stores, so that it waits on memory once rather than once per round.  Every lane of the
// block must reach the ballots.  With new_step, block 0 also sets the step size
// (step_dt_body), which no other block of this launch reads.
constexpr int kCellRounds = 4;
template <class Pol>
__global__ void __launch_bounds__(256) k_hash_fluid(Dev<Pol> d, const int *solids_moved, int new_step) {
  constexpr int U = kCellRounds;
  const int lane = threadIdx.x & 63;
  const int s0 = (blockIdx.x * 4 + (int)(threadIdx.x >> 6)) * (64 * U) + lane;  // round u: s0 + 64 u
  if (new_step && blockIdx.x == 0) step_dt_body(d, (int)threadIdx.x, (int)blockDim.x);
  // the solids' cell ranges start empty when the solids are sorted again (their bounds write the
  // rest); otherwise they stay as they are
  if (d.ns > 0 && (solids_moved == nullptr || *solids_moved != 0))
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < d.total_cells; q += gridDim.x * blockDim.x)
      d.s_range[q] = make_int2(0, 0);
  double px[U], py[U];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int s = min(s0 + 64 * u, d.nf - 1);
    px[u] = d.fa.x[s];
    py[u] = d.fa.y[s];
  }
  int key[U];
  unsigned fl = 0;
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int s = s0 + 64 * u;
    key[u] = -1 - lane;  // distinct for the lanes past the end
    if (s < d.nf) {
      const auto &cp = d.cases[case_of(s, d.f_case_off, d.f_case_blk, d.ncase)];
      const int sx = axis_sub(px[u], cp.cxmin, cp.inv_cdx, cp.nx, &fl);
      const int sy = axis_sub(py[u], cp.cymin, cp.inv_cdy, cp.ny, &fl);
      key[u] = 4 * cp.cell_off + sy + sx * (2 * cp.ny);
    }
  }
  if (fl) atomicOr(d.status, kStOutOfGrid);
  // runs of equal keys: a run starts where the key differs from the previous lane's
  int head[U], base[U];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int prev = __shfl_up(key[u], 1, 64);
    const unsigned long long heads = __ballot((lane == 0 || prev != key[u]) ? 1 : 0);
    const unsigned long long above = lane == 63 ? 0ull : heads >> (lane + 1);
    head[u] = 63 - __builtin_clzll(heads & ((2ull << lane) - 1ull));  // this lane's run start
    base[u] = 0;
    if (s0 + 64 * u < d.nf && head[u] == lane)
      base[u] = atomicAdd(&d.f_cnt[key[u]], above ? __builtin_ctzll(above) + 1 : 64 - lane);
  }
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int s = s0 + 64 * u, b = __shfl(base[u], head[u], 64);
    if (s < d.nf) {
      d.f_hkey[s] = key[u];
      d.f_hrank[s] = b + lane - head[u];
    }
  }
}

// Counting sort, step 2: the sum of each tile of kScanTile key counts (read in rounds of
// kSortThreads consecutive keys).
constexpr int kScanItems = 8, kScanTile = kSortThreads * kScanItems;
__global__ void __launch_bounds__(kSortThreads) k_count_tiles(const int *cnt, int nk, int *tsum) {
  __shared__ int tmp[4];
  const int k0 = blockIdx.x * kScanTile + (int)threadIdx.x;
  int v = 0;
#pragma unroll
  for (int q = 0; q < kScanItems; ++q)
    if (k0 + q * kSortThreads < nk) v += cnt[k0 + q * kSortThreads];
  int tot;
  block_exscan256(v, tmp, &tot);
  if (threadIdx.x == 0) tsum[blockIdx.x] = tot;
}

// Exclusive scan of n values in one block of 1024 threads; the total goes to out[n].
__global__ void __launch_bounds__(1024) k_exscan_small(const int *in, int n, int *out) {
  __shared__ int part[1024];
  const int t = threadIdx.x, per = (n + 1023) / 1024, a = min(t * per, n), b = min(a + per, n);
  int s = 0;
  for (int k = a; k < b; ++k) s += in[k];
  part[t] = s;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {  // Hillis-Steele inclusive scan
    const int v = t >= o ? part[t - o] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  int run = part[t] - s;
  for (int k = a; k < b; ++k) {
    out[k] = run;
    run += in[k];
  }
  if (t == 1023) out[n] = part[1023];
}

// Counting sort, step 3: each key's [start, end) slot range (empty: [0, 0)), in rounds of
// kSortThreads consecutive keys; the counts are cleared for the next step.  The tile's first
// slot comes from toff (the scanned tile sums), or, with up to kApplyPrefixMax tiles and
// toff == nullptr, from the block's own sum of the tile sums before it (no scan launch).
constexpr int kApplyPrefixMax = 4096;
__global__ void __launch_bounds__(kSortThreads) k_count_apply(int *cnt, int nk, const int *tsum, const int *toff,
                                                               int2 *range) {
  __shared__ int tmp[4];
  int run;
  if (toff != nullptr) {
    run = toff[blockIdx.x];
  } else {
    int before = 0;
    for (int k = (int)threadIdx.x; k < (int)blockIdx.x; k += kSortThreads) before += tsum[k];
    before = wave_sum(before);
    if ((threadIdx.x & 63) == 0) tmp[threadIdx.x >> 6] = before;
    __syncthreads();
    run = tmp[0] + tmp[1] + tmp[2] + tmp[3];
    __syncthreads();  // tmp is reused by the scans below
  }
  for (int q = 0; q < kScanItems; ++q) {
    const int k = blockIdx.x * kScanTile + q * kSortThreads + (int)threadIdx.x;
    const int c = k < nk ? cnt[k] : 0;
    int tot;
    const int st = run + block_exscan256(c, tmp, &tot);
    if (k < nk) {
      range[k] = c > 0 ? make_int2(st, st + c) : make_int2(0, 0);
      cnt[k] = 0;
    }
    run += tot;
  }
}

// Counting sort, step 4: every slot to its place in atomic order, with its original index
// and key (so f_pkey holds the sorted keys).  kCellRounds rounds of 256 slots per block,
// the loads of all rounds first.
template <class Pol>
__global__ void __launch_bounds__(256) k_count_scatter(Dev<Pol> d) {
  constexpr int U = kCellRounds;
  const int s0 = blockIdx.x * (256 * U) + (int)threadIdx.x;  // round u: s0 + 256 u
  int k[U], r[U], o[U];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int s = min(s0 + 256 * u, d.nf - 1);
    k[u] = d.f_hkey[s];
    r[u] = d.f_hrank[s];
    o[u] = d.f_orig[s];
  }
  int p[U];
#pragma unroll
  for (int u = 0; u < U; ++u) p[u] = d.f_srange[k[u]].x + r[u];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int s = s0 + 256 * u;
    if (s < d.nf) {
      d.f_perm[p[u]] = s;
      d.f_porig[p[u]] = o[u];
      d.f_pkey[p[u]] = k[u];
    }
  }
}

// fillSolidCellID (src/pair.cu:29-36); solids stay in original order, the key array is the
// sort input directly.  Also refreshes the M32 cell-relative position view.
template <class Pol>
__global__ void __launch_bounds__(256) k_hash_solid(Dev<Pol> d, const double *x, const double *y, int *keys,
                                                    int *cxy_out, int update_view, int *moved) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= d.ns) return;
  const int c = case_of(i, d.s_case_off, d.s_case_blk, d.ncase);
  const auto &cp = d.cases[c];
  unsigned fl = 0;
  const int cx = axis_cell(x[i], cp.cxmin, cp.inv_cdx, cp.nx, &fl);
  const int cy = axis_cell(y[i], cp.cymin, cp.inv_cdy, cp.ny, &fl);
  const int key = cp.cell_off + cy + cx * cp.ny;
  if (moved) {  // keys holds last step's keys: one atomic per wave with a change
    const unsigned long long ch = __ballot(keys[i] != key ? 1 : 0);
    if (ch != 0ull && (int)(threadIdx.x & 63) == __builtin_ctzll(__ballot(1))) atomicOr(moved, 1);
  }
  keys[i] = key;
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
  const int c = case_of(i, d.w_case_off, d.w_case_blk, d.ncase);
  const auto &cp = d.cases[c];
  unsigned fl = 0;
  const int cx = axis_cell(x[i], cp.cxmin, cp.inv_cdx, cp.nx, &fl);
  const int cy = axis_cell(y[i], cp.cymin, cp.inv_cdy, cp.ny, &fl);
  keys[i] = cp.cell_off + cy + cx * cp.ny;
  if (fl) atomicOr(d.status, kStOutOfGrid);
}

// The counting sort's last step fused with the gather of the fluid state into sorted order
// (the zip permute of Fluid_List::sort, src/sort.cu:10-20 / include/particle.cuh:93-96): a
// thread per scattered slot finds the slot's final place in its key and writes the particle
// there; also rebuilds the views and resets the per-case maximum speed slots and the path
// lengths of this step.
template <class Pol>
__global__ void __launch_bounds__(256) k_permute_fluid(Dev<Pol> d) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < d.ncase * kSpread) d.umax_slot[(size_t)i * kSlotStride] = 0ull;
  if (i < d.ncase) {
    d.umax2[i] = 0ull;
    d.umax[i] = 0;
  }
  if (i >= d.nf) return;
  // counting sort, step 5: inside each key, slots in original-index order -- a slot's place is
  // its key's start plus the number of the key's slots with a smaller original index -- so the
  // order is (sub-column, sub-row, original index) whatever order the atomics took
  const int k = d.f_pkey[i], o = d.f_porig[i], s = d.f_perm[i];
  // the particle's state is loaded before the rank loop, which does not need it
  const double x = d.fa.x[s], y = d.fa.y[s], u = d.fa.u[s], v = d.fa.v[s], rho = d.fa.rho[s];
  const auto p = d.fa.p[s];
  const int8_t phase = d.f_phase[s], cls = d.f_cls[s];
  const int2 r = d.f_srange[k];
  int rank = 0;
  for_run<4>(r.x, r.y, [&](int q) { return d.f_porig[q]; }, [&](int, int po) { rank += po < o ? 1 : 0; });
  const int t = r.x + rank;
  d.fb.x[t] = x;
  d.fb.y[t] = y;
  d.fb.u[t] = u;
  d.fb.v[t] = v;
  d.fb.rho[t] = rho;
  d.fb.p[t] = p;
  d.fb.cum[t] = 0.f;
  d.f_phase2[t] = phase;
  d.f_cls2[t] = cls;
  d.f_orig[t] = o;
  const int c = case_of(t, d.f_case_off, d.f_case_blk, d.ncase);
  const auto &cp = d.cases[c];
  unsigned fl = 0;
  const int sx = axis_sub(x, cp.cxmin, cp.inv_cdx, cp.nx, &fl);
  const int sy = axis_sub(y, cp.cymin, cp.inv_cdy, cp.ny, &fl);
  d.f_cxy[t] = cxy_pack(sx, sy);
  const int cx = sx >> 1, cy = sy >> 1;
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
  const int c = case_of(t, d.w_case_off, d.w_case_blk, d.ncase);
  const auto &cp = d.cases[c];
  unsigned fl = 0;
  const int cx = axis_cell(x, cp.cxmin, cp.inv_cdx, cp.nx, &fl);
  const int cy = axis_cell(y, cp.cymin, cp.inv_cdy, cp.ny, &fl);
  d.w_cxy[t] = cxy_pack(cx, cy);
  d.w_vx[t] = view_x<Pol>(x, cx, cp);
  d.w_vy[t] = view_y<Pol>(y, cy, cp);
}

__global__ void __launch_bounds__(256) k_clear_ranges(int2 *r0, int n0, int2 *r1, int n1) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n0) r0[i] = make_int2(0, 0);
  if (r1 && i < n1) r1[i] = make_int2(0, 0);
}

// Cell [start, end) from the sorted keys (replaces reduce_by_key + exclusive_scan +
// updateCell*, src/sort.cu:16-19, 59-137).  Empty cells keep [0, 0).
__global__ void __launch_bounds__(256) k_cell_bounds(const int *keys, int n, int2 *range, const int *only_if) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || (only_if != nullptr && *only_if == 0)) return;
  const int k = keys[i];
  if (i == 0 || keys[i - 1] != k) range[k].x = i;
  if (i == n - 1 || keys[i + 1] != k) range[k].y = i + 1;
}

// The reference keeps at most CELL_SIZE = 64 particles per cell (include/global.h:9).
template <class Pol>
__global__ void __launch_bounds__(256) k_cell_audit(Dev<Pol> d, unsigned *status) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  bool over = false;
  if (i < d.total_cells) {
    const int c = case_of_cell(i, d);
    const auto &cp = d.cases[c];
    const int q = i - cp.cell_off, jx = q / cp.ny, jy = q - jx * cp.ny;
    int2 r0, r1;
    fluid_cell_runs(d, cp, jx, jy, r0, r1);
    const int2 b = d.s_range[i], w = d.w_range[i];
    over = (r0.y - r0.x) + (r1.y - r1.x) + (b.y - b.x) + (w.y - w.x) > 64;
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
  const int c = case_of(i, d.w_case_off, d.w_case_blk, d.ncase);
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
// Each wall particle has a group of LANES lanes that split its fluid runs: 8 when the walls
// are few (tens of thousands, which alone would not fill the chip), fewer as they grow.
// Whatever the group size, the sums are formed in one order -- eight partial sums over the
// runs r = k, k + 8, k + 16 (k = 0..7), combined as an 8-lane butterfly -- so a wall's state
// does not depend on how many walls share the launch (e.g. a case alone or in a batch).
constexpr int kWallPartials = 8;
template <class Pol, int kWallLanes>
__device__ __forceinline__ void wall_update_block(const Dev<Pol> &d, int bid) {
  using R = typename Pol::R;
  constexpr int NP = kWallPartials / kWallLanes;  // partial sums per lane
  static_assert(NP * kWallLanes == kWallPartials, "group size divides the partial count");
  const int gid = bid * 256 + (int)threadIdx.x;
  const int i = gid / kWallLanes, l = gid % kWallLanes;
  const bool active = i < d.nw;
  R S[NP], U[NP], V[NP], P[NP], nearest = 0;
#pragma unroll
  for (int j = 0; j < NP; ++j) S[j] = U[j] = V[j] = P[j] = 0;
  int phase = 0, near_run = kCellRuns;  // run holding the lane's nearest fluid (none yet)
  int c = 0;
  if (active) {
    c = case_of(i, d.w_case_off, d.w_case_blk, d.ncase);
    const auto &cp = d.cases[c];
    const int pk = d.w_cxy[i], cx = cxy_x(pk), cy = cxy_y(pk);
    const R xi = d.w_vx[i], yi = d.w_vy[i];
    const auto &F = d.fa;
    // the 3x3 cells' 18 fluid runs: partial sum k = l + j * LANES takes runs k, k + 8, ...
#pragma unroll
    for (int j = 0; j < NP; ++j) {
      for (int r = l + j * kWallLanes; r < kCellRuns; r += kWallPartials) {
        int2 rr;
        R bx, by;
        if (!cell_run(d, cp, cx, cy, r, xi, yi, rr, bx, by)) continue;
        for (int s = rr.x; s < rr.y; ++s) {
          const R dx = F.vx[s] - bx, dy = F.vy[s] - by;
          const R r2 = dx * dx + dy * dy;
          if (!(r2 < cp.rc2)) continue;
          const R w = (kexp<R>(-(r2 * cp.hinv2)) - cp.eR) * cp.k2c;
          S[j] += w;
          U[j] += F.vu[s] * w;
          V[j] += F.vv[s] * w;
          P[j] += (F.p[s] - F.vrho[s] * (cp.gx * dx + cp.gy * dy)) * w;
          // first strict minimum of the serial scan: ties go to the earlier run
          if (near_run == kCellRuns || r2 < nearest || (r2 == nearest && r < near_run)) {
            near_run = r;
            nearest = r2;
            phase = d.f_phase[s];
          }
        }
      }
    }
  }
  // the butterfly's levels 4, 2, ... that pair partials of one lane, then those across lanes
#pragma unroll
  for (int h = NP / 2; h > 0; h >>= 1) {
#pragma unroll
    for (int j = 0; j < h; ++j) {
      S[j] += S[j + h];
      U[j] += U[j + h];
      V[j] += V[j + h];
      P[j] += P[j + h];
    }
  }
  const R Ssum = lanes_sum<kWallLanes>(S[0]), Usum = lanes_sum<kWallLanes>(U[0]);
  const R Vsum = lanes_sum<kWallLanes>(V[0]), Psum = lanes_sum<kWallLanes>(P[0]);
  // the nearest fluid of the whole scan: smallest r^2, ties to the earlier run (the scan order
  // of the cells, so the first strict minimum of a serial scan)
#pragma unroll
  for (int o = kWallLanes / 2; o > 0; o >>= 1) {
    const R n2 = __shfl_xor(nearest, o, kWallLanes);
    const int q2 = __shfl_xor(near_run, o, kWallLanes), p2 = __shfl_xor(phase, o, kWallLanes);
    if (q2 != kCellRuns && (near_run == kCellRuns || n2 < nearest || (n2 == nearest && q2 < near_run))) {
      nearest = n2;
      near_run = q2;
      phase = p2;
    }
  }
  if (!active || l != 0) return;
  const auto &cp = d.cases[c];
  R u = 0, v = 0, p = 0;
  if (Ssum != R(0)) {
    u = Usum / Ssum;
    v = Vsum / Ssum;
    p = Psum / Ssum;
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
template <class Pol, int kWallLanes>
__global__ void __launch_bounds__(256) k_wall_update(Dev<Pol> d) {
  wall_update_block<Pol, kWallLanes>(d, (int)blockIdx.x);
}

// ======================================================================================
// Tiles over the half-cell grid
// ======================================================================================
// Once per step (after the hash), the fluid slots of every cell column are cut into tiles
// of at most BLOCK lanes.  Both sub-columns of a cell column have the same candidate
// sub-columns (the 3x3 cells' six), so a tile takes the lanes of both over a band of R
// sub-rows: half the rows of a one-sub-column tile of as many lanes, so about half the
// candidates to stage per lane.  R is the cell column's largest band height (at most
// kBandMax) whose bands all hold at most BLOCK lanes; a band with none gets no tile, and one
// with more (a single overfull sub-row pair) is split into pieces of BLOCK lanes.  A cell
// column whose bands would need more tiles than its budget (2 per BLOCK slots of each
// sub-column) is cut per sub-column instead: chunks of BLOCK consecutive slots, a chunk whose
// slots span more than plan_span(BLOCK) sub-rows (a sparse sub-column, whose neighbours would
// overflow the LDS budget) split into kMaxPieces tiles of equal slot counts.  A tile's lanes
// are at most two slot runs: count0 lanes of the even sub-column from start, then the rest of
// the odd one from start1.  Every tile gets a plan: the candidate sub-columns
// [sc0, sc0 + nsc) and sub-rows [r0, r0 + nr) of its lanes -- the 3x3 cells of the tile's
// cell column over its rows -- as slot spans, and whether any solid or wall particle lies in
// those cells.  The row table f_lb gives, per sub-column, the first
// slot of every sub-row or later (2 ny + 1 entries, the last one the sub-column's end), so a
// row window [ra, rb] of a sub-column is the slot range [lb(ra), lb(rb + 1)).  The pair
// kernel and every shift pass reuse the plans: a block loads its plan, copies the spans and
// their row tables into LDS, and its lanes read their candidates there.  A tile whose spans
// exceed the kernel's LDS capacity reads them from global memory instead.
//
// Pruning.  A lane at position (tx, ty) in sub-cell units needs, in sub-column sc, only the
// rows that meet the disc of radius Rp around it, Rp = rc (1 + 1e-5) + D, where D bounds how
// far any candidate has moved since the hash (neighbours are binned by their hashed
// positions): the largest path length since the hash among the tile's candidates (0 right
// after the hash).  Every skipped candidate therefore lies beyond rc at the current
// positions -- by a margin far above the rounding of either precision -- and would have added
// exact zeros to every sum.
constexpr int kTileCols = 6;  // sub-columns of the 3x3 cells of a sub-column's cell column
constexpr int kBandMax = 48;  // sub-rows of a band at most (bounds the rows a tile stages)
// LDS budget of a tile of B slots: candidate particles and sub-rows.  A dense tile in one
// sub-column has about 6 B + 140 candidates (six sub-columns over its rows and three more on
// either side) over about B / 3.75 + 6 rows; a tile beyond the budget reads global memory.
__host__ __device__ constexpr int plan_cap(int B) { return 7 * B + 256; }
__host__ __device__ constexpr int plan_rows(int B) { return B / 2 + 32; }
// Sub-rows a tile of B slots may span: a dense sub-column holds about 3.9 slots per sub-row,
// so a full chunk spans about B / 3.9; a tile within the span has at most about
// 6 (plan_span + 7) 3.9 candidates, inside plan_cap.
__host__ __device__ constexpr int plan_span(int B) { return 9 * B / 32 + 2; }
constexpr int kMaxPieces = 2;

// Tiles of the chunk [start, start + count) of one sub-column.
__device__ __forceinline__ int chunk_pieces(const int *f_cxy, int start, int count, int B) {
  const int rows = cxy_y(f_cxy[start + count - 1]) - cxy_y(f_cxy[start]) + 1;
  return min(min((rows + plan_span(B) - 1) / plan_span(B), kMaxPieces), count);
}

// Lanes of a cell column in its sub-rows [a, b), from the row tables of its two sub-columns.
__device__ __forceinline__ int band_lanes(const int *lb0, const int *lb1, int a, int b) {
  return (lb0[b] - lb0[a]) + (lb1[b] - lb1[a]);
}
__device__ __forceinline__ int wave_max_int(int v) {
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
  return v;
}
// Band height of a cell column of ny2 sub-rows for tiles of B lanes (all lanes of a wave call
// it and get it): B over its densest sub-row pair, or the largest of the 8 heights above that
// whose bands all hold at most B (tested at once, 8 lanes per height).
__device__ __forceinline__ int band_height(const int *lb0, const int *lb1, int ny2, int B, int lane) {
  int dm = 0;
  for (int r = lane; r < ny2; r += 64) dm = max(dm, band_lanes(lb0, lb1, r, r + 1));
  dm = wave_max_int(dm);
  const int R0 = max(1, min(kBandMax, B / max(dm, 1))), h = R0 + 1 + (lane & 7);
  int m = h <= kBandMax ? 0 : B + 1;
  if (h <= kBandMax)
    for (int a = (lane >> 3) * h; a < ny2; a += 8 * h) m = max(m, band_lanes(lb0, lb1, a, min(a + h, ny2)));
  for (int o = 8; o < 64; o <<= 1) m = max(m, __shfl_xor(m, o, 64));  // over the lanes of height h
  const unsigned long long ok = __ballot(lane < 8 && m <= B);
  return ok ? R0 + 64 - __builtin_clzll(ok) : R0;
}
// Tiles of the bands of R sub-rows (all lanes of a wave; lane 0's sum is the count).
__device__ __forceinline__ int band_tiles(const int *lb0, const int *lb1, int ny2, int R, int B, int lane) {
  int t = 0;
  for (int a = lane * R; a < ny2; a += 64 * R) t += (band_lanes(lb0, lb1, a, min(a + R, ny2)) + B - 1) / B;
  return wave_sum(t);
}
template <int BLOCK>
struct TileRows {
  static constexpr int value = plan_rows(BLOCK);
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

struct TilePlan {
  int c, start, count;   // case, slot of lane 0, lanes
  int start1, count0;    // lanes [count0, count) are the slots from start1 on (odd sub-column)
  int sx, cx;            // even sub-column and cell column of the tile
  int sc0, nsc, r0, nr;  // candidate sub-columns and sub-rows
  int cy_ref;            // cell row of the M32 shift frame
  int sw;                // wall (bit 0) / solid (bit 1) particles lie in the tile's 3x3 cells
  int total;             // candidate particles (sum of the span lengths)
  int lo[kTileCols], len[kTileCols], row[kTileCols];  // slot span; f_lb index of (sc0 + a, r0)
  int gsz, gmul;         // staging: threads per span B / nsc; ceil(2^20 / gsz) (t / gsz = t gmul >> 20)
};

// Slot and sub-column of lane t of a tile.
__device__ __forceinline__ int lane_slot(const TilePlan &P, int t) {
  return t < P.count0 ? P.start + t : P.start1 + (t - P.count0);
}
__device__ __forceinline__ int lane_subcol(const TilePlan &P, int t) { return P.sx + (t < P.count0 ? 0 : 1); }

// The tiles of one block size: counts per global sub-column (a cell column's at its even
// one), their exclusive scan (G + 1 entries, the total last), how each cell column is cut and
// the plans.  The planning kernels take two jobs (by blockIdx.y), so the tiles of the pair and
// the shift block sizes are planned by the same launches.
struct PlanJob {
  int B;
  int *cnt, *off;
  int *band;  // per even global sub-column: the cell column's band height, 0 when cut per sub-column
  TilePlan *plans;
};
struct PlanJobs {
  PlanJob j[2];
};

// Tiles of block size B in the sub-column whose slots are [s0, s0 + n) when it is cut on its
// own (all lanes of a wave call it and get the count).
template <class Pol>
__device__ __forceinline__ int count_tiles(const Dev<Pol> &d, int s0, int n, int B, int lane) {
  const int nch = (n + B - 1) / B;
  int t = 0;
  for (int k = lane; k < nch; k += 64) t += chunk_pieces(d.f_cxy, s0 + k * B, min(B, n - k * B), B);
  return wave_sum(t);
}
// Tiles of block size B of the cell column whose sub-columns have the row tables lb0, lb1 (ny2
// sub-rows): in bands (*R = their height) while that needs at most 2 tiles per B slots of each
// sub-column, else per sub-column (*R = 0).  All lanes of a wave call it and get the count.
template <class Pol>
__device__ __forceinline__ int cell_column_tiles(const Dev<Pol> &d, const int *lb0, const int *lb1, int ny2, int B,
                                                 int lane, int *R) {
  const int n0 = lb0[ny2] - lb0[0], n1 = lb1[ny2] - lb1[0];
  const int budget = kMaxPieces * ((n0 + B - 1) / B + (n1 + B - 1) / B);
  const int h = band_height(lb0, lb1, ny2, B, lane);
  const int nb = band_tiles(lb0, lb1, ny2, h, B, lane);
  if (nb <= budget) {
    *R = h;
    return nb;
  }
  *R = 0;
  return count_tiles(d, lb0[0], n0, B, lane) + count_tiles(d, lb1[0], n1, B, lane);
}

// Wall row table w_lb of every cell column (once, after the walls are sorted; a thread per
// even global sub-column): the suffix minimum of the non-empty cells' starts, with the
// column's end past its last non-empty cell.
template <class Pol>
__global__ void __launch_bounds__(256) k_wall_rows(Dev<Pol> d, int G, const int *col_case) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= G) return;
  const auto &cp = d.cases[col_case[g]];
  const int sx = g - cp.col_off;
  if (sx & 1) return;
  const int qc = cp.cell_off + (sx >> 1) * cp.ny;
  int *lb = d.w_lb + cp.swp_off + (sx >> 1) * (cp.ny + 1);
  int end = 0;
  for (int k = 0; k < cp.ny; ++k) {
    const int2 r = d.w_range[qc + k];
    if (r.y > r.x) end = max(end, r.y);
  }
  int run = end;
  for (int k = cp.ny - 1; k >= 0; --k) {
    const int2 r = d.w_range[qc + k];
    if (r.y > r.x) run = min(run, r.x);
    lb[k] = run;
  }
  lb[cp.ny] = end;
}

// Row table of sub-column sx of case cp (all threads of a block of NW waves): f_lb[k] for row
// k = the suffix minimum of the non-empty rows' starts, with the sub-column's end past the last
// non-empty row.
template <class Pol, int NW, class CP>
__device__ __forceinline__ void subcol_row_table(const Dev<Pol> &d, const CP &cp, int sx, int *s_min, int *s_end) {
  constexpr int NT = 64 * NW;
  const int t = threadIdx.x, w = t >> 6, lane = t & 63;
  const int ny2 = 2 * cp.ny;
  const int kb = 4 * cp.cell_off + sx * ny2;
  int *lb = d.f_lb + cp.lb_off + sx * (ny2 + 1);
  // each thread takes a contiguous run of rows [k0, k1), held in registers when it has at most
  // kR of them (up to 2048 sub-rows with four waves, 1024 with one), else read again in the
  // backward pass
  constexpr int kR = NW == 1 ? 16 : 8;
  const int per = (ny2 + NT - 1) / NT, k0 = min(t * per, ny2), k1 = min(k0 + per, ny2);
  int lmin = 0x7fffffff, end = 0;
  int2 r[kR];
  for (int ka = k0; ka < k1; ka += kR) {
#pragma unroll
    for (int u = 0; u < kR; ++u) r[u] = ka + u < k1 ? d.f_srange[kb + ka + u] : make_int2(0, 0);
#pragma unroll
    for (int u = 0; u < kR; ++u)
      if (r[u].y > r[u].x) {
        lmin = min(lmin, r[u].x);
        end = max(end, r[u].y);
      }
  }
  for (int o = 32; o > 0; o >>= 1) end = max(end, __shfl_xor(end, o, 64));
  for (int o = 1; o < 64; o <<= 1) {  // suffix minimum over the wave's lanes
    const int u = __shfl_down(lmin, o, 64);
    if (lane + o < 64) lmin = min(lmin, u);
  }
  if (lane == 0) {
    s_min[w] = lmin;
    s_end[w] = end;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NW; ++k) end = max(end, s_end[k]);
  int run = __shfl_down(lmin, 1, 64);
  if (lane == 63) run = 0x7fffffff;
  for (int k = w + 1; k < NW; ++k) run = min(run, s_min[k]);
  run = min(run, end);  // rows past the last non-empty one: the end
  if (per <= kR) {
#pragma unroll
    for (int u = kR - 1; u >= 0; --u)
      if (k0 + u < k1) {
        if (r[u].y > r[u].x) run = min(run, r[u].x);
        lb[k0 + u] = run;
      }
  } else {
    for (int ka = k1 - 1; ka >= k0; --ka) {
      const int2 q = d.f_srange[kb + ka];
      if (q.y > q.x) run = min(run, q.x);
      lb[ka] = run;
    }
  }
  if (t == 0) lb[ny2] = end;
  __syncthreads();  // s_min, s_end free again; the table visible to the block
}

// Row tables of every cell column's two sub-columns (one block of NW waves per cell column,
// at its even global sub-column g = 2 blockIdx.x), then its tile counts for the njobs block
// sizes (how it is cut and how many tiles, at g; 0 at g + 1).  The block also audits the cell
// column's cells against the reference's 64-entry cell capacity (include/global.h:9) and
// builds its f_swp (when the batch has solids).  With four waves (a few thousand sub-columns)
// the waves build each row table together, then wave 0 counts the tiles while waves 1 and 2
// audit and index the cell column; with one (many small sub-columns, a batch of cases) the
// wave does all of it.
template <class Pol, int NW>
__device__ __forceinline__ void subcol_rows_block(const Dev<Pol> &d, int G, const int *col_case, const PlanJobs &jobs,
                                                  int njobs, int *reset_moved, int bid) {
  const int g = 2 * bid, t = threadIdx.x, w = t >> 6, lane = t & 63;
  if (g >= G) return;
  if (reset_moved != nullptr && g == 0 && t == 0) *reset_moved = 0;  // read by the earlier kernels only
  __shared__ int s_min[NW], s_end[NW];
  const auto &cp = d.cases[col_case[g]];
  const int sx = g - cp.col_off, ny2 = 2 * cp.ny;
  subcol_row_table<Pol, NW>(d, cp, sx, s_min, s_end);
  subcol_row_table<Pol, NW>(d, cp, sx + 1, s_min, s_end);
  if (w == 0) {  // the cell column's tiles
    const int *lb0 = d.f_lb + cp.lb_off + sx * (ny2 + 1), *lb1 = lb0 + (ny2 + 1);
    for (int j = 0; j < njobs; ++j) {
      int R;
      const int nt = cell_column_tiles(d, lb0, lb1, ny2, jobs.j[j].B, lane, &R);
      if (lane == 0) {
        jobs.j[j].cnt[g] = nt;
        jobs.j[j].band[g] = R;
        jobs.j[j].cnt[g + 1] = jobs.j[j].band[g + 1] = 0;
      }
    }
  }
  if (w == 1 % NW) {  // audits the cell column
    const int qc = cp.cell_off + (sx >> 1) * cp.ny, kc = 4 * cp.cell_off + sx * ny2;
    bool over = false;
    for (int k = lane; k < cp.ny; k += 64) {
      const int2 a = d.f_srange[kc + 2 * k], b = d.f_srange[kc + 2 * k + 1];
      const int2 e = d.f_srange[kc + ny2 + 2 * k], f = d.f_srange[kc + ny2 + 2 * k + 1];
      const int2 so = d.s_range[qc + k], wr = d.w_range[qc + k];
      const int n = max(a.y - a.x, 0) + max(b.y - b.x, 0) + max(e.y - e.x, 0) + max(f.y - f.x, 0) +
                    (so.y - so.x) + (wr.y - wr.x);
      over = over || n > 64;
    }
    if (__ballot(over ? 1 : 0) != 0ull && lane == 0) atomicOr(d.status, kStCellOver64);
  }
  if (w == 2 % NW && d.ns > 0) {  // ... and builds its f_swp
    constexpr int kU = 8;
    const int qc = cp.cell_off + (sx >> 1) * cp.ny;
    int *swp = d.f_swp + cp.swp_off + (sx >> 1) * (cp.ny + 1);
    const int pc = (cp.ny + 63) >> 6, c0 = min(lane * pc, cp.ny), c1 = min(c0 + pc, cp.ny);
    int n = 0;
    for (int ka = c0; ka < c1; ka += kU) {
      int2 a[kU];
#pragma unroll
      for (int u = 0; u < kU; ++u) a[u] = ka + u < c1 ? d.s_range[qc + ka + u] : make_int2(0, 0);
#pragma unroll
      for (int u = 0; u < kU; ++u) n += a[u].y > a[u].x ? 1 : 0;
    }
    int acc = wave_incl_sum(n, lane) - n;
    for (int ka = c0; ka < c1; ka += kU) {
      int2 a[kU];
#pragma unroll
      for (int u = 0; u < kU; ++u) a[u] = ka + u < c1 ? d.s_range[qc + ka + u] : make_int2(0, 0);
#pragma unroll
      for (int u = 0; u < kU; ++u)
        if (ka + u < c1) {
          acc += a[u].y > a[u].x ? 1 : 0;
          swp[ka + u + 1] = acc;
        }
    }
    if (lane == 0) swp[0] = 0;
  }
}
template <class Pol, int NW>
__global__ void __launch_bounds__(64 * NW) k_subcol_rows(Dev<Pol> d, int G, const int *col_case, PlanJobs jobs,
                                                         int njobs, int *reset_moved) {
  subcol_rows_block<Pol, NW>(d, G, col_case, jobs, njobs, reset_moved, (int)blockIdx.x);
}

// Tiles per cell column, counted at its even global sub-column (0 at the odd one), and how it
// is cut (one wave per global sub-column).
template <class Pol>
__global__ void __launch_bounds__(256) k_tile_count(Dev<Pol> d, int G, const int *col_case, PlanJobs jobs) {
  const int g = blockIdx.x * 4 + (int)(threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (g >= G) return;
  const PlanJob &J = jobs.j[blockIdx.y];
  const auto &cp = d.cases[col_case[g]];
  const int sx = g - cp.col_off, ny2 = 2 * cp.ny;
  if (sx & 1) {
    if (lane == 0) J.cnt[g] = J.band[g] = 0;
    return;
  }
  const int *lb0 = d.f_lb + cp.lb_off + sx * (ny2 + 1), *lb1 = lb0 + (ny2 + 1);
  int R;
  const int t = cell_column_tiles(d, lb0, lb1, ny2, J.B, lane, &R);
  if (lane == 0) {
    J.cnt[g] = t;
    J.band[g] = R;
  }
}

// Exclusive scan of the per-sub-column tile counts (one block of 1024 threads per job); the
// total goes to off[G].  Used when there are more than kPlanPrefixMax sub-columns; with fewer,
// k_tile_plan sums the counts before each of its blocks itself and this launch is skipped.
constexpr int kPlanPrefixMax = 4096;
__global__ void __launch_bounds__(1024) k_tile_scan(PlanJobs jobs, int G) {
  // wave w scans the contiguous segment [w seg, (w + 1) seg) in chunks of 64 (coalesced, four
  // chunks' loads in flight), after a block scan of the segments' totals
  constexpr int kU = 4;
  __shared__ int part[16];
  const int *cnt = jobs.j[blockIdx.x].cnt;
  int *off = jobs.j[blockIdx.x].off;
  const int t = threadIdx.x, w = t >> 6, lane = t & 63;
  const int seg = (G + 15) / 16, a = min(w * seg, G), b = min(a + seg, G);
  int s = 0;
  for (int k = a + lane; k < b; k += 64 * kU) {
    int v[kU];
#pragma unroll
    for (int u = 0; u < kU; ++u) v[u] = k + 64 * u < b ? cnt[k + 64 * u] : 0;
#pragma unroll
    for (int u = 0; u < kU; ++u) s += v[u];
  }
  s = wave_sum(s);
  if (lane == 0) part[w] = s;
  __syncthreads();
  int run = 0, total = 0;
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    if (k < w) run += part[k];
    total += part[k];
  }
  for (int k = a; k < b; k += 64 * kU) {
    int v[kU];
#pragma unroll
    for (int u = 0; u < kU; ++u) v[u] = k + 64 * u + lane < b ? cnt[k + 64 * u + lane] : 0;
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const int inc = wave_incl_sum(v[u], lane);
      if (k + 64 * u + lane < b) off[k + 64 * u + lane] = run + inc - v[u];
      run += __shfl(inc, 63, 64);
    }
  }
  if (t == 0) off[G] = total;
}

// Plans of the tiles of size B (one wave per global sub-column; the wave of a cell column's even
// sub-column plans it, a lane per band or chunk).
template <class Pol>
__global__ void __launch_bounds__(256) k_tile_plan(Dev<Pol> d, int G, const int *col_case, PlanJobs jobs,
                                                   int scanned) {
  const int w = threadIdx.x >> 6, g0 = blockIdx.x * 4, g = g0 + w, lane = threadIdx.x & 63;
  const int *cnt = jobs.j[blockIdx.y].cnt;
  int *off = jobs.j[blockIdx.y].off;
  int tile;
  if (scanned) {  // off holds the scan of the counts (many sub-columns)
    if (g >= G) return;
    tile = off[g];
  } else {
    // the sub-column's first tile: the block sums the counts of the sub-columns before its own
    // four (at most kPlanPrefixMax), each wave adds those of the block's earlier ones
    __shared__ int s_part[4];
    int before = 0;
    for (int k = threadIdx.x; k < g0; k += 256) before += cnt[k];
    before = wave_sum(before);
    if (lane == 0) s_part[w] = before;
    __syncthreads();
    tile = s_part[0] + s_part[1] + s_part[2] + s_part[3];
    for (int q = g0; q < g && q < G; ++q) tile += cnt[q];
    if (g >= G) return;
    if (lane == 0) {
      off[g] = tile;
      if (g == G - 1) off[G] = tile + cnt[g];  // the tile count
    }
  }
  const int B = jobs.j[blockIdx.y].B;
  TilePlan *plans = jobs.j[blockIdx.y].plans;
  const int c = col_case[g];
  const auto &cp = d.cases[c];
  const int sx = g - cp.col_off, cx = sx >> 1, ny2 = 2 * cp.ny;
  if (sx & 1) return;  // a cell column is planned at its even sub-column
  const int *lbc = d.f_lb + cp.lb_off;
  const int *lb0 = lbc + sx * (ny2 + 1), *lb1 = lb0 + (ny2 + 1);
  const int sc0 = max(2 * cx - 2, 0), sc1 = min(2 * cx + 3, 2 * cp.nx - 1), nsc = sc1 - sc0 + 1;
  const int wc0 = max(cx - 1, 0), wc1 = min(cx + 1, cp.nx - 1);
  const int *swp = d.f_swp + cp.swp_off, *wlb = d.w_lb + cp.swp_off;
  // the plan of a tile whose lanes lie in sub-rows [sy0, sy1]
  auto plan = [&](TilePlan &P, int start, int count0, int start1, int count, int sy0, int sy1) {
    const int r0 = max(2 * (sy0 >> 1) - 2, 0), r1 = min(2 * (sy1 >> 1) + 3, ny2 - 1), nr = r1 - r0 + 1;
    int total = 0;
#pragma unroll
    for (int a = 0; a < kTileCols; ++a) {
      int lo = 0, len = 0, row = 0;
      if (a < nsc) {
        row = (sc0 + a) * (ny2 + 1) + r0;
        lo = lbc[row];
        len = lbc[row + nr] - lo;
        row += cp.lb_off;
      }
      P.lo[a] = lo;
      P.len[a] = len;
      P.row[a] = row;
      total += len;
    }
    // wall (bit 0) and solid (bit 1) particles in the 3x3 cells of the tile's lanes
    const int wr0 = max((sy0 >> 1) - 1, 0), wr1 = min((sy1 >> 1) + 1, cp.ny - 1);
    int nw = 0, ns = 0;
    for (int col = wc0; col <= wc1; ++col) {
      const int q = col * (cp.ny + 1);
      nw += wlb[q + wr1 + 1] - wlb[q + wr0];
      if (d.ns > 0) ns += swp[q + wr1 + 1] - swp[q + wr0];
    }
    P.c = c;
    P.start = start;
    P.count = count;
    P.start1 = start1;
    P.count0 = count0;
    P.sx = sx;
    P.cx = cx;
    P.sc0 = sc0;
    P.nsc = nsc;
    P.r0 = r0;
    P.nr = nr;
    P.cy_ref = (sy0 + sy1) >> 2;
    P.sw = (nw > 0 ? 1 : 0) | (ns > 0 ? 2 : 0);
    P.total = total;
    P.gsz = B / nsc;
    P.gmul = ((1 << 20) + P.gsz - 1) / P.gsz;
  };
  const int R = jobs.j[blockIdx.y].band[g];
  if (R > 0) {  // bands of R sub-rows over both sub-columns, a lane per band
    const int nb = (ny2 + R - 1) / R;
    for (int k0 = 0; k0 < nb; k0 += 64) {
      const int k = k0 + lane, a = min(k * R, ny2), b = min(a + R, ny2);
      const int n = k < nb ? band_lanes(lb0, lb1, a, b) : 0, n0 = k < nb ? lb0[b] - lb0[a] : 0;
      const int np = (n + B - 1) / B;
      const int incl = wave_incl_sum(np, lane);
      for (int q = 0; q < np; ++q) {  // lanes [q B, q B + cnt) of the band: even sub-column first
        const int u0 = q * B, cnt = min(B, n - u0), c0 = max(0, min(n0 - u0, cnt));
        const int s1 = lb1[a] + max(0, u0 - n0);
        plan(plans[tile + incl - np + q], u0 < n0 ? lb0[a] + u0 : s1, c0, s1, cnt, a, b - 1);
      }
      tile += __shfl(incl, 63, 64);
    }
  } else {  // each sub-column on its own, a lane per chunk
    for (int h = 0; h < 2; ++h) {
      const int *lb = h ? lb1 : lb0;
      const int s0 = lb[0], n = lb[ny2] - s0, nch = (n + B - 1) / B;
      for (int k0 = 0; k0 < nch; k0 += 64) {
        const int k = k0 + lane, cstart = s0 + k * B, ccount = k < nch ? min(B, n - k * B) : 0;
        const int np = k < nch ? chunk_pieces(d.f_cxy, cstart, ccount, B) : 0;
        const int incl = wave_incl_sum(np, lane);
        for (int q = 0; q < np; ++q) {
          const int start = cstart + q * ccount / np, count = cstart + (q + 1) * ccount / np - start;
          plan(plans[tile + incl - np + q], start, h ? 0 : count, start, count, cxy_y(d.f_cxy[start]),
               cxy_y(d.f_cxy[start + count - 1]));
        }
        tile += __shfl(incl, 63, 64);
      }
    }
  }
}

// Blocks are dealt to the 8 XCDs round-robin, and each XCD has its own L2.  Giving XCD x the
// contiguous range [x per, (x + 1) per) of the n tiles makes the tiles an XCD runs at the
// same time neighbours, so the halo they stage is read from its L2.  The grid is sized for
// the largest possible tile count; a block with no tile gets n.
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_tile(int b, int n) {
  const int per = (n + kXcds - 1) / kXcds, k = b / kXcds;
  return k < per ? (b % kXcds) * per + k : n;
}

// Per-case maximum of non-negative double bit patterns: one atomic per block (a tile lies in
// one case), into slot blockIdx.x % kSpread of the case.  Every thread must call it.
template <int BLOCK>
__device__ __forceinline__ void reduce_case_max(unsigned long long *slots, int c, bool valid, unsigned long long v,
                                                unsigned long long *s_red) {
  const int t = threadIdx.x;
  const unsigned long long w = wave_max_u64(valid ? v : 0ull);
  if ((t & 63) == 0) s_red[t >> 6] = w;
  __syncthreads();
  if (t == 0) {
    unsigned long long m = 0ull;
    for (int k = 0; k < BLOCK / 64; ++k) m = s_red[k] > m ? s_red[k] : m;
    if (m != 0ull) atomicMax(slots + (size_t)(c * kSpread + (int)(blockIdx.x % kSpread)) * kSlotStride, m);
  }
}

// umax2[c] = max over the case's slots, and the capped speed umax[c]; one wave per case.
template <class Pol>
__global__ void __launch_bounds__(64) k_umax_final(Dev<Pol> d) {
  const int c = blockIdx.x;
  if (c >= d.ncase) return;
  const unsigned long long m = wave_max_u64(d.umax_slot[(size_t)(c * kSpread + (int)threadIdx.x) * kSlotStride]);
  if (threadIdx.x == 0) {
    d.umax2[c] = m;
    d.u2_last[c] = __longlong_as_double((long long)m);
    d.umax[c] = fmin((typename Pol::R)sqrt(__longlong_as_double((long long)m)), d.cases[c].umax_cap);
  }
}

// Copies a tile into LDS: the row tables (lb[a][k] = LDS index of the first candidate of row
// r0 + k or later of sub-column sc0 + a, k = 0..nr) and, through ld / st, the candidates:
// span a occupies LDS [base[a], base[a] + len[a]) in slot order.  Each span is served by a
// contiguous group of threads (its rows and its particles, coalesced), so every element's
// span is known without a search; the loads of U particles are issued before any is
// stored.  st(l, v, a) receives the span too.
template <int BLOCK, int U, class Ld, class St>
__device__ __forceinline__ void tile_stage(const TilePlan &P, const int *f_lb, int (*lb)[TileRows<BLOCK>::value + 1],
                                           Ld &&ld, St &&st) {
  const int t = threadIdx.x, n = P.nsc, nr = P.nr, gsz = P.gsz;
  const int a = min((int)(((unsigned)t * (unsigned)P.gmul) >> 20), n - 1), k0 = t - a * gsz;
  const int stride = a == n - 1 ? BLOCK - a * gsz : gsz;
  int base = 0;
  for (int q = 0; q < a; ++q) base += P.len[q];
  const int lo = P.lo[a], len = P.len[a], row = P.row[a];
  for (int k = k0; k <= nr; k += stride) lb[a][k] = f_lb[row + k] - lo + base;
  for (int k = k0; k < len; k += U * stride) {
    decltype(ld(0)) v[U];
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (k + u * stride < len) v[u] = ld(lo + k + u * stride);
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (k + u * stride < len) st(base + k + u * stride, v[u], a);
  }
}

// The walls of a tile's 3x3 cell columns (cx - 1 .. cx + 1) over its cell rows
// [wcr0, wcr0 + nwr): one span per column in the wall order (walls are sorted by cell once,
// and w_lb gives each column's first wall of a cell row or later).
struct WallSpans {
  int wcr0, nwr, n;
  int q[3], lo[3], len[3];
};
template <class Pol, class CP>
__device__ __forceinline__ WallSpans wall_spans(const Dev<Pol> &d, const TilePlan &P, const CP &cp) {
  WallSpans W;
  W.wcr0 = P.r0 >> 1;
  W.nwr = ((P.r0 + P.nr - 1) >> 1) - W.wcr0 + 1;
  W.n = 0;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const int col = P.cx - 1 + a;
    W.q[a] = W.lo[a] = W.len[a] = 0;
    if ((P.sw & 1) && col >= 0 && col < cp.nx) {
      W.q[a] = cp.swp_off + col * (cp.ny + 1) + W.wcr0;
      W.lo[a] = d.w_lb[W.q[a]];
      W.len[a] = d.w_lb[W.q[a] + W.nwr] - W.lo[a];
    }
    W.n += W.len[a];
  }
  return W;
}
// Copies the wall spans to LDS from index base on: the row table (wlb[a][k] = LDS index of the
// first wall of cell row wcr0 + k or later in column a) and, through st(l, g), the walls.
template <int BLOCK, int WROWS, class Pol, class St>
__device__ __forceinline__ void wall_stage(const Dev<Pol> &d, const WallSpans &W, int base, int (*wlb)[WROWS + 1],
                                           St &&st) {
  const int t = threadIdx.x;
  if (t <= W.nwr) {
    int b = base;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      wlb[a][t] = W.len[a] > 0 ? d.w_lb[W.q[a] + t] - W.lo[a] + b : b;
      b += W.len[a];
    }
  }
  const int e1 = W.len[0], e2 = W.len[0] + W.len[1];
  for (int e = t; e < W.n; e += BLOCK)  // (selects, not indexing: the spans stay in registers)
    st(base + e, e < e1 ? W.lo[0] + e : (e < e2 ? W.lo[1] + e - e1 : W.lo[2] + e - e2));
}

// A lane's position relative to its hashed sub-cell, in sub-cell units.
struct LanePos {
  float fx, fy;
};
template <class CP>
__device__ __forceinline__ LanePos lane_pos(double x, double y, int sx, int sy, const CP &cp) {
  const double tx = (x - cp.cxmin) * cp.inv_cdx * 2.0 - (double)sx;
  const double ty = (y - cp.cymin) * cp.inv_cdy * 2.0 - (double)sy;
  return {(float)fmin(fmax(tx, -8.0), 9.0), (float)fmin(fmax(ty, -8.0), 9.0)};
}
// The same from an M32 view (the offset from the origin of the lane's cell (cx, cy)): float
// arithmetic, whose rounding stays far inside the pruning margin.
template <class CP>
__device__ __forceinline__ LanePos lane_pos_view(float xi, float yi, int cx, int cy, int sx, int sy, const CP &cp) {
  const float tx = fmaf(xi, cp.inv_ssx, (float)(2 * cx - sx)), ty = fmaf(yi, cp.inv_ssy, (float)(2 * cy - sy));
  return {fminf(fmaxf(tx, -8.f), 9.f), fminf(fmaxf(ty, -8.f), 9.f)};
}

// Sub-column offset of the q-th window a lane visits: 0, -1, +1, -2, +2, -3, +3 (the 3x3
// cells of a sub-column span offsets -3 .. +3).
constexpr int kWindowOrder = 7;
__device__ __forceinline__ int window_offset(int q) { return (q & 1) ? -((q + 1) >> 1) : (q >> 1); }

// Visits the candidate row windows of a lane of sub-column sx: f(sc, ra, rb) for each
// sub-column sc that meets the lane's disc of radius R, with the sub-rows [ra, rb] that do,
// clipped to the 3x3 cells of the lane's cell (the tile's cell column, row cy) and to the
// tile's candidate range; sub-columns in window_offset order (PAIRED) or from left to right.
template <bool PAIRED, class CP, class F>
__device__ __forceinline__ void for_windows(const TilePlan &P, const CP &cp, int sx, int sy, int cy, LanePos lp,
                                            float R, F &&f) {
  const float R2 = R * R;
  const int a_lo = max((int)floorf(lp.fx - R * cp.inv_ssx), P.sc0 - sx);
  const int a_hi = min((int)floorf(lp.fx + R * cp.inv_ssx), P.sc0 + P.nsc - 1 - sx);
  const int rmin = max(2 * cy - 2, P.r0), rmax = min(2 * cy + 3, P.r0 + P.nr - 1);
  for (int q = PAIRED ? 0 : a_lo; q < (PAIRED ? kWindowOrder : a_hi + 1); ++q) {
    const int a = PAIRED ? window_offset(q) : q;
    if (PAIRED && (a < a_lo || a > a_hi)) continue;
    const float g = fmaxf(fmaxf((float)a - lp.fx, lp.fx - (float)(a + 1)), 0.f) * cp.ssx;
    if (g >= R) continue;
    const float hr = fminf(__builtin_amdgcn_sqrtf(R2 - g * g) * cp.inv_ssy, 8.f);
    const int ra = max(sy + (int)floorf(lp.fy - hr), rmin);
    const int rb = min(sy + (int)floorf(lp.fy + hr), rmax);
    if (ra <= rb) f(sx + a, ra, rb);
  }
}

// The row window [ra, rb] of a lane of sub-column sx in candidate sub-column sc0 + sa
// (for_windows for one sub-column); false when the sub-column holds no candidate of the lane.
template <class CP>
__device__ __forceinline__ bool span_window(const TilePlan &P, const CP &cp, int sx, int sa, int sy, int cy,
                                            LanePos lp, float R, int &ra, int &rb) {
  if (sa < 0 || sa >= P.nsc) return false;
  const int a = P.sc0 + sa - sx;
  const float R2 = R * R;
  if (a < (int)floorf(lp.fx - R * cp.inv_ssx) || a > (int)floorf(lp.fx + R * cp.inv_ssx)) return false;
  const float g = fmaxf(fmaxf((float)a - lp.fx, lp.fx - (float)(a + 1)), 0.f) * cp.ssx;
  if (g >= R) return false;
  const float hr = fminf(__builtin_amdgcn_sqrtf(R2 - g * g) * cp.inv_ssy, 8.f);
  ra = max(sy + (int)floorf(lp.fy - hr), max(2 * cy - 2, P.r0));
  rb = min(sy + (int)floorf(lp.fy + hr), min(2 * cy + 3, P.r0 + P.nr - 1));
  return ra <= rb;
}

// Shift windows over the passes of a step.  A step's shift passes move particles by a tiny
// fraction of rc, and every pass of a step uses the same tiles, so the same staged LDS
// order.  The first pass (kShiftRecord, when a step has more than one) widens its pruning
// radius by s rc and records, per lane and candidate sub-column, the run of LDS indices from
// the first to the last candidate within rc (1 + s).  Its path lengths restart: the lanes
// store the path since that pass (and keep the one since the hash with the runs).  A later
// pass (kShiftCached) walks only those runs while the lane's path since the first pass plus
// the largest one among the tile's staged candidates stays within s rc: every candidate
// within rc then was within rc (1 + s) at the first pass, so it lies in a run, and the runs
// keep the sub-column and slot order, so the non-zero terms and their order are those of the
// full windows and the sums are the same bit for bit.  Any other lane searches the full
// windows, with the tile's path bound rebuilt from the first pass's bound.  The last pass
// writes the path since the hash again.  kShiftPlain is the single pass of a one-pass step.
// When the pair kernel runs on the shift's tiles it records the runs itself, at the hash's
// positions (with the lane's path since the hash and the tile's staged path bound, as the
// first pass would), and every pass of the step is a kShiftCached one: the paths then count
// from the hash, which bounds the paths since the recording.
constexpr int kShiftPlain = 0, kShiftRecord = 1, kShiftCached = 2;
constexpr double kShiftSkin = 0.02;

// A lane's recorded candidate runs, in the packing of f_wc: up to 5 runs of (11-bit start,
// 7-bit length) in three words; ok turns false at a sixth run, a run past 127 slots or a start
// past 2047 (the lane then searches its full windows).
struct RunRec {
  unsigned w0 = 0u, w1 = 0u, w2 = 0u;
  int n = 0;
  bool ok = true;
  __device__ __forceinline__ void put(int f, int len) {
    if (n >= 5 || len > 127 || f > 2047) {
      ok = false;
      return;
    }
    const unsigned st = (unsigned)f, ln = (unsigned)len;
    if (n == 0) w0 |= st | ln << 22;
    if (n == 1) {
      w0 |= st << 11;
      w1 |= ln << 22;
    }
    if (n == 2) w1 |= st, w2 |= ln << 11;
    if (n == 3) w1 |= st << 11, w2 |= ln << 18;
    if (n == 4) w2 |= st | ln << 25;
    ++n;
  }
  // the f_wc entry; path: the path since the hash at the recording
  __device__ __forceinline__ uint4 entry(float path) const {
    return uint4{w0 | (ok ? 0x80000000u : 0u), w1, w2, __float_as_uint(path)};
  }
};

// Pruning radius: rc (1 + 1e-5) plus the tile's displacement bound D; a tile that is not
// staged searches its lanes' whole 3x3 cells.
template <class CP>
__device__ __forceinline__ float prune_radius(const CP &cp, bool staged, float D) {
  if (staged) return (float)cp.rcut_prune + D;
  return (float)(16.0 * fmax(cp.cdx, cp.cdy));
}

// Largest path length among the staged candidates of the block (after the staging barrier).
template <int BLOCK>
__device__ __forceinline__ float staged_bound(const float *s_dmax) {
  float D = 0.f;
  for (int k = 0; k < BLOCK / 64; ++k) D = fmaxf(D, s_dmax[k]);
  return D * (1.0f + 1e-6f);
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

// Calls f(s, true) for s in [lo, hi) U at a time, then covers the remainder with one more
// group of U calls f(s, s < hi): the caller makes a call with false add exact zeros, and
// may read up to U - 1 entries past hi.
template <int U, class F>
__device__ __forceinline__ void for_range_masked(int lo, int hi, F &&f) {
  int s = lo;
  for (const int last = hi - (U - 1); s < last; s += U) {
#pragma unroll
    for (int k = 0; k < U; ++k) f(s + k, true);
  }
  if (s < hi) {
#pragma unroll
    for (int k = 0; k < U; ++k) f(s + k, s + k < hi);
  }
}

// Calls f(s, true) for s in [lo, hi): four at a time, then the lane's last 0 to 3 by a pair and
// a single under branches, so no call is masked (a wave runs the pair if any lane has two left,
// the single if any has one).
template <class F>
__device__ __forceinline__ void for_range_4_2_1(int lo, int hi, F &&f) {
  int s = lo;
#pragma unroll 1
  for (const int last = hi - 3; s < last; s += 4) {
    f(s, true);
    f(s + 1, true);
    f(s + 2, true);
    f(s + 3, true);
  }
  if (s + 1 < hi) {
    f(s, true);
    f(s + 1, true);
    s += 2;
  }
  if (s < hi) f(s, true);
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
  // w'(r^2) of a candidate known to lie inside the cut-off
  __device__ __forceinline__ R w_in(R r2) const {
    if constexpr (sizeof(R) == 4)
      return __builtin_amdgcn_exp2f(r2 * nhl2) * kdc;
    else
      return exp(-(r2 * hinv2)) * kdc;
  }
};

// ======================================================================================
// FLUID: Riemann pair loop fused with the explicit update
// ======================================================================================

template <class R>
struct PairSums {
  R ax = 0, ay = 0, drho = 0, L00 = 0, L01 = 0, L10 = 0, L11 = 0, col = 0;
  R fsr = R(1e30);  // smallest r^2 of a FREE_SURFACE neighbour inside the cut-off (liquid lanes)
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
  // the density weights rho_i / (rho_i + rho_j) and rho_j / (rho_i + rho_j); the factors that
  // depend on particle i alone (-2 / rho_i, 2 rho_i, p_i / c) are formed once per lane
  const R isum = krcp<R>(rhoi + rhoj);
  const R ai = rhoi * isum, aj = rhoj * isum;
  R ps = ai * pj + aj * pi;
  const R du = q * inv_r;
  if (du >= R(0)) ps += (rhoi * aj) * du * fmin(R(3) * du, cs);
  const R t = (pi * csinv - pj * csinv) * isum;
  const R s = ps * ((R(-2) * rhoi_inv) * w);
  A.ax += s * dx;
  A.ay += s * dy;
  A.drho += ((R(2) * rhoi) * w) * (aj * q - t * rn);
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

// LDS capacity (particles) of a tile of BLOCK slots: the planner's budget.  The M32 pair kernel
// (25 B of LDS per candidate) stages five eighths of it: a dense 256-lane tile over both
// sub-columns of a cell column stages about 870 candidates, and at 37 KB per block with its
// registers held to 104 (no spills) four blocks share a CU (16 waves instead of 12: 6%
// faster).  FP64 candidates take 49 B, so P64 keeps the full budget.
template <class Pol, int BLOCK>
struct PairCap {
  static constexpr int value = Pol::MIXED ? plan_cap(BLOCK) * 5 / 8 : plan_cap(BLOCK);
};
// The shift stages three quarters of the pair's budget: tiles over both sub-columns of a cell
// column stage about 3.4 candidates a lane (a dense 256-lane tile ~870), and the smaller LDS
// lets 8 blocks of 256 share a CU.
template <int BLOCK>
struct ShiftCap {
  static constexpr int value = plan_cap(BLOCK) * 3 / 4;
};

// findFluidPair (src/fluid.cu:193-254) + fluidUpdate (src/fluid.cu:256-287), one tile per
// block.  Reads the current buffers (fa) of every neighbour and writes particle i's updated
// state into the alternate buffers (fb), so neighbours never see a half-updated state.  Also
// reduces the per-case max |u|^2 of the updated velocities (getMaxSpeed,
// src/fluid.cu:429-440) and advances every particle's path length since the hash.  With
// record (the shift runs on the same tiles), it also records the shift passes' candidate runs
// (f_wc, tile_d0) from its windows widened by the skin.
template <class Pol, bool VISCID, int BLOCK>
__global__ void __launch_bounds__(BLOCK, (Pol::MIXED ? 4 : 1))
    k_fluid_pair_update(Dev<Pol> d, const TilePlan *plans, const int *ntiles, int record) {
  using R = typename Pol::R;
  using R2 = typename V2<R>::T;
  constexpr int CAP = PairCap<Pol, BLOCK>::value;
  // walls of the tile's 3x3 cell columns over its cell rows follow the fluid candidates
  // (a tile's fluid candidates rarely fill CAP: the walls use what they leave, so that the
  // block stays within a fifth of the CU's LDS at 128 slots)
  constexpr int WCAP = 0, WROWS = TileRows<BLOCK>::value / 2 + 2;
  // fluid candidates, padding, walls, padding (the padding is far away, finite, without flags:
  // read by masked tails)
  __shared__ R2 s_pos[CAP + WCAP + 8], s_vel[CAP + WCAP + 8], s_rp[CAP + WCAP + 8];
  __shared__ uint8_t s_fl[CAP + WCAP + 8];  // phase | class << 4
  __shared__ int s_lb[kTileCols][TileRows<BLOCK>::value + 1];
  __shared__ int s_wlb[3][WROWS + 1];  // LDS index of the first wall of cell row wcr0 + k or later
  __shared__ unsigned long long s_red[BLOCK / 64];
  __shared__ float s_dmax[BLOCK / 64];
  const int nt = *ntiles, tile = xcd_tile(blockIdx.x, nt);
  if (tile >= nt) return;
  const TilePlan &P = plans[tile];
  const int t = threadIdx.x, c = P.c;
  const bool valid = t < P.count;
  const int i = valid ? lane_slot(P, t) : 0, lsx = lane_subcol(P, t);  // the lane's slot, sub-column
  const auto &A = d.fa;
  // the lane's own state, loaded ahead of the staging so that the loads overlap it
  int pk = 0, phi = 0, cls0 = 0;
  R xi = R(0), yi = R(0), ui = R(0), vi = R(0), rhoi = R(1), pi = R(0);
  double x0 = 0.0, y0 = 0.0, u0 = 0.0, v0 = 0.0, r0 = 0.0;
  float cum0 = 0.f;
  if (valid) {
    pk = d.f_cxy[i];
    xi = A.vx[i];
    yi = A.vy[i];
    ui = A.vu[i];
    vi = A.vv[i];
    rhoi = A.vrho[i];
    pi = A.p[i];
    x0 = A.x[i];
    y0 = A.y[i];
    u0 = A.u[i];
    v0 = A.v[i];
    r0 = A.rho[i];
    phi = d.f_phase[i];
    cls0 = d.f_cls[i];
    cum0 = A.cum[i];
  }
  const bool staged = P.total <= CAP && P.nr <= TileRows<BLOCK>::value;
  int dmax = 0;  // bit pattern of the largest path length (non-negative floats order as ints)
  if (staged) {
    struct Ld {
      R2 pos, vel, rp;
      int fl;
      int cum;
      int pk;
    };
    const auto &cq = d.cases[c];
    const R fdx = cq.cdx_r, fdy = cq.cdy_r;
    const int gcx = P.cx, gcy = P.cy_ref;
    tile_stage<BLOCK, 4>(
        P, d.f_lb, s_lb,
        [&](int g) {
          return Ld{R2{ld32(A.vx, g), ld32(A.vy, g)}, R2{ld32(A.vu, g), ld32(A.vv, g)},
                    R2{ld32(A.vrho, g), ld32(A.p, g)}, (int)ld32(d.f_phase, g) | ((int)ld32(d.f_cls, g) << 4),
                    ld32(reinterpret_cast<const int *>(A.cum), g), Pol::MIXED ? ld32(d.f_cxy, g) : 0};
        },
        [&](int l, const Ld &v, int) {
          if constexpr (Pol::MIXED)  // tile frame
            s_pos[l] = R2{v.pos.x + (float)((cxy_x(v.pk) >> 1) - gcx) * fdx,
                          v.pos.y + (float)((cxy_y(v.pk) >> 1) - gcy) * fdy};
          else
            s_pos[l] = v.pos;
          s_vel[l] = v.vel;
          s_rp[l] = v.rp;
          s_fl[l] = (uint8_t)v.fl;
          dmax = max(dmax, v.cum);
        });
    if (t < 4) {
      s_pos[P.total + t] = R2{R(1e18), R(1e18)};
      s_vel[P.total + t] = R2{R(0), R(0)};
      s_rp[P.total + t] = R2{R(1), R(0)};
      s_fl[P.total + t] = 0;
    }
  }
  // walls (fluidVirtCalculate: rho_j = rho0[wall phase]; a wall is never FREE_SURFACE)
  const auto &cw = d.cases[c];
  const WallSpans W = wall_spans(d, P, cw);
  const int wbase = P.total + 4;
  const bool wstaged = staged && W.n > 0 && P.total + W.n <= CAP + WCAP && W.nwr <= WROWS;
  if (wstaged) {
    const R fdx = cw.cdx_r, fdy = cw.cdy_r;
    wall_stage<BLOCK, WROWS>(d, W, wbase, s_wlb, [&](int l, int g) {
      const R wx = ld32(d.w_vx, g), wy = ld32(d.w_vy, g);
      const int phj = ld32(d.w_phase, g);
      if constexpr (Pol::MIXED) {  // tile frame
        const int wk = ld32(d.w_cxy, g);
        s_pos[l] = R2{wx + (float)(cxy_x(wk) - P.cx) * fdx, wy + (float)(cxy_y(wk) - P.cy_ref) * fdy};
      } else {
        s_pos[l] = R2{wx, wy};
      }
      s_vel[l] = R2{ld32(d.w_u, g), ld32(d.w_v, g)};
      s_rp[l] = R2{cw.rho0[phj], ld32(d.w_p, g)};
      s_fl[l] = (uint8_t)phj;
    });
    if (t < 4) {
      s_pos[wbase + W.n + t] = R2{R(1e18), R(1e18)};
      s_vel[wbase + W.n + t] = R2{R(0), R(0)};
      s_rp[wbase + W.n + t] = R2{R(1), R(0)};
      s_fl[wbase + W.n + t] = 0;
    }
  }
  for (int o = 32; o > 0; o >>= 1) dmax = max(dmax, __shfl_xor(dmax, o, 64));
  if ((t & 63) == 0) s_dmax[t >> 6] = __int_as_float(dmax);
  __syncthreads();
  unsigned long long my_u2 = 0ull;
  if (valid) {
    const auto &cp = d.cases[c];
    const int sy = cxy_y(pk), cx = P.cx, cy = sy >> 1;
    LanePos lp;
    if constexpr (Pol::MIXED)
      lp = lane_pos_view(xi, yi, cx, cy, lsx, sy, cp);
    else
      lp = lane_pos(x0, y0, lsx, sy, cp);
    const bool liq = phi == 0;
    const R cs = cp.c[phi], csinv = cp.cinv[phi], mu8 = cp.mu8[phi];
    const R rhoi_inv = krcp<R>(rhoi);
    const LoopConsts<R> K(cp);
    const R eps_h2 = cp.eps_h2, qrc2 = cp.quarter_rc2;
    // LDS index of the lane's own slot (in the span of its sub-column)
    int own = i;
    if (staged) {
      int b = 0;
      for (int a = 0; a < lsx - P.sc0; ++a) b += P.len[a];
      own = b + (i - P.lo[lsx - P.sc0]);
    }
    PairSums<R> S;
    // fluidFluidCalculate on one candidate (w = 0 outside the cut-off: exact zero terms);
    // INSIDE: the candidate is known to lie inside the cut-off.  The reference's sticky near
    // flag (src/fluid.cu:67-73) ends at HAS_NEAR (2) if a FREE_SURFACE neighbour lies within
    // half the cut-off, else HAS (1) if one lies within it: the smallest such r^2 decides.
    auto fluid_j = [&](auto inside, bool self, R dx, R dy, R r2, R uj, R vj, R rhoj, R pj, int fl, bool in) {
      constexpr bool INSIDE = decltype(inside)::value;
      R w;
      bool fs = liq & !self & ((fl >> 4) == 1);
      if constexpr (INSIDE) {
        w = K.w_in(r2);
      } else {
        const R wk = K.w(r2);
        w = in ? wk : R(0);
        fs = fs & in & (r2 < K.rc2);
      }
      riemann_pair<R, VISCID>(dx, dy, r2, w, ui, vi, uj, vj, rhoi, rhoi_inv, rhoj, pi, pj, cs, csinv, mu8, eps_h2, S);
      if (!(liq && (fl & 0xf) == 1)) S.col += r2 * w;
      S.fsr = (fs && r2 < S.fsr) ? r2 : S.fsr;
    };
    using Cut = std::false_type;
    using Inside = std::true_type;
    const int *lbc = d.f_lb + cp.lb_off;
    const int ny21 = 2 * cp.ny + 1;
    // rows [qa, qb] of sub-column sc from global memory, whose views share the cell origin
    // behind (bx, by)
    auto rows = [&](int sc, int qa, int qb, R bx, R by) {  // global memory
      for_range<2>(lbc[sc * ny21 + qa], lbc[sc * ny21 + qb + 1], [&](int s) {
        const R dx = A.vx[s] - bx, dy = A.vy[s] - by;
        fluid_j(Cut{}, s == i, dx, dy, dx * dx + dy * dy, A.vu[s], A.vv[s], A.vrho[s], A.p[s],
                (int)d.f_phase[s] | ((int)d.f_cls[s] << 4), true);
      });
    };
    const float Dst = staged_bound<BLOCK>(s_dmax), Rp = prune_radius(cp, staged, Dst);
    // a recording for the shift passes (record): the runs of every lane of a staged tile, the
    // tile's staged path bound with them; a lane of a tile that is not staged has none
    if (record) {
      if (t == 0) d.tile_d0[tile] = Dst;
      if (!staged) d.f_wc[i] = uint4{0u, 0u, 0u, 0u};
    }
    // the lane in the tile frame (M32 staged positions are re-based on it)
    const R yg = Pol::MIXED ? yi + (R)(cy - P.cy_ref) * cp.cdy_r : yi;
    auto staged_j = [&](auto inside, int l, bool in) {
      const R2 pj = s_pos[l], vj = s_vel[l], rp = s_rp[l];
      const R dx = pj.x - xi, dy = pj.y - yg;
      fluid_j(inside, l == own, dx, dy, dx * dx + dy * dy, vj.x, vj.y, rp.x, rp.y, s_fl[l], in);
    };
    if (staged) {
      // staged windows (one run each): first mark the candidates inside the cut-off (32-bit
      // mask per window), then evaluate the Riemann terms for those only, windows in
      // for_windows order and slots in order -- the same non-zero terms in the same order as
      // a pass over every candidate.  The windows at offsets -k and +k share one loop over a
      // 64-bit mask: a lane near one side of its sub-column has more terms on that side and
      // fewer on the other, so a wave runs the pair for its lane with the most terms in both
      // together instead of each window for its own longest lane.
      unsigned mk[kWindowOrder];
      int w0[kWindowOrder];
      // in-range mask of the run [s0, s1) of at most 32 candidates; with REC also the mask ms
      // of the candidates within rc (1 + s) (the recorded runs).  A longer run is evaluated
      // whole (direct), in its place in the window order.
      RunRec rr;
      const R rl2 = cp.rl2_pair;
      auto direct = [&](int s0, int s1) {
        for_range_masked<2>(s0, s1, [&](int l, bool in) { staged_j(Cut{}, l, in); });
      };
      auto mark = [&](auto rec, int s0, int s1, unsigned &ms) -> unsigned {
        constexpr bool REC = decltype(rec)::value;
        ms = 0u;
        // 4 candidates a round (then 2, then 1), their bits gathered in a nibble placed at
        // once, so that the bit positions cost no instruction per candidate
        auto r2of = [&](const R2 &pj) {
          const R dx = pj.x - xi, dy = pj.y - yg;
          return dx * dx + dy * dy;
        };
        const R rc2 = K.rc2;
        unsigned m = 0u;
        int s = s0, sh = 0;
#pragma unroll 1
        for (const int last = s1 - 3; s < last; s += 4, sh += 4) {
          const R2 p0 = s_pos[s], p1 = s_pos[s + 1], p2 = s_pos[s + 2], p3 = s_pos[s + 3];
          const R q0 = r2of(p0), q1 = r2of(p1), q2 = r2of(p2), q3 = r2of(p3);
          m |= ((q0 < rc2 ? 1u : 0u) | (q1 < rc2 ? 2u : 0u) | (q2 < rc2 ? 4u : 0u) | (q3 < rc2 ? 8u : 0u)) << sh;
          if (REC) ms |= ((q0 < rl2 ? 1u : 0u) | (q1 < rl2 ? 2u : 0u) | (q2 < rl2 ? 4u : 0u) | (q3 < rl2 ? 8u : 0u)) << sh;
        }
        if (s + 1 < s1) {
          const R q0 = r2of(s_pos[s]), q1 = r2of(s_pos[s + 1]);
          m |= ((q0 < rc2 ? 1u : 0u) | (q1 < rc2 ? 2u : 0u)) << sh;
          if (REC) ms |= ((q0 < rl2 ? 1u : 0u) | (q1 < rl2 ? 2u : 0u)) << sh;
          s += 2;
          sh += 2;
        }
        if (s < s1) {
          const R q0 = r2of(s_pos[s]);
          m |= (q0 < rc2 ? 1u : 0u) << sh;
          if (REC) ms |= (q0 < rl2 ? 1u : 0u) << sh;
        }
        return m;
      };
      // the marked candidates of window (ma, sa0) and then of window (mb, sb0), in slot order;
      // the LDS reads of the next candidate are issued before the current one is summed, so
      // they overlap its Riemann term instead of stalling the lane
      struct Cand {
        R2 p, v, r;
        int fl, l;
      };
      auto eval2 = [&](unsigned ma, int sa0, unsigned mb, int sb0) {
        unsigned long long m = (unsigned long long)ma | ((unsigned long long)mb << 32);
        if (m == 0ull) return;
        const int sb = sb0 - 32;
        auto next = [&]() {
          const int b = __builtin_ctzll(m);
          m &= m - 1ull;
          const int l = (b < 32 ? sa0 : sb) + b;
          return Cand{s_pos[l], s_vel[l], s_rp[l], (int)s_fl[l], l};
        };
        auto use = [&](const Cand &q) {
          const R dx = q.p.x - xi, dy = q.p.y - yg;
          fluid_j(Inside{}, q.l == own, dx, dy, dx * dx + dy * dy, q.v.x, q.v.y, q.r.x, q.r.y, q.fl, true);
        };
        Cand cur = next();
        // unrolled by two, so that every other round reads into the registers the previous one
        // used instead of copying nxt into cur
#pragma unroll 2
        while (m) {
          const Cand nxt = next();
          use(cur);
          cur = nxt;
        }
        use(cur);
      };
      auto eval = [&](unsigned m, int s0) { eval2(m, s0, 0u, 0); };
      const int off = lsx - P.sc0;  // tile sub-column of the lane's own sub-column
      // the windows are marked left to right (offsets -3 .. +3; window q has offset
      // window_offset(q)), the order in which a recording keeps its runs for the shift passes;
      // a recording widens them by the skin
      const float Rpw = record ? Rp + cp.skin_grow : Rp;
      unsigned dmask = 0u;  // windows evaluated whole
#pragma unroll
      for (int k = 0; k < kWindowOrder; ++k) {
        const int o = k - 3, q = o < 0 ? -2 * o - 1 : 2 * o;
        mk[q] = 0u;
        w0[q] = 0;
        const int sa = off + o;
        int ra, rb;
        if (!span_window(P, cp, lsx, sa, sy, cy, lp, Rpw, ra, rb)) continue;
        const int s0 = s_lb[sa][ra - P.r0], s1 = s_lb[sa][rb + 1 - P.r0];
        w0[q] = s0;
        if (s1 - s0 > 32) {  // evaluated whole below; its end in place of the mask
          mk[q] = (unsigned)s1;
          dmask |= 1u << q;
          rr.ok = false;
          continue;
        }
        unsigned ms;
        if (record) {
          mk[q] = mark(std::true_type{}, s0, s1, ms);
          if (ms != 0u) {
            const int f = __builtin_ctz(ms);
            rr.put(s0 + f, 32 - __builtin_clz(ms) - f);
          }
        } else {
          mk[q] = mark(std::false_type{}, s0, s1, ms);
        }
      }
      if (record) d.f_wc[i] = rr.entry(cum0);
      // windows in the order 0, -1, +1, ... (those at -k and +k in one loop unless one of them
      // is evaluated whole)
      auto evalq = [&](int q) {
        if ((dmask >> q) & 1u)
          direct(w0[q], (int)mk[q]);
        else
          eval(mk[q], w0[q]);
      };
      evalq(0);
#pragma unroll
      for (int q = 1; q < kWindowOrder; q += 2) {
        if (((dmask >> q) & 3u) != 0u) {
          evalq(q);
          evalq(q + 1);
        } else {
          eval2(mk[q], w0[q], mk[q + 1], w0[q + 1]);
        }
      }
      if (wstaged) {  // walls of the 3x3 cells, cell column by cell column
        const int lr0 = max(cy - 1, 0) - W.wcr0, lr1 = min(cy + 1, cp.ny - 1) - W.wcr0;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          const int s0 = s_wlb[a][lr0], s1 = s_wlb[a][lr1 + 1];
          if (s1 - s0 > 32) {
            direct(s0, s1);
          } else {
            unsigned ms;
            eval(mark(std::false_type{}, s0, s1, ms), s0);
          }
        }
      }
    }
    for_windows<true>(P, cp, lsx, sy, cy, lp, Rp, [&](int sc, int ra, int rb) {
      if (staged) return;
      const R bx = xi - cell_shift_x<Pol>((sc >> 1) - cx, cp);
      if constexpr (Pol::MIXED) {
        // M32 views are relative to their own cell: one run per cell row
        for (int cr = ra >> 1; cr <= (rb >> 1); ++cr)
          rows(sc, max(ra, 2 * cr), min(rb, 2 * cr + 1), bx, yi - cell_shift_y<Pol>(cr - cy, cp));
      } else {
        rows(sc, ra, rb, bx, yi);
      }
    });
    for (int a = -1; a <= 1 && (P.sw & 1) && !wstaged; ++a) {
      const int jx = cx + a;
      if (jx < 0 || jx >= cp.nx) continue;
      const R bx = xi - cell_shift_x<Pol>(a, cp);
      for (int b = -1; b <= 1; ++b) {
        const int jy = cy + b;
        if (jy < 0 || jy >= cp.ny) continue;
        const R by = yi - cell_shift_y<Pol>(b, cp);
        const int2 rw = d.w_range[cp.cell_off + jx * cp.ny + jy];
        for_range<4>(rw.x, rw.y, [&](int s) {  // fluidVirtCalculate: rho_j = rho0[wall phase]
          const R dx = ld32(d.w_vx, s) - bx, dy = ld32(d.w_vy, s) - by;
          const R r2 = dx * dx + dy * dy;
          if (!(r2 < K.rc2)) return;
          const R w = K.w(r2);
          const int phj = d.w_phase[s];
          riemann_pair<R, VISCID>(dx, dy, r2, w, ui, vi, d.w_u[s], d.w_v[s], rhoi, rhoi_inv, cp.rho0[phj], pi,
                                  d.w_p[s], cs, csinv, mu8, eps_h2, S);
          if (!(liq && phj == 1)) S.col += r2 * w;
        });
      }
    }
    for (int a = -1; a <= 1 && (P.sw & 2); ++a) {
      const int jx = cx + a;
      if (jx < 0 || jx >= cp.nx) continue;
      const R bx = xi - cell_shift_x<Pol>(a, cp);
      for (int b = -1; b <= 1; ++b) {
        const int jy = cy + b;
        if (jy < 0 || jy >= cp.ny) continue;
        const R by = yi - cell_shift_y<Pol>(b, cp);
        const int2 rs = d.s_range[cp.cell_off + jx * cp.ny + jy];
        for_range<2>(rs.x, rs.y, [&](int k) {  // fluidSolidCalculate: rho_j = rho_i
          const int j = d.s_sorted[k];
          const R dx = d.s_vx[j] - bx, dy = d.s_vy[j] - by;
          const R r2 = dx * dx + dy * dy;
          if (!(r2 < K.rc2)) return;
          const R w = K.w(r2);
          riemann_pair<R, VISCID>(dx, dy, r2, w, ui, vi, (R)d.su[j], (R)d.sv[j], rhoi, rhoi_inv, rhoi, pi, d.s_p[j],
                                  cs, csinv, mu8, eps_h2, S);
        });
      }
    }
    d.f_col[i] = S.col;
    const int near = S.fsr < qrc2 ? 2 : (S.fsr < K.rc2 ? 1 : 0);
    d.f_near[i] = (int8_t)near;
    // the fluid class of this step (correctionUpdate, src/fluid.cu:369-427), a function of
    // phase, colour and near flag only; it is set only where the case shifts
    int cls = cls0;
    if (cp.K > 0) {
      if (phi == 1)
        cls = 3;  // gas: INNER
      else if (S.col < R(1.5))
        cls = near == 0 ? 0 : 1;  // SPLASH / FREE_SURFACE
      else
        cls = (S.col < R(2.0) && near == 2) ? 2 : 3;  // VICINITY / INNER
    }
    d.f_cls2[i] = (int8_t)cls;
    d.f_L[4 * i + 0] = S.L00;
    d.f_L[4 * i + 1] = S.L01;
    d.f_L[4 * i + 2] = S.L10;
    d.f_L[4 * i + 3] = S.L11;
    // fluidUpdate, integrated in double
    double ax = (double)S.ax, ay = (double)S.ay;
    if (cp.steady) {
      ax += -0.1 * u0;
      ay += -0.1 * v0;
    }
    const double dt = d.dtv[c];
    double x = x0 + u0 * dt, y = y0 + v0 * dt;
    const double u = u0 + (ax + cp.gx_d) * dt, v = v0 + (ay + cp.gy_d) * dt;
    double rho = r0 + (double)S.drho * dt;
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
    const double mx = x - x0, my = y - y0;
    B.cum[i] = (float)((double)cum0 + sqrt(mx * mx + my * my));
    if (isnan(rho) || isnan(x) || isnan(y) || isnan(u2)) atomicOr(d.status, kStNaN);
  }
  // per-case max |u|^2 (non-negative doubles order like their bit patterns)
  reduce_case_max<BLOCK>(d.umax_slot, c, valid, my_u2, s_red);
}

// ======================================================================================
// SHIFT: gradient pass fused with the shift update (ping-pong positions)
// ======================================================================================

// correctionCalculate (src/fluid.cu:316-367) + correctionUpdate / shiftDisplacementCalculate
// (src/fluid.cu:369-427), one tile per block.  The gradient sums over all species with no
// phase filter; the class comes from the pair kernel.  A pass moves the positions the
// neighbour loops read (the views, and for P64 the positions themselves) and accumulates the
// step's shift per particle; the case's last pass applies the accumulated shift to the
// velocities (u += L s is linear in s) and to the M32 masters.  For M32 the staged fluid
// positions are re-based on the tile frame (the tile's cell column, and the cell row
// cy_ref), in units scaled so that the weight is 2^-r'^2; the frame costs at most a few
// float ulps of a tile height in the gradient of a small correction.
// A lane's own state for a shift pass (loaded ahead of the staging so that the loads overlap
// it): packed sub-cell, class, position, path length and the shift accumulated so far.  The
// position is the view the neighbour loops read (for P64 the position itself); M32 reads its
// double masters, which stay at the pair kernel's result until then, on the case's last pass
// only.  A pass after a recording one also reads the lane's runs (the path since the hash on
// the last pass only), and the sub-cell only where it needs it before the loops: walls or
// solids in the tile, or the last pass (pk = -1 marks it unread).
template <class Pol>
struct ShiftOwn {
  using R = typename Pol::R;
  using R2 = typename V2<R>::T;
  int pk = 0, cls = 3;
  double x0 = 0.0, y0 = 0.0;
  R vx = R(0), vy = R(0);
  R2 sacc{R(0), R(0)};
  float cum0 = 0.f, cum_hash = 0.f;
  uint3 wc{0u, 0u, 0u};
};
template <class Pol, int MODE>
__device__ __forceinline__ ShiftOwn<Pol> shift_own(const Dev<Pol> &d, const TilePlan &P, int pass, int K) {
  ShiftOwn<Pol> o;
  const int t = threadIdx.x;
  if (t < P.count) {
    const int i = lane_slot(P, t);
    const bool lastp = pass == K - 1;
    o.pk = (MODE != kShiftCached || lastp || P.sw) ? d.f_cxy[i] : -1;
    if (!Pol::MIXED || lastp) {
      o.x0 = d.fa.x[i];
      o.y0 = d.fa.y[i];
    }
    if constexpr (Pol::MIXED) {
      o.vx = d.fa.vx[i];
      o.vy = d.fa.vy[i];
    }
    o.cum0 = d.fa.cum[i];
    if (pass < K) {
      o.cls = d.f_cls[i];
      if (pass > 0) o.sacc = {d.f_sacc[2 * i], d.f_sacc[2 * i + 1]};
      if constexpr (MODE == kShiftCached) {
        o.wc = *reinterpret_cast<const uint3 *>(d.f_wc + i);
        if (lastp) o.cum_hash = __uint_as_float(d.f_wc[i].w);
      }
    }
  }
  return o;
}

// The tile's lanes after the staging barrier: gradient sums over the staged candidates (or
// global memory for a tile that is not staged), walls and solids, then the shift update.
// Reads only the LDS it is given; stores only the lanes' own results.  D bounds the staged
// candidates' paths since the hash; a cached pass also passes Dsp, their paths since the
// step's first pass, and runs whether a staged tile may walk the recorded runs.
template <class Pol, int BLOCK, int MODE, class LB, int WROWS>
__device__ __forceinline__ void shift_tile(const Dev<Pol> &d, const TilePlan &P, int pass, const ShiftOwn<Pol> &own,
                                           const typename V2<typename Pol::R>::T *s_pos,
                                           const LB (*s_lb)[TileRows<BLOCK>::value + 1], const int (*s_wlb)[WROWS + 1],
                                           float D, float Dsp, bool runs, bool staged, bool wstaged,
                                           const WallSpans &W) {
  using R = typename Pol::R;
  using R2 = typename V2<R>::T;
  const int t = threadIdx.x, c = P.c;
  if (t >= P.count) return;
  const int i = lane_slot(P, t), lsx = lane_subcol(P, t);  // the lane's slot and sub-column
  const auto &A = d.fa;
  const auto &B = d.fb;
  const auto &cp = d.cases[c];
  const bool active = pass < cp.K, last = pass == cp.K - 1;
  const int cls = own.cls;
  const double x0 = own.x0, y0 = own.y0;
  const R2 sacc = own.sacc;
  const float cum0 = own.cum0;
  if (!active) {  // this case has fewer passes than the batch: carry the moving fields over
    B.vx[i] = A.vx[i];
    B.vy[i] = A.vy[i];
    B.cum[i] = cum0;
    return;
  }
  bool cached = false;
  if constexpr (MODE == kShiftCached) cached = runs && (own.wc.x >> 31) && cum0 + Dsp <= cp.skin_ok;
  const int pk = own.pk >= 0 ? own.pk : (cached ? 0 : d.f_cxy[i]);
  const int sy = cxy_y(pk), cx = P.cx, cy = sy >> 1;
  // current position: the stored view (M32: cell-relative float, advanced by each pass)
  const R xi = Pol::MIXED ? own.vx : (R)x0, yi = Pol::MIXED ? own.vy : (R)y0;
  LanePos lp;
  if constexpr (Pol::MIXED)
    lp = lane_pos_view(xi, yi, cx, cy, lsx, sy, cp);
  else
    lp = lane_pos(x0, y0, lsx, sy, cp);
  const LoopConsts<R> K(cp);
  R gx = 0, gy = 0;
  auto add = [&](R dx, R dy) {
    const R w = K.w(dx * dx + dy * dy);
    gx += dx * w;
    gy += dy * w;
  };
  const float Rp = prune_radius(cp, staged, D) + ((MODE == kShiftRecord && staged) ? cp.skin_grow : 0.f);
  const int *lbc = d.f_lb + cp.lb_off;
  const int ny21 = 2 * cp.ny + 1;
  // staged fluid: M32 in the scaled tile frame (w = 2^-r'^2 inside the cut-off, sums
  // rescaled below), the lane's own entry giving its position there
  R gxs = 0, gys = 0;
  R xs = xi, ys = yi;
  if (staged) {
    int own = i - P.lo[lsx - P.sc0];
    for (int a = 0; a < lsx - P.sc0; ++a) own += P.len[a];
    const R2 o = s_pos[own];
    xs = o.x;
    ys = o.y;
  }
  const R rc2s = cp.rc2s;
  // one staged candidate (fluid or wall); w = 0 outside the cut-off; returns r^2 (scaled)
  auto staged_term = [&](int l) {
    const R2 pj = s_pos[l];
    const R dx = pj.x - xs, dy = pj.y - ys, r2 = dx * dx + dy * dy;
    if constexpr (Pol::MIXED) {
      const R e = __builtin_amdgcn_exp2f(-r2);
      const R w = r2 < rc2s ? e : R(0);
      gxs += dx * w;
      gys += dy * w;
    } else {
      const R w = K.w(r2);
      gx += dx * w;
      gy += dy * w;
    }
    return r2;
  };
  auto staged_add = [&](int l, bool) { (void)staged_term(l); };
  // first pass of several: the runs to record (see kShiftRecord)
  RunRec rr;
  rr.ok = staged;
  if (cached) {
    const unsigned x = own.wc.x, y = own.wc.y, z = own.wc.z;
    const int s0 = x & 0x7ff, s1 = (x >> 11) & 0x7ff, s2 = y & 0x7ff, s3 = (y >> 11) & 0x7ff, s4 = z & 0x7ff;
    for_range_4_2_1(s0, s0 + (int)((x >> 22) & 0x7f), staged_add);
    for_range_4_2_1(s1, s1 + (int)((y >> 22) & 0x7f), staged_add);
    for_range_4_2_1(s2, s2 + (int)((z >> 11) & 0x7f), staged_add);
    for_range_4_2_1(s3, s3 + (int)((z >> 18) & 0x7f), staged_add);
    for_range_4_2_1(s4, s4 + (int)(z >> 25), staged_add);
  } else {
    const R rl2 = cp.rl2_rec;
    for_windows<false>(P, cp, lsx, sy, cy, lp, Rp, [&](int sc, int ra, int rb) {
      if (staged) {
        const int sa = sc - P.sc0;
        const int lo = s_lb[sa][ra - P.r0], hi = s_lb[sa][rb + 1 - P.r0];
        if constexpr (MODE == kShiftRecord) {
          // first and last candidate within rc (1 + s)
          if (hi - lo <= 32) {  // from the mask of those candidates, 4 bits placed at a time
            unsigned mk = 0u;
            int l = lo, sh = 0;
#pragma unroll 1
            for (const int last = hi - 3; l < last; l += 4, sh += 4) {
              unsigned nib = staged_term(l) < rl2 ? 1u : 0u;
              nib |= staged_term(l + 1) < rl2 ? 2u : 0u;
              nib |= staged_term(l + 2) < rl2 ? 4u : 0u;
              nib |= staged_term(l + 3) < rl2 ? 8u : 0u;
              mk |= nib << sh;
            }
            if (l + 1 < hi) {
              unsigned nib = staged_term(l) < rl2 ? 1u : 0u;
              nib |= staged_term(l + 1) < rl2 ? 2u : 0u;
              mk |= nib << sh;
              l += 2;
              sh += 2;
            }
            if (l < hi) mk |= (staged_term(l) < rl2 ? 1u : 0u) << sh;
            if (mk != 0u) {
              const int f = __builtin_ctz(mk);
              rr.put(lo + f, 32 - __builtin_clz(mk) - f);
            }
          } else {  // f follows l until one is seen
            int f = lo, la = lo;
            bool seen = false;
            for_range_4_2_1(lo, hi, [&](int q, bool) {
              const bool in = staged_term(q) < rl2;
              f = seen ? f : q;
              seen = seen || in;
              la = in ? q : la;
            });
            if (seen) rr.put(f, la - f + 1);
          }
        } else {
          for_range_4_2_1(lo, hi, staged_add);
        }
      } else {  // global views, relative to their own cell: one run per cell row
        const R bx = xi - cell_shift_x<Pol>((sc >> 1) - cx, cp);
        for (int cr = ra >> 1; cr <= (rb >> 1); ++cr) {
          const R by = yi - cell_shift_y<Pol>(cr - cy, cp);
          const int lo = lbc[sc * ny21 + max(ra, 2 * cr)], hi = lbc[sc * ny21 + min(rb, 2 * cr + 1) + 1];
          for_range<4>(lo, hi, [&](int s) { add(A.vx[s] - bx, A.vy[s] - by); });
        }
      }
    });
  }
  if constexpr (MODE == kShiftRecord) d.f_wc[i] = rr.entry(cum0);
  // walls of the 3x3 cells (cell column by cell column, rows in order), then solids
  if (wstaged) {
    const int lr0 = max(cy - 1, 0) - W.wcr0, lr1 = min(cy + 1, cp.ny - 1) - W.wcr0;
#pragma unroll
    for (int a = 0; a < 3; ++a) for_range_4_2_1(s_wlb[a][lr0], s_wlb[a][lr1 + 1], staged_add);
  }
  for (int a = -1; a <= 1 && P.sw; ++a) {
    const int jx = cx + a;
    if (jx < 0 || jx >= cp.nx) continue;
    const R bx = xi - cell_shift_x<Pol>(a, cp);
    for (int b = -1; b <= 1 && (P.sw & 1) && !wstaged; ++b) {
      const int jy = cy + b;
      if (jy < 0 || jy >= cp.ny) continue;
      const R by = yi - cell_shift_y<Pol>(b, cp);
      const int2 rw = d.w_range[cp.cell_off + jx * cp.ny + jy];
      for_range<4>(rw.x, rw.y, [&](int s) { add(ld32(d.w_vx, s) - bx, ld32(d.w_vy, s) - by); });
    }
  }
  for (int a = -1; a <= 1 && (P.sw & 2); ++a) {
    const int jx = cx + a;
    if (jx < 0 || jx >= cp.nx) continue;
    const R bx = xi - cell_shift_x<Pol>(a, cp);
    for (int b = -1; b <= 1; ++b) {
      const int jy = cy + b;
      if (jy < 0 || jy >= cp.ny) continue;
      const R by = yi - cell_shift_y<Pol>(b, cp);
      const int2 rs = d.s_range[cp.cell_off + jx * cp.ny + jy];
      for_range<2>(rs.x, rs.y, [&](int k) {
        const int j = d.s_sorted[k];
        add(d.s_vx[j] - bx, d.s_vy[j] - by);
      });
    }
  }
  if constexpr (Pol::MIXED) {
    gx += gxs * cp.kdc_s;
    gy += gys * cp.kdc_s;
  }
  const R umax = d.umax[c];
  const R shift_k = (R)(cp.shift_ch * d.dtv[c]);  // particle_shift_coef h dt
  const R dxs = -shift_k * gx, dys = -shift_k * gy;
  R sxv = dxs, syv = dys;
  if (cls == 0) {  // SPLASH: no shift
    sxv = 0;
    syv = 0;
  } else if (cls == 1 || cls == 2) {  // FREE_SURFACE / VICINITY: tangential part only
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
    sxv = (R(1) - nx * nx) * dxs - nx * ny * dys;
    syv = -ny * nx * dxs + (R(1) - ny * ny) * dys;
  }
  sxv *= umax;
  syv *= umax;
  const R2 sn{sacc.x + sxv, sacc.y + syv};
  double x = 0.0, y = 0.0;
  if constexpr (Pol::MIXED) {
    // the views move by this pass's shift in float; the masters take the accumulated shift
    // in double on the last pass
    B.vx[i] = xi + sxv;
    B.vy[i] = yi + syv;
    if (last) {
      x = x0 + (double)sn.x;
      y = y0 + (double)sn.y;
    }
  } else {
    x = x0 + sxv;
    y = y0 + syv;
    B.x[i] = x;
    B.y[i] = y;
  }
  // path length: this pass moved the particle by (sxv, syv); the passes after a recording
  // one count from it and the last adds the path since the hash it recorded
  const float sx2 = (float)sxv, sy2 = (float)syv, step = __builtin_amdgcn_sqrtf(sx2 * sx2 + sy2 * sy2);
  if constexpr (MODE == kShiftPlain)
    B.cum[i] = cum0 + step;
  else if constexpr (MODE == kShiftRecord)
    B.cum[i] = last ? cum0 + step : step;
  else
    B.cum[i] = last ? (own.cum_hash + cum0) + step : cum0 + step;
  if (last) {
    if constexpr (Pol::MIXED) {
      A.x[i] = x;
      A.y[i] = y;
    }
    const R *L = d.f_L + 4 * i;
    A.u[i] += (double)(L[0] * sn.x + L[1] * sn.y);
    A.v[i] += (double)(L[2] * sn.x + L[3] * sn.y);
  } else {
    d.f_sacc[2 * i] = sn.x;
    d.f_sacc[2 * i + 1] = sn.y;
  }
}


// correctionCalculate (src/fluid.cu:316-367) + correctionUpdate / shiftDisplacementCalculate
// (src/fluid.cu:369-427), one tile per block: stage the tile, then shift_tile.  MODE is
// kShiftPlain, kShiftRecord or kShiftCached (the step's one pass, its first of several, a
// later one).  The cached passes are held to 8 waves per SIMD (64 registers, a few spilled:
// 6% faster than 7), the others to 7.
template <class Pol, int BLOCK, int MODE>
__global__ void __launch_bounds__(BLOCK, (MODE == kShiftCached ? 8 : 7))
    k_fluid_shift(Dev<Pol> d, const TilePlan *plans, const int *ntiles, int pass, int use_runs) {
  using R = typename Pol::R;
  using R2 = typename V2<R>::T;
  constexpr int CAP = ShiftCap<BLOCK>::value;
  // walls of the tile's 3x3 cell columns over its cell rows follow the fluid candidates
  // (a tile's fluid candidates rarely fill CAP, so the walls also use what they leave)
  constexpr int WCAP = BLOCK / 2, WROWS = TileRows<BLOCK>::value / 2 + 2;
  // fluid candidates, padding, walls, padding (the padding is far away: read by masked tails)
  __shared__ R2 s_pos[CAP + WCAP + 8];
  __shared__ int s_lb[kTileCols][TileRows<BLOCK>::value + 1];
  __shared__ int s_wlb[3][WROWS + 1];  // LDS index of the first wall of cell row wcr0 + k or later
  __shared__ float s_dmax[BLOCK / 64];
  const int nt = *ntiles, tile = xcd_tile(blockIdx.x, nt);
  if (tile >= nt) return;
  const TilePlan &P = plans[tile];
  const int t = threadIdx.x, c = P.c;
  const auto &A = d.fa;
  const auto &cp = d.cases[c];
  const bool staged = P.total <= CAP && P.nr <= TileRows<BLOCK>::value;
  const ShiftOwn<Pol> own = shift_own<Pol, MODE>(d, P, pass, cp.K);
  int dmax = 0;  // bit pattern of the largest path length (non-negative floats order as ints)
  if (staged) {
    struct Ld {
      R2 pos;
      int pk;
      int cum;
    };
    const R fdy = cp.cdy_r, gs = cp.gscale;
    // this thread's span lies in one cell column: its x offset in the tile frame, scaled
    const int sa = min((int)(((unsigned)t * (unsigned)P.gmul) >> 20), P.nsc - 1);
    const R xo = (R)(((P.sc0 + sa) >> 1) - P.cx) * cp.cdx_r * gs;
    const int gcy = P.cy_ref;
    tile_stage<BLOCK, 4>(
        P, d.f_lb, s_lb,
        [&](int g) {
          return Ld{R2{ld32(A.vx, g), ld32(A.vy, g)}, Pol::MIXED ? ld32(d.f_cxy, g) : 0,
                    ld32(reinterpret_cast<const int *>(A.cum), g)};
        },
        [&](int l, const Ld &v, int) {
          if constexpr (Pol::MIXED)  // tile frame, scaled units
            s_pos[l] = R2{fmaf(v.pos.x, gs, xo), (v.pos.y + (float)(((v.pk & 0xffff) >> 1) - gcy) * fdy) * gs};
          else
            s_pos[l] = v.pos;
          dmax = max(dmax, v.cum);
        });
    if (t < 4) s_pos[P.total + t] = R2{R(1e18), R(1e18)};
  }
  // walls: staged after the fluid candidates' padding
  const WallSpans W = wall_spans(d, P, cp);
  const int wbase = P.total + 4;
  const bool wstaged = staged && W.n > 0 && P.total + W.n <= CAP + WCAP && W.nwr <= WROWS;
  if (wstaged) {
    const R fdx = cp.cdx_r, fdy = cp.cdy_r, gs = cp.gscale;
    wall_stage<BLOCK, WROWS>(d, W, wbase, s_wlb, [&](int l, int g) {
      const R wx = ld32(d.w_vx, g), wy = ld32(d.w_vy, g);
      if constexpr (Pol::MIXED) {  // tile frame, scaled units
        const int wk = ld32(d.w_cxy, g);
        s_pos[l] = R2{(wx + (float)(cxy_x(wk) - P.cx) * fdx) * gs, (wy + (float)(cxy_y(wk) - P.cy_ref) * fdy) * gs};
      } else {
        s_pos[l] = R2{wx, wy};
      }
    });
    if (t < 4) s_pos[wbase + W.n + t] = R2{R(1e18), R(1e18)};
  }
  for (int o = 32; o > 0; o >>= 1) dmax = max(dmax, __shfl_xor(dmax, o, 64));
  if ((t & 63) == 0) s_dmax[t >> 6] = __int_as_float(dmax);
  __syncthreads();
  float D = staged_bound<BLOCK>(s_dmax), Dsp = 0.f;
  if constexpr (MODE == kShiftRecord) {
    if (t == 0) d.tile_d0[tile] = D;
  } else if constexpr (MODE == kShiftCached) {  // the staged paths count from the first pass
    Dsp = D;
    D = d.tile_d0[tile] + Dsp;
  }
  shift_tile<Pol, BLOCK, MODE, int, WROWS>(d, P, pass, own, s_pos, s_lb, s_wlb, D, Dsp, staged && use_runs, staged,
                                           wstaged, W);
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

// A lane's entries e = e0, e0 + kSolidLanes, ... < e1 of a reference list, U at a time: the U
// neighbour indices are read together, then the U neighbours' fields through them, then
// use(e, j, v) runs in entry order -- two waits per U entries instead of two per entry.
template <int U, class Ld, class F>
__device__ __forceinline__ void for_refs(const int *ref_j, int e0, int e1, Ld &&load, F &&use) {
  for (int b = e0; b < e1; b += U * kSolidLanes) {
    int jj[U];
#pragma unroll
    for (int u = 0; u < U; ++u) jj[u] = ref_j[min(b + u * kSolidLanes, e1 - 1)];
    decltype(load(0, 0)) v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) v[u] = load(min(b + u * kSolidLanes, e1 - 1), jj[u]);
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (b + u * kSolidLanes < e1) use(b + u * kSolidLanes, jj[u], v[u]);
  }
}

// solidCoupleUpdate (src/solid.cu:355-444): Shepard pressure from the fluid, normal from the
// solid neighbours in the current cells; and the check that every reference-list neighbour
// still lies within the 3x3 current cells.
template <class Pol>
__device__ __forceinline__ void solid_couple_block(const Dev<Pol> &d, int bid) {
  using R = typename Pol::R;
  const int gid = bid * 256 + (int)threadIdx.x;
  const int i = gid / kSolidLanes, l = gid % kSolidLanes;
  R nx = 0, ny = 0, P = 0, S = 0;
  int left = 0;
  int c = 0;
  // a case without fluid is not coupled (coupleCalculation returns, src/solid.cu:494-499):
  // its solids keep their pressure and normal, also in a batch with fluid in other cases
  bool active = i < d.ns;
  if (active) {
    c = case_of(i, d.s_case_off, d.s_case_blk, d.ncase);
    active = d.f_case_off[c + 1] > d.f_case_off[c];
  }
  if (active) {
    const auto &cp = d.cases[c];
    const int pk = d.s_cxy[i], cx = cxy_x(pk), cy = cxy_y(pk);
    const R xi = d.s_vx[i], yi = d.s_vy[i];
    const LoopConsts<R> K(cp);
    const R eR = cp.eR, k2c = cp.k2c, hinv2 = cp.hinv2;
    const auto &F = d.fa;
    // fluid: the 3x3 cells' 18 runs spread over the group's lanes (as in the force kernel)
    for (int r = l; r < kCellRuns; r += kSolidLanes) {
      int2 rr;
      R bx, by;
      if (!cell_run(d, cp, cx, cy, r, xi, yi, rr, bx, by)) continue;
      for_run<4>(rr.x, rr.y, [&](int s) { return R2c<R>{F.vx[s], F.vy[s], F.p[s]}; }, [&](int, const R2c<R> &f) {
        const R dx = f.x - bx, dy = f.y - by;
        const R r2 = dx * dx + dy * dy;
        if (r2 < K.rc2) {
          const R w = (kexp<R>(-(r2 * hinv2)) - eR) * k2c;
          S += w;
          P += f.p * w;
        }
      });
    }
    // solids: one cell of the 3x3 per lane
    if (l < 9) {
      const int a = l / 3 - 1, b = l % 3 - 1, jx = cx + a, jy = cy + b;
      if (jx >= 0 && jx < cp.nx && jy >= 0 && jy < cp.ny) {
        const R bx = xi - cell_shift_x<Pol>(a, cp), by = yi - cell_shift_y<Pol>(b, cp);
        const int2 rs = d.s_range[cp.cell_off + jx * cp.ny + jy];
        for (int k = rs.x; k < rs.y; ++k) {
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
template <class Pol>
__global__ void __launch_bounds__(256) k_solid_couple(Dev<Pol> d) {
  solid_couple_block<Pol>(d, (int)blockIdx.x);
}
// The solid coupling and the wall update in one launch (they read the fluid only): blocks below
// wall_blocks update walls, the rest couple solids.
template <class Pol, int kWallLanes>
__global__ void __launch_bounds__(256) k_couple_wall(Dev<Pol> d, int wall_blocks) {
  if ((int)blockIdx.x < wall_blocks)
    wall_update_block<Pol, kWallLanes>(d, (int)blockIdx.x);
  else
    solid_couple_block<Pol>(d, (int)blockIdx.x - wall_blocks);
}
// The cell columns' row tables (k_subcol_rows, 4 waves a cell column), the wall update and the
// solid coupling in one launch at the end of a step's cell stage: the last two read the sorted
// fluid and its sub-cell ranges, which the row tables leave as they are, and nothing they write
// is read by the tables or the tile counts.  Blocks below row_blocks build the tables, the next
// wall_blocks update walls, the rest (if any) couple solids.
template <class Pol, int kWallLanes>
__global__ void __launch_bounds__(256) k_rows_wall(Dev<Pol> d, int G, const int *col_case, PlanJobs jobs, int njobs,
                                                   int *reset_moved, int row_blocks, int wall_blocks) {
  const int b = (int)blockIdx.x;
  if (b < row_blocks)
    subcol_rows_block<Pol, 4>(d, G, col_case, jobs, njobs, reset_moved, b);
  else if (b < row_blocks + wall_blocks)
    wall_update_block<Pol, kWallLanes>(d, b - row_blocks);
  else
    solid_couple_block<Pol>(d, b - row_blocks - wall_blocks);
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
__device__ __forceinline__ void solid_deform_body(const Dev<Pol> &d, int sub, int gid) {
  using R = typename Pol::R;
  const int i = gid / kSolidLanes, l = gid % kSolidLanes;
  const bool active = i < d.ns;
  int c = 0;
  R h00 = 0, h01 = 0, h10 = 0, h11 = 0;
  if (active) {
    c = case_of(i, d.s_case_off, d.s_case_blk, d.ncase);
    if (sub < d.cases[c].S) {
      const double dix = d.sx[i] - d.sX[i], diy = d.sy[i] - d.sY[i];
      struct Nb {
        R gX, gY;
        double x, X, y, Y;
      };
      for_refs<4>(
          d.s_ref_j, d.s_ref_off[i] + l, d.s_ref_off[i + 1],
          [&](int e, int j) { return Nb{d.s_ref_g[2 * e], d.s_ref_g[2 * e + 1], d.sx[j], d.sX[j], d.sy[j], d.sY[j]}; },
          [&](int, int, const Nb &n) {
            const R ux = (R)((n.x - n.X) - dix), uy = (R)((n.y - n.Y) - diy);
            h00 += ux * n.gX;
            h01 += ux * n.gY;
            h10 += uy * n.gX;
            h11 += uy * n.gY;
          });
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

template <class Pol>
__global__ void __launch_bounds__(256) k_solid_deform(Dev<Pol> d, int sub) {
  solid_deform_body(d, sub, (int)(blockIdx.x * blockDim.x + threadIdx.x));
}

// solidForceCalculate (src/solid.cu:307-353: solidFluidCouple :211-255, solidSolidForce
// :257-305) fused with solidUpdate (:446-471); positions and velocities ping-pong.
template <class Pol>
__device__ __forceinline__ void solid_force_body(const Dev<Pol> &d, int sub, int gid) {
  using R = typename Pol::R;
  const int i = gid / kSolidLanes, l = gid % kSolidLanes;
  const bool active = i < d.ns;
  int c = 0;
  bool run = false;
  R cpx = 0, cpy = 0, ax = 0, ay = 0;
  if (active) {
    c = case_of(i, d.s_case_off, d.s_case_blk, d.ncase);
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
      // the 3x3 cells' 18 fluid runs (one per sub-column of a cell) go to the group's lanes,
      // each lane summing its runs in slot order: the lanes read their runs' ranges together
      // instead of the group walking the cells one after another
      for (int r = l; r < kCellRuns; r += kSolidLanes) {
        int2 rr;
        R bx, by;
        if (!cell_run(d, cp, cx, cy, r, xiv, yiv, rr, bx, by)) continue;
        for_run<4>(rr.x, rr.y, [&](int s) { return FluidNb<R>::load(F, d.f_phase, s); }, [&](int, const FluidNb<R> &f) {
          const R dx = f.x - bx, dy = f.y - by;
          const R r2 = dx * dx + dy * dy;
          if (!(r2 < K.rc2)) return;
          const R w = K.w(r2);
          const R inv_r = r2 > R(0) ? krsqrt<R>(r2) : R(0);
          R ps = R(0.5) * (f.p + psi);
          const R du = ((uiv - f.u) * dx + (viv - f.v) * dy) * inv_r;
          if (du >= R(0)) ps += R(0.5) * f.rho * du * fmin(R(3) * du, f.phase ? c1 : c0);
          const R t = R(-2) * ps * rr_inv * w;
          cpx += t * dx;
          cpy += t * dy;
        });
      }
      // solid -> solid over the reference-configuration list
      const R *Fi = d.s_F + 4 * i, *Li = d.s_L + 4 * i, *PLi = d.s_PL + 4 * i;
      const Mat2<R> FL = mmul(Mat2<R>{Fi[0], Fi[1], Fi[2], Fi[3]}, Mat2<R>{Li[0], Li[1], Li[2], Li[3]});
      const R pl0 = PLi[0], pl1 = PLi[1], pl2 = PLi[2], pl3 = PLi[3];
      const R h = cp.h, eps_h2 = cp.eps_h2, alpha = cp.alpha, beta = cp.beta, cs = cp.c_s;
      struct Nb {
        R gX, gY, p0, p1, p2, p3;
        double x, y, u, v;
      };
      for_refs<4>(
          d.s_ref_j, d.s_ref_off[i] + l, d.s_ref_off[i + 1],
          [&](int e, int j) {
            const R *PLj = d.s_PL + 4 * j;
            return Nb{d.s_ref_g[2 * e], d.s_ref_g[2 * e + 1], PLj[0], PLj[1], PLj[2], PLj[3],
                      d.sx[j], d.sy[j], d.su[j], d.sv[j]};
          },
          [&](int, int, const Nb &n) {
            ax += ((n.p0 + pl0) * n.gX + (n.p1 + pl1) * n.gY) * rr_inv;
            ay += ((n.p2 + pl2) * n.gX + (n.p3 + pl3) * n.gY) * rr_inv;
            const R dx = (R)(n.x - xi), dy = (R)(n.y - yi);
            const R du = (R)(n.u - ui), dv = (R)(n.v - vi);
            const R nu = h * (du * dx + dv * dy) / (dx * dx + dy * dy + eps_h2);
            const R vis = alpha * nu * cs - beta * nu * nu;
            ax += vis * (FL.a * n.gX + FL.b * n.gY);
            ay += vis * (FL.c * n.gX + FL.d * n.gY);
          });
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
  const double dt_s = d.dtv[c] / (double)cp.S;  // solid sub-step (src/solid.cu:456)
  d.s_cp[2 * i] = cpx;
  d.s_cp[2 * i + 1] = cpy;
  double tax = (double)ax + (double)cpx, tay = (double)ay + (double)cpy;
  double x = xi, y = yi, u = ui, v = vi;
  if (d.s_bc[i] == 1) {  // FIXED_PARTICLE
    u = cp.fixed_u;
    v = cp.fixed_v;
    x += u * dt_s;
    y += v * dt_s;
    tax = 0.0;
    tay = 0.0;
  } else {
    if (cp.steady) {
      tax += -0.1 * ui;
      tay += -0.1 * vi;
    }
    x += u * dt_s;
    y += v * dt_s;
    u += (tax + cp.gx_d) * dt_s;
    v += (tay + cp.gy_d) * dt_s;
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

template <class Pol>
__global__ void __launch_bounds__(256) k_solid_force_update(Dev<Pol> d, int sub) {
  solid_force_body(d, sub, (int)(blockIdx.x * blockDim.x + threadIdx.x));
}

// ---- all sub-steps of a fluid step in one launch ----------------------------------------
// A few thousand solids take a few microseconds per phase, so S sub-steps issued as 2 S
// launches are bound by launch latency.  k_solid_substeps runs them in one launch whose
// blocks all stay resident (the host sizes the grid to at most half of what the device holds
// at once, so two contexts stepping concurrently still fit) and meet at a grid barrier
// between the phases.  Every wait is bounded: a block that waits longer than spin_limit polls
// raises SPHNJ_ST_SYNC_TIMEOUT, sets the abort word (which ends every other block's waits)
// and goes on, so every block exits whatever happens.  The last block out resets the barrier
// words for the next launch.
constexpr int kSubstepBlock = 512;
constexpr int kSyncLine = 32;  // barrier words one 128-byte line apart
struct GridSync {
  unsigned *words;  // [k * kSyncLine]: k < 8 arrivals of XCD group k, 8 groups done, 9 blocks
                    // that have left, 10 abort
  __device__ unsigned *group(int k) const { return words + k * kSyncLine; }
  __device__ unsigned *top() const { return words + 8 * kSyncLine; }
  __device__ unsigned *done() const { return words + 9 * kSyncLine; }
  __device__ unsigned *abort() const { return words + 10 * kSyncLine; }
};

// The e-th grid barrier of the launch (e = 1, 2, ...).  Blocks are counted per XCD group
// (block b runs on XCD b mod 8), and the last block of a group counts the group once on the
// top word, so no word takes more than a group's arrivals.  Each wave waits for its own
// global stores; then one thread per block writes the CU's XCD L2 back (the release every
// XCD needs to see them) before arriving, and invalidates its caches after the wait.
__device__ __forceinline__ void grid_sync(const GridSync &g, unsigned e, int spin_limit, unsigned *status) {
  __builtin_amdgcn_s_waitcnt(0x0F70);  // s_waitcnt vmcnt(0): this wave's stores have landed
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned G = gridDim.x, x = blockIdx.x % 8u;
    const unsigned groups = G < 8u ? G : 8u, in_group = (G - x + 7u) / 8u;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    const unsigned old = __hip_atomic_fetch_add(g.group((int)x), 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (old == e * in_group - 1u) {  // the group's last arrival counts the group
      __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "agent");
      __hip_atomic_fetch_add(g.top(), 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    int spins = 0;
    while (__hip_atomic_load(g.top(), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < e * groups) {
      if (__hip_atomic_load(g.abort(), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) break;
      if (++spins > spin_limit) {
        __hip_atomic_fetch_or(status, kStSyncTimeout, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(g.abort(), 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  }
  __syncthreads();
}

// d0 is the state as the sub-steps begin; d1 the same with every ping-pong pair of solid
// buffers swapped (odd sub-steps read what even ones wrote).  The host swaps its own pointers
// S times after the launch, as after 2 S launches.
template <class Pol>
__global__ void __launch_bounds__(kSubstepBlock) k_solid_substeps(Dev<Pol> d0, Dev<Pol> d1, int S, GridSync g,
                                                                  int spin_limit) {
  const int n = d0.ns * kSolidLanes, stride = (int)(gridDim.x * blockDim.x);
  unsigned e = 0;
  for (int k = 0; k < S; ++k) {
    const Dev<Pol> &d = (k & 1) ? d1 : d0;
    for (int gid = blockIdx.x * blockDim.x + threadIdx.x; gid - (int)threadIdx.x < n; gid += stride)
      solid_deform_body(d, k, gid);
    grid_sync(g, ++e, spin_limit, d0.status);
    for (int gid = blockIdx.x * blockDim.x + threadIdx.x; gid - (int)threadIdx.x < n; gid += stride)
      solid_force_body(d, k, gid);
    if (k + 1 < S) grid_sync(g, ++e, spin_limit, d0.status);
  }
  // the last block out resets the barrier words (every other block has made its last read)
  __syncthreads();
  if (threadIdx.x == 0 &&
      __hip_atomic_fetch_add(g.done(), 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) == gridDim.x - 1) {
    for (int k = 0; k < 8; ++k) __hip_atomic_store(g.group(k), 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(g.top(), 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(g.abort(), 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(g.done(), 0u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// ---- init: reference-configuration neighbour lists --------------------------------------

// Count of reference neighbours (|X_j - X_i|^2 < rc^2, 9 cells of the reference hash).
template <class Pol>
__global__ void __launch_bounds__(256) k_ref_count(Dev<Pol> d, const int *cxy, const int2 *range,
                                                   const int *sorted, const double *rc2_d, int *count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= d.ns) return;
  const int c = case_of(i, d.s_case_off, d.s_case_blk, d.ncase);
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
  const int c = case_of(i, d.s_case_off, d.s_case_blk, d.ncase);
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
  const int c = case_of(i, d.f_case_off, d.f_case_blk, d.ncase);
  const auto &cp = d.cases[c];
  const auto &A = d.fa;
  const int pk = d.f_cxy[i], cx = cxy_x(pk) >> 1, cy = cxy_y(pk) >> 1;
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
      int2 run[2];
      fluid_cell_runs(d, cp, jx, jy, run[0], run[1]);
      for (int h = 0; h < 2; ++h) {
        for (int s = run[h].x; s < run[h].y; ++s) {
          const R dx = A.vx[s] - bx, dy = A.vy[s] - by;
          if (dx * dx + dy * dy < cp.rc2) {
            if (out && n < lim) out[n] = (0 << 28) | (d.f_orig[s] - cp.f_off);
            ++n;
          }
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

// step_dt_body as a launch of its own (a step without fluid).
template <class Pol>
__global__ void __launch_bounds__(64) k_step_dt(Dev<Pol> d) {
  step_dt_body(d, (int)threadIdx.x, 64);
}

// ---- probes of the reference's global outputs (src/output.cu:283-582) ----------------------
// One case.  Per-block partials in a fixed order (the host combines them block by block, so a
// probe's result does not depend on scheduling):
//   [0], [1] fluid force on the solids, sum couple_dudt rho_ref d^2 (outputTotalForce)
//   [2] largest x of a liquid particle (outputEdge's frontier), -inf if none
//   [3] largest y of a liquid particle within d/2 of x = xh (outputEdge's heights), -inf if none
//   [4], [5] largest von Mises stress of the elastic / auxetic solids, with the stress
//       F^-1 P (outputVonMisesStress)
//   [6] sum of the solid pressures (outputMeanPressure)
constexpr int kProbeParts = 7, kProbeBlocks = 256;
struct ProbeArgs {
  int flo, fhi, slo, shi;
  double d2, rho_s, rho_aux, xh, half_d;
};
template <class Pol>
__global__ void __launch_bounds__(256) k_probe_partial(Dev<Pol> d, ProbeArgs a, double *part) {
  __shared__ double sm[kProbeParts][256];
  const int t = threadIdx.x;
  double v[kProbeParts] = {0.0, 0.0, -INFINITY, -INFINITY, 0.0, 0.0, 0.0};
  for (int i = a.flo + blockIdx.x * 256 + t; i < a.fhi; i += gridDim.x * 256) {
    if (d.f_phase[i] != 0) continue;
    const double x = d.fa.x[i];
    v[2] = fmax(v[2], x);
    if (fabs(x - a.xh) < a.half_d) v[3] = fmax(v[3], d.fa.y[i]);
  }
  for (int i = a.slo + blockIdx.x * 256 + t; i < a.shi; i += gridDim.x * 256) {
    const bool aux = d.s_mat[i] == 1;
    const double rr = aux ? a.rho_aux : a.rho_s;
    v[0] += (double)d.s_cp[2 * i] * rr * a.d2;
    v[1] += (double)d.s_cp[2 * i + 1] * rr * a.d2;
    const double f0 = d.s_F[4 * i], f1 = d.s_F[4 * i + 1], f2 = d.s_F[4 * i + 2], f3 = d.s_F[4 * i + 3];
    const double p0 = d.s_P[4 * i], p1 = d.s_P[4 * i + 1], p2 = d.s_P[4 * i + 2], p3 = d.s_P[4 * i + 3];
    const double idet = 1.0 / (f0 * f3 - f1 * f2);  // F^-1 = [[f3, -f1], [-f2, f0]] / det F
    const double s00 = (f3 * p0 - f1 * p2) * idet, s01 = (f3 * p1 - f1 * p3) * idet;
    const double s11 = (-f2 * p1 + f0 * p3) * idet;
    const double vm = sqrt(0.5 * (s00 - s11) * (s00 - s11) + 3.0 * s01 * s01);
    v[aux ? 5 : 4] = fmax(v[aux ? 5 : 4], vm);
    v[6] += (double)d.s_p[i];
  }
#pragma unroll
  for (int k = 0; k < kProbeParts; ++k) sm[k][t] = v[k];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) {
      sm[0][t] += sm[0][t + w];
      sm[1][t] += sm[1][t + w];
      sm[2][t] = fmax(sm[2][t], sm[2][t + w]);
      sm[3][t] = fmax(sm[3][t], sm[3][t + w]);
      sm[4][t] = fmax(sm[4][t], sm[4][t + w]);
      sm[5][t] = fmax(sm[5][t], sm[5][t + w]);
      sm[6][t] += sm[6][t + w];
    }
    __syncthreads();
  }
  if (t < kProbeParts) part[blockIdx.x * kProbeParts + t] = sm[t][0];
}

// The lowest index among the case's solids whose reference position lies within d/2 of
// (X, Y) (outputTankBeamEnd, outputEndPosition, outputMiddlePosition, outputWedgeDisp find
// their point this way), into *best (INT_MAX when none); the caller reads its position.
template <class Pol>
__global__ void __launch_bounds__(256) k_probe_point(Dev<Pol> d, int slo, int shi, double X, double Y, double r2,
                                                     int *best) {
  const int i = slo + blockIdx.x * 256 + (int)threadIdx.x;
  if (i >= shi) return;
  const double dX = d.sX[i] - X, dY = d.sY[i] - Y;
  if (dX * dX + dY * dY < r2) atomicMin(best, i);
}

}  // namespace sphnj
