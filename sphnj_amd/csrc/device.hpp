// Device-side types and helpers of libsphnj (gfx950, wave64).
//
// Precision policies: P64 computes everything in double; M32 computes pair sums, rates,
// pressure views and stress in float on cell-relative float coordinates, and keeps the
// absolute positions, velocities and densities (the integrated state) in double.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace sphnj {

struct P64 {
  using R = double;
  static constexpr bool MIXED = false;
};
struct M32 {
  using R = float;
  static constexpr bool MIXED = true;
};

// ---- per-case constants (one entry per case of a batch) ---------------------------------
template <class R>
struct CaseDev {
  // kernel (Kernel::initiate, src/kernel.cpp:25-41)
  R rc2, hinv2, kdc, k2c, eR, h, eps_h2;  // eps_h2 = 0.01 h^2
  R nhl2;  // -hinv2 * log2(e): exp(-r^2/h^2) = 2^(r^2 nhl2)
  // scaled units of the M32 shift loop: x' = x sqrt(hinv2 log2 e), so exp(-r^2/h^2) = 2^-r'^2
  float gscale, rc2s, kdc_s;  // scale, rc^2 scaled, kd c32 / scale
  R quarter_rc2;
  // fluid per phase: pair-side constants in R
  R c[2], cinv[2], mu8[2], rho0[2];
  // fluid per phase: integration-side constants in double
  double rho0_d[2], rho0_inv_d[2], gam_d[2], ginv_d[2], rho2p_d[2], p2rho_d[2], rho_min_d[2];
  double p_min_d, p_bg_d;
  R p_min;
  R gx, gy;
  double gx_d, gy_d, dt;
  double shift_ch;  // particle_shift_coef * h (times the step's dt: the shift scale)
  // adaptive step size (cfl = 0: every step takes dt): C h, the larger sound speed, and the
  // solid cap S C h / c_s (0 without solids)
  double cfl_h, c_max, dt_solid_cap;
  int adaptive;
  R umax_cap;    // 0.1 * min(c_L, c_G)
  // domain / cell grid (Domain::initiate, src/domain.cpp:37-48)
  double cxmin, cymin, cdx, cdy, inv_cdx, inv_cdy;
  R cdx_r, cdy_r;
  // half-cell ("sub-cell") grid of the fluid order: sizes and the pruning cut-off rc (1 + 1e-5)
  double rcut_prune;
  // shift windows kept over a step's passes (skin s rc, s = kShiftSkin): the recording radius
  // as the first pass compares it (M32 scaled r'^2, P64 r^2) of rc (1 + s), the widening of
  // that pass's pruning radius, and the largest path the lane and its candidates may have
  // taken since that pass for the windows to hold every neighbour within rc; rl2_pair is the
  // recording radius as the pair kernel compares it (unscaled r^2), when it records the runs
  R rl2_rec, rl2_pair;
  float skin_grow, skin_ok;
  float ssx, ssy, inv_ssx, inv_ssy;
  double dxmin, dymin, dxmax, dymax;
  int nx, ny, cell_off, ncell;
  int col_off, lb_off;  // first global sub-column; offset in the fluid row table f_lb
  int swp_off;          // offset of the case's cell columns in f_swp
  int steady, viscid;
  int K, S;  // shift passes and solid sub-steps of this case
  // solid (SolidProperties, src/properties.cpp:69-89,118-161)
  R lambda, mu_s, c_s, alpha, beta;
  R rho_s_inv, rho_aux_inv;
  double fixed_u, fixed_v;
  R aux_mod, yield_stress, harden_strain, harden_coef;
  int plastic;
  // particle offsets of this case in the concatenated species arrays
  int f_off, f_n, s_off, s_n, w_off, w_n;
};

// ---- device buffers --------------------------------------------------------------------
// Fluid arrays are in cell-sorted slot order and re-permuted every step.  "view" arrays are
// what neighbour loops read: for P64 they alias the double masters; for M32 they are float
// copies with positions relative to the particle's own cell origin.  FluidBuf holds the
// fields that ping-pong between two buffers (permute, pair+update, shift passes).
template <class Pol>
struct FluidBuf {
  using R = typename Pol::R;
  double *x, *y, *u, *v, *rho;  // masters
  R *p;
  R *vx, *vy, *vu, *vv, *vrho;  // views
  float *cum;  // path length since the hash (sum of step lengths): bounds |x - x at hash|
};

template <class Pol>
struct Dev {
  using R = typename Pol::R;
  const CaseDev<R> *cases;
  int ncase;
  int total_cells;
  const int *f_case_off, *s_case_off, *w_case_off;  // ncase + 1 each
  const int *f_case_blk, *s_case_blk, *w_case_blk;  // case of slot 256 k, per species
  int nf, ns, nw;
  // fluid
  FluidBuf<Pol> fa, fb;  // cur (a) / alt (b)
  int8_t *f_phase, *f_phase2, *f_cls, *f_cls2;  // cur / alt (swap in the permute only)
  int *f_orig;           // slot -> global original index (sorted values)
  int *f_cxy;            // slot -> (sx << 16) | sy, sub-cell coordinates (cx = sx >> 1)
  // counting sort on the sub-cell keys: per key (4 total_cells + 1) the slots hashed there this
  // step (zero between steps); per slot of the order being replaced its key and its rank among
  // its key's slots in atomic order; per slot in atomic order the old slot, original index and
  // key (atomic order only permutes the slots inside a key, so f_pkey holds the sorted keys)
  int *f_cnt;
  int *f_hkey, *f_hrank;
  int *f_perm, *f_porig, *f_pkey;
  R *f_col;
  int8_t *f_near;
  R *f_L;  // 4 per slot
  // shift displacement accumulated over the passes of a step (2 per slot; velocities and
  // M32 masters take it on the case's last pass)
  R *f_sacc;
  // per slot, the first shift pass of a step records the lane's candidate runs for the later
  // passes: up to 5 runs of the staged LDS order (11-bit start, 7-bit length) packed in x, y
  // and z, bit 31 of x set when they are complete, and the lane's path length since the hash
  // at that pass (float bits) in w; per tile of the shift, the largest such path among its
  // staged candidates
  uint4 *f_wc;
  float *tile_d0;
  int2 *f_srange;  // per global sub-cell (4 per cell): [start, end) slot range, empty (0, 0)
  // per sub-column, 2 ny + 1 entries: first slot of sub-row k or later; the sub-column's end
  int *f_lb;
  // per cell column, ny + 1 entries: cells among rows [0, k) that hold a solid particle
  int *f_swp;
  // max |u|^2 per case as the bit pattern of a non-negative double: blocks reduce into
  // kSpread slots per case (slot s at umax_slot[(c kSpread + s) kSlotStride], one cache line
  // each, so no address takes more than a few hundred atomics); umax2[c] is their maximum
  // non-zero when a solid's cell key changed since the last sort of the solids (set by the
  // solid hash, cleared after the step's cell stage): the solids' sort, cell ranges and their
  // clearing run only then (a body moves by far less than a cell per step)
  int *s_moved;
  unsigned long long *umax_slot, *umax2;
  // per case: max |u|^2 after the last fluid update (the input speeds before the first), the
  // step size of the current step and the simulated time at its end
  double *u2_last, *dtv, *tsim;
  R *umax;  // per case: min(sqrt(umax2), umax_cap), the speed the shift scales with
  // solid (original order, never permuted)
  double *sx, *sy, *su, *sv;     // cur
  double *sx2, *sy2, *su2, *sv2; // alt (sub-step ping-pong)
  double *sX, *sY;
  int8_t *s_mat, *s_bc;
  R *s_vx, *s_vy, *s_vx2, *s_vy2;  // views (M32 cell-relative; P64 alias sx/sy)
  R *s_p, *s_n;                    // pressure, normal (2 per particle)
  R *s_F, *s_L, *s_P, *s_PL, *s_M; // 4 per particle
  R *s_Lref, *s_Bt;                // 4 per particle: B^-1 (or I) and B of the reference configuration
  int8_t *s_degen;                 // 1 where det(B) == 0 (L = I, F = F~)
  double *s_a;                     // du_dt, 2 per particle
  R *s_cp;                         // couple_dudt, 2 per particle
  int *s_key, *s_cxy, *s_sorted;   // current hash, (cell, orig)-sorted index
  int2 *s_range;
  int *s_ref_off, *s_ref_j;        // reference-configuration lists (global solid indices)
  R *s_ref_g;                      // 2 per entry: w'(dX) dX, w'(dX) dY
  // wall (sorted once at init)
  double *wx, *wy;
  R *w_vx, *w_vy;
  R *w_u, *w_v, *w_p, *w_rho, *w_n;
  int8_t *w_phase, *w_bc;
  int *w_orig, *w_cxy;
  int2 *w_range;
  // per cell column (case offset swp_off), ny + 1 entries: first wall slot of cell row k or
  // later, the column's end last (walls are sorted by cell once)
  int *w_lb;
  // status word (SPHNJ_ST_* bits)
  unsigned *status;
};

// ---- helpers ----------------------------------------------------------------------------

__device__ __forceinline__ int case_of(int i, const int *off, int ncase) {
  int lo = 0, hi = ncase;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (off[mid] <= i)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

// The same from a table of the case of every 256th slot (blk[k] = case_of(256 k)): one load
// and a short forward walk instead of a binary search (cases hold consecutive slot ranges, so
// the walk passes only cases ending inside the 256 slots, empty ones included).
__device__ __forceinline__ int case_of(int i, const int *off, const int *blk, int ncase) {
  if (ncase == 1) return 0;
  int c = blk[i >> 8];
  while (off[c + 1] <= i) ++c;
  return c;
}

constexpr int kSpread = 64, kSlotStride = 16;

// One axis of findPositionCell (src/pair.cu:8-18): truncation toward zero, clamp outside [0,n).
__device__ __forceinline__ int axis_cell(double x, double x0, double inv, int n, unsigned *fl) {
  const double t = (x - x0) * inv;
  if (!(t > -1.0)) {
    *fl |= 1u;
    return 0;
  }
  if (!(t < (double)n)) {
    *fl |= 1u;
    return n - 1;
  }
  return (int)t;
}

// Sub-cell index along one axis (0 .. 2n-1): 2 * cell + the half of the cell the coordinate
// lies in; a clamped coordinate takes the half on the side it left by.
__device__ __forceinline__ int axis_sub(double x, double x0, double inv, int n, unsigned *fl) {
  const double t = (x - x0) * inv;
  if (!(t > -1.0)) {
    *fl |= 1u;
    return 0;
  }
  if (!(t < (double)n)) {
    *fl |= 1u;
    return 2 * n - 1;
  }
  const int c = (int)t;
  const int h = min(max((int)(2.0 * t) - 2 * c, 0), 1);
  return 2 * c + h;
}

template <class R>
__device__ __forceinline__ R kexp(R x);
template <>
__device__ __forceinline__ double kexp<double>(double x) {
  return exp(x);
}
template <>
__device__ __forceinline__ float kexp<float>(float x) {
  // v_exp_f32 computes 2^x; arguments here lie in [-R^2 log2 e, 0]
  return __builtin_amdgcn_exp2f(x * 1.4426950408889634f);
}

template <class R>
__device__ __forceinline__ R krsqrt(R x);
template <>
__device__ __forceinline__ double krsqrt<double>(double x) {
  return 1.0 / sqrt(x);
}
template <>
__device__ __forceinline__ float krsqrt<float>(float x) {
  return __builtin_amdgcn_rsqf(x);
}

template <class R>
__device__ __forceinline__ R krcp(R x);
template <>
__device__ __forceinline__ double krcp<double>(double x) {
  return 1.0 / x;
}
template <>
__device__ __forceinline__ float krcp<float>(float x) {
  return __builtin_amdgcn_rcpf(x);
}

__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long v) {
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long w = __shfl_xor(v, o, 64);
    v = w > v ? w : v;
  }
  return v;
}

__device__ __forceinline__ float wave_max_f(float v) {
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
  return v;
}

__device__ __forceinline__ unsigned wave_or(unsigned v) {
  for (int o = 32; o > 0; o >>= 1) v |= __shfl_xor(v, o, 64);
  return v;
}

// Element i of an array whose base is wave-uniform, addressed by a 32-bit byte offset (the
// arrays here stay far below 4 GiB), so the load takes the base from scalar registers and
// needs no 64-bit address arithmetic per lane.
template <class T>
__device__ __forceinline__ T ld32(const T *base, int i) {
  return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + (unsigned)i * (unsigned)sizeof(T));
}

__device__ __forceinline__ int wave_sum(int v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// inclusive prefix sum over the lanes of a wave
__device__ __forceinline__ int wave_incl_sum(int v, int lane) {
  for (int o = 1; o < 64; o <<= 1) {
    const int u = __shfl_up(v, o, 64);
    if (lane >= o) v += u;
  }
  return v;
}

}  // namespace sphnj
