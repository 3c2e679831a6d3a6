// libsphnj context: device memory, step orchestration, hipGraph capture, and the C ABI of
// include/sphnj.h.  Replaces Sim (include/sim.cuh), DeviceCalculation::calculation
// (src/calculation.cu:20-48) and the __constant__ parameter / pointer tables
// (src/memory_copy.cu:5-12) with an opaque per-context engine.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "host.hpp"
#include "radix_sort.hpp"
#include "sph_kernels.hpp"

using namespace sphnj;

namespace {

#define HIPC(expr)                                                                  \
  do {                                                                              \
    hipError_t e_ = (expr);                                                         \
    if (e_ != hipSuccess) {                                                         \
      err = std::string(#expr) + ": " + hipGetErrorString(e_);                      \
      return SPHNJ_E_HIP;                                                           \
    }                                                                               \
  } while (0)

#define LAUNCH_CHECK()                                                              \
  do {                                                                              \
    hipError_t e_ = hipGetLastError();                                              \
    if (e_ != hipSuccess) {                                                         \
      err = std::string("kernel launch: ") + hipGetErrorString(e_);                 \
      return SPHNJ_E_HIP;                                                           \
    }                                                                               \
  } while (0)

inline int nblk(int n, int b) { return n <= 0 ? 0 : (n + b - 1) / b; }

// Host copy of one uploaded case.
struct Staged {
  bool uploaded = false;
  std::vector<double> fx, fy, fu, fv, frho, fp;
  std::vector<int8_t> fph;
  std::vector<double> sx, sy, su, sv, sX, sY;
  std::vector<int8_t> smat, sbc;
  std::vector<double> wx, wy;
  std::vector<int8_t> wbc;
  // state a step carries over, restored from a snapshot (empty: the initial values of
  // include/host_data.cuh -- class FREE_SURFACE, max_strain -yield_strain I, p_s = 0, n_s = 0)
  std::vector<int8_t> fcls;
  std::vector<double> sM, sp, sn;
};

// The FP64 reductions of sphnj_energy: 7 per-case constants and 256 blocks of 5 partials.
constexpr int kEnergyConsts = 7, kEnergyBlocks = 256;

// ---- the allocation-free entry point's scratch ------------------------------------------
// sphnj_energy_noalloc keeps its constants and partials here instead of allocating them: one
// array per GPU (every GPU the library runs on holds its own copy until the process ends),
// shared by every context on that GPU, so calls that overlap on one GPU race.  Host code
// never passes `scratch` itself to a kernel -- on the host the name is only a stand-in -- but
// the device address hipGetSymbolAddress returns for the current device, looked up once per
// device by scratch_address() when a context is initialised.
constexpr int kScratchDoubles = 8 + kEnergyBlocks * 5;  // 10,304 bytes
static __device__ __align__(256) double scratch[kScratchDoubles];

std::mutex g_scratch_mu;
std::map<int, double *> g_scratch_ptr;  // device -> address of its `scratch`

int scratch_address(double **out, std::string *err) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) {
    *err = "hipGetDevice failed";
    return SPHNJ_E_HIP;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  auto it = g_scratch_ptr.find(dev);
  if (it == g_scratch_ptr.end()) {
    double *p = nullptr;
    const hipError_t e = hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      *err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
      return SPHNJ_E_HIP;
    }
    it = g_scratch_ptr.emplace(dev, p).first;
  }
  *out = it->second;
  return SPHNJ_OK;
}

struct EngineBase {
  std::string err;
  hipStream_t stream = nullptr;
  bool use_graph = true;
  bool persistent_solid = false;  // S solid sub-steps as one launch (k_solid_substeps)
  int shift_runs = 1;             // shift passes walk the recorded candidate runs
  int blk_pair = 256, blk_shift = 256, blk_stream = 256;
  // set by enqueue_step around its cell stage: the row-table launch also runs the wall update
  // and the solid coupling (k_rows_wall)
  bool rows_wall = false;
  // the pair kernel records the shift passes' candidate runs (set around the fluid stage when
  // the pair and the shift share their tiles), and f_wc holds such runs for the current hash
  // and tiles: the step's first shift pass then walks them instead of recording
  bool pair_rec = false, pair_runs = false;
  // per-stage timing: HIP events recorded on the context stream around each stage while on
  bool timing = false;
  struct StageRec {
    int stage, launches;
    hipEvent_t a, b;
  };
  std::vector<StageRec> recs;
  std::vector<hipEvent_t> ev_pool;
  hipEvent_t take_event() {
    if (!ev_pool.empty()) {
      hipEvent_t e = ev_pool.back();
      ev_pool.pop_back();
      return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
  }
  void reset_times() {
    for (auto &r : recs) {
      ev_pool.push_back(r.a);
      ev_pool.push_back(r.b);
    }
    recs.clear();
  }
  virtual ~EngineBase() {
    reset_times();
    for (hipEvent_t e : ev_pool) (void)hipEventDestroy(e);
  }
  virtual int init(const std::vector<Derived> &D, const std::vector<Staged> &S) = 0;
  virtual int stage(int st) = 0;
  virtual int step(int n) = 0;
  virtual int download(int c, const sphnj_fluid_out *f, const sphnj_solid_out *s, const sphnj_wall_out *w) = 0;
  virtual int status(uint32_t *flags) = 0;
  virtual int counts(int c, int *nf, int *ns, int *nw, int *ncell) const = 0;
  virtual int sorted(int c, int species, int *keys, int *orig, int *start) = 0;
  virtual int neighbors(int c, int *counts, int *list, int64_t cap, int64_t *total) = 0;
  virtual int reflist(int c, int *off, int *list, int64_t cap) = 0;
  virtual int energy(int c, double out[6]) = 0;
  virtual int energy_in(int c, double out[6], double *consts, double *part) = 0;
  virtual int probe_global(int c, double xh, double out[SPHNJ_PROBE_COUNT]) = 0;
  virtual int sim_time(int c, double *t, double *dt) = 0;
  virtual int probe_solid(int c, double X, double Y, double pos[2], int *index) = 0;
  virtual int stage_time(int st, int *launches, double *ms) = 0;
  virtual int autotune() = 0;
  double tune_ms[2][3] = {{0, 0, 0}, {0, 0, 0}};  // [pair, shift][64, 128, 256]: ms per launch
  virtual void drop_graphs() = 0;
};

template <class Pol>
struct Engine : EngineBase {
  using R = typename Pol::R;
  Dev<Pol> d{};
  std::vector<CaseDev<R>> hc;
  std::vector<int> f_off, s_off, w_off;  // ncase + 1
  std::vector<Derived> Dh;
  int ncase = 0, total_cells = 0, S = 1, K = 0;
  // fluid tiles: global sub-columns G, row table size; per tile size 64 / 128 / 256: tile
  // counts per sub-column, their exclusive scan (+ total), plans, and whether the plans
  // match the current hash
  int G = 0;
  int64_t lb_total = 0, swp_total = 0;
  int *col_case = nullptr;
  int *tile_cnt[3] = {nullptr, nullptr, nullptr}, *tile_off[3] = {nullptr, nullptr, nullptr};
  int *tile_band[3] = {nullptr, nullptr, nullptr};
  TilePlan *plans[3] = {nullptr, nullptr, nullptr};
  bool plan_ok[3] = {false, false, false};
  static int bidx(int B) { return B == 64 ? 0 : B == 128 ? 1 : 2; }
  // tiles of B slots never span sub-columns: at most one partial tile per sub-column
  int max_tiles(int B) const { return kMaxPieces * (nblk(d.nf, B) + G); }
  int grid_tiles(int B) const { return (max_tiles(B) + kXcds - 1) / kXcds * kXcds; }
  bool viscid = false;
  // the persistent solid kernel: barrier words and the grid (0 = the device cannot hold the
  // solids' blocks resident at half its capacity: the sub-steps fall back to 2 S launches)
  GridSync gsync{};
  int solid_grid = 0;
  // probe workspace, allocated once: per-block partials and the point probe's index
  double *probe_part = nullptr;
  int *probe_best = nullptr;
  static constexpr int kSpinLimit = 1 << 22;  // polls of ~0.1 us: a stuck barrier gives up in < 1 s
  RadixSorter ssort;
  int *s_sorted_keys = nullptr;
  int *cnt_tsum = nullptr, *cnt_toff = nullptr;  // counting sort: tile sums and their scan
  std::vector<void *> allocs;
  struct GraphEntry {
    hipGraphExec_t exec;
    Dev<Pol> post;
  };
  std::map<std::array<const void *, 6>, GraphEntry> graphs;

  ~Engine() override {
    drop_graphs();
    for (void *p : allocs) (void)hipFree(p);
  }

  void drop_graphs() override {
    for (auto &kv : graphs) (void)hipGraphExecDestroy(kv.second.exec);
    graphs.clear();
  }

  template <class T>
  int dalloc(T **p, size_t n) {
    void *q = nullptr;
    HIPC(hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T)));
    allocs.push_back(q);
    *p = (T *)q;
    return SPHNJ_OK;
  }
  template <class T>
  int h2d(T *dst, const T *src, size_t n) {
    if (n) HIPC(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyHostToDevice, stream));
    return SPHNJ_OK;
  }
  template <class T>
  int d2h(T *dst, const T *src, size_t n) {
    if (n) HIPC(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyDeviceToHost, stream));
    return SPHNJ_OK;
  }

  int sorter_alloc(RadixSorter &s, int n, int max_key) {
    s.n = n;
    s.nblocks = nblk(n, kSortTile);
    s.passes = radix_passes_for(max_key);
    int rc;
    if ((rc = dalloc(&s.k1, n)) || (rc = dalloc(&s.v0, n)) || (rc = dalloc(&s.v1, n)) ||
        (rc = dalloc(&s.hist, (size_t)kRadix * std::max(s.nblocks, 1))) || (rc = dalloc(&s.dtot, kRadix)))
      return rc;
    return SPHNJ_OK;
  }

  void fill_case(const Derived &D, CaseDev<R> &c) {
    std::memset(&c, 0, sizeof(c));
    c.rc2 = (R)D.rc2;
    c.quarter_rc2 = (R)(0.25 * D.rc2);
    c.hinv2 = (R)D.hinv2;
    c.nhl2 = (R)(-D.hinv2 * 1.4426950408889634);
    c.kdc = (R)D.kdc;
    {
      const double sc = std::sqrt(D.hinv2 * 1.4426950408889634);
      c.gscale = (float)sc;
      c.rc2s = (float)(D.rc2 * D.hinv2 * 1.4426950408889634);
      c.kdc_s = (float)(D.kdc / sc);
    }
    c.k2c = (R)D.k2c;
    c.eR = (R)D.eR;
    c.h = (R)D.h;
    c.eps_h2 = (R)(0.01 * D.h2);
    for (int k = 0; k < 2; ++k) {
      c.c[k] = (R)D.c[k];
      c.cinv[k] = (R)(1.0 / D.c[k]);
      c.mu8[k] = D.viscid ? (R)(8.0 * D.mu[k]) : (R)0;
      c.rho0[k] = (R)D.rho0[k];
      c.rho0_d[k] = D.rho0[k];
      c.rho0_inv_d[k] = D.rho0_inv[k];
      c.gam_d[k] = D.gamma[k];
      c.ginv_d[k] = D.gamma_inv[k];
      c.rho2p_d[k] = D.rho2p[k];
      c.p2rho_d[k] = D.p2rho[k];
      c.rho_min_d[k] = D.rho_min[k];
    }
    c.p_min_d = D.p_min;
    c.p_bg_d = D.p_bg;
    c.p_min = (R)D.p_min;
    c.gx = (R)D.gx;
    c.gy = (R)D.gy;
    c.gx_d = D.gx;
    c.gy_d = D.gy;
    c.dt = D.dt;
    c.shift_ch = D.shift_coef * D.h;
    c.adaptive = D.cfl > 0.0 ? 1 : 0;
    c.cfl_h = D.cfl_h;
    c.c_max = D.c_max;
    c.dt_solid_cap = 0.0;  // set by init for cases with solids
    c.umax_cap = (R)D.umax_cap;
    c.cxmin = D.cxmin;
    c.cymin = D.cymin;
    c.cdx = D.cdx;
    c.cdy = D.cdy;
    c.inv_cdx = D.inv_cdx;
    c.inv_cdy = D.inv_cdy;
    c.cdx_r = (R)D.cdx;
    c.cdy_r = (R)D.cdy;
    c.rcut_prune = std::sqrt(D.rc2) * (1.0 + 1e-5);
    {
      const double g = (1.0 + kShiftSkin) * (1.0 + kShiftSkin), rc = std::sqrt(D.rc2);
      c.rl2_rec = Pol::MIXED ? (R)(D.rc2 * D.hinv2 * 1.4426950408889634 * g) : (R)(D.rc2 * g);
      c.rl2_pair = (R)(D.rc2 * g);
      c.skin_grow = (float)(kShiftSkin * rc);
      c.skin_ok = (float)(kShiftSkin * rc * (1.0 - 1e-3));
    }
    c.ssx = (float)(0.5 * D.cdx);
    c.ssy = (float)(0.5 * D.cdy);
    c.inv_ssx = (float)(2.0 * D.inv_cdx);
    c.inv_ssy = (float)(2.0 * D.inv_cdy);
    c.dxmin = D.dxmin;
    c.dymin = D.dymin;
    c.dxmax = D.dxmax;
    c.dymax = D.dymax;
    c.nx = D.nx;
    c.ny = D.ny;
    c.ncell = D.nx * D.ny;
    c.steady = D.steady;
    c.viscid = D.viscid;
    c.K = D.K;
    c.S = D.S;
    c.lambda = (R)D.lambda;
    c.mu_s = (R)D.mu_s;
    c.c_s = (R)D.c_s;
    c.alpha = (R)D.alpha;
    c.beta = (R)D.beta;
    c.rho_s_inv = (R)(1.0 / D.rho_s);
    c.rho_aux_inv = (R)(1.0 / D.rho_aux);
    c.fixed_u = D.fixed_u;
    c.fixed_v = D.fixed_v;
    c.aux_mod = (R)D.aux_modulus;
    c.yield_stress = (R)D.yield_stress;
    c.harden_strain = (R)D.harden_strain;
    c.harden_coef = (R)D.harden_coef;
    c.plastic = D.plastic;
  }

  // ---- init -----------------------------------------------------------------------------
  int init(const std::vector<Derived> &D, const std::vector<Staged> &St) override {
    int rc;
    Dh = D;
    ncase = (int)D.size();
    S = 1;
    K = 0;
    for (int c = 0; c < ncase; ++c) {  // launches cover the largest count; kernels skip per case
      K = std::max(K, D[c].K);
      if (!St[c].sx.empty()) S = std::max(S, D[c].S);
    }
    f_off.assign(ncase + 1, 0);
    s_off.assign(ncase + 1, 0);
    w_off.assign(ncase + 1, 0);
    hc.resize(ncase);
    total_cells = 0;
    int64_t cells64 = 0;
    for (int c = 0; c < ncase; ++c) cells64 += (int64_t)D[c].nx * D[c].ny;
    if (4 * cells64 >= (int64_t)0x7fffffff || ncase > 32767) {
      err = "cell grids too large for 32-bit sub-cell keys (4 x cells >= 2^31) or more than 32767 cases";
      return SPHNJ_E_UNSUPPORTED;
    }
    G = 0;
    lb_total = 0;
    swp_total = 0;
    std::vector<int> colc;
    for (int c = 0; c < ncase; ++c) {
      fill_case(D[c], hc[c]);
      if (!St[c].sx.empty()) hc[c].dt_solid_cap = D[c].dt_solid_cap;
      viscid = viscid || D[c].viscid;
      hc[c].cell_off = total_cells;
      total_cells += hc[c].ncell;
      hc[c].col_off = G;
      hc[c].lb_off = (int)lb_total;
      G += 2 * hc[c].nx;
      lb_total += (int64_t)(2 * hc[c].nx) * (2 * hc[c].ny + 1);
      hc[c].swp_off = (int)swp_total;
      swp_total += (int64_t)hc[c].nx * (hc[c].ny + 1);
      colc.insert(colc.end(), 2 * hc[c].nx, c);
      f_off[c + 1] = f_off[c] + (int)St[c].fx.size();
      s_off[c + 1] = s_off[c] + (int)St[c].sx.size();
      w_off[c + 1] = w_off[c] + (int)St[c].wx.size();
      hc[c].f_off = f_off[c];
      hc[c].f_n = f_off[c + 1] - f_off[c];
      hc[c].s_off = s_off[c];
      hc[c].s_n = s_off[c + 1] - s_off[c];
      hc[c].w_off = w_off[c];
      hc[c].w_n = w_off[c + 1] - w_off[c];
    }
    if (lb_total >= (int64_t)0x7fffffff) {
      err = "fluid row table too large for 32-bit indices";
      return SPHNJ_E_UNSUPPORTED;
    }
    const int nf = f_off[ncase], ns = s_off[ncase], nw = w_off[ncase], nc = total_cells;
    d.ncase = ncase;
    d.total_cells = nc;
    d.nf = nf;
    d.ns = ns;
    d.nw = nw;
    CaseDev<R> *cases;
    int *fco, *sco, *wco;
    if ((rc = dalloc(&cases, ncase)) || (rc = dalloc(&fco, ncase + 1)) || (rc = dalloc(&sco, ncase + 1)) ||
        (rc = dalloc(&wco, ncase + 1)))
      return rc;
    if ((rc = h2d(cases, hc.data(), ncase)) || (rc = h2d(fco, f_off.data(), ncase + 1)) ||
        (rc = h2d(sco, s_off.data(), ncase + 1)) || (rc = h2d(wco, w_off.data(), ncase + 1)))
      return rc;
    d.cases = cases;
    d.f_case_off = fco;
    d.s_case_off = sco;
    d.w_case_off = wco;
    {  // case of every 256th slot, per species (the binary search of case_of, on the host)
      auto table = [&](const std::vector<int> &off, int n, const int **out) -> int {
        std::vector<int> t(std::max(nblk(n, 256), 1), 0);
        for (size_t k = 0; k < t.size(); ++k) {
          const int i = (int)k * 256;
          int lo = 0, hi = ncase;
          while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (off[mid] <= i)
              lo = mid;
            else
              hi = mid;
          }
          t[k] = lo;
        }
        int *dev = nullptr, r;
        if ((r = dalloc(&dev, t.size())) || (r = h2d(dev, t.data(), t.size()))) return r;
        HIPC(hipStreamSynchronize(stream));  // t is freed on return
        *out = dev;
        return SPHNJ_OK;
      };
      if ((rc = table(f_off, nf, &d.f_case_blk)) || (rc = table(s_off, ns, &d.s_case_blk)) ||
          (rc = table(w_off, nw, &d.w_case_blk)))
        return rc;
    }

    // ---- fluid
    for (FluidBuf<Pol> *B : {&d.fa, &d.fb}) {
      if ((rc = dalloc(&B->x, nf)) || (rc = dalloc(&B->y, nf)) || (rc = dalloc(&B->u, nf)) ||
          (rc = dalloc(&B->v, nf)) || (rc = dalloc(&B->rho, nf)) || (rc = dalloc(&B->p, nf)) ||
          (rc = dalloc(&B->cum, nf)))
        return rc;
      if constexpr (Pol::MIXED) {
        if ((rc = dalloc(&B->vx, nf)) || (rc = dalloc(&B->vy, nf)) || (rc = dalloc(&B->vu, nf)) ||
            (rc = dalloc(&B->vv, nf)) || (rc = dalloc(&B->vrho, nf)))
          return rc;
      } else {
        B->vx = B->x;
        B->vy = B->y;
        B->vu = B->u;
        B->vv = B->v;
        B->vrho = B->rho;
      }
    }
    if ((rc = dalloc(&d.f_phase, nf)) || (rc = dalloc(&d.f_phase2, nf)) || (rc = dalloc(&d.f_cls, nf)) ||
        (rc = dalloc(&d.f_cls2, nf)) || (rc = dalloc(&d.f_orig, nf)) || (rc = dalloc(&d.f_cxy, nf)) ||
        (rc = dalloc(&d.f_hkey, nf)) || (rc = dalloc(&d.f_hrank, nf)) || (rc = dalloc(&d.f_perm, nf)) ||
        (rc = dalloc(&d.f_cnt, 4 * (size_t)nc + 1)) || (rc = dalloc(&d.f_porig, nf)) || (rc = dalloc(&d.f_pkey, nf)) ||
        (rc = dalloc(&cnt_tsum, (size_t)nblk(4 * nc, kScanTile) + 1)) ||
        (rc = dalloc(&cnt_toff, (size_t)nblk(4 * nc, kScanTile) + 1)) ||
        (rc = dalloc(&d.f_col, nf)) || (rc = dalloc(&d.f_near, nf)) || (rc = dalloc(&d.f_L, 4 * (size_t)nf)) ||
        (rc = dalloc(&d.f_sacc, 2 * (size_t)nf)) || (rc = dalloc(&d.f_wc, nf)) ||
        (rc = dalloc(&d.tile_d0, (size_t)max_tiles(64))) ||
        (rc = dalloc(&d.f_srange, 4 * (size_t)nc)) || (rc = dalloc(&d.umax2, ncase)) || (rc = dalloc(&d.umax, ncase)) ||
        (rc = dalloc(&d.umax_slot, (size_t)ncase * kSpread * kSlotStride)) || (rc = dalloc(&d.status, 1)) ||
        (rc = dalloc(&d.u2_last, ncase)) || (rc = dalloc(&d.dtv, ncase)) || (rc = dalloc(&d.tsim, ncase)) ||
        (rc = dalloc(&d.s_moved, 1)) ||
        (rc = dalloc(&d.f_lb, (size_t)lb_total)) || (rc = dalloc(&d.f_swp, (size_t)swp_total)) ||
        (rc = dalloc(&col_case, G)))
      return rc;
    if ((rc = h2d(col_case, colc.data(), G))) return rc;
    {  // step size state: the input speeds, dt, time 0
      std::vector<double> u2(ncase, 0.0), dt0(ncase), t0(ncase, 0.0);
      for (int c = 0; c < ncase; ++c) {
        for (size_t i = 0; i < St[c].fu.size(); ++i)
          u2[c] = std::max(u2[c], St[c].fu[i] * St[c].fu[i] + St[c].fv[i] * St[c].fv[i]);
        dt0[c] = D[c].dt;
      }
      if ((rc = h2d(d.u2_last, u2.data(), ncase)) || (rc = h2d(d.dtv, dt0.data(), ncase)) ||
          (rc = h2d(d.tsim, t0.data(), ncase)))
        return rc;
      HIPC(hipStreamSynchronize(stream));
    }
    for (int b = 0; b < 3; ++b) {
      const int B = 64 << b;
      if ((rc = dalloc(&tile_cnt[b], G)) || (rc = dalloc(&tile_off[b], G + 1)) || (rc = dalloc(&tile_band[b], G)) ||
          (rc = dalloc(&plans[b], max_tiles(B))))
        return rc;
    }
    {
      std::vector<double> x, y, u, v, rho;
      std::vector<R> p;
      std::vector<int8_t> ph;
      for (int c = 0; c < ncase; ++c) {
        const Staged &s = St[c];
        x.insert(x.end(), s.fx.begin(), s.fx.end());
        y.insert(y.end(), s.fy.begin(), s.fy.end());
        u.insert(u.end(), s.fu.begin(), s.fu.end());
        v.insert(v.end(), s.fv.begin(), s.fv.end());
        rho.insert(rho.end(), s.frho.begin(), s.frho.end());
        for (double q : s.fp) p.push_back((R)q);
        ph.insert(ph.end(), s.fph.begin(), s.fph.end());
      }
      std::vector<int> orig(nf);
      std::iota(orig.begin(), orig.end(), 0);
      std::vector<int8_t> cls, nearv(nf, 0);
      for (int c = 0; c < ncase; ++c) {
        const Staged &s = St[c];
        if (s.fcls.empty())
          cls.insert(cls.end(), s.fx.size(), (int8_t)SPHNJ_FREE_SURFACE);
        else
          cls.insert(cls.end(), s.fcls.begin(), s.fcls.end());
      }
      std::vector<R> zero4(4 * (size_t)nf, (R)0), zero1(nf, (R)0);
      if ((rc = h2d(d.fa.x, x.data(), nf)) || (rc = h2d(d.fa.y, y.data(), nf)) || (rc = h2d(d.fa.u, u.data(), nf)) ||
          (rc = h2d(d.fa.v, v.data(), nf)) || (rc = h2d(d.fa.rho, rho.data(), nf)) ||
          (rc = h2d(d.fa.p, p.data(), nf)) || (rc = h2d(d.f_phase, ph.data(), nf)) ||
          (rc = h2d(d.f_cls, cls.data(), nf)) || (rc = h2d(d.f_orig, orig.data(), nf)) ||
          (rc = h2d(d.f_near, nearv.data(), nf)) || (rc = h2d(d.f_col, zero1.data(), nf)) ||
          (rc = h2d(d.f_L, zero4.data(), 4 * (size_t)nf)))
        return rc;
      HIPC(hipStreamSynchronize(stream));
    }
    HIPC(hipMemsetAsync(d.f_cnt, 0, (4 * (size_t)nc + 1) * sizeof(int), stream));  // zero between steps
    HIPC(hipMemsetAsync(d.f_srange, 0, 4 * (size_t)nc * sizeof(int2), stream));

    // ---- solid
    if ((rc = dalloc(&d.sx, ns)) || (rc = dalloc(&d.sy, ns)) || (rc = dalloc(&d.su, ns)) || (rc = dalloc(&d.sv, ns)) ||
        (rc = dalloc(&d.sx2, ns)) || (rc = dalloc(&d.sy2, ns)) || (rc = dalloc(&d.su2, ns)) ||
        (rc = dalloc(&d.sv2, ns)) || (rc = dalloc(&d.sX, ns)) || (rc = dalloc(&d.sY, ns)) ||
        (rc = dalloc(&d.s_mat, ns)) || (rc = dalloc(&d.s_bc, ns)) || (rc = dalloc(&d.s_p, ns)) ||
        (rc = dalloc(&d.s_n, 2 * (size_t)ns)) || (rc = dalloc(&d.s_F, 4 * (size_t)ns)) ||
        (rc = dalloc(&d.s_L, 4 * (size_t)ns)) || (rc = dalloc(&d.s_P, 4 * (size_t)ns)) ||
        (rc = dalloc(&d.s_PL, 4 * (size_t)ns)) || (rc = dalloc(&d.s_M, 4 * (size_t)ns)) ||
        (rc = dalloc(&d.s_Lref, 4 * (size_t)ns)) || (rc = dalloc(&d.s_Bt, 4 * (size_t)ns)) ||
        (rc = dalloc(&d.s_degen, ns)) || (rc = dalloc(&d.s_a, 2 * (size_t)ns)) ||
        (rc = dalloc(&d.s_cp, 2 * (size_t)ns)) || (rc = dalloc(&d.s_key, ns)) || (rc = dalloc(&d.s_cxy, ns)) ||
        (rc = dalloc(&d.s_range, nc)) || (rc = dalloc(&d.s_ref_off, ns + 1)))
      return rc;
    HIPC(hipMemsetAsync(d.s_range, 0, (size_t)nc * sizeof(int2), stream));  // no solid: stays empty
    if ((rc = dalloc(&probe_part, (size_t)kProbeBlocks * kProbeParts)) || (rc = dalloc(&probe_best, 1))) return rc;
    if (ns > 0) {
      unsigned *words;
      if ((rc = dalloc(&words, 11 * kSyncLine))) return rc;
      HIPC(hipMemsetAsync(words, 0, 11 * kSyncLine * sizeof(unsigned), stream));
      gsync = GridSync{words};
      int dev = 0, cus = 0, per_cu = 0;
      HIPC(hipGetDevice(&dev));
      HIPC(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
      HIPC(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_solid_substeps<Pol>, kSubstepBlock, 0));
      const int want = nblk(ns * kSolidLanes, kSubstepBlock), fit = per_cu * cus / 2;
      solid_grid = fit >= 1 ? std::min(want, fit) : 0;
    }
    // the first cell stage sorts the solids: every key differs from -1, and the flag starts raised
    HIPC(hipMemsetAsync(d.s_key, 0xff, std::max(ns, 1) * sizeof(int), stream));
    HIPC(hipMemsetAsync(d.s_moved, 1, sizeof(int), stream));
    if constexpr (Pol::MIXED) {
      if ((rc = dalloc(&d.s_vx, ns)) || (rc = dalloc(&d.s_vy, ns)) || (rc = dalloc(&d.s_vx2, ns)) ||
          (rc = dalloc(&d.s_vy2, ns)))
        return rc;
    } else {
      d.s_vx = d.sx;
      d.s_vy = d.sy;
      d.s_vx2 = d.sx2;
      d.s_vy2 = d.sy2;
    }
    {
      std::vector<double> x, y, u, v, X, Y;
      std::vector<int8_t> mat, bc;
      std::vector<R> I4, Z4, M4, Z2, Z1;
      for (int c = 0; c < ncase; ++c) {
        const Staged &s = St[c];
        x.insert(x.end(), s.sx.begin(), s.sx.end());
        y.insert(y.end(), s.sy.begin(), s.sy.end());
        u.insert(u.end(), s.su.begin(), s.su.end());
        v.insert(v.end(), s.sv.begin(), s.sv.end());
        X.insert(X.end(), s.sX.begin(), s.sX.end());
        Y.insert(Y.end(), s.sY.begin(), s.sY.end());
        mat.insert(mat.end(), s.smat.begin(), s.smat.end());
        bc.insert(bc.end(), s.sbc.begin(), s.sbc.end());
        for (size_t i = 0; i < s.sx.size(); ++i) {  // include/host_data.cuh:77-83
          const R ys = (R)(-D[c].yield_strain);
          if (s.sM.empty())
            M4.insert(M4.end(), {ys, (R)0, (R)0, ys});
          else
            M4.insert(M4.end(), {(R)s.sM[4 * i], (R)s.sM[4 * i + 1], (R)s.sM[4 * i + 2], (R)s.sM[4 * i + 3]});
          Z1.push_back(s.sp.empty() ? (R)0 : (R)s.sp[i]);
          Z2.push_back(s.sn.empty() ? (R)0 : (R)s.sn[2 * i]);
          Z2.push_back(s.sn.empty() ? (R)0 : (R)s.sn[2 * i + 1]);
        }
      }
      for (int i = 0; i < ns; ++i) {
        I4.insert(I4.end(), {(R)1, (R)0, (R)0, (R)1});
        Z4.insert(Z4.end(), {(R)0, (R)0, (R)0, (R)0});
      }

      std::vector<double> Z2d(2 * (size_t)ns, 0.0);
      if ((rc = h2d(d.sx, x.data(), ns)) || (rc = h2d(d.sy, y.data(), ns)) || (rc = h2d(d.su, u.data(), ns)) ||
          (rc = h2d(d.sv, v.data(), ns)) || (rc = h2d(d.sX, X.data(), ns)) || (rc = h2d(d.sY, Y.data(), ns)) ||
          (rc = h2d(d.s_mat, mat.data(), ns)) || (rc = h2d(d.s_bc, bc.data(), ns)) ||
          (rc = h2d(d.s_F, I4.data(), 4 * (size_t)ns)) || (rc = h2d(d.s_L, I4.data(), 4 * (size_t)ns)) ||
          (rc = h2d(d.s_P, Z4.data(), 4 * (size_t)ns)) || (rc = h2d(d.s_PL, Z4.data(), 4 * (size_t)ns)) ||
          (rc = h2d(d.s_M, M4.data(), 4 * (size_t)ns)) || (rc = h2d(d.s_n, Z2.data(), 2 * (size_t)ns)) ||
          (rc = h2d(d.s_p, Z1.data(), ns)) || (rc = h2d(d.s_a, Z2d.data(), 2 * (size_t)ns)) ||
          (rc = h2d(d.s_cp, Z2.data(), 2 * (size_t)ns)))
        return rc;
      HIPC(hipStreamSynchronize(stream));
    }
    if ((rc = sorter_alloc(ssort, ns, std::max(nc - 1, 0)))) return rc;

    // ---- wall
    if ((rc = dalloc(&d.wx, nw)) || (rc = dalloc(&d.wy, nw)) || (rc = dalloc(&d.w_u, nw)) || (rc = dalloc(&d.w_v, nw)) ||
        (rc = dalloc(&d.w_p, nw)) || (rc = dalloc(&d.w_rho, nw)) || (rc = dalloc(&d.w_n, 2 * (size_t)nw)) ||
        (rc = dalloc(&d.w_phase, nw)) || (rc = dalloc(&d.w_bc, nw)) || (rc = dalloc(&d.w_orig, nw)) ||
        (rc = dalloc(&d.w_cxy, nw)) || (rc = dalloc(&d.w_range, nc)) || (rc = dalloc(&d.w_lb, (size_t)swp_total)))
      return rc;
    if constexpr (Pol::MIXED) {
      if ((rc = dalloc(&d.w_vx, nw)) || (rc = dalloc(&d.w_vy, nw))) return rc;
    } else {
      d.w_vx = d.wx;
      d.w_vy = d.wy;
    }
    HIPC(hipMemsetAsync(d.status, 0, sizeof(unsigned), stream));
    {
      std::vector<double> x, y;
      std::vector<int8_t> bc;
      std::vector<R> rho0, z;
      for (int c = 0; c < ncase; ++c) {
        x.insert(x.end(), St[c].wx.begin(), St[c].wx.end());
        y.insert(y.end(), St[c].wy.begin(), St[c].wy.end());
        bc.insert(bc.end(), St[c].wbc.begin(), St[c].wbc.end());
        rho0.insert(rho0.end(), St[c].wx.size(), (R)D[c].rho0[0]);  // include/host_data.cuh:111
      }
      z.assign(2 * (size_t)nw, (R)0);
      std::vector<int8_t> ph(nw, 0);
      double *x0, *y0;
      int8_t *bc0;
      int *keys;
      RadixSorter ws;
      if ((rc = dalloc(&x0, nw)) || (rc = dalloc(&y0, nw)) || (rc = dalloc(&bc0, nw)) || (rc = dalloc(&keys, nw)) ||
          (rc = sorter_alloc(ws, nw, std::max(nc - 1, 0))))
        return rc;
      if ((rc = h2d(x0, x.data(), nw)) || (rc = h2d(y0, y.data(), nw)) || (rc = h2d(bc0, bc.data(), nw)) ||
          (rc = h2d(d.w_rho, rho0.data(), nw)) || (rc = h2d(d.w_u, z.data(), nw)) || (rc = h2d(d.w_v, z.data(), nw)) ||
          (rc = h2d(d.w_p, z.data(), nw)) || (rc = h2d(d.w_n, z.data(), 2 * (size_t)nw)) ||
          (rc = h2d(d.w_phase, ph.data(), nw)))
        return rc;
      if (nw > 0) {
        k_hash_wall<Pol><<<nblk(nw, 256), 256, 0, stream>>>(d, x0, y0, keys);
        LAUNCH_CHECK();
        int *sk, *sv;
        HIPC(radix_sort_pairs(ws, keys, &sk, &sv, stream));
        k_permute_wall<Pol><<<nblk(nw, 256), 256, 0, stream>>>(d, sv, x0, y0, bc0);
        LAUNCH_CHECK();
        k_clear_ranges<<<nblk(nc, 256), 256, 0, stream>>>(d.w_range, nc, nullptr, 0);
        k_cell_bounds<<<nblk(nw, 256), 256, 0, stream>>>(sk, nw, d.w_range, nullptr);
        k_wall_normal<Pol><<<nblk(nw, 256), 256, 0, stream>>>(d);
        LAUNCH_CHECK();
      } else {
        k_clear_ranges<<<nblk(nc, 256), 256, 0, stream>>>(d.w_range, nc, nullptr, 0);
        LAUNCH_CHECK();
      }
      k_wall_rows<Pol><<<nblk(G, 256), 256, 0, stream>>>(d, G, col_case);
      LAUNCH_CHECK();
      HIPC(hipStreamSynchronize(stream));
    }

    // ---- solid reference-configuration lists (built from X, Y)
    {
      std::vector<int> off(ns + 1, 0);
      if (ns > 0) {
        int *keys, *cxy, *cnt;
        int2 *range;
        double *kdc, *hinv2, *rc2;
        RadixSorter rs;
        if ((rc = dalloc(&keys, ns)) || (rc = dalloc(&cxy, ns)) || (rc = dalloc(&cnt, ns)) ||
            (rc = dalloc(&range, nc)) || (rc = dalloc(&kdc, ncase)) || (rc = dalloc(&hinv2, ncase)) ||
            (rc = dalloc(&rc2, ncase)) || (rc = sorter_alloc(rs, ns, std::max(nc - 1, 0))))
          return rc;
        std::vector<double> hk, hh, hr;
        for (auto &q : D) {
          hk.push_back(q.kdc);
          hh.push_back(q.hinv2);
          hr.push_back(q.rc2);
        }
        if ((rc = h2d(kdc, hk.data(), ncase)) || (rc = h2d(hinv2, hh.data(), ncase)) || (rc = h2d(rc2, hr.data(), ncase)))
          return rc;
        k_hash_solid<Pol><<<nblk(ns, 256), 256, 0, stream>>>(d, d.sX, d.sY, keys, cxy, 0, nullptr);
        LAUNCH_CHECK();
        int *sk, *sv;
        HIPC(radix_sort_pairs(rs, keys, &sk, &sv, stream));
        k_clear_ranges<<<nblk(nc, 256), 256, 0, stream>>>(range, nc, nullptr, 0);
        k_cell_bounds<<<nblk(ns, 256), 256, 0, stream>>>(sk, ns, range, nullptr);
        k_ref_count<Pol><<<nblk(ns, 256), 256, 0, stream>>>(d, cxy, range, sv, rc2, cnt);
        LAUNCH_CHECK();
        std::vector<int> hcnt(ns);
        if ((rc = d2h(hcnt.data(), cnt, ns))) return rc;
        HIPC(hipStreamSynchronize(stream));
        for (int i = 0; i < ns; ++i) off[i + 1] = off[i] + hcnt[i];
        if ((rc = h2d(d.s_ref_off, off.data(), ns + 1))) return rc;
        if ((rc = dalloc(&d.s_ref_j, off[ns])) || (rc = dalloc(&d.s_ref_g, 2 * (size_t)off[ns]))) return rc;
        k_ref_fill<Pol><<<nblk(ns, 256), 256, 0, stream>>>(d, cxy, range, sv, kdc, hinv2, rc2);
        LAUNCH_CHECK();
        HIPC(hipStreamSynchronize(stream));
      } else {
        if ((rc = h2d(d.s_ref_off, off.data(), 1))) return rc;
        if ((rc = dalloc(&d.s_ref_j, 1)) || (rc = dalloc(&d.s_ref_g, 2))) return rc;
      }
    }
    // hash and order the fluid and solid once, so every stage finds valid cells
    if ((rc = stage_cells(false))) return rc;
    HIPC(hipStreamSynchronize(stream));
    return SPHNJ_OK;
  }

  // ---- stages -----------------------------------------------------------------------------
  void swap_fluid_all() {
    std::swap(d.fa, d.fb);
  }

  // new_step: a step begins (its size is set here: in the fluid hash, or on its own without
  // fluid)
  int stage_cells(bool new_step) {
    pair_runs = false;  // a new hash: new slots and tiles
    const int nf = d.nf, ns = d.ns, nc = d.total_cells;
    // every species' cell ranges start empty (the permute writes the fluid's, k_cell_bounds the
    // solids')
    // The solids' cell ranges start empty (k_cell_bounds writes the non-empty ones): the fluid
    // hash clears them; with no fluid, a clear launch does, with the fluid's ranges (otherwise
    // the counting sort writes every one of those).
    // With fluid and at most kSmallSortN solids, the solids are sorted again (and their ranges
    // cleared and rebuilt) only when a solid's cell changed: the solid hash runs first and
    // raises d.s_moved, which k_subcol_rows lowers at the end of the stage.
    const bool track = nf > 0 && ns > 0 && ns <= kSmallSortN;
    int *moved = track ? d.s_moved : nullptr;
    if (ns > 0) {
      k_hash_solid<Pol><<<nblk(ns, blk_stream), blk_stream, 0, stream>>>(d, d.sx, d.sy, d.s_key, d.s_cxy, 1, moved);
      LAUNCH_CHECK();
    }
    if (nf == 0) {
      if (new_step) k_step_dt<Pol><<<1, 64, 0, stream>>>(d);
      k_clear_ranges<<<nblk(4 * nc, 256), 256, 0, stream>>>(d.s_range, nc, d.f_srange, 4 * nc);
      LAUNCH_CHECK();
    }
    if (nf > 0) {  // counting sort on the sub-cell keys, then the permute
      const int nk = 4 * nc, nkb = nblk(nk, kScanTile);
      k_hash_fluid<Pol><<<nblk(nf, 256 * kCellRounds), 256, 0, stream>>>(d, moved, new_step ? 1 : 0);
      k_count_tiles<<<nkb, kSortThreads, 0, stream>>>(d.f_cnt, nk, cnt_tsum);
      if (nkb > kApplyPrefixMax) {
        k_exscan_small<<<1, 1024, 0, stream>>>(cnt_tsum, nkb, cnt_toff);
        k_count_apply<<<nkb, kSortThreads, 0, stream>>>(d.f_cnt, nk, cnt_tsum, cnt_toff, d.f_srange);
      } else {
        k_count_apply<<<nkb, kSortThreads, 0, stream>>>(d.f_cnt, nk, cnt_tsum, nullptr, d.f_srange);
      }
      k_count_scatter<Pol><<<nblk(nf, 256 * kCellRounds), 256, 0, stream>>>(d);
      LAUNCH_CHECK();
      k_permute_fluid<Pol><<<nblk(std::max(nf, ncase * kSpread), blk_stream), blk_stream, 0, stream>>>(d);
      LAUNCH_CHECK();
      swap_fluid_all();
      std::swap(d.f_phase, d.f_phase2);
      std::swap(d.f_cls, d.f_cls2);
    }
    if (ns > 0) {
      HIPC(radix_sort_pairs(ssort, d.s_key, &s_sorted_keys, &d.s_sorted, stream, moved));
      k_cell_bounds<<<nblk(ns, 256), 256, 0, stream>>>(s_sorted_keys, ns, d.s_range, moved);
    }
    plan_ok[0] = plan_ok[1] = plan_ok[2] = false;
    if (nf > 0) {  // row tables, the cell audit and the tile counts of both block sizes
      PlanJobs jobs{};
      const int nj = plan_jobs(blk_pair, blk_shift, jobs);
      if (G <= kPlanPrefixMax && rows_wall)
        launch_rows_wall(jobs, nj, moved);
      else if (G <= kPlanPrefixMax)
        k_subcol_rows<Pol, 4><<<G / 2, 256, 0, stream>>>(d, G, col_case, jobs, nj, moved);
      else
        k_subcol_rows<Pol, 1><<<G / 2, 64, 0, stream>>>(d, G, col_case, jobs, nj, moved);
      LAUNCH_CHECK();
      int rc;
      if ((rc = ensure_plans(blk_pair, blk_shift, true))) return rc;
    } else {
      k_cell_audit<Pol><<<nblk(nc, 256), 256, 0, stream>>>(d, d.status);
      LAUNCH_CHECK();
    }
    return SPHNJ_OK;
  }

  // Tile plans of size B for the current hash (the row table and tile counts are built by
  // the CELLS stage).
  // The planning jobs of block sizes B1 and B2 (0 = none) whose plans are not current.
  int plan_jobs(int B1, int B2, PlanJobs &jobs) const {
    int nj = 0;
    for (int B : {B1, B2}) {
      if (B == 0) continue;
      const int b = bidx(B);
      if (plan_ok[b] || (nj == 1 && jobs.j[0].B == B)) continue;
      jobs.j[nj++] = PlanJob{B, tile_cnt[b], tile_off[b], tile_band[b], plans[b]};
    }
    return nj;
  }
  // counted: k_subcol_rows has just written these jobs' tile counts
  int ensure_plans(int B1, int B2 = 0, bool counted = false) {
    if (d.nf == 0) return SPHNJ_OK;
    PlanJobs jobs{};
    const int nj = plan_jobs(B1, B2, jobs);
    if (nj == 0) return SPHNJ_OK;
    if (!counted) k_tile_count<Pol><<<dim3(nblk(G, 4), nj), 256, 0, stream>>>(d, G, col_case, jobs);
    const int scanned = G > kPlanPrefixMax;
    if (scanned) k_tile_scan<<<nj, 1024, 0, stream>>>(jobs, G);
    k_tile_plan<Pol><<<dim3(nblk(G, 4), nj), 256, 0, stream>>>(d, G, col_case, jobs, scanned);
    LAUNCH_CHECK();
    for (int k = 0; k < nj; ++k) plan_ok[bidx(jobs.j[k].B)] = true;
    return SPHNJ_OK;
  }

  template <int B>
  void pair_at() {
    const int b = bidx(B);
    const int *nt = tile_off[b] + G;
    const int rec = pair_rec ? 1 : 0;
    if (viscid)
      k_fluid_pair_update<Pol, true, B><<<grid_tiles(B), B, 0, stream>>>(d, plans[b], nt, rec);
    else
      k_fluid_pair_update<Pol, false, B><<<grid_tiles(B), B, 0, stream>>>(d, plans[b], nt, rec);
  }
  int launch_pair(int blk) {
    int rc;
    if ((rc = ensure_plans(blk))) return rc;
    if (blk == 64)
      pair_at<64>();
    else if (blk == 128)
      pair_at<128>();
    else
      pair_at<256>();
    LAUNCH_CHECK();
    return SPHNJ_OK;
  }
  // the step's first pass records the candidate runs of the later ones (kShiftRecord) unless
  // the pair kernel has recorded them (pair_runs: then every pass walks them); shift_runs 0
  // searches the full windows in every pass, 2 records but never walks the runs
  template <int B>
  void shift_at(int pass) {
    const int b = bidx(B);
    const int *nt = tile_off[b] + G;
    const dim3 g(grid_tiles(B));
    if (shift_runs == 0 || K == 1)
      k_fluid_shift<Pol, B, kShiftPlain><<<g, B, 0, stream>>>(d, plans[b], nt, pass, 0);
    else if (pass == 0 && !pair_runs)
      k_fluid_shift<Pol, B, kShiftRecord><<<g, B, 0, stream>>>(d, plans[b], nt, pass, 0);
    else
      k_fluid_shift<Pol, B, kShiftCached><<<g, B, 0, stream>>>(d, plans[b], nt, pass, shift_runs == 1);
  }
  int launch_shift(int blk, int pass) {
    int rc;
    if ((rc = ensure_plans(blk))) return rc;
    if (blk == 64)
      shift_at<64>(pass);
    else if (blk == 128)
      shift_at<128>(pass);
    else
      shift_at<256>(pass);
    LAUNCH_CHECK();
    return SPHNJ_OK;
  }

  // Autotuner: times the pair kernel and a step's K shift passes at each candidate block size
  // (= tile size in slots) on the current state with HIP events and keeps the fastest
  // (tune_ms holds ms per launch, per pass for the shift).  The pair kernel only writes the
  // alternate buffers and the per-step transients; the shift passes move the views and path
  // lengths (rebuilt by the next hash) and a case's last pass also updates u, v and x, y in
  // place, which are saved before and restored after.
  int autotune() override {
    if (d.nf == 0) return SPHNJ_OK;
    pair_runs = false;  // the shift family times its recording pass; its runs are not kept
    const int nf = d.nf;
    double *su = nullptr, *sv = nullptr, *sx_ = nullptr, *sy_ = nullptr;
    int8_t *scls = nullptr;
    int rc;
    if ((rc = dalloc(&su, nf)) || (rc = dalloc(&sv, nf)) || (rc = dalloc(&scls, nf)) || (rc = dalloc(&sx_, nf)) ||
        (rc = dalloc(&sy_, nf)))
      return rc;
    HIPC(hipMemcpyAsync(su, d.fa.u, nf * sizeof(double), hipMemcpyDeviceToDevice, stream));
    HIPC(hipMemcpyAsync(sv, d.fa.v, nf * sizeof(double), hipMemcpyDeviceToDevice, stream));
    HIPC(hipMemcpyAsync(sx_, d.fa.x, nf * sizeof(double), hipMemcpyDeviceToDevice, stream));
    HIPC(hipMemcpyAsync(sy_, d.fa.y, nf * sizeof(double), hipMemcpyDeviceToDevice, stream));
    HIPC(hipMemcpyAsync(scls, d.f_cls, nf, hipMemcpyDeviceToDevice, stream));
    hipEvent_t a = take_event(), b = take_event();
    if (!a || !b) {
      err = "hipEventCreate";
      return SPHNJ_E_HIP;
    }
    const int cand[3] = {64, 128, 256};
    // the shift family runs a step's K passes as the stage does (the first records, the later
    // ones walk the runs), buffer swaps included; their parity is restored below
    int swaps = 0;
    auto run = [&](int fam, int bs) -> int {
      if (fam == 0) return launch_pair(bs);
      for (int k = 0; k < K; ++k) {
        const int r = launch_shift(bs, k);
        if (r) return r;
        swap_shift_buffers();
        ++swaps;
      }
      return SPHNJ_OK;
    };
    for (int fam = 0; fam < 2; ++fam) {
      if (fam == 1 && K == 0) break;
      double best = 1e300;
      int best_b = 256;
      for (int bs : cand) {
        const int warm = fam == 0 ? 2 : 1, reps = fam == 0 ? 5 : 2;
        for (int w = 0; w < warm; ++w)
          if ((rc = run(fam, bs))) return rc;
        HIPC(hipEventRecord(a, stream));
        for (int r = 0; r < reps; ++r)
          if ((rc = run(fam, bs))) return rc;
        HIPC(hipEventRecord(b, stream));
        HIPC(hipEventSynchronize(b));
        float ms = 0.f;
        HIPC(hipEventElapsedTime(&ms, a, b));
        tune_ms[fam][bs == 64 ? 0 : bs == 128 ? 1 : 2] = ms / reps / (fam == 0 ? 1 : K);
        if (ms < best) {
          best = ms;
          best_b = bs;
        }
      }
      if (fam == 0)
        blk_pair = best_b;
      else
        blk_shift = best_b;
    }
    ev_pool.push_back(a);
    ev_pool.push_back(b);
    if (swaps & 1) swap_shift_buffers();
    HIPC(hipMemcpyAsync(d.fa.u, su, nf * sizeof(double), hipMemcpyDeviceToDevice, stream));
    HIPC(hipMemcpyAsync(d.fa.v, sv, nf * sizeof(double), hipMemcpyDeviceToDevice, stream));
    HIPC(hipMemcpyAsync(d.fa.x, sx_, nf * sizeof(double), hipMemcpyDeviceToDevice, stream));
    HIPC(hipMemcpyAsync(d.fa.y, sy_, nf * sizeof(double), hipMemcpyDeviceToDevice, stream));
    HIPC(hipMemcpyAsync(d.f_cls, scls, nf, hipMemcpyDeviceToDevice, stream));
    HIPC(hipStreamSynchronize(stream));
    for (void *q : {(void *)su, (void *)sv, (void *)scls, (void *)sx_, (void *)sy_}) {
      (void)hipFree(q);
      allocs.erase(std::remove(allocs.begin(), allocs.end(), q), allocs.end());
    }
    pair_runs = false;
    drop_graphs();
    return SPHNJ_OK;
  }

  int stage(int st) override {
    const int nf = d.nf, ns = d.ns, nw = d.nw;
    switch (st) {
      case SPHNJ_STAGE_CELLS:  // a step begins: its size first
        return stage_cells(true);
      case SPHNJ_STAGE_COUPLE:
        if (nf > 0 && ns > 0) {
          k_solid_couple<Pol><<<nblk(ns * kSolidLanes, 256), 256, 0, stream>>>(d);
          LAUNCH_CHECK();
        }
        return SPHNJ_OK;
      case SPHNJ_STAGE_WALL:
        if (nw > 0) {
          launch_wall(false);
          LAUNCH_CHECK();
        }
        return SPHNJ_OK;
      case SPHNJ_STAGE_FLUID:
        if (nf > 0) {
          // the pair kernel records the shift's runs when both use the same tiles
          pair_rec = shift_runs == 1 && K > 1 && blk_pair == blk_shift;
          const int rc = launch_pair(blk_pair);
          pair_runs = pair_rec;
          pair_rec = false;
          if (rc) return rc;
          k_umax_final<Pol><<<ncase, 64, 0, stream>>>(d);
          LAUNCH_CHECK();
          swap_fluid_all();
          std::swap(d.f_cls, d.f_cls2);  // the class of this step
        }
        return SPHNJ_OK;
      case SPHNJ_STAGE_SOLID:
        if (ns > 0) {
          const bool one = persistent_solid && solid_grid > 0;
          if (one) {
            Dev<Pol> d1 = d;
            std::swap(d1.sx, d1.sx2);
            std::swap(d1.sy, d1.sy2);
            std::swap(d1.su, d1.su2);
            std::swap(d1.sv, d1.sv2);
            std::swap(d1.s_vx, d1.s_vx2);
            std::swap(d1.s_vy, d1.s_vy2);
            k_solid_substeps<Pol><<<solid_grid, kSubstepBlock, 0, stream>>>(d, d1, S, gsync, kSpinLimit);
            LAUNCH_CHECK();
          }
          for (int k = 0; k < S; ++k) {
            if (!one) {
              k_solid_deform<Pol><<<nblk(ns * kSolidLanes, 256), 256, 0, stream>>>(d, k);
              k_solid_force_update<Pol><<<nblk(ns * kSolidLanes, 256), 256, 0, stream>>>(d, k);
              LAUNCH_CHECK();
            }
            std::swap(d.sx, d.sx2);
            std::swap(d.sy, d.sy2);
            std::swap(d.su, d.su2);
            std::swap(d.sv, d.sv2);
            std::swap(d.s_vx, d.s_vx2);
            std::swap(d.s_vy, d.s_vy2);
          }
        }
        return SPHNJ_OK;
      case SPHNJ_STAGE_SHIFT:
        if (nf > 0) {
          for (int k = 0; k < K; ++k) {
            const int rc = launch_shift(blk_shift, k);
            if (rc) return rc;
            swap_shift_buffers();
          }
        }
        return SPHNJ_OK;
    }
    err = "unknown stage";
    return SPHNJ_E_ARG;
  }

  // after a shift pass: views and path lengths ping-pong (P64 views are the positions)
  void swap_shift_buffers() {
    std::swap(d.fa.vx, d.fb.vx);
    std::swap(d.fa.vy, d.fb.vy);
    std::swap(d.fa.cum, d.fb.cum);
    if (!Pol::MIXED) {
      std::swap(d.fa.x, d.fb.x);
      std::swap(d.fa.y, d.fb.y);
    }
  }

  int launches_of(int st) const {
    if (st == SPHNJ_STAGE_SHIFT) return d.nf > 0 ? K : 0;
    if (st == SPHNJ_STAGE_FLUID) return d.nf > 0 ? 2 : 0;
    if (st == SPHNJ_STAGE_SOLID) return d.ns > 0 ? ((persistent_solid && solid_grid > 0) ? 1 : 2 * S) : 0;
    return 1;
  }

  // The wall update (lanes per wall particle: about half a million threads in all, 1 to 8 per
  // particle); with_couple puts the solid coupling's blocks after the walls' in the same launch.
  template <int WL>
  void launch_wall_at(bool with_couple) {
    const int wb = nblk(d.nw * WL, 256);
    if (with_couple)
      k_couple_wall<Pol, WL><<<wb + nblk(d.ns * kSolidLanes, 256), 256, 0, stream>>>(d, wb);
    else
      k_wall_update<Pol, WL><<<wb, 256, 0, stream>>>(d);
  }
  // The row tables with the wall update and the solid coupling (k_rows_wall).
  template <int WL>
  void launch_rows_wall_at(const PlanJobs &jobs, int nj, int *moved) {
    const int wb = d.nw > 0 ? nblk(d.nw * WL, 256) : 0, cb = d.ns > 0 ? nblk(d.ns * kSolidLanes, 256) : 0;
    k_rows_wall<Pol, WL><<<G / 2 + wb + cb, 256, 0, stream>>>(d, G, col_case, jobs, nj, moved, G / 2, wb);
  }
  void launch_rows_wall(const PlanJobs &jobs, int nj, int *moved) {
    if (d.nw <= (1 << 16))
      launch_rows_wall_at<8>(jobs, nj, moved);
    else if (d.nw <= (1 << 17))
      launch_rows_wall_at<4>(jobs, nj, moved);
    else if (d.nw <= (1 << 18))
      launch_rows_wall_at<2>(jobs, nj, moved);
    else
      launch_rows_wall_at<1>(jobs, nj, moved);
  }
  void launch_wall(bool with_couple) {
    if (d.nw <= (1 << 16))
      launch_wall_at<8>(with_couple);
    else if (d.nw <= (1 << 17))
      launch_wall_at<4>(with_couple);
    else if (d.nw <= (1 << 18))
      launch_wall_at<2>(with_couple);
    else
      launch_wall_at<1>(with_couple);
  }

  int enqueue_step() {
    // untimed steps run the wall update and the solid coupling in the cell stage's launch of the
    // row tables where that launch gives a cell column 4 waves, else the coupling and the wall
    // update as one launch (both read the fluid only)
    const bool rw = !timing && d.nf > 0 && (d.nw > 0 || d.ns > 0) && G <= kPlanPrefixMax;
    const bool fuse = !rw && !timing && d.nf > 0 && d.ns > 0 && d.nw > 0;
    for (int st = SPHNJ_STAGE_CELLS; st <= SPHNJ_STAGE_SHIFT; ++st) {
      if (rw && (st == SPHNJ_STAGE_COUPLE || st == SPHNJ_STAGE_WALL)) continue;
      if (fuse && st == SPHNJ_STAGE_COUPLE) {
        launch_wall(true);
        LAUNCH_CHECK();
        continue;
      }
      if (fuse && st == SPHNJ_STAGE_WALL) continue;
      hipEvent_t a = nullptr, b = nullptr;
      if (timing) {
        a = take_event();
        b = take_event();
        if (!a || !b) {
          err = "hipEventCreate";
          return SPHNJ_E_HIP;
        }
        HIPC(hipEventRecord(a, stream));
      }
      rows_wall = rw && st == SPHNJ_STAGE_CELLS;
      const int rc = stage(st);
      rows_wall = false;
      if (rc) return rc;
      if (timing) {
        HIPC(hipEventRecord(b, stream));
        recs.push_back({st, launches_of(st), a, b});
      }
    }
    return SPHNJ_OK;
  }

  int stage_time(int st, int *launches, double *ms) override {
    HIPC(hipStreamSynchronize(stream));
    int n = 0;
    double t = 0.0;
    for (auto &r : recs) {
      if (r.stage != st) continue;
      float f = 0.f;
      HIPC(hipEventElapsedTime(&f, r.a, r.b));
      t += f;
      n += r.launches;
    }
    *launches = n;
    *ms = t;
    return SPHNJ_OK;
  }

  std::array<const void *, 6> state_key() const {
    return {d.fa.x, d.fa.u, d.f_phase, d.sx, d.fa.vx, d.s_vx};
  }

  int step(int n) override {
    for (int s = 0; s < n; ++s) {
      if (!use_graph || timing) {
        const int rc = enqueue_step();
        if (rc) return rc;
        continue;
      }
      const auto key = state_key();
      auto it = graphs.find(key);
      if (it == graphs.end()) {
        hipGraph_t g;
        HIPC(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        const int rc = enqueue_step();
        hipGraph_t gtmp = nullptr;
        const hipError_t ec = hipStreamEndCapture(stream, &gtmp);
        if (rc) {
          if (gtmp) (void)hipGraphDestroy(gtmp);
          return rc;
        }
        if (ec != hipSuccess) {
          err = std::string("hipStreamEndCapture: ") + hipGetErrorString(ec);
          return SPHNJ_E_HIP;
        }
        g = gtmp;
        hipGraphExec_t ex;
        const hipError_t ei = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ei != hipSuccess) {
          err = std::string("hipGraphInstantiate: ") + hipGetErrorString(ei);
          return SPHNJ_E_HIP;
        }
        it = graphs.emplace(key, GraphEntry{ex, d}).first;
      }
      HIPC(hipGraphLaunch(it->second.exec, stream));
      d = it->second.post;
    }
    return SPHNJ_OK;
  }

  // ---- inspection -----------------------------------------------------------------------
  int status(uint32_t *flags) override {
    unsigned h = 0;
    int rc;
    if ((rc = d2h(&h, d.status, 1))) return rc;
    HIPC(hipStreamSynchronize(stream));
    *flags = h;
    return SPHNJ_OK;
  }

  int counts(int c, int *nf, int *ns, int *nw, int *ncell) const override {
    if (nf) *nf = f_off[c + 1] - f_off[c];
    if (ns) *ns = s_off[c + 1] - s_off[c];
    if (nw) *nw = w_off[c + 1] - w_off[c];
    if (ncell) *ncell = hc[c].ncell;
    return SPHNJ_OK;
  }

  template <class T>
  static void put_scatter(T *dst, const std::vector<T> &src, const std::vector<int> &orig, int lo, int n, int w) {
    if (!dst) return;
    for (int t = 0; t < n; ++t) {
      const int o = orig[t] - lo;
      for (int k = 0; k < w; ++k) dst[(size_t)w * o + k] = src[(size_t)w * t + k];
    }
  }
  template <class T>
  static void put_conv(double *dst, const std::vector<T> &src, const std::vector<int> &orig, int lo, int n, int w) {
    if (!dst) return;
    for (int t = 0; t < n; ++t) {
      const int o = orig[t] - lo;
      for (int k = 0; k < w; ++k) dst[(size_t)w * o + k] = (double)src[(size_t)w * t + k];
    }
  }
  template <class T>
  static void put_int(int32_t *dst, const std::vector<T> &src, const std::vector<int> &orig, int lo, int n) {
    if (!dst) return;
    for (int t = 0; t < n; ++t) dst[orig[t] - lo] = (int32_t)src[t];
  }

  int download(int c, const sphnj_fluid_out *f, const sphnj_solid_out *s, const sphnj_wall_out *w) override {
    int rc;
    if (f) {
      const int lo = f_off[c], n = f_off[c + 1] - lo;
      std::vector<double> x(n), y(n), u(n), v(n), rho(n);
      std::vector<R> p(n), col(n), L(4 * (size_t)n);
      std::vector<int8_t> ph(n), cls(n), nr(n);
      std::vector<int> orig(n);
      if ((rc = d2h(x.data(), d.fa.x + lo, n)) || (rc = d2h(y.data(), d.fa.y + lo, n)) ||
          (rc = d2h(u.data(), d.fa.u + lo, n)) || (rc = d2h(v.data(), d.fa.v + lo, n)) ||
          (rc = d2h(rho.data(), d.fa.rho + lo, n)) || (rc = d2h(p.data(), d.fa.p + lo, n)) ||
          (rc = d2h(col.data(), d.f_col + lo, n)) || (rc = d2h(L.data(), d.f_L + 4 * (size_t)lo, 4 * (size_t)n)) ||
          (rc = d2h(ph.data(), d.f_phase + lo, n)) || (rc = d2h(cls.data(), d.f_cls + lo, n)) ||
          (rc = d2h(nr.data(), d.f_near + lo, n)) || (rc = d2h(orig.data(), d.f_orig + lo, n)))
        return rc;
      HIPC(hipStreamSynchronize(stream));
      put_scatter(f->x, x, orig, lo, n, 1);
      put_scatter(f->y, y, orig, lo, n, 1);
      put_scatter(f->u, u, orig, lo, n, 1);
      put_scatter(f->v, v, orig, lo, n, 1);
      put_scatter(f->rho, rho, orig, lo, n, 1);
      put_conv(f->p, p, orig, lo, n, 1);
      put_conv(f->color, col, orig, lo, n, 1);
      put_conv(f->du_dx, L, orig, lo, n, 4);
      put_int(f->phase, ph, orig, lo, n);
      put_int(f->boundary_type, cls, orig, lo, n);
      put_int(f->near, nr, orig, lo, n);
    }
    if (s) {
      const int lo = s_off[c], n = s_off[c + 1] - lo;
      std::vector<int> orig(n);
      std::iota(orig.begin(), orig.end(), lo);
      std::vector<double> x(n), y(n), u(n), v(n), X(n), Y(n), a(2 * (size_t)n);
      std::vector<R> p(n), nn(2 * (size_t)n), F(4 * (size_t)n), L(4 * (size_t)n), P(4 * (size_t)n),
          M(4 * (size_t)n), cp(2 * (size_t)n);
      if ((rc = d2h(x.data(), d.sx + lo, n)) || (rc = d2h(y.data(), d.sy + lo, n)) || (rc = d2h(u.data(), d.su + lo, n)) ||
          (rc = d2h(v.data(), d.sv + lo, n)) || (rc = d2h(X.data(), d.sX + lo, n)) || (rc = d2h(Y.data(), d.sY + lo, n)) ||
          (rc = d2h(a.data(), d.s_a + 2 * (size_t)lo, 2 * (size_t)n)) || (rc = d2h(p.data(), d.s_p + lo, n)) ||
          (rc = d2h(nn.data(), d.s_n + 2 * (size_t)lo, 2 * (size_t)n)) ||
          (rc = d2h(F.data(), d.s_F + 4 * (size_t)lo, 4 * (size_t)n)) ||
          (rc = d2h(L.data(), d.s_L + 4 * (size_t)lo, 4 * (size_t)n)) ||
          (rc = d2h(P.data(), d.s_P + 4 * (size_t)lo, 4 * (size_t)n)) ||
          (rc = d2h(M.data(), d.s_M + 4 * (size_t)lo, 4 * (size_t)n)) ||
          (rc = d2h(cp.data(), d.s_cp + 2 * (size_t)lo, 2 * (size_t)n)))
        return rc;
      HIPC(hipStreamSynchronize(stream));
      put_scatter(s->x, x, orig, lo, n, 1);
      put_scatter(s->y, y, orig, lo, n, 1);
      put_scatter(s->u, u, orig, lo, n, 1);
      put_scatter(s->v, v, orig, lo, n, 1);
      put_scatter(s->X, X, orig, lo, n, 1);
      put_scatter(s->Y, Y, orig, lo, n, 1);
      put_scatter(s->du_dt, a, orig, lo, n, 2);
      put_conv(s->pressure, p, orig, lo, n, 1);
      put_conv(s->normal, nn, orig, lo, n, 2);
      put_conv(s->deformation, F, orig, lo, n, 4);
      put_conv(s->correction, L, orig, lo, n, 4);
      put_conv(s->stress, P, orig, lo, n, 4);
      put_conv(s->max_strain, M, orig, lo, n, 4);
      put_conv(s->couple_dudt, cp, orig, lo, n, 2);
    }
    if (w) {
      const int lo = w_off[c], n = w_off[c + 1] - lo;
      std::vector<double> x(n), y(n);
      std::vector<R> u(n), v(n), p(n), rho(n), nn(2 * (size_t)n);
      std::vector<int8_t> ph(n), bc(n);
      std::vector<int> orig(n);
      if ((rc = d2h(x.data(), d.wx + lo, n)) || (rc = d2h(y.data(), d.wy + lo, n)) || (rc = d2h(u.data(), d.w_u + lo, n)) ||
          (rc = d2h(v.data(), d.w_v + lo, n)) || (rc = d2h(p.data(), d.w_p + lo, n)) ||
          (rc = d2h(rho.data(), d.w_rho + lo, n)) || (rc = d2h(nn.data(), d.w_n + 2 * (size_t)lo, 2 * (size_t)n)) ||
          (rc = d2h(ph.data(), d.w_phase + lo, n)) || (rc = d2h(bc.data(), d.w_bc + lo, n)) ||
          (rc = d2h(orig.data(), d.w_orig + lo, n)))
        return rc;
      HIPC(hipStreamSynchronize(stream));
      put_scatter(w->x, x, orig, lo, n, 1);
      put_scatter(w->y, y, orig, lo, n, 1);
      put_conv(w->u, u, orig, lo, n, 1);
      put_conv(w->v, v, orig, lo, n, 1);
      put_conv(w->p, p, orig, lo, n, 1);
      put_conv(w->rho, rho, orig, lo, n, 1);
      put_conv(w->normal, nn, orig, lo, n, 2);
      put_int(w->phase, ph, orig, lo, n);
      put_int(w->bc, bc, orig, lo, n);
    }
    return SPHNJ_OK;
  }

  int sorted(int c, int species, int *keys, int *orig, int *start) override {
    int rc;
    const auto &cp = hc[c];
    int lo, n;
    std::vector<int> k, o;
    // fluid: one start per sub-cell (4 per cell); solid and wall: one per cell
    const int nb = species == SPHNJ_FLUID ? 4 * cp.ncell : cp.ncell;
    std::vector<int2> rg(nb);
    if (species == SPHNJ_FLUID) {
      lo = f_off[c];
      n = f_off[c + 1] - lo;
      k.resize(n);
      o.resize(n);
      if (n && (rc = d2h(k.data(), d.f_pkey + lo, n))) return rc;  // the scatter's keys are the sorted keys
      if ((rc = d2h(o.data(), d.f_orig + lo, n)) || (rc = d2h(rg.data(), d.f_srange + 4 * (size_t)cp.cell_off, nb)))
        return rc;
    } else if (species == SPHNJ_SOLID) {
      lo = s_off[c];
      n = s_off[c + 1] - lo;
      k.resize(n);
      o.resize(n);
      if (n && s_sorted_keys && ((rc = d2h(k.data(), s_sorted_keys + lo, n)) || (rc = d2h(o.data(), d.s_sorted + lo, n))))
        return rc;
      if ((rc = d2h(rg.data(), d.s_range + cp.cell_off, cp.ncell))) return rc;
    } else {
      lo = w_off[c];
      n = w_off[c + 1] - lo;
      std::vector<int> cxy(n);
      o.resize(n);
      k.resize(n);
      if ((rc = d2h(cxy.data(), d.w_cxy + lo, n)) || (rc = d2h(o.data(), d.w_orig + lo, n)) ||
          (rc = d2h(rg.data(), d.w_range + cp.cell_off, cp.ncell)))
        return rc;
      HIPC(hipStreamSynchronize(stream));
      for (int t = 0; t < n; ++t) k[t] = cp.cell_off + (int)((uint32_t)cxy[t] >> 16) * cp.ny + (cxy[t] & 0xffff);
    }
    HIPC(hipStreamSynchronize(stream));
    for (int t = 0; t < n; ++t) {
      if (keys) keys[t] = k[t];
      if (orig) orig[t] = o[t] - lo;
    }
    if (start) {
      int next = n;
      start[nb] = n;
      for (int q = nb - 1; q >= 0; --q) {
        if (rg[q].y > rg[q].x) next = rg[q].x - lo;
        start[q] = next;
      }
    }
    return SPHNJ_OK;
  }

  int neighbors(int c, int *counts_out, int *list, int64_t cap, int64_t *total) override {
    int rc;
    const int lo = f_off[c], n = f_off[c + 1] - lo;
    if (n == 0) {
      if (total) *total = 0;
      return SPHNJ_OK;
    }
    int *dcnt, *doff, *dlist;
    std::vector<void *> tmp;
    auto talloc = [&](int **p, size_t m) -> int {
      void *q = nullptr;
      HIPC(hipMalloc(&q, std::max<size_t>(m, 1) * sizeof(int)));
      tmp.push_back(q);
      *p = (int *)q;
      return SPHNJ_OK;
    };
    auto cleanup = [&]() {
      for (void *q : tmp) (void)hipFree(q);
    };
    if ((rc = talloc(&dcnt, n)) || (rc = talloc(&doff, n))) {
      cleanup();
      return rc;
    }
    k_neighbors<Pol><<<nblk(n, 256), 256, 0, stream>>>(d, lo, lo + n, nullptr, dcnt, nullptr, nullptr);
    std::vector<int> cnt(n), orig(n);
    if ((rc = d2h(cnt.data(), dcnt, n)) || (rc = d2h(orig.data(), d.f_orig + lo, n))) {
      cleanup();
      return rc;
    }
    hipError_t e = hipStreamSynchronize(stream);
    if (e != hipSuccess) {
      cleanup();
      err = hipGetErrorString(e);
      return SPHNJ_E_HIP;
    }
    // CSR in original order
    std::vector<int64_t> ooff(n + 1, 0);
    std::vector<int> cnt_o(n);
    for (int t = 0; t < n; ++t) cnt_o[orig[t] - lo] = cnt[t];
    for (int i = 0; i < n; ++i) ooff[i + 1] = ooff[i] + cnt_o[i];
    const int64_t tot = ooff[n];
    if (total) *total = tot;
    if (counts_out) std::copy(cnt_o.begin(), cnt_o.end(), counts_out);
    if (list && tot > 0) {
      if (tot > 0x7fffffff) {
        cleanup();
        err = "neighbour list too large";
        return SPHNJ_E_ARG;
      }
      std::vector<int> soff(n);
      for (int t = 0; t < n; ++t) soff[t] = (int)ooff[orig[t] - lo];
      if ((rc = talloc(&dlist, (size_t)tot)) || (rc = h2d(doff, soff.data(), n))) {
        cleanup();
        return rc;
      }
      k_neighbors<Pol><<<nblk(n, 256), 256, 0, stream>>>(d, lo, lo + n, doff, nullptr, dlist, dcnt);
      std::vector<int> hl((size_t)tot);
      if ((rc = d2h(hl.data(), dlist, (size_t)tot))) {
        cleanup();
        return rc;
      }
      e = hipStreamSynchronize(stream);
      if (e != hipSuccess) {
        cleanup();
        err = hipGetErrorString(e);
        return SPHNJ_E_HIP;
      }
      std::copy(hl.begin(), hl.begin() + std::min<int64_t>(cap, tot), list);
    }
    cleanup();
    return SPHNJ_OK;
  }

  int reflist(int c, int *off, int *list, int64_t cap) override {
    int rc;
    const int lo = s_off[c], n = s_off[c + 1] - lo;
    std::vector<int> o(n + 1);
    if ((rc = d2h(o.data(), d.s_ref_off + lo, n + 1))) return rc;
    HIPC(hipStreamSynchronize(stream));
    const int base = o[0], tot = o[n] - base;
    if (off)
      for (int i = 0; i <= n; ++i) off[i] = o[i] - base;
    if (list && tot > 0) {
      std::vector<int> l(tot);
      if ((rc = d2h(l.data(), d.s_ref_j + base, tot))) return rc;
      HIPC(hipStreamSynchronize(stream));
      for (int e = 0; e < tot && e < cap; ++e) list[e] = l[e] - lo;
    }
    return SPHNJ_OK;
  }

  int sim_time(int c, double *t, double *dt) override {
    double v[2];
    int rc;
    if ((rc = d2h(v, d.tsim + c, 1)) || (rc = d2h(v + 1, d.dtv + c, 1))) return rc;
    HIPC(hipStreamSynchronize(stream));
    if (t) *t = v[0];
    if (dt) *dt = v[1];
    return SPHNJ_OK;
  }

  int probe_global(int c, double xh, double out[SPHNJ_PROBE_COUNT]) override {
    const Derived &D = Dh[c];
    const ProbeArgs a{f_off[c], f_off[c + 1], s_off[c], s_off[c + 1], D.d * D.d, D.rho_s, D.rho_aux, xh, 0.5 * D.d};
    std::vector<double> hp((size_t)kProbeBlocks * kProbeParts);
    k_probe_partial<Pol><<<kProbeBlocks, 256, 0, stream>>>(d, a, probe_part);
    LAUNCH_CHECK();
    int rc;
    if ((rc = d2h(hp.data(), probe_part, hp.size()))) return rc;
    HIPC(hipStreamSynchronize(stream));
    double v[kProbeParts] = {0.0, 0.0, -INFINITY, -INFINITY, 0.0, 0.0, 0.0};
    for (int b = 0; b < kProbeBlocks; ++b) {  // block order: the same sums every call
      const double *q = hp.data() + (size_t)b * kProbeParts;
      v[0] += q[0];
      v[1] += q[1];
      v[2] = std::max(v[2], q[2]);
      v[3] = std::max(v[3], q[3]);
      v[4] = std::max(v[4], q[4]);
      v[5] = std::max(v[5], q[5]);
      v[6] += q[6];
    }
    const int ns = s_off[c + 1] - s_off[c];
    out[SPHNJ_PROBE_FORCE_X] = v[0];
    out[SPHNJ_PROBE_FORCE_Y] = v[1];
    out[SPHNJ_PROBE_FRONT] = std::max(v[2], 0.0) + 0.5 * D.d;
    out[SPHNJ_PROBE_HEIGHT] = std::max(v[3], -0.5 * D.d) + 0.5 * D.d;
    out[SPHNJ_PROBE_VM_ELASTIC] = v[4];
    out[SPHNJ_PROBE_VM_AUXETIC] = v[5];
    out[SPHNJ_PROBE_VM_MAX] = std::max(v[4], v[5]);
    out[SPHNJ_PROBE_MEAN_PRESSURE] = ns > 0 ? v[6] / ns : 0.0;
    return SPHNJ_OK;
  }

  int probe_solid(int c, double X, double Y, double pos[2], int *index) override {
    const Derived &D = Dh[c];
    const int lo = s_off[c], hi = s_off[c + 1];
    const int none = INT_MAX;
    int rc, best = none;
    if ((rc = h2d(probe_best, &none, 1))) return rc;
    if (hi > lo) {
      k_probe_point<Pol><<<nblk(hi - lo, 256), 256, 0, stream>>>(d, lo, hi, X, Y, 0.25 * D.d * D.d, probe_best);
      LAUNCH_CHECK();
    }
    if ((rc = d2h(&best, probe_best, 1))) return rc;
    HIPC(hipStreamSynchronize(stream));
    if (best == none) {
      if (index) *index = -1;
      err = "no solid particle has its reference position within d/2 of the probe point";
      return SPHNJ_E_ARG;
    }
    double xy[2];
    if ((rc = d2h(xy, d.sx + best, 1)) || (rc = d2h(xy + 1, d.sy + best, 1))) return rc;
    HIPC(hipStreamSynchronize(stream));
    pos[0] = xy[0];
    pos[1] = xy[1];
    if (index) *index = best - lo;
    return SPHNJ_OK;
  }

  int energy(int c, double out[6]) override {
    void *q1 = nullptr, *q2 = nullptr;
    HIPC(hipMalloc(&q1, kEnergyConsts * sizeof(double)));
    if (hipMalloc(&q2, (size_t)kEnergyBlocks * 5 * sizeof(double)) != hipSuccess) {
      (void)hipFree(q1);
      err = "hipMalloc";
      return SPHNJ_E_HIP;
    }
    const int rc = energy_in(c, out, (double *)q1, (double *)q2);
    (void)hipFree(q1);
    (void)hipFree(q2);
    return rc;
  }

  // The reductions of energy() into caller-provided device memory: kEnergyConsts doubles at
  // dc, kEnergyBlocks x 5 partials at part (every one rewritten by each call).
  int energy_in(int c, double out[6], double *dc, double *part) override {
    int rc;
    const Derived &D = Dh[c];
    const double consts[kEnergyConsts] = {D.d * D.d, D.rho0[0], D.rho0[1], D.rho_s, D.rho_aux, D.lambda, D.mu_s};
    const int G = kEnergyBlocks;
    std::vector<double> hp((size_t)G * 5);
    rc = h2d(dc, consts, kEnergyConsts);
    if (!rc) {
      k_energy_partial<Pol><<<G, 256, 0, stream>>>(d, f_off[c], f_off[c + 1], s_off[c], s_off[c + 1], dc, part);
      rc = d2h(hp.data(), part, (size_t)G * 5);
    }
    const hipError_t e = hipStreamSynchronize(stream);
    if (rc) return rc;
    if (e != hipSuccess) {
      err = hipGetErrorString(e);
      return SPHNJ_E_HIP;
    }
    double ke = 0, mass = 0, m2 = 0, kes = 0, ep = 0;
    for (int b = 0; b < G; ++b) {
      ke += hp[5 * b];
      mass += hp[5 * b + 1];
      m2 = std::max(m2, hp[5 * b + 2]);
      kes += hp[5 * b + 3];
      ep += hp[5 * b + 4];
    }
    out[0] = ke;
    out[1] = mass;
    out[2] = std::sqrt(m2);
    out[3] = kes;
    out[4] = ep;
    out[5] = (double)(f_off[c + 1] - f_off[c]);
    return SPHNJ_OK;
  }
};

}  // namespace

struct sphnj_ctx {
  int precision = 0;
  int ncase = 0;
  std::vector<sphnj_params> P;
  std::vector<Derived> D;
  std::vector<Staged> staged;
  // per case, kept after init for snapshots: solid material and bc, wall bc (input order)
  std::vector<std::vector<int32_t>> smat, sbc;
  std::unique_ptr<EngineBase> eng;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double last_ms = 0.0;
  bool inited = false;
  bool use_graph = true;
  bool persistent_solid = false;
  int shift_runs = 1;
  double *scratch_ptr = nullptr;  // this device's `scratch` (sphnj_energy_noalloc)
  std::string err;
  int fail(int code, const std::string &m) {
    err = m;
    return code;
  }
  int from_engine(int code) {
    if (code && eng) err = eng->err;
    return code;
  }
};

extern "C" {

int sphnj_create(const sphnj_params *params, int precision, sphnj_ctx **out) {
  return sphnj_batch_create(params, 1, precision, out);
}

int sphnj_batch_create(const sphnj_params *pc, int32_t ncase, int precision, sphnj_ctx **out) {
  if (!pc || ncase < 1 || !out) {
    set_global_error("sphnj_batch_create: bad argument");
    return SPHNJ_E_ARG;
  }
  *out = nullptr;
  if (precision != SPHNJ_PREC_P64 && precision != SPHNJ_PREC_M32) {
    set_global_error("unknown precision");
    return SPHNJ_E_ARG;
  }
  std::vector<Derived> D(ncase);
  for (int c = 0; c < ncase; ++c) {
    const int rc = sphnj_params_validate(&pc[c]);
    if (rc) return rc;
    std::string e;
    if (derive(pc[c], &D[c], &e)) {
      set_global_error(e);
      return SPHNJ_E_PARAM;
    }
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
    set_global_error("no GPU visible");
    return SPHNJ_E_NODEVICE;
  }
  auto *c = new sphnj_ctx();
  c->precision = precision;
  c->ncase = ncase;
  c->P.assign(pc, pc + ncase);
  c->D = D;
  c->staged.resize(ncase);
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
    set_global_error("cannot create HIP stream/events");
    delete c;
    return SPHNJ_E_HIP;
  }
  *out = c;
  return SPHNJ_OK;
}

int sphnj_upload(sphnj_ctx *c, int32_t ci, const sphnj_fluid_in *f, const sphnj_solid_in *s, const sphnj_wall_in *w) {
  if (!c) return SPHNJ_E_ARG;
  if (ci < 0 || ci >= c->ncase) return c->fail(SPHNJ_E_ARG, "case index out of range");
  if (c->inited) return c->fail(SPHNJ_E_STATE, "upload after init");
  Staged &S = c->staged[ci];
  S = Staged();
  if (f && f->n > 0) {
    if (!f->x || !f->y || !f->u || !f->v || !f->rho || !f->p || !f->phase)
      return c->fail(SPHNJ_E_ARG, "fluid input has a null array");
    S.fx.assign(f->x, f->x + f->n);
    S.fy.assign(f->y, f->y + f->n);
    S.fu.assign(f->u, f->u + f->n);
    S.fv.assign(f->v, f->v + f->n);
    S.frho.assign(f->rho, f->rho + f->n);
    S.fp.assign(f->p, f->p + f->n);
    for (int i = 0; i < f->n; ++i) {
      if (f->phase[i] != 0 && f->phase[i] != 1) return c->fail(SPHNJ_E_ARG, "fluid phase must be 0 or 1");
      S.fph.push_back((int8_t)f->phase[i]);
    }
  }
  if (s && s->n > 0) {
    if (!s->x || !s->y || !s->u || !s->v || !s->X || !s->Y || !s->material)
      return c->fail(SPHNJ_E_ARG, "solid input has a null array");
    S.sx.assign(s->x, s->x + s->n);
    S.sy.assign(s->y, s->y + s->n);
    S.su.assign(s->u, s->u + s->n);
    S.sv.assign(s->v, s->v + s->n);
    S.sX.assign(s->X, s->X + s->n);
    S.sY.assign(s->Y, s->Y + s->n);
    for (int i = 0; i < s->n; ++i) {
      if (s->material[i] < 0 || s->material[i] > 2) return c->fail(SPHNJ_E_ARG, "unknown solid material");
      S.smat.push_back((int8_t)s->material[i]);
      const int b = s->bc ? s->bc[i] : 0;
      S.sbc.push_back((int8_t)(b == 1 ? 1 : 0));
    }
  }
  if (w && w->n > 0) {
    if (!w->x || !w->y || !w->bc) return c->fail(SPHNJ_E_ARG, "wall input has a null array");
    S.wx.assign(w->x, w->x + w->n);
    S.wy.assign(w->y, w->y + w->n);
    for (int i = 0; i < w->n; ++i) S.wbc.push_back((int8_t)w->bc[i]);
  }
  S.uploaded = true;
  return SPHNJ_OK;
}

int sphnj_init(sphnj_ctx *c) {
  if (!c) return SPHNJ_E_ARG;
  if (c->inited) return c->fail(SPHNJ_E_STATE, "already initialised");
  for (int i = 0; i < c->ncase; ++i)
    if (!c->staged[i].uploaded) return c->fail(SPHNJ_E_STATE, "case " + std::to_string(i) + " not uploaded");
  if (c->precision == SPHNJ_PREC_P64)
    c->eng.reset(new Engine<P64>());
  else
    c->eng.reset(new Engine<M32>());
  c->eng->stream = c->stream;
  c->eng->use_graph = c->use_graph;
  c->eng->persistent_solid = c->persistent_solid;
  c->eng->shift_runs = c->shift_runs;
  {
    std::string e;
    if (scratch_address(&c->scratch_ptr, &e) != SPHNJ_OK) {
      c->eng.reset();
      return c->fail(SPHNJ_E_HIP, e);
    }
  }
  const int rc = c->eng->init(c->D, c->staged);
  if (rc) {
    c->err = c->eng->err;
    c->eng.reset();
    return rc;
  }
  c->smat.assign(c->ncase, {});
  c->sbc.assign(c->ncase, {});
  for (int i = 0; i < c->ncase; ++i) {
    c->smat[i].assign(c->staged[i].smat.begin(), c->staged[i].smat.end());
    c->sbc[i].assign(c->staged[i].sbc.begin(), c->staged[i].sbc.end());
  }
  c->staged.clear();
  c->staged.shrink_to_fit();
  c->inited = true;
  return SPHNJ_OK;
}

int sphnj_step(sphnj_ctx *c, int32_t n) {
  if (!c || n < 0) return SPHNJ_E_ARG;
  if (!c->inited) return c->fail(SPHNJ_E_STATE, "step before init");
  if (hipEventRecord(c->ev0, c->stream) != hipSuccess) return c->fail(SPHNJ_E_HIP, "hipEventRecord");
  const int rc = c->eng->step(n);
  if (rc) return c->from_engine(rc);
  if (hipEventRecord(c->ev1, c->stream) != hipSuccess) return c->fail(SPHNJ_E_HIP, "hipEventRecord");
  return SPHNJ_OK;
}

int sphnj_sync(sphnj_ctx *c) {
  if (!c) return SPHNJ_E_ARG;
  const hipError_t e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return c->fail(SPHNJ_E_HIP, hipGetErrorString(e));
  return SPHNJ_OK;
}

int sphnj_last_step_ms(sphnj_ctx *c, double *ms) {
  if (!c || !ms) return SPHNJ_E_ARG;
  if (hipEventSynchronize(c->ev1) != hipSuccess) return c->fail(SPHNJ_E_HIP, "hipEventSynchronize");
  float f = 0.f;
  if (hipEventElapsedTime(&f, c->ev0, c->ev1) != hipSuccess) return c->fail(SPHNJ_E_HIP, "hipEventElapsedTime");
  *ms = (double)f;
  return SPHNJ_OK;
}

int sphnj_stage(sphnj_ctx *c, int32_t st) {
  if (!c) return SPHNJ_E_ARG;
  if (!c->inited) return c->fail(SPHNJ_E_STATE, "stage before init");
  if (st < SPHNJ_STAGE_CELLS || st > SPHNJ_STAGE_SHIFT) return c->fail(SPHNJ_E_ARG, "unknown stage");
  return c->from_engine(c->eng->stage(st));
}

int sphnj_download(sphnj_ctx *c, int32_t ci, const sphnj_fluid_out *f, const sphnj_solid_out *s,
                   const sphnj_wall_out *w) {
  if (!c) return SPHNJ_E_ARG;
  if (!c->inited) return c->fail(SPHNJ_E_STATE, "download before init");
  if (ci < 0 || ci >= c->ncase) return c->fail(SPHNJ_E_ARG, "case index out of range");
  return c->from_engine(c->eng->download(ci, f, s, w));
}

int sphnj_status(sphnj_ctx *c, uint32_t *flags) {
  if (!c || !flags) return SPHNJ_E_ARG;
  if (!c->inited) {
    *flags = 0;
    return SPHNJ_OK;
  }
  return c->from_engine(c->eng->status(flags));
}

int sphnj_counts(const sphnj_ctx *c, int32_t ci, int32_t *nf, int32_t *ns, int32_t *nw, int32_t *ncell) {
  if (!c || ci < 0 || ci >= c->ncase) return SPHNJ_E_ARG;
  if (c->inited) return c->eng->counts(ci, nf, ns, nw, ncell);
  const Staged &S = c->staged[ci];
  if (nf) *nf = (int32_t)S.fx.size();
  if (ns) *ns = (int32_t)S.sx.size();
  if (nw) *nw = (int32_t)S.wx.size();
  if (ncell) *ncell = c->D[ci].nx * c->D[ci].ny;
  return SPHNJ_OK;
}

const char *sphnj_last_error(const sphnj_ctx *c) { return c ? c->err.c_str() : "null context"; }

void sphnj_destroy(sphnj_ctx *c) {
  if (!c) return;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  c->eng.reset();
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

int sphnj_set_graph(sphnj_ctx *c, int enable) {
  if (!c) return SPHNJ_E_ARG;
  c->use_graph = enable != 0;
  if (c->eng) {
    c->eng->use_graph = c->use_graph;
    if (!c->use_graph) c->eng->drop_graphs();
  }
  return SPHNJ_OK;
}

int sphnj_set_persistent_solid(sphnj_ctx *c, int enable) {
  if (!c) return SPHNJ_E_ARG;
  c->persistent_solid = enable != 0;
  if (c->eng) {
    if (c->eng->persistent_solid != c->persistent_solid) c->eng->drop_graphs();
    c->eng->persistent_solid = c->persistent_solid;
  }
  return SPHNJ_OK;
}

int sphnj_set_shift_runs(sphnj_ctx *c, int mode) {
  if (!c || mode < 0 || mode > 2) return SPHNJ_E_ARG;
  c->shift_runs = mode;
  if (c->eng) {
    if (c->eng->shift_runs != mode) c->eng->drop_graphs();
    c->eng->shift_runs = mode;
    c->eng->pair_runs = false;
  }
  return SPHNJ_OK;
}

int sphnj_get_block(const sphnj_ctx *c, int32_t fam, int32_t *block) {
  if (!c || !block || !c->eng) return SPHNJ_E_ARG;
  *block = fam == 0 ? c->eng->blk_pair : fam == 1 ? c->eng->blk_shift : c->eng->blk_stream;
  return SPHNJ_OK;
}

int sphnj_set_block(sphnj_ctx *c, int32_t fam, int32_t block) {
  if (!c || !c->eng) return SPHNJ_E_ARG;
  if (block != 64 && block != 128 && block != 256) return c->fail(SPHNJ_E_ARG, "block must be 64, 128 or 256");
  if (fam == 0)
    c->eng->blk_pair = block;
  else if (fam == 1)
    c->eng->blk_shift = block;
  else
    c->eng->blk_stream = block;
  c->eng->pair_runs = false;
  c->eng->drop_graphs();
  return SPHNJ_OK;
}

int sphnj_set_stage_timing(sphnj_ctx *c, int enable) {
  if (!c || !c->eng) return SPHNJ_E_ARG;
  const hipError_t e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return c->fail(SPHNJ_E_HIP, hipGetErrorString(e));
  c->eng->timing = enable != 0;
  c->eng->reset_times();
  return SPHNJ_OK;
}

int sphnj_stage_time(sphnj_ctx *c, int32_t stage, int32_t *launches, double *total_ms) {
  if (!c || !c->eng || !launches || !total_ms) return SPHNJ_E_ARG;
  if (stage < SPHNJ_STAGE_CELLS || stage > SPHNJ_STAGE_SHIFT) return c->fail(SPHNJ_E_ARG, "unknown stage");
  int n = 0;
  double t = 0;
  const int rc = c->eng->stage_time(stage, &n, &t);
  if (rc) return c->from_engine(rc);
  *launches = n;
  *total_ms = t;
  return SPHNJ_OK;
}

int sphnj_autotune(sphnj_ctx *c) {
  if (!c) return SPHNJ_E_ARG;
  if (!c->inited) return c->fail(SPHNJ_E_STATE, "autotune before init");
  return c->from_engine(c->eng->autotune());
}

int sphnj_autotune_report(const sphnj_ctx *c, int32_t family, double ms_per_launch[3]) {
  if (!c || !c->eng || !ms_per_launch || family < 0 || family > 1) return SPHNJ_E_ARG;
  for (int k = 0; k < 3; ++k) ms_per_launch[k] = c->eng->tune_ms[family][k];
  return SPHNJ_OK;
}

int sphnj_stage_sorted(sphnj_ctx *c, int32_t ci, int32_t species, int32_t *keys, int32_t *orig, int32_t *start) {
  if (!c || ci < 0 || ci >= c->ncase || species < 0 || species > 2) return SPHNJ_E_ARG;
  if (!c->inited) return c->fail(SPHNJ_E_STATE, "before init");
  return c->from_engine(c->eng->sorted(ci, species, keys, orig, start));
}

int sphnj_stage_neighbors(sphnj_ctx *c, int32_t ci, int32_t *counts, int32_t *list, int64_t cap, int64_t *total) {
  if (!c || ci < 0 || ci >= c->ncase) return SPHNJ_E_ARG;
  if (!c->inited) return c->fail(SPHNJ_E_STATE, "before init");
  return c->from_engine(c->eng->neighbors(ci, counts, list, cap, total));
}

int sphnj_solid_reflist(sphnj_ctx *c, int32_t ci, int32_t *off, int32_t *list, int64_t cap) {
  if (!c || ci < 0 || ci >= c->ncase) return SPHNJ_E_ARG;
  if (!c->inited) return c->fail(SPHNJ_E_STATE, "before init");
  return c->from_engine(c->eng->reflist(ci, off, list, cap));
}

int sphnj_energy(sphnj_ctx *c, int32_t ci, double out[6]) {
  if (!c || !out || ci < 0 || ci >= c->ncase) return SPHNJ_E_ARG;
  if (!c->inited) return c->fail(SPHNJ_E_STATE, "before init");
  return c->from_engine(c->eng->energy(ci, out));
}

// ---- snapshots ------------------------------------------------------------------------
// Layout (little-endian): "SPHNJSNP", uint32 version, uint32 sizeof(sphnj_params), the
// parameter struct, int64 step, int32 nf ns nw; fluid x y u v rho p (double), phase class
// (int32); solid x y u v X Y p (double), normal (2 per), max_strain (4 per) (double),
// material bc (int32); wall x y (double), bc (int32); uint64 FNV-1a hash of everything
// before it.
}  // extern "C"

namespace {
constexpr char kSnapMagic[8] = {'S', 'P', 'H', 'N', 'J', 'S', 'N', 'P'};
constexpr uint32_t kSnapVersion = 1;

struct SnapWriter {
  std::FILE *f;
  uint64_t h = 1469598103934665603ull;
  bool ok = true;
  void put(const void *p, size_t n) {
    const auto *b = static_cast<const unsigned char *>(p);
    for (size_t i = 0; i < n; ++i) h = (h ^ b[i]) * 1099511628211ull;
    ok = ok && std::fwrite(p, 1, n, f) == n;
  }
  template <class T>
  void vec(const std::vector<T> &v) {
    if (!v.empty()) put(v.data(), v.size() * sizeof(T));
  }
};
struct SnapReader {
  std::FILE *f;
  uint64_t h = 1469598103934665603ull;
  bool ok = true;
  void get(void *p, size_t n) {
    ok = ok && std::fread(p, 1, n, f) == n;
    if (!ok) return;
    const auto *b = static_cast<const unsigned char *>(p);
    for (size_t i = 0; i < n; ++i) h = (h ^ b[i]) * 1099511628211ull;
  }
  template <class T>
  void vec(std::vector<T> &v, size_t n) {
    v.resize(n);
    if (n) get(v.data(), n * sizeof(T));
  }
};

struct SnapHeader {
  sphnj_params P;
  int64_t step = 0;
  int32_t n[3] = {0, 0, 0};
};

// Reads and checks the header; leaves the reader at the particle arrays.
int snap_header(SnapReader &r, SnapHeader *hd, std::string *err) {
  char magic[8];
  uint32_t ver = 0, psize = 0;
  r.get(magic, 8);
  r.get(&ver, 4);
  r.get(&psize, 4);
  if (!r.ok || std::memcmp(magic, kSnapMagic, 8) != 0) {
    *err = "not a sphnj snapshot";
    return SPHNJ_E_IO;
  }
  if (ver != kSnapVersion || psize != sizeof(sphnj_params)) {
    *err = "snapshot written by another version of the library";
    return SPHNJ_E_IO;
  }
  r.get(&hd->P, sizeof(sphnj_params));
  r.get(&hd->step, 8);
  r.get(hd->n, 12);
  if (!r.ok || hd->n[0] < 0 || hd->n[1] < 0 || hd->n[2] < 0) {
    *err = "truncated snapshot header";
    return SPHNJ_E_IO;
  }
  return SPHNJ_OK;
}
}  // namespace

extern "C" {

int sphnj_snapshot_write(sphnj_ctx *c, int32_t ci, int64_t step, const char *path) {
  if (!c || !path || ci < 0 || ci >= c->ncase) return SPHNJ_E_ARG;
  if (!c->inited) return c->fail(SPHNJ_E_STATE, "before init");
  int32_t n[3];
  sphnj_counts(c, ci, &n[0], &n[1], &n[2], nullptr);
  const size_t nf = n[0], ns = n[1], nw = n[2];
  std::vector<double> fx(nf), fy(nf), fu(nf), fv(nf), fr(nf), fp(nf);
  std::vector<int32_t> fph(nf), fcl(nf);
  std::vector<double> sx(ns), sy(ns), su(ns), sv(ns), sX(ns), sY(ns), sp(ns), sn(2 * ns), sM(4 * ns);
  std::vector<double> wx(nw), wy(nw);
  std::vector<int32_t> wbc(nw);
  sphnj_fluid_out fo{};
  fo.x = fx.data(); fo.y = fy.data(); fo.u = fu.data(); fo.v = fv.data(); fo.rho = fr.data(); fo.p = fp.data();
  fo.phase = fph.data(); fo.boundary_type = fcl.data();
  sphnj_solid_out so{};
  so.x = sx.data(); so.y = sy.data(); so.u = su.data(); so.v = sv.data(); so.X = sX.data(); so.Y = sY.data();
  so.pressure = sp.data(); so.normal = sn.data(); so.max_strain = sM.data();
  sphnj_wall_out wo{};
  wo.x = wx.data(); wo.y = wy.data(); wo.bc = wbc.data();
  int rc = sphnj_download(c, ci, &fo, &so, &wo);
  if (rc) return rc;
  std::FILE *f = std::fopen(path, "wb");
  if (!f) return c->fail(SPHNJ_E_IO, std::string("cannot open ") + path);
  SnapWriter w{f};
  const uint32_t ver = kSnapVersion, psize = sizeof(sphnj_params);
  w.put(kSnapMagic, 8);
  w.put(&ver, 4);
  w.put(&psize, 4);
  w.put(&c->P[ci], sizeof(sphnj_params));
  w.put(&step, 8);
  w.put(n, 12);
  for (auto *v : {&fx, &fy, &fu, &fv, &fr, &fp}) w.vec(*v);
  w.vec(fph);
  w.vec(fcl);
  for (auto *v : {&sx, &sy, &su, &sv, &sX, &sY, &sp, &sn, &sM}) w.vec(*v);
  w.vec(c->smat[ci]);
  w.vec(c->sbc[ci]);
  w.vec(wx);
  w.vec(wy);
  w.vec(wbc);
  const uint64_t hash = w.h;
  w.put(&hash, 8);
  const bool closed = std::fclose(f) == 0;
  if (!w.ok || !closed) return c->fail(SPHNJ_E_IO, std::string("error writing ") + path);
  return SPHNJ_OK;
}

int sphnj_snapshot_info(const char *path, sphnj_params *params, int64_t *step, int32_t counts[3]) {
  if (!path) return SPHNJ_E_ARG;
  std::FILE *f = std::fopen(path, "rb");
  if (!f) {
    set_global_error(std::string("cannot open ") + path);
    return SPHNJ_E_IO;
  }
  SnapReader r{f};
  SnapHeader hd;
  std::string err;
  const int rc = snap_header(r, &hd, &err);
  std::fclose(f);
  if (rc) {
    set_global_error(err);
    return rc;
  }
  if (params) *params = hd.P;
  if (step) *step = hd.step;
  if (counts) std::memcpy(counts, hd.n, sizeof(hd.n));
  return SPHNJ_OK;
}

int sphnj_snapshot_load(sphnj_ctx *c, int32_t ci, const char *path) {
  if (!c || !path) return SPHNJ_E_ARG;
  if (ci < 0 || ci >= c->ncase) return c->fail(SPHNJ_E_ARG, "case index out of range");
  if (c->inited) return c->fail(SPHNJ_E_STATE, "load after init");
  std::FILE *f = std::fopen(path, "rb");
  if (!f) return c->fail(SPHNJ_E_IO, std::string("cannot open ") + path);
  SnapReader r{f};
  SnapHeader hd;
  std::string err;
  int rc = snap_header(r, &hd, &err);
  if (rc) {
    std::fclose(f);
    return c->fail(rc, err);
  }
  const size_t nf = hd.n[0], ns = hd.n[1], nw = hd.n[2];
  std::vector<double> fx, fy, fu, fv, fr, fp, sx, sy, su, sv, sX, sY, sp, sn, sM, wx, wy;
  std::vector<int32_t> fph, fcl, smat, sbc, wbc;
  for (auto *v : {&fx, &fy, &fu, &fv, &fr, &fp}) r.vec(*v, nf);
  r.vec(fph, nf);
  r.vec(fcl, nf);
  for (auto *v : {&sx, &sy, &su, &sv, &sX, &sY, &sp}) r.vec(*v, ns);
  r.vec(sn, 2 * ns);
  r.vec(sM, 4 * ns);
  r.vec(smat, ns);
  r.vec(sbc, ns);
  r.vec(wx, nw);
  r.vec(wy, nw);
  r.vec(wbc, nw);
  const uint64_t expect = r.h;
  uint64_t hash = 0;
  r.get(&hash, 8);
  char extra;
  const bool tail = std::fread(&extra, 1, 1, f) == 1;
  std::fclose(f);
  if (!r.ok || hash != expect || tail) return c->fail(SPHNJ_E_IO, "snapshot truncated or damaged (hash mismatch)");
  if (std::memcmp(&hd.P, &c->P[ci], sizeof(sphnj_params)) != 0)
    return c->fail(SPHNJ_E_PARAM, "snapshot parameters differ from the context's for this case");
  const sphnj_fluid_in fi{hd.n[0], fx.data(), fy.data(), fu.data(), fv.data(), fr.data(), fp.data(), fph.data()};
  const sphnj_solid_in si{hd.n[1], sx.data(), sy.data(), su.data(), sv.data(), sX.data(), sY.data(), smat.data(),
                          sbc.data()};
  const sphnj_wall_in wi{hd.n[2], wx.data(), wy.data(), wbc.data()};
  if ((rc = sphnj_upload(c, ci, &fi, &si, &wi))) return rc;
  Staged &S = c->staged[ci];
  S.fcls.assign(fcl.begin(), fcl.end());
  S.sM = std::move(sM);
  S.sp = std::move(sp);
  S.sn = std::move(sn);
  return SPHNJ_OK;
}

int sphnj_energy_noalloc(sphnj_ctx *c, int32_t ci, double out[6]) {
  if (!c || !out || ci < 0 || ci >= c->ncase) return SPHNJ_E_ARG;
  if (!c->inited) return c->fail(SPHNJ_E_STATE, "before init");
  const int need = kEnergyConsts + kEnergyBlocks * 5;
  if (need > kScratchDoubles)
    return c->fail(SPHNJ_E_UNSUPPORTED, "energy reductions need " + std::to_string(need) +
                                             " doubles of scratch, more than the " +
                                             std::to_string(kScratchDoubles) + " the library holds");
  if (!c->scratch_ptr) return c->fail(SPHNJ_E_STATE, "no scratch address for this device");
  return c->from_engine(c->eng->energy_in(ci, out, c->scratch_ptr, c->scratch_ptr + 8));
}

int sphnj_time(sphnj_ctx *c, int32_t ci, double *time, double *last_dt) {
  if (!c || ci < 0 || ci >= c->ncase) return SPHNJ_E_ARG;
  if (!c->inited) return c->fail(SPHNJ_E_STATE, "before init");
  return c->from_engine(c->eng->sim_time(ci, time, last_dt));
}

int sphnj_probe_global(sphnj_ctx *c, int32_t ci, double height_x, double out[SPHNJ_PROBE_COUNT]) {
  if (!c || !out || ci < 0 || ci >= c->ncase) return SPHNJ_E_ARG;
  if (!c->inited) return c->fail(SPHNJ_E_STATE, "before init");
  return c->from_engine(c->eng->probe_global(ci, height_x, out));
}

int sphnj_probe_solid(sphnj_ctx *c, int32_t ci, double X, double Y, double pos[2], int32_t *index) {
  if (!c || !pos || ci < 0 || ci >= c->ncase) return SPHNJ_E_ARG;
  if (!c->inited) return c->fail(SPHNJ_E_STATE, "before init");
  return c->from_engine(c->eng->probe_solid(ci, X, Y, pos, index));
}

}  // extern "C"
