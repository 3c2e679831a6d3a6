// TEST INFRASTRUCTURE ONLY -- the FP64 CPU oracle of SPHNJ's per-step particle pipeline.
//
// An independent restatement of the reference numerics (niji127/SPHNJ), written from the
// reference sources read as text.  Each function cites the reference file:line it follows.
// It is the checker of libsphnj (tests/, smoke) and the CPU baseline of bench.py; the product
// never calls it.
//
// Conventions fixed here and followed by libsphnj (documented choices, not reference facts):
//  * Every step hashes fluid and solid particles and orders fluids by (cell, quadrant of the
//    cell, original index) and solids by (cell, original index); walls likewise, once.  The
//    reference rebuilds its per-cell lists every step from the previous lists
//    (src/pair.cu:66-181) and sorts every sort_every_step steps (src/sort.cu:139-150); the set
//    of particles listed in a cell is the same, only the order inside a cell differs.
//  * Neighbour iteration: the 9 cells in the order of Sim::createCell (src/sim.cu:175-193,
//    x-major: (ix-1,iy-1), (ix-1,iy), (ix-1,iy+1), (ix,iy-1), ...), and inside a cell fluid,
//    then solid, then wall particles (the layout of Cell::particle_list, src/sort.cu:59-92).
//  * Walls are static; they are hashed and ordered once at init.
//  * A particle whose cell index falls outside [0,n) on an axis is clamped to the edge cell and
//    flags SPHNJ_ST_OUT_OF_GRID (the reference indexes out of range there).
//  * Quirk fixed: the fluid-wall colour filter reads the wall's phase (the reference reads
//    fluid.phase_type at the wall index, src/fluid.cu:179).
//  * Solid-solid pairs come from reference-configuration neighbour lists built at init from the
//    reference positions; the reference filters its current 3x3 cells by reference distance
//    (src/solid.cu:184-199).  The sets agree while every listed neighbour stays within the 3x3
//    current cells; SPHNJ_ST_REFLIST_LEFT flags a step where one does not.
#include "sphnj_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

thread_local std::string g_err;

struct Consts {  // restates Kernel/FluidProperties/SolidProperties/Domain::initiate in double
  double h, d, rc2, h2, hinv2, kdc, k2c, eR;
  double rho0[2], c[2], gam[2], p2rho[2], rho2p[2], mu[2], rho_min[2];
  double p_min, p_bg, shift_coef, umax_cap, gx, gy;
  int steady, viscid;
  double lambda, mu_s, c_s, alpha, beta, rho_s, rho_aux, fixed_u, fixed_v;
  double aux_mod, yield_strain, yield_stress, harden_strain, harden_coef;
  int plastic;
  double cxmin, cymin, cdx, cdy, inv_cdx, inv_cdy;
  int nx, ny;
  double dxmin, dymin, dxmax, dymax;
  double dt, dt_s;
  int S, K;
  double cfl, cfl_h, c_max, dt_solid_cap;  // adaptive step size (extension; cfl = 0: off)
};

bool make_consts(const sphnj_params &p, Consts *k) {
  Consts &c = *k;
  // Kernel::initiate, src/kernel.cpp:25-41 (tg_coef_2 = 5.551921, include/global.h:12)
  c.h = p.smoothing_length;
  c.d = p.particle_diameter;
  c.h2 = c.h * c.h;
  c.hinv2 = 1.0 / c.h2;
  const double R = p.impact_length_hsml_ratio;
  c.rc2 = R * R * c.h2;
  const double r = c.h / c.d, rm3 = 1.0 / (r * r * r);
  const double c32 = p.kernel_2d_coef != 0.0 ? p.kernel_2d_coef : 2.120729;  // COEF_3D_TO_2D
  c.eR = std::exp(-(R * R));
  c.kdc = (2.0 / 5.551921 * rm3 / c.h2) * c32;
  c.k2c = (1.0 / 5.551921 * rm3) * c32;
  // FluidProperties::initiate, src/properties.cpp:53-67
  const double cc[2] = {p.liquid_sound_speed, p.gas_sound_speed};
  const double rr[2] = {p.liquid_density, p.gas_density};
  const double gg[2] = {p.liquid_gamma, p.gas_gamma};
  const double mm[2] = {p.liquid_viscosity, p.gas_viscosity};
  for (int i = 0; i < 2; ++i) {
    c.c[i] = cc[i];
    c.rho0[i] = rr[i];
    c.gam[i] = gg[i];
    c.mu[i] = mm[i];
    c.p2rho[i] = gg[i] / (cc[i] * cc[i]) / rr[i];
    c.rho2p[i] = 1.0 / c.p2rho[i];
    c.rho_min[i] = 0.6 * rr[i];
  }
  c.p_min = p.min_pressure;
  c.p_bg = p.background_pressure;
  c.shift_coef = p.particle_shift_coef;
  c.umax_cap = 0.1 * std::min(cc[0], cc[1]);  // src/fluid.cu:438-439
  c.gx = 0.0;
  c.gy = -p.gravity;  // src/physics.cpp:12-14
  c.steady = p.is_steady != 0;
  c.viscid = p.is_viscid != 0;
  // SolidProperties::initiate, src/properties.cpp:69-89
  const double E = p.youngs_modulus, nu = p.poisson_ratio;
  c.lambda = p.is_2D_strain ? E * nu / (1.0 - 2.0 * nu) / (1.0 + nu) : E * nu / (1.0 - nu) / (1.0 + nu);
  c.mu_s = E / (1.0 + nu) / 2.0;
  c.rho_s = p.solid_density;
  c.c_s = std::sqrt(E / (3.0 * (1.0 - 2.0 * nu)) / c.rho_s);
  c.alpha = p.arti_vis_alpha;
  c.beta = p.arti_vis_beta;
  c.rho_aux = p.auxetic_density;
  c.fixed_u = p.solid_fixed_speed_x;
  c.fixed_v = p.solid_fixed_speed_y;
  c.yield_stress = p.yield_stress;
  c.harden_strain = p.harden_strain;
  c.harden_coef = p.harden_coef;
  c.plastic = p.is_plastic != 0;
  // src/properties.cpp:155-158
  if (p.yield_strain != 0.0) {
    c.yield_strain = p.yield_strain;
    c.aux_mod = p.yield_stress / p.yield_strain;
  } else {
    c.aux_mod = p.auxetic_modulus;
    c.yield_strain = p.auxetic_modulus != 0.0 ? p.yield_stress / p.auxetic_modulus : 0.0;
  }
  // Domain::initiate(Kernel), src/domain.cpp:37-48
  const double cs = R * c.h;
  const double ex = p.cell_xmax - p.cell_xmin, ey = p.cell_ymax - p.cell_ymin;
  if (!(ex / cs >= 1.0) || !(ey / cs >= 1.0) || ex / cs > 32767.0 || ey / cs > 32767.0) return false;
  c.nx = (int)(ex / cs);
  c.ny = (int)(ey / cs);
  c.cxmin = p.cell_xmin;
  c.cymin = p.cell_ymin;
  c.cdx = ex / (double)c.nx;
  c.cdy = ey / (double)c.ny;
  c.inv_cdx = 1.0 / c.cdx;
  c.inv_cdy = 1.0 / c.cdy;
  c.dxmin = p.domain_xmin;
  c.dymin = p.domain_ymin;
  c.dxmax = p.domain_xmax;
  c.dymax = p.domain_ymax;
  c.dt = p.dt;
  c.S = p.solid_sub_step > 0 ? p.solid_sub_step : 1;
  c.K = p.shift_per_step > 0 ? p.shift_per_step : 0;
  c.dt_s = p.dt / (double)c.S;
  c.cfl = p.adaptive_cfl > 0.0 ? p.adaptive_cfl : 0.0;
  c.cfl_h = c.cfl * c.h;
  c.c_max = std::max(cc[0], cc[1]);
  c.dt_solid_cap = (double)c.S * c.cfl_h / c.c_s;
  return true;
}

// One axis of findPositionCell (src/pair.cu:8-18): truncation toward zero; outside [0, n)
// is clamped and reported.  *half is the half of the cell the particle lies in (0 lower, 1
// upper; clamped particles take the side they left by).
inline int axis_cell(double x, double x0, double inv, int n, unsigned *flags, int *half = nullptr) {
  const double t = (x - x0) * inv;
  if (!(t > -1.0)) {
    *flags |= SPHNJ_ST_OUT_OF_GRID;
    if (half) *half = 0;
    return 0;
  }
  if (!(t < (double)n)) {
    *flags |= SPHNJ_ST_OUT_OF_GRID;
    if (half) *half = 1;
    return n - 1;
  }
  const int ix = (int)t;
  if (half) {
    const int h2 = (int)(2.0 * t) - 2 * ix;
    *half = h2 < 0 ? 0 : (h2 > 1 ? 1 : h2);
  }
  return ix;
}

struct CellIndex {  // one species: per-particle cell, sorted list, cell starts
  std::vector<int> cx, cy, key, sub, sorted, start;
  // fluids: slot order of the library, x-major over the half-cell grid, with sub-cell keys
  // and starts (4 ncell + 1)
  std::vector<int> xkey, xsorted, xstart;
};

struct Case {
  sphnj_params P;
  Consts k;
  int cell_off = 0;
  int ncell = 0;
  // fluid, original order
  std::vector<double> fx, fy, fu, fv, frho, fp, fcol, fL;
  std::vector<int> fph, fcls, fnear;
  // solid
  std::vector<double> sx, sy, su, sv, sX, sY, sp, sn, sF, sLc, sP, sM, sa, scp;
  std::vector<int> smat, sbc;
  std::vector<int> ref_off, ref_list;
  // wall
  std::vector<double> wx, wy, wu, wv, wp, wrho, wn;
  std::vector<int> wph, wbc;
  CellIndex cf, cs, cw;
  bool uploaded = false;
  // step size: the current step's, the time at its end, and max |u|^2 after the last fluid
  // update (the input speeds before the first)
  double dt_cur = 0.0, t_sim = 0.0, u2_last = 0.0;
};

struct Ctx {
  std::vector<Case> cases;
  unsigned flags = 0;
  bool inited = false;
};

// Stable counting sort of 0..n-1 by key (0 <= key < nb): order and bin starts (nb + 1).
void counting_sort(const std::vector<int> &key, int nb, std::vector<int> *order, std::vector<int> *start) {
  const int n = (int)key.size();
  start->assign(nb + 1, 0);
  for (int i = 0; i < n; ++i) (*start)[key[i] + 1]++;
  for (int q = 0; q < nb; ++q) (*start)[q + 1] += (*start)[q];
  std::vector<int> fill(start->begin(), start->end() - 1);
  order->assign(n, 0);
  for (int i = 0; i < n; ++i) (*order)[fill[key[i]]++] = i;
}

// Hash one species and order it stably over original order: by (cell, original index) for
// solids and walls; by (cell, quadrant, original index) for fluids, the quadrant being
// 2 hx + hy from the halves of the cell the particle lies in.  For fluids it also forms the
// library's slot order, x-major over the half-cell grid (sub-cell (2cx + hx, 2cy + hy)), in
// which a cell's fluids appear in the same (quadrant, original index) order.
void hash_species(const Case &c, const std::vector<double> &x, const std::vector<double> &y, CellIndex *ci,
                  unsigned *flags, bool quadrants = false) {
  const Consts &k = c.k;
  const int n = (int)x.size();
  ci->cx.assign(n, 0);
  ci->cy.assign(n, 0);
  ci->key.assign(n, 0);
  ci->sub.assign(n, 0);
  ci->xkey.assign(quadrants ? n : 0, 0);
  unsigned fl = 0;
  for (int i = 0; i < n; ++i) {
    int hx = 0, hy = 0;
    const int ix = axis_cell(x[i], k.cxmin, k.inv_cdx, k.nx, &fl, &hx);
    const int iy = axis_cell(y[i], k.cymin, k.inv_cdy, k.ny, &fl, &hy);
    ci->cx[i] = ix;
    ci->cy[i] = iy;
    ci->key[i] = iy + ix * k.ny;  // src/pair.cu:17
    ci->sub[i] = quadrants ? ci->key[i] * 4 + hx * 2 + hy : ci->key[i];
    if (quadrants) ci->xkey[i] = (2 * iy + hy) + (2 * ix + hx) * (2 * k.ny);
  }
  *flags |= fl;
  std::vector<int> st;
  counting_sort(ci->sub, quadrants ? 4 * c.ncell : c.ncell, &ci->sorted, &st);
  ci->start.assign(c.ncell + 1, 0);
  for (int q = 0; q <= c.ncell; ++q) ci->start[q] = st[quadrants ? 4 * q : q];
  if (quadrants) counting_sort(ci->xkey, 4 * c.ncell, &ci->xsorted, &ci->xstart);
}

inline double grad_w(const Consts &k, double r2) { return std::exp(-(r2 * k.hinv2)) * k.kdc; }
inline double shep_w(const Consts &k, double r2) { return (std::exp(-(r2 * k.hinv2)) - k.eR) * k.k2c; }

// Visit the 9 neighbour cells of (ix, iy) in Sim::createCell order (src/sim.cu:175-193).
template <class F>
inline void for_cells(const Consts &k, int ix, int iy, F &&f) {
  for (int a = -1; a <= 1; ++a) {
    const int jx = ix + a;
    if (jx < 0 || jx >= k.nx) continue;
    for (int b = -1; b <= 1; ++b) {
      const int jy = iy + b;
      if (jy < 0 || jy >= k.ny) continue;
      f(jy + jx * k.ny);
    }
  }
}

// pressureToDensity / densityToPressure (src/fluid.cu:6-23) plus the background-pressure
// extension (zero reproduces the reference).
inline double p_to_rho(const Consts &k, double p, int ph) {
  return k.rho0[ph] * std::pow((p - k.p_bg) * k.p2rho[ph] + 1.0, 1.0 / k.gam[ph]);
}
inline double rho_to_p(const Consts &k, double rho, int ph) {
  const double p = (std::pow(rho / k.rho0[ph], k.gam[ph]) - 1.0) * k.rho2p[ph] + k.p_bg;
  return std::max(p, k.p_min);
}

// ---- walls ---------------------------------------------------------------------------------

// getVirtNormal (src/virt.cu:8-48), over the wall cells instead of all walls (same set: every
// wall lies in the cell box).
void wall_normals(Case &c) {
  const Consts &k = c.k;
  const int n = (int)c.wx.size();
  c.wn.assign(2 * n, 0.0);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    double nx = 0, ny = 0;
    const double xi = c.wx[i], yi = c.wy[i];
    for_cells(k, c.cw.cx[i], c.cw.cy[i], [&](int q) {
      for (int s = c.cw.start[q]; s < c.cw.start[q + 1]; ++s) {
        const int j = c.cw.sorted[s];
        const double dx = c.wx[j] - xi, dy = c.wy[j] - yi;
        const double r2 = dx * dx + dy * dy;
        if (r2 < k.rc2) {
          const double w = grad_w(k, r2);
          nx -= dx * w;
          ny -= dy * w;
        }
      }
    });
    const double nn = std::sqrt(nx * nx + ny * ny);
    if (nn < 0.1) {  // EPS_FOR_SUM, include/global.h:6
      c.wn[2 * i] = 0.0;
      c.wn[2 * i + 1] = 0.0;
    } else {
      c.wn[2 * i] = nx / nn;
      c.wn[2 * i + 1] = ny / nn;
    }
  }
}

// getVirtData + virtDataCalculate (src/virt.cu:61-149)
void wall_update(Case &c) {
  const Consts &k = c.k;
  const int n = (int)c.wx.size();
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    const double xi = c.wx[i], yi = c.wy[i];
    double S = 0, U = 0, V = 0, Pp = 0, nearest = 0;
    int phase = SPHNJ_LIQUID;
    bool found = false;
    for_cells(k, c.cw.cx[i], c.cw.cy[i], [&](int q) {
      for (int s = c.cf.start[q]; s < c.cf.start[q + 1]; ++s) {
        const int j = c.cf.sorted[s];
        const double dx = c.fx[j] - xi, dy = c.fy[j] - yi;
        const double r2 = dx * dx + dy * dy;
        if (!(r2 < k.rc2)) continue;
        const double w = shep_w(k, r2);
        S += w;
        U += c.fu[j] * w;
        V += c.fv[j] * w;
        Pp += (c.fp[j] - c.frho[j] * (k.gx * dx + k.gy * dy)) * w;
        if (!found || r2 < nearest) {
          found = true;
          nearest = r2;
          phase = c.fph[j];
        }
      }
    });
    double u = 0, v = 0, p = 0;
    if (S != 0.0) {
      u = U / S;
      v = V / S;
      p = Pp / S;
    }
    const double nx = c.wn[2 * i], ny = c.wn[2 * i + 1];
    if (c.wbc[i] == SPHNJ_SLIP) {
      const double un = u * nx + v * ny;
      u = u - 2.0 * un * nx;
      v = v - 2.0 * un * ny;
    } else if (c.wbc[i] == SPHNJ_NO_SLIP) {
      u = -u;
      v = -v;
    }
    c.wu[i] = u;
    c.wv[i] = v;
    c.wp[i] = p;
    c.wph[i] = phase;
    c.wrho[i] = p_to_rho(k, p, phase);
  }
}

// ---- fluid ---------------------------------------------------------------------------------

// Riemann pair terms shared by fluidFluidCalculate / fluidSolidCalculate / fluidVirtCalculate
// (src/fluid.cu:25-191).  Accumulates a, Drho, L; returns nothing else.
struct PairAcc {
  double ax = 0, ay = 0, drho = 0, L00 = 0, L01 = 0, L10 = 0, L11 = 0, col = 0;
  int near = SPHNJ_NO_PARTICLE;
};

inline void riemann(const Consts &k, double dx, double dy, double r2, double w, double ui, double vi, double uj,
                    double vj, double rhoi, double rhoj, double pi, double pj, double cs, int phi, PairAcc &A) {
  double ex = 0, ey = 0;
  if (r2 > 0.0) {  // Eigen normalized(): zero stays zero
    const double nrm = std::sqrt(r2);
    ex = dx / nrm;
    ey = dy / nrm;
  }
  double ps = (rhoi * pj + rhoj * pi) / (rhoi + rhoj);
  const double du = (ui - uj) * ex + (vi - vj) * ey;
  if (du >= 0.0) ps += rhoi * rhoj * du / (rhoi + rhoj) * std::min(3.0 * du, cs);
  const double vsx = (rhoi * ui + rhoj * uj) / (rhoi + rhoj) + (pi - pj) / (rhoi + rhoj) / cs * ex;
  const double vsy = (rhoi * vi + rhoj * vj) / (rhoi + rhoj) + (pi - pj) / (rhoi + rhoj) / cs * ey;
  const double s = -2.0 * ps / rhoi;
  A.ax += s * dx * w;
  A.ay += s * dy * w;
  A.drho += 2.0 * rhoi * ((ui - vsx) * dx + (vi - vsy) * dy) * w;
  A.L00 += (uj - ui) * dx * w;
  A.L01 += (uj - ui) * dy * w;
  A.L10 += (vj - vi) * dx * w;
  A.L11 += (vj - vi) * dy * w;
  if (k.viscid) {
    const double vis = ((uj - ui) * dx + (vj - vi) * dy) / (r2 + 0.01 * k.h2);
    const double t = 8.0 * k.mu[phi] * vis / rhoi;
    A.ax += t * dx * w;
    A.ay += t * dy * w;
  }
}

// findFluidPair + fluidUpdate (src/fluid.cu:193-287)
void fluid_calc(Case &c) {
  const Consts &k = c.k;
  const int n = (int)c.fx.size();
  std::vector<double> nx_(n), ny_(n), nu_(n), nv_(n), nr_(n), np_(n);
#pragma omp parallel for schedule(dynamic, 256)
  for (int i = 0; i < n; ++i) {
    const double xi = c.fx[i], yi = c.fy[i], ui = c.fu[i], vi = c.fv[i], rhoi = c.frho[i], pi = c.fp[i];
    const int phi = c.fph[i];
    const double cs = k.c[phi];
    PairAcc A;
    for_cells(k, c.cf.cx[i], c.cf.cy[i], [&](int q) {
      for (int s = c.cf.start[q]; s < c.cf.start[q + 1]; ++s) {  // fluidFluidCalculate
        const int j = c.cf.sorted[s];
        const double dx = c.fx[j] - xi, dy = c.fy[j] - yi;
        const double r2 = dx * dx + dy * dy;
        if (!(r2 < k.rc2)) continue;
        const double w = grad_w(k, r2);
        riemann(k, dx, dy, r2, w, ui, vi, c.fu[j], c.fv[j], rhoi, c.frho[j], pi, c.fp[j], cs, phi, A);
        const int phj = c.fph[j];
        if (!(phi == SPHNJ_LIQUID && phj == SPHNJ_GAS)) A.col += r2 * w;
        if (c.fcls[j] == SPHNJ_FREE_SURFACE && phi == SPHNJ_LIQUID && j != i) {
          if (r2 < 0.25 * k.rc2)
            A.near = SPHNJ_HAS_NEAR_PARTICLE;
          else if (A.near == SPHNJ_NO_PARTICLE)
            A.near = SPHNJ_HAS_PARTICLE;
        }
      }
      for (int s = c.cs.start[q]; s < c.cs.start[q + 1]; ++s) {  // fluidSolidCalculate
        const int j = c.cs.sorted[s];
        const double dx = c.sx[j] - xi, dy = c.sy[j] - yi;
        const double r2 = dx * dx + dy * dy;
        if (!(r2 < k.rc2)) continue;
        const double w = grad_w(k, r2);
        riemann(k, dx, dy, r2, w, ui, vi, c.su[j], c.sv[j], rhoi, rhoi, pi, c.sp[j], cs, phi, A);
      }
      for (int s = c.cw.start[q]; s < c.cw.start[q + 1]; ++s) {  // fluidVirtCalculate
        const int j = c.cw.sorted[s];
        const double dx = c.wx[j] - xi, dy = c.wy[j] - yi;
        const double r2 = dx * dx + dy * dy;
        if (!(r2 < k.rc2)) continue;
        const double w = grad_w(k, r2);
        const int phj = c.wph[j];
        riemann(k, dx, dy, r2, w, ui, vi, c.wu[j], c.wv[j], rhoi, k.rho0[phj], pi, c.wp[j], cs, phi, A);
        if (!(phi == SPHNJ_LIQUID && phj == SPHNJ_GAS)) A.col += r2 * w;
      }
    });
    c.fcol[i] = A.col;
    c.fnear[i] = A.near;
    c.fL[4 * i + 0] = A.L00;
    c.fL[4 * i + 1] = A.L01;
    c.fL[4 * i + 2] = A.L10;
    c.fL[4 * i + 3] = A.L11;
    // fluidUpdate (src/fluid.cu:256-287)
    double ax = A.ax, ay = A.ay;
    if (k.steady) {
      ax += -0.1 * ui;
      ay += -0.1 * vi;
    }
    const double dt = c.dt_cur;
    double x = xi + ui * dt, y = yi + vi * dt;
    const double u = ui + (ax + k.gx) * dt, v = vi + (ay + k.gy) * dt;
    double rho = rhoi + A.drho * dt;
    rho = std::max(rho, k.rho_min[phi]);
    const double p = rho_to_p(k, rho, phi);
    x = std::min(std::max(x, k.dxmin), k.dxmax);
    y = std::min(std::max(y, k.dymin), k.dymax);
    nx_[i] = x;
    ny_[i] = y;
    nu_[i] = u;
    nv_[i] = v;
    nr_[i] = rho;
    np_[i] = p;
  }
  c.fx.swap(nx_);
  c.fy.swap(ny_);
  c.fu.swap(nu_);
  c.fv.swap(nv_);
  c.frho.swap(nr_);
  c.fp.swap(np_);
  double m2 = 0.0;  // the speed the next step's size is taken from
  for (int i = 0; i < n; ++i) m2 = std::max(m2, c.fu[i] * c.fu[i] + c.fv[i] * c.fv[i]);
  c.u2_last = m2;
}

// The step size of the step that begins (extension adaptive_cfl; the reference's SimTime
// steps by the fixed dt, src/sim_time.cpp:27-55): dt, or min(dt, C h / (c_max + U),
// S C h / c_s) with U the largest fluid speed after the last update, the last term only
// with solids.
void step_size(Case &c) {
  const Consts &k = c.k;
  double dt = k.dt;
  if (k.cfl > 0.0) {
    dt = std::min(dt, k.cfl_h / (k.c_max + std::sqrt(c.u2_last)));
    if (!c.sx.empty()) dt = std::min(dt, k.dt_solid_cap);
  }
  c.dt_cur = dt;
  c.t_sim += dt;
}

// fluidCorrection (src/fluid.cu:442-459): getMaxSpeed once, then shift_per_step passes of
// correctionCalculate (:316-367) + correctionUpdate (:416-427) / shiftDisplacementCalculate (:369-414).
void fluid_shift(Case &c, unsigned *flags) {
  const Consts &k = c.k;
  const int n = (int)c.fx.size();
  if (n == 0) return;
  double m2 = 0.0;
  for (int i = 0; i < n; ++i) m2 = std::max(m2, c.fu[i] * c.fu[i] + c.fv[i] * c.fv[i]);
  const double umax = std::min(std::sqrt(m2), k.umax_cap);
  std::vector<double> gx(n), gy(n);
  for (int pass = 0; pass < k.K; ++pass) {
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < n; ++i) {
      const double xi = c.fx[i], yi = c.fy[i];
      double Gx = 0, Gy = 0;
      for_cells(k, c.cf.cx[i], c.cf.cy[i], [&](int q) {
        for (int s = c.cf.start[q]; s < c.cf.start[q + 1]; ++s) {
          const int j = c.cf.sorted[s];
          const double dx = c.fx[j] - xi, dy = c.fy[j] - yi, r2 = dx * dx + dy * dy;
          if (r2 < k.rc2) {
            const double w = grad_w(k, r2);
            Gx += dx * w;
            Gy += dy * w;
          }
        }
        for (int s = c.cs.start[q]; s < c.cs.start[q + 1]; ++s) {
          const int j = c.cs.sorted[s];
          const double dx = c.sx[j] - xi, dy = c.sy[j] - yi, r2 = dx * dx + dy * dy;
          if (r2 < k.rc2) {
            const double w = grad_w(k, r2);
            Gx += dx * w;
            Gy += dy * w;
          }
        }
        for (int s = c.cw.start[q]; s < c.cw.start[q + 1]; ++s) {
          const int j = c.cw.sorted[s];
          const double dx = c.wx[j] - xi, dy = c.wy[j] - yi, r2 = dx * dx + dy * dy;
          if (r2 < k.rc2) {
            const double w = grad_w(k, r2);
            Gx += dx * w;
            Gy += dy * w;
          }
        }
      });
      gx[i] = Gx;
      gy[i] = Gy;
    }
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
      const double coef = k.shift_coef * k.h * c.dt_cur;
      const double dxs = -coef * gx[i], dys = -coef * gy[i];
      double sx = dxs, sy = dys;
      int cls;
      bool project = false;
      if (c.fph[i] == SPHNJ_GAS) {
        cls = SPHNJ_INNER;
      } else if (c.fcol[i] < 1.5) {
        if (c.fnear[i] == SPHNJ_NO_PARTICLE) {
          cls = SPHNJ_SPLASH;
          sx = 0.0;
          sy = 0.0;
        } else {
          cls = SPHNJ_FREE_SURFACE;
          project = true;
        }
      } else if (c.fcol[i] < 2.0 && c.fnear[i] == SPHNJ_HAS_NEAR_PARTICLE) {
        cls = SPHNJ_VICINITY;
        project = true;
      } else {
        cls = SPHNJ_INNER;
      }
      if (project) {
        double nx = 0, ny = 0;
        const double g2 = gx[i] * gx[i] + gy[i] * gy[i];
        if (g2 > 0.0) {
          const double g = std::sqrt(g2);
          nx = gx[i] / g;
          ny = gy[i] / g;
        }
        if (std::sqrt(nx * nx + ny * ny) < 0.1) {
          nx = 0.0;
          ny = 0.0;
        }
        sx = (1.0 - nx * nx) * dxs - nx * ny * dys;
        sy = -ny * nx * dxs + (1.0 - ny * ny) * dys;
      }
      c.fcls[i] = cls;
      sx *= umax;
      sy *= umax;
      c.fx[i] += sx;
      c.fy[i] += sy;
      c.fu[i] += c.fL[4 * i + 0] * sx + c.fL[4 * i + 1] * sy;
      c.fv[i] += c.fL[4 * i + 2] * sx + c.fL[4 * i + 3] * sy;
    }
  }
  for (int i = 0; i < n; ++i)
    if (std::isnan(c.fx[i]) || std::isnan(c.fy[i]) || std::isnan(c.frho[i])) *flags |= SPHNJ_ST_NAN;
}

// ---- solid ---------------------------------------------------------------------------------

struct M2 {
  double a, b, c, d;  // [[a b][c d]]
};
inline M2 mul(const M2 &x, const M2 &y) {
  return {x.a * y.a + x.b * y.c, x.a * y.b + x.b * y.d, x.c * y.a + x.d * y.c, x.c * y.b + x.d * y.d};
}
inline M2 tr(const M2 &x) { return {x.a, x.c, x.b, x.d}; }
inline double det(const M2 &x) { return x.a * x.d - x.b * x.c; }
// Eigen's 2x2 inverse: cofactors times 1/det.
inline M2 inv(const M2 &x) {
  const double id = 1.0 / det(x);
  return {x.d * id, -x.b * id, -x.c * id, x.a * id};
}

// plasticElasticFunction (src/solid.cu:80-108)
double plastic_elastic(const Consts &k, double strain, double *max_strain) {
  double ts = strain > -0.5 ? std::sqrt(1.0 + 2.0 * strain) - 1.0 : -1.0;
  if (ts > *max_strain) return k.aux_mod * (ts - *max_strain) - k.yield_stress;
  if (ts > -k.harden_strain) {
    if (*max_strain > ts && k.plastic) *max_strain = ts;
    return -k.yield_stress;
  }
  if (*max_strain > ts && k.plastic) *max_strain = ts;
  const double e = std::fabs(-k.harden_strain - ts);
  return -(k.harden_coef * e * e + k.yield_stress);
}

// solidStressCalculation (src/solid.cu:131-167) + the neo-Hookean extension, for F = I + H.
// The Green strain E = (F^T F - I) / 2 is formed as (H + H^T + H^T H) / 2 and J - 1 as
// tr H + det H: the same values without the cancellation of F^T F - I, which would otherwise
// bury a stiff solid's small strains (|H| ~ 1e-6) under rounding.
M2 stress(const Consts &k, const M2 &H, M2 *M, int mat) {
  const M2 F = {1.0 + H.a, H.b, H.c, 1.0 + H.d};
  const M2 HtH = mul(tr(H), H);
  const M2 E = {H.a + 0.5 * HtH.a, 0.5 * (H.b + H.c + HtH.b), 0.5 * (H.c + H.b + HtH.c), H.d + 0.5 * HtH.d};
  if (mat == SPHNJ_LINEAR_ELASTIC) {  // linearStressCalculation: St Venant-Kirchhoff
    const double trE = E.a + E.d;
    const M2 S = {2.0 * k.mu_s * E.a + k.lambda * trE, 2.0 * k.mu_s * E.b, 2.0 * k.mu_s * E.c,
                  2.0 * k.mu_s * E.d + k.lambda * trE};
    return mul(F, S);
  }
  if (mat == SPHNJ_AUXETIC) {  // auxeticStressCalculation
    const double m = 0.5 * (E.a + E.d);
    const double rad = std::sqrt(0.25 * (E.a - E.d) * (E.a - E.d) + E.b * E.c);
    const double ev[2] = {m + rad, m - rad};
    const double al = 0.5 * std::atan2(E.b + E.c, E.a - E.d);
    const M2 R = {std::cos(al), -std::sin(al), std::sin(al), std::cos(al)};
    const M2 Ri = inv(R);
    M2 Mp = mul(mul(Ri, *M), R);
    double s0 = plastic_elastic(k, ev[0], &Mp.a);
    double s1 = plastic_elastic(k, ev[1], &Mp.d);
    *M = mul(mul(R, Mp), Ri);
    const M2 Sg = {s0, 0.0, 0.0, s1};
    return mul(mul(mul(F, R), Sg), Ri);
  }
  if (mat == SPHNJ_NEO_HOOKEAN) {  // P = mu (F - F^-T) + lambda ln(J) F^-T
    const M2 FiT = tr(inv(F));
    const double lj = std::log1p(H.a + H.d + det(H));  // ln J, J - 1 = tr H + det H
    return {k.mu_s * (F.a - FiT.a) + k.lambda * lj * FiT.a, k.mu_s * (F.b - FiT.b) + k.lambda * lj * FiT.b,
            k.mu_s * (F.c - FiT.c) + k.lambda * lj * FiT.c, k.mu_s * (F.d - FiT.d) + k.lambda * lj * FiT.d};
  }
  return {0, 0, 0, 0};
}

inline double rho_ref(const Consts &k, int mat) { return mat == SPHNJ_AUXETIC ? k.rho_aux : k.rho_s; }

// solidCoupleUpdate (src/solid.cu:355-444) + the reference-list check.
void solid_couple(Case &c, unsigned *flags) {
  const Consts &k = c.k;
  const int n = (int)c.sx.size();
  unsigned fl = 0;
#pragma omp parallel for schedule(static) reduction(| : fl)
  for (int i = 0; i < n; ++i) {
    const double xi = c.sx[i], yi = c.sy[i];
    double nx = 0, ny = 0, P = 0, S = 0;
    for_cells(k, c.cs.cx[i], c.cs.cy[i], [&](int q) {
      for (int s = c.cf.start[q]; s < c.cf.start[q + 1]; ++s) {
        const int j = c.cf.sorted[s];
        const double dx = c.fx[j] - xi, dy = c.fy[j] - yi, r2 = dx * dx + dy * dy;
        if (r2 < k.rc2) {
          const double w = shep_w(k, r2);
          S += w;
          P += c.fp[j] * w;
        }
      }
      for (int s = c.cs.start[q]; s < c.cs.start[q + 1]; ++s) {
        const int j = c.cs.sorted[s];
        const double dx = c.sx[j] - xi, dy = c.sy[j] - yi, r2 = dx * dx + dy * dy;
        if (r2 < k.rc2) {
          const double w = grad_w(k, r2);
          nx -= dx * w;
          ny -= dy * w;
        }
      }
    });
    const double nn = std::sqrt(nx * nx + ny * ny);
    if (nn < 0.1) {
      c.sn[2 * i] = 0.0;
      c.sn[2 * i + 1] = 0.0;
    } else {
      c.sn[2 * i] = nx / nn;
      c.sn[2 * i + 1] = ny / nn;
    }
    c.sp[i] = S != 0.0 ? std::max(P / S, k.p_min) : 0.0;
    for (int e = c.ref_off[i]; e < c.ref_off[i + 1]; ++e) {
      const int j = c.ref_list[e];
      if (std::abs(c.cs.cx[j] - c.cs.cx[i]) > 1 || std::abs(c.cs.cy[j] - c.cs.cy[i]) > 1) fl |= SPHNJ_ST_REFLIST_LEFT;
    }
  }
  *flags |= fl;
}

// solidCalculation (src/solid.cu:473-492): S x (solidDeformationCalculate :169-209,
// solidForceCalculate :307-353, solidUpdate :446-471).
void solid_substeps(Case &c, unsigned *flags) {
  const Consts &k = c.k;
  const int n = (int)c.sx.size();
  if (n == 0) return;
  std::vector<double> nx_(n), ny_(n), nu_(n), nv_(n);
  const double dt_s = c.dt_cur / (double)k.S;  // src/solid.cu:456
  for (int sub = 0; sub < k.S; ++sub) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
      // F~ = sum (x_j - x_i) dX^T w' splits as B + D with B = sum dX dX^T w' and
      // D = sum (d_j - d_i) dX^T w', d = x - X the displacement; with L = B^-1 the deformation
      // gradient F~ L is I + D L, and with det B = 0 (L = I) it is B + D.  H = F - I is kept
      // apart from the identity (see stress()).
      M2 D = {0, 0, 0, 0}, B = {0, 0, 0, 0};
      const double dix = c.sx[i] - c.sX[i], diy = c.sy[i] - c.sY[i];
      for (int e = c.ref_off[i]; e < c.ref_off[i + 1]; ++e) {
        const int j = c.ref_list[e];
        const double dX = c.sX[j] - c.sX[i], dY = c.sY[j] - c.sY[i];
        const double w = grad_w(k, dX * dX + dY * dY);
        const double ux = (c.sx[j] - c.sX[j]) - dix, uy = (c.sy[j] - c.sY[j]) - diy;
        D.a += ux * dX * w;
        D.b += ux * dY * w;
        D.c += uy * dX * w;
        D.d += uy * dY * w;
        B.a += dX * dX * w;
        B.b += dX * dY * w;
        B.c += dY * dX * w;
        B.d += dY * dY * w;
      }
      const bool degenerate = det(B) == 0.0;
      const M2 L = degenerate ? M2{1, 0, 0, 1} : inv(B);
      const M2 H = degenerate ? M2{B.a + D.a - 1.0, B.b + D.b, B.c + D.c, B.d + D.d - 1.0} : mul(D, L);
      const M2 F = {1.0 + H.a, H.b, H.c, 1.0 + H.d};
      M2 Ms = {c.sM[4 * i], c.sM[4 * i + 1], c.sM[4 * i + 2], c.sM[4 * i + 3]};
      const M2 P = stress(k, H, &Ms, c.smat[i]);
      const double *fp = &F.a, *lp = &L.a, *pp = &P.a, *mp = &Ms.a;
      for (int t = 0; t < 4; ++t) {
        c.sF[4 * i + t] = fp[t];
        c.sLc[4 * i + t] = lp[t];
        c.sP[4 * i + t] = pp[t];
        c.sM[4 * i + t] = mp[t];
      }
    }
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
      const double xi = c.sx[i], yi = c.sy[i], ui = c.su[i], vi = c.sv[i];
      const double rr = rho_ref(k, c.smat[i]);
      double cx = 0, cy = 0;  // solidFluidCouple (src/solid.cu:211-255)
      for_cells(k, c.cs.cx[i], c.cs.cy[i], [&](int q) {
        for (int s = c.cf.start[q]; s < c.cf.start[q + 1]; ++s) {
          const int j = c.cf.sorted[s];
          const double dx = c.fx[j] - xi, dy = c.fy[j] - yi, r2 = dx * dx + dy * dy;
          if (!(r2 < k.rc2)) continue;
          const double w = grad_w(k, r2);
          double ex = 0, ey = 0;
          if (r2 > 0.0) {
            const double nrm = std::sqrt(r2);
            ex = dx / nrm;
            ey = dy / nrm;
          }
          const double rhoj = c.frho[j], rhoi = rhoj, pj = c.fp[j], pi = c.sp[i];
          const double cs = k.c[c.fph[j]];
          double ps = (rhoi * pj + rhoj * pi) / (rhoi + rhoj);
          const double du = (ui - c.fu[j]) * ex + (vi - c.fv[j]) * ey;
          if (du >= 0.0) ps += rhoi * rhoj * du / (rhoi + rhoj) * std::min(3.0 * du, cs);
          const double t = -2.0 * ps / rr;
          cx += t * dx * w;
          cy += t * dy * w;
        }
      });
      const M2 Fi = {c.sF[4 * i], c.sF[4 * i + 1], c.sF[4 * i + 2], c.sF[4 * i + 3]};
      const M2 Li = {c.sLc[4 * i], c.sLc[4 * i + 1], c.sLc[4 * i + 2], c.sLc[4 * i + 3]};
      const M2 Pi = {c.sP[4 * i], c.sP[4 * i + 1], c.sP[4 * i + 2], c.sP[4 * i + 3]};
      const M2 PLi = mul(Pi, Li), FLi = mul(Fi, Li);
      double ax = 0, ay = 0;  // solidSolidForce (src/solid.cu:257-305)
      for (int e = c.ref_off[i]; e < c.ref_off[i + 1]; ++e) {
        const int j = c.ref_list[e];
        const double dX = c.sX[j] - c.sX[i], dY = c.sY[j] - c.sY[i];
        const double w = grad_w(k, dX * dX + dY * dY);
        const M2 Pj = {c.sP[4 * j], c.sP[4 * j + 1], c.sP[4 * j + 2], c.sP[4 * j + 3]};
        const M2 Lj = {c.sLc[4 * j], c.sLc[4 * j + 1], c.sLc[4 * j + 2], c.sLc[4 * j + 3]};
        const M2 PLj = mul(Pj, Lj);
        const M2 Sm = {PLj.a + PLi.a, PLj.b + PLi.b, PLj.c + PLi.c, PLj.d + PLi.d};
        ax += (Sm.a * dX + Sm.b * dY) * w / rr;
        ay += (Sm.c * dX + Sm.d * dY) * w / rr;
        const double dx = c.sx[j] - xi, dy = c.sy[j] - yi;
        const double du = c.su[j] - ui, dv = c.sv[j] - vi;
        const double nu = k.h * (du * dx + dv * dy) / (dx * dx + dy * dy + 0.01 * k.h2);
        const double vis = k.alpha * nu * k.c_s - k.beta * nu * nu;
        // F_i.adjoint().transpose() == F_i for real matrices (src/solid.cu:304)
        ax += vis * (FLi.a * dX + FLi.b * dY) * w;
        ay += vis * (FLi.c * dX + FLi.d * dY) * w;
      }
      c.scp[2 * i] = cx;
      c.scp[2 * i + 1] = cy;
      double ax_ = ax + cx, ay_ = ay + cy;
      double x = xi, y = yi, u = ui, v = vi;
      if (c.sbc[i] == SPHNJ_FIXED) {  // solidUpdate
        u = k.fixed_u;
        v = k.fixed_v;
        x += u * dt_s;
        y += v * dt_s;
        ax_ = 0.0;
        ay_ = 0.0;
      } else {
        if (k.steady) {
          ax_ += -0.1 * ui;
          ay_ += -0.1 * vi;
        }
        x += u * dt_s;
        y += v * dt_s;
        u += (ax_ + k.gx) * dt_s;
        v += (ay_ + k.gy) * dt_s;
      }
      c.sa[2 * i] = ax_;
      c.sa[2 * i + 1] = ay_;
      nx_[i] = x;
      ny_[i] = y;
      nu_[i] = u;
      nv_[i] = v;
    }
    c.sx = nx_;
    c.sy = ny_;
    c.su = nu_;
    c.sv = nv_;
  }
  for (int i = 0; i < n; ++i)
    if (std::isnan(c.sx[i]) || std::isnan(c.sy[i])) *flags |= SPHNJ_ST_NAN;
}

void build_reflists(Case &c, unsigned *flags) {
  const Consts &k = c.k;
  const int n = (int)c.sX.size();
  CellIndex ref;
  hash_species(c, c.sX, c.sY, &ref, flags);
  c.ref_off.assign(n + 1, 0);
  c.ref_list.clear();
  for (int i = 0; i < n; ++i) {
    for_cells(k, ref.cx[i], ref.cy[i], [&](int q) {
      for (int s = ref.start[q]; s < ref.start[q + 1]; ++s) {
        const int j = ref.sorted[s];
        const double dX = c.sX[j] - c.sX[i], dY = c.sY[j] - c.sY[i];
        if (dX * dX + dY * dY < k.rc2) c.ref_list.push_back(j);
      }
    });
    c.ref_off[i + 1] = (int)c.ref_list.size();
  }
}

void stage(Ctx &C, int st) {
  for (auto &c : C.cases) {
    switch (st) {
      case SPHNJ_STAGE_CELLS:  // a step begins: its size first
        step_size(c);
        hash_species(c, c.fx, c.fy, &c.cf, &C.flags, true);
        hash_species(c, c.sx, c.sy, &c.cs, &C.flags);
        {
          // the reference's per-cell list holds at most CELL_SIZE = 64 entries
          for (int q = 0; q < c.ncell; ++q) {
            const int tot = (c.cf.start[q + 1] - c.cf.start[q]) + (c.cs.start[q + 1] - c.cs.start[q]) +
                            (c.cw.start[q + 1] - c.cw.start[q]);
            if (tot > 64) C.flags |= SPHNJ_ST_CELL_OVER_64;
          }
        }
        break;
      case SPHNJ_STAGE_COUPLE:
        if (!c.fx.empty() && !c.sx.empty()) solid_couple(c, &C.flags);
        break;
      case SPHNJ_STAGE_WALL:
        wall_update(c);
        break;
      case SPHNJ_STAGE_FLUID:
        fluid_calc(c);
        break;
      case SPHNJ_STAGE_SOLID:
        solid_substeps(c, &C.flags);
        break;
      case SPHNJ_STAGE_SHIFT:
        fluid_shift(c, &C.flags);
        break;
    }
  }
}

template <class T>
void put(T *dst, const std::vector<T> &src) {
  if (dst && !src.empty()) std::memcpy(dst, src.data(), src.size() * sizeof(T));
}
void put_i(int32_t *dst, const std::vector<int> &src) {
  if (dst)
    for (size_t i = 0; i < src.size(); ++i) dst[i] = src[i];
}

}  // namespace

struct oracle_ctx {
  Ctx C;
};

extern "C" {

const char *oracle_last_error(void) { return g_err.c_str(); }

int oracle_set_threads(int32_t n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
#else
  (void)n;
  return 1;
#endif
}

int oracle_create(const sphnj_params *pc, int32_t ncase, oracle_ctx **out) {
  if (!pc || ncase < 1 || !out) return SPHNJ_E_ARG;
  auto *o = new oracle_ctx();
  o->C.cases.resize(ncase);
  int off = 0;
  for (int c = 0; c < ncase; ++c) {
    Case &cs = o->C.cases[c];
    cs.P = pc[c];
    if (!make_consts(pc[c], &cs.k)) {
      g_err = "bad cell box";
      delete o;
      return SPHNJ_E_PARAM;
    }
    cs.cell_off = off;
    cs.ncell = cs.k.nx * cs.k.ny;
    off += cs.ncell;
  }
  *out = o;
  return SPHNJ_OK;
}

int oracle_upload(oracle_ctx *o, int32_t ci, const sphnj_fluid_in *f, const sphnj_solid_in *s,
                  const sphnj_wall_in *w) {
  if (!o || ci < 0 || ci >= (int)o->C.cases.size()) return SPHNJ_E_ARG;
  Case &c = o->C.cases[ci];
  const int nf = f ? f->n : 0, ns = s ? s->n : 0, nw = w ? w->n : 0;
  c.fx.assign(f ? f->x : nullptr, f ? f->x + nf : nullptr);
  c.fy.assign(f ? f->y : nullptr, f ? f->y + nf : nullptr);
  c.fu.assign(f ? f->u : nullptr, f ? f->u + nf : nullptr);
  c.fv.assign(f ? f->v : nullptr, f ? f->v + nf : nullptr);
  c.frho.assign(f ? f->rho : nullptr, f ? f->rho + nf : nullptr);
  c.fp.assign(f ? f->p : nullptr, f ? f->p + nf : nullptr);
  c.fph.assign(nf, 0);
  for (int i = 0; i < nf; ++i) c.fph[i] = f->phase[i];
  c.fcls.assign(nf, SPHNJ_FREE_SURFACE);  // include/host_data.cuh:39
  c.fnear.assign(nf, SPHNJ_NO_PARTICLE);
  c.fcol.assign(nf, 0.0);
  c.fL.assign(4 * (size_t)nf, 0.0);
  c.sx.assign(s ? s->x : nullptr, s ? s->x + ns : nullptr);
  c.sy.assign(s ? s->y : nullptr, s ? s->y + ns : nullptr);
  c.su.assign(s ? s->u : nullptr, s ? s->u + ns : nullptr);
  c.sv.assign(s ? s->v : nullptr, s ? s->v + ns : nullptr);
  c.sX.assign(s ? s->X : nullptr, s ? s->X + ns : nullptr);
  c.sY.assign(s ? s->Y : nullptr, s ? s->Y + ns : nullptr);
  c.smat.assign(ns, 0);
  c.sbc.assign(ns, 0);
  for (int i = 0; i < ns; ++i) {
    c.smat[i] = s->material[i];
    c.sbc[i] = s->bc ? s->bc[i] : 0;
  }
  c.sp.assign(ns, 0.0);
  c.sn.assign(2 * (size_t)ns, 0.0);
  c.sF.assign(4 * (size_t)ns, 0.0);
  c.sLc.assign(4 * (size_t)ns, 0.0);
  c.sP.assign(4 * (size_t)ns, 0.0);
  c.sM.assign(4 * (size_t)ns, 0.0);
  for (int i = 0; i < ns; ++i) {  // include/host_data.cuh:77-83
    c.sF[4 * i] = c.sF[4 * i + 3] = 1.0;
    c.sLc[4 * i] = c.sLc[4 * i + 3] = 1.0;
    c.sM[4 * i] = c.sM[4 * i + 3] = -c.k.yield_strain;
  }
  c.sa.assign(2 * (size_t)ns, 0.0);
  c.scp.assign(2 * (size_t)ns, 0.0);
  c.wx.assign(w ? w->x : nullptr, w ? w->x + nw : nullptr);
  c.wy.assign(w ? w->y : nullptr, w ? w->y + nw : nullptr);
  c.wbc.assign(nw, 0);
  for (int i = 0; i < nw; ++i) c.wbc[i] = w->bc[i];
  c.wu.assign(nw, 0.0);
  c.wv.assign(nw, 0.0);
  c.wp.assign(nw, 0.0);
  c.wrho.assign(nw, c.k.rho0[0]);  // include/host_data.cuh:111
  c.wph.assign(nw, SPHNJ_LIQUID);
  c.wn.assign(2 * (size_t)nw, 0.0);
  c.u2_last = 0.0;
  for (int i = 0; i < nf; ++i) c.u2_last = std::max(c.u2_last, c.fu[i] * c.fu[i] + c.fv[i] * c.fv[i]);
  c.dt_cur = c.k.dt;
  c.t_sim = 0.0;
  c.uploaded = true;
  return SPHNJ_OK;
}

int oracle_init(oracle_ctx *o) {
  if (!o) return SPHNJ_E_ARG;
  for (auto &c : o->C.cases) {
    if (!c.uploaded) {
      g_err = "case not uploaded";
      return SPHNJ_E_STATE;
    }
    hash_species(c, c.wx, c.wy, &c.cw, &o->C.flags);
    wall_normals(c);
    build_reflists(c, &o->C.flags);
    // empty fluid / solid indices until the first CELLS stage
    hash_species(c, c.fx, c.fy, &c.cf, &o->C.flags, true);
    hash_species(c, c.sx, c.sy, &c.cs, &o->C.flags);
  }
  o->C.inited = true;
  return SPHNJ_OK;
}

int oracle_stage(oracle_ctx *o, int32_t st) {
  if (!o || st < SPHNJ_STAGE_CELLS || st > SPHNJ_STAGE_SHIFT) return SPHNJ_E_ARG;
  if (!o->C.inited) return SPHNJ_E_STATE;
  stage(o->C, st);
  return SPHNJ_OK;
}

int oracle_step(oracle_ctx *o, int32_t nsteps) {
  if (!o || nsteps < 0) return SPHNJ_E_ARG;
  if (!o->C.inited) return SPHNJ_E_STATE;
  for (int s = 0; s < nsteps; ++s)
    for (int st = SPHNJ_STAGE_CELLS; st <= SPHNJ_STAGE_SHIFT; ++st) stage(o->C, st);
  return SPHNJ_OK;
}

int oracle_download(oracle_ctx *o, int32_t ci, const sphnj_fluid_out *f, const sphnj_solid_out *s,
                    const sphnj_wall_out *w) {
  if (!o || ci < 0 || ci >= (int)o->C.cases.size()) return SPHNJ_E_ARG;
  const Case &c = o->C.cases[ci];
  if (f) {
    put(f->x, c.fx); put(f->y, c.fy); put(f->u, c.fu); put(f->v, c.fv);
    put(f->rho, c.frho); put(f->p, c.fp);
    put_i(f->phase, c.fph); put_i(f->boundary_type, c.fcls); put_i(f->near, c.fnear);
    put(f->color, c.fcol); put(f->du_dx, c.fL);
  }
  if (s) {
    put(s->x, c.sx); put(s->y, c.sy); put(s->u, c.su); put(s->v, c.sv);
    put(s->X, c.sX); put(s->Y, c.sY);
    put(s->pressure, c.sp); put(s->normal, c.sn);
    put(s->deformation, c.sF); put(s->correction, c.sLc); put(s->stress, c.sP);
    put(s->max_strain, c.sM); put(s->du_dt, c.sa); put(s->couple_dudt, c.scp);
  }
  if (w) {
    put(w->x, c.wx); put(w->y, c.wy); put(w->u, c.wu); put(w->v, c.wv);
    put(w->p, c.wp); put(w->rho, c.wrho); put(w->normal, c.wn);
    put_i(w->phase, c.wph); put_i(w->bc, c.wbc);
  }
  return SPHNJ_OK;
}

int oracle_status(oracle_ctx *o, uint32_t *flags) {
  if (!o || !flags) return SPHNJ_E_ARG;
  *flags = o->C.flags;
  return SPHNJ_OK;
}

int oracle_stage_sorted(oracle_ctx *o, int32_t ci, int32_t species, int32_t *keys, int32_t *orig,
                        int32_t *cell_start) {
  if (!o || ci < 0 || ci >= (int)o->C.cases.size()) return SPHNJ_E_ARG;
  const Case &c = o->C.cases[ci];
  const CellIndex *x = species == SPHNJ_FLUID ? &c.cf : species == SPHNJ_SOLID ? &c.cs : &c.cw;
  const int n = (int)x->sorted.size();
  if (species == SPHNJ_FLUID) {  // the library's slot order: sub-cells, x-major
    for (int s = 0; s < n; ++s) {
      if (keys) keys[s] = 4 * c.cell_off + x->xkey[x->xsorted[s]];
      if (orig) orig[s] = x->xsorted[s];
    }
    if (cell_start)
      for (int q = 0; q <= 4 * c.ncell; ++q) cell_start[q] = x->xstart[q];
    return SPHNJ_OK;
  }
  for (int s = 0; s < n; ++s) {
    if (keys) keys[s] = c.cell_off + x->key[x->sorted[s]];
    if (orig) orig[s] = x->sorted[s];
  }
  if (cell_start)
    for (int q = 0; q <= c.ncell; ++q) cell_start[q] = x->start[q];
  return SPHNJ_OK;
}

int oracle_stage_neighbors(oracle_ctx *o, int32_t ci, int32_t *counts, int32_t *list, int64_t cap, int64_t *total) {
  if (!o || ci < 0 || ci >= (int)o->C.cases.size()) return SPHNJ_E_ARG;
  const Case &c = o->C.cases[ci];
  const Consts &k = c.k;
  int64_t t = 0;
  const int n = (int)c.fx.size();
  for (int i = 0; i < n; ++i) {
    int cnt = 0;
    const double xi = c.fx[i], yi = c.fy[i];
    auto emit = [&](int sp, int j) {
      if (list && t < cap) list[t] = (sp << 28) | j;
      ++t;
      ++cnt;
    };
    for_cells(k, c.cf.cx[i], c.cf.cy[i], [&](int q) {
      for (int s = c.cf.start[q]; s < c.cf.start[q + 1]; ++s) {
        const int j = c.cf.sorted[s];
        const double dx = c.fx[j] - xi, dy = c.fy[j] - yi;
        if (dx * dx + dy * dy < k.rc2) emit(SPHNJ_FLUID, j);
      }
      for (int s = c.cs.start[q]; s < c.cs.start[q + 1]; ++s) {
        const int j = c.cs.sorted[s];
        const double dx = c.sx[j] - xi, dy = c.sy[j] - yi;
        if (dx * dx + dy * dy < k.rc2) emit(SPHNJ_SOLID, j);
      }
      for (int s = c.cw.start[q]; s < c.cw.start[q + 1]; ++s) {
        const int j = c.cw.sorted[s];
        const double dx = c.wx[j] - xi, dy = c.wy[j] - yi;
        if (dx * dx + dy * dy < k.rc2) emit(SPHNJ_WALL, j);
      }
    });
    if (counts) counts[i] = cnt;
  }
  if (total) *total = t;
  return SPHNJ_OK;
}

int oracle_solid_reflist(oracle_ctx *o, int32_t ci, int32_t *offsets, int32_t *list, int64_t cap) {
  if (!o || ci < 0 || ci >= (int)o->C.cases.size()) return SPHNJ_E_ARG;
  const Case &c = o->C.cases[ci];
  put_i(offsets, c.ref_off);
  if (list)
    for (size_t e = 0; e < c.ref_list.size() && (int64_t)e < cap; ++e) list[e] = c.ref_list[e];
  return SPHNJ_OK;
}

int oracle_energy(oracle_ctx *o, int32_t ci, double out[6]) {
  if (!o || !out || ci < 0 || ci >= (int)o->C.cases.size()) return SPHNJ_E_ARG;
  const Case &c = o->C.cases[ci];
  const Consts &k = c.k;
  const double d2 = k.d * k.d;
  double ke = 0, mass = 0, m2 = 0, kes = 0, ep = 0;
  for (size_t i = 0; i < c.fx.size(); ++i) {
    const double u2 = c.fu[i] * c.fu[i] + c.fv[i] * c.fv[i];
    ke += 0.5 * k.rho0[c.fph[i]] * d2 * u2;
    mass += c.frho[i] * d2;
    m2 = std::max(m2, u2);
  }
  for (size_t i = 0; i < c.sx.size(); ++i) {  // src/output.cu:426-455
    const double u2 = c.su[i] * c.su[i] + c.sv[i] * c.sv[i];
    kes += 0.5 * rho_ref(k, c.smat[i]) * d2 * u2;
    const M2 F = {c.sF[4 * i], c.sF[4 * i + 1], c.sF[4 * i + 2], c.sF[4 * i + 3]};
    const M2 Et = mul(tr(F), F);
    const M2 E = {0.5 * (Et.a - 1.0), 0.5 * Et.b, 0.5 * Et.c, 0.5 * (Et.d - 1.0)};
    const double trE = E.a + E.d;
    ep += 0.5 * k.lambda * trE * trE * d2 + k.mu_s * (E.a * E.a + E.b * E.b + E.c * E.c + E.d * E.d) * d2;
  }
  out[0] = ke;
  out[1] = mass;
  out[2] = std::sqrt(m2);
  out[3] = kes;
  out[4] = ep;
  out[5] = (double)c.fx.size();
  return SPHNJ_OK;
}

// Output::outputTotalForce, outputEdge, outputVonMisesStress, outputMeanPressure
// (src/output.cu:283-305, 339-369, 503-544, 546-560) with the probe points as arguments.
int oracle_probe_global(oracle_ctx *o, int32_t ci, double height_x, double out[SPHNJ_PROBE_COUNT]) {
  if (!o || !out || ci < 0 || ci >= (int)o->C.cases.size()) return SPHNJ_E_ARG;
  const Case &c = o->C.cases[ci];
  const Consts &k = c.k;
  const double d = k.d, d2 = d * d;
  double fx = 0, fy = 0, front = 0.0, height = -0.5 * d, vm_el = 0, vm_aux = 0, psum = 0;
  for (size_t i = 0; i < c.fx.size(); ++i) {
    if (c.fph[i] != SPHNJ_LIQUID) continue;
    if (c.fx[i] > front) front = c.fx[i];
    if (c.fx[i] < height_x + 0.5 * d && c.fx[i] > height_x - 0.5 * d && c.fy[i] > height) height = c.fy[i];
  }
  const size_t ns = c.sx.size();
  for (size_t i = 0; i < ns; ++i) {
    const double rho = rho_ref(k, c.smat[i]);
    fx += c.scp[2 * i] * d2 * rho;
    fy += c.scp[2 * i + 1] * d2 * rho;
    const M2 F = {c.sF[4 * i], c.sF[4 * i + 1], c.sF[4 * i + 2], c.sF[4 * i + 3]};
    const M2 P = {c.sP[4 * i], c.sP[4 * i + 1], c.sP[4 * i + 2], c.sP[4 * i + 3]};
    const M2 S = mul(inv(F), P);
    const double m = std::sqrt(0.5 * (S.a - S.d) * (S.a - S.d) + 3.0 * S.b * S.b);
    if (c.smat[i] == SPHNJ_AUXETIC)
      vm_aux = std::max(vm_aux, m);
    else
      vm_el = std::max(vm_el, m);
    psum += c.sp[i];
  }
  out[SPHNJ_PROBE_FORCE_X] = fx;
  out[SPHNJ_PROBE_FORCE_Y] = fy;
  out[SPHNJ_PROBE_FRONT] = front + 0.5 * d;
  out[SPHNJ_PROBE_HEIGHT] = height + 0.5 * d;
  out[SPHNJ_PROBE_VM_ELASTIC] = vm_el;
  out[SPHNJ_PROBE_VM_AUXETIC] = vm_aux;
  out[SPHNJ_PROBE_VM_MAX] = std::max(vm_el, vm_aux);
  out[SPHNJ_PROBE_MEAN_PRESSURE] = ns ? psum / (double)ns : 0.0;
  return SPHNJ_OK;
}

// The point probes of Output::outputTankBeamEnd etc. (src/output.cu:457-478): the first solid
// in index order whose reference position lies within d/2 of (X, Y).
int oracle_probe_solid(oracle_ctx *o, int32_t ci, double X, double Y, double pos[2], int32_t *index) {
  if (!o || !pos || ci < 0 || ci >= (int)o->C.cases.size()) return SPHNJ_E_ARG;
  const Case &c = o->C.cases[ci];
  const double r = 0.5 * c.k.d;
  for (size_t i = 0; i < c.sx.size(); ++i) {
    if (std::hypot(c.sX[i] - X, c.sY[i] - Y) < r) {
      pos[0] = c.sx[i];
      pos[1] = c.sy[i];
      if (index) *index = (int32_t)i;
      return SPHNJ_OK;
    }
  }
  if (index) *index = -1;
  g_err = "no solid particle within d/2 of the probe point";
  return SPHNJ_E_ARG;
}

int oracle_time(oracle_ctx *o, int32_t ci, double *time, double *last_dt) {
  if (!o || ci < 0 || ci >= (int)o->C.cases.size()) return SPHNJ_E_ARG;
  if (time) *time = o->C.cases[ci].t_sim;
  if (last_dt) *last_dt = o->C.cases[ci].dt_cur;
  return SPHNJ_OK;
}

int oracle_derived(oracle_ctx *o, double out[15]) {
  if (!o || !out) return SPHNJ_E_ARG;
  const Consts &k = o->C.cases[0].k;
  const double v[15] = {k.rc2, k.hinv2, k.kdc, k.k2c, k.eR, (double)k.nx, (double)k.ny, k.cdx, k.cdy,
                        k.lambda, k.mu_s, k.c_s, k.yield_strain, k.aux_mod, k.umax_cap};
  std::memcpy(out, v, sizeof(v));
  return SPHNJ_OK;
}

void oracle_destroy(oracle_ctx *o) { delete o; }

}  // extern "C"
