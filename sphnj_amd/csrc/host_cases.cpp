// Case generators: the counterpart of cases/<name>/particle_generate.cpp + cases/<name>/config/
// input.ini in the reference.  Every generator builds a square lattice of particles at cell
// centres ((i + 1/2) dx, (j + 1/2) dx) as the reference generators do
// (cases/dam_break/particle_generate.cpp:75-113), in i-outer (x) / j-inner (y) order, with
// dummy wall layers around the tank, and writes the matching parameter set.
//
// Cases (options are "key=value" pairs separated by ';'):
//   dam_break       liquid column at the left wall of a closed tank, optional air phase.
//                   tank_w tank_h col_w col_h dx hfac layers air cfl steps K shift gas_rho gas_c
//                   c0 bg margin bc
//   gate_fsi        dam break against an elastic gate clamped at its top (Antoci-type).
//                   dx hfac layers tank_w tank_h col_w col_h gate_t gate_h E nu rho_s material
//                   S K clamp_rows margin cfl steps strain
//   wedge_entry     elastic V-wedge falling into water with air above.
//                   dx hfac layers tank_w tank_h depth wedge_w deadrise thick v0 E nu rho_s S K
//                   air gap margin cfl steps
//   ensemble_member member `index` of a seeded ensemble of liquid dam breaks.
//                   seed index layers K steps
//   ref_dam_break, ref_water_tank, ref_beam_vibrate, ref_circle_entry, ref_beam_dam,
//   ref_wedge_entry  the reference's six scenarios (cases/<name>), as its generators and
//                   input.ini files define them.  dx steps; ref_circle_entry and
//                   ref_wedge_entry also height (the body's start above the water),
//                   ref_wedge_entry fixed_v
// Initial pressure is hydrostatic p = rho0 g (level - y) (as in
// cases/water_tank/particle_generate.cpp:96-97), continuous across the air/water surface, and
// the density is its inverse Tait value rho = rho0 ((p - p_bg) B^-1 + 1)^(1/gamma).
#include "host.hpp"

#include <cmath>
#include <cstring>
#include <map>
#include <sstream>
#include <sys/stat.h>

using namespace sphnj;

namespace {

struct Opts {
  std::map<std::string, std::string> kv;
  explicit Opts(const char *s) {
    if (!s) return;
    std::string all(s), item;
    std::istringstream is(all);
    while (std::getline(is, item, ';')) {
      auto e = item.find('=');
      if (e == std::string::npos) continue;
      std::string k = item.substr(0, e), v = item.substr(e + 1);
      auto tr = [](std::string &x) {
        x.erase(0, x.find_first_not_of(" \t"));
        x.erase(x.find_last_not_of(" \t") + 1);
      };
      tr(k);
      tr(v);
      kv[k] = v;
    }
  }
  double d(const char *k, double def) const {
    auto it = kv.find(k);
    return it == kv.end() ? def : std::strtod(it->second.c_str(), nullptr);
  }
  long i(const char *k, long def) const {
    auto it = kv.find(k);
    return it == kv.end() ? def : std::strtol(it->second.c_str(), nullptr, 10);
  }
  unsigned long long u(const char *k, unsigned long long def) const {
    auto it = kv.find(k);
    return it == kv.end() ? def : std::strtoull(it->second.c_str(), nullptr, 10);
  }
};

const double kG = 9.81;

double tait_density(double p, double rho0, double c, double gamma, double bg) {
  const double p2rho = gamma / (c * c) / rho0;
  return rho0 * std::pow((p - bg) * p2rho + 1.0, 1.0 / gamma);
}

void base_params(sphnj_params *P, double dx, double hfac, double tank_w, double tank_h, double margin,
                 int layers) {
  sphnj_params_default(P);
  P->gravity = kG;
  P->particle_diameter = dx;
  P->smoothing_length = hfac * dx;
  P->impact_length_hsml_ratio = 3.0;
  const double m = std::fmax(margin, (layers + 1) * dx);
  P->cell_xmin = -m;
  P->cell_ymin = -m;
  P->cell_xmax = tank_w + m;
  P->cell_ymax = tank_h + m;
  P->domain_xmin = 0.0;
  P->domain_ymin = 0.0;
  P->domain_xmax = tank_w;
  P->domain_ymax = tank_h;
  // the reference rejects a parameter set with neither auxetic_modulus nor yield_strain
  // (src/properties.cpp:150-154), even without solids; give the reference INI's value.
  P->auxetic_modulus = 2.0e8;
  P->istart = 0;
  P->result_interval = 1000000;
  P->sort_every_step = 1;
}

void add_fluid(sphnj_particles *ps, double x, double y, double rho, double p, int phase) {
  ps->fx.push_back(x);
  ps->fy.push_back(y);
  ps->fu.push_back(0.0);
  ps->fv.push_back(0.0);
  ps->frho.push_back(rho);
  ps->fp.push_back(p);
  ps->fphase.push_back(phase);
  if (phase == SPHNJ_LIQUID)
    ps->n_liquid++;
  else
    ps->n_gas++;
}

void add_wall(sphnj_particles *ps, double x, double y, int bc) {
  ps->wx.push_back(x);
  ps->wy.push_back(y);
  ps->wbc.push_back(bc);
}

void add_solid(sphnj_particles *ps, double x, double y, double u, double v, int mat, int bc) {
  ps->sx.push_back(x);
  ps->sy.push_back(y);
  ps->su.push_back(u);
  ps->sv.push_back(v);
  ps->sX.push_back(x);
  ps->sY.push_back(y);
  ps->smat.push_back(mat);
  ps->sbc.push_back(bc);
}

// Liquid column [0,col_w] x [0,col_h] in a tank [0,tank_w] x [0,tank_h], walls all round.
int dam_break(const Opts &o, sphnj_params *P, sphnj_particles *ps) {
  const double tank_w = o.d("tank_w", 4.0), tank_h = o.d("tank_h", 3.0);
  const double col_w = o.d("col_w", 1.0), col_h = o.d("col_h", 2.0);
  const double dx = o.d("dx", 0.02), hfac = o.d("hfac", 1.3);
  const int layers = (int)o.i("layers", 4);
  const int air = (int)o.i("air", 0);
  const double cfl = o.d("cfl", 0.25);
  const double c0 = o.d("c0", 10.0 * std::sqrt(2.0 * kG * col_h));
  const double gas_rho = o.d("gas_rho", 1.0), gas_c = o.d("gas_c", c0), bg = o.d("bg", 0.0);
  const int bc = (int)o.i("bc", SPHNJ_SLIP);
  if (!(dx > 0) || !(tank_w > 0) || !(tank_h > 0) || col_w > tank_w || col_h > tank_h || layers < 0) {
    set_global_error("dam_break: bad geometry options");
    return SPHNJ_E_ARG;
  }
  base_params(P, dx, hfac, tank_w, tank_h, o.d("margin", 0.1), layers);
  P->liquid_density = 1000.0;
  P->liquid_sound_speed = c0;
  P->liquid_gamma = 7.0;
  P->gas_density = gas_rho;
  P->gas_sound_speed = gas_c;
  P->gas_gamma = 1.4;
  P->background_pressure = bg;
  P->particle_shift_coef = o.d("shift", 1.0);
  P->shift_per_step = (int)o.i("K", 10);
  P->solid_sub_step = 1;
  P->dt = cfl * P->smoothing_length / std::fmax(c0, gas_c);
  P->iend = (int)o.i("steps", 2000);

  const int nx = (int)std::lround(tank_w / dx), ny = (int)std::lround(tank_h / dx);
  const int ncx = (int)std::lround(col_w / dx), ncy = (int)std::lround(col_h / dx);
  const double p_surf = air ? gas_rho * kG * (tank_h - col_h) : 0.0;
  for (int i = -layers; i < nx + layers; ++i) {
    for (int j = -layers; j < ny + layers; ++j) {
      const double x = (i + 0.5) * dx, y = (j + 0.5) * dx;
      if (i < 0 || i >= nx || j < 0 || j >= ny) {
        add_wall(ps, x, y, bc);
      } else if (i < ncx && j < ncy) {
        const double p = bg + p_surf + 1000.0 * kG * (col_h - y);
        add_fluid(ps, x, y, tait_density(p, 1000.0, c0, 7.0, bg), p, SPHNJ_LIQUID);
      } else if (air) {
        const double p = bg + gas_rho * kG * (tank_h - y);
        add_fluid(ps, x, y, tait_density(p, gas_rho, gas_c, 1.4, bg), p, SPHNJ_GAS);
      }
    }
  }
  return SPHNJ_OK;
}

// Dam break against an elastic gate: the gate occupies x in [col_w, col_w + gate_t],
// y in [0, gate_h]; a rigid dummy wall of the same thickness continues above it; the
// top clamp_rows rows of the gate are FIXED (u = 0).
int gate_fsi(const Opts &o, sphnj_params *P, sphnj_particles *ps) {
  const double dx = o.d("dx", 2.5e-4), hfac = o.d("hfac", 1.3);
  const int layers = (int)o.i("layers", 4);
  const double tank_w = o.d("tank_w", 0.6), tank_h = o.d("tank_h", 0.4);
  const double col_w = o.d("col_w", 0.1), col_h = o.d("col_h", 0.14);
  const double gate_t = o.d("gate_t", 0.005), gate_h = o.d("gate_h", 0.079);
  const double E = o.d("E", 12.0e6), nu = o.d("nu", 0.4), rho_s = o.d("rho_s", 1100.0);
  const int mat = (int)o.i("material", SPHNJ_LINEAR_ELASTIC);
  const int S = (int)o.i("S", 10), K = (int)o.i("K", 1), clamp_rows = (int)o.i("clamp_rows", 4);
  const double cfl = o.d("cfl", 0.25);
  if (!(dx > 0) || col_w + gate_t > tank_w || gate_h > tank_h || col_h > tank_h || S < 1) {
    set_global_error("gate_fsi: bad geometry options");
    return SPHNJ_E_ARG;
  }
  base_params(P, dx, hfac, tank_w, tank_h, o.d("margin", 0.01), layers);
  const double c0 = 10.0 * std::sqrt(2.0 * kG * col_h);
  P->liquid_density = 1000.0;
  P->liquid_sound_speed = c0;
  P->gas_sound_speed = c0;
  P->gas_density = 1.0;
  P->particle_shift_coef = o.d("shift", 1.0);
  P->shift_per_step = K;
  P->solid_sub_step = S;
  P->solid_density = rho_s;
  P->youngs_modulus = E;
  P->poisson_ratio = nu;
  P->is_2D_strain = (int)o.i("strain", 1);
  const double Ks = E / (3.0 * (1.0 - 2.0 * nu));
  const double cs = std::sqrt(Ks / rho_s);
  const double h = P->smoothing_length;
  P->dt = std::fmin(cfl * h / c0, S * cfl * h / cs);
  P->iend = (int)o.i("steps", 1000);

  const int nx = (int)std::lround(tank_w / dx), ny = (int)std::lround(tank_h / dx);
  const int ic0 = (int)std::lround(col_w / dx), igt = (int)std::lround(gate_t / dx);
  const int jg = (int)std::lround(gate_h / dx), jc = (int)std::lround(col_h / dx);
  for (int i = -layers; i < nx + layers; ++i) {
    for (int j = -layers; j < ny + layers; ++j) {
      const double x = (i + 0.5) * dx, y = (j + 0.5) * dx;
      const bool frame = i < 0 || i >= nx || j < 0 || j >= ny;
      const bool in_gate_col = i >= ic0 && i < ic0 + igt;
      if (frame || (in_gate_col && j >= jg)) {
        add_wall(ps, x, y, SPHNJ_SLIP);
      } else if (in_gate_col) {
        add_solid(ps, x, y, 0.0, 0.0, mat, j >= jg - clamp_rows ? SPHNJ_FIXED : SPHNJ_UNCONSTRAINED);
      } else if (i < ic0 && j < jc) {
        const double p = 1000.0 * kG * (col_h - y);
        add_fluid(ps, x, y, tait_density(p, 1000.0, c0, 7.0, 0.0), p, SPHNJ_LIQUID);
      }
    }
  }
  return SPHNJ_OK;
}

// Elastic V-wedge (deadrise angle b, width w, wall thickness t measured normal to the hull)
// with its apex `gap` above still water of depth `depth`, moving down at v0.
int wedge_entry(const Opts &o, sphnj_params *P, sphnj_particles *ps) {
  const double dx = o.d("dx", 1.0e-3), hfac = o.d("hfac", 1.3);
  const int layers = (int)o.i("layers", 4);
  const double tank_w = o.d("tank_w", 3.0), tank_h = o.d("tank_h", 1.5), depth = o.d("depth", 1.0);
  const double ww = o.d("wedge_w", 0.6), beta = o.d("deadrise", 30.0) * M_PI / 180.0;
  const double thick = o.d("thick", 0.012), v0 = o.d("v0", 6.0), gap = o.d("gap", 2.0 * dx);
  const double E = o.d("E", 70.0e9), nu = o.d("nu", 0.3), rho_s = o.d("rho_s", 2700.0);
  const int S = (int)o.i("S", 10), K = (int)o.i("K", 1), air = (int)o.i("air", 1);
  const double cfl = o.d("cfl", 0.25);
  const double xc = 0.5 * tank_w, ya = depth + gap;
  const double ytop = ya + 0.5 * ww * std::tan(beta) + thick / std::cos(beta);
  if (!(dx > 0) || ww > tank_w || ytop > tank_h || depth > tank_h || S < 1) {
    set_global_error("wedge_entry: bad geometry options");
    return SPHNJ_E_ARG;
  }
  base_params(P, dx, hfac, tank_w, tank_h, o.d("margin", 0.1), layers);
  const double c0 = 10.0 * std::sqrt(2.0 * kG * depth);
  const double gas_rho = 1.0;
  P->liquid_density = 1000.0;
  P->liquid_sound_speed = c0;
  P->gas_density = gas_rho;
  P->gas_sound_speed = c0;
  P->gas_gamma = 1.4;
  P->particle_shift_coef = o.d("shift", 1.0);
  P->shift_per_step = K;
  P->solid_sub_step = S;
  P->solid_density = rho_s;
  P->youngs_modulus = E;
  P->poisson_ratio = nu;
  P->is_2D_strain = 1;
  const double cs = std::sqrt(E / (3.0 * (1.0 - 2.0 * nu)) / rho_s);
  const double h = P->smoothing_length;
  P->dt = std::fmin(cfl * h / c0, S * cfl * h / cs);
  P->iend = (int)o.i("steps", 2000);

  const int nx = (int)std::lround(tank_w / dx), ny = (int)std::lround(tank_h / dx);
  const int nd = (int)std::lround(depth / dx);
  const double tv = thick / std::cos(beta);  // vertical thickness of the hull
  const double p_surf = air ? gas_rho * kG * (tank_h - depth) : 0.0;
  for (int i = -layers; i < nx + layers; ++i) {
    for (int j = -layers; j < ny + layers; ++j) {
      const double x = (i + 0.5) * dx, y = (j + 0.5) * dx;
      if (i < 0 || i >= nx || j < 0 || j >= ny) {
        add_wall(ps, x, y, SPHNJ_SLIP);
        continue;
      }
      const double ax = std::fabs(x - xc);
      const double yo = ya + ax * std::tan(beta);
      if (ax <= 0.5 * ww && y >= yo && y < yo + tv) {
        add_solid(ps, x, y, 0.0, -v0, SPHNJ_LINEAR_ELASTIC, SPHNJ_UNCONSTRAINED);
      } else if (j < nd) {
        const double p = p_surf + 1000.0 * kG * (depth - y);
        add_fluid(ps, x, y, tait_density(p, 1000.0, c0, 7.0, 0.0), p, SPHNJ_LIQUID);
      } else if (air) {
        const double p = gas_rho * kG * (tank_h - y);
        add_fluid(ps, x, y, tait_density(p, gas_rho, c0, 1.4, 0.0), p, SPHNJ_GAS);
      }
    }
  }
  return SPHNJ_OK;
}

uint64_t splitmix64(uint64_t *s) {
  uint64_t z = (*s += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
double unif(uint64_t *s) { return (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0); }

// Ensemble member: a SplitMix64 stream per member index drawn from the master seed.
//   N ~ round(exp U[ln 5e3, ln 2e5]); aspect a ~ U[0.5, 2]; column H = 1, W = H / a;
//   dx = sqrt(W H / N); tank width ~ U[2, 5] W; tank height 1.5 H; c0 = 10 sqrt(2 g H).
int ensemble_member(const Opts &o, sphnj_params *P, sphnj_particles *ps) {
  const unsigned long long seed = o.u("seed", 20261202ULL);
  const long index = o.i("index", 0);
  uint64_t st = (uint64_t)seed ^ (0xD1B54A32D192ED03ULL * (uint64_t)(index + 1));
  const double N = std::round(std::exp(std::log(5.0e3) + unif(&st) * (std::log(2.0e5) - std::log(5.0e3))));
  const double a = 0.5 + 1.5 * unif(&st);
  const double H = 1.0, W = H / a;
  const double dx = std::sqrt(W * H / N);
  const double tw_raw = (2.0 + 3.0 * unif(&st)) * W;
  // tank edges on the particle lattice
  const double tank_w = std::lround(tw_raw / dx) * dx, tank_h = std::lround(1.5 * H / dx) * dx;
  const double col_w = std::lround(W / dx) * dx, col_h = std::lround(H / dx) * dx;
  std::ostringstream os;
  os.precision(17);
  os << "tank_w=" << tank_w << ";tank_h=" << tank_h << ";col_w=" << col_w << ";col_h=" << col_h << ";dx=" << dx
     << ";layers=" << o.i("layers", 4) << ";K=" << o.i("K", 10) << ";steps=" << o.i("steps", 1000)
     << ";c0=" << 10.0 * std::sqrt(2.0 * kG * H);
  Opts sub(os.str().c_str());
  return dam_break(sub, P, ps);
}

// ---- the reference's own scenarios (cases/<name>/particle_generate.cpp + config/input.ini) -----
// Each reproduces the reference generator's lattice, loop order (which fixes the original
// index of every particle), phases, materials, initial fields and boundary flags, and the
// parameters of its input.ini.  Positions are exact doubles (the reference writes its solid
// and wall files with 6 significant digits).  `dx` rescales a case (h, and dt at a fixed
// acoustic CFL number, follow it); `steps` sets iend.  Where a shipped input.ini does not fit
// its generator, the deviation is named at the generator.

// The parameters every reference input.ini shares: defaults, then the [solid]/[auxetic]
// block the INIs repeat (auxetic_modulus 2e8 keeps the reference's validation satisfied).
void ref_base(sphnj_params *P) {
  sphnj_params_default(P);
  P->auxetic_density = 636.0;
  P->auxetic_modulus = 2.0e8;
  P->auxetic_poisson_ratio = -0.2;
  P->yield_stress = 0.7e7;
  P->arti_vis_alpha = 2.5;
  P->arti_vis_beta = 2.5;
  P->impact_length_hsml_ratio = 3.0;
  P->thread_num = 32;
  P->istart = 0;
}
void ref_fluids(sphnj_params *P, double rl, double cl, double rg, double cg) {
  P->liquid_density = rl;
  P->liquid_sound_speed = cl;
  P->liquid_gamma = 7.0;
  P->liquid_viscosity = 1.01e-3;
  P->gas_density = rg;
  P->gas_sound_speed = cg;
  P->gas_gamma = 1.4;
  P->gas_viscosity = 17.9e-6;
}
void ref_boxes(sphnj_params *P, double cx0, double cx1, double cy0, double cy1, double dx0, double dx1, double dy0,
               double dy1) {
  P->cell_xmin = cx0;
  P->cell_xmax = cx1;
  P->cell_ymin = cy0;
  P->cell_ymax = cy1;
  P->domain_xmin = dx0;
  P->domain_xmax = dx1;
  P->domain_ymin = dy0;
  P->domain_ymax = dy1;
}
// h = 1.2 dx (every reference INI); dt scaled from the INI's dt at its own h
void ref_kernel_time(sphnj_params *P, const Opts &o, double dx_ref, double dt_ref, double h_ref_ini, int iend,
                     int result, int S, int K) {
  const double dx = o.d("dx", dx_ref);
  P->particle_diameter = dx;
  P->smoothing_length = 1.2 * dx;
  P->dt = dt_ref * (1.2 * dx) / h_ref_ini;
  P->iend = (int)o.i("steps", iend);
  P->result_interval = result;
  P->sort_every_step = 50;
  P->solid_sub_step = S;
  P->shift_per_step = K;
}

// A case run at a coarser spacing than its reference keeps its wall layers, which can then
// reach past the INI's cell box: widen the box to hold every particle with a spacing to
// spare (at the reference spacing every INI box already does, and is kept as it is).
void fit_cell_box(sphnj_params *P, const sphnj_particles *ps) {
  const double m = P->particle_diameter;
  auto widen = [&](const std::vector<double> &x, const std::vector<double> &y) {
    for (size_t i = 0; i < x.size(); ++i) {
      P->cell_xmin = std::fmin(P->cell_xmin, x[i] - m);
      P->cell_xmax = std::fmax(P->cell_xmax, x[i] + m);
      P->cell_ymin = std::fmin(P->cell_ymin, y[i] - m);
      P->cell_ymax = std::fmax(P->cell_ymax, y[i] + m);
    }
  };
  widen(ps->fx, ps->fy);
  widen(ps->sx, ps->sy);
  widen(ps->wx, ps->wy);
}

// cases/dam_break: liquid block 2.0 x 1.0 in a 5.366 x 2.5 tank, gas elsewhere, both at
// p = 0 and rho = rho0; 8 wall layers (particle_generate.cpp:57-113, input.ini).
int ref_dam_break(const Opts &o, sphnj_params *P, sphnj_particles *ps) {
  ref_base(P);
  P->gravity = 1.0;
  P->particle_shift_coef = 1.0;
  P->min_pressure = -1.013e5;
  ref_fluids(P, 1.0, 20.0, 1.0e-3, 20.0);
  ref_boxes(P, -0.2, 5.57, -0.2, 2.7, 0.0, 5.37, 0.0, 2.5);
  ref_kernel_time(P, o, 5.0e-3, 5.0e-5, 0.006, 200000, 2000, 1, 10);
  const double dx = P->particle_diameter, xmax = 5.366, ymax = 2.5;
  const int L = 8, nx = (int)(xmax / dx), ny = (int)(ymax / dx);
  for (int i = -L; i < nx + L; ++i)
    for (int j = -L; j < ny + L; ++j) {
      const double x = (i + 0.5) * dx, y = (j + 0.5) * dx;
      if (x < 0.0 || x > xmax || y < 0.0 || y > ymax)
        add_wall(ps, x, y, SPHNJ_SLIP);
      else if (x < 2.0 && y < 1.0)
        add_fluid(ps, x, y, 1.0, 0.0, SPHNJ_LIQUID);
      else
        add_fluid(ps, x, y, 1.0e-3, 0.0, SPHNJ_GAS);
    }
  if (P->particle_diameter != 5.0e-3) fit_cell_box(P, ps);
  return SPHNJ_OK;
}

// cases/water_tank: a 1 x 2 m water column at hydrostatic pressure (rho0 g (2 - y), Tait
// density at c = 200) on an elastic floor plate 0.05 m thick that spans the side walls, all
// particles of a column i generated together (particle_generate.cpp:57-111).  The plate's
// rows are plain 7-column solids (UNCONSTRAINED), as the reference writes them.
int ref_water_tank(const Opts &o, sphnj_params *P, sphnj_particles *ps) {
  ref_base(P);
  P->gravity = 9.81;
  P->particle_shift_coef = 1.0;
  P->is_viscid = 1;
  ref_fluids(P, 1000.0, 200.0, 1.29, 200.0);
  P->liquid_viscosity = 1.0;
  P->solid_density = 2700.0;
  P->youngs_modulus = 67.5e9;
  P->poisson_ratio = 0.34;
  ref_boxes(P, -0.1, 1.1, -0.1, 2.1, 0.0, 1.0, 0.0, 2.0);
  ref_kernel_time(P, o, 2.5e-3, 2.0e-6, 0.003, 500000, 500, 20, 10);
  const double dx = P->particle_diameter, xmax = 1.0, ymax = 2.0;
  const int L = 8, nx = (int)std::lround(xmax / dx), ny = (int)std::lround(ymax / dx);
  const int nys = (int)std::lround(0.05 / dx);
  const double rho0 = 1000.0, c0 = 200.0;
  for (int i = -L; i < nx + L; ++i) {
    for (int j = 0; j < ny + L; ++j) {
      const double x = (i + 0.5) * dx, y = (j + 0.5) * dx;
      if (!(y > 0.0)) continue;
      if (x < 0.0 || x > xmax || y > ymax) {
        add_wall(ps, x, y, SPHNJ_SLIP);
      } else {
        const double p = rho0 * kG * (2.0 - y);
        add_fluid(ps, x, y, tait_density(p, rho0, c0, 7.0, 0.0), p, SPHNJ_LIQUID);
      }
    }
    for (int j = 0; j < nys; ++j) add_solid(ps, (i + 0.5) * dx, -(j + 0.5) * dx, 0.0, 0.0, SPHNJ_LINEAR_ELASTIC, 0);
  }
  if (P->particle_diameter != 2.5e-3) fit_cell_box(P, ps);
  return SPHNJ_OK;
}

// cases/beam_vibrate: a free 0.244 x 0.024 m beam (4 columns behind x = 0) with the initial
// transverse velocity of a cantilever's first mode, v(x) = xi c_s f(x) / f(L), xi = 0.01,
// c_s = sqrt(K / rho), f the mode shape at kL = 1.875 (particle_generate.cpp:56-96); no fluid,
// no walls.  The reference writes 7-column solids, so no particle is held.
int ref_beam_vibrate(const Opts &o, sphnj_params *P, sphnj_particles *ps) {
  ref_base(P);
  P->gravity = 0.0;
  P->particle_shift_coef = 1.0;
  ref_fluids(P, 1.0, 20.0, 1.0e-3, 20.0);
  P->solid_density = 1000.0;
  P->youngs_modulus = 2.0e6;
  P->poisson_ratio = 0.4;
  ref_boxes(P, -0.1, 0.3, -0.05, 0.05, 0.0, 0.24, -0.012, 0.012);
  ref_kernel_time(P, o, 1.0e-3, 1.0e-5, 0.0012, 100000, 100, 1, 1);
  const double dx = P->particle_diameter, L_beam = 0.24, ymin = -0.012;
  const int L = 4, nx = (int)std::lround(L_beam / dx), ny = (int)std::lround(0.024 / dx);
  const double kL = 1.875, k = kL / L_beam;
  auto mode = [&](double x) {
    return (std::cos(kL) + std::cosh(kL)) * (std::cosh(k * x) - std::cos(k * x)) +
           (std::sin(kL) - std::sinh(kL)) * (std::sinh(k * x) - std::sin(k * x));
  };
  const double cs = std::sqrt(2.0e6 / (3.0 * (1.0 - 2.0 * 0.4)) / 1000.0);
  const double fL = mode(L_beam);
  for (int i = -L; i < nx; ++i)
    for (int j = 0; j < ny; ++j) {
      const double x = (i + 0.5) * dx, y = (j + 0.5) * dx + ymin;
      add_solid(ps, x, y, 0.0, 0.01 * cs * mode(x) / fL, SPHNJ_LINEAR_ELASTIC, 0);
    }
  if (P->particle_diameter != 1.0e-3) fit_cell_box(P, ps);
  return SPHNJ_OK;
}

// cases/circle_entry: a ring (radius 0.05 m; layers of 4, 10 and 4 dx from the inside out:
// elastic, auxetic, elastic) falling at 50 m/s into water 0.3 m deep under air, tank 0.3 x
// 0.5 m (particle_generate.cpp:56-211, with its pi = 3.1415926).  Its input.ini is the wedge's
// (particle_diameter 0.01, smoothing_length 0.012, dt 1e-6, no auxetic block) while the
// generator places particles 1e-3 apart; here h = 1.2 dx, dt keeps the INI's dt c / h = 0.125,
// and the auxetic block is the other INIs'.
int ref_circle_entry(const Opts &o, sphnj_params *P, sphnj_particles *ps) {
  ref_base(P);
  P->gravity = 9.81;
  P->particle_shift_coef = 2.0;
  ref_fluids(P, 1000.0, 1500.0, 1.27, 340.0);
  P->solid_density = 2700.0;
  P->youngs_modulus = 67.5e9;
  P->poisson_ratio = 0.34;
  ref_boxes(P, -0.1, 3.1, -0.1, 3.6, 0.0, 3.0, 0.0, 3.5);
  ref_kernel_time(P, o, 1.0e-3, 1.0e-6, 0.012, 10000000, 1000, 10, 10);
  const double Pi = 3.1415926, dx = P->particle_diameter;
  const double xmax = 0.3, ymax = 0.5, radius = 0.05, middle = 0.15, level = 0.3;
  const double shift = level + o.d("height", 0.1);  // ring centre above the water (0.1 m)
  const int L = 8, nx = (int)std::lround(xmax / dx), ny = (int)std::lround(ymax / dx);
  for (int i = -L; i < nx + L; ++i)
    for (int j = -L; j < ny + L; ++j) {
      const double x = (i + 0.5) * dx, y = (j + 0.5) * dx;
      if (x < 0.0 || x > xmax || y < 0.0 || y > ymax) {
        add_wall(ps, x, y, SPHNJ_SLIP);
        continue;
      }
      if (std::hypot(x - middle, y - shift) <= radius + dx) continue;
      if (y < level)
        add_fluid(ps, x, y, 1000.0, 0.0, SPHNJ_LIQUID);
      else
        add_fluid(ps, x, y, 1.27, 0.0, SPHNJ_GAS);
    }
  const double t_in = 0.004, t_mid = 0.01, t_out = 0.004, total = t_in + t_mid + t_out;
  auto ring = [&](double r0, double thick, int mat) {
    const int layers = (int)std::lround(thick / dx);
    for (int i = 0; i < layers; ++i) {
      const double r = r0 + (i + 0.5) * dx;
      const int n = (int)std::lround(2.0 * Pi * r / dx);
      const double dth = 2.0 * Pi * r / (double)n;
      for (int j = 0; j < n; ++j) {
        const double th = (double)j * dth / r;
        add_solid(ps, middle + r * std::cos(th), shift + r * std::sin(th), 0.0, -50.0, mat, 0);
      }
    }
  };
  ring(radius - total, t_in, SPHNJ_LINEAR_ELASTIC);
  ring(radius - t_out - t_mid, t_mid, SPHNJ_AUXETIC);
  ring(radius - t_out, t_out, SPHNJ_LINEAR_ELASTIC);
  if (P->particle_diameter != 1.0e-3) fit_cell_box(P, ps);
  return SPHNJ_OK;
}

// cases/beam_dam: water 0.1 m wide and 0.14 m deep at rest (p = 0, rho = rho0) right of an
// elastic plate (x in [-0.005, 0], y in [0, 0.15]), air elsewhere in the 0.4 x 0.2 m tank
// (particle_generate.cpp:57-135).  The plate is written as 7-column solids (UNCONSTRAINED).
int ref_beam_dam(const Opts &o, sphnj_params *P, sphnj_particles *ps) {
  ref_base(P);
  P->gravity = 9.81;
  P->particle_shift_coef = 1.0;
  ref_fluids(P, 1000.0, 20.0, 1.0, 20.0);
  P->solid_density = 1100.0;
  P->youngs_modulus = 7.8e6;
  P->poisson_ratio = 0.47;
  ref_boxes(P, -0.31, 0.11, -0.01, 0.21, -0.3, 0.1, 0.0, 0.2);
  ref_kernel_time(P, o, 1.25e-3, 5.0e-6, 0.0015, 1000000, 1000, 10, 1);
  const double dx = P->particle_diameter, xmin = -0.3, xmax = 0.1, ymin = 0.0, ymax = 0.2;
  const double sl = -0.005, sr = 0.0, su = 0.15, sd = 0.0, level = 0.14;
  const int L = 8, nx = (int)std::lround((xmax - xmin) / dx), ny = (int)std::lround((ymax - ymin) / dx);
  for (int i = -L; i < nx + L; ++i)
    for (int j = -L; j < ny + L; ++j) {
      const double x = (i + 0.5) * dx + xmin, y = (j + 0.5) * dx + ymin;
      if (!(x < sl || x > sr || y < sd || y > su)) continue;
      if (x < xmin || x > xmax || y < ymin || y > ymax)
        add_wall(ps, x, y, SPHNJ_SLIP);
      else if (x > sr && y < level)
        add_fluid(ps, x, y, 1000.0, 0.0, SPHNJ_LIQUID);
      else
        add_fluid(ps, x, y, 1.0, 0.0, SPHNJ_GAS);
    }
  const int nxs = (int)std::lround((sr - sl) / dx), nys = (int)std::lround((su - sd) / dx);
  for (int i = 0; i < nxs; ++i)
    for (int j = 0; j < nys; ++j)
      add_solid(ps, (i + 0.5) * dx + sl, (j + 0.5) * dx + sd, 0.0, 0.0, SPHNJ_LINEAR_ELASTIC, 0);
  if (P->particle_diameter != 1.25e-3) fit_cell_box(P, ps);
  return SPHNJ_OK;
}

// cases/wedge_entry: a 10 deg wedge 1.2 m wide, 0.08 m thick (its lower 0.04 m auxetic, the
// rest aluminium) whose apex starts 0.5 m above water 2.0 m deep, at 30 m/s, in a 3 x 3.5 m
// tank; the aluminium's two end columns and its top centre particles are FIXED
// (particle_generate.cpp:62-203).  The input.ini (particle_diameter 0.01, smoothing_length
// 0.012) does not fit the generator's 2.5e-3 spacing: h = 1.2 dx, and dt keeps the INI's
// dt c / h = 0.125.  The FIXED particles move at solid_fixed_speed_y = fixed_v (default -30):
// the INI leaves it at 0, while the wedge_disp probe (src/output.cu:307-337) subtracts a
// prescribed descent of 30 m/s.
int ref_wedge_entry(const Opts &o, sphnj_params *P, sphnj_particles *ps) {
  ref_base(P);
  P->gravity = 9.81;
  P->particle_shift_coef = 1.0;
  ref_fluids(P, 1000.0, 1500.0, 1.27, 340.0);
  P->solid_density = 2700.0;
  P->youngs_modulus = 67.5e9;
  P->poisson_ratio = 0.34;
  P->solid_fixed_speed_y = o.d("fixed_v", -30.0);
  ref_boxes(P, -0.1, 3.1, -0.1, 3.6, 0.0, 3.0, 0.0, 3.5);
  ref_kernel_time(P, o, 2.5e-3, 1.0e-6, 0.012, 40000, 100, 10, 1);
  const double Pi = 3.1415926, dx = P->particle_diameter;
  const double xmax = 3.0, ymax = 3.5, thick = 0.08, half_L = 0.6, middle = 1.5, level = 2.0;
  const double shift = level + o.d("height", 0.5);  // apex above the water (0.5 m)
  const double ta = std::tan(10.0 * Pi / 180.0);
  auto in_wedge = [&](double x, double y) {
    if (y < shift || x < middle - half_L || x > middle + half_L) return false;
    if (y > ta * half_L + shift + thick) return false;
    const double yb = ta * std::fabs(x - middle) + shift;
    return y > yb && y < yb + thick;
  };
  const int L = 8, nx = (int)std::lround(xmax / dx), ny = (int)std::lround(ymax / dx);
  for (int i = -L; i < nx + L; ++i)
    for (int j = -L; j < ny + L; ++j) {
      const double x = (i + 0.5) * dx, y = (j + 0.5) * dx;
      if (x < 0.0 || x > xmax || y < 0.0 || y > ymax) {
        add_wall(ps, x, y, SPHNJ_SLIP);
        continue;
      }
      if (in_wedge(x, y)) continue;
      if (y < level)
        add_fluid(ps, x, y, 1000.0, 0.0, SPHNJ_LIQUID);
      else
        add_fluid(ps, x, y, 1.27, 0.0, SPHNJ_GAS);
    }
  const int nxs = (int)std::lround(2.0 * half_L / dx), nys = (int)std::lround(thick / dx);
  for (int i = 0; i < nxs; ++i)
    for (int j = 0; j < nys; ++j) {
      const double x = (i + 0.5) * dx + middle - half_L;
      const double y = ta * std::fabs(x - middle) + shift + (j + 0.5) * dx;
      const int mat = j * dx < 0.04 ? SPHNJ_AUXETIC : SPHNJ_LINEAR_ELASTIC;
      int bc = SPHNJ_UNCONSTRAINED;
      if (mat == SPHNJ_LINEAR_ELASTIC) {
        if (x < middle - half_L + dx || x > middle + half_L - dx) bc = SPHNJ_FIXED;
        if (x < middle + dx && x > middle - dx && j == nys - 1) bc = SPHNJ_FIXED;
      }
      add_solid(ps, x, y, 0.0, -30.0, mat, bc);
    }
  if (P->particle_diameter != 2.5e-3) fit_cell_box(P, ps);
  return SPHNJ_OK;
}

int mkdir_p(const std::string &d) {
  std::string acc;
  std::istringstream is(d);
  std::string part;
  if (!d.empty() && d[0] == '/') acc = "/";
  while (std::getline(is, part, '/')) {
    if (part.empty()) continue;
    acc += part + "/";
    mkdir(acc.c_str(), 0755);
  }
  struct stat st;
  return stat(d.c_str(), &st) == 0 ? 0 : -1;
}

}  // namespace

extern "C" {

int sphnj_case_create(const char *name, const char *options, sphnj_params *params, sphnj_particles **out) {
  if (!name || !params || !out) return SPHNJ_E_ARG;
  *out = nullptr;
  Opts o(options);
  auto *ps = new sphnj_particles();
  int rc;
  const std::string n(name);
  if (n == "dam_break")
    rc = dam_break(o, params, ps);
  else if (n == "gate_fsi")
    rc = gate_fsi(o, params, ps);
  else if (n == "wedge_entry")
    rc = wedge_entry(o, params, ps);
  else if (n == "ensemble_member")
    rc = ensemble_member(o, params, ps);
  else if (n == "ref_dam_break")
    rc = ref_dam_break(o, params, ps);
  else if (n == "ref_water_tank")
    rc = ref_water_tank(o, params, ps);
  else if (n == "ref_beam_vibrate")
    rc = ref_beam_vibrate(o, params, ps);
  else if (n == "ref_circle_entry")
    rc = ref_circle_entry(o, params, ps);
  else if (n == "ref_beam_dam")
    rc = ref_beam_dam(o, params, ps);
  else if (n == "ref_wedge_entry")
    rc = ref_wedge_entry(o, params, ps);
  else {
    set_global_error("unknown case " + n);
    rc = SPHNJ_E_ARG;
  }
  if (rc != SPHNJ_OK) {
    delete ps;
    return rc;
  }
  *out = ps;
  return SPHNJ_OK;
}

int sphnj_case_write(const sphnj_params *P, const sphnj_particles *ps, const char *dir) {
  if (!P || !ps || !dir) return SPHNJ_E_ARG;
  const std::string d(dir);
  if (mkdir_p(d + "/config") || mkdir_p(d + "/data")) {
    set_global_error("cannot create " + d);
    return SPHNJ_E_IO;
  }
  FILE *f = std::fopen((d + "/config/input.ini").c_str(), "w");
  if (!f) {
    set_global_error("cannot write input.ini");
    return SPHNJ_E_IO;
  }
  auto D = [&](const char *k, double v) { std::fprintf(f, "%s = %.17g\n", k, v); };
  auto I = [&](const char *k, long v) { std::fprintf(f, "%s = %ld\n", k, v); };
  auto B = [&](const char *k, int v) { std::fprintf(f, "%s = %s\n", k, v ? "true" : "false"); };
  std::fprintf(f, "[physics]\n");
  D("gravity", P->gravity);
  B("is_steady", P->is_steady);
  std::fprintf(f, "\n[fluid]\n");
  D("particle_shift_coef", P->particle_shift_coef);
  B("is_viscid", P->is_viscid);
  D("min_pressure", P->min_pressure);
  D("liquid_density", P->liquid_density);
  D("liquid_sound_speed", P->liquid_sound_speed);
  D("liquid_gamma", P->liquid_gamma);
  D("liquid_viscosity", P->liquid_viscosity);
  D("gas_density", P->gas_density);
  D("gas_sound_speed", P->gas_sound_speed);
  D("gas_gamma", P->gas_gamma);
  D("gas_viscosity", P->gas_viscosity);
  D("background_pressure", P->background_pressure);
  std::fprintf(f, "\n[solid]\n");
  D("solid_density", P->solid_density);
  D("youngs_modulus", P->youngs_modulus);
  D("poisson_ratio", P->poisson_ratio);
  D("arti_vis_alpha", P->arti_vis_alpha);
  D("arti_vis_beta", P->arti_vis_beta);
  D("solid_fixed_speed_x", P->solid_fixed_speed_x);
  D("solid_fixed_speed_y", P->solid_fixed_speed_y);
  B("is_2D_strain", P->is_2D_strain);
  B("is_plastic", P->is_plastic);
  std::fprintf(f, "\n[auxetic]\n");
  D("auxetic_density", P->auxetic_density);
  D("auxetic_modulus", P->auxetic_modulus);
  D("auxetic_poisson_ratio", P->auxetic_poisson_ratio);
  D("yield_strain", P->yield_strain);
  D("yield_stress", P->yield_stress);
  D("harden_strain", P->harden_strain);
  D("harden_coef", P->harden_coef);
  std::fprintf(f, "\n[domain]\n");
  D("cell_xmin", P->cell_xmin);
  D("cell_xmax", P->cell_xmax);
  D("cell_ymin", P->cell_ymin);
  D("cell_ymax", P->cell_ymax);
  D("domain_xmin", P->domain_xmin);
  D("domain_xmax", P->domain_xmax);
  D("domain_ymin", P->domain_ymin);
  D("domain_ymax", P->domain_ymax);
  std::fprintf(f, "\n[time]\n");
  D("dt", P->dt);
  I("istart", P->istart);
  I("iend", P->iend);
  I("result_interval", P->result_interval);
  I("sort_every_step", P->sort_every_step);
  I("solid_sub_step", P->solid_sub_step);
  I("shift_per_step", P->shift_per_step);
  std::fprintf(f, "\n[kernel]\n");
  D("smoothing_length", P->smoothing_length);
  D("particle_diameter", P->particle_diameter);
  D("impact_length_hsml_ratio", P->impact_length_hsml_ratio);
  I("thread_num", P->thread_num);
  D("kernel_2d_coef", P->kernel_2d_coef);
  D("adaptive_cfl", P->adaptive_cfl);
  std::fclose(f);
  return sphnj_particles_write(ps, (d + "/data").c_str());
}

}  // extern "C"
