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
  std::fclose(f);
  return sphnj_particles_write(ps, (d + "/data").c_str());
}

}  // extern "C"
