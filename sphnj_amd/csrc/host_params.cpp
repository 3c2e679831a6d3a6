// INI parameter surface and derived constants.
//
// Config semantics follow include/config.h:114-159 and src/config.cpp:59-129 of the reference;
// keys and defaults follow the readParameter functions of src/{physics,properties,domain,
// sim_time,kernel}.cpp.  Values are parsed at double precision (the reference parses float).
#include "host.hpp"

#include <cctype>
#include <cmath>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <vector>

namespace sphnj {

static thread_local std::string g_error;
void set_global_error(const std::string &msg) { g_error = msg; }

namespace {

// The INI text is first split into physical lines, then grouped into entries: a line whose
// code part (the text before the first '#') holds '=' opens an entry, and the lines after it
// that carry a value but no '=' extend that entry's value.  The grouping rules reproduce the
// observable behaviour of the reference's Config reader (src/config.cpp:59-129):
//  * the key is the text before the first '=', the value the text after it, both stripped of
//    surrounding white space;
//  * a line that is entirely white space closes the open entry;
//  * a line whose code part holds '=' closes it and opens the next one;
//  * a comment-only line neither closes the entry nor adds to it;
//  * any other line is appended to the value after a '\n';
//  * lines before the first entry, and lines that open none, are ignored;
//  * a key given twice keeps its last value.
// The reference re-reads its stream and can append a file's last line twice when the file
// does not end in a newline; every line is taken once here.
const char *const kSpace = " \n\t\v\r\f";

std::string strip(const std::string &s) {
  const auto b = s.find_first_not_of(kSpace);
  if (b == std::string::npos) return std::string();
  const auto e = s.find_last_not_of(kSpace);
  return s.substr(b, e - b + 1);
}

bool all_space(const std::string &s) { return s.find_first_not_of(kSpace) == std::string::npos; }

std::string code_part(const std::string &s) { return s.substr(0, s.find('#')); }

std::map<std::string, std::string> parse_ini(std::istream &is) {
  std::vector<std::string> lines;
  for (std::string l; std::getline(is, l);) lines.push_back(l);

  std::map<std::string, std::string> kv;
  size_t i = 0;
  while (i < lines.size()) {
    const std::string head = code_part(lines[i++]);
    const auto eq = head.find('=');
    if (eq == std::string::npos) continue;
    std::string value = strip(head.substr(eq + 1));
    for (; i < lines.size(); ++i) {
      if (all_space(lines[i])) {
        ++i;  // the blank line is consumed with the entry it closes
        break;
      }
      const std::string code = code_part(lines[i]);
      if (code.find('=') != std::string::npos) break;  // the next entry starts here
      const std::string piece = strip(code);
      if (!piece.empty()) value += "\n" + piece;
    }
    kv[strip(head.substr(0, eq))] = strip(value);
  }
  return kv;
}

// string_as_T<T>: istringstream extraction; a failed extraction yields 0 (C++11 rule).
double as_double(const std::string &s) {
  std::istringstream ist(s);
  double t = 0.0;
  ist >> t;
  return t;
}
int32_t as_int(const std::string &s) {
  std::istringstream ist(s);
  int t = 0;
  ist >> t;
  return (int32_t)t;
}
// string_as_T<bool> (include/config.h:122-138).
int32_t as_bool(const std::string &s) {
  std::string up = s;
  for (auto &ch : up) ch = (char)std::toupper((unsigned char)ch);
  if (up == "FALSE" || up == "F" || up == "NO" || up == "N" || up == "0" || up == "NONE") return 0;
  return 1;
}

struct Reader {
  const std::map<std::string, std::string> &kv;
  void d(const char *k, double &v) const {
    auto it = kv.find(k);
    if (it != kv.end()) v = as_double(it->second);
  }
  void i(const char *k, int32_t &v) const {
    auto it = kv.find(k);
    if (it != kv.end()) v = as_int(it->second);
  }
  void b(const char *k, int32_t &v) const {
    auto it = kv.find(k);
    if (it != kv.end()) v = as_bool(it->second);
  }
};

}  // namespace

int derive(const sphnj_params &p, Derived *o, std::string *err) {
  Derived &D = *o;
  std::memset(&D, 0, sizeof(D));
  // Kernel::initiate (src/kernel.cpp:25-41); C3 = tg_coef_2 (include/global.h:12).
  const double C3 = 5.551921;
  D.h = p.smoothing_length;
  D.d = p.particle_diameter;
  D.R = p.impact_length_hsml_ratio;
  D.h2 = D.h * D.h;
  D.hinv2 = 1.0 / D.h2;
  D.rc2 = D.R * D.R * D.h2;
  const double ratio = D.h / D.d;
  const double ratio_m3 = 1.0 / (ratio * ratio * ratio);
  D.eR = std::exp(-(D.R * D.R));
  D.k2 = 1.0 / C3 * ratio_m3;
  D.kd = 2.0 / C3 * ratio_m3 / D.h2;
  D.c32 = p.kernel_2d_coef != 0.0 ? p.kernel_2d_coef : 2.120729;
  D.kdc = D.kd * D.c32;
  D.k2c = D.k2 * D.c32;
  // FluidProperties::initiate (src/properties.cpp:53-67).
  const double c[2] = {p.liquid_sound_speed, p.gas_sound_speed};
  const double r0[2] = {p.liquid_density, p.gas_density};
  const double g[2] = {p.liquid_gamma, p.gas_gamma};
  const double mu[2] = {p.liquid_viscosity, p.gas_viscosity};
  for (int k = 0; k < 2; ++k) {
    D.c[k] = c[k];
    D.rho0[k] = r0[k];
    D.gamma[k] = g[k];
    D.mu[k] = mu[k];
    D.p2rho[k] = g[k] / (c[k] * c[k]) / r0[k];
    D.rho2p[k] = 1.0 / D.p2rho[k];
    D.gamma_inv[k] = 1.0 / g[k];
    D.rho0_inv[k] = 1.0 / r0[k];
    D.rho_min[k] = (1.0 - 0.4) * r0[k];
  }
  D.p_min = p.min_pressure;
  D.p_bg = p.background_pressure;
  D.shift_coef = p.particle_shift_coef;
  // getMaxSpeed cap (src/fluid.cu:438-439).
  D.umax_cap = 0.1 * std::fmin(c[0], c[1]);
  // Physics::readParameter (src/physics.cpp:9-25).
  D.gx = 0.0;
  D.gy = -p.gravity;
  D.steady = p.is_steady ? 1 : 0;
  D.viscid = p.is_viscid ? 1 : 0;
  // SolidProperties::initiate (src/properties.cpp:69-89) and the yield coupling (:150-158).
  const double E = p.youngs_modulus, nu = p.poisson_ratio;
  if (p.is_2D_strain)
    D.lambda = E * nu / (1.0 - 2.0 * nu) / (1.0 + nu);
  else
    D.lambda = E * nu / (1.0 - nu) / (1.0 + nu);
  D.mu_s = E / (1.0 + nu) / 2.0;
  D.K_s = E / (3.0 * (1.0 - 2.0 * nu));
  D.rho_s = p.solid_density;
  D.c_s = std::sqrt(D.K_s / D.rho_s);
  D.alpha = p.arti_vis_alpha;
  D.beta = p.arti_vis_beta;
  D.rho_aux = p.auxetic_density;
  D.fixed_u = p.solid_fixed_speed_x;
  D.fixed_v = p.solid_fixed_speed_y;
  D.yield_stress = p.yield_stress;
  D.harden_strain = p.harden_strain;
  D.harden_coef = p.harden_coef;
  D.plastic = p.is_plastic ? 1 : 0;
  if (p.yield_strain != 0.0) {
    D.yield_strain = p.yield_strain;
    D.aux_modulus = p.yield_stress / p.yield_strain;
  } else {
    D.aux_modulus = p.auxetic_modulus;
    D.yield_strain = p.auxetic_modulus != 0.0 ? p.yield_stress / p.auxetic_modulus : 0.0;
  }
  // Domain::initiate(Kernel) (src/domain.cpp:37-48).
  const double cell_size = D.R * D.h;
  D.cxmin = p.cell_xmin;
  D.cymin = p.cell_ymin;
  const double ex = p.cell_xmax - p.cell_xmin, ey = p.cell_ymax - p.cell_ymin;
  const double fnx = ex / cell_size, fny = ey / cell_size;
  if (!(fnx >= 1.0) || !(fny >= 1.0) || fnx > 32767.0 || fny > 32767.0) {
    if (err) *err = "cell box must hold between 1 and 32767 cells of size R*h per axis";
    return SPHNJ_E_PARAM;
  }
  D.nx = (int)fnx;
  D.ny = (int)fny;
  D.cdx = ex / (double)D.nx;
  D.cdy = ey / (double)D.ny;
  D.inv_cdx = 1.0 / D.cdx;
  D.inv_cdy = 1.0 / D.cdy;
  D.dxmin = p.domain_xmin;
  D.dymin = p.domain_ymin;
  D.dxmax = p.domain_xmax;
  D.dymax = p.domain_ymax;
  // SimTime (src/sim_time.cpp) and the solid sub-step dt (src/solid.cu:456).
  D.dt = p.dt;
  D.S = p.solid_sub_step > 0 ? p.solid_sub_step : 1;
  D.K = p.shift_per_step > 0 ? p.shift_per_step : 0;
  D.dt_s = p.dt / (double)D.S;
  // adaptive step size (extension): dt_n = min(dt, C h / (c_max + U), S C h / c_s)
  D.cfl = p.adaptive_cfl > 0.0 ? p.adaptive_cfl : 0.0;
  D.cfl_h = D.cfl * D.h;
  D.c_max = std::fmax(c[0], c[1]);
  D.dt_solid_cap = D.c_s > 0.0 ? (double)D.S * D.cfl_h / D.c_s : 0.0;
  return SPHNJ_OK;
}

}  // namespace sphnj

using namespace sphnj;

extern "C" {

int sphnj_abi_version(void) { return SPHNJ_ABI_VERSION; }

const char *sphnj_global_error(void) { return g_error.c_str(); }

int sphnj_params_default(sphnj_params *o) {
  if (!o) return SPHNJ_E_ARG;
  std::memset(o, 0, sizeof(*o));
  o->gravity = 9.81;
  o->is_steady = 0;
  o->particle_shift_coef = 0.4;
  o->is_viscid = 0;
  o->min_pressure = -1.013e5;
  o->liquid_sound_speed = 1500.0;
  o->liquid_density = 1000.0;
  o->liquid_gamma = 7.0;
  o->liquid_viscosity = 1.01e-3;
  o->gas_sound_speed = 340.0;
  o->gas_density = 1.29;
  o->gas_gamma = 1.4;
  o->gas_viscosity = 17.9e-6;
  o->solid_density = 2700.0;
  o->youngs_modulus = 67.5e9;
  o->poisson_ratio = 0.34;
  o->arti_vis_alpha = 2.5;
  o->arti_vis_beta = 2.5;
  o->solid_fixed_speed_x = 0.0;
  o->solid_fixed_speed_y = 0.0;
  o->is_2D_strain = 0;
  o->is_plastic = 0;
  o->auxetic_density = 636.0;
  o->auxetic_modulus = 0.0;
  o->auxetic_poisson_ratio = -0.2;
  o->yield_strain = 0.0;
  o->yield_stress = 2.0e7;
  o->harden_strain = 0.5;
  o->harden_coef = 100.0;
  o->dt = 0.0;
  o->istart = 0;
  o->iend = 0;
  o->solid_sub_step = 1;
  o->result_interval = 1;
  o->sort_every_step = 1;
  o->shift_per_step = 1;
  o->smoothing_length = 0.0;
  o->particle_diameter = 0.0;
  o->impact_length_hsml_ratio = 3.0;
  o->thread_num = 32;
  o->background_pressure = 0.0;
  o->kernel_2d_coef = 0.0;
  o->adaptive_cfl = 0.0;
  return SPHNJ_OK;
}

int sphnj_params_from_ini(const char *path, sphnj_params *o) {
  if (!path || !o) {
    set_global_error("sphnj_params_from_ini: null argument");
    return SPHNJ_E_ARG;
  }
  std::ifstream in(path);
  if (!in) {
    set_global_error(std::string("cannot open ") + path);
    return SPHNJ_E_IO;
  }
  sphnj_params_default(o);
  const auto kv = parse_ini(in);
  Reader r{kv};
  r.d("gravity", o->gravity);
  r.b("is_steady", o->is_steady);
  r.d("particle_shift_coef", o->particle_shift_coef);
  r.b("is_viscid", o->is_viscid);
  r.d("min_pressure", o->min_pressure);
  r.d("liquid_sound_speed", o->liquid_sound_speed);
  r.d("liquid_density", o->liquid_density);
  r.d("liquid_gamma", o->liquid_gamma);
  r.d("liquid_viscosity", o->liquid_viscosity);
  r.d("gas_sound_speed", o->gas_sound_speed);
  r.d("gas_density", o->gas_density);
  r.d("gas_gamma", o->gas_gamma);
  r.d("gas_viscosity", o->gas_viscosity);
  r.d("solid_density", o->solid_density);
  r.d("youngs_modulus", o->youngs_modulus);
  r.d("poisson_ratio", o->poisson_ratio);
  r.d("arti_vis_alpha", o->arti_vis_alpha);
  r.d("arti_vis_beta", o->arti_vis_beta);
  r.d("solid_fixed_speed_x", o->solid_fixed_speed_x);
  r.d("solid_fixed_speed_y", o->solid_fixed_speed_y);
  r.b("is_2D_strain", o->is_2D_strain);
  r.b("is_plastic", o->is_plastic);
  r.d("auxetic_density", o->auxetic_density);
  r.d("auxetic_modulus", o->auxetic_modulus);
  r.d("auxetic_poisson_ratio", o->auxetic_poisson_ratio);
  r.d("yield_strain", o->yield_strain);
  r.d("yield_stress", o->yield_stress);
  r.d("harden_strain", o->harden_strain);
  r.d("harden_coef", o->harden_coef);
  r.d("cell_xmin", o->cell_xmin);
  r.d("cell_ymin", o->cell_ymin);
  r.d("cell_xmax", o->cell_xmax);
  r.d("cell_ymax", o->cell_ymax);
  r.d("domain_xmin", o->domain_xmin);
  r.d("domain_ymin", o->domain_ymin);
  r.d("domain_xmax", o->domain_xmax);
  r.d("domain_ymax", o->domain_ymax);
  r.d("dt", o->dt);
  r.i("istart", o->istart);
  r.i("iend", o->iend);
  r.i("solid_sub_step", o->solid_sub_step);
  r.i("result_interval", o->result_interval);
  r.i("sort_every_step", o->sort_every_step);
  r.i("shift_per_step", o->shift_per_step);
  r.d("smoothing_length", o->smoothing_length);
  r.d("particle_diameter", o->particle_diameter);
  r.d("impact_length_hsml_ratio", o->impact_length_hsml_ratio);
  r.i("thread_num", o->thread_num);
  r.d("background_pressure", o->background_pressure);
  r.d("kernel_2d_coef", o->kernel_2d_coef);
  r.d("adaptive_cfl", o->adaptive_cfl);
  return SPHNJ_OK;
}

int sphnj_params_validate(const sphnj_params *p) {
  if (!p) return SPHNJ_E_ARG;
  std::string why;
  // Kernel::readParameter (src/kernel.cpp:16-20)
  if (p->smoothing_length == 0.0 || p->particle_diameter == 0.0) why += "failed to read kernel information; ";
  // SolidProperties::readParameter (src/properties.cpp:150-154)
  if (p->auxetic_modulus == 0.0 && p->yield_strain == 0.0) why += "failed to read solid information; ";
  // Domain::readParameter (src/domain.cpp:20-29)
  if (p->cell_xmax - p->cell_xmin == 0.0 || p->cell_ymax - p->cell_ymin == 0.0) why += "failed to read cell information; ";
  if (p->domain_xmax - p->domain_xmin == 0.0 || p->domain_ymax - p->domain_ymin == 0.0)
    why += "failed to read domain information; ";
  // SimTime::readParameter (src/sim_time.cpp:18-22)
  if (p->dt == 0.0 || p->istart - p->iend == 0) why += "failed to read time information; ";
  if (!why.empty()) {
    set_global_error("FAILED TO READ INPUT.INI: " + why);
    return SPHNJ_E_PARAM;
  }
  Derived D;
  std::string err;
  if (derive(*p, &D, &err) != SPHNJ_OK) {
    set_global_error(err);
    return SPHNJ_E_PARAM;
  }
  return SPHNJ_OK;
}

int sphnj_params_grid(const sphnj_params *p, int32_t *nx, int32_t *ny, double *cdx, double *cdy) {
  if (!p) return SPHNJ_E_ARG;
  Derived D;
  std::string err;
  int rc = derive(*p, &D, &err);
  if (rc != SPHNJ_OK) {
    set_global_error(err);
    return rc;
  }
  if (nx) *nx = D.nx;
  if (ny) *ny = D.ny;
  if (cdx) *cdx = D.cdx;
  if (cdy) *cdy = D.cdy;
  return SPHNJ_OK;
}

}  // extern "C"
