// Host-side internals shared by the C-ABI translation units (not part of the public ABI).
#pragma once

#include "../../include/sphnj.h"

#include <cstdint>
#include <string>
#include <vector>

// A host particle set: the contents of data/{fluid,solid,virt}_particle.dat, held as
// structure-of-arrays in file (= original index) order.
struct sphnj_particles {
  int32_t n_liquid = 0, n_gas = 0;  // fluid_particle.dat header
  std::vector<double> fx, fy, fu, fv, frho, fp;
  std::vector<int32_t> fphase;
  std::vector<double> sx, sy, su, sv, sX, sY;
  std::vector<int32_t> smat, sbc;
  std::vector<double> wx, wy;
  std::vector<int32_t> wbc;
};

namespace sphnj {

// Thread-local message for calls that have no context.
void set_global_error(const std::string &msg);

// Derived constants of one parameter set, all in double precision.  Restates Kernel::initiate
// (src/kernel.cpp:25-41), FluidProperties::initiate (src/properties.cpp:53-67),
// SolidProperties::initiate/readParameter (src/properties.cpp:69-89,150-158) and
// Domain::initiate(Kernel) (src/domain.cpp:37-48).
struct Derived {
  // kernel
  double h, d, R, rc2, h2, hinv2;
  double kd;   // kernel_differential = 2/C3 (h/d)^-3 h^-2
  double k2;   // kernel_coefficient_2 = 1/C3 (h/d)^-3
  double eR;   // kernel_coefficient_1 = exp(-R^2)
  double c32;  // COEF_3D_TO_2D (or the kernel_2d_coef override)
  double kdc;  // kd * c32 : gradient weight w'(r^2) = exp(-r^2/h^2) * kdc
  double k2c;  // k2 * c32 : Shepard weight w(r^2) = (exp(-r^2/h^2) - eR) * k2c
  // fluid, per phase (0 liquid, 1 gas)
  double rho0[2], rho0_inv[2], c[2], gamma[2], gamma_inv[2], p2rho[2], rho2p[2], mu[2], rho_min[2];
  double p_min, p_bg, shift_coef, umax_cap;
  // physics
  double gx, gy;
  int steady, viscid;
  // solid
  double lambda, mu_s, K_s, c_s, alpha, beta, rho_s, rho_aux, fixed_u, fixed_v;
  double aux_modulus, yield_strain, yield_stress, harden_strain, harden_coef;
  int plastic;
  // domain
  double cxmin, cymin, cdx, cdy, inv_cdx, inv_cdy;
  int nx, ny;
  double dxmin, dymin, dxmax, dymax;
  // time
  double dt, dt_s;
  int S, K;
  // adaptive step size (0 = off): C, C h, the larger phase sound speed, S C h / c_s
  double cfl, cfl_h, c_max, dt_solid_cap;
};

int derive(const sphnj_params &p, Derived *out, std::string *err);

}  // namespace sphnj
