// Particle .dat files: data/fluid_particle.dat, data/solid_particle.dat, data/virt_particle.dat.
//
// Formats and checks follow the reference loaders:
//   header/line-count check: FluidProperties/SolidProperties/VirtualProperties::readNumber
//     (src/properties.cpp:7-33, 91-116, 164-189): every line after the header is one particle;
//   row parsing: Fluid_Host_List/Solid_Host_List/Virtual_Host_List::addList
//     (include/host_data.cuh:24-46, 64-87, 99-115): fluid "x y u v rho p phase", solid
//     "x y u v X Y material [bc]" (a missing bc column reads as 0 = UNCONSTRAINED), wall "x y bc".
// A missing file counts as zero particles of that species only when it is absent; the reference
// treats an unopenable file as an error, and so does this reader.
#include "host.hpp"

#include <cstdio>
#include <fstream>
#include <sstream>

using namespace sphnj;

namespace {

int read_lines(const std::string &path, std::string *header, std::vector<std::string> *rows) {
  std::ifstream in(path);
  if (!in) {
    set_global_error("failed to open " + path);
    return SPHNJ_E_IO;
  }
  std::getline(in, *header);
  std::string s;
  while (std::getline(in, s)) rows->push_back(s);
  return SPHNJ_OK;
}

// istringstream extraction; a failed extraction leaves 0 (C++11), as in addList.
template <class T>
T take(std::istringstream &is) {
  T v = T(0);
  is >> v;
  return v;
}

}  // namespace

extern "C" {

int sphnj_particles_read(const char *dir, sphnj_particles **out) {
  if (!dir || !out) return SPHNJ_E_ARG;
  *out = nullptr;
  auto *ps = new sphnj_particles();
  const std::string d(dir);
  std::string header;
  std::vector<std::string> rows;

  int rc = read_lines(d + "/fluid_particle.dat", &header, &rows);
  if (rc) { delete ps; return rc; }
  {
    std::istringstream h(header);
    ps->n_liquid = take<int>(h);
    ps->n_gas = take<int>(h);
    const long total = (long)ps->n_liquid + ps->n_gas;
    if (total < 0) { delete ps; set_global_error("fluid particle number should >= 0"); return SPHNJ_E_COUNT; }
    if (total > 0 && (long)rows.size() != total) {
      delete ps;
      set_global_error("mismatch fluid particle numbers");
      return SPHNJ_E_COUNT;
    }
    if (total == 0) rows.clear();
    for (const auto &r : rows) {
      std::istringstream is(r);
      ps->fx.push_back(take<double>(is));
      ps->fy.push_back(take<double>(is));
      ps->fu.push_back(take<double>(is));
      ps->fv.push_back(take<double>(is));
      ps->frho.push_back(take<double>(is));
      ps->fp.push_back(take<double>(is));
      ps->fphase.push_back(take<int>(is));
    }
  }
  rows.clear();
  rc = read_lines(d + "/solid_particle.dat", &header, &rows);
  if (rc) { delete ps; return rc; }
  {
    std::istringstream h(header);
    const long n = take<long>(h);
    if (n < 0) { delete ps; set_global_error("solid particle number should >= 0"); return SPHNJ_E_COUNT; }
    if (n > 0 && (long)rows.size() != n) {
      delete ps;
      set_global_error("mismatch solid particle numbers");
      return SPHNJ_E_COUNT;
    }
    if (n == 0) rows.clear();
    for (const auto &r : rows) {
      std::istringstream is(r);
      ps->sx.push_back(take<double>(is));
      ps->sy.push_back(take<double>(is));
      ps->su.push_back(take<double>(is));
      ps->sv.push_back(take<double>(is));
      ps->sX.push_back(take<double>(is));
      ps->sY.push_back(take<double>(is));
      ps->smat.push_back(take<int>(is));
      ps->sbc.push_back(take<int>(is));
    }
  }
  rows.clear();
  rc = read_lines(d + "/virt_particle.dat", &header, &rows);
  if (rc) { delete ps; return rc; }
  {
    std::istringstream h(header);
    const long n = take<long>(h);
    if (n < 0) { delete ps; set_global_error("virtual particle number should >= 0"); return SPHNJ_E_COUNT; }
    if (n > 0 && (long)rows.size() != n) {
      delete ps;
      set_global_error("mismatch virtual particle numbers");
      return SPHNJ_E_COUNT;
    }
    if (n == 0) rows.clear();
    for (const auto &r : rows) {
      std::istringstream is(r);
      ps->wx.push_back(take<double>(is));
      ps->wy.push_back(take<double>(is));
      ps->wbc.push_back(take<int>(is));
    }
  }
  *out = ps;
  return SPHNJ_OK;
}

int sphnj_particles_write(const sphnj_particles *ps, const char *dir) {
  if (!ps || !dir) return SPHNJ_E_ARG;
  const std::string d(dir);
  FILE *f = std::fopen((d + "/fluid_particle.dat").c_str(), "w");
  if (!f) { set_global_error("cannot write " + d + "/fluid_particle.dat"); return SPHNJ_E_IO; }
  std::fprintf(f, "%d %d\n", ps->n_liquid, ps->n_gas);
  for (size_t i = 0; i < ps->fx.size(); ++i)
    std::fprintf(f, "%.17g %.17g %.17g %.17g %.17g %.17g %d\n", ps->fx[i], ps->fy[i], ps->fu[i], ps->fv[i],
                 ps->frho[i], ps->fp[i], ps->fphase[i]);
  std::fclose(f);
  f = std::fopen((d + "/solid_particle.dat").c_str(), "w");
  if (!f) { set_global_error("cannot write " + d + "/solid_particle.dat"); return SPHNJ_E_IO; }
  std::fprintf(f, "%zu\n", ps->sx.size());
  for (size_t i = 0; i < ps->sx.size(); ++i)
    std::fprintf(f, "%.17g %.17g %.17g %.17g %.17g %.17g %d %d\n", ps->sx[i], ps->sy[i], ps->su[i], ps->sv[i],
                 ps->sX[i], ps->sY[i], ps->smat[i], ps->sbc[i]);
  std::fclose(f);
  f = std::fopen((d + "/virt_particle.dat").c_str(), "w");
  if (!f) { set_global_error("cannot write " + d + "/virt_particle.dat"); return SPHNJ_E_IO; }
  std::fprintf(f, "%zu\n", ps->wx.size());
  for (size_t i = 0; i < ps->wx.size(); ++i) std::fprintf(f, "%.17g %.17g %d\n", ps->wx[i], ps->wy[i], ps->wbc[i]);
  std::fclose(f);
  return SPHNJ_OK;
}

int sphnj_particles_views(const sphnj_particles *ps, sphnj_fluid_in *f, sphnj_solid_in *s, sphnj_wall_in *w,
                          int32_t *n_liquid) {
  if (!ps) return SPHNJ_E_ARG;
  if (f) {
    f->n = (int32_t)ps->fx.size();
    f->x = ps->fx.data(); f->y = ps->fy.data(); f->u = ps->fu.data(); f->v = ps->fv.data();
    f->rho = ps->frho.data(); f->p = ps->fp.data(); f->phase = ps->fphase.data();
  }
  if (s) {
    s->n = (int32_t)ps->sx.size();
    s->x = ps->sx.data(); s->y = ps->sy.data(); s->u = ps->su.data(); s->v = ps->sv.data();
    s->X = ps->sX.data(); s->Y = ps->sY.data(); s->material = ps->smat.data(); s->bc = ps->sbc.data();
  }
  if (w) {
    w->n = (int32_t)ps->wx.size();
    w->x = ps->wx.data(); w->y = ps->wy.data(); w->bc = ps->wbc.data();
  }
  if (n_liquid) *n_liquid = ps->n_liquid;
  return SPHNJ_OK;
}

void sphnj_particles_destroy(sphnj_particles *ps) { delete ps; }

}  // extern "C"
