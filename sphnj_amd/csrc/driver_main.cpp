// sphnj -- host driver over libsphnj, the counterpart of the reference's main
// (src/main.cu:14-40), Sim::initiate (src/sim.cu) and DeviceCalculation::calculation
// (src/calculation.cu:20-48), keeping its surfaces: <case>/config/input.ini and
// <case>/data/{fluid,solid,virt}_particle.dat in, <case>/out/*.vtk out every
// result_interval steps, "total time=...ms" on stdout.
//
//   sphnj run <case_dir> [--precision p64|m32] [--steps N] [--no-output] [--no-graph]
//   sphnj gen <case_name> <case_dir> [options]      (writes input.ini + data/*.dat)
//   sphnj vtk <case_dir> <out_dir>                  (the input particles as VTK, no GPU)
#include <cerrno>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <sys/stat.h>
#include <vector>

#include "sphnj.h"

namespace {

int usage() {
  std::fprintf(stderr,
               "usage: sphnj run <case_dir> [--precision p64|m32] [--steps N] [--no-output] [--no-graph]\n"
               "       sphnj gen <case_name> <case_dir> [key=value;...]\n"
               "       sphnj vtk <case_dir> <out_dir>\n");
  return 2;
}

// Legacy-VTK POLYDATA of one species (the reference writes one file per field,
// src/output.cu:7-248; here one file carries every field, with true names, at 17 significant
// digits so a reader gets the doubles back).  Returns 0, or SPHNJ_E_IO with a message on
// stderr when the file cannot be opened or written.
int write_vtk(const std::string &path, int n, const double *x, const double *y,
              const std::vector<std::pair<std::string, const double *>> &scalars,
              const std::vector<std::pair<std::string, const int32_t *>> &ints) {
  FILE *f = std::fopen(path.c_str(), "w");
  if (!f) {
    std::fprintf(stderr, "cannot open %s for writing\n", path.c_str());
    return SPHNJ_E_IO;
  }
  std::fprintf(f, "# vtk DataFile Version 3.0\nsphnj\nASCII\nDATASET POLYDATA\nPOINTS %d double\n", n);
  for (int i = 0; i < n; ++i) std::fprintf(f, "%.17g %.17g 0\n", x[i], y[i]);
  std::fprintf(f, "POINT_DATA %d\n", n);
  for (auto &s : scalars) {
    std::fprintf(f, "SCALARS %s double 1\nLOOKUP_TABLE default\n", s.first.c_str());
    for (int i = 0; i < n; ++i) std::fprintf(f, "%.17g\n", s.second[i]);
  }
  for (auto &s : ints) {
    std::fprintf(f, "SCALARS %s int 1\nLOOKUP_TABLE default\n", s.first.c_str());
    for (int i = 0; i < n; ++i) std::fprintf(f, "%d\n", s.second[i]);
  }
  const bool bad = std::ferror(f) != 0;
  if (std::fclose(f) != 0 || bad) {
    std::fprintf(stderr, "error writing %s\n", path.c_str());
    return SPHNJ_E_IO;
  }
  return 0;
}

int make_dir(const std::string &d) {
  if (mkdir(d.c_str(), 0755) != 0 && errno != EEXIST) {
    std::fprintf(stderr, "cannot create %s\n", d.c_str());
    return SPHNJ_E_IO;
  }
  return 0;
}

int output(sphnj_ctx *ctx, const std::string &dir, int step) {
  int nf, ns, nw;
  sphnj_counts(ctx, 0, &nf, &ns, &nw, nullptr);
  std::vector<double> fx(nf), fy(nf), fu(nf), fv(nf), fr(nf), fp(nf);
  std::vector<int32_t> fph(nf), fcls(nf);
  std::vector<double> sx(ns), sy(ns), su(ns), sv(ns), sp(ns), sP(4 * (size_t)ns);
  std::vector<double> wx(nw), wy(nw), wp(nw);
  sphnj_fluid_out fo{};
  fo.x = fx.data(); fo.y = fy.data(); fo.u = fu.data(); fo.v = fv.data(); fo.rho = fr.data(); fo.p = fp.data();
  fo.phase = fph.data(); fo.boundary_type = fcls.data();
  sphnj_solid_out so{};
  so.x = sx.data(); so.y = sy.data(); so.u = su.data(); so.v = sv.data(); so.pressure = sp.data(); so.stress = sP.data();
  sphnj_wall_out wo{};
  wo.x = wx.data(); wo.y = wy.data(); wo.p = wp.data();
  int rc = sphnj_download(ctx, 0, &fo, &so, &wo);
  if (rc) return rc;
  const std::string t = std::to_string(step);
  if (nf && (rc = write_vtk(dir + "/fluid-" + t + ".vtk", nf, fx.data(), fy.data(),
                            {{"density", fr.data()}, {"pressure", fp.data()}, {"u", fu.data()}, {"v", fv.data()}},
                            {{"phase", fph.data()}, {"boundary_type", fcls.data()}})))
    return rc;
  if (ns) {
    std::vector<double> pxx(ns), pyy(ns);
    for (int i = 0; i < ns; ++i) {
      pxx[i] = sP[4 * i];
      pyy[i] = sP[4 * i + 3];
    }
    if ((rc = write_vtk(dir + "/solid-" + t + ".vtk", ns, sx.data(), sy.data(),
                        {{"pressure", sp.data()}, {"u", su.data()}, {"v", sv.data()}, {"P_xx", pxx.data()},
                         {"P_yy", pyy.data()}},
                        {})))
      return rc;
  }
  if (nw && (rc = write_vtk(dir + "/virt-" + t + ".vtk", nw, wx.data(), wy.data(), {{"pressure", wp.data()}}, {})))
    return rc;
  std::printf("Outputing Data... --step=%d\n", step);
  return 0;
}

// The input particles of a case as VTK (fluid, solid, virt at step "input"), through the
// same writer as the run's output.
int vtk(int argc, char **argv) {
  if (argc < 4) return usage();
  sphnj_particles *ps = nullptr;
  if (sphnj_particles_read((std::string(argv[2]) + "/data").c_str(), &ps)) {
    std::fprintf(stderr, "FAILED TO READ PARTICLES DATA: %s\n", sphnj_global_error());
    return 1;
  }
  sphnj_fluid_in f;
  sphnj_solid_in s;
  sphnj_wall_in w;
  sphnj_particles_views(ps, &f, &s, &w, nullptr);
  const std::string out = argv[3];
  int rc = make_dir(out);
  if (!rc && f.n)
    rc = write_vtk(out + "/fluid-input.vtk", f.n, f.x, f.y,
                   {{"density", f.rho}, {"pressure", f.p}, {"u", f.u}, {"v", f.v}}, {{"phase", f.phase}});
  if (!rc && s.n)
    rc = write_vtk(out + "/solid-input.vtk", s.n, s.x, s.y, {{"u", s.u}, {"v", s.v}, {"X", s.X}, {"Y", s.Y}},
                   {{"material", s.material}});
  if (!rc && w.n) rc = write_vtk(out + "/virt-input.vtk", w.n, w.x, w.y, {}, {{"boundary_type", w.bc}});
  sphnj_particles_destroy(ps);
  return rc ? 1 : 0;
}

int run(int argc, char **argv) {
  if (argc < 3) return usage();
  const std::string dir = argv[2];
  int prec = SPHNJ_PREC_P64, steps = -1;
  bool out = true, graph = true;
  for (int i = 3; i < argc; ++i) {
    if (!std::strcmp(argv[i], "--precision") && i + 1 < argc) {
      prec = !std::strcmp(argv[++i], "m32") ? SPHNJ_PREC_M32 : SPHNJ_PREC_P64;
    } else if (!std::strcmp(argv[i], "--steps") && i + 1 < argc) {
      steps = std::atoi(argv[++i]);
    } else if (!std::strcmp(argv[i], "--no-output")) {
      out = false;
    } else if (!std::strcmp(argv[i], "--no-graph")) {
      graph = false;
    } else {
      return usage();
    }
  }
  sphnj_params P;
  if (sphnj_params_from_ini((dir + "/config/input.ini").c_str(), &P)) {
    std::fprintf(stderr, "%s\n", sphnj_global_error());
    return 1;
  }
  if (sphnj_params_validate(&P)) {
    std::fprintf(stderr, "%s\n", sphnj_global_error());
    return 1;
  }
  sphnj_particles *ps = nullptr;
  auto t0 = std::chrono::steady_clock::now();
  if (sphnj_particles_read((dir + "/data").c_str(), &ps)) {
    std::fprintf(stderr, "FAILED TO READ PARTICLES DATA: %s\n", sphnj_global_error());
    return 1;
  }
  auto t1 = std::chrono::steady_clock::now();
  std::printf("time=%lldms\n", (long long)std::chrono::duration_cast<std::chrono::milliseconds>(t1 - t0).count());
  sphnj_fluid_in f;
  sphnj_solid_in s;
  sphnj_wall_in w;
  sphnj_particles_views(ps, &f, &s, &w, nullptr);
  std::printf("fluid number:%d\nsolid number:%d\nvirtual number:%d\n", f.n, s.n, w.n);
  sphnj_ctx *ctx = nullptr;
  int rc = sphnj_create(&P, prec, &ctx);
  if (!rc) rc = sphnj_upload(ctx, 0, &f, &s, &w);
  if (!rc) rc = sphnj_set_graph(ctx, graph);
  if (!rc) rc = sphnj_init(ctx);
  sphnj_particles_destroy(ps);
  if (rc) {
    std::fprintf(stderr, "error %d: %s\n", rc, ctx ? sphnj_last_error(ctx) : sphnj_global_error());
    sphnj_destroy(ctx);
    return 1;
  }
  const std::string odir = dir + "/out";
  if (out && make_dir(odir)) {
    sphnj_destroy(ctx);
    return 1;
  }
  const int n = steps >= 0 ? steps : P.iend - P.istart;
  const int every = P.result_interval > 0 ? P.result_interval : 1;
  auto s0 = std::chrono::steady_clock::now();
  for (int i = 0; i < n && !rc;) {
    const int chunk = out ? std::min(every - (i % every), n - i) : n - i;
    if (out && i % every == 0) rc = output(ctx, odir, P.istart + i);
    if (!rc) rc = sphnj_step(ctx, chunk);
    i += chunk;
  }
  if (!rc) rc = sphnj_sync(ctx);
  auto s1 = std::chrono::steady_clock::now();
  if (rc) {
    std::fprintf(stderr, "error %d: %s\n", rc, sphnj_last_error(ctx));
    sphnj_destroy(ctx);
    return 1;
  }
  std::printf("total time=%lldms\n", (long long)std::chrono::duration_cast<std::chrono::milliseconds>(s1 - s0).count());
  if (out) rc = output(ctx, odir, P.istart + n);
  uint32_t st = 0;
  sphnj_status(ctx, &st);
  if (st) std::printf("status flags=0x%x\n", st);
  sphnj_destroy(ctx);
  return rc ? 1 : 0;
}

int gen(int argc, char **argv) {
  if (argc < 4) return usage();
  sphnj_params P;
  sphnj_particles *ps = nullptr;
  if (sphnj_case_create(argv[2], argc > 4 ? argv[4] : "", &P, &ps) || sphnj_case_write(&P, ps, argv[3])) {
    std::fprintf(stderr, "%s\n", sphnj_global_error());
    return 1;
  }
  sphnj_fluid_in f;
  sphnj_solid_in s;
  sphnj_wall_in w;
  sphnj_particles_views(ps, &f, &s, &w, nullptr);
  std::printf("%s: fluid %d solid %d wall %d -> %s\n", argv[2], f.n, s.n, w.n, argv[3]);
  sphnj_particles_destroy(ps);
  return 0;
}

}  // namespace

int main(int argc, char **argv) {
  if (argc < 2) return usage();
  if (!std::strcmp(argv[1], "run")) return run(argc, argv);
  if (!std::strcmp(argv[1], "gen")) return gen(argc, argv);
  if (!std::strcmp(argv[1], "vtk")) return vtk(argc, argv);
  return usage();
}
