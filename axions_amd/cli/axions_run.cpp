/*
 * axions_run - the reference programs' main() on the GPU library.
 *
 * Each of the reference's evolution programs is a main() with file-scope
 * constants, a TimeStep loop and text output files.  axions_run keeps those
 * constants (same names, same defaults), reads the same input files and
 * writes the same output files, with the lattice work done by libaxions.so
 * through the C ABI of include/axions/axions.h:
 *
 *   --program evolution      evolution.cpp: Data/ic.txt (5 columns, one
 *                            level if stationary_ic else two), per-step
 *                            Data/valsPerLoop_<tag>.txt "E meanG [L]",
 *                            GifData/gifStringPosData_<frame>.txt,
 *                            Data/finalField.txt (finalOut)
 *                            (evolution.cpp:17-59, 75-133, 723-1030, 1404-1457)
 *   --program mpi_evolution  mpi_evolution.cpp: ic_type random (seeded) /
 *                            "stationary data" / data, per-step energy to
 *                            Data/valsPerLoop_<outTag>_nx<nx>_seed<seed>.txt
 *                            (calcEnergy), crossing frames to
 *                            GifData/gifStringPosData_<outTag>_nx<nx>_seed<seed>_<frame>.txt
 *                            until a frame is empty (mpi_evolution.cpp:17-60,
 *                            120-121, 160-312, 795-899)
 *   --program moore          mooreEvol.cpp: Data/moore_ic.txt (7 columns),
 *                            Data/valsPerLoop.txt "E meanG",
 *                            GifData/gifData_<frame>.txt (|phi1|^2 |phi2|^2),
 *                            Data/finalField.txt (mooreEvol.cpp:19-91, 697-731)
 *   --program leapfrog       leapfrog.cpp: ic.txt (2 columns), valsPerLoop.txt
 *                            (E per step), GifData/gifData_<frame>.txt (|phi|^2),
 *                            finalField.txt (leapfrog.cpp:74-112, 323-386)
 *   --program sor            singleSemiAnalytic.cpp: straight-string radial
 *                            profile to Data/SOR_Fields.txt "phi A" per point
 *                            (ic_type 1 starts from that file)
 *                            (singleSemiAnalytic.cpp:19-32, 54-142)
 *   --program straight_ic    straight_ic.cpp: SOR_Fields.txt (SORnx rows) ->
 *                            ic.txt, 5 columns "Re phi, Im phi, dx g Ax,
 *                            dy g Ay, dz g Az" (straight_ic.cpp:20-31, 53-150)
 *
 * Constants are set with name=value arguments (e.g. nx=256 nt=1001 g=1);
 * --dir is the directory the reference's dir_path would be (default "."),
 * --tag the output tag evolution.cpp asks for on stdin (asked the same way
 * when absent), --digits the significant digits written (default 6, the
 * reference's ostream default; 17 round-trips doubles).
 *
 * Differences, all deliberate: mpi_evolution is periodic (the MPI program
 * has no other boundary); the string-length column(s)
 * of evolution.cpp (uninitialised there: its chaining, evolution.cpp:1229-1330,
 * is commented out) hold the chained length of ic_w_evolution.cpp:961-1087
 * (total, or inner/outer with splitLength) via ax_string_length; all gauged
 * runs use the library's shared sin/cos; straight_ic reports the number of
 * sites beyond its profile once instead of one line per site.
 */
#include <errno.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <sys/types.h>

#include <chrono>
#include <cmath>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "axions/axions.h"

namespace {

struct Opt {
  std::string program = "evolution", dir = ".", tag;
  int digits = 6;
  std::map<std::string, std::string> kv;
};

[[noreturn]] void die(const std::string &m) {
  fprintf(stderr, "axions_run: %s\n", m.c_str());
  exit(2);
}

void check(int rc, const char *what) {
  if (rc != AX_OK) die(std::string(what) + ": " + ax_strerror(rc) + " (" + ax_last_error() + ")");
}

struct Consts {
  std::map<std::string, std::string> v;
  double d(const std::string &k) const { return strtod(v.at(k).c_str(), nullptr); }
  int i(const std::string &k) const { return (int)strtol(v.at(k).c_str(), nullptr, 10); }
  bool b(const std::string &k) const { return v.at(k) == "true" || v.at(k) == "1"; }
  std::string s(const std::string &k) const { return v.at(k); }
};

/* the reference programs' file-scope constants */
std::map<std::string, std::string> defaults(const std::string &prog) {
  if (prog == "evolution")   /* evolution.cpp:23-57 */
    return {{"nx", "101"}, {"ny", "101"}, {"nz", "101"}, {"nt", "1001"}, {"dx", "0.5"}, {"dy", "0.5"},
            {"dz", "0.5"}, {"dt", "0.05"}, {"lambda", "1"}, {"eta", "1"}, {"g", "0"}, {"damped_nt", "200"},
            {"dampFac", "0.5"}, {"alpha", "3"}, {"scaling", "1"}, {"makeStringPosGif", "true"},
            {"saveFreq", "5"}, {"stationary_ic", "true"}, {"stringDetect", "true"}, {"detectBuffer", "true"},
            {"finalOut", "false"}, {"xyBC", "periodic"}, {"zBC", "periodic"}, {"splitLength", "false"},
            {"precision", "64"}};
  if (prog == "mpi_evolution")   /* mpi_evolution.cpp:17-60 */
    return {{"nx", "1001"}, {"ny", "1001"}, {"nz", "1001"}, {"nt", "2501"}, {"dx", "0.25"}, {"dy", "0.25"},
            {"dz", "0.25"}, {"dt", "0.05"}, {"lambda", "1"}, {"eta", "1"}, {"g", "0"}, {"damped_nt", "200"},
            {"dampFac", "0.5"}, {"alpha", "3"}, {"scaling", "1"}, {"makeStringPosGif", "true"},
            {"saveFreq", "5"}, {"outTag", "global_PRS_dx0p25"}, {"ic_type", "random"}, {"seed", "42"},
            {"stringDetect", "true"}, {"detectBuffer", "true"}, {"calcEnergy", "false"}, {"precision", "64"}};
  if (prog == "moore")   /* mooreEvol.cpp:19-41 */
    return {{"nx", "201"}, {"ny", "201"}, {"nz", "51"}, {"nt", "6001"}, {"dx", "0.7"}, {"dy", "0.7"},
            {"dz", "0.7"}, {"dt", "0.3"}, {"eta1", "1"}, {"eta2", "1"}, {"lambda1", "2"}, {"lambda2", "2"},
            {"lambda12", "0"}, {"q1", "2"}, {"q2", "1"}, {"g", ""}, {"makeGif", "true"}, {"saveFreq", "10"},
            {"xyBC", "fixed"}, {"precision", "64"}};
  if (prog == "leapfrog")   /* leapfrog.cpp:74-84 */
    return {{"nx", "200"}, {"ny", "200"}, {"nz", "200"}, {"nt", "6001"}, {"dx", "0.25"}, {"dy", "0.25"},
            {"dz", "0.25"}, {"dt", "0.1"}, {"makeGif", "false"}, {"saveFreq", "60"}, {"xyBC", "neumann"},
            {"precision", "64"}};
  if (prog == "sor")   /* singleSemiAnalytic.cpp:19-32 */
    return {{"nx", "30001"}, {"h", "0.01"}, {"w", "1.5"}, {"maxIter", "2000000"}, {"tol", "1e-3"},
            {"ic_type", "0"}, {"n", "1"}, {"lambda", "1"}, {"eta", "1"}, {"g", "0"}};
  if (prog == "straight_ic")   /* straight_ic.cpp:20-31 */
    return {{"nx", "101"}, {"ny", "101"}, {"nz", "51"}, {"dx", "0.5"}, {"dy", "0.5"}, {"dz", "0.5"},
            {"n", "1"}, {"g", "0.5"}, {"SORnx", "20001"}, {"SORa", "0.01"}, {"precision", "64"}};
  die("unknown --program " + prog);
}

/* whitespace-separated doubles from a text file, as `ic >> a >> b ...` reads */
std::vector<double> read_numbers(const std::string &path, size_t want) {
  FILE *f = fopen(path.c_str(), "rb");
  if (!f) die("cannot open " + path);
  std::vector<char> buf;
  char chunk[1 << 16];
  size_t got;
  while ((got = fread(chunk, 1, sizeof chunk, f)) > 0) buf.insert(buf.end(), chunk, chunk + got);
  fclose(f);
  buf.push_back(0);
  std::vector<double> out;
  out.reserve(want);
  char *p = buf.data();
  while (out.size() < want) {
    char *e;
    double x = strtod(p, &e);
    if (e == p) break;
    out.push_back(x);
    p = e;
  }
  if (out.size() < want) die(path + ": expected " + std::to_string(want) + " numbers, read " + std::to_string(out.size()));
  return out;
}

/* per-site rows of ncol columns (x-major, z fastest) -> [col][site] */
std::vector<double> columns(const std::vector<double> &rows, size_t V, int ncol, size_t off = 0) {
  std::vector<double> c((size_t)ncol * V);
  for (size_t s = 0; s < V; ++s)
    for (int k = 0; k < ncol; ++k) c[(size_t)k * V + s] = rows[off + s * ncol + k];
  return c;
}

struct Out {
  FILE *f = nullptr;
  int digits;
  Out(const std::string &path, int dg) : digits(dg) {
    f = fopen(path.c_str(), "w");
    if (!f) die("cannot write " + path);
  }
  ~Out() {
    if (f) fclose(f);
  }
  /* like `ostream << double` with precision(digits): %g */
  void num(double x, bool last) { fprintf(f, last ? "%.*g\n" : "%.*g ", digits, x); }
};


struct Run {
  Opt o;
  Consts C;
  ax_params p{};
  ax_lattice *lat = nullptr;
  size_t V = 0;
  std::string D(const std::string &rel) const { return o.dir + "/" + rel; }

  void set_common(int model) {
    ax_default_params(&p);
    p.model = model;
    p.nx = C.i("nx");
    p.ny = C.i("ny");
    p.nz = C.i("nz");
    p.dx = C.d("dx");
    p.dy = C.d("dy");
    p.dz = C.d("dz");
    p.dt = C.d("dt");
    p.precision = C.i("precision");
    V = (size_t)p.nx * p.ny * p.nz;
  }

  void load_state(const std::vector<double> &host) {
    check(ax_set_state(lat, host.data(), host.size()), "ax_set_state");
  }

  /* ic with ncomp columns per site: one level (copied to both) or two */
  void load_ic(const std::string &path, int ncomp, bool two_levels) {
    const size_t levels = two_levels ? 2 : 1;
    std::vector<double> rows = read_numbers(path, levels * V * ncomp);
    std::vector<double> host((size_t)ncomp * 2 * V);
    for (int t = 0; t < 2; ++t) {
      const std::vector<double> c = columns(rows, V, ncomp, (two_levels ? t : 0) * V * ncomp);
      for (int k = 0; k < ncomp; ++k)
        memcpy(&host[((size_t)k * 2 + t) * V], &c[(size_t)k * V], V * sizeof(double));
    }
    load_state(host);
  }

  /* the newest level, as finalField writes it (tPast after the last step) */
  std::vector<double> newest(int ncomp) {
    std::vector<double> host((size_t)ncomp * 2 * V);
    check(ax_get_state(lat, host.data(), host.size()), "ax_get_state");
    int32_t ts = 0;
    ax_get_timestep(lat, &ts);
    const int cur = (ts + 1) % 2;
    std::vector<double> out((size_t)ncomp * V);
    for (int k = 0; k < ncomp; ++k) memcpy(&out[(size_t)k * V], &host[((size_t)k * 2 + cur) * V], V * sizeof(double));
    return out;
  }

  void write_rows(const std::string &path, const std::vector<double> &cols, int ncomp) {
    Out f(path, o.digits);
    for (size_t s = 0; s < V; ++s)
      for (int k = 0; k < ncomp; ++k) f.num(cols[(size_t)k * V + s], k == ncomp - 1);
  }

  std::vector<double> crossings() {
    size_t n = 0;
    check(ax_crossings(lat, nullptr, 0, &n), "ax_crossings");
    std::vector<double> xyz(3 * n + 3);
    check(ax_crossings(lat, xyz.data(), n, &n), "ax_crossings");
    xyz.resize(3 * n);
    return xyz;
  }

  size_t frame(const std::string &path, const std::vector<double> &xyz) {
    const size_t n = xyz.size() / 3;
    Out f(path, o.digits);
    for (size_t i = 0; i < n; ++i) {
      f.num(xyz[3 * i], false);
      f.num(xyz[3 * i + 1], false);
      f.num(xyz[3 * i + 2], true);
    }
    return n;
  }

  void evolution() {
    set_common(AX_MODEL_AH);
    p.lambda = C.d("lambda");
    p.eta = C.d("eta");
    p.g = C.d("g");
    p.damped_nt = C.i("damped_nt");
    p.dampFac = C.d("dampFac");
    p.alpha = C.d("alpha");
    p.scaling = C.d("scaling");
    p.theta_damp_form = 1;   /* evolution.cpp:371-373 */
    p.energy_form = 0;       /* evolution.cpp:403-427 */
    /* evolution.cpp:52-53: "periodic", "neumann", "absorbing" (x, y only);
     * any other string is fixed */
    const std::string xy = C.s("xyBC"), zb = C.s("zBC");
    p.boundary = xy == "periodic" ? AX_BC_PERIODIC : xy == "neumann" ? AX_BC_NEUMANN
                                                   : xy == "absorbing" ? AX_BC_ABSORBING : AX_BC_FIXED;
    p.boundary_z = zb == "periodic" ? AX_BC_PERIODIC : zb == "neumann" ? AX_BC_NEUMANN : AX_BC_FIXED;
    const int nt = C.i("nt"), save = C.i("saveFreq");
    const bool detect = C.b("stringDetect"), buffer = C.b("detectBuffer"), gif = C.b("makeStringPosGif");
    const bool stationary = C.b("stationary_ic"), fin = C.b("finalOut"), split = C.b("splitLength");
    if (o.tag.empty()) {
      printf("Enter a tag for output files: ");
      fflush(stdout);
      char t[256];
      if (scanf("%255s", t) != 1) die("no tag");
      o.tag = t;
    }
    check(ax_create(&p, &lat), "ax_create");
    load_ic(D("Data/ic.txt"), 5, !stationary);
    Out vals(D("Data/valsPerLoop_" + o.tag + ".txt"), o.digits);
    int frameNo = 0;
    for (int n = 0; n < nt; ++n) {
      const bool active = detect && (!buffer || n >= p.damped_nt);
      ax_diag d;
      check(ax_diagnostics(lat, 0, &d), "ax_diagnostics");
      double len[3] = {0, 0, 0};
      if (active) {
        const std::vector<double> xyz = crossings();
        if (gif && n % save == 0) frame(D("GifData/gifStringPosData_" + std::to_string(frameNo++) + ".txt"), xyz);
        check(ax_string_length(xyz.data(), xyz.size() / 3, p.nx, p.ny, p.nz, p.dx, p.dy, p.dz,
                               p.boundary == AX_BC_PERIODIC, p.boundary_z == AX_BC_PERIODIC, len), "ax_string_length");
      }
      vals.num(d.energy, false);
      vals.num(d.gauss_mean, !detect);
      if (detect && !split) vals.num(len[0], true);
      if (detect && split) {
        vals.num(len[1], false);
        vals.num(len[2], true);
      }
      check(ax_step(lat, 1), "ax_step");
    }
    if (fin) write_rows(D("Data/finalField.txt"), newest(5), 5);
  }

  void mpi_evolution() {
    set_common(AX_MODEL_AH);
    p.lambda = C.d("lambda");
    p.eta = C.d("eta");
    p.g = C.d("g");
    p.damped_nt = C.i("damped_nt");
    p.dampFac = C.d("dampFac");
    p.alpha = C.d("alpha");
    p.scaling = C.d("scaling");
    p.theta_damp_form = 0;   /* mpi_evolution.cpp:490-497 */
    p.energy_form = 1;       /* mpi_evolution.cpp:513-530 */
    p.boundary = AX_BC_PERIODIC;
    const int nt = C.i("nt"), save = C.i("saveFreq"), seed = C.i("seed");
    const bool detect = C.b("stringDetect"), buffer = C.b("detectBuffer"), gif = C.b("makeStringPosGif");
    const bool energy = C.b("calcEnergy");
    const std::string ic = C.s("ic_type"), tag = C.s("outTag") + "_nx" + std::to_string(p.nx) + "_seed" +
                                                    std::to_string(seed);
    check(ax_create(&p, &lat), "ax_create");
    if (ic == "random") check(ax_init_random_phase(lat, (uint32_t)seed, 0), "ax_init_random_phase");
    else if (ic == "stationary data") load_ic(D("Data/ic.txt"), 5, false);
    else if (ic == "data") load_ic(D("Data/ic.txt"), 5, true);
    else die("ic_type " + ic);
    Out vals(D("Data/valsPerLoop_" + tag + ".txt"), o.digits);
    int frameNo = 0;
    bool exist = true;
    for (int n = 0; n < nt; ++n) {
      if (energy) {
        ax_diag d;
        check(ax_diagnostics(lat, 0, &d), "ax_diagnostics");
        vals.num(d.energy, true);
      }
      if (detect && n % save == 0 && gif && exist && (!buffer || n >= p.damped_nt)) {
        if (frame(D("GifData/gifStringPosData_" + tag + "_" + std::to_string(frameNo++) + ".txt"), crossings()) == 0)
          exist = false;   /* mpi_evolution.cpp:854 */
      }
      check(ax_step(lat, 1), "ax_step");
    }
  }

  void moore() {
    set_common(AX_MODEL_MOORE);
    p.eta1 = C.d("eta1");
    p.eta2 = C.d("eta2");
    p.lambda1 = C.d("lambda1");
    p.lambda2 = C.d("lambda2");
    p.lambda12 = C.d("lambda12");
    p.q1 = C.i("q1");
    p.q2 = C.i("q2");
    const std::string gs = C.s("g");
    /* mooreEvol.cpp:36: g = sqrt(0.5*lambda1*pow(eta1,2)/(q1*q1 + q2*q2)) */
    p.g = gs.empty() ? sqrt(0.5 * p.lambda1 * (p.eta1 * p.eta1) / (p.q1 * p.q1 + p.q2 * p.q2)) : strtod(gs.c_str(), nullptr);
    /* mooreEvol.cpp:42: "neumann" copies the x/y faces, z stays fixed;
     * absorbing is not implemented there for this model */
    const std::string xy = C.s("xyBC");
    if (xy == "absorbing") die("moore: xyBC=absorbing is not defined for the Moore model (mooreEvol.cpp:42)");
    p.boundary = xy == "periodic" ? AX_BC_PERIODIC : xy == "neumann" ? AX_BC_NEUMANN : AX_BC_FIXED;
    p.boundary_z = xy == "periodic" ? AX_BC_PERIODIC : AX_BC_FIXED;
    const int nt = C.i("nt"), save = C.i("saveFreq");
    const bool gif = C.b("makeGif");
    check(ax_create(&p, &lat), "ax_create");
    load_ic(D("Data/moore_ic.txt"), 7, false);
    Out vals(D("Data/valsPerLoop.txt"), o.digits);
    const double inner = (double)((p.nx - 2) * (p.ny - 2) * (p.nz - 2));
    int frameNo = 0;
    for (int n = 0; n < nt; ++n) {
      ax_diag d;
      check(ax_diagnostics(lat, 0, &d), "ax_diagnostics");
      vals.num(d.energy, false);
      /* mooreEvol.cpp:697 divides the per-site mean once more */
      vals.num(p.boundary != AX_BC_PERIODIC ? d.gauss_mean / inner : d.gauss_mean, true);
      check(ax_step(lat, 1), "ax_step");
      if (gif && n % save == 0) {
        const std::vector<double> f = newest(7);
        Out g(D("GifData/gifData_" + std::to_string(frameNo++) + ".txt"), o.digits);
        for (size_t s = 0; s < V; ++s) {
          g.num(f[s] * f[s] + f[V + s] * f[V + s], false);
          g.num(f[2 * V + s] * f[2 * V + s] + f[3 * V + s] * f[3 * V + s], true);
        }
      }
    }
    write_rows(D("Data/finalField.txt"), newest(7), 7);
  }

  void sor() {
    ax_sor_params sp;
    ax_sor_default_params(&sp);
    sp.nx = C.i("nx");
    sp.h = C.d("h");
    sp.w = C.d("w");
    sp.max_iter = C.i("maxIter");
    sp.tol = C.d("tol");
    sp.n = C.d("n");
    sp.lambda = C.d("lambda");
    sp.eta = C.d("eta");
    sp.g = C.d("g");
    std::vector<double> init;
    if (C.i("ic_type") == 1) init = read_numbers(D("Data/SOR_Fields.txt"), 2 * (size_t)sp.nx);
    std::vector<double> prof(2 * (size_t)sp.nx);
    int32_t iters = 0;
    double res = 0;
    check(ax_sor_profile(&sp, init.empty() ? nullptr : init.data(), prof.data(), &iters, &res), "ax_sor_profile");
    Out f(D("Data/SOR_Fields.txt"), o.digits);
    for (int j = 0; j < sp.nx; ++j) {
      f.num(prof[2 * (size_t)j], false);
      f.num(prof[2 * (size_t)j + 1], true);
    }
    if (iters == sp.max_iter) printf("Maximum number of iterations reached\n");
    printf("Number of iterations: %d\n", iters);
  }

  void straight_ic() {
    ax_default_params(&p);                 /* the lattice is only written, not stepped */
    p.nx = C.i("nx");
    p.ny = C.i("ny");
    p.nz = C.i("nz");
    p.dx = C.d("dx");
    p.dy = C.d("dy");
    p.dz = C.d("dz");
    p.precision = C.i("precision");
    V = (size_t)p.nx * p.ny * p.nz;
    p.g = C.d("g");
    p.boundary = AX_BC_FIXED;
    p.boundary_z = AX_BC_FIXED;
    const int sn = C.i("SORnx");
    const std::vector<double> prof = read_numbers(D("SOR_Fields.txt"), 2 * (size_t)sn);
    check(ax_create(&p, &lat), "ax_create");
    int64_t off = 0;
    check(ax_init_straight_string(lat, prof.data(), sn, C.d("SORa"), C.d("n"), p.g, &off),
          "ax_init_straight_string");
    if (off) printf("Off straight string solution grid: %lld sites\n", (long long)off);
    write_rows(D("ic.txt"), newest(5), 5);
  }

  void leapfrog() {
    set_common(AX_MODEL_GLOBAL);
    p.boundary = C.s("xyBC") == "periodic" ? AX_BC_PERIODIC : AX_BC_NEUMANN_XY;
    const int nt = C.i("nt"), save = C.i("saveFreq");
    const bool gif = C.b("makeGif");
    check(ax_create(&p, &lat), "ax_create");
    std::vector<double> rows = read_numbers(D("ic.txt"), 2 * V);
    std::vector<double> host(4 * V, 0.0);   /* phi from the file, phit = 0 */
    const std::vector<double> c = columns(rows, V, 2);
    memcpy(host.data(), c.data(), 2 * V * sizeof(double));
    load_state(host);
    Out vals(D("valsPerLoop.txt"), o.digits);
    int frameNo = 0;
    for (int n = 0; n < nt; ++n) {
      check(ax_step(lat, 1), "ax_step");
      ax_diag d;
      check(ax_diagnostics(lat, 0, &d), "ax_diagnostics");
      vals.num(d.energy, true);
      if (gif && n % save == 0) {
        std::vector<double> st(4 * V);
        check(ax_get_state(lat, st.data(), st.size()), "ax_get_state");
        Out g(D("GifData/gifData_" + std::to_string(frameNo++) + ".txt"), o.digits);
        for (size_t s = 0; s < V; ++s) g.num(st[s] * st[s] + st[V + s] * st[V + s], true);
      }
    }
    std::vector<double> st(4 * V);
    check(ax_get_state(lat, st.data(), st.size()), "ax_get_state");
    write_rows(D("finalField.txt"), st, 2);
  }
};

int mkdir_p(const std::string &path) {
  std::string cur;
  for (size_t i = 0; i <= path.size(); ++i) {
    if (i == path.size() || path[i] == '/') {
      if (!cur.empty() && mkdir(cur.c_str(), 0755) != 0 && errno != EEXIST) return -1;
    }
    if (i < path.size()) cur += path[i];
  }
  return 0;
}

}  // namespace

int main(int argc, char **argv) {
  Run r;
  for (int i = 1; i < argc; ++i) {
    std::string a = argv[i];
    auto val = [&](const char *name) -> std::string {
      if (i + 1 >= argc) die(std::string("missing value for ") + name);
      return argv[++i];
    };
    if (a == "--program") r.o.program = val("--program");
    else if (a == "--dir") r.o.dir = val("--dir");
    else if (a == "--tag") r.o.tag = val("--tag");
    else if (a == "--digits") r.o.digits = atoi(val("--digits").c_str());
    else if (a == "-h" || a == "--help") {
      printf("usage: axions_run --program evolution|mpi_evolution|moore|leapfrog|sor|straight_ic [--dir D] [--tag T] "
             "[--digits N] [name=value ...]\n");
      return 0;
    } else if (a.find('=') != std::string::npos) r.o.kv[a.substr(0, a.find('='))] = a.substr(a.find('=') + 1);
    else die("unknown argument " + a);
  }
  r.C.v = defaults(r.o.program);
  for (auto &kv : r.o.kv) {
    if (!r.C.v.count(kv.first)) die("program " + r.o.program + " has no constant " + kv.first);
    r.C.v[kv.first] = kv.second;
  }
  for (const char *sub : {"Data", "GifData"})
    if (mkdir_p(r.o.dir + "/" + sub) != 0) die("cannot create " + r.o.dir + "/" + sub);
  const auto t0 = std::chrono::steady_clock::now();
  if (r.o.program == "evolution") r.evolution();
  else if (r.o.program == "mpi_evolution") r.mpi_evolution();
  else if (r.o.program == "moore") r.moore();
  else if (r.o.program == "sor") r.sor();
  else if (r.o.program == "straight_ic") r.straight_ic();
  else r.leapfrog();
  if (r.lat) {
    check(ax_sync(r.lat), "ax_sync");
    ax_destroy(r.lat);
  }
  const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  printf("Time taken: %.3fs\n", s);
  return 0;
}
