// kifmm_tt0: the reference's test driver (fmm3d_mpi/tt0.cpp:76-581) on one GPU,
// written against the public C ABI (include/kifmm.h) only.
//
//   kifmm_tt0 -options_file <file> [-option [value] ...]
//
// Options come from the file and the command line (the command line wins), as
// PETSc reads them: -numsrc, -numtrg, -numchk, -check_all, -kt, -distribution,
// -trgs_and_srcs_coinc, -skip_initial_sort, -do_2nd_eval, -dump_src_coord,
// -log_summary and their preload_ forms, -fmm3d_np, -fmm3d_let3d_ptsMax,
// -fmm3d_let3d_maxLevel, plus -kifmm_precision fp64|mixed and
// -kifmm_host_only (tree and lists only, no GPU).
//
// The run is the reference's: a dry run with the preload_ sizes, then the
// final run; points and densities from drand48 seeded with the rank (0), the
// same streams tt0 draws on rank 0 (tt0.cpp:153, 211-298, 351-362, 473-476);
// the sources sorted in Morton order at depth 30 with duplicates removed
// (tt0.cpp:302-326); setup, evaluate (twice with -do_2nd_eval), and the
// check against the direct sum with the line "Relative error is: %e".
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/kifmm.h"

namespace {

struct Opts {
  std::string file;
  std::map<std::string, std::string> cli;  // lower-case name -> value ("" for a flag)

  static std::string lower(std::string s) {
    for (auto& c : s) c = char(std::tolower(static_cast<unsigned char>(c)));
    return s;
  }
  // PetscOptionsGetString semantics: command line first, then the file.
  bool get(const std::string& name, std::string* v) const {
    auto it = cli.find(lower(name));
    if (it != cli.end()) { *v = it->second; return true; }
    if (file.empty()) return false;
    char buf[1024];
    const int rc = kifmm_options_lookup(file.c_str(), name.c_str(), buf, sizeof(buf));
    if (rc < 0) {
      std::fprintf(stderr, "options file %s: %s\n", file.c_str(), kifmm_last_error(nullptr));
      std::exit(1);
    }
    if (rc == 0) return false;
    *v = buf;
    return true;
  }
  bool has(const std::string& name) const { std::string v; return get(name, &v); }
  // with the preload_ prefix while preloading (PetscOptionsGetInt(pre, ...))
  bool get_long(const char* pre, const std::string& name, long* out) const {
    std::string v;
    const std::string full = pre ? "-" + std::string(pre) + name.substr(1) : name;
    if (!get(full, &v)) return false;
    char* end = nullptr;
    *out = std::strtol(v.c_str(), &end, 10);
    return end && *end == 0 && !v.empty();
  }
  bool has_pre(const char* pre, const std::string& name) const {
    return has(pre ? "-" + std::string(pre) + name.substr(1) : name);
  }
};

// Normal deviate of tt0.cpp:39-73 (three uniforms in the unit ball, two of the
// three results cached for the next calls).
double gaussian(double mean, double sd) {
  static double t1 = 0, t2 = 0;
  if (t1) { const double t = t1; t1 = 0; return mean + sd * t; }
  if (t2) { const double t = t2; t2 = 0; return mean + sd * t; }
  double x1, x2, x3, r;
  do {
    x1 = 2 * drand48() - 1;
    x2 = 2 * drand48() - 1;
    x3 = 2 * drand48() - 1;
    r = x1 * x1 + x2 * x2 + x3 * x3;
  } while (r >= 1);
  r = std::sqrt(-2.0 * std::log(r) / r);
  t1 = r * x2;
  t2 = r * x3;
  return mean + sd * r * x1;
}

// Sources of the named distribution (tt0.cpp:211-298); false if unknown.  lc:
// the centre of the "tt1" distribution, drawn first for every distribution
// (tt0.cpp:203-207) and shared with the "tt1" targets.
bool make_sources(const std::string& dist, long n, std::vector<double>& p, double (&lc)[3]) {
  p.assign(size_t(3 * n), 0.0);
  const double R = 0.25, r = 0.499 - R;
  for (int d = 0; d < 3; d++) lc[d] = 0.5 + R * (2.0 * drand48() - 1.0);
  if (dist == "tt1") {
    for (long k = 0; k < n; k++)
      for (int d = 0; d < 3; d++) p[3 * k + d] = lc[d] + r * (2.0 * drand48() - 1.0);
  } else if (dist == "uniform") {
    for (long k = 0; k < 3 * n; k++) p[k] = drand48();
  } else if (dist == "normal" || dist == "logNormal") {
    const bool ln = dist == "logNormal";
    for (long k = 0; k < n; k++) {
      double t[3];
      for (;;) {
        for (int d = 0; d < 3; d++) t[d] = ln ? std::exp(gaussian(-1.75, 1.0)) : gaussian(0.5, 0.16);
        bool ok = true;
        for (int d = 0; d < 3; d++) ok = ok && t[d] < 1.0 && (ln || t[d] >= 0);
        if (ok) break;
      }
      for (int d = 0; d < 3; d++) p[3 * k + d] = t[d];
    }
  } else if (dist == "sphereUniformAngles" || dist == "ellipseUniformAngles") {
    const double a = dist == "sphereUniformAngles" ? 1.0 : 0.25, rr = 0.49;
    for (long k = 0; k < n; k++) {
      const double phi = 2 * M_PI * drand48(), th = M_PI * drand48();
      p[3 * k] = 0.5 + a * rr * std::sin(th) * std::cos(phi);
      p[3 * k + 1] = 0.5 + a * rr * std::sin(th) * std::sin(phi);
      p[3 * k + 2] = 0.5 + rr * std::cos(th);
    }
  } else {
    return false;
  }
  return true;
}

// Morton order at depth 30 (Dendro TreeNode order: z, y, x bits interleaved, z
// most significant), duplicates of a depth-30 cell removed, each point moved to
// its cell centre (tt0.cpp:302-326).
void morton_dedup(std::vector<double>& p) {
  const int L = 30;
  const size_t n = p.size() / 3;
  struct Cell { unsigned x, y, z; unsigned __int128 key; };
  std::vector<Cell> c(n);
  for (size_t i = 0; i < n; i++) {
    Cell& e = c[i];
    e.x = unsigned(std::ldexp(p[3 * i], L));
    e.y = unsigned(std::ldexp(p[3 * i + 1], L));
    e.z = unsigned(std::ldexp(p[3 * i + 2], L));
    unsigned __int128 k = 0;
    for (int b = L - 1; b >= 0; b--)
      k = (k << 3) | ((e.z >> b) & 1u) << 2 | ((e.y >> b) & 1u) << 1 | ((e.x >> b) & 1u);
    e.key = k;
  }
  std::sort(c.begin(), c.end(), [](const Cell& a, const Cell& b) { return a.key < b.key; });
  c.erase(std::unique(c.begin(), c.end(), [](const Cell& a, const Cell& b) { return a.key == b.key; }), c.end());
  p.resize(3 * c.size());
  for (size_t i = 0; i < c.size(); i++) {
    p[3 * i] = std::ldexp(c[i].x + 0.5, -L);
    p[3 * i + 1] = std::ldexp(c[i].y + 0.5, -L);
    p[3 * i + 2] = std::ldexp(c[i].z + 0.5, -L);
  }
}

// PETSc's ASCII MATLAB view of a Vec (the -dump_src_coord file of tt0.cpp:334-348).
void dump_matlab(const char* path, const std::vector<double>& v) {
  FILE* f = std::fopen(path, "w");
  if (!f) { std::fprintf(stderr, "cannot write %s\n", path); std::exit(1); }
  std::fprintf(f, "Vec_0 = [\n");
  for (double x : v) std::fprintf(f, "%.17e\n", x);  // round-trips exactly
  std::fprintf(f, "];\n");
  std::fclose(f);
}

int die(kifmm_handle h, const char* what, int rc) {
  std::fprintf(stderr, "%s failed (%d): %s\n", what, rc, kifmm_last_error(h));
  return 1;
}

double secs_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

int main(int argc, char** argv) {
  Opts o;
  for (int i = 1; i < argc; i++) {
    if (argv[i][0] != '-') { std::fprintf(stderr, "unexpected argument %s\n", argv[i]); return 1; }
    std::string name = argv[i], val;
    if (i + 1 < argc && !(argv[i + 1][0] == '-' && !std::isdigit(static_cast<unsigned char>(argv[i + 1][1])) &&
                          argv[i + 1][1] != '.'))
      val = argv[++i];
    if (name == "-options_file") o.file = val;
    else o.cli[Opts::lower(name)] = val;
  }
  std::string dist, prec_s = "fp64";
  long kt = 0;
  if (!o.get("-distribution", &dist)) { std::fprintf(stderr, "option -distribution missing\n"); return 1; }
  if (!o.get_long(nullptr, "-kt", &kt)) { std::fprintf(stderr, "option -kt missing\n"); return 1; }
  long np = 0, pm = 0, ml = 0;
  if (!o.get_long(nullptr, "-fmm3d_np", &np) || !o.get_long(nullptr, "-fmm3d_let3d_ptsMax", &pm) ||
      !o.get_long(nullptr, "-fmm3d_let3d_maxLevel", &ml)) {
    std::fprintf(stderr, "options -fmm3d_np, -fmm3d_let3d_ptsMax and -fmm3d_let3d_maxLevel are required\n");
    return 1;
  }
  o.get("-kifmm_precision", &prec_s);
  if (prec_s != "fp64" && prec_s != "mixed") { std::fprintf(stderr, "-kifmm_precision must be fp64 or mixed\n"); return 1; }
  const int prec = prec_s == "mixed" ? KIFMM_PREC_MIXED : KIFMM_PREC_FP64;
  const bool host_only = o.has("-kifmm_host_only");

  for (int preloading = 1; preloading >= 0; preloading--) {
    const char* pre = preloading ? "preload_" : nullptr;
    std::printf("\n%s\n", preloading ? "*** Dry run (preloading) ***" : "*** Final run ***");
    srand48(0);  // the rank (tt0.cpp:153)
    long numsrc = 0;
    if (!o.get_long(pre, "-numsrc", &numsrc) || numsrc < 0) {
      std::fprintf(stderr, "option -%snumsrc missing\n", pre ? pre : "");
      return 1;
    }
    std::printf("Using numsrc=%ld\n", numsrc);
    std::vector<double> src;
    double lc[3];
    if (!make_sources(dist, numsrc, src, lc)) { std::fprintf(stderr, "Invalid distribution of sources selected\n"); return 1; }
    std::printf("Using %s distribution for sources\n", dist.c_str());
    if (!o.has("-skip_initial_sort")) {
      std::printf("Sorting the sources in morton order and removing duplicates\n");
      morton_dedup(src);
    }
    const long ns = long(src.size() / 3);
    std::printf("Total number of sources on all processors: %ld\n", ns);
    std::string dump;
    if (o.get(pre ? "-preload_dump_src_coord" : "-dump_src_coord", &dump) && !dump.empty()) {
      std::printf("Dumping source coordinates to file\n");
      dump_matlab(dump.c_str(), src);
    }
    // normals: three uniforms per point, normalised (tt0.cpp:351-362)
    std::vector<double> nor(size_t(3 * ns));
    for (long k = 0; k < ns; k++) {
      double* q = &nor[size_t(3 * k)];
      for (int d = 0; d < 3; d++) q[d] = drand48();
      const double nn = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
      for (int d = 0; d < 3; d++) q[d] /= nn;
    }
    std::vector<double> trg;
    const bool coinc = o.has("-trgs_and_srcs_coinc");
    if (coinc) {
      std::printf("Targets and sources coincide\n");
      trg = src;
    } else {
      long numtrg = 0;
      if (!o.get_long(pre, "-numtrg", &numtrg) || numtrg < 0) {
        std::fprintf(stderr, "option -%snumtrg missing\n", pre ? pre : "");
        return 1;
      }
      std::printf("Using numtrg=%ld\n", numtrg);
      trg.resize(size_t(3 * numtrg));
      if (dist == "tt1") {
        // tt1 targets around the sources' centre (tt0.cpp:389-396)
        const double r = 0.499 - 0.25;
        for (long k = 0; k < numtrg; k++)
          for (int d = 0; d < 3; d++) trg[size_t(3 * k + d)] = lc[d] + r * (2.0 * drand48() - 1.0);
      } else if (dist == "uniform") {
        for (auto& v : trg) v = drand48();
      } else {
        std::fprintf(stderr, "Invalid distribution of sources/targets selected\n");
        return 1;
      }
    }
    const long nt = long(trg.size() / 3);
    std::printf("Total number of targets on all processors: %ld\n", nt);

    const double coefs[2] = {1.0, 0.25};  // tt0.cpp:179-183
    const double ctr[3] = {0.5, 0.5, 0.5};
    kifmm_handle h = nullptr;
    int rc = kifmm_create(int(kt), coefs, 2, int(np), int(pm), int(ml), ctr, 1, prec, &h);
    if (rc) return die(nullptr, "create", rc);
    if ((rc = kifmm_set_points(h, ns, src.data(), nor.data(), nt, trg.data()))) return die(h, "set_points", rc);
    auto t0 = std::chrono::steady_clock::now();
    rc = host_only ? kifmm_setup_host(h) : kifmm_setup(h);
    if (rc) return die(h, "setup", rc);
    std::printf("FMM setup %.3f s\n", secs_since(t0));
    if (host_only) {
      int64_t info[8];
      kifmm_tree_info(h, info);
      std::printf("Tree: %lld nodes, depth %lld, lists U %lld V %lld W %lld X %lld\n", (long long)info[0],
                  (long long)info[1], (long long)info[2], (long long)info[3], (long long)info[4], (long long)info[5]);
      kifmm_destroy(h);
      continue;
    }
    const int sdof = kifmm_src_dof(h), tdof = kifmm_trg_dof(h);
    std::printf("srcDOF=%d trgDOF=%d\n", sdof, tdof);
    std::vector<double> den(size_t(ns) * sdof), val(size_t(nt) * tdof, 0.0);
    for (auto& v : den) v = drand48();  // tt0.cpp:473-476
    kifmm_set_timing(h, o.has("-log_summary") ? 1 : 0);
    t0 = std::chrono::steady_clock::now();
    if ((rc = kifmm_evaluate(h, den.data(), val.data()))) return die(h, "evaluate", rc);
    std::printf("FMM evaluate %.3f ms\n", secs_since(t0) * 1e3);
    if (o.has("-do_2nd_eval")) {
      t0 = std::chrono::steady_clock::now();
      if ((rc = kifmm_evaluate(h, den.data(), val.data()))) return die(h, "evaluate", rc);
      std::printf("FMM evaluate (2nd) %.3f ms\n", secs_since(t0) * 1e3);
    }
    if (o.has("-log_summary")) {
      float ms[KIFMM_NPHASE];
      kifmm_phase_times(h, ms, KIFMM_NPHASE);
      std::printf("Event                Time (ms)\n");
      for (int p = 0; p < KIFMM_NPHASE; p++) std::printf("%-20s %10.4f\n", kifmm_phase_name(p), ms[p]);
    }
    long numchk = 0;
    if (o.has_pre(pre, "-check_all")) {
      numchk = nt;
    } else if (!o.get_long(pre, "-numchk", &numchk)) {
      std::fprintf(stderr, "option -%snumchk missing\n", pre ? pre : "");
      return 1;
    }
    if (numchk > 0) {
      std::printf("Peeking globally %ld target points (out of %ld ) to check potential\n", numchk, nt);
      double rerr = 0;
      if ((rc = kifmm_check(h, den.data(), val.data(), int(std::min(numchk, nt)), 0, &rerr))) return die(h, "check", rc);
      std::printf("Relative error is: %e\n", rerr);
    } else {
      std::printf("No points to check potential, thus skipping check.\n");
    }
    kifmm_destroy(h);
  }
  return 0;
}
