// topopt_3d: the crate's CLI for the 3D isogeometric SIMP cantilever
// (rust/src/bin/topopt_3d.rs) on libicc.  Same arguments, progress lines and
// result file (benchmarks/topopt_result.json when ./benchmarks exists, else
// ./topopt_result.json; keys nelx, nely, nelz, solver, compliance, densities).
//
// Usage: topopt_3d [nelx nely nelz] [max_iter]   (default 24 12 6, 40 iterations)
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <sys/stat.h>
#include <vector>

#include "../../include/icc.h"

namespace {

using clk = std::chrono::steady_clock;

double ms_since(clk::time_point t) {
  return std::chrono::duration<double, std::milli>(clk::now() - t).count();
}

struct Progress {
  clk::time_point t_iter;
};

void report(int64_t it, double c, double vol, double change, void* user) {
  auto* p = static_cast<Progress*>(user);
  std::printf("  iter %3lld | compliance %12.4f | volfrac %.4f | change %.4e | %.0f ms\n", (long long)it, c, vol,
              change, ms_since(p->t_iter));
  std::fflush(stdout);
  p->t_iter = clk::now();
}

int fail(const char* what, int rc) {
  std::fprintf(stderr, "topopt_3d: %s failed (%d): %s\n", what, rc, icc_last_error());
  return 1;
}

void write_list(FILE* f, const std::vector<double>& v) {
  std::fputc('[', f);
  for (size_t i = 0; i < v.size(); ++i) std::fprintf(f, "%s%.17g", i ? ", " : "", v[i]);
  std::fputc(']', f);
}

}  // namespace

int main(int argc, char** argv) {
  std::vector<long long> a;
  for (int i = 1; i < argc; ++i) {
    char* end = nullptr;
    long long v = std::strtoll(argv[i], &end, 10);
    if (!end || *end != '\0' || v < 1) {
      std::fprintf(stderr, "topopt_3d: arguments must be positive integers: [nelx nely nelz] [max_iter]\n");
      return 2;
    }
    a.push_back(v);
  }
  icc_topopt_config cfg;
  icc_topopt_config_default(&cfg);  // CantileverTopOptConfig::default()
  if (a.size() >= 3) {
    cfg.nel[0] = a[0];
    cfg.nel[1] = a[1];
    cfg.nel[2] = a[2];
  } else {
    cfg.nel[0] = 24;
    cfg.nel[1] = 12;
    cfg.nel[2] = 6;
  }
  cfg.max_iter = a.size() >= 4 ? a[3] : 40;
  std::printf("3D isogeometric topology optimization: %lldx%lldx%lld elements, p = 2, volfrac %g, penal %g, rmin %g\n",
              (long long)cfg.nel[0], (long long)cfg.nel[1], (long long)cfg.nel[2], cfg.volfrac, cfg.penal, cfg.rmin);
  const auto t0 = clk::now();
  icc_ctx* ctx = nullptr;
  icc_topopt* t = nullptr;
  int rc = icc_ctx_create(0, &ctx);
  if (rc != ICC_OK) return fail("icc_ctx_create", rc);
  rc = icc_topopt_create(ctx, &cfg, &t);
  if (rc != ICC_OK) return fail("icc_topopt_create", rc);
  int64_t sz[2];
  icc_topopt_sizes(t, sz);
  std::printf("setup: %lld DOFs, %.0f ms\n", (long long)sz[1], ms_since(t0));
  const int64_t mi = cfg.max_iter;
  std::vector<double> comp(mi), change(mi), dens(sz[0]);
  std::vector<int64_t> its(mi);
  int64_t n = 0;
  Progress prog{clk::now()};
  rc = icc_topopt_run(t, report, &prog, comp.data(), change.data(), its.data(), dens.data(), &n);
  if (rc != ICC_OK) return fail("icc_topopt_run", rc);
  comp.resize(n);
  std::printf("finished %lld iterations in %.2f s (PCG iterations per solve: [", (long long)n, ms_since(t0) / 1e3);
  for (int64_t i = 0; i < n; ++i) std::printf("%s%lld", i ? ", " : "", (long long)its[i]);
  std::printf("])\n");
  struct stat sb;
  const std::string out =
      (stat("benchmarks", &sb) == 0 && S_ISDIR(sb.st_mode)) ? "benchmarks/topopt_result.json" : "topopt_result.json";
  FILE* f = std::fopen(out.c_str(), "w");
  if (!f) {
    std::fprintf(stderr, "topopt_3d: cannot write %s\n", out.c_str());
    return 1;
  }
  std::fprintf(f, "{\"nelx\": %lld, \"nely\": %lld, \"nelz\": %lld, ", (long long)cfg.nel[0], (long long)cfg.nel[1],
               (long long)cfg.nel[2]);
  std::fprintf(f, "\"solver\": \"immersed_catmull_clark_amd (MI355X, B-spline p = 2, SIMP, OC)\", \"compliance\": ");
  write_list(f, comp);
  std::fprintf(f, ", \"densities\": ");
  write_list(f, dens);
  std::fprintf(f, "}");
  std::fclose(f);
  std::printf("saved %s\n", out.c_str());
  icc_topopt_destroy(t);
  icc_ctx_destroy(ctx);
  return 0;
}
