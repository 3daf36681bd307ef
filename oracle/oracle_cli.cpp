/*
 * CPU ORACLE command line — test infrastructure only.  Mirrors the flags of
 * the reference's verifycurves (src/define_args.h:9-55, src/main.cpp:14-99) so
 * the tests can produce oracle certificates for any input file.  Extra flags:
 * --obj_segment_type (import an OBJ directly and round it through fp32 BCC as
 * obj2bcc does, src/obj2bcc.cpp:27-77).
 */
#include "oracle.h"

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

static bool starts(const std::string &s, const char *p) { return s.rfind(p, 0) == 0; }

int main(int argc, char **argv) {
  std::string input = "data/glove.bcc", output = "results/out.txt", method = "barneshut";
  std::string obj_type;
  bool discretize = true;
  double beta_init = 2.0, beta_limit = 10.0;
  for (int k = 1; k < argc; ++k) {
    std::string a = argv[k];
    while (starts(a, "-")) a = a.substr(1);
    std::string key = a, val;
    bool has = false;
    size_t eq = a.find('=');
    if (eq != std::string::npos) { key = a.substr(0, eq); val = a.substr(eq + 1); has = true; }
    auto next = [&]() -> std::string {
      if (has) return val;
      if (k + 1 >= argc) { std::fprintf(stderr, "missing value for %s\n", key.c_str()); std::exit(2); }
      return argv[++k];
    };
    if (key == "input") input = next();
    else if (key == "output") output = next();
    else if (key == "method") method = next();
    else if (key == "obj_segment_type") obj_type = next();
    else if (key == "barnes_hut_init_beta") beta_init = std::atof(next().c_str());
    else if (key == "barnes_hut_beta_limit") beta_limit = std::atof(next().c_str());
    else if (key == "discretize") discretize = has ? (val == "true" || val == "1") : true;
    else if (key == "nodiscretize") discretize = false;
    else { std::fprintf(stderr, "unknown flag %s\n", argv[k]); return 2; }
  }
  bool ds;
  if (method == "directsum") ds = true;
  else if (method == "barneshut") ds = false;
  else { std::fprintf(stderr, "Invalid method: '%s'.\n", method.c_str()); return 3; }
  orc_model *m = orc_model_new();
  int rc;
  if (!obj_type.empty()) {
    int t = obj_type == "C0" ? ORC_CATMULLROM : obj_type == "BS" ? ORC_BSPLINE
          : obj_type == "BZ" ? ORC_BEZIER : ORC_POLYLINE;
    rc = orc_model_load_obj(m, input.c_str(), t);
    if (rc == 0) {
      std::string tmp = output + ".tmp.bcc";
      rc = orc_model_save_bcc(m, tmp.c_str());
      if (rc == 0) rc = orc_model_load_bcc(m, tmp.c_str());
      std::remove(tmp.c_str());
    }
  } else {
    rc = orc_model_load_bcc(m, input.c_str());
  }
  if (rc) { std::fprintf(stderr, "import failed: %s\n", orc_last_error()); return 4; }
  auto t0 = std::chrono::steady_clock::now();
  int64_t cap = 1 << 22;
  std::vector<int32_t> trip(3 * cap);
  double me = 0, mx = 0;
  int64_t nc = 0;
  int64_t nt = orc_certificate(m, ds, discretize, beta_init, beta_limit, trip.data(), cap, &me,
                               &mx, nullptr, nullptr, 0, &nc);
  auto t1 = std::chrono::steady_clock::now();
  if (nt < 0) { std::fprintf(stderr, "compute failed: %s\n", orc_last_error()); return 5; }
  std::printf("Computation runtime: %g ms. This includes all 3 stages.\n",
              1e-3 * std::chrono::duration_cast<std::chrono::microseconds>(t1 - t0).count());
  FILE *f = std::fopen(output.c_str(), "w");
  if (!f) return 6;
  std::fprintf(f, "%d\n", orc_model_ncurves(m));
  for (int64_t k = 0; k < nt; ++k) std::fprintf(f, "%d,%d,%d\n", trip[3 * k], trip[3 * k + 1], trip[3 * k + 2]);
  std::fclose(f);
  orc_model_free(m);
  return 0;
}
