// Certificate driver of the drop-in API (reference
// src/linking_number_certificate.cpp): potential-link search, discretization,
// the linking-number stage on the MI355X engine, rounding and file I/O.
#include "linking_number_certificate.h"

#include <dlfcn.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <iostream>
#include <map>
#include <mutex>
#include <set>
#include <sstream>

#include "host_internal.h"

namespace VerifyCurves {

// ---------------------------------------------------------------- engine
namespace {
std::mutex g_engine_mu;
std::map<int, flk_ctx *> g_engines;
}  // namespace

flk_ctx *engine(int device) {
  std::lock_guard<std::mutex> lk(g_engine_mu);
  auto it = g_engines.find(device);
  if (it != g_engines.end()) return it->second;
  flk_ctx *ctx = nullptr;
  const int32_t rc = flk_create(device, &ctx);
  if (rc != FLK_OK)
    throw std::runtime_error(std::string("MI355X engine unavailable: ") + flk_last_error(nullptr));
  g_engines[device] = ctx;
  return ctx;
}

void engine_check(flk_ctx *ctx, int32_t rc, const char *what) {
  if (rc != FLK_OK) throw std::runtime_error(std::string(what) + ": " + flk_last_error(ctx));
}

// ---------------------------------------------------------------- helpers
namespace {

// Learned-criterion weight file: "FLKMLP2\0", 5,057 + 24 float32 (little
// endian) of the hard-swish network.  The default lives beside the package's
// lib/ directory.  "FLKMLP1\0" files hold an earlier SiLU network's weights.
void load_mlp_weights(flk_ctx *ctx, std::string path) {
  if (path.empty()) {
    Dl_info info;
    if (dladdr((void *)&load_mlp_weights, &info) && info.dli_fname) {
      std::string lib = info.dli_fname;
      const size_t slash = lib.rfind('/');
      path = (slash == std::string::npos ? std::string(".") : lib.substr(0, slash)) +
             "/../mlp_criterion.bin";
    }
  }
  std::ifstream f(path, std::ios::binary);
  if (!f) throw std::runtime_error("cannot open MLP weight file " + path);
  char magic[8];
  std::vector<float> buf(FLK_MLP_NPARAM + 2 * FLK_MLP_NFEAT);
  f.read(magic, 8);
  f.read(reinterpret_cast<char *>(buf.data()), sizeof(float) * buf.size());
  if (f && std::memcmp(magic, "FLKMLP1\0", 8) == 0)
    throw std::runtime_error("MLP weight file of the earlier SiLU network, retrain: " + path);
  if (!f || std::memcmp(magic, "FLKMLP2\0", 8) != 0)
    throw std::runtime_error("not an MLP weight file: " + path);
  engine_check(ctx, flk_mlp_set_weights(ctx, buf.data(), buf.data() + FLK_MLP_NPARAM),
               "flk_mlp_set_weights");
}

// Uploads the start points of every curve (the polyline the linking-number
// stage integrates; reference compute_link.cpp:204-205).
void upload(flk_ctx *ctx, const std::vector<Curve> &curves) {
  std::vector<int64_t> off(curves.size() + 1, 0);
  for (size_t i = 0; i < curves.size(); ++i) off[i + 1] = off[i] + (int64_t)curves[i].get_q().size();
  std::vector<double> xyz(3 * (size_t)off.back());
#pragma omp parallel for schedule(dynamic)
  for (size_t i = 0; i < curves.size(); ++i) {
    const Curve &c = curves[i];
    if (c.get_q().empty()) continue;
    if (c.get_segment_type() == SegmentType::Polyline) {
      std::copy(c.get_q()[0].data(), c.get_q()[0].data() + 3 * c.get_q().size(), xyz.data() + 3 * off[i]);
    } else if (c.is_looped()) {
      const auto s = c.GetStartPoints();
      std::copy(s[0].data(), s[0].data() + 3 * s.size(), xyz.data() + 3 * off[i]);
    } else {
      for (size_t k = 0; k < c.get_q().size(); ++k)
        for (int d = 0; d < 3; ++d) xyz[3 * (off[i] + k) + d] = c.get_q()[k][d];
    }
  }
  static_assert(sizeof(Point3) == 3 * sizeof(double), "Point3 must be 3 packed doubles");
  engine_check(ctx, flk_set_curves(ctx, xyz.data(), off.data(), (int32_t)curves.size()),
               "flk_set_curves");
}

// Potential links of the curves (their exact boxes computed here), or, with
// curves == nullptr, of the engine's current n curves (device boxes).
std::vector<std::pair<int, int>> potential_links(flk_ctx *ctx, const std::vector<Curve> *curves, int32_t n) {
  const bool device_boxes = curves == nullptr;
  std::vector<double> boxes;
  if (!device_boxes) {
    boxes.resize(6 * (size_t)n);
#pragma omp parallel for schedule(dynamic)
    for (int32_t i = 0; i < n; ++i) {
      (*curves)[i].ComputeCurveBoundingBox();
      const Box3d b = (*curves)[i].get_last_curve_bounding_box();
      for (int d = 0; d < 3; ++d) {
        boxes[6 * i + d] = b.min()[d];
        boxes[6 * i + 3 + d] = b.max()[d];
      }
    }
  }
  const int32_t *p = nullptr;
  int64_t m = 0;
  engine_check(ctx, flk_potential_links(ctx, device_boxes ? nullptr : boxes.data(), n, &p, &m),
               "flk_potential_links");
  std::vector<std::pair<int, int>> out(m);
  for (int64_t k = 0; k < m; ++k) out[k] = {p[2 * k], p[2 * k + 1]};
  return out;
}

// Discretize (reference src/discretize.cpp:102-281) on the device: the
// engine bisects the pieces, evaluates the polyline vertices and keeps them as
// its curve set; this prints the reference's progress and failure messages
// and returns the looped flag of each output polyline.
std::vector<uint8_t> discretize_curves(flk_ctx *ctx, const std::vector<Curve> &in,
                                       const std::vector<std::pair<int, int>> &links, bool verbose) {
  const int32_t nc = (int32_t)in.size();
  std::vector<int64_t> off(nc + 1, 0);
  std::vector<int32_t> types(nc);
  std::vector<uint8_t> looped(nc);
  for (int32_t c = 0; c < nc; ++c) {
    off[c + 1] = off[c] + (int64_t)in[c].get_q().size();
    types[c] = static_cast<int32_t>(in[c].get_segment_type());
    looped[c] = in[c].is_looped() ? 1 : 0;
  }
  std::vector<double> ctrl(3 * (size_t)off[nc]);
#pragma omp parallel for schedule(dynamic)
  for (int32_t c = 0; c < nc; ++c)
    if (!in[c].get_q().empty())
      std::copy(in[c].get_q()[0].data(), in[c].get_q()[0].data() + 3 * in[c].get_q().size(),
                ctrl.data() + 3 * off[c]);
  std::vector<int32_t> pairs(2 * links.size());
  for (size_t k = 0; k < links.size(); ++k) {
    pairs[2 * k] = links[k].first;
    pairs[2 * k + 1] = links[k].second;
  }
  const int64_t *voff = nullptr;
  flk_disc_info info;
  const int32_t rc = flk_discretize(ctx, ctrl.data(), off.data(), types.data(), looped.data(), nc, pairs.data(),
                                    (int64_t)links.size(), nullptr, &voff, &info);
  if (verbose)
    for (int32_t r = 0; r < info.rounds && r < 55; ++r)
      if (rc == FLK_OK || r + 1 < info.rounds || info.small_curve >= 0 || info.rounds == 55)
        std::cout << "Iteration " << r + 1 << ", " << info.overlaps[r] << " overlaps." << std::endl;
  if (rc != FLK_OK) {
    const std::string msg = flk_last_error(ctx);
    if (info.small_curve >= 0) {
      std::cerr << "Discretize: Smallest input segment, (" << info.small_curve << "," << info.small_segment
                << ") has a size of " << info.small_size
                << ", which is smaller than eps * avg coordinate. Consider combining them into one vertex."
                << std::endl;
    } else if (msg.find("Smallest segment") != std::string::npos) {
      std::cout << "Discretize: Smallest segment is smaller than eps * avg coordinate. Perhaps curves are too close."
                << std::endl;
    } else if (msg.find("max recursion") != std::string::npos) {
      std::cout << "Max recursion (curves are too close)." << std::endl;
      std::cout << "Number of overlaps: " << info.overlaps[54] << std::endl;
    }
    throw std::runtime_error(msg);
  }
  if (verbose) {
    std::cout << "Finished after " << info.rounds << " iteration(s)." << std::endl;
    std::cout << "Original segment count: " << info.segments << ", discretized count: " << info.pieces
              << std::endl;
  }
  // an open input curve with pieces stays open (its last end point closes the list)
  std::vector<uint8_t> out(nc);
  for (int32_t c = 0; c < nc; ++c) out[c] = (in[c].is_looped() || voff[c + 1] == voff[c]) ? 1 : 0;
  return out;
}

void save_entries(const std::string &filename, int size,
                  const std::vector<LinkingNumberCertificate::Entry> &e, bool omit_size) {
  std::ofstream f(filename);
  if (!omit_size) f << size << "\n";
  for (const auto &t : e) f << std::get<0>(t) << "," << std::get<1>(t) << "," << std::get<2>(t) << "\n";
}

// Sorted by column (first index) then row, as a column-major sparse matrix
// iterates.
void sort_entries(std::vector<LinkingNumberCertificate::Entry> &e) {
  std::sort(e.begin(), e.end(), [](const LinkingNumberCertificate::Entry &a,
                                   const LinkingNumberCertificate::Entry &b) {
    return std::get<0>(a) != std::get<0>(b) ? std::get<0>(a) < std::get<0>(b)
                                            : std::get<1>(a) < std::get<1>(b);
  });
}

}  // namespace

// ---------------------------------------------------------------- class
LinkingNumberCertificate::LinkingNumberCertificate() {}

LinkingNumberCertificate::LinkingNumberCertificate(const std::string &filename) {
  ImportFromFile(filename);
}

void LinkingNumberCertificate::ComputeFromModel(const Model &model, bool force_direct_sum,
                                                bool discretize, double beta_init,
                                                double beta_limit) {
  ComputeFromModel(model, force_direct_sum, discretize, beta_init, beta_limit, EngineOptions());
}

void LinkingNumberCertificate::ComputeFromModel(const Model &model, bool force_direct_sum,
                                                bool discretize, double beta_init,
                                                double beta_limit, const EngineOptions &opt) {
  const bool v = opt.verbose;
  if (v) std::cout << "Starting Potential Link Search." << std::endl;
  flk_ctx *ctx = engine(opt.device);
  const std::vector<Curve> &in = model.GetCurves();
  const int nc = (int)in.size();
  std::vector<std::pair<int, int>> pairs;
  std::vector<uint8_t> looped(nc);
  bool all_pl = true;
  for (const auto &c : in) all_pl = all_pl && c.get_segment_type() == SegmentType::Polyline;
  if (discretize) {
    auto first = potential_links(ctx, &in, nc);
    if (v) std::cout << "Starting Discretization." << std::endl;
    looped = discretize_curves(ctx, in, first, v);
    pairs = potential_links(ctx, nullptr, nc);
  } else {
    upload(ctx, in);
    pairs = potential_links(ctx, all_pl ? nullptr : &in, nc);
    for (int c = 0; c < nc; ++c) looped[c] = in[c].is_looped() ? 1 : 0;
    if (v) std::cout << "Skipping Discretization." << std::endl;
  }
  for (auto &p : pairs)
    if (!looped[p.first] || !looped[p.second])
      throw std::runtime_error("Curve::GetStartPoints() does not support unlooped curves.");

  flk_params prm;
  flk_default_params(&prm);
  prm.method = force_direct_sum ? FLK_DIRECTSUM : FLK_BARNESHUT;
  prm.criterion = opt.learned_criterion ? FLK_CRIT_MLP : FLK_CRIT_BETA;
  prm.beta_init = beta_init;
  prm.beta_limit = beta_limit;
  prm.fallback_tol = opt.fallback_tol;
  prm.bh_direct_threshold = opt.bh_direct_threshold;
  prm.leaf_size = opt.leaf_size;
  prm.mlp_log_tol = opt.mlp_log_tol;
  if (opt.learned_criterion && !force_direct_sum) load_mlp_weights(ctx, opt.mlp_weights);
  if (v) std::cout << (force_direct_sum ? "Starting Direct Summation." : "Starting Barnes–Hut.") << std::endl;
  std::vector<int32_t> flat(2 * pairs.size());
  for (size_t k = 0; k < pairs.size(); ++k) {
    flat[2 * k] = pairs[k].first;
    flat[2 * k + 1] = pairs[k].second;
  }
  std::vector<double> lk(pairs.size()), err(pairs.size());
  std::vector<uint32_t> flags(pairs.size());
  if (opt.autotune) {
    flk_tune t;
    engine_check(ctx, flk_autotune(ctx, flat.data(), (int64_t)pairs.size(), &prm, &t), "flk_autotune");
    if (v)
      std::cout << "Autotuned shapes: direct sum " << t.ds_rows_per_lane << " rows per lane, "
                << t.ds_blocks_per_cu << " blocks per CU (0 = all resident), " << t.ds_tasks_per_wave
                << " tasks per wave; Barnes–Hut " << t.bh_items_per_thread << " items per thread." << std::endl;
  }
  engine_check(ctx, flk_linking_numbers(ctx, flat.data(), (int64_t)pairs.size(), &prm, lk.data(),
                                        err.data(), flags.data()),
               "flk_linking_numbers");

  // Rounding (reference linking_number_certificate.cpp:185-228): nearest
  // integer, kept when |round| >= 0.25, a NaN reported and kept.
  entries_.clear();
  double sum = 0, mx = 0;
  for (size_t k = 0; k < pairs.size(); ++k) {
    const double r = std::round(lk[k]);
    const double e = std::fabs(lk[k] - r);
    sum += e;
    mx = std::max(mx, e);
    if (!(std::fabs(r) < 0.25)) {
      if (std::isnan(r)) {
        std::cerr << "Linking number (" << pairs[k].first << "," << pairs[k].second << ") is NaN."
                  << std::endl;
        entries_.emplace_back(pairs[k].first, pairs[k].second, std::numeric_limits<int>::min());
      } else {
        entries_.emplace_back(pairs[k].first, pairs[k].second, (int)r);
      }
    }
  }
  mean_err_ = pairs.empty() ? 0.0 : sum / (double)pairs.size();
  max_err_ = mx;
  if (v && !pairs.empty()) {
    std::cout << "Linking-Number Computation, Mean Off-Integer Error: " << mean_err_ << std::endl;
    std::cout << "Linking-Number Computation, Max  Off-Integer Error: " << max_err_ << std::endl;
    flk_stats st;
    flk_get_stats(ctx, &st);
    if (st.fallbacks)
      std::cout << "FP64 direct-sum judge recomputed " << st.fallbacks << " pair(s)." << std::endl;
  }
  sort_entries(entries_);
  size_ = nc;
}

void LinkingNumberCertificate::ExportToFile(const std::string &filename) const {
  save_entries(filename, size_, entries_, false);
}

// "n" then "col,row,value" lines (reference linking_number_certificate.cpp:402-430);
// repeated entries add up.
void LinkingNumberCertificate::ImportFromFile(const std::string &filename) {
  std::ifstream f(filename);
  std::string line;
  std::map<std::pair<int, int>, int> acc;
  size_ = 0;
  if (std::getline(f, line)) {
    size_ = std::atoi(line.c_str());
    while (std::getline(f, line)) {
      std::stringstream ls(line);
      std::string a, b, c;
      if (std::getline(ls, a, ',') && std::getline(ls, b, ',') && std::getline(ls, c, ','))
        acc[{std::atoi(a.c_str()), std::atoi(b.c_str())}] += std::atoi(c.c_str());
    }
  }
  entries_.clear();
  for (auto &kv : acc) entries_.emplace_back(kv.first.first, kv.first.second, kv.second);
  sort_entries(entries_);
}

bool LinkingNumberCertificate::Equals(const LinkingNumberCertificate &other) const {
  if (size_ != other.size_) return false;
  std::map<std::pair<int, int>, long long> d;
  for (auto &e : entries_) d[{std::get<0>(e), std::get<1>(e)}] += std::get<2>(e);
  for (auto &e : other.entries_) d[{std::get<0>(e), std::get<1>(e)}] -= std::get<2>(e);
  for (auto &kv : d)
    if (kv.second != 0) return false;
  return true;
}

// Differences and the participating curves (reference
// linking_number_certificate.cpp:520-561).  As in the reference, the curve list
// names curves with a nonzero entry in this certificate, not in the difference.
void LinkingNumberCertificate::ExportDiffsToFiles(const std::string &pairs_file,
                                                  const std::string &curves_file,
                                                  const LinkingNumberCertificate &other) const {
  if (size_ != other.size_) {
    std::ostringstream msg;
    msg << "Unable to compare: The two models have different numbers of curves: " << size_
        << " vs. " << other.size_ << ".";
    std::cerr << msg.str() << std::endl;
    std::ofstream(pairs_file) << msg.str() << std::endl;
    std::ofstream(curves_file) << msg.str() << std::endl;
    return;
  }
  std::map<std::pair<int, int>, long long> d;
  for (auto &e : entries_) d[{std::get<0>(e), std::get<1>(e)}] += std::get<2>(e);
  for (auto &e : other.entries_) d[{std::get<0>(e), std::get<1>(e)}] -= std::get<2>(e);
  std::vector<Entry> diff;
  for (auto &kv : d)
    if (kv.second != 0) diff.emplace_back(kv.first.first, kv.first.second, (int)kv.second);
  sort_entries(diff);
  if (pairs_file != "[none]") save_entries(pairs_file, size_, diff, true);
  if (curves_file != "[none]") {
    std::set<int> active;
    for (auto &e : entries_)
      if (std::get<2>(e) != 0) {
        active.insert(std::get<0>(e));
        active.insert(std::get<1>(e));
      }
    std::ofstream f(curves_file);
    for (int i : active)
      if (i >= 0 && i < size_) f << i << "\n";
  }
}

}  // namespace VerifyCurves
