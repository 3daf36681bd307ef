/*
 * linking_number_certificate.h — drop-in for the reference's
 * include/linking_number_certificate.h:26-133.
 *
 * ComputeFromModel runs potential-link search, discretization and the
 * linking-number stage on the MI355X engine (libflk.so, include/flk.h).  The
 * certificate is the sparse lower-triangular integer matrix of pairwise
 * linking numbers, kept as sorted (col, row, value) entries with row > col
 * (the reference keeps an Eigen::SparseMatrix<int>).
 */
#pragma once
#include <cstdint>
#include <string>
#include <tuple>
#include <vector>

#include "model.h"

namespace VerifyCurves {

/* Engine options beyond the reference's arguments (defaults reproduce the
 * reference; see include/flk.h for their meaning). */
struct EngineOptions {
  bool learned_criterion = false;  // MLP opening criterion instead of the beta test
  double fallback_tol = 0.05;      // FP64 direct-sum judge for pairs off-integer by more
  int64_t bh_direct_threshold = 0; // Barnes–Hut pairs with n1*n2 below this use direct sum
  int leaf_size = 1;               // Barnes–Hut leaf bucket size
  int device = 0;
  bool verbose = true;             // print the reference's progress lines
  double mlp_log_tol = -6.0;       // learned criterion: far field below this log10 error
  std::string mlp_weights;         // weight file; empty = mlp_criterion.bin of the package
  bool autotune = false;           // time candidate kernel shapes on this model's pairs first
};

class LinkingNumberCertificate {
 public:
  LinkingNumberCertificate();
  explicit LinkingNumberCertificate(const std::string &filename);
  void ComputeFromModel(const Model &model, bool force_direct_sum = false, bool discretize = true,
                        double barnes_hut_init_beta = 2.0, double barnes_hut_beta_limit = 10.0);
  void ComputeFromModel(const Model &model, bool force_direct_sum, bool discretize,
                        double barnes_hut_init_beta, double barnes_hut_beta_limit,
                        const EngineOptions &options);
  void ExportToFile(const std::string &filename) const;
  void ImportFromFile(const std::string &filename);
  void ExportDiffsToFiles(const std::string &out_diff_links_filename,
                          const std::string &out_diff_curves_filename,
                          const LinkingNumberCertificate &other) const;
  bool Equals(const LinkingNumberCertificate &other) const;
  bool operator==(const LinkingNumberCertificate &other) const { return Equals(other); }
  bool operator!=(const LinkingNumberCertificate &other) const { return !Equals(other); }

  /* (col i, row j, value) entries with i < j, sorted by (i, j), zeros pruned. */
  typedef std::tuple<int, int, int> Entry;
  const std::vector<Entry> &entries() const { return entries_; }
  int size() const { return size_; }
  double last_mean_off_integer_error() const { return mean_err_; }
  double last_max_off_integer_error() const { return max_err_; }

 private:
  int size_ = 0;
  std::vector<Entry> entries_;
  double mean_err_ = 0, max_err_ = 0;
};

}  // namespace VerifyCurves
