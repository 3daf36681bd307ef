// verifycurves: command-line drop-in for the reference tool (src/main.cpp,
// src/define_args.h).  Same flags and defaults, same output files and
// messages; the linking-number stage runs on the MI355X engine.
#include <chrono>
#include <iostream>

#include "flags.h"
#include "linking_number_certificate.h"
#include "model.h"

using namespace VerifyCurves;

static double ms_since(std::chrono::steady_clock::time_point t0) {
  return 1e-3 * std::chrono::duration_cast<std::chrono::microseconds>(
                    std::chrono::steady_clock::now() - t0).count();
}

int main(int argc, char **argv) {
  Flags F;
  F.add("input", "data/glove.bcc", "input model in the BCC file format");
  F.add("output", "results/out.txt", "output certificate file");
  F.add("comparewith", "[none]", "existing certificate to compare against");
  F.add("diffpairs", "[none]", "output file for inconsistent loop pairs");
  F.add("diffcurves", "[none]", "output file for loops in inconsistent pairs");
  F.add("method", "barneshut", "'barneshut' or 'directsum' (case sensitive)");
  F.add("discretize", "true", "discretize splines before computing (false: polyline input)", true);
  F.add("barnes_hut_init_beta", "2.0", "initial Barnes-Hut beta");
  F.add("barnes_hut_beta_limit", "10.0", "maximum Barnes-Hut beta");
  // engine options (no reference counterpart)
  F.add("criterion", "beta", "Barnes-Hut opening criterion: 'beta' or 'mlp' (learned)");
  F.add("fallback_tol", "0.05", "FP64 direct-sum judge for pairs farther than this from an integer");
  F.add("bh_direct_threshold", "0", "Barnes-Hut pairs with n1*n2 below this use direct summation");
  F.add("leaf_size", "1", "Barnes-Hut leaf bucket size");
  F.add("device", "0", "GPU device index");
  F.add("mlp_log_tol", "-6", "learned criterion: accept far field below this predicted log10 error");
  F.add("mlp_weights", "", "learned criterion weight file (default: the package's mlp_criterion.bin)");
  F.add("autotune", "false", "time candidate kernel launch shapes on this model's pairs first", true);
  try {
    F.parse(argc, argv);
  } catch (const std::exception &e) {
    std::cerr << "ERROR: " << e.what() << std::endl;
    return 1;
  }
  Model model;
  {
    std::cout << "Importing from " << F.str("input") << "." << std::endl;
    const auto t0 = std::chrono::steady_clock::now();
    model.ImportFromFile(F.str("input"));
    std::cout << "Import runtime: " << ms_since(t0) << " ms." << std::endl;
  }
  bool ds;
  if (F.str("method") == "directsum") ds = true;
  else if (F.str("method") == "barneshut") ds = false;
  else {
    std::cerr << "Invalid method: '" << F.str("method") << "'." << std::endl;
    throw std::runtime_error("Invalid method encountered.");
  }
  EngineOptions opt;
  if (F.str("criterion") == "mlp") opt.learned_criterion = true;
  else if (F.str("criterion") != "beta") throw std::runtime_error("Invalid criterion.");
  opt.fallback_tol = F.num("fallback_tol");
  opt.bh_direct_threshold = (int64_t)F.num("bh_direct_threshold");
  opt.leaf_size = (int)F.num("leaf_size");
  opt.device = (int)F.num("device");
  opt.mlp_log_tol = F.num("mlp_log_tol");
  opt.mlp_weights = F.str("mlp_weights");
  opt.autotune = F.flag("autotune");
  LinkingNumberCertificate certificate;
  {
    std::cout << "Computing the certificate using the " << (ds ? "Direct Summation" : "Barnes–Hut")
              << " Method." << std::endl;
    const auto t0 = std::chrono::steady_clock::now();
    certificate.ComputeFromModel(model, ds, F.flag("discretize"), F.num("barnes_hut_init_beta"),
                                 F.num("barnes_hut_beta_limit"), opt);
    std::cout << "Computation runtime: " << ms_since(t0) << " ms. This includes all 3 stages."
              << std::endl;
  }
  {
    std::cout << "Exporting the certificate to " << F.str("output") << "." << std::endl;
    const auto t0 = std::chrono::steady_clock::now();
    certificate.ExportToFile(F.str("output"));
    std::cout << "Export runtime: " << ms_since(t0) << " ms." << std::endl;
  }
  if (F.str("comparewith") != "[none]" &&
      (F.str("diffpairs") != "[none]" || F.str("diffcurves") != "[none]")) {
    LinkingNumberCertificate reference(F.str("comparewith"));
    certificate.ExportDiffsToFiles(F.str("diffpairs"), F.str("diffcurves"), reference);
    std::cout << "Compared and exported the difference between the model and the reference "
                 "certificate."
              << std::endl;
  }
  return 0;
}
