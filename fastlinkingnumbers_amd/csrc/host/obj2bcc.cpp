// obj2bcc: converts Blender/Houdini OBJ line data into a BCC curve collection
// (reference src/obj2bcc.cpp).
#include <chrono>
#include <iostream>

#include "flags.h"
#include "model.h"

using namespace VerifyCurves;

int main(int argc, char **argv) {
  Flags F;
  F.add("input", "data/houdiniobj.obj", "OBJ input file (.obj)");
  F.add("output", "data/houdiniobj.bcc", "BCC output file (.bcc)");
  F.add("segment_type", "PL", "C0, BS, BZ or PL (case sensitive)");
  try {
    F.parse(argc, argv);
  } catch (const std::exception &e) {
    std::cerr << "ERROR: " << e.what() << std::endl;
    return 1;
  }
  SegmentType t;
  const std::string s = F.str("segment_type");
  if (s == "C0") t = SegmentType::CatmullRom;
  else if (s == "BS") t = SegmentType::BSpline;
  else if (s == "BZ") t = SegmentType::Bezier;
  else if (s == "PL") t = SegmentType::Polyline;
  else throw std::runtime_error("Invalid segment_type. (Must be 'C0', 'BS', 'BZ', or 'PL'.)");
  Model model;
  auto t0 = std::chrono::steady_clock::now();
  std::cout << "Importing from " << F.str("input") << "." << std::endl;
  model.ImportFromObjFile(F.str("input"), t);
  std::cout << "Import runtime: "
            << 1e-3 * std::chrono::duration_cast<std::chrono::microseconds>(
                          std::chrono::steady_clock::now() - t0).count()
            << " ms." << std::endl;
  t0 = std::chrono::steady_clock::now();
  std::cout << "Exporting to " << F.str("output") << "." << std::endl;
  model.ExportToFile(F.str("output"));
  std::cout << "Export runtime: "
            << 1e-3 * std::chrono::duration_cast<std::chrono::microseconds>(
                          std::chrono::steady_clock::now() - t0).count()
            << " ms." << std::endl;
  return 0;
}
