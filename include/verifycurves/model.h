/*
 * model.h — drop-in for the reference's include/model.h: a curve collection
 * with BCC import/export (fp32 on disk, model.cpp:266-452) and OBJ import
 * (model.cpp:113-264).
 */
#pragma once
#include <string>
#include <vector>

#include "curve.h"

namespace VerifyCurves {

class Model {
 public:
  Model();
  void ImportFromFile(const std::string &filename);
  void ExportToFile(const std::string &filename) const;
  std::vector<Curve> &GetCurves() { return curves_; }
  const std::vector<Curve> &GetCurves() const { return curves_; }
  void ImportFromObjFile(const std::string &filename, SegmentType segment_type);

 private:
  void ImportBccFile(const std::string &filename);
  void ExportBccFile(const std::string &filename) const;
  std::vector<Curve> curves_;
};

}  // namespace VerifyCurves
