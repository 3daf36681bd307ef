// Curve collections and their file formats (reference src/model.cpp).
//
// BCC: a 64-byte header ("BCC", 0x44 = 4-byte ints and floats, curve type
// C0/BS/BZ/PL, 3 dimensions, up axis, curve and point counts, 40 bytes of
// info), then per curve an int32 control-point count (negative = closed loop)
// and that many xyz float triples (reference model.cpp:266-452).
#include "model.h"

#include <cctype>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>

namespace VerifyCurves {

namespace {

#pragma pack(push, 1)
struct BccHeader {
  char sign[3];
  unsigned char bytes;
  char type[2];
  char dims;
  char up;
  uint64_t ncurves;
  uint64_t npoints;
  char info[40];
};
#pragma pack(pop)
static_assert(sizeof(BccHeader) == 64, "BCC header is 64 bytes");

std::string lower_ext(const std::string &name) {
  const size_t dot = name.rfind('.');
  if (dot == std::string::npos) throw std::out_of_range("file name has no extension: " + name);
  std::string e = name.substr(dot);
  for (auto &c : e) c = (char)std::tolower((unsigned char)c);
  return e;
}

std::string trim(const std::string &s) {
  const char *ws = " \t\v";
  const size_t b = s.find_first_not_of(ws);
  if (b == std::string::npos) return "";
  const size_t e = s.find_last_not_of(ws);
  return s.substr(b, e + 1 - b);
}

}  // namespace

Model::Model() {}

void Model::ImportFromFile(const std::string &filename) {
  if (lower_ext(filename) != ".bcc") {
    std::cerr << "File extension for import must be .bcc: " << filename << std::endl;
    throw std::runtime_error("Input model file extension error.");
  }
  ImportBccFile(filename);
}

void Model::ExportToFile(const std::string &filename) const {
  if (lower_ext(filename) != ".bcc") {
    std::cerr << "File extension for export must be .bcc: " << filename << std::endl;
    throw std::runtime_error("Output model file extension error.");
  }
  ExportBccFile(filename);
}

void Model::ImportBccFile(const std::string &filename) {
  FILE *f = std::fopen(filename.c_str(), "rb");
  if (!f) {
    std::cerr << "Error opening input file: " << filename << "\n" << std::endl;
    throw std::runtime_error("Unable to load BCC file.");
  }
  auto fail = [&](const std::string &why) {
    std::fclose(f);
    std::cerr << why << ": " << filename << "\n" << std::endl;
    throw std::runtime_error("Unable to load BCC file.");
  };
  BccHeader h;
  if (std::fread(&h, sizeof h, 1, f) != 1) fail("Incomplete header");
  if (h.sign[0] != 'B' || h.sign[1] != 'C' || h.sign[2] != 'C') fail("Invalid header");
  if (h.bytes != 0x44) fail("Invalid bytecount. We only support 32-bit integers and floats.");
  if (h.dims != 3) fail("Invalid dimensions. We only support data in 3D Cartesian space.");
  SegmentType type;
  if (h.type[0] == 'C' && h.type[1] == '0') type = SegmentType::CatmullRom;
  else if (h.type[0] == 'B' && h.type[1] == 'S') type = SegmentType::BSpline;
  else if (h.type[0] == 'B' && h.type[1] == 'Z') type = SegmentType::Bezier;
  else if (h.type[0] == 'P' && h.type[1] == 'L') type = SegmentType::Polyline;
  else fail("Invalid valid curve segment type. We only support C0, BS, BZ, and PL.");
  curves_.assign(h.ncurves, Curve());
  std::vector<float> buf;
  for (size_t i = 0; i < curves_.size(); ++i) {
    int32_t cnt;
    if (std::fread(&cnt, sizeof cnt, 1, f) != 1) fail("Cannot read curve control point count");
    curves_[i].set_looped(cnt < 0);
    curves_[i].set_segment_type(type);
    if (cnt < 0) cnt = -cnt;
    if (cnt == 0) continue;
    buf.resize(3 * (size_t)cnt);
    if (std::fread(buf.data(), sizeof(float), buf.size(), f) != buf.size())
      fail("Invalid number of control points read");
    auto &q = curves_[i].get_q();
    q.resize(cnt);
    for (int32_t k = 0; k < cnt; ++k) q[k] = Point3(buf[3 * k], buf[3 * k + 1], buf[3 * k + 2]);
  }
  std::fclose(f);
}

void Model::ExportBccFile(const std::string &filename) const {
  FILE *f = std::fopen(filename.c_str(), "wb");
  if (!f) {
    std::cerr << "Error creating output file: " << filename << std::endl;
    throw std::runtime_error("Unable to export BCC file.");
  }
  BccHeader h;
  std::memset(&h, 0, sizeof h);
  h.sign[0] = 'B'; h.sign[1] = 'C'; h.sign[2] = 'C';
  h.bytes = 0x44;
  h.type[0] = 'C'; h.type[1] = '0';
  if (!curves_.empty()) {
    switch (curves_[0].get_segment_type()) {
      case SegmentType::BSpline: h.type[0] = 'B'; h.type[1] = 'S'; break;
      case SegmentType::Bezier: h.type[0] = 'B'; h.type[1] = 'Z'; break;
      case SegmentType::Polyline: h.type[0] = 'P'; h.type[1] = 'L'; break;
      default: break;
    }
  }
  h.dims = 3;
  h.up = 1;
  h.ncurves = curves_.size();
  for (const auto &c : curves_) h.npoints += c.get_q().size();
  bool ok = std::fwrite(&h, sizeof h, 1, f) == 1;
  std::vector<float> buf;
  for (const auto &c : curves_) {
    int32_t cnt = (int32_t)c.get_q().size();
    if (c.is_looped()) cnt = -cnt;
    ok = ok && std::fwrite(&cnt, sizeof cnt, 1, f) == 1;
    buf.resize(3 * c.get_q().size());
    for (size_t k = 0; k < c.get_q().size(); ++k)
      for (int d = 0; d < 3; ++d) buf[3 * k + d] = (float)c.get_q()[k][d];
    if (!buf.empty()) ok = ok && std::fwrite(buf.data(), sizeof(float), buf.size(), f) == buf.size();
  }
  std::fclose(f);
  if (!ok) throw std::runtime_error("Unable to export BCC file.");
}

// OBJ 'v' and 'l' records (reference model.cpp:124-264).  A line whose first
// vertex is not the previous line's last vertex starts a new curve (Blender
// writes one 'l' per edge); Houdini's '\' continuations are joined; a curve
// closing back onto its first vertex drops the duplicate.  Every curve is a
// closed loop of the requested segment type.
void Model::ImportFromObjFile(const std::string &filename, SegmentType segment_type) {
  if (lower_ext(filename) != ".obj") {
    std::cerr << "File extension must be OBJ: " << filename << std::endl;
    throw std::runtime_error("Import OBJ model: file extension error.");
  }
  std::ifstream file(filename);
  if (!file.is_open()) {
    std::cerr << "Unable to open file: " << filename << std::endl;
    throw std::runtime_error("Import OBJ model: unable to open file.");
  }
  std::vector<Point3> verts;
  Curve cur;
  cur.set_segment_type(segment_type);
  auto &q = cur.get_q();
  int first_vid = -1, last_vid = -1;
  auto finish = [&]() {
    if (q.size() > 1) {
      if (last_vid == first_vid || q.back().isApprox(q.front())) q.pop_back();
      cur.set_looped(true);
      curves_.push_back(cur);
      return true;
    }
    return false;
  };
  std::string line;
  long lines = 0;
  while (std::getline(file, line)) {
    if (++lines % 100000 == 0)
      std::cout << "Processed the first " << lines << " lines of " << filename << "." << std::endl;
    line = trim(line);
    if (line.empty() || line[0] == '#') continue;
    while (!line.empty() && line.back() == '\\') {
      std::string more;
      if (!std::getline(file, more)) break;
      line.back() = ' ';
      line += ' ' + trim(more);
    }
    if (line.size() < 2 || !(line[1] == ' ' || line[1] == '\t' || line[1] == '\v')) continue;
    std::istringstream ss(line.substr(1));
    if (line[0] == 'v') {
      double x, y, z;
      if (!(ss >> x >> y >> z)) {
        std::cerr << "Obj input: Invalid vertex (vid " << verts.size() + 1 << ")." << std::endl;
        throw std::runtime_error("Obj input: Invalid vertex.");
      }
      verts.emplace_back(x, y, z);
    } else if (line[0] == 'l') {
      std::string tok;
      bool first = true;
      while (ss >> tok) {
        const int vid = std::stoi(tok.substr(0, tok.find('/')));
        if (vid < 1 || vid > (int)verts.size()) {
          std::cerr << "Obj import: Illegal vid: " << vid << std::endl;
          throw std::runtime_error("Obj Import: illegal vertex ID.");
        }
        if (first) {
          first = false;
          if (vid != last_vid) {
            if (finish()) {
              q.clear();
              last_vid = -1;
            }
            q.push_back(verts[vid - 1]);
            last_vid = first_vid = vid;
          }
        } else {
          q.push_back(verts[vid - 1]);
          last_vid = vid;
        }
      }
    }
  }
  finish();
  std::cout << "Successfully imported " << curves_.size() << " curves." << std::endl;
}

}  // namespace VerifyCurves
