// Source and receiver positions from the SPS/RPS text files named in the
// parameter file (SeisFAT2D src/geometry/geometry.hpp:6-22).  Each non-comment
// line holds "x, z" in metres.
#pragma once

#include <string>
#include <vector>

namespace seisfat {

class Geometry {
 public:
  int nsrc = 0;
  int nrec = 0;

  float* xsrc = nullptr;
  float* zsrc = nullptr;
  float* xrec = nullptr;
  float* zrec = nullptr;

  std::string parameters;

  // Reads the files named by the SPS and RPS keys
  // (src/geometry/geometry.cpp:3-42).  Throws std::invalid_argument on a
  // missing file or a malformed line.
  void set_parameters();

  Geometry() = default;
  Geometry(const Geometry&) = delete;
  Geometry& operator=(const Geometry&) = delete;
  ~Geometry();

 private:
  std::vector<float> store_;
};

}  // namespace seisfat
