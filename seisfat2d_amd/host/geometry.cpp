#include "geometry.hpp"

#include <stdexcept>

#include "admin.hpp"

namespace seisfat {

namespace {

void parse_points(const std::string& path, std::vector<float>& xs, std::vector<float>& zs) {
  std::vector<std::string> lines;
  import_text_file(path, lines);
  for (const std::string& line : lines) {
    const std::vector<std::string> f = split(line, ',');
    if (f.size() < 2) throw std::invalid_argument("Error: malformed line '" + line + "' in " + path);
    try {
      xs.push_back(std::stof(f[0]));
      zs.push_back(std::stof(f[1]));
    } catch (const std::exception&) {
      throw std::invalid_argument("Error: malformed line '" + line + "' in " + path);
    }
  }
}

}  // namespace

void Geometry::set_parameters() {
  std::vector<float> xs, zs, xr, zr;
  parse_points(catch_parameter("SPS", parameters), xs, zs);
  parse_points(catch_parameter("RPS", parameters), xr, zr);
  nsrc = (int)xs.size();
  nrec = (int)xr.size();
  store_.assign(2 * (size_t)nsrc + 2 * (size_t)nrec, 0.0f);
  xsrc = store_.data();
  zsrc = xsrc + nsrc;
  xrec = zsrc + nsrc;
  zrec = xrec + nrec;
  for (int i = 0; i < nsrc; ++i) {
    xsrc[i] = xs[i];
    zsrc[i] = zs[i];
  }
  for (int i = 0; i < nrec; ++i) {
    xrec[i] = xr[i];
    zrec[i] = zr[i];
  }
}

Geometry::~Geometry() = default;

}  // namespace seisfat
