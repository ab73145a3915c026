// Per-shot eikonal modeling driver (SeisFAT2D src/modeling_main.cpp).
//
//   seisfat2d_modeling <parameters.txt> [--batch] [--quiet] [--check]
//                      [--schedule plain|tiled|tiled_diag]
//
// Reads modeling_type (0 = Eikonal_ISO, 1 = Eikonal_ANI) and solves every
// source of the SPS file, writing one
// <modeling_output_folder><type>_nStations<nrec>_shot_<k>.bin per shot, as the
// reference does.  --batch solves all shots in one batched GPU call instead of
// the per-shot loop (same output files); --check validates the parameter,
// geometry and model files without touching the GPU.
#include <chrono>
#include <cmath>
#include <cstring>
#include <iostream>
#include <memory>
#include <vector>

#include "admin.hpp"
#include "modeling.hpp"

using namespace seisfat;

namespace {

int check_only(const std::string& file, int type) {
  const int nx = std::stoi(catch_parameter("x_samples", file));
  const int nz = std::stoi(catch_parameter("z_samples", file));
  const float dx = std::stof(catch_parameter("x_spacing", file));
  const float dz = std::stof(catch_parameter("z_spacing", file));
  if (nx < 2 || nz < 2 || !(dx > 0.0f) || !(dz > 0.0f)) throw std::invalid_argument("Error: bad model dimensions");
  std::vector<float> s((size_t)nx * nz);
  import_binary_float(catch_parameter("slowness_file", file), s.data(), s.size());
  for (float v : s)
    if (!(v > 0.0f) || !std::isfinite(v)) throw std::invalid_argument("Error: slowness must be positive and finite");
  if (type == 1) {
    const std::string folder = catch_parameter("Cijkl_folder", file);
    for (const char* c : {"C11", "C13", "C15", "C33", "C35", "C55"})
      import_binary_float(folder + c + ".bin", s.data(), s.size());
  }
  Geometry g;
  g.parameters = file;
  g.set_parameters();
  const float xmax = (nx - 1) * dx, zmax = (nz - 1) * dz;
  for (int i = 0; i < g.nsrc; ++i)
    if (!(g.xsrc[i] >= 0 && g.xsrc[i] <= xmax && g.zsrc[i] >= 0 && g.zsrc[i] <= zmax))
      throw std::invalid_argument("Error: source " + std::to_string(i + 1) + " outside the model");
  for (int i = 0; i < g.nrec; ++i)
    if (!(g.xrec[i] >= 0 && g.xrec[i] <= xmax && g.zrec[i] >= 0 && g.zrec[i] <= zmax))
      throw std::invalid_argument("Error: receiver " + std::to_string(i + 1) + " outside the model");
  std::cout << "ok: model " << nz << " x " << nx << ", " << g.nsrc << " sources, " << g.nrec << " receivers\n";
  return 0;
}

void run_batch(Modeling& m) {
  const int nsrc = m.geometry->nsrc, nrec = m.geometry->nrec;
  std::vector<seisfat_shot> shots(nsrc);
  std::vector<float> rx((size_t)nsrc * nrec), rz((size_t)nsrc * nrec);
  for (int k = 0; k < nsrc; ++k) {
    shots[k] = seisfat_shot{m.geometry->xsrc[k], m.geometry->zsrc[k], 0, m.nx, k * nrec, nrec};
    for (int r = 0; r < nrec; ++r) {
      rx[(size_t)k * nrec + r] = m.geometry->xrec[r];
      rz[(size_t)k * nrec + r] = m.geometry->zrec[r];
    }
  }
  const int type = m.modeling_type == "eikonal_ani" ? SEISFAT_ANI : SEISFAT_ISO;
  seisfat_batch_t b = nullptr;
  check_status(seisfat_batch_create(m.handle(), type, m.schedule, nsrc, shots.data(), nsrc * nrec, rx.data(),
                                    rz.data(), &b),
               "seisfat_batch_create");
  std::vector<float> picks((size_t)nsrc * nrec);
  int rc = seisfat_batch_run(b, nullptr);
  if (rc == SEISFAT_OK) rc = seisfat_batch_picks(b, picks.data());
  seisfat_batch_destroy(b);
  check_status(rc, "batched solve");
  for (int k = 0; k < nsrc; ++k) {
    const std::string file = m.data_folder + m.modeling_type + "_nStations" + std::to_string(nrec) + "_shot_" +
                             std::to_string(k + 1) + ".bin";
    export_binary_float(file, picks.data() + (size_t)k * nrec, (size_t)nrec);
  }
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 2) {
    std::cerr << "usage: " << argv[0]
              << " <parameters.txt> [--batch] [--quiet] [--check] [--schedule plain|tiled|tiled_diag]\n";
    return 2;
  }
  bool batch = false, quiet = false, check = false;
  int schedule = SEISFAT_SCHED_TILED;
  for (int a = 2; a < argc; ++a) {
    if (!std::strcmp(argv[a], "--batch")) batch = true;
    else if (!std::strcmp(argv[a], "--quiet")) quiet = true;
    else if (!std::strcmp(argv[a], "--check")) check = true;
    else if (!std::strcmp(argv[a], "--schedule") && a + 1 < argc) {
      const std::string v = argv[++a];
      schedule = v == "plain" ? SEISFAT_SCHED_PLAIN : v == "tiled_diag" ? SEISFAT_SCHED_TILED_DIAG : SEISFAT_SCHED_TILED;
    } else {
      std::cerr << "unknown option " << argv[a] << "\n";
      return 2;
    }
  }
  try {
    const std::string file = argv[1];
    const int type = std::stoi(catch_parameter("modeling_type", file));
    if (type != 0 && type != 1) throw std::invalid_argument("Error: modeling_type must be 0 (ISO) or 1 (ANI)");
    if (check) return check_only(file, type);

    std::unique_ptr<Modeling> modeling;
    if (type == 0)
      modeling.reset(new Eikonal_ISO());
    else
      modeling.reset(new Eikonal_ANI());
    modeling->parameters = file;
    modeling->schedule = schedule;
    modeling->set_parameters();

    const auto ti = std::chrono::steady_clock::now();
    if (batch) {
      run_batch(*modeling);
    } else {
      for (int shot = 0; shot < modeling->geometry->nsrc; ++shot) {
        modeling->srcId = shot;
        modeling->set_shot_point();
        if (!quiet) modeling->show_information();
        modeling->time_propagation();
        modeling->export_seismogram();
      }
    }
    const auto tf = std::chrono::steady_clock::now();
    std::cout << "\nRun time: " << std::chrono::duration<double>(tf - ti).count() << " s." << std::endl;
  } catch (const std::exception& ex) {
    std::cerr << ex.what() << std::endl;
    return 1;
  }
  return 0;
}
