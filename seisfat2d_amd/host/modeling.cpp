#include "modeling.hpp"

#include <cmath>
#include <iostream>
#include <stdexcept>

#include "admin.hpp"

namespace seisfat {

void check_status(int rc, const char* what) {
  if (rc != SEISFAT_OK)
    throw std::runtime_error(std::string(what) + " failed (" + std::to_string(rc) + "): " + seisfat_last_error());
}

Modeling::~Modeling() {
  if (h_) seisfat_eik_destroy(h_);
  delete geometry;
}

void Modeling::set_parameters() {
  nx = std::stoi(catch_parameter("x_samples", parameters));
  nz = std::stoi(catch_parameter("z_samples", parameters));
  dx = std::stof(catch_parameter("x_spacing", parameters));
  dz = std::stof(catch_parameter("z_spacing", parameters));
  data_folder = catch_parameter_or("modeling_output_folder", parameters, "");
  Cijkl_folder = catch_parameter_or("Cijkl_folder", parameters, "");
  if (nx < 2 || nz < 2) throw std::invalid_argument("Error: x_samples and z_samples must be at least 2");

  nPoints = nx * nz;
  nb = 3;
  nxx = nx + 2 * nb;
  nzz = nz + 2 * nb;
  matsize = nxx * nzz;
  nBlocks = (matsize + 255) / 256;

  delete geometry;
  geometry = new Geometry();
  geometry->parameters = parameters;
  geometry->set_parameters();
  seis_.assign(geometry->nrec, 0.0f);
  seismogram = seis_.data();

  std::vector<float> s((size_t)nPoints);
  import_binary_float(catch_parameter("slowness_file", parameters), s.data(), s.size());
  S_.assign((size_t)matsize, 0.0f);
  T_.assign((size_t)matsize, 0.0f);
  S = S_.data();
  T = T_.data();
  expand_boundary(s.data(), S);

  if (h_) seisfat_eik_destroy(h_);
  h_ = nullptr;
  check_status(seisfat_eik_create(nz, nx, dz, dx, &h_), "seisfat_eik_create");
  copy_slowness_to_device();
  set_conditions();
}

void Modeling::set_shot_point() {
  if (srcId < 0 || srcId >= geometry->nsrc) throw std::out_of_range("srcId outside the source list");
  sx = geometry->xsrc[srcId];
  sz = geometry->zsrc[srcId];
}

void Modeling::show_information() {
  std::cout << "-------------------------------------------------------------------\n";
  std::cout << " SeisFAT2D eikonal (MI355X)\n";
  std::cout << "-------------------------------------------------------------------\n\n";
  std::cout << "Model dimensions: (z = " << (nz - 1) * dz << ", x = " << (nx - 1) * dx << ") m\n\n";
  std::cout << "Running shot " << srcId + 1 << " of " << geometry->nsrc << " in total\n\n";
  std::cout << "Current shot position: (z = " << sz << ", x = " << sx << ") m\n\n";
  std::cout << modeling_name << "\n";
}

void Modeling::initialization() {
  sIdx = (int)((sx + 0.5f * dx) / dx) + nb;
  sIdz = (int)((sz + 0.5f * dz) / dz) + nb;
}

void Modeling::eikonal_solver() {
  const int nrec = geometry ? geometry->nrec : 0;
  check_status(seisfat_eik_solve(h_, solve_type_, schedule, sx, sz, nullptr, nrec ? geometry->xrec : nullptr,
                                 nrec ? geometry->zrec : nullptr, nrec, nrec ? seismogram : nullptr),
               "seisfat_eik_solve");
  d_T = seisfat_eik_device_T(h_);
}

void Modeling::compute_seismogram() { check_status(seisfat_eik_copy_T(h_, T), "seisfat_eik_copy_T"); }

void Modeling::export_seismogram() {
  compute_seismogram();
  const std::string file = data_folder + modeling_type + "_nStations" + std::to_string(geometry->nrec) + "_shot_" +
                           std::to_string(srcId + 1) + ".bin";
  export_binary_float(file, seismogram, (size_t)geometry->nrec);
}

void Modeling::copy_slowness_to_device() {
  check_status(seisfat_eik_set_slowness_padded(h_, S), "seisfat_eik_set_slowness_padded");
  d_S = seisfat_eik_device_S(h_);
}

void Modeling::expand_boundary(const float* input, float* output) const {
  for (int j = 0; j < nx; ++j)
    for (int i = 0; i < nz; ++i) output[(i + nb) + (size_t)(j + nb) * nzz] = input[i + (size_t)j * nz];
  for (int i = 0; i < nb; ++i)
    for (int j = nb; j < nxx - nb; ++j) {
      output[i + (size_t)j * nzz] = output[nb + (size_t)j * nzz];
      output[(nzz - i - 1) + (size_t)j * nzz] = output[(nzz - nb - 1) + (size_t)j * nzz];
    }
  for (int i = 0; i < nzz; ++i)
    for (int j = 0; j < nb; ++j) {
      output[i + (size_t)j * nzz] = output[i + (size_t)nb * nzz];
      output[i + (size_t)(nxx - j - 1) * nzz] = output[i + (size_t)(nxx - nb - 1) * nzz];
    }
}

void Modeling::reduce_boundary(const float* input, float* output) const {
  for (int j = 0; j < nx; ++j)
    for (int i = 0; i < nz; ++i) output[i + (size_t)j * nz] = input[(i + nb) + (size_t)(j + nb) * nzz];
}

void Eikonal_ISO::set_conditions() {
  modeling_type = "eikonal_iso";
  modeling_name = "Modeling type: Eikonal isotropic solver";
  solve_type_ = SEISFAT_ISO;
}

void Eikonal_ISO::time_propagation() {
  initialization();
  eikonal_solver();
}

void Eikonal_ANI::set_conditions() {
  modeling_type = "eikonal_ani";
  modeling_name = "Modeling type: Anisotropic eikonal solver";
  solve_type_ = SEISFAT_ANI;
  static const char* names[6] = {"C11", "C13", "C15", "C33", "C35", "C55"};
  std::vector<std::vector<float>> c(6, std::vector<float>((size_t)nPoints));
  const float* ptrs[6];
  for (int e = 0; e < 6; ++e) {
    import_binary_float(Cijkl_folder + names[e] + ".bin", c[e].data(), c[e].size());
    ptrs[e] = c[e].data();
  }
  check_status(seisfat_eik_set_stiffness(h_, ptrs), "seisfat_eik_set_stiffness");
}

void Eikonal_ANI::time_propagation() {
  initialization();
  eikonal_solver();
}

}  // namespace seisfat
