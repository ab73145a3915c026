// Modeling plugins of the eikonal path, over the C ABI of seisfat2d_eik.
//
// Keeps SeisFAT2D's class interface (src/modeling/modeling.cuh:17-79,
// src/modeling/eikonal_iso.cuh, src/modeling/eikonal_ani.cuh): the same public
// members and call sequence
//   m->parameters = file; m->set_parameters();
//   per shot: m->srcId = k; m->set_shot_point(); m->time_propagation();
//             m->export_seismogram();   (or compute_seismogram() / read d_T)
// so the reference's drivers (modeling, tomography, migration table builders)
// can call it unchanged.  Differences: the solve runs on the MI355X through
// seisfat_eik_solve, picks are computed on the device, errors throw
// std::runtime_error, and the object owns and frees its memory.
#pragma once

#include <string>
#include <vector>

#include "../../include/seisfat2d_eik.h"
#include "geometry.hpp"

namespace seisfat {

class Modeling {
 public:
  float dx = 0.0f, dz = 0.0f;
  int nxx = 0, nzz = 0, matsize = 0;
  int nx = 0, nz = 0, nb = 3, nPoints = 0;
  int srcId = 0, recId = 0, sIdx = 0, sIdz = 0;
  int nBlocks = 0;

  float sx = 0.0f, sz = 0.0f;

  // Padded slowness on the host.  Callers may edit interior values and push
  // them with copy_slowness_to_device() (src/inversion/tomography_iso.cpp:88-91).
  float* S = nullptr;
  // Padded traveltimes on the host, valid after compute_seismogram().
  float* T = nullptr;
  // Padded traveltimes of the last solve on the device (read by table builders).
  const float* d_T = nullptr;
  // Padded slowness on the device (src/modeling/modeling.cuh:171), as last
  // pushed by set_parameters() / copy_slowness_to_device(); read only.
  const float* d_S = nullptr;
  // Receiver picks of the last solve (geometry->nrec values).
  float* seismogram = nullptr;

  Geometry* geometry = nullptr;

  std::string parameters;
  std::string data_folder;
  std::string modeling_type;
  std::string modeling_name;
  std::string Cijkl_folder;

  // Sweep schedule of the solves (SEISFAT_SCHED_*).
  int schedule = SEISFAT_SCHED_TILED;

  // Reads x_samples, z_samples, x_spacing, z_spacing, slowness_file,
  // modeling_output_folder, Cijkl_folder, SPS and RPS, pads the model and
  // uploads it (src/modeling/modeling.cu:3-102).
  void set_parameters();
  // sx, sz of shot srcId (src/modeling/modeling.cu:104-108).
  void set_shot_point();
  void show_information();
  // Source node of (sx, sz) (src/modeling/modeling.cu:112-113); the device
  // initialisation runs inside eikonal_solver().
  void initialization();
  // One solve for the current (sx, sz): initialisation, one pass of four
  // sweeps and the receiver picks (src/modeling/modeling.cu:110-159).
  void eikonal_solver();
  // Copies the last solve's traveltimes into T; the picks are already in
  // seismogram (src/modeling/modeling.cu:161-194).
  void compute_seismogram();
  // compute_seismogram() then writes
  // <modeling_output_folder><modeling_type>_nStations<nrec>_shot_<srcId+1>.bin
  // (src/modeling/modeling.cu:196-202).
  void export_seismogram();
  // Pushes the whole padded S, border included, to the device
  // (src/modeling/modeling.cu:281-284).
  void copy_slowness_to_device();

  void expand_boundary(const float* input, float* output) const;
  void reduce_boundary(const float* input, float* output) const;

  virtual void set_conditions() = 0;
  virtual void time_propagation() = 0;

  Modeling() = default;
  Modeling(const Modeling&) = delete;
  Modeling& operator=(const Modeling&) = delete;
  virtual ~Modeling();

  seisfat_eik_t handle() const { return h_; }

 protected:
  int solve_type_ = SEISFAT_ISO;
  seisfat_eik_t h_ = nullptr;
  std::vector<float> S_, T_, seis_;
};

// Isotropic plugin (src/modeling/eikonal_iso.cu).
class Eikonal_ISO : public Modeling {
 public:
  void set_conditions() override;
  void time_propagation() override;
};

// Anisotropic (2-D TTI qP) plugin (src/modeling/eikonal_ani.cu): reads
// <Cijkl_folder>C11.bin, C13, C15, C33, C35, C55.
class Eikonal_ANI : public Modeling {
 public:
  void set_conditions() override;
  void time_propagation() override;
};

// Throws std::runtime_error with seisfat_last_error() when rc != 0.
void check_status(int rc, const char* what);

}  // namespace seisfat
