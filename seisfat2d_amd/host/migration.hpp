// Kirchhoff depth migration driver: SeisFAT2D's Migration class and its four
// plugins IDKDM, ADKDM, IDLSKDM, ADLSKDM (src/migration/migration.cuh,
// KDM.cu, LSKDM.cu, IDKDM.cu, ADKDM.cu, IDLSKDM.cu, ADLSKDM.cu) over the C
// ABI of seisfat2d_kdm.h.
//
// Same parameter keys, call sequence and outputs:
//   m.parameters = file; m.set_parameters(); m.kirchhoff_depth_migration();
//   m.export_outputs();
// set_parameters reads the migration keys of src/migration/migration.cu:3-35
// (time_samples, time_spacing, max_frequency, mig_aperture, mig_max_angle,
// mig_angle_spacing, mig_max_offset, mig_max_iteration, mig_smooth_lambda,
// mig_data_folder, mig_data_prefix, mig_tables_folder, mig_seismic_folder,
// mig_residuo_folder, anisotropy) and the modeling keys.  The migration
// solves every source and receiver table with the batched eikonal solver and
// hands them to the operators in device memory (also writing
// eikonal_{src,rec}_<k>.bin as the reference does, unless write_tables is
// false), reads mig_data_prefix<k>.bin (nt x nrec samples per source), and
// writes the reference's result files.  Differences: the trace loop is one
// pair of launches over all traces instead of a launch per trace; the wavelet
// convolutions are direct double sums instead of FFTW products; the image
// adjoint adds each point's contributions in a fixed order without atomics.
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "../../include/seisfat2d_kdm.h"
#include "modeling.hpp"

namespace seisfat {

class Migration {
 public:
  // 0 IDKDM, 1 ADKDM, 2 IDLSKDM, 3 ADLSKDM (the reference's mig_type)
  int type = 0;
  std::string parameters;

  int nt = 0, nw = 0, max_it = 0, nang = 0, nCMP = 0, m_samples = 0;
  float dt = 0.0f, da = 0.0f, fmax = 0.0f, max_angle = 0.0f, aperture = 0.0f, max_offset = 0.0f;
  float smooth = 0.0f, ds = 0.0f, dr = 0.0f, dCMP = 0.0f, minCMP = 0.0f, maxCMP = 0.0f;
  bool anisotropy = false;
  bool write_tables = true;

  std::string input_data_folder, input_data_prefix, tables_folder, seismic_folder, residuo_folder;
  std::string domain, migType, output_path;

  std::vector<float> wavelet;   // Ricker (src/migration/migration.cu:54-73)
  std::vector<float> h_model;   // the result (m_samples)
  std::vector<float> residuo;   // LS residual norms

  std::unique_ptr<Modeling> modeling;

  // The traces the migration visits, in order (src/migration/KDM.cu:27-53).
  std::vector<int> tsrc, trec, tcid;
  std::vector<float> tcmp;

  double tables_s = 0.0;  // wall time of the table solves (and files)

  void set_parameters();
  void kirchhoff_depth_migration();
  void export_outputs();

  Migration() = default;
  Migration(const Migration&) = delete;
  Migration& operator=(const Migration&) = delete;
  ~Migration();

 private:
  seisfat_kdm_t kdm_ = nullptr;
  float* d_data_ = nullptr;   // ntr x nt traces
  float* d_conv_ = nullptr;   // ntr x nt traces after a convolution
  float* d_model_ = nullptr;  // m_samples
  int ntr_ = 0;

  void set_wavelet();
  void set_gathers();
  void set_migration();
  void set_traces();
  void set_travel_times();
  void read_seismic(std::vector<float>& traces) const;
  void forward(const std::vector<float>& model, std::vector<float>& traces);  // L then forward convolution
  void adjoint(const std::vector<float>& traces, std::vector<float>& model);  // adjoint convolution then L^T
  void least_squares(const std::vector<float>& dobs);
};

}  // namespace seisfat
