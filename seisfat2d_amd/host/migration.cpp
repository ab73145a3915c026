#include "migration.hpp"

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <fstream>
#include <iostream>
#include <stdexcept>

#include "admin.hpp"

namespace seisfat {

namespace {

constexpr float EPS = 1e-6f;  // src/migration/migration.cuh:7

void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

template <class T>
void dfree(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

}  // namespace

Migration::~Migration() {
  if (kdm_) seisfat_kdm_destroy(kdm_);
  dfree(d_data_);
  dfree(d_conv_);
  dfree(d_model_);
}

void Migration::set_parameters() {
  const std::string& f = parameters;
  if (type < 0 || type > 3) throw std::invalid_argument("Error: mig_type must be 0, 1, 2 or 3");
  nt = std::stoi(catch_parameter("time_samples", f));
  dt = std::stof(catch_parameter("time_spacing", f));
  da = std::stof(catch_parameter("mig_angle_spacing", f));
  max_angle = std::stof(catch_parameter("mig_max_angle", f));
  fmax = std::stof(catch_parameter("max_frequency", f));
  aperture = std::stof(catch_parameter("mig_aperture", f));
  max_offset = std::stof(catch_parameter("mig_max_offset", f));
  input_data_folder = catch_parameter("mig_data_folder", f);
  input_data_prefix = catch_parameter("mig_data_prefix", f);
  tables_folder = catch_parameter("mig_tables_folder", f);
  seismic_folder = catch_parameter("mig_seismic_folder", f);
  anisotropy = str2bool(catch_parameter_or("anisotropy", f, "false"));
  // the least-squares keys (LSKDM.cu:33, migration.cu:8-9) are needed by types 2 and 3 only
  const bool ls = type >= 2;
  max_it = std::stoi(ls ? catch_parameter("mig_max_iteration", f) : catch_parameter_or("mig_max_iteration", f, "0"));
  residuo_folder = ls ? catch_parameter("mig_residuo_folder", f) : catch_parameter_or("mig_residuo_folder", f, "");
  smooth = std::stof(ls ? catch_parameter("mig_smooth_lambda", f) : catch_parameter_or("mig_smooth_lambda", f, "0"));
  if (nt < 1 || nt > 16384) throw std::invalid_argument("Error: time_samples must lie in [1, 16384]");
  if (!(dt > 0.0f) || !(da > 0.0f)) throw std::invalid_argument("Error: time_spacing and mig_angle_spacing must be positive");
  if (ls && max_it < 0) throw std::invalid_argument("Error: mig_max_iteration must be >= 0");

  modeling.reset(anisotropy ? (Modeling*)new Eikonal_ANI() : (Modeling*)new Eikonal_ISO());
  modeling->parameters = f;
  modeling->set_parameters();

  set_wavelet();
  set_gathers();
  set_migration();
  set_traces();

  const seisfat_kdm_params p = {(type == 1 || type == 3) ? SEISFAT_KDM_ANGLE : SEISFAT_KDM_IMAGE,
                                nt, dt, aperture, da, nang, nCMP};
  check_status(seisfat_kdm_create(modeling->handle(), &p, modeling->geometry->nsrc, modeling->geometry->nrec, &kdm_),
               "seisfat_kdm_create");
  check_status(seisfat_kdm_set_traces(kdm_, ntr_, tsrc.data(), trec.data(), tcmp.data(), tcid.data()),
               "seisfat_kdm_set_traces");
  const size_t nd = (size_t)std::max(ntr_, 1) * nt;
  hip_check(hipMalloc((void**)&d_data_, nd * sizeof(float)), "hipMalloc traces");
  hip_check(hipMalloc((void**)&d_conv_, nd * sizeof(float)), "hipMalloc traces");
  hip_check(hipMalloc((void**)&d_model_, (size_t)m_samples * sizeof(float)), "hipMalloc model");
  h_model.assign(m_samples, 0.0f);
}

void Migration::set_wavelet() {
  nw = seisfat_ricker(dt, fmax, nullptr, 0);
  if (nw < 0) check_status(nw, "seisfat_ricker");
  if (nw == 0) throw std::invalid_argument("Error: max_frequency and time_spacing give an empty wavelet");
  wavelet.assign(nw, 0.0f);
  const int rc = seisfat_ricker(dt, fmax, wavelet.data(), nw);
  if (rc < 0) check_status(rc, "seisfat_ricker");
}

void Migration::set_gathers() {
  const Geometry& g = *modeling->geometry;
  nang = (int)(max_angle / da) + 1;
  // dCMP: half the smaller of the source and receiver spacings; a geometry
  // with one source (receiver) takes the other spacing (the reference would
  // read xsrc[1] past the end)
  ds = g.nsrc > 1 ? g.xsrc[1] - g.xsrc[0] : 0.0f;
  dr = g.nrec > 1 ? g.xrec[1] - g.xrec[0] : 0.0f;
  if (g.nsrc > 1 && g.nrec > 1) dCMP = 0.5f * std::min(ds, dr);
  else if (g.nsrc > 1) dCMP = 0.5f * ds;
  else if (g.nrec > 1) dCMP = 0.5f * dr;
  else throw std::invalid_argument("Error: one source and one receiver give no CMP spacing");
  if (!(dCMP > 0.0f)) throw std::invalid_argument("Error: sources and receivers must be listed with increasing x");
  minCMP = 1e6f;
  maxCMP = -1e6f;  // the reference leaves it unset
  for (int s = 0; s < g.nsrc; ++s)
    for (int r = 0; r < g.nrec; ++r) {
      const float cmp = 0.5f * (g.xsrc[s] + g.xrec[r]);
      minCMP = cmp < minCMP ? cmp : minCMP;
      maxCMP = cmp > maxCMP ? cmp : maxCMP;
    }
  nCMP = (int)((maxCMP - minCMP) / dCMP) + 1;
}

void Migration::set_migration() {
  const int nz = modeling->nz, nx = modeling->nx;
  const bool angle = type == 1 || type == 3;
  const bool ls = type >= 2;
  static const char* names[4] = {"IDKDM", "ADKDM", "IDLSKDM", "ADLSKDM"};
  migType = names[type];
  domain = angle ? "Angle Domain" : "Image Domain";
  m_samples = angle ? nz * nang * nCMP : nz * nx;
  output_path = seismic_folder + migType + "_result_" + std::to_string(nz) + "x" +
                (angle ? std::to_string(nCMP) + "x" + std::to_string(nang) : std::to_string(nx)) +
                (ls ? "_iteration_" + std::to_string(max_it) : std::string()) + ".bin";
}

void Migration::set_traces() {
  const Geometry& g = *modeling->geometry;
  tsrc.clear();
  trec.clear();
  tcmp.clear();
  tcid.clear();
  for (int s = 0; s < g.nsrc; ++s) {
    const float sx = g.xsrc[s];
    for (int r = 0; r < g.nrec; ++r) {
      const float rx = g.xrec[r];
      if (std::fabs(sx - rx) < max_offset) {
        const float cmp = 0.5f * (sx + rx);
        tsrc.push_back(s);
        trec.push_back(r);
        tcmp.push_back(cmp);
        tcid.push_back((int)((cmp - minCMP) / dCMP));
      }
    }
  }
  ntr_ = (int)tsrc.size();
}

// Migration::set_src_travel_times / set_rec_travel_times
// (src/migration/migration.cu:116-171): whole-model solves from every source
// and receiver position, 64 per batch, each table handed to the operators from
// device memory and (write_tables) written to <mig_tables_folder>.
void Migration::set_travel_times() {
  const auto t0 = std::chrono::steady_clock::now();
  const Geometry& g = *modeling->geometry;
  const int type_e = modeling->modeling_type == "eikonal_ani" ? SEISFAT_ANI : SEISFAT_ISO;
  std::vector<float> table(write_tables ? (size_t)modeling->matsize : 0);
  for (int which = 0; which < 2; ++which) {
    const float* x = which ? g.xrec : g.xsrc;
    const float* z = which ? g.zrec : g.zsrc;
    const int n = which ? g.nrec : g.nsrc;
    const std::string prefix = which ? "eikonal_rec_" : "eikonal_src_";
    for (int k0 = 0; k0 < n; k0 += 64) {
      const int m = std::min(64, n - k0);
      std::vector<seisfat_shot> shots(m);
      for (int q = 0; q < m; ++q) shots[q] = seisfat_shot{x[k0 + q], z[k0 + q], 0, modeling->nx, 0, 0};
      seisfat_batch_t b = nullptr;
      check_status(seisfat_batch_create(modeling->handle(), type_e, modeling->schedule, m, shots.data(), 0, nullptr,
                                        nullptr, &b),
                   "seisfat_batch_create");
      int rc = seisfat_batch_run(b, nullptr);
      if (rc == SEISFAT_OK) rc = seisfat_batch_status(b);
      for (int q = 0; q < m && rc == SEISFAT_OK; ++q) {
        int nzz = 0, nxx = 0;
        const float* dT = seisfat_batch_device_T(b, q, &nzz, &nxx);
        rc = dT ? seisfat_kdm_set_table(kdm_, which, k0 + q, dT, nullptr) : SEISFAT_EINVAL;
        if (rc == SEISFAT_OK && write_tables) {
          rc = seisfat_batch_copy_T(b, q, table.data());
          if (rc == SEISFAT_OK)
            export_binary_float(tables_folder + prefix + std::to_string(k0 + q + 1) + ".bin", table.data(),
                                table.size());
        }
      }
      const hipError_t e = hipDeviceSynchronize();  // the table kernels read the batch's grids
      seisfat_batch_destroy(b);
      check_status(rc, "traveltime tables");
      hip_check(e, "traveltime tables");
    }
  }
  tables_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// Every migrated trace in order: sample tId of trace (s, r) is sample tId of
// receiver r in mig_data_prefix<s+1>.bin (src/migration/KDM.cu:16-17,47-48).
void Migration::read_seismic(std::vector<float>& traces) const {
  const Geometry& g = *modeling->geometry;
  traces.assign((size_t)ntr_ * nt, 0.0f);
  std::vector<float> seismic((size_t)nt * g.nrec);
  int loaded = -1;
  for (int t = 0; t < ntr_; ++t) {
    if (tsrc[t] != loaded) {
      loaded = tsrc[t];
      import_binary_float(input_data_folder + input_data_prefix + std::to_string(loaded + 1) + ".bin",
                          seismic.data(), seismic.size());
    }
    std::copy_n(seismic.data() + (size_t)trec[t] * nt, nt, traces.data() + (size_t)t * nt);
  }
}

// traces = forward_convolution(L model) (perform_forward per trace, then the
// host convolution; src/migration/LSKDM.cu:87-91,175-179)
void Migration::forward(const std::vector<float>& model, std::vector<float>& traces) {
  const size_t nd = (size_t)ntr_ * nt;
  hip_check(hipMemcpy(d_model_, model.data(), model.size() * sizeof(float), hipMemcpyHostToDevice), "model upload");
  hip_check(hipMemset(d_data_, 0, nd * sizeof(float)), "trace clear");
  check_status(seisfat_kdm_forward(kdm_, d_model_, d_data_, nullptr), "seisfat_kdm_forward");
  check_status(seisfat_trace_convolve(0, ntr_, nt, wavelet.data(), nw, d_data_, d_conv_, nullptr),
               "seisfat_trace_convolve");
  traces.resize(nd);
  hip_check(hipMemcpy(traces.data(), d_conv_, nd * sizeof(float), hipMemcpyDeviceToHost), "trace download");
}

// model = L^T adjoint_convolution(traces) (src/migration/KDM.cu:50-52)
void Migration::adjoint(const std::vector<float>& traces, std::vector<float>& model) {
  const size_t nd = (size_t)ntr_ * nt;
  hip_check(hipMemcpy(d_data_, traces.data(), nd * sizeof(float), hipMemcpyHostToDevice), "trace upload");
  check_status(seisfat_trace_convolve(1, ntr_, nt, wavelet.data(), nw, d_data_, d_conv_, nullptr),
               "seisfat_trace_convolve");
  hip_check(hipMemset(d_model_, 0, (size_t)m_samples * sizeof(float)), "model clear");
  check_status(seisfat_kdm_adjoint(kdm_, d_conv_, d_model_, nullptr), "seisfat_kdm_adjoint");
  model.resize(m_samples);
  hip_check(hipMemcpy(model.data(), d_model_, (size_t)m_samples * sizeof(float), hipMemcpyDeviceToHost),
            "model download");
}

void Migration::kirchhoff_depth_migration() {
  set_travel_times();
  std::vector<float> dobs;
  read_seismic(dobs);
  if (type < 2) {
    adjoint(dobs, h_model);  // KDM::kirchhoff_depth_migration (src/migration/KDM.cu:3-55)
  } else {
    least_squares(dobs);
  }
}

// LSKDM::kirchhoff_depth_migration (src/migration/LSKDM.cu:3-226): the
// reference's host loops in float32 in its order (residual sums in trace
// order, model sums in index order), ADLSKDM::regularization for the angle
// domain (src/migration/ADLSKDM.cu:13-41).
void Migration::least_squares(const std::vector<float>& dobs) {
  const size_t m = (size_t)m_samples, nd = dobs.size();
  std::vector<float> gradient(m), direction(m, 0.0f), gradient_old(m, 0.0f), dcal;
  h_model.assign(m, 0.0f);
  residuo.clear();
  int iteration = 0;
  while (true) {
    // compute_gradient
    forward(h_model, dcal);
    float residuals = 0.0f;
    for (size_t i = 0; i < nd; ++i) {
      dcal[i] = (std::fabs(dobs[i]) > EPS) ? dobs[i] - dcal[i] : 0.0f;
      residuals += dcal[i] * dcal[i];
    }
    adjoint(dcal, gradient);
    // compute_residuals
    ++iteration;
    residuo.push_back(std::sqrt(residuals));
    std::cout << "Iteration " << iteration - 1 << " residuo: " << residuo.back() << "\n";
    if (iteration > max_it) break;
    // compute_direction
    float beta_num = 0.0f, beta_den = 0.0f, alpha_num = 0.0f;
    for (size_t i = 0; i < m; ++i) {
      beta_num += gradient[i] * gradient[i];
      beta_den += gradient_old[i] * gradient_old[i];
    }
    const float beta = beta_num / (beta_den + EPS);
    for (size_t i = 0; i < m; ++i) {
      direction[i] = gradient[i] + beta * direction[i];
      alpha_num += gradient[i] * direction[i];
      gradient_old[i] = gradient[i];
    }
    // compute_stepLength
    forward(direction, dcal);
    float alpha_den = 0.0f;
    for (size_t i = 0; i < nd; ++i) alpha_den += dcal[i] * dcal[i];
    const float alpha = alpha_num / (alpha_den + EPS);
    // regularization
    if (type == 3) {
      if (nang < 2) throw std::invalid_argument("Error: angle-domain smoothing needs two angle bins");
      const int nz = modeling->nz;
      for (int c = 0; c < nCMP; ++c)
        for (int z = 0; z < nz; ++z)
          for (int a = 0; a < nang; ++a) {
            const size_t t = z + (size_t)a * nz + (size_t)c * nang * nz;
            const size_t fw = t + nz, bw = t - nz;
            const float mf = a == nang - 1 ? h_model[bw] : h_model[fw];
            const float mb = a == 0 ? h_model[fw] : h_model[bw];
            direction[t] += smooth * (mf - 2.0f * h_model[t] + mb) / (da * da);
          }
    }
    // update_model
    for (size_t i = 0; i < m; ++i) h_model[i] = h_model[i] + alpha * direction[i];
  }
}

void Migration::export_outputs() {
  if (type >= 2) {
    const std::string path = residuo_folder + migType + "_convergence_" + std::to_string(max_it) + "_iterations.txt";
    std::ofstream res(path);
    if (!res) throw std::runtime_error("Error: cannot write " + path);
    for (float r : residuo) res << r << "\n";
    std::cout << "Text file " << path << " was successfully written." << std::endl;
  }
  export_binary_float(output_path, h_model.data(), h_model.size());
}

}  // namespace seisfat
