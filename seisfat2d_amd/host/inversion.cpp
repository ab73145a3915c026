#include "inversion.hpp"

#include <algorithm>
#include <cmath>
#include <fstream>
#include <iostream>
#include <stdexcept>

#include "admin.hpp"

namespace seisfat {

// ---------------------------------------------------------------------------
// Operators of the reference's optimisation

namespace {

// Coefficients of the order-k difference: [-1, 1] convolved with itself k
// times, by the recurrence of src/inversion/inversion.cpp:231-247.
std::vector<float> difference_coefficients(int order) {
  std::vector<float> df(std::max(order, 0) + 1, 0.0f);
  if (order <= 0) {
    df[0] = 1.0f;
    return df;
  }
  std::vector<float> d1(order + 2, 0.0f), d2(order + 2, 0.0f);
  df[0] = -1.0f;
  df[1] = 1.0f;
  for (int it = 1; it < order; ++it)
    for (int q = 0; q <= order; ++q) {
      d2[q] = df[q];
      d1[q + 1] = df[q];
      df[q] = d1[q] - d2[q];
    }
  return df;
}

// The first key of `keys` the parameter file holds (the reference's code and
// its sample parameter file name several keys differently).
std::string first_of(const std::vector<std::string>& keys, const std::string& file) {
  for (const auto& k : keys) {
    const std::string v = catch_parameter_or(k, file, "\x01");
    if (v != "\x01") return v;
  }
  std::string msg = "Error: none of the keys";
  for (const auto& k : keys) msg += " " + k;
  throw std::invalid_argument(msg + " found in " + file);
}

}  // namespace

void tikhonov_apply(int order, const double* x, long long n, double* y) {
  const std::vector<float> df = difference_coefficients(order);
  const long long rows = n - std::max(order, 0);
  for (long long r = 0; r < rows; ++r) {
    double acc = 0.0;
    for (size_t q = 0; q < df.size(); ++q) acc += (double)df[q] * x[r + (long long)q];
    y[r] = acc;
  }
}

void tikhonov_adjoint(int order, const double* y, long long n, double* x) {
  const std::vector<float> df = difference_coefficients(order);
  const long long rows = n - std::max(order, 0);
  std::fill(x, x + n, 0.0);
  for (long long r = 0; r < rows; ++r)
    for (size_t q = 0; q < df.size(); ++q) x[r + (long long)q] += (double)df[q] * y[r];
}

void smooth_model(const float* in, float* out, int nz, int nx, float stdv, int samples) {
  // the model inside a zero border of `samples` cells (inversion.cpp:351-366)
  const int anx = nx + 2 * samples, anz = nz + 2 * samples;
  std::vector<float> aux((size_t)anx * anz, 0.0f);
  for (int j = 0; j < nx; ++j)
    for (int i = 0; i < nz; ++i) aux[(i + samples) + (size_t)(j + samples) * anz] = in[i + (size_t)j * nz];
  // normalised Gaussian kernel, a unit delta for stdv = 0 (inversion.cpp:396-420)
  const int mid = samples / 2;
  std::vector<float> kernel((size_t)samples * samples, 0.0f);
  kernel[mid + (size_t)mid * samples] = 1.0f;
  if (stdv != 0.0f) {
    float sum = 0.0f;
    for (int x = -mid; x <= mid; ++x)
      for (int z = -mid; z <= mid; ++z) {
        const size_t k = (size_t)(z + mid) + (size_t)(x + mid) * samples;
        const float r = std::sqrt((float)(x * x + z * z));
        kernel[k] = (float)(1.0 / (M_PI * (double)stdv) * (double)std::exp(-((r * r) / (2.0f * stdv * stdv))));
        sum += kernel[k];
      }
    for (float& k : kernel) k /= sum;
  }
  // nodes the whole kernel covers are convolved, the others keep their value
  // (inversion.cpp:422-441); only the model region is read back
#pragma omp parallel for schedule(static)
  for (int j = 0; j < nx; ++j)
    for (int i = 0; i < nz; ++i) {
      const int ja = j + samples, ia = i + samples;
      float v = aux[ia + (size_t)ja * anz];
      if (ja >= mid && ja < anx - mid && ia >= mid && ia < anz - mid) {
        float acc = 0.0f;
        for (int xk = 0; xk < samples; ++xk)
          for (int zk = 0; zk < samples; ++zk)
            acc += aux[(ia - mid + zk) + (size_t)(ja - mid + xk) * anz] * kernel[zk + (size_t)xk * samples];
        v = acc;
      }
      out[i + (size_t)j * nz] = v;
    }
}

// ---------------------------------------------------------------------------
// Inversion

Inversion::~Inversion() {
  if (batch_) seisfat_batch_destroy(batch_);
}

void Inversion::set_parameters() {
  max_iteration = std::stoi(first_of({"max_iteration", "inv_max_it"}, parameters));
  tk_order = std::stoi(catch_parameter("tk_order", parameters));
  tk_param = std::stof(catch_parameter("tk_param", parameters));
  obs_data_folder = first_of({"obs_data_folder", "inv_data_folder"}, parameters);
  obs_data_prefix = first_of({"obs_data_prefix", "inv_data_prefix"}, parameters);
  smoother_stdv = std::stof(catch_parameter("gaussian_filter_stdv", parameters));
  smoother_samples = std::stoi(catch_parameter("gaussian_filter_samples", parameters));
  convergence_map_folder = first_of({"convergence_folder", "inv_residuo_folder"}, parameters);
  estimated_model_folder = first_of({"inversion_output_folder", "inv_output_folder"}, parameters);
  smooth_model_per_iteration = str2bool(catch_parameter("smooth_per_iteration", parameters));
  gradient_folder = catch_parameter_or("inv_gradient_folder", parameters, "");
  if (max_iteration < 0) throw std::invalid_argument("Error: the iteration count must not be negative");
  if (tk_order < 0) throw std::invalid_argument("Error: tk_order must not be negative");
  if (smooth_model_per_iteration && (smoother_samples < 1 || smoother_samples % 2 == 0 || smoother_stdv < 0.0f))
    throw std::invalid_argument("Error: gaussian_filter_samples must be a positive odd number, stdv >= 0");
  set_modeling_type();
  if (tk_order >= modeling->nPoints) throw std::invalid_argument("Error: tk_order exceeds the model size");
}

void Inversion::interior(std::vector<float>& s) const {
  const Modeling& m = *modeling;
  s.resize((size_t)m.nPoints);
  m.reduce_boundary(m.S, s.data());
}

void Inversion::import_obsData() {
  Geometry& g = *modeling->geometry;
  n_data = g.nsrc * g.nrec;
  n_model = modeling->nPoints;
  dobs.assign((size_t)n_data, 0.0f);
  dcal.assign((size_t)n_data, 0.0f);
  for (int k = 0; k < g.nsrc; ++k)
    import_binary_float(obs_data_folder + obs_data_prefix + std::to_string(k + 1) + ".bin",
                        dobs.data() + (size_t)k * g.nrec, (size_t)g.nrec);
  // every shot on the whole model with every receiver, in one batch
  std::vector<seisfat_shot> shots(g.nsrc);
  std::vector<float> rx((size_t)n_data), rz((size_t)n_data);
  for (int k = 0; k < g.nsrc; ++k) {
    shots[k] = seisfat_shot{g.xsrc[k], g.zsrc[k], 0, modeling->nx, k * g.nrec, g.nrec};
    std::copy(g.xrec, g.xrec + g.nrec, rx.begin() + (size_t)k * g.nrec);
    std::copy(g.zrec, g.zrec + g.nrec, rz.begin() + (size_t)k * g.nrec);
  }
  if (batch_) seisfat_batch_destroy(batch_);
  batch_ = nullptr;
  check_status(seisfat_batch_create(modeling->handle(), SEISFAT_ISO, modeling->schedule, g.nsrc, shots.data(),
                                    n_data, rx.data(), rz.data(), &batch_),
               "seisfat_batch_create");
  interior(s0_);
}

void Inversion::solve_picks(std::vector<float>& picks) {
  picks.resize((size_t)n_data);
  check_status(seisfat_batch_run(batch_, nullptr), "seisfat_batch_run");
  check_status(seisfat_batch_picks(batch_, picks.data()), "seisfat_batch_picks");
}

void Inversion::forward_modeling() {
  show_information();
  if (iteration < max_iteration) {
    gradient.assign((size_t)n_model, 0.0);
    double J = 0.0;
    check_status(seisfat_batch_misfit_gradient(batch_, dobs.data(), &J, gradient.data(), dcal.data()),
                 "seisfat_batch_misfit_gradient");
  } else {
    solve_picks(dcal);
  }
}

void Inversion::show_information() {
  std::cout << "\nInversion type: " << inversion_method << "\n\n";
  if (iteration == max_iteration) {
    std::cout << "-------- Checking final residuo --------\n\n";
  } else {
    std::cout << "-------- Computing iteration " << iteration + 1 << " of " << max_iteration << " --------\n\n";
    if (iteration > 0) std::cout << "Previous residuo: " << residuo.back() << "\n\n";
  }
}

void Inversion::check_convergence() {
  // src/inversion/inversion.cpp:179-198
  float square_difference = 0.0f;
  for (int i = 0; i < n_data; i++) square_difference += powf(dobs[i] - dcal[i], 2.0f);
  residuo.push_back(sqrtf(square_difference));
  if (iteration >= max_iteration) {
    std::cout << "Final residuo: " << residuo.back() << "\n";
    converged = true;
  } else {
    iteration += 1;
    converged = false;
  }
}

double Inversion::objective(const std::vector<float>& s, const std::vector<float>& picks) const {
  double data = 0.0;
  for (int i = 0; i < n_data; ++i) {
    const double r = (double)picks[i] - (double)dobs[i];
    data += 0.5 * r * r;
  }
  if (tk_param == 0.0f) return data;
  std::vector<double> m((size_t)n_model), y((size_t)(n_model - tk_order));
  for (int c = 0; c < n_model; ++c) m[c] = coverage[c] * ((double)s[c] - (double)s0_[c]);
  tikhonov_apply(tk_order, m.data(), n_model, y.data());
  double reg = 0.0;
  for (double v : y) reg += v * v;
  const double tk = tk_param;
  return data + 0.5 * tk * tk * reg;
}

void Inversion::optimization() {
  const Modeling& M = *modeling;
  // ray coverage: J^T 1 of the recorded solve (the column sums of the
  // reference's G, src/inversion/tomography_iso.cpp:34-38)
  std::vector<float> ones((size_t)n_data, 1.0f);
  coverage.assign((size_t)n_model, 0.0);
  check_status(seisfat_batch_vjp(batch_, ones.data(), coverage.data()), "seisfat_batch_vjp");
  std::vector<float> s;
  interior(s);
  // gradient of Phi: dJ/ds + tk^2 R o D^T D (R o (s - s0))
  std::vector<double> g(gradient);
  if (tk_param != 0.0f) {
    std::vector<double> m((size_t)n_model), y((size_t)(n_model - tk_order)), x((size_t)n_model);
    for (int c = 0; c < n_model; ++c) m[c] = coverage[c] * ((double)s[c] - (double)s0_[c]);
    tikhonov_apply(tk_order, m.data(), n_model, y.data());
    tikhonov_adjoint(tk_order, y.data(), n_model, x.data());
    const double tk2 = (double)tk_param * (double)tk_param;
    for (int c = 0; c < n_model; ++c) g[c] += tk2 * coverage[c] * x[c];
  }
  // direction: the negative gradient, smoothed like the reference's update
  std::vector<float> p((size_t)n_model);
  for (int c = 0; c < n_model; ++c) p[c] = (float)(-g[c]);
  if (smooth_model_per_iteration) {
    std::vector<float> ps((size_t)n_model);
    smooth_model(p.data(), ps.data(), M.nz, M.nx, smoother_stdv, smoother_samples);
    p.swap(ps);
  }
  double gp = 0.0;
  for (int c = 0; c < n_model; ++c) gp += g[c] * (double)p[c];
  if (!(gp < 0.0)) {  // the smoothed direction is not a descent direction: plain gradient
    gp = 0.0;
    for (int c = 0; c < n_model; ++c) {
      p[c] = (float)(-g[c]);
      gp += g[c] * (double)p[c];
    }
  }
  phi_before = objective(s, dcal);
  float pmax = 0.0f;
  double smean = 0.0;
  for (int c = 0; c < n_model; ++c) {
    pmax = std::max(pmax, std::fabs(p[c]));
    smean += s[c];
  }
  smean /= n_model;
  dS.assign((size_t)n_model, 0.0f);
  alpha = 0.0f;
  trials = 0;
  phi_after = phi_before;
  if (pmax > 0.0f && gp < 0.0) {
    // first trial: the last accepted step doubled; at first, a 1 % change of
    // the mean slowness where the direction is largest; at most 5 %
    const float cap = (float)(0.05 * smean) / pmax;
    float a = alpha_next_ > 0.0f ? std::min(alpha_next_, cap) : (float)(0.01 * smean) / pmax;
    std::vector<float> picks, st((size_t)n_model), step((size_t)n_model);
    trial_.assign(M.S, M.S + M.matsize);
    for (trials = 1; trials <= 12; ++trials, a *= 0.5f) {
      for (int c = 0; c < n_model; ++c) {
        step[c] = a * p[c];
        st[c] = s[c] + step[c];
      }
      bool ok = true;
      for (int j = 0; j < M.nx && ok; ++j)
        for (int i = 0; i < M.nz; ++i) {
          const float v = st[i + (size_t)j * M.nz];
          if (!(v > 0.0f) || !std::isfinite(v)) {
            ok = false;
            break;
          }
          trial_[(i + M.nb) + (size_t)(j + M.nb) * M.nzz] = v;
        }
      if (!ok) continue;  // a step past zero slowness: shorter
      check_status(seisfat_eik_set_slowness_padded(M.handle(), trial_.data()), "seisfat_eik_set_slowness_padded");
      solve_picks(picks);
      const double phi = objective(st, picks);
      if (phi <= phi_before + 1e-4 * (double)a * gp) {
        alpha = a;
        phi_after = phi;
        dS = step;
        break;
      }
    }
  }
  alpha_next_ = alpha > 0.0f ? 2.0f * alpha : 0.25f * alpha_next_;
  if (alpha == 0.0f) {
    // no acceptable step: the device model goes back to S, and the update is 0
    check_status(seisfat_eik_set_slowness_padded(M.handle(), M.S), "seisfat_eik_set_slowness_padded");
  }
  if (!gradient_folder.empty()) {
    const std::string it = std::to_string(iteration);
    std::ofstream(gradient_folder + "gradient_" + it + ".bin", std::ios::binary)
        .write(reinterpret_cast<const char*>(gradient.data()), (std::streamsize)(gradient.size() * sizeof(double)));
    std::ofstream(gradient_folder + "coverage_" + it + ".bin", std::ios::binary)
        .write(reinterpret_cast<const char*>(coverage.data()), (std::streamsize)(coverage.size() * sizeof(double)));
    export_binary_float(gradient_folder + "update_" + it + ".bin", dS.data(), dS.size());
    std::ofstream step(gradient_folder + "step_" + it + ".txt");
    step.precision(17);
    step << "alpha " << alpha << "\ntrials " << trials << "\nphi_before " << phi_before << "\nphi_after "
         << phi_after << "\ngp " << gp << "\n";
  }
  std::cout << "step " << alpha << " after " << trials << " trial(s): objective " << phi_before << " -> "
            << phi_after << "\n";
}

void Inversion::export_results() {
  const std::string path =
      convergence_map_folder + inversion_name + "_convergence_" + std::to_string(iteration) + "_iterations.txt";
  export_estimated_models();
  std::ofstream resFile(path, std::ios::out);
  if (!resFile.is_open()) throw std::invalid_argument("Error: " + path + " could not be opened!");
  for (float r : residuo) resFile << r << "\n";
  std::cout << "Text file " << path << " was successfully written." << std::endl;
}

// ---------------------------------------------------------------------------
// Tomography_ISO

void Tomography_ISO::set_modeling_type() {
  modeling.reset(new Eikonal_ISO());
  modeling->parameters = parameters;
  modeling->set_parameters();
  inversion_name = "tomography_iso";
  inversion_method = "Isotropic First-Arrival Tomography";
  dS.assign((size_t)modeling->nPoints, 0.0f);
}

void Tomography_ISO::model_update() {
  Modeling& m = *modeling;
  for (int index = 0; index < m.nPoints; index++) {
    const int i = index % m.nz;
    const int j = index / m.nz;
    m.S[(i + m.nb) + (size_t)(j + m.nb) * m.nzz] += dS[index];
  }
  m.copy_slowness_to_device();
}

void Tomography_ISO::export_estimated_models() {
  std::vector<float> vp;
  interior(vp);
  for (float& v : vp) v = 1.0f / v;
  export_binary_float(estimated_model_folder + inversion_name + "_final_model_vp_" + std::to_string(modeling->nz) +
                          "x" + std::to_string(modeling->nx) + ".bin",
                      vp.data(), vp.size());
}

}  // namespace seisfat
