// First-arrival traveltime tomography on the adjoint-state gradient.
//
// Keeps the interface and main loop of SeisFAT2D's Inversion /
// Tomography_ISO (src/inversion/inversion.hpp, inversion.cpp,
// tomography_iso.cpp, src/inversion_main.cpp):
//
//   inv->parameters = file; inv->set_parameters(); inv->import_obsData();
//   while (true) {
//     inv->forward_modeling(); inv->check_convergence();
//     if (inv->converged) break;
//     inv->optimization(); inv->model_update();
//   }
//   inv->export_results();
//
// with the same parameter keys, observed-data files, residual history,
// Tikhonov operator (set_regularization_matrix), Gaussian smoothing of the
// update (model_smoothing) and output files.  What differs is the step: the
// reference traces rays through the traveltimes to build a sparse sensitivity
// G and solves the Tikhonov-regularised normal equations by 10 CGLS
// iterations (inversion.cpp:95-177,264-345), which does not compile as
// shipped; here the exact gradient of the discrete single-pass solve
// (seisfat_batch_misfit_gradient, all shots in one batch) gives a steepest-
// descent direction of
//
//   Phi(s) = 1/2 |picks(s) - dobs|^2 + 1/2 |tk_param * D (R o (s - s0))|^2
//
// (D the reference's order-tk_order difference operator, R the ray coverage
// G^T 1 -- here J^T 1 of the recorded solve -- as in the reference's
// regularisation rows, s0 the starting model), smoothed like the reference's
// update, and the step length comes from a backtracking (Armijo) line search
// on Phi with one batched forward solve per trial, so every accepted step
// lowers Phi.
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "modeling.hpp"

namespace seisfat {

class Inversion {
 public:
  std::string parameters;

  int iteration = 0;
  int max_iteration = 0;
  bool converged = false;

  int tk_order = 0;
  float tk_param = 0.0f;
  int smoother_samples = 0;
  float smoother_stdv = 0.0f;
  bool smooth_model_per_iteration = false;

  int n_data = 0;
  int n_model = 0;
  std::vector<float> dobs, dcal;  // observed and computed picks, shot after shot
  std::vector<float> dS;          // model update of the last optimization()
  std::vector<float> residuo;     // |dobs - dcal| per forward_modeling()

  std::string obs_data_folder, obs_data_prefix;
  std::string convergence_map_folder, estimated_model_folder;
  std::string gradient_folder;  // optional: per-iteration gradient, coverage, update and step files
  std::string inversion_name, inversion_method;

  // the last optimization(): misfit gradient dJ/ds, coverage J^T 1, step
  // length, the objective before and after it, and the trials it took
  std::vector<double> gradient, coverage;
  float alpha = 0.0f;
  double phi_before = 0.0, phi_after = 0.0;
  int trials = 0;

  std::unique_ptr<Modeling> modeling;

  // Reads the inversion keys (the code's names and the parameter file's:
  // max_iteration | inv_max_it, obs_data_folder | inv_data_folder, ...) and
  // sets up the modeling object.
  void set_parameters();
  // <obs_data_folder><obs_data_prefix><k>.bin, k = 1..nsrc, nrec floats each.
  void import_obsData();
  // All shots in one batched solve; the misfit gradient too unless this is
  // the final check (iteration == max_iteration).
  void forward_modeling();
  void show_information();
  // residuo.push_back(|dobs - dcal|); converged once iteration reaches
  // max_iteration (src/inversion/inversion.cpp:179-198).
  void check_convergence();
  // The regularised, smoothed steepest-descent step and its line search.
  void optimization();
  virtual void model_update() = 0;
  // <convergence_folder><name>_convergence_<it>_iterations.txt and the
  // estimated model (src/inversion/inversion.cpp:449-463).
  void export_results();

  virtual void set_modeling_type() = 0;
  virtual void export_estimated_models() = 0;

  Inversion() = default;
  Inversion(const Inversion&) = delete;
  Inversion& operator=(const Inversion&) = delete;
  virtual ~Inversion();

 protected:
  seisfat_batch_t batch_ = nullptr;
  std::vector<float> s0_;     // starting interior slowness (nz*nx)
  std::vector<float> trial_;  // padded slowness of a line-search trial
  float alpha_next_ = 0.0f;

  void interior(std::vector<float>& s) const;     // the padded S's interior
  double objective(const std::vector<float>& s, const std::vector<float>& picks) const;
  void solve_picks(std::vector<float>& picks);
};

// Isotropic tomography (src/inversion/tomography_iso.cpp).
class Tomography_ISO : public Inversion {
 public:
  void set_modeling_type() override;
  // S[(i+nb) + (j+nb)*nzz] += dS[i + j*nz] over the interior, then
  // copy_slowness_to_device() (src/inversion/tomography_iso.cpp:77-92).
  void model_update() override;
  // <inversion_output_folder><name>_final_model_vp_<nz>x<nx>.bin
  // (src/inversion/tomography_iso.cpp:94-105).
  void export_estimated_models() override;
};

// The reference's Tikhonov operator (src/inversion/inversion.cpp:209-262):
// the order-k forward difference of the flattened model (identity for k = 0).
void tikhonov_apply(int order, const double* x, long long n, double* y);
void tikhonov_adjoint(int order, const double* y, long long n, double* x);
// Inversion::model_smoothing (src/inversion/inversion.cpp:347-447).
void smooth_model(const float* in, float* out, int nz, int nx, float stdv, int samples);

}  // namespace seisfat
