// First-arrival tomography driver (SeisFAT2D src/inversion_main.cpp).
//
//   seisfat2d_inversion <parameters.txt>
//
// inversion_type (or inv_type) selects the plugin: 0 = Tomography_ISO.  The
// reference's Tomography_VTI (1) is not provided: it builds on stiffness
// accessors the reference's modeling classes do not have
// (src/inversion/tomography_vti.cpp:21,150), so it never compiled.
#include <chrono>
#include <iostream>
#include <memory>

#include "admin.hpp"
#include "inversion.hpp"

using namespace seisfat;

int main(int argc, char** argv) {
  if (argc < 2) {
    std::cerr << "usage: " << argv[0] << " <parameters.txt>\n";
    return 2;
  }
  try {
    const std::string file = argv[1];
    std::string t = catch_parameter_or("inversion_type", file, "");
    if (t.empty()) t = catch_parameter_or("inv_type", file, "0");
    const int type = std::stoi(t);
    if (type != 0) throw std::invalid_argument("Error: inversion_type must be 0 (isotropic tomography)");
    std::unique_ptr<Inversion> inversion(new Tomography_ISO());
    inversion->parameters = file;
    inversion->set_parameters();
    inversion->import_obsData();
    const auto ti = std::chrono::steady_clock::now();
    while (true) {
      inversion->forward_modeling();
      inversion->check_convergence();
      if (inversion->converged) break;
      inversion->optimization();
      inversion->model_update();
    }
    const auto tf = std::chrono::steady_clock::now();
    inversion->export_results();
    std::cout << "\nRun time: " << std::chrono::duration<double>(tf - ti).count() << " s." << std::endl;
  } catch (const std::exception& ex) {
    std::cerr << ex.what() << std::endl;
    return 1;
  }
  return 0;
}
