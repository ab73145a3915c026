// Traveltime-table driver: the table stage of SeisFAT2D's migration
// (src/migration/KDM.cu:3-7 -> Migration::set_src_travel_times /
// set_rec_travel_times, src/migration/migration.cu:116-171).
//
//   seisfat2d_tables <parameters.txt> [--chunk N] [--src-only | --rec-only]
//
// Reads the modeling keys plus mig_tables_folder and anisotropy (as
// src/migration/migration.cu:22-27 does) and writes
// <mig_tables_folder>eikonal_src_<k>.bin for every source of the SPS file and
// eikonal_rec_<k>.bin for every receiver of the RPS file.
#include <chrono>
#include <cstring>
#include <iostream>
#include <memory>

#include "admin.hpp"
#include "modeling.hpp"
#include "tables.hpp"

using namespace seisfat;

int main(int argc, char** argv) {
  if (argc < 2) {
    std::cerr << "usage: " << argv[0] << " <parameters.txt> [--chunk N] [--src-only | --rec-only]\n";
    return 2;
  }
  int chunk = 64;
  bool src = true, rec = true;
  for (int a = 2; a < argc; ++a) {
    if (!std::strcmp(argv[a], "--chunk") && a + 1 < argc) chunk = std::stoi(argv[++a]);
    else if (!std::strcmp(argv[a], "--src-only")) rec = false;
    else if (!std::strcmp(argv[a], "--rec-only")) src = false;
    else {
      std::cerr << "unknown option " << argv[a] << "\n";
      return 2;
    }
  }
  try {
    const std::string file = argv[1];
    const bool ani = str2bool(catch_parameter_or("anisotropy", file, "false"));
    std::unique_ptr<Modeling> modeling(ani ? (Modeling*)new Eikonal_ANI() : (Modeling*)new Eikonal_ISO());
    modeling->parameters = file;
    modeling->set_parameters();
    TravelTimeTables tables;
    tables.modeling = modeling.get();
    tables.chunk = chunk;
    tables.set_parameters();
    const auto ti = std::chrono::steady_clock::now();
    if (src) tables.set_src_travel_times();
    if (rec) tables.set_rec_travel_times();
    const auto tf = std::chrono::steady_clock::now();
    std::cout << tables.tables_written << " tables, batched solves " << tables.solve_ms << " ms on the device\n";
    std::cout << "\nRun time: " << std::chrono::duration<double>(tf - ti).count() << " s." << std::endl;
  } catch (const std::exception& ex) {
    std::cerr << ex.what() << std::endl;
    return 1;
  }
  return 0;
}
