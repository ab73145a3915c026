// Traveltime tables for Kirchhoff migration: SeisFAT2D's
// Migration::set_src_travel_times / set_rec_travel_times
// (src/migration/migration.cu:116-134,153-171) on the batched GPU solver.
//
// The reference solves one shot per source (and, by reciprocity, one per
// receiver) with the per-shot modeling call and writes each padded
// traveltime grid to <tables_folder>eikonal_src_<k>.bin (eikonal_rec_<k>.bin),
// k counted from 1, nzz*nxx float32 z-fastest, which KDM / LSKDM read back
// (src/migration/KDM.cu:18,41).  Here the positions are solved in batches of
// whole-model shots (same bits as the per-shot call), and each table is
// written as its batch finishes.
#pragma once

#include <string>

#include "modeling.hpp"

namespace seisfat {

class TravelTimeTables {
 public:
  Modeling* modeling = nullptr;  // set_parameters() done; not owned
  std::string tables_folder;     // mig_tables_folder (src/migration/migration.cu:22)
  int chunk = 64;                // positions per batched solve

  // Reads mig_tables_folder from the modeling object's parameter file.
  void set_parameters();
  // eikonal_src_<k>.bin for every source, eikonal_rec_<k>.bin for every receiver.
  void set_src_travel_times();
  void set_rec_travel_times();

  // Number of tables written by the calls so far, and the device time of
  // their batched solves (HIP events around each batch's sweeps, ms).
  int tables_written = 0;
  double solve_ms = 0.0;

 private:
  void build(const float* x, const float* z, int n, const std::string& prefix);
};

}  // namespace seisfat
