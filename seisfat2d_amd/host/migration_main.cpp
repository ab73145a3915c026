// Kirchhoff depth migration driver (SeisFAT2D src/migration_main.cpp).
//
//   seisfat2d_migration <parameters.txt> [--no-table-files]
//
// mig_type selects the plugin: 0 IDKDM, 1 ADKDM, 2 IDLSKDM, 3 ADLSKDM.  The
// traveltime tables are solved on the device and also written to
// mig_tables_folder as the reference does, unless --no-table-files.
#include <chrono>
#include <cstring>
#include <iostream>

#include "admin.hpp"
#include "migration.hpp"

using namespace seisfat;

int main(int argc, char** argv) {
  if (argc < 2) {
    std::cerr << "usage: " << argv[0] << " <parameters.txt> [--no-table-files]\n";
    return 2;
  }
  bool files = true;
  for (int a = 2; a < argc; ++a) {
    if (!std::strcmp(argv[a], "--no-table-files")) files = false;
    else {
      std::cerr << "unknown option " << argv[a] << "\n";
      return 2;
    }
  }
  try {
    const std::string file = argv[1];
    Migration migration;
    migration.type = std::stoi(catch_parameter("mig_type", file));
    migration.parameters = file;
    migration.write_tables = files;
    migration.set_parameters();
    const auto ti = std::chrono::steady_clock::now();
    migration.kirchhoff_depth_migration();
    const auto tf = std::chrono::steady_clock::now();
    migration.export_outputs();
    std::cout << migration.migType << ": " << migration.tsrc.size() << " traces, tables " << migration.tables_s
              << " s\n";
    std::cout << "\nRun time: " << std::chrono::duration<double>(tf - ti).count() << " s." << std::endl;
  } catch (const std::exception& ex) {
    std::cerr << ex.what() << std::endl;
    return 1;
  }
  return 0;
}
