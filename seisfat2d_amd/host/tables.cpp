#include "tables.hpp"

#include <algorithm>
#include <stdexcept>
#include <vector>

#include "admin.hpp"

namespace seisfat {

void TravelTimeTables::set_parameters() {
  if (!modeling) throw std::invalid_argument("TravelTimeTables: no modeling object");
  tables_folder = catch_parameter("mig_tables_folder", modeling->parameters);
}

void TravelTimeTables::set_src_travel_times() {
  build(modeling->geometry->xsrc, modeling->geometry->zsrc, modeling->geometry->nsrc, "eikonal_src_");
}

void TravelTimeTables::set_rec_travel_times() {
  build(modeling->geometry->xrec, modeling->geometry->zrec, modeling->geometry->nrec, "eikonal_rec_");
}

void TravelTimeTables::build(const float* x, const float* z, int n, const std::string& prefix) {
  if (!modeling || !modeling->handle()) throw std::invalid_argument("TravelTimeTables: modeling not set up");
  if (chunk < 1) throw std::invalid_argument("TravelTimeTables: chunk must be positive");
  const int type = modeling->modeling_type == "eikonal_ani" ? SEISFAT_ANI : SEISFAT_ISO;
  std::vector<float> table((size_t)modeling->matsize);
  for (int k0 = 0; k0 < n; k0 += chunk) {
    const int m = std::min(chunk, n - k0);
    std::vector<seisfat_shot> shots(m);
    for (int q = 0; q < m; ++q) shots[q] = seisfat_shot{x[k0 + q], z[k0 + q], 0, modeling->nx, 0, 0};
    seisfat_batch_t b = nullptr;
    check_status(seisfat_batch_create(modeling->handle(), type, modeling->schedule, m, shots.data(), 0, nullptr,
                                      nullptr, &b),
                 "seisfat_batch_create");
    int rc = seisfat_batch_run(b, nullptr);
    if (rc == SEISFAT_OK) rc = seisfat_batch_status(b);
    float ms = 0.0f;
    if (rc == SEISFAT_OK) rc = seisfat_batch_sweep_ms(b, &ms);
    for (int q = 0; q < m && rc == SEISFAT_OK; ++q) {
      rc = seisfat_batch_copy_T(b, q, table.data());
      if (rc == SEISFAT_OK)
        export_binary_float(tables_folder + prefix + std::to_string(k0 + q + 1) + ".bin", table.data(),
                            table.size());
    }
    seisfat_batch_destroy(b);
    check_status(rc, "traveltime tables");
    solve_ms += ms;
    tables_written += m;
  }
}

}  // namespace seisfat
