// Sanitizer driver of the CPU oracle (tests/test_sanitizers.py builds it with
// AddressSanitizer and UndefinedBehaviorSanitizer, any report fatal): every C
// entry point of oracle/oracle.cpp on small inputs, including the edge cases
// the reference leaves undefined (a constant stiffness component, sources and
// receivers on the model's last row and column).
#include "../../oracle/oracle.cpp"

#include <cstdio>
#include <random>

int main() {
  const int nz = 23, nx = 31, nzz = nz + 6, nxx = nx + 6;
  const float dx = 10.0f, dz = 3.0f;
  std::mt19937 gen(7);
  std::uniform_real_distribution<float> U(1.0f / 6000.0f, 1.0f / 1500.0f);
  std::vector<float> s((size_t)nz * nx), S((size_t)nzz * nxx), T(S.size()), T2(S.size()), red(s.size());
  for (float& v : s) v = U(gen);
  orc_expand_boundary_f32(s.data(), S.data(), nz, nx);
  orc_reduce_boundary_f32(S.data(), red.data(), nz, nx);
  int sIdx = 0, sIdz = 0;
  orc_init_f32(T.data(), S.data(), nzz, nxx, dx, dz, (nx - 1) * dx, (nz - 1) * dz, &sIdx, &sIdz);
  orc_fsm_f32(T.data(), S.data(), nzz, nxx, dx, dz, 0, 1, 1);
  orc_solve_iso_f32(T2.data(), S.data(), nzz, nxx, dx, dz, 0.0f, 0.0f, 1, 20);
  std::vector<double> s64(s.begin(), s.end()), S64(S.size()), T64(S.size());
  orc_expand_boundary_f64(s64.data(), S64.data(), nz, nx);
  orc_solve_iso_f64(T64.data(), S64.data(), nzz, nxx, dx, dz, 100.0, 30.0, 0, 1);
  long long counts[7];
  orc_branch_counts_f32(S.data(), nzz, nxx, dx, dz, 150.0f, 20.0f, counts);
  orc_sweep_counted_f32(T.data(), S.data(), nzz, nxx, dx, dz, 3, counts);
  // windows
  const float wsx[2] = {50.0f, 200.0f}, wsz[2] = {0.0f, (nz - 1) * dz};
  const int wx0[2] = {0, 12}, ww[2] = {nx, 9};
  const long long toff[2] = {0, (long long)nzz * nxx};
  std::vector<float> Tw((size_t)nzz * (nxx + 15));
  orc_solve_windows_f32(s.data(), nz, nx, dx, dz, 2, wsx, wsz, wx0, ww, toff, Tw.data(), 2);
  // stiffness: one constant component (the reference's 0/0)
  std::vector<float> c(S.size(), 0.0f);
  std::vector<uint16_t> q[6];
  float mn[6], mx[6];
  for (int e = 0; e < 6; ++e) {
    q[e].resize(S.size());
    for (size_t k = 0; k < c.size(); ++k) c[k] = (e == 2 || e == 4) ? 0.0f : 1e9f * (1.0f + 0.01f * (float)(k % 7) + e);
    orc_compress(c.data(), q[e].data(), (long long)c.size(), &mx[e], &mn[e]);
  }
  std::vector<float> Sw(S.size());
  orc_solve_ani_f32(T.data(), S.data(), Sw.data(), nzz, nxx, dx, dz, (nx - 1) * dx, 0.0f, q[0].data(), q[1].data(),
                    q[2].data(), q[3].data(), q[4].data(), q[5].data(), mn, mx);
  // picks on the last row and column, adjoints, batch misfit, rays
  const float rx[3] = {0.0f, (nx - 1) * dx, 123.4f}, rz[3] = {(nz - 1) * dz, 0.0f, 17.0f};
  float pk[3], obs[3] = {0.1f, 0.2f, 0.3f};
  orc_picks_f32(T2.data(), nzz, dx, dz, rx, rz, 3, pk);
  std::vector<double> g(S.size(), 0.0);
  orc_adjoint_f32(S.data(), nzz, nxx, dx, dz, 150.0f, 10.0f, rx, rz, 3, obs, g.data(), pk);
  const double rxd[3] = {rx[0], rx[1], rx[2]}, rzd[3] = {rz[0], rz[1], rz[2]}, obsd[3] = {0.1, 0.2, 0.3};
  double pkd[3];
  std::vector<double> gd(S.size(), 0.0);
  orc_adjoint_f64(S64.data(), nzz, nxx, dx, dz, 150.0, 10.0, rxd, rzd, 3, obsd, gd.data(), pkd);
  const int roff[2] = {0, 2}, nrec[2] = {2, 1};
  const float brx[3] = {10.0f, 200.0f, 150.0f}, brz[3] = {0.0f, 5.0f, 9.0f};
  std::vector<double> bg(s.size());
  float bpk[3];
  orc_batch_misfit_f32(s.data(), nz, nx, dx, dz, 2, wsx, wsz, wx0, ww, roff, nrec, brx, brz, obs, bg.data(), bpk, 2);
  std::vector<int> orec(4096), ocell(4096);
  std::vector<float> olen(4096);
  orc_solve_iso_f32(T2.data(), S.data(), nzz, nxx, dx, dz, 150.0f, 0.0f, 0, 1);
  const float rrx[2] = {20.0f, 290.0f}, rrz[2] = {60.0f, 3.0f};
  orc_ray_paths(T2.data(), nz, nx, dx, dz, 150.0f, 0.0f, rrx, rrz, 2, 100000, orec.data(), ocell.data(), olen.data(),
                (int)orec.size());
  std::printf("oracle sanitizer run ok (%lld A-branch updates)\n", counts[3]);
  return 0;
}
