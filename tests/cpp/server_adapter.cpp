// Drives the serving runtime through the C++ adapter (lpm_gpu::Server over
// lpmg_server_*): one simple-cubic solve submitted twice as concurrent
// requests; both results must agree bit for bit and are printed as
// displacements, one particle per line.
#include <cstdio>
#include <cstdlib>
#include <string>

#include "lpm_gpu.hpp"

using namespace lpm_gpu;

int main(int argc, char** argv) {
  const int n = argc > 1 ? std::atoi(argv[1]) : 6;
  const int n_steps = argc > 2 ? std::atoi(argv[2]) : 3;
  const double radius = 0.1, total = 0.05;
  std::vector<std::array<double, 3>> xyz;
  for (int k = 0; k < n; ++k)
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) xyz.push_back({2 * radius * i, 2 * radius * j, 2 * radius * k});
  const double zmax = 2 * radius * (n - 1);
  std::vector<int> top, bottom;
  for (int p = 0; p < (int)xyz.size(); ++p) {
    if (xyz[p][2] < 0.5 * radius) bottom.push_back(p);
    if (xyz[p][2] > zmax - 0.5 * radius) top.push_back(p);
  }
  // the same solve as a request to the serving runtime, next to a second copy
  std::vector<double> flat, params{0.0, 0, 70e3, 0.33, 1.0}, out0(3 * xyz.size()), out1(3 * xyz.size());
  for (auto& q : xyz) flat.insert(flat.end(), q.begin(), q.end());
  std::vector<int32_t> gsize{(int32_t)bottom.size(), (int32_t)top.size()}, gidx(bottom.begin(), bottom.end());
  gidx.insert(gidx.end(), top.begin(), top.end());
  lpmg_bc bcs[4] = {{0, 0, LPMG_RELATIVE, 'x', {0, 0, 0}}, {0, 0, LPMG_RELATIVE, 'y', {0, 0, 0}},
                    {0, 0, LPMG_RELATIVE, 'z', {0, 0, 0}}, {0, 1, LPMG_RELATIVE, 'z', {total / n_steps, 0, 0}}};
  lpmg_request r{};
  r.lattice = LPMG_SC3D; r.radius = radius; r.ptype = LPMG_ELASTIC;
  r.n = (int)xyz.size(); r.xyz = flat.data();
  r.model = 0; r.n_params = 5; r.params = params.data();
  r.n_groups = 2; r.group_size = gsize.data(); r.group_idx = gidx.data();
  r.n_bc = 4; r.bc = bcs; r.n_steps = n_steps;
  r.undamaged = -1; r.max_iter = 30; r.tol = 1e-11; r.max_halvings = 8;
  const double o[8] = {1, 1e-24, 0.0, 0, 0, 16, 0, 0};
  for (int k = 0; k < 8; ++k) r.opts[k] = o[k];
  lpmg_out_field f0{LPMG_F_XYZ, (int64_t)out0.size(), out0.data()}, f1{LPMG_F_XYZ, (int64_t)out1.size(), out1.data()};
  Server srv(2);
  r.n_out = 1; r.out = &f0;
  const int64_t t0 = srv.submit(r);
  r.out = &f1;
  const int64_t t1 = srv.submit(r);
  srv.wait(t0);
  srv.wait(t1);
  for (size_t k = 0; k < out0.size(); ++k)
    if (out0[k] != out1[k]) { std::fprintf(stderr, "served copies differ\n"); return 1; }
  for (size_t p = 0; p < xyz.size(); ++p)
    std::printf("%.17g %.17g %.17g\n", out0[3 * p] - xyz[p][0], out0[3 * p + 1] - xyz[p][1], out0[3 * p + 2] - xyz[p][2]);
  return 0;
}
