// Drives the GPU path through the C++ adapter (include/lpm_gpu.hpp) the way
// LPM-CPP's examples drive the reference (examples/ex1_isoelasticity_3d_sc.cpp:
// build a lattice, set the material, define BC groups and load steps, then
// SolverStatic::solveProblem).  Prints the final displacement field.
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "lpm_gpu.hpp"

using namespace lpm_gpu;

int main(int argc, char** argv) {
  const int n = argc > 1 ? std::atoi(argv[1]) : 6;
  const int n_steps = argc > 2 ? std::atoi(argv[2]) : 3;
  const double radius = 0.1, total = 0.05;
  UnitCell cell(LatticeType::SimpleCubic3D, radius);
  std::vector<std::array<double, 3>> xyz;
  for (int k = 0; k < n; ++k)
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) xyz.push_back({2 * radius * i, 2 * radius * j, 2 * radius * k});
  std::array<double, 6> box{0, 2 * radius * (n - 1), 0, 2 * radius * (n - 1), 0, 2 * radius * (n - 1)};
  Assembly<2> ass(xyz, box, cell, ParticleType::Elastic);
  ass.setElastic(0.0, false, 70e3, 0.33, 1.0);
  std::vector<int> top, bottom;
  for (int p = 0; p < (int)xyz.size(); ++p) {
    if (xyz[p][2] < 0.5 * radius) bottom.push_back(p);
    if (xyz[p][2] > box[5] - 0.5 * radius) top.push_back(p);
  }
  std::vector<LoadStep<2>> load;
  for (int s = 0; s < n_steps; ++s) {
    LoadStep<2> step;
    step.dispBCs.push_back(DispBC<2>(bottom, LoadMode::Relative, 'x', 0.0));
    step.dispBCs.push_back(DispBC<2>(bottom, LoadMode::Relative, 'y', 0.0));
    step.dispBCs.push_back(DispBC<2>(bottom, LoadMode::Relative, 'z', 0.0));
    step.dispBCs.push_back(DispBC<2>(top, LoadMode::Relative, 'z', total / n_steps));
    load.push_back(step);
  }
  ass.updateGeometry();
  ass.updateForceState();
  SolverStatic<2> solv(-1, ass, StiffnessMode::Analytical, SolverMode::CG, "unused.dump", 30, 1e-11);
  solv.opts = {1, 1e-24, 0.0, 0, 0, 16, 0, 0};
  int start = 0;
  solv.solveProblem(load, start);
  auto now = ass.positions();
  for (size_t p = 0; p < xyz.size(); ++p)
    std::printf("%.17g %.17g %.17g\n", now[p][0] - xyz[p][0], now[p][1] - xyz[p][1], now[p][2] - xyz[p][2]);
  return 0;
}
