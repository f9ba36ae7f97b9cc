// Drives cyclic fatigue loading through the C++ adapter (include/lpm_gpu.hpp)
// the way examples/ex5_isoelasticity_2d_fatigue_gage.cpp drives the reference:
// FatigueHCF particles of a 2D hexagonal plate, setParticleProperty, nonlocal
// neighbours, a static step to the spectrum's first value, then
// SolverFatigue::solveProblemCyclic under LoadUniaxialDisp.
//
// Input (whitespace-separated, from argv[1]): N; N lines x y z; N types; for
// the left, top and bottom groups a count then the indices; then undamaged
// type, nonlocal L, cutoff ratio, tau, d_max and the number of simulated cycles.
// Output: one line per particle, "ux uy damage".
#include <cstdio>
#include <fstream>

#include "lpm_gpu.hpp"

using namespace lpm_gpu;

int main(int argc, char** argv) {
  if (argc < 2) return 2;
  std::ifstream in(argv[1]);
  int n = 0;
  in >> n;
  std::vector<std::array<double, 3>> xyz(n);
  for (auto& p : xyz) in >> p[0] >> p[1] >> p[2];
  std::vector<int32_t> type(n);
  for (auto& t : type) in >> t;
  std::vector<std::vector<int>> groups(3);
  for (auto& g : groups) {
    int m = 0;
    in >> m;
    g.resize(m);
    for (auto& v : g) in >> v;
  }
  int undamaged = 0, n_sim = 0;
  double L = 0, ratio = 0, tau = 0, d_max = 0;
  in >> undamaged >> L >> ratio >> tau >> d_max >> n_sim;
  if (!in) return 3;

  UnitCell cell(LatticeType::Hexagon2D, 0.12);
  std::array<double, 6> box{0.0, 5.08, 0.0, 14.7, 0.0, 4.8};
  Assembly<2> ass(xyz, box, cell, ParticleType::FatigueHCF);
  ass.setTypes(type);
  ass.setFatigueHCF(L, true, 71.7e3, 0.306, 2.003e-5, 2.833, 0.5, 1.0, 1.108);
  ass.searchNonlocalNeighbors(ratio);
  ass.updateGeometry();
  ass.updateForceState();

  SolverFatigue<2> solv(undamaged, ass, StiffnessMode::Analytical, SolverMode::CG, TimeMapMode::Linear, tau,
                        "unused.dump", 30, 1e-11);
  solv.opts = {1, 1e-24, 0.0, 0, 0, 16, 0, 0};
  std::vector<LoadStep<2>> load_static(1);
  load_static[0].dispBCs.push_back(DispBC<2>(groups[0], LoadMode::Relative, 'x', 0.0));
  load_static[0].dispBCs.push_back(DispBC<2>(groups[1], LoadMode::Relative, 'y', 0.0));
  load_static[0].dispBCs.push_back(DispBC<2>(groups[2], LoadMode::Relative, 'y', 0.0));
  int start = 0;
  solv.solveProblemStatic(load_static, start);

  const int dN = SolverFatigue<2>::func_dNdt_linear(tau);
  solv.load_spectrum.push_back(0.0);
  for (int c = 0; c < n_sim * dN; ++c) {
    solv.load_spectrum.push_back(d_max);
    solv.load_spectrum.push_back(0.0);
  }
  const int N = solv.solveProblemCyclic(FatigueLoadType::LoadUniaxialDisp, 1, groups);
  if (N != n_sim * dN) return 4;
  auto now = ass.positions();
  auto dmg = ass.field(LPMG_F_DAMAGE);
  for (int p = 0; p < n; ++p)
    std::printf("%.17g %.17g %.17g\n", now[p][0] - xyz[p][0], now[p][1] - xyz[p][1], dmg[p]);
  return 0;
}
