// Drives cyclic fatigue loading through the reference-named headers of the
// GPU adapter (include/lpmcpp) the way
// examples/ex5_isoelasticity_2d_fatigue_gage.cpp drives the reference:
// FatigueHCF particles of a 2D hexagonal plate, setParticleProperty, nonlocal
// neighbours, a static step to the spectrum's first value, then
// SolverFatigue::solveProblemCyclic under LoadUniaxialDisp.
//
// Input (whitespace-separated, from argv[1]): N; N lines x y z; N types; for
// the left, top and bottom groups a count then the indices; then undamaged
// type, nonlocal L, cutoff ratio, tau, d_max and the number of simulated cycles.
// Output: one line per particle, "ux uy damage".
#include "lpm.h"
#include "assembly.h"
#include "load_step_fatigue.h"
#include "solver_fatigue.h"

#include <fstream>

int main(int argc, char** argv) {
  if (argc < 2) return 2;
  std::ifstream in(argv[1]);
  int n = 0;
  in >> n;
  std::vector<std::array<double, NDIM>> xyz(n);
  for (auto& p : xyz) in >> p[0] >> p[1] >> p[2];
  std::vector<int> type(n);
  for (auto& t : type) in >> t;
  std::vector<std::vector<int>> idx(3);
  for (auto& g : idx) {
    int m = 0;
    in >> m;
    g.resize(m);
    for (auto& v : g) in >> v;
  }
  int undamaged = 0, n_sim = 0;
  double L = 0, ratio = 0, tau = 0, d_max = 0;
  in >> undamaged >> L >> ratio >> tau >> d_max >> n_sim;
  if (!in) return 3;

  const int n_layer = 2;
  UnitCell cell(LatticeType::Hexagon2D, 0.12);
  std::array<double, 2 * NDIM> box{0.0, 5.08, 0.0, 14.7, 0.0, 4.8};
  Assembly<n_layer> pt_ass{xyz, box, cell, ParticleType::FatigueHCF};
  std::vector<std::vector<Particle<n_layer>*>> groups(3);
  for (int g = 0; g < 3; ++g)
    for (int i : idx[g]) groups[g].push_back(pt_ass.pt_sys[i]);
  for (Particle<n_layer>* p1 : pt_ass.pt_sys) {
    p1->type = type[p1->id];
    ParticleFatigueHCF<n_layer>* ftpt = dynamic_cast<ParticleFatigueHCF<n_layer>*>(p1);
    ftpt->setParticleProperty(L, true, 71.7e3, 0.306, 2.003e-5, 2.833, 0.5, 1.0, 1.108);
  }
  pt_ass.searchNonlocalNeighbors(ratio);
  pt_ass.updateGeometry();
  pt_ass.updateForceState();

  SolverFatigue<n_layer> solv(undamaged, pt_ass, StiffnessMode::Analytical, SolverMode::CG, TimeMapMode::Linear, tau,
                              "unused.dump", 30, 1e-11);
  solv.opts = {1, 1e-24, 0.0, 0, 0, 16, 0, 0};
  std::vector<LoadStep<n_layer>> load_static(1);
  load_static[0].dispBCs.push_back(DispBC<n_layer>(groups[0], LoadMode::Relative, 'x', 0.0));
  load_static[0].dispBCs.push_back(DispBC<n_layer>(groups[1], LoadMode::Relative, 'y', 0.0));
  load_static[0].dispBCs.push_back(DispBC<n_layer>(groups[2], LoadMode::Relative, 'y', 0.0));
  int start = 0;
  solv.solveProblemStatic(load_static, start);

  const int dN = func_dNdt_linear(tau);
  solv.load_spectrum.push_back(0.0);
  for (int c = 0; c < n_sim * dN; ++c) {
    solv.load_spectrum.push_back(d_max);
    solv.load_spectrum.push_back(0.0);
  }
  const int N = solv.solveProblemCyclic(FatigueLoadType::LoadUniaxialDisp, 1, groups);
  if (N != n_sim * dN) return 4;
  const auto& sent = pt_ass.sends();
  std::fprintf(stderr, "host edits sent: types %lld, xyz %lld, Pex %lld, damage %lld, constraints %lld\n",
               (long long)sent[0], (long long)sent[1], (long long)sent[2], (long long)sent[3], (long long)sent[4]);
  for (Particle<n_layer>* p : pt_ass.pt_sys)
    std::printf("%.17g %.17g %.17g\n", p->xyz[0] - p->xyz_initial[0], p->xyz[1] - p->xyz_initial[1], p->damage);
  return 0;
}
