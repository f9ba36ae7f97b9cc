// A driver written the way LPM-CPP's examples drive the reference
// (examples/ex1_isoelasticity_3d_sc.cpp): only the include path differs
// (include/lpmcpp instead of LPM-CPP's include/).  It builds a simple-cubic
// cuboid with createCuboidSC3D, walks pt_sys to form the boundary groups and
// set every particle's material through dynamic_cast, loads the top face in
// displacement steps through DispBC over Particle* groups, solves with
// SolverStatic::solveProblem and reads the result back from the particle views
// and Stiffness's public CSR arrays.
//
// Usage: sc_cube_adapter <n_steps> <out_dir>; writes u.bin (N x 3 displacements
// from p->xyz - p->xyz_initial), pin.bin (p->Pin), Kp.bin, IK.bin, JK.bin,
// K.bin (solv.stiffness.*) and the dump file the solver appends per step.
#include "lpm.h"
#include "particle.h"
#include "assembly.h"
#include "utilities.h"
#include "stiffness.h"
#include "load_step.h"
#include "solver_static.h"

template <typename T>
static void dump(const std::string& path, const T* data, size_t n) {
  FILE* f = std::fopen(path.c_str(), "wb");
  std::fwrite(data, sizeof(T), n, f);
  std::fclose(f);
}

int main(int argc, char** argv) {
  if (argc < 3) return 2;
  const int n_steps = std::atoi(argv[1]);
  const std::string out = argv[2];

  const int n_layer = 2;
  double radius = 0.1;
  UnitCell cell(LatticeType::SimpleCubic3D, radius);
  int eulerflag = 0;
  double angles[] = {PI / 180.0 * 0.0, PI / 180.0 * 0.0, PI / 180.0 * 0.0};
  double* R_matrix = createRMatrix(eulerflag, angles);
  std::array<double, 2 * NDIM> box{0.0, 1.0, -0.1, 1.0, -0.1, 1.6};
  std::vector<std::array<double, NDIM>> sc_xyz = createCuboidSC3D(box, cell, R_matrix);
  Assembly<n_layer> pt_ass{sc_xyz, box, cell, ParticleType::Elastic};

  bool is_plane_stress = false;
  double E0 = 69e3, mu0 = 0.3, critical_bstrain = 1;
  double step_size = 0.002;
  int max_iter = 30, start_index = 0;
  double tol_iter = 1e-7;
  int undamaged_pt_type = -1;

  std::vector<Particle<n_layer>*> top_group, bottom_group;
  for (Particle<n_layer>* p1 : pt_ass.pt_sys) {
    if (p1->xyz[2] > box[5] - 2 * radius) {
      top_group.push_back(p1);
      p1->type = 1;
    }
    if (p1->xyz[2] < box[4] + 2 * radius) {
      bottom_group.push_back(p1);
      p1->type = 2;
    }
    ParticleElastic<n_layer>* elpt = dynamic_cast<ParticleElastic<n_layer>*>(p1);
    elpt->setParticleProperty(0.0, is_plane_stress, E0, mu0, critical_bstrain);
  }

  std::vector<LoadStep<n_layer>> load;
  for (int i = 0; i < n_steps; i++) {
    LoadStep<n_layer> step;
    step.dispBCs.push_back(DispBC<n_layer>(bottom_group, LoadMode::Relative, 'x', 0.0));
    step.dispBCs.push_back(DispBC<n_layer>(bottom_group, LoadMode::Relative, 'y', 0.0));
    step.dispBCs.push_back(DispBC<n_layer>(bottom_group, LoadMode::Relative, 'z', 0.0));
    step.dispBCs.push_back(DispBC<n_layer>(top_group, LoadMode::Relative, 'z', step_size));
    load.push_back(step);
  }

  pt_ass.updateGeometry();
  pt_ass.updateForceState();

  SolverStatic<n_layer> solv{undamaged_pt_type, pt_ass, StiffnessMode::Analytical, SolverMode::CG,
                             out + "/sc_position_3d.dump", max_iter, tol_iter};
  solv.solveProblem(load, start_index);

  const size_t n = pt_ass.pt_sys.size();
  std::vector<double> u(3 * n), pin(3 * n);
  for (Particle<n_layer>* p : pt_ass.pt_sys)
    for (int k = 0; k < 3; ++k) {
      u[3 * p->id + k] = p->xyz[k] - p->xyz_initial[k];
      pin[3 * p->id + k] = p->Pin[k];
    }
  dump(out + "/u.bin", u.data(), u.size());
  dump(out + "/pin.bin", pin.data(), pin.size());
  const int dim = pt_ass.pt_sys[0]->cell.dim;
  const long long nnz = solv.stiffness.K_pointer[n];
  dump(out + "/Kp.bin", solv.stiffness.K_pointer, n + 1);
  dump(out + "/IK.bin", solv.stiffness.IK, dim * n + 1);
  dump(out + "/JK.bin", solv.stiffness.JK, nnz);
  dump(out + "/K.bin", solv.stiffness.K_global, nnz);
  std::printf("%zu %lld %d\n", n, nnz, start_index);
  return 0;
}
