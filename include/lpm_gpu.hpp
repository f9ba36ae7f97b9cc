// lpm_gpu.hpp -- header-only C++17 adapter that keeps LPM-CPP's class and method
// names (UnitCell, Assembly, Stiffness, Solver, SolverStatic, SolverFatigue,
// LoadStep and the fatigue load steps, DispBC, ForceBC) and forwards every call
// to the C ABI of liblpmg.so (lpm_gpu.h).
//
// A caller of the reference switches by including this header instead of
// assembly.h / stiffness.h / solver_static.h and linking -llpmg.  Differences a
// caller sees:
//   * particle groups are index vectors (std::vector<int>), not Particle* lists;
//   * Stiffness's public IK/JK/K_pointer/K_global/residual are host copies made
//     by Stiffness::download() (the matrix itself stays on the GPU);
//   * errors throw lpm_gpu::Error instead of calling exit();
//   * SolverMode::CG and SolverMode::PARDISO both run the GPU (P)CG; the
//     StiffnessMode is always the analytic tangent (the finite-difference mode
//     of stiffness.h:144-195 is not on the GPU path);
//   * the solvers write no dump files (Assembly::writeDump); callers read the
//     fields they need with Assembly::field().
#pragma once

#include <array>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "lpm_gpu.h"

namespace lpm_gpu {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& what) : std::runtime_error(what), code(c) {}
};

inline int check(int rc, const char* where) {
  if (rc < 0) throw Error(rc, std::string(where) + ": " + lpmg_last_error());
  return rc;
}

// include/lpm.h:67-108
enum class LoadMode : char { Relative, Absolute };
enum class SolverMode : char { CG, PARDISO };
enum class StiffnessMode : char { Analytical, FiniteDifference };
enum class ParticleType : char { Elastic, ElasticDamage, J2Plasticity, FatigueHCF };
enum class TimeMapMode : char { Linear, Exponential };
// include/load_step_fatigue.h:13-19
enum class FatigueLoadType : char { LoadUniaxialDisp, LoadUniaxialForce, LoadDogBoneDisp, LoadCTForce };
enum class LatticeType : char { Square2D, Hexagon2D, SimpleCubic3D, FCC3D, BCC3D_1Slip, BCC3D_2Slip, BCC3D_3Slip };

inline int lattice_id(LatticeType t) {
  switch (t) {
    case LatticeType::Square2D: return LPMG_SQUARE2D;
    case LatticeType::Hexagon2D: return LPMG_HEXAGON2D;
    case LatticeType::SimpleCubic3D: return LPMG_SC3D;
    case LatticeType::FCC3D: return LPMG_FCC3D;
    default: return LPMG_BCC3D;
  }
}

// UnitCell (include/unit_cell.h:7-27)
class UnitCell {
 public:
  LatticeType lattice;
  int dim = 0, nneighbors = 0;
  double radius = 0, particle_volume = 0;
  std::array<double, 9> el_mapping{};
  std::vector<double> neighbor_cutoff;
  UnitCell(LatticeType p_lattice, double p_radius) : lattice(p_lattice), radius(p_radius) {
    double out[16];
    check(lpmg_unit_cell(lattice_id(lattice), radius, out), "UnitCell");
    dim = (int)out[0];
    nneighbors = (int)out[1];
    particle_volume = out[2];
    neighbor_cutoff = {out[3], out[4]};
    for (int k = 0; k < 9; ++k) el_mapping[k] = out[5 + k];
  }
};

// DispBC / ForceBC (include/boundary.h:9-44), groups as particle indices
template <int nlayer>
class DispBC {
 public:
  std::vector<int> group;
  LoadMode load_mode;
  char flag;
  double step;
  DispBC(std::vector<int> p_group, LoadMode m, char f, double s) : group(std::move(p_group)), load_mode(m), flag(f), step(s) {}
  void cutHalf() { step /= 2.0; }
};

template <int nlayer>
class ForceBC {
 public:
  std::vector<int> group;
  LoadMode load_mode;
  double fx, fy, fz;
  ForceBC(std::vector<int> p_group, LoadMode m, double x, double y, double z)
      : group(std::move(p_group)), load_mode(m), fx(x), fy(y), fz(z) {}
  void cutHalf() { fx /= 2.0, fy /= 2.0, fz /= 2.0; }
};

// LoadStep (include/load_step.h:9-28)
template <int nlayer>
class LoadStep {
 public:
  std::vector<DispBC<nlayer>> dispBCs;
  std::vector<ForceBC<nlayer>> forceBCs;
  void loadCutHalf() {
    for (auto& d : dispBCs) d.cutHalf();
    for (auto& f : forceBCs) f.cutHalf();
  }
};

// The fatigue load steps (include/load_step_fatigue.h:21-85)
template <int nlayer>
class LoadUniaxialForce : public LoadStep<nlayer> {
 public:
  LoadUniaxialForce(double fy, const std::vector<int>& left, const std::vector<int>& top, const std::vector<int>& bottom) {
    this->dispBCs.push_back(DispBC<nlayer>(left, LoadMode::Relative, 'x', 0.0));
    this->forceBCs.push_back(ForceBC<nlayer>(top, LoadMode::Relative, 0.0, fy, 0.0));
    this->dispBCs.push_back(DispBC<nlayer>(bottom, LoadMode::Relative, 'y', 0.0));
  }
};

template <int nlayer>
class LoadUniaxialDisp : public LoadStep<nlayer> {
 public:
  LoadUniaxialDisp(double dy, const std::vector<int>& left, const std::vector<int>& top, const std::vector<int>& bottom) {
    this->dispBCs.push_back(DispBC<nlayer>(left, LoadMode::Relative, 'x', 0.0));
    this->dispBCs.push_back(DispBC<nlayer>(top, LoadMode::Relative, 'y', dy));
    this->dispBCs.push_back(DispBC<nlayer>(bottom, LoadMode::Relative, 'y', 0.0));
  }
};

template <int nlayer>
class LoadDogBoneDisp : public LoadStep<nlayer> {
 public:
  LoadDogBoneDisp(double x, double y, double z, const std::vector<int>& group1, const std::vector<int>& group2) {
    this->dispBCs.push_back(DispBC<nlayer>(group1, LoadMode::Relative, 'x', x));
    this->dispBCs.push_back(DispBC<nlayer>(group1, LoadMode::Relative, 'y', y));
    this->dispBCs.push_back(DispBC<nlayer>(group1, LoadMode::Relative, 'z', z));
    this->dispBCs.push_back(DispBC<nlayer>(group2, LoadMode::Relative, 'x', 0.0));
    this->dispBCs.push_back(DispBC<nlayer>(group2, LoadMode::Relative, 'y', 0.0));
    this->dispBCs.push_back(DispBC<nlayer>(group2, LoadMode::Relative, 'z', 0.0));
  }
};

template <int nlayer>
class LoadCTForce : public LoadStep<nlayer> {
 public:
  LoadCTForce(double fy, const std::vector<int>& mid, const std::vector<int>& top, const std::vector<int>& bottom) {
    this->dispBCs.push_back(DispBC<nlayer>(mid, LoadMode::Relative, 'y', 0));
    this->dispBCs.push_back(DispBC<nlayer>(top, LoadMode::Relative, 'x', 0.0));
    this->dispBCs.push_back(DispBC<nlayer>(bottom, LoadMode::Relative, 'x', 0.0));
    this->forceBCs.push_back(ForceBC<nlayer>(top, LoadMode::Relative, 0.0, fy, 0.0));
    this->forceBCs.push_back(ForceBC<nlayer>(bottom, LoadMode::Relative, 0.0, -fy, 0.0));
  }
};

// Assembly (include/assembly.h:40-71): the particle system, resident on the GPU
template <int nlayer>
class Assembly {
  static_assert(nlayer == 2, "the reference supports two neighbour layers");

 public:
  ParticleType ptype;
  int nparticle = 0;
  std::array<double, 6> box{};
  lpmg_handle h = nullptr;

  Assembly(std::vector<std::array<double, 3>>& p_xyz, std::array<double, 6>& p_box, UnitCell& p_cell,
           const ParticleType& p_ptype)
      : ptype(p_ptype), box(p_box) {
    std::vector<double> xyz;
    xyz.reserve(3 * p_xyz.size());
    for (auto& p : p_xyz) xyz.insert(xyz.end(), p.begin(), p.end());
    check(lpmg_create(lattice_id(p_cell.lattice), p_cell.radius, (int)ptype, (int)p_xyz.size(), xyz.data(), &h),
          "Assembly");
    nparticle = (int)p_xyz.size();
  }
  // Assembly(dump, bond) with the bond list already parsed (assembly.h:83-95, 390-424)
  Assembly(std::vector<std::array<double, 3>>& p_xyz, const std::vector<int32_t>& p1, const std::vector<int32_t>& p2,
           const std::vector<int32_t>& layer, UnitCell& p_cell, const ParticleType& p_ptype)
      : ptype(p_ptype) {
    std::vector<double> xyz;
    for (auto& p : p_xyz) xyz.insert(xyz.end(), p.begin(), p.end());
    check(lpmg_create_with_bonds(lattice_id(p_cell.lattice), p_cell.radius, (int)ptype, (int)p_xyz.size(),
                                 xyz.data(), (int64_t)p1.size(), p1.data(), p2.data(), layer.data(), &h),
          "Assembly");
    nparticle = (int)p_xyz.size();
  }
  Assembly(const Assembly&) = delete;
  Assembly& operator=(const Assembly&) = delete;
  ~Assembly() { lpmg_destroy(h); }

  // updateGeometry and updateForceState are one fused device pass; the reference
  // always calls them as a pair (assembly.h:132-191, solver.h:58-59, 82-83).
  void updateGeometry() { pending_geometry_ = true; }
  void updateForceState() {
    check(lpmg_forward(h), "updateForceState");
    pending_geometry_ = false;
  }
  void flushGeometry() {
    if (pending_geometry_) updateForceState();
  }
  void updateStateVar() { check(lpmg_update_state_var(h), "updateStateVar"); }
  void storeStateVar() { check(lpmg_store_state(h), "storeStateVar"); }
  void resetStateVar(bool reset_xyz) { check(lpmg_reset_state(h, reset_xyz ? 1 : 0), "resetStateVar"); }
  void searchNonlocalNeighbors(double cutoff_ratio) { check(lpmg_search_nonlocal(h, cutoff_ratio), "searchNonlocalNeighbors"); }

  // setParticleProperty for all particles (or the listed ones)
  void setElastic(double nonlocal_L, bool plane_stress, double E, double mu, double crit, const std::vector<int>* idx = nullptr) {
    const double p[5] = {nonlocal_L, plane_stress ? 1.0 : 0.0, E, mu, crit};
    material(LPMG_MAT_ELASTIC, p, idx);
  }
  void setElasticCubic(double nonlocal_L, double C11, double C12, double C44, double crit, const std::vector<int>* idx = nullptr) {
    const double p[5] = {nonlocal_L, C11, C12, C44, crit};
    material(LPMG_MAT_ELASTIC_CUBIC, p, idx);
  }
  void setElasticDamage(double nonlocal_L, bool plane_stress, double E, double mu, double k0, double k1, double ct_ratio,
                        double d_thres, const std::vector<int>* idx = nullptr) {
    const double p[8] = {nonlocal_L, plane_stress ? 1.0 : 0.0, E, mu, k0, k1, ct_ratio, d_thres};
    material(LPMG_MAT_DAMAGE, p, idx);
  }
  void setJ2(double nonlocal_L, bool plane_stress, double E, double mu, double sigmay, double xi, double H, double A,
             double crit, const std::vector<int>* idx = nullptr) {
    const double p[9] = {nonlocal_L, plane_stress ? 1.0 : 0.0, E, mu, sigmay, xi, H, A, crit};
    material(LPMG_MAT_J2, p, idx);
  }
  // ParticleFatigueHCF::setParticleProperty (particle_fatigue_hcf.h:162-206)
  void setFatigueHCF(double nonlocal_L, bool plane_stress, double E, double mu, double A, double B, double d,
                     double d_thres, double f_lmt_ratio, const std::vector<int>* idx = nullptr) {
    const double p[9] = {nonlocal_L, plane_stress ? 1.0 : 0.0, E, mu, A, B, d, d_thres, f_lmt_ratio};
    material(LPMG_MAT_FATIGUE, p, idx);
  }
  void setTypes(const std::vector<int32_t>& type) { check(lpmg_set_types(h, type.data()), "setTypes"); }

  std::vector<double> field(int f) const {
    std::vector<double> out((size_t)lpmg_field_len(h, f));
    check(lpmg_get_field(h, f, out.data()), "field");
    return out;
  }
  std::vector<std::array<double, 3>> positions() const {
    auto v = field(LPMG_F_XYZ);
    std::vector<std::array<double, 3>> out(v.size() / 3);
    for (size_t i = 0; i < out.size(); ++i) out[i] = {v[3 * i], v[3 * i + 1], v[3 * i + 2]};
    return out;
  }

  // group id of a particle list (cached by content)
  int group(const std::vector<int>& idx) {
    for (size_t g = 0; g < groups_.size(); ++g)
      if (groups_[g] == idx) return (int)g;
    check(lpmg_define_group(h, (int)idx.size(), idx.data()), "group");
    groups_.push_back(idx);
    return (int)groups_.size() - 1;
  }

 private:
  bool pending_geometry_ = false;
  std::vector<std::vector<int>> groups_;
  void material(int model, const double* p, const std::vector<int>* idx) {
    check(lpmg_set_material(h, idx ? (int)idx->size() : 0, idx ? idx->data() : nullptr, model, p), "setParticleProperty");
  }
};

// Stiffness (include/stiffness.h:18-52): the matrix stays on the GPU; download()
// fills the reference's upper-triangular 1-based CSR arrays on the host.
template <int nlayer>
class Stiffness {
 public:
  std::vector<int64_t> IK, JK, K_pointer;
  std::vector<double> residual, K_global;
  explicit Stiffness(Assembly<nlayer>& a) : ass_(a) {}
  void reset() {}
  void calcStiffness() {
    ass_.flushGeometry();
    check(lpmg_assemble(ass_.h, precision), "calcStiffness");
  }
  void calcStiffness2D() { calcStiffness(); }
  void calcStiffness3D() { calcStiffness(); }
  void updateStiffnessDispBC() {}  // applied inside lpmg_assemble (stiffness.h:363-438)
  void download() {
    int64_t s[8];
    check(lpmg_sizes(ass_.h, s), "sizes");
    K_pointer.resize(s[0] + 1);
    IK.resize(s[1] * s[0] + 1);
    JK.resize(s[3]);
    K_global.resize(s[3]);
    check(lpmg_export_csr(ass_.h, K_pointer.data(), IK.data(), JK.data(), K_global.data()), "download");
    residual = ass_.field(LPMG_F_RESIDUAL);
  }
  int precision = 0;  // 1: FP32 copy for the mixed-precision inner solve

 private:
  Assembly<nlayer>& ass_;
};

// Solver (include/solver.h:16-52)
template <int nlayer>
class Solver {
 public:
  int max_NR_iter;
  double tol_NR_iter;
  SolverMode sol_mode;
  std::string dumpFile;
  Assembly<nlayer>& ass;
  Stiffness<nlayer> stiffness;
  // opts of lpmg_solve_linear: [precond, rtol, atol, maxit, precision, ...]
  std::array<double, 8> opts{1, 1e-12, 1e-12, 0, 0, 16, 0, 0};
  std::vector<double> last_stats = std::vector<double>(4, 0.0);

  Solver(Assembly<nlayer>& p_ass, const StiffnessMode&, const SolverMode& p_sol_mode, const std::string& p_dumpFile,
         const int& p_niter, const double& p_tol)
      : max_NR_iter(p_niter), tol_NR_iter(p_tol), sol_mode(p_sol_mode), dumpFile(p_dumpFile), ass(p_ass), stiffness(p_ass) {}

  void setMixedPrecision(bool on, double inner_rtol_sq = 1e-4) {
    opts[4] = on ? 1 : 0;
    stiffness.precision = on ? 1 : 0;
    if (on) opts[1] = inner_rtol_sq;
  }
  // Multigrid-preconditioned CG instead of Jacobi (large 2D systems)
  void setMultigrid(bool on) { opts[0] = on ? 2 : 1; }
  void updateDisplacementBC(LoadStep<nlayer>& ls) {
    for (auto& bc : ls.dispBCs)
      check(lpmg_disp_bc(ass.h, ass.group(bc.group), bc.load_mode == LoadMode::Relative ? LPMG_RELATIVE : LPMG_ABSOLUTE,
                         bc.flag, bc.step), "updateDisplacementBC");
  }
  void updateForceBC(LoadStep<nlayer>& ls) {
    for (auto& bc : ls.forceBCs)
      check(lpmg_force_bc(ass.h, ass.group(bc.group), bc.load_mode == LoadMode::Relative ? LPMG_RELATIVE : LPMG_ABSOLUTE,
                          bc.fx, bc.fy, bc.fz), "updateForceBC");
  }
  std::array<double, 2> updateRR() {
    std::array<double, 2> n{};
    check(lpmg_update_rr(ass.h, n.data()), "updateRR");
    return n;
  }
  void solveLinearSystem() {
    const int rc = lpmg_solve_linear(ass.h, opts.data(), last_stats.data());
    if (rc != -4) check(rc, "solveLinearSystem");  // a CG stopped at maxit keeps its update, as MKL's (solver.h:348-349)
  }
  void LPM_CG() { solveLinearSystem(); }
  void LPM_PARDISO() { solveLinearSystem(); }
  int NewtonIteration() {
    ass.flushGeometry();
    return check(lpmg_newton(ass.h, max_NR_iter, tol_NR_iter, opts.data(), last_stats.data()), "NewtonIteration");
  }
};

// SolverStatic (include/solver_static.h:20-152)
template <int nlayer>
class SolverStatic : public Solver<nlayer> {
 public:
  int undamaged_pt_type;
  SolverStatic(const int& p_undamaged, Assembly<nlayer>& p_ass, const StiffnessMode& sm, const SolverMode& so,
               const std::string& dump, const int& niter, const double& tol)
      : Solver<nlayer>(p_ass, sm, so, dump, niter, tol), undamaged_pt_type(p_undamaged) {}

  bool updateStaticDamage() { return check(lpmg_update_static_damage(this->ass.h, undamaged_pt_type), "updateStaticDamage") == 1; }

  bool solveProblemStep(LoadStep<nlayer>& load_step, double& /*dt*/) {
    this->updateForceBC(load_step);
    this->updateDisplacementBC(load_step);
    this->ass.flushGeometry();
    return check(lpmg_solve_step(this->ass.h, undamaged_pt_type, this->max_NR_iter, this->tol_NR_iter,
                                 this->opts.data(), this->last_stats.data()), "solveProblemStep") == 1;
  }

  // With opts[6] > 0 (the load-step predictor) each step's predictor scale
  // follows its size relative to the last converged step's across halvings.
  void solveProblem(std::vector<LoadStep<nlayer>>& load, int& start_index) {
    int n_step = (int)load.size();
    double dt = 1;
    const double predict = this->opts[6];
    std::vector<double> size(load.size(), 1.0);
    double last = 1.0;
    for (int i = 0; i < n_step; i++) {
      if (predict > 0) this->opts[6] = predict * size[i] / last;
      while (!solveProblemStep(load[i], dt)) {  // solver_static.h:78-87
        this->ass.resetStateVar(true);
        load[i].loadCutHalf();
        load.insert(load.begin() + i, load[i]);
        size[i] /= 2;
        size.insert(size.begin() + i, size[i]);
        ++n_step;
        if (predict > 0) this->opts[6] = predict * size[i] / last;
      }
      last = size[i];
    }
    this->opts[6] = predict;
    start_index += n_step;
  }
};

// SolverFatigue (include/solver_fatigue.h:18-303): cyclic loading of FatigueHCF
// particles; damage grows once per simulated cycle by the nonlocal damage rate
// times dN/dt, the real cycles the simulated one stands for.
template <int nlayer>
class SolverFatigue : public Solver<nlayer> {
 public:
  int undamaged_pt_type{0};
  TimeMapMode t_mode;
  double tau{0};
  std::vector<double> load_spectrum;

  SolverFatigue(const int& p_undamaged, Assembly<nlayer>& p_ass, const StiffnessMode& sm, const SolverMode& so,
                const TimeMapMode& p_t_mode, const double& p_tau, const std::string& dump, const int& niter,
                const double& tol)
      : Solver<nlayer>(p_ass, sm, so, dump, niter, tol), undamaged_pt_type(p_undamaged), t_mode(p_t_mode), tau(p_tau) {}

  // solver_fatigue.h:41-63 (throws instead of exit(1) for a missing file)
  void readLoad(const std::string& loadFile) {
    FILE* f = std::fopen(loadFile.c_str(), "r");
    if (!f) throw Error(-1, "'" + loadFile + "' does not exist");
    double v;
    while (std::fscanf(f, "%lf", &v) > 0) load_spectrum.push_back(v);
    std::fclose(f);
  }

  // solver_fatigue.h:65-73
  static int func_dNdt_linear(const double& p_tau) { return (int)(1 / p_tau); }
  static int func_dNdt_exponental(const double& p_tau, const double& t) { return (int)(1 / p_tau * std::exp(t / p_tau)); }

  bool updateFatigueDamage(double dNdt) {
    return check(lpmg_update_fatigue_damage(this->ass.h, undamaged_pt_type, dNdt), "updateFatigueDamage") == 1;
  }

  // solver_fatigue.h:128-180; pt_groups are the load step's particle lists
  std::vector<LoadStep<nlayer>> generateLoadCycle(FatigueLoadType ltype, int N, int n_interval,
                                                  const std::vector<std::vector<int>>& g) {
    const double f_0 = load_spectrum.at(2 * N), f_1 = load_spectrum.at(2 * N + 1), f_2 = load_spectrum.at(2 * N + 2);
    std::vector<LoadStep<nlayer>> cyc;
    for (const double inc : {(f_1 - f_0) / n_interval, (f_2 - f_1) / n_interval})
      for (int i = 0; i < n_interval; ++i) {
        if (ltype == FatigueLoadType::LoadUniaxialDisp) cyc.push_back(LoadUniaxialDisp<nlayer>(inc, g.at(0), g.at(1), g.at(2)));
        if (ltype == FatigueLoadType::LoadUniaxialForce) cyc.push_back(LoadUniaxialForce<nlayer>(inc, g.at(0), g.at(1), g.at(2)));
        if (ltype == FatigueLoadType::LoadDogBoneDisp) cyc.push_back(LoadDogBoneDisp<nlayer>(0, inc, 0, g.at(0), g.at(1)));
        if (ltype == FatigueLoadType::LoadCTForce) cyc.push_back(LoadCTForce<nlayer>(inc, g.at(0), g.at(1), g.at(2)));
      }
    return cyc;
  }

  // solver_fatigue.h:277-301
  bool solveProblemStep(LoadStep<nlayer>& load_step) {
    this->updateForceBC(load_step);
    this->updateDisplacementBC(load_step);
    this->stiffness.reset();
    this->stiffness.calcStiffness();
    return this->NewtonIteration() < this->max_NR_iter;
  }

  // solver_fatigue.h:220-241
  void solveProblemOneCycle(std::vector<LoadStep<nlayer>>& load, double dNdt) {
    for (auto& step : load) {
      solveProblemStep(step);
      this->ass.updateStateVar();
    }
    updateFatigueDamage(dNdt);
    this->ass.storeStateVar();
    this->ass.updateGeometry();
    this->ass.updateForceState();
  }

  // solver_fatigue.h:182-218; returns the real cycle count reached
  int solveProblemCyclic(FatigueLoadType loadType, int n_interval, const std::vector<std::vector<int>>& pt_groups) {
    const double n_cycles = 0.5 * ((double)load_spectrum.size() - 1);
    double t = 0, dNdt = 1;
    int N = 0;
    do {
      auto cyc = generateLoadCycle(loadType, N, n_interval, pt_groups);
      dNdt = (t_mode == TimeMapMode::Linear) ? func_dNdt_linear(tau) : func_dNdt_exponental(tau, t++);
      solveProblemOneCycle(cyc, dNdt);
      N += (int)dNdt;
    } while (N < n_cycles);
    return N;
  }

  // solver_fatigue.h:243-275
  void solveProblemStatic(std::vector<LoadStep<nlayer>>& load, int& start_index) {
    int n_step = (int)load.size();
    for (int i = 0; i < n_step; i++) {
      while (!solveProblemStep(load[i])) {
        this->ass.resetStateVar(true);
        load[i].loadCutHalf();
        load.insert(load.begin() + i, load[i]);
        ++n_step;
      }
      this->ass.updateStateVar();
      this->ass.storeStateVar();
    }
    start_index += n_step;
  }
};

// Many solves at once on one GPU (lpmg_server_*; no reference counterpart).
// Requests are lpmg_request structs; their output buffers must outlive wait().
class Server {
 public:
  explicit Server(int max_active = 16, int64_t mem_budget = 0, int64_t fuse_max_rows = 50000) {
    check(lpmg_server_create(max_active, mem_budget, fuse_max_rows, &s_), "lpmg_server_create");
  }
  ~Server() { lpmg_server_destroy(s_); }
  Server(const Server&) = delete;
  Server& operator=(const Server&) = delete;
  int64_t submit(const lpmg_request& r) {
    int64_t t = 0;
    check(lpmg_server_submit(s_, &r, &t), "lpmg_server_submit");
    return t;
  }
  // false on timeout; throws on a failed request; stats[LPMG_RS_N]
  bool wait(int64_t ticket, std::array<double, LPMG_RS_N>* stats = nullptr, double timeout_s = -1.0) {
    std::array<double, LPMG_RS_N> st{};
    const int rc = lpmg_server_wait(s_, ticket, timeout_s, st.data());
    if (rc == 1) return false;
    if (stats) *stats = st;
    check(rc, "request");
    return true;
  }
  lpmg_server handle() const { return s_; }

 private:
  lpmg_server s_ = nullptr;
};

}  // namespace lpm_gpu
