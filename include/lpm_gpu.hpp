// lpm_gpu.hpp -- header-only C++17 adapter that keeps LPM-CPP's classes, members
// and call sequence (UnitCell, Particle and its four material classes,
// Assembly::pt_sys, DispBC/ForceBC over Particle* groups, LoadStep and the
// fatigue load steps, Stiffness with its public CSR arrays, Solver,
// SolverStatic, SolverFatigue) and forwards the work to the C ABI of liblpmg.so
// (lpm_gpu.h).
//
// The particle system lives on the GPU.  pt_sys holds host views of it: every
// Particle carries the reference's public fields (id, type, nb, nconn,
// nconn_largeq, xyz, xyz_initial, xyz_last, Pin, Pex, damage, damage_visual,
// damage_last, Ddot_local, Ddot_nonlocal, state_var, stress, dLe_total,
// TdLe_total, cs_sum*, disp_constraint, nonlocal_L, neighbors, conns,
// neighbors_nonlocal, cell).  The views are refreshed from the device after
// every call that changes the system (forward pass, solve step, state update,
// restore), and what a caller writes into them -- type, xyz (moveTo/moveBy),
// Pex, damage, disp_constraint -- is sent to the device before the next such
// call; setParticleProperty calls are collected and sent grouped by their
// parameters.  Stiffness::IK/JK/K_pointer/K_global/residual are filled from
// the device's assembly (the reference's 1-based upper-triangular CSR, int64
// indices) after every assembly the caller can observe.
//
// A reference driver switches by putting include/lpmcpp (reference-named
// headers that include this one) on the include path and linking -llpmg;
// INTEGRATION.md lists what differs: the Bond objects of bond_layers are not
// mirrored, the finite-difference StiffnessMode and SolverMode::PARDISO run the
// GPU's analytic tangent and PCG, and errors throw lpm_gpu::Error instead of
// calling exit().
#pragma once

#include <algorithm>
#include <array>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
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

using MKL_INT = long long;  // the reference's MKL_INT under ILP64

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

// UnitCell (include/unit_cell.h:7-122): the values come from the library,
// the shell and AFEM counts are the reference's per-lattice data
class UnitCell {
 public:
  LatticeType lattice;
  int dim = 0;
  int nneighbors = 0, nneighbors1 = 0, nneighbors2 = 0;
  int nneighbors_AFEM = 0, nneighbors_AFEM1 = 0, nneighbors_AFEM2 = 0;
  double radius = 0;
  double particle_volume = 0;
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
    static const int shells[5][5] = {{4, 4, 16, 12, 12},    // setSquare
                                     {6, 6, 30, 18, 18},    // setHexagon
                                     {6, 12, 60, 24, 54},   // setSimpleCubic
                                     {12, 6, 60, 54, 24},   // setFCC
                                     {8, 6, 40, 34, 24}};   // setBCC
    const int* s = shells[lattice_id(lattice)];
    nneighbors1 = s[0], nneighbors2 = s[1], nneighbors_AFEM = s[2], nneighbors_AFEM1 = s[3], nneighbors_AFEM2 = s[4];
  }
};

template <int nlayer>
class Assembly;

// Particle (include/particle.h:13-69): a host view of one particle of an Assembly
template <int nlayer>
class Particle {
 public:
  int id{0};
  int type{0};
  int nconn_largeq{0};
  int nb{0}, nconn{0};
  double nonlocal_L{0.};
  double damage_last{0.}, damage{0.}, damage_visual{0.};
  double Ddot_local{0}, Ddot_nonlocal{0};
  UnitCell cell;
  std::vector<double> state_var, state_var_last;
  std::array<int, 3> disp_constraint{0, 0, 0};
  std::array<double, nlayer> TdLe_total{}, dLe_total{}, cs_sumx{}, cs_sumy{}, cs_sumz{};
  std::array<double, 3> xyz{}, xyz_initial{}, xyz_last{};
  std::array<double, 3> Pin{}, Pex{};
  std::vector<Particle<nlayer>*> neighbors;
  std::vector<Particle<nlayer>*> conns;
  std::vector<Particle<nlayer>*> neighbors_nonlocal;
  std::vector<double> stress, strain;

  Particle(double p_x, double p_y, double p_z, const UnitCell& p_cell, int p_type) : type(p_type), cell(p_cell) {
    xyz = {p_x, p_y, p_z};
    xyz_initial = xyz;
    xyz_last = xyz;
    stress.assign(6, 0.0);
    strain.assign(6, 0.0);
  }
  virtual ~Particle() = default;

  // moved on the host; the device follows at the next call that uses positions
  void moveTo(const double& x, const double& y, const double& z) { xyz = {x, y, z}; }
  void moveTo(const std::array<double, 3>& p) { xyz = p; }
  void moveBy(const std::array<double, 3>& d) {
    for (int k = 0; k < 3; ++k) xyz[k] += d[k];
  }
  // Particle::distanceTo (particle.h:276-280)
  double distanceTo(Particle<nlayer>* A) {
    return std::sqrt(std::pow((xyz[0] - A->xyz[0]), 2) + std::pow((xyz[1] - A->xyz[1]), 2) +
                     std::pow((xyz[2] - A->xyz[2]), 2));
  }
  bool operator==(const Particle<nlayer>& other) { return id == other.id; }

  // no reference counterpart: the system this view belongs to and its row in it
  Assembly<nlayer>* owner() const { return ass_; }
  int index() const { return idx_; }

 protected:
  void queue(int model, std::initializer_list<double> p);
  Assembly<nlayer>* ass_ = nullptr;
  int idx_ = 0;
  friend class Assembly<nlayer>;
};

// ParticleElastic::setParticleProperty (particle_elastic.h:70-140)
template <int nlayer>
class ParticleElastic : public Particle<nlayer> {
 public:
  double critical_bstrain{0};
  ParticleElastic(double x, double y, double z, const UnitCell& c, int t) : Particle<nlayer>(x, y, z, c, t) {
    this->state_var = {0};
    this->state_var_last = {0};
  }
  void setParticleProperty(double p_nonlocalL, bool is_plane_stress, double p_E, double p_mu, double p_critical_bstrain) {
    critical_bstrain = p_critical_bstrain;
    this->nonlocal_L = p_nonlocalL;
    this->queue(LPMG_MAT_ELASTIC, {p_nonlocalL, is_plane_stress ? 1.0 : 0.0, p_E, p_mu, p_critical_bstrain});
  }
  void setParticleProperty(double p_nonlocalL, double p_C11, double p_C12, double p_C44, double p_critical_bstrain) {
    critical_bstrain = p_critical_bstrain;
    this->nonlocal_L = p_nonlocalL;
    this->queue(LPMG_MAT_ELASTIC_CUBIC, {p_nonlocalL, p_C11, p_C12, p_C44, p_critical_bstrain});
  }
};

// ParticleElasticDamage::setParticleProperty (particle_elastic_damage.h:175-250)
template <int nlayer>
class ParticleElasticDamage : public Particle<nlayer> {
 public:
  double k0{0}, k1{0};
  double damage_threshold{0}, comp_tensile_ratio{0};
  ParticleElasticDamage(double x, double y, double z, const UnitCell& c, int t) : Particle<nlayer>(x, y, z, c, t) {
    this->state_var = std::vector<double>(1, 0.);
    this->state_var_last = std::vector<double>(1, 0.);
  }
  void setParticleProperty(double p_nonlocalL, bool is_plane_stress, double p_E, double p_mu, double p_k0, double p_k1,
                           double p_ct_ratio, double p_d_thres) {
    set_(p_k0, p_k1, p_ct_ratio, p_d_thres, p_nonlocalL);
    this->queue(LPMG_MAT_DAMAGE, {p_nonlocalL, is_plane_stress ? 1.0 : 0.0, p_E, p_mu, p_k0, p_k1, p_ct_ratio, p_d_thres});
  }
  void setParticleProperty(double p_nonlocalL, double p_C11, double p_C12, double p_C44, double p_k0, double p_k1,
                           double p_ct_ratio, double p_d_thres) {
    set_(p_k0, p_k1, p_ct_ratio, p_d_thres, p_nonlocalL);
    this->queue(LPMG_MAT_DAMAGE_CUBIC, {p_nonlocalL, p_C11, p_C12, p_C44, p_k0, p_k1, p_ct_ratio, p_d_thres});
  }

 private:
  void set_(double a, double b, double ct, double thr, double L) {
    k0 = a, k1 = b, comp_tensile_ratio = ct, damage_threshold = thr;
    this->nonlocal_L = L;
  }
};

// ParticleJ2Plasticity::setParticleProperty (particle_j2plasticity.h:154-198)
template <int nlayer>
class ParticleJ2Plasticity : public Particle<nlayer> {
 public:
  double E{0}, mu{0}, sigmay{0}, xi{0}, H{0}, A{0};
  double damage_threshold{0}, critical_bstrain{0};
  ParticleJ2Plasticity(double x, double y, double z, const UnitCell& c, int t) : Particle<nlayer>(x, y, z, c, t) {
    this->state_var = std::vector<double>(7, 0.);
    this->state_var_last = std::vector<double>(7, 0.);
  }
  void setParticleProperty(double p_nonlocalL, bool is_plane_stress, double p_E, double p_mu, double p_sigmay,
                           double p_xi, double p_H, double p_A, double p_critical_bstrain) {
    E = p_E, mu = p_mu, sigmay = p_sigmay, xi = p_xi, H = p_H, A = p_A, critical_bstrain = p_critical_bstrain;
    this->nonlocal_L = p_nonlocalL;
    this->queue(LPMG_MAT_J2, {p_nonlocalL, is_plane_stress ? 1.0 : 0.0, p_E, p_mu, p_sigmay, p_xi, p_H, p_A,
                              p_critical_bstrain});
  }
};

// ParticleFatigueHCF::setParticleProperty (particle_fatigue_hcf.h:162-206)
template <int nlayer>
class ParticleFatigueHCF : public Particle<nlayer> {
 public:
  double E{0}, mu{0}, fatigue_limit_ratio{1.108};
  double fatigue_A{0}, fatigue_B{0}, fatigue_d{0}, damage_threshold{0};
  ParticleFatigueHCF(double x, double y, double z, const UnitCell& c, int t) : Particle<nlayer>(x, y, z, c, t) {
    this->state_var = {0, INT_MAX, 0};
    this->state_var_last = {0, INT_MAX, 0};
  }
  void setParticleProperty(double p_nonlocalL, bool is_plane_stress, double p_E, double p_mu, double p_A, double p_B,
                           double p_d, double p_damage_threshold, double p_fatigue_limit_ratio) {
    E = p_E, mu = p_mu, fatigue_A = p_A, fatigue_B = p_B, fatigue_d = p_d;
    damage_threshold = p_damage_threshold, fatigue_limit_ratio = p_fatigue_limit_ratio;
    this->nonlocal_L = p_nonlocalL;
    this->queue(LPMG_MAT_FATIGUE, {p_nonlocalL, is_plane_stress ? 1.0 : 0.0, p_E, p_mu, p_A, p_B, p_d,
                                   p_damage_threshold, p_fatigue_limit_ratio});
  }
};

// DispBC / ForceBC (include/boundary.h:9-44)
template <int nlayer>
class DispBC {
 public:
  std::vector<Particle<nlayer>*> group;
  LoadMode load_mode;
  char flag;
  double step;
  DispBC(std::vector<Particle<nlayer>*> p_group, LoadMode p_load_mode, char p_flag, double p_step)
      : group{p_group}, load_mode{p_load_mode}, flag{p_flag}, step{p_step} {}
  void cutHalf() { step /= 2.0; }
};

template <int nlayer>
class ForceBC {
 public:
  std::vector<Particle<nlayer>*> group;
  LoadMode load_mode;
  double fx, fy, fz;
  ForceBC(std::vector<Particle<nlayer>*> p_group, LoadMode p_load_mode, double p_fx, double p_fy, double p_fz)
      : group{p_group}, load_mode{p_load_mode}, fx{p_fx}, fy{p_fy}, fz{p_fz} {}
  void cutHalf() { fx /= 2.0, fy /= 2.0, fz /= 2.0; }
};

// LoadStep (include/load_step.h:9-28)
template <int nlayer>
class LoadStep {
 public:
  std::vector<DispBC<nlayer>> dispBCs;
  std::vector<ForceBC<nlayer>> forceBCs;
  LoadStep() {}
  void loadCutHalf() {
    for (DispBC<nlayer>& d : dispBCs) d.cutHalf();
    for (ForceBC<nlayer>& f : forceBCs) f.cutHalf();
  }
};

// The fatigue load steps (include/load_step_fatigue.h:21-85)
template <int nlayer>
class LoadUniaxialForce : public LoadStep<nlayer> {
 public:
  LoadUniaxialForce(double fy, std::vector<Particle<nlayer>*>& left, std::vector<Particle<nlayer>*>& top,
                    std::vector<Particle<nlayer>*>& bottom) {
    this->dispBCs.push_back(DispBC<nlayer>(left, LoadMode::Relative, 'x', 0.0));
    this->forceBCs.push_back(ForceBC<nlayer>(top, LoadMode::Relative, 0.0, fy, 0.0));
    this->dispBCs.push_back(DispBC<nlayer>(bottom, LoadMode::Relative, 'y', 0.0));
  }
};

template <int nlayer>
class LoadUniaxialDisp : public LoadStep<nlayer> {
 public:
  LoadUniaxialDisp(double dy, std::vector<Particle<nlayer>*>& left, std::vector<Particle<nlayer>*>& top,
                   std::vector<Particle<nlayer>*>& bottom) {
    this->dispBCs.push_back(DispBC<nlayer>(left, LoadMode::Relative, 'x', 0.0));
    this->dispBCs.push_back(DispBC<nlayer>(top, LoadMode::Relative, 'y', dy));
    this->dispBCs.push_back(DispBC<nlayer>(bottom, LoadMode::Relative, 'y', 0.0));
  }
};

template <int nlayer>
class LoadDogBoneDisp : public LoadStep<nlayer> {
 public:
  LoadDogBoneDisp(double x, double y, double z, std::vector<Particle<nlayer>*>& group1,
                  std::vector<Particle<nlayer>*>& group2) {
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
  LoadCTForce(double fy, std::vector<Particle<nlayer>*>& mid, std::vector<Particle<nlayer>*>& top,
              std::vector<Particle<nlayer>*>& bottom) {
    this->dispBCs.push_back(DispBC<nlayer>(mid, LoadMode::Relative, 'y', 0));
    this->dispBCs.push_back(DispBC<nlayer>(top, LoadMode::Relative, 'x', 0.0));
    this->dispBCs.push_back(DispBC<nlayer>(bottom, LoadMode::Relative, 'x', 0.0));
    this->forceBCs.push_back(ForceBC<nlayer>(top, LoadMode::Relative, 0.0, fy, 0.0));
    this->forceBCs.push_back(ForceBC<nlayer>(bottom, LoadMode::Relative, 0.0, -fy, 0.0));
  }
};

// Assembly (include/assembly.h:40-71): the particle system, resident on the GPU,
// with pt_sys as its host views
template <int nlayer>
class Assembly {
  static_assert(nlayer == 2, "the reference supports two neighbour layers");

 public:
  ParticleType ptype;
  int nparticle = 0;
  std::vector<Particle<nlayer>*> pt_sys;
  std::array<double, 6> box{};

  // Assembly(p_xyz, box, cell, ptype) (assembly.h:97-106): createParticles,
  // createBonds, updateConnections
  Assembly(std::vector<std::array<double, 3>>& p_xyz, std::array<double, 6>& p_box, UnitCell& p_cell,
           const ParticleType& p_ptype)
      : ptype(p_ptype), box(p_box) {
    std::vector<int> type(p_xyz.size(), 0);
    create(p_xyz, type, nullptr, p_cell);
  }
  // Assembly(dumpFile, cell, ptype) (assembly.h:73-80): readDump, createBonds, updateConnections
  Assembly(const std::string& dumpFile, UnitCell& p_cell, const ParticleType& p_ptype) : ptype(p_ptype) {
    std::vector<std::array<double, 3>> xyz;
    std::vector<int> type;
    readDump(dumpFile, xyz, type);
    create(xyz, type, nullptr, p_cell);
  }
  // Assembly(dumpFile, bondFile, cell, ptype) (assembly.h:82-95): readDump, readBond, updateConnections
  Assembly(const std::string& dumpFile, const std::string& bondFile, UnitCell& p_cell, const ParticleType& p_ptype)
      : ptype(p_ptype) {
    std::vector<std::array<double, 3>> xyz;
    std::vector<int> type;
    readDump(dumpFile, xyz, type);
    std::array<std::vector<int32_t>, 3> bonds;
    readBond(bondFile, bonds);
    create(xyz, type, &bonds, p_cell);
  }
  Assembly(const Assembly&) = delete;
  Assembly& operator=(const Assembly&) = delete;
  ~Assembly() {
    for (Particle<nlayer>* p : pt_sys) delete p;
    lpmg_destroy(h);
  }

  // updateGeometry and updateForceState are one fused device pass (geometry
  // from the current bond damage, then forces with the damage models' bond
  // damage update); the reference always calls them as a pair
  // (assembly.h:132-191, solver.h:58-59, 82-83).  The pass runs exactly once
  // per pair: a second one would see the raised bond damage and treat bonds
  // of fully damaged particles as broken, which the reference does only at
  // its next updateGeometry.
  void updateGeometry() { pending_forward_ = true; }
  void updateForceState() {
    pending_forward_ = false;
    run([&] { return lpmg_forward(h); }, "updateForceState");
  }
  void updateStateVar() { run([&] { return lpmg_update_state_var(h); }, "updateStateVar"); }
  void storeStateVar() { run([&] { return lpmg_store_state(h); }, "storeStateVar"); }
  void resetStateVar(bool reset_xyz) { run([&] { return lpmg_reset_state(h, reset_xyz ? 1 : 0); }, "resetStateVar"); }
  void searchNonlocalNeighbors(double cutoff_ratio) {
    run([&] { return lpmg_search_nonlocal(h, cutoff_ratio); }, "searchNonlocalNeighbors");
    std::vector<int32_t> ptr(nparticle + 1);
    check(lpmg_export_nonlocal(h, ptr.data(), nullptr), "searchNonlocalNeighbors");
    std::vector<int32_t> idx(std::max<int32_t>(ptr[nparticle], 1));
    check(lpmg_export_nonlocal(h, nullptr, idx.data()), "searchNonlocalNeighbors");
    for (int i = 0; i < nparticle; ++i) {
      auto& v = pt_sys[i]->neighbors_nonlocal;
      v.clear();
      for (int t = ptr[i]; t < ptr[i + 1]; ++t) v.push_back(pt_sys[idx[t]]);
    }
  }

  std::map<int, Particle<nlayer>*> toMap() {
    std::map<int, Particle<nlayer>*> map;
    for (Particle<nlayer>* pt : pt_sys) map[pt->id] = pt;
    return map;
  }

  // Assembly::writeDump (assembly.h:307-331): appended, 5 significant digits
  void writeDump(const std::string& dumpFile, int step) {
    FILE* fpt = std::fopen(dumpFile.c_str(), "a+");
    if (!fpt) throw Error(-1, "cannot open '" + dumpFile + "'");
    writeHeader(fpt, step);
    std::fprintf(fpt, "ITEM: ATOMS id type x y z dx dy dz s11 s22 s33 s23 s13 s12 damage damage_visual state_var[0] \n");
    for (auto pt : pt_sys)
      std::fprintf(fpt, "%d %d %.4e %.4e %.4e %.4e %.4e %.4e %.4e %.4e %.4e %.4e %.4e %.4e %.4e %.4e %.4e\n", pt->id,
                   pt->type, pt->xyz[0], pt->xyz[1], pt->xyz[2], pt->xyz[0] - pt->xyz_initial[0],
                   pt->xyz[1] - pt->xyz_initial[1], pt->xyz[2] - pt->xyz_initial[2], pt->stress[0], pt->stress[1],
                   pt->stress[2], pt->stress[3], pt->stress[4], pt->stress[5], pt->damage, pt->damage_visual,
                   pt->state_var.empty() ? 0.0 : pt->state_var[0]);
    std::fclose(fpt);
  }
  // Assembly::writeConfigurationDump (assembly.h:283-305)
  void writeConfigurationDump(const std::string& dumpFile) {
    FILE* fpt = std::fopen(dumpFile.c_str(), "w+");
    if (!fpt) throw Error(-1, "cannot open '" + dumpFile + "'");
    writeHeader(fpt, 0);
    std::fprintf(fpt, "ITEM: ATOMS id type x y z\n");
    for (auto pt : pt_sys) std::fprintf(fpt, "%d %d %.4e %.4e %.4e \n", pt->id, pt->type, pt->xyz[0], pt->xyz[1], pt->xyz[2]);
    std::fclose(fpt);
  }
  // Assembly::writeBond (assembly.h:268-281): "bond-id layer p1 p2", bonds
  // numbered in the order they are written
  void writeBond(const std::string& bondFile) {
    FILE* fpt = std::fopen(bondFile.c_str(), "w+");
    if (!fpt) throw Error(-1, "cannot open '" + bondFile + "'");
    std::vector<int32_t> bptr(nparticle + 1), nb0(nparticle), p2(nbond_);
    check(lpmg_export_bonds(h, bptr.data(), nb0.data(), p2.data()), "writeBond");
    int id = 0;
    for (int i = 0; i < nparticle; ++i)
      for (int b = bptr[i]; b < bptr[i + 1]; ++b)
        std::fprintf(fpt, "%d %d %d %d\n", id++, b < bptr[i] + nb0[i] ? 0 : 1, pt_sys[i]->id, pt_sys[p2[b]]->id);
    std::fclose(fpt);
  }

  // ---- no reference counterpart
  lpmg_handle h = nullptr;

  // Host edits of the views to the device: types, positions, external
  // forces, damage and constraints that differ from what the device last
  // reported, and the collected setParticleProperty calls.
  void push() {
    flush_materials();
    std::vector<int32_t> t(nparticle);
    for (int i = 0; i < nparticle; ++i) t[i] = pt_sys[i]->type;
    if (t != sent_type_) {
      check(lpmg_set_types(h, t.data()), "types");
      sent_type_ = t;
      ++sends_[0];
    }
    sends_[1] += send3(LPMG_F_XYZ, &Particle<nlayer>::xyz, sent_xyz_);
    sends_[2] += send3(LPMG_F_PEX, &Particle<nlayer>::Pex, sent_pex_);
    std::vector<double> d(nparticle);
    for (int i = 0; i < nparticle; ++i) d[i] = pt_sys[i]->damage;
    if (!same(d, sent_damage_)) {
      check(lpmg_set_field(h, LPMG_F_DAMAGE, d.data()), "damage");
      sent_damage_ = d;
      ++sends_[3];
    }
    std::vector<double> c(3 * (size_t)nparticle);
    for (int i = 0; i < nparticle; ++i)
      for (int k = 0; k < 3; ++k) c[3 * i + k] = pt_sys[i]->disp_constraint[k];
    if (!same(c, sent_cons_)) {
      check(lpmg_set_field(h, LPMG_F_CONSTRAINT, c.data()), "disp_constraint");
      sent_cons_ = c;
      ++sends_[4];
    }
    if (pending_forward_) {
      check(lpmg_forward(h), "updateGeometry/updateForceState");
      pending_forward_ = false;
    }
  }

  // The device's state into the views.
  void pull() {
    const size_t n = nparticle;
    auto get = [&](int f) {
      std::vector<double> v((size_t)lpmg_field_len(h, f));
      check(lpmg_get_field(h, f, v.data()), "field");
      return v;
    };
    auto xyz = get(LPMG_F_XYZ), xyz0 = get(LPMG_F_XYZ_INITIAL), xyzl = get(LPMG_F_XYZ_LAST), pin = get(LPMG_F_PIN),
         pex = get(LPMG_F_PEX), st = get(LPMG_F_STRESS), dmg = get(LPMG_F_DAMAGE), dv = get(LPMG_F_DAMAGE_VISUAL),
         dl = get(LPMG_F_DAMAGE_LAST), sv = get(LPMG_F_STATE_VAR), ddl = get(LPMG_F_DDOT_LOCAL),
         ddn = get(LPMG_F_DDOT_NONLOCAL), dle = get(LPMG_F_DLE_TOTAL), tdle = get(LPMG_F_TDLE_TOTAL),
         css = get(LPMG_F_CS_SUM), cons = get(LPMG_F_CONSTRAINT);
    for (size_t i = 0; i < n; ++i) {
      Particle<nlayer>* p = pt_sys[i];
      for (int k = 0; k < 3; ++k) {
        p->xyz[k] = xyz[3 * i + k];
        p->xyz_initial[k] = xyz0[3 * i + k];
        p->xyz_last[k] = xyzl[3 * i + k];
        p->Pin[k] = pin[3 * i + k];
        p->Pex[k] = pex[3 * i + k];
        p->disp_constraint[k] = (int)cons[3 * i + k];
      }
      for (int k = 0; k < 6; ++k) p->stress[k] = st[6 * i + k];
      p->damage = dmg[i];
      p->damage_visual = dv[i];
      p->damage_last = dl[i];
      p->Ddot_local = ddl[i];
      p->Ddot_nonlocal = ddn[i];
      for (size_t k = 0; k < p->state_var.size() && k < 7; ++k) p->state_var[k] = sv[7 * i + k];
      for (int L = 0; L < nlayer; ++L) {
        p->dLe_total[L] = dle[2 * i + L];
        p->TdLe_total[L] = tdle[2 * i + L];
        p->cs_sumx[L] = css[(2 * i + L) * 3];
        p->cs_sumy[L] = css[(2 * i + L) * 3 + 1];
        p->cs_sumz[L] = css[(2 * i + L) * 3 + 2];
      }
    }
    sent_xyz_ = xyz;
    sent_pex_ = pex;
    sent_damage_ = dmg;
    sent_cons_ = cons;
  }

  // a device call that changes the system, with the views kept in step:
  // host edits go first, the device's state comes back after
  template <class F>
  int run(F&& f, const char* where) {
    push();
    const int rc = check(f(), where);
    pull();
    return rc;
  }

  // group id of a particle list (cached by content)
  int group(const std::vector<Particle<nlayer>*>& g) {
    std::vector<int> idx(g.size());
    for (size_t k = 0; k < g.size(); ++k) idx[k] = g[k]->index();
    auto it = groups_.find(idx);
    if (it != groups_.end()) return it->second;
    const int id = check(lpmg_define_group(h, (int)idx.size(), idx.data()), "group");
    groups_[idx] = id;
    return id;
  }

  void queue_material(int idx, int model, std::vector<double> params) {
    params.insert(params.begin(), (double)model);
    pending_[params].push_back(idx);
  }

  // no reference counterpart: how often push() sent types, positions,
  // external forces, damage and constraints to the device
  const std::array<int64_t, 5>& sends() const { return sends_; }

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

 private:
  bool pending_forward_ = false;
  int64_t nbond_ = 0;
  std::map<std::vector<int>, int> groups_;
  std::map<std::vector<double>, std::vector<int32_t>> pending_;
  std::vector<int32_t> sent_type_;
  std::vector<double> sent_xyz_, sent_pex_, sent_damage_, sent_cons_;

  void flush_materials() {
    for (auto& e : pending_) {
      const int model = (int)e.first[0];
      check(lpmg_set_material(h, (int)e.second.size(), e.second.data(), model, e.first.data() + 1),
            "setParticleProperty");
    }
    pending_.clear();
  }

  // bitwise equality (a NaN the device reported is not a host edit)
  static bool same(const std::vector<double>& a, const std::vector<double>& b) {
    return a.size() == b.size() && (a.empty() || std::memcmp(a.data(), b.data(), a.size() * sizeof(double)) == 0);
  }
  int send3(int f, std::array<double, 3> Particle<nlayer>::*m, std::vector<double>& sent) {
    std::vector<double> v(3 * (size_t)nparticle);
    for (int i = 0; i < nparticle; ++i)
      for (int k = 0; k < 3; ++k) v[3 * i + k] = (pt_sys[i]->*m)[k];
    if (same(v, sent)) return 0;
    check(lpmg_set_field(h, f, v.data()), "field");
    sent = v;
    return 1;
  }
  std::array<int64_t, 5> sends_{};

  void writeHeader(FILE* fpt, int step) {
    std::fprintf(fpt, "ITEM: TIMESTEP\n%d\nITEM: NUMBER OF ATOMS\n%d\nITEM: BOX BOUNDS pp pp pp\n", step, nparticle);
    std::fprintf(fpt, "%8.8f %8.8f\n", box[0], box[1]);
    std::fprintf(fpt, "%8.8f %8.8f\n", box[2], box[3]);
    std::fprintf(fpt, "%8.8f %8.8f\n", box[4], box[5]);
  }

  // Assembly::readDump (assembly.h:333-388): three lines skipped, the particle
  // count, one line, the box (three lines), one line, then "id type x y z" per
  // particle until the numbers end
  void readDump(const std::string& dumpFile, std::vector<std::array<double, 3>>& xyz, std::vector<int>& type) {
    FILE* fpt = std::fopen(dumpFile.c_str(), "r");
    if (!fpt) throw Error(-1, "'" + dumpFile + "' does not exist!");
    char line[500];
    for (int n = 0; n < 3; ++n)
      if (!std::fgets(line, sizeof line, fpt)) break;
    int np = 0;
    bool ok = std::fscanf(fpt, "%d\n", &np) == 1 && std::fgets(line, sizeof line, fpt);
    for (int k = 0; k < 3 && ok; ++k) ok = std::fscanf(fpt, "%lf %lf\n", &box[2 * k], &box[2 * k + 1]) == 2;
    ok = ok && std::fgets(line, sizeof line, fpt);
    if (!ok) {
      std::fclose(fpt);
      throw Error(-1, "'" + dumpFile + "' is not a dump file");
    }
    int id = 0, t = 0;
    double x[3];
    while (std::fscanf(fpt, "%d %d %lf %lf %lf", &id, &t, &x[0], &x[1], &x[2]) > 0) {
      xyz.push_back({x[0], x[1], x[2]});
      type.push_back(t);
      file_ids_.push_back(id);
    }
    std::fclose(fpt);
  }

  // Assembly::readBond (assembly.h:390-424): "bond-id layer p1-id p2-id" per bond
  void readBond(const std::string& bondFile, std::array<std::vector<int32_t>, 3>& out) {
    FILE* fpt = std::fopen(bondFile.c_str(), "r");
    if (!fpt) throw Error(-1, "'" + bondFile + "' does not exist!");
    std::map<int, int> row;
    for (size_t i = 0; i < file_ids_.size(); ++i) row[file_ids_[i]] = (int)i;
    int id, layer, a, b;
    while (std::fscanf(fpt, "%d %d %d %d", &id, &layer, &a, &b) == 4) {
      auto ia = row.find(a), ib = row.find(b);
      if (ia == row.end() || ib == row.end()) {
        std::fclose(fpt);
        throw Error(-1, "bond " + std::to_string(id) + " names a particle missing from the dump");
      }
      out[0].push_back(ia->second);
      out[1].push_back(ib->second);
      out[2].push_back(layer);
    }
    std::fclose(fpt);
  }
  std::vector<int> file_ids_;

  void create(const std::vector<std::array<double, 3>>& p_xyz, const std::vector<int>& type,
              const std::array<std::vector<int32_t>, 3>* bonds, const UnitCell& cell) {
    const int n = (int)p_xyz.size();
    std::vector<double> flat;
    flat.reserve(3 * (size_t)n);
    for (auto& p : p_xyz) flat.insert(flat.end(), p.begin(), p.end());
    if (bonds)
      check(lpmg_create_with_bonds(lattice_id(cell.lattice), cell.radius, (int)ptype, n, flat.data(),
                                   (int64_t)(*bonds)[0].size(), (*bonds)[0].data(), (*bonds)[1].data(),
                                   (*bonds)[2].data(), &h),
            "Assembly");
    else
      check(lpmg_create(lattice_id(cell.lattice), cell.radius, (int)ptype, n, flat.data(), &h), "Assembly");
    nparticle = n;
    int64_t s[8];
    check(lpmg_sizes(h, s), "Assembly");
    nbond_ = s[2];
    pt_sys.reserve(n);
    for (int i = 0; i < n; ++i) {
      const auto& p = p_xyz[i];
      Particle<nlayer>* pt = nullptr;
      if (ptype == ParticleType::Elastic) pt = new ParticleElastic<nlayer>(p[0], p[1], p[2], cell, type[i]);
      if (ptype == ParticleType::ElasticDamage) pt = new ParticleElasticDamage<nlayer>(p[0], p[1], p[2], cell, type[i]);
      if (ptype == ParticleType::J2Plasticity) pt = new ParticleJ2Plasticity<nlayer>(p[0], p[1], p[2], cell, type[i]);
      if (ptype == ParticleType::FatigueHCF) pt = new ParticleFatigueHCF<nlayer>(p[0], p[1], p[2], cell, type[i]);
      pt->ass_ = this;
      pt->idx_ = i;
      pt->id = file_ids_.empty() ? i : file_ids_[i];
      pt_sys.push_back(pt);
    }
    std::vector<int32_t> bptr(n + 1), nb0(n), p2(std::max<int64_t>(nbond_, 1));
    check(lpmg_export_bonds(h, bptr.data(), nb0.data(), p2.data()), "Assembly");
    std::vector<int32_t> cptr(n + 1), conns(std::max<int64_t>(s[4], 1)), ncl(n);
    check(lpmg_export_conns(h, cptr.data(), conns.data(), ncl.data()), "Assembly");
    for (int i = 0; i < n; ++i) {
      Particle<nlayer>* pt = pt_sys[i];
      pt->nb = bptr[i + 1] - bptr[i];
      for (int b = bptr[i]; b < bptr[i + 1]; ++b) pt->neighbors.push_back(pt_sys[p2[b]]);
      pt->nconn = cptr[i + 1] - cptr[i];
      pt->nconn_largeq = ncl[i];
      for (int c = cptr[i]; c < cptr[i + 1]; ++c) pt->conns.push_back(pt_sys[conns[c]]);
    }
    std::vector<int32_t> t(type.begin(), type.end());
    check(lpmg_set_types(h, t.data()), "Assembly");
    sent_type_ = t;
    pull();
  }
};

template <int nlayer>
void Particle<nlayer>::queue(int model, std::initializer_list<double> p) {
  if (!ass_) throw Error(-1, "setParticleProperty on a particle outside an Assembly");
  ass_->queue_material(idx_, model, std::vector<double>(p));
}

// Stiffness (include/stiffness.h:18-52): the matrix is assembled on the GPU;
// IK/JK/K_pointer/K_global/residual hold the reference's upper-triangular
// 1-based CSR of the last assembly (filled by download(), which the solvers
// call after every assembly a caller can observe)
template <int nlayer>
class Stiffness {
  StiffnessMode mode;

 public:
  MKL_INT *IK = nullptr, *JK = nullptr;
  MKL_INT* K_pointer = nullptr;
  double *residual = nullptr, *K_global = nullptr;
  int precision = 0;  // 1: FP32 copy for the mixed-precision inner solve (no reference counterpart)

  Stiffness(std::vector<Particle<nlayer>*>& pt_sys, StiffnessMode p_mode) : mode(p_mode) { initialize(pt_sys); }
  ~Stiffness() { release(); }
  Stiffness(const Stiffness&) = delete;
  Stiffness& operator=(const Stiffness&) = delete;

  // Stiffness::initialize (stiffness.h:54-71)
  void initialize(std::vector<Particle<nlayer>*>& pt_sys) {
    if (pt_sys.empty() || !pt_sys[0]->owner()) throw Error(-1, "Stiffness needs the particles of an Assembly");
    ass_ = pt_sys[0]->owner();
    int64_t s[8];
    check(lpmg_sizes(ass_->h, s), "Stiffness");
    release();
    n_ = s[0], dim_ = s[1], nnz_ = s[3];
    K_pointer = new MKL_INT[n_ + 1]{};
    IK = new MKL_INT[dim_ * n_ + 1]{};
    JK = new MKL_INT[std::max<int64_t>(nnz_, 1)]{};
    residual = new double[dim_ * n_]{};
    K_global = new double[std::max<int64_t>(nnz_, 1)]{};
    check(lpmg_export_csr(ass_->h, (int64_t*)K_pointer, nullptr, nullptr, nullptr), "Stiffness");
  }
  // Stiffness::reset (stiffness.h:73-89): the device assembles into fresh storage
  void reset(std::vector<Particle<nlayer>*>&) {}
  // calcStiffness2D/3D + updateStiffnessDispBC (stiffness.h:197-438) are one device pass
  void calcStiffness2D(std::vector<Particle<nlayer>*>&) { assemble(); }
  void calcStiffness3D(std::vector<Particle<nlayer>*>&) { assemble(); }
  void updateStiffnessDispBC(std::vector<Particle<nlayer>*>&) {}
  void assemble() {
    ass_->run([&] { return lpmg_assemble(ass_->h, precision); }, "calcStiffness");
    download();
  }
  // the device's last assembly into the public arrays
  void download() {
    check(lpmg_export_csr(ass_->h, (int64_t*)K_pointer, (int64_t*)IK, (int64_t*)JK, K_global), "Stiffness");
    check(lpmg_get_field(ass_->h, LPMG_F_RESIDUAL, residual), "Stiffness");
  }

 private:
  Assembly<nlayer>* ass_ = nullptr;
  int64_t n_ = 0, dim_ = 0, nnz_ = 0;
  void release() {
    delete[] IK;
    delete[] JK;
    delete[] residual;
    delete[] K_pointer;
    delete[] K_global;
    IK = JK = K_pointer = nullptr;
    residual = K_global = nullptr;
  }
};

// Solver (include/solver.h:16-52)
template <int nlayer>
class Solver {
 public:
  int problem_size{0}, max_NR_iter{0};
  double* disp = nullptr;  // kept for the reference's layout; the increments stay on the GPU
  double tol_NR_iter{0};
  std::string dumpFile;
  std::vector<double> reaction_force;
  SolverMode sol_mode;
  Stiffness<nlayer> stiffness;
  Assembly<nlayer>& ass;
  // no reference counterpart: opts of lpmg_solve_linear [precond, rtol, atol,
  // maxit, precision, -, predictor, -]; the default is Jacobi-PCG stopped where
  // the reference stops MKL's CG (dpar[0] = dpar[1] = 1e-12, solver.h:312-313)
  std::array<double, 8> opts{1, 1e-12, 1e-12, 0, 0, 16, 0, 0};
  std::vector<double> last_stats = std::vector<double>(4, 0.0);

  Solver(Assembly<nlayer>& p_ass, const StiffnessMode& p_stiff_mode, const SolverMode& p_sol_mode,
         const std::string& p_dumpFile, const int& p_niter, const double& p_tol)
      : max_NR_iter(p_niter), tol_NR_iter(p_tol), dumpFile(p_dumpFile), sol_mode(p_sol_mode),
        stiffness(p_ass.pt_sys, p_stiff_mode), ass(p_ass) {
    problem_size = ass.pt_sys[0]->cell.dim * (int)ass.pt_sys.size();
    disp = new double[problem_size]{};
  }
  ~Solver() { delete[] disp; }
  Solver(const Solver&) = delete;
  Solver& operator=(const Solver&) = delete;

  // no reference counterpart: FP32 inner solves inside the FP64 Newton iteration
  void setMixedPrecision(bool on, double inner_rtol_sq = 1e-4) {
    opts[4] = on ? 1 : 0;
    stiffness.precision = on ? 1 : 0;
    if (on) opts[1] = inner_rtol_sq;
  }
  // no reference counterpart: multigrid-preconditioned CG (large 2D systems);
  // smoothed: the smoothed-aggregation hierarchy (fewer iterations)
  void setMultigrid(bool on, bool smoothed = false) { opts[0] = on ? (smoothed ? 4 : 2) : 1; }
  // no reference counterpart: Jacobi-preconditioned BiCGSTAB inner solves
  void setBiCGSTAB(bool on) { opts[0] = on ? 3 : 1; }

  // Solver::updateDisplacementBC / updateForceBC (solver.h:128-188)
  void updateDisplacementBC(LoadStep<nlayer>& load_step) {
    ass.run([&] {
      for (DispBC<nlayer>& bc : load_step.dispBCs)
        check(lpmg_disp_bc(ass.h, ass.group(bc.group), mode(bc.load_mode), bc.flag, bc.step), "updateDisplacementBC");
      return 0;
    }, "updateDisplacementBC");
  }
  void updateForceBC(LoadStep<nlayer>& load_step) {
    ass.run([&] {
      for (ForceBC<nlayer>& bc : load_step.forceBCs)
        check(lpmg_force_bc(ass.h, ass.group(bc.group), mode(bc.load_mode), bc.fx, bc.fy, bc.fz), "updateForceBC");
      return 0;
    }, "updateForceBC");
  }
  // Solver::updateRR (solver.h:111-126)
  void updateRR() {
    double n[2];
    ass.run([&] { return lpmg_update_rr(ass.h, n); }, "updateRR");
    check(lpmg_get_field(ass.h, LPMG_F_RESIDUAL, stiffness.residual), "updateRR");
  }
  // Solver::solveLinearSystem (solver.h:92-109): the GPU (P)CG for both modes;
  // a CG stopped at its limit keeps its update, as MKL's does (solver.h:348-349)
  void solveLinearSystem() {
    ass.run([&] {
      const int rc = lpmg_solve_linear(ass.h, opts.data(), last_stats.data());
      return rc == -4 ? 0 : rc;
    }, "solveLinearSystem");
  }
  void LPM_CG() { solveLinearSystem(); }
  void LPM_PARDISO() { solveLinearSystem(); }
  // Solver::NewtonIteration (solver.h:54-90)
  int NewtonIteration() {
    return ass.run([&] { return lpmg_newton(ass.h, max_NR_iter, tol_NR_iter, opts.data(), last_stats.data()); },
                   "NewtonIteration");
  }

 protected:
  static int mode(LoadMode m) { return m == LoadMode::Relative ? LPMG_RELATIVE : LPMG_ABSOLUTE; }

 public:
};

// SolverStatic (include/solver_static.h:20-152)
template <int nlayer>
class SolverStatic : public Solver<nlayer> {
 public:
  int undamaged_pt_type{0};
  SolverStatic(const int& p_undamaged, Assembly<nlayer>& p_ass, const StiffnessMode& p_stiff_mode,
               const SolverMode& p_sol_mode, const std::string& p_dumpFile, const int& p_niter, const double& p_tol)
      : Solver<nlayer>(p_ass, p_stiff_mode, p_sol_mode, p_dumpFile, p_niter, p_tol), undamaged_pt_type(p_undamaged) {}

  // SolverStatic::updateStaticDamage (solver_static.h:31-62)
  bool updateStaticDamage() {
    return this->ass.run([&] { return lpmg_update_static_damage(this->ass.h, undamaged_pt_type); },
                         "updateStaticDamage") == 1;
  }

  // SolverStatic::solveProblemStep (solver_static.h:97-152): BCs, then
  // {assemble, Newton, state update, damage update, store} while new damage appears
  bool solveProblemStep(LoadStep<nlayer>& load_step, double& dt) {
    const bool ok = step_(load_step, dt);
    this->stiffness.download();
    return ok;
  }

  // SolverStatic::solveProblem (solver_static.h:64-95): a step that does not
  // converge restores the last converged state, is cut in half and runs twice;
  // every converged step appends its state to dumpFile.  With opts[6] > 0 (the
  // load-step predictor) each step's predictor scale follows its size relative
  // to the last converged step's across halvings.
  void solveProblem(std::vector<LoadStep<nlayer>>& load, int& start_index) {
    int n_step = (int)load.size();
    double dt = 1;
    const double predict = this->opts[6];
    std::vector<double> size(load.size(), 1.0);
    double last = 1.0;
    for (int i = 0; i < n_step; i++) {
      if (predict > 0) this->opts[6] = predict * size[i] / last;
      while (!step_(load[i], dt)) {
        this->ass.resetStateVar(true);
        load[i].loadCutHalf();
        load.insert(load.begin() + i, load[i]);
        size[i] /= 2;
        size.insert(size.begin() + i, size[i]);
        ++n_step;
        if (predict > 0) this->opts[6] = predict * size[i] / last;
      }
      last = size[i];
      if (write_dumps) this->ass.writeDump(this->dumpFile, i + start_index);
    }
    this->opts[6] = predict;
    this->stiffness.download();
    start_index += n_step;
  }
  bool write_dumps = true;  // no reference counterpart: false skips the per-step dump

 private:
  bool step_(LoadStep<nlayer>& load_step, double&) {
    this->updateForceBC(load_step);
    this->updateDisplacementBC(load_step);
    return this->ass.run([&] { return lpmg_solve_step(this->ass.h, undamaged_pt_type, this->max_NR_iter,
                                                      this->tol_NR_iter, this->opts.data(), this->last_stats.data()); },
                         "solveProblemStep") == 1;
  }
};

// func_dNdt_linear / func_dNdt_exponental (solver_fatigue.h:65-73)
inline int func_dNdt_linear(const double& tau) { return (int)(1 / tau); }
inline int func_dNdt_exponental(const double& tau, const double& t) { return (int)(1 / tau * std::exp(t / tau)); }

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

  SolverFatigue(const int& p_undamaged, Assembly<nlayer>& p_ass, const StiffnessMode& p_stiff_mode,
                const SolverMode& p_sol_mode, const TimeMapMode& p_t_mode, const double& p_tau,
                const std::string& p_dumpFile, const int& p_niter, const double& p_tol)
      : Solver<nlayer>(p_ass, p_stiff_mode, p_sol_mode, p_dumpFile, p_niter, p_tol), undamaged_pt_type(p_undamaged),
        t_mode(p_t_mode), tau(p_tau) {}

  // solver_fatigue.h:41-63 (throws instead of exit(1) for a missing file)
  void readLoad(const std::string& loadFile) {
    FILE* f = std::fopen(loadFile.c_str(), "r");
    if (!f) throw Error(-1, "'" + loadFile + "' does not exist");
    double v;
    while (std::fscanf(f, "%lf", &v) > 0) load_spectrum.push_back(v);
    std::fclose(f);
  }

  static int func_dNdt_linear(const double& p_tau) { return lpm_gpu::func_dNdt_linear(p_tau); }
  static int func_dNdt_exponental(const double& p_tau, const double& t) { return lpm_gpu::func_dNdt_exponental(p_tau, t); }

  // SolverFatigue::updateFatigueDamage (solver_fatigue.h:75-126)
  bool updateFatigueDamage(double dNdt) {
    return this->ass.run([&] { return lpmg_update_fatigue_damage(this->ass.h, undamaged_pt_type, dNdt); },
                         "updateFatigueDamage") == 1;
  }

  // solver_fatigue.h:128-180
  std::vector<LoadStep<nlayer>> generateLoadCycle(FatigueLoadType ltype, int N, int n_interval,
                                                  std::vector<std::vector<Particle<nlayer>*>> g) {
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
    this->stiffness.reset(this->ass.pt_sys);
    if (this->ass.pt_sys[0]->cell.dim == 2) this->stiffness.calcStiffness2D(this->ass.pt_sys);
    else this->stiffness.calcStiffness3D(this->ass.pt_sys);
    this->stiffness.updateStiffnessDispBC(this->ass.pt_sys);
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
  int solveProblemCyclic(FatigueLoadType loadType, int n_interval, std::vector<std::vector<Particle<nlayer>*>> pt_groups) {
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
