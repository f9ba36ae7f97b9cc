// Prints the points of one lattice generator of include/lpmcpp/utilities.h
// (LPM-CPP's createRMatrix + createPlateSQ2D / createPlateHEX2D /
// createCuboidSC3D / createCuboidFCC3D) as hexadecimal floats, one point per
// line, for comparison with lpmg.lattice.
//
// Usage: lattice_gen <sq|hex|sc|fcc> <eulerflag> <a1> <a2> <a3> <radius> <xmin> <xmax> <ymin> <ymax> <zmin> <zmax>
// (angles in degrees, converted as the examples do: PI / 180.0 * angle)
#include "lpm.h"
#include "utilities.h"

int main(int argc, char** argv) {
  if (argc < 13) return 2;
  const std::string kind = argv[1];
  const int flag = std::atoi(argv[2]);
  double angles[3];
  for (int k = 0; k < 3; ++k) angles[k] = PI / 180.0 * std::atof(argv[3 + k]);
  const double radius = std::atof(argv[6]);
  std::array<double, 2 * NDIM> box;
  for (int k = 0; k < 6; ++k) box[k] = std::atof(argv[7 + k]);
  double* R = createRMatrix(flag, angles);
  LatticeType lt = kind == "sq" ? LatticeType::Square2D : kind == "hex" ? LatticeType::Hexagon2D
                   : kind == "sc" ? LatticeType::SimpleCubic3D : LatticeType::FCC3D;
  UnitCell cell(lt, radius);
  std::vector<std::array<double, NDIM>> pts = kind == "sq" ? createPlateSQ2D(box, cell, R)
                                              : kind == "hex" ? createPlateHEX2D(box, cell, R)
                                              : kind == "sc" ? createCuboidSC3D(box, cell, R)
                                                             : createCuboidFCC3D(box, cell, R);
  for (int k = 0; k < 9; ++k) std::printf("%a%c", R[k], k == 8 ? '\n' : ' ');
  for (auto& p : pts) std::printf("%a %a %a\n", p[0], p[1], p[2]);
  return 0;
}
