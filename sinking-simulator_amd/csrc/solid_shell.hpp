// solid_shell.hpp -- C++ mirror of the reference's solid-shell element interface on top of
// the C ABI.
//
//   shipfem::solidshell mirrors sim::solidshell (engine/sim/solid_shell.hpp:111-126,
//   200-203, 327-370, 546-560, 714-719): the element constants and node ordering,
//   StructuralMaterial, elementStiffness, elementMass, elementRotation, internalForce,
//   criticalTimestep (element and mesh), HexMesh with pin / gather, makePlateMesh and
//   uniformPressureLoad.
//
// Same names, argument meanings and error behaviour: a degenerate or inverted element
// gives a zero stiffness and mass, the identity rotation and a zero stable step, and the
// mesh step is the smallest positive element step (zero when there is none). One
// difference: the library builds the solid-shell formulation only, so the
// Displacement and AssumedNaturalStrain controls throw std::invalid_argument.
#pragma once

#include "../../include/shipfem.h"

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace shipfem::solidshell {

inline constexpr int kNodes = 8;   // nodes per element; 0-3 on zeta = -1, 4-7 on zeta = +1
inline constexpr int kDof = 24;    // 3 translations each
inline constexpr int kEas = 7;     // enhanced assumed strain parameters
inline constexpr int kGauss = 8;   // 2x2x2 Gauss points

enum class Formulation { Displacement, AssumedNaturalStrain, SolidShell };
const char* name(Formulation form);

struct StructuralMaterial {
    std::string name = "steel_ah36";
    double density = 7850.0;
    double youngsModulus = 206.0e9;
    double poissonRatio = 0.30;
    double yieldStrength = 355.0e6;
};

// 24 x 24 row-major condensed stiffness in the global frame.
void elementStiffness(const double nodes[kDof], const StructuralMaterial& material,
                      Formulation form, double out[kDof * kDof]);
// Row-sum lumped mass: density times the integral of N_a.
void elementMass(const double nodes[kDof], double density, double out[kNodes]);
// Polar factor of the centre deformation gradient, column-major.
void elementRotation(const double rest[kDof], const double current[kDof], double out[9]);
// f = -R K (R^T x - X), the force the element applies to its nodes.
void internalForce(const double stiffness[kDof * kDof], const double rest[kDof],
                   const double current[kDof], double out[kDof]);
// safety * 2 / omega_max of the element with its lumped mass.
double criticalTimestep(const double nodes[kDof], const StructuralMaterial& material,
                        Formulation form, double safety = 0.9);

struct HexMesh {
    std::vector<double> position;       // 3 per node
    std::vector<std::uint32_t> index;   // 8 per element
    std::vector<std::uint8_t> fixed;    // 3 per node; non-zero pins that DOF
    std::vector<double> prescribed;     // 3 per node; the value a pinned DOF takes

    std::size_t nodeCount() const { return position.size() / 3; }
    std::size_t elementCount() const { return index.size() / 8; }

    void pin(std::size_t node, int axis, double value = 0.0);
    void gather(std::size_t element, const std::vector<double>& nodal, double out[kDof]) const;
};

// lx x ly x thickness, nx x ny in plane and nz through it; thickness index fastest.
HexMesh makePlateMesh(double lx, double ly, double thickness, int nx, int ny, int nz = 1);
double criticalTimestep(const HexMesh& mesh, const StructuralMaterial& material,
                        Formulation form, double safety = 0.9);
// Consistent nodal load of a pressure on every boundary +zeta face; 3 per node.
std::vector<double> uniformPressureLoad(const HexMesh& mesh, double pressure);

}  // namespace shipfem::solidshell
