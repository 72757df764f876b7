// solid_shell.cpp -- the C++ mirror of sim::solidshell over the C ABI.
#include "solid_shell.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <stdexcept>

namespace shipfem::solidshell {

namespace {

void requireSolidShell(Formulation form) {
    if (form != Formulation::SolidShell)
        throw std::invalid_argument(std::string("solidshell: formulation ") + name(form) +
                                    " is not built; only SolidShell is");
}

}  // namespace

const char* name(Formulation form) {
    switch (form) {
        case Formulation::Displacement: return "displacement";
        case Formulation::AssumedNaturalStrain: return "ans";
        case Formulation::SolidShell: return "solid-shell";
    }
    return "unknown";
}

void elementStiffness(const double nodes[kDof], const StructuralMaterial& material,
                      Formulation form, double out[kDof * kDof]) {
    requireSolidShell(form);
    if (shipfem_hex_stiffness(nodes, material.youngsModulus, material.poissonRatio, out) != SHIPFEM_OK)
        std::fill(out, out + kDof * kDof, 0.0);
}

void elementMass(const double nodes[kDof], double density, double out[kNodes]) {
    if (shipfem_hex_mass(nodes, density, out) != SHIPFEM_OK) std::fill(out, out + kNodes, 0.0);
}

void elementRotation(const double rest[kDof], const double current[kDof], double out[9]) {
    // the C entry point leaves the identity in `out` for a degenerate rest element
    shipfem_hex_rotation(rest, current, out);
}

void internalForce(const double stiffness[kDof * kDof], const double rest[kDof],
                   const double current[kDof], double out[kDof]) {
    shipfem_hex_force(stiffness, rest, current, out);
}

double criticalTimestep(const double nodes[kDof], const StructuralMaterial& material,
                        Formulation form, double safety) {
    requireSolidShell(form);
    double dt = 0.0;
    if (shipfem_hex_critical_timestep(nodes, material.youngsModulus, material.poissonRatio,
                                      material.density, safety, &dt) != SHIPFEM_OK)
        return 0.0;
    return dt;
}

void HexMesh::pin(std::size_t node, int axis, double value) {
    if (fixed.size() != nodeCount() * 3) fixed.assign(nodeCount() * 3, 0u);
    if (prescribed.size() != nodeCount() * 3) prescribed.assign(nodeCount() * 3, 0.0);
    fixed[node * 3 + static_cast<std::size_t>(axis)] = 1u;
    prescribed[node * 3 + static_cast<std::size_t>(axis)] = value;
}

void HexMesh::gather(std::size_t element, const std::vector<double>& nodal, double out[kDof]) const {
    for (int a = 0; a < kNodes; ++a) {
        const std::size_t n = index[element * kNodes + static_cast<std::size_t>(a)];
        for (int i = 0; i < 3; ++i) out[3 * a + i] = nodal[3 * n + static_cast<std::size_t>(i)];
    }
}

HexMesh makePlateMesh(double lx, double ly, double thickness, int nx, int ny, int nz) {
    HexMesh mesh;
    uint32_t N = 0, E = 0;
    if (shipfem_hex_plate_mesh(lx, ly, thickness, nx, ny, nz, &N, &E, nullptr, nullptr) != SHIPFEM_OK)
        return mesh;
    mesh.position.resize(3ull * N);
    mesh.index.resize(8ull * E);
    shipfem_hex_plate_mesh(lx, ly, thickness, nx, ny, nz, &N, &E, mesh.position.data(),
                           mesh.index.data());
    return mesh;
}

double criticalTimestep(const HexMesh& mesh, const StructuralMaterial& material,
                        Formulation form, double safety) {
    requireSolidShell(form);
    double smallest = std::numeric_limits<double>::infinity();
    for (std::size_t e = 0; e < mesh.elementCount(); ++e) {
        double nodes[kDof];
        mesh.gather(e, mesh.position, nodes);
        const double dt = criticalTimestep(nodes, material, form, safety);
        if (dt > 0.0) smallest = std::min(smallest, dt);
    }
    return std::isfinite(smallest) ? smallest : 0.0;
}

std::vector<double> uniformPressureLoad(const HexMesh& mesh, double pressure) {
    std::vector<double> load(mesh.position.size(), 0.0);
    if (shipfem_hex_pressure_load(mesh.position.data(), static_cast<uint32_t>(mesh.nodeCount()),
                                  mesh.index.data(), static_cast<uint32_t>(mesh.elementCount()),
                                  pressure, load.data()) != SHIPFEM_OK)
        std::fill(load.begin(), load.end(), 0.0);
    return load;
}

}  // namespace shipfem::solidshell
