// fem_gpu.hpp -- C++ mirror of the reference's tet FEM interface on top of the C ABI.
//
//   shipfem::fem    mirrors sim::fem   (engine/sim/fem.hpp:22-88): Material, TetMesh,
//                   makeBoxTetMesh, criticalTimestep, stepCpu, tetForces.
//   shipfem::gpu    mirrors gpu::FemGpuSolver (engine/gpu/fem_gpu.hpp:17-45):
//                   initialise / run / readback / deviceName / valid.
//
// Same names, argument meanings and error behaviour: initialise returns false with a
// reason (including "no device", which callers treat as a skip); run returns device
// milliseconds; readback copies positions and velocities into the mesh. One addition:
// TetMesh::restPosition, filled by computeRestState, because the device keeps rest
// coordinates and displacement apart (FP32 positions of a 10 m plate cannot carry a
// nanometre step).
#pragma once

#include "../../include/shipfem.h"

#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

namespace shipfem::fem {

struct Material {
    float youngsModulus = 210.0e9f;
    float poissonRatio = 0.30f;
    float density = 7850.0f;
    float lameLambda() const {
        return youngsModulus * poissonRatio / ((1.0f + poissonRatio) * (1.0f - 2.0f * poissonRatio));
    }
    float lameMu() const { return youngsModulus / (2.0f * (1.0f + poissonRatio)); }
    float waveSpeed() const { return std::sqrt(youngsModulus / density); }
};

struct TetMesh {
    std::vector<float> position, velocity, mass;
    std::vector<uint32_t> fixed;
    std::vector<uint32_t> index;
    std::vector<float> restInv, restVolume, tetForce;
    std::vector<uint32_t> adjacencyOffset, adjacencyEntry;
    std::vector<double> restPosition;  // rest coordinates in double (set by computeRestState)

    std::size_t nodeCount() const { return mass.size(); }
    std::size_t tetCount() const { return restVolume.size(); }
    void computeRestState(const Material& material);
    float shortestEdge() const;
    float totalMass() const;
};

TetMesh makeBoxTetMesh(float lx, float ly, float lz, int nx, int ny, int nz);
float criticalTimestep(const TetMesh& mesh, const Material& material, float safety = 0.5f);

// One explicit FP32 step on the host (the library's shipfem_cpu_step): element forces into
// tetForce, CSR gather, symplectic Euler. `damping` multiplies velocity each step.
void stepCpu(TetMesh& mesh, const Material& material, float dt, float gravity, float damping);

// The four nodal forces one tet contributes at the current positions
// (shipfem_cpu_tet_forces, displacement form when restPosition is set).
void tetForces(const TetMesh& mesh, std::size_t tet, float lambda, float mu, float out[12]);

}  // namespace shipfem::fem

namespace shipfem::gpu {

class FemGpuSolver {
public:
    FemGpuSolver() = default;
    FemGpuSolver(const FemGpuSolver&) = delete;
    FemGpuSolver& operator=(const FemGpuSolver&) = delete;
    ~FemGpuSolver();

    // `shaderDirectory` is accepted for signature compatibility and ignored: the
    // kernels are compiled into the library.
    bool initialise(const fem::TetMesh& mesh, const std::string& shaderDirectory, std::string& error);
    double run(int substeps, const fem::Material& material, float dt, float gravity, float damping);
    void readback(fem::TetMesh& mesh);
    const std::string& deviceName() const { return deviceName_; }
    bool valid() const { return solver_ != nullptr; }
    // The FP64 stable step of the mesh on the device (2/omega_max).
    double stableTimestep() const;

private:
    shipfem_solver* solver_ = nullptr;
    std::string deviceName_;
    std::vector<double> rest_;
};

}  // namespace shipfem::gpu
