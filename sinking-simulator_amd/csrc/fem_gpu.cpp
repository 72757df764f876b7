// fem_gpu.cpp -- the C++ mirror of sim::fem set-up and gpu::FemGpuSolver over the C ABI.
#include "fem_gpu.hpp"

#include <algorithm>
#include <limits>

namespace shipfem::fem {

TetMesh makeBoxTetMesh(float lx, float ly, float lz, int nx, int ny, int nz) {
    TetMesh mesh;
    shipfem_mesh* m = nullptr;
    if (shipfem_mesh_box(lx, ly, lz, nx, ny, nz, &m) != SHIPFEM_OK) return mesh;
    uint32_t N = 0, T = 0;
    shipfem_mesh_counts(m, &N, &T);
    const double* X = nullptr;
    const uint32_t* idx = nullptr;
    shipfem_mesh_arrays(m, &X, &idx, nullptr, nullptr);
    mesh.position.assign(X, X + 3ull * N);  // float, as the reference stores it
    mesh.index.assign(idx, idx + 4ull * T);
    shipfem_mesh_destroy(m);
    return mesh;
}

void TetMesh::computeRestState(const Material& material) {
    // The device evaluates the rest state itself (in FP64 from restPosition); these
    // host arrays keep the reference's fields filled for callers that read them.
    const std::size_t N = position.size() / 3, T = index.size() / 4;
    restPosition.assign(position.begin(), position.end());
    velocity.assign(N * 3, 0.0f);
    mass.assign(N, 0.0f);
    if (fixed.size() != N) fixed.assign(N, 0u);
    restInv.assign(T * 9, 0.0f);
    restVolume.assign(T, 0.0f);
    tetForce.assign(T * 12, 0.0f);
    std::vector<double> m(N, 0.0);
    for (std::size_t t = 0; t < T; ++t) {
        const uint32_t* v = &index[t * 4];
        double d[3][3];
        for (int c = 0; c < 3; ++c)
            for (int a = 0; a < 3; ++a) d[a][c] = restPosition[v[c + 1] * 3 + a] - restPosition[v[0] * 3 + a];
        const double det = d[0][0] * (d[1][1] * d[2][2] - d[1][2] * d[2][1]) -
                           d[0][1] * (d[1][0] * d[2][2] - d[1][2] * d[2][0]) +
                           d[0][2] * (d[1][0] * d[2][1] - d[1][1] * d[2][0]);
        const double id = std::fabs(det) > 1e-30 ? 1.0 / det : 0.0;
        const double inv[3][3] = {{(d[1][1] * d[2][2] - d[1][2] * d[2][1]) * id,
                                   (d[0][2] * d[2][1] - d[0][1] * d[2][2]) * id,
                                   (d[0][1] * d[1][2] - d[0][2] * d[1][1]) * id},
                                  {(d[1][2] * d[2][0] - d[1][0] * d[2][2]) * id,
                                   (d[0][0] * d[2][2] - d[0][2] * d[2][0]) * id,
                                   (d[0][2] * d[1][0] - d[0][0] * d[1][2]) * id},
                                  {(d[1][0] * d[2][1] - d[1][1] * d[2][0]) * id,
                                   (d[0][1] * d[2][0] - d[0][0] * d[2][1]) * id,
                                   (d[0][0] * d[1][1] - d[0][1] * d[1][0]) * id}};
        for (int c = 0; c < 3; ++c)
            for (int r = 0; r < 3; ++r) restInv[t * 9 + c * 3 + r] = static_cast<float>(inv[r][c]);
        restVolume[t] = static_cast<float>(std::fabs(det) / 6.0);
        for (int c = 0; c < 4; ++c) m[v[c]] += material.density * std::fabs(det) / 24.0;
    }
    for (std::size_t n = 0; n < N; ++n) mass[n] = static_cast<float>(m[n]);
    adjacencyOffset.assign(N + 1, 0u);
    for (uint32_t q : index) ++adjacencyOffset[q + 1];
    for (std::size_t n = 0; n < N; ++n) adjacencyOffset[n + 1] += adjacencyOffset[n];
    adjacencyEntry.resize(adjacencyOffset.back());
    std::vector<uint32_t> cursor(adjacencyOffset.begin(), adjacencyOffset.end() - 1);
    for (std::size_t t = 0; t < T; ++t)
        for (uint32_t c = 0; c < 4; ++c) adjacencyEntry[cursor[index[t * 4 + c]]++] = static_cast<uint32_t>(t) * 4 + c;
}

float TetMesh::shortestEdge() const {
    float s = std::numeric_limits<float>::infinity();
    for (std::size_t t = 0; t < index.size() / 4; ++t)
        for (int a = 0; a < 4; ++a)
            for (int b = a + 1; b < 4; ++b) {
                float d2 = 0;
                for (int k = 0; k < 3; ++k) {
                    const float d = position[index[t * 4 + a] * 3 + k] - position[index[t * 4 + b] * 3 + k];
                    d2 += d * d;
                }
                s = std::min(s, std::sqrt(d2));
            }
    return s;
}

float TetMesh::totalMass() const {
    float m = 0;
    for (float v : mass) m += v;
    return m;
}

float criticalTimestep(const TetMesh& mesh, const Material& material, float safety) {
    return safety * mesh.shortestEdge() / material.waveSpeed();
}

namespace {
const double* restOf(const TetMesh& mesh) {
    return mesh.restPosition.size() == mesh.position.size() ? mesh.restPosition.data() : nullptr;
}
}  // namespace

void stepCpu(TetMesh& mesh, const Material& material, float dt, float gravity, float damping) {
    const std::size_t N = mesh.nodeCount(), T = mesh.tetCount();
    if (mesh.tetForce.size() != 12 * T) mesh.tetForce.assign(12 * T, 0.0f);
    if (mesh.velocity.size() != 3 * N) mesh.velocity.assign(3 * N, 0.0f);
    shipfem_cpu_step(static_cast<uint32_t>(N), static_cast<uint32_t>(T), mesh.position.data(),
                     mesh.velocity.data(), mesh.mass.data(), mesh.fixed.size() == N ? mesh.fixed.data() : nullptr,
                     mesh.index.data(), mesh.restInv.data(), mesh.restVolume.data(), mesh.tetForce.data(),
                     mesh.adjacencyOffset.data(), mesh.adjacencyEntry.data(), restOf(mesh),
                     material.lameLambda(), material.lameMu(), dt, gravity, damping);
}

void tetForces(const TetMesh& mesh, std::size_t tet, float lambda, float mu, float out[12]) {
    shipfem_cpu_tet_forces(mesh.position.data(), restOf(mesh), mesh.index.data(), mesh.restInv.data(),
                           mesh.restVolume.data(), tet, lambda, mu, out);
}

}  // namespace shipfem::fem

namespace shipfem::gpu {

FemGpuSolver::~FemGpuSolver() {
    if (solver_) shipfem_solver_destroy(solver_);
}

bool FemGpuSolver::initialise(const fem::TetMesh& mesh, const std::string&, std::string& error) {
    if (solver_) {
        shipfem_solver_destroy(solver_);
        solver_ = nullptr;
    }
    char name[256] = {0};
    uint64_t hbm = 0;
    if (shipfem_device_query(name, sizeof name, &hbm) != SHIPFEM_OK) {
        error = name;
        return false;
    }
    deviceName_ = name;
    const std::size_t N = mesh.position.size() / 3, T = mesh.index.size() / 4;
    rest_ = mesh.restPosition.size() == N * 3 ? mesh.restPosition
                                              : std::vector<double>(mesh.position.begin(), mesh.position.end());
    std::vector<double> u0(N * 3), v0(N * 3, 0.0);
    for (std::size_t q = 0; q < N * 3; ++q) u0[q] = static_cast<double>(mesh.position[q]) - rest_[q];
    for (std::size_t q = 0; q < std::min(v0.size(), mesh.velocity.size()); ++q) v0[q] = mesh.velocity[q];
    std::vector<uint32_t> flags(N, 0u);
    for (std::size_t n = 0; n < std::min(N, mesh.fixed.size()); ++n)
        if (mesh.fixed[n]) flags[n] = SHIPFEM_PIN_ALL;
    shipfem_case_desc d{};
    d.node_count = static_cast<uint32_t>(N);
    d.tet_count = static_cast<uint32_t>(T);
    d.rest_position = rest_.data();
    d.index = mesh.index.data();
    d.node_flags = flags.data();
    d.initial_displacement = u0.data();
    d.initial_velocity = v0.data();
    d.material = shipfem_material{210.0e9, 0.30, 7850.0, 0, 0.0, 0.0};
    // the caller's lumped masses, node by node (TetMesh::mass, fem.hpp:42)
    std::vector<double> mass;
    if (mesh.mass.size() == N) {
        mass.assign(mesh.mass.begin(), mesh.mass.end());
        d.mass = mass.data();
    }
    char err[512] = {0};
    const shipfem_status s = shipfem_solver_create(&d, &solver_, err, sizeof err);
    if (s != SHIPFEM_OK) {
        error = std::string(shipfem_status_string(s)) + ": " + err;
        solver_ = nullptr;
        return false;
    }
    return true;
}

double FemGpuSolver::run(int substeps, const fem::Material& material, float dt, float gravity,
                         float damping) {
    if (!solver_) return 0.0;
    const shipfem_step_params p{dt, gravity, damping, material.youngsModulus, material.poissonRatio};
    if (shipfem_solver_set_params(solver_, &p) != SHIPFEM_OK) return 0.0;
    double ms = 0;
    shipfem_solver_run(solver_, substeps, 0, &ms);
    return ms;
}

void FemGpuSolver::readback(fem::TetMesh& mesh) {
    if (!solver_) return;
    const std::size_t N = rest_.size() / 3;
    std::vector<double> u(N * 3), v(N * 3);
    shipfem_state_view view{};
    view.displacement = u.data();
    view.velocity = v.data();
    if (shipfem_solver_readback(solver_, &view) != SHIPFEM_OK) return;
    mesh.position.resize(N * 3);
    mesh.velocity.resize(N * 3);
    for (std::size_t q = 0; q < N * 3; ++q) {
        mesh.position[q] = static_cast<float>(rest_[q] + u[q]);
        mesh.velocity[q] = static_cast<float>(v[q]);
    }
}

double FemGpuSolver::stableTimestep() const {
    double cfl = 0;
    if (solver_) shipfem_solver_timestep(solver_, &cfl, nullptr, nullptr);
    return cfl;
}

}  // namespace shipfem::gpu
