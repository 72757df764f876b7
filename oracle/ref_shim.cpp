// ref_shim.cpp -- extern "C" entry points over the reference's own CPU FEM and J2
// point law, so tests/golden/make_golden.py can record golden vectors from the
// compiled reference. TEST INFRASTRUCTURE ONLY.
//
// Built by oracle/Makefile together with the reference sources where they lie:
//   engine/sim/fem.cpp         (makeBoxTetMesh, computeRestState, tetForces, stepCpu)
//   engine/sim/plasticity.cpp  (update, regularisedFailureStrain, triaxialityFactor)
//   engine/sim/solid_shell.cpp (elementStiffness, elementMass, internalForce,
//                               elementRotation, criticalTimestep of the hex element)
// into oracle/_ref/libref.so. Nothing here restates the reference; it only moves
// arrays in and out of sim::fem::TetMesh and sim::plasticity::State.
#include "engine/sim/fem.hpp"
#include "engine/sim/plasticity.hpp"
#include "engine/sim/solid_shell.hpp"

#include <cstdint>
#include <cstring>

namespace {

sim::fem::TetMesh meshFrom(const float* position, int64_t N, const uint32_t* index, int64_t T) {
    sim::fem::TetMesh mesh;
    mesh.position.assign(position, position + N * 3);
    mesh.index.assign(index, index + T * 4);
    return mesh;
}

sim::StructuralMaterial structuralOf(double E, double nu, double rho) {
    sim::StructuralMaterial m;
    m.youngsModulus = E;
    m.poissonRatio = nu;
    m.density = rho;
    return m;
}

sim::solidshell::Formulation formOf(int form) {
    return form == 0   ? sim::solidshell::Formulation::Displacement
           : form == 1 ? sim::solidshell::Formulation::AssumedNaturalStrain
                       : sim::solidshell::Formulation::SolidShell;
}

sim::fem::Material materialOf(float E, float nu, float rho) {
    sim::fem::Material m;
    m.youngsModulus = E;
    m.poissonRatio = nu;
    m.density = rho;
    return m;
}

}  // namespace

extern "C" {

void ref_box_counts(int nx, int ny, int nz, int64_t* nodes, int64_t* tets) {
    const sim::fem::TetMesh m = sim::fem::makeBoxTetMesh(1.0f, 1.0f, 1.0f, nx, ny, nz);
    *nodes = static_cast<int64_t>(m.position.size() / 3);
    *tets = static_cast<int64_t>(m.index.size() / 4);
}

void ref_box_mesh(float lx, float ly, float lz, int nx, int ny, int nz, float* position,
                  uint32_t* index) {
    const sim::fem::TetMesh m = sim::fem::makeBoxTetMesh(lx, ly, lz, nx, ny, nz);
    std::memcpy(position, m.position.data(), m.position.size() * sizeof(float));
    std::memcpy(index, m.index.data(), m.index.size() * sizeof(uint32_t));
}

// computeRestState on the given positions; also returns criticalTimestep(safety).
float ref_rest_state(const float* position, int64_t N, const uint32_t* index, int64_t T, float E,
                     float nu, float rho, float safety, float* rest_inv, float* volume,
                     float* mass, uint32_t* adj_offset, uint32_t* adj_entry) {
    sim::fem::TetMesh m = meshFrom(position, N, index, T);
    const sim::fem::Material mat = materialOf(E, nu, rho);
    m.computeRestState(mat);
    std::memcpy(rest_inv, m.restInv.data(), m.restInv.size() * sizeof(float));
    std::memcpy(volume, m.restVolume.data(), m.restVolume.size() * sizeof(float));
    std::memcpy(mass, m.mass.data(), m.mass.size() * sizeof(float));
    std::memcpy(adj_offset, m.adjacencyOffset.data(), m.adjacencyOffset.size() * sizeof(uint32_t));
    std::memcpy(adj_entry, m.adjacencyEntry.data(), m.adjacencyEntry.size() * sizeof(uint32_t));
    return sim::fem::criticalTimestep(m, mat, safety);
}

// Rest state from `rest`, then positions replaced by `current` (the fem_spike
// deform-after-rest pattern, tools/fem_spike/main.cpp:147-159), then tetForces for
// every tet into out[12T].
void ref_tet_forces(const float* rest, const float* current, int64_t N, const uint32_t* index,
                    int64_t T, float E, float nu, float rho, float* out) {
    sim::fem::TetMesh m = meshFrom(rest, N, index, T);
    const sim::fem::Material mat = materialOf(E, nu, rho);
    m.computeRestState(mat);
    std::memcpy(m.position.data(), current, static_cast<size_t>(N) * 3 * sizeof(float));
    for (int64_t t = 0; t < T; ++t)
        sim::fem::tetForces(m, static_cast<size_t>(t), mat.lameLambda(), mat.lameMu(), &out[t * 12]);
}

// Rest state from `rest`, pins from `fixed`, positions := current, velocity :=
// velocity; `steps` calls of stepCpu; positions and velocities copied back.
void ref_step(const float* rest, float* position, float* velocity, const uint32_t* fixed, int64_t N,
              const uint32_t* index, int64_t T, float E, float nu, float rho, float dt,
              float gravity, float damping, int steps) {
    sim::fem::TetMesh m = meshFrom(rest, N, index, T);
    const sim::fem::Material mat = materialOf(E, nu, rho);
    m.computeRestState(mat);
    m.fixed.assign(fixed, fixed + N);
    std::memcpy(m.position.data(), position, static_cast<size_t>(N) * 3 * sizeof(float));
    std::memcpy(m.velocity.data(), velocity, static_cast<size_t>(N) * 3 * sizeof(float));
    for (int s = 0; s < steps; ++s) sim::fem::stepCpu(m, mat, dt, gravity, damping);
    std::memcpy(position, m.position.data(), static_cast<size_t>(N) * 3 * sizeof(float));
    std::memcpy(velocity, m.velocity.data(), static_cast<size_t>(N) * 3 * sizeof(float));
}

// plasticity::update with a linear hardening curve, no kinematic hardening and no
// failure. State in/out; returns the number of consistency evaluations.
int ref_j2_update(double E, double nu, double yield, double hardening, const double strain[6],
                  double eps_p[6], double* eqps, double stress[6], double* dissipation,
                  int* yielded) {
    sim::plasticity::Material mat;
    mat.youngsModulus = E;
    mat.poissonRatio = nu;
    mat.flow = sim::plasticity::linearHardening(yield, hardening);
    sim::plasticity::State st;
    for (int i = 0; i < 6; ++i) st.plasticStrain[i] = eps_p[i];
    st.equivalentPlasticStrain = *eqps;
    const sim::plasticity::Increment inc =
        sim::plasticity::update(mat, sim::plasticity::kNeverFails, strain, st, stress);
    for (int i = 0; i < 6; ++i) eps_p[i] = st.plasticStrain[i];
    *eqps = st.equivalentPlasticStrain;
    *dissipation = inc.dissipation;
    *yielded = inc.yielded ? 1 : 0;
    return inc.iterations;
}

// plasticity::update with a linear hardening curve and ductile failure: the given
// (already regularised) failure strain and the Rice-Tracey constants. State in/out
// including damage and the failed flag; returns 1 when the point failed in this call.
int ref_j2_damage_update(double E, double nu, double yield, double hardening,
                         double failure_strain, double sensitivity, double reference_eta,
                         double cutoff_eta, const double strain[6], double eps_p[6],
                         double* eqps, double* damage, int* failed, double stress[6],
                         double* dissipation, double normal[3], double* triaxiality,
                         int* yielded) {
    sim::plasticity::Material mat;
    mat.youngsModulus = E;
    mat.poissonRatio = nu;
    mat.flow = sim::plasticity::linearHardening(yield, hardening);
    mat.failure.triaxialitySensitivity = sensitivity;
    mat.failure.referenceTriaxiality = reference_eta;
    mat.failure.cutoffTriaxiality = cutoff_eta;
    sim::plasticity::State st;
    for (int i = 0; i < 6; ++i) st.plasticStrain[i] = eps_p[i];
    st.equivalentPlasticStrain = *eqps;
    st.damage = *damage;
    st.failed = *failed != 0;
    const sim::plasticity::Increment inc =
        sim::plasticity::update(mat, failure_strain, strain, st, stress);
    for (int i = 0; i < 6; ++i) eps_p[i] = st.plasticStrain[i];
    *eqps = st.equivalentPlasticStrain;
    *damage = st.damage;
    *failed = st.failed ? 1 : 0;
    *dissipation = inc.dissipation;
    *triaxiality = inc.triaxiality;
    *yielded = inc.yielded ? 1 : 0;
    for (int i = 0; i < 3; ++i) normal[i] = inc.failureNormal[i];
    return inc.failedNow ? 1 : 0;
}

double ref_regularised_failure_strain(double uniform, double fracture, double length,
                                      double thickness) {
    sim::plasticity::Failure f;
    f.uniformStrain = uniform;
    f.fractureStrain = fracture;
    return sim::plasticity::regularisedFailureStrain(f, length, thickness);
}

double ref_triaxiality_factor(double sensitivity, double reference_eta, double cutoff_eta,
                              double eta) {
    sim::plasticity::Failure f;
    f.triaxialitySensitivity = sensitivity;
    f.referenceTriaxiality = reference_eta;
    f.cutoffTriaxiality = cutoff_eta;
    return sim::plasticity::triaxialityFactor(f, eta);
}

// The solid-shell hex (8 nodes x xyz, node-major). form: 0 Displacement,
// 1 AssumedNaturalStrain, 2 SolidShell.
void ref_hex_stiffness(const double nodes[24], double E, double nu, int form, double out[576]) {
    sim::solidshell::elementStiffness(nodes, structuralOf(E, nu, 7850.0), formOf(form), out);
}

void ref_hex_mass(const double nodes[24], double rho, double out[8]) {
    sim::solidshell::elementMass(nodes, rho, out);
}

void ref_hex_force(const double stiffness[576], const double rest[24], const double current[24],
                   double out[24]) {
    sim::solidshell::internalForce(stiffness, rest, current, out);
}

void ref_hex_rotation(const double rest[24], const double current[24], double out[9]) {
    sim::solidshell::elementRotation(rest, current, out);
}

// elementPlasticUpdate (solid_shell.cpp:826-1029) along a path of K current
// configurations from a fresh state whose failure strain is `failure_strain` (<= 0: the
// point law's never-fails value). Per step k: force[24], stress[48] (8 points x 6),
// point state[8 x 9] (plastic strain 6, eqps, damage, failed), enhanced[7], and
// info[6] = {dissipation, iterations, converged, yielded points, failed points, torn}.
void ref_hex_plastic_path(const double rest[24], const double* currents, int K, double E,
                          double nu, double yield, double hardening, double failure_strain,
                          double sensitivity, double reference_eta, double cutoff_eta,
                          double* force, double* stress, double* state, double* enhanced,
                          double* info) {
    sim::plasticity::Material mat;
    mat.youngsModulus = E;
    mat.poissonRatio = nu;
    mat.flow = sim::plasticity::linearHardening(yield, hardening);
    mat.failure.triaxialitySensitivity = sensitivity;
    mat.failure.referenceTriaxiality = reference_eta;
    mat.failure.cutoffTriaxiality = cutoff_eta;
    sim::solidshell::ElementPlasticState st;
    st.failureStrain = failure_strain > 0 ? failure_strain : sim::plasticity::kNeverFails;
    for (int k = 0; k < K; ++k) {
        const sim::solidshell::PlasticUpdate u = sim::solidshell::elementPlasticUpdate(
            rest, currents + 24 * k, mat, sim::solidshell::Formulation::SolidShell, st,
            force + 24 * k, stress + 48 * k);
        for (int g = 0; g < 8; ++g) {
            double* q = state + 72 * k + 9 * g;
            for (int i = 0; i < 6; ++i) q[i] = st.point[g].plasticStrain[i];
            q[6] = st.point[g].equivalentPlasticStrain;
            q[7] = st.point[g].damage;
            q[8] = st.point[g].failed ? 1.0 : 0.0;
        }
        for (int i = 0; i < 7; ++i) enhanced[7 * k + i] = st.enhanced[i];
        double* in = info + 6 * k;
        in[0] = u.dissipation;
        in[1] = u.iterations;
        in[2] = u.converged ? 1.0 : 0.0;
        in[3] = u.yieldedPoints;
        in[4] = u.failedPoints;
        in[5] = st.torn ? 1.0 : 0.0;
    }
}

// initialisePlasticState's failure strain for an element (regularised by its size).
double ref_hex_failure_strain(const double nodes[24], double uniform, double fracture) {
    sim::plasticity::Material mat;
    mat.failure.uniformStrain = uniform;
    mat.failure.fractureStrain = fracture;
    sim::solidshell::ElementPlasticState st;
    sim::solidshell::initialisePlasticState(nodes, mat, st);
    return st.failureStrain;
}

// makePlateMesh counts, then arrays; uniformPressureLoad on it.
void ref_hex_plate_counts(double lx, double ly, double t, int nx, int ny, int nz, int64_t* nodes,
                          int64_t* elements) {
    const sim::solidshell::HexMesh m = sim::solidshell::makePlateMesh(lx, ly, t, nx, ny, nz);
    *nodes = static_cast<int64_t>(m.nodeCount());
    *elements = static_cast<int64_t>(m.elementCount());
}

void ref_hex_plate(double lx, double ly, double t, int nx, int ny, int nz, double pressure,
                   double* position, uint32_t* index, double* load) {
    const sim::solidshell::HexMesh m = sim::solidshell::makePlateMesh(lx, ly, t, nx, ny, nz);
    std::memcpy(position, m.position.data(), m.position.size() * sizeof(double));
    std::memcpy(index, m.index.data(), m.index.size() * sizeof(uint32_t));
    const std::vector<double> f = sim::solidshell::uniformPressureLoad(m, pressure);
    std::memcpy(load, f.data(), f.size() * sizeof(double));
}

double ref_hex_critical_timestep(const double nodes[24], double E, double nu, double rho,
                                 double safety) {
    return sim::solidshell::criticalTimestep(nodes, structuralOf(E, nu, rho),
                                             sim::solidshell::Formulation::SolidShell, safety);
}

}  // extern "C"
