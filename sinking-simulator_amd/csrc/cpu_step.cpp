// cpu_step.cpp -- the host FP32 explicit step of sim::fem (engine/sim/fem.hpp:82-86):
// the per-tet nodal forces and one symplectic-Euler step over the caller's TetMesh arrays.
//
// The element is the device element pass's: the deformation gradient is taken as
// F = I + G, G = Du Dm^-1 from corner displacements u = x - X when the rest coordinates
// are known (else G = Ds Dm^-1 - I from the positions), the rotation R = I + W by four
// Higham steps R <- (R + R^-T)/2 in perturbation form, the co-rotated small strain
// sym(W^T + G + W^T G), Hooke's law, and the forces -V (I + W) sigma Dm^-T of corners
// 1..3 with corner 0 taking minus their sum. No entry is formed as 1 + (small) and
// cancelled again, so an FP32 step of a long bar keeps its nanometre increments.
//
// The node pass is the reference's symplectic Euler (fem.cpp:196-214): a pinned node
// holds its position with zero velocity; otherwise f = sum of its corner forces in CSR
// order minus g m z, v <- (v + f dt / m) damping, x <- x + v dt. Single-threaded, as
// the reference's; the gather order is the CSR's, so the result is deterministic.
#include "../../include/shipfem.h"

#include <cmath>

namespace {

// Forces of one tet: G (row-major), Dm^-1 as B (row-major), rest volume V.
void elementForces(const float G[9], const float B[9], float V, float lambda, float mu, float f[12]) {
    // Higham's step on R = I + W: with a = adj(W), t = tr W, i2 = tr a, det(I + W) =
    // 1 + t + i2 + det W and inv(I + W) - I = (a - W - (i2 + det W) I) / det(I + W);
    // then W <- (W + (inv(I + W) - I)^T) / 2. A singular I + W takes inv = 0 (the
    // reference's |det| > 1e-30 guard), i.e. inv - I = -I.
    float W[9];
    for (int k = 0; k < 9; ++k) W[k] = G[k];
    for (int it = 0; it < 4; ++it) {
        const float a[9] = {W[4] * W[8] - W[5] * W[7], W[2] * W[7] - W[1] * W[8], W[1] * W[5] - W[2] * W[4],
                            W[5] * W[6] - W[3] * W[8], W[0] * W[8] - W[2] * W[6], W[2] * W[3] - W[0] * W[5],
                            W[3] * W[7] - W[4] * W[6], W[1] * W[6] - W[0] * W[7], W[0] * W[4] - W[1] * W[3]};
        const float t = W[0] + W[4] + W[8];
        const float i2 = a[0] + a[4] + a[8];
        const float detW = W[0] * a[0] + W[1] * a[3] + W[2] * a[6];
        const float den = 1.0f + (t + (i2 + detW));
        float E[9];  // inv(I + W) - I
        if (std::fabs(den) > 1e-30f) {
            const float r = 1.0f / den, sh = i2 + detW;
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) E[i * 3 + j] = (a[i * 3 + j] - W[i * 3 + j] - (i == j ? sh : 0.0f)) * r;
        } else {
            for (int k = 0; k < 9; ++k) E[k] = (k % 4 == 0) ? -1.0f : 0.0f;
        }
        float Wn[9];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) Wn[i * 3 + j] = 0.5f * (W[i * 3 + j] + E[j * 3 + i]);
        for (int k = 0; k < 9; ++k) W[k] = Wn[k];
    }
    // R^T F - I = W^T + G + W^T G
    float A[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            float wg = 0.0f;
            for (int k = 0; k < 3; ++k) wg += W[k * 3 + i] * G[k * 3 + j];
            A[i * 3 + j] = W[j * 3 + i] + G[i * 3 + j] + wg;
        }
    const float e00 = A[0], e11 = A[4], e22 = A[8];
    const float e01 = 0.5f * (A[1] + A[3]), e12 = 0.5f * (A[5] + A[7]), e02 = 0.5f * (A[2] + A[6]);
    const float lt = lambda * (e00 + e11 + e22), m2 = 2.0f * mu;
    const float S[9] = {m2 * e00 + lt, m2 * e01,      m2 * e02,
                        m2 * e01,      m2 * e11 + lt, m2 * e12,
                        m2 * e02,      m2 * e12,      m2 * e22 + lt};
    // P = (I + W) sigma; corner c = 1..3 gets -V (P Dm^-T)(:, c - 1), corner 0 minus their sum
    float P[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            float ws = 0.0f;
            for (int k = 0; k < 3; ++k) ws += W[i * 3 + k] * S[k * 3 + j];
            P[i * 3 + j] = S[i * 3 + j] + ws;
        }
    for (int c = 0; c < 3; ++c)
        for (int a = 0; a < 3; ++a) {
            float h = 0.0f;
            for (int k = 0; k < 3; ++k) h += P[a * 3 + k] * B[c * 3 + k];
            f[(c + 1) * 3 + a] = -V * h;
        }
    for (int a = 0; a < 3; ++a) f[a] = -(f[3 + a] + f[6 + a] + f[9 + a]);
}

// G of tet `t` from the caller's arrays; restInv is column-major (fem.cpp:10-16).
void gradient(const float* position, const double* rest, const uint32_t* index, const float* restInv,
              uint64_t t, float G[9], float B[9]) {
    const uint32_t* v = index + 4 * t;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) B[r * 3 + c] = restInv[9 * t + 3 * c + r];
    float D[9];  // columns: corner c + 1 relative to corner 0
    if (rest) {
        float u[4][3];
        for (int q = 0; q < 4; ++q)
            for (int a = 0; a < 3; ++a)
                u[q][a] = static_cast<float>(static_cast<double>(position[3ull * v[q] + a]) - rest[3ull * v[q] + a]);
        for (int a = 0; a < 3; ++a)
            for (int c = 0; c < 3; ++c) D[a * 3 + c] = u[c + 1][a] - u[0][a];
    } else {
        for (int a = 0; a < 3; ++a)
            for (int c = 0; c < 3; ++c) D[a * 3 + c] = position[3ull * v[c + 1] + a] - position[3ull * v[0] + a];
    }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            float s = 0.0f;
            for (int k = 0; k < 3; ++k) s += D[i * 3 + k] * B[k * 3 + j];
            G[i * 3 + j] = s - (rest ? 0.0f : (i == j ? 1.0f : 0.0f));
        }
}

}  // namespace

extern "C" {

shipfem_status shipfem_cpu_tet_forces(const float* position, const double* rest_position,
                                      const uint32_t* index, const float* rest_inv,
                                      const float* rest_volume, uint64_t tet, float lambda, float mu,
                                      float out[12]) {
    if (!position || !index || !rest_inv || !rest_volume || !out) return SHIPFEM_INVALID_ARG;
    float G[9], B[9];
    gradient(position, rest_position, index, rest_inv, tet, G, B);
    elementForces(G, B, rest_volume[tet], lambda, mu, out);
    return SHIPFEM_OK;
}

shipfem_status shipfem_cpu_step(uint32_t node_count, uint32_t tet_count, float* position, float* velocity,
                                const float* mass, const uint32_t* fixed, const uint32_t* index,
                                const float* rest_inv, const float* rest_volume, float* tet_force,
                                const uint32_t* adjacency_offset, const uint32_t* adjacency_entry,
                                const double* rest_position, float lambda, float mu, float dt,
                                float gravity, float damping) {
    if (!position || !velocity || !mass || !index || !rest_inv || !rest_volume || !tet_force ||
        !adjacency_offset || !adjacency_entry)
        return SHIPFEM_INVALID_ARG;
    if (adjacency_offset[node_count] > 4ull * tet_count) return SHIPFEM_INVALID_ARG;
    for (uint64_t t = 0; t < tet_count; ++t) {
        float G[9], B[9];
        gradient(position, rest_position, index, rest_inv, t, G, B);
        elementForces(G, B, rest_volume[t], lambda, mu, tet_force + 12 * t);
    }
    for (uint64_t n = 0; n < node_count; ++n) {
        float* v = velocity + 3 * n;
        if (fixed && fixed[n]) {
            v[0] = v[1] = v[2] = 0.0f;
            continue;
        }
        float f[3] = {0.0f, 0.0f, -gravity * mass[n]};
        for (uint32_t q = adjacency_offset[n]; q < adjacency_offset[n + 1]; ++q) {
            const uint32_t x = adjacency_entry[q];  // tet * 4 + corner
            const float* s = tet_force + 12ull * (x >> 2) + 3 * (x & 3u);
            for (int a = 0; a < 3; ++a) f[a] += s[a];
        }
        const float step = dt / mass[n];
        for (int a = 0; a < 3; ++a) {
            v[a] = (v[a] + f[a] * step) * damping;
            position[3 * n + a] += v[a] * dt;
        }
    }
    return SHIPFEM_OK;
}

}  // extern "C"
