// kernels.hip -- the explicit tet step on gfx950 (wave64).
//
// Per step, two passes over a tile table (tiles never span cases):
//   element pass: one lane per tet. Gathers the four nodes' displacement (float4),
//     forms F = I + G with G = Du Dm^-1, takes the rotation by Higham's iteration in
//     perturbation form R = I + W (four iterations, as fem.cpp:62-69, the last two
//     skipped once the second has converged below FP32 resolution of the strain),
//     the co-rotated strain eps = sym(W^T + G + W^T G) (no cancellation against I in
//     FP32), Hooke or a J2 radial return, P = R sigma, and the four nodal forces
//     -V P Dm^-T (fem.cpp:73-117). The slots variant writes them to the tet's own
//     48-byte slot; the tiled variant sums them per tile-local node in LDS.
//   node pass: one lane per node. Gathers the node's slots (or tile partials) in a
//     fixed order from a wave-transposed slab (coalesced index loads), adds ramped dead load,
//     gravity and sphere penalty contact, applies pins and the prescribed drive,
//     and advances v and u with Kahan compensation (fem.cpp:187-215, zone.cpp:1094-1200).
// No atomics anywhere: energy terms are FP64 per-wave sums (xor butterfly, the same
// order in every lane) accumulated into per-wave slots, summed in a fixed order by
// the reduce kernel on sample steps.
#include "shipfem_dev.hpp"
#include "hex_plastic.hpp"
#include "omega.hpp"

#include <hip/hip_ext.h>

namespace shipfem {
namespace {

// Plain launch, or with start/stop events carried on the dispatch itself (the events
// then bracket exactly the kernel's execution, with no marker packets in between).
#define SHIPFEM_LAUNCH(timing, kern, grid, block, lds, stream, ...)                           \
    do {                                                                                      \
        if (timing)                                                                           \
            hipExtLaunchKernelGGL(kern, grid, block, lds, stream, (timing)->start,             \
                                  (timing)->stop, 0, __VA_ARGS__);                             \
        else                                                                                  \
            hipLaunchKernelGGL(kern, grid, block, lds, stream, __VA_ARGS__);                   \
    } while (0)

__device__ __forceinline__ int remap_block(int b, int nb, int xcd) {
    if (!xcd) return b;
    // workgroups are dispatched round-robin over the 8 XCDs; give XCD x the
    // contiguous tile range [x q + min(x, r), ...) so its L2 sees neighbouring tets
    const int q = nb >> 3, r = nb & 7;
    const int x = b & 7, j = b >> 3;
    return x * q + (x < r ? x : r) + j;
}

// This block's tile: from the table, or with no table (a plan of one case) tile b of case 0
// starting at item b * block, which spares the kernel one dependent load before its first.
__device__ __forceinline__ Tile tile_of(const Tile* tiles, int block, int xcd) {
    const int b = remap_block(blockIdx.x, gridDim.x, xcd);
    if (!tiles) return Tile{0, static_cast<uint32_t>(b) * static_cast<uint32_t>(block)};
    return tiles[b];
}

// A load the wave issues on the scalar path (through the constant cache). Only for
// wave-uniform addresses of data that no lane writes during the kernel; the cache is
// invalidated at every kernel start, so values written by earlier launches are seen.
template <typename T>
__device__ __forceinline__ T ld_uniform(const T* p) {
    return *reinterpret_cast<const __attribute__((address_space(4))) T*>(reinterpret_cast<uintptr_t>(p));
}

// The same pointer in the global address space: its loads are vector-memory only (not
// flat), so the compiler can wait for one of several loads in flight.
template <typename T>
__device__ __forceinline__ const __attribute__((address_space(1))) T* gmem(const T* p) {
    return reinterpret_cast<const __attribute__((address_space(1))) T*>(reinterpret_cast<uintptr_t>(p));
}

// A load at a wave-uniform base plus a 32-bit byte offset per lane: the address stays in
// the load's scalar base + vector offset form, with no 64-bit address arithmetic per lane.
template <typename T>
__device__ __forceinline__ T ld_at(const T* base, uint32_t byte_offset) {
    return *gmem(reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + byte_offset));
}

// A wave-uniform pointer pinned in scalar registers: the compiler may not rewrite a set of
// such bases into one per-lane base plus running 64-bit increments.
template <typename T>
__device__ __forceinline__ const T* sgpr_base(const T* p) {
    uint64_t v = reinterpret_cast<uint64_t>(p);
    asm volatile("" : "+s"(v));
    return reinterpret_cast<const T*>(v);
}

// Vector-typed loads through the global address space (HIP's vector structs cannot be
// copied from an address-space-qualified reference, their native vectors can).
__device__ __forceinline__ float4 gload(const float4* p) {
    typedef float v4 __attribute__((ext_vector_type(4)));
    const v4 x = *gmem(reinterpret_cast<const v4*>(p));
    return make_float4(x.x, x.y, x.z, x.w);
}
__device__ __forceinline__ uint2 gload(const uint2* p) {
    typedef uint32_t v2 __attribute__((ext_vector_type(2)));
    const v2 x = *gmem(reinterpret_cast<const v2*>(p));
    return make_uint2(x.x, x.y);
}
__device__ __forceinline__ uint4 gload(const uint4* p) {
    typedef uint32_t v4 __attribute__((ext_vector_type(4)));
    const v4 x = *gmem(reinterpret_cast<const v4*>(p));
    return make_uint4(x.x, x.y, x.z, x.w);
}

__device__ __forceinline__ double wave_sum(double s) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    return s;
}

__device__ __forceinline__ double wave_max(double s) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s = fmax(s, __shfl_xor(s, off, 64));
    return s;
}

// ---------------------------------------------------------------- element math
// Forces of one tet from its corner displacements p0..p3, Dm^-1 (B, row-major) and
// rest volume V at step n. PLASTIC cases read and update the tet's J2 state. Adds
// V sigma:d eps_p to diss and, on sample steps, the strain energy 1/2 V sigma:(eps - eps_p)
// to se. With damage, a tet that tears releases its stored energy into fract and carries
// no stress from then on (plasticity.cpp:326, 457-477).
template <bool PLASTIC, bool DAMAGE>
__device__ __forceinline__ void element_force(const CaseDev& c, uint32_t e, int64_t n,
                                              const float4& p0,
                                              const float4& p1, const float4& p2, const float4& p3,
                                              const float B[9], float V, int sample, float f[12],
                                              double& diss, double& se, double& fract,
                                              bool& tore) {
    tore = false;
    // eqps grows with every plastic increment, so a +0 eqps marks a tet that has never
    // yielded: its plastic strain and damage are zero and are not read. (A state handed
    // over from elsewhere with plastic strain but zero eqps carries -0, which reads.)
    const bool plastic = PLASTIC && (c.features & kFeatPlastic);
    const uint32_t e4 = 4u * e;  // a plane's byte offset (Tpad < 2^30)
    const float eqps0 = plastic ? ld_at(c.pstate + 6ull * c.Tpad, e4) : 0.0f;
    const bool history = plastic && __float_as_uint(eqps0) != 0u;
    if (DAMAGE && (c.features & kFeatDamage) && history && ld_at(c.pstate + 7ull * c.Tpad, e4) >= 1.0f) {
        // torn: no stress, no strain energy, history frozen
#pragma unroll
        for (int k = 0; k < 12; ++k) f[k] = 0.0f;
        return;
    }
    // Du: columns are corner displacements relative to corner 0
    const float D[9] = {p1.x - p0.x, p2.x - p0.x, p3.x - p0.x,
                        p1.y - p0.y, p2.y - p0.y, p3.y - p0.y,
                        p1.z - p0.z, p2.z - p0.z, p3.z - p0.z};
    float G[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            G[i * 3 + j] = D[i * 3 + 0] * B[0 * 3 + j] + D[i * 3 + 1] * B[1 * 3 + j] +
                           D[i * 3 + 2] * B[2 * 3 + j];

    // R = I + W by R <- (R + R^-T)/2. With adj(W) the adjugate, t = tr W, i2 = tr adj(W):
    // det(I+W) = 1 + t + i2 + det W and inv(I+W) - I = (adj(W) - W - (i2 + det W) I)/det(I+W),
    // so no entry is formed as 1 + (small) and cancelled again.
    float W[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) W[k] = G[k];
#pragma unroll
    for (int it = 0; it < 4; ++it) {
        const float a[9] = {W[4] * W[8] - W[5] * W[7], W[2] * W[7] - W[1] * W[8], W[1] * W[5] - W[2] * W[4],
                            W[5] * W[6] - W[3] * W[8], W[0] * W[8] - W[2] * W[6], W[2] * W[3] - W[0] * W[5],
                            W[3] * W[7] - W[4] * W[6], W[1] * W[6] - W[0] * W[7], W[0] * W[4] - W[1] * W[3]};
        const float t = W[0] + W[4] + W[8];
        const float i2 = a[0] + a[4] + a[8];
        const float dW = W[0] * a[0] + W[1] * a[3] + W[2] * a[6];
        const float q = t + (i2 + dW);  // det(I + W) - 1
        const float den = 1.0f + q;
        const bool ok = fabsf(den) > 1e-30f;  // the reference's singular guard
        // hardware reciprocal (<= 1 ulp): Newton's iteration corrects all but the last
        // iteration's rounding, which this changes by at most an ulp
        const float id = ok ? __builtin_amdgcn_rcpf(den) : 0.0f;
        // W' = (W + E^T)/2 with E = (adj W - W - sh I)/den, or E = -I when singular.
        // fma(1/2, W, (id/2) x) rounds exactly as (W + id x)/2: halving is exact.
        const float hid = 0.5f * id, hnd = ok ? 0.0f : 0.5f;
        const float sh = i2 + dW;
        float Wn[9];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                const float x = a[j * 3 + i] - W[j * 3 + i];
                Wn[i * 3 + j] = i == j ? fmaf(0.5f, W[i * 3 + j], fmaf(hid, x - sh, -hnd))
                                       : fmaf(0.5f, W[i * 3 + j], hid * x);
            }
#pragma unroll
        for (int k = 0; k < 9; ++k) W[k] = Wn[k];
        // Quadratic convergence: after the second step the stretch error d left in
        // R = I + W satisfies 0 <= d_i and sum d_i ~ q, and the next step leaves d^2/2.
        // With |q| < 2^-18 the two remaining steps would move the co-rotated strain by
        // less than 2^-37, below FP32 resolution of any strain they could act on.
        if (it == 1 && fabsf(q) < 0x1p-18f) break;
    }

    // co-rotated small strain: R^T F - I = W^T + G + W^T G
    float A[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            A[i * 3 + j] = W[j * 3 + i] + G[i * 3 + j] +
                           (W[0 * 3 + i] * G[0 * 3 + j] + W[1 * 3 + i] * G[1 * 3 + j] +
                            W[2 * 3 + i] * G[2 * 3 + j]);
    const float e00 = A[0], e11 = A[4], e22 = A[8];
    const float e01 = 0.5f * (A[1] + A[3]), e12 = 0.5f * (A[5] + A[7]), e02 = 0.5f * (A[2] + A[6]);

    float s00, s11, s22, s01, s12, s02;
    if (plastic) {
        // J2 radial return, linear isotropic hardening (plasticity.cpp:319-455)
        const uint32_t tp = c.Tpad;
        float ep[7];
        ep[6] = eqps0;
        if (history) {
#pragma unroll
            for (int k = 0; k < 6; ++k) ep[k] = ld_at(sgpr_base(c.pstate + (uint64_t)k * tp), e4);
        } else {
#pragma unroll
            for (int k = 0; k < 6; ++k) ep[k] = 0.0f;
        }
        const float st[6] = {e00, e11, e22, 2.0f * e01, 2.0f * e12, 2.0f * e02};
        float el[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) el[k] = st[k] - ep[k];
        const float vol = el[0] + el[1] + el[2];
        const float pr = c.bulk * vol;
        float tr[6];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            tr[k] = 2.0f * c.mu * (el[k] - vol * (1.0f / 3.0f));
            tr[3 + k] = c.mu * el[3 + k];
        }
        // an elastic tet tests the yield surface by squared norms and takes no root
        const float nrm2 = tr[0] * tr[0] + tr[1] * tr[1] + tr[2] * tr[2] +
                           2.0f * (tr[3] * tr[3] + tr[4] * tr[4] + tr[5] * tr[5]);
        const float root23 = 0.81649658092772603f;
        const float radius = root23 * (c.yield + c.hard * fmaxf(ep[6], 0.0f));
        float sv[6];
        if (nrm2 > radius * radius) {
            const float nrm = sqrtf(nrm2);
            const float gamma = (nrm - radius) / (2.0f * c.mu + (2.0f / 3.0f) * c.hard);
            const float inv = 1.0f / nrm;
            float d = 0.0f;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const float n0 = tr[k] * inv, n1 = tr[3 + k] * inv;
                sv[k] = pr + tr[k] - 2.0f * c.mu * gamma * n0;
                sv[3 + k] = tr[3 + k] - 2.0f * c.mu * gamma * n1;
                const float dn = gamma * n0, ds = 2.0f * gamma * n1;
                d += sv[k] * dn + sv[3 + k] * ds;
                ep[k] += dn;
                ep[3 + k] += ds;
            }
            ep[6] += root23 * gamma;
            diss += (double)d * (double)V;
#pragma unroll
            for (int k = 0; k < 7; ++k) c.pstate[k * tp + e] = ep[k];
            if (DAMAGE && (c.features & kFeatDamage)) {
                // d damage = d eqps / (eps_f exp(-s (eta - eta_ref))), none at or below
                // the cutoff; eta = mean stress over von Mises of the returned stress
                const float q = 1.22474487139158905f * (nrm - 2.0f * c.mu * gamma);
                const float eta = q > 0.0f ? pr / q : 0.0f;
                float dmg = history ? c.pstate[7 * tp + e] : 0.0f;
                if (eta > c.fail_ecut)
                    dmg += root23 * gamma * __expf(c.fail_sens * (eta - c.fail_eref)) / c.fail_strain;
                if (dmg >= 1.0f) {
                    dmg = 1.0f;
                    tore = true;
                }
                c.pstate[7 * tp + e] = dmg;
            }
        } else {
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                sv[k] = pr + tr[k];
                sv[3 + k] = tr[3 + k];
            }
        }
        if (DAMAGE && tore) {
            // the stored energy leaves as fracture (the caller adds the half step of
            // work its forces still carry and zeroes them); the world-frame stress
            // R sigma R^T at the tear is kept for the failure plane (found on the host)
            float s = 0.0f;
#pragma unroll
            for (int k = 0; k < 6; ++k) s += sv[k] * (st[k] - ep[k]);
            fract += 0.5 * (double)s * (double)V;
            // row by row: (R S)_i. then (R S R^T)_ij for j >= i, Voigt slot of (i, j)
            const float S3[9] = {sv[0], sv[3], sv[5], sv[3], sv[1], sv[4], sv[5], sv[4], sv[2]};
            constexpr int kSlot[9] = {0, 3, 5, 3, 1, 4, 5, 4, 2};
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                float rs[3];
#pragma unroll
                for (int k = 0; k < 3; ++k)
                    rs[k] = S3[i * 3 + k] + (W[i * 3 + 0] * S3[0 * 3 + k] + W[i * 3 + 1] * S3[1 * 3 + k] +
                                             W[i * 3 + 2] * S3[2 * 3 + k]);
#pragma unroll
                for (int j = i; j < 3; ++j)
                    c.tear_stress[kSlot[i * 3 + j] * tp + e] =
                        rs[j] + (rs[0] * W[j * 3 + 0] + rs[1] * W[j * 3 + 1] + rs[2] * W[j * 3 + 2]);
            }
            c.tear_step[e] = (int32_t)(n + 1);
        }
        if (sample && !(DAMAGE && tore)) {
            float s = 0.0f;
#pragma unroll
            for (int k = 0; k < 6; ++k) s += sv[k] * (st[k] - ep[k]);
            se += 0.5 * (double)s * (double)V;
        }
        s00 = sv[0]; s11 = sv[1]; s22 = sv[2]; s01 = sv[3]; s12 = sv[4]; s02 = sv[5];
    } else {
        const float tr = e00 + e11 + e22;
        const float m2 = 2.0f * c.mu, lt = c.lam * tr;
        s00 = m2 * e00 + lt; s11 = m2 * e11 + lt; s22 = m2 * e22 + lt;
        s01 = m2 * e01; s12 = m2 * e12; s02 = m2 * e02;
        if (sample) {
            const float s = s00 * e00 + s11 * e11 + s22 * e22 + 2.0f * (s01 * e01 + s12 * e12 + s02 * e02);
            se += 0.5 * (double)s * (double)V;
        }
    }
    const float S[9] = {s00, s01, s02, s01, s11, s12, s02, s12, s22};
    // P = (I + W) sigma, H = P Dm^-T, f_c = -V H(:, c-1), f_0 = -(f_1 + f_2 + f_3)
    float P[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            P[i * 3 + j] = S[i * 3 + j] + (W[i * 3 + 0] * S[0 * 3 + j] + W[i * 3 + 1] * S[1 * 3 + j] +
                                           W[i * 3 + 2] * S[2 * 3 + j]);
#pragma unroll
    for (int cc = 0; cc < 3; ++cc)
#pragma unroll
        for (int a = 0; a < 3; ++a)
            f[(cc + 1) * 3 + a] = -V * (P[a * 3 + 0] * B[cc * 3 + 0] + P[a * 3 + 1] * B[cc * 3 + 1] +
                                        P[a * 3 + 2] * B[cc * 3 + 2]);
#pragma unroll
    for (int a = 0; a < 3; ++a) f[a] = -(f[3 + a] + f[6 + a] + f[9 + a]);
}

// Per-wave FP64 sums of the element pass into the case's wave slots. Every lane of the
// wave takes part (dead lanes add zero); a tail block's waves past the case's last wave
// hold only dead lanes and write nothing.
// A tet that tore in this element pass: against the shadow kinetic energy
// 1/2 m v(n-1/2).v(n+1/2) the potential it carried at u_n is U_e + 1/2 f_e.(u_n - u_(n-1))
// with f_e its intact nodal forces, so the half step of work dt/2 f_e.v(n-1/2) joins the
// fracture release before the forces are zeroed (v read as hi - compensation).
__device__ __forceinline__ void release_torn(const CaseDev& c, const uint32_t g[4], float f[12],
                                             double& fract) {
    double half = 0.0;
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const uint32_t k = a * c.Npad + g[q];
            half += (double)f[q * 3 + a] * ((double)c.v[k] - (double)c.cv[k]);
        }
    fract += 0.5 * c.dt * half;
#pragma unroll
    for (int k = 0; k < 12; ++k) f[k] = 0.0f;
}

template <bool PLASTIC, bool DAMAGE>
__device__ __forceinline__ void element_wave_sums(const CaseDev& c, uint32_t e, int sample,
                                                  double diss, double se, double fract) {
    const uint32_t w = e >> 6;
    const bool writer = (threadIdx.x & 63u) == 0 && w < c.elem_waves;
    // A wave none of whose tets yielded (or tore) adds +0 to its running sum, which leaves
    // the sum's bits as they are: it skips the butterfly and the store.
    if (PLASTIC && (c.features & kFeatPlastic) && __ballot(diss != 0.0)) {
        const double s = wave_sum(diss);
        if (writer) c.acc_e[kAccDiss * c.elem_waves + w] += s;
    }
    if (DAMAGE && (c.features & kFeatDamage) && __ballot(fract != 0.0)) {
        const double s = wave_sum(fract);
        if (writer) c.acc_e[kAccFract * c.elem_waves + w] += s;
    }
    if (sample) {
        const double s = wave_sum(se);
        if (writer) c.smp_e[kSmpStrain * c.elem_waves + w] = s;
    }
}

// ---------------------------------------------------------------- element pass, slots
template <int BLOCK, bool PLASTIC, bool DAMAGE>
__global__ __launch_bounds__(BLOCK) void k_elem(const CaseDev* __restrict__ cases,
                                                const Tile* __restrict__ tiles, int ntiles,
                                                int parity, int sample, int xcd) {
    const int b = remap_block(blockIdx.x, gridDim.x, xcd);
    const Tile tile = tiles[b];
    const CaseDev& c = cases[tile.case_index];
    if (c.step[parity] >= c.step_end) return;  // the case has finished: whole block idles

    const uint32_t e = tile.first + threadIdx.x;
    double diss = 0.0, se = 0.0, fract = 0.0;
    if (e < c.T) {
        const uint4 nd = c.conn[e];
        const float4 p0 = c.u4[nd.x], p1 = c.u4[nd.y], p2 = c.u4[nd.z], p3 = c.u4[nd.w];
        float B[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) B[k] = c.dmi[k * c.Tpad + e];
        float f[12];
        bool tore;
        element_force<PLASTIC, DAMAGE>(c, e, c.step[parity], p0, p1, p2, p3, B, c.vol[e], sample, f,
                                       diss, se, fract, tore);
        if (DAMAGE && tore) {
            const uint32_t g[4] = {nd.x, nd.y, nd.z, nd.w};
            release_torn(c, g, f, fract);
        }
        float4* out = reinterpret_cast<float4*>(c.force) + (size_t)e * 3;
        out[0] = make_float4(f[0], f[1], f[2], f[3]);
        out[1] = make_float4(f[4], f[5], f[6], f[7]);
        out[2] = make_float4(f[8], f[9], f[10], f[11]);
    }
    element_wave_sums<PLASTIC, DAMAGE>(c, e, sample, diss, se, fract);
}

// ---------------------------------------------------------------- element pass, hex
// One lane per solid-shell hex: f = -R K (R^T x - X) (solid_shell.cpp:735-766) with K the
// condensed elastic stiffness formed at rest (FP64), R the polar factor of the centre
// deformation gradient. K annihilates rigid translations, so the element works on the
// displacements of corners 1..7 relative to corner 0 and the force on corner 0 is minus
// the sum of the others: only the 21 x 21 block of K that couples corners 1..7 is read
// (231 packed doubles per element, the pass's dominant traffic). With R = I + W and
// corner offsets r_a = u_a - u_0 (each displacement taken as its FP32 value minus its
// Kahan compensation), R^T x_a - X_a - (R^T x_0 - X_0) = r_a + W^T (X_a - X_0 + r_a): every
// operand is small, so nothing cancels at FP32. The product with K and the strain energy
// 1/2 d.Kd are FP64; the nodal forces are written as FP32 slots [T][8][3].
__device__ __forceinline__ void polar_perturbation(float W[9]) {
    // R = I + W by R <- (R + R^-T)/2 in the same perturbation form as the tet pass, six
    // steps: quadratic convergence takes a stretch error below 0.5 under 2^-40 of it
#pragma unroll
    for (int it = 0; it < 6; ++it) {
        const float a[9] = {W[4] * W[8] - W[5] * W[7], W[2] * W[7] - W[1] * W[8], W[1] * W[5] - W[2] * W[4],
                            W[5] * W[6] - W[3] * W[8], W[0] * W[8] - W[2] * W[6], W[2] * W[3] - W[0] * W[5],
                            W[3] * W[7] - W[4] * W[6], W[1] * W[6] - W[0] * W[7], W[0] * W[4] - W[1] * W[3]};
        const float t = W[0] + W[4] + W[8];
        const float i2 = a[0] + a[4] + a[8];
        const float dW = W[0] * a[0] + W[1] * a[3] + W[2] * a[6];
        const float den = 1.0f + (t + (i2 + dW));
        const bool ok = fabsf(den) > 1e-30f;
        const float id = ok ? __builtin_amdgcn_rcpf(den) : 0.0f;
        const float hid = 0.5f * id, hnd = ok ? 0.0f : 0.5f;
        const float sh = i2 + dW;
        float Wn[9];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                const float x = a[j * 3 + i] - W[j * 3 + i];
                Wn[i * 3 + j] = i == j ? fmaf(0.5f, W[i * 3 + j], fmaf(hid, x - sh, -hnd))
                                       : fmaf(0.5f, W[i * 3 + j], hid * x);
            }
#pragma unroll
        for (int k = 0; k < 9; ++k) W[k] = Wn[k];
    }
}

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_elem_hex(const CaseDev* __restrict__ cases,
                                                    const Tile* __restrict__ tiles, int ntiles,
                                                    int parity, int sample, int xcd) {
    const int b = remap_block(blockIdx.x, gridDim.x, xcd);
    const Tile tile = tiles[b];
    const CaseDev& c = cases[tile.case_index];
    if (c.step[parity] >= c.step_end) return;  // the case has finished: whole block idles

    const uint32_t e = tile.first + threadIdx.x;
    double se = 0.0;
    if (e < c.T) {
        const uint32_t tp = c.Tpad, np = c.Npad;
        const uint4 q0 = c.conn[2 * e], q1 = c.conn[2 * e + 1];
        const uint32_t g[8] = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, q1.z, q1.w};
        float uh[8][3], uc[8][3];
#pragma unroll
        for (int a = 0; a < 8; ++a) {
            const float4 p = c.u4[g[a]];
            uh[a][0] = p.x;
            uh[a][1] = p.y;
            uh[a][2] = p.z;
#pragma unroll
            for (int i = 0; i < 3; ++i) uc[a][i] = c.cu[i * np + g[a]];
        }
        // offsets from corner 0: the FP32 difference, then the compensations' difference
        float r[7][3], H[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
        for (int a = 0; a < 7; ++a)
#pragma unroll
            for (int i = 0; i < 3; ++i) r[a][i] = (uh[a + 1][i] - uh[0][i]) - (uc[a + 1][i] - uc[0][i]);
#pragma unroll
        for (int a = 0; a < 7; ++a) {
            const float b0[3] = {c.hgeo[(3 * a) * tp + e], c.hgeo[(3 * a + 1) * tp + e],
                                 c.hgeo[(3 * a + 2) * tp + e]};
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j) H[i * 3 + j] = fmaf(r[a][i], b0[j], H[i * 3 + j]);
        }
        polar_perturbation(H);  // H now holds W, R = I + W
        double d[kHexDof];
#pragma unroll
        for (int a = 0; a < 7; ++a) {
            float y[3];
#pragma unroll
            for (int i = 0; i < 3; ++i) y[i] = c.hgeo[(kHexDof + 3 * a + i) * tp + e] + r[a][i];
#pragma unroll
            for (int i = 0; i < 3; ++i)
                d[3 * a + i] = (double)(r[a][i] + (H[0 * 3 + i] * y[0] + H[1 * 3 + i] * y[1] +
                                                   H[2 * 3 + i] * y[2]));
        }
        double kd[kHexDof];
#pragma unroll
        for (int i = 0; i < kHexDof; ++i) kd[i] = 0.0;
        const double* kp = c.hk + e;
#pragma unroll
        for (int i = 0; i < kHexDof; ++i)
#pragma unroll
            for (int j = i; j < kHexDof; ++j) {
                const double k = kp[(size_t)(i * kHexDof - i * (i - 1) / 2 + j - i) * tp];
                kd[i] = fma(k, d[j], kd[i]);
                if (j != i) kd[j] = fma(k, d[i], kd[j]);
            }
        if (sample) {
#pragma unroll
            for (int i = 0; i < kHexDof; ++i) se = fma(d[i], kd[i], se);
            se *= 0.5;
        }
        float f[24];
        double k0[3] = {0.0, 0.0, 0.0};
#pragma unroll
        for (int a = 0; a < 7; ++a)
#pragma unroll
            for (int i = 0; i < 3; ++i) k0[i] -= kd[3 * a + i];
#pragma unroll
        for (int a = 0; a < 8; ++a) {
            const float v[3] = {(float)(a == 0 ? k0[0] : kd[3 * a - 3]),
                                (float)(a == 0 ? k0[1] : kd[3 * a - 2]),
                                (float)(a == 0 ? k0[2] : kd[3 * a - 1])};
#pragma unroll
            for (int i = 0; i < 3; ++i)
                f[3 * a + i] = -(v[i] + (H[i * 3 + 0] * v[0] + H[i * 3 + 1] * v[1] + H[i * 3 + 2] * v[2]));
        }
        float4* out = reinterpret_cast<float4*>(c.force) + (size_t)e * 6;
#pragma unroll
        for (int k = 0; k < 6; ++k) out[k] = make_float4(f[4 * k], f[4 * k + 1], f[4 * k + 2], f[4 * k + 3]);
    }
    if (sample) {
        const double s = wave_sum(se);
        const uint32_t w = e >> 6;
        if ((threadIdx.x & 63u) == 0 && w < c.elem_waves) c.smp_e[kSmpStrain * c.elem_waves + w] = s;
    }
}

// ---------------------------------------------------------------- element pass, plastic hex
// One lane per solid-shell hex, FP64: elementPlasticUpdate (solid_shell.cpp:826-1029) with
// the strain operator formed from the element's rest offsets (hex_plastic.hpp). The lane
// forms the co-rotated corner offsets d in FP64 from the compensated displacement, runs the
// update on the element's point history in place, and writes -R times the internal force.
// The elastic pressure kappa tr(eps) of a thin element is where FP32 offsets would show: a
// 12 mm element displaced 7.5 mm carries 4e-8 of strain in them, 7 kPa at kappa = 172 GPa.

// polar_perturbation in FP64
__device__ __forceinline__ void polar_perturbation_d(double W[9]) {
#pragma unroll
    for (int it = 0; it < 6; ++it) {
        const double a[9] = {W[4] * W[8] - W[5] * W[7], W[2] * W[7] - W[1] * W[8], W[1] * W[5] - W[2] * W[4],
                             W[5] * W[6] - W[3] * W[8], W[0] * W[8] - W[2] * W[6], W[2] * W[3] - W[0] * W[5],
                             W[3] * W[7] - W[4] * W[6], W[1] * W[6] - W[0] * W[7], W[0] * W[4] - W[1] * W[3]};
        const double t = W[0] + W[4] + W[8];
        const double i2 = a[0] + a[4] + a[8];
        const double dW = W[0] * a[0] + W[1] * a[3] + W[2] * a[6];
        const double q = t + (i2 + dW);  // det(I + W) - 1
        const double den = 1.0 + q;
        const bool ok = fabs(den) > 1e-300;
        const double hid = ok ? 0.5 * hexp::recip(den) : 0.0, hnd = ok ? 0.0 : 0.5;
        const double sh = i2 + dW;
        double Wn[9];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                const double x = a[j * 3 + i] - W[j * 3 + i];
                Wn[i * 3 + j] = i == j ? fma(0.5, W[i * 3 + j], fma(hid, x - sh, -hnd))
                                       : fma(0.5, W[i * 3 + j], hid * x);
            }
#pragma unroll
        for (int k = 0; k < 9; ++k) W[k] = Wn[k];
        // From the second step on, the stretch error d left in I + W is positive
        // semi-definite with trace ~ det(I + W) - 1, and the next step leaves d^2/2: below
        // 2^-30 the step just taken has brought R to FP64 resolution.
        if (it >= 1 && fabs(q) < 0x1p-30) break;
    }
}

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_elem_hex_plastic(const CaseDev* __restrict__ cases,
                                                            const Tile* __restrict__ tiles,
                                                            int ntiles, int parity, int sample,
                                                            int xcd) {
    const int b = remap_block(blockIdx.x, gridDim.x, xcd);
    const Tile tile = tiles[b];
    const CaseDev& c = cases[tile.case_index];

    // each lane's base strains of its element's eight points and its element's rest
    // coefficients, lane-interleaved
    __shared__ double sh_eu[48 * BLOCK];
    __shared__ double sh_cx[kHexDof * BLOCK];
    __shared__ double sh_w[9 * BLOCK];  // the element's rotation I + W
    const uint32_t e = tile.first + threadIdx.x;
    const bool live = e < c.T;
    const uint32_t tp = c.Tpad, np = c.Npad;
    // One wave per SIMD leaves nothing to hide a load chain behind, so every load that does
    // not depend on another is issued first: the connectivity (which heads the gather), each
    // point's eqps and the centre gradients; then the corner gather; the step counter is
    // tested only then. Indices are clamped into range; a dead lane's values are discarded.
    const uint32_t ec = live ? e : c.T - 1;
    const uint4 q0 = gload(c.conn + 2 * ec), q1 = gload(c.conn + 2 * ec + 1);
    double eqps[8];
#pragma unroll
    for (int g = 0; g < 8; ++g) eqps[g] = *gmem(c.hpoint + (8 * g + 6) * tp + ec);
    double b0[7][3];
#pragma unroll
    for (int a = 0; a < 7; ++a)
#pragma unroll
        for (int j = 0; j < 3; ++j) b0[a][j] = *gmem(c.hgeod + (3 * a + j) * tp + ec);
    double cX[kHexDof], alpha0[7];
#pragma unroll
    for (int j = 0; j < kHexDof; ++j) cX[j] = *gmem(c.hgeod + (kHexDof + j) * tp + ec);
#pragma unroll
    for (int k = 0; k < 7; ++k) alpha0[k] = *gmem(c.henh + k * tp + ec);
    const double failE = c.hfail ? *gmem(c.hfail + ec) : c.fail_d;
    const uint32_t gi[8] = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, q1.z, q1.w};
    double ux[8][3];
#pragma unroll
    for (int a = 0; a < 8; ++a) {
        const float4 pv = gload(c.u4 + gi[a]);
        ux[a][0] = (double)pv.x - (double)*gmem(c.cu + gi[a]);
        ux[a][1] = (double)pv.y - (double)*gmem(c.cu + np + gi[a]);
        ux[a][2] = (double)pv.z - (double)*gmem(c.cu + 2 * np + gi[a]);
    }
    if (c.step[parity] >= c.step_end) return;  // the case has finished: whole block idles

    double diss = 0.0, se = 0.0, fract = 0.0;
    if (live) {
        const bool damage = (c.features & kFeatDamage) != 0;
        float4* out = reinterpret_cast<float4*>(c.force) + (size_t)e * 6;
        const hexp::Start start = hexp::startOf(eqps, c.hpoint + e, tp);
        if (damage && start.failed == 8) {
            // a torn element carries nothing and is not updated (zone.cpp:944-955)
#pragma unroll
            for (int k = 0; k < 6; ++k) out[k] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        } else {
            double r[7][3], W[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
            for (int a = 0; a < 7; ++a)
#pragma unroll
                for (int i = 0; i < 3; ++i) r[a][i] = ux[a + 1][i] - ux[0][i];
#pragma unroll
            for (int a = 0; a < 7; ++a)
#pragma unroll
                for (int i = 0; i < 3; ++i)
#pragma unroll
                    for (int j = 0; j < 3; ++j) W[i * 3 + j] = fma(r[a][i], b0[a][j], W[i * 3 + j]);
            polar_perturbation_d(W);
            // the co-rotated displacement d = r + W^T (X + r), taken to basis coefficients
            // (the map is linear): cd = cr + W^T (cX + cr)
            double cd[kHexDof];
            hexp::basis(&r[0][0], cd);
#pragma unroll
            for (int mm = 0; mm < 7; ++mm) {
                const double y[3] = {cX[3 * mm] + cd[3 * mm], cX[3 * mm + 1] + cd[3 * mm + 1],
                                     cX[3 * mm + 2] + cd[3 * mm + 2]};
#pragma unroll
                for (int i = 0; i < 3; ++i)
                    cd[3 * mm + i] += W[0 * 3 + i] * y[0] + W[1 * 3 + i] * y[1] + W[2 * 3 + i] * y[2];
            }
            const hexp::Material m{c.mu_d, c.lam_d + (2.0 / 3.0) * c.mu_d, c.yield_d, c.hard_d,
                                   failE, c.sens_d, c.eref_d, c.ecut_d, damage,
                                   c.hret_d[0], c.hret_d[1], c.hret_d[2], c.hret_d[3], c.hret_d[4]};
            double fi[kHexDof];
#pragma unroll
            for (int j = 0; j < kHexDof; ++j) sh_cx[j * BLOCK + threadIdx.x] = cX[j];
#pragma unroll
            for (int j = 0; j < 9; ++j) sh_w[j * BLOCK + threadIdx.x] = W[j];
            const hexp::Result res = hexp::update(m, cX, hexp::Strided{sh_cx + threadIdx.x, BLOCK}, cd, c.hpoint + e,
                                                  c.henh + e, tp, sample != 0, fi, sh_eu + threadIdx.x, BLOCK,
                                                  start, alpha0);
            diss = res.dissipation;
            fract = res.released;
            se = res.strain;
            if (damage && res.failedEnd == 8 && res.failedStart < 8)
                c.tear_step[e] = (int32_t)(c.step[parity] + 1);
            double k0[3] = {0.0, 0.0, 0.0};
#pragma unroll
            for (int a = 0; a < 7; ++a)
#pragma unroll
                for (int i = 0; i < 3; ++i) k0[i] -= fi[3 * a + i];
            float f[24];
#pragma unroll
            for (int j = 0; j < 9; ++j) W[j] = sh_w[j * BLOCK + threadIdx.x];
#pragma unroll
            for (int a = 0; a < 8; ++a) {
                const double v[3] = {a == 0 ? k0[0] : fi[3 * a - 3], a == 0 ? k0[1] : fi[3 * a - 2],
                                     a == 0 ? k0[2] : fi[3 * a - 1]};
#pragma unroll
                for (int i = 0; i < 3; ++i)
                    f[3 * a + i] = (float)-(v[i] + (W[i * 3 + 0] * v[0] + W[i * 3 + 1] * v[1] + W[i * 3 + 2] * v[2]));
            }
#pragma unroll
            for (int k = 0; k < 6; ++k) out[k] = make_float4(f[4 * k], f[4 * k + 1], f[4 * k + 2], f[4 * k + 3]);
        }
    }
    const uint32_t w = e >> 6;
    const bool writer = (threadIdx.x & 63u) == 0 && w < c.elem_waves;
    // a wave with no dissipation (or fracture) adds +0 to its running sum, which leaves the
    // sum's bits as they are: it skips the butterfly, the read and the store
    if (__ballot(diss != 0.0)) {
        const double s = wave_sum(diss);
        if (writer) c.acc_e[kAccDiss * c.elem_waves + w] += s;
    }
    if ((c.features & kFeatDamage) && __ballot(fract != 0.0)) {
        const double s = wave_sum(fract);
        if (writer) c.acc_e[kAccFract * c.elem_waves + w] += s;
    }
    if (sample) {
        const double s = wave_sum(se);
        if (writer) c.smp_e[kSmpStrain * c.elem_waves + w] = s;
    }
}

// ---------------------------------------------------------------- element pass, tiled
// One block per 256-tet tile. The tile's local nodes are numbered in ascending node id
// and each has a contiguous run of 4-kTile corner slots, in (tet, corner) order,
// starting at its offset lofs[l]. A tet's connectivity word carries, per corner, the
// local node (10 bits) and the corner's rank in that node's run (6 bits), so a tet
// lane writes each corner force straight into its slot and a node lane sums one
// contiguous run: the tile needs no separate node->corner map.
// LDS: corner-slot forces as an (x, y) float2 plane and a z plane of 4 kTile slots each (an
// 8-byte and a 4-byte access per slot cost less LDS time than three 4-byte ones), the run offsets
// [lmax + 1] (16-bit), then the tile's local nodes' displacement float4[lmax]. Every
// global load of the staging is issued before the first barrier, so its gathers share one
// latency. A tet that forms its Dm^-1 reads its corners' rest positions through the vector
// L1 (the tile's few KB of them stay cached) rather than from a staged LDS copy: it keeps
// the LDS traffic, most of it the tets' scattered accesses, to the displacements and slots.
constexpr uint32_t kSlots4 = 4 * kTile;  // corner slots of a tile
__host__ __device__ constexpr size_t tiled_lds_offsets_offset() { return 3 * kSlots4 * sizeof(float); }
__host__ __device__ constexpr size_t tiled_lds_nodes_offset(uint32_t lmax) {
    return tiled_lds_offsets_offset() + ((lmax + 1) * sizeof(uint16_t) + 15) / 16 * 16;
}

// The element pass of the 256-tet tile from `first`, at the case's step `n`. FORMED: each
// tet forms its Dm^-1 and V from the rest positions of its corners (the tile's, about its
// centroid); else it reads the stored planes. Its 256 threads (lane `lid`) all reach its two barriers; with `active`
// false (no tile for them) they do nothing else. With `nptr` the step is read from there
// instead, once the staging loads are in flight, and a finished case (step_end reached)
// returns before its first store.
template <bool PLASTIC, bool DAMAGE, bool FORMED>
__device__ __forceinline__ void tiled_elem_block(const CaseDev& c, uint32_t first, uint32_t lid, bool active,
                                                 int64_t n, uint32_t lmax, int sample, unsigned char* smem,
                                                 const int64_t* nptr = nullptr) {
    float2* sh_fxy = reinterpret_cast<float2*>(smem);                 // corner slots' x, y
    float* sh_fz = reinterpret_cast<float*>(smem + 2 * kSlots4 * sizeof(float));  // and z
    uint16_t* sh_lo = reinterpret_cast<uint16_t*>(smem + tiled_lds_offsets_offset());
    float4* sh_u = reinterpret_cast<float4*>(smem + tiled_lds_nodes_offset(lmax));

    const uint32_t tb = first / kTile;
    uint32_t p0 = 0, L = 1;
    if (active) {
        p0 = ld_uniform(c.tn_off + tb);
        L = ld_uniform(c.tn_off + tb + 1) - p0;
    }
    const uint32_t e = first + lid;
    const bool live = active && e < c.T;

    // Loads in the order their chains need them: the staging gather's node ids head
    // the longest chain (ids, then displacements), then the run offsets, then the
    // streamed per-tet planes, so waiting for the ids never waits for the stream.
    // Indices are clamped into range instead of branched on, so every load is issued
    // unconditionally; values a lane does not own are discarded. A lane stages local
    // nodes lid and lid + kTile; tiles with more (not seen in Kuhn-type meshes) stage the
    // rest afterwards.
    uint2 lc = make_uint2(0u, 0u);
    float B[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, V = 0.0f;
    if (active) {
        const uint32_t la = lid, lb = lid + kTile;
        const auto* lo = gmem(c.lofs + p0 + tb);
        const auto* tn = gmem(c.tnode + p0);
        const uint32_t ga = tn[min(la, L - 1)];
        const uint32_t gb = tn[min(lb, L - 1)];
        const uint32_t oa = lo[min(la, L)];
        const uint32_t ob = lo[min(lb, L)];
        const uint32_t ec = min(e, c.T - 1);
        lc = gload(c.lconn + ec);
        if (!FORMED) {
            const uint32_t ec4 = 4u * ec;  // byte offset into a per-tet plane
#pragma unroll
            for (int k = 0; k < 9; ++k) B[k] = ld_at(sgpr_base(c.dmi + (uint64_t)k * c.Tpad), ec4);
            V = ld_at(c.vol, ec4);
        }
        const float4 ua = gload(c.u4 + ga);
        const float4 ub = gload(c.u4 + gb);
        if (nptr) {
            n = ld_uniform(nptr);
            if (n >= c.step_end) return;  // the case has finished: the whole block idles
        }
        // lanes past the tile's nodes rewrite the last entry with the value it already
        // holds, so no store waits behind a branch
        sh_lo[min(la, L)] = static_cast<uint16_t>(oa);
        sh_lo[min(lb, L)] = static_cast<uint16_t>(ob);
        // a local node's displacement with its run offset in the fourth word: a tet's one
        // 16-byte read per corner gives both (a 12-byte LDS read costs about twice a 16-byte one)
        if (la < L) sh_u[la] = make_float4(ua.x, ua.y, ua.z, __uint_as_float(oa));
        if (lb < L) sh_u[lb] = make_float4(ub.x, ub.y, ub.z, __uint_as_float(ob));
        for (uint32_t l = lid + 2 * kTile; l <= L; l += kTile) sh_lo[l] = lo[l];
        for (uint32_t l = lid + 2 * kTile; l < L; l += kTile) {
            const float4 u = c.u4[c.tnode[p0 + l]];
            sh_u[l] = make_float4(u.x, u.y, u.z, __uint_as_float((uint32_t)lo[l]));
        }
    }
    __syncthreads();

    double diss = 0.0, se = 0.0, fract = 0.0;
    if (live) {
        const uint32_t w4[4] = {lc.x & 0xFFFFu, lc.x >> 16, lc.y & 0xFFFFu, lc.y >> 16};
        uint32_t ln[4], slot[4];
        float4 uq[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            ln[q] = w4[q] & 1023u;
            uq[q] = sh_u[ln[q]];
            slot[q] = __float_as_uint(uq[q].w) + (w4[q] >> 10);
        }
        // Dm^-1 and V from the rest positions: Dm's column k is corner k+1 minus corner 0;
        // the inverse by cofactors, a degenerate tet (|det| <= 1e-30) taking zero
        // (fem.cpp:42-56)
        if (FORMED) {
            const auto* xr = c.trest + p0;
            const float4 x0 = gload(xr + ln[0]), x1 = gload(xr + ln[1]), x2 = gload(xr + ln[2]), x3 = gload(xr + ln[3]);
            const float D[3][3] = {{x1.x - x0.x, x2.x - x0.x, x3.x - x0.x},
                                   {x1.y - x0.y, x2.y - x0.y, x3.y - x0.y},
                                   {x1.z - x0.z, x2.z - x0.z, x3.z - x0.z}};
            const float a00 = D[1][1] * D[2][2] - D[1][2] * D[2][1];
            const float a01 = D[0][2] * D[2][1] - D[0][1] * D[2][2];
            const float a02 = D[0][1] * D[1][2] - D[0][2] * D[1][1];
            const float det = D[0][0] * a00 + D[1][0] * a01 + D[2][0] * a02;
            const float id = fabsf(det) > 1e-30f ? __builtin_amdgcn_rcpf(det) : 0.0f;
            B[0] = a00 * id;
            B[1] = a01 * id;
            B[2] = a02 * id;
            B[3] = (D[1][2] * D[2][0] - D[1][0] * D[2][2]) * id;
            B[4] = (D[0][0] * D[2][2] - D[0][2] * D[2][0]) * id;
            B[5] = (D[0][2] * D[1][0] - D[0][0] * D[1][2]) * id;
            B[6] = (D[1][0] * D[2][1] - D[1][1] * D[2][0]) * id;
            B[7] = (D[0][1] * D[2][0] - D[0][0] * D[2][1]) * id;
            B[8] = (D[0][0] * D[1][1] - D[0][1] * D[1][0]) * id;
            V = fabsf(det) * (1.0f / 6.0f);
        }
        float f[12];
        bool tore;
        element_force<PLASTIC, DAMAGE>(c, e, n, uq[0], uq[1], uq[2], uq[3], B, V, sample, f, diss, se, fract,
                                       tore);
        if (DAMAGE && tore) {
            const uint32_t g[4] = {c.tnode[p0 + ln[0]], c.tnode[p0 + ln[1]], c.tnode[p0 + ln[2]],
                                   c.tnode[p0 + ln[3]]};
            release_torn(c, g, f, fract);
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            sh_fxy[slot[q]] = make_float2(f[q * 3], f[q * 3 + 1]);
            sh_fz[slot[q]] = f[q * 3 + 2];
        }
    }
    if (active) element_wave_sums<PLASTIC, DAMAGE>(c, e, sample, diss, se, fract);
    __syncthreads();

    // one lane per local node sums its run of corner slots in order, four per round
    // (all reads of a round in flight together)
    for (uint32_t l = lid; active && l < L; l += kTile) {
        const uint32_t q0 = sh_lo[l], q1 = sh_lo[l + 1];
        float fx = 0.0f, fy = 0.0f, fz = 0.0f;
        for (uint32_t q = q0; q < q1; q += 4) {
            float g[4][3];
#pragma unroll
            for (uint32_t k = 0; k < 4; ++k) {
                const uint32_t s = min(q + k, q1 - 1);
                const float2 xy = sh_fxy[s];
                g[k][0] = xy.x;
                g[k][1] = xy.y;
                g[k][2] = sh_fz[s];
            }
#pragma unroll
            for (uint32_t k = 0; k < 4; ++k)
                if (q + k < q1) {
                    fx += g[k][0];
                    fy += g[k][1];
                    fz += g[k][2];
                }
        }
        c.partial[p0 + l] = make_float4(fx, fy, fz, 0.0f);
    }
}

template <bool PLASTIC, bool DAMAGE, bool FORMED>
__global__ __launch_bounds__(kTile) void k_elem_tiled(const CaseDev* __restrict__ cases,
                                                      const Tile* __restrict__ tiles, int ntiles,
                                                      uint32_t lmax, int parity, int sample,
                                                      int xcd) {
    extern __shared__ __align__(16) unsigned char smem[];
    const Tile tile = tile_of(tiles, kTile, xcd);
    const CaseDev& c = cases[tile.case_index];
    tiled_elem_block<PLASTIC, DAMAGE, FORMED>(c, tile.first, threadIdx.x, true, 0, lmax, sample, smem,
                                              c.step + parity);
}

// ---------------------------------------------------------------- node pass
// The node pass of the nodes from `first` (a multiple of 64), lane `lid` taking node
// first + lid, at step `n`.
template <bool EXTRA, int VARIANT>
__device__ __forceinline__ void node_block(const CaseDev& c, uint32_t first, uint32_t lid, int64_t n, int parity,
                                           int sample) {
    const uint32_t i = first + lid;
    const uint32_t lane = lid & 63u;
    // tiles start on multiples of 64 nodes, so the wave index is wave-uniform
    const uint32_t w = __builtin_amdgcn_readfirstlane(i >> 6);
    // this wave's gather slab (fixed order: slots in CSR (tet, corner) order, tiled in
    // tile order). A block may overhang the case's last node: waves past its
    // node_waves slabs have no offsets (off holds node_waves + 1 entries) and gather
    // nothing.
    const uint32_t* off = VARIANT == kSlots ? c.ell_off : c.pell_off;
    const uint32_t* slab = VARIANT == kSlots ? c.ell : c.pell;
    uint32_t base = 0, width = 0;
    if (w < c.node_waves) {
        base = ld_uniform(off + w);
        width = (ld_uniform(off + w + 1) - base) >> 6;
    }
    const bool live = i < c.N;
    const double dt = c.dt;
    const float dtf = c.dtf;
    const uint32_t np = c.Npad;

    double a_ext = 0.0, a_con = 0.0, a_bnd = 0.0, a_damp = 0.0;
    double s_kin = 0.0, s_hext = 0.0, s_hcon = 0.0, s_hbnd = 0.0, s_hdamp = 0.0;
    double s_f[3] = {0.0, 0.0, 0.0};

    // The gather's slab indices go out first (they head the longest chain: index, then
    // data), then every load that does not depend on the gather, then the gathered
    // data; a wave then waits about two memory round trips. Rows are clamped into
    // the slab and padding lanes read node 0, so every load is issued unconditionally
    // and values a lane does not own are discarded. A lane's column ends in
    // sentinels, which add nothing.
    constexpr uint32_t kGather = 8;
    // a wave with an empty slab reads the case's first 64 flags instead and turns
    // them into sentinels (nflags holds Npad >= 64 entries)
    const auto* sl = width > 0 ? gmem(slab) + base + lane : gmem(c.nflags) + lane;
    const uint32_t last = width > 0 ? width - 1 : 0;
    const uint32_t none = width > 0 ? 0u : kEllSentinel;
    uint32_t idx[kGather];
#pragma unroll
    for (uint32_t k = 0; k < kGather; ++k) idx[k] = sl[min(k, last) * 64u] | none;
    const uint32_t ii = live ? i : 0u;
    const uint32_t flr = gmem(c.nflags)[ii];
    // m is needed on sample steps and by gravity, contact, the drive and damping; other
    // steps read the 1/m line again (already fetched) instead of the mass plane
    const bool need_m = sample || (c.features & (kFeatGravity | kFeatContact | kFeatDrive | kFeatDamp));
    const float m = gmem(need_m ? c.mass : c.invm)[ii];
    const float im = gmem(c.invm)[ii];
    const float4 u = gload(c.u4 + ii);
    // without a load plane the velocity plane is read in its place and never used
    const float* lplane = (c.features & kFeatLoad) ? c.load : c.v;
    float vv[3], cvv[3], cuu[3], ld[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        vv[a] = gmem(c.v)[a * np + ii];
        cvv[a] = gmem(c.cv)[a * np + ii];
        cuu[a] = gmem(c.cu)[a * np + ii];
        ld[a] = gmem(lplane)[a * np + ii];
    }
    const uint32_t fl = live ? flr : kPinAll;
    // the step counter is needed only now: the loads above went out without waiting for it
    const bool frozen = n >= c.step_end;
    if (first == 0 && lid == 0) c.step[parity ^ 1] = frozen ? n : n + 1;
    if (frozen) return;

    float fx = 0.0f, fy = 0.0f, fz = 0.0f;
    auto gather = [&](uint32_t q0) {
        float g[kGather][3];
#pragma unroll
        for (uint32_t k = 0; k < kGather; ++k) {
            const uint32_t x = idx[k] == kEllSentinel ? 0u : idx[k];
            if (VARIANT == kSlots) {
                const auto* s = gmem(c.force) + (size_t)(x >> 2) * 12u + (x & 3u) * 3u;
                g[k][0] = s[0];
                g[k][1] = s[1];
                g[k][2] = s[2];
            } else {
                const float4 pv = gload(c.partial + x);
                g[k][0] = pv.x;
                g[k][1] = pv.y;
                g[k][2] = pv.z;
            }
        }
#pragma unroll
        for (uint32_t k = 0; k < kGather; ++k)
            if (q0 + k < width && idx[k] != kEllSentinel) {
                fx += g[k][0];
                fy += g[k][1];
                fz += g[k][2];
            }
    };
    gather(0);  // straight-line, so its loads wait only for the indices
    for (uint32_t q0 = kGather; q0 < width; q0 += kGather) {
#pragma unroll
        for (uint32_t k = 0; k < kGather; ++k) idx[k] = sl[min(q0 + k, last) * 64u];
        gather(q0);
    }

    if ((fl & kPinAll) != kPinAll) {
        const double t = (double)n * dt;
        float fe[3] = {0.0f, 0.0f, 0.0f};
        if ((c.features & kFeatLoad) && (fl & kLoaded)) {
            const double sl = c.load_ramp > 0.0 ? fmin(fmax(t / c.load_ramp, 0.0), 1.0) : 1.0;
            const float slf = (float)sl;
#pragma unroll
            for (int a = 0; a < 3; ++a) fe[a] = slf * ld[a];
        }
        if (c.features & kFeatGravity) fe[2] -= c.gravity * m;

        float uu[3] = {u.x, u.y, u.z};
        float fc[3] = {0.0f, 0.0f, 0.0f};
        if (EXTRA && (c.features & kFeatContact) && (fl & kContact)) {
            // d = (X - c0) + (u - t v_ind); phi = |d|^2 - r^2 from the stored
            // |X - c0|^2 - r^2 so no term cancels at FP32
            float xc[3], wv[3];
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                xc[a] = c.xc[a * np + i];
                wv[a] = uu[a] - (float)(t * c.ind_vel[a]);
            }
            const float h = c.xc[3 * np + i];
            const float phi = h + 2.0f * (xc[0] * wv[0] + xc[1] * wv[1] + xc[2] * wv[2]) +
                              (wv[0] * wv[0] + wv[1] * wv[1] + wv[2] * wv[2]);
            if (phi < 0.0f) {
                const float r = c.ind_r;
                const float dist = sqrtf(fmaxf(r * r + phi, 0.0f));
                if (dist > 0.0f) {
                    const float gap = -phi / (r + dist);
                    const float k = c.ind_kdt2 * m * gap / dist;
#pragma unroll
                    for (int a = 0; a < 3; ++a) fc[a] = k * (xc[a] + wv[a]);
                }
            }
        }

        const float fint[3] = {fx, fy, fz};
        const float sd = (EXTRA && (c.features & kFeatDrive))
                             ? [&] {
                                   if (!(c.drive_ramp > 0.0)) return 1.0f;
                                   double s = fmin(fmax((t + dt) / c.drive_ramp, 0.0), 1.0);
                                   return (float)(s * s * (3.0 - 2.0 * s));
                               }()
                             : 1.0f;
        bool bad = false;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            if (fl & (kPinX << a)) {
                vv[a] = 0.0f;
                cvv[a] = 0.0f;
                continue;
            }
            const float vo = vv[a];
            const float ftot = fint[a] + fe[a] + fc[a];
            float vn, cvn;
            if (EXTRA && (c.features & kFeatDrive) && (fl & (kDriveX << a))) {
                const float vfree = vo + ftot * (im * dtf);
                const float ut = uu[a] - cuu[a];
                vn = (c.presc[a * np + i] * sd - ut) / dtf;
                cvn = 0.0f;
                // the constraint impulse m (vn - vfree) works at the mid-step velocity,
                // like every other impulse in the account
                const double J = (double)m * ((double)vn - (double)vfree);
                a_bnd += J * ((double)vo + (double)vn) * 0.5;
                if (sample) s_hbnd += J * (double)vn * 0.5;
            } else {
                // Kahan: the true velocity is vv - cvv
                const float y = ftot * (im * dtf) - cvv[a];
                const float tt = vo + y;
                cvn = (tt - vo) - y;
                vn = tt;
                if (EXTRA && (c.features & kFeatDamp)) {
                    // the damping impulse m (d - 1) v removes m (1 - d) v (vo + d v) / 2
                    const float d = c.damping;
                    const double J = (double)m * (1.0 - (double)d) * (double)vn;
                    a_damp += J * ((double)vo + (double)d * (double)vn) * 0.5;
                    if (sample) s_hdamp += J * (double)d * (double)vn * 0.5;
                    vn *= d;
                    cvn *= d;
                }
            }
            a_ext += (double)fe[a] * ((double)vo + (double)vn) * 0.5 * dt;
            a_con += (double)fc[a] * ((double)vo + (double)vn) * 0.5 * dt;
            if (sample) {
                s_kin += 0.5 * (double)m * (double)vo * (double)vn;
                s_hext += (double)fe[a] * (double)vn * 0.5 * dt;
                s_hcon += (double)fc[a] * (double)vn * 0.5 * dt;
            }
            // Kahan displacement update: the true displacement is uu - cuu
            const float y = vn * dtf - cuu[a];
            const float tt = uu[a] + y;
            cuu[a] = (tt - uu[a]) - y;
            uu[a] = tt;
            vv[a] = vn;
            cvv[a] = cvn;
            bad |= !isfinite(vn);
        }
        if (sample) {
#pragma unroll
            for (int a = 0; a < 3; ++a) s_f[a] = fc[a];
        }
        c.u4[i] = make_float4(uu[0], uu[1], uu[2], 0.0f);
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            c.v[a * np + i] = vv[a];
            c.cv[a * np + i] = cvv[a];
            c.cu[a * np + i] = cuu[a];
        }
        if (bad) c.status[0] = 1u;
    } else if (live) {
        // a fully pinned node keeps v = 0 (fem.cpp:197-200); its state never changes
    }

    const uint32_t nw = c.node_waves;
    const bool writer = lane == 0 && w < nw;
    // a running sum a wave adds only zeros to keeps its bits: such waves skip the butterfly
    if ((c.features & (kFeatLoad | kFeatGravity)) && __ballot(a_ext != 0.0)) {
        const double s = wave_sum(a_ext);
        if (writer) c.acc_n[kAccWext * nw + w] += s;
    }
    if (EXTRA) {
        if ((c.features & kFeatContact) && __ballot(a_con != 0.0)) {
            const double s = wave_sum(a_con);
            if (writer) c.acc_n[kAccWcon * nw + w] += s;
        }
        if ((c.features & kFeatDrive) && __ballot(a_bnd != 0.0)) {
            const double s = wave_sum(a_bnd);
            if (writer) c.acc_n[kAccWbnd * nw + w] += s;
        }
        if ((c.features & kFeatDamp) && __ballot(a_damp != 0.0)) {
            const double s = wave_sum(a_damp);
            if (writer) c.acc_n[kAccDamp * nw + w] += s;
        }
    }
    if (sample) {
        const double vals[kSmpNodeSlots] = {s_kin, s_hext, s_hcon, s_f[0], s_f[1], s_f[2],
                                            s_hbnd, s_hdamp};
#pragma unroll
        for (int k = 0; k < kSmpNodeSlots; ++k) {
            const double s = wave_sum(vals[k]);
            if (writer) c.smp_n[k * nw + w] = s;
        }
    }
}

template <int BLOCK, bool EXTRA, int VARIANT>
__global__ __launch_bounds__(BLOCK) void k_node(const CaseDev* __restrict__ cases,
                                                const Tile* __restrict__ tiles, int ntiles,
                                                int parity, int sample, int xcd) {
    const Tile tile = tile_of(tiles, BLOCK, xcd);
    const CaseDev& c = cases[tile.case_index];
    node_block<EXTRA, VARIANT>(c, tile.first, threadIdx.x, ld_uniform(c.step + parity), parity, sample);
}

// ---------------------------------------------------------------- FP64 reductions
constexpr int kReduceBlock = 256;

// One block per case, 1024 threads. Every thread first folds its strided share of all
// slot arrays in one pass (independent loads in flight together), then each quantity
// goes through the same fixed tree. The order depends only on the case's wave count.
constexpr int kRecordBlock = 1024;
constexpr int kQ = kAccElemSlots + kSmpElemSlots + kAccNodeSlots + kSmpNodeSlots;

// The energy record of the case's step n0 from its per-wave slots, by its 1024 threads
// (lane `tid`) through `sh` (1024 doubles): every thread folds its strided share, each wave
// folds its 64 lanes by the xor butterfly, and one lane per quantity sums the 16 waves'
// partials in wave order: a fixed tree for every quantity, with two barriers instead of one
// per tree level and quantity. Ends with a barrier, so `sh` may be reused.
__device__ __forceinline__ void reduce_record(const CaseDev& c, int64_t n0, uint32_t tid, double* sh) {
    const uint32_t ew = c.elem_waves, nw = c.node_waves;
    // q: element accumulators, the element sample, node accumulators, node samples
    constexpr int QS = kAccElemSlots, QN = QS + kSmpElemSlots, QM = QN + kAccNodeSlots;
    double q[kQ];
#pragma unroll
    for (int k = 0; k < kQ; ++k) q[k] = 0.0;
    for (uint32_t w = tid; w < ew; w += kRecordBlock) {
#pragma unroll
        for (int k = 0; k < kAccElemSlots; ++k) q[k] += c.acc_e[k * ew + w];
        q[QS] += c.smp_e[kSmpStrain * ew + w];
    }
    for (uint32_t w = tid; w < nw; w += kRecordBlock) {
#pragma unroll
        for (int k = 0; k < kAccNodeSlots; ++k) q[QN + k] += c.acc_n[k * nw + w];
#pragma unroll
        for (int k = 0; k < kSmpNodeSlots; ++k) q[QM + k] += c.smp_n[k * nw + w];
    }
    constexpr uint32_t kWaves = kRecordBlock / 64;
    static_assert(kWaves * kQ + kQ <= kRecordBlock, "record partials fit the shared buffer");
#pragma unroll
    for (int k = 0; k < kQ; ++k) q[k] = wave_sum(q[k]);
    if ((tid & 63u) == 0)
#pragma unroll
        for (int k = 0; k < kQ; ++k) sh[(tid >> 6) * kQ + k] = q[k];
    __syncthreads();
    if (tid < (uint32_t)kQ) {
        double t = 0.0;
        for (uint32_t w = 0; w < kWaves; ++w) t += sh[w * kQ + tid];
        sh[kWaves * kQ + tid] = t;
    }
    __syncthreads();
    if (tid == 0) {
        double tot[kQ];
#pragma unroll
        for (int k = 0; k < kQ; ++k) tot[k] = sh[kWaves * kQ + k];
        const int32_t h = *c.hist_count;
        if (h < c.hist_cap) {
            const double* sn = tot + QM;
            EnergyRecordDev& r = c.hist[h];
            r.step = n0;
            r.time = (double)n0 * c.dt;
            r.strain = tot[QS];
            r.kinetic = sn[kSmpKin];
            r.dissipation = tot[kAccDiss];
            r.fracture = tot[kAccFract];
            r.w_ext = tot[QN + kAccWext] - sn[kSmpHalfExt];
            r.w_contact = tot[QN + kAccWcon] - sn[kSmpHalfCon];
            r.w_boundary = tot[QN + kAccWbnd] - sn[kSmpHalfBnd];
            r.damping_loss = tot[QN + kAccDamp] - sn[kSmpHalfDamp];
            r.contact_force[0] = sn[kSmpFx];
            r.contact_force[1] = sn[kSmpFy];
            r.contact_force[2] = sn[kSmpFz];
            *c.hist_count = h + 1;
        }
    }
    __syncthreads();
}

__global__ __launch_bounds__(kRecordBlock) void k_reduce(const CaseDev* __restrict__ cases,
                                                          int parity) {
    __shared__ double sh[kRecordBlock];
    const CaseDev& c = cases[blockIdx.x];
    const int64_t n0 = c.step[parity];
    if (c.step[parity ^ 1] != n0 + 1) return;  // the case did not step: no record
    reduce_record(c, n0, threadIdx.x, sh);
}

// ---------------------------------------------------------------- a small case resident
// One 1024-thread block per case of at most four tiles (1024 tets) and 2048 nodes runs
// `steps` whole steps in one launch: the element pass of tile g by threads [256 g, 256 g +
// 256) (each group its own LDS stage), a barrier, the node pass of 256-node tiles g, g + 4
// by the same groups, a barrier; on the sampled first step the energy record. Each step
// runs the separate passes' own code in their order, so a case ends with the bits the
// two-launch step gives it. The step counter is carried in a register: within the launch
// it is never read back from memory the launch has written.
constexpr int kResidentBlock = 4 * kTile;

template <bool PLASTIC, bool DAMAGE, bool FORMED, bool EXTRA>
__global__ __launch_bounds__(kResidentBlock) void k_resident(const CaseDev* __restrict__ cases, uint32_t lmax,
                                                            uint32_t tile_lds, int steps, int parity0,
                                                            int sample_first) {
    extern __shared__ __align__(16) unsigned char smem[];
    const CaseDev& c = cases[blockIdx.x];
    const uint32_t tid = threadIdx.x, g = tid / kTile, lid = tid % kTile;
    const uint32_t tiles = c.elem_tiles, node_tiles = (c.N + kTile - 1) / kTile;
    int64_t n = ld_uniform(c.step + parity0);  // written by earlier launches only
    for (int k = 0; k < steps; ++k) {
        if (n >= c.step_end) break;  // the case has finished (uniform)
        const int parity = parity0 ^ (k & 1);
        const int sample = k == 0 ? sample_first : 0;
        tiled_elem_block<PLASTIC, DAMAGE, FORMED>(c, g * kTile, lid, g < tiles, n, lmax, sample,
                                                  smem + (size_t)g * tile_lds);
        __syncthreads();  // every tile's partials written
        for (uint32_t t = g; t < node_tiles; t += 4) node_block<EXTRA, kTiled>(c, t * kTile, lid, n, parity, sample);
        __syncthreads();  // u, v, the sums and the step counter written
        if (sample) reduce_record(c, n, tid, reinterpret_cast<double*>(smem));
        ++n;
    }
}


// ---------------------------------------------------------------- FP64 stable step
// The element bound of omega.hpp on the stored Dm^-1 planes.
__device__ double element_lambda_max(const float* dmi, uint32_t tp, uint32_t e, double lam,
                                     double mu) {
    double B[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) B[k] = (double)dmi[k * tp + e];
    return lambda_max_dminv(B, lam, mu);
}

__global__ __launch_bounds__(256) void k_omega(const CaseDev* __restrict__ cases,
                                               const Tile* __restrict__ tiles) {
    const Tile tile = tiles[blockIdx.x];
    const CaseDev& c = cases[tile.case_index];
    const uint32_t e = tile.first + threadIdx.x;
    double l = 0.0;
    if (e < c.T) l = element_lambda_max(c.dmi, c.Tpad, e, c.lam_d, c.mu_d);
    l = wave_max(l);
    if ((threadIdx.x & 63u) == 0 && (e >> 6) < c.elem_waves) c.omega_w[e >> 6] = l;
}

__global__ __launch_bounds__(kReduceBlock) void k_omega_max(const CaseDev* __restrict__ cases,
                                                             double* out) {
    __shared__ double sh[kReduceBlock];
    const CaseDev& c = cases[blockIdx.x];
    double m = 0.0;
    for (uint32_t k = threadIdx.x; k < c.elem_waves; k += kReduceBlock) m = fmax(m, c.omega_w[k]);
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int s = kReduceBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = sh[0];
}

// Step counter and status of every case, for the host to poll with one copy.
__global__ __launch_bounds__(256) void k_progress(const CaseDev* __restrict__ cases, int n,
                                                  int64_t* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const CaseDev& c = cases[i];
    const int64_t a = c.step[0], b = c.step[1];
    out[2 * i] = a > b ? a : b;
    out[2 * i + 1] = static_cast<int64_t>(c.status[0]);
}

// PLASTIC / DAMAGE instantiations only add code gated per case by its feature bits, so
// a case rounds identically in whichever instantiation a mix selects.
template <int BLOCK>
void elem_launch(const CaseDev* cases, const Tile* tiles, int ntiles, int parity, int sample,
                 int xcd, bool plastic, bool damage, hipStream_t s, const KernelTiming* t) {
    if (damage)
        SHIPFEM_LAUNCH(t, (k_elem<BLOCK, true, true>), dim3(ntiles), dim3(BLOCK), 0, s, cases, tiles,
                       ntiles, parity, sample, xcd);
    else if (plastic)
        SHIPFEM_LAUNCH(t, (k_elem<BLOCK, true, false>), dim3(ntiles), dim3(BLOCK), 0, s, cases, tiles,
                       ntiles, parity, sample, xcd);
    else
        SHIPFEM_LAUNCH(t, (k_elem<BLOCK, false, false>), dim3(ntiles), dim3(BLOCK), 0, s, cases,
                       tiles, ntiles, parity, sample, xcd);
}

template <int BLOCK, int VARIANT>
void node_launch(const CaseDev* cases, const Tile* tiles, int ntiles, int parity, int sample,
                 int xcd, bool extra, hipStream_t s, const KernelTiming* t) {
    if (extra)
        SHIPFEM_LAUNCH(t, (k_node<BLOCK, true, VARIANT>), dim3(ntiles), dim3(BLOCK), 0, s, cases,
                       tiles, ntiles, parity, sample, xcd);
    else
        SHIPFEM_LAUNCH(t, (k_node<BLOCK, false, VARIANT>), dim3(ntiles), dim3(BLOCK), 0, s, cases,
                       tiles, ntiles, parity, sample, xcd);
}

template <int BLOCK>
void node_launch_v(const CaseDev* cases, const Tile* tiles, int ntiles, int parity, int sample,
                   int xcd, bool extra, int variant, hipStream_t s, const KernelTiming* t) {
    if (variant == kTiled)
        node_launch<BLOCK, kTiled>(cases, tiles, ntiles, parity, sample, xcd, extra, s, t);
    else
        node_launch<BLOCK, kSlots>(cases, tiles, ntiles, parity, sample, xcd, extra, s, t);
}

}  // namespace

void launch_elem(const CaseDev* cases, const Tile* tiles, int ntiles, int block, int parity,
                 int sample, int xcd, uint32_t feat_union, hipStream_t stream,
                 const KernelTiming* t) {
    const bool plastic = (feat_union & kFeatPlastic) != 0;
    const bool damage = (feat_union & kFeatDamage) != 0;
    switch (block) {
        case 64: elem_launch<64>(cases, tiles, ntiles, parity, sample, xcd, plastic, damage, stream, t); break;
        case 128: elem_launch<128>(cases, tiles, ntiles, parity, sample, xcd, plastic, damage, stream, t); break;
        case 512: elem_launch<512>(cases, tiles, ntiles, parity, sample, xcd, plastic, damage, stream, t); break;
        default: elem_launch<256>(cases, tiles, ntiles, parity, sample, xcd, plastic, damage, stream, t); break;
    }
}

void launch_node(const CaseDev* cases, const Tile* tiles, int ntiles, int block, int parity,
                 int sample, int xcd, uint32_t feat_union, int variant, hipStream_t stream,
                 const KernelTiming* t) {
    const bool extra = (feat_union & (kFeatContact | kFeatDrive | kFeatDamp)) != 0;
    switch (block) {
        case 64: node_launch_v<64>(cases, tiles, ntiles, parity, sample, xcd, extra, variant, stream, t); break;
        case 128: node_launch_v<128>(cases, tiles, ntiles, parity, sample, xcd, extra, variant, stream, t); break;
        case 512: node_launch_v<512>(cases, tiles, ntiles, parity, sample, xcd, extra, variant, stream, t); break;
        default: node_launch_v<256>(cases, tiles, ntiles, parity, sample, xcd, extra, variant, stream, t); break;
    }
}

void launch_elem_hex(const CaseDev* cases, const Tile* tiles, int ntiles, int block, int parity,
                     int sample, int xcd, hipStream_t stream, const KernelTiming* t) {
    switch (block) {
        case 64:
            SHIPFEM_LAUNCH(t, (k_elem_hex<64>), dim3(ntiles), dim3(64), 0, stream, cases, tiles, ntiles,
                           parity, sample, xcd);
            break;
        case 128:
            SHIPFEM_LAUNCH(t, (k_elem_hex<128>), dim3(ntiles), dim3(128), 0, stream, cases, tiles, ntiles,
                           parity, sample, xcd);
            break;
        case 512:
            SHIPFEM_LAUNCH(t, (k_elem_hex<512>), dim3(ntiles), dim3(512), 0, stream, cases, tiles, ntiles,
                           parity, sample, xcd);
            break;
        default:
            SHIPFEM_LAUNCH(t, (k_elem_hex<256>), dim3(ntiles), dim3(256), 0, stream, cases, tiles, ntiles,
                           parity, sample, xcd);
            break;
    }
}

void launch_elem_hex_plastic(const CaseDev* cases, const Tile* tiles, int ntiles, int block,
                             int parity, int sample, int xcd, hipStream_t stream,
                             const KernelTiming* t) {
    switch (block) {
        case 64:
            SHIPFEM_LAUNCH(t, (k_elem_hex_plastic<64>), dim3(ntiles), dim3(64), 0, stream, cases, tiles,
                           ntiles, parity, sample, xcd);
            break;
        case 128:
            SHIPFEM_LAUNCH(t, (k_elem_hex_plastic<128>), dim3(ntiles), dim3(128), 0, stream, cases, tiles,
                           ntiles, parity, sample, xcd);
            break;
        default:
            SHIPFEM_LAUNCH(t, (k_elem_hex_plastic<256>), dim3(ntiles), dim3(256), 0, stream, cases, tiles,
                           ntiles, parity, sample, xcd);
            break;
    }
}

void launch_elem_tiled(const CaseDev* cases, const Tile* tiles, int ntiles, uint32_t lmax,
                       int parity, int sample, int xcd, uint32_t feat_union, bool formed, hipStream_t stream,
                       const KernelTiming* t) {
    // the local nodes' displacements after the slots and run offsets
    const size_t lds = tiled_lds_nodes_offset(lmax) + 16 * (size_t)lmax;
#define SHIPFEM_TILED(P, D, F)                                                                         \
    SHIPFEM_LAUNCH(t, (k_elem_tiled<P, D, F>), dim3(ntiles), dim3(kTile), lds, stream, cases, tiles, ntiles, \
                   lmax, parity, sample, xcd)
    if (feat_union & kFeatDamage) {
        if (formed) SHIPFEM_TILED(true, true, true); else SHIPFEM_TILED(true, true, false);
    } else if (feat_union & kFeatPlastic) {
        if (formed) SHIPFEM_TILED(true, false, true); else SHIPFEM_TILED(true, false, false);
    } else {
        if (formed) SHIPFEM_TILED(false, false, true); else SHIPFEM_TILED(false, false, false);
    }
#undef SHIPFEM_TILED
}

size_t resident_lds(uint32_t lmax) {
    const size_t tile = tiled_lds_nodes_offset(lmax) + 16 * (size_t)lmax;
    return std::max<size_t>(4 * ((tile + 15) / 16 * 16), 8 * kRecordBlock);
}

void launch_resident(const CaseDev* cases, int ncases, uint32_t lmax, int steps, int parity0, int sample_first,
                     uint32_t feat_union, bool formed, hipStream_t stream) {
    const uint32_t tile = (uint32_t)((tiled_lds_nodes_offset(lmax) + 16 * (size_t)lmax + 15) / 16 * 16);
    const size_t lds = resident_lds(lmax);
    const bool extra = (feat_union & (kFeatContact | kFeatDrive | kFeatDamp)) != 0;
#define SHIPFEM_RES(P, D, F, X)                                                                           \
    hipLaunchKernelGGL((k_resident<P, D, F, X>), dim3(ncases), dim3(kResidentBlock), lds, stream, cases, lmax, \
                       tile, steps, parity0, sample_first)
#define SHIPFEM_RES_X(P, D, F) \
    if (extra) SHIPFEM_RES(P, D, F, true); else SHIPFEM_RES(P, D, F, false)
#define SHIPFEM_RES_F(P, D) \
    if (formed) { SHIPFEM_RES_X(P, D, true); } else { SHIPFEM_RES_X(P, D, false); }
    if (feat_union & kFeatDamage) {
        SHIPFEM_RES_F(true, true);
    } else if (feat_union & kFeatPlastic) {
        SHIPFEM_RES_F(true, false);
    } else {
        SHIPFEM_RES_F(false, false);
    }
#undef SHIPFEM_RES_F
#undef SHIPFEM_RES_X
#undef SHIPFEM_RES
}

void launch_reduce(const CaseDev* cases, int ncases, int parity, hipStream_t stream) {
    hipLaunchKernelGGL(k_reduce, dim3(ncases), dim3(kRecordBlock), 0, stream, cases, parity);
}

void launch_omega(const CaseDev* cases, const Tile* tiles, int ntiles, hipStream_t stream) {
    hipLaunchKernelGGL(k_omega, dim3(ntiles), dim3(256), 0, stream, cases, tiles);
}

void launch_omega_max(const CaseDev* cases, int ncases, double* out, hipStream_t stream) {
    hipLaunchKernelGGL(k_omega_max, dim3(ncases), dim3(kReduceBlock), 0, stream, cases, out);
}

void launch_progress(const CaseDev* cases, int ncases, int64_t* out, hipStream_t stream) {
    hipLaunchKernelGGL(k_progress, dim3((ncases + 255) / 256), dim3(256), 0, stream, cases, ncases,
                       out);
}

}  // namespace shipfem
