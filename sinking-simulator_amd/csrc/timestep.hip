// timestep.hip -- the FP64 stable step of a tet mesh on the device, without a solver.
//
// criticalTimestep (engine/sim/fem.cpp:183-185) takes a safety factor times the shortest
// edge over the bar wave speed. This is the bound the solver steps with instead: the
// largest lumped-mass element frequency omega_e^2 = 4 lambda_max(B^T C B) / rho (omega.hpp),
// over every tet, and cfl = 2 / omega_max. The device forms each tet's Dm^-1 from the rest
// coordinates in FP64 (the reference's |det| > 1e-30 guard: a degenerate tet takes zero),
// reduces lambda_max to one maximum per 64-tet wave, then folds the wave maxima in one block.
// Maxima are exact in any order, so the result does not depend on the launch.
#include "../../include/shipfem.h"
#include "omega.hpp"
#include "shipfem_host.hpp"

#include <cmath>
#include <mutex>
#include <string>

// Device scratch of shipfem_stable_timestep_noalloc: one copy per GPU, held until the process
// ends. In host code `scratch` only names the symbol; its device address comes from
// hipGetSymbolAddress, once per device (scratchOnDevice below).
constexpr uint64_t kScratchBytes = 32ull << 20;
static __device__ __align__(256) char scratch[kScratchBytes];

namespace shipfem {
namespace {

constexpr int kOmegaBlock = 256;

__device__ __forceinline__ double wave_max64(double s) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s = fmax(s, __shfl_xor(s, off, 64));
    return s;
}

__global__ __launch_bounds__(kOmegaBlock) void k_rest_omega(const double* __restrict__ X,
                                                            const uint4* __restrict__ index, uint32_t T,
                                                            double lam, double mu,
                                                            double* __restrict__ wave) {
    const uint32_t e = blockIdx.x * kOmegaBlock + threadIdx.x;
    double l = 0.0;
    if (e < T) {
        const uint4 q = index[e];
        const uint32_t v[4] = {q.x, q.y, q.z, q.w};
        double D[3][3];  // column c: X_{c+1} - X_0
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int a = 0; a < 3; ++a) D[a][c] = X[3ull * v[c + 1] + a] - X[3ull * v[0] + a];
        const double a00 = D[1][1] * D[2][2] - D[1][2] * D[2][1];
        const double a01 = D[0][2] * D[2][1] - D[0][1] * D[2][2];
        const double a02 = D[0][1] * D[1][2] - D[0][2] * D[1][1];
        const double det = D[0][0] * a00 + D[1][0] * a01 + D[2][0] * a02;
        const double id = fabs(det) > 1e-30 ? 1.0 / det : 0.0;
        const double B[9] = {a00 * id,
                             a01 * id,
                             a02 * id,
                             (D[1][2] * D[2][0] - D[1][0] * D[2][2]) * id,
                             (D[0][0] * D[2][2] - D[0][2] * D[2][0]) * id,
                             (D[0][2] * D[1][0] - D[0][0] * D[1][2]) * id,
                             (D[1][0] * D[2][1] - D[1][1] * D[2][0]) * id,
                             (D[0][1] * D[2][0] - D[0][0] * D[2][1]) * id,
                             (D[0][0] * D[1][1] - D[0][1] * D[1][0]) * id};
        l = lambda_max_dminv(B, lam, mu);
    }
    l = wave_max64(l);
    if ((threadIdx.x & 63u) == 0 && e < T) wave[e >> 6] = l;
}

__global__ __launch_bounds__(kOmegaBlock) void k_fold_max(const double* __restrict__ wave, uint32_t n,
                                                          double* __restrict__ out) {
    __shared__ double sh[kOmegaBlock];
    double m = 0.0;
    for (uint32_t k = threadIdx.x; k < n; k += kOmegaBlock) m = fmax(m, wave[k]);
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int s = kOmegaBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = sh[0];
}

uint64_t up256(uint64_t b) { return (b + 255) / 256 * 256; }

// Device memory one evaluation uses: rest coordinates [3N] double, connectivity [4T],
// the wave maxima [ceil(T / 64)] and the result, each 256-byte aligned.
struct StepLayout {
    uint64_t X = 0, index = 0, wave = 0, out = 0, total = 0;
    StepLayout(uint64_t N, uint64_t T) {
        index = up256(24 * N);
        wave = index + up256(16 * T);
        out = wave + up256(8 * ((T + 63) / 64));
        total = out + 256;
    }
};

std::string checkArgs(const double* X, const uint32_t* index, uint32_t N, uint32_t T, double E, double nu,
                      double rho, const double* cfl) {
    if (!X || !index || !cfl) return "rest_position, index and cfl are required";
    if (N == 0 || T == 0) return "empty mesh (node_count and tet_count must be positive)";
    if (!(E > 0) || !(nu > -1.0 && nu < 0.5) || !(rho > 0))
        return "material out of range (E > 0, -1 < nu < 0.5, density > 0)";
    for (uint64_t q = 0; q < 4ull * T; ++q)
        if (index[q] >= N)
            return "tet " + std::to_string(q / 4) + " references node " + std::to_string(index[q]) +
                   " >= node_count";
    return {};
}

// Uploads, evaluates and reads back on `stream` using `base` (StepLayout(N, T).total bytes).
hipError_t evaluate(char* base, const double* X, const uint32_t* index, uint32_t N, uint32_t T, double E,
                    double nu, double rho, double* cfl, hipStream_t stream) {
    const StepLayout L(N, T);
    hipError_t e = hipMemcpyAsync(base + L.X, X, 24ull * N, hipMemcpyHostToDevice, stream);
    if (e == hipSuccess) e = hipMemcpyAsync(base + L.index, index, 16ull * T, hipMemcpyHostToDevice, stream);
    if (e != hipSuccess) return e;
    const double lam = E * nu / ((1 + nu) * (1 - 2 * nu)), mu = E / (2 * (1 + nu));
    const uint32_t blocks = (T + kOmegaBlock - 1) / kOmegaBlock, waves = (T + 63) / 64;
    auto* wave = reinterpret_cast<double*>(base + L.wave);
    auto* out = reinterpret_cast<double*>(base + L.out);
    hipLaunchKernelGGL(k_rest_omega, dim3(blocks), dim3(kOmegaBlock), 0, stream,
                       reinterpret_cast<const double*>(base + L.X), reinterpret_cast<const uint4*>(base + L.index),
                       T, lam, mu, wave);
    hipLaunchKernelGGL(k_fold_max, dim3(1), dim3(kOmegaBlock), 0, stream, wave, waves, out);
    e = hipGetLastError();
    double lmax = 0.0;
    if (e == hipSuccess) e = hipMemcpyAsync(&lmax, out, 8, hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e != hipSuccess) return e;
    *cfl = lmax > 0 ? 2.0 / std::sqrt(4.0 * lmax / rho) : INFINITY;
    return hipSuccess;
}

// The current device's copy of `scratch`, looked up on the first call on that device (with
// it current) and kept per device.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
char* g_scratch[kMaxDevices] = {};

hipError_t scratchOnDevice(char** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mutex);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        g_scratch[dev] = static_cast<char*>(p);
    }
    *out = g_scratch[dev];
    return hipSuccess;
}

}  // namespace
}  // namespace shipfem

using namespace shipfem;

extern "C" {

shipfem_status shipfem_stable_timestep(const double* rest_position, const uint32_t* index, uint32_t node_count,
                                       uint32_t tet_count, double youngs_modulus, double poisson_ratio,
                                       double density, double* cfl, char* err, size_t err_len) {
    const std::string why = checkArgs(rest_position, index, node_count, tet_count, youngs_modulus, poisson_ratio,
                                      density, cfl);
    if (!why.empty()) {
        setError(err, err_len, why);
        return SHIPFEM_INVALID_ARG;
    }
    std::string dev;
    if (!deviceAvailable(&dev)) {
        setError(err, err_len, dev);
        return SHIPFEM_NO_DEVICE;
    }
    const StepLayout L(node_count, tet_count);
    char* base = nullptr;
    hipStream_t stream = nullptr;
    hipError_t e = hipStreamCreateWithFlags(&stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&base), L.total);
    if (e == hipSuccess)
        e = evaluate(base, rest_position, index, node_count, tet_count, youngs_modulus, poisson_ratio, density, cfl,
                     stream);
    if (base) (void)hipFree(base);
    if (stream) (void)hipStreamDestroy(stream);
    if (e != hipSuccess) {
        setError(err, err_len, std::string("HIP: ") + hipGetErrorString(e));
        return statusOf(e);
    }
    return SHIPFEM_OK;
}

shipfem_status shipfem_stable_timestep_noalloc(const double* rest_position, const uint32_t* index,
                                               uint32_t node_count, uint32_t tet_count, double youngs_modulus,
                                               double poisson_ratio, double density, double* cfl, char* err,
                                               size_t err_len) {
    const std::string why = checkArgs(rest_position, index, node_count, tet_count, youngs_modulus, poisson_ratio,
                                      density, cfl);
    if (!why.empty()) {
        setError(err, err_len, why);
        return SHIPFEM_INVALID_ARG;
    }
    const StepLayout L(node_count, tet_count);
    if (L.total > kScratchBytes) {
        setError(err, err_len, "the mesh needs " + std::to_string(L.total) + " bytes of device scratch (24 per node, " +
                                   "16 per tet, 8 per 64 tets), more than the " + std::to_string(kScratchBytes) +
                                   " this entry point holds; use shipfem_stable_timestep");
        return SHIPFEM_INVALID_ARG;
    }
    std::string dev;
    if (!deviceAvailable(&dev)) {
        setError(err, err_len, dev);
        return SHIPFEM_NO_DEVICE;
    }
    char* base = nullptr;
    hipError_t e = scratchOnDevice(&base);
    // every wave slot and the result are written before they are read: nothing to clear
    if (e == hipSuccess)
        e = evaluate(base, rest_position, index, node_count, tet_count, youngs_modulus, poisson_ratio, density, cfl,
                     nullptr);
    if (e != hipSuccess) {
        setError(err, err_len, std::string("HIP: ") + hipGetErrorString(e));
        return statusOf(e);
    }
    return SHIPFEM_OK;
}

}  // extern "C"
