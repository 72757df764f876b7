// shipfem_dev.hpp -- device-visible case record and kernel launch interface.
//
// One CaseDev per case lives in device memory; the element and node kernels read it
// through a tile table whose entries never span two cases, so every case's
// arithmetic and every per-wave reduction is the same whatever else is in the launch.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace shipfem {

constexpr uint32_t kWave = 64;
constexpr uint32_t kEllSentinel = 0xFFFFFFFFu;

// node flag bits (mirror include/shipfem.h)
constexpr uint32_t kPinX = 0x01, kPinY = 0x02, kPinZ = 0x04, kPinAll = 0x07;
constexpr uint32_t kLoaded = 0x08, kContact = 0x10;
constexpr uint32_t kDriveX = 0x20, kDriveY = 0x40, kDriveZ = 0x80;

// case feature bits: which optional terms the node/element passes evaluate
enum Feature : uint32_t {
    kFeatPlastic = 1u << 0,
    kFeatLoad = 1u << 1,
    kFeatContact = 1u << 2,
    kFeatDrive = 1u << 3,
    kFeatDamp = 1u << 4,
    kFeatGravity = 1u << 5,
    kFeatDamage = 1u << 6,  // ductile damage and element deletion (needs kFeatPlastic)
};

// per-wave FP64 accumulator slots (running sums, read-modify-write every step)
enum { kAccDiss = 0, kAccFract = 1, kAccElemSlots = 2 };
enum { kAccWext = 0, kAccWcon = 1, kAccWbnd = 2, kAccDamp = 3, kAccNodeSlots = 4 };
// per-wave FP64 sample partials (written on sample steps only)
enum { kSmpStrain = 0, kSmpElemSlots = 1 };
enum { kSmpKin = 0, kSmpHalfExt = 1, kSmpHalfCon = 2, kSmpFx = 3, kSmpFy = 4, kSmpFz = 5,
       kSmpHalfBnd = 6, kSmpHalfDamp = 7, kSmpNodeSlots = 8 };

// Same layout as shipfem_energy_record.
struct EnergyRecordDev {
    int64_t step;
    double time;
    double strain, kinetic, dissipation, w_ext, w_contact, w_boundary, damping_loss;
    double fracture;
    double contact_force[3];
};

// Solid-shell hex: the 21 x 21 block of the condensed stiffness that couples corners
// 1..7, packed upper triangle row by row (entry (i, j), i <= j, at i*21 - i(i-1)/2 + j - i).
constexpr int kHexDof = 21;
constexpr int kHexPacked = kHexDof * (kHexDof + 1) / 2;  // 231
constexpr int kHexGeo = 42;  // centre gradients of corners 1..7, then X_a - X_0
// Plastic solid-shell hex: per element 8 points x (plastic strain 6, eqps, damage) plus the 7
// enhanced parameters.
constexpr int kHexPoint = 8 * 8;  // 64

struct CaseDev {
    uint32_t N, T;            // nodes, elements (tets, or hexes when corners == 8)
    uint32_t Npad, Tpad;      // plane strides (multiples of 64)
    uint32_t node_waves, elem_waves;
    uint32_t features;
    uint32_t corners;         // 4: tet, 8: solid-shell hex
    // ---- element arrays
    const uint4* conn;   // [T] node ids (tet), or [2T] (hex: corners 0-3, then 4-7)
    const float* dmi;    // [9][Tpad] Dm^-1, row-major entry k = 3*row + col
    const float* vol;    // [Tpad]
    float* force;        // [T][3 corners] per-element nodal forces, reference slot layout
                         // (the slot of corner entry x = corners * element + corner is 3x)
    const double* hk;    // hex: [kHexPacked][Tpad] stiffness block of corners 1..7 (FP64)
    const float* hgeo;   // hex: [kHexGeo][Tpad] d N_a/dX at the centre (21: 3(a-1) + j),
                         // then rest X_a - X_0 (21: 21 + 3(a-1) + i), corners a = 1..7
    const double* hgeod; // plastic hex: [kHexGeo][Tpad] FP64, d N_a/dX at the centre as in
                         // hgeo, then the rest offsets' coefficients over the monomials
                         // (xi, eta, zeta, xi eta, eta zeta, zeta xi, xi eta zeta), 3 m + i
    double* hpoint;      // plastic hex: [kHexPoint][Tpad], entry 8 g + k
    double* henh;        // plastic hex: [7][Tpad] enhanced parameters (warm start)
    const double* hfail; // plastic hex: [Tpad] each element's failure strain, or null (fail_d)
    float* pstate;       // [7 or 8][Tpad] plastic strain (6, eng. shears), eqps, damage
    int32_t* tear_step;  // [Tpad] step + 1 at which the tet tore, 0 intact (kFeatDamage)
    float* tear_stress;  // [6][Tpad] world-frame stress at the tear (kFeatDamage)
    // ---- node arrays
    float4* u4;          // [N] displacement (xyz, w = 0), gathered by the element pass
    float* cu;           // [3][Npad] Kahan compensation of u
    float* v;            // [3][Npad] half-step velocity
    float* cv;           // [3][Npad] Kahan compensation of v
    const float* invm;   // [Npad]
    const float* mass;   // [Npad]
    const uint32_t* nflags;  // [Npad]
    const float* load;   // [3][Npad] nodal load at full ramp (kFeatLoad)
    const float* xc;     // [4][Npad] X - c0 (3 planes) and |X - c0|^2 - r^2 (kFeatContact)
    const float* presc;  // [3][Npad] prescribed displacement (kFeatDrive)
    const uint32_t* ell_off;  // [node_waves + 1] first entry of each wave's slab
    const uint32_t* ell;      // wave slabs, entry e of lane l at ell_off[w] + 64 e + l
    // ---- tiled element pass: 256-tet tiles staged through LDS
    const uint2* lconn;       // [T] per corner (16 bits): tile-local node | rank << 10, the
                              // corner's place in that node's run of corner slots
    const uint32_t* tn_off;   // [elem_tiles + 1] first (tile, local node) slot of each tile
    const uint32_t* tnode;    // [P] global node id of each (tile, local node) slot
    const float4* trest;      // [P] its rest position relative to the tile's centroid
    const uint16_t* lofs;     // [P + elem_tiles] per tile: L+1 run offsets into its
                              // 4 kTile corner slots
    float4* partial;          // [P] force sum of each (tile, local node)
    const uint32_t* pell_off; // [node_waves + 1] node -> partial slots, wave slabs
    const uint32_t* pell;
    uint32_t elem_tiles, P, lmax, pad3;
    // ---- scalars
    double dt;
    double load_ramp, drive_ramp;
    double ind_vel[3];
    float dtf, lam, mu, bulk;
    float yield, hard;
    float gravity, damping;
    float ind_r, ind_kdt2;   // radius, kappa / dt^2
    float fail_strain;       // regularised eps_f (kFeatDamage)
    float fail_sens;         // Rice-Tracey sensitivity
    float fail_eref, fail_ecut;  // reference and cutoff triaxiality
    float pad1[2];
    double lam_d, mu_d;       // FP64 Lame constants for the stable-step evaluation
    double yield_d, hard_d;   // FP64 flow curve and failure constants (plastic hexes)
    double fail_d, sens_d, eref_d, ecut_d;
    double hret_d[5];         // plastic hexes: the return's constants (hexp::Material's derived)
    // ---- FP64 accounting
    double* acc_e;  // [kAccElemSlots][elem_waves]
    double* acc_n;  // [kAccNodeSlots][node_waves]
    double* smp_e;  // [kSmpElemSlots][elem_waves]
    double* smp_n;  // [kSmpNodeSlots][node_waves]
    double* omega_w;  // [elem_waves] per-wave max of lambda_max(B^T C B) (set-up)
    // ---- time and status
    int64_t* step;    // [2] ping-pong step counter
    int64_t step_end; // stop advancing at this step
    uint32_t* status; // [1] nonzero: non-finite velocity seen
    EnergyRecordDev* hist;
    int32_t hist_cap;
    int32_t pad2;
    int32_t* hist_count;
};

// One block of a pass: which case and the first item (tet or node) it covers.
struct Tile {
    int32_t case_index;
    uint32_t first;
};

// Element-pass variants. kSlots: one lane per tet writes its four corner forces to a
// 48-byte slot (the reference's tetForce layout) and the node pass gathers slots
// through the CSR node map. kTiled: a 256-tet tile stages its nodes' displacement in
// LDS and sums corner forces per local node in LDS; the node pass gathers one partial
// per (tile, node).
enum Variant : int { kSlots = 0, kTiled = 1 };
constexpr int kTile = 256;  // tets per tile of the tiled variant (fixed: it sets the sum order)

struct LaunchShape {
    int elem_block = 256;
    int node_block = 256;
    int xcd_remap = 1;
    int variant = kTiled;
    int split = 1;   // a batch: up to four parts on as many streams
    int formed = 0;    // tiled pass: form Dm^-1 and V from the tile's rest positions
    int resident = 1;  // plans of small tiled cases step whole chunks in one launch
};

// Start/stop events recorded by the kernel dispatch itself (hipExtLaunchKernel), so
// their elapsed time is the kernel's execution alone.
struct KernelTiming {
    hipEvent_t start = nullptr, stop = nullptr;
};

// Kernel launchers (kernels.hip). All launch asynchronously on `stream`; a non-null
// `timing` times the launch (never used under graph capture).
// For the tiled variant the tiles are 256-tet tiles, `block` is ignored and `lmax` sizes
// the LDS node stage. launch_elem_tiled and launch_node take a null `tiles` for a plan of
// one case: block b then covers items [b * block, (b + 1) * block) of cases[0].
void launch_elem(const CaseDev* cases, const Tile* tiles, int ntiles, int block, int parity,
                 int sample, int xcd, uint32_t feat_union, hipStream_t stream,
                 const KernelTiming* timing = nullptr);
// The solid-shell hex element pass: one lane per element, `block` threads per block.
void launch_elem_hex(const CaseDev* cases, const Tile* tiles, int ntiles, int block, int parity,
                     int sample, int xcd, hipStream_t stream, const KernelTiming* timing = nullptr);
// The plastic solid-shell hex element pass: one lane per element, `block` (64, 128 or 256;
// larger is launched as 256) threads per block.
void launch_elem_hex_plastic(const CaseDev* cases, const Tile* tiles, int ntiles, int block,
                             int parity, int sample, int xcd, hipStream_t stream,
                             const KernelTiming* timing = nullptr);
void launch_elem_tiled(const CaseDev* cases, const Tile* tiles, int ntiles, uint32_t lmax,
                       int parity, int sample, int xcd, uint32_t feat_union, bool formed, hipStream_t stream,
                       const KernelTiming* timing = nullptr);
void launch_node(const CaseDev* cases, const Tile* tiles, int ntiles, int block, int parity,
                 int sample, int xcd, uint32_t feat_union, int variant, hipStream_t stream,
                 const KernelTiming* timing = nullptr);
void launch_reduce(const CaseDev* cases, int ncases, int parity, hipStream_t stream);
// Small cases (at most 4 tiles and 2048 nodes each), one 1024-thread block per case: `steps`
// whole steps in one launch, the first sampled when sample_first (k_resident).
void launch_resident(const CaseDev* cases, int ncases, uint32_t lmax, int steps, int parity0, int sample_first,
                     uint32_t feat_union, bool formed, hipStream_t stream);
// Dynamic LDS bytes that launch takes.
size_t resident_lds(uint32_t lmax);
// FP64 largest eigenvalue of B^T C B per element (the lumped-mass element frequency
// is omega^2 = 4 lambda / rho), reduced to a per-wave max, then to one value per case.
void launch_omega(const CaseDev* cases, const Tile* tiles, int ntiles, hipStream_t stream);
void launch_omega_max(const CaseDev* cases, int ncases, double* out, hipStream_t stream);
// out[2i] = current step of case i, out[2i+1] = its status word.
void launch_progress(const CaseDev* cases, int ncases, int64_t* out, hipStream_t stream);

}  // namespace shipfem
