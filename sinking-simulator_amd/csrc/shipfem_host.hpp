// shipfem_host.hpp -- host-side runtime: case preparation, device arenas, step plans.
#pragma once

#include "../../include/shipfem.h"
#include "shipfem_dev.hpp"

#include <cstdint>
#include <string>
#include <vector>

struct shipfem_mesh {
    enum Kind { Box, Panel } kind = Box;
    struct {
        double lx, ly, lz;
        int32_t nx, ny, nz;
    } box{};
    std::vector<double> X;        // [3N]
    std::vector<uint32_t> index;  // [4T]
    std::vector<uint32_t> flags;  // [N]
    std::vector<double> load;     // [3N]
};

namespace shipfem {

double uniform01(uint64_t seed, uint64_t counter);

inline uint32_t roundUp64(uint64_t n) { return static_cast<uint32_t>((n + 63) / 64 * 64); }

// A case after host preparation: rest state in FP64, stored as the FP32 SoA planes
// the kernels read, plus the wave-transposed CSR slab of the node gather.
struct HostCase {
    uint32_t N = 0, T = 0, Npad = 0, Tpad = 0, node_waves = 0, elem_waves = 0;
    uint32_t features = 0;
    // Internal orders: internal tet k is the caller's tet tperm[k], internal node i the
    // caller's node nperm[i]. Every array below is in the internal orders.
    std::vector<uint32_t> tperm, nperm;
    std::vector<double> X;        // [3N] rest positions
    uint32_t corners = 4;         // 4: tet, 8: solid-shell hex
    std::vector<uint32_t> conn;   // [corners T]
    std::vector<double> hk;       // hex: [kHexPacked][Tpad]
    std::vector<double> hgeod;    // plastic hex: [kHexGeo][Tpad]
    std::vector<double> hfail;    // plastic hex with per-element failure: [Tpad] eps_f
    std::vector<float> hgeo;      // hex: [kHexGeo][Tpad]
    double hex_omega = 0;         // hex: largest element frequency (FP64, host)
    std::vector<double> hvol;     // hex: [T] element volumes (host only)
    std::vector<float> dmi;       // [9][Tpad]
    std::vector<float> vol;       // [Tpad]
    std::vector<float> mass, invm;  // [Npad]
    std::vector<uint32_t> nflags;   // [Npad]
    std::vector<float> load;        // [3][Npad]
    std::vector<float> xc;          // [4][Npad]
    std::vector<float> presc;       // [3][Npad]
    std::vector<uint32_t> ell_off;  // [node_waves + 1]
    std::vector<uint32_t> ell;
    // tiled element pass (256-tet tiles)
    uint32_t elem_tiles = 0, P = 0, lmax = 0;
    std::vector<uint32_t> lconn;    // [2T] per corner: local node | rank in its run << 10
    std::vector<uint32_t> tn_off;   // [elem_tiles + 1]
    std::vector<uint32_t> tnode;    // [P]
    std::vector<uint16_t> lofs;     // [P + elem_tiles]
    std::vector<float> trest;       // [4P] each (tile, local node)'s rest position relative to
                                    // the tile's centroid (FP32 of the FP64 difference), w = 0
    std::vector<uint32_t> pell_off; // [node_waves + 1]
    std::vector<uint32_t> pell;
    std::vector<float> u4;          // [4N]
    std::vector<float> cu, v, cv;   // [3][Npad]
    // material and loading
    double E = 0, nu = 0, rho = 0, lam = 0, mu = 0, bulk = 0, yield = 0, hard = 0;
    double fail_strain = 0, fail_sens = 0, fail_eref = 0, fail_ecut = 0;  // kFeatDamage
    double load_ramp = 0, drive_ramp = 0, gravity = 0, damping = 1;
    shipfem_indenter indenter{};
    double dt_request = 0, dt_safety = 0.9;
    int32_t sample_every = 100;
    int64_t max_steps = 0;
    int32_t hist_cap = 0;
    double edge_dt = 0;  // the reference's formula at safety 0.5, for the record
    // sqrt(min m_caller / m_density) over free nodes with mass, when the caller gives the
    // lumped masses (1 otherwise): the factor on the density-based stable step
    double cfl_scale = 1.0;
    double total_mass = 0, total_volume = 0;
};

// The caller's ids of an internal tet and node.
inline uint32_t externalTet(const HostCase& hc, uint32_t t) { return hc.tperm.empty() ? t : hc.tperm[t]; }
inline uint32_t externalNode(const HostCase& hc, uint32_t n) { return hc.nperm.empty() ? n : hc.nperm[n]; }

// Validates and prepares; returns an empty string on success, else the reason. The
// case is laid out in the internal tet and node orders; results go back through
// externalTet / externalNode.
std::string prepareCase(const shipfem_case_desc& d, HostCase& hc, int32_t hist_cap = -1);
// The same for a solid-shell hex case: index holds 8 node ids per element (tet_count
// elements), kept in the caller's orders.
std::string prepareHexCase(const shipfem_case_desc& d, HostCase& hc, int32_t hist_cap = -1);
// Hex rest state (hex.cpp): the packed stiffness block, centre gradients and rest offsets
// of every element, nodal masses accumulated into `mass`, and hc.hex_omega.
std::string hexRestState(const shipfem_case_desc& d, HostCase& hc, std::vector<double>& mass);
// FP64 solid-shell element (hex.cpp).
bool hexStiffness(const double* x, double E, double nu, double K[576]);
bool hexMass(const double* x, double rho, double m[8]);
bool hexCentreGradient(const double* x, double b0[8][3]);
double hexOmegaMax(const double K[576], const double m[8]);
// Polar factor of the centre deformation gradient (row-major); identity and false for a
// rest element whose centre Jacobian is singular.
bool hexRotation(const double* X, const double* x, double R[3][3]);
void hexForce(const double K[576], const double* X, const double* x, double out[24]);

// Device bytes of a prepared case (its arena), and the admission upper bound.
uint64_t arenaBytes(const HostCase& hc);

// The HBM bytes one element pass and one node pass of a case must move (a step that is not
// a sample step): every plane the pass reads, once, and every plane it writes, once. The
// J2 history a yielding tet writes back is not counted (an elastic tet writes none), nor
// the per-wave FP64 sums (8 bytes per 64 items). `variant`: kTiled or kSlots.
struct PassBytes {
    uint64_t elem_read = 0, elem_write = 0, node_read = 0, node_write = 0;
};
// The bytes one step's passes must move; `yielded` tets (or, for a plastic hex case,
// integration points) with eqps not +0 also have their plastic strain and damage read.
// `formed`: the tiled pass forms Dm^-1 and V from the tiles' rest positions.
PassBytes compulsoryBytes(const HostCase& hc, int variant, uint64_t yielded = 0, bool formed = false);

// One case's device memory: a single arena holding every plane of the case.
struct DeviceCase {
    char* arena = nullptr;
    uint64_t bytes = 0;
    bool owns = false;  // allocated by hipMalloc (else borrowed from a pool)
    CaseDev rec{};      // device pointers into the arena, scalars filled in
    double cfl = 0, dt = 0, omega = 0;
};

// Tets of a plastic tet case whose eqps is not +0, or integration points of a plastic hex
// case whose eqps is not +0 (read from the device).
hipError_t yieldedTets(const DeviceCase& dc, const HostCase& hc, hipStream_t s, uint64_t* count);

// Lays out `hc` in the arena at `base` (which must hold arenaBytes(hc)) and fills
// rec's pointers and scalars (dt excluded). Uploads on `stream`.
hipError_t uploadCase(const HostCase& hc, char* base, DeviceCase& dc, hipStream_t stream);
// Sets dt (and everything derived from it) in the record.
void setTimestep(DeviceCase& dc, const HostCase& hc, double dt);

// A set of cases stepping together.
struct Plan {
    std::vector<CaseDev> recs;     // host copies
    CaseDev* d_recs = nullptr;     // device array
    // Element and node tiles: the cases the tiled pass runs first (elem_split and node_split
    // tiles), then those on the slots pass (every case of a SLOTS plan, a hex plan's cases,
    // and tet cases whose mesh cannot be tiled). A case is in one group only, so each runs
    // the same kernel and sum order whatever else is in the plan.
    std::vector<Tile> elem_tiles, node_tiles, omega_tiles;
    size_t elem_split = 0, node_split = 0;
    uint32_t feat_tiled = 0, feat_slots = 0;
    Tile* d_elem = nullptr;
    Tile* d_node = nullptr;
    Tile* d_omega = nullptr;
    double* d_omega_out = nullptr;
    LaunchShape shape;
    uint32_t feat_union = 0;
    uint32_t lmax = 0;  // largest tile node count over the plan's cases (tiled LDS stage)
    bool hex = false;   // the plan's cases are solid-shell hexes (all or none)
    bool resident = false;  // every case small enough for one block: whole chunks in one launch
    int parity = 0;
    int chunk = 0;
    hipGraphExec_t graph = nullptr;
    hipGraph_t graph_src = nullptr;
    size_t tile_cap = 0, case_cap = 0;

    ~Plan();
    // (Re)builds tiles and the device record array for `recs` with `shape`.
    hipError_t build(hipStream_t stream);
    // Uploads the host record copies only (after a scalar or feature change), and takes the
    // groups' feature sets from them.
    hipError_t uploadRecords(hipStream_t stream);
    bool inTiledGroup(const CaseDev& r) const;
    void refreshFeatures();
    void dropGraph();
    // The element and node launches of one step at `parity` (no reduce).
    void launchStep(int parity, int sample, hipStream_t stream);
    void launchElem(int parity, int sample, hipStream_t stream,
                    const KernelTiming* timing = nullptr);
    void launchNode(int parity, int sample, hipStream_t stream,
                    const KernelTiming* timing = nullptr);
    // One step as plain launches.
    void stepPlain(bool sample, hipStream_t stream);
    // Captures `chunk` (even) steps starting at parity 0, sample on the first.
    hipError_t capture(int chunk, hipStream_t stream);
    // Evaluates the FP64 stable step of every case: returns omega_max per case.
    hipError_t omegaMax(std::vector<double>& out, hipStream_t stream);
};

// Copies a case's state back to the host view.
hipError_t readbackCase(const DeviceCase& dc, const HostCase& hc, shipfem_state_view* view,
                        hipStream_t stream);
hipError_t readEnergy(const DeviceCase& dc, shipfem_energy_record* out, int32_t max,
                      int32_t* count, hipStream_t stream);
// The case's tears (kFeatDamage), ordered by step then tet, normals from the stored
// world-frame stress.
hipError_t readTears(const DeviceCase& dc, const HostCase& hc, std::vector<shipfem_tear>& out,
                     hipStream_t stream);
// Largest principal direction of a symmetric stress (Voigt, tensor shears), largest
// component positive.
void maxPrincipalDirection(const double s6[6], double out[3]);

bool shapeFrom(const shipfem_launch_config* cfg, LaunchShape& out);
// The default launch of a hex plan (plastic: the plastic pass's block).
void hexShape(LaunchShape& sh, bool plastic);
void setError(char* err, size_t len, const std::string& msg);
shipfem_status statusOf(hipError_t e);
bool deviceAvailable(std::string* why);

}  // namespace shipfem

// One case on its own stream (the opaque shipfem_solver of the C ABI).
struct shipfem_solver {
    shipfem::HostCase hc;
    shipfem::DeviceCase dc;
    shipfem::Plan plan;
    hipStream_t stream = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int64_t host_step = 0;
    ~shipfem_solver() {
        plan.dropGraph();
        if (dc.owns && dc.arena) (void)hipFree(dc.arena);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (stream) (void)hipStreamDestroy(stream);
    }
};
