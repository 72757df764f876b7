// api.cpp -- the single-case solver (gpu::FemGpuSolver's replacement) and the
// many-case batch, behind the C ABI of include/shipfem.h.
#include "shipfem_host.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <new>

using namespace shipfem;

namespace {

const char* kStatusText[] = {"ok", "no usable device", "invalid argument", "out of memory",
                             "unstable (non-finite state)", "device error", "busy",
                             "not found", "unsupported"};

// Advances all cases of `p` by `steps`, replaying the captured chunk when the
// cases sit on a chunk boundary. Samples (energy records) fall on steps n with
// n % S == 0, in both modes.
hipError_t advance(Plan& p, int64_t& host_step, int64_t steps, int mode, int S, hipStream_t s) {
    while (steps > 0) {
        if (mode == 0 && p.parity == 0 && host_step % S == 0 && steps >= S) {
            if (!p.graph || p.chunk != S) {
                const hipError_t e = p.capture(S, s);
                if (e != hipSuccess) return e;
            }
            const hipError_t e = hipGraphLaunch(p.graph, s);
            if (e != hipSuccess) return e;
            host_step += S;
            steps -= S;
        } else {
            p.stepPlain(host_step % S == 0, s);
            ++host_step;
            --steps;
        }
    }
    return hipGetLastError();
}

// The board's name: the runtime's marketing name, else the PCI board's product name
// from sysfs (some boxes leave the runtime's empty), else the PCI address.
std::string boardName(const hipDeviceProp_t& prop) {
    if (prop.name[0]) return prop.name;
    char bus[32] = {0};
    std::snprintf(bus, sizeof bus, "%04x:%02x:%02x.0", prop.pciDomainID, prop.pciBusID, prop.pciDeviceID);
    for (const char* leaf : {"product_name", "product_number"}) {
        std::FILE* f = std::fopen((std::string("/sys/bus/pci/devices/") + bus + "/" + leaf).c_str(), "r");
        if (!f) continue;
        char line[128] = {0};
        const bool got = std::fgets(line, sizeof line, f) != nullptr;
        std::fclose(f);
        std::string v = got ? line : "";
        while (!v.empty() && (v.back() == '\n' || v.back() == ' ')) v.pop_back();
        if (!v.empty()) return v;
    }
    return std::string("AMD GPU at PCI ") + bus;
}

double omegaToCfl(double lambda_max, double rho) { return 2.0 / std::sqrt(4.0 * lambda_max / rho); }

bool blockOk(int x) { return x == 64 || x == 128 || x == 256 || x == 512; }

}  // namespace

namespace shipfem {
// The default launch of a hex plan: blocks dealt round-robin over the XCDs. Measured
// against XCD-contiguous blocks: the elastic plate's passes 0.518 + 0.087 ms against
// 0.541 + 0.094 ms (1.2M elements). The plastic pass is bound by FP64 arithmetic and its
// cost per element varies (yielding elements iterate): 64-element blocks balance it best,
// 0.573 against 0.606 ms on a 720,000-element punched zone.
void hexShape(LaunchShape& sh, bool plastic) {
    if (plastic) sh.elem_block = 64;
    sh.xcd_remap = 0;
}

// A launch configuration from the C ABI: zero fields keep the defaults. Returns false
// for a value the kernels are not built for.
bool shapeFrom(const shipfem_launch_config* cfg, LaunchShape& out) {
    LaunchShape sh;
    if (cfg) {
        if (cfg->elem_block) sh.elem_block = cfg->elem_block;
        if (cfg->node_block) sh.node_block = cfg->node_block;
        if (cfg->elem_block || cfg->node_block) sh.xcd_remap = cfg->xcd_remap ? 1 : 0;
        if (cfg->variant != SHIPFEM_VARIANT_TILED && cfg->variant != SHIPFEM_VARIANT_SLOTS) return false;
        sh.variant = cfg->variant == SHIPFEM_VARIANT_SLOTS ? kSlots : kTiled;
        if (cfg->split != 0 && cfg->split != -1) return false;
        sh.split = cfg->split == 0 ? 1 : 0;
        if (cfg->rest_from_positions != 0 && cfg->rest_from_positions != 1) return false;
        sh.formed = cfg->rest_from_positions;
        if (cfg->resident != 0 && cfg->resident != -1) return false;
        sh.resident = cfg->resident == 0 ? 1 : 0;
    }
    if (!blockOk(sh.elem_block) || !blockOk(sh.node_block)) return false;
    out = sh;
    return true;
}
}  // namespace shipfem

struct shipfem_batch {
    std::vector<HostCase> hcs;
    std::vector<DeviceCase> dcs;
    Plan plan;  // every case: profiling, the tuner's reference, one-stream runs
    // With shape.split and two or more cases: the cases in `groups` parts, each advanced by
    // its own plan on its own stream (part[0] on `stream`, part[g] on `side[g - 1]`), forked
    // from and joined back into `stream`. Every plan keeps the same parity. Four parts, one per
    // hardware queue the runtime gives a process by default: on the 64-case batch of 4.06M
    // tets they step in 63-65 us against 65-69 us for two parts (bit-identical states).
    static constexpr int kGroups = 4;
    Plan part[kGroups];
    int groups = 0;
    bool split = false;
    char* arena = nullptr;
    char* backup = nullptr;
    uint64_t bytes = 0;
    hipStream_t stream = nullptr, side[kGroups - 1] = {};
    hipEvent_t e0 = nullptr, e1 = nullptr, fork = nullptr, join[kGroups - 1] = {};
    int64_t host_step = 0;
    int S = 100;
    ~shipfem_batch() {
        plan.dropGraph();
        for (Plan& h : part) h.dropGraph();
        if (arena) (void)hipFree(arena);
        if (backup) (void)hipFree(backup);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (fork) (void)hipEventDestroy(fork);
        for (hipEvent_t j : join)
            if (j) (void)hipEventDestroy(j);
        for (hipStream_t t : side)
            if (t) (void)hipStreamDestroy(t);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {
// Splits the batch's cases into parts of nearly equal element count (largest first, each
// to the lightest part) and builds their plans, when the shape asks for it.
hipError_t buildParts(shipfem_batch* b) {
    b->split = b->plan.shape.split && b->plan.recs.size() >= 2;
    b->groups = b->split ? static_cast<int>(std::min<size_t>(shipfem_batch::kGroups, b->plan.recs.size())) : 0;
    if (!b->split) {
        for (Plan& h : b->part) h.dropGraph();
        return hipSuccess;
    }
    std::vector<size_t> order(b->plan.recs.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(),
                     [&](size_t x, size_t y) { return b->plan.recs[x].T > b->plan.recs[y].T; });
    uint64_t load[shipfem_batch::kGroups] = {};
    std::vector<size_t> members[shipfem_batch::kGroups];
    for (size_t i : order) {
        int h = 0;
        for (int g = 1; g < b->groups; ++g)
            if (load[g] < load[h]) h = g;
        members[h].push_back(i);
        load[h] += b->plan.recs[i].T;
    }
    hipError_t e = hipSuccess;
    for (int h = 0; h < b->groups && e == hipSuccess; ++h) {
        std::sort(members[h].begin(), members[h].end());
        Plan& p = b->part[h];
        p.recs.clear();
        for (size_t i : members[h]) p.recs.push_back(b->plan.recs[i]);
        p.shape = b->plan.shape;
        p.parity = b->plan.parity;
        e = p.build(b->stream);
    }
    return e;
}

// Advances every case `steps` steps: on `stream` alone, or the parts side by side.
hipError_t batchAdvance(shipfem_batch* b, int64_t steps) {
    if (!b->split) {
        const hipError_t e = advance(b->plan, b->host_step, steps, 0, b->S, b->stream);
        for (Plan& h : b->part) h.parity = b->plan.parity;
        return e;
    }
    const int64_t start = b->host_step;
    hipError_t e = hipEventRecord(b->fork, b->stream);
    for (int g = 1; g < b->groups && e == hipSuccess; ++g) e = hipStreamWaitEvent(b->side[g - 1], b->fork, 0);
    for (int g = 0; g < b->groups && e == hipSuccess; ++g) {
        int64_t at = start;
        e = advance(b->part[g], at, steps, 0, b->S, g ? b->side[g - 1] : b->stream);
        b->host_step = at;
    }
    for (int g = 1; g < b->groups && e == hipSuccess; ++g) {
        e = hipEventRecord(b->join[g - 1], b->side[g - 1]);
        if (e == hipSuccess) e = hipStreamWaitEvent(b->stream, b->join[g - 1], 0);
    }
    b->plan.parity = b->part[0].parity;
    return e;
}
}  // namespace

extern "C" {

const char* shipfem_status_string(shipfem_status s) {
    const int i = static_cast<int>(s);
    return (i >= 0 && i < static_cast<int>(sizeof(kStatusText) / sizeof(kStatusText[0]))) ? kStatusText[i]
                                                                                            : "unknown";
}

shipfem_status shipfem_device_query(char* name, size_t name_len, uint64_t* hbm_bytes) {
    std::string why;
    if (!deviceAvailable(&why)) {
        setError(name, name_len, why);
        return SHIPFEM_NO_DEVICE;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, 0) != hipSuccess) return SHIPFEM_DEVICE_ERROR;
    setError(name, name_len, boardName(prop) + " (" + prop.gcnArchName + ")");
    if (hbm_bytes) *hbm_bytes = prop.totalGlobalMem;
    return SHIPFEM_OK;
}

namespace {
// arenaBytes with every optional plane present, no node in more than `maxdeg` elements (the
// width of a wave's gather slabs) and at most P (tile, node) slots.
uint64_t footprintBound(uint32_t N, uint32_t T, int32_t plastic, int32_t cap, uint64_t maxdeg, uint64_t P) {
    const uint64_t Np = roundUp64(N), Tp = roundUp64(T), nw = Np / 64, ew = Tp / 64;
    const uint64_t nt = (T + kTile - 1) / kTile;
    const uint64_t pstate = plastic == 0 ? 0 : plastic == 1 ? 28ull * Tp : (32ull + 4 + 24) * Tp;
    uint64_t b = 16ull * T + 40ull * Tp + 48ull * T + pstate + 16ull * N +
                 36ull * Np + 12ull * Np + 40ull * Np + 4ull * (nw + 1) + 4ull * 64 * maxdeg * nw +
                 8ull * T + 4ull * (nt + 1) + 4ull * P + 2ull * (P + nt) +
                 16ull * P + 4ull * (nw + 1) + 4ull * 64 * maxdeg * nw +
                 8ull * ((kAccElemSlots + kSmpElemSlots + 1) * ew + (kAccNodeSlots + kSmpNodeSlots) * nw) +
                 24 + sizeof(EnergyRecordDev) * (uint64_t)std::max(cap, 1);
    return b + 256ull * 40;  // each of the arena's planes rounds up to 256 bytes
}
}  // namespace

uint64_t shipfem_case_footprint(uint32_t N, uint32_t T, int32_t plastic, int32_t cap) {
    // Kuhn-type meshes: node degree <= 32, at most two (tile, node) slots per tet (they hold
    // about 0.75)
    return footprintBound(N, T, plastic, cap, 32, 2ull * T);
}

uint64_t shipfem_case_footprint_mesh(const shipfem_case_desc* d, int32_t cap) {
    if (!d || !d->index || d->node_count == 0 || d->tet_count == 0) return 0;
    // the mesh's own largest node degree bounds every wave's slab width; a tile holds at
    // most 4 slots per tet
    std::vector<uint32_t> deg(d->node_count, 0u);
    uint64_t maxdeg = 0;
    for (uint64_t q = 0; q < 4ull * d->tet_count; ++q)
        if (d->index[q] < d->node_count) maxdeg = std::max<uint64_t>(maxdeg, ++deg[d->index[q]]);
    const shipfem_material& m = d->material;
    const int32_t kind = m.plastic ? (m.failure_strain > 0.0 ? 2 : 1) : 0;
    return footprintBound(d->node_count, d->tet_count, kind, cap, maxdeg, 4ull * d->tet_count);
}

namespace {

// One case on its own stream: prepares (tet or hex), uploads, builds the plan and sets
// dt from the largest element frequency (tets: FP64 on the device; hexes: FP64 power
// iteration on the host during preparation).
shipfem_status createSolver(const shipfem_case_desc* desc, bool hex, shipfem_solver** out,
                            char* err, size_t err_len) {
    if (!desc || !out) {
        setError(err, err_len, "null argument");
        return SHIPFEM_INVALID_ARG;
    }
    *out = nullptr;
    std::string why;
    if (!deviceAvailable(&why)) {
        setError(err, err_len, why);
        return SHIPFEM_NO_DEVICE;
    }
    auto* s = new (std::nothrow) shipfem_solver();
    if (!s) return SHIPFEM_OUT_OF_MEMORY;
    const std::string msg = hex ? prepareHexCase(*desc, s->hc) : prepareCase(*desc, s->hc);
    if (!msg.empty()) {
        setError(err, err_len, msg);
        delete s;
        return SHIPFEM_INVALID_ARG;
    }
    hipError_t e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&s->e0);
    if (e == hipSuccess) e = hipEventCreate(&s->e1);
    const uint64_t bytes = arenaBytes(s->hc);
    char* base = nullptr;
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&base), bytes);
    if (e == hipSuccess) {
        s->dc.owns = true;
        e = uploadCase(s->hc, base, s->dc, s->stream);
        if (e != hipSuccess) s->dc.arena = base;
    } else if (base) {
        (void)hipFree(base);
    }
    if (e == hipSuccess) {
        s->plan.recs = {s->dc.rec};
        if (hex) hexShape(s->plan.shape, desc->material.plastic != 0);
        e = s->plan.build(s->stream);
    }
    std::vector<double> lmax;
    if (e == hipSuccess && !hex) e = s->plan.omegaMax(lmax, s->stream);
    if (e != hipSuccess) {
        setError(err, err_len, std::string("HIP: ") + hipGetErrorString(e));
        const shipfem_status st = statusOf(e);
        delete s;
        return st;
    }
    if (hex) {
        s->dc.omega = s->hc.hex_omega;
        s->dc.cfl = 2.0 / s->hc.hex_omega * s->hc.cfl_scale;
    } else {
        s->dc.omega = std::sqrt(4.0 * lmax[0] / s->hc.rho);
        s->dc.cfl = omegaToCfl(lmax[0], s->hc.rho) * s->hc.cfl_scale;
    }
    setTimestep(s->dc, s->hc, s->hc.dt_request > 0 ? s->hc.dt_request : s->hc.dt_safety * s->dc.cfl);
    s->plan.recs[0] = s->dc.rec;
    e = s->plan.uploadRecords(s->stream);
    if (e != hipSuccess) {
        setError(err, err_len, std::string("HIP: ") + hipGetErrorString(e));
        delete s;
        return statusOf(e);
    }
    *out = s;
    return SHIPFEM_OK;
}

}  // namespace

shipfem_status shipfem_solver_create(const shipfem_case_desc* desc, shipfem_solver** out,
                                     char* err, size_t err_len) {
    return createSolver(desc, false, out, err, err_len);
}

shipfem_status shipfem_hex_solver_create(const shipfem_case_desc* desc, shipfem_solver** out,
                                         char* err, size_t err_len) {
    return createSolver(desc, true, out, err, err_len);
}

shipfem_status shipfem_solver_timestep(const shipfem_solver* s, double* cfl, double* dt,
                                       double* edge_dt) {
    if (!s) return SHIPFEM_INVALID_ARG;
    if (cfl) *cfl = s->dc.cfl;
    if (dt) *dt = s->dc.dt;
    if (edge_dt) *edge_dt = s->hc.edge_dt;
    return SHIPFEM_OK;
}

shipfem_status shipfem_solver_set_params(shipfem_solver* s, const shipfem_step_params* p) {
    if (!s || !p) return SHIPFEM_INVALID_ARG;
    if (!(p->dt > 0) || !std::isfinite(p->gravity)) return SHIPFEM_INVALID_ARG;
    const double damping = p->damping == 0.0 ? 1.0 : p->damping;
    if (!(damping > 0 && damping <= 1.0)) return SHIPFEM_INVALID_ARG;
    HostCase& hc = s->hc;
    if (p->youngs_modulus > 0) hc.E = p->youngs_modulus;
    if (p->poisson_ratio > 0 && p->poisson_ratio < 0.5) hc.nu = p->poisson_ratio;
    hc.lam = hc.E * hc.nu / ((1 + hc.nu) * (1 - 2 * hc.nu));
    hc.mu = hc.E / (2 * (1 + hc.nu));
    hc.bulk = hc.E / (3 * (1 - 2 * hc.nu));
    hc.gravity = p->gravity;
    hc.damping = damping;
    CaseDev& r = s->dc.rec;
    r.lam = static_cast<float>(hc.lam);
    r.mu = static_cast<float>(hc.mu);
    r.bulk = static_cast<float>(hc.bulk);
    r.lam_d = hc.lam;
    r.mu_d = hc.mu;
    r.gravity = static_cast<float>(hc.gravity);
    r.damping = static_cast<float>(hc.damping);
    hc.features = (hc.features & ~(kFeatGravity | kFeatDamp)) | (hc.gravity != 0 ? kFeatGravity : 0) |
                  (hc.damping != 1.0 ? kFeatDamp : 0);
    r.features = hc.features;
    setTimestep(s->dc, hc, p->dt);
    s->plan.recs[0] = r;
    s->plan.dropGraph();
    return statusOf(s->plan.uploadRecords(s->stream));
}

shipfem_status shipfem_solver_run(shipfem_solver* s, int32_t steps, int32_t mode, double* device_ms) {
    if (!s || steps < 0 || (mode != 0 && mode != 1)) return SHIPFEM_INVALID_ARG;
    hipError_t e = hipEventRecord(s->e0, s->stream);
    if (e == hipSuccess) e = advance(s->plan, s->host_step, steps, mode, s->hc.sample_every, s->stream);
    if (e == hipSuccess) e = hipEventRecord(s->e1, s->stream);
    if (e == hipSuccess) e = hipEventSynchronize(s->e1);
    float ms = 0;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, s->e0, s->e1);
    if (e != hipSuccess) return statusOf(e);
    if (device_ms) *device_ms = ms;
    uint32_t status = 0;
    e = hipMemcpy(&status, s->dc.rec.status, 4, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return statusOf(e);
    return status ? SHIPFEM_UNSTABLE : SHIPFEM_OK;
}

shipfem_status shipfem_solver_readback(shipfem_solver* s, shipfem_state_view* view) {
    if (!s || !view) return SHIPFEM_INVALID_ARG;
    return statusOf(readbackCase(s->dc, s->hc, view, s->stream));
}

shipfem_status shipfem_solver_energy(shipfem_solver* s, shipfem_energy_record* out, int32_t max,
                                     int32_t* count) {
    if (!s) return SHIPFEM_INVALID_ARG;
    return statusOf(readEnergy(s->dc, out, max, count, s->stream));
}

namespace {
shipfem_status copyTears(const std::vector<shipfem_tear>& tears, shipfem_tear* out, uint32_t max,
                         uint32_t* count) {
    if (count) *count = static_cast<uint32_t>(tears.size());
    if (out)
        for (uint32_t k = 0; k < max && k < tears.size(); ++k) out[k] = tears[k];
    return SHIPFEM_OK;
}
}  // namespace

shipfem_status shipfem_solver_tears(shipfem_solver* s, shipfem_tear* out, uint32_t max,
                                    uint32_t* count) {
    if (!s || (!out && max)) return SHIPFEM_INVALID_ARG;
    std::vector<shipfem_tear> tears;
    const hipError_t e = readTears(s->dc, s->hc, tears, s->stream);
    if (e != hipSuccess) return statusOf(e);
    return copyTears(tears, out, max, count);
}

shipfem_status shipfem_solver_tet_forces(shipfem_solver* s, float* out) {
    if (!s || !out) return SHIPFEM_INVALID_ARG;
    // only the slots variant keeps per-tet force slots; the tiled one sums in LDS (a hex
    // case always writes slots)
    if (s->plan.shape.variant != kSlots && !s->plan.hex && s->hc.elem_tiles > 0) return SHIPFEM_UNSUPPORTED;
    const HostCase& hc = s->hc;
    const uint64_t per = 3ull * hc.corners;
    std::vector<float> f(per * hc.T);
    hipError_t e = hipMemcpyAsync(f.data(), s->dc.rec.force, 4 * f.size(), hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) return statusOf(e);
    for (uint32_t t = 0; t < hc.T; ++t)  // back in the caller's element order (corners keep theirs)
        std::memcpy(out + per * externalTet(hc, t), f.data() + per * t, 4 * per);
    return SHIPFEM_OK;
}

shipfem_status shipfem_solver_set_launch(shipfem_solver* s, const shipfem_launch_config* cfg) {
    if (!s || !cfg) return SHIPFEM_INVALID_ARG;
    LaunchShape sh;
    if (!shapeFrom(cfg, sh)) return SHIPFEM_INVALID_ARG;
    // a captured chunk starts at parity 0; changing the launch keeps the cadence
    s->plan.shape = sh;
    const int parity = s->plan.parity;
    const hipError_t e = s->plan.build(s->stream);
    s->plan.parity = parity;
    return statusOf(e);
}

namespace {
// Per-kernel timing of `steps` plain steps of a plan: every element and node launch carries
// its own start/stop events (recorded by the dispatch), so the elapsed time is the
// kernel's execution. Samples fall on host steps that are multiples of S, as in a run.
hipError_t profilePlan(Plan& p, int64_t& host_step, int S, int steps, hipStream_t stream, double* elem_ms,
                       double* node_ms) {
    std::vector<hipEvent_t> ev(4 * static_cast<size_t>(steps), nullptr);
    hipError_t e = hipSuccess;
    for (auto& x : ev)
        if (e == hipSuccess) e = hipEventCreate(&x);
    for (int k = 0; k < steps && e == hipSuccess; ++k) {
        const int sample = host_step % S == 0 ? 1 : 0;
        const KernelTiming te{ev[4 * k], ev[4 * k + 1]}, tn{ev[4 * k + 2], ev[4 * k + 3]};
        p.launchElem(p.parity, sample, stream, &te);
        p.launchNode(p.parity, sample, stream, &tn);
        if (sample) launch_reduce(p.d_recs, static_cast<int>(p.recs.size()), p.parity, stream);
        e = hipGetLastError();
        p.parity ^= 1;
        ++host_step;
    }
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    double te = 0, tn = 0;
    for (int k = 0; k < steps && e == hipSuccess; ++k) {
        float a = 0, b = 0;
        e = hipEventElapsedTime(&a, ev[4 * k], ev[4 * k + 1]);
        if (e == hipSuccess) e = hipEventElapsedTime(&b, ev[4 * k + 2], ev[4 * k + 3]);
        te += a;
        tn += b;
    }
    for (auto& x : ev)
        if (x) (void)hipEventDestroy(x);
    if (e == hipSuccess) {
        if (elem_ms) *elem_ms = te / steps;
        if (node_ms) *node_ms = tn / steps;
    }
    return e;
}

void addBytes(const PassBytes& b, uint64_t out[4]) {
    out[0] += b.elem_read;
    out[1] += b.elem_write;
    out[2] += b.node_read;
    out[3] += b.node_write;
}
}  // namespace

shipfem_status shipfem_solver_profile(shipfem_solver* s, int32_t steps, double* elem_ms,
                                      double* node_ms) {
    if (!s || steps < 1) return SHIPFEM_INVALID_ARG;
    return statusOf(profilePlan(s->plan, s->host_step, s->hc.sample_every, steps, s->stream, elem_ms, node_ms));
}

shipfem_status shipfem_solver_traffic(const shipfem_solver* s, uint64_t out[4]) {
    if (!s || !out) return SHIPFEM_INVALID_ARG;
    for (int k = 0; k < 4; ++k) out[k] = 0;
    uint64_t yielded = 0;
    const hipError_t e = yieldedTets(s->dc, s->hc, s->stream, &yielded);
    if (e != hipSuccess) return statusOf(e);
    addBytes(compulsoryBytes(s->hc, s->plan.shape.variant, yielded, s->plan.shape.formed != 0), out);
    return SHIPFEM_OK;
}

shipfem_status shipfem_solver_footprint(const shipfem_solver* s, uint64_t* bytes) {
    if (!s || !bytes) return SHIPFEM_INVALID_ARG;
    *bytes = s->dc.bytes;
    return SHIPFEM_OK;
}

void shipfem_solver_destroy(shipfem_solver* s) {
    if (s) (void)hipStreamSynchronize(s->stream);
    delete s;
}

// ---------------------------------------------------------------- batch

namespace {

shipfem_status createBatch(const shipfem_case_desc* cases, uint32_t n,
                           const shipfem_launch_config* cfg, bool hex, shipfem_batch** out,
                           char* err, size_t err_len) {
    if (!cases || !out || n == 0) {
        setError(err, err_len, "null argument or empty batch");
        return SHIPFEM_INVALID_ARG;
    }
    *out = nullptr;
    std::string why;
    if (!deviceAvailable(&why)) {
        setError(err, err_len, why);
        return SHIPFEM_NO_DEVICE;
    }
    auto* b = new (std::nothrow) shipfem_batch();
    if (!b) return SHIPFEM_OUT_OF_MEMORY;
    b->hcs.resize(n);
    b->dcs.resize(n);
    for (uint32_t i = 0; i < n; ++i) {
        const std::string msg = hex ? prepareHexCase(cases[i], b->hcs[i]) : prepareCase(cases[i], b->hcs[i]);
        if (!msg.empty()) {
            setError(err, err_len, "case " + std::to_string(i) + ": " + msg);
            delete b;
            return SHIPFEM_INVALID_ARG;
        }
    }
    b->S = (cfg && cfg->chunk > 0) ? cfg->chunk : b->hcs[0].sample_every;
    if (b->S < 2 || (b->S & 1)) {
        setError(err, err_len, "chunk must be an even number >= 2");
        delete b;
        return SHIPFEM_INVALID_ARG;
    }
    if (!shapeFrom(cfg, b->plan.shape)) {
        setError(err, err_len, "launch config: blocks must be 64/128/256/512, variant TILED or SLOTS");
        delete b;
        return SHIPFEM_INVALID_ARG;
    }
    if (hex && !(cfg && (cfg->elem_block || cfg->node_block)))
        hexShape(b->plan.shape, cases[0].material.plastic != 0);
    std::vector<uint64_t> off(n + 1, 0);
    for (uint32_t i = 0; i < n; ++i) off[i + 1] = off[i] + (arenaBytes(b->hcs[i]) + 4095) / 4096 * 4096;
    b->bytes = off[n];
    hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    for (hipStream_t& t : b->side)
        if (e == hipSuccess) e = hipStreamCreateWithFlags(&t, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&b->e0);
    if (e == hipSuccess) e = hipEventCreate(&b->e1);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming);
    for (hipEvent_t& j : b->join)
        if (e == hipSuccess) e = hipEventCreateWithFlags(&j, hipEventDisableTiming);
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&b->arena), b->bytes);
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&b->backup), b->bytes);
    for (uint32_t i = 0; i < n && e == hipSuccess; ++i) {
        e = uploadCase(b->hcs[i], b->arena + off[i], b->dcs[i], b->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(b->stream);  // host planes stay valid
    }
    if (e == hipSuccess) {
        b->plan.recs.clear();
        for (auto& d : b->dcs) b->plan.recs.push_back(d.rec);
        e = b->plan.build(b->stream);
    }
    std::vector<double> lmax;
    if (e == hipSuccess && !hex) e = b->plan.omegaMax(lmax, b->stream);
    if (e == hipSuccess) {
        for (uint32_t i = 0; i < n; ++i) {
            HostCase& hc = b->hcs[i];
            DeviceCase& dc = b->dcs[i];
            dc.omega = hex ? hc.hex_omega : std::sqrt(4.0 * lmax[i] / hc.rho);
            dc.cfl = (hex ? 2.0 / hc.hex_omega : omegaToCfl(lmax[i], hc.rho)) * hc.cfl_scale;
            setTimestep(dc, hc, hc.dt_request > 0 ? hc.dt_request : hc.dt_safety * dc.cfl);
            b->plan.recs[i] = dc.rec;
        }
        e = b->plan.uploadRecords(b->stream);
    }
    if (e == hipSuccess) e = buildParts(b);
    if (e == hipSuccess)
        e = hipMemcpyAsync(b->backup, b->arena, b->bytes, hipMemcpyDeviceToDevice, b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    if (e != hipSuccess) {
        setError(err, err_len, std::string("HIP: ") + hipGetErrorString(e));
        const shipfem_status st = statusOf(e);
        delete b;
        return st;
    }
    // the prepared host planes are no longer needed once on the device
    for (auto& hc : b->hcs) {
        std::vector<float>().swap(hc.dmi);
        std::vector<uint32_t>().swap(hc.ell);
        std::vector<uint32_t>().swap(hc.conn);
        std::vector<uint32_t>().swap(hc.lconn);
        std::vector<uint32_t>().swap(hc.tnode);
        std::vector<float>().swap(hc.trest);
        std::vector<uint32_t>().swap(hc.pell);
        std::vector<double>().swap(hc.hk);
        std::vector<float>().swap(hc.hgeo);
        std::vector<double>().swap(hc.hgeod);
    }
    *out = b;
    return SHIPFEM_OK;
}

}  // namespace

shipfem_status shipfem_batch_create(const shipfem_case_desc* cases, uint32_t n,
                                    const shipfem_launch_config* cfg, shipfem_batch** out,
                                    char* err, size_t err_len) {
    return createBatch(cases, n, cfg, false, out, err, err_len);
}

shipfem_status shipfem_hex_batch_create(const shipfem_case_desc* cases, uint32_t n,
                                        const shipfem_launch_config* cfg, shipfem_batch** out,
                                        char* err, size_t err_len) {
    // one launch runs one element pass: the batch is all elastic or all plastic
    for (uint32_t i = 1; cases && i < n; ++i)
        if ((cases[i].material.plastic != 0) != (cases[0].material.plastic != 0)) {
            setError(err, err_len, "case " + std::to_string(i) +
                                       ": a hex batch is all elastic or all plastic");
            if (out) *out = nullptr;
            return SHIPFEM_INVALID_ARG;
        }
    return createBatch(cases, n, cfg, true, out, err, err_len);
}

shipfem_status shipfem_batch_run(shipfem_batch* b, int32_t steps, double* device_ms) {
    if (!b || steps < 0) return SHIPFEM_INVALID_ARG;
    hipError_t e = hipEventRecord(b->e0, b->stream);
    if (e == hipSuccess) e = batchAdvance(b, steps);
    if (e == hipSuccess) e = hipEventRecord(b->e1, b->stream);
    if (e == hipSuccess) e = hipEventSynchronize(b->e1);
    float ms = 0;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, b->e0, b->e1);
    if (e != hipSuccess) return statusOf(e);
    if (device_ms) *device_ms = ms;
    return SHIPFEM_OK;
}

shipfem_status shipfem_batch_profile(shipfem_batch* b, int32_t steps, double* elem_ms, double* node_ms) {
    if (!b || steps < 1) return SHIPFEM_INVALID_ARG;
    // a captured chunk starts at parity 0 on a chunk boundary: profile whole chunks
    if (b->host_step % b->S != 0 || b->plan.parity != 0 || steps % b->S != 0) return SHIPFEM_INVALID_ARG;
    const hipError_t e = profilePlan(b->plan, b->host_step, b->S, steps, b->stream, elem_ms, node_ms);
    for (Plan& h : b->part) h.parity = b->plan.parity;
    return statusOf(e);
}

shipfem_status shipfem_batch_footprint(const shipfem_batch* b, uint64_t* bytes) {
    if (!b || !bytes) return SHIPFEM_INVALID_ARG;
    *bytes = b->bytes;
    return SHIPFEM_OK;
}

shipfem_status shipfem_batch_traffic(const shipfem_batch* b, uint64_t out[4]) {
    if (!b || !out) return SHIPFEM_INVALID_ARG;
    for (int k = 0; k < 4; ++k) out[k] = 0;
    for (size_t i = 0; i < b->hcs.size(); ++i) {
        uint64_t yielded = 0;
        const hipError_t e = yieldedTets(b->dcs[i], b->hcs[i], b->stream, &yielded);
        if (e != hipSuccess) return statusOf(e);
        addBytes(compulsoryBytes(b->hcs[i], b->plan.shape.variant, yielded, b->plan.shape.formed != 0), out);
    }
    return SHIPFEM_OK;
}

shipfem_status shipfem_batch_set_config(shipfem_batch* b, const shipfem_launch_config* cfg) {
    if (!b || !cfg) return SHIPFEM_INVALID_ARG;
    LaunchShape sh;
    if (!shapeFrom(cfg, sh)) return SHIPFEM_INVALID_ARG;
    if (b->host_step % b->S != 0 || b->plan.parity != 0) return SHIPFEM_INVALID_ARG;
    b->plan.shape = sh;
    hipError_t e = b->plan.build(b->stream);
    if (e == hipSuccess) e = buildParts(b);
    return statusOf(e);
}

shipfem_status shipfem_batch_tune(shipfem_batch* b, int32_t steps_per_trial,
                                  shipfem_launch_config* best, double* best_ms) {
    if (!b || steps_per_trial < 1) return SHIPFEM_INVALID_ARG;
    if (b->host_step % b->S != 0 || b->plan.parity != 0) return SHIPFEM_INVALID_ARG;
    // snapshot the live state; every trial starts from it and it is restored at the end
    char* snap = nullptr;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&snap), b->bytes);
    if (e == hipSuccess) e = hipMemcpyAsync(snap, b->arena, b->bytes, hipMemcpyDeviceToDevice, b->stream);
    if (e != hipSuccess) {
        if (snap) (void)hipFree(snap);
        return statusOf(e);
    }
    const int64_t start = b->host_step;
    const int trial = std::max<int>(b->S, (steps_per_trial + b->S - 1) / b->S * b->S);
    auto timeShape = [&](const LaunchShape& sh, double& ms) -> hipError_t {
        hipError_t r = hipMemcpyAsync(b->arena, snap, b->bytes, hipMemcpyDeviceToDevice, b->stream);
        b->host_step = start;
        b->plan.parity = 0;
        b->plan.shape = sh;
        if (r == hipSuccess) r = b->plan.build(b->stream);
        if (r == hipSuccess) r = buildParts(b);
        if (r == hipSuccess) r = batchAdvance(b, b->S);  // warm-up
        if (r == hipSuccess) r = hipEventRecord(b->e0, b->stream);
        if (r == hipSuccess) r = batchAdvance(b, trial);
        if (r == hipSuccess) r = hipEventRecord(b->e1, b->stream);
        if (r == hipSuccess) r = hipEventSynchronize(b->e1);
        float f = 0;
        if (r == hipSuccess) r = hipEventElapsedTime(&f, b->e0, b->e1);
        ms = f / trial;
        return r;
    };
    // coordinate descent: one stream or two, XCD remap, element block, node block
    LaunchShape cur = b->plan.shape;
    double curMs = 0;
    e = timeShape(cur, curMs);
    const int blocks[] = {64, 128, 256, 512};
    for (int pass = -1; pass < 3 && e == hipSuccess; ++pass) {
        if (pass == -1 && b->plan.recs.size() < 2) continue;
        // the tiled variant's element block is its 256-tet tile: not a free parameter (the
        // hex pass always takes one lane per element, so for hexes it is)
        if (pass == 1 && cur.variant == kTiled && !b->plan.hex) continue;
        std::vector<LaunchShape> cands;
        for (int v : (pass <= 0 ? std::vector<int>{0, 1} : std::vector<int>(blocks, blocks + 4))) {
            LaunchShape c = cur;
            if (pass == -1) c.split = v;
            else if (pass == 0) c.xcd_remap = v;
            else if (pass == 1) c.elem_block = v;
            else c.node_block = v;
            cands.push_back(c);
        }
        for (const LaunchShape& c : cands) {
            if (c.elem_block == cur.elem_block && c.node_block == cur.node_block &&
                c.xcd_remap == cur.xcd_remap && c.split == cur.split)
                continue;
            double ms = 0;
            e = timeShape(c, ms);
            if (e != hipSuccess) break;
            if (ms < curMs) {
                cur = c;
                curMs = ms;
            }
        }
    }
    if (e == hipSuccess) e = hipMemcpyAsync(b->arena, snap, b->bytes, hipMemcpyDeviceToDevice, b->stream);
    b->host_step = start;
    b->plan.parity = 0;
    b->plan.shape = cur;
    if (e == hipSuccess) e = b->plan.build(b->stream);
    if (e == hipSuccess) e = buildParts(b);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    (void)hipFree(snap);
    if (e != hipSuccess) return statusOf(e);
    if (best) {
        best->elem_block = cur.elem_block;
        best->node_block = cur.node_block;
        best->xcd_remap = cur.xcd_remap;
        best->chunk = b->S;
        best->variant = cur.variant == kSlots ? SHIPFEM_VARIANT_SLOTS : SHIPFEM_VARIANT_TILED;
        best->split = cur.split ? 0 : -1;
        best->rest_from_positions = cur.formed;
        best->resident = cur.resident ? 0 : -1;
    }
    if (best_ms) *best_ms = curMs;
    return SHIPFEM_OK;
}

shipfem_status shipfem_batch_readback(shipfem_batch* b, uint32_t i, shipfem_state_view* view) {
    if (!b || !view || i >= b->dcs.size()) return SHIPFEM_INVALID_ARG;
    return statusOf(readbackCase(b->dcs[i], b->hcs[i], view, b->stream));
}

shipfem_status shipfem_batch_energy(shipfem_batch* b, uint32_t i, shipfem_energy_record* out,
                                    int32_t max, int32_t* count) {
    if (!b || i >= b->dcs.size()) return SHIPFEM_INVALID_ARG;
    return statusOf(readEnergy(b->dcs[i], out, max, count, b->stream));
}

shipfem_status shipfem_batch_tears(shipfem_batch* b, uint32_t i, shipfem_tear* out, uint32_t max,
                                   uint32_t* count) {
    if (!b || i >= b->dcs.size() || (!out && max)) return SHIPFEM_INVALID_ARG;
    std::vector<shipfem_tear> tears;
    const hipError_t e = readTears(b->dcs[i], b->hcs[i], tears, b->stream);
    if (e != hipSuccess) return statusOf(e);
    return copyTears(tears, out, max, count);
}

shipfem_status shipfem_batch_reset(shipfem_batch* b) {
    if (!b) return SHIPFEM_INVALID_ARG;
    hipError_t e = hipMemcpyAsync(b->arena, b->backup, b->bytes, hipMemcpyDeviceToDevice, b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    b->host_step = 0;
    b->plan.parity = 0;
    for (Plan& h : b->part) h.parity = 0;
    return statusOf(e);
}

void shipfem_batch_destroy(shipfem_batch* b) {
    if (b) (void)hipStreamSynchronize(b->stream);
    delete b;
}

}  // extern "C"
