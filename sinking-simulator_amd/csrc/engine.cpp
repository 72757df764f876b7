// engine.cpp -- step plans: tile tables over one or many cases, plain launches and
// hipGraph capture of a chunk of steps.
#include "shipfem_host.hpp"

#include <algorithm>

namespace shipfem {

namespace {

// Appends the tiles of case i: `block` items each over its elements or nodes.
void tilesOf(const std::vector<CaseDev>& recs, size_t i, int block, bool nodes, std::vector<Tile>& out) {
    const uint32_t count = nodes ? recs[i].N : recs[i].T;
    for (uint32_t first = 0; first < count; first += static_cast<uint32_t>(block))
        out.push_back(Tile{static_cast<int32_t>(i), first});
}

void tilesFor(const std::vector<CaseDev>& recs, int block, bool nodes, std::vector<Tile>& out) {
    out.clear();
    for (size_t i = 0; i < recs.size(); ++i) tilesOf(recs, i, block, nodes, out);
}

}  // namespace

Plan::~Plan() {
    dropGraph();
    if (d_recs) (void)hipFree(d_recs);
    if (d_elem) (void)hipFree(d_elem);
    if (d_node) (void)hipFree(d_node);
    if (d_omega) (void)hipFree(d_omega);
    if (d_omega_out) (void)hipFree(d_omega_out);
}

void Plan::dropGraph() {
    if (graph) (void)hipGraphExecDestroy(graph);
    if (graph_src) (void)hipGraphDestroy(graph_src);
    graph = nullptr;
    graph_src = nullptr;
    chunk = 0;
}

hipError_t Plan::build(hipStream_t s) {
    dropGraph();
    hex = !recs.empty() && recs[0].corners == 8;
    // one element kind per plan, and hex plans all elastic or all plastic
    for (const CaseDev& r : recs)
        if ((r.corners == 8) != hex || (hex && ((r.features ^ recs[0].features) & kFeatPlastic)))
            return hipErrorInvalidValue;
    // the plastic hex pass keeps 48 doubles a lane in LDS: at most 256 lanes a block
    if (hex && (recs[0].features & kFeatPlastic)) shape.elem_block = std::min(shape.elem_block, 256);
    // the tiled group first, then the slots group
    elem_tiles.clear();
    node_tiles.clear();
    lmax = 0;
    for (int group = 0; group < 2; ++group) {
        for (size_t i = 0; i < recs.size(); ++i) {
            if (inTiledGroup(recs[i]) != (group == 0)) continue;
            tilesOf(recs, i, group == 0 ? kTile : shape.elem_block, false, elem_tiles);
            tilesOf(recs, i, shape.node_block, true, node_tiles);
            if (group == 0) lmax = std::max(lmax, recs[i].lmax);
        }
        if (group == 0) {
            elem_split = elem_tiles.size();
            node_split = node_tiles.size();
        }
    }
    tilesFor(recs, 256, false, omega_tiles);
    refreshFeatures();
    // a plan of small tiled cases steps whole chunks in one block per case (k_resident)
    resident = shape.resident && !hex && !recs.empty() && elem_split == elem_tiles.size() &&
               node_split == node_tiles.size();
    for (const CaseDev& r : recs) resident = resident && r.elem_tiles > 0 && r.elem_tiles <= 4 && r.N <= 2048;
    if (resident) {
        int dev = 0, maxLds = 0;
        resident = hipGetDevice(&dev) == hipSuccess &&
                   hipDeviceGetAttribute(&maxLds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess &&
                   resident_lds(lmax) <= static_cast<size_t>(maxLds);
    }
    hipError_t e = hipSuccess;
    size_t need = std::max({elem_tiles.size(), node_tiles.size(), omega_tiles.size()});
    if (need > tile_cap || !d_elem) {
        if (d_elem) (void)hipFree(d_elem);
        if (d_node) (void)hipFree(d_node);
        if (d_omega) (void)hipFree(d_omega);
        d_elem = d_node = d_omega = nullptr;
        tile_cap = std::max<size_t>(need, 1);
        e = hipMalloc(reinterpret_cast<void**>(&d_elem), sizeof(Tile) * tile_cap);
        if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&d_node), sizeof(Tile) * tile_cap);
        if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&d_omega), sizeof(Tile) * tile_cap);
        if (e != hipSuccess) return e;
    }
    if (recs.size() > case_cap || !d_recs) {
        if (d_recs) (void)hipFree(d_recs);
        if (d_omega_out) (void)hipFree(d_omega_out);
        d_recs = nullptr;
        d_omega_out = nullptr;
        case_cap = std::max<size_t>(recs.size(), 1);
        e = hipMalloc(reinterpret_cast<void**>(&d_recs), sizeof(CaseDev) * case_cap);
        if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&d_omega_out), sizeof(double) * case_cap);
        if (e != hipSuccess) return e;
    }
    e = hipMemcpyAsync(d_elem, elem_tiles.data(), sizeof(Tile) * elem_tiles.size(), hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
        e = hipMemcpyAsync(d_node, node_tiles.data(), sizeof(Tile) * node_tiles.size(), hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
        e = hipMemcpyAsync(d_omega, omega_tiles.data(), sizeof(Tile) * omega_tiles.size(), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = uploadRecords(s);
    // host-side staging buffers must outlive the async copies
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    return e;
}

bool Plan::inTiledGroup(const CaseDev& r) const { return !hex && shape.variant == kTiled && r.elem_tiles > 0; }

void Plan::refreshFeatures() {
    feat_tiled = feat_slots = 0;
    for (const CaseDev& r : recs) (inTiledGroup(r) ? feat_tiled : feat_slots) |= r.features;
    feat_union = feat_tiled | feat_slots;
}

hipError_t Plan::uploadRecords(hipStream_t s) {
    // a record's features choose the instantiations its group's launches take
    refreshFeatures();
    if (recs.empty()) return hipSuccess;
    hipError_t e = hipMemcpyAsync(d_recs, recs.data(), sizeof(CaseDev) * recs.size(),
                                  hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    return e;
}

void Plan::launchStep(int par, int sample, hipStream_t s) {
    launchElem(par, sample, s);
    launchNode(par, sample, s);
}

namespace {
// Timing of a phase issued as two launches: the start event on the first, the stop on the
// second (one launch takes both).
void splitTiming(const KernelTiming* t, bool both, KernelTiming& first, KernelTiming& second) {
    if (!t) return;
    first = *t;
    second = *t;
    if (both) {
        first.stop = nullptr;
        second.start = nullptr;
    }
}
}  // namespace

void Plan::launchElem(int par, int sample, hipStream_t s, const KernelTiming* t) {
    const int n1 = static_cast<int>(elem_split), n2 = static_cast<int>(elem_tiles.size() - elem_split);
    KernelTiming t1, t2;
    splitTiming(t, n1 > 0 && n2 > 0, t1, t2);
    const KernelTiming* p1 = t ? &t1 : nullptr;
    const KernelTiming* p2 = t ? &t2 : nullptr;
    // a plan of one case passes no tile table: tile b is item b * block of case 0
    if (n1 > 0)
        launch_elem_tiled(d_recs, recs.size() == 1 ? nullptr : d_elem, n1, lmax, par, sample, shape.xcd_remap,
                          feat_tiled, shape.formed != 0, s, p1);
    if (n2 == 0) return;
    if (hex && (feat_slots & kFeatPlastic))
        launch_elem_hex_plastic(d_recs, d_elem + n1, n2, shape.elem_block, par, sample, shape.xcd_remap, s, p2);
    else if (hex)
        launch_elem_hex(d_recs, d_elem + n1, n2, shape.elem_block, par, sample, shape.xcd_remap, s, p2);
    else
        launch_elem(d_recs, d_elem + n1, n2, shape.elem_block, par, sample, shape.xcd_remap, feat_slots, s, p2);
}

void Plan::launchNode(int par, int sample, hipStream_t s, const KernelTiming* t) {
    // the tiled group gathers tile partials; the slots group (and every hex case) slots
    const int n1 = static_cast<int>(node_split), n2 = static_cast<int>(node_tiles.size() - node_split);
    KernelTiming t1, t2;
    splitTiming(t, n1 > 0 && n2 > 0, t1, t2);
    const bool one = recs.size() == 1;
    if (n1 > 0)
        launch_node(d_recs, one ? nullptr : d_node, n1, shape.node_block, par, sample, shape.xcd_remap, feat_tiled,
                    kTiled, s, t ? &t1 : nullptr);
    if (n2 > 0)
        launch_node(d_recs, one ? nullptr : d_node + n1, n2, shape.node_block, par, sample, shape.xcd_remap,
                    feat_slots, kSlots, s, t ? &t2 : nullptr);
}

void Plan::stepPlain(bool sample, hipStream_t s) {
    if (resident) {
        launch_resident(d_recs, static_cast<int>(recs.size()), lmax, 1, parity, sample ? 1 : 0, feat_tiled,
                         shape.formed != 0, s);
    } else {
        launchStep(parity, sample ? 1 : 0, s);
        if (sample) launch_reduce(d_recs, static_cast<int>(recs.size()), parity, s);
    }
    parity ^= 1;
}

hipError_t Plan::capture(int steps, hipStream_t s) {
    dropGraph();
    if (steps < 2 || (steps & 1)) return hipErrorInvalidValue;
    hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return e;
    if (resident) {
        launch_resident(d_recs, static_cast<int>(recs.size()), lmax, steps, 0, 1, feat_tiled, shape.formed != 0, s);
    } else {
        int par = 0;
        for (int k = 0; k < steps; ++k) {
            const int sample = k == 0 ? 1 : 0;
            launchStep(par, sample, s);
            if (sample) launch_reduce(d_recs, static_cast<int>(recs.size()), par, s);
            par ^= 1;
        }
    }
    e = hipStreamEndCapture(s, &graph_src);
    if (e != hipSuccess) return e;
    e = hipGraphInstantiate(&graph, graph_src, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(graph_src);
        graph_src = nullptr;
        graph = nullptr;
        return e;
    }
    chunk = steps;
    return hipSuccess;
}

hipError_t Plan::omegaMax(std::vector<double>& out, hipStream_t s) {
    launch_omega(d_recs, d_omega, static_cast<int>(omega_tiles.size()), s);
    launch_omega_max(d_recs, static_cast<int>(recs.size()), d_omega_out, s);
    out.assign(recs.size(), 0.0);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
        e = hipMemcpyAsync(out.data(), d_omega_out, sizeof(double) * out.size(), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    return e;
}

}  // namespace shipfem
