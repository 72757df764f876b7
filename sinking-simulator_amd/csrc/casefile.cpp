// casefile.cpp -- the case wire format: one self-checking little-endian blob per case,
// written to disk or handed to the server. A header (magic, version, counts, which
// optional arrays are present and whether the elements are solid-shell hexes, payload
// size, FNV-1a 64 of the payload), then the scalar block, then the arrays in a fixed order.
#include "shipfem.h"
#include "shipfem_host.hpp"

#include <cstdio>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <vector>

namespace {

constexpr char kMagic[8] = {'S', 'H', 'I', 'P', 'F', 'E', 'M', 'C'};
constexpr uint32_t kVersion = 1;

enum Present : uint32_t {
    kHasFlags = 1u << 0,
    kHasLoad = 1u << 1,
    kHasPrescribed = 1u << 2,
    kHasU0 = 1u << 3,
    kHasV0 = 1u << 4,
    kHexElements = 1u << 5,  // solid-shell hexes: 8 node ids per element
    kHasMass = 1u << 6,      // the caller's lumped nodal masses
    kHasFracture = 1u << 7,  // hexes: the local fracture strain (8 bytes after the arrays)
};
constexpr uint32_t kPresentMask = 0xFFu;

struct Header {
    char magic[8];
    uint32_t version;
    uint32_t node_count, tet_count;
    uint32_t present;
    uint64_t payload;   // bytes after the header
    uint64_t checksum;  // FNV-1a 64 of the payload
    uint64_t reserved[3];
};
static_assert(sizeof(Header) == 64, "header layout");

// Scalars of shipfem_case_desc, each at a fixed width.
struct Scalars {
    double E, nu, rho;
    int32_t plastic, pad0;
    double yield, hardening, failure_strain, sensitivity, eta_ref, eta_cut;
    int32_t indenter_enabled, pad1;
    double radius, center[3], velocity[3], penalty;
    double load_ramp, drive_ramp, gravity, damping, dt, dt_safety;
    int32_t sample_every, pad2;
    int64_t max_steps;
};
static_assert(sizeof(Scalars) == 216, "scalar block layout");

uint64_t fnv1a(const unsigned char* p, uint64_t n) {
    uint64_t h = 1469598103934665603ull;
    for (uint64_t i = 0; i < n; ++i) {
        h ^= p[i];
        h *= 1099511628211ull;
    }
    return h;
}

uint32_t presentOf(const shipfem_case_desc& d, bool hex) {
    return (d.node_flags ? kHasFlags : 0u) | (d.load ? kHasLoad : 0u) |
           (d.prescribed ? kHasPrescribed : 0u) | (d.initial_displacement ? kHasU0 : 0u) |
           (d.initial_velocity ? kHasV0 : 0u) | (hex ? kHexElements : 0u) | (d.mass ? kHasMass : 0u) |
           (d.fracture_strain != 0.0 ? kHasFracture : 0u);
}

uint64_t payloadBytes(uint64_t N, uint64_t T, uint32_t present) {
    uint64_t b = sizeof(Scalars) + 24 * N + ((present & kHexElements) ? 32 : 16) * T;
    if (present & kHasFlags) b += 4 * N;
    for (uint32_t bit : {kHasLoad, kHasPrescribed, kHasU0, kHasV0})
        if (present & bit) b += 24 * N;
    if (present & kHasMass) b += 8 * N;
    if (present & kHasFracture) b += 8;
    return b;
}

uint64_t encodedSize(const shipfem_case_desc* d, bool hex) {
    if (!d) return 0;
    return sizeof(Header) + payloadBytes(d->node_count, d->tet_count, presentOf(*d, hex));
}

shipfem_status encode(const shipfem_case_desc* d, void* buf, uint64_t len, bool hex);
shipfem_status save(const shipfem_case_desc* d, const char* path, bool hex);

}  // namespace

// Decoded case: owns the arrays a shipfem_case_desc borrows.
struct shipfem_case_blob {
    std::vector<double> X, load, prescribed, u0, v0, mass;
    std::vector<uint32_t> index, flags;
    uint32_t corners = 4;
};

namespace {

shipfem_status encode(const shipfem_case_desc* d, void* buf, uint64_t len, bool hex) {
    if (!d || !buf || !d->rest_position || !d->index || d->node_count == 0 || d->tet_count == 0)
        return SHIPFEM_INVALID_ARG;
    const uint64_t need = encodedSize(d, hex);
    if (len < need) return SHIPFEM_INVALID_ARG;
    const uint64_t N = d->node_count, T = d->tet_count;
    Header h{};
    std::memcpy(h.magic, kMagic, 8);
    h.version = kVersion;
    h.node_count = d->node_count;
    h.tet_count = d->tet_count;
    h.present = presentOf(*d, hex);
    h.payload = need - sizeof(Header);
    const shipfem_material& m = d->material;
    const shipfem_indenter& ind = d->indenter;
    Scalars s{};
    s.E = m.youngs_modulus;
    s.nu = m.poisson_ratio;
    s.rho = m.density;
    s.plastic = m.plastic;
    s.yield = m.yield_strength;
    s.hardening = m.hardening_modulus;
    s.failure_strain = m.failure_strain;
    s.sensitivity = m.triaxiality_sensitivity;
    s.eta_ref = m.reference_triaxiality;
    s.eta_cut = m.cutoff_triaxiality;
    s.indenter_enabled = ind.enabled;
    s.radius = ind.radius;
    for (int a = 0; a < 3; ++a) {
        s.center[a] = ind.center[a];
        s.velocity[a] = ind.velocity[a];
    }
    s.penalty = ind.penalty;
    s.load_ramp = d->load_ramp;
    s.drive_ramp = d->drive_ramp;
    s.gravity = d->gravity;
    s.damping = d->damping;
    s.dt = d->dt;
    s.dt_safety = d->dt_safety;
    s.sample_every = d->sample_every;
    s.max_steps = d->max_steps;
    unsigned char* p = static_cast<unsigned char*>(buf) + sizeof(Header);
    unsigned char* const body = p;
    auto put = [&p](const void* src, uint64_t bytes) {
        std::memcpy(p, src, bytes);
        p += bytes;
    };
    put(&s, sizeof s);
    put(d->rest_position, 24 * N);
    put(d->index, (hex ? 32 : 16) * T);
    if (d->node_flags) put(d->node_flags, 4 * N);
    if (d->load) put(d->load, 24 * N);
    if (d->prescribed) put(d->prescribed, 24 * N);
    if (d->initial_displacement) put(d->initial_displacement, 24 * N);
    if (d->initial_velocity) put(d->initial_velocity, 24 * N);
    if (d->mass) put(d->mass, 8 * N);
    if (d->fracture_strain != 0.0) put(&d->fracture_strain, 8);
    h.checksum = fnv1a(body, h.payload);
    std::memcpy(buf, &h, sizeof h);
    return SHIPFEM_OK;
}

shipfem_status save(const shipfem_case_desc* d, const char* path, bool hex) {
    if (!d || !path) return SHIPFEM_INVALID_ARG;
    std::vector<unsigned char> buf(encodedSize(d, hex));
    const shipfem_status st = encode(d, buf.data(), buf.size(), hex);
    if (st != SHIPFEM_OK) return st;
    const std::string tmp = std::string(path) + ".tmp";
    FILE* f = std::fopen(tmp.c_str(), "wb");
    if (!f) return SHIPFEM_INVALID_ARG;
    const bool ok = std::fwrite(buf.data(), 1, buf.size(), f) == buf.size();
    if (std::fclose(f) != 0 || !ok || std::rename(tmp.c_str(), path) != 0) {
        std::remove(tmp.c_str());
        return SHIPFEM_INVALID_ARG;
    }
    return SHIPFEM_OK;
}

}  // namespace

extern "C" {

uint64_t shipfem_case_encoded_size(const shipfem_case_desc* d) { return encodedSize(d, false); }
uint64_t shipfem_hex_case_encoded_size(const shipfem_case_desc* d) { return encodedSize(d, true); }

shipfem_status shipfem_case_encode(const shipfem_case_desc* d, void* buf, uint64_t len) {
    return encode(d, buf, len, false);
}

shipfem_status shipfem_hex_case_encode(const shipfem_case_desc* d, void* buf, uint64_t len) {
    return encode(d, buf, len, true);
}

shipfem_status shipfem_case_decode(const void* buf, uint64_t len, shipfem_case_blob** out,
                                   shipfem_case_desc* d) {
    if (!buf || !out || !d || len < sizeof(Header)) return SHIPFEM_INVALID_ARG;
    *out = nullptr;
    Header h;
    std::memcpy(&h, buf, sizeof h);
    if (std::memcmp(h.magic, kMagic, 8) != 0 || h.version != kVersion || h.node_count == 0 ||
        h.tet_count == 0 || (h.present & ~kPresentMask) != 0)
        return SHIPFEM_INVALID_ARG;
    const uint64_t N = h.node_count, T = h.tet_count;
    if (h.payload != payloadBytes(N, T, h.present) || len < sizeof(Header) + h.payload)
        return SHIPFEM_INVALID_ARG;
    const unsigned char* p = static_cast<const unsigned char*>(buf) + sizeof(Header);
    if (fnv1a(p, h.payload) != h.checksum) return SHIPFEM_INVALID_ARG;
    std::unique_ptr<shipfem_case_blob> b(new (std::nothrow) shipfem_case_blob());
    if (!b) return SHIPFEM_OUT_OF_MEMORY;
    Scalars s;
    std::memcpy(&s, p, sizeof s);
    p += sizeof s;
    auto take = [&p](auto& vec, uint64_t count) {
        vec.resize(count);
        std::memcpy(vec.data(), p, count * sizeof(vec[0]));
        p += count * sizeof(vec[0]);
    };
    take(b->X, 3 * N);
    b->corners = (h.present & kHexElements) ? 8 : 4;
    take(b->index, b->corners * T);
    if (h.present & kHasFlags) take(b->flags, N);
    if (h.present & kHasLoad) take(b->load, 3 * N);
    if (h.present & kHasPrescribed) take(b->prescribed, 3 * N);
    if (h.present & kHasU0) take(b->u0, 3 * N);
    if (h.present & kHasV0) take(b->v0, 3 * N);
    if (h.present & kHasMass) take(b->mass, N);
    double fracture = 0.0;
    if (h.present & kHasFracture) {
        std::memcpy(&fracture, p, 8);
        p += 8;
    }
    for (uint32_t v : b->index)
        if (v >= N) return SHIPFEM_INVALID_ARG;
    *d = shipfem_case_desc{};
    d->node_count = h.node_count;
    d->tet_count = h.tet_count;
    d->rest_position = b->X.data();
    d->index = b->index.data();
    d->node_flags = b->flags.empty() ? nullptr : b->flags.data();
    d->load = b->load.empty() ? nullptr : b->load.data();
    d->prescribed = b->prescribed.empty() ? nullptr : b->prescribed.data();
    d->initial_displacement = b->u0.empty() ? nullptr : b->u0.data();
    d->initial_velocity = b->v0.empty() ? nullptr : b->v0.data();
    d->mass = b->mass.empty() ? nullptr : b->mass.data();
    d->fracture_strain = fracture;
    shipfem_material& m = d->material;
    m.youngs_modulus = s.E;
    m.poisson_ratio = s.nu;
    m.density = s.rho;
    m.plastic = s.plastic;
    m.yield_strength = s.yield;
    m.hardening_modulus = s.hardening;
    m.failure_strain = s.failure_strain;
    m.triaxiality_sensitivity = s.sensitivity;
    m.reference_triaxiality = s.eta_ref;
    m.cutoff_triaxiality = s.eta_cut;
    shipfem_indenter& ind = d->indenter;
    ind.enabled = s.indenter_enabled;
    ind.radius = s.radius;
    for (int a = 0; a < 3; ++a) {
        ind.center[a] = s.center[a];
        ind.velocity[a] = s.velocity[a];
    }
    ind.penalty = s.penalty;
    d->load_ramp = s.load_ramp;
    d->drive_ramp = s.drive_ramp;
    d->gravity = s.gravity;
    d->damping = s.damping;
    d->dt = s.dt;
    d->dt_safety = s.dt_safety;
    d->sample_every = s.sample_every;
    d->max_steps = s.max_steps;
    *out = b.release();
    return SHIPFEM_OK;
}

void shipfem_case_blob_destroy(shipfem_case_blob* b) { delete b; }

uint32_t shipfem_case_blob_corners(const shipfem_case_blob* b) { return b ? b->corners : 0; }

shipfem_status shipfem_case_save(const shipfem_case_desc* d, const char* path) {
    return save(d, path, false);
}

shipfem_status shipfem_hex_case_save(const shipfem_case_desc* d, const char* path) {
    return save(d, path, true);
}

shipfem_status shipfem_case_load(const char* path, shipfem_case_blob** out, shipfem_case_desc* d) {
    if (!path || !out || !d) return SHIPFEM_INVALID_ARG;
    FILE* f = std::fopen(path, "rb");
    if (!f) return SHIPFEM_NOT_FOUND;
    std::vector<unsigned char> buf;
    unsigned char chunk[1 << 16];
    size_t n;
    while ((n = std::fread(chunk, 1, sizeof chunk, f)) > 0) buf.insert(buf.end(), chunk, chunk + n);
    std::fclose(f);
    return shipfem_case_decode(buf.data(), buf.size(), out, d);
}

}  // extern "C"
