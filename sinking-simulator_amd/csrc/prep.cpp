// prep.cpp -- case preparation (FP64 rest state on the host, FP32 SoA planes for the
// device), the per-case device arena, upload and readback.
#include "shipfem_host.hpp"
#include "hex_plastic.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>

namespace shipfem {

namespace {

constexpr uint32_t kKnownFlags = SHIPFEM_PIN_ALL | SHIPFEM_LOADED | SHIPFEM_CONTACT |
                                 SHIPFEM_DRIVE_X | SHIPFEM_DRIVE_Y | SHIPFEM_DRIVE_Z;

std::string fmt(const char* what, double v) { return std::string(what) + " (" + std::to_string(v) + ")"; }

// Splits a double into a float and the Kahan compensation c = hi - x (the kernels
// take the true value as hi - c).
inline void split(double x, float& hi, float& c) {
    hi = static_cast<float>(x);
    c = static_cast<float>(static_cast<double>(hi) - x);
}

struct Layout {
    uint64_t off = 0;
    uint64_t take(uint64_t bytes) {
        const uint64_t o = off;
        off += (bytes + 255) / 256 * 256;
        return o;
    }
};

// Offsets of every array in a case arena; shared by arenaBytes and uploadCase.
struct ArenaMap {
    uint64_t conn, dmi, vol, force, hk, hgeo, hgeod, hpoint, henh, hfail, pstate, tear_step, tear_stress, u4, cu, v, cv, invm, mass, nflags, load, xc, presc,
        ell_off, ell, lconn, tn_off, tnode, trest, lofs, partial, pell_off, pell, acc_e, acc_n,
        smp_e, smp_n, omega_w, step, status, hist_count, hist, total;
};

ArenaMap mapArena(const HostCase& hc) {
    Layout L;
    ArenaMap m{};
    const uint64_t T = hc.T, N = hc.N, Tp = hc.Tpad, Np = hc.Npad;
    m.conn = L.take(4ull * hc.corners * T);
    m.dmi = L.take(4ull * hc.dmi.size());
    m.vol = L.take(4ull * hc.vol.size());
    m.force = L.take(12ull * hc.corners * T);
    m.hk = L.take(8ull * hc.hk.size());
    m.hgeo = L.take(4ull * hc.hgeo.size());
    m.hgeod = L.take(8ull * hc.hgeod.size());
    const bool hexPlastic = hc.corners == 8 && (hc.features & kFeatPlastic);
    m.hpoint = L.take(hexPlastic ? 8ull * kHexPoint * Tp : 0);
    m.henh = L.take(hexPlastic ? 8ull * 7 * Tp : 0);
    m.hfail = L.take(8ull * hc.hfail.size());
    const bool dmg = (hc.features & kFeatDamage) != 0;
    m.pstate = L.take((hc.features & kFeatPlastic) && !hexPlastic ? (dmg ? 32 : 28) * Tp : 0);
    m.tear_step = L.take(dmg ? 4 * Tp : 0);
    m.tear_stress = L.take(dmg ? 24 * Tp : 0);
    m.u4 = L.take(16 * N);
    m.cu = L.take(12 * Np);
    m.v = L.take(12 * Np);
    m.cv = L.take(12 * Np);
    m.invm = L.take(4 * Np);
    m.mass = L.take(4 * Np);
    m.nflags = L.take(4 * Np);
    m.load = L.take((hc.features & kFeatLoad) ? 12 * Np : 0);
    m.xc = L.take((hc.features & kFeatContact) ? 16 * Np : 0);
    m.presc = L.take((hc.features & kFeatDrive) ? 12 * Np : 0);
    m.ell_off = L.take(4 * (uint64_t)hc.ell_off.size());
    m.ell = L.take(4 * (uint64_t)hc.ell.size());
    m.lconn = L.take(4ull * hc.lconn.size());
    m.tn_off = L.take(4 * (uint64_t)hc.tn_off.size());
    m.tnode = L.take(4 * (uint64_t)hc.tnode.size());
    m.trest = L.take(4 * (uint64_t)hc.trest.size());
    m.lofs = L.take(2 * (uint64_t)hc.lofs.size());
    m.partial = L.take(16 * (uint64_t)hc.P);
    m.pell_off = L.take(4 * (uint64_t)hc.pell_off.size());
    m.pell = L.take(4 * (uint64_t)hc.pell.size());
    m.acc_e = L.take(8ull * kAccElemSlots * hc.elem_waves);
    m.acc_n = L.take(8ull * kAccNodeSlots * hc.node_waves);
    m.smp_e = L.take(8ull * kSmpElemSlots * hc.elem_waves);
    m.smp_n = L.take(8ull * kSmpNodeSlots * hc.node_waves);
    m.omega_w = L.take(8ull * hc.elem_waves);
    m.step = L.take(16);
    m.status = L.take(4);
    m.hist_count = L.take(4);
    m.hist = L.take(sizeof(EnergyRecordDev) * (uint64_t)hc.hist_cap);
    m.total = L.off;
    return m;
}

}  // namespace

void setError(char* err, size_t len, const std::string& msg) {
    if (!err || len == 0) return;
    const size_t n = std::min(len - 1, msg.size());
    std::memcpy(err, msg.data(), n);
    err[n] = '\0';
}

shipfem_status statusOf(hipError_t e) {
    if (e == hipSuccess) return SHIPFEM_OK;
    if (e == hipErrorOutOfMemory) return SHIPFEM_OUT_OF_MEMORY;
    if (e == hipErrorNoDevice || e == hipErrorInvalidDevice) return SHIPFEM_NO_DEVICE;
    return SHIPFEM_DEVICE_ERROR;
}

bool deviceAvailable(std::string* why) {
    int n = 0;
    const hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n < 1) {
        if (why) *why = std::string("no HIP device (") + hipGetErrorString(e) + ")";
        return false;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, 0) != hipSuccess) {
        if (why) *why = "device 0 properties unavailable";
        return false;
    }
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
        if (why) *why = std::string("device 0 is ") + prop.gcnArchName + ", this build is gfx950 only";
        return false;
    }
    return true;
}

namespace {

// The library's internal tet order: recursive coordinate bisection of the tet centroids
// into leaves of exactly kTile tets (the remainder leaf last), each leaf in the caller's
// order. A 256-tet tile of the tiled element pass is then a compact block of cells, not
// a strip one cell wide, so it touches fewer nodes: fewer partial sums to gather, fewer
// displacements to stage. Each split is at a multiple of kTile, at that rank along the
// longest axis of the range's centroid box, ties broken by tet id, so the order depends
// only on the mesh. perm[k] is the caller's id of internal tet k.
std::vector<uint32_t> spatialTetOrder(const double* X, const uint32_t* index, uint32_t T) {
    std::vector<uint32_t> perm(T);
    for (uint32_t t = 0; t < T; ++t) perm[t] = t;
    if (T <= static_cast<uint32_t>(kTile)) return perm;
    std::vector<double> cent(3ull * T);
    for (uint64_t t = 0; t < T; ++t)
        for (int a = 0; a < 3; ++a) {
            double c = 0;
            for (int q = 0; q < 4; ++q) c += X[index[4 * t + q] * 3ull + a];
            cent[3 * t + a] = 0.25 * c;
        }
    std::vector<std::pair<uint32_t, uint32_t>> stack{{0u, T}};
    while (!stack.empty()) {
        const auto [b, e] = stack.back();
        stack.pop_back();
        const uint32_t n = e - b;
        if (n <= static_cast<uint32_t>(kTile)) {
            std::sort(perm.begin() + b, perm.begin() + e);
            continue;
        }
        double lo[3], hi[3];
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::numeric_limits<double>::infinity();
            hi[a] = -lo[a];
        }
        for (uint32_t k = b; k < e; ++k)
            for (int a = 0; a < 3; ++a) {
                lo[a] = std::min(lo[a], cent[3ull * perm[k] + a]);
                hi[a] = std::max(hi[a], cent[3ull * perm[k] + a]);
            }
        int ax = 0;
        for (int a = 1; a < 3; ++a)
            if (hi[a] - lo[a] > hi[ax] - lo[ax]) ax = a;
        // a multiple of kTile near the middle, never past the last whole leaf
        const uint32_t whole = n / kTile;
        const uint32_t half = static_cast<uint32_t>(std::lround(n / (2.0 * kTile)));
        const uint32_t left = std::min(whole, std::max<uint32_t>(1u, half)) * kTile;
        std::nth_element(perm.begin() + b, perm.begin() + b + left, perm.begin() + e,
                         [&](uint32_t x, uint32_t y) {
                             const double cx = cent[3ull * x + ax], cy = cent[3ull * y + ax];
                             return cx < cy || (cx == cy && x < y);
                         });
        stack.push_back({b + left, e});
        stack.push_back({b, b + left});
    }
    return perm;
}

// The internal node order: nodes by first appearance in the internal tet order (corner
// order within a tet), nodes in no tet last in the caller's order. A tile's new nodes
// are then consecutive, so the node pass's waves read their partials side by side.
std::vector<uint32_t> tileNodeOrder(const uint32_t* index, const std::vector<uint32_t>& tperm,
                                    uint32_t N) {
    std::vector<uint32_t> nperm;
    nperm.reserve(N);
    std::vector<uint8_t> seen(N, 0);
    for (uint32_t t : tperm)
        for (int c = 0; c < 4; ++c) {
            const uint32_t g = index[4ull * t + c];
            if (!seen[g]) {
                seen[g] = 1;
                nperm.push_back(g);
            }
        }
    for (uint32_t n = 0; n < N; ++n)
        if (!seen[n]) nperm.push_back(n);
    return nperm;
}

std::string prepareInternal(const shipfem_case_desc& d, HostCase& hc, int32_t hist_cap,
                            std::vector<uint32_t>&& tperm, std::vector<uint32_t>&& nperm,
                            uint32_t corners) {
    const uint32_t N = d.node_count, T = d.tet_count;
    if (N == 0 || T == 0) return "empty mesh (node_count and tet_count must be positive)";
    if (!d.rest_position || !d.index) return "rest_position and index are required";
    if (T > 0xFFFFFFFFu / corners - 64) return "element count too large for 32-bit slot entries";
    const shipfem_material& mt = d.material;
    if (!(mt.youngs_modulus > 0)) return fmt("youngs_modulus must be positive", mt.youngs_modulus);
    if (!(mt.poisson_ratio > -1.0 && mt.poisson_ratio < 0.5))
        return fmt("poisson_ratio must lie in (-1, 0.5)", mt.poisson_ratio);
    if (!(mt.density > 0)) return fmt("density must be positive", mt.density);
    if (mt.plastic && !(mt.yield_strength > 0)) return fmt("yield_strength must be positive", mt.yield_strength);
    if (mt.plastic && !(mt.hardening_modulus >= 0)) return fmt("hardening_modulus must be >= 0", mt.hardening_modulus);
    const bool failure = mt.plastic && mt.failure_strain > 0.0;
    if (failure) {
        if (!std::isfinite(mt.failure_strain)) return fmt("failure_strain must be finite", mt.failure_strain);
        if (!std::isfinite(mt.triaxiality_sensitivity) || !std::isfinite(mt.reference_triaxiality) ||
            !std::isfinite(mt.cutoff_triaxiality))
            return "triaxiality constants must be finite";
    }
    if (!(d.fracture_strain >= 0.0) || !std::isfinite(d.fracture_strain))
        return fmt("fracture_strain must be finite and >= 0", d.fracture_strain);
    if (corners == 4 && d.fracture_strain > 0.0)
        return "fracture_strain resolves per element for solid-shell hexes only; a tet case takes "
               "material.failure_strain as given (see shipfem_regularised_failure_strain)";
    const double damping = d.damping == 0.0 ? 1.0 : d.damping;
    if (!(damping > 0 && damping <= 1.0)) return fmt("damping must lie in (0, 1]", d.damping);
    if (!(d.dt >= 0) || !std::isfinite(d.dt)) return fmt("dt must be >= 0", d.dt);
    const double safety = d.dt_safety == 0.0 ? 0.9 : d.dt_safety;
    if (!(safety > 0 && safety <= 4.0)) return fmt("dt_safety must lie in (0, 4]", d.dt_safety);
    const int32_t S = d.sample_every == 0 ? 100 : d.sample_every;
    if (S < 2 || (S & 1)) return "sample_every must be an even number >= 2";
    if (d.max_steps < 0) return "max_steps must be >= 0";
    if (!(d.load_ramp >= 0) || !(d.drive_ramp >= 0)) return "ramp times must be >= 0";
    if (!std::isfinite(d.gravity)) return "gravity must be finite";
    const shipfem_indenter& ind = d.indenter;
    const double penalty = ind.penalty == 0.0 ? 0.1 : ind.penalty;
    if (ind.enabled) {
        if (!(ind.radius > 0)) return fmt("indenter radius must be positive", ind.radius);
        if (!(penalty > 0 && penalty <= 1.0)) return fmt("indenter penalty must lie in (0, 1]", ind.penalty);
    }

    hc = HostCase();
    hc.corners = corners;
    hc.tperm = std::move(tperm);
    hc.nperm = std::move(nperm);
    hc.N = N;
    hc.T = T;
    hc.Npad = roundUp64(N);
    hc.Tpad = roundUp64(T);
    hc.node_waves = hc.Npad / kWave;
    hc.elem_waves = hc.Tpad / kWave;
    hc.E = mt.youngs_modulus;
    hc.nu = mt.poisson_ratio;
    hc.rho = mt.density;
    hc.lam = hc.E * hc.nu / ((1 + hc.nu) * (1 - 2 * hc.nu));
    hc.mu = hc.E / (2 * (1 + hc.nu));
    hc.bulk = hc.E / (3 * (1 - 2 * hc.nu));
    hc.yield = mt.yield_strength;
    hc.hard = mt.hardening_modulus;
    hc.fail_strain = failure ? mt.failure_strain : 0.0;
    hc.fail_sens = mt.triaxiality_sensitivity;
    hc.fail_eref = mt.reference_triaxiality;
    hc.fail_ecut = mt.cutoff_triaxiality;
    hc.load_ramp = d.load_ramp;
    hc.drive_ramp = d.drive_ramp;
    hc.gravity = d.gravity;
    hc.damping = damping;
    hc.indenter = ind;
    hc.indenter.penalty = penalty;
    hc.dt_request = d.dt;
    hc.dt_safety = safety;
    hc.sample_every = S;
    hc.max_steps = d.max_steps;
    hc.hist_cap = hist_cap > 0 ? hist_cap : 4096;

    // flags
    hc.nflags.assign(hc.Npad, SHIPFEM_PIN_ALL);  // pad lanes behave as pinned
    for (uint32_t n = 0; n < N; ++n) {
        const uint32_t f = d.node_flags ? d.node_flags[n] : 0u;
        if (f & ~kKnownFlags) return "node " + std::to_string(externalNode(hc, n)) + " has unknown flag bits";
        hc.nflags[n] = f & ~(SHIPFEM_LOADED | SHIPFEM_CONTACT);  // derived below
    }
    for (uint64_t q = 0; q < uint64_t(corners) * T; ++q)
        if (d.index[q] >= N) return "element " + std::to_string(q / corners) + " references node " +
                                    std::to_string(d.index[q]) + " >= node_count";

    // rest state in FP64: Dm^-1 (row-major planes), volume, lumped mass (tets); the
    // condensed stiffness, centre gradients and mass of every element (hexes)
    hc.X.assign(d.rest_position, d.rest_position + 3ull * N);
    hc.conn.assign(d.index, d.index + uint64_t(corners) * T);
    std::vector<double> mass(N, 0.0);
    const double* X = d.rest_position;
    double shortest = std::numeric_limits<double>::infinity();
    if (corners == 8) {
        const std::string why = hexRestState(d, hc, mass);
        if (!why.empty()) return why;
    } else {
        hc.dmi.assign(9ull * hc.Tpad, 0.0f);
        hc.vol.assign(hc.Tpad, 0.0f);
    }
    for (uint32_t t = 0; t < T && corners == 4; ++t) {
        const uint32_t* v = &d.index[4ull * t];
        double Dm[3][3];  // Dm[row][col], col c = X_{c+1} - X_0
        double longest2 = 0;
        for (int c = 0; c < 3; ++c)
            for (int a = 0; a < 3; ++a) Dm[a][c] = X[v[c + 1] * 3ull + a] - X[v[0] * 3ull + a];
        for (int a = 0; a < 4; ++a)
            for (int b = a + 1; b < 4; ++b) {
                double d2 = 0;
                for (int k = 0; k < 3; ++k) {
                    const double e = X[v[a] * 3ull + k] - X[v[b] * 3ull + k];
                    d2 += e * e;
                }
                longest2 = std::max(longest2, d2);
                shortest = std::min(shortest, std::sqrt(d2));
            }
        const double det = Dm[0][0] * (Dm[1][1] * Dm[2][2] - Dm[1][2] * Dm[2][1]) -
                           Dm[0][1] * (Dm[1][0] * Dm[2][2] - Dm[1][2] * Dm[2][0]) +
                           Dm[0][2] * (Dm[1][0] * Dm[2][1] - Dm[1][1] * Dm[2][0]);
        if (!(std::fabs(det) > 1e-12 * longest2 * std::sqrt(longest2)))
            return "tet " + std::to_string(externalTet(hc, t)) + " is degenerate (|det Dm| = " +
                   std::to_string(det) + ")";
        const double id = 1.0 / det;
        const double inv[9] = {(Dm[1][1] * Dm[2][2] - Dm[1][2] * Dm[2][1]) * id,
                               (Dm[0][2] * Dm[2][1] - Dm[0][1] * Dm[2][2]) * id,
                               (Dm[0][1] * Dm[1][2] - Dm[0][2] * Dm[1][1]) * id,
                               (Dm[1][2] * Dm[2][0] - Dm[1][0] * Dm[2][2]) * id,
                               (Dm[0][0] * Dm[2][2] - Dm[0][2] * Dm[2][0]) * id,
                               (Dm[0][2] * Dm[1][0] - Dm[0][0] * Dm[1][2]) * id,
                               (Dm[1][0] * Dm[2][1] - Dm[1][1] * Dm[2][0]) * id,
                               (Dm[0][1] * Dm[2][0] - Dm[0][0] * Dm[2][1]) * id,
                               (Dm[0][0] * Dm[1][1] - Dm[0][1] * Dm[1][0]) * id};
        for (int k = 0; k < 9; ++k) hc.dmi[k * (uint64_t)hc.Tpad + t] = static_cast<float>(inv[k]);
        const double V = std::fabs(det) / 6.0;
        hc.vol[t] = static_cast<float>(V);
        hc.total_volume += V;
        for (int c = 0; c < 4; ++c) mass[v[c]] += hc.rho * V * 0.25;
    }
    hc.edge_dt = corners == 4 ? 0.5 * shortest / std::sqrt(hc.E / hc.rho) : 0.0;
    if (d.mass) {
        // the caller's lumped masses (TetMesh::mass). With m >= s m_density at every node
        // that carries stiffness, x.Kx / x.Mx <= (x.Kx / x.M_density x) / s, so the largest
        // frequency is at most the density bound over sqrt(s).
        double smin = std::numeric_limits<double>::infinity();
        for (uint32_t n = 0; n < N; ++n) {
            const double mc = d.mass[n];
            if (!std::isfinite(mc) || mc < 0)
                return fmt(("mass of node " + std::to_string(externalNode(hc, n)) + " must be finite and >= 0").c_str(), mc);
            const bool held = (hc.nflags[n] & SHIPFEM_PIN_ALL) == SHIPFEM_PIN_ALL;
            if (mass[n] > 0 && !held) {
                if (!(mc > 0))
                    return "node " + std::to_string(externalNode(hc, n)) + " is free but its mass is zero";
                smin = std::min(smin, mc / mass[n]);
            }
            mass[n] = mc;
        }
        if (std::isfinite(smin)) hc.cfl_scale = std::sqrt(smin);
    }
    hc.mass.assign(hc.Npad, 0.0f);
    hc.invm.assign(hc.Npad, 0.0f);
    for (uint32_t n = 0; n < N; ++n) {
        hc.total_mass += mass[n];
        hc.mass[n] = static_cast<float>(mass[n]);
        if (mass[n] > 0) hc.invm[n] = static_cast<float>(1.0 / mass[n]);
        else hc.nflags[n] |= SHIPFEM_PIN_ALL;  // a node in no tet carries no mass: held
    }

    // loads, contact candidates, prescribed drive
    bool anyLoad = false, anyContact = false, anyDrive = false;
    if (d.load) {
        hc.load.assign(3ull * hc.Npad, 0.0f);
        for (uint32_t n = 0; n < N; ++n)
            for (int a = 0; a < 3; ++a) {
                const double f = d.load[n * 3ull + a];
                if (!std::isfinite(f))
                    return "load of node " + std::to_string(externalNode(hc, n)) + " is not finite";
                if (f != 0.0) {
                    hc.load[a * (uint64_t)hc.Npad + n] = static_cast<float>(f);
                    hc.nflags[n] |= SHIPFEM_LOADED;
                    anyLoad = true;
                }
            }
        if (!anyLoad) hc.load.clear();
    }
    if (ind.enabled) {
        bool flagged = false;
        for (uint32_t n = 0; n < N && d.node_flags; ++n) flagged |= (d.node_flags[n] & SHIPFEM_CONTACT) != 0;
        hc.xc.assign(4ull * hc.Npad, 0.0f);
        for (uint32_t n = 0; n < N; ++n) {
            if (flagged && !(d.node_flags[n] & SHIPFEM_CONTACT)) continue;
            if ((hc.nflags[n] & SHIPFEM_PIN_ALL) == SHIPFEM_PIN_ALL) continue;
            hc.nflags[n] |= SHIPFEM_CONTACT;
            double r2 = 0;
            for (int a = 0; a < 3; ++a) {
                const double x = X[n * 3ull + a] - ind.center[a];
                hc.xc[a * (uint64_t)hc.Npad + n] = static_cast<float>(x);
                r2 += x * x;
            }
            hc.xc[3ull * hc.Npad + n] = static_cast<float>(r2 - ind.radius * ind.radius);
            anyContact = true;
        }
        if (!anyContact) hc.xc.clear();
    }
    if (d.prescribed) {
        hc.presc.assign(3ull * hc.Npad, 0.0f);
        for (uint32_t n = 0; n < N; ++n)
            for (int a = 0; a < 3; ++a)
                if (hc.nflags[n] & (SHIPFEM_DRIVE_X << a)) {
                    hc.presc[a * (uint64_t)hc.Npad + n] = static_cast<float>(d.prescribed[n * 3ull + a]);
                    anyDrive = true;
                }
        if (!anyDrive) hc.presc.clear();
    }
    hc.features = (mt.plastic ? kFeatPlastic : 0) | (failure ? kFeatDamage : 0) |
                  (anyLoad ? kFeatLoad : 0) |
                  (anyContact ? kFeatContact : 0) | (anyDrive ? kFeatDrive : 0) |
                  (damping != 1.0 ? kFeatDamp : 0) | (d.gravity != 0.0 ? kFeatGravity : 0);

    // CSR node -> (element*corners + corner) by counting sort, (element, corner) order
    // (fem.cpp:147-157, zone.cpp:978-980), then wave-transposed: wave w holds width_w
    // rows of 64 lanes
    const uint64_t CT = uint64_t(corners) * T;
    std::vector<uint32_t> off(N + 1, 0);
    for (uint64_t q = 0; q < CT; ++q) ++off[d.index[q] + 1];
    for (uint32_t n = 0; n < N; ++n) off[n + 1] += off[n];
    std::vector<uint32_t> ent(CT), cur(off.begin(), off.end() - 1);
    for (uint64_t q = 0; q < CT; ++q) ent[cur[d.index[q]]++] = static_cast<uint32_t>(q);
    hc.ell_off.assign(hc.node_waves + 1, 0);
    uint64_t total = 0;
    for (uint32_t w = 0; w < hc.node_waves; ++w) {
        uint32_t width = 0;
        for (uint32_t l = 0; l < kWave; ++l) {
            const uint32_t n = w * kWave + l;
            if (n < N && (hc.nflags[n] & SHIPFEM_PIN_ALL) != SHIPFEM_PIN_ALL)
                width = std::max(width, off[n + 1] - off[n]);
        }
        hc.ell_off[w] = static_cast<uint32_t>(total);
        total += uint64_t(width) * kWave;
        if (total > 0xFFFFFFF0ull) return "node gather slab exceeds 32-bit indexing";
    }
    hc.ell_off[hc.node_waves] = static_cast<uint32_t>(total);
    hc.ell.assign(total, kEllSentinel);
    for (uint32_t w = 0; w < hc.node_waves; ++w)
        for (uint32_t l = 0; l < kWave; ++l) {
            const uint32_t n = w * kWave + l;
            if (n >= N || (hc.nflags[n] & SHIPFEM_PIN_ALL) == SHIPFEM_PIN_ALL) continue;
            for (uint32_t q = off[n]; q < off[n + 1]; ++q)
                hc.ell[hc.ell_off[w] + (q - off[n]) * kWave + l] = ent[q];
        }

    // 256-tet tiles of the tiled element pass: each tile's unique nodes (one partial
    // slot each) in ascending internal id, so a tile's new nodes (consecutive ids) hold
    // consecutive partial slots. Each local node owns a contiguous run of the tile's
    // corner slots, in (tet, corner) order, from its offset lofs[l]; a tet's
    // connectivity packs, per corner, the local node (10 bits) and the corner's rank in
    // that run (6 bits). Hex cases have no tiles: their element pass writes slots.
    if (corners == 4) {
        const uint32_t nt = (T + kTile - 1) / kTile;
        hc.elem_tiles = nt;
        hc.lconn.assign(2ull * T, 0u);
        hc.tn_off.assign(nt + 1, 0u);
        hc.tnode.clear();
        hc.trest.clear();
        hc.lofs.clear();
        std::vector<uint32_t> stamp(N, 0xFFFFFFFFu), loc(N, 0u), deg(N, 0u), seen(N, 0u);
        std::vector<uint32_t> nodes;
        std::vector<uint16_t> cnt;
        for (uint32_t b = 0; b < nt; ++b) {
            const uint32_t t0 = b * kTile, t1 = std::min<uint32_t>(T, t0 + kTile);
            nodes.clear();
            for (uint32_t t = t0; t < t1; ++t)
                for (int c = 0; c < 4; ++c) {
                    const uint32_t g = d.index[4ull * t + c];
                    if (stamp[g] != b) {
                        stamp[g] = b;
                        deg[g] = 0;
                        seen[g] = 0;
                        nodes.push_back(g);
                    }
                    ++deg[g];
                }
            std::sort(nodes.begin(), nodes.end());
            const uint32_t L = static_cast<uint32_t>(nodes.size());
            bool fits = true;
            for (uint32_t g : nodes) fits = fits && deg[g] <= 64;
            if (!fits) {
                // a corner's rank in its node's run has 6 bits: a node in more than 64 tets of
                // one tile cannot be tiled. The case keeps no tiles and every plan runs it on
                // the slots variant (the reference's scheme, any node degree).
                hc.elem_tiles = 0;
                hc.P = hc.lmax = 0;
                std::vector<uint32_t>().swap(hc.lconn);
                std::vector<uint32_t>().swap(hc.tn_off);
                std::vector<uint32_t>().swap(hc.tnode);
                std::vector<float>().swap(hc.trest);
                std::vector<uint16_t>().swap(hc.lofs);
                break;
            }
            for (uint32_t l = 0; l < L; ++l) loc[nodes[l]] = l;
            hc.tn_off[b + 1] = hc.tn_off[b] + L;
            hc.lmax = std::max(hc.lmax, L);
            hc.tnode.insert(hc.tnode.end(), nodes.begin(), nodes.end());
            // the tile's rest positions about its centroid: FP32 keeps ~1e-7 of a tile's
            // extent, so edge vectors formed from them keep ~1e-6 of an edge
            double ctr[3] = {0.0, 0.0, 0.0};
            for (uint32_t g : nodes)
                for (int a = 0; a < 3; ++a) ctr[a] += d.rest_position[3ull * g + a];
            for (int a = 0; a < 3; ++a) ctr[a] /= static_cast<double>(L);
            for (uint32_t g : nodes) {
                for (int a = 0; a < 3; ++a) hc.trest.push_back(static_cast<float>(d.rest_position[3ull * g + a] - ctr[a]));
                hc.trest.push_back(0.0f);
            }
            cnt.assign(L + 1, 0);
            for (uint32_t l = 0; l < L; ++l) cnt[l + 1] = static_cast<uint16_t>(cnt[l] + deg[nodes[l]]);
            hc.lofs.insert(hc.lofs.end(), cnt.begin(), cnt.end());
            for (uint32_t t = t0; t < t1; ++t) {
                uint32_t wv[4];
                for (int c = 0; c < 4; ++c) {
                    const uint32_t g = d.index[4ull * t + c];
                    wv[c] = loc[g] | (seen[g]++ << 10);
                }
                hc.lconn[2ull * t] = wv[0] | (wv[1] << 16);
                hc.lconn[2ull * t + 1] = wv[2] | (wv[3] << 16);
            }
        }
    }
    if (corners == 4 && hc.elem_tiles > 0) {
        const uint32_t nt = hc.elem_tiles;
        hc.P = hc.tn_off[nt];
        // node -> its partial slots in tile order, wave-transposed like the CSR slab
        std::vector<uint32_t> poff(N + 1, 0u);
        for (uint32_t p = 0; p < hc.P; ++p) ++poff[hc.tnode[p] + 1];
        for (uint32_t n = 0; n < N; ++n) poff[n + 1] += poff[n];
        std::vector<uint32_t> plist(hc.P), pcur(poff.begin(), poff.end() - 1);
        for (uint32_t p = 0; p < hc.P; ++p) plist[pcur[hc.tnode[p]]++] = p;
        hc.pell_off.assign(hc.node_waves + 1, 0u);
        uint64_t ptotal = 0;
        for (uint32_t w = 0; w < hc.node_waves; ++w) {
            uint32_t width = 0;
            for (uint32_t l = 0; l < kWave; ++l) {
                const uint32_t n = w * kWave + l;
                if (n < N && (hc.nflags[n] & SHIPFEM_PIN_ALL) != SHIPFEM_PIN_ALL)
                    width = std::max(width, poff[n + 1] - poff[n]);
            }
            hc.pell_off[w] = static_cast<uint32_t>(ptotal);
            ptotal += uint64_t(width) * kWave;
            if (ptotal > 0xFFFFFFF0ull) return "partial gather slab exceeds 32-bit indexing";
        }
        hc.pell_off[hc.node_waves] = static_cast<uint32_t>(ptotal);
        hc.pell.assign(ptotal, kEllSentinel);
        for (uint32_t w = 0; w < hc.node_waves; ++w)
            for (uint32_t l = 0; l < kWave; ++l) {
                const uint32_t n = w * kWave + l;
                if (n >= N || (hc.nflags[n] & SHIPFEM_PIN_ALL) == SHIPFEM_PIN_ALL) continue;
                for (uint32_t q = poff[n]; q < poff[n + 1]; ++q)
                    hc.pell[hc.pell_off[w] + (q - poff[n]) * kWave + l] = plist[q];
            }
    }

    // initial state, split into float + Kahan compensation
    hc.u4.assign(4ull * N, 0.0f);
    hc.cu.assign(3ull * hc.Npad, 0.0f);
    hc.v.assign(3ull * hc.Npad, 0.0f);
    hc.cv.assign(3ull * hc.Npad, 0.0f);
    for (uint32_t n = 0; n < N; ++n)
        for (int a = 0; a < 3; ++a) {
            if (d.initial_displacement)
                split(d.initial_displacement[n * 3ull + a], hc.u4[n * 4ull + a],
                      hc.cu[a * (uint64_t)hc.Npad + n]);
            if (d.initial_velocity && !(hc.nflags[n] & (SHIPFEM_PIN_X << a)))
                split(d.initial_velocity[n * 3ull + a], hc.v[a * (uint64_t)hc.Npad + n],
                      hc.cv[a * (uint64_t)hc.Npad + n]);
        }
    return {};
}

// A copy of the caller's case in the internal orders.
struct PermutedCase {
    std::vector<double> X, load, presc, u0, v0, mass;
    std::vector<uint32_t> index, flags;
};

}  // namespace

std::string prepareCase(const shipfem_case_desc& d, HostCase& hc, int32_t hist_cap) {
    const uint32_t N = d.node_count, T = d.tet_count;
    if (N == 0 || T == 0) return "empty mesh (node_count and tet_count must be positive)";
    if (!d.rest_position || !d.index) return "rest_position and index are required";
    for (uint64_t q = 0; q < 4ull * T; ++q)
        if (d.index[q] >= N) return "tet " + std::to_string(q / 4) + " references node " +
                                    std::to_string(d.index[q]) + " >= node_count";
    std::vector<uint32_t> tperm = spatialTetOrder(d.rest_position, d.index, T);
    std::vector<uint32_t> nperm = tileNodeOrder(d.index, tperm, N);
    std::vector<uint32_t> ninv(N);
    for (uint32_t i = 0; i < N; ++i) ninv[nperm[i]] = i;
    PermutedCase pc;
    shipfem_case_desc dp = d;
    pc.X.resize(3ull * N);
    for (uint32_t i = 0; i < N; ++i)
        for (int a = 0; a < 3; ++a) pc.X[3ull * i + a] = d.rest_position[3ull * nperm[i] + a];
    dp.rest_position = pc.X.data();
    pc.index.resize(4ull * T);
    for (uint32_t k = 0; k < T; ++k)
        for (int c = 0; c < 4; ++c) pc.index[4ull * k + c] = ninv[d.index[4ull * tperm[k] + c]];
    dp.index = pc.index.data();
    if (d.node_flags) {
        pc.flags.resize(N);
        for (uint32_t i = 0; i < N; ++i) pc.flags[i] = d.node_flags[nperm[i]];
        dp.node_flags = pc.flags.data();
    }
    auto nodeVec = [&](const double* src, std::vector<double>& dst) -> const double* {
        if (!src) return nullptr;
        dst.resize(3ull * N);
        for (uint32_t i = 0; i < N; ++i)
            for (int a = 0; a < 3; ++a) dst[3ull * i + a] = src[3ull * nperm[i] + a];
        return dst.data();
    };
    dp.load = nodeVec(d.load, pc.load);
    dp.prescribed = nodeVec(d.prescribed, pc.presc);
    dp.initial_displacement = nodeVec(d.initial_displacement, pc.u0);
    dp.initial_velocity = nodeVec(d.initial_velocity, pc.v0);
    if (d.mass) {
        pc.mass.resize(N);
        for (uint32_t i = 0; i < N; ++i) pc.mass[i] = d.mass[nperm[i]];
        dp.mass = pc.mass.data();
    }
    return prepareInternal(dp, hc, hist_cap, std::move(tperm), std::move(nperm), 4);
}

std::string prepareHexCase(const shipfem_case_desc& d, HostCase& hc, int32_t hist_cap) {
    return prepareInternal(d, hc, hist_cap, {}, {}, 8);
}

uint64_t arenaBytes(const HostCase& hc) { return mapArena(hc).total; }

PassBytes compulsoryBytes(const HostCase& hc, int variant, uint64_t yielded, bool formed) {
    PassBytes b;
    if (hc.elem_tiles == 0) variant = kSlots;  // a case without tiles always runs slots
    const uint64_t T = hc.T, N = hc.N, Np = hc.Npad, P = hc.P, nt = hc.elem_tiles;
    const uint32_t f = hc.features;
    uint64_t freeNodes = 0;
    for (uint32_t n = 0; n < hc.N && n < hc.nflags.size(); ++n)
        freeNodes += (hc.nflags[n] & SHIPFEM_PIN_ALL) != SHIPFEM_PIN_ALL;
    // a tet pass reads eqps of every tet and the plastic strain (and damage) of those that
    // have yielded
    const uint64_t history = (f & kFeatPlastic) ? 4 * T + ((f & kFeatDamage) ? 28 : 24) * yielded : 0;
    if (hc.corners == 8) {
        const bool plastic = (f & kFeatPlastic) != 0;
        // connectivity, the stiffness block (elastic) or the FP64 rest geometry and the point
        // history with the enhanced parameters (plastic), and the displacement with its
        // compensation of every node; the 24-float force slots out (and the enhanced
        // parameters, plastic)
        // (plastic: every point's eqps, and the other seven words of each point that has
        // yielded, `yielded` counting points)
        b.elem_read = 32 * T + 28 * N + (hc.hfail.empty() ? 0 : 8 * T) +
                      (plastic ? (8ull * kHexGeo + 8ull * (8 + 7)) * T + 56 * yielded
                               : (8ull * kHexPacked + 4ull * kHexGeo) * T);
        b.elem_write = 96 * T + (plastic ? 56 * T : 0);
    } else if (variant == kSlots) {
        b.elem_read = 16 * T + 36 * T + 4 * T + 16 * N + history;
        b.elem_write = 48 * T;
    } else {
        // tile offsets, each tile's node ids and run offsets, 16-bit local connectivity, Dm^-1
        // and V (or, formed in the pass, each tile node's rest position), every node's
        // displacement once; one float4 partial per (tile, node)
        b.elem_read = 4 * (nt + 1) + 4 * P + 2 * (P + nt) + 8 * T + (formed ? 16 * P : 40 * T) + 16 * N + history;
        b.elem_write = 16 * P;
    }
    const bool slots = hc.corners == 8 || variant == kSlots;
    const uint64_t slab = slots ? (hc.ell_off.empty() ? 0 : hc.ell_off.back()) : (hc.pell_off.empty() ? 0 : hc.pell_off.back());
    const uint64_t gathered = slots ? 12ull * hc.corners * T : 16 * P;
    const bool needMass = (f & (kFeatGravity | kFeatContact | kFeatDrive | kFeatDamp)) != 0;
    // the gather slab and its offsets, the gathered forces, flags, 1/m (and m), u, its
    // compensation, v and its compensation, the optional load, contact and drive planes;
    // a free node writes u, v and both compensations
    b.node_read = 4 * (hc.node_waves + 1) + 4 * slab + gathered + 4 * Np + 4 * Np + (needMass ? 4 * Np : 0) +
                  16 * N + 36 * Np + ((f & kFeatLoad) ? 12 * Np : 0) + ((f & kFeatContact) ? 16 * Np : 0) +
                  ((f & kFeatDrive) ? 12 * Np : 0);
    b.node_write = 52 * freeNodes;
    return b;
}

hipError_t yieldedTets(const DeviceCase& dc, const HostCase& hc, hipStream_t s, uint64_t* count) {
    *count = 0;
    if (hc.corners == 8 && (hc.features & kFeatPlastic) && dc.rec.hpoint && hc.T > 0) {
        // a hex counts its integration points: eqps of point g is plane 8 g + 6
        std::vector<uint64_t> q(hc.T);
        for (int g = 0; g < 8; ++g) {
            hipError_t e = hipMemcpyAsync(q.data(), dc.rec.hpoint + (8ull * g + 6) * hc.Tpad, 8ull * hc.T,
                                          hipMemcpyDeviceToHost, s);
            if (e == hipSuccess) e = hipStreamSynchronize(s);
            if (e != hipSuccess) return e;
            for (uint64_t v : q) *count += v != 0u;
        }
        return hipSuccess;
    }
    if (hc.corners != 4 || !(hc.features & kFeatPlastic) || !dc.rec.pstate || hc.T == 0) return hipSuccess;
    std::vector<uint32_t> q(hc.T);
    hipError_t e = hipMemcpyAsync(q.data(), dc.rec.pstate + 6ull * hc.Tpad, 4ull * hc.T, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return e;
    for (uint32_t v : q) *count += v != 0u;  // eqps bits not +0
    return hipSuccess;
}

hipError_t uploadCase(const HostCase& hc, char* base, DeviceCase& dc, hipStream_t s) {
    const ArenaMap m = mapArena(hc);
    dc.arena = base;
    dc.bytes = m.total;
    CaseDev& r = dc.rec;
    r = CaseDev{};
    r.N = hc.N;
    r.T = hc.T;
    r.Npad = hc.Npad;
    r.Tpad = hc.Tpad;
    r.node_waves = hc.node_waves;
    r.elem_waves = hc.elem_waves;
    r.features = hc.features;
    r.corners = hc.corners;
    r.conn = reinterpret_cast<const uint4*>(base + m.conn);
    r.hk = hc.hk.empty() ? nullptr : reinterpret_cast<const double*>(base + m.hk);
    r.hgeo = hc.hgeo.empty() ? nullptr : reinterpret_cast<const float*>(base + m.hgeo);
    const bool hexPlastic = hc.corners == 8 && (hc.features & kFeatPlastic);
    r.hgeod = hexPlastic ? reinterpret_cast<const double*>(base + m.hgeod) : nullptr;
    r.hpoint = hexPlastic ? reinterpret_cast<double*>(base + m.hpoint) : nullptr;
    r.henh = hexPlastic ? reinterpret_cast<double*>(base + m.henh) : nullptr;
    r.hfail = hc.hfail.empty() ? nullptr : reinterpret_cast<const double*>(base + m.hfail);
    r.dmi = reinterpret_cast<const float*>(base + m.dmi);
    r.vol = reinterpret_cast<const float*>(base + m.vol);
    r.force = reinterpret_cast<float*>(base + m.force);
    r.pstate = (hc.features & kFeatPlastic) && !hexPlastic ? reinterpret_cast<float*>(base + m.pstate)
                                                           : nullptr;
    r.tear_step = (hc.features & kFeatDamage) ? reinterpret_cast<int32_t*>(base + m.tear_step) : nullptr;
    r.tear_stress = (hc.features & kFeatDamage) ? reinterpret_cast<float*>(base + m.tear_stress) : nullptr;
    r.u4 = reinterpret_cast<float4*>(base + m.u4);
    r.cu = reinterpret_cast<float*>(base + m.cu);
    r.v = reinterpret_cast<float*>(base + m.v);
    r.cv = reinterpret_cast<float*>(base + m.cv);
    r.invm = reinterpret_cast<const float*>(base + m.invm);
    r.mass = reinterpret_cast<const float*>(base + m.mass);
    r.nflags = reinterpret_cast<const uint32_t*>(base + m.nflags);
    r.load = (hc.features & kFeatLoad) ? reinterpret_cast<const float*>(base + m.load) : nullptr;
    r.xc = (hc.features & kFeatContact) ? reinterpret_cast<const float*>(base + m.xc) : nullptr;
    r.presc = (hc.features & kFeatDrive) ? reinterpret_cast<const float*>(base + m.presc) : nullptr;
    r.ell_off = reinterpret_cast<const uint32_t*>(base + m.ell_off);
    r.ell = reinterpret_cast<const uint32_t*>(base + m.ell);
    r.lconn = reinterpret_cast<const uint2*>(base + m.lconn);
    r.tn_off = reinterpret_cast<const uint32_t*>(base + m.tn_off);
    r.tnode = reinterpret_cast<const uint32_t*>(base + m.tnode);
    r.trest = reinterpret_cast<const float4*>(base + m.trest);
    r.lofs = reinterpret_cast<const uint16_t*>(base + m.lofs);
    r.partial = reinterpret_cast<float4*>(base + m.partial);
    r.pell_off = reinterpret_cast<const uint32_t*>(base + m.pell_off);
    r.pell = reinterpret_cast<const uint32_t*>(base + m.pell);
    r.elem_tiles = hc.elem_tiles;
    r.P = hc.P;
    r.lmax = hc.lmax;
    r.load_ramp = hc.load_ramp;
    r.drive_ramp = hc.drive_ramp;
    for (int a = 0; a < 3; ++a) r.ind_vel[a] = hc.indenter.velocity[a];
    r.lam = static_cast<float>(hc.lam);
    r.mu = static_cast<float>(hc.mu);
    r.bulk = static_cast<float>(hc.bulk);
    r.yield = static_cast<float>(hc.yield);
    r.hard = static_cast<float>(hc.hard);
    r.fail_strain = static_cast<float>(hc.fail_strain);
    r.fail_sens = static_cast<float>(hc.fail_sens);
    r.fail_eref = static_cast<float>(hc.fail_eref);
    r.fail_ecut = static_cast<float>(hc.fail_ecut);
    r.gravity = static_cast<float>(hc.gravity);
    r.damping = static_cast<float>(hc.damping);
    r.ind_r = static_cast<float>(hc.indenter.radius);
    r.lam_d = hc.lam;
    r.mu_d = hc.mu;
    r.yield_d = hc.yield;
    r.hard_d = hc.hard;
    r.fail_d = hc.fail_strain;
    r.sens_d = hc.fail_sens;
    r.eref_d = hc.fail_eref;
    r.ecut_d = hc.fail_ecut;
    {
        const hexp::Material hm = hexp::derive(hexp::Material{hc.mu, hc.lam + (2.0 / 3.0) * hc.mu, hc.yield, hc.hard});
        const double d[5] = {hm.igam, hm.thb0, hm.i2mu, hm.i3k, hm.imu};
        std::copy(d, d + 5, r.hret_d);
    }
    r.acc_e = reinterpret_cast<double*>(base + m.acc_e);
    r.acc_n = reinterpret_cast<double*>(base + m.acc_n);
    r.smp_e = reinterpret_cast<double*>(base + m.smp_e);
    r.smp_n = reinterpret_cast<double*>(base + m.smp_n);
    r.omega_w = reinterpret_cast<double*>(base + m.omega_w);
    r.step = reinterpret_cast<int64_t*>(base + m.step);
    r.step_end = hc.max_steps > 0 ? hc.max_steps : std::numeric_limits<int64_t>::max();
    r.status = reinterpret_cast<uint32_t*>(base + m.status);
    r.hist = reinterpret_cast<EnergyRecordDev*>(base + m.hist);
    r.hist_cap = hc.hist_cap;
    r.hist_count = reinterpret_cast<int32_t*>(base + m.hist_count);

    hipError_t e = hipMemsetAsync(base, 0, m.total, s);
    auto up = [&](uint64_t off, const void* src, uint64_t bytes) {
        if (e == hipSuccess && bytes) e = hipMemcpyAsync(base + off, src, bytes, hipMemcpyHostToDevice, s);
    };
    up(m.conn, hc.conn.data(), 4ull * hc.conn.size());
    up(m.dmi, hc.dmi.data(), 4ull * hc.dmi.size());
    up(m.vol, hc.vol.data(), 4ull * hc.vol.size());
    up(m.hk, hc.hk.data(), 8ull * hc.hk.size());
    up(m.hgeo, hc.hgeo.data(), 4ull * hc.hgeo.size());
    up(m.hgeod, hc.hgeod.data(), 8ull * hc.hgeod.size());
    up(m.hfail, hc.hfail.data(), 8ull * hc.hfail.size());
    up(m.u4, hc.u4.data(), 4ull * hc.u4.size());
    up(m.cu, hc.cu.data(), 4ull * hc.cu.size());
    up(m.v, hc.v.data(), 4ull * hc.v.size());
    up(m.cv, hc.cv.data(), 4ull * hc.cv.size());
    up(m.invm, hc.invm.data(), 4ull * hc.invm.size());
    up(m.mass, hc.mass.data(), 4ull * hc.mass.size());
    up(m.nflags, hc.nflags.data(), 4ull * hc.nflags.size());
    up(m.load, hc.load.data(), 4ull * hc.load.size());
    up(m.xc, hc.xc.data(), 4ull * hc.xc.size());
    up(m.presc, hc.presc.data(), 4ull * hc.presc.size());
    up(m.ell_off, hc.ell_off.data(), 4ull * hc.ell_off.size());
    up(m.ell, hc.ell.data(), 4ull * hc.ell.size());
    up(m.lconn, hc.lconn.data(), 4ull * hc.lconn.size());
    up(m.tn_off, hc.tn_off.data(), 4ull * hc.tn_off.size());
    up(m.tnode, hc.tnode.data(), 4ull * hc.tnode.size());
    up(m.trest, hc.trest.data(), 4ull * hc.trest.size());
    up(m.lofs, hc.lofs.data(), 2ull * hc.lofs.size());
    up(m.pell_off, hc.pell_off.data(), 4ull * hc.pell_off.size());
    up(m.pell, hc.pell.data(), 4ull * hc.pell.size());
    return e;
}

void setTimestep(DeviceCase& dc, const HostCase& hc, double dt) {
    // the kernels use dt in FP32 for the update and in FP64 for time and work: keep
    // the two identical by taking dt at float precision
    dt = static_cast<double>(static_cast<float>(dt));
    dc.dt = dt;
    dc.rec.dt = dt;
    dc.rec.dtf = static_cast<float>(dt);
    dc.rec.ind_kdt2 = static_cast<float>(hc.indenter.penalty / (dt * dt));
}

hipError_t readbackCase(const DeviceCase& dc, const HostCase& hc, shipfem_state_view* view,
                        hipStream_t s) {
    const CaseDev& r = dc.rec;
    const uint64_t N = hc.N, Np = hc.Npad;
    std::vector<float> u4(4 * N), pl(9 * Np);
    int64_t step[2] = {0, 0};
    uint32_t status = 0;
    hipError_t e = hipMemcpyAsync(u4.data(), r.u4, 16 * N, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(pl.data(), r.cu, 12 * Np, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(pl.data() + 3 * Np, r.v, 12 * Np, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(pl.data() + 6 * Np, r.cv, 12 * Np, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(step, r.step, 16, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(&status, r.status, 4, hipMemcpyDeviceToHost, s);
    std::vector<float> ps;
    std::vector<double> hp;  // plastic hex points
    if (r.hpoint && (view->eqps || view->plastic_strain || view->damage)) {
        hp.resize(uint64_t(kHexPoint) * hc.Tpad);
        if (e == hipSuccess)
            e = hipMemcpyAsync(hp.data(), r.hpoint, 8 * hp.size(), hipMemcpyDeviceToHost, s);
    }
    const bool dmg = (hc.features & kFeatDamage) != 0;
    if (r.pstate && (view->eqps || view->plastic_strain || view->damage)) {
        ps.resize((dmg ? 8 : 7) * (uint64_t)hc.Tpad);
        if (e == hipSuccess)
            e = hipMemcpyAsync(ps.data(), r.pstate, 4 * ps.size(), hipMemcpyDeviceToHost, s);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return e;
    const float* cu = pl.data();
    const float* v = pl.data() + 3 * Np;
    const float* cv = pl.data() + 6 * Np;
    // node and tet results go back in the caller's orders
    for (uint32_t n = 0; n < N; ++n) {
        const uint64_t x = externalNode(hc, n);
        for (int a = 0; a < 3; ++a) {
            const double u = (double)u4[n * 4ull + a] - (double)cu[a * Np + n];
            if (view->displacement) view->displacement[x * 3 + a] = u;
            if (view->velocity) view->velocity[x * 3 + a] = (double)v[a * Np + n] - (double)cv[a * Np + n];
            if (view->position) view->position[x * 3 + a] = static_cast<float>(hc.X[n * 3ull + a] + u);
        }
    }
    const uint64_t Tp = hc.Tpad;
    const uint32_t T = hc.T;
    if (!hp.empty()) {
        // a hex element reports its largest point eqps, its points' mean plastic strain and
        // its smallest point damage (1: every point has failed, the element is torn)
        for (uint32_t t = 0; t < T; ++t) {
            double q = 0, dmin = 1.0, ep[6] = {0, 0, 0, 0, 0, 0};
            for (int g = 0; g < 8; ++g) {
                q = std::max(q, hp[(8 * g + 6) * Tp + t]);
                dmin = std::min(dmin, hp[(8 * g + 7) * Tp + t]);
                for (int k = 0; k < 6; ++k) ep[k] += 0.125 * hp[(8 * g + k) * Tp + t];
            }
            if (view->eqps) view->eqps[t] = static_cast<float>(q);
            if (view->damage) view->damage[t] = dmg ? static_cast<float>(dmin) : 0.0f;
            if (view->plastic_strain)
                for (int k = 0; k < 6; ++k) view->plastic_strain[t * 6ull + k] = static_cast<float>(ep[k]);
        }
        view->step = std::max(step[0], step[1]);
        view->time = static_cast<double>(view->step) * dc.dt;
        view->status = status;
        return hipSuccess;
    }
    if (view->eqps)
        for (uint32_t t = 0; t < T; ++t) view->eqps[externalTet(hc, t)] = ps.empty() ? 0.0f : ps[6 * Tp + t];
    if (view->plastic_strain)
        for (uint32_t t = 0; t < T; ++t)
            for (int k = 0; k < 6; ++k)
                view->plastic_strain[externalTet(hc, t) * 6ull + k] = ps.empty() ? 0.0f : ps[k * Tp + t];
    if (view->damage)
        for (uint32_t t = 0; t < T; ++t)
            view->damage[externalTet(hc, t)] = dmg && !ps.empty() ? ps[7 * Tp + t] : 0.0f;
    view->step = std::max(step[0], step[1]);
    view->time = static_cast<double>(view->step) * dc.dt;
    view->status = status;
    return hipSuccess;
}

hipError_t readEnergy(const DeviceCase& dc, shipfem_energy_record* out, int32_t max,
                      int32_t* count, hipStream_t s) {
    int32_t n = 0;
    hipError_t e = hipMemcpyAsync(&n, dc.rec.hist_count, 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return e;
    n = std::min(n, dc.rec.hist_cap);
    const int32_t k = out ? std::min(n, max) : 0;
    if (k > 0) {
        static_assert(sizeof(EnergyRecordDev) == sizeof(shipfem_energy_record), "record layout");
        e = hipMemcpyAsync(out, dc.rec.hist, sizeof(EnergyRecordDev) * k, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
    }
    if (count) *count = n;
    return e;
}

hipError_t readTears(const DeviceCase& dc, const HostCase& hc, std::vector<shipfem_tear>& out,
                     hipStream_t s) {
    out.clear();
    if (!(hc.features & kFeatDamage)) return hipSuccess;
    const uint64_t Tp = hc.Tpad;
    std::vector<int32_t> step(Tp);
    std::vector<float> st(6 * Tp);
    hipError_t e = hipMemcpyAsync(step.data(), dc.rec.tear_step, 4 * Tp, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess)
        e = hipMemcpyAsync(st.data(), dc.rec.tear_stress, 24 * Tp, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return e;
    for (uint32_t t = 0; t < hc.T; ++t) {
        if (step[t] == 0) continue;
        shipfem_tear tear{};
        tear.tet = externalTet(hc, t);
        tear.step = static_cast<int64_t>(step[t]) - 1;
        double v[6];
        for (int k = 0; k < 6; ++k) v[k] = st[k * Tp + t];
        if (hc.corners == 4) maxPrincipalDirection(v, tear.normal);  // a hex reports no plane
        out.push_back(tear);
    }
    std::sort(out.begin(), out.end(), [](const shipfem_tear& a, const shipfem_tear& b) {
        return a.step < b.step || (a.step == b.step && a.tet < b.tet);
    });
    return hipSuccess;
}

// Largest-eigenvalue direction of a symmetric 3x3 (Voigt, tensor shears) by cyclic
// Jacobi in FP64; sign fixed so the largest component is positive.
void maxPrincipalDirection(const double s6[6], double out[3]) {
    double a[9] = {s6[0], s6[3], s6[5], s6[3], s6[1], s6[4], s6[5], s6[4], s6[2]};
    double v[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    for (int sweep = 0; sweep < 50; ++sweep) {
        const double off = a[1] * a[1] + a[2] * a[2] + a[5] * a[5];
        const double diag = a[0] * a[0] + a[4] * a[4] + a[8] * a[8];
        if (off == 0.0 || off <= 1e-30 * diag) break;
        for (int p = 0; p < 2; ++p)
            for (int q = p + 1; q < 3; ++q) {
                const double apq = a[p * 3 + q];
                if (apq == 0.0) continue;
                const double th = (a[q * 3 + q] - a[p * 3 + p]) / (2.0 * apq);
                const double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), sn = t * c;
                for (int k = 0; k < 3; ++k) {
                    const double x = a[k * 3 + p], y = a[k * 3 + q];
                    a[k * 3 + p] = c * x - sn * y;
                    a[k * 3 + q] = sn * x + c * y;
                }
                for (int k = 0; k < 3; ++k) {
                    const double x = a[p * 3 + k], y = a[q * 3 + k];
                    a[p * 3 + k] = c * x - sn * y;
                    a[q * 3 + k] = sn * x + c * y;
                }
                for (int k = 0; k < 3; ++k) {
                    const double x = v[k * 3 + p], y = v[k * 3 + q];
                    v[k * 3 + p] = c * x - sn * y;
                    v[k * 3 + q] = sn * x + c * y;
                }
            }
    }
    int best = 0;
    for (int i = 1; i < 3; ++i)
        if (a[i * 3 + i] > a[best * 3 + best]) best = i;
    for (int i = 0; i < 3; ++i) out[i] = v[i * 3 + best];
    double norm = std::sqrt(out[0] * out[0] + out[1] * out[1] + out[2] * out[2]);
    if (!(norm > 0.0)) {
        out[0] = 1.0;
        out[1] = out[2] = 0.0;
        return;
    }
    int largest = 0;
    for (int i = 1; i < 3; ++i)
        if (std::fabs(out[i]) > std::fabs(out[largest])) largest = i;
    if (out[largest] < 0.0) norm = -norm;
    for (int i = 0; i < 3; ++i) out[i] /= norm;
}

}  // namespace shipfem
