// handoff.cpp -- a solve's state in and out of the device: export in the shape
// zone::Solver::adopt takes (engine/gpu/zone_gpu.hpp:113-126, zone.hpp:963-969), adopt a
// state computed elsewhere, bitwise checkpoint and restore, and the Tier-1 edge drive
// (coupling::applyEdgeDrive, engine/sim/coupling.cpp:110-125).
#include "shipfem.h"
#include "shipfem_host.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

using namespace shipfem;

namespace {

// The true value is hi - c (the kernels' Kahan convention).
void splitValue(double x, float& hi, float& c) {
    hi = static_cast<float>(x);
    c = static_cast<float>(static_cast<double>(hi) - x);
}

hipError_t toHost(void* dst, const void* src, uint64_t bytes, hipStream_t s) {
    return bytes ? hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s) : hipSuccess;
}
hipError_t toDevice(void* dst, const void* src, uint64_t bytes, hipStream_t s) {
    return bytes ? hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s) : hipSuccess;
}

bool plasticHex(const HostCase& hc) { return hc.corners == 8 && (hc.features & kFeatPlastic); }

uint32_t pstatePlanes(const HostCase& hc) {
    if (!(hc.features & kFeatPlastic) || hc.corners == 8) return 0;
    return (hc.features & kFeatDamage) ? 8 : 7;
}

// The evolving part of a case's arena, as (device pointer, bytes) pieces in a fixed
// order; shared by checkpoint and restore.
struct Piece {
    void* dev;
    uint64_t bytes;
};

std::vector<Piece> evolvingState(const shipfem_solver* s) {
    const HostCase& hc = s->hc;
    const CaseDev& r = s->dc.rec;
    const uint64_t Np = hc.Npad, Tp = hc.Tpad;
    const bool dmg = (hc.features & kFeatDamage) != 0;
    return {
        {r.u4, 16ull * hc.N},
        {r.cu, 12 * Np},
        {r.v, 12 * Np},
        {r.cv, 12 * Np},
        {r.pstate, 4ull * pstatePlanes(hc) * Tp},
        {r.hpoint, r.hpoint ? 8ull * kHexPoint * Tp : 0},
        {r.henh, r.henh ? 8ull * 7 * Tp : 0},
        {r.tear_step, dmg ? 4 * Tp : 0},
        {r.tear_stress, dmg ? 24 * Tp : 0},
        {r.acc_e, 8ull * kAccElemSlots * hc.elem_waves},
        {r.acc_n, 8ull * kAccNodeSlots * hc.node_waves},
        {r.status, 4},
        {r.hist_count, 4},
        {r.hist, sizeof(EnergyRecordDev) * static_cast<uint64_t>(r.hist_cap)},
    };
}

struct CheckpointHeader {
    char magic[8];  // "SHIPFEM1"
    uint32_t N, T, Npad, Tpad, node_waves, elem_waves, features;
    int32_t hist_cap;
    int64_t step;
    double dt;
    uint64_t payload;
};

// Both step counter slots hold n: whichever parity the next launch reads, it sees n.
hipError_t setStep(shipfem_solver* s, int64_t n) {
    const int64_t step[2] = {n, n};
    hipError_t e = toDevice(s->dc.rec.step, step, 16, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);  // `step` is local
    s->host_step = n;
    s->plan.parity = 0;
    return e;
}

}  // namespace

extern "C" {

shipfem_status shipfem_solver_export(shipfem_solver* s, shipfem_handoff* out) {
    if (!s || !out) return SHIPFEM_INVALID_ARG;
    const HostCase& hc = s->hc;
    const CaseDev& r = s->dc.rec;
    const uint64_t N = hc.N, T = hc.T, Np = hc.Npad, Tp = hc.Tpad;
    const hipStream_t st = s->stream;
    std::vector<float> u4(4 * N), pl(9 * Np), ps(static_cast<uint64_t>(pstatePlanes(hc)) * Tp);
    std::vector<double> ae(kAccElemSlots * (uint64_t)hc.elem_waves), an(kAccNodeSlots * (uint64_t)hc.node_waves);
    const bool dmg = (hc.features & kFeatDamage) != 0;
    std::vector<int32_t> tstep(dmg ? Tp : 0);
    std::vector<float> tstress(dmg ? 6 * Tp : 0);
    int64_t step[2] = {0, 0};
    hipError_t e = toHost(u4.data(), r.u4, 16 * N, st);
    if (e == hipSuccess) e = toHost(pl.data(), r.cu, 12 * Np, st);
    if (e == hipSuccess) e = toHost(pl.data() + 3 * Np, r.v, 12 * Np, st);
    if (e == hipSuccess) e = toHost(pl.data() + 6 * Np, r.cv, 12 * Np, st);
    if (e == hipSuccess) e = toHost(ps.data(), r.pstate, 4 * ps.size(), st);
    if (e == hipSuccess) e = toHost(ae.data(), r.acc_e, 8 * ae.size(), st);
    if (e == hipSuccess) e = toHost(an.data(), r.acc_n, 8 * an.size(), st);
    if (e == hipSuccess) e = toHost(tstep.data(), r.tear_step, 4 * tstep.size(), st);
    if (e == hipSuccess) e = toHost(tstress.data(), r.tear_stress, 4 * tstress.size(), st);
    const bool hexp = plasticHex(hc);
    std::vector<double> hp(hexp ? (uint64_t)kHexPoint * Tp : 0), he(hexp ? 7ull * Tp : 0);
    if (e == hipSuccess) e = toHost(hp.data(), r.hpoint, 8 * hp.size(), st);
    if (e == hipSuccess) e = toHost(he.data(), r.henh, 8 * he.size(), st);
    if (e == hipSuccess) e = toHost(step, r.step, 16, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return statusOf(e);

    const float* cu = pl.data();
    const float* v = pl.data() + 3 * Np;
    const float* cv = pl.data() + 6 * Np;
    // node and tet arrays go out in the caller's orders
    for (uint32_t n = 0; n < N; ++n) {
        const uint64_t x = externalNode(hc, n);
        for (int a = 0; a < 3; ++a) {
            const double u = (double)u4[n * 4ull + a] - (double)cu[a * Np + n];
            if (out->position) out->position[x * 3 + a] = hc.X[n * 3ull + a] + u;
            if (out->displacement) out->displacement[x * 3 + a] = u;
            if (out->velocity) out->velocity[x * 3 + a] = (double)v[a * Np + n] - (double)cv[a * Np + n];
        }
    }
    if (out->plastic && !hexp)
        for (uint32_t t = 0; t < T; ++t)
            for (uint32_t k = 0; k < 8; ++k)
                out->plastic[externalTet(hc, t) * 8ull + k] = k < pstatePlanes(hc) ? (double)ps[k * Tp + t] : 0.0;
    if (hexp)
        for (uint32_t t = 0; t < T; ++t) {
            const uint64_t x = externalTet(hc, t);
            double eq = 0.0, dmin = 1.0;
            for (uint32_t g = 0; g < 8; ++g) {
                eq = std::max(eq, hp[(8ull * g + 6) * Tp + t]);
                dmin = std::min(dmin, hp[(8ull * g + 7) * Tp + t]);
                if (out->hex_points)
                    for (uint32_t k = 0; k < 8; ++k) out->hex_points[x * 64 + 8 * g + k] = hp[(8ull * g + k) * Tp + t];
            }
            if (out->plastic) {
                for (uint32_t k = 0; k < 6; ++k) out->plastic[x * 8 + k] = 0.0;
                out->plastic[x * 8 + 6] = eq;
                out->plastic[x * 8 + 7] = dmin;
            }
            if (out->hex_enhanced)
                for (uint32_t k = 0; k < 7; ++k) out->hex_enhanced[x * 7 + k] = he[k * Tp + t];
        }
    uint32_t torn = 0;
    for (uint32_t t = 0; t < T; ++t) {
        const uint64_t x = externalTet(hc, t);
        const bool is_torn = dmg && tstep[t] != 0;
        torn += is_torn ? 1u : 0u;
        if (out->tear_step) out->tear_step[x] = is_torn ? static_cast<int64_t>(tstep[t]) - 1 : -1;
        if (out->tear_stress)
            for (int k = 0; k < 6; ++k) out->tear_stress[x * 6 + k] = is_torn ? (double)tstress[k * Tp + t] : 0.0;
    }
    // the account's running sums, wave by wave in order
    auto total = [](const std::vector<double>& a, uint32_t slot, uint32_t waves) {
        double x = 0.0;
        for (uint32_t w = 0; w < waves; ++w) x += a[(uint64_t)slot * waves + w];
        return x;
    };
    out->steps = std::max(step[0], step[1]);
    out->time = static_cast<double>(out->steps) * s->dc.dt;
    out->dissipation = total(ae, kAccDiss, hc.elem_waves);
    out->fracture = total(ae, kAccFract, hc.elem_waves);
    out->w_ext = total(an, kAccWext, hc.node_waves);
    out->w_contact = total(an, kAccWcon, hc.node_waves);
    out->w_boundary = total(an, kAccWbnd, hc.node_waves);
    out->damping_loss = total(an, kAccDamp, hc.node_waves);
    out->work = out->w_ext + out->w_contact + out->w_boundary;
    const double* iv = hc.indenter.velocity;
    out->penetration = hc.indenter.enabled
                           ? std::sqrt(iv[0] * iv[0] + iv[1] * iv[1] + iv[2] * iv[2]) * out->time
                           : 0.0;
    out->torn = torn;
    return SHIPFEM_OK;
}

shipfem_status shipfem_solver_adopt(shipfem_solver* s, const shipfem_handoff* in) {
    if (!s || !in || !(in->position || in->displacement) || !in->velocity || in->steps < 0)
        return SHIPFEM_INVALID_ARG;
    const HostCase& hc = s->hc;
    const CaseDev& r = s->dc.rec;
    const double want = static_cast<double>(in->steps) * s->dc.dt;
    if (!(std::fabs(in->time - want) <= 1e-9 * std::max(std::fabs(want), s->dc.dt)))
        return SHIPFEM_INVALID_ARG;
    const uint64_t N = hc.N, T = hc.T, Np = hc.Npad, Tp = hc.Tpad;
    std::vector<float> u4(4 * N, 0.0f), pl(9 * Np, 0.0f);
    float* cu = pl.data();
    float* v = pl.data() + 3 * Np;
    float* cv = pl.data() + 6 * Np;
    for (uint32_t n = 0; n < N; ++n) {
        const uint64_t e = externalNode(hc, n);
        for (int a = 0; a < 3; ++a) {
            const double x = in->displacement ? in->displacement[e * 3 + a]
                                              : in->position[e * 3 + a] - hc.X[n * 3ull + a];
            const double vel = in->velocity[e * 3 + a];
            if (!std::isfinite(x) || !std::isfinite(vel)) return SHIPFEM_INVALID_ARG;
            splitValue(x, u4[n * 4ull + a], cu[a * Np + n]);
            splitValue(vel, v[a * Np + n], cv[a * Np + n]);
        }
    }
    // a plastic hex's point history and enhanced parameters, checked before anything is sent
    const bool hexp = plasticHex(hc) && in->hex_points;
    std::vector<double> hp(hexp ? (uint64_t)kHexPoint * Tp : 0, 0.0), he(hexp ? 7ull * Tp : 0, 0.0);
    if (hexp)
        for (uint32_t t = 0; t < T; ++t) {
            const uint64_t x = externalTet(hc, t);
            for (uint32_t q = 0; q < (uint32_t)kHexPoint; ++q) {
                const double val = in->hex_points[x * 64 + q];
                if (!std::isfinite(val)) return SHIPFEM_INVALID_ARG;
                hp[q * Tp + t] = val;
            }
            // the pass reads a point's plastic strain and damage only when its eqps is not +0:
            // a point handed over with zero eqps and something else not zero keeps eqps as -0
            for (uint32_t g = 0; g < 8; ++g) {
                bool any = false;
                for (uint32_t k = 0; k < 8; ++k) any = any || (k != 6 && hp[(8ull * g + k) * Tp + t] != 0.0);
                if (any && hp[(8ull * g + 6) * Tp + t] == 0.0) hp[(8ull * g + 6) * Tp + t] = -0.0;
            }
            if (in->hex_enhanced)
                for (uint32_t k = 0; k < 7; ++k) {
                    const double val = in->hex_enhanced[x * 7 + k];
                    if (!std::isfinite(val)) return SHIPFEM_INVALID_ARG;
                    he[k * Tp + t] = val;
                }
        }
    const hipStream_t st = s->stream;
    hipError_t e = toDevice(r.u4, u4.data(), 16 * N, st);
    if (e == hipSuccess) e = toDevice(r.cu, cu, 12 * Np, st);
    if (e == hipSuccess) e = toDevice(r.v, v, 12 * Np, st);
    if (e == hipSuccess) e = toDevice(r.cv, cv, 12 * Np, st);
    const uint32_t planes = pstatePlanes(hc);
    const bool dmg = (hc.features & kFeatDamage) != 0;
    std::vector<float> ps, tstress;
    std::vector<int32_t> tstep;
    if (in->plastic && planes) {
        ps.assign(planes * Tp, 0.0f);
        for (uint32_t t = 0; t < T; ++t)
            for (uint32_t k = 0; k < planes; ++k)
                ps[k * Tp + t] = static_cast<float>(in->plastic[externalTet(hc, t) * 8ull + k]);
        // the element pass reads a tet's plastic strain and damage only when its eqps is not
        // +0 (zero eqps means never yielded for any state the device makes); history handed
        // over with zero eqps and something else not zero keeps its eqps as -0 so it is read
        for (uint32_t t = 0; t < T; ++t) {
            bool any = false;
            for (uint32_t k = 0; k < planes; ++k) any = any || (k != 6 && ps[k * Tp + t] != 0.0f);
            if (any && ps[6 * Tp + t] == 0.0f) ps[6 * Tp + t] = -0.0f;
        }
        if (e == hipSuccess) e = toDevice(r.pstate, ps.data(), 4 * ps.size(), st);
        if (dmg) {
            // a torn tet handed over without its tear step is recorded as torn at `steps`
            tstep.assign(Tp, 0);
            tstress.assign(6 * Tp, 0.0f);
            for (uint32_t t = 0; t < T; ++t) {
                if (!(ps[7 * Tp + t] >= 1.0f)) continue;
                const uint64_t x = externalTet(hc, t);
                const int64_t at = in->tear_step && in->tear_step[x] >= 0 ? in->tear_step[x] : in->steps;
                tstep[t] = static_cast<int32_t>(at + 1);
                if (in->tear_stress)
                    for (int k = 0; k < 6; ++k) tstress[k * Tp + t] = static_cast<float>(in->tear_stress[x * 6 + k]);
            }
            if (e == hipSuccess) e = toDevice(r.tear_step, tstep.data(), 4 * Tp, st);
            if (e == hipSuccess) e = toDevice(r.tear_stress, tstress.data(), 24 * Tp, st);
        }
    }
    if (hexp) {
        if (e == hipSuccess) e = toDevice(r.hpoint, hp.data(), 8 * hp.size(), st);
        if (e == hipSuccess) e = toDevice(r.henh, he.data(), 8 * he.size(), st);
        if (dmg) {
            // an element whose eight points have failed is torn; one handed over without its
            // tear step is recorded as torn at `steps` (the hex pass keeps no tear stress)
            tstep.assign(Tp, 0);
            tstress.assign(6 * Tp, 0.0f);
            for (uint32_t t = 0; t < T; ++t) {
                bool torn = true;
                for (uint32_t g = 0; g < 8; ++g) torn = torn && hp[(8ull * g + 7) * Tp + t] >= 1.0;
                if (!torn) continue;
                const uint64_t x = externalTet(hc, t);
                const int64_t at = in->tear_step && in->tear_step[x] >= 0 ? in->tear_step[x] : in->steps;
                tstep[t] = static_cast<int32_t>(at + 1);
            }
            if (e == hipSuccess) e = toDevice(r.tear_step, tstep.data(), 4 * Tp, st);
            if (e == hipSuccess) e = toDevice(r.tear_stress, tstress.data(), 24 * Tp, st);
        }
    }
    // running sums continue from the handed-over totals: each sits in the first wave
    std::vector<double> ae(kAccElemSlots * (uint64_t)hc.elem_waves, 0.0);
    std::vector<double> an(kAccNodeSlots * (uint64_t)hc.node_waves, 0.0);
    if (!ae.empty()) {
        ae[(uint64_t)kAccDiss * hc.elem_waves] = in->dissipation;
        ae[(uint64_t)kAccFract * hc.elem_waves] = in->fracture;
    }
    if (!an.empty()) {
        an[(uint64_t)kAccWext * hc.node_waves] = in->w_ext;
        an[(uint64_t)kAccWcon * hc.node_waves] = in->w_contact;
        an[(uint64_t)kAccWbnd * hc.node_waves] = in->w_boundary;
        an[(uint64_t)kAccDamp * hc.node_waves] = in->damping_loss;
    }
    const uint32_t zero[2] = {0u, 0u};
    if (e == hipSuccess) e = toDevice(r.acc_e, ae.data(), 8 * ae.size(), st);
    if (e == hipSuccess) e = toDevice(r.acc_n, an.data(), 8 * an.size(), st);
    if (e == hipSuccess) e = toDevice(r.status, &zero[0], 4, st);
    if (e == hipSuccess) e = toDevice(r.hist_count, &zero[1], 4, st);
    if (e == hipSuccess) e = setStep(s, in->steps);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    return statusOf(e);
}

shipfem_status shipfem_solver_checkpoint(shipfem_solver* s, void* buf, uint64_t* len) {
    if (!s || !len) return SHIPFEM_INVALID_ARG;
    const std::vector<Piece> pieces = evolvingState(s);
    uint64_t payload = 0;
    for (const Piece& p : pieces) payload += p.bytes;
    const uint64_t need = sizeof(CheckpointHeader) + payload;
    if (!buf) {
        *len = need;
        return SHIPFEM_OK;
    }
    if (*len < need) return SHIPFEM_INVALID_ARG;
    const HostCase& hc = s->hc;
    CheckpointHeader h{};
    std::memcpy(h.magic, "SHIPFEM1", 8);
    h.N = hc.N;
    h.T = hc.T;
    h.Npad = hc.Npad;
    h.Tpad = hc.Tpad;
    h.node_waves = hc.node_waves;
    h.elem_waves = hc.elem_waves;
    h.features = hc.features;
    h.hist_cap = s->dc.rec.hist_cap;
    h.dt = s->dc.dt;
    h.payload = payload;
    int64_t step[2] = {0, 0};
    char* out = static_cast<char*>(buf) + sizeof(CheckpointHeader);
    hipError_t e = toHost(step, s->dc.rec.step, 16, s->stream);
    for (const Piece& p : pieces) {
        if (e == hipSuccess) e = toHost(out, p.dev, p.bytes, s->stream);
        out += p.bytes;
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) return statusOf(e);
    h.step = std::max(step[0], step[1]);
    std::memcpy(buf, &h, sizeof h);
    *len = need;
    return SHIPFEM_OK;
}

shipfem_status shipfem_solver_restore(shipfem_solver* s, const void* buf, uint64_t len) {
    if (!s || !buf || len < sizeof(CheckpointHeader)) return SHIPFEM_INVALID_ARG;
    CheckpointHeader h;
    std::memcpy(&h, buf, sizeof h);
    const HostCase& hc = s->hc;
    const std::vector<Piece> pieces = evolvingState(s);
    uint64_t payload = 0;
    for (const Piece& p : pieces) payload += p.bytes;
    const bool same = std::memcmp(h.magic, "SHIPFEM1", 8) == 0 && h.N == hc.N && h.T == hc.T &&
                      h.Npad == hc.Npad && h.Tpad == hc.Tpad && h.node_waves == hc.node_waves &&
                      h.elem_waves == hc.elem_waves && h.features == hc.features &&
                      h.hist_cap == s->dc.rec.hist_cap && h.dt == s->dc.dt && h.payload == payload &&
                      len >= sizeof h + payload && h.step >= 0;
    if (!same) return SHIPFEM_INVALID_ARG;
    const char* in = static_cast<const char*>(buf) + sizeof h;
    hipError_t e = hipSuccess;
    for (const Piece& p : pieces) {
        if (e == hipSuccess) e = toDevice(p.dev, in, p.bytes, s->stream);
        in += p.bytes;
    }
    if (e == hipSuccess) e = setStep(s, h.step);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    return statusOf(e);
}

uint64_t shipfem_apply_edge_drive(uint32_t node_count, const double* displacement,
                                  const uint8_t* driven, uint32_t* node_flags, double* prescribed) {
    if (!displacement || !driven || !node_flags || !prescribed) return 0;
    uint64_t written = 0;
    for (uint64_t d = 0; d < 3ull * node_count; ++d) {
        if (!driven[d]) continue;
        const uint32_t axis = static_cast<uint32_t>(d % 3);
        uint32_t& f = node_flags[d / 3];
        f = (f & ~(SHIPFEM_PIN_X << axis)) | (SHIPFEM_DRIVE_X << axis);
        prescribed[d] = displacement[d];
        ++written;
    }
    return written;
}

}  // extern "C"
