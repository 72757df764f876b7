// server.cpp -- resident serving loop with continuous batching.
//
// Requests (damage-case solves of uneven size) arrive at any time. Host preparation
// runs on a small worker pool; one submission thread owns the device. The device runs
// rounds: `chunk` steps of every resident case in one element launch and one node
// launch per step over a shared tile table, replayed from a captured graph. Two rounds
// are kept in flight on the run stream: while the device runs one, the submission
// thread retires the cases the previous round finished (read back on a side stream),
// admits prepared cases in arrival order while their arena fits the device pool
// (first-fit, coalescing) and captures the next round. A case's step count after each
// round is known when the round is launched, so a finishing case is left out of the
// next round before its last round has even run. A case's arithmetic and its per-wave
// reductions never depend on which other cases share the launch, so its result is
// bit-identical alone and in the mix.
#include "shipfem_host.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <thread>
#include <unordered_map>

using namespace shipfem;

namespace {

double nowSeconds() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// First-fit allocator over one device buffer; free ranges coalesce on release.
class DevicePool {
public:
    hipError_t init(uint64_t bytes) {
        size_ = bytes;
        hipError_t e = hipMalloc(reinterpret_cast<void**>(&base_), bytes);
        if (e == hipSuccess) free_[0] = bytes;
        return e;
    }
    ~DevicePool() {
        if (base_) (void)hipFree(base_);
    }
    // offset or UINT64_MAX
    uint64_t alloc(uint64_t bytes) {
        bytes = (bytes + 4095) / 4096 * 4096;
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            const uint64_t off = it->first, len = it->second;
            free_.erase(it);
            if (len > bytes) free_[off + bytes] = len - bytes;
            used_ += bytes;
            return off;
        }
        return UINT64_MAX;
    }
    void release(uint64_t off, uint64_t bytes) {
        bytes = (bytes + 4095) / 4096 * 4096;
        used_ -= bytes;
        auto it = free_.emplace(off, bytes).first;
        auto next = std::next(it);
        if (next != free_.end() && it->first + it->second == next->first) {
            it->second += next->second;
            free_.erase(next);
        }
        if (it != free_.begin()) {
            auto prev = std::prev(it);
            if (prev->first + prev->second == it->first) {
                prev->second += it->second;
                free_.erase(it);
            }
        }
    }
    char* base() const { return base_; }
    uint64_t size() const { return size_; }
    uint64_t used() const { return used_; }

private:
    char* base_ = nullptr;
    uint64_t size_ = 0, used_ = 0;
    std::map<uint64_t, uint64_t> free_;
};

struct Request {
    uint64_t id = 0;
    double run_to = 0, submit_s = 0, admit_s = 0;
    // owned copies of the caller's arrays
    std::vector<double> X, load, presc, u0, v0, mass;
    std::vector<uint32_t> index, flags;
    shipfem_case_desc desc{};
    bool hex = false;  // a solid-shell case: 8 node ids per element
    HostCase hc;
    std::string error;
};

// Rounds run one element kind at a time: tets, elastic hexes, plastic hexes.
enum { kKindTet = 0, kKindHex = 1, kKindHexPlastic = 2, kKinds = 3 };
int kindOf(const HostCase& hc) {
    return hc.corners != 8 ? kKindTet : (hc.features & kFeatPlastic) ? kKindHexPlastic : kKindHex;
}

struct Active {
    std::unique_ptr<Request> req;
    DeviceCase dc;
    uint64_t off = 0;
    int64_t steps = 0;
    int64_t last = 0;        // step reached at the end of the last completed round
    int64_t predicted = 0;   // step it reaches when every launched round has run
    uint64_t last_round = 0; // the last launched round it belongs to (0: none)
    int32_t left_out = 0;    // consecutive rounds of its kind it sat out (SHORTEST)
    bool dead = false;       // reported non-finite: joins no further round
};

// One launched round: its plan, members, timing events and progress buffers.
struct Round {
    uint64_t id = 0;
    int kind = 0, plan = 0;
    std::vector<Active*> members;
    hipEvent_t start = nullptr, stop = nullptr, done = nullptr;
    int64_t* d_prog = nullptr;  // [2 * members] device
    int64_t* h_prog = nullptr;  // [2 * members] pinned host
    size_t cap = 0;
};

// A retired case's state kept for shipfem_server_fetch, every array in the caller's node
// and element orders (as readbackCase writes them).
struct Kept {
    std::vector<double> u, v;
    std::vector<float> position, eqps, plastic_strain, damage;
    int64_t step = 0;
    double time = 0;
    uint32_t status = 0;
};

}  // namespace

struct shipfem_server {
    shipfem_server_opts opts{};
    DevicePool pool;
    hipStream_t run = nullptr;   // rounds
    hipStream_t side = nullptr;  // uploads, stable steps, capture, readback

    std::mutex m;
    std::condition_variable cv_raw, cv_ready, cv_done;
    std::deque<std::unique_ptr<Request>> raw, ready;
    std::deque<shipfem_result> results;
    std::unordered_map<uint64_t, bool> keep;
    std::unordered_map<uint64_t, Kept> kept;
    std::vector<std::thread> workers;
    std::thread loop;
    bool stop = false;
    uint64_t next_id = 1;
    shipfem_server_stats st{};

    // owned by the loop thread
    std::vector<std::unique_ptr<Active>> active;
    // per element kind, two plans used in turn, so a round's graph is never rebuilt while
    // the kind's previous round is in flight
    Plan plans[kKinds][2];
    std::vector<uint64_t> plan_sig[kKinds][2];  // request ids each graph was captured for
    int next_plan[kKinds] = {0, 0, 0};
    int next_kind = 0;
    Plan omega_plan;                    // stable steps of tet newcomers
    Round slots[2];
    std::deque<int> inflight;           // slot indices, oldest first
    uint64_t rounds_launched = 0;

    void workerMain();
    void loopMain();
    // SHORTEST: keeps the members a round takes, least remaining work first up to the
    // round's work, starved cases first; counts the rounds the others sit out.
    void pickShortest(std::vector<Active*>& members);
    bool admit();
    void finish(Active& a, hipError_t fault);
    hipError_t launchRound(const std::vector<Active*>& members, int kind);
    hipError_t completeOldest();
    void failAll(hipError_t e);
    // Device objects: waits for both streams, then frees (safe on partial creation).
    void release();
};

void shipfem_server::release() {
    if (run) (void)hipStreamSynchronize(run);
    if (side) (void)hipStreamSynchronize(side);
    for (auto& k : plans)
        for (Plan& p : k) p.dropGraph();
    active.clear();
    for (Round& r : slots) {
        if (r.d_prog) (void)hipFree(r.d_prog);
        if (r.h_prog) (void)hipHostFree(r.h_prog);
        if (r.start) (void)hipEventDestroy(r.start);
        if (r.stop) (void)hipEventDestroy(r.stop);
        if (r.done) (void)hipEventDestroy(r.done);
        r = Round{};
    }
    if (run) (void)hipStreamDestroy(run);
    if (side) (void)hipStreamDestroy(side);
    run = side = nullptr;
}

void shipfem_server::workerMain() {
    for (;;) {
        std::unique_ptr<Request> r;
        {
            std::unique_lock<std::mutex> lk(m);
            cv_raw.wait(lk, [&] { return stop || !raw.empty(); });
            if (stop) return;
            r = std::move(raw.front());
            raw.pop_front();
        }
        r->error = r->hex ? prepareHexCase(r->desc, r->hc, 512) : prepareCase(r->desc, r->hc, 512);
        // the raw copies are no longer needed; the prepared case holds what it uses
        std::vector<double>().swap(r->load);
        std::vector<double>().swap(r->presc);
        std::vector<double>().swap(r->u0);
        std::vector<double>().swap(r->v0);
        std::vector<double>().swap(r->mass);
        std::vector<uint32_t>().swap(r->flags);
        {
            std::lock_guard<std::mutex> lk(m);
            if (!r->error.empty()) {
                shipfem_result res{};
                res.request_id = r->id;
                res.status = SHIPFEM_INVALID_ARG;
                res.submit_s = r->submit_s;
                res.done_s = nowSeconds();
                results.push_back(res);
                ++st.failed;
                cv_done.notify_all();
            } else {
                ready.push_back(std::move(r));
                cv_ready.notify_all();
            }
        }
    }
}

// Moves prepared requests into free slots while their arenas fit. Arrival order:
// the head of the queue waits for room rather than being overtaken. Runs while rounds
// are in flight: it only touches free pool memory and the side stream.
bool shipfem_server::admit() {
    std::vector<std::unique_ptr<Request>> take;
    std::vector<uint64_t> offs;
    {
        std::lock_guard<std::mutex> lk(m);
        while (!ready.empty() && active.size() + take.size() < opts.max_active) {
            const uint64_t bytes = arenaBytes(ready.front()->hc);
            if (bytes > pool.size()) {
                // can never fit: refuse it rather than block the queue behind it
                shipfem_result res{};
                res.request_id = ready.front()->id;
                res.status = SHIPFEM_BUSY;
                res.submit_s = ready.front()->submit_s;
                res.done_s = nowSeconds();
                results.push_back(res);
                ++st.refused;
                ++st.failed;
                ready.pop_front();
                cv_done.notify_all();
                continue;
            }
            const uint64_t off = pool.alloc(bytes);
            if (off == UINT64_MAX) break;
            take.push_back(std::move(ready.front()));
            ready.pop_front();
            offs.push_back(off);
        }
    }
    if (take.empty()) return false;
    // upload, then the FP64 stable step of the tet newcomers in one pass (a hex case's
    // comes from its host set-up)
    std::vector<std::unique_ptr<Active>> add(take.size());
    omega_plan.recs.clear();
    hipError_t e = hipSuccess;
    for (size_t k = 0; k < take.size(); ++k) {
        add[k] = std::make_unique<Active>();
        add[k]->off = offs[k];
        if (e == hipSuccess) e = uploadCase(take[k]->hc, pool.base() + offs[k], add[k]->dc, side);
        if (e == hipSuccess) e = hipStreamSynchronize(side);
        if (!take[k]->hex) omega_plan.recs.push_back(add[k]->dc.rec);
    }
    std::vector<double> lmax;
    if (e == hipSuccess && !omega_plan.recs.empty()) e = omega_plan.build(side);
    if (e == hipSuccess && !omega_plan.recs.empty()) e = omega_plan.omegaMax(lmax, side);
    size_t tet_index = 0;
    const double now = nowSeconds();
    size_t admitted = 0;
    for (size_t k = 0; k < take.size(); ++k) {
        Active& a = *add[k];
        HostCase& hc = take[k]->hc;
        if (e != hipSuccess) {
            std::lock_guard<std::mutex> lk(m);
            pool.release(a.off, arenaBytes(hc));
            shipfem_result res{};
            res.request_id = take[k]->id;
            res.status = statusOf(e);
            res.submit_s = take[k]->submit_s;
            res.done_s = now;
            results.push_back(res);
            ++st.failed;
            continue;
        }
        a.dc.cfl = (take[k]->hex ? 2.0 / hc.hex_omega : 2.0 / std::sqrt(4.0 * lmax[tet_index++] / hc.rho)) *
                   hc.cfl_scale;
        setTimestep(a.dc, hc, hc.dt_request > 0 ? hc.dt_request : hc.dt_safety * a.dc.cfl);
        a.steps = std::max<int64_t>(1, static_cast<int64_t>(std::ceil(take[k]->run_to / a.dc.dt - 1e-9)));
        a.dc.rec.step_end = a.steps;
        take[k]->admit_s = now;
        a.req = std::move(take[k]);
        active.push_back(std::move(add[k]));
        ++admitted;
    }
    {
        std::lock_guard<std::mutex> lk(m);
        st.admitted += admitted;
        st.active = active.size();
    }
    cv_done.notify_all();
    if (e != hipSuccess) (void)hipGetLastError();
    return admitted > 0;
}

// Reads a retired case back (side stream), reports it and returns its pool memory.
// The case belongs to no round still in flight.
void shipfem_server::finish(Active& a, hipError_t fault) {
    HostCase& hc = a.req->hc;
    shipfem_result res{};
    res.request_id = a.req->id;
    res.status = fault != hipSuccess ? statusOf(fault) : (a.dead ? SHIPFEM_UNSTABLE : SHIPFEM_OK);
    res.submit_s = a.req->submit_s;
    res.admit_s = a.req->admit_s;
    res.node_count = hc.N;
    res.tet_count = hc.T;
    Kept k;
    shipfem_state_view view{};
    std::vector<shipfem_energy_record> rec;
    int32_t nrec = 0;
    bool keeping = false;
    {
        std::lock_guard<std::mutex> lk(m);
        keeping = keep.count(res.request_id) != 0;
    }
    if (fault == hipSuccess) {
        k.u.resize(3ull * hc.N);
        k.v.resize(3ull * hc.N);
        k.eqps.resize(hc.T);
        k.damage.resize(hc.T);
        view.displacement = k.u.data();
        view.velocity = k.v.data();
        view.eqps = k.eqps.data();
        view.damage = k.damage.data();
        if (keeping) {
            // one readback gives every kept field in the caller's orders
            k.position.resize(3ull * hc.N);
            k.plastic_strain.resize(6ull * hc.T);
            view.position = k.position.data();
            view.plastic_strain = k.plastic_strain.data();
        }
        hipError_t e = readbackCase(a.dc, hc, &view, side);
        rec.resize(static_cast<size_t>(a.dc.rec.hist_cap));
        if (e == hipSuccess) e = readEnergy(a.dc, rec.data(), a.dc.rec.hist_cap, &nrec, side);
        if (e != hipSuccess) res.status = statusOf(e);
        res.steps = view.step;
        res.sim_time = view.step * a.dc.dt;
        for (uint32_t n = 0; n < hc.N; ++n) {
            const double* u = &k.u[3ull * n];
            res.max_displacement = std::max(res.max_displacement, std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]));
        }
        for (float q : k.eqps) res.max_eqps = std::max(res.max_eqps, static_cast<double>(q));
        if ((hc.features & kFeatDamage) && e == hipSuccess)
            for (uint32_t t = 0; t < hc.T; ++t)  // damage is in the caller's order, vol internal
                if (k.damage[externalTet(hc, t)] >= 1.0f) {
                    ++res.torn_elements;
                    res.torn_volume += hc.corners == 8 ? hc.hvol[t] : hc.vol[t];
                }
        for (int32_t r = 0; r < nrec; ++r) {
            const double* f = rec[r].contact_force;
            res.peak_contact_force = std::max(res.peak_contact_force, std::sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]));
        }
        if (nrec > 0) res.last = rec[nrec - 1];
    }
    res.done_s = nowSeconds();
    std::lock_guard<std::mutex> lk(m);
    pool.release(a.off, arenaBytes(hc));
    auto it = keep.find(res.request_id);
    if (it != keep.end()) {
        if (fault == hipSuccess) {
            k.step = view.step;
            k.time = view.time;
            k.status = view.status;
            kept[res.request_id] = std::move(k);
        }
        keep.erase(it);
    }
    results.push_back(res);
    if (res.status == SHIPFEM_OK) ++st.completed;
    else ++st.failed;
}

// Captures (or reuses) the graph for `members` (all of one element kind) on the kind's plan
// not used by its round in flight, then queues the round, its progress read-out and its
// events on the run stream.
hipError_t shipfem_server::launchRound(const std::vector<Active*>& members, int kind) {
    const int R = opts.chunk;
    const int pi = next_plan[kind];
    Plan& p = plans[kind][pi];
    std::vector<uint64_t> sig;
    sig.reserve(members.size());
    for (const Active* a : members) sig.push_back(a->req->id);
    hipError_t e = hipSuccess;
    if (sig != plan_sig[kind][pi] || !p.graph) {
        plan_sig[kind][pi].clear();
        p.recs.clear();
        for (const Active* a : members) p.recs.push_back(a->dc.rec);
        p.parity = 0;
        e = p.build(side);
        if (e == hipSuccess) e = p.capture(R, side);
        if (e != hipSuccess) return e;
        plan_sig[kind][pi] = std::move(sig);
    }
    int slot = 0;
    while (std::find(inflight.begin(), inflight.end(), slot) != inflight.end()) ++slot;
    Round& r = slots[slot];
    const size_t n = members.size();
    if (n > r.cap) {
        if (r.d_prog) (void)hipFree(r.d_prog);
        if (r.h_prog) (void)hipHostFree(r.h_prog);
        r.d_prog = nullptr;
        r.h_prog = nullptr;
        r.cap = 0;
        const size_t cap = std::max<size_t>(2 * n, 64);
        e = hipMalloc(reinterpret_cast<void**>(&r.d_prog), 16 * cap);
        if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&r.h_prog), 16 * cap);
        if (e != hipSuccess) return e;
        r.cap = cap;
    }
    r.id = ++rounds_launched;
    r.kind = kind;
    r.plan = pi;
    r.members = members;
    e = hipEventRecord(r.start, run);
    if (e == hipSuccess) e = hipGraphLaunch(p.graph, run);
    if (e == hipSuccess) e = hipEventRecord(r.stop, run);
    if (e == hipSuccess) {
        launch_progress(p.d_recs, static_cast<int>(n), r.d_prog, run);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(r.h_prog, r.d_prog, 16 * n, hipMemcpyDeviceToHost, run);
    if (e == hipSuccess) e = hipEventRecord(r.done, run);
    if (e != hipSuccess) return e;
    for (Active* a : members) {
        a->predicted = std::min(a->predicted + R, a->steps);
        a->last_round = r.id;
    }
    inflight.push_back(slot);
    next_plan[kind] ^= 1;
    return hipSuccess;
}

// Waits for the oldest round in flight, accounts it and retires what it finished.
hipError_t shipfem_server::completeOldest() {
    Round& r = slots[inflight.front()];
    hipError_t e = hipEventSynchronize(r.done);
    float ms = 0;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, r.start, r.stop);
    if (e != hipSuccess) return e;
    inflight.pop_front();
    // element-steps actually advanced (finished cases idle in a round)
    double done_steps = 0;
    for (size_t i = 0; i < r.members.size(); ++i) {
        Active* a = r.members[i];
        done_steps += static_cast<double>(a->req->hc.T) * static_cast<double>(r.h_prog[2 * i] - a->last);
        a->last = r.h_prog[2 * i];
        if (r.h_prog[2 * i + 1] != 0) a->dead = true;
    }
    // retire cases whose last round this was and that will join no other
    for (size_t i = active.size(); i-- > 0;) {
        Active& a = *active[i];
        const bool over = a.dead || a.predicted >= a.steps;
        if (over && a.last_round <= r.id) {
            finish(a, hipSuccess);
            active.erase(active.begin() + static_cast<std::ptrdiff_t>(i));
        }
    }
    std::lock_guard<std::mutex> lk(m);
    st.rounds += 1;
    st.device_ms += ms;
    st.element_steps += done_steps;
    st.active = active.size();
    cv_done.notify_all();
    return hipSuccess;
}

// A device error fails every resident case; the server keeps serving.
void shipfem_server::failAll(hipError_t e) {
    (void)hipStreamSynchronize(run);
    (void)hipStreamSynchronize(side);
    (void)hipGetLastError();
    inflight.clear();
    for (auto& a : active) finish(*a, e);
    active.clear();
    for (auto& k : plan_sig)
        for (auto& sg : k) sg.clear();
    std::lock_guard<std::mutex> lk(m);
    st.active = 0;
    cv_done.notify_all();
}

// SHORTEST's defaults: about 3 M tets at a 64-step chunk fill the device for a round (a
// smaller round leaves CUs idle), and a case waits at most 4 rounds, so the largest keeps
// at least a fifth of the rate it would have under ALL.
constexpr double kShortestRoundWork = 2e8;
constexpr int32_t kShortestStarveRounds = 4;

void shipfem_server::pickShortest(std::vector<Active*>& members) {
    if (members.empty()) return;
    const double budget = opts.round_work > 0 ? opts.round_work : kShortestRoundWork;
    const int32_t starve = opts.starve_rounds > 0 ? opts.starve_rounds : kShortestStarveRounds;
    auto left = [](const Active* a) { return static_cast<double>(a->req->hc.T) * static_cast<double>(a->steps - a->predicted); };
    std::stable_sort(members.begin(), members.end(), [&](const Active* x, const Active* y) {
        const bool sx = x->left_out >= starve, sy = y->left_out >= starve;
        if (sx != sy) return sx;
        if (left(x) != left(y)) return left(x) < left(y);
        return x->req->id < y->req->id;
    });
    std::vector<Active*> take;
    double work = 0;
    for (Active* a : members) {
        const double w = static_cast<double>(a->req->hc.T) *
                         static_cast<double>(std::min<int64_t>(opts.chunk, a->steps - a->predicted));
        if (take.empty() || a->left_out >= starve || work + w <= budget) {
            take.push_back(a);
            work += w;
            a->left_out = 0;
        } else {
            ++a->left_out;
        }
    }
    members.swap(take);
}

void shipfem_server::loopMain() {
    for (;;) {
        bool stopping = false;
        {
            std::unique_lock<std::mutex> lk(m);
            if (active.empty() && inflight.empty())
                cv_ready.wait(lk, [&] { return stop || !ready.empty(); });
            stopping = stop;
        }
        if (stopping) {
            // launch nothing more; let the rounds in flight end before the streams go
            while (!inflight.empty())
                if (completeOldest() != hipSuccess) break;
            return;
        }
        admit();
        bool launched = false;
        if (inflight.size() < 2) {
            // the next element kind in turn that has cases to advance
            std::vector<Active*> members;
            int kind = next_kind;
            for (int t = 0; t < kKinds && members.empty(); ++t) {
                kind = (next_kind + t) % kKinds;
                for (auto& a : active)
                    if (!a->dead && a->predicted < a->steps && kindOf(a->req->hc) == kind)
                        members.push_back(a.get());
            }
            if (opts.policy == SHIPFEM_ROUNDS_SHORTEST) pickShortest(members);
            if (!members.empty()) {
                next_kind = (kind + 1) % kKinds;
                const hipError_t e = launchRound(members, kind);
                if (e != hipSuccess) {
                    failAll(e);
                    continue;
                }
                launched = true;
            }
        }
        if (!inflight.empty() && (inflight.size() == 2 || !launched)) {
            const hipError_t e = completeOldest();
            if (e != hipSuccess) failAll(e);
        } else if (!launched) {
            // resident cases wait for nothing and nothing fits yet: brief back-off
            std::this_thread::sleep_for(std::chrono::milliseconds(1));
        }
    }
}

extern "C" {

shipfem_status shipfem_server_create(const shipfem_server_opts* opts, shipfem_server** out,
                                     char* err, size_t err_len) {
    if (!out) return SHIPFEM_INVALID_ARG;
    *out = nullptr;
    std::string why;
    if (!deviceAvailable(&why)) {
        setError(err, err_len, why);
        return SHIPFEM_NO_DEVICE;
    }
    auto* s = new (std::nothrow) shipfem_server();
    if (!s) return SHIPFEM_OUT_OF_MEMORY;
    if (opts) s->opts = *opts;
    if (s->opts.max_active == 0) s->opts.max_active = 256;
    if (s->opts.chunk == 0) s->opts.chunk = 64;
    if (s->opts.policy != SHIPFEM_ROUNDS_ALL && s->opts.policy != SHIPFEM_ROUNDS_SHORTEST) {
        setError(err, err_len, "policy must be SHIPFEM_ROUNDS_ALL or SHIPFEM_ROUNDS_SHORTEST");
        delete s;
        return SHIPFEM_INVALID_ARG;
    }
    if (s->opts.chunk < 2 || (s->opts.chunk & 1)) {
        setError(err, err_len, "chunk must be an even number >= 2");
        delete s;
        return SHIPFEM_INVALID_ARG;
    }
    if (!shapeFrom(&s->opts.launch, s->plans[0][0].shape)) {
        setError(err, err_len, "launch config: blocks must be 64/128/256/512, variant TILED or SLOTS");
        delete s;
        return SHIPFEM_INVALID_ARG;
    }
    if (s->opts.memory_budget == 0) {
        size_t freeb = 0, total = 0;
        (void)hipMemGetInfo(&freeb, &total);
        s->opts.memory_budget = static_cast<uint64_t>(0.6 * static_cast<double>(total));
    }
    hipError_t e = hipStreamCreateWithFlags(&s->run, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&s->side, hipStreamNonBlocking);
    for (Round& r : s->slots) {
        if (e == hipSuccess) e = hipEventCreate(&r.start);
        if (e == hipSuccess) e = hipEventCreate(&r.stop);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&r.done, hipEventDisableTiming);
    }
    if (e == hipSuccess) e = s->pool.init(s->opts.memory_budget);
    if (e != hipSuccess) {
        setError(err, err_len, std::string("HIP: ") + hipGetErrorString(e));
        const shipfem_status st = statusOf(e);
        s->release();
        delete s;
        return st;
    }
    // the hex kinds take the hex plans' defaults unless the caller chose a launch
    const bool chosen = s->opts.launch.elem_block || s->opts.launch.node_block;
    for (int k = 0; k < kKinds; ++k)
        for (int i = 0; i < 2; ++i) {
            s->plans[k][i].shape = s->plans[0][0].shape;
            if (k != kKindTet && !chosen) hexShape(s->plans[k][i].shape, k == kKindHexPlastic);
        }
    s->omega_plan.shape = s->plans[0][0].shape;
    s->st.bytes_budget = s->opts.memory_budget;
    const unsigned nworkers = std::max(1u, std::min(8u, std::thread::hardware_concurrency() / 2));
    for (unsigned w = 0; w < nworkers; ++w) s->workers.emplace_back([s] { s->workerMain(); });
    s->loop = std::thread([s] { s->loopMain(); });
    *out = s;
    return SHIPFEM_OK;
}

namespace {

shipfem_status submitCase(shipfem_server* s, const shipfem_case_desc* d, double run_to,
                          uint64_t* request_id, bool hex) {
    if (!s || !d || !(run_to > 0) || !d->rest_position || !d->index || d->node_count == 0 ||
        d->tet_count == 0)
        return SHIPFEM_INVALID_ARG;
    // a tet case is refused at once when its footprint bound exceeds the budget; a hex
    // case is measured after its set-up (a result with SHIPFEM_BUSY when it can never fit)
    if (!hex) {
        const uint64_t bound = shipfem_case_footprint_mesh(d, 512);
        if (bound > s->opts.memory_budget) {
            std::lock_guard<std::mutex> lk(s->m);
            ++s->st.refused;
            return SHIPFEM_BUSY;
        }
    }
    auto r = std::make_unique<Request>();
    const uint64_t N = d->node_count, T = d->tet_count;
    r->hex = hex;
    r->X.assign(d->rest_position, d->rest_position + 3 * N);
    r->index.assign(d->index, d->index + (hex ? 8 : 4) * T);
    if (d->node_flags) r->flags.assign(d->node_flags, d->node_flags + N);
    if (d->load) r->load.assign(d->load, d->load + 3 * N);
    if (d->prescribed) r->presc.assign(d->prescribed, d->prescribed + 3 * N);
    if (d->initial_displacement) r->u0.assign(d->initial_displacement, d->initial_displacement + 3 * N);
    if (d->initial_velocity) r->v0.assign(d->initial_velocity, d->initial_velocity + 3 * N);
    if (d->mass) r->mass.assign(d->mass, d->mass + N);
    r->desc = *d;
    r->desc.rest_position = r->X.data();
    r->desc.index = r->index.data();
    r->desc.node_flags = r->flags.empty() ? nullptr : r->flags.data();
    r->desc.load = r->load.empty() ? nullptr : r->load.data();
    r->desc.prescribed = r->presc.empty() ? nullptr : r->presc.data();
    r->desc.initial_displacement = r->u0.empty() ? nullptr : r->u0.data();
    r->desc.initial_velocity = r->v0.empty() ? nullptr : r->v0.data();
    r->desc.mass = r->mass.empty() ? nullptr : r->mass.data();
    // every round samples its first step: records land on multiples of the chunk
    r->desc.sample_every = s->opts.chunk;
    r->run_to = run_to;
    r->submit_s = nowSeconds();
    std::lock_guard<std::mutex> lk(s->m);
    r->id = s->next_id++;
    if (request_id) *request_id = r->id;
    s->raw.push_back(std::move(r));
    ++s->st.submitted;
    s->cv_raw.notify_one();
    return SHIPFEM_OK;
}

}  // namespace

shipfem_status shipfem_server_submit(shipfem_server* s, const shipfem_case_desc* d, double run_to,
                                     uint64_t* request_id) {
    return submitCase(s, d, run_to, request_id, false);
}

shipfem_status shipfem_server_submit_hex(shipfem_server* s, const shipfem_case_desc* d,
                                         double run_to, uint64_t* request_id) {
    return submitCase(s, d, run_to, request_id, true);
}

shipfem_status shipfem_server_poll(shipfem_server* s, shipfem_result* out, uint32_t max,
                                   uint32_t* count) {
    if (!s || (!out && max)) return SHIPFEM_INVALID_ARG;
    std::lock_guard<std::mutex> lk(s->m);
    uint32_t n = 0;
    while (n < max && !s->results.empty()) {
        out[n++] = s->results.front();
        s->results.pop_front();
    }
    if (count) *count = n;
    return SHIPFEM_OK;
}

shipfem_status shipfem_server_submit_encoded(shipfem_server* s, const void* buf, uint64_t len,
                                             double run_to, uint64_t* request_id) {
    if (!s) return SHIPFEM_INVALID_ARG;
    shipfem_case_blob* blob = nullptr;
    shipfem_case_desc d{};
    const shipfem_status st = shipfem_case_decode(buf, len, &blob, &d);
    if (st != SHIPFEM_OK) return st;
    // submit copies the case, so the decoded arrays can go at once
    const shipfem_status r = shipfem_case_blob_corners(blob) == 8
                                 ? shipfem_server_submit_hex(s, &d, run_to, request_id)
                                 : shipfem_server_submit(s, &d, run_to, request_id);
    shipfem_case_blob_destroy(blob);
    return r;
}

shipfem_status shipfem_server_keep(shipfem_server* s, uint64_t id) {
    if (!s) return SHIPFEM_INVALID_ARG;
    std::lock_guard<std::mutex> lk(s->m);
    s->keep[id] = true;
    return SHIPFEM_OK;
}

shipfem_status shipfem_server_fetch(shipfem_server* s, uint64_t id, shipfem_state_view* view) {
    if (!s || !view) return SHIPFEM_INVALID_ARG;
    std::lock_guard<std::mutex> lk(s->m);
    auto it = s->kept.find(id);
    if (it == s->kept.end()) return SHIPFEM_NOT_FOUND;
    const Kept& k = it->second;
    if (view->displacement) std::copy(k.u.begin(), k.u.end(), view->displacement);
    if (view->velocity) std::copy(k.v.begin(), k.v.end(), view->velocity);
    if (view->position) std::copy(k.position.begin(), k.position.end(), view->position);
    if (view->eqps) std::copy(k.eqps.begin(), k.eqps.end(), view->eqps);
    if (view->plastic_strain) std::copy(k.plastic_strain.begin(), k.plastic_strain.end(), view->plastic_strain);
    if (view->damage) std::copy(k.damage.begin(), k.damage.end(), view->damage);
    view->step = k.step;
    view->time = k.time;
    view->status = k.status;
    s->kept.erase(it);
    return SHIPFEM_OK;
}

shipfem_status shipfem_server_drain(shipfem_server* s, double timeout_s) {
    if (!s) return SHIPFEM_INVALID_ARG;
    std::unique_lock<std::mutex> lk(s->m);
    const bool ok = s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), [&] {
        return s->st.completed + s->st.failed >= s->st.submitted;
    });
    return ok ? SHIPFEM_OK : SHIPFEM_BUSY;
}

shipfem_status shipfem_server_stats_get(shipfem_server* s, shipfem_server_stats* out) {
    if (!s || !out) return SHIPFEM_INVALID_ARG;
    std::lock_guard<std::mutex> lk(s->m);
    *out = s->st;
    out->queued = s->raw.size() + s->ready.size();
    out->bytes_in_use = s->pool.used();
    return SHIPFEM_OK;
}

double shipfem_now(void) { return nowSeconds(); }

void shipfem_server_destroy(shipfem_server* s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->m);
        s->stop = true;
    }
    s->cv_raw.notify_all();
    s->cv_ready.notify_all();
    for (auto& w : s->workers) w.join();
    if (s->loop.joinable()) s->loop.join();
    s->release();
    delete s;
}

}  // extern "C"
