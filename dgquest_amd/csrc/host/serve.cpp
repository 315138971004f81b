// serve.cpp -- a request server for many registers on one GPU (dgq_serve.h).
//
// Threads:
//   planners  take arriving requests, plan them (planner.cpp, the same plans
//             the QuEST front end runs) and draw their sample uniforms;
//   lanes     one HIP stream each.  A lane admits planned requests in arrival
//             order when their device memory fits the arena (the HBM budget),
//             launches them, and retires them when their completion event
//             fires.  Batch lanes take registers of up to batch_max_qubits,
//             grouped by size so one launch per fused sweep serves the whole
//             group (dq_job lists of the engine); large lanes take one larger
//             register at a time and fall back to batches when none waits.
// Device memory comes from one arena reserved at start; a request that does
// not fit waits for a retirement to free memory.  Within a class admission
// is in arrival order (no overtaking); the batch class holds at most
// lanes x depth x kBatchStateCap bytes, so it cannot starve large requests.
//
// The reference has no serving loop (one register per process through
// src/QuEST.c); the per-register work is the same fused sweeps, the P(0)
// reduction of src/cpp/QuEST_gpu_kernel.cu:769-807 and MT19937 draws
// (src/mt19937ar.c).
#include "dgq_serve.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <set>
#include <stdexcept>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "mt19937.hpp"
#include "planner.hpp"
#include "tune.hpp"

namespace dgq {
namespace serve {

thread_local char g_err[512] = "";

int fail(int code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

inline uint64_t align_up(uint64_t x, uint64_t a) { return (x + a - 1) / a * a; }

// ---------------------------------------------------------------------------
// arena: best-fit allocation of offsets in [0, cap), free blocks coalesced
// ---------------------------------------------------------------------------
class Arena {
public:
    static constexpr uint64_t kAlign = 4096;
    explicit Arena(uint64_t cap = 0) { reset(cap); }
    void reset(uint64_t cap) {
        by_off_.clear();
        by_size_.clear();
        cap_ = cap / kAlign * kAlign;
        used_ = 0;
        if (cap_) put(0, cap_);
    }
    bool alloc(uint64_t bytes, uint64_t *off) {
        bytes = align_up(bytes ? bytes : 1, kAlign);
        auto it = by_size_.lower_bound({bytes, 0});
        if (it == by_size_.end()) return false;
        const uint64_t sz = it->first, o = it->second;
        by_size_.erase(it);
        by_off_.erase(o);
        if (sz > bytes) put(o + bytes, sz - bytes);
        used_ += bytes;
        *off = o;
        return true;
    }
    void release(uint64_t off, uint64_t bytes) {
        bytes = align_up(bytes ? bytes : 1, kAlign);
        used_ -= bytes;
        uint64_t o = off, sz = bytes;
        auto next = by_off_.lower_bound(o);
        if (next != by_off_.end() && next->first == o + sz) {
            sz += next->second;
            by_size_.erase({next->second, next->first});
            next = by_off_.erase(next);
        }
        if (next != by_off_.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == o) {
                o = prev->first;
                sz += prev->second;
                by_size_.erase({prev->second, prev->first});
                by_off_.erase(prev);
            }
        }
        put(o, sz);
    }
    uint64_t used() const { return used_; }
    uint64_t cap() const { return cap_; }
    size_t free_blocks() const { return by_off_.size(); }
    uint64_t largest_free() const { return by_size_.empty() ? 0 : by_size_.rbegin()->first; }

private:
    void put(uint64_t o, uint64_t sz) {
        by_off_[o] = sz;
        by_size_.insert({sz, o});
    }
    std::map<uint64_t, uint64_t> by_off_;
    std::set<std::pair<uint64_t, uint64_t>> by_size_;
    uint64_t cap_ = 0, used_ = 0;
};

// ---------------------------------------------------------------------------
// requests and launch groups
// ---------------------------------------------------------------------------
using Clock = std::chrono::steady_clock;

struct Req {
    uint64_t id = 0;
    int n = 0;
    uint32_t shots = 0, seed = 0;
    bool recheck = false;
    std::vector<Gate> gates;
    std::vector<int> l2p;
    int init_kind = DQ_INIT_CLASSICAL;
    uint64_t init_index = 0;
    double init_value = 0.0;
    Plan plan, plan64;
    std::vector<double> u;
    dgs_result res;
    std::vector<double> p0;         // logical qubit order, total last
    std::vector<uint64_t> samples;  // logical basis-state indices
    bool done = false;
};

// One launch group: a batch of registers of one size, or one register.
struct Item {
    std::vector<Req *> reqs;
    int n = 0, prec = DQ_F32;
    uint32_t shots = 0;
    bool recheck = false;
    uint64_t off = 0, bytes = 0;  // arena allocation
    std::vector<uint64_t> st_off;
    uint64_t st64_off = 0, ws_off = 0, ws_bytes = 0, ro_off = 0, ro_bytes = 0, res_off = 0, res_bytes = 0;
    uint64_t samp_off = 0, p064_off = 0, samp64_off = 0, cmp_off = 0, cmpws_off = 0, cmpws_bytes = 0;
    void *ev = nullptr;
    unsigned char *host = nullptr;  // pinned copy of the result area
    size_t host_cap = 0;
};

size_t elem_bytes(int prec) { return prec == DQ_F64 ? 16 : 8; }

uint64_t apply_ws(const std::vector<const Plan *> &plans) {
    uint32_t nb = 0, np = 0, no = 0;
    for (const Plan *p : plans) {
        nb += (uint32_t)p->blocks.size();
        np += (uint32_t)p->phases.size();
        no += (uint32_t)p->ops.size();
    }
    return dq_apply_ws_bytes(nb, np, no, nb);
}

// device layout of a launch group, relative to its arena offset
void layout(Item &it) {
    const uint32_t count = (uint32_t)it.reqs.size();
    uint64_t off = 0;
    it.st_off.resize(count);
    for (uint32_t i = 0; i < count; i++) {
        it.st_off[i] = off;
        off += align_up((uint64_t)elem_bytes(it.prec) << it.n, 256);
    }
    if (it.recheck) {
        it.st64_off = off;
        off += align_up((uint64_t)16 << it.n, 256);
    }
    std::vector<const Plan *> plans;
    for (Req *r : it.reqs) plans.push_back(&r->plan);
    it.ws_bytes = apply_ws(plans);
    if (it.recheck) it.ws_bytes = std::max<uint64_t>(it.ws_bytes, apply_ws({&it.reqs[0]->plan64}));
    it.ws_off = off;
    off += align_up(it.ws_bytes, 256);
    it.ro_bytes = dq_readout_ws_bytes(it.n, count, it.shots);
    it.ro_off = off;
    off += align_up(it.ro_bytes, 256);
    // result area: p0 | samples | (p0 fp64 | samples fp64 | comparison)
    it.res_off = off;
    uint64_t r = 0;
    r += align_up((uint64_t)count * (it.n + 1) * 8, 256);
    it.samp_off = r;
    r += align_up((uint64_t)count * it.shots * 8, 256);
    if (it.recheck) {
        it.p064_off = r;
        r += align_up((uint64_t)(it.n + 1) * 8, 256);
        it.samp64_off = r;
        r += align_up((uint64_t)it.shots * 8, 256);
        it.cmp_off = r;
        r += 256;
    }
    it.res_bytes = r;
    off += r;
    if (it.recheck) {
        it.cmpws_off = off;
        it.cmpws_bytes = dq_compare_ws_bytes(it.n);
        off += align_up(it.cmpws_bytes, 256);
    }
    it.bytes = off;
}

double ms_between(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

uint64_t splitmix64(uint64_t x) {
    x += 0x9e3779b97f4a7c15ull;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    return x ^ (x >> 31);
}

// Copy a request in (the caller's arrays may go away).
int make_req(const dgs_request &q, int prec, double recheck_fraction, Req &r) {
    if (q.n_qubits < 1 || q.n_qubits > 40) return fail(DQ_ERR_ARG, "request: n_qubits %d out of range", q.n_qubits);
    if (q.n_gates && !q.gates) return fail(DQ_ERR_ARG, "request: null gate list");
    if (q.init_kind != DQ_INIT_CLASSICAL && q.init_kind != DQ_INIT_PLUS)
        return fail(DQ_ERR_ARG, "request: init_kind must be DQ_INIT_CLASSICAL or DQ_INIT_PLUS");
    if (q.init_kind == DQ_INIT_CLASSICAL && (q.init_index >> q.n_qubits))
        return fail(DQ_ERR_ARG, "request: init_index out of range");
    if (q.shots > (1u << 20)) return fail(DQ_ERR_ARG, "request: at most 2^20 samples");
    const uint64_t full = (q.n_qubits == 64) ? ~0ull : ((1ull << q.n_qubits) - 1);
    r.n = q.n_qubits;
    r.shots = q.shots;
    r.seed = q.seed;
    r.init_kind = q.init_kind;
    r.init_index = q.init_index;
    r.init_value = q.init_value;
    r.gates.resize(q.n_gates);
    for (uint32_t i = 0; i < q.n_gates; i++) {
        const dgq_gate_rec &g = q.gates[i];
        if ((g.tmask | g.cmask) & ~full) return fail(DQ_ERR_ARG, "request: gate %u acts outside the register", i);
        if (const char *why = gate_record_error(g.type, g.tmask, g.cmask, g.op))
            return fail(DQ_ERR_ARG, "request: gate %u (type %d): %s", i, g.type, why);
        Gate &d = r.gates[i];
        d.type = g.type;
        d.tmask = g.tmask;
        d.cmask = g.cmask;
        memcpy(d.op, g.op, sizeof d.op);
    }
    r.l2p.resize(r.n);
    uint64_t seen = 0;
    for (int i = 0; i < r.n; i++) {
        r.l2p[i] = q.l2p ? q.l2p[i] : i;
        if (r.l2p[i] < 0 || r.l2p[i] >= r.n || ((seen >> r.l2p[i]) & 1))
            return fail(DQ_ERR_ARG, "request: l2p is not a permutation");
        seen |= 1ull << r.l2p[i];
    }
    if (prec == DQ_F64 || (q.flags & DGS_NO_RECHECK)) r.recheck = false;
    else if (q.flags & DGS_RECHECK) r.recheck = true;
    else r.recheck = (double)(splitmix64(((uint64_t)q.seed << 8) ^ (uint64_t)q.n_qubits) % 1000000ull) <
                     recheck_fraction * 1e6;
    memset(&r.res, 0, sizeof r.res);
    r.res.n_qubits = r.n;
    return DQ_OK;
}

// The fusion shape for an n-qubit register: the configured one when k is
// set, else the tuning table's entry for (prec, n).
TuneParams tune_of(const dgs_config &c, int prec, int n) {
    TuneParams t = tune_for(prec, n);
    if (c.k > 0) {
        t.k = c.k;
        if (c.R > 0) t.R = c.R;
        if (c.W >= 0) t.W = c.W;
    }
    if (c.max_ops > 0) t.max_ops = c.max_ops;
    return t;
}

// The request's circuit, then the qubit moves that undo its relabelling
// swaps: samples walk the CDF in logical index order, so the final state must
// be stored in logical order.
void plan_one(const Req &r, int prec, const TuneParams &tp, Plan &out) {
    std::vector<int> l2p = r.l2p;
    std::vector<Plan> passes;
    std::vector<Gate> extra;
    std::string err;
    if (!plan_restore_mapping(r.n, prec, l2p, tp, passes, extra, err)) throw std::runtime_error(err);
    if (extra.empty()) {
        plan_circuit(r.gates, r.n, prec, r.init_kind, r.init_index, r.init_value, tp, out);
    } else {
        std::vector<Gate> g = r.gates;
        g.insert(g.end(), extra.begin(), extra.end());
        plan_circuit(g, r.n, prec, r.init_kind, r.init_index, r.init_value, tp, out);
    }
    for (const Plan &p : passes) append_plan(out, p);
}

void plan_req(Req &r, int prec, const dgs_config &cfg) {
    plan_one(r, prec, tune_of(cfg, prec, r.n), r.plan);
    if (r.recheck) plan_one(r, DQ_F64, tune_of(cfg, DQ_F64, r.n), r.plan64);
    for (int q = 0; q < r.n; q++) r.l2p[q] = q;
    r.u.resize(r.shots);
    Mt19937 g;
    const uint32_t key = r.seed;
    g.init_by_array(&key, 1);
    for (uint32_t s = 0; s < r.shots; s++) r.u[s] = g.res53();
}

// logical index of a physical basis state: logical bit q = physical bit l2p[q]
uint64_t to_logical(uint64_t phys, const std::vector<int> &l2p) {
    uint64_t v = 0;
    for (size_t q = 0; q < l2p.size(); q++) v |= ((phys >> l2p[q]) & 1ull) << q;
    return v;
}

// Fill a request's results from the host copy of its group's result area.
void collect(const Item &it, size_t i, Req &r) {
    const int n = it.n;
    const double *p0 = (const double *)(it.host) + i * (n + 1);
    const uint64_t *sm = (const uint64_t *)(it.host + it.samp_off) + i * it.shots;
    bool ident = true;
    for (int q = 0; q < n; q++) ident &= r.l2p[q] == q;
    r.p0.resize(n + 1);
    // P(q = 0) of the normalised state: the FP32 norm drifts off 1 by up to
    // ~1e-6, which would otherwise bias every <Z_q> by as much
    const double tot = p0[n];
    const double inv = tot > 0.0 ? 1.0 / tot : 0.0;
    double sz = 0.0;
    for (int q = 0; q < n; q++) {
        r.p0[q] = p0[r.l2p[q]] * inv;
        sz += 2.0 * r.p0[q] - 1.0;
    }
    r.p0[n] = tot;
    r.res.sum_z = sz;
    r.res.total_prob = tot;
    r.samples.resize(it.shots);
    for (uint32_t s = 0; s < it.shots; s++) r.samples[s] = ident ? sm[s] : to_logical(sm[s], r.l2p);
    r.res.batch_size = (int)it.reqs.size();
    if (it.recheck) {
        const double *q64 = (const double *)(it.host + it.p064_off);
        const uint64_t *s64 = (const uint64_t *)(it.host + it.samp64_off);
        const double *cmp = (const double *)(it.host + it.cmp_off);
        const double inv64 = q64[n] > 0.0 ? 1.0 / q64[n] : 0.0;
        double sz64 = 0.0;
        for (int q = 0; q < n; q++) sz64 += 2.0 * q64[r.l2p[q]] * inv64 - 1.0;
        uint32_t mis = 0;
        for (uint32_t s = 0; s < it.shots; s++) mis += s64[s] != sm[s];
        r.res.rechecked = 1;
        // cmp = re <a|b>, im <a|b>, <a|a>, <b|b>, max |a_i - b_i> (a: FP64, b: served)
        const double ov = cmp[0] * cmp[0] + cmp[1] * cmp[1];
        const double nn = cmp[2] * cmp[3];
        r.res.fidelity_raw = ov;
        r.res.norm_fp64 = cmp[2];
        r.res.fidelity = nn > 0.0 ? ov / nn : 0.0;
        r.res.max_abs_err = cmp[4];
        r.res.sum_z_fp64 = sz64;
        r.res.shot_mismatches = mis;
    }
}

// Enqueue a launch group on `stream`: sweeps, read-out (and the FP64 recheck),
// then the copy of the result area to it.host and the completion event.
int issue(Item &it, unsigned char *base, void *stream) {
    const uint32_t count = (uint32_t)it.reqs.size();
    std::vector<dq_block> blocks;
    std::vector<dq_phase> phases;
    std::vector<dq_op> ops;
    std::vector<uint32_t> boff(count);
    uint32_t nl = 0;
    for (uint32_t i = 0; i < count; i++) {
        const Plan &p = it.reqs[i]->plan;
        boff[i] = (uint32_t)blocks.size();
        const uint32_t pb = (uint32_t)phases.size(), ob = (uint32_t)ops.size();
        for (dq_block b : p.blocks) {
            b.phase_begin += pb;
            blocks.push_back(b);
        }
        for (dq_phase ph : p.phases) {
            ph.op_begin += ob;
            phases.push_back(ph);
        }
        ops.insert(ops.end(), p.ops.begin(), p.ops.end());
        nl = std::max<uint32_t>(nl, (uint32_t)p.blocks.size());
    }
    std::vector<dq_job> jobs;
    std::vector<uint32_t> jb(1, 0);
    for (uint32_t l = 0; l < nl; l++) {
        for (uint32_t i = 0; i < count; i++) {
            if (l >= it.reqs[i]->plan.blocks.size()) continue;
            dq_job j;
            memset(&j, 0, sizeof j);
            j.state = base + it.st_off[i];
            j.block = boff[i] + l;
            jobs.push_back(j);
        }
        jb.push_back((uint32_t)jobs.size());
    }
    int rc = dq_apply_blocks(it.prec, blocks.data(), (uint32_t)blocks.size(), phases.data(), (uint32_t)phases.size(),
                             ops.data(), (uint32_t)ops.size(), jobs.data(), jb.data(), nl, base + it.ws_off,
                             it.ws_bytes, stream);
    if (rc) return rc;
    std::vector<const void *> states(count);
    std::vector<double> u((size_t)count * it.shots);
    for (uint32_t i = 0; i < count; i++) {
        states[i] = base + it.st_off[i];
        if (it.shots) memcpy(u.data() + (size_t)i * it.shots, it.reqs[i]->u.data(), it.shots * sizeof(double));
    }
    unsigned char *res = base + it.res_off;
    rc = dq_readout_many(it.prec, it.n, states.data(), count, u.data(), it.shots, (double *)res,
                         (uint64_t *)(res + it.samp_off), base + it.ro_off, it.ro_bytes, stream);
    if (rc) return rc;
    if (it.recheck) {
        const Plan &p = it.reqs[0]->plan64;
        std::vector<uint32_t> jb64(p.blocks.size() + 1);
        std::vector<dq_job> j64(p.blocks.size());
        for (size_t l = 0; l < p.blocks.size(); l++) {
            memset(&j64[l], 0, sizeof j64[l]);
            j64[l].state = base + it.st64_off;
            j64[l].block = (uint32_t)l;
            jb64[l] = (uint32_t)l;
        }
        jb64[p.blocks.size()] = (uint32_t)p.blocks.size();
        rc = dq_apply_blocks(DQ_F64, p.blocks.data(), (uint32_t)p.blocks.size(), p.phases.data(),
                             (uint32_t)p.phases.size(), p.ops.data(), (uint32_t)p.ops.size(), j64.data(), jb64.data(),
                             (uint32_t)p.blocks.size(), base + it.ws_off, it.ws_bytes, stream);
        if (rc) return rc;
        const void *s64 = base + it.st64_off;
        rc = dq_readout_many(DQ_F64, it.n, &s64, 1, it.reqs[0]->u.data(), it.shots, (double *)(res + it.p064_off),
                             (uint64_t *)(res + it.samp64_off), base + it.ro_off, it.ro_bytes, stream);
        if (rc) return rc;
        rc = dq_compare_states(s64, DQ_F64, base + it.st_off[0], it.prec, it.n, (double *)(res + it.cmp_off),
                               base + it.cmpws_off, it.cmpws_bytes, stream);
        if (rc) return rc;
    }
    rc = dq_memcpy_async(it.host, res, it.res_bytes, DQ_D2H, stream);
    if (rc) return rc;
    return dq_event_record(it.ev, stream);
}

// ---------------------------------------------------------------------------
// the server
// ---------------------------------------------------------------------------
constexpr uint64_t kBatchStateCap = 1ull << 30;  // state bytes per batch

struct Lane {
    int idx = 0;
    bool batch = false;
    void *stream = nullptr;
    std::deque<Item *> inflight;
    std::vector<Item *> spare;  // retired groups, reused with their event and pinned buffer
    std::thread th;
};

class Server {
public:
    explicit Server(const dgs_config &c) : cfg_(c), t0_(Clock::now()) {}
    ~Server() { shutdown(); }

    int start() {
        int rc = dq_device_init(cfg_.device);
        if (rc) return fail(rc, "dq_device_init: %s", dq_last_error());
        void *p = nullptr;
        rc = dq_malloc(&p, cfg_.hbm_budget);
        if (rc) return fail(rc, "arena of %llu bytes: %s", (unsigned long long)cfg_.hbm_budget, dq_last_error());
        base_ = (unsigned char *)p;
        arena_.reset(cfg_.hbm_budget);
        const int nlanes = cfg_.batch_lanes + cfg_.large_lanes;
        lanes_.resize(nlanes);
        for (int i = 0; i < nlanes; i++) {
            Lane &L = lanes_[i];
            L.idx = i;
            L.batch = i < cfg_.batch_lanes;
            rc = dq_stream_create(&L.stream);
            if (rc) return fail(rc, "stream: %s", dq_last_error());
        }
        for (int i = 0; i < cfg_.planner_threads; i++) planners_.emplace_back([this] { planner_main(); });
        for (auto &L : lanes_) L.th = std::thread([this, &L] { lane_main(L); });
        started_ = true;
        return DQ_OK;
    }

    void shutdown() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            if (stopping_) return;
            stopping_ = true;
        }
        cv_plan_.notify_all();
        cv_work_.notify_all();
        for (auto &t : planners_) t.join();
        for (auto &L : lanes_)
            if (L.th.joinable()) L.th.join();
        for (auto &L : lanes_) {
            for (Item *it : L.spare) free_item(it);
            if (L.stream) dq_stream_destroy(L.stream);
        }
        if (base_) dq_free(base_);
        base_ = nullptr;
        for (auto &kv : reqs_) delete kv.second;
        reqs_.clear();
    }

    double now_ms() const { return ms_between(t0_, Clock::now()); }
    Clock::time_point t0() const { return t0_; }

    int submit(const dgs_request &q, uint64_t *id, double t_arrive_ms) {
        std::unique_ptr<Req> r(new Req());
        int rc = make_req(q, cfg_.prec, cfg_.recheck_fraction, *r);
        if (rc) return rc;
        std::lock_guard<std::mutex> lk(mu_);
        if (stopping_) return fail(DQ_ERR_ARG, "server is shutting down");
        r->id = ++next_id_;
        r->res.id = r->id;
        r->res.t_arrive_ms = t_arrive_ms < 0 ? now_ms() : t_arrive_ms;
        if (id) *id = r->id;
        Req *p = r.release();
        reqs_[p->id] = p;
        incoming_.push_back(p);
        unfinished_++;
        cv_plan_.notify_one();
        return DQ_OK;
    }

    int wait(uint64_t id, dgs_result *res, double *p0, uint64_t *samples) {
        std::unique_lock<std::mutex> lk(mu_);
        auto it = reqs_.find(id);
        if (it == reqs_.end()) return fail(DQ_ERR_ARG, "dgs_wait: unknown request %llu", (unsigned long long)id);
        Req *r = it->second;
        cv_done_.wait(lk, [&] { return r->done; });
        if (res) *res = r->res;
        if (p0 && !r->p0.empty()) memcpy(p0, r->p0.data(), r->p0.size() * sizeof(double));
        if (samples && !r->samples.empty()) memcpy(samples, r->samples.data(), r->samples.size() * sizeof(uint64_t));
        reqs_.erase(it);
        delete r;
        return DQ_OK;
    }

    int poll(uint64_t id) {
        std::lock_guard<std::mutex> lk(mu_);
        auto it = reqs_.find(id);
        if (it == reqs_.end()) return fail(DQ_ERR_ARG, "dgs_poll: unknown request %llu", (unsigned long long)id);
        return it->second->done ? 0 : 1;
    }

    void stats(double out[8]) {
        std::lock_guard<std::mutex> lk(mu_);
        out[0] = (double)st_done_;
        out[1] = (double)st_batches_;
        out[2] = (double)st_batched_regs_;
        out[3] = (double)st_peak_;
        out[4] = (double)arena_.cap();
        out[5] = (double)st_mem_waits_;
        out[6] = (double)st_rechecks_;
        out[7] = st_plan_ms_;
    }

private:
    bool batchable(const Req *r) const { return r->n <= cfg_.batch_max_qubits && !r->recheck; }

    // Finish a request (lock held).
    void finish(Req *r, int status, const char *msg) {
        r->res.status = status;
        r->res.t_done_ms = now_ms();
        if (status && msg) fprintf(stderr, "dgq_serve: request %llu failed: %s\n", (unsigned long long)r->id, msg);
        r->done = true;
        r->plan = Plan();
        r->plan64 = Plan();
        r->gates.clear();
        r->gates.shrink_to_fit();
        st_done_++;
        if (r->res.rechecked) st_rechecks_++;
        unfinished_--;
        cv_done_.notify_all();
    }

    void planner_main() {
        for (;;) {
            Req *r;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_plan_.wait(lk, [&] { return stopping_ || !incoming_.empty(); });
                if (incoming_.empty()) return;
                r = incoming_.front();
                incoming_.pop_front();
                planning_++;
            }
            const auto a = Clock::now();
            int status = DQ_OK;
            std::string msg;
            try {
                plan_req(*r, cfg_.prec, cfg_);
            } catch (const std::exception &e) {
                status = DQ_ERR_ARG;
                msg = e.what();
            }
            uint64_t need = 0;
            if (!status) {
                Item probe;
                probe.reqs.push_back(r);
                probe.n = r->n;
                probe.prec = cfg_.prec;
                probe.shots = r->shots;
                probe.recheck = r->recheck;
                layout(probe);
                need = probe.bytes;
            }
            const double pms = ms_between(a, Clock::now());
            {
                std::lock_guard<std::mutex> lk(mu_);
                planning_--;
                st_plan_ms_ += pms;
                if (status) {
                    finish(r, status, msg.c_str());
                } else if (align_up(need, Arena::kAlign) > arena_.cap()) {
                    char b[160];
                    snprintf(b, sizeof b, "needs %llu bytes, more than the whole budget of %llu",
                             (unsigned long long)need, (unsigned long long)arena_.cap());
                    finish(r, DQ_ERR_NOMEM, b);
                } else {
                    planned_[r->id] = r;
                }
            }
            cv_work_.notify_all();
        }
    }

    Item *new_item(Lane &L) {
        Item *it;
        if (!L.spare.empty()) {
            it = L.spare.back();
            L.spare.pop_back();
        } else {
            it = new Item();
            if (dq_event_create(&it->ev)) {
                delete it;
                return nullptr;
            }
        }
        it->reqs.clear();
        it->recheck = false;
        return it;
    }

    void free_item(Item *it) {
        if (it->ev) dq_event_destroy(it->ev);
        if (it->host) dq_host_free(it->host);
        delete it;
    }

    // Try to allocate the group's memory (lock held).
    bool place(Item &it) {
        layout(it);
        uint64_t off;
        if (!arena_.alloc(it.bytes, &off)) return false;
        it.off = off;
        st_peak_ = std::max<uint64_t>(st_peak_, arena_.used());
        return true;
    }

    // Admission (lock held): the next launch group for this lane, or nullptr.
    // Requests that cannot be batched (too large, or rechecked in FP64) run
    // one per group on large lanes; with no large lanes the batch lanes take
    // them too, so every request is admitted by some lane.
    Item *take(Lane &L) {
        if (!L.batch || cfg_.large_lanes == 0) {
            for (auto &kv : planned_) {
                Req *r = kv.second;
                if (batchable(r)) continue;
                Item *it = new_item(L);
                if (!it) return nullptr;
                it->reqs.push_back(r);
                it->n = r->n;
                it->prec = cfg_.prec;
                it->shots = r->shots;
                it->recheck = r->recheck;
                if (!place(*it)) {
                    L.spare.push_back(it);
                    st_mem_waits_++;
                    break;  // no overtaking inside the class; a batch may still go
                }
                const uint64_t id = r->id;
                planned_.erase(id);
                return it;
            }
        }
        Req *head = nullptr;
        for (auto &kv : planned_)
            if (batchable(kv.second)) { head = kv.second; break; }
        if (!head) return nullptr;
        Item *it = new_item(L);
        if (!it) return nullptr;
        it->n = head->n;
        it->prec = cfg_.prec;
        it->shots = head->shots;
        const uint64_t sb = (uint64_t)elem_bytes(cfg_.prec) << head->n;
        for (auto &kv : planned_) {
            Req *r = kv.second;
            if (!batchable(r) || r->n != head->n || r->shots != head->shots) continue;
            if ((int)it->reqs.size() >= cfg_.batch_max) break;
            if (!it->reqs.empty() && (it->reqs.size() + 1) * sb > kBatchStateCap) break;
            it->reqs.push_back(r);
        }
        while (!place(*it)) {
            if (it->reqs.size() == 1) {
                L.spare.push_back(it);
                st_mem_waits_++;
                return nullptr;
            }
            it->reqs.resize(it->reqs.size() / 2);
        }
        for (Req *r : it->reqs) planned_.erase(r->id);
        st_batches_++;
        st_batched_regs_ += it->reqs.size();
        return it;
    }

    // Launch a group (no lock held); false when the engine refused it.
    bool launch(Lane &L, Item *it) {
        if (it->host_cap < it->res_bytes) {
            if (it->host) dq_host_free(it->host);
            it->host = nullptr;
            it->host_cap = 0;
            size_t cap = 64 * 1024;
            while (cap < it->res_bytes) cap *= 2;
            void *h = nullptr;
            if (dq_host_alloc(&h, cap)) return fail_item(L, it, DQ_ERR_NOMEM, dq_last_error());
            it->host = (unsigned char *)h;
            it->host_cap = cap;
        }
        const double t = now_ms();
        for (Req *r : it->reqs) r->res.t_start_ms = t;
        int rc = issue(*it, base_ + it->off, L.stream);
        if (rc) {
            dq_stream_sync(L.stream);  // let whatever was enqueued drain before the memory is reused
            return fail_item(L, it, rc, dq_last_error());
        }
        return true;
    }

    bool fail_item(Lane &L, Item *it, int rc, const char *msg) {
        std::string m(msg ? msg : "");
        std::lock_guard<std::mutex> lk(mu_);
        arena_.release(it->off, it->bytes);
        for (Req *r : it->reqs) finish(r, rc, m.c_str());
        it->reqs.clear();
        L.spare.push_back(it);
        cv_work_.notify_all();
        return false;
    }

    void retire(Lane &L, Item *it) {
        for (size_t i = 0; i < it->reqs.size(); i++) collect(*it, i, *it->reqs[i]);
        {
            std::lock_guard<std::mutex> lk(mu_);
            arena_.release(it->off, it->bytes);
            for (Req *r : it->reqs) finish(r, DQ_OK, nullptr);
            it->reqs.clear();
            L.spare.push_back(it);
        }
        cv_work_.notify_all();
    }

    void lane_main(Lane &L) {
        dq_device_init(cfg_.device);
        const int depth = cfg_.lane_depth > 0 ? cfg_.lane_depth : 1;
        for (;;) {
            while (!L.inflight.empty()) {
                const int q = dq_event_query(L.inflight.front()->ev);
                if (q == 1) break;
                Item *it = L.inflight.front();
                L.inflight.pop_front();
                if (q < 0) {
                    fail_item(L, it, q, dq_last_error());
                    continue;
                }
                retire(L, it);
            }
            if ((int)L.inflight.size() < depth) {
                Item *it;
                {
                    std::lock_guard<std::mutex> lk(mu_);
                    it = take(L);
                }
                if (it) {
                    if (launch(L, it)) L.inflight.push_back(it);
                    continue;
                }
            }
            std::unique_lock<std::mutex> lk(mu_);
            if (!L.inflight.empty()) {
                // poll the oldest launch; new work or freed memory wakes us earlier
                cv_work_.wait_for(lk, std::chrono::microseconds(100));
                continue;
            }
            if (stopping_ && unfinished_ == 0) return;
            cv_work_.wait_for(lk, std::chrono::milliseconds(2));
        }
    }

    dgs_config cfg_;
    Clock::time_point t0_;
    std::mutex mu_;
    std::condition_variable cv_plan_, cv_work_, cv_done_;
    std::deque<Req *> incoming_;
    std::map<uint64_t, Req *> planned_;
    std::unordered_map<uint64_t, Req *> reqs_;
    uint64_t next_id_ = 0;
    int planning_ = 0;
    long unfinished_ = 0;
    bool stopping_ = false, started_ = false;
    Arena arena_;
    unsigned char *base_ = nullptr;
    std::vector<Lane> lanes_;
    std::vector<std::thread> planners_;
    // counters
    uint64_t st_done_ = 0, st_batches_ = 0, st_batched_regs_ = 0, st_peak_ = 0, st_mem_waits_ = 0, st_rechecks_ = 0;
    double st_plan_ms_ = 0.0;
};

int check_config(const dgs_config &c) {
    if (c.prec != DQ_F32 && c.prec != DQ_F64) return fail(DQ_ERR_ARG, "config: prec must be DQ_F32 or DQ_F64");
    if (c.hbm_budget < (1u << 20)) return fail(DQ_ERR_ARG, "config: hbm_budget below 1 MiB");
    if (c.planner_threads < 1 || c.planner_threads > 64) return fail(DQ_ERR_ARG, "config: planner_threads 1..64");
    if (c.batch_lanes < 0 || c.large_lanes < 0 || c.batch_lanes + c.large_lanes < 1 || c.batch_lanes + c.large_lanes > 16)
        return fail(DQ_ERR_ARG, "config: 1..16 lanes");
    if (c.lane_depth < 1 || c.lane_depth > 8) return fail(DQ_ERR_ARG, "config: lane_depth 1..8");
    if (c.batch_max < 1 || c.batch_max > 4096) return fail(DQ_ERR_ARG, "config: batch_max 1..4096");
    if (!(c.recheck_fraction >= 0.0 && c.recheck_fraction <= 1.0)) return fail(DQ_ERR_ARG, "config: recheck_fraction 0..1");
    return DQ_OK;
}

void sleep_until(Clock::time_point t) {
    const auto now = Clock::now();
    if (t > now) std::this_thread::sleep_until(t);
}

}  // namespace serve
}  // namespace dgq

using namespace dgq;
using namespace dgq::serve;

struct dgs_server {
    std::unique_ptr<Server> s;
};

extern "C" {

void dgs_config_default(dgs_config *c) {
    if (!c) return;
    memset(c, 0, sizeof *c);
    c->device = 0;
    c->prec = DQ_F32;
    c->hbm_budget = 200000000000ull;
    c->planner_threads = 4;
    c->batch_lanes = 1;
    c->large_lanes = 3;
    c->lane_depth = 2;
    c->batch_max_qubits = 22;
    c->batch_max = 256;
    c->recheck_fraction = 0.01;
    c->k = c->R = c->max_ops = 0;
    c->W = -1;
}

const char *dgs_last_error(void) { return serve::g_err; }

dgs_server *dgs_create(const dgs_config *cfg) {
    if (!cfg) {
        fail(DQ_ERR_ARG, "dgs_create: null config");
        return nullptr;
    }
    if (check_config(*cfg)) return nullptr;
    dgs_server *h = new dgs_server();
    h->s.reset(new Server(*cfg));
    if (h->s->start()) {
        const std::string m = serve::g_err;
        delete h;
        snprintf(serve::g_err, sizeof serve::g_err, "%s", m.c_str());
        return nullptr;
    }
    return h;
}

void dgs_destroy(dgs_server *s) { delete s; }

int dgs_submit(dgs_server *s, const dgs_request *req, uint64_t *id) {
    if (!s || !req) return fail(DQ_ERR_ARG, "dgs_submit: null argument");
    return s->s->submit(*req, id, -1.0);
}

int dgs_wait(dgs_server *s, uint64_t id, dgs_result *res, double *p0, uint64_t *samples) {
    if (!s) return fail(DQ_ERR_ARG, "dgs_wait: null server");
    return s->s->wait(id, res, p0, samples);
}

int dgs_poll(dgs_server *s, uint64_t id) {
    if (!s) return fail(DQ_ERR_ARG, "dgs_poll: null server");
    return s->s->poll(id);
}

int dgs_stats(dgs_server *s, double out[8]) {
    if (!s || !out) return fail(DQ_ERR_ARG, "dgs_stats: null argument");
    s->s->stats(out);
    return DQ_OK;
}

int dgs_replay(dgs_server *s, const dgs_request *reqs, uint32_t count, const double *arrive_s, dgs_result *res,
               uint64_t *samples, double *p0) {
    if (!s || (count && (!reqs || !res))) return fail(DQ_ERR_ARG, "dgs_replay: null argument");
    for (uint32_t i = 1; arrive_s && i < count; i++)
        if (arrive_s[i] < arrive_s[i - 1]) return fail(DQ_ERR_ARG, "dgs_replay: arrivals must not decrease");
    Server &S = *s->s;
    std::vector<uint64_t> ids(count, 0);
    std::vector<int> rcs(count, 0);
    const auto start = Clock::now();
    const double start_ms = S.now_ms();
    for (uint32_t i = 0; i < count; i++) {
        const double a = arrive_s ? arrive_s[i] : 0.0;
        sleep_until(start + std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double>(a)));
        rcs[i] = S.submit(reqs[i], &ids[i], start_ms + 1e3 * a);
    }
    size_t off = 0, poff = 0;
    int first = DQ_OK;
    for (uint32_t i = 0; i < count; i++) {
        memset(&res[i], 0, sizeof res[i]);
        if (rcs[i]) {
            res[i].status = rcs[i];
            if (!first) first = rcs[i];
        } else {
            S.wait(ids[i], &res[i], p0 ? p0 + poff : nullptr, samples ? samples + off : nullptr);
        }
        off += reqs[i].shots;
        poff += (size_t)reqs[i].n_qubits + 1;
    }
    return first;
}

int dgs_run_plain(const dgs_config *cfg, const dgs_request *reqs, uint32_t count, const double *arrive_s,
                  dgs_result *res, uint64_t *samples, double *p0) {
    if (!cfg || (count && (!reqs || !res))) return fail(DQ_ERR_ARG, "dgs_run_plain: null argument");
    if (cfg->prec != DQ_F32 && cfg->prec != DQ_F64) return fail(DQ_ERR_ARG, "dgs_run_plain: bad precision");
    int rc = dq_device_init(cfg->device);
    if (rc) return fail(rc, "dq_device_init: %s", dq_last_error());
    void *stream = nullptr;
    rc = dq_stream_create(&stream);
    if (rc) return fail(rc, "stream: %s", dq_last_error());
    void *host = nullptr;
    size_t host_cap = 0;
    // Device memory for the largest request, allocated once before the clock
    // starts and reused by every request: a one-at-a-time loop that keeps its
    // register memory.  (Allocating and freeing per request, as createQureg /
    // destroyQureg would, timed the HIP allocator too: single requests then
    // stalled for up to seconds and the loop's rate varied 2x between runs.)
    size_t dev_cap = 0;
    for (uint32_t i = 0; i < count; i++) {
        Req r;
        dgs_request q = reqs[i];
        q.flags = DGS_NO_RECHECK;
        if (make_req(q, cfg->prec, 0.0, r)) continue;  // reported in the timed loop
        try {
            plan_req(r, cfg->prec, *cfg);
        } catch (const std::exception &) {
            continue;
        }
        Item it;
        it.reqs.push_back(&r);
        it.n = r.n;
        it.prec = cfg->prec;
        it.shots = r.shots;
        layout(it);
        dev_cap = std::max(dev_cap, it.bytes);
    }
    void *dev = nullptr;
    if (dev_cap) {
        rc = dq_malloc(&dev, dev_cap);
        if (rc) {
            dq_stream_destroy(stream);
            return fail(rc, "dgs_run_plain: device memory: %s", dq_last_error());
        }
    }
    const auto start = Clock::now();
    size_t soff = 0, poff = 0;
    int first = DQ_OK;
    for (uint32_t i = 0; i < count; i++) {
        const double a = arrive_s ? arrive_s[i] : 0.0;
        const auto arrive = start + std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double>(a));
        sleep_until(arrive);
        dgs_result &R = res[i];
        memset(&R, 0, sizeof R);
        const double t_start = ms_between(start, Clock::now());
        Req r;
        dgs_request q = reqs[i];
        q.flags = DGS_NO_RECHECK;
        rc = make_req(q, cfg->prec, 0.0, r);
        if (!rc) {
            try {
                plan_req(r, cfg->prec, *cfg);
            } catch (const std::exception &e) {
                rc = fail(DQ_ERR_ARG, "plan: %s", e.what());
            }
        }
        Item it;
        if (!rc) {
            it.reqs.push_back(&r);
            it.n = r.n;
            it.prec = cfg->prec;
            it.shots = r.shots;
            layout(it);
            if (host_cap < it.res_bytes) {
                if (host) dq_host_free(host);
                host = nullptr;
                host_cap = 0;
                rc = dq_host_alloc(&host, it.res_bytes);
                if (!rc) host_cap = it.res_bytes;
            }
        }
        if (!rc && it.bytes > dev_cap) rc = fail(DQ_ERR_ARG, "dgs_run_plain: request %u larger than planned", i);
        if (!rc) rc = dq_event_create(&it.ev);
        if (!rc) {
            it.host = (unsigned char *)host;
            rc = issue(it, (unsigned char *)dev, stream);
            if (!rc) rc = dq_event_sync(it.ev);
        }
        if (!rc) {
            collect(it, 0, r);
            R = r.res;
            if (samples && r.shots) memcpy(samples + soff, r.samples.data(), r.shots * sizeof(uint64_t));
            if (p0) memcpy(p0 + poff, r.p0.data(), r.p0.size() * sizeof(double));
        }
        if (it.ev) dq_event_destroy(it.ev);
        it.ev = nullptr;
        it.host = nullptr;
        dq_stream_sync(stream);
        R.id = i + 1;
        R.status = rc;
        R.n_qubits = reqs[i].n_qubits;
        R.batch_size = 1;
        R.t_arrive_ms = 1e3 * a;
        R.t_start_ms = t_start;
        R.t_done_ms = ms_between(start, Clock::now());
        if (rc && !first) first = rc;
        soff += reqs[i].shots;
        poff += (size_t)reqs[i].n_qubits + 1;
        rc = DQ_OK;
    }
    if (host) dq_host_free(host);
    if (dev) dq_free(dev);
    dq_stream_destroy(stream);
    return first;
}

int dgs_uniforms(uint32_t seed, uint32_t count, double *out) {
    if (count && !out) return fail(DQ_ERR_ARG, "dgs_uniforms: null output");
    Mt19937 g;
    g.init_by_array(&seed, 1);
    for (uint32_t s = 0; s < count; s++) out[s] = g.res53();
    return DQ_OK;
}

int dgs_arena_selftest(uint32_t seed, int iters) {
    const uint64_t pages = 256, P = Arena::kAlign;
    Arena a(pages * P);
    std::vector<int> owner(pages, -1);
    struct Blk { uint64_t off, bytes; };
    std::vector<Blk> live;
    Mt19937 g;
    g.init_by_array(&seed, 1);
    auto longest_run = [&]() {
        uint64_t best = 0, cur = 0;
        for (uint64_t p = 0; p < pages; p++) {
            cur = owner[p] < 0 ? cur + 1 : 0;
            best = std::max(best, cur);
        }
        return best;
    };
    auto runs = [&]() {
        size_t r = 0;
        for (uint64_t p = 0; p < pages; p++)
            if (owner[p] < 0 && (p == 0 || owner[p - 1] >= 0)) r++;
        return r;
    };
    for (int it = 0; it < iters; it++) {
        if (live.empty() || g.next_u32() % 3 != 0) {
            const uint64_t bytes = 1 + (uint64_t)(g.next_u32() % (24 * P));
            const uint64_t np = (bytes + P - 1) / P;
            uint64_t off = 0;
            const bool ok = a.alloc(bytes, &off);
            const bool fits = longest_run() >= np;
            if (ok != fits) return 1;
            if (ok) {
                if (off % P || off / P + np > pages) return 2;
                for (uint64_t p = off / P; p < off / P + np; p++) {
                    if (owner[p] >= 0) return 3;
                    owner[p] = it;
                }
                live.push_back({off, bytes});
            }
        } else {
            const size_t k = g.next_u32() % live.size();
            const Blk b = live[k];
            live.erase(live.begin() + (long)k);
            a.release(b.off, b.bytes);
            for (uint64_t p = b.off / P; p < b.off / P + (b.bytes + P - 1) / P; p++) owner[p] = -1;
        }
        uint64_t used = 0;
        for (uint64_t p = 0; p < pages; p++) used += owner[p] >= 0 ? P : 0;
        if (used != a.used()) return 4;
        if (a.free_blocks() != runs()) return 5;
        if (a.largest_free() != longest_run() * P) return 6;
    }
    return 0;
}

}  // extern "C"
