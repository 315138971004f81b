// paging.cpp -- see paging.hpp.
#include "paging.hpp"

#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <string>

#include "dag.hpp"
#include "planner.hpp"

namespace dgq {

namespace {

constexpr int kMaxRemap = 2;       // high <-> piece qubit pairs per section
constexpr int kGroupMax = 1 << kMaxRemap;
constexpr int kBuffers = 2 * kGroupMax;

void check(int rc, const char *what) {
    if (rc != DQ_OK) engine_fatal(what, rc);
}

uint64_t spread(uint64_t v, uint64_t mask) {  // deposit the low bits of v into the set bits of mask
    uint64_t r = 0;
    for (int i = 0; mask; i++) {
        const int p = ctz64(mask);
        mask &= mask - 1;
        r |= ((v >> i) & 1ull) << p;
    }
    return r;
}

uint64_t map_mask(uint64_t m, const std::vector<int> &pos) {
    uint64_t r = 0;
    while (m) {
        const int p = ctz64(m);
        m &= m - 1;
        r |= 1ull << pos[p];
    }
    return r;
}

// Timing events of one unit of work (a group of pieces).
struct UnitEv {
    void *in0 = nullptr, *in1 = nullptr, *c0 = nullptr, *c1 = nullptr, *o0 = nullptr, *o1 = nullptr;
};

struct EvPool {
    std::vector<void *> evs;
    size_t next = 0;
    void *get() {
        if (next == evs.size()) {
            void *e = nullptr;
            check(dq_event_create(&e), "paging: event");
            evs.push_back(e);
        }
        return evs[next++];
    }
    ~EvPool() {
        for (void *e : evs) dq_event_destroy(e);
    }
};

double interval_overlap(double a0, double a1, const std::vector<std::pair<double, double>> &u) {
    double s = 0;
    for (const auto &iv : u) {
        const double lo = std::max(a0, iv.first), hi = std::min(a1, iv.second);
        if (hi > lo) s += hi - lo;
    }
    return s;
}

// One section: the pieces stream through the device, each new piece gathered
// from the pieces of its group when `remap` moves qubits, `plan` run on it
// (if any), and written back.
void run_section(QuregState *s, const std::vector<Gate> &gates, const std::vector<std::pair<int, int>> &remap,
                 bool init, int ik, uint64_t iidx, double ival) {
    PagedStore &P = *s->paged;
    const int nl = P.nl;
    const uint64_t lmask = (1ull << nl) - 1;
    Plan PL;
    const bool compute = !gates.empty() || init;
    if (compute) {
        plan_circuit(gates, P.n, P.prec, init ? ik : DQ_INIT_LOAD, init ? (iidx & lmask) : 0, init ? ival : 0.0,
                     s->tune, PL, nl);
    }
    const bool have_blocks = !PL.blocks.empty();
    const int b = (int)remap.size();
    if (!have_blocks && b == 0 && !init) return;
    const int G = 1 << b;
    // pair i moves piece-index bit hpos[i] <-> piece qubit lpos[i]; a group's
    // pieces are numbered by their bits at hpos (bit i of the number = hpos[i])
    uint64_t hbits = 0, lbits = 0;
    int hpos[kMaxRemap] = {0}, lpos[kMaxRemap] = {0};
    for (int i = 0; i < b; i++) {
        hpos[i] = remap[i].first - nl;
        lpos[i] = remap[i].second;
        hbits |= 1ull << hpos[i];
        lbits |= 1ull << lpos[i];
    }
    auto at_h = [&](uint64_t v) {
        uint64_t r = 0;
        for (int i = 0; i < b; i++) r |= ((v >> i) & 1ull) << hpos[i];
        return r;
    };
    auto at_l = [&](uint64_t v) {
        uint64_t r = 0;
        for (int i = 0; i < b; i++) r |= ((v >> i) & 1ull) << lpos[i];
        return r;
    };
    const uint64_t npieces = 1ull << P.K;

    // units of work: lists of (new) pieces that share device buffers
    struct Unit {
        std::vector<uint64_t> pieces;
        bool load;
    };
    std::vector<Unit> units;
    std::vector<uint8_t> nzero = P.zero;
    if (init && ik == DQ_INIT_CLASSICAL) {
        for (uint64_t c = 0; c < npieces; c++) nzero[c] = 1;
        const uint64_t ci = iidx >> nl;
        units.push_back({{ci}, false});
        P.st.pieces_skipped += (long long)npieces - 1;  // every other piece is zero
    } else if (init) {  // every piece generates its amplitudes
        for (uint64_t c = 0; c < npieces; c += kGroupMax) {
            Unit u{{}, false};
            for (uint64_t k = c; k < std::min<uint64_t>(npieces, c + kGroupMax); k++) u.pieces.push_back(k);
            units.push_back(u);
        }
    } else if (b == 0) {  // no qubit moves: the non-zero pieces, kGroupMax per unit
        Unit u{{}, true};
        for (uint64_t c = 0; c < npieces; c++) {
            if (P.zero[c]) {
                P.st.pieces_skipped++;
                continue;
            }
            u.pieces.push_back(c);
            if ((int)u.pieces.size() == kGroupMax) {
                units.push_back(u);
                u.pieces.clear();
            }
        }
        if (!u.pieces.empty()) units.push_back(u);
    } else {
        for (uint64_t g = 0; g < npieces; g++) {
            if (g & hbits) continue;
            bool all_zero = true;
            for (int k = 0; k < G; k++) all_zero &= P.zero[g | at_h((uint64_t)k)] != 0;
            Unit u{{}, true};
            for (int j = 0; j < G; j++) u.pieces.push_back(g | at_h((uint64_t)j));
            if (all_zero) {
                P.st.pieces_skipped += G;
                continue;  // stays zero
            }
            units.push_back(u);
        }
    }

    // runs of one gathered piece: free bits above the lowest moved piece bit
    int lmin = nl;
    for (const auto &pr : remap) lmin = std::min(lmin, pr.second);
    uint64_t free_hi = 0;
    for (int q = lmin + 1; q < nl; q++)
        if (!((lbits >> q) & 1)) free_hi |= 1ull << q;
    const int nfree = popc64(free_hi);
    const size_t run_bytes = (b ? (size_t)1 << lmin : (size_t)1 << nl) * P.elem;

    EvPool pool;
    void *t_start = pool.get();
    check(dq_event_record(t_start, P.s_in), "paging: event record");
    std::vector<UnitEv> uev(units.size());
    void *last_out[2] = {nullptr, nullptr};
    for (size_t ui = 0; ui < units.size(); ui++) {
        const Unit &u = units[ui];
        const int set = (int)(ui & 1);
        UnitEv &E = uev[ui];
        E.in0 = pool.get(); E.in1 = pool.get(); E.c0 = pool.get(); E.c1 = pool.get(); E.o0 = pool.get(); E.o1 = pool.get();
        if (last_out[set]) check(dq_stream_wait_event(P.s_in, last_out[set]), "paging: wait");
        check(dq_event_record(E.in0, P.s_in), "paging: event record");
        if (u.load) {
            const uint64_t g = u.pieces[0] & ~hbits;  // group base (when qubits move)
            for (size_t j = 0; j < u.pieces.size(); j++) {
                unsigned char *dst = (unsigned char *)P.dbuf[set * kGroupMax + j];
                if (b == 0) {
                    check(dq_memcpy_async(dst, P.host[u.pieces[j]], P.piece_bytes, DQ_H2D, P.s_in), "paging: copy in");
                    P.st.bytes_in += P.piece_bytes;
                    continue;
                }
                // new piece j, element x: from old piece k = x's bits at lpos, at the
                // old element with those bits replaced by j's (other bits equal)
                for (int k = 0; k < G; k++) {
                    const uint64_t o = g | at_h((uint64_t)k);
                    const uint64_t dfix = at_l((uint64_t)k), sfix = at_l((uint64_t)j);
                    for (uint64_t f = 0; f < (1ull << nfree); f++) {
                        const uint64_t fb = spread(f, free_hi);
                        unsigned char *d = dst + (size_t)(fb | dfix) * P.elem;
                        if (P.zero[o]) {
                            int lg = 0;
                            while (((size_t)1 << lg) * P.elem < run_bytes) lg++;
                            check(dq_init_state(d, lg, P.prec, DQ_INIT_PLUS, 0, 0.0, P.s_in), "paging: zero fill");
                        } else {
                            check(dq_memcpy_async(d, P.host[o] + (size_t)(fb | sfix) * P.elem, run_bytes, DQ_H2D, P.s_in),
                                  "paging: copy in");
                            P.st.bytes_in += run_bytes;
                        }
                    }
                }
            }
        }
        check(dq_event_record(E.in1, P.s_in), "paging: event record");
        check(dq_stream_wait_event(P.s_comp, E.in1), "paging: wait");
        check(dq_event_record(E.c0, P.s_comp), "paging: event record");
        if (have_blocks) {
            const uint32_t nb = (uint32_t)PL.blocks.size(), nj = (uint32_t)u.pieces.size();
            std::vector<dq_job> jobs((size_t)nb * nj);
            std::vector<uint32_t> begin(nb + 1);
            for (uint32_t i = 0; i < nb; i++) {
                begin[i] = i * nj;
                for (uint32_t j = 0; j < nj; j++) {
                    dq_job &jb = jobs[(size_t)i * nj + j];
                    jb.state = P.dbuf[set * kGroupMax + j];
                    jb.block = i;
                    jb.pad = 0;
                    jb.base_hi = u.pieces[j] << nl;
                }
            }
            begin[nb] = nb * nj;
            const size_t need = dq_apply_ws_bytes(nb, (uint32_t)PL.phases.size(), (uint32_t)PL.ops.size(), nb * nj);
            if (need > P.ws_cap) {
                check(dq_device_sync(), "paging: sync");
                if (P.d_ws) dq_free(P.d_ws);
                P.ws_cap = std::max(need, (size_t)1 << 20);
                check(dq_malloc(&P.d_ws, P.ws_cap), "paging: workspace");
            }
            check(dq_apply_blocks(P.prec, PL.blocks.data(), nb, PL.phases.data(), (uint32_t)PL.phases.size(),
                                  PL.ops.empty() ? nullptr : PL.ops.data(), (uint32_t)PL.ops.size(), jobs.data(),
                                  begin.data(), nb, P.d_ws, P.ws_cap, P.s_comp),
                  "paging: sweep");
        }
        check(dq_event_record(E.c1, P.s_comp), "paging: event record");
        check(dq_stream_wait_event(P.s_out, E.c1), "paging: wait");
        check(dq_event_record(E.o0, P.s_out), "paging: event record");
        for (size_t j = 0; j < u.pieces.size(); j++) {
            check(dq_memcpy_async(P.host[u.pieces[j]], P.dbuf[set * kGroupMax + j], P.piece_bytes, DQ_D2H, P.s_out),
                  "paging: copy out");
            P.st.bytes_out += P.piece_bytes;
            nzero[u.pieces[j]] = 0;
        }
        check(dq_event_record(E.o1, P.s_out), "paging: event record");
        last_out[set] = E.o1;
        P.st.pieces_run += (long long)u.pieces.size();
    }
    check(dq_stream_sync(P.s_out), "paging: sync");
    check(dq_stream_sync(P.s_comp), "paging: sync");
    check(dq_stream_sync(P.s_in), "paging: sync");
    P.zero = nzero;
    P.st.sections++;

    // copy / compute overlap from the event timeline
    std::vector<std::pair<double, double>> comp;
    std::vector<std::pair<double, double>> copies;
    for (const UnitEv &E : uev) {
        auto at = [&](void *e) {
            float ms = 0;
            check(dq_event_elapsed_ms(t_start, e, &ms), "paging: event time");
            return (double)ms;
        };
        comp.push_back({at(E.c0), at(E.c1)});
        copies.push_back({at(E.in0), at(E.in1)});
        copies.push_back({at(E.o0), at(E.o1)});
    }
    for (const auto &iv : comp) P.st.compute_ms += iv.second - iv.first;
    for (const auto &iv : copies) {
        P.st.copy_ms += iv.second - iv.first;
        P.st.hidden_ms += interval_overlap(iv.first, iv.second, comp);
    }
}

}  // namespace

size_t device_budget() {
    if (const char *e = getenv("DGQ_HBM_BUDGET")) {
        const double v = atof(e);
        if (v > 0) return (size_t)v;
    }
    char name[128];
    int cus = 0;
    size_t total = 0, freem = 0;
    if (dq_device_info(name, sizeof name, &cus, &total, &freem) != DQ_OK) return (size_t)-1;
    const size_t margin = (size_t)4 << 30;
    return freem > 2 * margin ? freem - margin : freem / 2;
}

PagedStore *paged_create(int n, int prec, size_t budget) {
    const size_t elem = prec == DQ_F64 ? 16 : 8;
    if ((elem << n) <= budget) return nullptr;
    // pieces: kBuffers device buffers within 3/4 of the budget
    int nl = n - 1;
    while (nl > 10 && (size_t)kBuffers * (elem << nl) > budget / 4 * 3) nl--;
    if (n - nl > 24) nl = n - 24;
    PagedStore *P = new PagedStore();
    P->n = n;
    P->nl = nl;
    P->K = n - nl;
    P->prec = prec;
    P->elem = elem;
    P->piece_bytes = elem << nl;
    const uint64_t np = 1ull << P->K;
    P->host.assign(np, nullptr);
    P->zero.assign(np, 1);
    for (uint64_t c = 0; c < np; c++) {
        void *h = nullptr;
        const int rc = dq_host_alloc(&h, P->piece_bytes);
        if (rc != DQ_OK) {
            for (uint64_t k = 0; k < c; k++) dq_host_free(P->host[k]);
            delete P;
            engine_fatal("createQureg: pinned host memory for a paged register", rc);
        }
        P->host[c] = (unsigned char *)h;
    }
    P->dbuf.assign(kBuffers, nullptr);
    for (int i = 0; i < kBuffers; i++) check(dq_malloc(&P->dbuf[i], P->piece_bytes), "createQureg: device piece buffers");
    check(dq_stream_create(&P->s_in), "createQureg: stream");
    check(dq_stream_create(&P->s_comp), "createQureg: stream");
    check(dq_stream_create(&P->s_out), "createQureg: stream");
    return P;
}

void paged_destroy(PagedStore *P) {
    if (!P) return;
    dq_device_sync();
    for (void *h : P->host) dq_host_free(h);
    for (void *d : P->dbuf) dq_free(d);
    if (P->d_ws) dq_free(P->d_ws);
    dq_stream_destroy(P->s_in);
    dq_stream_destroy(P->s_comp);
    dq_stream_destroy(P->s_out);
    delete P;
}

void paged_flush(QuregState *s) {
    PagedStore &P = *s->paged;
    if (s->pending.empty() && s->init_kind == DQ_INIT_LOAD) return;
    const auto t0 = std::chrono::steady_clock::now();
    P.st = PagedStats();
    const int n = P.n, nl = P.nl;
    std::vector<int> pos(n);
    for (int i = 0; i < n; i++) pos[i] = i;
    bool init = s->init_kind != DQ_INIT_LOAD;
    const int ik = s->init_kind;
    uint64_t iidx = s->init_index;
    const double ival = s->init_value;
    std::vector<std::pair<int, int>> remap;
    Dag d(n, true);
    for (const auto &g : s->pending) d.add(g);
    while (!d.empty()) {
        uint64_t local = 0;
        for (int p = 0; p < n; p++)
            if (pos[p] < nl) local |= 1ull << p;
        std::vector<Gate> bg;
        uint64_t m = local;
        d.select(bg, m, nl, (size_t)-1);
        if (!bg.empty()) {
            for (auto &g : bg) {
                g.tmask = map_mask(g.tmask, pos);
                g.cmask = map_mask(g.cmask, pos);
            }
            run_section(s, bg, remap, init, ik, iidx, ival);
            remap.clear();
            init = false;
            continue;
        }
        // stuck: every front gate has a target among the high qubits
        const uint64_t ft = d.front_targets();
        std::vector<int> need;
        uint64_t busy = 0;  // piece positions the front gates target
        for (uint64_t t = ft; t; t &= t - 1) {
            const int p = pos[ctz64(t)];
            if (p >= nl) need.push_back(p);
            else busy |= 1ull << p;
        }
        std::sort(need.begin(), need.end());
        if ((int)remap.size() >= kMaxRemap) {  // move what is decided before deciding more
            run_section(s, {}, remap, init, ik, iidx, ival);
            remap.clear();
            init = false;
        }
        std::vector<int> cand;
        for (int l = nl - 1; l >= nl - kMaxRemap; l--) {
            bool used = false;
            for (const auto &pr : remap) used |= pr.second == l;
            if (!used) cand.push_back(l);
        }
        std::stable_sort(cand.begin(), cand.end(), [&](int a, int c) {
            return ((busy >> a) & 1) < ((busy >> c) & 1);
        });
        const int take = std::min({(int)need.size(), (int)cand.size(), kMaxRemap - (int)remap.size()});
        for (int i = 0; i < take; i++) {
            const int h = need[i], l = cand[i];
            for (int p = 0; p < n; p++) {
                if (pos[p] == h) pos[p] = l;
                else if (pos[p] == l) pos[p] = h;
            }
            if (((iidx >> h) & 1) != ((iidx >> l) & 1)) iidx ^= (1ull << h) | (1ull << l);
            remap.push_back({h, l});
            P.st.remaps++;
        }
    }
    if (init || !remap.empty()) run_section(s, {}, remap, init, ik, iidx, ival);
    for (int q = 0; q < n; q++) s->l2p[q] = pos[s->l2p[q]];
    for (int q = 0; q < n; q++) s->p2l[s->l2p[q]] = q;
    s->pending.clear();
    s->init_kind = DQ_INIT_LOAD;
    s->probs_dirty = true;
    s->det_base = false;
    P.st.wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

void paged_probs(QuregState *s) {
    PagedStore &P = *s->paged;
    const int n = P.n, nl = P.nl;
    const uint64_t np = 1ull << P.K;
    const size_t pws = dq_prob_ws_bytes(nl);
    const size_t res_bytes = (size_t)np * (nl + 1) * sizeof(double);
    const size_t need = 2 * pws + res_bytes + 256;
    if (need > P.ws_cap) {
        check(dq_device_sync(), "paging: sync");
        if (P.d_ws) dq_free(P.d_ws);
        P.ws_cap = need;
        check(dq_malloc(&P.d_ws, P.ws_cap), "paging: workspace");
    }
    unsigned char *w = (unsigned char *)P.d_ws;
    double *d_res = (double *)(w + 2 * pws);
    std::vector<void *> evs;
    void *last[2] = {nullptr, nullptr};
    int slot = 0;
    for (uint64_t c = 0; c < np; c++) {
        if (P.zero[c]) continue;
        if (last[slot]) check(dq_stream_wait_event(P.s_in, last[slot]), "paging: wait");
        check(dq_memcpy_async(P.dbuf[slot], P.host[c], P.piece_bytes, DQ_H2D, P.s_in), "paging: copy in");
        void *e_in = nullptr, *e_done = nullptr;
        check(dq_event_create(&e_in), "paging: event");
        check(dq_event_create(&e_done), "paging: event");
        evs.push_back(e_in);
        evs.push_back(e_done);
        check(dq_event_record(e_in, P.s_in), "paging: event record");
        check(dq_stream_wait_event(P.s_comp, e_in), "paging: wait");
        check(dq_prob_zero_all_async(P.dbuf[slot], nl, P.prec, d_res + c * (nl + 1), w + slot * pws, pws, P.s_comp),
              "paging: probabilities");
        check(dq_event_record(e_done, P.s_comp), "paging: event record");
        last[slot] = e_done;
        slot ^= 1;
    }
    std::vector<double> res((size_t)np * (nl + 1), 0.0);
    check(dq_stream_sync(P.s_comp), "paging: sync");
    check(dq_memcpy_async(res.data(), d_res, res_bytes, DQ_D2H, P.s_comp), "paging: copy out");
    check(dq_stream_sync(P.s_comp), "paging: sync");
    for (void *e : evs) dq_event_destroy(e);
    s->p0.assign(n + 1, 0.0);
    for (uint64_t c = 0; c < np; c++) {
        if (P.zero[c]) continue;
        const double *r = &res[(size_t)c * (nl + 1)];
        for (int q = 0; q < nl; q++) s->p0[q] += r[q];
        const double tot = r[nl];
        for (int h = nl; h < n; h++)
            if (!((c >> (h - nl)) & 1)) s->p0[h] += tot;
        s->p0[n] += tot;
    }
    s->probs_dirty = false;
}

static uint64_t phys_index(const QuregState *s, uint64_t i) {
    uint64_t p = 0;
    for (int q = 0; q < s->n; q++) p |= ((i >> q) & 1ull) << s->l2p[q];
    return p;
}

void paged_read(QuregState *s, uint64_t start, uint64_t count, double *out) {
    PagedStore &P = *s->paged;
    const uint64_t mask = (1ull << P.nl) - 1;
    for (uint64_t k = 0; k < count; k++) {
        const uint64_t p = phys_index(s, start + k);
        const uint64_t c = p >> P.nl, off = p & mask;
        if (P.zero[c]) {
            out[2 * k] = out[2 * k + 1] = 0.0;
        } else if (P.prec == DQ_F64) {
            const double *a = (const double *)P.host[c] + 2 * off;
            out[2 * k] = a[0];
            out[2 * k + 1] = a[1];
        } else {
            const float *a = (const float *)P.host[c] + 2 * off;
            out[2 * k] = a[0];
            out[2 * k + 1] = a[1];
        }
    }
}

void paged_write(QuregState *s, uint64_t start, uint64_t count, const double *re, const double *im) {
    PagedStore &P = *s->paged;
    const uint64_t mask = (1ull << P.nl) - 1;
    for (uint64_t k = 0; k < count; k++) {
        const uint64_t p = phys_index(s, start + k);
        const uint64_t c = p >> P.nl, off = p & mask;
        if (P.zero[c]) {
            memset(P.host[c], 0, P.piece_bytes);
            P.zero[c] = 0;
        }
        if (P.prec == DQ_F64) {
            double *a = (double *)P.host[c] + 2 * off;
            a[0] = re[k];
            a[1] = im[k];
        } else {
            float *a = (float *)P.host[c] + 2 * off;
            a[0] = (float)re[k];
            a[1] = (float)im[k];
        }
    }
    s->probs_dirty = true;
    s->det_base = false;
}

}  // namespace dgq
