// tune.cpp -- see tune.hpp.
#include "tune.hpp"

#include <dlfcn.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

namespace dgq {
namespace {

struct Entry {
    int prec, n;
    TuneParams t;
};

struct CostEntry {
    int prec, k, R, W;
    CostModel c;
};

std::mutex g_mu;
bool g_loaded = false;
std::vector<Entry> g_table;
std::vector<CostEntry> g_cost;

std::string default_path() {
    if (const char *e = getenv("DGQ_TUNE_TABLE")) return e;
    Dl_info info;
    if (dladdr((void *)&default_path, &info) && info.dli_fname) {
        std::string p = info.dli_fname;  // .../dgquest_amd/lib/libQuEST_fXX.so
        size_t s = p.rfind('/');
        if (s != std::string::npos) {
            p = p.substr(0, s);
            s = p.rfind('/');
            if (s != std::string::npos) return p.substr(0, s) + "/tune_table.txt";
        }
    }
    return "tune_table.txt";
}

bool shape_ok(int prec, int k, int R, int W) {
    return (prec == DQ_F32 || prec == DQ_F64) && k >= 1 && R >= 1 && R <= DQ_MAX_REG_BITS && W >= 0 &&
           W <= DQ_MAX_WAVE_BITS;
}

int load_locked(const std::string &path) {
    g_table.clear();
    g_cost.clear();
    g_loaded = true;
    FILE *f = fopen((path.empty() ? default_path() : path).c_str(), "r");
    if (!f) return -1;
    char line[512];
    while (fgets(line, sizeof line, f)) {
        char *h = strchr(line, '#');
        if (h) *h = 0;
        CostEntry ce;
        if (sscanf(line, " cost %d %d %d %d %lf %lf %lf %lf", &ce.prec, &ce.k, &ce.R, &ce.W, &ce.c.sweep, &ce.c.phase,
                   &ce.c.op, &ce.c.launch) == 8) {
            if (shape_ok(ce.prec, ce.k, ce.R, ce.W) && ce.c.sweep > 0 && ce.c.phase >= 0 && ce.c.op >= 0 &&
                ce.c.launch >= 0)
                g_cost.push_back(ce);
            continue;
        }
        Entry e;
        double ms;
        if (sscanf(line, "%d %d %d %d %d %d %lf", &e.prec, &e.n, &e.t.k, &e.t.R, &e.t.W, &e.t.max_ops, &ms) != 7)
            continue;
        if ((e.prec != DQ_F32 && e.prec != DQ_F64) || e.n < 1 || e.n > 64 || e.t.k < 1 || e.t.R < 1 ||
            e.t.R > DQ_MAX_REG_BITS || e.t.W < 0 || e.t.W > DQ_MAX_WAVE_BITS || e.t.max_ops < 1)
            continue;
        g_table.push_back(e);
    }
    fclose(f);
    return (int)g_table.size();
}

}  // namespace

int tune_load(const std::string &path) {
    std::lock_guard<std::mutex> lk(g_mu);
    return load_locked(path);
}

TuneParams tune_for(int prec, int n) {
    TuneParams t;
    if (const char *e = getenv("DGQ_TUNE")) {
        int k, R, W, mo;
        if (sscanf(e, "%d,%d,%d,%d", &k, &R, &W, &mo) == 4) {
            t.k = k;
            t.R = R;
            t.W = W;
            t.max_ops = mo;
        }
    } else {
        std::lock_guard<std::mutex> lk(g_mu);
        if (!g_loaded) load_locked("");
        const Entry *best = nullptr;
        for (const Entry &e : g_table) {
            if (e.prec != prec) continue;
            if (!best) { best = &e; continue; }
            const int d = abs(e.n - n), db = abs(best->n - n);
            if (d < db || (d == db && e.n > best->n)) best = &e;
        }
        if (best) t = best->t;
    }
    if (const char *e = getenv("DGQ_FUSE")) t.fuse_diagonals = atoi(e) != 0;
    return t;
}

CostModel tune_cost(int prec, const TuneParams &tp) {
    {
        std::lock_guard<std::mutex> lk(g_mu);
        if (!g_loaded) load_locked("");
        for (const CostEntry &e : g_cost)
            if (e.prec == prec && e.k == tp.k && e.R == tp.R && e.W == tp.W) return e.c;
    }
    // built-in estimate: a sweep at 5.5 TB/s; each further phase and each op
    // cost what they cost on the 13-qubit FP64 tile (2 waves per SIMD and
    // workgroup), the op cost halved for FP32 (half the bytes per amplitude,
    // the same instructions)
    CostModel c;
    c.sweep = 1.0 / 5.5;
    c.phase = 0.022;
    c.op = prec == DQ_F64 ? 0.0004 : 0.0008;
    c.launch = 0.006;
    return c;
}

std::vector<TuneParams> tune_candidates(int prec) {
    std::vector<TuneParams> out;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        if (!g_loaded) load_locked("");
        for (const CostEntry &e : g_cost) {
            if (e.prec != prec) continue;
            TuneParams t;
            t.k = e.k;
            t.R = e.R;
            t.W = e.W;
            out.push_back(t);
        }
    }
    if (out.empty()) {
        static const int built_in[][3] = {{3, 3, 0}, {4, 4, 0}, {5, 4, 1}, {6, 4, 2}, {6, 3, 3}, {7, 4, 3}};
        for (const auto &b : built_in) {
            TuneParams t;
            t.k = b[0];
            t.R = b[1];
            t.W = b[2];
            out.push_back(t);
        }
    }
    return out;
}

double predict_plan_ms(const Plan &P, int prec, const CostModel &c) {
    const double amp = prec == DQ_F64 ? 16.0 : 8.0;
    double ms = 0.0;
    for (const dq_block &b : P.blocks) {
        const double gb = (b.init_kind == DQ_INIT_LOAD ? 2.0 : 1.0) * amp * std::ldexp(1.0, (int)b.n_qubits) * 1e-9;
        uint32_t ops = 0;
        for (uint32_t p = b.phase_begin; p < b.phase_begin + b.phase_count; p++) ops += P.phases[p].op_count;
        ms += c.launch + gb * (c.sweep + c.phase * (double)(b.phase_count - 1) + c.op * (double)ops);
    }
    return ms;
}

}  // namespace dgq
