// tune.cpp -- see tune.hpp.
#include "tune.hpp"

#include <dlfcn.h>

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

std::mutex g_mu;
bool g_loaded = false;
std::vector<Entry> g_table;

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

int load_locked(const std::string &path) {
    g_table.clear();
    g_loaded = true;
    FILE *f = fopen((path.empty() ? default_path() : path).c_str(), "r");
    if (!f) return -1;
    char line[512];
    while (fgets(line, sizeof line, f)) {
        char *h = strchr(line, '#');
        if (h) *h = 0;
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

}  // namespace dgq
