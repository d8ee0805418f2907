// Tuning table loader (tune.h).
#include "tune.h"

#include "../../include/gmmhip.h"

#include <dlfcn.h>

#include <cmath>
#include <cstdlib>
#include <fstream>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

namespace gmmhip_tune {
namespace {

std::mutex g_mu;
std::vector<Entry> g_table;
bool g_loaded = false;

std::string default_path() {
    const char *env = std::getenv("GMMHIP_TUNE_FILE");
    if (env && *env) return env;
    Dl_info info;
    if (dladdr((void *)&load, &info) && info.dli_fname) {
        std::string lib = info.dli_fname;
        const size_t slash = lib.rfind('/');
        const std::string dir = slash == std::string::npos ? "." : lib.substr(0, slash);
        return dir + "/../tuning/gfx950.tune";
    }
    return "gmm_amd/tuning/gfx950.tune";
}

int parse_prec(const std::string &s) {
    if (s == "fp32") return GMMHIP_PREC_FP32;
    if (s == "bf16x3") return GMMHIP_PREC_BF16X3;
    if (s == "fp16x3") return GMMHIP_PREC_FP16X3;
    return -1;
}

int load_locked(const char *path) {
    g_table.clear();
    g_loaded = true;
    const std::string p = path ? std::string(path) : default_path();
    std::ifstream f(p);
    if (!f) return -1;
    std::string line;
    while (std::getline(f, line)) {
        const size_t hash = line.find('#');
        if (hash != std::string::npos) line.resize(hash);
        std::istringstream ss(line);
        std::string kind, prec;
        Entry e{};
        if (!(ss >> kind >> prec >> e.D >> e.M >> e.ll_frames >> e.chunk_frames)) continue;
        if (!(ss >> e.ms)) e.ms = 0.0;
        if (!(ss >> e.comp_tile)) e.comp_tile = 128;
        if (e.comp_tile != 128 && e.comp_tile != 256) continue;
        e.path = kind == "estep" ? GMMHIP_TUNE_ESTEP : kind == "batch" ? GMMHIP_TUNE_BATCH : -1;
        e.prec = parse_prec(prec);
        if (e.path < 0 || e.prec < 0 || e.D <= 0 || e.M <= 0) continue;
        if (e.ll_frames != 64 && e.ll_frames != 128 && e.ll_frames != 256) continue;
        if (e.chunk_frames < 64) continue;
        g_table.push_back(e);
    }
    return (int)g_table.size();
}

}  // namespace

int load(const char *path) {
    std::lock_guard<std::mutex> g(g_mu);
    return load_locked(path);
}

bool lookup(int path, int prec, int D, int M, Entry *out) {
    std::lock_guard<std::mutex> g(g_mu);
    if (!g_loaded) load_locked(nullptr);
    const Entry *best = nullptr;
    double bestd = 0.0;
    for (const Entry &e : g_table) {
        if (e.path != path || e.prec != prec) continue;
        // same D first (distance below 1e3 beats any other D), then log M
        const double dM = std::fabs(std::log((double)e.M / M));
        const double dD = std::fabs(std::log((double)e.D / D));
        const double d = (e.D == D ? 0.0 : 1e3 + dD * 10.0) + dM;
        if (!best || d < bestd) {
            best = &e;
            bestd = d;
        }
    }
    if (!best) return false;
    *out = *best;
    return true;
}

}  // namespace gmmhip_tune
