// getLogLike -- per-file mean or per-frame log-likelihood of feature files
// under a GMM, on the gfx950 accumulator.  Options follow the reference tool
// (projects/GetLogLikeGMM/main.cpp:262-361): -i files, --data-T, --in-GMM/-g,
// -L, --for-each/-f with -o DIR, -x EXT (.logL) and -T, written with saveMat
// (matrices/IOMats.h:168-232); without --for-each each file's mean
// log-likelihood is printed as "<file>\t<mean>".
#include <cstdio>
#include <iostream>
#include <string>

#include "../host/feature_io.h"
#include "../host/gmm_estimator.h"
#include "cli_args.h"

using namespace gmm;

static std::string out_path(const std::string &in, const std::string &dir, const std::string &ext) {
    std::string base = in;
    const size_t sl = base.find_last_of("/\\");
    if (sl != std::string::npos) base = base.substr(sl + 1);
    const size_t dot = base.rfind('.');
    if (dot != std::string::npos) base = base.substr(0, dot);
    std::string d = dir.empty() ? "./" : dir;
    if (d.back() != '/') d += '/';
    return d + base + ext;
}

int main(int argc, char **argv) {
    CliArgs a;
    a.option("help", 'h', true);
    a.option("in-file", 'i', false);
    a.option("in-GMM", 'g', false);
    a.option("for-each", 'f', true, "0");
    a.option("outDir", 'o', false, "./");
    a.option("outExt", 'x', false, ".logL");
    a.option("load-txt", 'L', true, "0");
    a.option("save-txt", 'T', true, "0");
    a.option("data-T", 0, false, "0");
    a.option("SSE", 0, true, "0");
    a.option("CUDA", 0, true, "0");
    a.option("GPU-id", 'G', false, "-1");
    a.option("numThrd", 0, false, "1");
    a.option("verbosity", 'v', true, "0");
    try {
        a.parse(argc, argv);
        if (a.b("help") || a.all("in-file").empty() || !a.has("in-GMM")) {
            std::cout << "usage: getLogLike -i FEATS [-i ...] --data-T {0,1,2} --in-GMM M.gmm [-L] "
                         "[--for-each -o DIR [-x .logL] [-T]]\n";
            return a.b("help") ? 0 : 1;
        }
        GMModel gmm;
        if (gmm.Load(a.str("in-GMM").c_str(), a.b("load-txt")) != 0)
            throw std::runtime_error("GMM model [" + a.str("in-GMM") + "] not found: " + gmm.LastError());
        GMMStatsEstimator_HIP<float> est(a.i("GPU-id"));
        est._verbosity = a.i("verbosity");
        est.insertModel(gmm);
        std::vector<float> ll;
        for (const std::string &fn : a.all("in-file")) {
            Features f;
            std::string err;
            const int rc = LoadFeatures(fn.c_str(), a.i("data-T"), f, &err);
            if (rc == -1) {
                std::cerr << "WARNING: Param file [" << fn << "] not found -> skipped!" << std::endl;
                continue;
            }
            if (rc != 0) throw std::runtime_error(err);
            if (f.N < 1) continue;
            ll.resize(f.N);
            est.compVecsLogLikeMT(f.rows.data(), f.N, f.D, ll.data());
            if (est._fNumStabilityTroubles) {
                std::cerr << "WARNING: numerical stability troubles detected [file> " << fn << "]"
                          << std::endl;
                est._fNumStabilityTroubles = false;
            }
            if (a.b("for-each")) {
                SaveMat(out_path(fn, a.str("outDir"), a.str("outExt")).c_str(), a.b("save-txt"),
                        ll.data(), 1, f.N);
            } else {
                float mean = 0.0f;
                for (float v : ll) mean += v;
                mean /= f.N;
                std::cout << fn << "\t" << mean << "\n";
            }
        }
    } catch (std::exception &e) {
        std::cerr << e.what() << std::endl;
        return 1;
    }
    return 0;
}
