// trainGMM -- training and adaptation tool over the gfx950 accumulator.
// Option names follow the reference tool (projects/Model_ADAPT/Model_ADAPT.cpp
// :278-401, trainer/OL_ModelAdapt.cpp:2032-2135); the hot-path subset is
// supported: --type 0 (from scratch; --full-cov for full covariances),
// 1 (MLLR), 2 (MAP/EM), 3 (fMLLR), 4 (MLLR+MAP); --robust (-R) for the
// robust covariance estimate (from scratch and MAP/EM).
// --file-by-file (-f) reads the feature files at every accumulation in groups
// of --prm-buffer GB (the next group is read while the device works on the
// current one); by default all frames stay resident in HBM across the EM
// iterations.  --SSE/--CUDA are accepted and ignored: statistics always run
// on the GPU.
#include <algorithm>
#include <cstdio>
#include <iostream>
#include <string>

#include "../host/feature_io.h"
#include "../host/model_adapt.h"
#include "cli_args.h"

using namespace gmm;

static int precision_of(const std::string &s) {
    if (s.empty() || s == "fp32") return GMMHIP_PREC_FP32;
    if (s == "bf16x3") return GMMHIP_PREC_BF16X3;
    if (s == "fp16x3") return GMMHIP_PREC_FP16X3;
    throw std::runtime_error("unknown --precision " + s);
}

int main(int argc, char **argv) {
    CliArgs a;
    a.option("help", 'h', true);
    a.option("in-file", 'i', false);
    a.option("data-T", 0, false, "0");
    a.option("in-UBM", 'u', false);
    a.option("load-txt", 'L', true, "0");
    a.option("save-txt", 'T', true, "0");
    a.option("outFile", 'o', false);
    a.option("type", 't', false, "2");
    a.option("nummix", 'g', false, "0");
    a.option("it-EM", 0, false, "0");
    a.option("it-MAP", 0, false, "0");
    a.option("tau", 'r', false, "0");
    a.option("mean", 'M', true, "0");
    a.option("var", 'V', true, "0");
    a.option("weight", 'W', true, "0");
    a.option("min-var", 0, false, "1e-6");
    a.option("mllr-init", 0, true, "0");
    a.option("it-mllr", 0, false, "1");
    a.option("it-fMLLR", 0, false, "20");
    a.option("apply-trn", 'a', true, "0");
    a.option("save-trn", 's', true, "0");
    a.option("mllr-Qval", 0, true, "0");
    a.option("mat-ext", 'X', false, ".mat");
    a.option("SSE", 0, true, "0");
    a.option("CUDA", 0, true, "0");
    a.option("GPU-id", 'G', false, "-1");
    a.option("NB-GPU", 0, false, "8");
    a.option("numThrd", 0, false, "1");
    a.option("precision", 0, false, "fp32");
    a.option("full-cov", 0, true, "0");
    a.option("robust", 'R', true, "0");
    a.option("file-by-file", 'f', true, "0");
    a.option("prm-buffer", 0, false, "0.005");
    a.option("dwnsmp", 0, false, "1");
    a.option("verbosity", 'v', true, "0");
    try {
        a.parse(argc, argv);
        if (a.b("help") || a.all("in-file").empty() || !a.has("outFile")) {
            std::cout << "usage: trainGMM -i FEATS [-i ...] --data-T {0,1,2} -o OUT.gmm [-T] "
                         "--type {0,1,2,3,4} [--in-UBM UBM.gmm [-L]] [--nummix K [--full-cov]] [--robust] [--it-EM n] "
                         "[--it-MAP n] [--tau r] [--mean] [--var] [--weight] [--mllr-init] "
                         "[--it-mllr n] [--apply-trn] [--save-trn] [--precision fp32|fp16x3|bf16x3]\n";
            return a.b("help") ? 0 : 1;
        }
        Features feats;
        const int dtype = a.i("data-T");
        const unsigned dwnsmp = (unsigned)std::max(1, a.i("dwnsmp"));
        const bool by_file = a.b("file-by-file");
        if (!by_file) {
            // all frames resident (in host memory, then in HBM across the EM
            // iterations)
            for (const std::string &fn : a.all("in-file")) {
                Features f;
                std::string err;
                if (LoadFeatures(fn.c_str(), dtype, f, &err) != 0) throw std::runtime_error(err);
                if (feats.D && f.D != feats.D)
                    throw std::runtime_error("feature dimensions differ across files");
                feats.D = f.D;
                for (unsigned i = 0; i < f.N; i += dwnsmp) {
                    feats.data.insert(feats.data.end(), f.data.begin() + (size_t)i * f.D,
                                      f.data.begin() + (size_t)(i + 1) * f.D);
                    feats.N++;
                }
            }
            feats.rows.resize(feats.N);
            for (unsigned i = 0; i < feats.N; i++) feats.rows[i] = &feats.data[(size_t)i * feats.D];
        }

        CModelAdapt ad;
        CModelAdapt::OPTIONS o;
        o._adaptType = a.i("type");
        o._iterCountEM = a.i("it-EM");
        o._iterCountMAP = a.i("it-MAP");
        o._tau = a.i("tau");
        o._AdaptMeanBool = a.b("mean");
        o._AdaptVarBool = a.b("var");
        o._AdaptWeightBool = a.b("weight");
        o._min_var = (float)a.f("min-var");
        o._mllrUBMInit = a.b("mllr-init");
        o._iterMLLR = a.i("it-mllr");
        o._iter_fMLLR = a.i("it-fMLLR");
        o._applyTransforms = a.b("apply-trn");
        o._saveTransforms = a.b("save-trn");
        o._getQvalMLLR = a.b("mllr-Qval");
        o._transformFileExt = a.str("mat-ext");
        o._cudaGPUid = a.i("GPU-id");
        o._verbosity = a.i("verbosity");
        o._precision = precision_of(a.str("precision"));
        o._robustCov = a.b("robust");
        ad.SetOptions(o);
        if (by_file)
            ad.InsertFiles(a.all("in-file"), dtype, dwnsmp, (float)a.f("prm-buffer"));
        else
            ad.InsertData(feats.rows.data(), feats.N, feats.D);
        if (o._adaptType == FROM_SCRATCH) {
            if (a.i("nummix") < 1) throw std::runtime_error("--type 0 needs --nummix");
            ad.TrainFromScratch((unsigned)a.i("nummix"), a.b("full-cov"));
        } else {
            if (!a.has("in-UBM")) throw std::runtime_error("--in-UBM is required for adaptation");
            GMModel ubm;
            if (ubm.Load(a.str("in-UBM").c_str(), a.b("load-txt")) != 0)
                throw std::runtime_error("UBM [" + a.str("in-UBM") + "] not loaded: " + ubm.LastError());
            ad.AdaptModel(&ubm);
        }
        if (ad.Save(a.str("outFile").c_str(), a.b("save-txt")) != 0) {
            std::cerr << "Missing client model or transformation matrix." << std::endl;
            return 1;
        }
    } catch (std::exception &e) {
        std::cerr << e.what() << std::endl;
        return 1;
    }
    return 0;
}
