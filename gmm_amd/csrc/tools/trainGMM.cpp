// trainGMM -- training and adaptation tool over the gfx950 accumulator.
// Option names follow the reference tool (projects/Model_ADAPT/Model_ADAPT.cpp
// :278-401, trainer/OL_ModelAdapt.cpp:2032-2135); the hot-path subset is
// supported: --type 0 (from scratch; --full-cov for full covariances),
// 1 (MLLR), 2 (MAP/EM), 3 (fMLLR), 4 (MLLR+MAP); --robust (-R) for the
// robust covariance estimate (from scratch and MAP/EM); --split-X,
// --split-th, --ll-Th and --tmp-N for training from scratch
// (OL_ModelAdapt.cpp:2044-2056, :2099-2100).
// --file-by-file (-f) reads the feature files at every accumulation in groups
// of --prm-buffer GB (the next group is read while the device works on the
// current one); by default all frames stay resident in HBM across the EM
// iterations.  --SSE/--CUDA are accepted and ignored: statistics always run
// on the GPU.
//
// One model per class (Model_ADAPT.cpp:97-240, :403-476), written to
// OUTDIR/CLASS.EXT (--outDir, --outExt):
//   MLF mode    -S classes -i labels.mlf -e .ext [-I dir] [--sample-period p]:
//               every class line "CLASS elem1 elem2 ..." collects the time
//               intervals of its elements from the MLF, per file
//               (MLF_Parser::parseClasses / parseMLF,
//               tools/OL_MLFParser.cpp:57-163); file = dir/ + the MLF path
//               + ext.  SVES and RAW files need --sample-period (HTK files
//               carry theirs).
//   CLASS mode  -S classes [-e .ext] [-I dir]: "CLASS file1 file2 ..." trains
//               each class from whole files dir/fileK.ext.
#include <algorithm>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

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

// Where the frames of one model come from.
struct DataSource {
    std::vector<std::string> files;  // whole files
    file_2_ranges ranges;            // or intervals of files (MLF mode)
    bool use_ranges = false;
};

// parseClasses (OL_MLFParser.cpp:57-87): "CLASS elem ..." per line; classes
// in file order, element -> classes, and class -> elements in line order.
static void parse_classes(const std::string &fn, std::vector<std::string> &classes,
                          std::map<std::string, std::vector<std::string>> &elem2class,
                          std::map<std::string, std::vector<std::string>> &class2elem) {
    std::ifstream f(fn);
    if (!f) throw std::runtime_error("parseClasses(): file " + fn + " could not be opened for reading");
    std::string line;
    while (std::getline(f, line)) {
        std::istringstream ls(line);
        std::string cls, item;
        if (!(ls >> cls)) continue;
        while (ls >> item) {
            elem2class[item].push_back(cls);
            class2elem[cls].push_back(item);
        }
        classes.push_back(cls);
    }
}

// fixPath (OL_MLFParser.cpp:90-112): drops the leading quote / '*' / slash
// characters of an MLF file line and the extension of its last component,
// keeping the directories between.
static std::string fix_path(const std::string &p) {
    size_t j = 0;
    while (j < p.size() && (p[j] == '"' || p[j] == ' ' || p[j] == '*' || p[j] == '/' || p[j] == '\\')) j++;
    size_t slash = p.find_last_of('/');
    if (slash == std::string::npos) slash = p.find_last_of('\\');
    std::string out;
    if (slash != std::string::npos && slash > j) out = p.substr(j, slash - j) + "/";
    std::string leaf = p.substr(slash == std::string::npos ? 0 : slash + 1);
    const size_t dot = leaf.rfind('.');
    if (dot != std::string::npos && dot > 0) leaf.erase(dot);
    return out + leaf;
}

// parseMLF (OL_MLFParser.cpp:115-163): a quoted line names the file, a label
// line "begin end element ..." adds [begin, end) to every class of element.
static void parse_mlf(const std::string &fn, const std::map<std::string, std::vector<std::string>> &elem2class,
                      std::map<std::string, file_2_ranges> &class2frames) {
    std::ifstream f(fn);
    if (!f) throw std::runtime_error("MLF_Parser::parse(): file " + fn + " could not be opened for reading");
    std::string line, file;
    while (std::getline(f, line)) {
        if (line.empty() || line[0] == '\r' || line[0] == '.' || line[0] == '#') continue;
        if (line[0] == '"') {
            file = fix_path(line);
            continue;
        }
        std::istringstream ls(line);
        unsigned int b = 0, e = 0;
        std::string elem;
        if (!(ls >> b >> e >> elem)) continue;
        auto it = elem2class.find(elem);
        if (it == elem2class.end()) continue;
        for (const std::string &cls : it->second) class2frames[cls][file].push_back({b, e});
    }
}

int main(int argc, char **argv) {
    CliArgs a;
    a.option("help", 'h', true);
    a.option("in-file", 'i', false);
    a.option("in-list", 'I', false);
    a.option("in-classes", 'S', false);
    a.option("inExt", 'e', false);
    a.option("sample-period", 0, false, "0");
    a.option("outDir", 'D', false, "./");
    a.option("outExt", 'x', false, ".gmm");
    a.option("data-T", 0, false, "0");
    a.option("in-UBM", 'u', false);
    a.option("load-txt", 'L', true, "0");
    a.option("save-txt", 'T', true, "0");
    a.option("outFile", 'o', false);
    a.option("type", 't', false, "2");
    a.option("nummix", 'g', false, "0");
    a.option("split-X", 0, false, "0");
    a.option("split-th", 0, false, "0");
    a.option("ll-Th", 0, false, "0");
    a.option("tmp-N", 0, false, "0");
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
        const bool classes = a.given("in-classes");
        const bool mlf = classes && a.given("in-file");
        if (a.b("help") || (!classes && (a.all("in-file").empty() || !a.given("outFile")))) {
            std::cout << "usage: trainGMM -i FEATS [-i ...] --data-T {0,1,2} -o OUT.gmm [-T] "
                         "--type {0,1,2,3,4} [--in-UBM UBM.gmm [-L]] [--nummix K [--full-cov] [--split-X f] "
                         "[--split-th f] [--ll-Th t] [--tmp-N n]] [--robust] [--it-EM n] "
                         "[--it-MAP n] [--tau r] [--mean] [--var] [--weight] [--mllr-init] "
                         "[--it-mllr n] [--apply-trn] [--save-trn] [--precision fp32|fp16x3|bf16x3]\n"
                         "       trainGMM -S CLASSES [-i LABELS.mlf -e .EXT [--sample-period P]] [-I DIR] "
                         "[-D OUTDIR] [-x .gmm] ... (one model per class)\n";
            return a.b("help") ? 0 : 1;
        }
        const int dtype = a.i("data-T");
        const unsigned dwnsmp = (unsigned)std::max(1, a.i("dwnsmp"));
        const bool by_file = a.b("file-by-file");
        const unsigned sample_period = (unsigned)std::max(0, a.i("sample-period"));
        if (mlf && dtype != 1 && sample_period == 0)
            throw std::runtime_error("LoadConfiguration(): you have to specify \"sample-period\" for SVES files!");
        if (mlf && !a.given("inExt"))
            throw std::runtime_error("LoadConfiguration(): you have to specify \"inExt\" for MLF mode!");
        if (classes && a.all("in-list").size() > 1)
            throw std::runtime_error("LoadConfiguration(): \"in-list\" in MLF | CLASS mode may be specified only once!");

        // the models to produce: (output file, data)
        std::vector<std::pair<std::string, DataSource>> jobs;
        if (classes) {
            std::vector<std::string> names;
            std::map<std::string, std::vector<std::string>> elem2class, class2elem;
            for (const std::string &fn : a.all("in-classes")) parse_classes(fn, names, elem2class, class2elem);
            const std::string prefix = a.given("in-list") && !a.str("in-list").empty() ? a.str("in-list") + "/" : "";
            const std::string ext = a.given("inExt") ? a.str("inExt") : (mlf ? "" : ".prm");
            std::map<std::string, file_2_ranges> class2frames;
            if (mlf)
                for (const std::string &fn : a.all("in-file")) parse_mlf(fn, elem2class, class2frames);
            for (const std::string &cls : names) {
                DataSource ds;
                if (mlf) {
                    ds.use_ranges = true;
                    for (const auto &kv : class2frames[cls]) ds.ranges[prefix + kv.first + ext] = kv.second;
                } else {
                    for (const std::string &e : class2elem[cls]) ds.files.push_back(prefix + e + ext);
                }
                jobs.emplace_back(a.str("outDir") + "/" + cls + a.str("outExt"), std::move(ds));
            }
        } else {
            DataSource ds;
            ds.files = a.all("in-file");
            jobs.emplace_back(a.str("outFile"), std::move(ds));
        }

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
        o._splitXpercAtOnce = (float)a.f("split-X");
        o._splitTh = (float)a.f("split-th");
        o._llTh = (float)a.f("ll-Th");
        o._saveModelAfterNSplits = a.i("tmp-N");
        if (o._adaptType == FROM_SCRATCH && a.i("nummix") < 1) throw std::runtime_error("--type 0 needs --nummix");
        if (o._adaptType != FROM_SCRATCH && !a.given("in-UBM"))
            throw std::runtime_error("--in-UBM is required for adaptation");

        int failed = 0;
        for (auto &job : jobs) {
            const std::string &out = job.first;
            DataSource &ds = job.second;
            CModelAdapt ad;
            ad.SetOptions(o);
            Features feats;
            try {
                if (by_file) {
                    if (ds.use_ranges)
                        ad.InsertRanges(ds.ranges, dtype, sample_period, (float)a.f("prm-buffer"));
                    else
                        ad.InsertFiles(ds.files, dtype, dwnsmp, (float)a.f("prm-buffer"));
                } else {
                    // all frames resident (in host memory, then in HBM
                    // across the EM iterations); unreadable files are
                    // skipped as fillAdaptationWithData skips them
                    std::vector<std::pair<std::string, const element_ranges *>> items;
                    if (ds.use_ranges)
                        for (const auto &kv : ds.ranges) items.emplace_back(kv.first, &kv.second);
                    else
                        for (const std::string &fn : ds.files) items.emplace_back(fn, nullptr);
                    for (const auto &it : items) {
                        Features f;
                        std::string err;
                        const int rc = it.second
                                           ? LoadFeatureRanges(it.first.c_str(), dtype, *it.second, sample_period,
                                                               f, &err)
                                           : LoadFeatures(it.first.c_str(), dtype, f, &err);
                        if (rc == -1 && classes) {
                            if (o._verbosity > 1)
                                std::cout << "File " << it.first << " could not be opened -> skipped" << std::endl;
                            continue;
                        }
                        if (rc != 0) throw std::runtime_error(err);
                        if (f.N == 0) continue;
                        if (feats.D && f.D != feats.D)
                            throw std::runtime_error("feature dimensions differ across files");
                        feats.D = f.D;
                        const unsigned step = it.second ? 1 : dwnsmp;
                        for (unsigned i = 0; i < f.N; i += step) {
                            feats.data.insert(feats.data.end(), f.data.begin() + (size_t)i * f.D,
                                              f.data.begin() + (size_t)(i + 1) * f.D);
                            feats.N++;
                        }
                    }
                    if (feats.N == 0) throw std::runtime_error("no feature vectors");
                    feats.rows.resize(feats.N);
                    for (unsigned i = 0; i < feats.N; i++) feats.rows[i] = &feats.data[(size_t)i * feats.D];
                    ad.InsertData(feats.rows.data(), feats.N, feats.D);
                }
            } catch (std::runtime_error &e) {
                if (!classes) throw;
                std::cerr << " WARNING: none data, adaptation skipped! [" << out << ": " << e.what() << "]"
                          << std::endl;
                continue;
            }
            if (o._adaptType == FROM_SCRATCH) {
                ad.TrainFromScratch((unsigned)a.i("nummix"), a.b("full-cov"), out.c_str(), a.b("save-txt"));
            } else {
                GMModel ubm;
                if (ubm.Load(a.str("in-UBM").c_str(), a.b("load-txt")) != 0)
                    throw std::runtime_error("UBM [" + a.str("in-UBM") + "] not loaded: " + ubm.LastError());
                ad.AdaptModel(&ubm);
            }
            if (ad.Save(out.c_str(), a.b("save-txt")) != 0) {
                std::cerr << "Missing client model or transformation matrix." << std::endl;
                failed++;
            }
        }
        if (failed) return 1;
    } catch (std::exception &e) {
        std::cerr << e.what() << std::endl;
        return 1;
    }
    return 0;
}
