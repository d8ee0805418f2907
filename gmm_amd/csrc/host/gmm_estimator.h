// GMMStatsEstimator_HIP<TS>: the drop-in for the reference's CUDA
// accumulator GMMStatsEstimator_CUDA<TS> (trainer/OL_GMMStatsEstimator_CUDA.h:
// 39-107 of GMMTeam/GMM).  Same method names, arguments and accumulate-into
// semantics; every statistic is computed by the gfx950 library through the C
// ABI (include/gmmhip.h).  There is no CPU path: the reference falls back to
// its SSE estimator without a device or below 8 frames per GPU
// (OL_GMMStatsEstimator_CUDA.cpp:55-60, :402-408); here the constructor throws
// without a device and every frame count runs on the GPU.  Errors of the C
// ABI are rethrown as std::runtime_error (std::logic_error for misuse), the
// exception classes the reference throws.
#pragma once

#include <algorithm>
#include <cstdint>
#include <future>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/gmmhip.h"
#include "feature_io.h"
#include "gmm_model.h"
#include "gmm_stats.h"

namespace gmm {

template <typename TS>
class GMMStatsEstimator_HIP {
public:
    // public knobs of GMMStatsEstimator (OL_GMMStatsEstimator.h:61-71)
    float _minLogLike = -1e20f;
    float _minGamma = 1e-4f;
    float _min_var = 1e-6f;
    unsigned int _numThreads = 1;
    unsigned int _NAccBlocks_CUDA = 8;  // accepted, unused (chunking is automatic)
    float _memoryBuffDataGB = 0.0f;
    float _memoryBuffDataGB_GPU = 0.0f;
    int _verbosity = 1;
    bool _fNumStabilityTroubles = false;
    int _precision = GMMHIP_PREC_FP32;

    explicit GMMStatsEstimator_HIP(int GPU_id = -1) {
        int n = 0;
        gmmhip_device_count(&n);
        if (n == 0) throw std::runtime_error("GMMStatsEstimator_HIP: no HIP device");
        dev_ = (GPU_id >= 0 && GPU_id < n) ? GPU_id : 0;
        gmmhip_opts o;
        gmmhip_default_opts(&o);
        check(gmmhip_create(dev_, &o, &h_));
    }
    ~GMMStatsEstimator_HIP() { gmmhip_destroy(h_); }
    GMMStatsEstimator_HIP(const GMMStatsEstimator_HIP &) = delete;
    GMMStatsEstimator_HIP &operator=(const GMMStatsEstimator_HIP &) = delete;

    const std::string &getEstimationType() const { return type_; }

    // Rearranges the model (zero-weight components last), computes its
    // constants and uploads it.  first=false adds it to the current group.
    void insertModel(GMModel &model, bool first = true) {
        if (!first && !group_.empty() && model.GetDimension() != group_[0].D)
            throw std::runtime_error("Models (GMMs) have to have same dimensions");
        if (first) group_.clear();
        const unsigned nzero = model.RearrangeMixtures();
        Packed p;
        p.D = model.GetDimension();
        p.Malloc = model.GetNumberOfMixtures();
        p.M = p.Malloc - (int)nzero;
        p.w.assign(model.weights().begin(), model.weights().begin() + p.M);
        p.mu.assign(model.means().begin(), model.means().begin() + (size_t)p.M * p.D);
        p.full = model.GetFullCovStatus();
        const size_t dv = model.GetDimVar();
        p.var.assign(model.vars().begin(), model.vars().begin() + (size_t)p.M * dv);
        // the diagonal form rejects variances below min_var
        // (OL_GMMStatsEstimator.cpp:279-285); the full form rejects matrices
        // that are not positive definite when they are uploaded
        if (!p.full)
            for (float v : p.var)
                if (v < _min_var) throw std::runtime_error("insertModel(): model contains very low variances!");
        group_.push_back(std::move(p));
        select(group_.size() - 1);
    }

    void setModelToBeUsed(unsigned int n) {
        if (n >= group_.size()) throw std::out_of_range("setModelToBeUsed(): no such model");
        select(n);
    }

    // auxStats: the robust auxiliary statistics (accumulateAuxStats) are
    // added too, into statistics allocated with them.
    void setAccFlags(bool mixProbsOnly, bool meanStats, bool varStats, bool varFull = false,
                     bool auxStats = false) {
        flagsSet_ = true;
        fFullVar_ = varFull;
        fMixProbOnly_ = mixProbsOnly;
        fMean_ = meanStats && !mixProbsOnly;
        fVar_ = varStats && !mixProbsOnly;
        fAux_ = auxStats && !mixProbsOnly;
    }
    void resetAccFlags() { flagsSet_ = false; }
    void setSameDataFlag() { sameData_ = true; }

    // Gaussian-selection mask (logLikeMask, OL_GMMStatsEstimator.h:73-114):
    // "fill" stores every frame's log-likelihood and posteriors before the
    // cut; "use" restricts frame i to maskIdxs[i].  Both apply to the frames
    // of each accumulateStatsMT / compVecsLogLikeMT call, frame i of the
    // call in row i.  The reference's CUDA estimator refuses fill and
    // ignores use; here both run on the device.
    struct logLikeMask {
        unsigned int N = 0;    // >= NSamples
        unsigned int dim = 0;  // >= number of components
        float *logLikes = nullptr;  // N
        float **gammas = nullptr;   // N x dim
        std::vector<std::vector<unsigned int>> maskIdxs;  // N lists
    };
    void mask2fill(logLikeMask &mask) {
        fillMask_ = true;
        useMask_ = false;
        mask_ = &mask;
    }
    void mask2use(logLikeMask &mask) {
        useMask_ = true;
        fillMask_ = false;
        mask_ = &mask;
    }
    logLikeMask *withdrawMask() {
        fillMask_ = useMask_ = false;
        logLikeMask *m = mask_;
        mask_ = nullptr;
        return m;
    }
    // estimateMaskIdxs (OL_GMMStatsEstimator.cpp:383-394): the components
    // whose filled posterior exceeds minGamma.
    static void estimateMaskIdxs(logLikeMask &mask, float minGamma) {
        mask.maskIdxs.assign(mask.N, {});
        for (unsigned i = 0; i < mask.N; i++)
            for (unsigned j = 0; j < mask.dim; j++)
                if (mask.gammas[i][j] > minGamma) mask.maskIdxs[i].push_back(j);
    }

    // Adds the statistics of the N frames into gmms (allocating it for the
    // model's components if it is not allocated yet).
    void accumulateStatsMT(float **vectors, unsigned int NSamples, unsigned int dim, GMMStats<TS> &gmms) {
        const Packed &p = current();
        if (!flagsSet_) {
            fMixProbOnly_ = false;
            fMean_ = gmms.isAllocated() ? gmms.hasMeans() : true;
            fVar_ = gmms.isAllocated() ? gmms.hasVars() : false;
            fAux_ = gmms.isAllocated() ? gmms.hasAux() : false;
        }
        if (!gmms.isAllocated()) {
            gmms.alloc(p.D, p.Malloc, fMean_, fVar_, p.full, fAux_);
            gmms.reset();
        }
        if ((int)dim != p.D || gmms.getDim() != dim)
            throw std::logic_error("accumulateStatsMT(): Inconsistent feature & statistic dimensions!");
        if ((fMean_ && !gmms.hasMeans()) || (fVar_ && !gmms.hasVars()) || (fAux_ && !gmms.hasAux()))
            throw std::logic_error("accumulateStatsMT(): statistics not allocated for the requested flags");
        // a full-covariance model produces full second-order statistics
        // (the diagonal ones would be their diagonal; not offered); a
        // diagonal model gives full ones on request (setAccFlags' varFull,
        // fMLLR)
        if (fVar_ && p.full && !gmms.hasFullCov())
            throw std::logic_error("accumulateStatsMT(): a full-covariance model needs full statistics");
        if (NSamples == 0) return;
        upload(vectors, NSamples, dim);
        syncOpts();
        MaskScope ms(*this, NSamples);
        const int flags = (fMean_ ? GMMHIP_ACC_MEAN : 0) | (fVar_ ? GMMHIP_ACC_VAR : 0) |
                          (fVar_ && !p.full && gmms.hasFullCov() ? GMMHIP_ACC_FULLVAR : 0);
        const size_t vs = gmms.hasVars() ? gmms.getVarSize() : (size_t)p.D;
        std::vector<double> n(p.M, 0.0), F(fMean_ ? (size_t)p.M * p.D : 0), S(fVar_ ? (size_t)p.M * vs : 0);
        double tot = 0.0;
        uint64_t nf = 0;
        check(gmmhip_accumulate(h_, flags, n.data(), fMean_ ? F.data() : nullptr,
                                fVar_ ? S.data() : nullptr, &tot, &nf));
        TS *gn = gmms.probData();
        TS *gF = gmms.meanData();
        TS *gS = gmms.varData();
        for (int m = 0; m < p.M; m++) {
            gn[m] += (TS)n[m];
            for (int d = 0; d < p.D && fMean_; d++) gF[(size_t)m * p.D + d] += (TS)F[(size_t)m * p.D + d];
            for (size_t d = 0; d < vs && fVar_; d++) gS[(size_t)m * vs + d] += (TS)S[(size_t)m * vs + d];
        }
        if (fAux_) {
            const size_t W = (size_t)p.D + 2;
            std::vector<double> aux((size_t)p.M * W, 0.0);
            check(gmmhip_accumulate_aux(h_, aux.data()));
            TS *gA = gmms.auxData();
            for (size_t i = 0; i < aux.size(); i++) gA[i] += (TS)aux[i];
        }
        ms.fill();
        gmms._totalAccSamples += nf;
        gmms._totLogLike += (TS)tot;
        sameData_ = false;
    }

    // accumulateStatsMT over a file list (OL_GMMStatsEstimator.cpp:641-707):
    // files (prm_type: 0 SVES, 1 HTK, 2 RAW; every dwnsmp-th frame) are
    // grouped until a group exceeds _memoryBuffDataGB (0: one file per
    // group) and every group is accumulated; a file that cannot be opened or
    // holds no frames is skipped with a warning, a malformed one throws.
    // The next group is read by a second thread while the device accumulates
    // the current one.
    void accumulateStatsMT(const std::vector<std::string> &files, GMMStats<TS> &gmms, int prm_type,
                           unsigned int dwnsmp = 1) {
        const unsigned step = dwnsmp > 1 ? dwnsmp : 1;
        accumulateGroups(files.size(), gmms, [&files, prm_type, step](size_t i, Features &f, std::string &name,
                                                                      std::string &err) {
            name = files[i];
            const int rc = LoadFeatures(name.c_str(), prm_type, f, &err);
            if (rc == 0 && step > 1 && f.N > 0) {
                Features k;
                k.D = f.D;
                for (unsigned t = 0; t < f.N; t += step) {
                    k.data.insert(k.data.end(), f.data.begin() + (size_t)t * f.D,
                                  f.data.begin() + (size_t)(t + 1) * f.D);
                    k.N++;
                }
                f = std::move(k);
            }
            return rc;
        });
    }

    // accumulateStatsMT over frame intervals of files
    // (OL_GMMStatsEstimator.cpp:712-780): per file, the frames of its
    // intervals [x0, x1) in sample-period units (LoadFeatureRanges; SVES and
    // RAW take samplePeriod, HTK its header's), grouped and accumulated as
    // above, in file-name order.
    void accumulateStatsMT(const file_2_ranges &file2ranges, GMMStats<TS> &gmms, int prm_type,
                           unsigned int samplePeriod) {
        std::vector<const file_2_ranges::value_type *> items;
        for (const auto &kv : file2ranges) items.push_back(&kv);
        accumulateGroups(items.size(), gmms, [&items, prm_type, samplePeriod](size_t i, Features &f,
                                                                              std::string &name,
                                                                              std::string &err) {
            name = items[i]->first;
            return LoadFeatureRanges(name.c_str(), prm_type, items[i]->second, samplePeriod, f, &err);
        });
    }

    // Per-frame log-likelihoods of the N frames into outLogLikes[N].
    void compVecsLogLikeMT(float **vecs, unsigned int NSamples, unsigned int dim, float *outLogLikes) {
        const Packed &p = current();
        if ((int)dim != p.D) throw std::logic_error("compVecsLogLikeMT(): Inconsistent dimensions!");
        if (NSamples == 0) return;
        upload(vecs, NSamples, dim);
        syncOpts();
        MaskScope ms(*this, NSamples);
        check(gmmhip_loglikes(h_, outLogLikes));
        ms.fill();
        for (unsigned n = 0; n < NSamples; n++)
            if (outLogLikes[n] <= _minLogLike) _fNumStabilityTroubles = true;
        sameData_ = false;
    }

    GMMStatsEstimator_HIP<TS> *getNewInstance() { return new GMMStatsEstimator_HIP<TS>(dev_); }

    gmmhip_t handle() const { return h_; }

private:
    struct Packed {
        int D = 0, M = 0, Malloc = 0;
        bool full = false;
        std::vector<float> w, mu, var;  // var: [M][D] or [M][D][D]
    };
    // Installs the "use" mask for the frames of one call (removed again on
    // exit) and copies the "fill" outputs after it.
    struct MaskScope {
        GMMStatsEstimator_HIP &e;
        unsigned int N;
        bool installed = false;
        MaskScope(GMMStatsEstimator_HIP &est, unsigned int n) : e(est), N(n) {
            if (!e.mask_) return;
            if (e.mask_->N < N || (e.useMask_ && e.mask_->maskIdxs.size() < N))
                throw std::logic_error("accumulateStatsMT(): logLikeMask improperly allocated!");
            if (!e.useMask_) return;
            const Packed &p = e.current();
            std::vector<uint64_t> off(N + 1, 0);
            std::vector<int> idx;
            for (unsigned i = 0; i < N; i++) {
                for (unsigned m : e.mask_->maskIdxs[i])
                    if ((int)m < p.M) idx.push_back((int)m);  // inactive components never count
                off[i + 1] = idx.size();
            }
            if (idx.empty()) idx.push_back(0);
            e.check(gmmhip_set_mask(e.h_, N, off.data(), idx.data()));
            installed = true;
        }
        void fill() {
            if (!e.mask_ || !e.fillMask_) return;
            const Packed &p = e.current();
            if (e.mask_->dim < (unsigned)p.Malloc)
                throw std::logic_error("accumulateStatsMT(): logLikeMask improperly allocated!");
            std::vector<float> post((size_t)N * p.M);
            e.check(gmmhip_posteriors(e.h_, e.mask_->logLikes, post.data()));
            for (unsigned i = 0; i < N; i++) {
                std::copy(post.begin() + (size_t)i * p.M, post.begin() + (size_t)(i + 1) * p.M,
                          e.mask_->gammas[i]);
                std::fill(e.mask_->gammas[i] + p.M, e.mask_->gammas[i] + e.mask_->dim, 0.0f);
            }
        }
        ~MaskScope() {
            if (installed) gmmhip_set_mask(e.h_, 0, nullptr, nullptr);
        }
    };

    // The grouping loop of the file overloads: load(i, features, name, err)
    // reads item i (0, or -1 to skip with a warning, or -2 to throw).
    template <typename Load>
    void accumulateGroups(size_t count, GMMStats<TS> &gmms, Load load) {
        const double giga = 1024.0 * 1024.0 * 1024.0;
        size_t next = 0;
        auto load_group = [&load, count, giga, this](size_t &pos) {
            Features g;
            while (pos < count) {
                Features f;
                std::string err, name;
                const int rc = load(pos++, f, name, err);
                if (rc == -1) {
                    std::cerr << "\n WARNING: Param file [" << name << "] not found -> skipped!" << std::endl;
                    continue;
                }
                if (rc != 0) throw std::runtime_error(err);
                if (f.N < 1) {
                    std::cerr << "\n WARNING: Param file [" << name << "] empty" << std::endl;
                    continue;
                }
                if (g.D && f.D != g.D)
                    throw std::logic_error("accumulateStatsMT(): feature dimensions differ across files");
                g.D = f.D;
                g.data.insert(g.data.end(), f.data.begin(), f.data.end());
                g.N += f.N;
                if (!(_memoryBuffDataGB > (double)g.N * g.D * sizeof(float) / giga)) break;
            }
            return g;
        };
        std::future<Features> pending = std::async(std::launch::async, load_group, std::ref(next));
        for (;;) {
            Features cur = pending.get();
            if (cur.N == 0) break;
            if (next < count)
                pending = std::async(std::launch::async, load_group, std::ref(next));
            else
                pending = std::async(std::launch::deferred, [] { return Features(); });
            cur.rows.resize(cur.N);
            for (unsigned i = 0; i < cur.N; i++) cur.rows[i] = &cur.data[(size_t)i * cur.D];
            sameData_ = false;
            accumulateStatsMT(cur.rows.data(), cur.N, cur.D, gmms);
        }
    }

    gmmhip_t h_ = nullptr;
    int dev_ = 0;
    std::string type_ = "GPU HIP";
    std::vector<Packed> group_;
    size_t cur_ = 0;
    bool flagsSet_ = false, fMixProbOnly_ = false, fMean_ = true, fVar_ = false, fFullVar_ = false,
         fAux_ = false;
    bool sameData_ = false;
    bool haveFrames_ = false;
    bool fillMask_ = false, useMask_ = false;
    logLikeMask *mask_ = nullptr;

    const Packed &current() const {
        if (group_.empty()) throw std::logic_error("accumulateStatsMT(): None model inserted!");
        return group_[cur_];
    }
    void select(size_t n) {
        const Packed &p = group_[n];
        syncOpts();
        if (p.full)
            check(gmmhip_set_model_full(h_, p.M, p.D, p.w.data(), p.mu.data(), p.var.data()));
        else
            check(gmmhip_set_model(h_, p.M, p.D, p.w.data(), p.mu.data(), p.var.data()));
        cur_ = n;
    }
    void syncOpts() {
        gmmhip_opts o;
        gmmhip_default_opts(&o);
        o.min_loglike = _minLogLike;
        o.min_gamma = _minGamma;
        o.min_var = _min_var;
        o.precision = _precision;
        o.verbosity = _verbosity;
        check(gmmhip_set_opts(h_, &o));
    }
    void upload(float **vecs, unsigned int N, unsigned int dim) {
        if (sameData_ && haveFrames_) return;
        check(gmmhip_set_frames(h_, vecs, N, (int)dim));
        haveFrames_ = true;
    }
    void check(int rc) const {
        if (rc == GMMHIP_OK) return;
        const std::string msg = std::string(h_ ? gmmhip_last_error(h_) : "") + " [gmmhip status " +
                                std::to_string(rc) + "]";
        if (rc == GMMHIP_ERR_ARG || rc == GMMHIP_ERR_STATE) throw std::logic_error(msg);
        throw std::runtime_error(msg);
    }
};

}  // namespace gmm
