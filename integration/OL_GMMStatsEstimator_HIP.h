// GMMStatsEstimator_HIP<TS>: the gfx950 accumulator as a GMMTeam/GMM
// estimator.  This file belongs to the reference tree: it is compiled there,
// next to trainer/OL_GMMStatsEstimator_CUDA.h (GMM/lib_bmssr/trainer/
// OL_GMMStatsEstimator_CUDA.h:39-56), against the reference's own GMModel,
// GMMStats and GMMStatsEstimator_SSE, and links libgmmhip.so through the C
// ABI of include/gmmhip.h.  It needs the reference's dependencies (Boost
// threads for the SSE base class, ACML for the rest of the trainer library),
// so it builds only inside the reference tree.
//
// Overridden, as the CUDA estimator overrides them: insertModel,
// setModelToBeUsed, accumulateStatsMT(float**, ...), compVecsLogLikeMT and
// getNewInstance.  Inherited unchanged: the file-list and file-to-ranges
// accumulateStatsMT overloads (OL_GMMStatsEstimator.cpp:641-780), which load
// frames and call the virtual accumulateStatsMT(float**, ...) above, so they
// run on the device too; setAccFlags / resetAccFlags / setSameDataFlag; the
// mask calls (mask2fill / mask2use / withdrawMask).
//
// Differences from GMMStatsEstimator_CUDA, all deliberate:
//  - no SSE fallback: the constructor throws without a HIP device and every
//    frame count N >= 1 runs on the device (the CUDA estimator switches to
//    SSE below 8 frames per GPU, OL_GMMStatsEstimator_CUDA.cpp:344, :402);
//  - one device per estimator (GPU_id < 0 takes device 0; the CUDA estimator
//    splits the frames over every visible GPU when GPU_id is -1);
//  - the logLikeMask works: "fill" receives every frame's log-likelihood and
//    its posteriors before the cut, "use" restricts each frame to its
//    component list (the CUDA estimator throws on fill and ignores use,
//    OL_GMMStatsEstimator_CUDA.cpp:352-353, :421-422);
//  - statistics are summed in FP64 on the device and added into GMMStats<TS>
//    once per call (the CUDA estimator sums in FP32).
//
// One line has to be added to model/OL_GMMStats.h beside the CUDA friend
// (model/OL_GMMStats.h:116), because the two running totals are private:
//     template <typename TS> friend class GMMStatsEstimator_HIP;
#ifndef _OL_STATS_EST_HIP_
#define _OL_STATS_EST_HIP_

#include "trainer/OL_GMMStatsEstimator.h"
#include "trainer/OL_GMMStatsEstimator_SSE.h"

#include "gmmhip.h"

#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

template <typename TS>
class GMMStatsEstimator_HIP : public GMMStatsEstimator_SSE<TS, float> {
    typedef GMMStatsEstimator_SSE<TS, float> SSE;
    typedef GMMStatsEstimator<TS, float, float> Base;

public:
    // E-step arithmetic (GMMHIP_PREC_*): FP32 by default, the precision of
    // the CUDA estimator; the split-operand forms trade ~1e-6 relative
    // posteriors for speed.
    int _hipPrecision;

    GMMStatsEstimator_HIP(int GPU_id = -1) : SSE(), _hipPrecision(GMMHIP_PREC_FP32), h_(NULL), dev_(0),
                                             framesN_(0), framesD_(0) {
        int count = 0;
        if (gmmhip_device_count(&count) != GMMHIP_OK || count == 0)
            throw std::runtime_error("GMMStatsEstimator_HIP(): no HIP device found");
        dev_ = (GPU_id >= 0 && GPU_id < count) ? GPU_id : 0;
        gmmhip_opts o;
        options(o);
        check(gmmhip_create(dev_, &o, &h_), "GMMStatsEstimator_HIP()");
        this->_estimType = std::string("GPU HIP");
    }

    virtual ~GMMStatsEstimator_HIP() {
        if (h_ != NULL) gmmhip_destroy(h_);
    }

    // The base class computes the constants, moves the zero-weight components
    // to the end (GMModel::RearrangeMixtures) and keeps the model in its
    // group; the active components (_nummix of them) go to the device.
    virtual void insertModel(GMModel &model, bool first = true) {
        SSE::insertModel(model, first);
        upload_model();
    }

    // The device holds one model at a time: switching re-uploads it
    // (a few hundred KB for 2048 x 60).
    virtual void setModelToBeUsed(unsigned int n) {
        SSE::setModelToBeUsed(n);
        upload_model();
    }

    using SSE::accumulateStatsMT;

    // Adds the statistics of the NSamples frames into gmms: occupancies,
    // first- and second-order sums (diagonal, or packed upper triangles for
    // full statistics), robust auxiliary statistics, the frame count and the
    // total log-likelihood.  What is accumulated follows setAccFlags or, if
    // it was not called, the allocation of gmms (setAccFlags(gmms)).
    virtual void accumulateStatsMT(float **vectors, unsigned int NSamples, unsigned int dim,
                                   GMMStats<TS> &gmms) {
        if (!this->_modelIn) throw std::logic_error("accumulateStatsMT(): None model inserted!");
        if (!this->_flagsSet) this->setAccFlags(gmms);
        this->checkProperAllocation(gmms);
        if (dim != gmms.getDim() || dim != this->_dim)
            throw std::logic_error("accumulateStatsMT(): Inconsistent feature & statistic dimensions!");
        if (this->_mask != NULL && (this->_mask->N < NSamples || this->_mask->dim < this->_nummixAlloc))
            throw std::logic_error("accumulateStatsMT(): logLikeMask improperly allocated! \n\t");
        if (NSamples == 0) {
            this->_fSameDataProvided = false;
            return;
        }
        frames(vectors, NSamples, dim);
        sync_options();
        MaskScope scope(*this, NSamples);

        const unsigned M = this->_nummix, D = this->_dim;
        const bool full_model = this->_model->GetFullCovStatus();
        const bool any = !this->_fMixProbOnly;
        const bool fmean = any && this->_fMean;
        const bool fvar = any && this->_fVar;
        const bool ffull = fvar && this->_fFullVar;
        // the device's second-order layout: D per component, or the packed
        // upper triangle D(D+1)/2 (full model, or full statistics requested)
        const bool packed = fvar && (full_model || ffull);
        const size_t vs = packed ? (size_t)D * (D + 1) / 2 : (size_t)D;
        int flags = 0;
        if (fmean) flags |= GMMHIP_ACC_MEAN;
        if (fvar) flags |= GMMHIP_ACC_VAR;
        if (ffull && !full_model) flags |= GMMHIP_ACC_FULLVAR;

        std::vector<double> n(M, 0.0), F(fmean ? (size_t)M * D : 0), S(fvar ? (size_t)M * vs : 0);
        double tot = 0.0;
        uint64_t nf = 0;
        const int rc = gmmhip_accumulate(h_, flags, n.data(), fmean ? F.data() : NULL,
                                         fvar ? S.data() : NULL, &tot, &nf);
        if (rc == GMMHIP_ERR_NAN) throw std::runtime_error("accumulateStatsMT(): overall log-likelihood = NaN!");
        check(rc, "accumulateStatsMT()");

        bool aM, aV, aFV, aA;
        gmms.getAllocStatus(aM, aV, aFV, aA);
        for (unsigned m = 0; m < M; m++) {
            typename GMMStats<TS>::mixStats &ms = gmms[m];
            ms.mixProb += (TS)n[m];
            if (fmean)
                for (unsigned d = 0; d < D; d++) ms.mean[d] += (TS)F[(size_t)m * D + d];
            if (!fvar) continue;
            const double *s = &S[(size_t)m * vs];
            if (ffull) {
                // packed upper triangle, row d holding columns d..D-1, as
                // GMMStats stores full second-order statistics
                for (size_t k = 0; k < vs; k++) ms.var[k] += (TS)s[k];
            } else {
                // diagonal statistics: into the diagonal of a packed block
                // when gmms holds full ones (OL_GMMStatsEstimator_CUDA.cpp:
                // 623-629), else one per dimension
                size_t diag = 0;  // index of (d, d) in the packed triangle
                for (unsigned d = 0; d < D; d++) {
                    const double v = packed ? s[diag] : s[d];
                    ms.var[aFV ? diag : d] += (TS)v;
                    diag += D - d;
                }
            }
        }
        if (any && this->_fAux) {
            const size_t W = (size_t)D + 2;
            std::vector<double> aux((size_t)M * W, 0.0);
            check(gmmhip_accumulate_aux(h_, aux.data()), "accumulateStatsMT()");
            for (unsigned m = 0; m < M; m++) {
                typename GMMStats<TS>::mixStats &ms = gmms[m];
                for (unsigned d = 0; d < D; d++) ms.aux[d] += (TS)aux[m * W + d];
                ms.aux2 += (TS)aux[m * W + D];
                ms.aux3 += (TS)aux[m * W + D + 1];
            }
        }
        scope.fill();
        gmms._totalAccSamples += (unsigned int)nf;
        gmms._totLogLike += (TS)tot;
        this->_fSameDataProvided = false;
    }

    // Per-frame log-likelihoods of the NSamples frames into outLogLikes.
    virtual void compVecsLogLikeMT(float **vecs, unsigned int NSamples, unsigned int dim,
                                   float *outLogLikes) {
        if (!this->_modelIn) throw std::logic_error("compVecsLogLikeMT(): None model inserted!");
        if (dim != this->_dim) throw std::logic_error("compVecsLogLikeMT(): Inconsistent dimensions!");
        if (this->_mask != NULL && (this->_mask->N < NSamples || this->_mask->dim < this->_nummixAlloc))
            throw std::logic_error("compVecsLogLikeMT(): logLikeMask improperly allocated! \n\t");
        if (NSamples == 0) {
            this->_fSameDataProvided = false;
            return;
        }
        frames(vecs, NSamples, dim);
        sync_options();
        MaskScope scope(*this, NSamples);
        check(gmmhip_loglikes(h_, outLogLikes), "compVecsLogLikeMT()");
        scope.fill();
        double sum = 0.0;
        for (unsigned t = 0; t < NSamples; t++) {
            sum += outLogLikes[t];
            if (outLogLikes[t] <= this->_minLogLike) this->_fNumStabilityTroubles = true;
        }
        if (std::isnan(sum)) throw std::runtime_error("compVecsLogLikeMT(): overall log-likelihood = NaN!");
        this->_fSameDataProvided = false;
    }

    virtual Base *getNewInstance() {
        GMMStatsEstimator_HIP<TS> *e = new GMMStatsEstimator_HIP<TS>(dev_);
        e->_hipPrecision = _hipPrecision;
        return e;
    }

    // the library handle, for calls this class does not wrap
    // (gmmhip_batch_adapt, gmmhip_kernel_times, ...)
    gmmhip_t handle() const { return h_; }

protected:
    // Installs the "use" mask for the frames of one call and removes it on
    // exit; fill() copies the "fill" outputs (log-likelihood and _nummix
    // posteriors per frame, as the CPU estimators store them,
    // OL_GMMStatsEstimator.cpp:1168-1171).
    struct MaskScope {
        GMMStatsEstimator_HIP &e;
        unsigned int N;
        bool installed;
        MaskScope(GMMStatsEstimator_HIP &est, unsigned int n) : e(est), N(n), installed(false) {
            if (e._mask == NULL || !e._useMask) return;
            if (e._mask->maskIdxs.size() < N)
                throw std::logic_error("accumulateStatsMT(): logLikeMask improperly allocated! \n\t");
            std::vector<uint64_t> off(N + 1, 0);
            std::vector<int> idx;
            for (unsigned t = 0; t < N; t++) {
                const std::vector<unsigned int> &l = e._mask->maskIdxs[t];
                for (size_t k = 0; k < l.size(); k++)
                    if (l[k] < e._nummix) idx.push_back((int)l[k]);  // zero-weight components never count
                off[t + 1] = idx.size();
            }
            if (idx.empty()) idx.push_back(0);
            e.check(gmmhip_set_mask(e.h_, N, &off[0], &idx[0]), "mask2use");
            installed = true;
        }
        void fill() {
            if (e._mask == NULL || !e._fillMask) return;
            const unsigned M = e._nummix;
            std::vector<float> post((size_t)N * M);
            e.check(gmmhip_posteriors(e.h_, e._mask->logLikes, &post[0]), "mask2fill");
            for (unsigned t = 0; t < N; t++)
                for (unsigned m = 0; m < M; m++) e._mask->gammas[t][m] = post[(size_t)t * M + m];
        }
        ~MaskScope() {
            if (installed) gmmhip_set_mask(e.h_, 0, NULL, NULL);
        }
    };

    gmmhip_t h_;
    int dev_;
    uint64_t framesN_;
    unsigned int framesD_;

    void options(gmmhip_opts &o) const {
        gmmhip_default_opts(&o);
        o.min_loglike = this->_minLogLike;
        o.min_gamma = this->_minGamma;
        o.min_var = this->_min_var;
        o.precision = _hipPrecision;
        o.verbosity = this->_verbosity > 3 ? 1 : 0;
    }
    // the public knobs may change between calls (updateOptionsOnGPU,
    // OL_GMMStatsEstimator_CUDA.cpp:114-123)
    void sync_options() {
        gmmhip_opts o;
        options(o);
        check(gmmhip_set_opts(h_, &o), "options");
    }

    // Active components of the current model, gathered into row-major arrays.
    void upload_model() {
        GMModel &g = *this->_model;
        const unsigned M = this->_nummix, D = this->_dim;
        const bool full = g.GetFullCovStatus();
        const size_t dv = full ? (size_t)D * D : (size_t)D;
        std::vector<float> w(M), mu((size_t)M * D), var((size_t)M * dv);
        for (unsigned m = 0; m < M; m++) {
            w[m] = g.GetMixtureWeight(m);
            const float *a = g.GetMixtureMean(m);
            const float *v = g.GetMixtureVar(m);
            for (unsigned d = 0; d < D; d++) mu[(size_t)m * D + d] = a[d];
            for (size_t k = 0; k < dv; k++) var[m * dv + k] = v[k];
        }
        sync_options();
        this->_modelIn = false;
        if (full)
            check(gmmhip_set_model_full(h_, (int)M, (int)D, &w[0], &mu[0], &var[0]), "insertModel()");
        else
            check(gmmhip_set_model(h_, (int)M, (int)D, &w[0], &mu[0], &var[0]), "insertModel()");
        this->_modelIn = true;
    }

    // Frames stay resident between calls; setSameDataFlag() skips the upload
    // when the count and dimension match the resident set.
    void frames(float **vecs, unsigned int N, unsigned int dim) {
        if (this->_fSameDataProvided && framesN_ == N && framesD_ == dim) return;
        check(gmmhip_set_frames(h_, vecs, N, (int)dim), "upload");
        framesN_ = N;
        framesD_ = dim;
    }

    void check(int rc, const char *where) const {
        if (rc == GMMHIP_OK) return;
        std::string msg = std::string(where) + ": " + (h_ != NULL ? gmmhip_last_error(h_) : "") +
                          " [gmmhip status " + std::to_string(rc) + "]";
        if (rc == GMMHIP_ERR_ARG || rc == GMMHIP_ERR_STATE) throw std::logic_error(msg);
        throw std::runtime_error(msg);
    }
};

#endif
