// Trainer / adaptation driver over the HIP estimator: the subset of
// CModelAdapt (trainer/OL_ModelAdapt.h:140-180 of GMMTeam/GMM) on the
// GMM hot path -- EM from a UBM or from scratch (Gaussian splitting; diagonal
// or full covariance; optionally robust covariances), MAP adaptation, global
// MLLR (+ MLLR then MAP), fMLLR -- with the reference's update rules.  All statistics come
// from GMMStatsEstimator_HIP.
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "gmm_estimator.h"
#include "gmm_model.h"
#include "gmm_stats.h"

namespace gmm {

enum AdaptType { FROM_SCRATCH = 0, MLLR = 1, MAP = 2, fMLLR = 3, MLLR_MAP = 4 };

// MLLR transform W = [A | b], D x (D+1), row-major FP64.
struct MLLRTransform {
    int D = 0;
    std::vector<double> W;
    std::vector<double> Q;  // per-row auxiliary-function values (optional)
};

class CModelAdapt {
public:
    // CModelAdapt::OPTIONS (OL_ModelAdapt.h:66-138), hot-path subset.
    struct OPTIONS {
        int _adaptType = MAP;
        int _iterCountEM = 0;
        int _iterCountMAP = 0;
        bool _AdaptMeanBool = false;
        bool _AdaptVarBool = false;
        bool _AdaptWeightBool = false;
        int _tau = 0;
        float _min_var = 1e-6f;
        bool _mllrUBMInit = false;
        int _iterMLLR = 1;
        bool _applyTransforms = false;
        bool _saveTransforms = false;
        bool _getQvalMLLR = false;
        bool _saveMixOcc = false;
        int _verbosity = 0;
        int _cudaGPUid = -1;
        unsigned int _numThreads = 1;
        int _precision = GMMHIP_PREC_FP32;
        bool _robustCov = false;  // robust covariances from the auxiliary statistics
        int _iter_fMLLR = 20;     // row-update sweeps of fMLLR
        std::string _transformFileExt = ".mat";
        // TrainFromScratch (OL_ModelAdapt.h:69-80, :212-229): split-X, the
        // fraction of the non-zero components split at once per round
        // (those heavier than split-th times the heaviest); ll-Th, stop when
        // the mean log-likelihood gains less than this over the better of
        // the two previous rounds; tmp-N, save a backup model every N rounds.
        float _splitXpercAtOnce = 0.0f;
        float _splitTh = 0.0f;
        float _llTh = 0.0f;
        int _saveModelAfterNSplits = 0;
    };

    CModelAdapt() = default;
    void SetOptions(const OPTIONS &opt);
    const OPTIONS &GetOptions() const { return opt_; }

    // Frames are borrowed (row pointers), as in InsertData.
    void InsertData(float **vectors, unsigned int N, unsigned int D);
    // Feature files read file by file at every accumulation (the reference's
    // InsertData(CFileList) + --file-by-file): prm_type 0 SVES, 1 HTK, 2 RAW,
    // every dwnsmp-th frame, groups of up to buffGB gigabytes.  The files
    // are scanned once here for their frame count and dimension.
    void InsertFiles(const std::vector<std::string> &files, int prm_type, unsigned int dwnsmp,
                     float buffGB);
    // Frame intervals of files, read at every accumulation (the reference's
    // InsertData(file_2_ranges*) of the MLF mode, OL_ModelAdapt.cpp:321-328,
    // accumulated by the file-to-ranges accumulateStatsMT, :440-443).
    void InsertRanges(const file_2_ranges &ranges, int prm_type, unsigned int samplePeriod,
                      float buffGB);

    // MAP / MLLR / MLLR_MAP adaptation of UBM (AdaptModel, :1163-1247).
    void AdaptModel(GMModel *UBM);
    // EM from scratch with binary splitting up to nummix components
    // (TrainFromScratch, :673-821), diagonal or full covariance.  FileName
    // names the backups of --tmp-N (FileName.itNNNNN.GNNNNN.gmm.bckup).
    void TrainFromScratch(unsigned int nummix, bool fullcov = false, const char *FileName = nullptr,
                          bool save_txt = false);

    GMModel *GetClientModel() { return model_.get(); }
    const MLLRTransform *GetTransform() const { return W_.get(); }
    // Saves the adapted model, or the transform when no model was produced
    // or --save-trn was given (Save/SaveTRNMatrices, :1900-1989).
    int Save(const char *filename, bool save_txt);

    // Exposed update rules (reference names), operating on _stats.
    unsigned int UpdateModelEM(GMModel *model, GMMStats<double> &stats, bool upweight = true,
                               bool upmean = true, bool upvar = true);
    void UpdateModelMAP(GMModel *model, GMModel *UBM);
    void AccumulateStats(GMModel *model, GMMStats<double> &stats, bool mixProbsOnly = false,
                         bool meanStats = true, bool varStats = false, bool auxStats = false);

    // Robust covariance pieces (OL_ModelAdapt.cpp:468-588): independent
    // frames per component from the auxiliary statistics, the covariance
    // multipliers, and the in-place rewrite of the second-order statistics.
    static std::vector<double> GetNIndFrames(const GMMStats<double> &stats);
    static float getDiagCovMultiplier(double nind);
    static float getOffDiagCovMultiplier(int D, double nind);
    static std::vector<double> MakeRobustStats(GMMStats<double> &stats);
    GMMStats<double> &Stats() { return stats_; }
    void SetNumMix(int M) { nummix_ = M; }

    // MLLR pieces (OL_ModelAdapt.cpp:1417-1612).
    void CheckInvertibility() const;
    static void prepareKiGiInitMatrices(const GMModel &UBM, float min_var, std::vector<double> &G,
                                        std::vector<double> &k);
    void ComputeKiGi(const GMModel &model, int row, const double *Ginit, const double *kinit,
                     double *G, double *k) const;
    MLLRTransform ComputeW_MLLR(GMModel *model, const std::vector<double> &Ginit,
                                const std::vector<double> &kinit, bool wantQ);
    static std::unique_ptr<GMModel> AdaptModelMLLR(const GMModel &model, const MLLRTransform &W);
    // fMLLR (OL_ModelAdapt.cpp:1614-1840): statistics with full second-order
    // sums under model, then the row-by-row constrained MLLR estimate.
    MLLRTransform ComputeW_fMLLR(GMModel *model);
    // fMLLR rows from given statistics (full second-order, packed upper).
    static MLLRTransform EstimateW_fMLLR(const GMModel &model, const GMMStats<double> &stats,
                                         int iters);
    static void ComposeMLLR(MLLRTransform &Wsum, const MLLRTransform &Wnew);

private:
    OPTIONS opt_;
    float **data_ = nullptr;
    unsigned int N_ = 0, D_ = 0;
    std::vector<std::string> files_;  // file-by-file mode when non-empty
    file_2_ranges ranges_;            // file-by-file interval mode when non-empty
    unsigned int samplePeriod_ = 0;
    int fileType_ = 2;
    unsigned int dwnsmp_ = 1;
    float buffGB_ = 0.0f;
    bool dataChanged_ = true;
    std::unique_ptr<GMMStatsEstimator_HIP<double>> est_;
    GMMStats<double> stats_;
    std::unique_ptr<GMModel> model_;
    std::unique_ptr<MLLRTransform> W_;
    GMModel *UBM_ = nullptr;
    int nummix_ = 0;

    void initEstimator();
    bool hasData() const { return data_ || !files_.empty() || !ranges_.empty(); }
    std::unique_ptr<GMModel> AdaptModelMAP(GMModel *ubm);
    void IterEM(GMModel *model, int iters);
    void IterMAP(GMModel *model, int iters);
    MLLRTransform IterW_MLLR(GMModel *UBM, int iters, bool wantQ);
};

// FP64 Cholesky solve of the symmetric positive definite system G w = k
// (n x n, row-major).  Throws std::runtime_error if G is not SPD.
void spd_solve(int n, const double *G, const double *k, double *w);

}  // namespace gmm
