#include "model_adapt.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <limits>
#include <stdexcept>

namespace gmm {

namespace {
constexpr double MIN_IND_SAMPLES = 2.0;  // OL_ModelAdapt.h:54
}

void CModelAdapt::SetOptions(const OPTIONS &opt) {
    opt_ = opt;
    if (opt_._min_var < 1.17549435e-38f) opt_._min_var = 1.17549435e-38f;  // CheckOptions :2141-2143
    // CheckOptions :2191-2212: thresholds clamped into range
    if (opt_._llTh < 0.0f) opt_._llTh = 0.0f;
    opt_._splitXpercAtOnce = std::min(1.0f, std::max(0.0f, opt_._splitXpercAtOnce));
    opt_._splitTh = std::min(1.0f, std::max(0.0f, opt_._splitTh));
}

void CModelAdapt::InsertData(float **vectors, unsigned int N, unsigned int D) {
    data_ = vectors;
    N_ = N;
    D_ = D;
    files_.clear();
    ranges_.clear();
    dataChanged_ = true;
}

void CModelAdapt::InsertFiles(const std::vector<std::string> &files, int prm_type,
                              unsigned int dwnsmp, float buffGB) {
    unsigned int N = 0, D = 0;
    for (const std::string &fn : files) {
        Features f;
        std::string err;
        const int rc = LoadFeatures(fn.c_str(), prm_type, f, &err);
        if (rc == -1 || (rc == 0 && f.N == 0)) continue;
        if (rc != 0) throw std::runtime_error(err);
        if (D && f.D != D) throw std::logic_error("InsertData(): feature dimensions differ across files");
        D = f.D;
        N += dwnsmp > 1 ? (f.N + dwnsmp - 1) / dwnsmp : f.N;
    }
    if (N == 0) throw std::runtime_error("InsertData(): no feature vectors in the given files");
    ranges_.clear();
    files_ = files;
    fileType_ = prm_type;
    dwnsmp_ = dwnsmp;
    buffGB_ = buffGB;
    data_ = nullptr;
    N_ = N;
    D_ = D;
    dataChanged_ = true;
}

void CModelAdapt::InsertRanges(const file_2_ranges &ranges, int prm_type, unsigned int samplePeriod,
                               float buffGB) {
    unsigned int N = 0, D = 0;
    for (const auto &kv : ranges) {
        Features f;
        std::string err;
        const int rc = LoadFeatureRanges(kv.first.c_str(), prm_type, kv.second, samplePeriod, f, &err);
        if (rc == -1 || (rc == 0 && f.N == 0)) continue;
        if (rc != 0) throw std::runtime_error(err);
        if (D && f.D != D) throw std::logic_error("InsertData(): feature dimensions differ across files");
        D = f.D;
        N += f.N;
    }
    if (N == 0) throw std::runtime_error("InsertData(): no feature vectors in the given intervals");
    files_.clear();
    ranges_ = ranges;
    fileType_ = prm_type;
    samplePeriod_ = samplePeriod;
    buffGB_ = buffGB;
    data_ = nullptr;
    N_ = N;
    D_ = D;
    dataChanged_ = true;
}

void CModelAdapt::initEstimator() {
    if (!est_) {
        est_.reset(new GMMStatsEstimator_HIP<double>(opt_._cudaGPUid));
        est_->_min_var = opt_._min_var;
        est_->_precision = opt_._precision;
    }
    est_->_verbosity = opt_._verbosity;
    est_->_numThreads = opt_._numThreads;
}

// AccumulateStats (OL_ModelAdapt.cpp:411-466): insert the model (which
// rearranges it), set the flags and add the statistics of the inserted data.
void CModelAdapt::AccumulateStats(GMModel *model, GMMStats<double> &stats, bool mixProbsOnly,
                                  bool meanStats, bool varStats, bool auxStats) {
    if (!hasData() || N_ == 0) throw std::logic_error("AccumulateStats(): Insert data first!");
    initEstimator();
    if (model) est_->insertModel(*model);
    est_->setAccFlags(mixProbsOnly, meanStats, varStats, stats.hasFullCov(), auxStats);
    if (!files_.empty() || !ranges_.empty()) {
        est_->_memoryBuffDataGB = buffGB_;
        if (!files_.empty())
            est_->accumulateStatsMT(files_, stats, fileType_, dwnsmp_);
        else
            est_->accumulateStatsMT(ranges_, stats, fileType_, samplePeriod_);
        dataChanged_ = true;
        return;
    }
    if (!dataChanged_) est_->setSameDataFlag();
    est_->accumulateStatsMT(data_, N_, D_, stats);
    dataChanged_ = false;
    if (opt_._verbosity)
        std::cout << "[#frames accumulated = " << stats.getTotAccSamples() << "]" << std::endl;
}

// GetNIndFrames (OL_ModelAdapt.cpp:468-535): per component
//   MII = (1/D) sum_d aux[d] / aux3 / std_d   (std: global deviation from the
//         summed statistics)
//   nind = 1 + (aux2 N_glob / sum aux2 - 1) (MII / 1.12)^3,
// 0 for a component with n = 0 or aux3 = 0.
std::vector<double> CModelAdapt::GetNIndFrames(const GMMStats<double> &stats) {
    const unsigned D = stats.getDim(), M = stats.getMixNum();
    if (!stats.hasAux() || !stats.hasMeans() || !stats.hasVars())
        throw std::logic_error("GetNIndFrames(): robust estimation needs auxiliary, first- and "
                               "second-order statistics");
    const bool full = stats.hasFullCov();
    std::vector<double> gstd(D, 0.0), gmean(D, 0.0), nind(M, 0.0);
    double gprob = 0.0, sumNg = 0.0;
    for (unsigned m = 0; m < M; m++) {
        gprob += stats.getProb(m);
        const double *F = stats.getMean(m), *S = stats.getVar(m);
        for (unsigned d = 0; d < D; d++) {
            gmean[d] += F[d];
            gstd[d] += S[full ? (size_t)d * D + d - (size_t)d * (d + 1) / 2 : d];
        }
        sumNg += stats.getAux2(m);
    }
    for (unsigned d = 0; d < D; d++) {
        gmean[d] /= gprob;
        gstd[d] = std::sqrt(gstd[d] / gprob - gmean[d] * gmean[d]);
        if (gstd[d] == 0)
            throw std::runtime_error("GetNIndFrames(): global variance equal zero!");
    }
    for (unsigned m = 0; m < M; m++) {
        const double *a = stats.getAux(m);
        const double a3 = stats.getAux3(m);
        if (stats.getProb(m) == 0 || a3 == 0) continue;
        double v = 0.0;
        for (unsigned d = 0; d < D; d++) v += a[d] / a3 / gstd[d];
        v /= D;
        v = std::pow(v / 1.12, 3);
        nind[m] = 1 + (stats.getAux2(m) * gprob / sumNg - 1) * v;
    }
    return nind;
}

// getDiagCovMultiplier / getOffDiagCovMultiplier (OL_ModelAdapt.cpp:537-548).
float CModelAdapt::getDiagCovMultiplier(double nind) {
    if (nind < 2) return (float)std::pow(std::pow(1.0 + 1.0 / (2.0 - 1.25), 3.5), 3 - nind);
    return (float)std::pow(1.0 + 1.0 / (nind - 1.25), 3.5);
}

float CModelAdapt::getOffDiagCovMultiplier(int D, double nind) {
    if (nind < 2) return 0.0f;
    return (float)(1 - std::pow(((double)D / ((double)D - 1 + nind)), 1.4));
}

// MakeRobustStats (OL_ModelAdapt.cpp:550-588): S <- dc S + (1 - dc) F_k F_kk / n
// on the diagonal (dc = diagonal multiplier x nind / (nind - 1), float) and
// with the off-diagonal multiplier on the off-diagonal entries; (1 - dc) is
// a float difference, as in the reference.  Returns the independent-frame
// counts.
std::vector<double> CModelAdapt::MakeRobustStats(GMMStats<double> &stats) {
    std::vector<double> nind = GetNIndFrames(stats);
    const int D = (int)stats.getDim();
    const size_t VS = stats.getVarSize();
    for (unsigned j = 0; j < stats.getMixNum(); j++) {
        float dc = getDiagCovMultiplier(nind[j]);
        dc = (float)((double)dc * (nind[j] / (nind[j] - 1)));
        const double *F = stats.getMean(j);
        const double n = stats.getProb(j);
        double *S = stats.varData() + (size_t)j * VS;
        if (stats.hasFullCov()) {
            const float odc = getOffDiagCovMultiplier(D, nind[j]);
            size_t p = 0;
            for (int k = 0; k < D; k++)
                for (int kk = k; kk < D; kk++, p++) {
                    if (kk == k)
                        S[p] = dc * S[p] + (double)(1.0f - dc) * std::pow(F[k], 2.0) / n;
                    else
                        S[p] = odc * S[p] + (double)(1.0f - odc) * F[k] * F[kk] / n;
                }
        } else {
            for (int k = 0; k < D; k++) S[k] = dc * S[k] + (double)(1.0f - dc) * std::pow(F[k], 2.0) / n;
        }
    }
    return nind;
}

// UpdateModelEM (OL_ModelAdapt.cpp:823-949); with _robustCov the statistics
// are first made robust and the independent-frame counts replace the
// occupancies in the minimum-data test.  The full-covariance branch
// (:906-925) sets cov[k][kk] = cov[kk][k] = (float)(S/n - mean_k * mean_kk)
// with the product of the float means taken in float and no floor.
unsigned int CModelAdapt::UpdateModelEM(GMModel *model, GMMStats<double> &stats, bool upweight,
                                        bool upmean, bool upvar) {
    const int D = model->GetDimension();
    const int nummix = std::min<int>(model->GetNumberOfMixtures(), (int)stats.getMixNum());
    std::vector<double> nind;
    if (opt_._robustCov) nind = MakeRobustStats(stats);
    std::vector<float> mean(D);
    bool wchange = false;
    unsigned int nzero = 0;
    double gcorr = 0.0;
    for (int j = 0; j < nummix; j++) {
        if (model->GetMixtureWeight(j) == 0.0f) {
            ++nzero;
            continue;
        }
        const double np = stats.getProb(j);
        if (opt_._saveMixOcc) model->SetMixOcc(j, (float)np);
        const double ni = opt_._robustCov ? nind[j] : np;
        if (ni < MIN_IND_SAMPLES) {
            model->ResetMixture(j);
            model->SetMixWeight(j, 0.0f);
            wchange = true;
            continue;
        }
        if (upweight) {
            model->SetMixWeight(j, (float)(np / (double)stats.getTotAccSamples()));
            wchange = true;
        }
        gcorr += model->GetMixtureWeight(j);
        if (upmean || upvar) {
            const double *F = stats.getMean(j);
            const double *S = stats.getVar(j);
            if (!F || (upvar && !S))
                throw std::logic_error("UpdateModelEM(): statistics not allocated");
            for (int k = 0; k < D; k++) mean[k] = (float)(F[k] / np);
            if (upvar && model->GetFullCovStatus()) {
                if (!stats.hasFullCov())
                    throw std::logic_error("UpdateModelEM(): full-covariance model needs full statistics");
                if (!opt_._robustCov && ni < D) {
                    std::cerr << " [WARNING: covariance matrix for Gaussian with idx=" << j
                              << " is likely to be ill conditioned (not enough data!)]" << std::endl;
                    std::cerr << " [         try to turn on --robust when working with "
                                 "full-covariance matrices]" << std::endl;
                }
                float *cov = model->GetMixtureVar(j);
                size_t p = 0;
                for (int k = 0; k < D; k++)
                    for (int kk = k; kk < D; kk++, p++) {
                        const float mm = mean[k] * mean[kk];
                        const float v = (float)(S[p] / np - (double)mm);
                        cov[(size_t)k * D + kk] = v;
                        cov[(size_t)kk * D + k] = v;
                    }
            } else if (upvar) {
                float *var = model->GetMixtureVar(j);
                for (int k = 0; k < D; k++) {
                    float v = (float)(S[k] / np - std::pow((double)mean[k], 2));
                    if (v < opt_._min_var) v = opt_._min_var;
                    var[k] = v;
                }
            }
            if (upmean) std::memcpy(model->GetMixtureMean(j), mean.data(), sizeof(float) * D);
        }
    }
    if (wchange && gcorr > 0.0)
        for (int j = 0; j < nummix; j++)
            model->SetMixWeight(j, (float)(model->GetMixtureWeight(j) / gcorr));
    return (unsigned int)(nummix - (int)nzero);
}

// UpdateModelMAP (OL_ModelAdapt.cpp:993-1106), non-robust.  Full
// covariance (:1079-1085): cov[k][kk] = cov[kk][k] = (float)(alfa S/n +
// (1-alfa)(ubm cov[k][kk] + ubm mu_k ubm mu_kk) - mu_k mu_kk), products of
// float means taken in float, no floor.
void CModelAdapt::UpdateModelMAP(GMModel *model, GMModel *UBM) {
    if (opt_._tau == 0) {
        if (UpdateModelEM(model, stats_, opt_._AdaptWeightBool, opt_._AdaptMeanBool,
                          opt_._AdaptVarBool) == 0) {
            std::cerr << "\n [WARNING: All Gaussians discarded - keeping the initial/global model]"
                      << std::endl;
            *model = *UBM;
        }
        return;
    }
    if (opt_._robustCov) MakeRobustStats(stats_);  // :1037-1040
    const int nummix = model->GetNumberOfMixtures();
    const int D = model->GetDimension();
    const bool full = stats_.hasVars() && stats_.hasFullCov();
    if (opt_._AdaptVarBool && (full != UBM->GetFullCovStatus() || full != model->GetFullCovStatus()))
        throw std::logic_error("UpdateModelMAP(): full/diagonal statistics do not match the models");
    std::vector<float> am(D), av((size_t)model->GetDimVar());
    double mixcorr = 0.0;
    for (int j = 0; j < nummix; j++) {
        if (UBM->GetMixtureWeight(j) == 0) continue;
        const double np = stats_.getProb(j);
        if (opt_._saveMixOcc) model->SetMixOcc(j, (float)np);
        if (!(np > 0.0)) continue;
        const float *um = UBM->GetMixtureMean(j);
        const float *uv = UBM->GetMixtureVar(j);
        const double alfa = np / (np + opt_._tau);
        if (opt_._AdaptWeightBool && stats_.getTotAccSamples() > 0) {
            const float aw = (float)(alfa * (np / (double)stats_.getTotAccSamples()) +
                                     (1 - alfa) * UBM->GetMixtureWeight(j));
            model->SetMixWeight(j, aw);
            mixcorr += aw;
        }
        if (opt_._AdaptMeanBool || opt_._AdaptVarBool) {
            const double *F = stats_.getMean(j);
            const double *S = stats_.getVar(j);
            for (int k = 0; k < D; k++) am[k] = (float)(alfa * F[k] / np + (1 - alfa) * um[k]);
            if (opt_._AdaptVarBool && !full) {
                for (int k = 0; k < D; k++) {
                    av[k] = (float)(alfa * S[k] / np + (1 - alfa) * (uv[k] + um[k] * um[k]) -
                                    am[k] * am[k]);
                    if (av[k] < opt_._min_var) av[k] = opt_._min_var;
                }
                std::memcpy(model->GetMixtureVar(j), av.data(), sizeof(float) * D);
            } else if (opt_._AdaptVarBool) {
                size_t p = 0;  // packed upper index of (k, kk)
                for (int k = 0; k < D; k++)
                    for (int kk = k; kk < D; kk++, p++) {
                        const float v = (float)(alfa * S[p] / np +
                                                (1 - alfa) * (uv[(size_t)k * D + kk] + um[k] * um[kk]) -
                                                am[k] * am[kk]);
                        av[(size_t)k * D + kk] = v;
                        av[(size_t)kk * D + k] = v;
                    }
                std::memcpy(model->GetMixtureVar(j), av.data(), sizeof(float) * av.size());
            }
            std::memcpy(model->GetMixtureMean(j), am.data(), sizeof(float) * D);
        }
    }
    if (opt_._AdaptWeightBool && mixcorr != 0)
        for (int j = 0; j < nummix; j++)
            model->SetMixWeight(j, model->GetMixtureWeight(j) / (float)mixcorr);
}

void CModelAdapt::IterEM(GMModel *model, int iters) {
    const bool weightOnly = opt_._AdaptWeightBool && !opt_._AdaptMeanBool && !opt_._AdaptVarBool;
    for (int i = 0; i < iters; i++) {
        stats_.reset();
        AccumulateStats(model, stats_, weightOnly, opt_._AdaptMeanBool, opt_._AdaptVarBool,
                        opt_._robustCov);
        GMModel tmp(*model);
        UpdateModelMAP(model, &tmp);
    }
}

void CModelAdapt::IterMAP(GMModel *model, int iters) {
    const bool weightOnly = opt_._AdaptWeightBool && !opt_._AdaptMeanBool && !opt_._AdaptVarBool;
    for (int i = 0; i < iters; i++) {
        stats_.reset();
        AccumulateStats(model, stats_, weightOnly, opt_._AdaptMeanBool, opt_._AdaptVarBool,
                        opt_._robustCov);
        UpdateModelMAP(model, UBM_);
    }
}

std::unique_ptr<GMModel> CModelAdapt::AdaptModelMAP(GMModel *ubm) {
    std::unique_ptr<GMModel> model(new GMModel(*ubm));
    if (opt_._iterCountEM == 0 && opt_._iterCountMAP == 0)
        std::cerr << "AdaptModel(): WARNING: iterCountEM = iterCountMAP = 0 => nothing to do "
                     "(speaker model == UBM)!"
                  << std::endl;
    if (opt_._iterCountEM) IterEM(model.get(), opt_._iterCountEM);
    if (opt_._iterCountMAP) IterMAP(model.get(), opt_._iterCountMAP);
    return model;
}

// ------------------------------------------------------------------- MLLR
void CModelAdapt::CheckInvertibility() const {
    const int D = (int)D_;
    if (nummix_ < D + 1)
        throw std::runtime_error(
            "CheckInvertibility(): In order to use MLLR, number of Gaussians in the model has to be "
            "greater than the dimension of feature vectors (use MAP adaptation)!");
    int active = 0;
    double occ = 0.0;
    for (int m = 0; m < nummix_; m++)
        if (stats_.getProb(m) > 0) {
            ++active;
            occ += stats_.getProb(m);
        }
    if (active < D + 1)
        throw std::runtime_error(
            "CheckInvertibility(): To ensure invertibility of matrix Gi, count of active mixtures has "
            "to be >= dimension of feature vectors (increase amount of data!)!");
    if (occ * D < (double)D * (D + 1))
        throw std::runtime_error(
            "CheckInvertibility(): Number of parameters to be estimated is higher than the number of "
            "available data (increase amount of data!)!");
}

// prepareKiGiInitMatrices (OL_ModelAdapt.cpp:1484-1523): every UBM component
// contributes one virtual frame at its mean with its variance.
void CModelAdapt::prepareKiGiInitMatrices(const GMModel &UBM, float min_var, std::vector<double> &G,
                                          std::vector<double> &k) {
    const int M = UBM.GetNumberOfMixtures(), D = UBM.GetDimension(), E = D + 1;
    G.assign((size_t)D * E * E, 0.0);
    k.assign((size_t)D * E, 0.0);
    std::vector<float> vd(D);
    for (int m = 0; m < M; m++) {
        // a full-covariance UBM contributes its diagonal (GetMixtureVarDiag, :1502)
        std::memcpy(vd.data(), UBM.GetMixtureVarDiag(m), sizeof(float) * D);
        const float *var = vd.data(), *mean = UBM.GetMixtureMean(m);
        for (int i = 0; i < D; i++) {
            const double cG = (var[i] < min_var) ? 1.0 / min_var : 1.0 / var[i];
            const double ck = mean[i] * cG;
            double *Gi = &G[(size_t)i * E * E], *ki = &k[(size_t)i * E];
            for (int a = 0; a < D; a++) {
                ki[a] += ck * mean[a];
                Gi[(size_t)D * E + a] += cG * mean[a];
                Gi[(size_t)a * E + D] += cG * mean[a];
                for (int b = 0; b < D; b++) Gi[(size_t)a * E + b] += cG * mean[a] * mean[b];
                Gi[(size_t)a * E + a] += cG * var[a];
            }
            Gi[(size_t)E * E - 1] += cG;
            ki[D] += ck;
        }
    }
}

// ComputeKiGi_MLLR (OL_ModelAdapt.cpp:1525-1581) for row i.
void CModelAdapt::ComputeKiGi(const GMModel &model, int i, const double *Ginit, const double *kinit,
                              double *G, double *k) const {
    const int D = model.GetDimension(), E = D + 1;
    if (Ginit && kinit) {
        std::memcpy(G, Ginit, sizeof(double) * E * E);
        std::memcpy(k, kinit, sizeof(double) * E);
    } else {
        std::memset(G, 0, sizeof(double) * E * E);
        std::memset(k, 0, sizeof(double) * E);
    }
    for (int m = 0; m < nummix_; m++) {
        const double np = stats_.getProb(m);
        if (!(np > 0)) continue;
        const float *mm = model.GetMixtureMean(m);
        const double vi = model.GetMixtureVar(m)[model.GetFullCovStatus() ? (size_t)i * D + i : i];
        const double cm = stats_.getMean(m)[i] / vi, cv = np / vi;
        for (int a = 0; a < D; a++) {
            k[a] += cm * mm[a];
            const double cva = cv * mm[a];
            for (int b = a; b < D; b++) G[(size_t)a * E + b] += cva * mm[b];
            G[(size_t)a * E + D] += cva;
        }
        k[D] += cm;
        G[(size_t)D * E + D] += cv;
    }
    for (int a = 0; a < E; a++)
        for (int b = a + 1; b < E; b++) G[(size_t)b * E + a] = G[(size_t)a * E + b];
}

void spd_solve(int n, const double *G, const double *k, double *w) {
    std::vector<double> L((size_t)n * n, 0.0);
    for (int i = 0; i < n; i++) {
        for (int j = 0; j <= i; j++) {
            double s = G[(size_t)i * n + j];
            for (int p = 0; p < j; p++) s -= L[(size_t)i * n + p] * L[(size_t)j * n + p];
            if (i == j) {
                if (!(s > 0.0)) throw std::runtime_error("MLLR: matrix G_i is not positive definite");
                L[(size_t)i * n + i] = std::sqrt(s);
            } else {
                L[(size_t)i * n + j] = s / L[(size_t)j * n + j];
            }
        }
    }
    std::vector<double> y(n);
    for (int i = 0; i < n; i++) {
        double s = k[i];
        for (int p = 0; p < i; p++) s -= L[(size_t)i * n + p] * y[p];
        y[i] = s / L[(size_t)i * n + i];
    }
    for (int i = n - 1; i >= 0; i--) {
        double s = y[i];
        for (int p = i + 1; p < n; p++) s -= L[(size_t)p * n + i] * w[p];
        w[i] = s / L[(size_t)i * n + i];
    }
}

// ComputeW_MLLR (OL_ModelAdapt.cpp:1417-1482): statistics of the model, then
// w_i = G_i^-1 k_i per row (FP64; Cholesky instead of the reference's
// unpivoted Gauss-Jordan inverse, same solution for the SPD G_i).
MLLRTransform CModelAdapt::ComputeW_MLLR(GMModel *model, const std::vector<double> &Ginit,
                                         const std::vector<double> &kinit, bool wantQ) {
    stats_.reset();
    AccumulateStats(model, stats_, false, true, false);
    CheckInvertibility();
    const int D = (int)D_, E = D + 1;
    MLLRTransform W;
    W.D = D;
    W.W.assign((size_t)D * E, 0.0);
    std::vector<double> G((size_t)E * E), k(E), w(E);
    for (int i = 0; i < D; i++) {
        ComputeKiGi(*model, i, Ginit.empty() ? nullptr : &Ginit[(size_t)i * E * E],
                    kinit.empty() ? nullptr : &kinit[(size_t)i * E], G.data(), k.data());
        spd_solve(E, G.data(), k.data(), w.data());
        std::memcpy(&W.W[(size_t)i * E], w.data(), sizeof(double) * E);
        if (wantQ) {
            double Q = 0.0;
            for (int a = 0; a < E; a++) {
                double gw = 0.0;
                for (int b = 0; b < E; b++) gw += G[(size_t)a * E + b] * w[b];
                Q += w[a] * (k[a] - 0.5 * gw);
            }
            W.Q.push_back(Q);
        }
    }
    return W;
}

// AdaptModelMLLR (OL_ModelAdapt.cpp:1583-1612): mu' = A mu + b.
std::unique_ptr<GMModel> CModelAdapt::AdaptModelMLLR(const GMModel &model, const MLLRTransform &W) {
    std::unique_ptr<GMModel> up(new GMModel(model));
    const int D = W.D, E = D + 1;
    std::vector<float> src(D);
    for (int m = 0; m < up->GetNumberOfMixtures(); m++) {
        float *mean = up->GetMixtureMean(m);
        std::memcpy(src.data(), mean, sizeof(float) * D);
        for (int a = 0; a < D; a++) {
            double v = 0.0;
            for (int b = 0; b < D; b++) v += W.W[(size_t)a * E + b] * src[b];
            mean[a] = (float)(v + W.W[(size_t)a * E + D]);
        }
    }
    return up;
}

// LU factorisation with partial pivoting in place (row-major n x n); returns
// the determinant and fills perm; a zero pivot gives 0.
static double lu_factor(int n, std::vector<double> &A, std::vector<int> &perm) {
    perm.resize(n);
    for (int i = 0; i < n; i++) perm[i] = i;
    double det = 1.0;
    for (int k = 0; k < n; k++) {
        int p = k;
        for (int i = k + 1; i < n; i++)
            if (std::fabs(A[(size_t)i * n + k]) > std::fabs(A[(size_t)p * n + k])) p = i;
        if (A[(size_t)p * n + k] == 0.0) return 0.0;
        if (p != k) {
            for (int j = 0; j < n; j++) std::swap(A[(size_t)k * n + j], A[(size_t)p * n + j]);
            std::swap(perm[k], perm[p]);
            det = -det;
        }
        det *= A[(size_t)k * n + k];
        for (int i = k + 1; i < n; i++) {
            const double l = A[(size_t)i * n + k] / A[(size_t)k * n + k];
            A[(size_t)i * n + k] = l;
            for (int j = k + 1; j < n; j++) A[(size_t)i * n + j] -= l * A[(size_t)k * n + j];
        }
    }
    return det;
}

// Inverse from an LU factorisation (column by column).
static std::vector<double> lu_inverse(int n, const std::vector<double> &LU,
                                      const std::vector<int> &perm) {
    std::vector<double> inv((size_t)n * n), y(n);
    for (int c = 0; c < n; c++) {
        for (int i = 0; i < n; i++) {
            double s = perm[i] == c ? 1.0 : 0.0;
            for (int j = 0; j < i; j++) s -= LU[(size_t)i * n + j] * y[j];
            y[i] = s;
        }
        for (int i = n - 1; i >= 0; i--) {
            double s = y[i];
            for (int j = i + 1; j < n; j++) s -= LU[(size_t)i * n + j] * inv[(size_t)j * n + c];
            inv[(size_t)i * n + c] = s / LU[(size_t)i * n + i];
        }
    }
    return inv;
}

// ComputeKiGi_fMLLR (OL_ModelAdapt.cpp:1793-1840) + ComputeW_fMLLR /
// EstimateAb_fmllr (:1614-1791).  Row i of W = [A | b] maximises
// beta log|det A| + w_i k_i - w_i G_i w_i^T / 2 with the cofactors p_i of
// the A of the previous sweep: w = (p / f + k) G^-1 for the roots f of
// beta f^2 - (p G^-1 k) f - p G^-1 p^T = 0, the root with the larger
// sum w (k - G w / 2) (the reference's criterion) kept.  The reference
// inverts with unpivoted Gauss-Jordan; here G by LU with partial pivoting
// (G is SPD) and A likewise.
MLLRTransform CModelAdapt::EstimateW_fMLLR(const GMModel &model, const GMMStats<double> &st,
                                           int iters) {
    const int D = model.GetDimension(), E = D + 1, M = (int)st.getMixNum();
    if (!st.hasMeans() || !st.hasVars() || !st.hasFullCov())
        throw std::logic_error("ComputeW_fMLLR(): needs first and full second-order statistics");
    std::vector<double> G((size_t)D * E * E, 0.0), Gi((size_t)D * E * E), K((size_t)D * E, 0.0);
    std::vector<float> vd(D);
    for (int m = 0; m < M; m++) {
        const double np = st.getProb(m);
        if (!(np > 0)) continue;
        std::memcpy(vd.data(), model.GetMixtureVarDiag(m), sizeof(float) * D);
        const float *mu = model.GetMixtureMean(m);
        const double *F = st.getMean(m), *S = st.getVar(m);
        for (int i = 0; i < D; i++) {
            const double cm = (double)(mu[i] / vd[i]), cv = (double)(1.0f / vd[i]);
            double *g = &G[(size_t)i * E * E], *k = &K[(size_t)i * E];
            size_t p = 0;
            for (int a = 0; a < D; a++) {
                k[a] += cm * F[a];
                for (int b = a; b < D; b++, p++) g[(size_t)a * E + b] += cv * S[p];
                g[(size_t)a * E + D] += cv * F[a];
            }
            k[D] += cm * np;
            g[(size_t)D * E + D] += cv * np;
        }
    }
    std::vector<int> perm;
    for (int i = 0; i < D; i++) {
        double *g = &G[(size_t)i * E * E];
        for (int a = 0; a < E; a++)
            for (int b = a + 1; b < E; b++) g[(size_t)b * E + a] = g[(size_t)a * E + b];
        std::vector<double> lu(g, g + (size_t)E * E);
        if (lu_factor(E, lu, perm) == 0.0)
            throw std::runtime_error("ComputeW_fMLLR(): matrix G_i is singular");
        const std::vector<double> inv = lu_inverse(E, lu, perm);
        std::copy(inv.begin(), inv.end(), Gi.begin() + (size_t)i * E * E);
    }
    std::vector<double> A((size_t)D * D, 0.0), b(D, 0.0), p(E), u(E), t(E), w[2], gw(E);
    w[0].resize(E);
    w[1].resize(E);
    for (int d = 0; d < D; d++) A[(size_t)d * D + d] = 1.0;
    const double beta = (double)(float)st.getTotAccSamples();
    for (int it = 0; it < iters; it++) {
        std::vector<double> lu(A);
        const double det = lu_factor(D, lu, perm);
        if (det == 0.0) throw std::runtime_error("ComputeW_fMLLR(): Matrix invertion failed!");
        std::vector<double> cof = lu_inverse(D, lu, perm);  // inv(A) det(A): columns = cofactor rows
        for (double &c : cof) c *= det;
        for (int i = 0; i < D; i++) {
            const double *g = &G[(size_t)i * E * E], *gi = &Gi[(size_t)i * E * E];
            const double *k = &K[(size_t)i * E];
            for (int a = 0; a < D; a++) p[a] = cof[(size_t)a * D + i];
            p[D] = 0.0;
            double bb = 0.0, cc = 0.0;
            for (int c = 0; c < E; c++) {
                double s = 0.0;
                for (int a = 0; a < E; a++) s += p[a] * gi[(size_t)a * E + c];
                u[c] = s;
                bb += s * k[c];
                cc += s * p[c];
            }
            bb = -bb;
            cc = -cc;
            const double sq = std::sqrt(bb * bb - 4 * beta * cc);
            if (!(sq >= 0)) throw std::runtime_error("EstimateAb_fmllr(): Discriminant < 0!");
            const double f[2] = {(-bb + sq) / (2 * beta), (-bb - sq) / (2 * beta)};
            double Q[2];
            for (int j = 0; j < 2; j++) {
                if (f[j] == 0) throw std::runtime_error("EstimateAb_fmllr(): Dividing by zero!");
                for (int c = 0; c < E; c++) t[c] = p[c] / f[j] + k[c];
                for (int c = 0; c < E; c++) {
                    double s = 0.0;
                    for (int a = 0; a < E; a++) s += t[a] * gi[(size_t)a * E + c];
                    w[j][c] = s;
                }
                Q[j] = 0.0;
                for (int a = 0; a < E; a++) {
                    double s = 0.0;
                    for (int c = 0; c < E; c++) s += g[(size_t)a * E + c] * w[j][c];
                    Q[j] += w[j][a] * (k[a] - 0.5 * s);
                }
            }
            const int idx = (Q[0] > Q[1]) ? 0 : 1;
            for (int c = 0; c < D; c++) A[(size_t)i * D + c] = w[idx][c];
            b[i] = w[idx][D];
        }
    }
    MLLRTransform W;
    W.D = D;
    W.W.resize((size_t)D * E);
    for (int i = 0; i < D; i++) {
        for (int c = 0; c < D; c++) W.W[(size_t)i * E + c] = A[(size_t)i * D + c];
        W.W[(size_t)i * E + D] = b[i];
    }
    return W;
}

MLLRTransform CModelAdapt::ComputeW_fMLLR(GMModel *model) {
    stats_.reset();
    AccumulateStats(model, stats_, false, true, true);
    CheckInvertibility();
    return EstimateW_fMLLR(*model, stats_, opt_._iter_fMLLR);
}

// UpdateMLLRMatrix (OL_ModelAdapt.cpp:1319-1340): Wsum <- Wnew o Wsum.
void CModelAdapt::ComposeMLLR(MLLRTransform &Wsum, const MLLRTransform &Wnew) {
    const int D = Wsum.D, E = D + 1;
    std::vector<double> out((size_t)D * E, 0.0);
    for (int a = 0; a < D; a++) {
        for (int b = 0; b < E; b++) {
            double s = 0.0;
            for (int c = 0; c < D; c++) s += Wnew.W[(size_t)a * E + c] * Wsum.W[(size_t)c * E + b];
            out[(size_t)a * E + b] = s;
        }
        out[(size_t)a * E + D] += Wnew.W[(size_t)a * E + D];
    }
    Wsum.W.swap(out);
}

MLLRTransform CModelAdapt::IterW_MLLR(GMModel *UBM, int iters, bool wantQ) {
    std::vector<double> Gi, ki;
    if (opt_._mllrUBMInit) prepareKiGiInitMatrices(*UBM, opt_._min_var, Gi, ki);
    MLLRTransform W = ComputeW_MLLR(UBM, Gi, ki, wantQ && iters <= 1);
    if (iters > 1) {
        std::unique_ptr<GMModel> m = AdaptModelMLLR(*UBM, W);
        for (int i = 1; i < iters; i++) {
            MLLRTransform Wt = ComputeW_MLLR(m.get(), Gi, ki, wantQ && i == iters - 1);
            m = AdaptModelMLLR(*m, Wt);
            ComposeMLLR(W, Wt);
            if (i == iters - 1) W.Q = Wt.Q;
        }
    }
    return W;
}

void CModelAdapt::AdaptModel(GMModel *UBM) {
    if (!UBM) throw std::logic_error("AdaptModel(): UBM not given!");
    if (!hasData()) throw std::logic_error("AdaptModel(): Insert adaptation data first!");
    // statistics as the reference allocates them (OL_ModelAdapt.cpp:281-298):
    // full second-order statistics for a full UBM, none for MLLR
    const bool fullV = UBM->GetFullCovStatus();
    UBM_ = UBM;
    nummix_ = UBM->GetNumberOfMixtures();
    const int D = UBM->GetDimension();
    if ((unsigned)D != D_) throw std::logic_error("AdaptModel(): data and UBM dimensions differ");
    model_.reset();
    W_.reset();
    switch (opt_._adaptType) {
    case MAP:
        stats_.alloc(D, nummix_, opt_._AdaptMeanBool || opt_._AdaptVarBool, opt_._AdaptVarBool, fullV,
                     opt_._robustCov);
        model_ = AdaptModelMAP(UBM);
        break;
    case MLLR: {
        stats_.alloc(D, nummix_, true, false);
        W_.reset(new MLLRTransform(IterW_MLLR(UBM, opt_._iterMLLR, opt_._getQvalMLLR)));
        if (opt_._applyTransforms) model_ = AdaptModelMLLR(*UBM, *W_);
        break;
    }
    case fMLLR:
        // full second-order statistics under the (diagonal or full) UBM
        stats_.alloc(D, nummix_, true, true, true);
        W_.reset(new MLLRTransform(ComputeW_fMLLR(UBM)));
        break;
    case MLLR_MAP: {
        stats_.alloc(D, nummix_, true, opt_._AdaptVarBool, fullV, opt_._robustCov);
        MLLRTransform W = IterW_MLLR(UBM, opt_._iterMLLR, false);
        std::unique_ptr<GMModel> mm = AdaptModelMLLR(*UBM, W);
        model_ = AdaptModelMAP(mm.get());
        break;
    }
    default:
        throw std::runtime_error("AdaptModel(): adaptation type not supported by this build");
    }
}

// Leading eigenpair of a symmetric matrix (getEigValVecs' ssyevr('V', 'I',
// 'U', il = iu = D), OL_ModelAdapt.cpp:72-160) by cyclic Jacobi rotations
// in double.  LAPACK leaves the eigenvector's sign open; it is fixed so that
// the entry of largest magnitude (first on a tie) is negative, the choice
// under which the reference's example/myModelFC.gmm is reproduced.
static void leading_eigenpair(int D, const float *cov, double &lam, std::vector<double> &v) {
    std::vector<double> A((size_t)D * D), V((size_t)D * D, 0.0);
    for (int i = 0; i < D * D; i++) A[i] = cov[i];
    for (int i = 0; i < D; i++) V[(size_t)i * D + i] = 1.0;
    auto a = [&](int i, int j) -> double & { return A[(size_t)i * D + j]; };
    for (int sweep = 0; sweep < 100; sweep++) {
        double off = 0.0;
        for (int p = 0; p < D; p++)
            for (int q = p + 1; q < D; q++) off += a(p, q) * a(p, q);
        if (off < 1e-30) break;
        for (int p = 0; p < D; p++)
            for (int q = p + 1; q < D; q++) {
                if (std::fabs(a(p, q)) < 1e-300) continue;
                const double theta = (a(q, q) - a(p, p)) / (2.0 * a(p, q));
                const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < D; k++) {
                    const double kp = a(k, p), kq = a(k, q);
                    a(k, p) = c * kp - s * kq;
                    a(k, q) = s * kp + c * kq;
                }
                for (int k = 0; k < D; k++) {
                    const double pk = a(p, k), qk = a(q, k);
                    a(p, k) = c * pk - s * qk;
                    a(q, k) = s * pk + c * qk;
                }
                for (int k = 0; k < D; k++) {
                    double &vp = V[(size_t)k * D + p], &vq = V[(size_t)k * D + q];
                    const double tp = vp, tq = vq;
                    vp = c * tp - s * tq;
                    vq = s * tp + c * tq;
                }
            }
    }
    int best = 0;
    for (int i = 1; i < D; i++)
        if (a(i, i) > a(best, best)) best = i;
    lam = a(best, best);
    int imax = 0;
    for (int i = 1; i < D; i++)
        if (std::fabs(V[(size_t)i * D + best]) > std::fabs(V[(size_t)imax * D + best])) imax = i;
    const double sg = V[(size_t)imax * D + best] > 0.0 ? -1.0 : 1.0;
    v.resize(D);
    for (int i = 0; i < D; i++) v[i] = sg * V[(size_t)i * D + best];
}

// splitGaussian (OL_ModelAdapt.cpp:614-660): component m2 halves its weight
// with the empty slot mn, which takes its covariance; the two means move
// apart by +-0.5 sqrt(var) (diagonal) or +-sqrt(lambda) v along the leading
// eigenpair (full covariance).
static void split_component(GMModel *model, int m2, int mn) {
    const int D = model->GetDimension();
    std::memcpy(model->GetMixtureVar(mn), model->GetMixtureVar(m2), sizeof(float) * model->GetDimVar());
    const float half = model->GetMixtureWeight(m2) / 2.0f;
    model->SetMixWeight(mn, half);
    model->SetMixWeight(m2, half);
    float *mi = model->GetMixtureMean(m2), *mv = model->GetMixtureMean(mn);
    const float *vi = model->GetMixtureVar(m2);
    std::vector<float> sh(D);
    if (model->GetFullCovStatus()) {
        double lam = 0.0;
        std::vector<double> ev;
        leading_eigenpair(D, vi, lam, ev);
        if (!(lam >= 0.0)) throw std::runtime_error("splitGaussian(): covariance matrix not positive definite!");
        for (int d = 0; d < D; d++) sh[d] = (float)(std::sqrt(lam) * ev[d]);
    } else {
        for (int d = 0; d < D; d++) sh[d] = (float)(std::sqrt(vi[d]) * 0.5);
    }
    for (int d = 0; d < D; d++) {
        mv[d] = mi[d] + sh[d];
        mi[d] -= sh[d];
    }
}

// splitMultipleGaussians + GetGaussStats (OL_ModelAdapt.cpp:590-611,
// :951-990).  The scan runs from the last component down: the heaviest
// component is the highest-index one among equal maxima, empty slots are
// taken from the highest index down, and the split order is by weight,
// descending, lower index first among equal weights.  ceil(split-X x
// non-zero) < 2 splits the heaviest component only; otherwise components are
// split in that order while they weigh at least split-th x the heaviest and
// empty slots remain.  Returns false when there was no empty slot.
static bool split_multiple(GMModel *model, float splitX, float splitTh) {
    const int M = model->GetNumberOfMixtures();
    int imax = 0;
    float wm = 0.0f;
    std::vector<int> zero, order;
    for (int m = M - 1; m >= 0; m--) {
        const float w = model->GetMixtureWeight(m);
        if (wm < w) {
            wm = w;
            imax = m;
        }
        if (w == 0.0f)
            zero.push_back(m);
        else
            order.push_back(m);
    }
    if (zero.empty()) return false;
    const unsigned nsplit = (unsigned)std::ceil(splitX * (float)order.size());
    if (nsplit < 2) {
        split_component(model, imax, zero.front());
        return true;
    }
    // order holds indices descending; a stable sort by weight keeps the
    // insertion rule of GetGaussStats (an equal weight goes before the
    // entries already placed, i.e. lower index first)
    std::reverse(order.begin(), order.end());
    std::vector<float> w0(M);
    for (int m = 0; m < M; m++) w0[m] = model->GetMixtureWeight(m);
    std::stable_sort(order.begin(), order.end(), [&w0](int a, int b) { return w0[a] > w0[b]; });
    const float th = splitTh * model->GetMixtureWeight(imax);
    size_t z = 0;
    for (unsigned n = 0; n < order.size(); n++) {
        if (model->GetMixtureWeight(order[n]) < th || z == zero.size() || n >= nsplit) break;
        split_component(model, order[n], zero[z++]);
    }
    return true;
}

// TrainFromScratch (OL_ModelAdapt.cpp:663-821): EM rounds of --it-EM
// iterations, a split after each round, up to nummix components.  --ll-Th
// (:738-741, :781-797) needs at least two EM iterations and stops from the
// fourth round on once a round gains less than the threshold over the better
// of the two rounds before it, keeping the best of those models.  The mean
// log-likelihood of the last EM iteration is computed at every verbosity
// here; the reference computes it only when verbose (:766-769), so at
// verbosity 0 its check compares -FLT_MAX with itself and stops at the
// fourth round whatever the data.
void CModelAdapt::TrainFromScratch(unsigned int nummix, bool fullcov, const char *FileName,
                                   bool save_txt) {
    if (!hasData() || N_ <= 3)
        throw std::runtime_error("TrainFromScratch(): Not enough data to model training.");
    const int D = (int)D_;
    // Global statistics: a single component with a broad variance has
    // posterior 1 for every finite frame.  The first pass finds the mean; the
    // second is centred on it, so E[x^2] - mu^2 does not cancel.
    GMModel g(D, 1, fullcov);
    std::vector<float> z(D, 0.0f), v((size_t)g.GetDimVar(), fullcov ? 0.0f : 1e4f);
    if (fullcov)
        for (int d = 0; d < D; d++) v[(size_t)d * D + d] = 1e4f;
    g.SetMixParam(0, z.data(), v.data(), 1.0f);
    const bool robust = opt_._robustCov;
    GMMStats<double> gs(D, 1, true, true, fullcov, robust);
    gs.reset();
    AccumulateStats(&g, gs, false, true, false);
    for (int d = 0; d < D; d++) z[d] = (float)(gs.getMean(0)[d] / gs.getProb(0));
    g.SetMixParam(0, z.data(), v.data(), 1.0f);
    gs.reset();
    AccumulateStats(&g, gs, false, true, true, robust);
    GMModel init(D, 1, fullcov);
    init.SetMixWeight(0, 1.0f);
    if (UpdateModelEM(&init, gs) == 0)
        throw std::runtime_error("trainFromScratch(): amount of data too low to estimate a model");
    model_.reset(new GMModel(D, (int)nummix, fullcov));
    model_->SetMixParam(0, init.GetMixtureMean(0), init.GetMixtureVar(0), init.GetMixtureWeight(0));
    nummix_ = (int)nummix;
    stats_.alloc(D, nummix, true, true, fullcov, robust);
    const unsigned int maxIter = nummix < 2 ? 0 : nummix;
    int iterEM = opt_._iterCountEM;
    if (iterEM < 2 && opt_._llTh > 0.0f) {
        std::cout << " (num of EM iteartions < 2 -> it-EM set to 2)" << std::endl;
        iterEM = 2;
    }
    float ll[3] = {-std::numeric_limits<float>::max(), -std::numeric_limits<float>::max(),
                   -std::numeric_limits<float>::max()};
    GMModel m2, m3;
    unsigned int nonzero = 1;
    GMModel *model = model_.get();
    for (unsigned int it = 0; it < maxIter; it++) {
        for (int e = 0; e < iterEM; e++) {
            stats_.reset();
            AccumulateStats(model, stats_, false, true, true, robust);
            ll[0] = (float)(stats_.getTotLogLike() / (double)stats_.getTotAccSamples());
            if (opt_._verbosity) std::cout << "     [LL = " << ll[0] << "]" << std::endl;
            if (nonzero > 1) nonzero = UpdateModelEM(model, stats_);
            if (nonzero < 2) {
                if (nonzero < 1) *model = init;
                break;
            }
        }
        if (nonzero == 0 || nonzero == nummix) break;
        if (opt_._llTh > 0.0f) {
            const float mx = std::max(ll[1], ll[2]);
            if (it > 2 && ll[0] - mx < opt_._llTh) {
                if (mx > ll[0]) *model = (mx == ll[1]) ? m2 : m3;
                break;
            }
            m3 = m2;
            m2 = *model;
            ll[2] = ll[1];
            ll[1] = ll[0];
        }
        if (opt_._saveModelAfterNSplits > 0 && (it + 1) % (unsigned)opt_._saveModelAfterNSplits == 0 &&
            FileName) {
            char tag[64];
            std::snprintf(tag, sizeof(tag), ".it%05u.G%05u.gmm.bckup", it, nonzero);
            if (model->Save((std::string(FileName) + tag).c_str(), save_txt) != 0)
                throw std::runtime_error(model->LastError());
        }
        if (it < maxIter - 1) {
            if (!split_multiple(model, opt_._splitXpercAtOnce, opt_._splitTh)) break;
            nonzero += 1;
        }
    }
    model->RearrangeMixtures();
}

int CModelAdapt::Save(const char *filename, bool save_txt) {
    if (W_ && (!model_ || opt_._saveTransforms)) {
        std::string out(filename);
        const size_t dot = out.rfind('.');
        if (dot != std::string::npos) out.erase(dot);
        out += opt_._transformFileExt;
        std::ofstream f(out);
        if (!f) throw std::runtime_error("Save(): Unable to open output file " + out);
        f.precision(10);
        f << (opt_._adaptType == fMLLR ? " fMLLR" : " MLLR") << "\n 1\t" << W_->D << " "
          << W_->D + 1 << "\n";
        for (int a = 0; a < W_->D; a++) {
            for (int b = 0; b <= W_->D; b++) f << " " << W_->W[(size_t)a * (W_->D + 1) + b];
            f << "\n";
        }
    }
    if (model_) {
        if (model_->Save(filename, save_txt) != 0) throw std::runtime_error(model_->LastError());
        return 0;
    }
    return W_ ? 0 : 1;
}

}  // namespace gmm
