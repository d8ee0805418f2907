// GMM container (diagonal or full covariance) and the reference's `.gmm`
// model file format.
//
// Mirrors GMModel (model/OL_GMModel.h:69-136 of GMMTeam/GMM): same accessor
// names, the same TXT layout (WriteTXT/ReadTXT, OL_GMModel.cpp:475-547,
// :627-753; full covariance: " F" after the weight and D rows of the
// diagonal element followed by the row) and BIN layout "GMM-DB-VN"
// (WriteBIN/ReadBIN, :551-624, :757-853; full covariance: negative
// dimension and D*D variances per component), both writers skipping
// zero-weight components, and the same RearrangeMixtures order (zero-weight
// components moved to the end in reversed order, :857-885), which fixes the
// component indexing of every statistics buffer.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace gmm {

class GMModel {
public:
    GMModel() = default;
    GMModel(int dim, int nummix, bool full_cov = false);

    int GetDimension() const { return dim_; }
    int GetNumberOfMixtures() const { return (int)w_.size(); }
    // valid after RearrangeMixtures(), as in the reference
    int GetNumberOfNZWMixtures() const { return nzw_; }
    bool GetFullCovStatus() const { return full_; }
    // variances per component: D, or D*D for a full-covariance model
    int GetDimVar() const { return full_ ? dim_ * dim_ : dim_; }

    float GetMixtureWeight(int m) const { return w_[m]; }
    const float *GetMixtureWeights() const { return w_.data(); }
    void SetMixWeight(int m, float w) { w_[m] = w; }
    float *GetMixtureMean(int m) { return &mu_[(size_t)m * dim_]; }
    const float *GetMixtureMean(int m) const { return &mu_[(size_t)m * dim_]; }
    float *GetMixtureVar(int m) { return &var_[(size_t)m * GetDimVar()]; }
    const float *GetMixtureVar(int m) const { return &var_[(size_t)m * GetDimVar()]; }
    // diagonal of the covariance (a copy for full-covariance models, valid
    // until the next call)
    const float *GetMixtureVarDiag(int m) const;
    float GetMixtureOccup(int m) const { return occ_[m]; }
    void SetMixOcc(int m, float occ) { occ_[m] = occ; }
    void SetMixParam(int m, const float *mean, const float *var, float weight);
    const std::vector<float> &NormCoef() const { return norm_; }

    // GMMMixture::Reset (OL_GMModel.cpp:73-83): mean = 0; a diagonal
    // component gets every variance byte set to 0x01, a full one the
    // identity.
    void ResetMixture(int m);

    // Moves zero-weight components to the end (their order reversed) and
    // returns how many there are; GetNumberOfNZWMixtures() is then valid.
    unsigned int RearrangeMixtures();

    // 0 on success, non-zero on failure (message in LastError()).
    int Load(const char *filename, bool txt);
    int Save(const char *filename, bool txt) const;
    int ReadTXT(const char *filename);
    int WriteTXT(const char *filename) const;
    int ReadBIN(const char *filename);
    int WriteBIN(const char *filename) const;
    const std::string &LastError() const { return err_; }

    // Flat storage (row-major [M][D], variances [M][GetDimVar()]) for the
    // device layer.
    const std::vector<float> &weights() const { return w_; }
    const std::vector<float> &means() const { return mu_; }
    const std::vector<float> &vars() const { return var_; }

private:
    int dim_ = 0;
    int nzw_ = 0;
    bool full_ = false;
    std::vector<float> w_, mu_, var_, occ_, norm_;
    mutable std::vector<float> vdiag_;
    mutable std::string err_;
    int fail(const std::string &msg) const {
        err_ = msg;
        return 1;
    }
};

}  // namespace gmm
