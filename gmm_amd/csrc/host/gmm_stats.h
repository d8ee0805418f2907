// Sufficient-statistics container, the counterpart of GMMStats<T>
// (model/OL_GMMStats.h:44-148 of GMMTeam/GMM): per component the occupancy
// n = sum g, first order F = sum g x and second order S = sum g x^2
// (diagonal) or the packed upper triangle of sum g x x^T (full: D(D+1)/2
// per component, row k holding columns k..D-1, OL_GMMStatsEstimator.cpp:
// 1358-1364),
// the robust auxiliary statistics when allocated (per component
// [aux (D) | aux2 | aux3]: sum h |x_t - x_t-1|, sum sqrt g, sum h over frame
// pairs; OL_GMMStats.h:53-55),
// plus the number of accumulated frames and the total log-likelihood.
// Estimators ADD into it (the caller resets), operator+= is the reduction
// (OL_GMMStats.cpp:99-142), and save()/load() keep the reference's TXT/BIN
// layout (OL_GMMStats.cpp:699-776).  Storage is flat ([M], [M][D]) so the
// device layer can accumulate into it directly.
#pragma once

#include <cstdint>
#include <cstring>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace gmm {

template <typename T>
class GMMStats {
public:
    GMMStats() = default;
    GMMStats(unsigned int dim, unsigned int nummix, bool allocMeans = true, bool allocVars = false,
             bool allocFullVars = false, bool allocAux = false) {
        alloc(dim, nummix, allocMeans, allocVars, allocFullVars, allocAux);
    }

    // alloc does not initialise values -> reset()
    void alloc(unsigned int dim, unsigned int nummix, bool allocMeans = true,
               bool allocVars = false, bool allocFullVars = false, bool allocAux = false) {
        dim_ = dim;
        M_ = nummix;
        allM_ = allocMeans;
        allV_ = allocVars;
        allFV_ = allocVars && allocFullVars;
        allA_ = allocAux;
        varsize_ = allFV_ ? dim * (dim + 1) / 2 : dim;
        n_.assign(nummix, T(0));
        F_.assign(allocMeans ? (size_t)nummix * dim : 0, T(0));
        S_.assign(allocVars ? (size_t)nummix * varsize_ : 0, T(0));
        A_.assign(allocAux ? (size_t)nummix * (dim + 2) : 0, T(0));
        allocated_ = true;
    }
    void reset() {
        std::fill(n_.begin(), n_.end(), T(0));
        std::fill(F_.begin(), F_.end(), T(0));
        std::fill(S_.begin(), S_.end(), T(0));
        std::fill(A_.begin(), A_.end(), T(0));
        _totalAccSamples = 0;
        _totLogLike = T(0);
    }
    bool isAllocated() const { return allocated_; }
    bool hasMeans() const { return allM_; }
    bool hasVars() const { return allV_; }
    bool hasFullCov() const { return allFV_; }
    bool hasAux() const { return allA_; }
    unsigned int getVarSize() const { return varsize_; }
    unsigned int getMixNum() const { return M_; }
    unsigned int getDim() const { return dim_; }
    uint64_t getTotAccSamples() const { return _totalAccSamples; }
    T getTotLogLike() const { return _totLogLike; }
    T getProb(unsigned int m) const { return n_[m]; }
    const T *getMean(unsigned int m) const { return allM_ ? &F_[(size_t)m * dim_] : nullptr; }
    const T *getVar(unsigned int m) const { return allV_ ? &S_[(size_t)m * varsize_] : nullptr; }

    // robust statistics of component m: aux[D], then aux2, aux3
    const T *getAux(unsigned int m) const { return allA_ ? &A_[(size_t)m * (dim_ + 2)] : nullptr; }
    T getAux2(unsigned int m) const { return A_[(size_t)m * (dim_ + 2) + dim_]; }
    T getAux3(unsigned int m) const { return A_[(size_t)m * (dim_ + 2) + dim_ + 1]; }
    T *auxData() { return allA_ ? A_.data() : nullptr; }

    T *probData() { return n_.data(); }
    T *meanData() { return allM_ ? F_.data() : nullptr; }
    T *varData() { return allV_ ? S_.data() : nullptr; }

    // Only statistics allocated on the left are summed (as in the reference).
    GMMStats &operator+=(const GMMStats &r) {
        if (!allocated_) {
            *this = r;
            return *this;
        }
        if (r.M_ != M_ || r.dim_ != dim_ || (allV_ && r.allV_ && r.varsize_ != varsize_))
            throw std::logic_error("GMMStats::operator+=: shape mismatch");
        for (size_t i = 0; i < n_.size(); i++) n_[i] += r.n_[i];
        if (allM_ && r.allM_)
            for (size_t i = 0; i < F_.size(); i++) F_[i] += r.F_[i];
        if (allV_ && r.allV_)
            for (size_t i = 0; i < S_.size(); i++) S_[i] += r.S_[i];
        if (allA_ && r.allA_)
            for (size_t i = 0; i < A_.size(); i++) A_[i] += r.A_[i];
        _totalAccSamples += r._totalAccSamples;
        _totLogLike += r._totLogLike;
        return *this;
    }

    // Header "saveMProbs sMeans sVars sFullVars dim varsize nummix", then per
    // component the occupancy, the first-order and the second-order rows.
    void save(const char *filename, bool txt, bool saveMProbs = true, bool saveMeans = true,
              bool saveVars = true) const {
        const bool sM = saveMeans && allM_, sV = saveVars && allV_;
        std::ofstream f(filename, txt ? std::ios::out : (std::ios::out | std::ios::binary));
        if (!f) throw std::runtime_error(std::string("GMMStats::save: cannot open ") + filename);
        if (txt) {
            f.precision(17);
            f << (int)saveMProbs << " " << (int)sM << " " << (int)sV << " " << (int)(sV && allFV_)
              << " " << dim_ << " " << varsize_ << " " << M_ << "\n\n";
            for (unsigned m = 0; m < M_; m++) {
                if (saveMProbs) f << n_[m] << "\n";
                if (sM) {
                    for (unsigned d = 0; d < dim_; d++) f << " " << F_[(size_t)m * dim_ + d];
                    f << "\n";
                }
                if (sV) {
                    for (unsigned d = 0; d < varsize_; d++) f << " " << S_[(size_t)m * varsize_ + d];
                    f << "\n";
                }
                f << "\n";
            }
        } else {
            const int32_t hdr[7] = {(int32_t)saveMProbs, (int32_t)sM, (int32_t)sV,
                                    (int32_t)(sV && allFV_), (int32_t)dim_, (int32_t)varsize_,
                                    (int32_t)M_};
            f.write(reinterpret_cast<const char *>(hdr), sizeof(hdr));
            for (unsigned m = 0; m < M_; m++) {
                if (saveMProbs) f.write(reinterpret_cast<const char *>(&n_[m]), sizeof(T));
                if (sM) f.write(reinterpret_cast<const char *>(&F_[(size_t)m * dim_]), sizeof(T) * dim_);
                if (sV)
                    f.write(reinterpret_cast<const char *>(&S_[(size_t)m * varsize_]), sizeof(T) * varsize_);
            }
        }
        if (!f) throw std::runtime_error(std::string("GMMStats::save: write failed ") + filename);
    }

    void load(const char *filename, bool txt) {
        std::ifstream f(filename, txt ? std::ios::in : (std::ios::in | std::ios::binary));
        if (!f) throw std::runtime_error(std::string("GMMStats::load: cannot open ") + filename);
        int32_t hdr[7];
        if (txt) {
            for (int i = 0; i < 7; i++) f >> hdr[i];
        } else {
            f.read(reinterpret_cast<char *>(hdr), sizeof(hdr));
        }
        const bool fv = hdr[3] != 0;
        if (!f || hdr[4] <= 0 || hdr[5] != (fv ? hdr[4] * (hdr[4] + 1) / 2 : hdr[4]))
            throw std::runtime_error("GMMStats::load: bad header");
        alloc(hdr[4], hdr[6], hdr[1] != 0, hdr[2] != 0, fv);
        reset();
        for (unsigned m = 0; m < M_; m++) {
            if (txt) {
                if (hdr[0]) f >> n_[m];
                if (allM_)
                    for (unsigned d = 0; d < dim_; d++) f >> F_[(size_t)m * dim_ + d];
                if (allV_)
                    for (unsigned d = 0; d < varsize_; d++) f >> S_[(size_t)m * varsize_ + d];
            } else {
                if (hdr[0]) f.read(reinterpret_cast<char *>(&n_[m]), sizeof(T));
                if (allM_) f.read(reinterpret_cast<char *>(&F_[(size_t)m * dim_]), sizeof(T) * dim_);
                if (allV_)
                    f.read(reinterpret_cast<char *>(&S_[(size_t)m * varsize_]), sizeof(T) * varsize_);
            }
        }
        if (!f) throw std::runtime_error(std::string("GMMStats::load: truncated ") + filename);
    }

    // saveAux (OL_GMMStats.cpp:778-812): text, "dim nummix", a blank line,
    // then per component aux2, aux3 and the row of aux.
    void saveAux(const char *filename) const {
        std::ofstream f(filename);
        if (!f) throw std::runtime_error(std::string("saveAux(): Unable to open file ") + filename);
        f << dim_ << " " << M_ << "\n\n";
        if (allA_)
            for (unsigned m = 0; m < M_; m++) {
                f << getAux2(m) << "\n" << getAux3(m) << "\n";
                for (unsigned d = 0; d < dim_; d++) f << " " << A_[(size_t)m * (dim_ + 2) + d];
                f << "\n";
            }
        if (!f) throw std::runtime_error(std::string("saveAux(): Exception writing file ") + filename);
    }

    uint64_t _totalAccSamples = 0;
    T _totLogLike = T(0);

private:
    bool allocated_ = false, allM_ = false, allV_ = false, allFV_ = false, allA_ = false;
    unsigned int dim_ = 0, M_ = 0, varsize_ = 0;
    std::vector<T> n_, F_, S_, A_;
};

}  // namespace gmm
