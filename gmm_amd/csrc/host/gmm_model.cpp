#include "gmm_model.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>

namespace gmm {

GMModel::GMModel(int dim, int nummix, bool full_cov) {
    dim_ = dim;
    full_ = full_cov;
    w_.assign(nummix, 0.0f);
    occ_.assign(nummix, -1.0f);
    mu_.assign((size_t)nummix * dim, 0.0f);
    var_.assign((size_t)nummix * GetDimVar(), 0.0f);
    norm_.assign(dim, 1.0f);
    for (int m = 0; m < nummix; m++) ResetMixture(m);
}

void GMModel::SetMixParam(int m, const float *mean, const float *var, float weight) {
    std::memcpy(GetMixtureMean(m), mean, sizeof(float) * dim_);
    std::memcpy(GetMixtureVar(m), var, sizeof(float) * GetDimVar());
    w_[m] = weight;
}

const float *GMModel::GetMixtureVarDiag(int m) const {
    if (!full_) return GetMixtureVar(m);
    vdiag_.resize(dim_);
    for (int d = 0; d < dim_; d++) vdiag_[d] = GetMixtureVar(m)[(size_t)d * dim_ + d];
    return vdiag_.data();
}

void GMModel::ResetMixture(int m) {
    std::memset(GetMixtureMean(m), 0, sizeof(float) * dim_);
    if (full_) {
        float *v = GetMixtureVar(m);
        std::memset(v, 0, sizeof(float) * GetDimVar());
        for (int d = 0; d < dim_; d++) v[(size_t)d * dim_ + d] = 1.0f;
    } else {
        std::memset(GetMixtureVar(m), 1, sizeof(float) * dim_);
    }
}

unsigned int GMModel::RearrangeMixtures() {
    const int M = GetNumberOfMixtures();
    std::vector<int> order(M);
    int front = 0, back = M - 1;
    for (int m = 0; m < M; m++) {
        if (w_[m] > 0.0f)
            order[front++] = m;
        else
            order[back--] = m;
    }
    std::vector<float> w(M), occ(M), mu(mu_.size()), var(var_.size());
    const size_t dv = GetDimVar();
    for (int i = 0; i < M; i++) {
        const int m = order[i];
        w[i] = w_[m];
        occ[i] = occ_[m];
        std::memcpy(&mu[(size_t)i * dim_], &mu_[(size_t)m * dim_], sizeof(float) * dim_);
        std::memcpy(&var[(size_t)i * dv], &var_[(size_t)m * dv], sizeof(float) * dv);
    }
    w_.swap(w);
    occ_.swap(occ);
    mu_.swap(mu);
    var_.swap(var);
    nzw_ = front;
    return (unsigned int)(M - front);
}

int GMModel::Load(const char *filename, bool txt) { return txt ? ReadTXT(filename) : ReadBIN(filename); }

int GMModel::Save(const char *filename, bool txt) const {
    return txt ? WriteTXT(filename) : WriteBIN(filename);
}

int GMModel::WriteTXT(const char *filename) const {
    FILE *f = std::fopen(filename, "w");
    if (!f) return fail(std::string("cannot open output file ") + filename);
    std::fprintf(f, "%d\n", dim_);
    for (int d = 0; d < dim_; d++) std::fprintf(f, "%15e\n", norm_[d]);
    int M = 0;
    for (float w : w_) M += w > 0;
    std::fprintf(f, "%d\n", M);
    for (int m = 0; m < GetNumberOfMixtures(); m++) {
        if (w_[m] == 0) continue;
        std::fprintf(f, "%15e", w_[m]);
        if (full_) std::fprintf(f, " F");
        if (occ_[m] != -1) std::fprintf(f, " <MIX_OCCUP> %15e", occ_[m]);
        std::fprintf(f, "\n\n");
        for (int d = 0; d < dim_; d++) std::fprintf(f, "%15e\n", GetMixtureMean(m)[d]);
        std::fprintf(f, "\n");
        const float *v = GetMixtureVar(m);
        if (full_) {
            // each row: the diagonal element, then the row
            for (int i = 0; i < dim_; i++) {
                std::fprintf(f, "%15e", v[(size_t)i * dim_ + i]);
                for (int j = 0; j < dim_; j++) std::fprintf(f, " %15e", v[(size_t)i * dim_ + j]);
                std::fprintf(f, "\n");
            }
        } else {
            for (int d = 0; d < dim_; d++) std::fprintf(f, "%15e\n", v[d]);
        }
        std::fprintf(f, "\n");
    }
    if (std::fclose(f) != 0) return fail(std::string("cannot close ") + filename);
    return 0;
}

int GMModel::ReadTXT(const char *filename) {
    std::ifstream in(filename);
    if (!in) return fail(std::string("cannot open file ") + filename);
    std::vector<std::string> lines;
    for (std::string ln; std::getline(in, ln);) lines.push_back(ln);
    size_t pos = 0;
    auto next = [&](std::string &out) {
        if (pos >= lines.size()) return false;
        out = lines[pos++];
        return true;
    };
    std::string ln;
    if (!next(ln)) return fail("truncated model file (dimension)");
    const int D = std::atoi(ln.c_str());
    if (D <= 0) return fail("bad dimension in model file");
    std::vector<float> norm(D);
    for (int d = 0; d < D; d++) {
        if (!next(ln)) return fail("truncated model file (normalisation vector)");
        norm[d] = std::strtof(ln.c_str(), nullptr);
    }
    if (!next(ln)) return fail("truncated model file (number of components)");
    const int M = std::atoi(ln.c_str());
    if (M < 0) return fail("bad number of components");
    std::vector<float> w(M), occ(M, -1.0f), mu((size_t)M * D), var;
    bool full = false;
    for (int m = 0; m < M; m++) {
        if (!next(ln)) return fail("truncated model file (weight)");
        std::istringstream ws(ln);
        std::string tok;
        ws >> w[m];
        bool fc = false;
        if (ws >> tok) {
            if (tok == "F") {
                fc = true;
                if (ws >> tok && tok == "<MIX_OCCUP>") ws >> occ[m];
            } else if (tok == "<MIX_OCCUP>") {
                ws >> occ[m];
            }
        }
        if (m == 0) {
            full = fc;
            var.assign((size_t)M * (full ? D * D : D), 0.0f);
        } else if (fc != full) {
            return fail("model file mixes diagonal and full-covariance components");
        }
        if (!next(ln)) return fail("truncated model file");
        for (int d = 0; d < D; d++) {
            if (!next(ln)) return fail("truncated model file (mean)");
            mu[(size_t)m * D + d] = std::strtof(ln.c_str(), nullptr);
        }
        if (!next(ln)) return fail("truncated model file");
        if (full) {
            // D rows of 1 + D numbers: the diagonal element, then the row
            for (int i = 0; i < D; i++) {
                if (!next(ln)) return fail("truncated model file (covariance)");
                std::istringstream rs(ln);
                float diag;
                if (!(rs >> diag)) return fail("bad covariance row in model file");
                for (int j = 0; j < D; j++)
                    if (!(rs >> var[((size_t)m * D + i) * D + j]))
                        return fail("bad covariance row in model file");
            }
        } else {
            for (int d = 0; d < D; d++) {
                if (!next(ln)) return fail("truncated model file (variance)");
                var[(size_t)m * D + d] = std::strtof(ln.c_str(), nullptr);
            }
        }
        next(ln);  // blank line after each component (may be missing at EOF)
    }
    dim_ = D;
    full_ = full;
    norm_.swap(norm);
    w_.swap(w);
    occ_.swap(occ);
    mu_.swap(mu);
    var_.swap(var);
    nzw_ = 0;
    return 0;
}

int GMModel::WriteBIN(const char *filename) const {
    FILE *f = std::fopen(filename, "wb");
    if (!f) return fail(std::string("cannot open output file ") + filename);
    bool ok = std::fwrite("GMM-DB-VN", 9, 1, f) == 1;
    // a full-covariance model stores its dimension negated
    const int16_t d16 = (int16_t)(full_ ? -dim_ : dim_);
    ok = ok && std::fwrite(&d16, sizeof(d16), 1, f) == 1;
    ok = ok && std::fwrite(norm_.data(), sizeof(float), dim_, f) == (size_t)dim_;
    std::vector<float> ws;
    for (float w : w_)
        if (w > 0) ws.push_back(w);
    const int16_t m16 = (int16_t)ws.size();
    ok = ok && std::fwrite(&m16, sizeof(m16), 1, f) == 1;
    ok = ok && std::fwrite(ws.data(), sizeof(float), ws.size(), f) == ws.size();
    for (int m = 0; m < GetNumberOfMixtures() && ok; m++) {
        if (w_[m] == 0) continue;
        ok = std::fwrite(GetMixtureMean(m), sizeof(float), dim_, f) == (size_t)dim_ &&
             std::fwrite(GetMixtureVar(m), sizeof(float), GetDimVar(), f) == (size_t)GetDimVar();
    }
    if (std::fclose(f) != 0) ok = false;
    return ok ? 0 : fail(std::string("error writing ") + filename);
}

int GMModel::ReadBIN(const char *filename) {
    FILE *f = std::fopen(filename, "rb");
    if (!f) return fail(std::string("cannot open file ") + filename);
    char id[9];
    int16_t d16 = 0, m16 = 0;
    bool ok = std::fread(id, 9, 1, f) == 1 && std::memcmp(id, "GMM-DB-VN", 9) == 0;
    if (!ok) {
        std::fclose(f);
        return fail("incorrect file format (not GMM-DB-VN)");
    }
    ok = std::fread(&d16, sizeof(d16), 1, f) == 1;
    const bool full = d16 < 0;
    const int D = full ? -d16 : d16;
    const size_t DV = full ? (size_t)D * D : (size_t)D;
    std::vector<float> norm(D > 0 ? D : 0);
    ok = ok && D > 0 && std::fread(norm.data(), sizeof(float), D, f) == (size_t)D;
    ok = ok && std::fread(&m16, sizeof(m16), 1, f) == 1 && m16 >= 0;
    const int M = ok ? m16 : 0;
    std::vector<float> w(M), mu((size_t)M * D), var((size_t)M * DV);
    ok = ok && std::fread(w.data(), sizeof(float), M, f) == (size_t)M;
    for (int m = 0; m < M && ok; m++)
        ok = std::fread(&mu[(size_t)m * D], sizeof(float), D, f) == (size_t)D &&
             std::fread(&var[(size_t)m * DV], sizeof(float), DV, f) == DV;
    std::fclose(f);
    if (!ok) return fail(std::string("truncated or corrupt model file ") + filename);
    dim_ = D;
    full_ = full;
    norm_.swap(norm);
    w_.swap(w);
    occ_.assign(M, -1.0f);
    mu_.swap(mu);
    var_.swap(var);
    nzw_ = 0;
    return 0;
}

}  // namespace gmm
