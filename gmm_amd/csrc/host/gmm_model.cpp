#include "gmm_model.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>

namespace gmm {

GMModel::GMModel(int dim, int nummix, bool full_cov) {
    if (full_cov) err_ = "full-covariance models are not supported";
    dim_ = dim;
    w_.assign(nummix, 0.0f);
    occ_.assign(nummix, -1.0f);
    mu_.assign((size_t)nummix * dim, 0.0f);
    var_.assign((size_t)nummix * dim, 0.0f);
    norm_.assign(dim, 1.0f);
    for (int m = 0; m < nummix; m++) ResetMixture(m);
}

void GMModel::SetMixParam(int m, const float *mean, const float *var, float weight) {
    std::memcpy(GetMixtureMean(m), mean, sizeof(float) * dim_);
    std::memcpy(GetMixtureVar(m), var, sizeof(float) * dim_);
    w_[m] = weight;
}

void GMModel::ResetMixture(int m) {
    std::memset(GetMixtureMean(m), 0, sizeof(float) * dim_);
    std::memset(GetMixtureVar(m), 1, sizeof(float) * dim_);
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
    for (int i = 0; i < M; i++) {
        const int m = order[i];
        w[i] = w_[m];
        occ[i] = occ_[m];
        std::memcpy(&mu[(size_t)i * dim_], &mu_[(size_t)m * dim_], sizeof(float) * dim_);
        std::memcpy(&var[(size_t)i * dim_], &var_[(size_t)m * dim_], sizeof(float) * dim_);
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
        if (occ_[m] != -1) std::fprintf(f, " <MIX_OCCUP> %15e", occ_[m]);
        std::fprintf(f, "\n\n");
        for (int d = 0; d < dim_; d++) std::fprintf(f, "%15e\n", GetMixtureMean(m)[d]);
        std::fprintf(f, "\n");
        for (int d = 0; d < dim_; d++) std::fprintf(f, "%15e\n", GetMixtureVar(m)[d]);
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
    std::vector<float> w(M), occ(M, -1.0f), mu((size_t)M * D), var((size_t)M * D);
    for (int m = 0; m < M; m++) {
        if (!next(ln)) return fail("truncated model file (weight)");
        std::istringstream ws(ln);
        std::string tok;
        ws >> w[m];
        if (ws >> tok) {
            if (tok == "F") return fail("full-covariance models are not supported");
            if (tok == "<MIX_OCCUP>") ws >> occ[m];
        }
        if (!next(ln)) return fail("truncated model file");
        for (int d = 0; d < D; d++) {
            if (!next(ln)) return fail("truncated model file (mean)");
            mu[(size_t)m * D + d] = std::strtof(ln.c_str(), nullptr);
        }
        if (!next(ln)) return fail("truncated model file");
        for (int d = 0; d < D; d++) {
            if (!next(ln)) return fail("truncated model file (variance)");
            var[(size_t)m * D + d] = std::strtof(ln.c_str(), nullptr);
        }
        next(ln);  // blank line after each component (may be missing at EOF)
    }
    dim_ = D;
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
    const int16_t d16 = (int16_t)dim_;
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
             std::fwrite(GetMixtureVar(m), sizeof(float), dim_, f) == (size_t)dim_;
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
    if (ok && d16 < 0) {
        std::fclose(f);
        return fail("full-covariance models are not supported");
    }
    const int D = d16;
    std::vector<float> norm(D > 0 ? D : 0);
    ok = ok && D > 0 && std::fread(norm.data(), sizeof(float), D, f) == (size_t)D;
    ok = ok && std::fread(&m16, sizeof(m16), 1, f) == 1 && m16 >= 0;
    const int M = ok ? m16 : 0;
    std::vector<float> w(M), mu((size_t)M * D), var((size_t)M * D);
    ok = ok && std::fread(w.data(), sizeof(float), M, f) == (size_t)M;
    for (int m = 0; m < M && ok; m++)
        ok = std::fread(&mu[(size_t)m * D], sizeof(float), D, f) == (size_t)D &&
             std::fread(&var[(size_t)m * D], sizeof(float), D, f) == (size_t)D;
    std::fclose(f);
    if (!ok) return fail(std::string("truncated or corrupt model file ") + filename);
    dim_ = D;
    norm_.swap(norm);
    w_.swap(w);
    occ_.assign(M, -1.0f);
    mu_.swap(mu);
    var_.swap(var);
    nzw_ = 0;
    return 0;
}

}  // namespace gmm
