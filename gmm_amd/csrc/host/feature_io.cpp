#include "feature_io.h"

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <stdexcept>

namespace gmm {

int LoadRaw(const char *filename, Features &out, std::string *err) {
    std::ifstream f(filename, std::ios::binary);
    if (!f) {
        if (err) *err = std::string("cannot open ") + filename;
        return -1;
    }
    int32_t hdr[2] = {0, 0};
    f.read(reinterpret_cast<char *>(hdr), sizeof(hdr));
    if (!f || hdr[1] <= 0) {
        if (err) *err = std::string("bad RAW header in ") + filename;
        return -1;
    }
    uint64_t n = hdr[0] > 0 ? (uint64_t)hdr[0] : 0;
    const unsigned int D = (unsigned int)hdr[1];
    if (n == 0) {
        f.seekg(0, std::ios::end);
        const uint64_t sz = (uint64_t)f.tellg() - 8;
        n = sz / D / sizeof(float);
        f.seekg(8, std::ios::beg);
    }
    out.N = (unsigned int)n;
    out.D = D;
    out.data.resize((size_t)n * D);
    f.read(reinterpret_cast<char *>(out.data.data()), sizeof(float) * out.data.size());
    if (!f) {
        if (err) *err = std::string("truncated RAW file ") + filename;
        return -1;
    }
    out.rows.resize(n);
    for (uint64_t i = 0; i < n; i++) out.rows[i] = &out.data[(size_t)i * D];
    return 0;
}

static void make_rows(Features &out) {
    out.rows.resize(out.N);
    for (unsigned int i = 0; i < out.N; i++) out.rows[i] = &out.data[(size_t)i * out.D];
}

int LoadSVES(const char *filename, Features &out, std::string *err) {
    std::ifstream f(filename, std::ios::binary);
    char id[11];
    int32_t D = 0, N = 0;
    if (!f || !f.read(id, 11) || std::string(id, 11) != "SV-ES-PARAM") {
        if (err) *err = std::string("not an SVES file: ") + filename;
        return -1;
    }
    f.read(reinterpret_cast<char *>(&D), 4);
    f.read(reinterpret_cast<char *>(&N), 4);
    if (!f || D <= 0 || N < 0) {
        if (err) *err = std::string("bad SVES header in ") + filename;
        return -1;
    }
    f.seekg((std::streamoff)sizeof(float) * N, std::ios::cur);  // confidence factors
    out.N = (unsigned)N;
    out.D = (unsigned)D;
    out.data.resize((size_t)N * D);
    f.read(reinterpret_cast<char *>(out.data.data()), sizeof(float) * out.data.size());
    if (!f) {
        if (err) *err = std::string("truncated SVES file ") + filename;
        return -1;
    }
    make_rows(out);
    return 0;
}

static uint32_t be32(const unsigned char *b) {
    return ((uint32_t)b[0] << 24) | ((uint32_t)b[1] << 16) | ((uint32_t)b[2] << 8) | b[3];
}

int LoadHTK(const char *filename, Features &out, std::string *err) {
    std::ifstream f(filename, std::ios::binary);
    unsigned char h[12];
    if (!f || !f.read(reinterpret_cast<char *>(h), 12)) {
        if (err) *err = std::string("cannot read HTK header of ") + filename;
        return -1;
    }
    const uint32_t n = be32(h);
    const int sampsize = (int16_t)(((uint16_t)h[8] << 8) | h[9]);
    const unsigned D = (unsigned)(sampsize / 4);
    if (D == 0) {
        if (err) *err = std::string("bad HTK sample size in ") + filename;
        return -1;
    }
    out.N = n;
    out.D = D;
    out.data.resize((size_t)n * D);
    std::vector<unsigned char> buf(sizeof(float) * out.data.size());
    f.read(reinterpret_cast<char *>(buf.data()), buf.size());
    if (!f) {
        if (err) *err = std::string("truncated HTK file ") + filename;
        return -1;
    }
    for (size_t i = 0; i < out.data.size(); i++) {
        const uint32_t u = be32(&buf[4 * i]);
        float v;
        std::memcpy(&v, &u, 4);
        out.data[i] = v;
    }
    make_rows(out);
    return 0;
}

int LoadFeatures(const char *filename, int type, Features &out, std::string *err) {
    switch (type) {
    case 0: return LoadSVES(filename, out, err);
    case 1: return LoadHTK(filename, out, err);
    case 2: return LoadRaw(filename, out, err);
    default:
        if (err) *err = "Param::Load(): Unknown input data type specified";
        return -1;
    }
}

int SaveRaw(const char *filename, const float *x, unsigned int N, unsigned int D) {
    std::ofstream f(filename, std::ios::binary);
    if (!f) return -1;
    const int32_t hdr[2] = {(int32_t)N, (int32_t)D};
    f.write(reinterpret_cast<const char *>(hdr), sizeof(hdr));
    f.write(reinterpret_cast<const char *>(x), sizeof(float) * (size_t)N * D);
    return f ? 0 : -1;
}

void SaveMat(const char *filename, bool txt, const float *mat, unsigned int x, unsigned int y,
             bool include_head) {
    std::ofstream f(filename, txt ? std::ios::out : (std::ios::out | std::ios::binary));
    if (!f) throw std::runtime_error(std::string("saveMat(): file ") + filename +
                                     " could not be opened for writing");
    if (txt) {
        if (include_head) f << x << " " << y << "\n";
        for (unsigned int i = 0; i < x; i++) {
            for (unsigned int j = 0; j < y; j++) f << mat[(size_t)i * y + j] << " ";
            f << "\n";
        }
        f << std::endl;
    } else {
        if (include_head) {
            const int32_t h[2] = {(int32_t)x, (int32_t)y};
            f.write(reinterpret_cast<const char *>(h), sizeof(h));
        }
        f.write(reinterpret_cast<const char *>(mat), sizeof(float) * (size_t)x * y);
    }
    if (!f) throw std::runtime_error(std::string("saveMat(): Exception writing file ") + filename);
}

}  // namespace gmm
