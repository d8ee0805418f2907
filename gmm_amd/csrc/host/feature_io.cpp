#include "feature_io.h"

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <iostream>
#include <cstring>
#include <fstream>
#include <stdexcept>

namespace gmm {

static void make_rows(Features &out);

static int fail(std::string *err, int rc, const std::string &msg) {
    if (err) *err = msg;
    return rc;
}

int LoadRaw(const char *filename, Features &out, std::string *err) {
    out = Features();
    std::ifstream f(filename, std::ios::binary);
    if (!f) return fail(err, -1, std::string("cannot open ") + filename);
    int32_t hdr[2] = {0, 0};
    f.read(reinterpret_cast<char *>(hdr), sizeof(hdr));
    if (!f || hdr[1] < 0) return fail(err, -2, std::string("bad RAW header in ") + filename);
    if (hdr[1] == 0) return 0;  // no dimension: empty
    uint64_t n = hdr[0] > 0 ? (uint64_t)hdr[0] : 0;
    const unsigned int D = (unsigned int)hdr[1];
    if (n == 0) {
        f.seekg(0, std::ios::end);
        const uint64_t sz = (uint64_t)f.tellg() - 8;
        n = sz / D / sizeof(float);
        f.seekg(8, std::ios::beg);
    }
    if (n == 0) return 0;
    out.N = (unsigned int)n;
    out.D = D;
    out.data.resize((size_t)n * D);
    f.read(reinterpret_cast<char *>(out.data.data()), sizeof(float) * out.data.size());
    if (!f) {
        out = Features();
        return fail(err, -2, std::string("truncated RAW file ") + filename);
    }
    make_rows(out);
    return 0;
}

static void make_rows(Features &out) {
    out.rows.resize(out.N);
    for (unsigned int i = 0; i < out.N; i++) out.rows[i] = &out.data[(size_t)i * out.D];
}

int LoadSVES(const char *filename, Features &out, std::string *err) {
    out = Features();
    std::ifstream f(filename, std::ios::binary);
    if (!f) return fail(err, -1, std::string("cannot open ") + filename);
    char id[11];
    int32_t D = 0, N = 0;
    if (!f.read(id, 11) || std::string(id, 11) != "SV-ES-PARAM")
        return fail(err, -2, std::string("Param::Load(): file's ") + filename +
                                 " format different from SV-ES-PARAM");
    f.read(reinterpret_cast<char *>(&D), 4);
    f.read(reinterpret_cast<char *>(&N), 4);
    if (!f || D < 0 || N < 0) return fail(err, -2, std::string("bad SVES header in ") + filename);
    if (D == 0 || N == 0) return 0;  // empty
    f.seekg((std::streamoff)sizeof(float) * N, std::ios::cur);  // confidence factors
    out.N = (unsigned)N;
    out.D = (unsigned)D;
    out.data.resize((size_t)N * D);
    f.read(reinterpret_cast<char *>(out.data.data()), sizeof(float) * out.data.size());
    if (!f) {
        out = Features();
        return fail(err, -2, std::string("truncated SVES file ") + filename);
    }
    make_rows(out);
    return 0;
}

static uint32_t be32(const unsigned char *b) {
    return ((uint32_t)b[0] << 24) | ((uint32_t)b[1] << 16) | ((uint32_t)b[2] << 8) | b[3];
}

int LoadHTK(const char *filename, Features &out, std::string *err) {
    out = Features();
    std::ifstream f(filename, std::ios::binary);
    if (!f) return fail(err, -1, std::string("cannot open ") + filename);
    unsigned char h[12];
    if (!f.read(reinterpret_cast<char *>(h), 12))
        return fail(err, -2, std::string("cannot read HTK header of ") + filename);
    const uint32_t n = be32(h);
    const uint32_t period = be32(h + 4);
    const int sampsize = (int16_t)(((uint16_t)h[8] << 8) | h[9]);
    if (sampsize < 0) return fail(err, -2, std::string("bad HTK sample size in ") + filename);
    const unsigned D = (unsigned)(sampsize / 4);
    if (D == 0 || n == 0) return 0;  // empty
    out.N = n;
    out.D = D;
    out.samplePeriod = period;
    out.data.resize((size_t)n * D);
    std::vector<unsigned char> buf(sizeof(float) * out.data.size());
    f.read(reinterpret_cast<char *>(buf.data()), buf.size());
    if (!f) {
        out = Features();
        return fail(err, -2, std::string("truncated HTK file ") + filename);
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
        return fail(err, -2, "Param::Load(): Unknown input data type specified");
    }
}

int LoadFeatureRanges(const char *filename, int type, const element_ranges &ranges,
                      unsigned int samplePeriod, Features &out, std::string *err) {
    Features all;
    const int rc = LoadFeatures(filename, type, all, err);
    out = Features();
    if (rc != 0) return rc;
    const unsigned p = type == 1 ? all.samplePeriod : samplePeriod;
    if (p == 0)
        return fail(err, -2, std::string("Param::Load(): ") + filename +
                                 ": set sample period first (not present in SVES / RAW files)");
    unsigned last = 0, maxend = 0;
    for (const std::vector<unsigned int> &r : ranges) {
        if (r.size() < 2) return fail(err, -2, std::string("Param::Load(): bad interval for ") + filename);
        const unsigned b = r[0] / p, e = r[1] / p;
        if (b >= e || b >= all.N)
            return fail(err, -2, std::string("Param::Load(): In file ") + filename +
                                     ": prm segment_begin >= (segment_end OR number_of_samples)");
        if (b < last)
            return fail(err, -2, "Param::Load(): Incorrect frame intervals (not ascending!) specified");
        maxend = std::max(maxend, e);
        const unsigned ec = std::min(e, all.N);
        out.data.insert(out.data.end(), all.data.begin() + (size_t)b * all.D,
                        all.data.begin() + (size_t)ec * all.D);
        out.N += ec - b;
        last = ec;
    }
    if (maxend > all.N)
        std::cerr << "WARNING: Param::Load(): file " << filename << " too short: " << maxend - all.N
                  << " frames missing" << std::endl;
    out.D = out.N ? all.D : 0;
    out.samplePeriod = all.samplePeriod;
    make_rows(out);
    return 0;
}

static void put_be32(unsigned char *b, uint32_t v) {
    b[0] = (unsigned char)(v >> 24);
    b[1] = (unsigned char)(v >> 16);
    b[2] = (unsigned char)(v >> 8);
    b[3] = (unsigned char)v;
}

int SaveSVES(const char *filename, const float *x, unsigned int N, unsigned int D, const float *cf) {
    std::ofstream f(filename, std::ios::binary | std::ios::trunc);
    if (!f) return -1;
    f.write("SV-ES-PARAM", 11);
    const int32_t hdr[2] = {(int32_t)D, (int32_t)N};
    f.write(reinterpret_cast<const char *>(hdr), sizeof(hdr));
    if (N > 0) {
        if (cf) {
            f.write(reinterpret_cast<const char *>(cf), sizeof(float) * (size_t)N);
        } else {
            const std::vector<float> ones(N, 1.0f);
            f.write(reinterpret_cast<const char *>(ones.data()), sizeof(float) * (size_t)N);
        }
        f.write(reinterpret_cast<const char *>(x), sizeof(float) * (size_t)N * D);
    }
    return f ? 0 : -1;
}

int SaveHTK(const char *filename, const float *x, unsigned int N, unsigned int D,
            unsigned int samplePeriod, int parmKind) {
    if (4ull * D > 32767) return -2;  // sampSize is a 16-bit field
    std::ofstream f(filename, std::ios::binary | std::ios::trunc);
    if (!f) return -1;
    unsigned char h[12];
    put_be32(h, N);
    put_be32(h + 4, samplePeriod);
    h[8] = (unsigned char)((4 * D) >> 8);
    h[9] = (unsigned char)(4 * D);
    h[10] = (unsigned char)((unsigned)parmKind >> 8);
    h[11] = (unsigned char)parmKind;
    f.write(reinterpret_cast<const char *>(h), 12);
    std::vector<unsigned char> buf(4 * (size_t)N * D);
    for (size_t i = 0; i < (size_t)N * D; i++) {
        uint32_t u;
        std::memcpy(&u, &x[i], 4);
        put_be32(&buf[4 * i], u);
    }
    f.write(reinterpret_cast<const char *>(buf.data()), buf.size());
    return f ? 0 : -1;
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
