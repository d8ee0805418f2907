// Feature and matrix files on the hot path's edges:
//   RAW features  int32 N, int32 D, float32[N*D]  (Param::LoadRaw,
//                 param/OL_Param.cpp:358-453; N < 1 means "from file size")
//   saveMat       TXT "x y\n" + rows / BIN int32 x, int32 y, float[x*y]
//                 (matrices/IOMats.h:168-232), the per-frame log-likelihood
//                 output of getLogLike.
#pragma once

#include <string>
#include <vector>

namespace gmm {

struct Features {
    unsigned int N = 0, D = 0;
    std::vector<float> data;        // row-major N x D
    std::vector<float *> rows;      // the reference's float** view
};

// 0 on success.
int LoadRaw(const char *filename, Features &out, std::string *err = nullptr);
// SVES: "SV-ES-PARAM", int32 D, int32 N, float CF[N], float[N*D]
// (Param::Save/Load, param/OL_Param.cpp:177-354).
int LoadSVES(const char *filename, Features &out, std::string *err = nullptr);
// HTK: big-endian int32 nSamples, int32 sampPeriod, int16 sampSize,
// int16 parmKind, big-endian float rows (Param::LoadHTK, :477-576).
int LoadHTK(const char *filename, Features &out, std::string *err = nullptr);
// --data-T 0 (SVES) / 1 (HTK) / 2 (RAW), Param::Load (:222-244).
int LoadFeatures(const char *filename, int type, Features &out, std::string *err = nullptr);
int SaveRaw(const char *filename, const float *x, unsigned int N, unsigned int D);
void SaveMat(const char *filename, bool txt, const float *mat, unsigned int x, unsigned int y,
             bool include_head = true);

}  // namespace gmm
