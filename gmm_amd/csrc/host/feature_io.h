// Feature and matrix files on the hot path's edges:
//   RAW features  int32 N, int32 D, float32[N*D]  (Param::LoadRaw,
//                 param/OL_Param.cpp:358-453; N < 1 means "from file size")
//   saveMat       TXT "x y\n" + rows / BIN int32 x, int32 y, float[x*y]
//                 (matrices/IOMats.h:168-232), the per-frame log-likelihood
//                 output of getLogLike.
#pragma once

#include <list>
#include <map>
#include <string>
#include <vector>

namespace gmm {

struct Features {
    unsigned int N = 0, D = 0;
    std::vector<float> data;        // row-major N x D
    std::vector<float *> rows;      // the reference's float** view
    unsigned int samplePeriod = 0;  // HTK header's sample period; 0 for SVES/RAW
};

// Frame intervals per feature file, the reference's file_2_ranges
// (tools/OL_MLFParser.h:53): file name -> list of [x0, x1) in sample-period
// units (HTK MLF times, 100 ns).  Ordered by file name, so every run visits
// the files in the same order.
typedef std::list<std::vector<unsigned int>> element_ranges;
typedef std::map<std::string, element_ranges> file_2_ranges;

// Loaders return 0 on success (an empty file gives N = D = 0, as Param
// leaves it), -1 when the file cannot be opened (callers skip it with a
// warning, as the reference does) and -2 for a malformed or truncated file
// or bad arguments (callers throw, as Param does).
int LoadRaw(const char *filename, Features &out, std::string *err = nullptr);
// SVES: "SV-ES-PARAM", int32 D, int32 N, float CF[N], float[N*D]
// (Param::Save/Load, param/OL_Param.cpp:177-354).
int LoadSVES(const char *filename, Features &out, std::string *err = nullptr);
// HTK: big-endian int32 nSamples, int32 sampPeriod, int16 sampSize,
// int16 parmKind, big-endian float rows (Param::LoadHTK, :477-576).
int LoadHTK(const char *filename, Features &out, std::string *err = nullptr);
// --data-T 0 (SVES) / 1 (HTK) / 2 (RAW), Param::Load (:222-244).
int LoadFeatures(const char *filename, int type, Features &out, std::string *err = nullptr);
// The frames of the intervals of one file (Param::Load / LoadHTK / LoadRaw
// with frameIdxs, param/OL_Param.cpp:581-1037): interval [x0, x1) covers
// frames [x0 / p, x1 / p) with p the HTK header's period, or samplePeriod
// for SVES and RAW (which carry none; 0 is an error there).  Intervals must
// be ascending; one with x0 / p >= x1 / p or starting past the end is an
// error (-2); one running past the end is clipped with a warning.  The
// reference's HTK variant (:897-1000) is followed for all three types: its
// SVES and RAW variants count every interval twice (:647 and :659-662,
// :788 and :800-803) and read the frames at shifted offsets (:699, :843).
int LoadFeatureRanges(const char *filename, int type, const element_ranges &ranges,
                      unsigned int samplePeriod, Features &out, std::string *err = nullptr);
int SaveRaw(const char *filename, const float *x, unsigned int N, unsigned int D);
// SVES writer (Param::Save, param/OL_Param.cpp:177-220); cf = NULL writes
// confidence factors of 1.
int SaveSVES(const char *filename, const float *x, unsigned int N, unsigned int D,
             const float *cf = nullptr);
// HTK writer, the layout LoadHTK reads (:477-576; IOscripts/writeHTK.m):
// big-endian nSamples, sampPeriod, sampSize = 4 D, parmKind, float rows.
int SaveHTK(const char *filename, const float *x, unsigned int N, unsigned int D,
            unsigned int samplePeriod, int parmKind = 9);
void SaveMat(const char *filename, bool txt, const float *mat, unsigned int x, unsigned int y,
             bool include_head = true);

}  // namespace gmm
