// Tuning table of the E-step kernels (internal): per (path, precision, D, M)
// the log-likelihood frame tile, the frames per statistics work item and the
// mixed statistics component tile that a timed sweep chose.  Text format,
// one entry per line:
//   <estep|batch> <fp32|bf16x3|fp16x3> D M ll_frames chunk_frames ms [comp_tile]
// (comp_tile 128 when absent).  '#' starts a comment.  Host-only code
// (tune.cpp).
#pragma once

namespace gmmhip_tune {

struct Entry {
    int path;          // GMMHIP_TUNE_ESTEP / GMMHIP_TUNE_BATCH
    int prec;          // GMMHIP_PREC_*
    int D, M;
    int ll_frames;     // 64, 128 or 256
    int chunk_frames;  // frames per statistics work item (upper bound)
    double ms;         // time measured by the sweep
    int comp_tile;     // mixed statistics components per block: 128 or 256
};

// (Re)load the table; path == nullptr: $GMMHIP_TUNE_FILE, else
// <library dir>/../tuning/gfx950.tune.  Returns the number of entries, or -1
// when the file cannot be read (the table is then empty).
int load(const char *path);

// Entry for (path, prec, D, M): exact match, else the same D with the
// nearest M (log scale), else the nearest (D, M) (log scale).  Loads the
// default table on first use.  False when the table holds no entry for
// (path, prec).
bool lookup(int path, int prec, int D, int M, Entry *e);

}  // namespace gmmhip_tune
