// Exercises the C++ estimator's Gaussian-selection API (mask2fill,
// estimateMaskIdxs, mask2use; OL_GMMStatsEstimator.h:73-114) on the device:
// fills the mask from the per-frame posteriors, derives the selections
// above the cut, accumulates with them, and prints what the Python test
// compares against the oracle.
//   mask_api <model.gmm (TXT)> <frames.raw>
#include <cstdio>
#include <vector>

#include "feature_io.h"
#include "gmm_estimator.h"
#include "gmm_model.h"
#include "gmm_stats.h"

int main(int argc, char **argv) {
    if (argc != 3) return 2;
    gmm::GMModel model;
    if (model.Load(argv[1], true) != 0) return 3;
    gmm::Features f;
    if (gmm::LoadRaw(argv[2], f) != 0) return 4;
    f.rows.resize(f.N);
    for (unsigned i = 0; i < f.N; i++) f.rows[i] = &f.data[(size_t)i * f.D];
    const unsigned M = model.GetNumberOfMixtures();
    gmm::GMMStatsEstimator_HIP<double> est(0);
    est.insertModel(model);
    using Mask = gmm::GMMStatsEstimator_HIP<double>::logLikeMask;
    Mask mask;
    mask.N = f.N;
    mask.dim = M;
    std::vector<float> ll(f.N), g((size_t)f.N * M);
    std::vector<float *> rows(f.N);
    for (unsigned i = 0; i < f.N; i++) rows[i] = &g[(size_t)i * M];
    mask.logLikes = ll.data();
    mask.gammas = rows.data();
    est.mask2fill(mask);
    std::vector<float> out(f.N);
    est.compVecsLogLikeMT(f.rows.data(), f.N, f.D, out.data());
    gmm::GMMStatsEstimator_HIP<double>::estimateMaskIdxs(mask, est._minGamma);
    est.mask2use(mask);
    gmm::GMMStats<double> st(f.D, M, true, true);
    st.reset();
    est.accumulateStatsMT(f.rows.data(), f.N, f.D, st);
    est.withdrawMask();
    std::printf("N %u M %u\n", f.N, M);
    for (unsigned i = 0; i < f.N; i++) {
        std::printf("LL %.9g", ll[i]);
        for (unsigned m = 0; m < M; m++) std::printf(" %.9g", g[(size_t)i * M + m]);
        std::printf("\n");
    }
    for (unsigned m = 0; m < M; m++) {
        std::printf("n %.17g", st.getProb(m));
        for (unsigned d = 0; d < f.D; d++) std::printf(" %.17g", st.getMean(m)[d]);
        std::printf("\n");
    }
    std::printf("tot %.17g\n", (double)st.getTotLogLike());
    return 0;
}
