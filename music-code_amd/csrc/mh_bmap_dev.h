// mh_bmap_dev.h -- device view of a bias map (mh_bmap) and its search,
// shared by the bias-map builder and the GCMC kernel.
#pragma once
#include "mh_internal.h"

namespace mh {

struct BiasRef {
    const int32_t *cube;    // [numtab] cubelet id (i * ny1 + j) * nz1 + k
    const int32_t *cindex;  // [ncubelets] 1-based list position, 0 = not biased
    const int32_t *guide;   // [2^gbits + 1]
    const double *cum;      // [numtab] cumulative probability
    const double *wn;       // [numtab] weight / total (bmap_getbiaswt)
    long long numtab;
    int ncub[3], gbits;
    double ncubelets, step[3], eff[3];
};

// first list position whose cumulative probability is >= key (key in
// [0, 1)): the cubelet bmap_getbiasindex (bmap.F90:575-636) returns
__device__ __forceinline__ long long bias_index(const BiasRef &B, double key) {
    const long long M = 1ll << B.gbits;
    long long b = (long long)(key * (double)M);
    if (b > M - 1) b = M - 1;
    long long lo = B.guide[b], hi = B.guide[b + 1];
    while (lo < hi) {
        const long long mid = (lo + hi) >> 1;
        if (B.cum[mid] >= key) hi = mid;
        else lo = mid + 1;
    }
    return lo;
}

inline BiasRef bias_ref(const mh_bmap *b) {
    BiasRef r{};
    r.cube = b->cube.as<int32_t>();
    r.cindex = b->cindex.as<int32_t>();
    r.guide = b->guide.as<int32_t>();
    r.cum = b->cum.as<double>();
    r.wn = b->wn.as<double>();
    r.numtab = b->numtab;
    for (int i = 0; i < 3; i++) {
        r.ncub[i] = b->ncub[i];
        r.step[i] = b->step[i];
        r.eff[i] = b->eff[i];
    }
    r.gbits = b->gbits;
    r.ncubelets = (double)b->ncubelets;
    return r;
}

}  // namespace mh
