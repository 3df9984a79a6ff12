// Device helpers shared by the Monte Carlo kernels (GCMC, HMC) of
// libmusic_hip: the reference RNG, cell folding, map lookups, reductions.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mh {

struct MapRef {
    const float *dense;   // [nx][ny][nz][8], slot 0 NaN where untabulated
    int n[3];
    double step[3];
    double ell[3];
};

struct Rng {
    int32_t ix, iy, first, iseed;
};

// rranf_NR (random.F90:121-152); am = (1 - 2^-24) 2^-31 (single-precision
// Nearest(1.0,-1.0)/CONST_IM, exact)
__device__ __forceinline__ double rranf(Rng &r) {
    if (r.first) {
        int32_t a = r.iseed < 0 ? -r.iseed : r.iseed;
        r.first = 0;
        r.iy = (int32_t)((888889999 ^ a) | 1);
        r.ix = (int32_t)(777755555 ^ a);
        r.iseed = a + 1;
    }
    uint32_t x = (uint32_t)r.ix;
    x ^= x << 13;
    x ^= x >> 17;
    x ^= x << 5;
    r.ix = (int32_t)x;
    int32_t k = r.iy / 127773;
    r.iy = 16807 * (r.iy - k * 127773) - 2836 * k;
    if (r.iy < 0) r.iy += 2147483647;
    int32_t v = (int32_t)((2147483647 & (r.ix ^ r.iy)) | 1);
    const double am = (double)(0.99999994039535522461f / 2147483648.0f);
    return __dmul_rn(am, (double)v);
}

// simcell_pbc of one coordinate (simcell.F90:521-547)
__device__ __forceinline__ double wrap_full(double x, double L) {
    return __dsub_rn(x, __dmul_rn(floor(x / L), L));
}

// simcell_maptouc + mapheader_getIndices along one axis (simcell.F90:634-668,
// mapheader.F90:277-299): 0-based lower / upper node (upper wraps to 0 at the
// last node, interpolate.F90:640-660) and the fractional offset t.
// simcell_maptouc on one axis (simcell.F90:634-668): the coordinate folded
// into [0, ell) with the reference's fmod form
__device__ __forceinline__ double maptouc(double c, double ell) {
    double f = c * (1.0 / ell);
    const double fr = f - trunc(f);  // == fmod(f, 1.0), exact
    f = (f >= 0.0) ? fr : fr + 1.0;
    return f * ell;
}

__device__ __forceinline__ bool cell_axis(double c, double ell, double st, int nb, int &i0, int &i1, double &t) {
    double f = c * (1.0 / ell);
    double fr = f - trunc(f);  // == fmod(f, 1.0), exact
    f = (f >= 0.0) ? fr : fr + 1.0;
    c = f * ell;
    const int idx = (int)(c / st + 1.0);
    if (idx < 1 || idx > nb) return false;
    t = (c - (double)(idx - 1) * st) / st;
    i0 = idx - 1;
    i1 = (idx == nb) ? 0 : idx;
    return true;
}

// corner and value-slot patterns of pothermite (interpolate.F90:725-835):
// bit i = upper node along the axis for corner i / derivative order for slot i
constexpr int kCX = 0xB2, kCY = 0xD4, kCZ = 0xE8;

// block reduction of NV doubles per lane, fixed order (lane tree then waves)
template <int NT, int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double *red /*[NT/64][NV]*/) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < NV; i++)
        for (int off = 32; off > 0; off >>= 1) v[i] += __shfl_xor(v[i], off);
    if (lane == 0)
        for (int i = 0; i < NV; i++) red[wave * NV + i] = v[i];
    __syncthreads();
    for (int i = 0; i < NV; i++) {
        double s = 0.0;
        for (int w = 0; w < NT / 64; w++) s += red[w * NV + i];
        v[i] = s;
    }
    __syncthreads();
}

}  // namespace mh
