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

// v from the lane CTRL names (DPP on both 32-bit halves); lanes with no
// source (and rows outside ROWS) read 0
template <int CTRL, int ROWS = 0xf>
__device__ __forceinline__ double dpp_f64(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, ROWS, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, ROWS, 0xf, false);
    return __hiloint2double(hi, lo);
}

// wavefront sum in registers, fixed order: row_shr 1, 2, 4, 8 leave each
// 16-lane row's sum in its lane 15; row_bcast15 and row_bcast31 carry the row
// sums up to lane 63, whose value is returned to every lane
__device__ __forceinline__ double wave_sum(double v) {
    v += dpp_f64<0x111>(v);
    v += dpp_f64<0x112>(v);
    v += dpp_f64<0x114>(v);
    v += dpp_f64<0x118>(v);
    v += dpp_f64<0x142, 0xa>(v);
    v += dpp_f64<0x143, 0xc>(v);
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), 63);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), 63);
    return __hiloint2double(hi, lo);
}

// block reduction of NV doubles per lane, fixed order (wave sums, then the
// waves in order)
template <int NT, int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double *red /*[NT/64][NV]*/) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < NV; i++) v[i] = wave_sum(v[i]);
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
