// mh_bmap.hip -- bias map for biased insertion and deletion (bmap_init +
// bmap_cubeinit, bmap.F90:69-403), built on the device from a pmap.
//
// A cubelet (i, j, k) of the map grid is biased when its lower-corner node
// is tabulated with V <= 30 kcal/mol (BMAP_DEFAULT_CUTOFF); its weight is
// exp(-V / kT_bias) with the exponent clamped to +-30.  Cubelets are listed
// in the reference's loop order (x outer, z inner) and the cumulative
// probabilities are the running sums over the total, both summed in that
// order.  When min(weight) / sum(weight) <= 1e-6 the energies above
// -log(1e-6 sum) / rti are lowered to it first (bmap.F90:283-295; when the
// ratio is larger the reference compares against an unset local, here
// nothing is clamped).
//
// Kernels:
//   k_bmap_energy   one thread per cubelet: V of the lower-corner node or NaN
//   k_bmap_weight   one thread per cubelet: the clamped Boltzmann weight, or -1
//   k_bmap_sum      one wavefront: the sum and minimum in list order
//   k_bmap_list     one wavefront: list positions (ballot + popcount), the
//                   running sums in list order (each lane's weight added in
//                   lane order, so the result is the sequential sum bit for bit)
//   k_bmap_prob     one thread per biased cubelet: cumulative probability and
//                   normalised weight (weight / total, bmap_getbiaswt)
//   k_bmap_guide    one thread per biased cubelet: a guide table g[b] = first
//                   position whose cumulative probability is >= b / 2^gbits,
//                   so a search reads two guide entries and bisects between them
//                   (the position found is that of bmap_getbiasindex,
//                   bmap.F90:575-636: the first cumulative probability >= key)
#include "mh_internal.h"
#include "mh_bmap_dev.h"

#include <cmath>
#include <vector>

using namespace mh;

namespace {

constexpr double kBiasUcut = 30.0;  // BMAP_DEFAULT_CUTOFF, bmap.F90:61
constexpr double kMaxExp = 30.0;    // default_MAX_EXP, defaults.F90:135

__global__ void k_bmap_energy(const int32_t *__restrict__ ptr, const float *__restrict__ vals, int ny, int nz,
                              int ny1, int nz1, long long ncub, double *__restrict__ u) {
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncub) return;
    const long long i = c / ((long long)ny1 * nz1), j = (c / nz1) % ny1, k = c % nz1;
    const int p = ptr[(i * ny + j) * nz + k];
    double v = NAN;
    if (p != 0) {
        const double e = (double)vals[(long long)(p - 1) * 8];
        if (e <= kBiasUcut) v = e;
    }
    u[c] = v;
}

__global__ void k_bmap_weight(const double *__restrict__ u, long long n, double rti, double umax,
                              double *__restrict__ wt) {
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    double v = u[c], w = -1.0;
    if (!isnan(v)) {
        if (v > umax) v = umax;
        double e = -rti * v;
        if (e > kMaxExp) e = kMaxExp;
        else if (e < -kMaxExp) e = -kMaxExp;
        w = exp(e);
    }
    wt[c] = w;
}

__device__ __forceinline__ double lane_value(double v, int l) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
    return __hiloint2double(hi, lo);
}

// out: sum, min, count of the weights >= 0, in list order
__global__ void __launch_bounds__(64) k_bmap_sum(const double *__restrict__ wt, long long n, double *out) {
    const int lane = threadIdx.x;
    double sum = 0.0, mn = 1.0e20;
    long long cnt = 0;
    for (long long base = 0; base < n; base += 64) {
        const long long c = base + lane;
        const double w = c < n ? wt[c] : -1.0;
#pragma unroll 8
        for (int l = 0; l < 64; l++) {
            const double wl = lane_value(w, l);
            if (wl >= 0.0) {
                sum = sum + wl;
                if (wl < mn) mn = wl;
                cnt++;
            }
        }
    }
    if (lane == 0) {
        out[0] = sum;
        out[1] = mn;
        out[2] = (double)cnt;
    }
}

// list of biased cubelets with their running sums; out: total, count
__global__ void __launch_bounds__(64) k_bmap_list(const double *__restrict__ wt, long long n,
                                                  int32_t *__restrict__ cube, double *__restrict__ run,
                                                  int32_t *__restrict__ cindex, double *out) {
    const int lane = threadIdx.x;
    double cum = 0.0;
    long long t0 = 0;
    const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (long long base = 0; base < n; base += 64) {
        const long long c = base + lane;
        const double w = c < n ? wt[c] : -1.0;
        const bool f = w >= 0.0;
        const unsigned long long bal = __ballot(f);
        const long long t = t0 + __popcll(bal & below);
        double mine = 0.0;
#pragma unroll 8
        for (int l = 0; l < 64; l++) {
            const double wl = lane_value(w, l);
            if (wl >= 0.0) {
                cum = cum + wl;
                if (lane == l) mine = cum;
            }
        }
        if (f) {
            cube[t] = (int32_t)c;
            run[t] = mine;
        }
        if (c < n) cindex[c] = f ? (int32_t)(t + 1) : 0;
        t0 += __popcll(bal);
    }
    if (lane == 0) {
        out[0] = cum;
        out[1] = (double)t0;
    }
}

__global__ void k_bmap_prob(const double *__restrict__ run, const int32_t *__restrict__ cube,
                            const double *__restrict__ wt, long long numtab, const double *__restrict__ tot,
                            double *__restrict__ cum, double *__restrict__ w, double *__restrict__ wn) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= numtab) return;
    const double cw = tot[0];
    const double x = wt[cube[t]];
    cum[t] = run[t] / cw;
    w[t] = x;
    wn[t] = x / cw;
}

__global__ void k_bmap_guide(const double *__restrict__ cum, long long numtab, int gbits,
                             int32_t *__restrict__ guide) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= numtab) return;
    const long long M = 1ll << gbits;
    // buckets b with b / M in (cum[t-1], cum[t]]; b / M and cum * M are exact
    const long long b0 = t == 0 ? 0 : (long long)floor(cum[t - 1] * (double)M) + 1;
    long long b1 = (long long)floor(cum[t] * (double)M);
    if (b1 > M - 1) b1 = M - 1;
    for (long long b = b0; b <= b1; b++) guide[b] = (int32_t)t;
    if (t == numtab - 1) guide[M] = (int32_t)t;
}

__global__ void k_bmap_search(BiasRef B, int n, const double *__restrict__ keys, int64_t *__restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = bias_index(B, keys[i]);
}

inline unsigned nblk(long long n, int t) { return (unsigned)((n + t - 1) / t); }

}  // namespace

extern "C" {

int mh_bmap_create(mh_ctx *ctx, const mh_map *pmap, double bias_temperature, mh_bmap **out) {
    MH_CHECK(ctx && pmap && out, MH_EINVAL, "mh_bmap_create: NULL argument");
    MH_CHECK(bias_temperature > 0.0, MH_EINVAL, "mh_bmap_create: bias temperature must be > 0");
    MH_CHECK(pmap->npnts[0] >= 2 && pmap->npnts[1] >= 2 && pmap->npnts[2] >= 2, MH_EINVAL,
             "mh_bmap_create: map grid too small");
    MH_CHECK(pmap->ptr.p && pmap->vals.p, MH_ESTATE, "mh_bmap_create: map has no compacted values");
    MH_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    std::unique_ptr<mh_bmap> b(new mh_bmap());
    b->ctx = ctx;
    for (int i = 0; i < 3; i++) {
        b->ncub[i] = pmap->npnts[i] - 1;
        b->step[i] = pmap->step[i];
        b->eff[i] = pmap->fund_ell[i];
    }
    const long long ncub = (long long)b->ncub[0] * b->ncub[1] * b->ncub[2];
    MH_CHECK(ncub < (1ll << 31), MH_ECAPACITY, "mh_bmap_create: more than 2^31 cubelets");
    b->ncubelets = ncub;
    DBuf u, wt, run, res;
    int rc;
    if ((rc = u.alloc(8 * ncub)) || (rc = wt.alloc(8 * ncub)) || (rc = res.alloc(64)) ||
        (rc = b->cindex.alloc(4 * ncub)))
        return rc;
    const double rti = 1.0 / (bias_temperature * (8.314e-3 / 4.184));
    hipLaunchKernelGGL(k_bmap_energy, dim3(nblk(ncub, 256)), dim3(256), 0, st, pmap->ptr.as<int32_t>(),
                       pmap->vals.as<float>(), pmap->npnts[1], pmap->npnts[2], b->ncub[1], b->ncub[2], ncub,
                       u.as<double>());
    hipLaunchKernelGGL(k_bmap_weight, dim3(nblk(ncub, 256)), dim3(256), 0, st, u.as<double>(), ncub, rti,
                       (double)INFINITY, wt.as<double>());
    hipLaunchKernelGGL(k_bmap_sum, dim3(1), dim3(64), 0, st, wt.as<double>(), ncub, res.as<double>());
    double h[3];
    MH_HIP(hipMemcpyAsync(h, res.p, 24, hipMemcpyDeviceToHost, st));
    MH_HIP(hipStreamSynchronize(st));
    MH_CHECK(h[2] > 0.0, MH_EINVAL, "mh_bmap_create: no cubelet of the map is below 30 kcal/mol");
    b->clamped = !(h[1] / h[0] > 1.0e-6);
    b->clamp_energy = INFINITY;
    if (b->clamped) {
        b->clamp_energy = -std::log(h[0] * 1.0e-6) / rti;
        hipLaunchKernelGGL(k_bmap_weight, dim3(nblk(ncub, 256)), dim3(256), 0, st, u.as<double>(), ncub, rti,
                           b->clamp_energy, wt.as<double>());
    }
    const long long numtab = (long long)h[2];
    b->numtab = numtab;
    if ((rc = b->cube.alloc(4 * numtab)) || (rc = run.alloc(8 * numtab)) || (rc = b->cum.alloc(8 * numtab)) ||
        (rc = b->wt.alloc(8 * numtab)) || (rc = b->wn.alloc(8 * numtab)))
        return rc;
    hipLaunchKernelGGL(k_bmap_list, dim3(1), dim3(64), 0, st, wt.as<double>(), ncub, b->cube.as<int32_t>(),
                       run.as<double>(), b->cindex.as<int32_t>(), res.as<double>());
    hipLaunchKernelGGL(k_bmap_prob, dim3(nblk(numtab, 256)), dim3(256), 0, st, run.as<double>(),
                       b->cube.as<int32_t>(), wt.as<double>(), numtab, res.as<double>(), b->cum.as<double>(),
                       b->wt.as<double>(), b->wn.as<double>());
    int gbits = 10;
    while (gbits < 20 && (1ll << gbits) < numtab) gbits++;
    b->gbits = gbits;
    if ((rc = b->guide.alloc(4 * ((1ll << gbits) + 1)))) return rc;
    hipLaunchKernelGGL(k_bmap_guide, dim3(nblk(numtab, 256)), dim3(256), 0, st, b->cum.as<double>(), numtab, gbits,
                       b->guide.as<int32_t>());
    MH_HIP(hipGetLastError());
    MH_HIP(hipMemcpyAsync(h, res.p, 16, hipMemcpyDeviceToHost, st));
    MH_HIP(hipStreamSynchronize(st));
    b->cum_wt = h[0];
    MH_CHECK((long long)h[1] == numtab, MH_EHIP, "mh_bmap_create: inconsistent cubelet count");
    ctx_retain(ctx);
    *out = b.release();
    return MH_OK;
}

int mh_bmap_destroy(mh_bmap *b) {
    if (!b) return MH_OK;
    mh_ctx *ctx = b->ctx;
    (void)hipSetDevice(ctx->device);
    delete b;
    ctx_release(ctx);
    return MH_OK;
}

int mh_bmap_info(const mh_bmap *b, int64_t *numtab, int64_t *ncubelets, double *cum_wt, double *clamp_energy) {
    MH_CHECK(b, MH_EINVAL, "mh_bmap_info: NULL bias map");
    if (numtab) *numtab = b->numtab;
    if (ncubelets) *ncubelets = b->ncubelets;
    if (cum_wt) *cum_wt = b->cum_wt;
    if (clamp_energy) *clamp_energy = b->clamp_energy;
    return MH_OK;
}

int mh_bmap_get(const mh_bmap *b, int32_t *cube, double *wt, double *cum_prob) {
    MH_CHECK(b, MH_EINVAL, "mh_bmap_get: NULL bias map");
    MH_HIP(hipSetDevice(b->ctx->device));
    const size_t n = (size_t)b->numtab;
    if (cube) MH_HIP(hipMemcpy(cube, b->cube.p, 4 * n, hipMemcpyDeviceToHost));
    if (wt) MH_HIP(hipMemcpy(wt, b->wt.p, 8 * n, hipMemcpyDeviceToHost));
    if (cum_prob) MH_HIP(hipMemcpy(cum_prob, b->cum.p, 8 * n, hipMemcpyDeviceToHost));
    return MH_OK;
}

int mh_bmap_search(const mh_bmap *b, int n, const double *keys, int64_t *index) {
    MH_CHECK(b && n >= 0 && (n == 0 || (keys && index)), MH_EINVAL, "mh_bmap_search: bad arguments");
    if (n == 0) return MH_OK;
    for (int i = 0; i < n; i++)
        MH_CHECK(keys[i] >= 0.0 && keys[i] < 1.0, MH_EINVAL, "mh_bmap_search: keys must lie in [0, 1)");
    MH_HIP(hipSetDevice(b->ctx->device));
    hipStream_t st = b->ctx->stream;
    DBuf dk, di;
    int rc;
    if ((rc = dk.alloc(8 * (size_t)n)) || (rc = di.alloc(8 * (size_t)n))) return rc;
    MH_HIP(hipMemcpyAsync(dk.p, keys, 8 * (size_t)n, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_bmap_search, dim3(nblk(n, 256)), dim3(256), 0, st, bias_ref(b), n, dk.as<double>(),
                       di.as<int64_t>());
    MH_HIP(hipGetLastError());
    MH_HIP(hipMemcpyAsync(index, di.p, 8 * (size_t)n, hipMemcpyDeviceToHost, st));
    MH_HIP(hipStreamSynchronize(st));
    return MH_OK;
}

}  // extern "C"
