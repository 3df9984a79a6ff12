// Device helpers shared by the map and GCMC kernels of libmusic_hip.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mh {

// tricubic Hermite at one point (interpolate_int + pothermite,
// interpolate.F90:104-158, 603-874); the 64-term sum runs corner-major in the
// reference order, FP32 table values widened to FP64.
static __device__ __forceinline__ double h0f(double x) { return x * x * (3.0 - 2.0 * x); }
static __device__ __forceinline__ double h1f(double x, double ds) { return ds * x * x * (x - 1.0); }

// A slanted fundamental cell (fundcell.F90:378-412): U maps slanted to
// Cartesian coordinates, S = U^-1, row-major; origin 0.
struct SlantCell {
    double U[9], S[9];
};

// matrix_m_mult_v (matrix.F90:137-150): row r = ((0 + m r1 v1) + m r2 v2) + m r3 v3
static __device__ __forceinline__ void mat_vec3(const double *m, const double *v, double *o) {
    for (int r = 0; r < 3; r++) {
        double s = 0.0;
        s = __dadd_rn(s, __dmul_rn(m[3 * r], v[0]));
        s = __dadd_rn(s, __dmul_rn(m[3 * r + 1], v[1]));
        s = __dadd_rn(s, __dmul_rn(m[3 * r + 2], v[2]));
        o[r] = s;
    }
}

static __device__ __forceinline__ bool hermite_point(const int32_t *__restrict__ ptr, const float *__restrict__ vals,
                              int nx, int ny, int nz, double sx, double sy, double sz,
                              double ex, double ey, double ez, double rx, double ry, double rz,
                              double &p, const SlantCell *sc = nullptr) {
    // simcell_maptouc (simcell.F90:634-668), origin 0: slanted cells fold in
    // their slanted basis, and without the periodic wrap of the upper corner
    // (interpolate.F90:675-679) a point in the last cubelet fails
    double c[3] = {rx, ry, rz};
    const double ell[3] = {ex, ey, ez};
    const double st[3] = {sx, sy, sz};
    const int nb[3] = {nx, ny, nz};
    int idx[3], idp[3];
    double tuv[3];
    if (sc) {
        double v[3];
        mat_vec3(sc->S, c, v);
        for (int d = 0; d < 3; d++) {
            double f = __dmul_rn(v[d], 1.0 / ell[d]);
            f = (f >= 0.0) ? fmod(f, 1.0) : __dadd_rn(fmod(f, 1.0), 1.0);
            v[d] = __dmul_rn(f, ell[d]);
        }
        mat_vec3(sc->U, v, c);
    } else {
        for (int d = 0; d < 3; d++) {
            double f = c[d] * (1.0 / ell[d]);
            f = (f >= 0.0) ? fmod(f, 1.0) : fmod(f, 1.0) + 1.0;
            c[d] = f * ell[d];
        }
    }
    for (int d = 0; d < 3; d++) {
        idx[d] = (int)(c[d] / st[d] + 1.0);  // mapheader_getIndices, Int()
        if (idx[d] < 1 || idx[d] > nb[d] || (sc && idx[d] == nb[d])) {
            p = 0.0;
            return false;
        }
        tuv[d] = (c[d] - (double)(idx[d] - 1) * st[d]) / st[d];
        idp[d] = (idx[d] == nb[d]) ? 1 : idx[d] + 1;
    }
#define P(i, j, k) ptr[(((long long)(i) - 1) * ny + ((j) - 1)) * nz + ((k) - 1)]
    const int ii = idx[0], jj = idx[1], kk = idx[2], ip = idp[0], jp = idp[1], kp = idp[2];
    int32_t mi[8] = {P(ii, jj, kk), P(ip, jj, kk), P(ii, jp, kk), P(ii, jj, kp),
                     P(ip, jp, kk), P(ip, jj, kp), P(ii, jp, kp), P(ip, jp, kp)};
#undef P
    for (int i = 0; i < 8; i++)
        if (mi[i] == 0) {
            p = 0.0;
            return false;
        }
    double hx[2][2], hy[2][2], hz[2][2];
    hx[0][0] = h0f(1 - tuv[0]); hx[1][0] = h0f(tuv[0]);
    hx[0][1] = h1f(1 - tuv[0], -sx); hx[1][1] = h1f(tuv[0], sx);
    hy[0][0] = h0f(1 - tuv[1]); hy[1][0] = h0f(tuv[1]);
    hy[0][1] = h1f(1 - tuv[1], -sy); hy[1][1] = h1f(tuv[1], sy);
    hz[0][0] = h0f(1 - tuv[2]); hz[1][0] = h0f(tuv[2]);
    hz[0][1] = h1f(1 - tuv[2], -sz); hz[1][1] = h1f(tuv[2], sz);
    // corner pattern and value-slot pattern (interpolate.F90:725-835)
    const int CX = 0xB2, CY = 0xD4, CZ = 0xE8;  // bit i: upper node along axis for corner i
    double sum = 0.0;
    for (int i = 0; i < 8; i++) {
        const float4 *r = reinterpret_cast<const float4 *>(vals + 8 * (long long)(mi[i] - 1));
        float4 a = r[0], b = r[1];
        const float f[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
        const int cx = (CX >> i) & 1, cy = (CY >> i) & 1, cz = (CZ >> i) & 1;
#pragma unroll
        for (int s = 0; s < 8; s++) {
            const int dx = (CX >> s) & 1, dy = (CY >> s) & 1, dz = (CZ >> s) & 1;
            sum = __dadd_rn(sum, __dmul_rn(__dmul_rn(__dmul_rn((double)f[s], hx[cx][dx]), hy[cy][dy]), hz[cz][dz]));
        }
    }
    p = sum;
    return true;
}


// A&S 7.1.26 erfc with single-precision coefficients (utils.F90:2090-2101)
static __device__ __forceinline__ double erfc_as64(double x) {
    const double a1 = (double)0.254829592f, a2 = (double)-0.284496736f, a3 = (double)1.421413741f,
                 a4 = (double)-1.453152027f, a5 = (double)1.061405429f, p = (double)0.3275911f;
    double t = 1.0 / __dadd_rn(1.0, __dmul_rn(p, x));
    double xsq = __dmul_rn(x, x);
    double tp = __dmul_rn(t, __dadd_rn(a1, __dmul_rn(t, __dadd_rn(a2, __dmul_rn(t, __dadd_rn(a3, __dmul_rn(t, __dadd_rn(a4, __dmul_rn(t, a5)))))))));
    return __dmul_rn(tp, exp(-xsq));
}

// hardware reciprocal (v_rcp_f32, within 1 ulp) for the FP32 pair paths; the
// correctly rounded __frcp_rn expands to a ten-instruction division
static __device__ __forceinline__ float rcp_fast(float x) { return __builtin_amdgcn_rcpf(x); }

static __device__ __forceinline__ float erfc_as32(float x) {
#pragma clang fp contract(fast)
    const float a1 = 0.254829592f, a2 = -0.284496736f, a3 = 1.421413741f, a4 = -1.453152027f,
                a5 = 1.061405429f, p = 0.3275911f;
    float t = rcp_fast(1.0f + p * x);
    return t * (a1 + t * (a2 + t * (a3 + t * (a4 + t * a5)))) * __expf(-x * x);
}

// fundcell_multMinImage orthorhombic single fold (fundcell.F90:676-690)
static __device__ __forceinline__ double fold64(double d, double L) {
    double h = 0.5 * L;
    if (d > h) d = d - L;
    else if (d < -h) d = d + L;
    return d;
}

}  // namespace mh
