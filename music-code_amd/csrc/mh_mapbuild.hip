// libmusic_hip: potential-map build kernels (LJ, Buckingham, Ewald real and
// reciprocal space) and the deterministic parallel compaction.
//
// Layout in HBM:
//   framework atoms        double [natoms][3] + int type + double q
//   full map (FP32)        float  [nodes][8]   scan order s = (x*ny + y)*nz + z
//   full map (FP64, opt.)  double [nodes][8]
//   Ewald reciprocal part  double [nodes][8]   node m = x + nx*(y + ny*z) (separable 3-stage DFT)
//   compact map            float  [nfsize][8]  32-byte records
//   ptr                    int32  [nx][ny][nz] z fastest, 1-based
//
// One workgroup of 256 lanes owns an 8x8x4 brick of grid nodes (one lane per
// node).  It culls the framework atoms against the brick's bounding box
// (lower bound of the minimum-image distance), keeps the survivors in LDS in
// their original order and streams them to every lane.
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "mh_device.h"
#include "mh_internal.h"

namespace mh {

constexpr int BX = 8, BY = 8, BZ = 4, NT = BX * BY * BZ;
constexpr int SEG = 512;   // atoms examined (and culled into LDS) per pass
// FP32 kernels: 480 keeps a workgroup's LDS (atom records + the FP64
// derivative sums) under 40 KB, i.e. four workgroups per CU
constexpr int SEG32 = 480;

constexpr double kPiR = 3.1415927;  // defaults.F90:33
constexpr double kTokcal = 332.064;                       // ewald.F90:60
constexpr double kVeryLrg = 1.0e10;                       // mapmaker.F90:57
constexpr double kLowCut2 = 1.0e-3 * 1.0e-3;              // ewald.F90:361-362

struct PairTab {
    int ntypes;
    double A[kMaxTypes], B[kMaxTypes], h2[kMaxTypes], l2[kMaxTypes];
    double bA[kMaxTypes], brho[kMaxTypes], bC[kMaxTypes], bD[kMaxTypes];
};

struct GridP {
    int nx, ny, nz;
    double step[3];
    double full[3];
    int natoms;
    double cull2;    // atoms farther than this from the brick are skipped
    double r2safe;   // FP32 path: pairs closer than this are evaluated in FP64
    double r2pot;    // FP32 Ewald: the potential term of pairs closer than this in FP64
    double kappa, charge1;
    float rc2f;      // FP32 Ewald real-space cutoff^2
    int refold;      // FP32 path: fold each separation again (cutoff near L/2)
    // slanted cell (FP64 kernel): atoms arrive in slanted coordinates S a,
    // a node p is slanted to S p, the difference folded on the full cell's
    // edges and unslanted by U (fundcell_multMinImage, fundcell.F90:640-709)
    int slanted;
    double U[9], S[9];
};

__device__ __forceinline__ double fold1(double d, double L) {
    // fundcell_multMinImage orthorhombic branch (fundcell.F90:676-690)
    double h = 0.5 * L;
    if (d > h) d = d - L;
    else if (d < -h) d = d + L;
    return d;
}

// lower bound of |d + kL| over d in [lo, hi], k in {-1,0,1}
__device__ __forceinline__ double ival_lb(double lo, double hi, double L) {
    double best = 1e300;
#pragma unroll
    for (int k = -1; k <= 1; k++) {
        double a = lo + k * L, b = hi + k * L;
        double d = (a > 0.0) ? a : ((b < 0.0) ? -b : 0.0);
        best = d < best ? d : best;
    }
    return best;
}

// ---------------------------------------------------------------- pairs
// lj_snglint with hot (lj.F90:652-733).  Returns 2 evaluated, 1 beyond
// hicut, 0 below locut.
__device__ __forceinline__ int lj_pair64(double A, double B, double h2, double l2, double sx,
                                         double sy, double sz, double &pot, double *h) {
    double r2 = __dadd_rn(__dadd_rn(__dmul_rn(sx, sx), __dmul_rn(sy, sy)), __dmul_rn(sz, sz));
    if (r2 > h2) return 1;
    if (r2 < l2) return 0;
    double r2i = 1.0 / r2;
    double r6i = __dmul_rn(__dmul_rn(r2i, r2i), r2i);
    double r12i = __dmul_rn(r6i, r6i);
    double nrg = __dsub_rn(__dmul_rn(A, r12i), __dmul_rn(B, r6i));
    pot = __dadd_rn(pot, nrg);
    double fac1 = __dmul_rn(__dsub_rn(__dmul_rn(__dmul_rn(12.0, A), r12i), __dmul_rn(__dmul_rn(6.0, B), r6i)), r2i);
    h[0] = __dadd_rn(h[0], -__dmul_rn(sx, fac1));
    h[1] = __dadd_rn(h[1], -__dmul_rn(sy, fac1));
    h[2] = __dadd_rn(h[2], -__dmul_rn(sz, fac1));
    double r10i = __dmul_rn(r12i, r2);
    double r16i = __dmul_rn(r10i, r6i);
    fac1 = __dsub_rn(__dmul_rn(__dmul_rn(168.0, A), r16i), __dmul_rn(__dmul_rn(48.0, B), r10i));
    h[3] = __dadd_rn(h[3], __dmul_rn(__dmul_rn(fac1, sx), sy));
    h[4] = __dadd_rn(h[4], __dmul_rn(__dmul_rn(fac1, sx), sz));
    h[5] = __dadd_rn(h[5], __dmul_rn(__dmul_rn(fac1, sy), sz));
    fac1 = __dmul_rn(__dadd_rn(__dmul_rn(__dmul_rn(-2688.0, A), r16i), __dmul_rn(__dmul_rn(480.0, B), r10i)), r2i);
    h[6] = __dadd_rn(h[6], __dmul_rn(__dmul_rn(__dmul_rn(fac1, sx), sy), sz));
    return 2;
}

// Buckingham closed form, radial derivatives f1 = V'/r, g2 = f1'/r, g3 = g2'/r
__device__ __forceinline__ int buck_pair64(double A, double rho, double C, double D, double h2,
                                           double l2, double sx, double sy, double sz,
                                           double &pot, double *h) {
    double r2 = sx * sx + sy * sy + sz * sz;
    if (r2 > h2) return 1;
    if (r2 < l2) return 0;
    double r = sqrt(r2);
    double ri = 1.0 / r, r2i = ri * ri, r6i = r2i * r2i * r2i, r8i = r6i * r2i;
    double e = A * exp(-r / rho), ir = 1.0 / rho;
    double v = e - C * r6i - D * r8i;
    double f1 = -e * ir * ri + 6.0 * C * r8i + 8.0 * D * r8i * r2i;
    double g2 = e * ir * (ir * r2i + r2i * ri) - 48.0 * C * r8i * r2i - 80.0 * D * r8i * r2i * r2i;
    double g3 = -e * ir * (ir * ir * r2i * ri + 3.0 * ir * r2i * r2i + 3.0 * r2i * r2i * ri) +
                480.0 * C * r6i * r6i + 960.0 * D * r8i * r6i;
    pot += v;
    h[0] += sx * f1;
    h[1] += sy * f1;
    h[2] += sz * f1;
    h[3] += g2 * sx * sy;
    h[4] += g2 * sx * sz;
    h[5] += g2 * sy * sz;
    h[6] += g3 * sx * sy * sz;
    return 2;
}

// A&S 7.1.26 erfc with single-precision coefficients (utils.F90:2090-2101)
__device__ __forceinline__ double erfc_ref64(double x) {
    const double a1 = (double)0.254829592f, a2 = (double)-0.284496736f, a3 = (double)1.421413741f,
                 a4 = (double)-1.453152027f, a5 = (double)1.061405429f, p = (double)0.3275911f;
    double t = 1.0 / __dadd_rn(1.0, __dmul_rn(p, x));
    double xsq = __dmul_rn(x, x);
    double tp = __dmul_rn(t, __dadd_rn(a1, __dmul_rn(t, __dadd_rn(a2, __dmul_rn(t, __dadd_rn(a3, __dmul_rn(t, __dadd_rn(a4, __dmul_rn(t, a5)))))))));
    return __dmul_rn(tp, exp(-xsq));
}

// ewald_rwaldHOT per atom (ewald.F90:874-916); h[0..6] = hot(2..8)
__device__ __forceinline__ void ewald_real_pair64(double kappa, double charge1, double q2,
                                                  double sx, double sy, double sz, double d2,
                                                  double &pot, double *h) {
    const double risqrtpi = 0.5641896;
    double dist = sqrt(d2);
    double kr = kappa * dist;
    double erfkr = erfc_ref64(kr);
    double ri1 = 1.0 / dist;
    pot = __dadd_rn(pot, __dmul_rn(__dmul_rn(__dmul_rn(charge1, q2), erfkr), ri1));
    double ri2 = ri1 * ri1, ri3 = ri2 * ri1;
    double ek2r2 = exp(-(kr * kr));
    double kappa3 = kappa * kappa * kappa;
    double temp = __dsub_rn(-(ri3 * erfkr), __dmul_rn(__dmul_rn(__dmul_rn(2.0 * kappa, ri2), risqrtpi), ek2r2));
    double qt = __dmul_rn(q2, temp);
    h[0] = __dadd_rn(h[0], __dmul_rn(__dmul_rn(sx, qt), charge1));
    h[1] = __dadd_rn(h[1], __dmul_rn(__dmul_rn(sy, qt), charge1));
    h[2] = __dadd_rn(h[2], __dmul_rn(__dmul_rn(sz, qt), charge1));
    double ri4 = ri3 * ri1, ri5 = ri4 * ri1;
    temp = __dadd_rn(__dadd_rn(__dmul_rn(3.0 * ri5, erfkr),
                               __dmul_rn(__dmul_rn(__dmul_rn(6.0 * kappa, risqrtpi), ri4), ek2r2)),
                     __dmul_rn(__dmul_rn(__dmul_rn(4.0 * kappa3, risqrtpi), ri2), ek2r2));
    double cq = __dmul_rn(charge1, q2);
    h[3] = __dadd_rn(h[3], __dmul_rn(__dmul_rn(__dmul_rn(cq, sx), sy), temp));
    h[4] = __dadd_rn(h[4], __dmul_rn(__dmul_rn(__dmul_rn(cq, sx), sz), temp));
    h[5] = __dadd_rn(h[5], __dmul_rn(__dmul_rn(__dmul_rn(cq, sy), sz), temp));
    double ri6 = ri5 * ri1, ri7 = ri6 * ri1;
    temp = __dsub_rn(__dsub_rn(__dsub_rn(-(15.0 * ri7 * erfkr),
                                         __dmul_rn(__dmul_rn(__dmul_rn(30.0 * kappa, risqrtpi), ri6), ek2r2)),
                               __dmul_rn(__dmul_rn(__dmul_rn(20.0 * kappa3, risqrtpi), ri4), ek2r2)),
                     __dmul_rn(__dmul_rn(__dmul_rn(8.0 * kappa3 * kappa * kappa, risqrtpi), ri2), ek2r2));
    h[6] = __dadd_rn(h[6], __dmul_rn(__dmul_rn(__dmul_rn(__dmul_rn(cq, sx), sy), sz), temp));
}

// FP32 Ewald real-space pair: same expression tree in single precision,
// returns the 8 terms (t[0] potential, t[1..7] derivatives)
// FP32 Ewald real-space pair (ewald_rwaldHOT terms, ewald.F90:875-916) with
// u = erfc(kr)/r and g = (2 kappa / sqrt(pi)) e^{-k^2 r^2}:
//   V = q u, (1/r) dV/dr = -q (u + g) / r^2,
//   d2 terms  q (3 (u + g) + 2 k^2 r^2 g) / r^4,
//   d3 term  -q (15 (u + g) + (10 k^2 r^2 + 4 k^4 r^4) g) / r^6;
// the potential goes straight into the FP64 sum (terms cancel strongly),
// the 7 derivatives accumulate into the FP32 tile sums with FMAs.
// Constants of the FP32 real-space pair, formed once per launch from kappa
// (workgroup-uniform, so they stay in SGPRs): kap2 = kappa^2, cexp =
// -kappa^2 log2(e) (exp(-k^2 r^2) = exp2(cexp r^2)), pk = p kappa of the
// A&S 7.1.26 argument, twoksp = 2 kappa / sqrt(pi) with the reference's
// rounded 1/sqrt(pi) = 0.5641896.
struct EwC {
    float kap2, cexp, pk, twoksp, a5;
};

__device__ __forceinline__ EwC ewc_make(double kappa) {
    EwC c;
    c.kap2 = (float)(kappa * kappa);
    c.cexp = (float)(-kappa * kappa * 1.4426950408889634);
    c.pk = (float)(0.3275911 * kappa);
    c.twoksp = (float)(2.0 * 0.5641896 * kappa);
    c.a5 = 1.061405429f;  // leading A&S coefficient, held in a register (one literal per VOP2)
    return c;
}

// q erfc(kr)/r (A&S 7.1.26, as erfc_ref64) and its derivatives at the
// separation (sx, sy, sz), r2 = |s|^2: the potential added to the FP64 node
// sum, the 7 derivatives to FP32 tile partials.  With s1 = u + g, g =
// (2 kappa / sqrt(pi)) q e^{-k^2 r^2}, w = 1/r^2:
//   d/dx = -s1 w x;  d2/dxdy = (3 s1 + 2 k^2 r^2 g) w^2 x y;
//   d3/dxdydz = -(15 s1 + (10 + 4 k^2 r^2) k^2 r^2 g) w^3 x y z.
__device__ __forceinline__ float ewald_real_acc32(const EwC &c, float qq, float sx, float sy, float sz, float r2,
                                                  float *hf) {
    const float a1 = 0.254829592f, a2 = -0.284496736f, a3 = 1.421413741f, a4 = -1.453152027f;
    const float ri1 = __frsqrt_rn(r2);
    const float r = r2 * ri1;
    const float tt = rcp_fast(fmaf(c.pk, r, 1.0f));
    const float k2r2 = c.kap2 * r2;
    const float ek = __builtin_amdgcn_exp2f(c.cexp * r2);
    const float qek = qq * ek;
    const float poly = tt * fmaf(tt, fmaf(tt, fmaf(tt, fmaf(tt, c.a5, a4), a3), a2), a1);
    const float pu = poly * qek;
    const float u = pu * ri1;  // q erfc(kr)/r
    const float g = c.twoksp * qek;
    const float s1 = fmaf(pu, ri1, g);
    const float w = ri1 * ri1;
    const float t1 = -s1 * w;
    hf[0] = fmaf(sx, t1, hf[0]);
    hf[1] = fmaf(sy, t1, hf[1]);
    hf[2] = fmaf(sz, t1, hf[2]);
    const float w2 = w * w;
    const float gk = g * k2r2;
    const float t2 = fmaf(2.0f, gk, 3.0f * s1) * w2;
    const float sxy = sx * sy;
    hf[3] = fmaf(t2, sxy, hf[3]);
    const float sxt2 = sx * t2;
    hf[4] = fmaf(sxt2, sz, hf[4]);
    const float syt2 = sy * t2;
    hf[5] = fmaf(syt2, sz, hf[5]);
    const float t3 = fmaf(fmaf(4.0f, k2r2, 10.0f), gk, 15.0f * s1) * (w2 * w);
    const float t3xy = t3 * sxy;
    hf[6] = fmaf(-t3xy, sz, hf[6]);
    return u;
}

// A near pair of the FP32 Ewald kernel in FP64: returns q erfc(kr)/r
// (A&S 7.1.26 as erfc_ref64) and adds the 7 derivative terms to col(i, v)
// one at a time, in the factored form of ewald_real_acc32 (few live values).
template <class Col>
__device__ __forceinline__ double ewald_near64(double kappa, double qq, double sx, double sy, double sz, double d2,
                                               Col col) {
    const double a1 = (double)0.254829592f, a2 = (double)-0.284496736f, a3 = (double)1.421413741f,
                 a4 = (double)-1.453152027f, a5 = (double)1.061405429f, p = (double)0.3275911f;
    const double r = sqrt(d2);
    const double kr = kappa * r;
    const double t = 1.0 / (1.0 + p * kr);
    const double ek = exp(-(kr * kr));
    const double ri = 1.0 / r;
    const double u = qq * (t * (a1 + t * (a2 + t * (a3 + t * (a4 + t * a5))))) * ek * ri;
    const double g = 2.0 * 0.5641896 * kappa * qq * ek;
    const double s1 = u + g;
    const double w = ri * ri;
    const double t1 = -s1 * w;
    col(0, sx * t1);
    col(1, sy * t1);
    col(2, sz * t1);
    const double k2r2 = kappa * kappa * d2;
    const double t2 = (3.0 * s1 + 2.0 * k2r2 * g) * (w * w);
    col(3, t2 * sx * sy);
    col(4, t2 * sx * sz);
    col(5, t2 * sy * sz);
    const double t3 = -(15.0 * s1 + (10.0 + 4.0 * k2r2) * k2r2 * g) * (w * w * w);
    col(6, t3 * sx * sy * sz);
    return u;
}

// q erfc(kr)/r of one pair in FP64 (the potential term of
// ewald_rwaldHOT, ewald.F90:874-916, as ewald_real_pair64)
__device__ __forceinline__ double ewald_real_pot64(double kappa, double charge1, double q2, double d2) {
    const double dist = sqrt(d2);
    const double erfkr = erfc_ref64(kappa * dist);
    const double ri1 = 1.0 / dist;
    return __dmul_rn(__dmul_rn(__dmul_rn(charge1, q2), erfkr), ri1);
}

// FP32 LJ pair with derivatives
__device__ __forceinline__ void lj_pair32(float A, float B, float sx, float sy, float sz, float r2,
                                          float &pot, float *h) {
    // with X = A/r^12, Y = B/r^6 and w = 1/r^2:
    //   V = X - Y;  dV/dx = -(12 X - 6 Y) w x;
    //   d2V/dxdy = (168 X - 48 Y) w^2 x y;  d3V/dxdydz = (480 Y - 2688 X) w^3 x y z
    const float w = rcp_fast(r2);
    const float w2 = w * w;
    const float r6i = w2 * w;
    const float X = A * (r6i * r6i);
    const float Y = B * r6i;
    pot = (pot + X) - Y;
    const float fac1 = fmaf(12.0f, X, -6.0f * Y) * w;
    h[0] = fmaf(-sx, fac1, h[0]);
    h[1] = fmaf(-sy, fac1, h[1]);
    h[2] = fmaf(-sz, fac1, h[2]);
    const float f2 = fmaf(168.0f, X, -48.0f * Y) * w2;
    const float f2x = f2 * sx;
    h[3] = fmaf(f2x, sy, h[3]);
    h[4] = fmaf(f2x, sz, h[4]);
    const float f2y = f2 * sy;
    h[5] = fmaf(f2y, sz, h[5]);
    const float f3 = fmaf(-2688.0f, X, 480.0f * Y) * (w2 * w);
    const float f3x = f3 * sx;
    h[6] = fmaf(f3x * sy, sz, h[6]);
}

__device__ __forceinline__ void kahan_add(double &s, double &c, double v) {
    double y = v - c;
    double t = s + y;
    c = (t - s) - y;
    s = t;
}

enum { K_LJ = 1, K_EWALD = 2, K_BUCK = 3 };

// ------------------------------------------------------- node kernels
// FP64 node kernel: reference arithmetic, atoms visited in their order
// (culled atoms are beyond every cutoff of the brick and contribute +0).
// Ewald: real space only; the reciprocal part is added from recip[].
template <int KIND>
__global__ void __launch_bounds__(NT) k_nodes64(GridP g, const PairTab *__restrict__ tab_g,
                                                 const double *__restrict__ xyz,
                                                 const int *__restrict__ typ,
                                                 const double *__restrict__ qv,
                                                 const double *__restrict__ recip,
                                                 float *__restrict__ out32,
                                                 double *__restrict__ out64) {
    __shared__ double s_x[SEG], s_y[SEG], s_z[SEG], s_q[SEG];
    __shared__ int s_t[SEG];
    __shared__ int s_wtot[NT / 64];
    __shared__ PairTab tab;
    const int tid = threadIdx.x;
    if (KIND != K_EWALD) {
        for (int i = tid; i < (int)(sizeof(PairTab) / sizeof(double)); i += NT)
            reinterpret_cast<double *>(&tab)[i] = reinterpret_cast<const double *>(tab_g)[i];
    }
    const int nbx = (g.nx + BX - 1) / BX, nby = (g.ny + BY - 1) / BY;
    const int b = blockIdx.x;
    const int bx = b % nbx, by = (b / nbx) % nby, bz = b / (nbx * nby);
    // lanes run along z first: the node values are stored in the map's scan
    // order (z fastest), four consecutive z of a column per 128-byte line
    const int z = bz * BZ + (tid % BZ), x = bx * BX + ((tid / BZ) % BX), y = by * BY + tid / (BZ * BX);
    const bool valid = (x < g.nx) && (y < g.ny) && (z < g.nz);
    double px = (double)x * g.step[0], py = (double)y * g.step[1], pz = (double)z * g.step[2];
    if (g.slanted) {
        const double pc[3] = {px, py, pz};
        double ps[3];
        mat_vec3(g.S, pc, ps);
        px = ps[0];
        py = ps[1];
        pz = ps[2];
    }
    // brick bounding box
    const double lo[3] = {(double)(bx * BX) * g.step[0], (double)(by * BY) * g.step[1], (double)(bz * BZ) * g.step[2]};
    const double hi[3] = {(double)min(bx * BX + BX - 1, g.nx - 1) * g.step[0],
                          (double)min(by * BY + BY - 1, g.ny - 1) * g.step[1],
                          (double)min(bz * BZ + BZ - 1, g.nz - 1) * g.step[2]};
    __syncthreads();
    double pot = 0.0, h[7] = {0, 0, 0, 0, 0, 0, 0};
    bool failed = false;
    for (int seg0 = 0; seg0 < g.natoms; seg0 += SEG) {
        const int segn = min(SEG, g.natoms - seg0);
        int cnt = 0;
        for (int c0 = 0; c0 < segn; c0 += NT) {
            int a = seg0 + c0 + tid;
            bool keep = false;
            double ax = 0, ay = 0, az = 0;
            if (c0 + tid < segn) {
                ax = xyz[3 * a];
                ay = xyz[3 * a + 1];
                az = xyz[3 * a + 2];
                double lx = ival_lb(lo[0] - ax, hi[0] - ax, g.full[0]);
                double ly = ival_lb(lo[1] - ay, hi[1] - ay, g.full[1]);
                double lz = ival_lb(lo[2] - az, hi[2] - az, g.full[2]);
                keep = (lx * lx + ly * ly + lz * lz) <= g.cull2;
            }
            unsigned long long msk = __ballot(keep);
            const int lane = tid & 63, wave = tid >> 6;
            if (lane == 0) s_wtot[wave] = __popcll(msk);
            __syncthreads();
            int off = cnt, add = 0;
            for (int w = 0; w < NT / 64; w++) {
                if (w < wave) off += s_wtot[w];
                add += s_wtot[w];
            }
            if (keep) {
                int pos = off + __popcll(msk & ((1ull << lane) - 1ull));
                s_x[pos] = ax;
                s_y[pos] = ay;
                s_z[pos] = az;
                s_t[pos] = typ[a];
                s_q[pos] = (KIND == K_EWALD) ? qv[a] : 0.0;
            }
            cnt += add;
            __syncthreads();
        }
        if (valid && !failed) {
            for (int j = 0; j < cnt; j++) {
                double sx = fold1(px - s_x[j], g.full[0]);
                double sy = fold1(py - s_y[j], g.full[1]);
                double sz = fold1(pz - s_z[j], g.full[2]);
                if (g.slanted) {
                    const double d[3] = {sx, sy, sz};
                    double c[3];
                    mat_vec3(g.U, d, c);
                    sx = c[0];
                    sy = c[1];
                    sz = c[2];
                }
                if (KIND == K_LJ) {
                    int t = s_t[j];
                    double hh[7] = {0, 0, 0, 0, 0, 0, 0};
                    int rc = lj_pair64(tab.A[t], tab.B[t], tab.h2[t], tab.l2[t], sx, sy, sz, pot, hh);
                    for (int i = 0; i < 7; i++) h[i] = __dadd_rn(h[i], hh[i]);
                    if (rc == 0) { failed = true; break; }
                } else if (KIND == K_BUCK) {
                    int t = s_t[j];
                    int rc = buck_pair64(tab.bA[t], tab.brho[t], tab.bC[t], tab.bD[t], tab.h2[t],
                                         tab.l2[t], sx, sy, sz, pot, h);
                    if (rc == 0) { failed = true; break; }
                } else {
                    double d2 = __dadd_rn(__dadd_rn(__dmul_rn(sx, sx), __dmul_rn(sy, sy)), __dmul_rn(sz, sz));
                    if (d2 < kLowCut2) { failed = true; break; }
                    ewald_real_pair64(g.kappa, g.charge1, s_q[j], sx, sy, sz, d2, pot, h);
                }
            }
        }
        __syncthreads();
    }
    if (!valid) return;
    // m: config_makegrid order (x fastest) of the reciprocal-space table;
    // s: the map's scan order (z fastest) of the node values
    const long long m = (long long)x + (long long)g.nx * ((long long)y + (long long)g.ny * z);
    const long long s = ((long long)x * g.ny + y) * g.nz + z;
    double o[8];
    if (KIND == K_EWALD) {
        if (failed) {
            o[0] = kVeryLrg;
            for (int i = 1; i < 8; i++) o[i] = 0.0;
        } else {
            const double *rp = recip + 8 * m;
            o[0] = __dmul_rn(__dadd_rn(pot, rp[0]), kTokcal);
            for (int i = 0; i < 7; i++) o[i + 1] = __dmul_rn(__dadd_rn(h[i], rp[i + 1]), kTokcal);
        }
    } else {
        o[0] = failed ? kVeryLrg : pot;
        for (int i = 0; i < 7; i++) o[i + 1] = h[i];
    }
    float4 *d32 = reinterpret_cast<float4 *>(out32 + 8 * s);
    d32[0] = make_float4((float)o[0], (float)o[1], (float)o[2], (float)o[3]);
    d32[1] = make_float4((float)o[4], (float)o[5], (float)o[6], (float)o[7]);
    if (out64) {
        double2 *d64 = reinterpret_cast<double2 *>(out64 + 8 * s);
        for (int i = 0; i < 4; i++) d64[i] = make_double2(o[2 * i], o[2 * i + 1]);
    }
}

// FP32 node kernel (LJ or Ewald real space).  Separations are formed
// relative to the brick centre c: sep = (p - c) + fold(c - a), both parts
// rounded to FP32 (|p - c| is under a brick half-diagonal, so the FP32
// separation of a near atom keeps ~1e-7 A absolute accuracy).  When a cutoff
// plus the brick half-diagonal reaches half the cell, each component is
// folded again (same single fold as the reference).  FP32 pair terms are
// summed in FP32 over tiles of 32 atoms and the tiles in FP64 (the Ewald
// potential term goes to the FP64 sum per pair: its terms cancel strongly).
// Pairs closer than sqrt(r2safe) and LJ pairs within rounding of hicut are
// evaluated with the FP64 routine (reference cutoff decisions).
__device__ __forceinline__ float foldf(float d, float L) {
    float h = 0.5f * L;
    d = (d > h) ? d - L : d;
    d = (d < -h) ? d + L : d;
    return d;
}

// a workgroup-uniform double kept in SGPRs (frees VGPRs in the pair loop)
__device__ __forceinline__ double uniform_double(double v) {
    const unsigned long long u = __double_as_longlong(v);
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)u);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(u >> 32));
    return __longlong_as_double(((unsigned long long)hi << 32) | lo);
}

// LJ pair coefficients of one listed atom (FP32, formed once per brick):
// with u = 1/r^6 and w = 1/r^2,
//   V      = (A u - B) u
//   dV/dx  = -(12 A u - 6 B) u w x
//   d2V/dxdy = (168 A u - 48 B) u w^2 x y
//   d3V/dxdydz = (480 B - 2688 A u) u w^3 x y z
// c1 = (A, -B, 12 A, -6 B), c2 = (168 A, -48 B, -2688 A, 480 B).
struct LJRec {
    float4 p;       // x, y, z relative to the brick centre, hicut^2 (1 - 1e-6)
    float4 c1, c2;  // pair coefficients
    float4 e;       // hicut^2 (1 + 1e-6), unused
};

// at most 128 VGPRs: four waves per SIMD hide the transcendental and LDS latency
template <int KIND>
__global__ void __launch_bounds__(NT, 4) k_nodes32(GridP g, const PairTab *__restrict__ tab_g,
                                                 const double *__restrict__ xyz,
                                                 const int *__restrict__ typ,
                                                 const double *__restrict__ qv,
                                                 const double *__restrict__ recip,
                                                 float *__restrict__ out32,
                                                 double *__restrict__ out64) {
    // culled atoms, one LDS record per atom: (x, y, z relative to the brick
    // centre, p0) with p0 = hicut^2 (1 - 1e-6) for LJ and the charge product
    // for Ewald; LJ adds its pair coefficients and hicut^2 (1 + 1e-6) with the
    // atom index (the FP64 path's band around the cutoff)
    __shared__ float4 s_a4[KIND == K_LJ ? 1 : SEG32];
    __shared__ LJRec s_lj[KIND == K_LJ ? SEG32 : 1];
    __shared__ int s_a[SEG32];
    __shared__ int s_wtot[NT / 64];
    // Ewald: FP64 sums of the 7 derivatives, one lane-private column per node
    // (keeps them out of the VGPR budget of the longer Ewald pair)
    __shared__ double s_acc[KIND == K_EWALD ? 7 : 1][NT];
    __shared__ PairTab tab;
    const int tid = threadIdx.x;
    if (KIND == K_LJ) {
        for (int i = tid; i < (int)(sizeof(PairTab) / sizeof(double)); i += NT)
            reinterpret_cast<double *>(&tab)[i] = reinterpret_cast<const double *>(tab_g)[i];
        __syncthreads();
    }
    const int nbx = (g.nx + BX - 1) / BX, nby = (g.ny + BY - 1) / BY;
    const int b = blockIdx.x;
    const int bx = b % nbx, by = (b / nbx) % nby, bz = b / (nbx * nby);
    // lanes run along z first: the node values are stored in the map's scan
    // order (z fastest), four consecutive z of a column per 128-byte line
    const int z = bz * BZ + (tid % BZ), x = bx * BX + ((tid / BZ) % BX), y = by * BY + tid / (BZ * BX);
    const bool valid = (x < g.nx) && (y < g.ny) && (z < g.nz);
    const double px = (double)x * g.step[0], py = (double)y * g.step[1], pz = (double)z * g.step[2];
    double lo[3] = {(double)(bx * BX) * g.step[0], (double)(by * BY) * g.step[1], (double)(bz * BZ) * g.step[2]};
    double hi[3] = {(double)min(bx * BX + BX - 1, g.nx - 1) * g.step[0],
                    (double)min(by * BY + BY - 1, g.ny - 1) * g.step[1],
                    (double)min(bz * BZ + BZ - 1, g.nz - 1) * g.step[2]};
    double cen[3] = {0.5 * (lo[0] + hi[0]), 0.5 * (lo[1] + hi[1]), 0.5 * (lo[2] + hi[2])};
    for (int d = 0; d < 3; d++) {  // brick bounds are workgroup-uniform: SGPRs
        lo[d] = uniform_double(lo[d]);
        hi[d] = uniform_double(hi[d]);
        cen[d] = uniform_double(cen[d]);
    }
    const float dx0 = (float)(px - cen[0]), dy0 = (float)(py - cen[1]), dz0 = (float)(pz - cen[2]);
    const float r2safe = (float)g.r2safe, r2pot = (float)g.r2pot;
    const EwC ewc = ewc_make(g.kappa);
    const float Lx = (float)g.full[0], Ly = (float)g.full[1], Lz = (float)g.full[2];
    const bool refold = g.refold != 0;
    // FP64 node sums: FP32 tile partials (32 atoms) and the FP64 near pairs;
    // the potential in a register, the derivatives in registers (LJ) or in
    // s_acc (Ewald)
    double S[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (KIND == K_EWALD)
        for (int i = 0; i < 7; i++) s_acc[i][tid] = 0.0;
    // (volatile: the Ewald columns stay in LDS, out of the registers of the
    // longer pair loop)
    volatile double *acc = &s_acc[0][tid];
    auto dsum = [&](int i, double v) {  // derivative column i (0..6)
        if (KIND == K_EWALD) acc[i * NT] = acc[i * NT] + v;
        else S[i + 1] += v;
    };
    __syncthreads();
    bool failed = false;
    for (int seg0 = 0; seg0 < g.natoms; seg0 += SEG32) {
        const int segn = min(SEG32, g.natoms - seg0);
        int cnt = 0;
        for (int c0 = 0; c0 < segn; c0 += NT) {
            int a = seg0 + c0 + tid;
            bool keep = false;
            double ax = 0, ay = 0, az = 0;
            if (c0 + tid < segn) {
                ax = xyz[3 * a];
                ay = xyz[3 * a + 1];
                az = xyz[3 * a + 2];
                double lx = ival_lb(lo[0] - ax, hi[0] - ax, g.full[0]);
                double ly = ival_lb(lo[1] - ay, hi[1] - ay, g.full[1]);
                double lz = ival_lb(lo[2] - az, hi[2] - az, g.full[2]);
                keep = (lx * lx + ly * ly + lz * lz) <= g.cull2;
            }
            unsigned long long m = __ballot(keep);
            const int lane = tid & 63, wave = tid >> 6;
            if (lane == 0) s_wtot[wave] = __popcll(m);
            __syncthreads();
            int off = cnt, add = 0;
            for (int w = 0; w < NT / 64; w++) {
                if (w < wave) off += s_wtot[w];
                add += s_wtot[w];
            }
            if (keep) {
                int pos = off + __popcll(m & ((1ull << lane) - 1ull));
                const float hx = (float)fold1(cen[0] - ax, g.full[0]);
                const float hy = (float)fold1(cen[1] - ay, g.full[1]);
                const float hz = (float)fold1(cen[2] - az, g.full[2]);
                s_a[pos] = a;
                if (KIND == K_LJ) {
                    const int t = typ[a];
                    const float A = (float)tab.A[t], B = (float)tab.B[t];
                    const float h2f = (float)tab.h2[t];
                    const float band = h2f * 1.0e-6f;
                    LJRec rec;
                    rec.p = make_float4(hx, hy, hz, h2f - band);
                    rec.c1 = make_float4(A, -B, 12.0f * A, -6.0f * B);
                    rec.c2 = make_float4(168.0f * A, -48.0f * B, -2688.0f * A, 480.0f * B);
                    rec.e = make_float4(h2f + band, 0.0f, 0.0f, 0.0f);
                    s_lj[pos] = rec;
                } else {
                    s_a4[pos] = make_float4(hx, hy, hz, (float)(g.charge1 * qv[a]));
                }
            }
            cnt += add;
            __syncthreads();
        }
        if (valid && !failed) {
            for (int j0 = 0; j0 < cnt && !failed; j0 += 32) {
                float hf[8] = {0, 0, 0, 0, 0, 0, 0, 0};
                const int j1 = min(cnt, j0 + 32);
                // one listed atom against the lane's node (unrolled by two
                // below: the LDS reads of a pair share one address)
                auto pair = [&](const int j, auto refold_c) {
                    const float4 a4 = (KIND == K_LJ) ? s_lj[j].p : s_a4[j];
                    float sx = dx0 + a4.x, sy = dy0 + a4.y, sz = dz0 + a4.z;
                    if (decltype(refold_c)::value) {
                        sx = foldf(sx, Lx);
                        sy = foldf(sy, Ly);
                        sz = foldf(sz, Lz);
                    }
                    const float r2 = fmaf(sx, sx, fmaf(sy, sy, sz * sz));
                    if (KIND == K_LJ) {
                        // FP32 pairs: r2safe <= r2 <= hicut^2 (1 - 1e-6); the
                        // near pairs and the band around hicut go to FP64
                        const float hhi = s_lj[j].e.x;
                        const bool in32 = (r2 <= a4.w) & !(r2 < r2safe);
                        const bool near = !in32 & (r2 <= hhi);
                        if (__builtin_expect(__any(near), 0)) {
                            if (near) {
                                const int a = s_a[j];
                                const int t = typ[a];
                                // node position recomputed here (not kept live in the loop)
                                const double ex = fold1((double)x * g.step[0] - xyz[3 * a], g.full[0]);
                                const double ey = fold1((double)y * g.step[1] - xyz[3 * a + 1], g.full[1]);
                                const double ez = fold1((double)z * g.step[2] - xyz[3 * a + 2], g.full[2]);
                                double pot = 0.0, h[7] = {0, 0, 0, 0, 0, 0, 0};
                                // no early exit from the tile: a failed node is VERY_LRG_NRG and its
                                // derivative columns (unlike the FP64 kernel's) sum every other pair
                                if (lj_pair64(tab.A[t], tab.B[t], tab.h2[t], tab.l2[t], ex, ey, ez, pot, h)) {
                                    S[0] += pot;
#pragma unroll
                                    for (int i = 0; i < 7; i++) dsum(i, h[i]);
                                } else {
                                    failed = true;
                                }
                            }
                        }
                        // w = 1/r^2, u = w^3: V = g0 u, first derivatives
                        // g1 w^4, second g2 w^5, third g3 u^2
                        const LJRec rc = s_lj[j];
                        const float w = in32 ? rcp_fast(r2) : 0.0f;
                        const float w2 = w * w;
                        const float u = w2 * w;
                        const float w4 = w2 * w2;
                        const float g0 = fmaf(rc.c1.x, u, rc.c1.y);
                        const float g1 = fmaf(rc.c1.z, u, rc.c1.w);
                        const float g2 = fmaf(rc.c2.x, u, rc.c2.y);
                        const float g3 = fmaf(rc.c2.z, u, rc.c2.w);
                        hf[0] = fmaf(g0, u, hf[0]);
                        const float f1 = g1 * w4;
                        hf[1] = fmaf(-sx, f1, hf[1]);
                        hf[2] = fmaf(-sy, f1, hf[2]);
                        hf[3] = fmaf(-sz, f1, hf[3]);
                        const float f2 = g2 * (w4 * w);
                        const float sxy = sx * sy;
                        hf[4] = fmaf(f2, sxy, hf[4]);
                        const float f2z = f2 * sz;
                        hf[5] = fmaf(f2z, sx, hf[5]);
                        hf[6] = fmaf(f2z, sy, hf[6]);
                        const float f3 = g3 * (u * u);
                        hf[7] = fmaf(f3 * sxy, sz, hf[7]);
                    } else {
                        if (r2 > g.rc2f) return;
                        // pairs within sqrt(r2safe): every term in FP64; within
                        // sqrt(r2pot): the potential term in FP64 (the largest
                        // terms, whose FP32 rounding dominates V), the
                        // derivatives in FP32; beyond: FP32
                        const bool full64 = r2 < r2safe;
                        const float u = ewald_real_acc32(ewc, full64 ? 0.0f : a4.w, sx, sy, sz, r2, hf + 1);
                        if (r2 < r2pot) {
                            const int a = s_a[j];
                            // node position recomputed here (not kept live in the loop)
                            const double ex = fold1((double)x * g.step[0] - xyz[3 * a], g.full[0]);
                            const double ey = fold1((double)y * g.step[1] - xyz[3 * a + 1], g.full[1]);
                            const double ez = fold1((double)z * g.step[2] - xyz[3 * a + 2], g.full[2]);
                            const double d2 = ex * ex + ey * ey + ez * ez;
                            // no early exit from the tile: a failed node is VERY_LRG_NRG
                            if (d2 < kLowCut2) {
                                failed = true;
                            } else if (full64) {
                                S[0] += ewald_near64(g.kappa, g.charge1 * qv[a], ex, ey, ez, d2, dsum);
                            } else {
                                S[0] += ewald_real_pot64(g.kappa, g.charge1, qv[a], d2);
                            }
                        } else {
                            S[0] += (double)u;
                        }
                    }
                                };
                // the refold test is workgroup-uniform: one loop per case
                auto tile = [&](auto refold_c) {
                    int j = j0;
#pragma unroll 1
                    for (; j + 2 <= j1; j += 2) {
                        pair(j, refold_c);
                        pair(j + 1, refold_c);
                    }
                    if (j < j1) pair(j, refold_c);
                };
                if (refold) tile(std::true_type());
                else tile(std::false_type());
                if (KIND == K_LJ) S[0] += (double)hf[0];
#pragma unroll
                for (int i = 1; i < 8; i++) dsum(i - 1, (double)hf[i]);
            }
        }
        __syncthreads();
    }
    if (!valid) return;
    // m: config_makegrid order (x fastest) of the reciprocal-space table;
    // s: the map's scan order (z fastest) of the node values
    const long long m = (long long)x + (long long)g.nx * ((long long)y + (long long)g.ny * z);
    const long long s = ((long long)x * g.ny + y) * g.nz + z;
    double o[8];
    if (KIND == K_EWALD)
        for (int i = 0; i < 7; i++) S[i + 1] = s_acc[i][tid];
    if (KIND == K_EWALD) {
        if (failed) {
            o[0] = kVeryLrg;
            for (int i = 1; i < 8; i++) o[i] = 0.0;
        } else {
            const double *rp = recip + 8 * m;
            for (int i = 0; i < 8; i++) o[i] = (S[i] + rp[i]) * kTokcal;
        }
    } else {
        o[0] = failed ? kVeryLrg : S[0];
        for (int i = 1; i < 8; i++) o[i] = S[i];
    }
    float4 *d32 = reinterpret_cast<float4 *>(out32 + 8 * s);
    d32[0] = make_float4((float)o[0], (float)o[1], (float)o[2], (float)o[3]);
    d32[1] = make_float4((float)o[4], (float)o[5], (float)o[6], (float)o[7]);
    if (out64) {
        double2 *d64 = reinterpret_cast<double2 *>(out64 + 8 * s);
        for (int i = 0; i < 4; i++) d64[i] = make_double2(o[2 * i], o[2 * i + 1]);
    }
}

// ------------------------------------------------ Ewald reciprocal part
// Parity mode: V(r) = Re sum_k c_k e^{ik.r}, with c_k = s_k + i si_k
// (ewald_skwaldHOT, ewald.F90:966-1005: realPart = s cos - si sin, which
// also feeds the second derivatives); the first and third derivative terms
// use imagPart = s sin - si cos = Im(conj(c_k) e^{ik.r}).  Physics mode:
// the framework's own potential Re sum conj(c_k) e^{ik.r} = s cos + si sin
// with dV/dx = -kx Im(.), d2V/dxdy = -kx ky Re(.), d3V = +kx ky kz Im(.),
// all from the conj(c) stage.  k lies on the lattice
// 2 pi (mx/Lx, my/Ly, mz/Lz) and the nodes on the grid lattice, so every
// output is a separable sum evaluated axis by axis:
//   stage Z: Zc_w(mx,my,z) = sum_mz c kz^w e^{i kz z}, Zb_w likewise with conj(c)
//   stage Y+X (per (y,z) line): Y(mx) = sum_my ky^v e^{i ky y} Z(mx,my),
//                               out(x) = sum_mx kx^u e^{i kx x} Y(mx)
// Table E[axis][m][n] = e^{i k(m) n step}, FP64.
__global__ void k_recip_z(int side, int nz, const double2 *__restrict__ cgrid,
                          const double2 *__restrict__ Ez, const double *__restrict__ kzv,
                          double2 *__restrict__ Zout) {
    // one lane per (mx, my, z); Zout[((mx*side+my)*nz + z)*4 + w]
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long n = (long long)side * side * nz;
    if (i >= n) return;
    int z = (int)(i % nz);
    int mxy = (int)(i / nz);
    double2 c0 = make_double2(0, 0), c1 = c0, b0 = c0, b1 = c0;
    for (int mz = 0; mz < side; mz++) {
        double2 c = cgrid[(long long)mxy * side + mz];
        double2 e = Ez[(long long)mz * nz + z];
        double kz = kzv[mz];
        // c e
        double re = c.x * e.x - c.y * e.y, im = c.x * e.y + c.y * e.x;
        // conj(c) e
        double bre = c.x * e.x + c.y * e.y, bim = c.x * e.y - c.y * e.x;
        c0.x += re; c0.y += im;
        c1.x += kz * re; c1.y += kz * im;
        b0.x += bre; b0.y += bim;
        b1.x += kz * bre; b1.y += kz * bim;
    }
    double2 *o = Zout + i * 4;
    o[0] = c0; o[1] = c1; o[2] = b0; o[3] = b1;
}

constexpr int MAXSIDE = 81;

__global__ void __launch_bounds__(256) k_recip_xy(int side, int nx, int ny, int nz,
                                                  const double2 *__restrict__ Zin,
                                                  const double2 *__restrict__ Ex,
                                                  const double2 *__restrict__ Ey,
                                                  const double *__restrict__ kxv,
                                                  const double *__restrict__ kyv, double charge1,
                                                  int physics, double *__restrict__ recip) {
    // block per (y, z) line
    __shared__ double2 Y[MAXSIDE * 8];
    const int line = blockIdx.x;
    const int y = line % ny, z = line / ny;
    for (int t = threadIdx.x; t < side * 8; t += blockDim.x) {
        int mx = t / 8, o = t % 8;
        // o: 0 Yc00, 1 Yc10, 2 Yc01, 3 Yc11, 4 Yb00, 5 Yb10, 6 Yb01, 7 Yb11
        int zsel = (o < 4 && !physics) ? ((o & 2) ? 1 : 0) : ((o & 2) ? 3 : 2);
        bool kyw = (o & 1);
        double2 acc = make_double2(0, 0);
        for (int my = 0; my < side; my++) {
            double2 zz = Zin[(((long long)mx * side + my) * nz + z) * 4 + zsel];
            double2 e = Ey[(long long)my * ny + y];
            double re = zz.x * e.x - zz.y * e.y, im = zz.x * e.y + zz.y * e.x;
            if (kyw) { re *= kyv[my]; im *= kyv[my]; }
            acc.x += re;
            acc.y += im;
        }
        Y[t] = acc;
    }
    __syncthreads();
    for (int x = threadIdx.x; x < nx; x += blockDim.x) {
        double V = 0, h2 = 0, h3 = 0, h4 = 0, h5 = 0, h6 = 0, h7 = 0, h8 = 0;
        for (int mx = 0; mx < side; mx++) {
            double2 e = Ex[(long long)mx * nx + x];
            double kx = kxv[mx];
            const double2 *Yp = &Y[mx * 8];
            // Re(Y e) and Im(Y e)
#define RE(q) (Yp[q].x * e.x - Yp[q].y * e.y)
#define IM(q) (Yp[q].x * e.y + Yp[q].y * e.x)
            V += RE(0);
            h5 += kx * RE(1);
            h6 += kx * RE(2);
            h7 += RE(3);
            h2 += kx * IM(4);
            h3 += IM(5);
            h4 += IM(6);
            h8 += kx * IM(7);
#undef RE
#undef IM
        }
        long long m = (long long)x + (long long)nx * ((long long)y + (long long)ny * z);
        double *o = recip + 8 * m;
        o[0] = V * charge1;
        o[1] = -h2 * charge1;
        o[2] = -h3 * charge1;
        o[3] = -h4 * charge1;
        o[4] = -h5 * charge1;
        o[5] = -h6 * charge1;
        o[6] = -h7 * charge1;
        o[7] = (physics ? h8 : -h8) * charge1;
    }
}

// ---------------------------------------------------------- compaction
// pmap_compact (pmap.F90:509-656) made order-independent: a node above the
// cutoff is numbered by the first below-cutoff node (in the reference's
// scan order i(x) outer, k(z) inner) that has it as a wrapped neighbour;
// each below-cutoff node takes 1 + (number of neighbours it owns) indices,
// assigned by an exclusive prefix sum in scan order.  Scan index
// s = (i*ny + j)*nz + k is also the ptr layout index.
__device__ __forceinline__ int wrapi(int a, int n) { return a >= n ? 0 : (a < 0 ? n - 1 : a); }

__global__ void k_below(long long n, int nx, int ny, int nz, const float *__restrict__ full,
                        float cut, uint8_t *__restrict__ below) {
    long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    below[s] = full[8 * s] < cut ? 1 : 0;
}

__global__ void k_owner(long long n, int nx, int ny, int nz, const uint8_t *__restrict__ below,
                        int32_t *__restrict__ owner) {
    long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    int k = (int)(s % nz), j = (int)((s / nz) % ny), i = (int)(s / ((long long)nz * ny));
    long long best = -1;
    if (!below[s]) {
        for (int ii = -1; ii <= 1; ii++)
            for (int jj = -1; jj <= 1; jj++)
                for (int kk = -1; kk <= 1; kk++) {
                    if (!ii && !jj && !kk) continue;
                    long long t = ((long long)wrapi(i + ii, nx) * ny + wrapi(j + jj, ny)) * nz + wrapi(k + kk, nz);
                    if (below[t] && (best < 0 || t < best)) best = t;
                }
    }
    owner[s] = (int32_t)best;
}

__global__ void __launch_bounds__(256) k_count(long long n, int nx, int ny, int nz,
                                               const uint8_t *__restrict__ below,
                                               const int32_t *__restrict__ owner,
                                               int32_t *__restrict__ cnt,
                                               unsigned long long *__restrict__ partial) {
    __shared__ unsigned long long sred[3][4];
    long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    int c = 0;
    unsigned long long tb = 0, tn = 0, tr = 0;
    if (s < n) {
        int k = (int)(s % nz), j = (int)((s / nz) % ny), i = (int)(s / ((long long)nz * ny));
        if (below[s]) {
            c = 1;
            tb = 1;
            for (int ii = -1; ii <= 1; ii++)
                for (int jj = -1; jj <= 1; jj++)
                    for (int kk = -1; kk <= 1; kk++) {
                        if (!ii && !jj && !kk) continue;
                        long long t = ((long long)wrapi(i + ii, nx) * ny + wrapi(j + jj, ny)) * nz + wrapi(k + kk, nz);
                        if (!below[t] && owner[t] == s) c++;
                    }
        } else {
            long long o = owner[s];
            if (o >= 0) tn = 1;
            if (o < 0 || o > s) tr = 1;  // still unassigned when the scan reaches s
        }
        cnt[s] = c;
    }
    // block reduction, one partial per block (no global atomics)
    for (int off = 32; off > 0; off >>= 1) {
        tb += __shfl_down(tb, off);
        tn += __shfl_down(tn, off);
        tr += __shfl_down(tr, off);
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
        sred[0][wave] = tb;
        sred[1][wave] = tn;
        sred[2][wave] = tr;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        unsigned long long v = 0;
        for (int w = 0; w < 4; w++) v += sred[threadIdx.x][w];
        partial[3 * (long long)blockIdx.x + threadIdx.x] = v;
    }
}

__global__ void __launch_bounds__(1024) k_tally(long long nblk, const unsigned long long *__restrict__ partial,
                                                unsigned long long *__restrict__ tallies) {
    __shared__ unsigned long long sh[3][1024];
    unsigned long long v[3] = {0, 0, 0};
    for (long long b = threadIdx.x; b < nblk; b += 1024)
        for (int q = 0; q < 3; q++) v[q] += partial[3 * b + q];
    for (int q = 0; q < 3; q++) sh[q][threadIdx.x] = v[q];
    __syncthreads();
    for (int off = 512; off > 0; off >>= 1) {
        if (threadIdx.x < (unsigned)off)
            for (int q = 0; q < 3; q++) sh[q][threadIdx.x] += sh[q][threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x < 3) tallies[threadIdx.x] = sh[threadIdx.x][0];
}

// exclusive scan, three phases: block sums, scan of block sums, add back
constexpr int SCAN_B = 1024;
__global__ void __launch_bounds__(SCAN_B) k_scan_blocks(long long n, const int32_t *__restrict__ in,
                                                         int32_t *__restrict__ out,
                                                         int32_t *__restrict__ bsum) {
    __shared__ int32_t sh[SCAN_B];
    long long i = (long long)blockIdx.x * SCAN_B + threadIdx.x;
    int32_t v = (i < n) ? in[i] : 0;
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < SCAN_B; off <<= 1) {
        int32_t t = (threadIdx.x >= (unsigned)off) ? sh[threadIdx.x - off] : 0;
        __syncthreads();
        sh[threadIdx.x] += t;
        __syncthreads();
    }
    if (i < n) out[i] = sh[threadIdx.x] - v;
    if (threadIdx.x == SCAN_B - 1) bsum[blockIdx.x] = sh[SCAN_B - 1];
}

__global__ void k_scan_add(long long n, int32_t *__restrict__ out, const int32_t *__restrict__ boff) {
    long long i = (long long)blockIdx.x * SCAN_B + threadIdx.x;
    if (i < n) out[i] += boff[blockIdx.x];
}

__global__ void k_assign(long long n, int nx, int ny, int nz, const uint8_t *__restrict__ below,
                         const int32_t *__restrict__ owner, const int32_t *__restrict__ base,
                         const float *__restrict__ full, int32_t *__restrict__ ptr,
                         float *__restrict__ cvals) {
    long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    int k = (int)(s % nz), j = (int)((s / nz) % ny), i = (int)(s / ((long long)nz * ny));
    if (below[s]) {
        long long idx = base[s];
        ptr[s] = (int32_t)(idx + 1);
        const float4 *src = reinterpret_cast<const float4 *>(full + 8 * s);
        float4 *dst = reinterpret_cast<float4 *>(cvals + 8 * idx);
        dst[0] = src[0];
        dst[1] = src[1];
        for (int ii = -1; ii <= 1; ii++)
            for (int jj = -1; jj <= 1; jj++)
                for (int kk = -1; kk <= 1; kk++) {
                    if (!ii && !jj && !kk) continue;
                    int a = wrapi(i + ii, nx), bb = wrapi(j + jj, ny), c = wrapi(k + kk, nz);
                    long long t = ((long long)a * ny + bb) * nz + c;
                    if (!below[t] && owner[t] == s) {
                        idx++;
                        ptr[t] = (int32_t)(idx + 1);
                        const float4 *s2 = reinterpret_cast<const float4 *>(full + 8 * t);
                        float4 *d2 = reinterpret_cast<float4 *>(cvals + 8 * idx);
                        d2[0] = s2[0];
                        d2[1] = s2[1];
                    }
                }
    } else if (owner[s] < 0) {
        ptr[s] = 0;
    }
}

__global__ void k_fill_i32(long long n, int32_t v, int32_t *__restrict__ p) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_fill_u64(int n, unsigned long long *__restrict__ p) {
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0ull;
}

// Scratch of mh_map_build_static: a file-scope device array (one copy per
// device, held until the process ends), carved as below; its device address
// per device comes from hipGetSymbolAddress (the array's name is a host
// stand-in, not a device address).
constexpr long long kStaticMaxNodes = 1LL << 21;
constexpr size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }
constexpr long long kStaticBlocks256 = kStaticMaxNodes / 256;
constexpr long long kStaticScan0 = kStaticMaxNodes / 1024 + 1;
constexpr size_t kOffFull32 = 0;
constexpr size_t kOffBelow = kOffFull32 + al256(sizeof(float) * 8 * kStaticMaxNodes);
constexpr size_t kOffOwner = kOffBelow + al256(kStaticMaxNodes);
constexpr size_t kOffCnt = kOffOwner + al256(sizeof(int32_t) * kStaticMaxNodes);
constexpr size_t kOffBase = kOffCnt + al256(sizeof(int32_t) * kStaticMaxNodes);
constexpr size_t kOffPartial = kOffBase + al256(sizeof(int32_t) * kStaticMaxNodes);
constexpr size_t kOffTallies = kOffPartial + al256(sizeof(unsigned long long) * 3 * kStaticBlocks256);
constexpr size_t kOffScan = kOffTallies + 256;
constexpr size_t kScanLevelBytes = al256(sizeof(int32_t) * kStaticScan0);
constexpr size_t kOffPairTab = kOffScan + 8 * kScanLevelBytes;
constexpr size_t kStaticScratchBytes = kOffPairTab + al256(sizeof(PairTab));
static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

}  // namespace mh

using namespace mh;

// device address of `scratch` per device, set once by static_scratch_init
static std::mutex g_scratch_mu;
static std::vector<unsigned char *> g_scratch_ptr;

// called by mh_ctx_create with `device` current
int mh::static_scratch_init(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if ((int)g_scratch_ptr.size() <= device) g_scratch_ptr.resize(device + 1, nullptr);
    if (g_scratch_ptr[device]) return MH_OK;
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        return MH_EHIP;
    }
    g_scratch_ptr[device] = (unsigned char *)p;
    return MH_OK;
}

static unsigned char *static_scratch(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return device < (int)g_scratch_ptr.size() ? g_scratch_ptr[device] : nullptr;
}

// ------------------------------------------------------------ host side
// exclusive prefix sum of in[0..n) into out, level buffers bsum[l], boff[l]
// of at least n / 1024^(l+1) + 1 entries; total (level 0) on the host
static int scan_exclusive_b(mh_ctx *ctx, int level, long long n, const int32_t *in, int32_t *out,
                            int32_t *const *bsum, int32_t *const *boff, int64_t *total) {
    hipStream_t st = ctx->stream;
    long long nb = (n + SCAN_B - 1) / SCAN_B;
    MH_CHECK(level < 4, MH_ECAPACITY, "scan: too many levels");
    hipLaunchKernelGGL(k_scan_blocks, dim3((unsigned)nb), dim3(SCAN_B), 0, st, n, in, out, bsum[level]);
    MH_HIP(hipGetLastError());
    if (nb > 1) {
        int64_t sub = 0;
        int rc;
        if ((rc = scan_exclusive_b(ctx, level + 1, nb, bsum[level], boff[level], bsum, boff, &sub))) return rc;
        hipLaunchKernelGGL(k_scan_add, dim3((unsigned)nb), dim3(SCAN_B), 0, st, n, out, boff[level]);
        MH_HIP(hipGetLastError());
    }
    if (level == 0) {
        // total = last exclusive value + last input
        int32_t lastx = 0, lasti = 0;
        MH_HIP(hipMemcpyAsync(&lastx, out + (n - 1), sizeof(int32_t), hipMemcpyDeviceToHost, st));
        MH_HIP(hipMemcpyAsync(&lasti, in + (n - 1), sizeof(int32_t), hipMemcpyDeviceToHost, st));
        MH_HIP(hipStreamSynchronize(st));
        *total = (int64_t)lastx + lasti;
    }
    return MH_OK;
}

static int scan_exclusive(mh_ctx *ctx, long long n, const int32_t *in, int32_t *out, int64_t *total) {
    int32_t *bsum[4] = {nullptr, nullptr, nullptr, nullptr}, *boff[4] = {nullptr, nullptr, nullptr, nullptr};
    long long m = n;
    int rc = MH_OK;
    for (int l = 0; l < 4 && m > 0; l++) {
        long long nb = (m + SCAN_B - 1) / SCAN_B;
        std::string lv = std::to_string(l);
        bsum[l] = (int32_t *)ws_get(ctx, "scan_bsum" + lv, sizeof(int32_t) * (nb + 1), &rc);
        if (rc) return rc;
        boff[l] = (int32_t *)ws_get(ctx, "scan_boff" + lv, sizeof(int32_t) * (nb + 1), &rc);
        if (rc) return rc;
        if (nb == 1) break;
        m = nb;
    }
    return scan_exclusive_b(ctx, 0, n, in, out, bsum, boff, total);
}

// compaction buffers (pmap_compact restated for the device)
struct CompactBufs {
    uint8_t *below;
    int32_t *owner, *cnt, *base;
    unsigned long long *tallies, *partial;
};

// below / owner / count / tallies / scan: returns the number of tabulated
// records (host) and leaves ptr zeroed
static int compact_count(mh_ctx *ctx, const int np[3], long long nodes, const float *full32, float cut,
                         const CompactBufs &b, int32_t *const *bsum, int32_t *const *boff, int32_t *ptr,
                         int64_t *total) {
    hipStream_t st = ctx->stream;
    unsigned gb = (unsigned)((nodes + 255) / 256);
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, st, 4, b.tallies);
    hipLaunchKernelGGL(k_fill_i32, dim3(gb), dim3(256), 0, st, nodes, 0, ptr);
    hipLaunchKernelGGL(k_below, dim3(gb), dim3(256), 0, st, nodes, np[0], np[1], np[2], full32, cut, b.below);
    hipLaunchKernelGGL(k_owner, dim3(gb), dim3(256), 0, st, nodes, np[0], np[1], np[2], b.below, b.owner);
    hipLaunchKernelGGL(k_count, dim3(gb), dim3(256), 0, st, nodes, np[0], np[1], np[2], b.below, b.owner, b.cnt,
                       b.partial);
    hipLaunchKernelGGL(k_tally, dim3(1), dim3(1024), 0, st, (long long)gb, b.partial, b.tallies);
    if (hipGetLastError() != hipSuccess) {
        set_error("map compaction: kernel launch failed");
        return MH_EHIP;
    }
    if (bsum) return scan_exclusive_b(ctx, 0, nodes, b.cnt, b.base, bsum, boff, total);
    return scan_exclusive(ctx, nodes, b.cnt, b.base, total);
}

// numbering and record copies; counts = acceptCut, acceptNeighbor, naccept, reject
static int compact_assign(mh_ctx *ctx, const int np[3], long long nodes, const float *full32,
                          const CompactBufs &b, int32_t *ptr, float *vals, int64_t total, int64_t counts[4]) {
    hipStream_t st = ctx->stream;
    unsigned gb = (unsigned)((nodes + 255) / 256);
    hipLaunchKernelGGL(k_assign, dim3(gb), dim3(256), 0, st, nodes, np[0], np[1], np[2], b.below, b.owner, b.base,
                       full32, ptr, vals);
    if (hipGetLastError() != hipSuccess) {
        set_error("map compaction: assign kernel launch failed");
        return MH_EHIP;
    }
    unsigned long long t[4];
    if (hipMemcpyAsync(t, b.tallies, sizeof t, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
        set_error("map compaction: synchronisation failed");
        return MH_EHIP;
    }
    counts[0] = (int64_t)t[0];
    counts[1] = (int64_t)t[1];
    counts[2] = total;
    counts[3] = (int64_t)t[2];
    return MH_OK;
}

static void grid_dims(const double eff[3], double gridspace, int np[3], double step[3]) {
    // config_makegrid (config.F90:765-786)
    for (int i = 0; i < 3; i++) {
        np[i] = (int)(eff[i] / gridspace) + 1;
        step[i] = eff[i] / (double)(np[i] - 1);
    }
}

// separable reciprocal tables and stages; writes recip[nodes][8]
static int build_recip(mh_ctx *ctx, const int np[3], const double step[3], double charge1,
                       double **recip_out) {
    const Ewald &ew = ctx->ew;
    const int K = ew.kmax, side = 2 * K + 1;
    MH_CHECK(side <= MAXSIDE, MH_EINVAL, "Ewald kmax too large for the reciprocal kernel");
    const long long nodes = (long long)np[0] * np[1] * np[2];
    int rc;
    double *recip = (double *)ws_get(ctx, "recip", sizeof(double) * 8 * nodes, &rc);
    if (rc) return rc;
    *recip_out = recip;
    // per-axis k values (ewald.F90:599-619: k = (m * (1/L)) * twopi) and
    // tables E[m][n] = exp(i k(m) * n * step)
    std::vector<double> kv[3];
    std::vector<double> E[3];
    for (int a = 0; a < 3; a++) {
        double il = 1.0 / ctx->full_ell[a];
        kv[a].resize(side);
        E[a].resize(2 * (size_t)side * np[a]);
        for (int mm = 0; mm < side; mm++) {
            double k = ((double)(mm - K) * il) * (2.0 * kPiR);
            kv[a][mm] = k;
            for (int n = 0; n < np[a]; n++) {
                double arg = k * ((double)n * step[a]);
                E[a][2 * ((size_t)mm * np[a] + n)] = std::cos(arg);
                E[a][2 * ((size_t)mm * np[a] + n) + 1] = std::sin(arg);
            }
        }
    }
    double *dE[3], *dk[3];
    for (int a = 0; a < 3; a++) {
        dE[a] = (double *)ws_get(ctx, "recipE" + std::to_string(a), sizeof(double) * E[a].size(), &rc);
        if (rc) return rc;
        dk[a] = (double *)ws_get(ctx, "recipk" + std::to_string(a), sizeof(double) * side, &rc);
        if (rc) return rc;
        MH_HIP(hipMemcpyAsync(dE[a], E[a].data(), sizeof(double) * E[a].size(), hipMemcpyHostToDevice, ctx->stream));
        MH_HIP(hipMemcpyAsync(dk[a], kv[a].data(), sizeof(double) * side, hipMemcpyHostToDevice, ctx->stream));
    }
    KTimer trec(ctx, "ewald_recip");
    long long nzs = (long long)side * side * np[2];
    double2 *Z = (double2 *)ws_get(ctx, "recipZ", sizeof(double2) * 4 * nzs, &rc);
    if (rc) return rc;
    hipLaunchKernelGGL(k_recip_z, dim3((unsigned)((nzs + 255) / 256)), dim3(256), 0, ctx->stream,
                       side, np[2], ew.ccoef.as<double2>(), (const double2 *)dE[2], dk[2], Z);
    MH_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_recip_xy, dim3((unsigned)((long long)np[1] * np[2])), dim3(256), 0,
                       ctx->stream, side, np[0], np[1], np[2], Z, (const double2 *)dE[0],
                       (const double2 *)dE[1], dk[0], dk[1], charge1, ew.sign_mode == MH_EWALD_PHYSICS ? 1 : 0,
                       recip);
    MH_HIP(hipGetLastError());
    return MH_OK;
}

// erfc(x) of A&S 7.1.26 bounded by 1.5e-7 + exp(-x^2) * t-polynomial; the
// FP32 real-space cutoff is the distance where erfc(kappa r)/r < 1e-9 r^-1
static double real_space_cutoff(double kappa) {
    // smallest x with erfc(x) <= 1e-9 is ~4.32
    return 4.32 / kappa;
}

// pair tables, grid parameters and the node kernel of one map: FP32 node
// values into full32 (and FP64 into o64 when not NULL); dt is device memory
// for the pair table
static int run_nodes(mh_ctx *ctx, int kind, int precision, double probe_q, const int np[3], const double step[3],
                     PairTab *dt, float *full32, double *o64, bool timed) {
    const LJTable &lj = ctx->lj;
    double *recip = nullptr;
    // pair tables
    PairTab tab;
    std::memset(&tab, 0, sizeof tab);
    double maxh2 = 0.0, maxl2 = 0.0;
    if (kind == MH_MAP_LJ) {
        tab.ntypes = lj.ntypes;
        for (int t = 0; t < lj.ntypes; t++) {
            tab.A[t] = lj.A[t];
            tab.B[t] = lj.B[t];
            tab.h2[t] = lj.hicut2[t];
            tab.l2[t] = lj.locut2[t];
            if (lj.on[t]) {
                maxh2 = fmax(maxh2, lj.hicut2[t]);
                maxl2 = fmax(maxl2, lj.locut2[t]);
            }
        }
    } else if (kind == MH_MAP_BUCK) {
        const BuckTable &bk = ctx->buck;
        tab.ntypes = bk.ntypes;
        for (int t = 0; t < bk.ntypes; t++) {
            tab.bA[t] = bk.A[t];
            tab.brho[t] = bk.rho[t];
            tab.bC[t] = bk.C[t];
            tab.bD[t] = bk.D[t];
            tab.h2[t] = bk.hicut2[t];
            tab.l2[t] = bk.locut2[t];
            maxh2 = fmax(maxh2, bk.hicut2[t]);
        }
    }
    if (hipMemcpyAsync(dt, &tab, sizeof tab, hipMemcpyHostToDevice, ctx->stream) != hipSuccess)
        return MH_EHIP;
    GridP g;
    g.nx = np[0];
    g.ny = np[1];
    g.nz = np[2];
    for (int i = 0; i < 3; i++) {
        g.step[i] = step[i];
        g.full[i] = ctx->full_ell[i];
    }
    g.natoms = ctx->natoms;
    g.kappa = ctx->ew.kappa;
    g.charge1 = probe_q;
    g.r2safe = 0.0;
    g.r2pot = 0.0;
    g.rc2f = 0.0f;
    g.refold = 0;
    g.slanted = ctx->slanted ? 1 : 0;
    for (int i = 0; i < 9; i++) {
        g.U[i] = ctx->U[i];
        g.S[i] = ctx->S[i];
    }
    // slanted cells run the FP64 kernel for either precision, every atom
    // staged for every brick (the brick bound is orthorhombic)
    const bool slanted = ctx->slanted;
    if (kind == MH_MAP_EWALD) {
        int rc;
        if ((rc = build_recip(ctx, np, step, probe_q, &recip))) return rc;
        if (precision == MH_PREC_FP32 && !slanted) {
            double rc_real = real_space_cutoff(ctx->ew.kappa);
            g.cull2 = rc_real * rc_real * (1.0 + 1e-6);
            g.rc2f = (float)(rc_real * rc_real);
            // pairs within 3 A are evaluated in FP64 and the potential
            // term of pairs within 5 A: with the FP32 erfc / exp /
            // reciprocal each near 1 ulp, the FP32 sum of the remaining
            // terms keeps V within 1e-5 of max(|V|, 1 kcal/mol)
            g.r2safe = 9.0;
            g.r2pot = 25.0;
        } else {
            g.cull2 = 1e300;  // the reference sums every atom (ewald.F90:874)
        }
    } else {
        g.cull2 = slanted ? 1e300 : maxh2 * (1.0 + 1e-9) + 1e-9;
        g.r2safe = fmax(2.25, 4.0 * maxl2);
    }
    {
        // brick half-diagonal; fold(c - a) + (p - c) equals the reference's
        // single fold of p - a whenever cutoff + half-diagonal < L/2
        double hd = 0.5 * sqrt(pow((BX - 1) * step[0], 2) + pow((BY - 1) * step[1], 2) +
                               pow((BZ - 1) * step[2], 2));
        double reach = sqrt(fmin(g.cull2, 1e200)) + hd;
        double half = 0.5 * fmin(ctx->full_ell[0], fmin(ctx->full_ell[1], ctx->full_ell[2]));
        g.refold = (reach >= half) ? 1 : 0;
    }
    const int nbx = (np[0] + BX - 1) / BX, nby = (np[1] + BY - 1) / BY, nbz = (np[2] + BZ - 1) / BZ;
    const long long nblk = (long long)nbx * nby * nbz;
    const bool fp32 = (precision == MH_PREC_FP32) && kind != MH_MAP_BUCK && !slanted;
    const double *axyz = slanted ? ctx->xyz_s.as<double>() : ctx->xyz.as<double>();
    static const char *knames[2][4] = {
        {"", "nodes64_lj", "nodes64_ewald_real", "nodes64_buck"},
        {"", "nodes32_lj", "nodes32_ewald_real", "nodes64_buck"}};
    KTimer tnode(timed ? ctx : nullptr, knames[fp32 ? 1 : 0][kind]);
    if (kind == MH_MAP_LJ) {
        if (fp32)
            hipLaunchKernelGGL(k_nodes32<K_LJ>, dim3((unsigned)nblk), dim3(NT), 0, ctx->stream, g, dt,
                               axyz, ctx->type.as<int>(), ctx->q.as<double>(),
                               (const double *)nullptr, full32, o64);
        else
            hipLaunchKernelGGL(k_nodes64<K_LJ>, dim3((unsigned)nblk), dim3(NT), 0, ctx->stream, g, dt,
                               axyz, ctx->type.as<int>(), ctx->q.as<double>(),
                               (const double *)nullptr, full32, o64);
    } else if (kind == MH_MAP_BUCK) {
        hipLaunchKernelGGL(k_nodes64<K_BUCK>, dim3((unsigned)nblk), dim3(NT), 0, ctx->stream, g, dt,
                           axyz, ctx->type.as<int>(), ctx->q.as<double>(),
                           (const double *)nullptr, full32, o64);
    } else {
        if (fp32)
            hipLaunchKernelGGL(k_nodes32<K_EWALD>, dim3((unsigned)nblk), dim3(NT), 0, ctx->stream, g, dt,
                               axyz, ctx->type.as<int>(), ctx->q.as<double>(),
                               recip, full32, o64);
        else
            hipLaunchKernelGGL(k_nodes64<K_EWALD>, dim3((unsigned)nblk), dim3(NT), 0, ctx->stream, g, dt,
                               axyz, ctx->type.as<int>(), ctx->q.as<double>(),
                               recip, full32, o64);
    }
    if (hipGetLastError() != hipSuccess) {
        set_error("mh_map_build: node kernel launch failed");
        return MH_EHIP;
    }
    tnode.stop();
    return MH_OK;
}

extern "C" int mh_map_build(mh_ctx *ctx, int kind, int precision, double probe_q,
                            double gridspace, double hipotcut_kJ, int keep_full, mh_map **out) {
    MH_CHECK(ctx && out, MH_EINVAL, "mh_map_build: bad arguments");
    MH_CHECK(ctx->natoms > 0, MH_ESTATE, "mh_map_build: upload the framework first");
    MH_CHECK(kind == MH_MAP_LJ || kind == MH_MAP_EWALD || kind == MH_MAP_BUCK, MH_EINVAL,
             "mh_map_build: unknown map kind");
    MH_CHECK(precision == MH_PREC_FP64 || precision == MH_PREC_FP32, MH_EINVAL,
             "mh_map_build: unknown precision");
    MH_CHECK(gridspace > 0.0, MH_EINVAL, "mh_map_build: gridspace must be > 0");
    MH_CHECK(kind != MH_MAP_LJ || ctx->lj_set, MH_ESTATE, "mh_map_build: call mh_ff_set_lj first");
    MH_CHECK(kind != MH_MAP_BUCK || ctx->buck_set, MH_ESTATE, "mh_map_build: call mh_ff_set_buck first");
    MH_CHECK(kind != MH_MAP_EWALD || ctx->ew.set, MH_ESTATE, "mh_map_build: call mh_ewald_set first");
    const LJTable &lj = ctx->lj;
    MH_CHECK(kind != MH_MAP_LJ || ctx->max_type < lj.ntypes, MH_EINVAL,
             "mh_map_build: framework atom type without LJ parameters");
    MH_CHECK(kind != MH_MAP_BUCK || ctx->max_type < ctx->buck.ntypes, MH_EINVAL,
             "mh_map_build: framework atom type without Buckingham parameters");
    MH_HIP(hipSetDevice(ctx->device));
    int np[3];
    double step[3];
    grid_dims(ctx->eff, gridspace, np, step);
    MH_CHECK(np[0] >= 3 && np[1] >= 3 && np[2] >= 3, MH_EINVAL,
             "mh_map_build: at least 3 grid nodes per axis are required");
    const long long nodes = (long long)np[0] * np[1] * np[2];
    MH_CHECK(nodes < (1LL << 31), MH_ECAPACITY, "mh_map_build: more than 2^31 grid nodes");
    mh_map *map = new mh_map();
    map->ctx = ctx;
    for (int i = 0; i < 3; i++) {
        map->npnts[i] = np[i];
        map->step[i] = step[i];
        map->fund_ell[i] = ctx->fund_ell[i];
    }
    map->slanted = ctx->slanted;
    for (int i = 0; i < 9; i++) {
        map->U[i] = ctx->U[i];
        map->S[i] = ctx->S[i];
    }
    int rc = MH_OK;
    auto fail = [&](int code) {
        delete map;
        return code;
    };
    // node values: owned by the map when kept, else context scratch
    DBuf full32k, full64k;
    float *full32 = nullptr;
    double *full64 = nullptr;
    const bool want64 = keep_full != 0;
    if (keep_full) {
        if ((rc = full32k.alloc(sizeof(float) * 8 * nodes))) return fail(rc);
        if ((rc = full64k.alloc(sizeof(double) * 8 * nodes))) return fail(rc);
        full32 = full32k.as<float>();
        full64 = full64k.as<double>();
    } else {
        full32 = (float *)ws_get(ctx, "full32", sizeof(float) * 8 * nodes, &rc);
        if (rc) return fail(rc);
    }
    double *o64 = want64 ? full64 : nullptr;
    PairTab *dt = (PairTab *)ws_get(ctx, "pairtab", sizeof(PairTab), &rc);
    if (rc) return fail(rc);
    if ((rc = run_nodes(ctx, kind, precision, probe_q, np, step, dt, full32, o64, true))) {
        if (rc == MH_EHIP) set_error("mh_map_build: node kernel launch failed");
        return fail(rc);
    }
    KTimer tcomp(ctx, "compaction");
    CompactBufs cb;
    cb.below = (uint8_t *)ws_get(ctx, "below", nodes, &rc);
    if (rc) return fail(rc);
    cb.owner = (int32_t *)ws_get(ctx, "owner", sizeof(int32_t) * nodes, &rc);
    if (rc) return fail(rc);
    cb.cnt = (int32_t *)ws_get(ctx, "cnt", sizeof(int32_t) * nodes, &rc);
    if (rc) return fail(rc);
    cb.base = (int32_t *)ws_get(ctx, "base", sizeof(int32_t) * nodes, &rc);
    if (rc) return fail(rc);
    cb.tallies = (unsigned long long *)ws_get(ctx, "tallies", 64, &rc);
    if (rc) return fail(rc);
    cb.partial = (unsigned long long *)ws_get(ctx, "cpartial",
                                              sizeof(unsigned long long) * 3 * ((nodes + 255) / 256), &rc);
    if (rc) return fail(rc);
    if ((rc = map->ptr.alloc(sizeof(int32_t) * nodes))) return fail(rc);
    const float cut = (float)(hipotcut_kJ / 4.184);  // mapmaker.F90:562, Real(cutoff,RSgl)
    int64_t total = 0;
    if ((rc = compact_count(ctx, np, nodes, full32, cut, cb, nullptr, nullptr, map->ptr.as<int32_t>(), &total)))
        return fail(rc);
    MH_CHECK(total < (1LL << 31), MH_ECAPACITY, "mh_map_build: more than 2^31 tabulated points");
    map->nfsize = total;
    if ((rc = map->vals.alloc(sizeof(float) * 8 * (total + 1)))) return fail(rc);
    if ((rc = compact_assign(ctx, np, nodes, full32, cb, map->ptr.as<int32_t>(), map->vals.as<float>(), total,
                             map->counts)))
        return fail(rc);
    tcomp.stop();
    if (keep_full) {
        map->full32.p = full32k.p;
        map->full32.n = full32k.n;
        full32k.p = nullptr;
        map->full64.p = full64k.p;
        map->full64.n = full64k.n;
        full64k.p = nullptr;
    }
    map->retained = true;
    ctx_retain(ctx);
    *out = map;
    return MH_OK;
}

// The allocation-free variant of mh_map_build for LJ and Buckingham maps:
// the caller's device buffers take the compacted map, the node values and
// compaction arrays live in the file-scope scratch array (kStaticMaxNodes
// grid nodes at most).  No hipMalloc / hipFree / allocating library call;
// counters and ptr are cleared by fill kernels.  Calls on one device share
// the scratch, so calls that overlap on one device race.
extern "C" int mh_map_build_static(mh_ctx *ctx, int kind, int precision, double gridspace, double hipotcut_kJ,
                                   int32_t *ptr_dev, float *vals_dev, int64_t vals_capacity, int npnts_out[3],
                                   int64_t *nfsize_out, int64_t counts_out[4]) {
    MH_CHECK(ctx && npnts_out, MH_EINVAL, "mh_map_build_static: bad arguments");
    MH_CHECK(ctx->natoms > 0, MH_ESTATE, "mh_map_build_static: upload the framework first");
    MH_CHECK(kind == MH_MAP_LJ || kind == MH_MAP_BUCK, MH_EINVAL,
             "mh_map_build_static: LJ and Buckingham maps only (Ewald maps need per-node reciprocal scratch)");
    MH_CHECK(precision == MH_PREC_FP64 || precision == MH_PREC_FP32, MH_EINVAL,
             "mh_map_build_static: unknown precision");
    MH_CHECK(gridspace > 0.0, MH_EINVAL, "mh_map_build_static: gridspace must be > 0");
    MH_CHECK(kind != MH_MAP_LJ || ctx->lj_set, MH_ESTATE, "mh_map_build_static: call mh_ff_set_lj first");
    MH_CHECK(kind != MH_MAP_BUCK || ctx->buck_set, MH_ESTATE, "mh_map_build_static: call mh_ff_set_buck first");
    MH_CHECK(kind != MH_MAP_LJ || ctx->max_type < ctx->lj.ntypes, MH_EINVAL,
             "mh_map_build_static: framework atom type without LJ parameters");
    MH_CHECK(kind != MH_MAP_BUCK || ctx->max_type < ctx->buck.ntypes, MH_EINVAL,
             "mh_map_build_static: framework atom type without Buckingham parameters");
    int np[3];
    double step[3];
    grid_dims(ctx->eff, gridspace, np, step);
    for (int i = 0; i < 3; i++) npnts_out[i] = np[i];
    if (!ptr_dev) return MH_OK;  // size query
    MH_CHECK(vals_dev && nfsize_out && counts_out, MH_EINVAL, "mh_map_build_static: bad output arguments");
    MH_CHECK(np[0] >= 3 && np[1] >= 3 && np[2] >= 3, MH_EINVAL,
             "mh_map_build_static: at least 3 grid nodes per axis are required");
    const long long nodes = (long long)np[0] * np[1] * np[2];
    MH_CHECK(nodes <= kStaticMaxNodes, MH_ECAPACITY,
             "mh_map_build_static: more grid nodes than the static scratch holds (2^21); use mh_map_build");
    MH_HIP(hipSetDevice(ctx->device));
    unsigned char *sc = static_scratch(ctx->device);
    MH_CHECK(sc, MH_ESTATE, "mh_map_build_static: no scratch address for this device (mh_ctx_create sets it)");
    float *full32 = (float *)(sc + kOffFull32);
    CompactBufs cb;
    cb.below = (uint8_t *)(sc + kOffBelow);
    cb.owner = (int32_t *)(sc + kOffOwner);
    cb.cnt = (int32_t *)(sc + kOffCnt);
    cb.base = (int32_t *)(sc + kOffBase);
    cb.partial = (unsigned long long *)(sc + kOffPartial);
    cb.tallies = (unsigned long long *)(sc + kOffTallies);
    int32_t *bsum[4], *boff[4];
    for (int l = 0; l < 4; l++) {
        bsum[l] = (int32_t *)(sc + kOffScan + (2 * l) * kScanLevelBytes);
        boff[l] = (int32_t *)(sc + kOffScan + (2 * l + 1) * kScanLevelBytes);
    }
    PairTab *dt = (PairTab *)(sc + kOffPairTab);
    int rc;
    if ((rc = run_nodes(ctx, kind, precision, 0.0, np, step, dt, full32, nullptr, false))) {
        if (rc == MH_EHIP) set_error("mh_map_build_static: node kernel launch failed");
        return rc;
    }
    const float cut = (float)(hipotcut_kJ / 4.184);  // mapmaker.F90:562, Real(cutoff,RSgl)
    int64_t total = 0;
    if ((rc = compact_count(ctx, np, nodes, full32, cut, cb, bsum, boff, ptr_dev, &total))) return rc;
    *nfsize_out = total;
    MH_CHECK(total <= vals_capacity, MH_ECAPACITY,
             "mh_map_build_static: more tabulated points than vals_capacity records (nfsize_out holds the count)");
    return compact_assign(ctx, np, nodes, full32, cb, ptr_dev, vals_dev, total, counts_out);
}
