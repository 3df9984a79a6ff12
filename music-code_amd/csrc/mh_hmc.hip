// libmusic_hip: hybrid MC bursts of flexible sorbates (HMC) in a fixed
// framework, many replicas per launch, one workgroup per replica.
//
// Per move (hybridmoves_integrate_boltz, hybridmoves.F90:186-241):
//   1. Gaussian velocities for every atom in the reference's draw order
//      (molecule, atom, x/y/z; Box-Muller pairs with the saved second value,
//      random.F90:197-217).  Lanes generate their atoms' draws in parallel:
//      the rranf_NR state is jumped ahead to the lane's pair with
//      precomputed powers of the xorshift matrix (GF(2)) and of 16807 mod
//      2^31 - 1, so every value is the one the serial stream gives.
//   2. Energy and forces of the start configuration.
//   3. md_steps NVE velocity-Verlet steps (velverlet_nveIntegrate,
//      velverlet.F90:148-288) on the principal and simulation-cell
//      coordinates; a failed evaluation or an energy above max_nrg_md ends
//      the move as rejected without an RNG draw.
//   4. Metropolis on dU + dKE / 4.184 (mcmoves_perturb, mcmoves.F90:371-426).
// Layout: site positions rp (principal) and r (cell) FP64 in LDS for the
// launch, with 32-bit fixed-point copies of r (pair arithmetic) and the
// neighbour-list bookkeeping; each lane owns sites s = tid + k NT, whose
// velocities and accelerations it keeps in global memory.  Forces per owned
// site: map (forcehermite, interpolate.F90:189-584), LJ against the sites of
// the other molecules in its neighbour list (minimum image), and the bonded
// terms that contain the site (each term is evaluated by the lanes of its
// atoms, its energy counted once by the lane of its first atom).
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "mh_device.h"
#include "mh_internal.h"
#include "mh_mc_common.h"

namespace mh {

constexpr int HS = MH_MAX_SITES;
constexpr int HT = 16;
constexpr int JLEV = 20;  // jump-ahead levels: up to 2^20 draws per move
// neighbour lists: partner sites within hicut + kSkin, rebuilt whenever a
// site has moved more than half the skin since the last build (pairs left out
// then stay beyond hicut)
constexpr float kSkin = 1.0f;
constexpr float kHalfSkin2 = 0.49f * 0.49f;

struct HDev {
    int ns, nst, nb, na, nt, nmd, prec, maxN;
    int st[HS], has_map[HS];
    double mass[HS], ainv[HS], sigma[HS];
    MapRef map[HS];
    double full[3], fund[3];
    // LJ of the site pair (a, b) (index a * HS + b): A, B, hicut^2 (-1 where
    // the type pair is switched off), locut^2; list radius^2 (-1 where off)
    float4 pp32[HS * HS];
    double pp64[HS * HS][4];
    float rl2[HS * HS];
    // fixed-point cell positions: 2^32 / L per axis, and its inverse (A per unit)
    double fxs[3];
    float fxu[3];
    int bi[HT][2];
    double bk[HT], br0[HT];
    int ai[HT][3];
    double ak[HT], at0[HT];
    int ti[HT][4];
    double tc[HT][6];
    double T, dt, rti, max_nrg_md, max_nrg0;
    double mmassi2;  // 1 / M / M, M = molecule mass
    uint32_t xsj[JLEV][32];     // columns of M^(2^k), M = xorshift(13, 17, 5) on GF(2)^32
    unsigned long long pmj[JLEV];  // 16807^(2^k) mod (2^31 - 1)
};

struct HRep {
    int *N, *err;
    int32_t *ix, *iy, *first, *iseed, *ghave;
    double *gz2, *energy, *dH;
    long long *att, *acc, *fail;
    double *rp, *r;             // [rep][maxN][ns][3]
    double *rp0, *r0;           // saved start configuration of the move
    double *v, *a;              // [rep][maxN][ns][3] velocities, accelerations
    uint16_t *nbl;              // [rep][cap][cap] neighbour lists, entry-major
    int cap;                    // maxN * nsites
};

constexpr double kScaleKE = 0.01;   // (A/ps)^2 amu -> kJ/mol, defaults.F90:76
constexpr double kScaleF = 418.4;   // kcal/mol/A / amu -> A/ps^2, defaults.F90:73
constexpr double kKjKb = 8.314e-3;

__device__ __forceinline__ uint32_t gf2_apply(const uint32_t (&col)[32], uint32_t x) {
    uint32_t y = 0;
#pragma unroll
    for (int b = 0; b < 32; b++)
        if (x & (1u << b)) y ^= col[b];
    return y;
}

// rranf_NR state advanced by n draws
__device__ __forceinline__ void rng_jump(const HDev &P, Rng &r, unsigned n) {
    uint32_t x = (uint32_t)r.ix;
    unsigned long long y = (unsigned long long)(uint32_t)r.iy;
    for (int k = 0; n && k < JLEV; k++, n >>= 1)
        if (n & 1) {
            x = gf2_apply(P.xsj[k], x);
            y = (y * P.pmj[k]) % 2147483647ULL;
        }
    r.ix = (int32_t)x;
    r.iy = (int32_t)y;
}

__device__ __forceinline__ void rng_first(Rng &r) {
    if (r.first) {
        int32_t a = r.iseed < 0 ? -r.iseed : r.iseed;
        r.first = 0;
        r.iy = (int32_t)((888889999 ^ a) | 1);
        r.ix = (int32_t)(777755555 ^ a);
        r.iseed = a + 1;
    }
}

// Box-Muller pair p of the move (draws 2p + 1, 2p + 2 after state S0):
// first (cos) or second (sin) value
__device__ double gauss_pair(const HDev &P, const Rng &S0, unsigned p, int second) {
    Rng s = S0;
    rng_jump(P, s, 2u * p);
    double r1 = rranf(s), r2 = rranf(s);
    if (r1 < 1.0e-10) r1 = 1.0e-10;
    if (r2 < 1.0e-10) r2 = 1.0e-10;
    const double twopi = 2.0 * 3.1415927;
    const double rad = sqrt(-2.0 * log(r1));
    return second ? rad * sin(twopi * r2) : rad * cos(twopi * r2);
}

// forcehermite (interpolate.F90:189-584) in two parts, so that other work
// can run while the corner records are in flight.  hermite_fetch: the cell of
// a position (false outside the tabulated range) and the loads of its eight
// corner records (global address space, so waits on LDS reads do not wait on
// them).
struct Corners {
    float4 rec[16];
    double t[3];
};

typedef float F4v __attribute__((ext_vector_type(4)));
typedef const F4v __attribute__((address_space(1))) *GlobalF4;

__device__ __forceinline__ bool hermite_fetch(const MapRef &m, const double *r, Corners &g) {
    int i0[3], i1[3];
    for (int d = 0; d < 3; d++)
        if (!cell_axis(r[d], m.ell[d], m.step[d], m.n[d], i0[d], i1[d], g.t[d])) return false;
#pragma unroll
    for (int c = 0; c < 8; c++) {
        const int cx = (kCX >> c) & 1, cy = (kCY >> c) & 1, cz = (kCZ >> c) & 1;
        const long long node = ((long long)(cx ? i1[0] : i0[0]) * m.n[1] + (cy ? i1[1] : i0[1])) * m.n[2] +
                               (cz ? i1[2] : i0[2]);
        const GlobalF4 q = (GlobalF4)(m.dense + 8 * node);
        const F4v lo = q[0], hi = q[1];
        g.rec[2 * c] = make_float4(lo.x, lo.y, lo.z, lo.w);
        g.rec[2 * c + 1] = make_float4(hi.x, hi.y, hi.z, hi.w);
    }
    return true;
}

// hermite_eval: p = V, -dV/dx, -dV/dy, -dV/dz from the fetched corners;
// false where a corner is untabulated
template <int PREC>
__device__ __forceinline__ bool hermite_eval(const MapRef &m, const Corners &g, double p[4]) {
    bool ok = true;
#pragma unroll
    for (int c = 0; c < 8; c++) ok = ok && g.rec[2 * c].x == g.rec[2 * c].x;
    if (!ok) return false;
    if (PREC == MH_PREC_FP64) {
        double h[3][2][2], hd[3][2][2];
        for (int d = 0; d < 3; d++) {
            const double st = m.step[d], u = g.t[d], w = 1 - g.t[d];
            h[d][0][0] = w * w * (3.0 - 2.0 * w);
            h[d][1][0] = u * u * (3.0 - 2.0 * u);
            h[d][0][1] = (-st) * w * w * (w - 1.0);
            h[d][1][1] = st * u * u * (u - 1.0);
            hd[d][0][0] = -(6.0 * w * (1.0 - w));
            hd[d][1][0] = 6.0 * u * (1.0 - u);
            hd[d][0][1] = -((-st) * w * (3.0 * w - 2.0));
            hd[d][1][1] = st * u * (3.0 * u - 2.0);
        }
        double s1 = 0.0, s2 = 0.0, s3 = 0.0, s4 = 0.0;
#pragma unroll
        for (int c = 0; c < 8; c++) {
            const int cx = (kCX >> c) & 1, cy = (kCY >> c) & 1, cz = (kCZ >> c) & 1;
            const float4 a = g.rec[2 * c], b = g.rec[2 * c + 1];
            const float f[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
#pragma unroll
            for (int k = 0; k < 8; k++) {
                const int dx = (kCX >> k) & 1, dy = (kCY >> k) & 1, dz = (kCZ >> k) & 1;
                const double fv = f[k], hi = h[0][cx][dx], hj = h[1][cy][dy], hk = h[2][cz][dz];
                const double di = hd[0][cx][dx], dj = hd[1][cy][dy], dk = hd[2][cz][dz];
                s1 = s1 + fv * hi * hj * hk;
                s2 = s2 + fv * di * hj * hk;
                s3 = s3 + fv * hi * dj * hk;
                s4 = s4 + fv * hi * hj * dk;
            }
        }
        p[0] = s1;
        p[1] = -s2 / m.step[0];
        p[2] = -s3 / m.step[1];
        p[3] = -s4 / m.step[2];
    } else {
#pragma clang fp contract(fast)
        float h[3][2][2], hd[3][2][2];
        for (int d = 0; d < 3; d++) {
            const float st = (float)m.step[d], u = (float)g.t[d], w = 1.0f - u;
            h[d][0][0] = w * w * (3.0f - 2.0f * w);
            h[d][1][0] = u * u * (3.0f - 2.0f * u);
            h[d][0][1] = (-st) * w * w * (w - 1.0f);
            h[d][1][1] = st * u * u * (u - 1.0f);
            hd[d][0][0] = -(6.0f * w * (1.0f - w));
            hd[d][1][0] = 6.0f * u * (1.0f - u);
            hd[d][0][1] = -((-st) * w * (3.0f * w - 2.0f));
            hd[d][1][1] = st * u * (3.0f * u - 2.0f);
        }
        double s1 = 0.0, s2 = 0.0, s3 = 0.0, s4 = 0.0;
#pragma unroll
        for (int c = 0; c < 8; c++) {
            const int cx = (kCX >> c) & 1, cy = (kCY >> c) & 1, cz = (kCZ >> c) & 1;
            const float4 a = g.rec[2 * c], b = g.rec[2 * c + 1];
            const float f[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
            float q1 = 0.0f, q2 = 0.0f, q3 = 0.0f, q4 = 0.0f;
#pragma unroll
            for (int k = 0; k < 8; k++) {
                const int dx = (kCX >> k) & 1, dy = (kCY >> k) & 1, dz = (kCZ >> k) & 1;
                const float hi = h[0][cx][dx], hj = h[1][cy][dy], hk = h[2][cz][dz];
                const float di = hd[0][cx][dx], dj = hd[1][cy][dy], dk = hd[2][cz][dz];
                q1 += f[k] * hi * hj * hk;
                q2 += f[k] * di * hj * hk;
                q3 += f[k] * hi * dj * hk;
                q4 += f[k] * hi * hj * dk;
            }
            s1 += q1;
            s2 += q2;
            s3 += q3;
            s4 += q4;
        }
        p[0] = s1;
        p[1] = -s2 / m.step[0];
        p[2] = -s3 / m.step[1];
        p[3] = -s4 / m.step[2];
    }
    return true;
}

__device__ __forceinline__ double dot3(const double *a, const double *b) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

// bonded terms containing atom `a` of molecule R (principal coordinates):
// force on a into f, energy of terms whose first atom is a into eb (stretch,
// bending, torsion)
__device__ void bonded_site(const HDev &P, const double *R, int a, double f[3], double (&eb)[3]) {
    for (int q = 0; q < P.nb; q++) {  // harmonic_getinteraction, harmonic.F90:130-160
        const int i = P.bi[q][0], j = P.bi[q][1];
        if (a != i && a != j) continue;
        double vr[3] = {R[3 * j] - R[3 * i], R[3 * j + 1] - R[3 * i + 1], R[3 * j + 2] - R[3 * i + 2]};
        const double r = sqrt(dot3(vr, vr));
        double vru[3] = {0.0, 0.0, 0.0};
        if (r != 0.0)
            for (int k = 0; k < 3; k++) vru[k] = vr[k] / r;
        const double k_ = P.bk[q], req = P.br0[q];
        const double fm = -k_ * (r - req);
        for (int k = 0; k < 3; k++) f[k] += (a == j) ? vru[k] * fm : vru[k] * fm * (-1.0);
        if (a == i) eb[0] += 0.5 * k_ * (r - req) * (r - req);
    }
    for (int q = 0; q < P.na; q++) {  // harmonica, harmonica.F90:100-160
        const int i = P.ai[q][0], j = P.ai[q][1], l = P.ai[q][2];
        if (a != i && a != j && a != l) continue;
        double v1[3], v2[3];
        for (int k = 0; k < 3; k++) {
            v1[k] = R[3 * i + k] - R[3 * j + k];
            v2[k] = R[3 * l + k] - R[3 * j + k];
        }
        const double r12 = dot3(v1, v1), r22 = dot3(v2, v2);
        const double r1 = sqrt(r12), r2 = sqrt(r22);
        const double adotb = dot3(v2, v1);
        const double ct = adotb / (r1 * r2);
        const double th = acos(ct);
        const double st = sqrt(1.0 - ct * ct);
        const double at = (fabs(st) < 1.0e-10) ? 0.0 : P.ak[q] * (th - P.at0[q]) / (r1 * r2 * st);
        for (int k = 0; k < 3; k++) {
            const double a1 = (v2[k] - v1[k] * (adotb / r12)) * at;
            const double a3 = (v1[k] - v2[k] * (adotb / r22)) * at;
            f[k] += (a == i) ? a1 : (a == l) ? a3 : (a1 + a3) * (-1.0);
        }
        if (a == i) eb[1] += 0.5 * P.ak[q] * (th - P.at0[q]) * (th - P.at0[q]);
    }
    for (int q = 0; q < P.nt; q++) {  // cosexpansion_getinteraction, cosexpansion.F90:88-196
        const int i1 = P.ti[q][0], i2 = P.ti[q][1], i3 = P.ti[q][2], i4 = P.ti[q][3];
        if (a != i1 && a != i2 && a != i3 && a != i4) continue;
        double va[3], vb[3], vc[3];
        for (int k = 0; k < 3; k++) {
            va[k] = R[3 * i2 + k] - R[3 * i1 + k];
            vb[k] = R[3 * i3 + k] - R[3 * i2 + k];
            vc[k] = R[3 * i4 + k] - R[3 * i3 + k];
        }
        const double aa = dot3(va, va), ab = dot3(va, vb), ac = dot3(va, vc), bb = dot3(vb, vb),
                     bc = dot3(vb, vc), cc = dot3(vc, vc);
        const double T1 = ab * bc - ac * bb, T2 = aa * bb - ab * ab, T3 = bb * cc - bc * bc;
        if (fabs(T2 * T3) < 1.0e-11) continue;
        const double inv = 1.0 / sqrt(T2 * T3), T1T2 = T1 / T2, T1T3 = T1 / T3;
        const double cs = -(T1 * inv);
        double term = 1.0, sum = 0.0, s = 0.0;
        for (int n = 0; n < 6; n++) {
            sum = sum + P.tc[q][n] * term;
            term = term * cs;
        }
        term = 1.0;
        for (int n = 1; n < 6; n++) {
            s = s + (double)n * P.tc[q][n] * term;
            term = term * cs;
        }
        const double ffac = -s;
        for (int k = 0; k < 3; k++) {
            const double da = ((va[k] * (T1T2 * bb)) + (vc[k] * bb) - (vb[k] * (bc + ab * T1T2))) * inv;
            const double db = ((vb[k] * (2 * ac + T1T2 * aa + T1T3 * cc)) - (va[k] * (bc + ab * T1T2)) -
                               (vc[k] * (ab + T1T3 * bc))) * inv;
            const double dc = ((va[k] * bb) + (vc[k] * (T1T3 * bb)) - (vb[k] * (ab + bc * T1T3))) * inv;
            double g;
            if (a == i1) g = da * (-1.0);
            else if (a == i2) g = da - db;
            else if (a == i3) g = db - dc;
            else g = dc;
            f[k] += g * ffac;
        }
        if (a == i1) eb[2] += sum;
    }
}

// fixed-point cell position of site s (x 2^32 / L, wrapped mod 2^32: the
// difference of two such coordinates read as int32 is the minimum image) and
// its atom index a
__device__ __forceinline__ void put_fx(const HDev &P, uint4 *ru, const double *r, int s, int a) {
    uint32_t u[3];
    for (int d = 0; d < 3; d++) u[d] = (uint32_t)(unsigned long long)(long long)rint(r[3 * s + d] * P.fxs[d]);
    ru[s] = make_uint4(u[0], u[1], u[2], (uint32_t)a);
}

// true where one of the lane's sites has moved more than half the skin (in
// principal coordinates, which are never wrapped) since its list was built
template <int NT>
__device__ __forceinline__ int moved_far(const double *rp, const float4 *nb, int nsite) {
    int moved = 0;
#pragma unroll 1
    for (int s = threadIdx.x; s < nsite; s += NT) {
        const float4 b = nb[s];
        const float dx = (float)rp[3 * s] - b.x, dy = (float)rp[3 * s + 1] - b.y, dz = (float)rp[3 * s + 2] - b.z;
        if (dx * dx + dy * dy + dz * dz > kHalfSkin2) moved = 1;
    }
    return moved;
}

// neighbour lists of the lane's sites: partner sites j of the other molecules,
// ascending, within the list radius of the site pair, stored entry-major
// (entry q of site s at q * cap + s); nb[s] = build position (principal) and
// the entry count
template <int NT>
__device__ void hmc_lists(const HDev &P, const uint4 *ru, const double *rp, int N, uint16_t *nbl, int cap,
                          float4 *nb) {
    const int ns = P.ns, nsite = N * ns;
    const float ux = P.fxu[0], uy = P.fxu[1], uz = P.fxu[2];
#pragma unroll 1
    for (int s = threadIdx.x; s < nsite; s += NT) {
        const int m = s / ns, a = s - m * ns, j0 = m * ns;
        const uint4 pi = ru[s];
        const float *rl2 = P.rl2 + a * HS;
        unsigned o = (unsigned)s;  // offset of entry c
        int c = 0;
#pragma unroll 4
        for (int j = 0; j < nsite; j++) {
            const uint4 pj = ru[j];
            const float x = (float)(int)(pi.x - pj.x) * ux, y = (float)(int)(pi.y - pj.y) * uy,
                        z = (float)(int)(pi.z - pj.z) * uz;
            if ((unsigned)(j - j0) >= (unsigned)ns && x * x + y * y + z * z <= rl2[pj.w]) {
                nbl[o] = (uint16_t)j;
                o += (unsigned)cap;
                c++;
            }
        }
        nb[s] = make_float4((float)rp[3 * s], (float)rp[3 * s + 1], (float)rp[3 * s + 2], __int_as_float(c));
    }
}

// f(j) for the entries of a neighbour list in order; the next group of eight
// entries is loaded while a group is processed
template <class Fn>
__device__ __forceinline__ void list_loop(const uint16_t *nbl, unsigned s, int n, unsigned cap, Fn &&f) {
    int q = 0;
    unsigned o = s;  // offset of entry q
    if (n >= 8) {
        int cur[8];
#pragma unroll
        for (int u = 0; u < 8; u++) cur[u] = nbl[o + u * cap];
        for (; q + 8 <= n; q += 8) {
            int nxt[8];
            const bool more = q + 16 <= n;
            o += 8 * cap;
            if (more) {
#pragma unroll
                for (int u = 0; u < 8; u++) nxt[u] = nbl[o + u * cap];
            }
#pragma unroll
            for (int u = 0; u < 8; u++) f(cur[u]);
            if (more) {
#pragma unroll
                for (int u = 0; u < 8; u++) cur[u] = nxt[u];
            }
        }
    }
    for (; q < n; q++, o += cap) f(nbl[o]);
}

// energy parts of a configuration: framework map, sorbate-sorbate LJ,
// stretch, bending, torsion
constexpr int kParts = 5;

// force on site s (map, sorbate-sorbate LJ over its neighbour list, bonded
// terms) into f; its energy shares added to ep; bad on a failed evaluation
template <int PREC>
__device__ __forceinline__ void site_force(const HDev &P, const double *rp, const double *r, const uint4 *ru,
                                           const uint16_t *nbl, int cap, int s, int n, double f[3],
                                           double (&ep)[kParts], double &bad) {
    const int ns = P.ns, m = s / ns, a = s - m * ns;
    const double *ri = r + 3 * s;
    f[0] = f[1] = f[2] = 0.0;
    // map corners requested first; in FP32 the bonded terms are evaluated
    // while they arrive (in FP64 that costs more registers than it saves).
    // The bonded terms are added last.
    Corners g;
    const bool mapped = P.has_map[a] != 0;
    const bool inside = mapped && hermite_fetch(P.map[a], ri, g);
    double fb[3] = {0.0, 0.0, 0.0}, eb[3] = {0.0, 0.0, 0.0};
    if (PREC != MH_PREC_FP64) bonded_site(P, rp + 3 * (m * ns), a, fb, eb);
    if (mapped) {
        double p[4];
        if (inside && hermite_eval<PREC>(P.map[a], g, p)) {
            ep[0] += p[0];
            for (int d = 0; d < 3; d++) f[d] += p[1 + d];
        } else {
            bad = 1.0;
        }
    }
    // sorbate-sorbate LJ against the listed partners in site order: the list
    // holds every partner within hicut, so the sums are those of the loop over
    // all sites of the other molecules that skips pairs beyond hicut and
    // switched-off type pairs
    double el = 0.0;
    if (PREC == MH_PREC_FP64) {
        const double *pp = P.pp64[a * HS];
        list_loop(nbl, (unsigned)s, n, (unsigned)cap, [&](int j) {
            const double *pr = pp + 4 * ru[j].w;
            const double sx = fold64(ri[0] - r[3 * j], P.full[0]);
            const double sy = fold64(ri[1] - r[3 * j + 1], P.full[1]);
            const double sz = fold64(ri[2] - r[3 * j + 2], P.full[2]);
            const double r2 = sx * sx + sy * sy + sz * sz;
            if (r2 > pr[2]) return;
            if (r2 < pr[3]) bad = 1.0;
            const double r2i = 1.0 / r2, r6i = r2i * r2i * r2i, r12i = r6i * r6i;
            el += pr[0] * r12i - pr[1] * r6i;
            const double fm = (12.0 * pr[0] * r12i - 6.0 * pr[1] * r6i) * r2i;
            f[0] += sx * fm;
            f[1] += sy * fm;
            f[2] += sz * fm;
        });
    } else {
#pragma clang fp contract(fast)
        const float ux = P.fxu[0], uy = P.fxu[1], uz = P.fxu[2];
        const uint4 pi = ru[s];
        const float4 *pp = P.pp32 + a * HS;
        float elf = 0.0f, fxf = 0.0f, fyf = 0.0f, fzf = 0.0f;
        bool badf = false;
        list_loop(nbl, (unsigned)s, n, (unsigned)cap, [&](int j) {
            const uint4 pj = ru[j];
            const float x = (float)(int)(pi.x - pj.x) * ux, y = (float)(int)(pi.y - pj.y) * uy,
                        z = (float)(int)(pi.z - pj.z) * uz;
            const float r2 = x * x + y * y + z * z;
            const float4 pr = pp[pj.w];
            const bool in = !(r2 > pr.z);
            badf |= in && r2 < pr.w;
            // A/r^12 - B/r^6 and its force factor (12 A/r^12 - 6 B/r^6)/r^2
            const float r2i = in ? rcp_fast(r2) : 0.0f, r6i = r2i * r2i * r2i;
            const float ta = pr.x * r6i * r6i, tb = pr.y * r6i, de = ta - tb;
            elf += de;
            const float fm = (de + ta) * (6.0f * r2i);
            fxf += x * fm;
            fyf += y * fm;
            fzf += z * fm;
        });
        if (badf) bad = 1.0;
        el = elf;
        f[0] += fxf;
        f[1] += fyf;
        f[2] += fzf;
    }
    ep[1] += 0.5 * el;  // each pair is seen from both of its sites
    if (PREC == MH_PREC_FP64) bonded_site(P, rp + 3 * (m * ns), a, fb, eb);
    for (int k = 0; k < 3; k++) ep[2 + k] += eb[k];
    for (int d = 0; d < 3; d++) f[d] += fb[d];
}

// energy and forces of the configuration.  With Fo the forces are stored
// there; otherwise each site's acceleration goes to ga and, after a
// velocity-Verlet step (step > 0), its velocity gv takes the second half
// kick.  ke = block total of m v^2 with the velocities after this call.
// Accelerations and velocities written for an evaluation that fails are not
// used (the move or run ends there).  With MD, parts = block totals of the
// energy parts and mom = total momentum sum m v; without, parts[0] = the
// total energy.
template <int NT, int PREC, bool MD>
__device__ void hmc_forces_parts(const HDev &P, const double *rp, const double *r, const uint4 *ru,
                                 const float4 *nb, int N, const uint16_t *nbl, int cap, double *gv, double *ga,
                                 double *Fo, int step, double (&parts)[kParts], int &fail, double &ke,
                                 double (&mom)[3], double *red) {
    const int ns = P.ns, nsite = N * ns;
    double ep[kParts] = {0.0, 0.0, 0.0, 0.0, 0.0}, bad = 0.0, k2 = 0.0, px = 0.0, py = 0.0, pz = 0.0;
#pragma unroll 1
    for (int s = threadIdx.x; s < nsite; s += NT) {
        double f[3];
        site_force<PREC>(P, rp, r, ru, nbl, cap, s, __float_as_int(nb[s].w), f, ep, bad);
        if (Fo) {
            for (int i = 0; i < 3; i++) Fo[3 * s + i] = f[i];
            continue;
        }
        const int at = s - (s / ns) * ns;
        double v[3];
        for (int i = 0; i < 3; i++) {
            const double acc = f[i] * (1.0 / P.mass[at]) * kScaleF;
            ga[3 * s + i] = acc;
            v[i] = gv[3 * s + i];
            if (step > 0) {
                v[i] = v[i] + acc * 0.5 * P.dt;
                gv[3 * s + i] = v[i];
            }
        }
        k2 = k2 + (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) * P.mass[at];
        if (MD) {
            px += P.mass[at] * v[0];
            py += P.mass[at] * v[1];
            pz += P.mass[at] * v[2];
        }
    }
    if (MD) {
        double w[kParts + 5] = {ep[0], ep[1], ep[2], ep[3], ep[4], bad, k2, px, py, pz};
        block_sum<NT, kParts + 5>(w, red);
        for (int k = 0; k < kParts; k++) parts[k] = w[k];
        fail = w[kParts] != 0.0;
        ke = w[kParts + 1];
        for (int d = 0; d < 3; d++) mom[d] = w[kParts + 2 + d];
    } else {
        double w[3] = {ep[0] + ep[1] + ep[2] + ep[3] + ep[4], bad, k2};
        block_sum<NT, 3>(w, red);
        parts[0] = w[0];
        fail = w[1] != 0.0;
        ke = w[2];
    }
}

// hmc_forces_parts with the total energy U
template <int NT, int PREC>
__device__ void hmc_forces(const HDev &P, const double *rp, const double *r, const uint4 *ru, const float4 *nb,
                           int N, const uint16_t *nbl, int cap, double *gv, double *ga, double *Fo, int step,
                           double &U, int &fail, double &ke, double *red) {
    double parts[kParts], mom[3];
    hmc_forces_parts<NT, PREC, false>(P, rp, r, ru, nb, N, nbl, cap, gv, ga, Fo, step, parts, fail, ke, mom, red);
    U = parts[0];
}

template <int NT, int PREC>
__global__ void __launch_bounds__(NT) k_hmc(const HDev *__restrict__ Pg, HRep R, int nmoves) {
    extern __shared__ double dyn[];
    __shared__ HDev P;
    __shared__ double red[NT / 64 * 4];
    __shared__ double s_z2;
    const int tid = threadIdx.x, rep = blockIdx.x;
    for (int i = tid; i < (int)(sizeof(HDev) / 4); i += NT)
        reinterpret_cast<int *>(&P)[i] = reinterpret_cast<const int *>(Pg)[i];
    __syncthreads();
    if (R.err[rep]) return;
    const int ns = P.ns, N = R.N[rep], nsite = N * ns;
    const long long off = (long long)rep * R.cap * 3;
    double *rp = dyn, *r = dyn + (long long)R.cap * 3;
    uint4 *ru = reinterpret_cast<uint4 *>(dyn + (long long)R.cap * 6);
    float4 *nb = reinterpret_cast<float4 *>(ru + R.cap);
    uint16_t *nbl = R.nbl + (long long)rep * R.cap * R.cap;
    double *gv = R.v + off, *ga = R.a + off;
    for (int i = tid; i < nsite * 3; i += NT) {
        rp[i] = R.rp[off + i];
        r[i] = R.r[off + i];
    }
    for (int i = tid; i < nsite; i += NT) put_fx(P, ru, R.r + off, i, i % ns);
    Rng rng{R.ix[rep], R.iy[rep], R.first[rep], R.iseed[rep]};
    rng_first(rng);
    int ghave = R.ghave[rep];
    double gz2 = R.gz2[rep], energy = R.energy[rep], lastdH = R.dH[rep];
    long long att = R.att[rep], acc = R.acc[rep], nfail = R.fail[rep];
    int err = 0;
    __syncthreads();
    hmc_lists<NT>(P, ru, rp, N, nbl, R.cap, nb);
    const double dt = P.dt, dt2 = dt * dt * 0.5;
    for (int mv = 0; mv < nmoves; mv++) {
        att++;
        // ---- Gaussian velocities (hybridmoves_boltzVel); start configuration saved
        const unsigned G = 3u * nsite, c = ghave ? 1u : 0u;
        const unsigned used = G >= c ? G - c : 0u;  // draws-consuming Gaussians
        const unsigned npairs = (used + 1u) / 2u;
#pragma unroll 1
        for (int s = tid; s < nsite; s += NT) {
            const int at = s - (s / ns) * ns;
            for (int i = 0; i < 3; i++) {
                const unsigned g = 3u * s + i;
                double z;
                if (c && g == 0) z = gz2;
                else {
                    const unsigned gg = g - c;
                    z = gauss_pair(P, rng, gg >> 1, gg & 1);
                    if ((used & 1u) && gg == used - 1u) s_z2 = gauss_pair(P, rng, gg >> 1, 1);
                }
                gv[3 * s + i] = 0.0 + P.sigma[at] * z;
                R.rp0[off + 3 * s + i] = rp[3 * s + i];
                R.r0[off + 3 * s + i] = r[3 * s + i];
            }
        }
        if (G < c) {  // no atoms: the saved value stays pending
        } else {
            ghave = (used & 1u) ? 1 : 0;
        }
        rng_jump(P, rng, 2u * npairs);
        // ---- energy and forces at the start (step 0) and after each NVE
        // velocity-Verlet step (velverlet_nveIntegrate)
        bool good = true;
        double U0 = 0.0, U1 = 0.0, ke0 = 0.0, ke1 = 0.0;
        for (int step = 0; step <= P.nmd; step++) {
            if (step > 0) {
#pragma unroll 1
                for (int s = tid; s < nsite; s += NT) {
                    for (int i = 0; i < 3; i++) {
                        const double vi = gv[3 * s + i], ai = ga[3 * s + i];
                        rp[3 * s + i] = rp[3 * s + i] + vi * dt + ai * dt2;
                        r[3 * s + i] = r[3 * s + i] + vi * dt + ai * dt2;
                        gv[3 * s + i] = vi + ai * 0.5 * dt;
                    }
                    put_fx(P, ru, r, s, s - (s / ns) * ns);
                }
            }
            if (__syncthreads_or(moved_far<NT>(rp, nb, nsite))) hmc_lists<NT>(P, ru, rp, N, nbl, R.cap, nb);
            int fl;
            double k2;
            hmc_forces<NT, PREC>(P, rp, r, ru, nb, N, nbl, R.cap, gv, ga, nullptr, step, U1, fl, k2, red);
            if (step == 0) {
                if (ghave && (used & 1u)) gz2 = s_z2;  // written before the block sums
                if (fl || U1 > P.max_nrg0) {
                    err = 1;  // the reference stops here (hybridmoves.F90:224-229)
                    break;
                }
                U0 = U1;
                ke0 = 0.5 * k2 * kScaleKE;
            } else if (fl || U1 > P.max_nrg_md) {
                good = false;
                break;
            }
            ke1 = 0.5 * k2 * kScaleKE;
        }
        if (err) break;
        bool accepted = false;
        if (good) {
            const double dH = (U1 - U0) + (ke1 - ke0) / 4.184;
            lastdH = dH;
            double u = -P.rti * dH;
            u = u + log(1.0);
            u = fmin(fmax(u, -30.0), 30.0);
            const double boltz = exp(u);
            accepted = boltz > rranf(rng);
        } else {
            nfail++;
            lastdH = NAN;
        }
        if (accepted) {
            acc++;
            energy = U1;
#pragma unroll 1
            for (int s = tid; s < nsite; s += NT) {
                for (int i = 0; i < 3; i++) r[3 * s + i] = wrap_full(r[3 * s + i], P.full[i]);
                put_fx(P, ru, r, s, s - (s / ns) * ns);
            }
        } else {
            energy = U0;
#pragma unroll 1
            for (int s = tid; s < nsite; s += NT) {
                for (int i = 0; i < 3; i++) {
                    rp[3 * s + i] = R.rp0[off + 3 * s + i];
                    r[3 * s + i] = R.r0[off + 3 * s + i];
                }
                put_fx(P, ru, r, s, s - (s / ns) * ns);
            }
        }
        __syncthreads();
    }
    for (int i = tid; i < nsite * 3; i += NT) {
        R.rp[off + i] = rp[i];
        R.r[off + i] = r[i];
    }
    if (tid == 0) {
        R.err[rep] = err;
        R.ix[rep] = rng.ix;
        R.iy[rep] = rng.iy;
        R.first[rep] = rng.first;
        R.iseed[rep] = rng.iseed;
        R.ghave[rep] = ghave;
        R.gz2[rep] = gz2;
        R.energy[rep] = energy;
        R.dH[rep] = lastdH;
        R.att[rep] = att;
        R.acc[rep] = acc;
        R.fail[rep] = nfail;
    }
}

// plain NVE molecular dynamics of every replica (md_doSim with
// velverlet_nveIntegrate, md.F90:592-623, velverlet.F90:148-288): nsteps
// velocity-Verlet steps from the current positions and velocities; per step
// the block totals go to trace[rep][step][MH_MD_NQ]: energy parts (map, LJ,
// stretch, bending, torsion; kcal/mol), sum m v^2 over sites, sum over
// molecules |sum m v|^2 / M (amu A^2/ps^2) and the total momentum (amu A/ps).
// A failed evaluation stops the replica with err 2 (the reference stops).
template <int NT, int PREC>
__global__ void __launch_bounds__(NT) k_md(const HDev *__restrict__ Pg, HRep R, int nsteps, double *trace) {
    extern __shared__ double dyn[];
    __shared__ HDev P;
    __shared__ double red[NT / 64 * 16];
    const int tid = threadIdx.x, rep = blockIdx.x;
    for (int i = tid; i < (int)(sizeof(HDev) / 4); i += NT)
        reinterpret_cast<int *>(&P)[i] = reinterpret_cast<const int *>(Pg)[i];
    __syncthreads();
    if (R.err[rep]) return;
    const int ns = P.ns, N = R.N[rep], nsite = N * ns;
    const long long off = (long long)rep * R.cap * 3;
    double *rp = dyn, *r = dyn + (long long)R.cap * 3;
    uint4 *ru = reinterpret_cast<uint4 *>(dyn + (long long)R.cap * 6);
    float4 *nb = reinterpret_cast<float4 *>(ru + R.cap);
    double *pm = reinterpret_cast<double *>(nb + R.cap);  // [cap][3] site momenta
    uint16_t *nbl = R.nbl + (long long)rep * R.cap * R.cap;
    double *gv = R.v + off, *ga = R.a + off;
    for (int i = tid; i < nsite * 3; i += NT) {
        rp[i] = R.rp[off + i];
        r[i] = R.r[off + i];
    }
    for (int i = tid; i < nsite; i += NT) put_fx(P, ru, R.r + off, i, i % ns);
    __syncthreads();
    hmc_lists<NT>(P, ru, rp, N, nbl, R.cap, nb);
    const double dt = P.dt, dt2 = dt * dt * 0.5;
    double parts[kParts], mom[3], k2;
    int fl, err = 0;
    // accelerations of the start configuration
    hmc_forces_parts<NT, PREC, true>(P, rp, r, ru, nb, N, nbl, R.cap, gv, ga, nullptr, 0, parts, fl, k2, mom, red);
    if (fl) err = 2;
    for (int step = 1; step <= nsteps && !err; step++) {
#pragma unroll 1
        for (int s = tid; s < nsite; s += NT) {
            for (int i = 0; i < 3; i++) {
                const double vi = gv[3 * s + i], ai = ga[3 * s + i];
                rp[3 * s + i] = rp[3 * s + i] + vi * dt + ai * dt2;
                r[3 * s + i] = r[3 * s + i] + vi * dt + ai * dt2;
                gv[3 * s + i] = vi + ai * 0.5 * dt;
            }
            put_fx(P, ru, r, s, s - (s / ns) * ns);
        }
        if (__syncthreads_or(moved_far<NT>(rp, nb, nsite))) hmc_lists<NT>(P, ru, rp, N, nbl, R.cap, nb);
        hmc_forces_parts<NT, PREC, true>(P, rp, r, ru, nb, N, nbl, R.cap, gv, ga, nullptr, step, parts, fl, k2,
                                         mom, red);
        // velverlet_nveIntegrate evaluates with max_nrg = 1000 (velverlet.F90:197-198)
        if (fl || parts[0] + parts[1] + parts[2] + parts[3] + parts[4] > P.max_nrg_md) {
            err = 2;
            break;
        }
        // molecular kinetic term sum_m |sum m v|^2 / M^2 (config_kineticTemp,
        // config.F90:2579-2607): site momenta in LDS, summed per molecule by
        // one lane
#pragma unroll 1
        for (int s = tid; s < nsite; s += NT) {
            const double ms = P.mass[s - (s / ns) * ns];
            for (int i = 0; i < 3; i++) pm[3 * s + i] = gv[3 * s + i] * ms;
        }
        __syncthreads();
        double kmol = 0.0;
#pragma unroll 1
        for (int m = tid; m < N; m += NT) {
            double q[3] = {0.0, 0.0, 0.0};
            for (int a = 0; a < ns; a++)
                for (int i = 0; i < 3; i++) q[i] = pm[3 * (m * ns + a) + i] + q[i];
            kmol = (q[0] * q[0] + q[1] * q[1] + q[2] * q[2]) * P.mmassi2 + kmol;
        }
        double w[1] = {kmol};
        block_sum<NT, 1>(w, red);
        if (tid == 0) {
            double *o = trace + ((long long)rep * nsteps + (step - 1)) * MH_MD_NQ;
            for (int k = 0; k < kParts; k++) o[k] = parts[k];
            o[5] = k2;
            o[6] = w[0];
            for (int d = 0; d < 3; d++) o[7 + d] = mom[d];
        }
    }
    for (int i = tid; i < nsite * 3; i += NT) {
        R.rp[off + i] = rp[i];
        R.r[off + i] = r[i];
    }
    if (tid == 0 && err) R.err[rep] = err;
}

// energy and forces of one replica's configuration (tests, checks)
template <int NT, int PREC>
__global__ void __launch_bounds__(NT) k_hmc_forces(const HDev *__restrict__ Pg, HRep R, int rep, double *F_out,
                                                   double *e_out) {
    extern __shared__ double dyn[];
    __shared__ HDev P;
    __shared__ double red[NT / 64 * 4];
    const int tid = threadIdx.x;
    for (int i = tid; i < (int)(sizeof(HDev) / 4); i += NT)
        reinterpret_cast<int *>(&P)[i] = reinterpret_cast<const int *>(Pg)[i];
    __syncthreads();
    const int ns = P.ns, N = R.N[rep], nsite = N * ns;
    const long long off = (long long)rep * R.cap * 3;
    double *rp = dyn, *r = dyn + (long long)R.cap * 3;
    uint4 *ru = reinterpret_cast<uint4 *>(dyn + (long long)R.cap * 6);
    float4 *nb = reinterpret_cast<float4 *>(ru + R.cap);
    uint16_t *nbl = R.nbl + (long long)rep * R.cap * R.cap;
    for (int i = tid; i < nsite * 3; i += NT) {
        rp[i] = R.rp[off + i];
        r[i] = R.r[off + i];
    }
    for (int i = tid; i < nsite; i += NT) put_fx(P, ru, R.r + off, i, i % ns);
    __syncthreads();
    hmc_lists<NT>(P, ru, rp, N, nbl, R.cap, nb);
    double U, k2;
    int fl;
    hmc_forces<NT, PREC>(P, rp, r, ru, nb, N, nbl, R.cap, nullptr, nullptr, F_out, 0, U, fl, k2, red);
    if (tid == 0) {
        e_out[0] = U;
        e_out[1] = fl ? 0.0 : 1.0;
    }
}

}  // namespace mh

using namespace mh;

struct mh_hmc {
    mh_ctx *ctx = nullptr;
    int nrep = 0, maxN = 0, ns = 0, prec = 0, threads = 512;
    size_t lds = 0;
    HDev host;
    DBuf dP;
    DBuf N, err, ix, iy, first, iseed, ghave, gz2, energy, dH, att, acc, fail, rp, r, rp0, r0, v, a, nbl;
    HRep rep() {
        HRep h;
        h.N = N.as<int>();
        h.err = err.as<int>();
        h.ix = ix.as<int32_t>();
        h.iy = iy.as<int32_t>();
        h.first = first.as<int32_t>();
        h.iseed = iseed.as<int32_t>();
        h.ghave = ghave.as<int32_t>();
        h.gz2 = gz2.as<double>();
        h.energy = energy.as<double>();
        h.dH = dH.as<double>();
        h.att = att.as<long long>();
        h.acc = acc.as<long long>();
        h.fail = fail.as<long long>();
        h.rp = rp.as<double>();
        h.r = r.as<double>();
        h.rp0 = rp0.as<double>();
        h.r0 = r0.as<double>();
        h.v = v.as<double>();
        h.a = a.as<double>();
        h.nbl = nbl.as<uint16_t>();
        h.cap = maxN * ns;
        return h;
    }
};

static uint32_t xs_step(uint32_t x) {
    x ^= x << 13;
    x ^= x >> 17;
    x ^= x << 5;
    return x;
}

static uint32_t gf2_apply_host(const uint32_t *col, uint32_t x) {
    uint32_t y = 0;
    for (int b = 0; b < 32; b++)
        if (x & (1u << b)) y ^= col[b];
    return y;
}

template <int NT, int PREC>
static int launch_hmc_nt(mh_hmc *h, int nmoves) {
    auto kern = k_hmc<NT, PREC>;
    (void)hipFuncSetAttribute(reinterpret_cast<const void *>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)h->lds);
    hipLaunchKernelGGL(kern, dim3(h->nrep), dim3(NT), h->lds, h->ctx->stream, h->dP.as<HDev>(), h->rep(), nmoves);
    return MH_OK;
}

template <int PREC>
static int launch_hmc(mh_hmc *h, int nmoves) {
    switch (h->threads) {
    case 128: return launch_hmc_nt<128, PREC>(h, nmoves);
    case 256: return launch_hmc_nt<256, PREC>(h, nmoves);
    case 1024: return launch_hmc_nt<1024, PREC>(h, nmoves);
    default: return launch_hmc_nt<512, PREC>(h, nmoves);
    }
}

// LDS per site: rp and r (FP64), fixed-point position, list bookkeeping
constexpr size_t kLdsPerSite = 6 * sizeof(double) + sizeof(uint4) + sizeof(float4);

static int plan_hmc(mh_hmc *h) {
    int dev_max = 65536;
    (void)hipDeviceGetAttribute(&dev_max, hipDeviceAttributeMaxSharedMemoryPerBlock, h->ctx->device);
    h->lds = (size_t)h->maxN * h->ns * kLdsPerSite;
    MH_CHECK((long long)h->lds + (long long)sizeof(HDev) + 8192 <= dev_max, MH_ECAPACITY,
             "mh_hmc: maxN * nsites too large for LDS (positions of a replica are held in LDS)");
    return MH_OK;
}

template <int NT, int PREC>
static void launch_md_nt(mh_hmc *h, int nsteps, double *dtrace, size_t lds) {
    auto kern = k_md<NT, PREC>;
    (void)hipFuncSetAttribute(reinterpret_cast<const void *>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)lds);
    hipLaunchKernelGGL(kern, dim3(h->nrep), dim3(NT), lds, h->ctx->stream, h->dP.as<HDev>(), h->rep(), nsteps,
                       dtrace);
}

template <int PREC>
static void launch_md(mh_hmc *h, int nsteps, double *dtrace, size_t lds) {
    switch (h->threads) {
    case 128: launch_md_nt<128, PREC>(h, nsteps, dtrace, lds); break;
    case 256: launch_md_nt<256, PREC>(h, nsteps, dtrace, lds); break;
    case 1024: launch_md_nt<1024, PREC>(h, nsteps, dtrace, lds); break;
    default: launch_md_nt<512, PREC>(h, nsteps, dtrace, lds); break;
    }
}

extern "C" {

int mh_hmc_create(mh_ctx *ctx, const mh_hmc_params *p, int nrep, const int32_t *seeds, int maxN, mh_hmc **out) {
    MH_CHECK(ctx && p && seeds && out, MH_EINVAL, "mh_hmc_create: bad arguments");
    MH_CHECK(!ctx->slanted, MH_EINVAL, "mh_hmc_create: slanted cells are not offered for HMC");
    MH_CHECK(nrep >= 1 && nrep <= 65535 && maxN >= 1, MH_EINVAL, "mh_hmc_create: bad nrep / maxN");
    MH_CHECK(p->nsites >= 1 && p->nsites <= HS && p->ntypes >= 1 && p->ntypes <= 8, MH_EINVAL,
             "mh_hmc_create: nsites in [1,8], ntypes in [1,8]");
    MH_CHECK(p->nbonds >= 0 && p->nbonds <= HT && p->nangles >= 0 && p->nangles <= HT && p->ntorsions >= 0 &&
                 p->ntorsions <= HT,
             MH_EINVAL, "mh_hmc_create: at most 16 bonds, angles and torsions");
    MH_CHECK(p->temperature > 0.0 && p->dt_ps > 0.0 && p->md_steps >= 1, MH_EINVAL,
             "mh_hmc_create: temperature, dt_ps and md_steps must be > 0");
    MH_CHECK(p->precision == MH_PREC_FP64 || p->precision == MH_PREC_FP32, MH_EINVAL, "mh_hmc_create: precision");
    MH_CHECK(ctx->fund_ell[0] > 0.0, MH_ESTATE, "mh_hmc_create: upload the framework (cell) first");
    MH_HIP(hipSetDevice(ctx->device));
    std::unique_ptr<mh_hmc> h(new mh_hmc());
    h->ctx = ctx;
    h->nrep = nrep;
    h->maxN = maxN;
    h->ns = p->nsites;
    h->prec = p->precision;
    // workgroup size measured on the C5 butane workload: FP32 is fastest at
    // 512 lanes, FP64 at 1024 once a replica holds more than 512 sites
    h->threads = (h->prec == MH_PREC_FP64 && maxN * p->nsites > 512) ? 1024 : 512;
    HDev &H = h->host;
    std::memset(&H, 0, sizeof H);
    H.ns = p->nsites;
    H.nst = p->ntypes;
    H.nb = p->nbonds;
    H.na = p->nangles;
    H.nt = p->ntorsions;
    H.nmd = p->md_steps;
    H.prec = p->precision;
    H.maxN = maxN;
    for (int a = 0; a < H.ns; a++) {
        MH_CHECK(p->site_type[a] >= 0 && p->site_type[a] < p->ntypes && p->site_mass[a] > 0.0, MH_EINVAL,
                 "mh_hmc_create: bad site type or mass");
        H.st[a] = p->site_type[a];
        H.mass[a] = p->site_mass[a];
        H.ainv[a] = kScaleF / p->site_mass[a];
        // hybridmoves_boltzVel (hybridmoves.F90:280)
        H.sigma[a] = std::sqrt((kKjKb * p->temperature) / (kScaleKE * p->site_mass[a]));
        if (p->site_map[a]) {
            H.has_map[a] = 1;
            int rc = map_dense(ctx, const_cast<mh_map *>(p->site_map[a]), &H.map[a].dense);
            if (rc) return rc;
            for (int i = 0; i < 3; i++) {
                H.map[a].n[i] = p->site_map[a]->npnts[i];
                H.map[a].step[i] = p->site_map[a]->step[i];
                H.map[a].ell[i] = ctx->fund_ell[i];
            }
        }
    }
    for (int a = 0; a < H.ns; a++)
        for (int b = 0; b < H.ns; b++) {
            const int t = H.st[a] * p->ntypes + H.st[b], q = a * HS + b;
            const bool on = p->lj_on[t] != 0;
            const double h2 = on ? p->lj_hicut2[t] : -1.0;
            H.pp64[q][0] = p->lj_A[t];
            H.pp64[q][1] = p->lj_B[t];
            H.pp64[q][2] = h2;
            H.pp64[q][3] = p->lj_locut2[t];
            H.pp32[q] = make_float4((float)p->lj_A[t], (float)p->lj_B[t], (float)h2, (float)p->lj_locut2[t]);
            const double rl = on ? std::sqrt(std::max(p->lj_hicut2[t], 0.0)) + kSkin : 0.0;
            H.rl2[q] = on ? (float)(rl * rl) : -1.0f;
        }
    auto inr = [&](int i) { return i >= 0 && i < p->nsites; };
    for (int q = 0; q < H.nb; q++) {
        MH_CHECK(inr(p->bond[q][0]) && inr(p->bond[q][1]), MH_EINVAL, "mh_hmc_create: bond atom out of range");
        H.bi[q][0] = p->bond[q][0];
        H.bi[q][1] = p->bond[q][1];
        H.bk[q] = p->bond_k[q];
        H.br0[q] = p->bond_r0[q];
    }
    for (int q = 0; q < H.na; q++) {
        for (int k = 0; k < 3; k++) {
            MH_CHECK(inr(p->angle[q][k]), MH_EINVAL, "mh_hmc_create: angle atom out of range");
            H.ai[q][k] = p->angle[q][k];
        }
        H.ak[q] = p->angle_k[q];
        H.at0[q] = p->angle_theta0[q];
    }
    for (int q = 0; q < H.nt; q++) {
        for (int k = 0; k < 4; k++) {
            MH_CHECK(inr(p->torsion[q][k]), MH_EINVAL, "mh_hmc_create: torsion atom out of range");
            H.ti[q][k] = p->torsion[q][k];
        }
        for (int k = 0; k < 6; k++) H.tc[q][k] = p->torsion_c[q][k];
    }
    for (int i = 0; i < 3; i++) {
        H.full[i] = ctx->full_ell[i];
        H.fund[i] = ctx->fund_ell[i];
        H.fxs[i] = 4294967296.0 / ctx->full_ell[i];
        H.fxu[i] = (float)(ctx->full_ell[i] / 4294967296.0);
    }
    double mmass = 0.0;
    for (int a = 0; a < H.ns; a++) mmass += H.mass[a];
    H.mmassi2 = 1.0 / mmass / mmass;
    H.T = p->temperature;
    H.dt = p->dt_ps;
    H.rti = 1.0 / ((8.314e-3 / 4.184) * p->temperature);
    H.max_nrg_md = p->max_nrg_md;
    H.max_nrg0 = p->max_nrg_start;
    // jump-ahead tables of rranf_NR
    for (int b = 0; b < 32; b++) H.xsj[0][b] = xs_step(1u << b);
    for (int k = 1; k < JLEV; k++)
        for (int b = 0; b < 32; b++) H.xsj[k][b] = gf2_apply_host(H.xsj[k - 1], H.xsj[k - 1][b]);
    H.pmj[0] = 16807ULL;
    for (int k = 1; k < JLEV; k++) H.pmj[k] = (H.pmj[k - 1] * H.pmj[k - 1]) % 2147483647ULL;
    int rc;
    if ((rc = plan_hmc(h.get()))) return rc;
    if ((rc = h->dP.alloc(sizeof(HDev)))) return rc;
    MH_HIP(hipMemcpy(h->dP.p, &H, sizeof(HDev), hipMemcpyHostToDevice));
    const size_t R = nrep, nc = (size_t)nrep * maxN * h->ns * 3;
    if ((rc = h->N.alloc(4 * R)) || (rc = h->err.alloc(4 * R)) || (rc = h->ix.alloc(4 * R)) ||
        (rc = h->iy.alloc(4 * R)) || (rc = h->first.alloc(4 * R)) || (rc = h->iseed.alloc(4 * R)) ||
        (rc = h->ghave.alloc(4 * R)) || (rc = h->gz2.alloc(8 * R)) || (rc = h->energy.alloc(8 * R)) ||
        (rc = h->dH.alloc(8 * R)) || (rc = h->att.alloc(8 * R)) || (rc = h->acc.alloc(8 * R)) ||
        (rc = h->fail.alloc(8 * R)) || (rc = h->rp.alloc(8 * nc)) || (rc = h->r.alloc(8 * nc)) ||
        (rc = h->rp0.alloc(8 * nc)) || (rc = h->r0.alloc(8 * nc)) || (rc = h->v.alloc(8 * nc)) ||
        (rc = h->a.alloc(8 * nc)) ||
        (rc = h->nbl.alloc(2 * R * (size_t)maxN * h->ns * maxN * h->ns)))
        return rc;
    std::vector<int32_t> zi(R, 0), ones(R, 1), sd(seeds, seeds + R);
    MH_HIP(hipMemcpy(h->first.p, ones.data(), 4 * R, hipMemcpyHostToDevice));
    MH_HIP(hipMemcpy(h->iseed.p, sd.data(), 4 * R, hipMemcpyHostToDevice));
    for (DBuf *b : {&h->N, &h->err, &h->ix, &h->iy, &h->ghave, &h->gz2, &h->energy, &h->dH, &h->att, &h->acc,
                    &h->fail})
        MH_HIP(hipMemset(b->p, 0, b->n));
    ctx_retain(ctx);
    *out = h.release();
    return MH_OK;
}

int mh_hmc_destroy(mh_hmc *h) {
    if (!h) return MH_OK;
    mh_ctx *ctx = h->ctx;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    delete h;
    ctx_release(ctx);
    return MH_OK;
}

int mh_hmc_set_threads(mh_hmc *h, int threads) {
    MH_CHECK(h && (threads == 128 || threads == 256 || threads == 512 || threads == 1024), MH_EINVAL,
             "mh_hmc_set_threads: threads must be 128, 256, 512 or 1024");
    const int old = h->threads;
    h->threads = threads;
    int rc = plan_hmc(h);
    if (rc) h->threads = old, (void)plan_hmc(h);
    return rc;
}

int mh_hmc_set_config(mh_hmc *h, int rep, int n, const double *rp) {
    MH_CHECK(h && rep >= 0 && rep < h->nrep && n >= 0 && n <= h->maxN && (n == 0 || rp), MH_EINVAL,
             "mh_hmc_set_config: bad arguments");
    MH_HIP(hipSetDevice(h->ctx->device));
    const size_t nc = (size_t)n * h->ns * 3, off = (size_t)rep * h->maxN * h->ns * 3;
    std::vector<double> r(nc);
    for (size_t o = 0; o < nc; o++) {
        const double L = h->host.full[o % 3];
        r[o] = rp[o] - std::floor(rp[o] / L) * L;
    }
    if (nc) {
        MH_HIP(hipMemcpy(h->rp.as<double>() + off, rp, 8 * nc, hipMemcpyHostToDevice));
        MH_HIP(hipMemcpy(h->r.as<double>() + off, r.data(), 8 * nc, hipMemcpyHostToDevice));
    }
    MH_HIP(hipMemcpy(h->N.as<int>() + rep, &n, 4, hipMemcpyHostToDevice));
    double e = 0.0;
    int ok = 1;
    int rc = mh_hmc_energy_forces(h, rep, &e, nullptr, &ok);
    if (rc) return rc;
    MH_HIP(hipMemcpy(h->energy.as<double>() + rep, &e, 8, hipMemcpyHostToDevice));
    return MH_OK;
}

int mh_hmc_run(mh_hmc *h, int nmoves) {
    MH_CHECK(h && nmoves >= 0, MH_EINVAL, "mh_hmc_run: bad arguments");
    if (!nmoves) return MH_OK;
    MH_HIP(hipSetDevice(h->ctx->device));
    KTimer t(h->ctx, h->prec == MH_PREC_FP32 ? "hmc32" : "hmc64");
    int rc = h->prec == MH_PREC_FP32 ? launch_hmc<MH_PREC_FP32>(h, nmoves) : launch_hmc<MH_PREC_FP64>(h, nmoves);
    if (rc) return rc;
    MH_HIP(hipGetLastError());
    MH_HIP(hipStreamSynchronize(h->ctx->stream));
    t.stop();
    std::vector<int> err(h->nrep);
    MH_HIP(hipMemcpy(err.data(), h->err.p, 4 * h->nrep, hipMemcpyDeviceToHost));
    for (int i = 0; i < h->nrep; i++)
        MH_CHECK(err[i] == 0, MH_ESTATE, "mh_hmc_run: a replica's start configuration failed to evaluate");
    return MH_OK;
}

int mh_hmc_get_state(mh_hmc *h, int32_t *N, double *energy, int64_t *att, int64_t *acc, int64_t *fail,
                     double *last_dH, int32_t *rng_ix, int32_t *rng_iy) {
    MH_CHECK(h, MH_EINVAL, "mh_hmc_get_state: h is NULL");
    MH_HIP(hipSetDevice(h->ctx->device));
    const size_t R = h->nrep;
    if (N) MH_HIP(hipMemcpy(N, h->N.p, 4 * R, hipMemcpyDeviceToHost));
    if (energy) MH_HIP(hipMemcpy(energy, h->energy.p, 8 * R, hipMemcpyDeviceToHost));
    if (att) MH_HIP(hipMemcpy(att, h->att.p, 8 * R, hipMemcpyDeviceToHost));
    if (acc) MH_HIP(hipMemcpy(acc, h->acc.p, 8 * R, hipMemcpyDeviceToHost));
    if (fail) MH_HIP(hipMemcpy(fail, h->fail.p, 8 * R, hipMemcpyDeviceToHost));
    if (last_dH) MH_HIP(hipMemcpy(last_dH, h->dH.p, 8 * R, hipMemcpyDeviceToHost));
    if (rng_ix) MH_HIP(hipMemcpy(rng_ix, h->ix.p, 4 * R, hipMemcpyDeviceToHost));
    if (rng_iy) MH_HIP(hipMemcpy(rng_iy, h->iy.p, 4 * R, hipMemcpyDeviceToHost));
    return MH_OK;
}

int mh_hmc_get_config(mh_hmc *h, int rep, int *n, double *rp, double *r) {
    MH_CHECK(h && rep >= 0 && rep < h->nrep && n, MH_EINVAL, "mh_hmc_get_config: bad arguments");
    MH_HIP(hipSetDevice(h->ctx->device));
    int hn = 0;
    MH_HIP(hipMemcpy(&hn, h->N.as<int>() + rep, 4, hipMemcpyDeviceToHost));
    *n = hn;
    const size_t nc = (size_t)hn * h->ns * 3, off = (size_t)rep * h->maxN * h->ns * 3;
    if (nc && rp) MH_HIP(hipMemcpy(rp, h->rp.as<double>() + off, 8 * nc, hipMemcpyDeviceToHost));
    if (nc && r) MH_HIP(hipMemcpy(r, h->r.as<double>() + off, 8 * nc, hipMemcpyDeviceToHost));
    return MH_OK;
}

int mh_hmc_set_velocities(mh_hmc *h, int rep, int n, const double *v) {
    MH_CHECK(h && rep >= 0 && rep < h->nrep && n >= 0 && (n == 0 || v), MH_EINVAL,
             "mh_hmc_set_velocities: bad arguments");
    MH_HIP(hipSetDevice(h->ctx->device));
    int hn = 0;
    MH_HIP(hipMemcpy(&hn, h->N.as<int>() + rep, 4, hipMemcpyDeviceToHost));
    MH_CHECK(n == hn, MH_EINVAL, "mh_hmc_set_velocities: n differs from the replica's molecule count");
    const size_t nc = (size_t)n * h->ns * 3, off = (size_t)rep * h->maxN * h->ns * 3;
    if (nc) MH_HIP(hipMemcpy(h->v.as<double>() + off, v, 8 * nc, hipMemcpyHostToDevice));
    return MH_OK;
}

int mh_hmc_get_velocities(mh_hmc *h, int rep, int *n, double *v) {
    MH_CHECK(h && rep >= 0 && rep < h->nrep && n, MH_EINVAL, "mh_hmc_get_velocities: bad arguments");
    MH_HIP(hipSetDevice(h->ctx->device));
    int hn = 0;
    MH_HIP(hipMemcpy(&hn, h->N.as<int>() + rep, 4, hipMemcpyDeviceToHost));
    *n = hn;
    const size_t nc = (size_t)hn * h->ns * 3, off = (size_t)rep * h->maxN * h->ns * 3;
    if (nc && v) MH_HIP(hipMemcpy(v, h->v.as<double>() + off, 8 * nc, hipMemcpyDeviceToHost));
    return MH_OK;
}

int mh_hmc_md(mh_hmc *h, int nsteps, double *trace) {
    MH_CHECK(h && nsteps >= 0 && (nsteps == 0 || trace), MH_EINVAL, "mh_hmc_md: bad arguments");
    if (!nsteps) return MH_OK;
    MH_HIP(hipSetDevice(h->ctx->device));
    const size_t lds = h->lds + (size_t)h->maxN * h->ns * 3 * sizeof(double);
    int dev_max = 65536;
    (void)hipDeviceGetAttribute(&dev_max, hipDeviceAttributeMaxSharedMemoryPerBlock, h->ctx->device);
    MH_CHECK((long long)lds + (long long)sizeof(HDev) + 8192 <= dev_max, MH_ECAPACITY,
             "mh_hmc_md: maxN * nsites too large for LDS");
    const size_t nt = (size_t)h->nrep * nsteps * MH_MD_NQ;
    DBuf dtrace;
    int rc;
    if ((rc = dtrace.alloc(8 * nt))) return rc;
    MH_HIP(hipMemsetAsync(dtrace.p, 0, 8 * nt, h->ctx->stream));
    KTimer t(h->ctx, h->prec == MH_PREC_FP32 ? "md32" : "md64");
    if (h->prec == MH_PREC_FP32) launch_md<MH_PREC_FP32>(h, nsteps, dtrace.as<double>(), lds);
    else launch_md<MH_PREC_FP64>(h, nsteps, dtrace.as<double>(), lds);
    MH_HIP(hipGetLastError());
    MH_HIP(hipStreamSynchronize(h->ctx->stream));
    t.stop();
    MH_HIP(hipMemcpy(trace, dtrace.p, 8 * nt, hipMemcpyDeviceToHost));
    std::vector<int> err(h->nrep);
    MH_HIP(hipMemcpy(err.data(), h->err.p, 4 * h->nrep, hipMemcpyDeviceToHost));
    for (int i = 0; i < h->nrep; i++)
        MH_CHECK(err[i] == 0, MH_ESTATE, "mh_hmc_md: a replica's energy evaluation failed (overlap or off-map site)");
    return MH_OK;
}

int mh_hmc_energy_forces(mh_hmc *h, int rep, double *e, double *F, int *ok) {
    MH_CHECK(h && rep >= 0 && rep < h->nrep && e && ok, MH_EINVAL, "mh_hmc_energy_forces: bad arguments");
    MH_HIP(hipSetDevice(h->ctx->device));
    DBuf dF, de;
    int rc;
    const size_t nc = (size_t)h->maxN * h->ns * 3;
    if ((rc = dF.alloc(8 * (nc + 1))) || (rc = de.alloc(16))) return rc;
    HRep R = h->rep();
    const HDev *P = h->dP.as<HDev>();
    auto go = [&](auto kern) {
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)h->lds);
        hipLaunchKernelGGL(kern, dim3(1), dim3(256), h->lds, h->ctx->stream, P, R, rep, dF.as<double>(),
                           de.as<double>());
    };
    if (h->prec == MH_PREC_FP32) go(k_hmc_forces<256, MH_PREC_FP32>);
    else go(k_hmc_forces<256, MH_PREC_FP64>);
    MH_HIP(hipGetLastError());
    MH_HIP(hipStreamSynchronize(h->ctx->stream));
    double hv[2];
    MH_HIP(hipMemcpy(hv, de.p, 16, hipMemcpyDeviceToHost));
    *e = hv[0];
    *ok = hv[1] != 0.0;
    int hn = 0;
    MH_HIP(hipMemcpy(&hn, h->N.as<int>() + rep, 4, hipMemcpyDeviceToHost));
    if (F && hn) MH_HIP(hipMemcpy(F, dF.p, 8 * (size_t)hn * h->ns * 3, hipMemcpyDeviceToHost));
    return MH_OK;
}

}  // extern "C"
