// libmusic_hip: device-resident GCMC of rigid sorbates (K9), Ewald
// structure-factor deltas (K8), map interpolation (K6) and sorbate-sorbate
// pair energies (K7) inside one persistent kernel per mh_gcmc_run call.
//
// One workgroup per replica (state point).  Every lane keeps an identical
// copy of the replica's RNG state, N, running energy and counters, so the
// reference's serial draw order (gcmc.F90:240-250, gcmcmoves.F90:270-364,
// mcmoves.F90:290-647, rigidmoves.F90:691-1570) is reproduced without
// communication; the energy of the trial is computed by all lanes in
// parallel (maps: one lane per site; pairs: lanes over the other molecules;
// Ewald: lanes over k vectors) and reduced in LDS in a fixed order.
//
// Layout in HBM (per replica r, stride by replica):
//   com[r][maxN][3], ang[r][maxN][3], pos[r][maxN][ns][3]  (FP64)
//   S[r][nk] (double2 running structure factor), dS[r][nk] scratch
//   scalars: N, energy, sumN, nsteps, att[4], acc[4], rng ix/iy/first/iseed
#include <cmath>
#include <cstring>
#include <vector>

#include "mh_device.h"
#include "mh_internal.h"

namespace mh {

constexpr int GS = MH_MAX_SITES;   // max sites
constexpr int GKMAX = 15;          // max kmax for the sorbate Ewald
constexpr double kTwoPiR = 2.0 * 3.1415927;  // defaults.F90:35
constexpr double kTokcalG = 332.064;

struct MapRef {
    const int32_t *ptr;
    const float *vals;
    int n[3];
    double step[3];
    double ell[3];
};

struct GDev {
    int ns;
    double ref[GS][3];
    int st[GS];
    double q[GS];
    int has_p[GS], has_e[GS];
    MapRef pm[GS], em[GS];
    int nst;
    double A[64], B[64], h2[64], l2[64];
    float Af[64], Bf[64], h2f[64], l2f[64];
    int on[64];
    int ewald, kmax, nk;
    double kappa, rc2;
    const double *wk;      // [nk]
    const int8_t *kidx;    // [nk][4] (mx, my, mz, pad)
    double full[3], fund[3];
    int nmoves;
    int mtype[4];
    double cumwt[4];
    double insdel, dtrans, drot, maxnrg, rti, molconst;
    int prec, maxN;
};

struct Rep {
    int *N, *err;
    int32_t *ix, *iy, *first, *iseed;
    double *B, *energy, *sumN;
    long long *nsteps, *att, *acc;
    double *com, *ang, *pos;  // [rep][maxN][3], [rep][maxN][3], [rep][maxN][ns][3]
    double2 *S, *dS;          // [rep][nk]
    int *tr_move, *tr_acc, *tr_N;
    double *tr_dE;
    int tr_cap;
    long long tr_base;        // step index at which the trace starts
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

// rigidcoords_toxyz + matrix_b2feuler + simcell_pbc
__device__ void site_coords(const GDev &P, const double com[3], const double ang[3], double *out) {
    if (P.ns == 1) {
        for (int i = 0; i < 3; i++) out[i] = __dsub_rn(com[i], __dmul_rn(floor(com[i] / P.full[i]), P.full[i]));
        return;
    }
    double sp, cp, st, ct, ss, cs;
    sincos(ang[0], &sp, &cp);
    sincos(ang[1], &st, &ct);
    sincos(ang[2], &ss, &cs);
    double m[3][3];
    m[0][0] = __dsub_rn(__dmul_rn(cs, cp), __dmul_rn(__dmul_rn(ct, sp), ss));
    m[0][1] = __dsub_rn(__dmul_rn(-ss, cp), __dmul_rn(__dmul_rn(ct, sp), cs));
    m[0][2] = __dmul_rn(st, sp);
    m[1][0] = __dadd_rn(__dmul_rn(cs, sp), __dmul_rn(__dmul_rn(ct, cp), ss));
    m[1][1] = __dadd_rn(__dmul_rn(-ss, sp), __dmul_rn(__dmul_rn(ct, cp), cs));
    m[1][2] = __dmul_rn(-st, cp);
    m[2][0] = __dmul_rn(ss, st);
    m[2][1] = __dmul_rn(cs, st);
    m[2][2] = ct;
    for (int a = 0; a < P.ns; a++)
        for (int i = 0; i < 3; i++) {
            double d = 0.0;
            d = __dadd_rn(d, __dmul_rn(m[i][0], P.ref[a][0]));
            d = __dadd_rn(d, __dmul_rn(m[i][1], P.ref[a][1]));
            d = __dadd_rn(d, __dmul_rn(m[i][2], P.ref[a][2]));
            double rp = __dadd_rn(com[i], d);
            out[3 * a + i] = __dsub_rn(rp, __dmul_rn(floor(rp / P.full[i]), P.full[i]));
        }
}

// map energy of one site (maps_int -> interpolate_int); returns false when
// a corner is untabulated
__device__ bool site_map_energy(const GDev &P, int a, const double *r, double &e) {
    e = 0.0;
    double p;
    if (P.has_p[a]) {
        const MapRef &m = P.pm[a];
        if (!hermite_point(m.ptr, m.vals, m.n[0], m.n[1], m.n[2], m.step[0], m.step[1], m.step[2],
                           m.ell[0], m.ell[1], m.ell[2], r[0], r[1], r[2], p))
            return false;
        e = __dadd_rn(e, p);
    }
    if (P.has_e[a]) {
        const MapRef &m = P.em[a];
        if (!hermite_point(m.ptr, m.vals, m.n[0], m.n[1], m.n[2], m.step[0], m.step[1], m.step[2],
                           m.ell[0], m.ell[1], m.ell[2], r[0], r[1], r[2], p))
            return false;
        e = __dadd_rn(e, __dmul_rn(p, P.q[a]));
    }
    return true;
}

// pair energy of a site pair (lj_snglint without derivatives + Ewald real
// space).  Returns false on an LJ locut violation.
template <int PREC>
__device__ __forceinline__ bool pair_energy(const GDev &P, int ta, int tb, double qa, double qb,
                                            double sx, double sy, double sz, double &e) {
    const int t = ta * P.nst + tb;
    if (PREC == MH_PREC_FP64) {
        double r2 = __dadd_rn(__dadd_rn(__dmul_rn(sx, sx), __dmul_rn(sy, sy)), __dmul_rn(sz, sz));
        if (P.on[t] && !(r2 > P.h2[t])) {
            if (r2 < P.l2[t]) return false;
            double r2i = 1.0 / r2;
            double r6i = __dmul_rn(__dmul_rn(r2i, r2i), r2i);
            double r12i = __dmul_rn(r6i, r6i);
            e = __dadd_rn(e, __dsub_rn(__dmul_rn(P.A[t], r12i), __dmul_rn(P.B[t], r6i)));
        }
        if (P.ewald && r2 < P.rc2 && qa != 0.0 && qb != 0.0) {
            double rr = sqrt(r2);
            e = __dadd_rn(e, __dmul_rn(__ddiv_rn(__dmul_rn(__dmul_rn(qa, qb), erfc_as64(P.kappa * rr)), rr), kTokcalG));
        }
    } else {
#pragma clang fp contract(fast)
        float fx = (float)sx, fy = (float)sy, fz = (float)sz;
        float r2 = fx * fx + fy * fy + fz * fz;
        if (P.on[t] && !(r2 > P.h2f[t])) {
            if (r2 < P.l2f[t]) return false;
            float r2i = __frcp_rn(r2);
            float r6i = r2i * r2i * r2i;
            e += (double)(P.Af[t] * (r6i * r6i) - P.Bf[t] * r6i);
        }
        if (P.ewald && r2 < (float)P.rc2 && qa != 0.0 && qb != 0.0) {
            float ri = __frsqrt_rn(r2);
            float rr = r2 * ri;
            e += (double)((float)(qa * qb * kTokcalG) * erfc_as32((float)P.kappa * rr) * ri);
        }
    }
    return true;
}

// e^{i m 2pi x / L} for m = 0..kmax; the k vector of index (mx,my,mz) uses
// E_x[mx] E_y[|my|] E_z[|mz|] with conjugates for negative m
template <int PREC>
__device__ __forceinline__ double2 expk(const double2 *Ex, const double2 *Ey, const double2 *Ez, int mx,
                                        int my, int mz) {
    double2 a = Ex[mx];
    double2 b = Ey[my < 0 ? -my : my];
    if (my < 0) b.y = -b.y;
    double2 c = Ez[mz < 0 ? -mz : mz];
    if (mz < 0) c.y = -c.y;
    double2 ab = make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
    return make_double2(ab.x * c.x - ab.y * c.y, ab.x * c.y + ab.y * c.x);
}

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

struct Shared {
    double posA[GS * 3], posB[GS * 3];   // trial (A) and old (B) site coordinates
    double comA[3], angA[3];
    double mapv[2][GS];
    int mapok[2][GS];
    double2 E[2][GS][3][GKMAX + 1];      // per position set, site, axis: e^{i m theta}
    double red[16 * 6];
};

// Energy terms of a move: set A (+) and set B (-) of one molecule against
// every other molecule (skip = its slot), reciprocal delta with dS written to
// dS[k].  modeA/modeB: 1 when the set is present.  rsign: +1 add A / -1 remove A.
template <int NT, int PREC>
__device__ void move_energy(const GDev &P, Shared &sh, const double *__restrict__ pos, int N, int skip,
                            bool hasA, bool hasB, double asign, const double2 *__restrict__ S,
                            double2 *__restrict__ dS, double out[6]) {
    const int tid = threadIdx.x;
    const int ns = P.ns;
    // maps (one lane per site and set)
    if (tid < 2 * GS) {
        int set = tid / GS, a = tid % GS;
        if (a < ns && ((set == 0 && hasA) || (set == 1 && hasB))) {
            double e;
            bool ok = site_map_energy(P, a, set == 0 ? &sh.posA[3 * a] : &sh.posB[3 * a], e);
            sh.mapv[set][a] = e;
            sh.mapok[set][a] = ok ? 1 : 0;
        }
    }
    // e^{i m theta} tables for the Ewald delta
    if (P.ewald) {
        const int ntab = 2 * GS * 3 * (P.kmax + 1);
        for (int i = tid; i < ntab; i += NT) {
            int m = i % (P.kmax + 1);
            int ax = (i / (P.kmax + 1)) % 3;
            int a = (i / (3 * (P.kmax + 1))) % GS;
            int set = i / (GS * 3 * (P.kmax + 1));
            if (a >= ns) continue;
            const double *pp = set == 0 ? sh.posA : sh.posB;
            double th = ((double)m / P.full[ax]) * kTwoPiR * pp[3 * a + ax];
            double s_, c_;
            if (PREC == MH_PREC_FP64) sincos(th, &s_, &c_);
            else {
                float sf, cf;
                __sincosf((float)th, &sf, &cf);
                s_ = sf;
                c_ = cf;
            }
            sh.E[set][a][ax][m] = make_double2(c_, s_);
        }
    }
    __syncthreads();
    // pairs: lanes over the other molecules
    double eA = 0.0, eB = 0.0, failA = 0.0;
    for (int j = tid; j < N; j += NT) {
        if (j == skip) continue;
        const double *pj = pos + (long long)j * ns * 3;
        for (int b = 0; b < ns; b++) {
            const double xb = pj[3 * b], yb = pj[3 * b + 1], zb = pj[3 * b + 2];
            const int tb = P.st[b];
            const double qb = P.q[b];
            for (int a = 0; a < ns; a++) {
                if (hasA) {
                    double sx = fold64(sh.posA[3 * a] - xb, P.full[0]);
                    double sy = fold64(sh.posA[3 * a + 1] - yb, P.full[1]);
                    double sz = fold64(sh.posA[3 * a + 2] - zb, P.full[2]);
                    if (!pair_energy<PREC>(P, P.st[a], tb, P.q[a], qb, sx, sy, sz, eA)) failA = 1.0;
                }
                if (hasB) {
                    double sx = fold64(sh.posB[3 * a] - xb, P.full[0]);
                    double sy = fold64(sh.posB[3 * a + 1] - yb, P.full[1]);
                    double sz = fold64(sh.posB[3 * a + 2] - zb, P.full[2]);
                    pair_energy<PREC>(P, P.st[a], tb, P.q[a], qb, sx, sy, sz, eB);
                }
            }
        }
    }
    // reciprocal delta: lanes over k
    double eK = 0.0;
    if (P.ewald) {
        for (int k = tid; k < P.nk; k += NT) {
            const int8_t *ki = P.kidx + 4 * k;
            double2 d = make_double2(0.0, 0.0);
            for (int a = 0; a < ns; a++) {
                if (P.q[a] == 0.0) continue;
                if (hasA) {
                    double2 e = expk<PREC>(sh.E[0][a][0], sh.E[0][a][1], sh.E[0][a][2], ki[0], ki[1], ki[2]);
                    d.x += asign * P.q[a] * e.x;
                    d.y += asign * P.q[a] * e.y;
                }
                if (hasB) {
                    double2 e = expk<PREC>(sh.E[1][a][0], sh.E[1][a][1], sh.E[1][a][2], ki[0], ki[1], ki[2]);
                    d.x -= P.q[a] * e.x;
                    d.y -= P.q[a] * e.y;
                }
            }
            double2 s = S[k];
            double nre = s.x + d.x, nim = s.y + d.y;
            eK += P.wk[k] * ((nre * nre + nim * nim) - (s.x * s.x + s.y * s.y));
            dS[k] = d;
        }
    }
    double v[4] = {eA, eB, eK, failA};
    block_sum<NT, 4>(v, sh.red);
    double mA = 0.0, mB = 0.0;
    int okA = 1;
    for (int a = 0; a < ns; a++) {
        if (hasA) {
            mA = __dadd_rn(mA, sh.mapv[0][a]);
            okA &= sh.mapok[0][a];
        }
        if (hasB) mB = __dadd_rn(mB, sh.mapv[1][a]);
    }
    out[0] = mA;
    out[1] = v[0];
    out[2] = mB;
    out[3] = v[1];
    out[4] = v[2];
    out[5] = (okA && v[3] == 0.0) ? 1.0 : 0.0;
}

template <int NT, int PREC>
__global__ void __launch_bounds__(NT) k_gcmc(const GDev *__restrict__ Pg, Rep R, int nsteps) {
    __shared__ Shared sh;
    __shared__ GDev P;
    const int rep = blockIdx.x;
    const int tid = threadIdx.x;
    for (int i = tid; i < (int)(sizeof(GDev) / 4); i += NT)
        reinterpret_cast<int *>(&P)[i] = reinterpret_cast<const int *>(Pg)[i];
    __syncthreads();
    if (R.err[rep]) return;
    const int ns = P.ns;
    Rng rng{R.ix[rep], R.iy[rep], R.first[rep], R.iseed[rep]};
    int N = R.N[rep];
    double energy = R.energy[rep], sumN = R.sumN[rep];
    long long nst = R.nsteps[rep];
    long long att[4], acc[4];
    for (int i = 0; i < 4; i++) {
        att[i] = R.att[4 * rep + i];
        acc[i] = R.acc[4 * rep + i];
    }
    const double B = R.B[rep];
    double *com = R.com + (long long)rep * P.maxN * 3;
    double *ang = R.ang + (long long)rep * P.maxN * 3;
    double *pos = R.pos + (long long)rep * P.maxN * ns * 3;
    double2 *S = R.S + (long long)rep * P.nk;
    double2 *dS = R.dS + (long long)rep * P.nk;
    int err = 0;
    for (int it = 0; it < nsteps; it++) {
        (void)rranf(rng);  // species (nspc = 1), gcmc.F90:242
        const double prob = rranf(rng);
        int mv = 0;
        for (int i = 0; i < P.nmoves; i++) {
            if (prob > P.cumwt[i]) mv++;
            else break;
        }
        if (mv >= P.nmoves) mv = P.nmoves - 1;
        const int mtype = P.mtype[mv];
        int accepted = 0;
        double de = 0.0;
        bool doit = true;
        int mol = -1;
        // ---- trial generation (identical in every lane)
        double tcom[3], tang[3] = {0.0, 0.0, 0.0};
        if (mtype == MH_MV_INSERT) {
            att[0]++;
            if (N >= P.maxN) {
                err = 2;  // capacity
                break;
            }
            for (int i = 0; i < 3; i++) tcom[i] = __dmul_rn(P.full[i], rranf(rng));
            if (ns != 1) {
                double costheta = __dsub_rn(1.0, __dmul_rn(rranf(rng), 2.0));
                tang[1] = acos(costheta);
                tang[0] = __dsub_rn(3.1415927, __dmul_rn(rranf(rng), kTwoPiR));
                tang[2] = __dsub_rn(3.1415927, __dmul_rn(rranf(rng), kTwoPiR));
            }
        } else {
            if (N == 0) {
                doit = false;
            } else {
                mol = (int)(rranf(rng) * N);
                if (mtype == MH_MV_DELETE) att[1]++;
                else att[mtype]++;
                for (int i = 0; i < 3; i++) {
                    tcom[i] = com[3 * mol + i];
                    tang[i] = ang[3 * mol + i];
                }
                if (mtype == MH_MV_TRANSLATE) {
                    for (int i = 0; i < 3; i++) {
                        double dv = __dmul_rn(__dsub_rn(__dmul_rn(rranf(rng), 2.0), 1.0), P.dtrans);
                        tcom[i] = __dadd_rn(tcom[i], dv);
                    }
                } else if (mtype == MH_MV_ROTATE) {
                    double dphi = __dmul_rn(__dsub_rn(__dmul_rn(rranf(rng), 2.0), 1.0), P.drot);
                    double dpsi = __dmul_rn(__dsub_rn(__dmul_rn(rranf(rng), 2.0), 1.0), P.drot);
                    double dcos = __dmul_rn(__dsub_rn(__dmul_rn(rranf(rng), 2.0), 1.0), P.drot);
                    if (ns != 1) {  // rigidmoves_rotperturb (rigidmoves.F90:1690-1707)
                        double phi = __dadd_rn(tang[0], dphi);
                        double psi = __dadd_rn(tang[2], dpsi);
                        double costheta = __dadd_rn(cos(tang[1]), dcos);
                        tang[0] = __dsub_rn(phi, __dmul_rn(round(phi / kTwoPiR), kTwoPiR));
                        tang[2] = __dsub_rn(psi, __dmul_rn(round(psi / kTwoPiR), kTwoPiR));
                        costheta = __dsub_rn(costheta, __dmul_rn(round(costheta / 2.0), 2.0));
                        tang[1] = acos(costheta);
                    }
                }
            }
        }
        if (doit) {
            const bool hasA = (mtype != MH_MV_DELETE);
            const bool hasB = (mtype != MH_MV_INSERT);
            if (tid == 0) {
                if (hasA) site_coords(P, tcom, tang, sh.posA);
                if (hasB)
                    for (int i = 0; i < 3 * ns; i++) sh.posB[i] = pos[(long long)mol * ns * 3 + i];
            }
            __syncthreads();
            double o[6];
            if (mtype == MH_MV_DELETE) {
                // the molecule's energy in the current system: its sites as set A
                if (tid == 0)
                    for (int i = 0; i < 3 * ns; i++) sh.posA[i] = sh.posB[i];
                __syncthreads();
                move_energy<NT, PREC>(P, sh, pos, N, mol, true, false, -1.0, S, dS, o);
            } else {
                move_energy<NT, PREC>(P, sh, pos, N, mol, hasA, hasB, 1.0, S, dS, o);
            }
            const double mA = o[0], pA = o[1], mB = o[2], pB = o[3], eK = o[4];
            bool ok = o[5] != 0.0;
            if (mtype == MH_MV_INSERT) {
                double U = __dadd_rn(mA, pA);
                if (P.ewald) U = U + eK + P.molconst;
                if (ok && U > P.maxnrg) ok = false;
                if (ok) {
                    double bias = 1.0 / P.insdel;
                    double u = -P.rti * U;
                    u = u + log(bias);
                    u = u + B;
                    u = fmin(fmax(u, -30.0), 30.0);
                    double boltz = exp(u) * (1.0 / (double)(float)(N + 1));
                    accepted = boltz > rranf(rng);
                    if (accepted) {
                        if (tid == 0) {
                            for (int i = 0; i < 3; i++) {
                                com[3 * N + i] = tcom[i];
                                ang[3 * N + i] = tang[i];
                            }
                            for (int i = 0; i < 3 * ns; i++) pos[(long long)N * ns * 3 + i] = sh.posA[i];
                        }
                        if (P.ewald)
                            for (int k = tid; k < P.nk; k += NT) {
                                double2 s = S[k], d = dS[k];
                                S[k] = make_double2(s.x + d.x, s.y + d.y);
                            }
                        N++;
                        energy += U;
                        acc[0]++;
                        de = U;
                    }
                }
            } else if (mtype == MH_MV_DELETE) {
                if (!ok) {
                    err = 3;  // an existing molecule failed to evaluate (mcmoves.F90:599-603)
                    break;
                }
                double U = __dadd_rn(mA, pA);
                if (P.ewald) U = U - eK + P.molconst;
                double bias = 1.0 * P.insdel;
                double u = -P.rti * (-U);
                u = u + log(bias);
                u = u + (-B);
                u = fmin(fmax(u, -30.0), 30.0);
                double boltz = exp(u) * (1.0 * N);
                accepted = boltz > rranf(rng);
                if (accepted) {
                    const int last = N - 1;
                    __syncthreads();  // all lanes done reading pos before the move
                    if (tid == 0) {
                        for (int i = 0; i < 3; i++) {
                            com[3 * mol + i] = com[3 * last + i];
                            ang[3 * mol + i] = ang[3 * last + i];
                        }
                        for (int i = 0; i < 3 * ns; i++)
                            pos[(long long)mol * ns * 3 + i] = pos[(long long)last * ns * 3 + i];
                    }
                    if (P.ewald)
                        for (int k = tid; k < P.nk; k += NT) {
                            double2 s = S[k], d = dS[k];
                            S[k] = make_double2(s.x + d.x, s.y + d.y);
                        }
                    N--;
                    energy -= U;
                    acc[1]++;
                    de = -U;
                }
            } else {
                double U1 = __dadd_rn(mA, pA), U0 = __dadd_rn(mB, pB);
                if (ok && U1 > P.maxnrg) ok = false;
                double dU = __dsub_rn(U1, U0);
                if (P.ewald) dU = dU + eK;
                if (ok) {
                    double u = -P.rti * dU;
                    u = u + log(1.0);
                    u = fmin(fmax(u, -30.0), 30.0);
                    double boltz = exp(u);
                    accepted = boltz > rranf(rng);
                    if (accepted) {
                        if (tid == 0) {
                            for (int i = 0; i < 3; i++) {
                                com[3 * mol + i] = tcom[i];
                                ang[3 * mol + i] = tang[i];
                            }
                            for (int i = 0; i < 3 * ns; i++) pos[(long long)mol * ns * 3 + i] = sh.posA[i];
                        }
                        if (P.ewald)
                            for (int k = tid; k < P.nk; k += NT) {
                                double2 s = S[k], d = dS[k];
                                S[k] = make_double2(s.x + d.x, s.y + d.y);
                            }
                        energy += dU;
                        acc[mtype]++;
                        de = dU;
                    }
                }
            }
        }
        if (tid == 0 && R.tr_cap > 0) {
            long long t = nst - R.tr_base;
            if (t >= 0 && t < R.tr_cap) {
                long long o = (long long)rep * R.tr_cap + t;
                R.tr_move[o] = mtype;
                R.tr_acc[o] = accepted;
                R.tr_N[o] = N;
                R.tr_dE[o] = de;
            }
        }
        sumN += N;
        nst++;
        __syncthreads();
    }
    if (tid == 0) {
        R.err[rep] = err;
        R.ix[rep] = rng.ix;
        R.iy[rep] = rng.iy;
        R.first[rep] = rng.first;
        R.iseed[rep] = rng.iseed;
        R.N[rep] = N;
        R.energy[rep] = energy;
        R.sumN[rep] = sumN;
        R.nsteps[rep] = nst;
        for (int i = 0; i < 4; i++) {
            R.att[4 * rep + i] = att[i];
            R.acc[4 * rep + i] = acc[i];
        }
    }
}

// Insertion energies of trial molecules (one workgroup per trial)
template <int NT, int PREC>
__global__ void __launch_bounds__(NT) k_trial(const GDev *__restrict__ Pg, Rep R, int rep, int n,
                                               const double *__restrict__ tcom, const double *__restrict__ tang,
                                               double2 *__restrict__ dSbuf, double *__restrict__ e,
                                               uint8_t *__restrict__ okout) {
    __shared__ Shared sh;
    __shared__ GDev P;
    const int tid = threadIdx.x;
    for (int i = tid; i < (int)(sizeof(GDev) / 4); i += NT)
        reinterpret_cast<int *>(&P)[i] = reinterpret_cast<const int *>(Pg)[i];
    __syncthreads();
    const int t = blockIdx.x;
    if (t >= n) return;
    const int N = R.N[rep];
    const double *pos = R.pos + (long long)rep * P.maxN * P.ns * 3;
    const double2 *S = R.S + (long long)rep * P.nk;
    if (tid == 0) site_coords(P, tcom + 3 * t, tang + 3 * t, sh.posA);
    __syncthreads();
    double o[6];
    move_energy<NT, PREC>(P, sh, pos, N, -1, true, false, 1.0, S, dSbuf + (long long)t * P.nk, o);
    if (tid == 0) {
        double U = __dadd_rn(o[0], o[1]);
        if (P.ewald) U = U + o[4] + P.molconst;
        bool ok = o[5] != 0.0 && !(U > P.maxnrg);
        e[t] = U;
        okout[t] = ok ? 1 : 0;
    }
}

// S(k) of every replica from the coordinates (FP64), and total energies
__global__ void k_sk_recompute(const GDev *__restrict__ Pg, Rep R) {
    const GDev &P = *Pg;
    const int rep = blockIdx.y;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= P.nk) return;
    const int N = R.N[rep];
    const double *pos = R.pos + (long long)rep * P.maxN * P.ns * 3;
    const int8_t *ki = P.kidx + 4 * k;
    double kv[3] = {((double)ki[0] / P.full[0]) * kTwoPiR, ((double)ki[1] / P.full[1]) * kTwoPiR,
                    ((double)ki[2] / P.full[2]) * kTwoPiR};
    double re = 0.0, im = 0.0;
    for (int m = 0; m < N; m++)
        for (int a = 0; a < P.ns; a++) {
            if (P.q[a] == 0.0) continue;
            const double *r = pos + ((long long)m * P.ns + a) * 3;
            double arg = kv[0] * r[0] + kv[1] * r[1] + kv[2] * r[2];
            double s_, c_;
            sincos(arg, &s_, &c_);
            re += P.q[a] * c_;
            im += P.q[a] * s_;
        }
    R.S[(long long)rep * P.nk + k] = make_double2(re, im);
}

// total energy per replica: sum over molecules m of (maps + pairs with j < m)
// + reciprocal + N * molconst; one workgroup per replica
template <int NT>
__global__ void __launch_bounds__(NT) k_total(const GDev *__restrict__ Pg, Rep R, double *__restrict__ out) {
    __shared__ GDev P;
    __shared__ double red[NT / 64 * 2];
    const int tid = threadIdx.x;
    for (int i = tid; i < (int)(sizeof(GDev) / 4); i += NT)
        reinterpret_cast<int *>(&P)[i] = reinterpret_cast<const int *>(Pg)[i];
    __syncthreads();
    const int rep = blockIdx.x;
    const int N = R.N[rep], ns = P.ns;
    const double *pos = R.pos + (long long)rep * P.maxN * ns * 3;
    double e = 0.0, bad = 0.0;
    for (int m = tid; m < N; m += NT) {
        const double *pm = pos + (long long)m * ns * 3;
        for (int a = 0; a < ns; a++) {
            double em;
            if (!site_map_energy(P, a, pm + 3 * a, em)) bad = 1.0;
            e += em;
        }
        for (int j = 0; j < m; j++) {
            const double *pj = pos + (long long)j * ns * 3;
            for (int a = 0; a < ns; a++)
                for (int b = 0; b < ns; b++) {
                    double sx = fold64(pm[3 * a] - pj[3 * b], P.full[0]);
                    double sy = fold64(pm[3 * a + 1] - pj[3 * b + 1], P.full[1]);
                    double sz = fold64(pm[3 * a + 2] - pj[3 * b + 2], P.full[2]);
                    if (!pair_energy<MH_PREC_FP64>(P, P.st[a], P.st[b], P.q[a], P.q[b], sx, sy, sz, e)) bad = 1.0;
                }
        }
    }
    if (P.ewald)
        for (int k = tid; k < P.nk; k += NT) {
            double2 s = R.S[(long long)rep * P.nk + k];
            e += P.wk[k] * (s.x * s.x + s.y * s.y);
        }
    double v[2] = {e, bad};
    block_sum<NT, 2>(v, red);
    if (tid == 0) out[rep] = (v[1] != 0.0) ? NAN : v[0] + (P.ewald ? N * P.molconst : 0.0);
}

}  // namespace mh

using namespace mh;

struct mh_gcmc {
    mh_ctx *ctx = nullptr;
    int nrep = 0, maxN = 0, ns = 0, nk = 0, prec = 0, threads = 256;
    GDev host;
    DBuf dP, wk, kidx;
    DBuf N, err, ix, iy, first, iseed, B, energy, sumN, nsteps, att, acc;
    DBuf com, ang, pos, S, dS;
    DBuf tr_move, tr_acc, tr_N, tr_dE;
    int tr_cap = 0;
    long long tr_base = 0;
    Rep rep() {
        Rep r;
        r.N = N.as<int>();
        r.err = err.as<int>();
        r.ix = ix.as<int32_t>();
        r.iy = iy.as<int32_t>();
        r.first = first.as<int32_t>();
        r.iseed = iseed.as<int32_t>();
        r.B = B.as<double>();
        r.energy = energy.as<double>();
        r.sumN = sumN.as<double>();
        r.nsteps = nsteps.as<long long>();
        r.att = att.as<long long>();
        r.acc = acc.as<long long>();
        r.com = com.as<double>();
        r.ang = ang.as<double>();
        r.pos = pos.as<double>();
        r.S = S.as<double2>();
        r.dS = dS.as<double2>();
        r.tr_move = tr_move.as<int>();
        r.tr_acc = tr_acc.as<int>();
        r.tr_N = tr_N.as<int>();
        r.tr_dE = tr_dE.as<double>();
        r.tr_cap = tr_cap;
        r.tr_base = tr_base;
        return r;
    }
};

static double erfc_as_host(double x) {
    const double a1 = (double)0.254829592f, a2 = (double)-0.284496736f, a3 = (double)1.421413741f,
                 a4 = (double)-1.453152027f, a5 = (double)1.061405429f, p = (double)0.3275911f;
    double t = 1.0 / (1.0 + p * x);
    return t * (a1 + t * (a2 + t * (a3 + t * (a4 + t * a5)))) * std::exp(-x * x);
}

template <int PREC>
static void launch_gcmc(mh_gcmc *g, int nsteps) {
    Rep r = g->rep();
    const GDev *P = g->dP.as<GDev>();
    dim3 grid(g->nrep);
    switch (g->threads) {
    case 64: hipLaunchKernelGGL((k_gcmc<64, PREC>), grid, dim3(64), 0, g->ctx->stream, P, r, nsteps); break;
    case 128: hipLaunchKernelGGL((k_gcmc<128, PREC>), grid, dim3(128), 0, g->ctx->stream, P, r, nsteps); break;
    case 512: hipLaunchKernelGGL((k_gcmc<512, PREC>), grid, dim3(512), 0, g->ctx->stream, P, r, nsteps); break;
    case 1024: hipLaunchKernelGGL((k_gcmc<1024, PREC>), grid, dim3(1024), 0, g->ctx->stream, P, r, nsteps); break;
    default: hipLaunchKernelGGL((k_gcmc<256, PREC>), grid, dim3(256), 0, g->ctx->stream, P, r, nsteps); break;
    }
}

extern "C" {

int mh_gcmc_create(mh_ctx *ctx, const mh_gcmc_params *p, int nrep, const double *fugacity_kPa,
                   const int32_t *seeds, int maxN, mh_gcmc **out) {
    MH_CHECK(ctx && p && fugacity_kPa && seeds && out, MH_EINVAL, "mh_gcmc_create: bad arguments");
    MH_CHECK(nrep >= 1 && nrep <= 65535, MH_EINVAL, "mh_gcmc_create: nrep out of range");
    MH_CHECK(p->nsites >= 1 && p->nsites <= GS, MH_EINVAL, "mh_gcmc_create: nsites out of range [1,8]");
    MH_CHECK(p->ntypes >= 1 && p->ntypes <= 8, MH_EINVAL, "mh_gcmc_create: ntypes out of range [1,8]");
    MH_CHECK(maxN >= 1 && maxN <= (1 << 20), MH_EINVAL, "mh_gcmc_create: maxN out of range");
    MH_CHECK(p->nmoves >= 1 && p->nmoves <= 4, MH_EINVAL, "mh_gcmc_create: nmoves out of range [1,4]");
    MH_CHECK(p->temperature > 0.0, MH_EINVAL, "mh_gcmc_create: temperature must be > 0");
    MH_CHECK(p->precision == MH_PREC_FP64 || p->precision == MH_PREC_FP32, MH_EINVAL,
             "mh_gcmc_create: bad precision");
    MH_CHECK(ctx->fund_ell[0] > 0.0, MH_ESTATE, "mh_gcmc_create: upload the framework (cell) first");
    if (p->ewald) {
        MH_CHECK(p->ewald_kmax >= 1 && p->ewald_kmax <= GKMAX, MH_EINVAL,
                 "mh_gcmc_create: ewald_kmax out of range [1,15]");
        MH_CHECK(p->ewald_kappa > 0.0, MH_EINVAL, "mh_gcmc_create: ewald_kappa must be > 0");
    }
    MH_HIP(hipSetDevice(ctx->device));
    mh_gcmc *g = new mh_gcmc();
    g->ctx = ctx;
    g->nrep = nrep;
    g->maxN = maxN;
    g->ns = p->nsites;
    g->prec = p->precision;
    GDev &H = g->host;
    std::memset(&H, 0, sizeof H);
    H.ns = p->nsites;
    for (int a = 0; a < H.ns; a++) {
        for (int i = 0; i < 3; i++) H.ref[a][i] = p->ref[a][i];
        H.st[a] = p->site_type[a];
        H.q[a] = p->site_q[a];
        if (p->site_type[a] < 0 || p->site_type[a] >= p->ntypes) {
            delete g;
            set_error("mh_gcmc_create: site_type out of range");
            return MH_EINVAL;
        }
        const mh_map *maps[2] = {p->site_pmap[a], p->site_emap[a]};
        for (int w = 0; w < 2; w++) {
            if (!maps[w]) continue;
            MapRef &m = w == 0 ? H.pm[a] : H.em[a];
            (w == 0 ? H.has_p : H.has_e)[a] = 1;
            m.ptr = maps[w]->ptr.as<int32_t>();
            m.vals = maps[w]->vals.as<float>();
            for (int i = 0; i < 3; i++) {
                m.n[i] = maps[w]->npnts[i];
                m.step[i] = maps[w]->step[i];
                m.ell[i] = ctx->fund_ell[i];
            }
        }
    }
    H.nst = p->ntypes;
    for (int t = 0; t < p->ntypes * p->ntypes; t++) {
        H.A[t] = p->lj_A[t];
        H.B[t] = p->lj_B[t];
        H.h2[t] = p->lj_hicut2[t];
        H.l2[t] = p->lj_locut2[t];
        H.Af[t] = (float)p->lj_A[t];
        H.Bf[t] = (float)p->lj_B[t];
        H.h2f[t] = (float)p->lj_hicut2[t];
        H.l2f[t] = (float)p->lj_locut2[t];
        H.on[t] = p->lj_on[t] ? 1 : 0;
    }
    for (int i = 0; i < 3; i++) {
        H.full[i] = ctx->full_ell[i];
        H.fund[i] = ctx->fund_ell[i];
    }
    H.ewald = p->ewald ? 1 : 0;
    H.kmax = p->ewald ? p->ewald_kmax : 0;
    H.kappa = p->ewald_kappa;
    H.rc2 = p->ewald_rcut * p->ewald_rcut;
    // k vectors of the half space (mx > 0, or mx = 0 & my > 0, or mx = my = 0 & mz > 0)
    std::vector<double> w;
    std::vector<int8_t> ki;
    const double V = H.full[0] * H.full[1] * H.full[2];
    if (H.ewald) {
        int K = H.kmax;
        for (int mx = 0; mx <= K; mx++)
            for (int my = -K; my <= K; my++)
                for (int mz = -K; mz <= K; mz++) {
                    if (mx == 0 && (my < 0 || (my == 0 && mz <= 0))) continue;
                    double kx = ((double)mx / H.full[0]) * kTwoPiR, ky = ((double)my / H.full[1]) * kTwoPiR,
                           kz = ((double)mz / H.full[2]) * kTwoPiR;
                    double k2 = kx * kx + ky * ky + kz * kz;
                    w.push_back((4.0 * M_PI / V) * std::exp(-k2 / (4.0 * H.kappa * H.kappa)) / k2 * kTokcalG);
                    ki.push_back((int8_t)mx);
                    ki.push_back((int8_t)my);
                    ki.push_back((int8_t)mz);
                    ki.push_back(0);
                }
        double mc = 0.0;
        for (int a = 0; a < H.ns; a++) mc -= H.kappa / std::sqrt(M_PI) * H.q[a] * H.q[a];
        for (int a = 0; a < H.ns; a++)
            for (int b = a + 1; b < H.ns; b++) {
                double d[3];
                for (int i = 0; i < 3; i++) d[i] = H.ref[a][i] - H.ref[b][i];
                double r = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
                mc -= H.q[a] * H.q[b] * (1.0 - erfc_as_host(H.kappa * r)) / r;
            }
        H.molconst = mc * kTokcalG;
    }
    H.nk = (int)w.size();
    g->nk = H.nk;
    H.nmoves = p->nmoves;
    double wsum = 0.0, wi = 0.0, wd = 0.0;
    for (int i = 0; i < p->nmoves; i++) wsum += p->move_weight[i];
    double c = 0.0;
    for (int i = 0; i < p->nmoves; i++) {
        H.mtype[i] = p->move_type[i];
        double wn = p->move_weight[i] / wsum;
        c = c + wn;
        H.cumwt[i] = c;
        if (p->move_type[i] == MH_MV_INSERT) wi = wn;
        if (p->move_type[i] == MH_MV_DELETE) wd = wn;
    }
    H.insdel = (wi > 0.0 && wd > 0.0) ? wi / wd : 1.0;  // mcmoveset.F90:424-430
    H.dtrans = p->delta_trans;
    H.drot = p->delta_rot;
    H.maxnrg = p->max_nrg;
    H.rti = 1.0 / ((8.314e-3 / 4.184) * p->temperature);
    H.prec = p->precision;
    H.maxN = maxN;
    int rc;
    auto fail = [&](int code) {
        delete g;
        return code;
    };
    if ((rc = g->wk.alloc(sizeof(double) * (H.nk + 1))) || (rc = g->kidx.alloc(4 * (H.nk + 1))))
        return fail(rc);
    if (H.nk) {
        if (hipMemcpy(g->wk.p, w.data(), sizeof(double) * H.nk, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(g->kidx.p, ki.data(), 4 * H.nk, hipMemcpyHostToDevice) != hipSuccess)
            return fail(MH_EHIP);
    }
    H.wk = g->wk.as<double>();
    H.kidx = g->kidx.as<int8_t>();
    if ((rc = g->dP.alloc(sizeof(GDev)))) return fail(rc);
    if (hipMemcpy(g->dP.p, &H, sizeof(GDev), hipMemcpyHostToDevice) != hipSuccess) return fail(MH_EHIP);
    const long long R = nrep;
    if ((rc = g->N.alloc(4 * R)) || (rc = g->err.alloc(4 * R)) || (rc = g->ix.alloc(4 * R)) ||
        (rc = g->iy.alloc(4 * R)) || (rc = g->first.alloc(4 * R)) || (rc = g->iseed.alloc(4 * R)) ||
        (rc = g->B.alloc(8 * R)) || (rc = g->energy.alloc(8 * R)) || (rc = g->sumN.alloc(8 * R)) ||
        (rc = g->nsteps.alloc(8 * R)) || (rc = g->att.alloc(32 * R)) || (rc = g->acc.alloc(32 * R)) ||
        (rc = g->com.alloc(sizeof(double) * 3 * R * maxN)) || (rc = g->ang.alloc(sizeof(double) * 3 * R * maxN)) ||
        (rc = g->pos.alloc(sizeof(double) * 3 * R * maxN * g->ns)) ||
        (rc = g->S.alloc(sizeof(double2) * R * (H.nk + 1))) || (rc = g->dS.alloc(sizeof(double2) * R * (H.nk + 1))) ||
        (rc = g->tr_move.alloc(4)) || (rc = g->tr_acc.alloc(4)) || (rc = g->tr_N.alloc(4)) || (rc = g->tr_dE.alloc(8)))
        return fail(rc);
    std::vector<int32_t> zi(R, 0), hiseed(seeds, seeds + R), ones(R, 1);
    std::vector<double> hB(R), zd(R, 0.0);
    std::vector<long long> zl(4 * R, 0);
    const double sivolume = V * 1.0e-30 * 6.0221367e23;
    for (int i = 0; i < nrep; i++) {
        MH_CHECK(fugacity_kPa[i] > 0.0, MH_EINVAL, "mh_gcmc_create: fugacity must be > 0");
        hB[i] = std::log(fugacity_kPa[i] * 1.0e3 * sivolume / (8.314 * p->temperature)) - std::log(1.0);
    }
    if (hipMemcpy(g->N.p, zi.data(), 4 * R, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(g->err.p, zi.data(), 4 * R, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(g->ix.p, zi.data(), 4 * R, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(g->iy.p, zi.data(), 4 * R, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(g->first.p, ones.data(), 4 * R, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(g->iseed.p, hiseed.data(), 4 * R, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(g->B.p, hB.data(), 8 * R, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(g->energy.p, zd.data(), 8 * R, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(g->sumN.p, zd.data(), 8 * R, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(g->nsteps.p, zl.data(), 8 * R, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(g->att.p, zl.data(), 32 * R, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(g->acc.p, zl.data(), 32 * R, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(g->S.p, 0, sizeof(double2) * R * (H.nk + 1)) != hipSuccess)
        return fail(MH_EHIP);
    ctx_retain(ctx);
    *out = g;
    return MH_OK;
}

int mh_gcmc_destroy(mh_gcmc *g) {
    if (!g) return MH_OK;
    mh_ctx *ctx = g->ctx;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    delete g;
    ctx_release(ctx);
    return MH_OK;
}

int mh_gcmc_set_threads(mh_gcmc *g, int threads) {
    MH_CHECK(g, MH_EINVAL, "mh_gcmc_set_threads: g is NULL");
    MH_CHECK(threads == 64 || threads == 128 || threads == 256 || threads == 512 || threads == 1024,
             MH_EINVAL, "mh_gcmc_set_threads: threads must be 64, 128, 256, 512 or 1024");
    g->threads = threads;
    return MH_OK;
}

int mh_gcmc_set_config(mh_gcmc *g, int rep, int n, const double *com, const double *ang) {
    MH_CHECK(g && rep >= 0 && rep < g->nrep, MH_EINVAL, "mh_gcmc_set_config: bad replica");
    MH_CHECK(n >= 0 && n <= g->maxN && (n == 0 || com), MH_EINVAL, "mh_gcmc_set_config: bad n");
    MH_HIP(hipSetDevice(g->ctx->device));
    const int ns = g->ns;
    std::vector<double> hc(3 * (size_t)n), ha(3 * (size_t)n, 0.0), hp(3 * (size_t)n * ns);
    const GDev &H = g->host;
    for (int m = 0; m < n; m++) {
        for (int i = 0; i < 3; i++) {
            hc[3 * m + i] = com[3 * m + i];
            if (ang) ha[3 * m + i] = ang[3 * m + i];
        }
        // host restatement of site_coords (rigidcoords_toxyz + simcell_pbc)
        if (ns == 1) {
            for (int i = 0; i < 3; i++) {
                double c = hc[3 * m + i];
                hp[3 * m + i] = c - std::floor(c / H.full[i]) * H.full[i];
            }
            continue;
        }
        double sp = std::sin(ha[3 * m]), cp = std::cos(ha[3 * m]), st = std::sin(ha[3 * m + 1]),
               ct = std::cos(ha[3 * m + 1]), ss = std::sin(ha[3 * m + 2]), cs = std::cos(ha[3 * m + 2]);
        double M[3][3] = {{cs * cp - ct * sp * ss, -ss * cp - ct * sp * cs, st * sp},
                          {cs * sp + ct * cp * ss, -ss * sp + ct * cp * cs, -st * cp},
                          {ss * st, cs * st, ct}};
        for (int a = 0; a < ns; a++)
            for (int i = 0; i < 3; i++) {
                double d = M[i][0] * H.ref[a][0] + M[i][1] * H.ref[a][1] + M[i][2] * H.ref[a][2];
                double rp = hc[3 * m + i] + d;
                hp[(3 * m * ns) + 3 * a + i] = rp - std::floor(rp / H.full[i]) * H.full[i];
            }
    }
    const long long off3 = (long long)rep * g->maxN * 3, offp = (long long)rep * g->maxN * ns * 3;
    if (n) {
        MH_HIP(hipMemcpy(g->com.as<double>() + off3, hc.data(), sizeof(double) * 3 * n, hipMemcpyHostToDevice));
        MH_HIP(hipMemcpy(g->ang.as<double>() + off3, ha.data(), sizeof(double) * 3 * n, hipMemcpyHostToDevice));
        MH_HIP(hipMemcpy(g->pos.as<double>() + offp, hp.data(), sizeof(double) * 3 * n * ns, hipMemcpyHostToDevice));
    }
    MH_HIP(hipMemcpy(g->N.as<int>() + rep, &n, sizeof(int), hipMemcpyHostToDevice));
    // structure factor and energy of every replica from their coordinates
    std::vector<double> e(g->nrep);
    int rc = mh_gcmc_total_energy(g, e.data(), 1);
    return rc;
}

int mh_gcmc_run(mh_gcmc *g, int nsteps) {
    MH_CHECK(g && nsteps >= 0, MH_EINVAL, "mh_gcmc_run: bad arguments");
    if (nsteps == 0) return MH_OK;
    MH_HIP(hipSetDevice(g->ctx->device));
    KTimer t(g->ctx, g->prec == MH_PREC_FP32 ? "gcmc32" : "gcmc64");
    if (g->prec == MH_PREC_FP32) launch_gcmc<MH_PREC_FP32>(g, nsteps);
    else launch_gcmc<MH_PREC_FP64>(g, nsteps);
    MH_HIP(hipGetLastError());
    MH_HIP(hipStreamSynchronize(g->ctx->stream));
    t.stop();
    std::vector<int> err(g->nrep);
    MH_HIP(hipMemcpy(err.data(), g->err.p, 4 * g->nrep, hipMemcpyDeviceToHost));
    for (int i = 0; i < g->nrep; i++) {
        MH_CHECK(err[i] != 2, MH_ECAPACITY, "mh_gcmc_run: a replica reached maxN molecules");
        MH_CHECK(err[i] == 0, MH_ESTATE, "mh_gcmc_run: an existing molecule failed to evaluate");
    }
    g->tr_base = g->tr_base;  // trace window stays anchored at the step set_trace saw
    return MH_OK;
}

int mh_gcmc_get_state(mh_gcmc *g, int32_t *N, double *energy, double *sumN, int64_t *nsteps, int64_t *att,
                      int64_t *acc, int32_t *rng_ix, int32_t *rng_iy) {
    MH_CHECK(g, MH_EINVAL, "mh_gcmc_get_state: g is NULL");
    MH_HIP(hipSetDevice(g->ctx->device));
    const size_t R = g->nrep;
    if (N) MH_HIP(hipMemcpy(N, g->N.p, 4 * R, hipMemcpyDeviceToHost));
    if (energy) MH_HIP(hipMemcpy(energy, g->energy.p, 8 * R, hipMemcpyDeviceToHost));
    if (sumN) MH_HIP(hipMemcpy(sumN, g->sumN.p, 8 * R, hipMemcpyDeviceToHost));
    if (nsteps) MH_HIP(hipMemcpy(nsteps, g->nsteps.p, 8 * R, hipMemcpyDeviceToHost));
    if (att) MH_HIP(hipMemcpy(att, g->att.p, 32 * R, hipMemcpyDeviceToHost));
    if (acc) MH_HIP(hipMemcpy(acc, g->acc.p, 32 * R, hipMemcpyDeviceToHost));
    if (rng_ix) MH_HIP(hipMemcpy(rng_ix, g->ix.p, 4 * R, hipMemcpyDeviceToHost));
    if (rng_iy) MH_HIP(hipMemcpy(rng_iy, g->iy.p, 4 * R, hipMemcpyDeviceToHost));
    return MH_OK;
}

int mh_gcmc_set_rng(mh_gcmc *g, int rep, int32_t ix, int32_t iy) {
    MH_CHECK(g && rep >= 0 && rep < g->nrep, MH_EINVAL, "mh_gcmc_set_rng: bad replica");
    MH_HIP(hipSetDevice(g->ctx->device));
    int zero = 0;
    MH_HIP(hipMemcpy(g->ix.as<int32_t>() + rep, &ix, 4, hipMemcpyHostToDevice));
    MH_HIP(hipMemcpy(g->iy.as<int32_t>() + rep, &iy, 4, hipMemcpyHostToDevice));
    MH_HIP(hipMemcpy(g->first.as<int32_t>() + rep, &zero, 4, hipMemcpyHostToDevice));
    return MH_OK;
}

int mh_gcmc_get_config(mh_gcmc *g, int rep, int *n, double *com, double *ang, double *r) {
    MH_CHECK(g && rep >= 0 && rep < g->nrep && n, MH_EINVAL, "mh_gcmc_get_config: bad arguments");
    MH_HIP(hipSetDevice(g->ctx->device));
    int hn = 0;
    MH_HIP(hipMemcpy(&hn, g->N.as<int>() + rep, 4, hipMemcpyDeviceToHost));
    *n = hn;
    const long long off3 = (long long)rep * g->maxN * 3, offp = (long long)rep * g->maxN * g->ns * 3;
    if (hn && com) MH_HIP(hipMemcpy(com, g->com.as<double>() + off3, sizeof(double) * 3 * hn, hipMemcpyDeviceToHost));
    if (hn && ang) MH_HIP(hipMemcpy(ang, g->ang.as<double>() + off3, sizeof(double) * 3 * hn, hipMemcpyDeviceToHost));
    if (hn && r) MH_HIP(hipMemcpy(r, g->pos.as<double>() + offp, sizeof(double) * 3 * hn * g->ns, hipMemcpyDeviceToHost));
    return MH_OK;
}

int mh_gcmc_total_energy(mh_gcmc *g, double *e, int refresh) {
    MH_CHECK(g && e, MH_EINVAL, "mh_gcmc_total_energy: bad arguments");
    MH_HIP(hipSetDevice(g->ctx->device));
    Rep r = g->rep();
    DBuf save, out;
    int rc;
    const size_t R = g->nrep;
    if (g->nk && !refresh) {
        if ((rc = save.alloc(sizeof(double2) * R * (g->nk + 1)))) return rc;
        MH_HIP(hipMemcpyAsync(save.p, g->S.p, sizeof(double2) * R * g->nk, hipMemcpyDeviceToDevice, g->ctx->stream));
    }
    if (g->nk) {
        hipLaunchKernelGGL(k_sk_recompute, dim3((g->nk + 127) / 128, g->nrep), dim3(128), 0, g->ctx->stream,
                           g->dP.as<GDev>(), r);
        MH_HIP(hipGetLastError());
    }
    if ((rc = out.alloc(8 * R))) return rc;
    hipLaunchKernelGGL(k_total<256>, dim3(g->nrep), dim3(256), 0, g->ctx->stream, g->dP.as<GDev>(), r,
                       out.as<double>());
    MH_HIP(hipGetLastError());
    MH_HIP(hipMemcpyAsync(e, out.p, 8 * R, hipMemcpyDeviceToHost, g->ctx->stream));
    if (g->nk && !refresh)
        MH_HIP(hipMemcpyAsync(g->S.p, save.p, sizeof(double2) * R * g->nk, hipMemcpyDeviceToDevice, g->ctx->stream));
    if (refresh) MH_HIP(hipMemcpyAsync(g->energy.p, out.p, 8 * R, hipMemcpyDeviceToDevice, g->ctx->stream));
    MH_HIP(hipStreamSynchronize(g->ctx->stream));
    return MH_OK;
}

int mh_gcmc_set_trace(mh_gcmc *g, int capacity) {
    MH_CHECK(g && capacity >= 0, MH_EINVAL, "mh_gcmc_set_trace: bad arguments");
    MH_HIP(hipSetDevice(g->ctx->device));
    const size_t n = (size_t)g->nrep * (capacity > 0 ? capacity : 1);
    int rc;
    if ((rc = g->tr_move.alloc(4 * n)) || (rc = g->tr_acc.alloc(4 * n)) || (rc = g->tr_N.alloc(4 * n)) ||
        (rc = g->tr_dE.alloc(8 * n)))
        return rc;
    MH_HIP(hipMemset(g->tr_move.p, 0xff, 4 * n));
    g->tr_cap = capacity;
    // the trace starts at the current step count of replica 0 (all replicas
    // advance together)
    long long ns0 = 0;
    MH_HIP(hipMemcpy(&ns0, g->nsteps.p, 8, hipMemcpyDeviceToHost));
    g->tr_base = ns0;
    return MH_OK;
}

int mh_gcmc_get_trace(mh_gcmc *g, int rep, int *move, int *accepted, int *N, double *dE, int *ntrace) {
    MH_CHECK(g && rep >= 0 && rep < g->nrep, MH_EINVAL, "mh_gcmc_get_trace: bad replica");
    MH_HIP(hipSetDevice(g->ctx->device));
    const size_t c = g->tr_cap, off = (size_t)rep * c;
    if (ntrace) *ntrace = g->tr_cap;
    if (!c) return MH_OK;
    if (move) MH_HIP(hipMemcpy(move, g->tr_move.as<int>() + off, 4 * c, hipMemcpyDeviceToHost));
    if (accepted) MH_HIP(hipMemcpy(accepted, g->tr_acc.as<int>() + off, 4 * c, hipMemcpyDeviceToHost));
    if (N) MH_HIP(hipMemcpy(N, g->tr_N.as<int>() + off, 4 * c, hipMemcpyDeviceToHost));
    if (dE) MH_HIP(hipMemcpy(dE, g->tr_dE.as<double>() + off, 8 * c, hipMemcpyDeviceToHost));
    return MH_OK;
}

int mh_trial_energy(mh_gcmc *g, int rep, int n, const double *com, const double *ang, double *e, uint8_t *ok) {
    MH_CHECK(g && rep >= 0 && rep < g->nrep && n >= 0 && (n == 0 || (com && e && ok)), MH_EINVAL,
             "mh_trial_energy: bad arguments");
    if (n == 0) return MH_OK;
    MH_HIP(hipSetDevice(g->ctx->device));
    DBuf dc, da, de, dok, dsb;
    int rc;
    if ((rc = dc.alloc(sizeof(double) * 3 * n)) || (rc = da.alloc(sizeof(double) * 3 * n)) ||
        (rc = de.alloc(sizeof(double) * n)) || (rc = dok.alloc(n)) ||
        (rc = dsb.alloc(sizeof(double2) * (size_t)n * (g->nk + 1))))
        return rc;
    std::vector<double> za(3 * (size_t)n, 0.0);
    MH_HIP(hipMemcpy(dc.p, com, sizeof(double) * 3 * n, hipMemcpyHostToDevice));
    MH_HIP(hipMemcpy(da.p, ang ? ang : za.data(), sizeof(double) * 3 * n, hipMemcpyHostToDevice));
    Rep r = g->rep();
    if (g->prec == MH_PREC_FP32)
        hipLaunchKernelGGL((k_trial<256, MH_PREC_FP32>), dim3(n), dim3(256), 0, g->ctx->stream, g->dP.as<GDev>(), r,
                           rep, n, dc.as<double>(), da.as<double>(), dsb.as<double2>(), de.as<double>(),
                           dok.as<uint8_t>());
    else
        hipLaunchKernelGGL((k_trial<256, MH_PREC_FP64>), dim3(n), dim3(256), 0, g->ctx->stream, g->dP.as<GDev>(), r,
                           rep, n, dc.as<double>(), da.as<double>(), dsb.as<double2>(), de.as<double>(),
                           dok.as<uint8_t>());
    MH_HIP(hipGetLastError());
    MH_HIP(hipStreamSynchronize(g->ctx->stream));
    MH_HIP(hipMemcpy(e, de.p, sizeof(double) * n, hipMemcpyDeviceToHost));
    MH_HIP(hipMemcpy(ok, dok.p, n, hipMemcpyDeviceToHost));
    return MH_OK;
}

}  // extern "C"
