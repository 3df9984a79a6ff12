// libmusic_hip: device-resident GCMC of rigid sorbates (K9), Ewald
// structure-factor deltas (K8), map interpolation (K6) and sorbate-sorbate
// pair energies (K7) inside one persistent kernel per mh_gcmc_run call.
//
// One workgroup per replica (state point).  Every lane keeps an identical
// copy of the replica's RNG state, N, running energy and counters, so the
// reference's serial draw order (gcmc.F90:240-250, gcmcmoves.F90:270-364,
// mcmoves.F90:290-647, rigidmoves.F90:691-1570) is reproduced without
// communication; the energy of a trial is computed by all lanes in parallel
// and reduced in a fixed order:
//   * maps: one lane per (position set, site, cell corner), the 8 corner
//     terms of the tricubic Hermite sum reduced by lane shuffles; the map is
//     read from an uncompacted node-major copy (one HBM round trip);
//   * pairs: lanes over (other molecule, its site, trial site) items;
//   * Ewald reciprocal space: lane t owns the k vectors t, t + NT, ...; their
//     running structure factor S(k), weight and the trial delta dS(k) stay in
//     registers for the whole launch (KPL = k vectors per lane), the per-site
//     e^{i m 2 pi x / L} tables are built in LDS by one lane per
//     (set, site, axis) with a complex recurrence.
// The replica's molecules (site positions, centres of mass, Euler angles) are
// held in LDS for the launch when they fit (LDSMOL), else read from HBM.
//
// Layout in HBM (per replica r, stride by replica):
//   com[r][maxN][3], ang[r][maxN][3], pos[r][maxN][ns][3]  (FP64)
//   S[r][nk] (double2 running structure factor), dS[r][nk] scratch
//   scalars: N, energy, sumN, sumN2, nsteps, att[4], acc[4], rng ix/iy/first/iseed
#include <cmath>
#include <cstring>
#include <vector>

#include "mh_device.h"
#include "mh_internal.h"
#include "mh_mc_common.h"
#include "mh_bmap_dev.h"

namespace mh {

constexpr int GS = MH_MAX_SITES;   // max sites
constexpr int GKMAX = 15;          // max kmax for the sorbate Ewald
constexpr double kTwoPiR = 2.0 * 3.1415927;  // defaults.F90:35
constexpr double kTokcalG = 332.064;


constexpr int GSPC = MH_MAX_SPECIES;  // max species of a mixture

// One sorbate species: its rigid model, maps and move set.
struct SpcDev {
    int ns;
    double ref[GS][3];
    int st[GS];       // site types (index into the common pair tables)
    double q[GS];
    int has_p[GS], has_e[GS];
    MapRef pm[GS], em[GS];
    int nmoves;
    int mtype[4];
    double cumwt[4];
    double insdel, dtrans, drot, maxnrg, molconst;
    int adj_trans, adj_rot;  // adjustable displacements (rigidmoves_adjustdelta*)
    double inv_ns, inv_nss;  // 1/ns, 1/ns^2 for the pair-item decomposition
    BiasRef bias;            // BINSERT / BDELETE
    // cavity-biased BINSERT / BDELETE (CAVITY_BIAS, MAP_CAVITY_BIAS): trials
    // per insertion (0: not cavity biased), centres from the bias map
    int cav_T, cav_map;
};

// A molecular system: the SpcDev base is the species of the move in
// progress (a copy of sp[s] in the kernel's LDS); the rest is shared by
// the species (cell, pair tables, Ewald, temperature).
struct GDev : SpcDev {
    int nst;
    double A[64], B[64], h2[64], l2[64];
    float Af[64], Bf[64], h2f[64], l2f[64];
    int on[64];
    int ewald, kmax, nk;
    double kappa, rc2;
    // plain Coulomb: 331.5 shield, hicut^2 (0: none), 1/hicut (0: none), locut^2
    int coul;
    int con[64];
    double cfac, ch2, cshift, cl2;
    // centre-of-mass cutoffs (ssbasic COMCUT@ / COMTRUNC@, ssbasic.F90:
    // 745-756), squared, per species pair: a molecule pair whose geometric
    // centres are further apart has no LJ (cc_lj) or plain-Coulomb (cc_cl)
    // term; 0 = no cutoff.  comcut: any of them set.
    int comcut;
    double cc_lj[GSPC][GSPC], cc_cl[GSPC][GSPC];
    const double *wk;      // [nk]
    const int8_t *kidx;    // [nk][4] (mx, my, mz, pad)
    double full[3], fund[3];
    double rti;
    int ncell[3];            // unit cells per axis of the simulation cell
    int prec, maxN;
    // species: molecules of species s occupy slots [s maxN, (s + 1) maxN) of
    // a replica, each slot nsmax sites (the batch's largest molecule)
    int nspc, nsmax;
    SpcDev sp[GSPC];
    // e^{i (kx x + ky y)} table per (set, site) in dynamic LDS at byte
    // offset exy_off: [2][ns][kmax + 1][2 kmax + 1] (double2, float2 in FP32)
    int use_exy, exy_off;
    // slanted simulation cell: unslant U / slant S matrices (fundcell.F90:
    // 378-412) and the simulation cell's lattice vectors (insertion points)
    int slanted;
    SlantCell sc;
    double lvec[3][3];
    // cavity list (cavitylist.F90): cav_n cubes per axis of edge cav_d over
    // the simulation cell, tabulated cubes (NULL: all), species whose sites
    // occupy cubes (bit s)
    int cav_on, cav_n[3], cav_mask;
    double cav_d[3];
    const uint8_t *cav_tab;
};

// rigidmoves_adjustdeltatrans / _adjustdeltarot (rigidmoves.F90:2586-2617):
// steer the cumulative acceptance ratio of the move towards 1/2, the
// displacement kept within [0.01, 2.0] (MIN/MAX_TRANS_DELTA, :216-217).
// The thresholds are default-real (float32) literals compared with the
// FP64 ratio succ/att (moves.F90:553-560), i.e. 0.49f and 0.51f widened.
__device__ __forceinline__ double adjust_delta(double d, double ratio) {
    if (ratio < (double)0.49f) return fmax(__dmul_rn(d, 0.95), 0.01);
    if (ratio > (double)0.51f) return fmin(__dmul_rn(d, 1.05), 2.0);
    return d;
}

// per replica; per-species quantities are [rep][GSPC](...) and a replica's
// molecules [rep][nspc maxN] slots (species-major) of nsmax sites
struct Rep {
    int *N, *err;             // N: [rep][GSPC]
    int32_t *ix, *iy, *first, *iseed;
    double *B, *energy, *sumN, *sumN2;  // B, sumN, sumN2: [rep][GSPC]
    long long *nsteps, *att, *acc;      // att, acc: [rep][GSPC][4]
    double *com, *ang, *pos;  // [rep][slots][3], [rep][slots][3], [rep][slots][nsmax][3]
    double *ctr;              // [rep][slots][3] geometric centres (centre-of-mass cutoffs)
    double *delta;            // [rep][GSPC][2] translate / rotate displacement in use
    double2 *S, *dS;          // [rep][nk]
    int *tr_move, *tr_acc, *tr_N;
    double *tr_dE;
    int tr_cap;
    long long tr_base;        // step index at which the trace starts
    // slot pools: system (GDev table index) of each replica, -1 = free slot;
    // remaining step budget per replica.  NULL for a single-system batch.
    const int *sys;
    int *budget;
    long long kstride;        // S(k) entries per replica
    // cavity bias: sites per cube [rep][cube_stride]; cav_prob running
    // averages and their update counts [rep][GSPC][cav_cap]
    int *cube;
    long long cube_stride;
    double *cav_avg;
    long long *cav_cnt;
    int cav_cap;
};




// row i of matrix_b2feuler (matrix.F90:383-403) from sin/cos of phi, theta, psi
__device__ __forceinline__ void euler_row(int i, double sp, double cp, double st, double ct, double ss,
                                          double cs, double m[3]) {
    if (i == 0) {
        m[0] = __dsub_rn(__dmul_rn(cs, cp), __dmul_rn(__dmul_rn(ct, sp), ss));
        m[1] = __dsub_rn(__dmul_rn(-ss, cp), __dmul_rn(__dmul_rn(ct, sp), cs));
        m[2] = __dmul_rn(st, sp);
    } else if (i == 1) {
        m[0] = __dadd_rn(__dmul_rn(cs, sp), __dmul_rn(__dmul_rn(ct, cp), ss));
        m[1] = __dadd_rn(__dmul_rn(-ss, sp), __dmul_rn(__dmul_rn(ct, cp), cs));
        m[2] = __dmul_rn(-st, cp);
    } else {
        m[0] = __dmul_rn(ss, st);
        m[1] = __dmul_rn(cs, st);
        m[2] = ct;
    }
}

// coordinate i of site a, unfolded: rigidcoords_toxyz (rigidcoords.F90:
// 254-276); site_coord adds the pbc
__device__ __forceinline__ double site_coord_unf(const GDev &P, int a, int i, const double com[3],
                                                 const double m[3]) {
    double d = 0.0;
    d = __dadd_rn(d, __dmul_rn(m[0], P.ref[a][0]));
    d = __dadd_rn(d, __dmul_rn(m[1], P.ref[a][1]));
    d = __dadd_rn(d, __dmul_rn(m[2], P.ref[a][2]));
    return __dadd_rn(com[i], d);
}

__device__ __forceinline__ double site_coord(const GDev &P, int a, int i, const double com[3], const double m[3]) {
    return wrap_full(site_coord_unf(P, a, i, com, m), P.full[i]);
}

// config_getMolecCTR (config.F90:1788-1802): coordinate i of the mean of the
// ns unfolded site positions rp[3 a + i], summed in site order
__device__ __forceinline__ double mol_centre(const double *rp, int ns, int i) {
    double c = 0.0;
    for (int a = 0; a < ns; a++) c = __dadd_rn(c, rp[3 * a + i]);
    return __ddiv_rn(c, (double)ns);
}

// simcell_pbc of one position in a slanted cell (fundcell.F90:640-709 basis):
// slanted coordinates folded into [0, ell), back to Cartesian
__device__ __forceinline__ void slant_pbc(const GDev &P, double r[3]) {
    double v[3];
    mat_vec3(P.sc.S, r, v);
    for (int i = 0; i < 3; i++) v[i] = wrap_full(v[i], P.full[i]);
    mat_vec3(P.sc.U, v, r);
}

// simcell_minimage of a site pair (fundcell_snglMinImage, fundcell.F90:
// 582-623): sep = a - b, folded once in the slanted basis when slanted
__device__ __forceinline__ void min_image(const GDev &P, const double *a, const double *b, double sep[3],
                                          bool slanted) {
    if (!slanted) {
        for (int i = 0; i < 3; i++) sep[i] = fold64(a[i] - b[i], P.full[i]);
        return;
    }
    double va[3], vb[3], d[3];
    mat_vec3(P.sc.S, a, va);
    mat_vec3(P.sc.S, b, vb);
    for (int i = 0; i < 3; i++) d[i] = fold64(va[i] - vb[i], P.full[i]);
    mat_vec3(P.sc.U, d, sep);
}

// all site coordinates of one molecule by one lane (host-side paths); unf
// (optional) receives them unfolded
__device__ void site_coords(const GDev &P, const double com[3], const double ang[3], double *out,
                            double *unf = nullptr) {
    if (P.slanted) {
        if (P.ns == 1) {
            for (int i = 0; i < 3; i++) out[i] = com[i];
        } else {
            double sp, cp, st, ct, ss, cs;
            sincos(ang[0], &sp, &cp);
            sincos(ang[1], &st, &ct);
            sincos(ang[2], &ss, &cs);
            for (int i = 0; i < 3; i++) {
                double m[3];
                euler_row(i, sp, cp, st, ct, ss, cs, m);
                for (int a = 0; a < P.ns; a++) {
                    double d = 0.0;
                    d = __dadd_rn(d, __dmul_rn(m[0], P.ref[a][0]));
                    d = __dadd_rn(d, __dmul_rn(m[1], P.ref[a][1]));
                    d = __dadd_rn(d, __dmul_rn(m[2], P.ref[a][2]));
                    out[3 * a + i] = __dadd_rn(com[i], d);
                }
            }
        }
        if (unf)
            for (int i = 0; i < 3 * P.ns; i++) unf[i] = out[i];
        for (int a = 0; a < P.ns; a++) slant_pbc(P, out + 3 * a);
        return;
    }
    if (P.ns == 1) {
        for (int i = 0; i < 3; i++) {
            out[i] = wrap_full(com[i], P.full[i]);
            if (unf) unf[i] = com[i];
        }
        return;
    }
    double sp, cp, st, ct, ss, cs;
    sincos(ang[0], &sp, &cp);
    sincos(ang[1], &st, &ct);
    sincos(ang[2], &ss, &cs);
    for (int i = 0; i < 3; i++) {
        double m[3];
        euler_row(i, sp, cp, st, ct, ss, cs, m);
        for (int a = 0; a < P.ns; a++) {
            const double u = site_coord_unf(P, a, i, com, m);
            out[3 * a + i] = wrap_full(u, P.full[i]);
            if (unf) unf[3 * a + i] = u;
        }
    }
}



// Hermite contribution of one cell corner: sum over the 8 stored values of
// f[s] hx hy hz (FP64 products in the reference order, or FP32).
template <int PREC>
__device__ __forceinline__ bool corner_term(const MapRef &m, const double *r, int corner, double &part,
                                            const SlantCell *sc = nullptr) {
    int i0[3], i1[3];
    double t[3];
    if (sc) {
        // simcell_maptouc in the slanted basis, then the node cell without
        // the periodic wrap of the upper corner (interpolate.F90:675-679)
        double v[3], c[3];
        mat_vec3(sc->S, r, v);
        for (int d = 0; d < 3; d++) {
            double f = __dmul_rn(v[d], 1.0 / m.ell[d]);
            f = (f >= 0.0) ? fmod(f, 1.0) : __dadd_rn(fmod(f, 1.0), 1.0);
            v[d] = __dmul_rn(f, m.ell[d]);
        }
        mat_vec3(sc->U, v, c);
        for (int d = 0; d < 3; d++) {
            const int idx = (int)(c[d] / m.step[d] + 1.0);
            if (idx < 1 || idx >= m.n[d]) return false;
            t[d] = (c[d] - (double)(idx - 1) * m.step[d]) / m.step[d];
            i0[d] = idx - 1;
            i1[d] = idx;
        }
    } else {
        for (int d = 0; d < 3; d++)
            if (!cell_axis(r[d], m.ell[d], m.step[d], m.n[d], i0[d], i1[d], t[d])) return false;
    }
    const int cx = (kCX >> corner) & 1, cy = (kCY >> corner) & 1, cz = (kCZ >> corner) & 1;
    const long long node = ((long long)(cx ? i1[0] : i0[0]) * m.n[1] + (cy ? i1[1] : i0[1])) * m.n[2] +
                           (cz ? i1[2] : i0[2]);
    const float4 *rec = reinterpret_cast<const float4 *>(m.dense + 8 * node);
    const float4 a = rec[0], b = rec[1];
    if (a.x != a.x) return false;  // untabulated node
    const float f[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
    if (PREC == MH_PREC_FP64) {
        double hx[2], hy[2], hz[2];
        hx[0] = cx ? h0f(t[0]) : h0f(1 - t[0]);
        hx[1] = cx ? h1f(t[0], m.step[0]) : h1f(1 - t[0], -m.step[0]);
        hy[0] = cy ? h0f(t[1]) : h0f(1 - t[1]);
        hy[1] = cy ? h1f(t[1], m.step[1]) : h1f(1 - t[1], -m.step[1]);
        hz[0] = cz ? h0f(t[2]) : h0f(1 - t[2]);
        hz[1] = cz ? h1f(t[2], m.step[2]) : h1f(1 - t[2], -m.step[2]);
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < 8; k++)
            s = __dadd_rn(s, __dmul_rn(__dmul_rn(__dmul_rn((double)f[k], hx[(kCX >> k) & 1]), hy[(kCY >> k) & 1]),
                                       hz[(kCZ >> k) & 1]));
        part = s;
    } else {
#pragma clang fp contract(fast)
        float hx[2], hy[2], hz[2];
        const float tx = cx ? (float)t[0] : (float)(1 - t[0]);
        const float ty = cy ? (float)t[1] : (float)(1 - t[1]);
        const float tz = cz ? (float)t[2] : (float)(1 - t[2]);
        hx[0] = tx * tx * (3.0f - 2.0f * tx);
        hx[1] = (cx ? (float)m.step[0] : -(float)m.step[0]) * tx * tx * (tx - 1.0f);
        hy[0] = ty * ty * (3.0f - 2.0f * ty);
        hy[1] = (cy ? (float)m.step[1] : -(float)m.step[1]) * ty * ty * (ty - 1.0f);
        hz[0] = tz * tz * (3.0f - 2.0f * tz);
        hz[1] = (cz ? (float)m.step[2] : -(float)m.step[2]) * tz * tz * (tz - 1.0f);
        float s = 0.0f;
#pragma unroll
        for (int k = 0; k < 8; k++) s += f[k] * hx[(kCX >> k) & 1] * hy[(kCY >> k) & 1] * hz[(kCZ >> k) & 1];
        part = s;
    }
    return true;
}

// map energy of one site by one lane (totals and trial energies off the hot
// loop): the 8 corner terms in corner order
__device__ bool site_map_energy(const GDev &P, const SpcDev &Q, int a, const double *r, double &e) {
    e = 0.0;
    double pm = 0.0, em = 0.0;
    for (int w = 0; w < 2; w++) {
        if (!(w == 0 ? Q.has_p[a] : Q.has_e[a])) continue;
        const MapRef &m = w == 0 ? Q.pm[a] : Q.em[a];
        double s = 0.0;
        for (int c = 0; c < 8; c++) {
            double part;
            if (!corner_term<MH_PREC_FP64>(m, r, c, part, P.slanted ? &P.sc : nullptr)) return false;
            s = __dadd_rn(s, part);
        }
        (w == 0 ? pm : em) = s;
    }
    if (Q.has_p[a]) e = __dadd_rn(e, pm);
    if (Q.has_e[a]) e = __dadd_rn(e, __dmul_rn(em, Q.q[a]));
    return true;
}

// plain Coulomb of a site pair, coulomb_snglint (coulomb.F90:271-306):
// (331.5 shield q1 q2)(1/r - 1/hicut) within hicut; false below locut.  FP64
// in both kernel precisions: without a cutoff thousands of +-1/r terms
// cancel, and FP32 pair terms would leave ~1e-5 of the net.
__device__ __forceinline__ bool coul_energy64(const GDev &P, int t, double qa, double qb, double r2, double &e,
                                              bool pair_on = true) {
    if (pair_on && P.con[t] && !(P.ch2 > 0.0 && r2 > P.ch2)) {
        if (r2 < P.cl2) return false;
        const double r = sqrt(r2);
        e = __dadd_rn(e, __dmul_rn(__dmul_rn(__dmul_rn(P.cfac, qa), qb), __dsub_rn(1.0 / r, P.cshift)));
    }
    return true;
}

// pair energy of a site pair (lj_snglint without derivatives + Ewald real
// space + plain Coulomb).  Returns false on an LJ or Coulomb locut violation.
// lj_on / cl_on: the molecule pair is within its centre-of-mass cutoffs.
__device__ __forceinline__ bool pair_energy64(const GDev &P, int t, double qa, double qb, double sx, double sy,
                                              double sz, double &e, bool lj_on = true, bool cl_on = true) {
    double r2 = __dadd_rn(__dadd_rn(__dmul_rn(sx, sx), __dmul_rn(sy, sy)), __dmul_rn(sz, sz));
    if (lj_on && P.on[t] && !(r2 > P.h2[t])) {
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
    return P.coul ? coul_energy64(P, t, qa, qb, r2, e, cl_on) : true;
}

__device__ __forceinline__ bool pair_energy32(const GDev &P, int t, float qq, float kappa, float rc2, float sx,
                                              float sy, float sz, float &e, bool lj_on = true) {
#pragma clang fp contract(fast)
    const float r2 = sx * sx + sy * sy + sz * sz;
    if (lj_on && P.on[t] && !(r2 > P.h2f[t])) {
        if (r2 < P.l2f[t]) return false;
        const float r2i = rcp_fast(r2);
        const float r6i = r2i * r2i * r2i;
        e += P.Af[t] * (r6i * r6i) - P.Bf[t] * r6i;
    }
    if (qq != 0.0f && r2 < rc2) {
        const float ri = __frsqrt_rn(r2);
        e += qq * erfc_as32(kappa * (r2 * ri)) * ri;
    }
    return true;
}


struct Shared {
    int N[GSPC];                         // molecules per species (species of the move: its count)
    int spc;                             // species of the move in progress
    double posA[GS * 3], posB[GS * 3];   // trial (A) and old (B) site coordinates
    double rpA[GS * 3];                  // trial sites unfolded (centre-of-mass cutoffs)
    double ctr[2][3];                    // geometric centres of sets A and B
    double sc[3][2];                     // sin, cos of the trial Euler angles
    double mapv[2][GS];
    int mapok[2][GS];
    double2 E[2][GS][3][GKMAX + 1];      // per position set, site, axis: e^{i m 2 pi x / L}
    float2 Ef[2][GS][3][GKMAX + 1];      // FP32 copy
    double red[16 * 6];
    unsigned long long cavbits[16];      // cavity trials that found a cavity, per wave
    double cavsel[7];                    // the chosen trial: centre, angles, bias
    double cavpc;                        // cav_prob(totmol) after its update
};

// k vectors owned by one lane, kept in registers for the launch
template <int KPL>
struct KReg {
    double2 S[KPL > 0 ? KPL : 1];
    double2 d[KPL > 0 ? KPL : 1];
    double w[KPL > 0 ? KPL : 1];
    int m[KPL > 0 ? KPL : 1];  // packed (mx, my, mz) bytes
};

__device__ __forceinline__ int kpack(const int8_t *ki) {
    return (int)(uint8_t)ki[0] | ((int)(uint8_t)ki[1] << 8) | ((int)(uint8_t)ki[2] << 16);
}

// k vector as (index into the e^{i(kx x + ky y)} table, |mz|, mz < 0)
__device__ __forceinline__ int kpack_xy(const int8_t *ki, int kmax) {
    const int mz = ki[2];
    return (ki[0] * (2 * kmax + 1) + ki[1] + kmax) | ((mz < 0 ? -mz : mz) << 10) | ((mz < 0) << 15);
}

template <typename T2>
__device__ __forceinline__ T2 expk(const T2 *Ex, const T2 *Ey, const T2 *Ez, int mx, int my, int mz) {
    T2 a = Ex[mx];
    T2 b = Ey[my < 0 ? -my : my];
    if (my < 0) b.y = -b.y;
    T2 c = Ez[mz < 0 ? -mz : mz];
    if (mz < 0) c.y = -c.y;
    T2 ab;
    ab.x = a.x * b.x - a.y * b.y;
    ab.y = a.x * b.y + a.y * b.x;
    T2 o;
    o.x = ab.x * c.x - ab.y * c.y;
    o.y = ab.x * c.y + ab.y * c.x;
    return o;
}

// delta of the structure factor at one k for the present position sets
template <int PREC>
__device__ __forceinline__ double2 k_delta(const GDev &P, const Shared &sh, int mx, int my, int mz, bool hasA,
                                           bool hasB, double asign) {
    if (PREC == MH_PREC_FP64) {
        double2 d = make_double2(0.0, 0.0);
        for (int a = 0; a < P.ns; a++) {
            if (P.q[a] == 0.0) continue;
            if (hasA) {
                double2 e = expk<double2>(sh.E[0][a][0], sh.E[0][a][1], sh.E[0][a][2], mx, my, mz);
                d.x += asign * P.q[a] * e.x;
                d.y += asign * P.q[a] * e.y;
            }
            if (hasB) {
                double2 e = expk<double2>(sh.E[1][a][0], sh.E[1][a][1], sh.E[1][a][2], mx, my, mz);
                d.x -= P.q[a] * e.x;
                d.y -= P.q[a] * e.y;
            }
        }
        return d;
    } else {
#pragma clang fp contract(fast)
        float dx = 0.0f, dy = 0.0f;
        for (int a = 0; a < P.ns; a++) {
            if (P.q[a] == 0.0) continue;
            const float qa = (float)P.q[a];
            if (hasA) {
                float2 e = expk<float2>(sh.Ef[0][a][0], sh.Ef[0][a][1], sh.Ef[0][a][2], mx, my, mz);
                dx += (float)asign * qa * e.x;
                dy += (float)asign * qa * e.y;
            }
            if (hasB) {
                float2 e = expk<float2>(sh.Ef[1][a][0], sh.Ef[1][a][1], sh.Ef[1][a][2], mx, my, mz);
                dx -= qa * e.x;
                dy -= qa * e.y;
            }
        }
        return make_double2((double)dx, (double)dy);
    }
}

// Energy terms of a move: set A (+) and set B (-) of one molecule of species
// sh.spc against every other molecule of every species (skip = its slot in
// its species, N = that species' count) and the reciprocal-space delta,
// whose dS(k) is left in K.d (registers) or dSg (KPL == 0).
// out: map A, pairs A, map B, pairs B, reciprocal delta, ok(A).
// Centre-of-mass cutoffs (P.comcut): ctrg = the replica's molecule centres
// [slots][3]; set A's centre is cA (a stored molecule) or the mean of
// sh.rpA (a trial), set B's is cB.
template <int NT, int PREC, int KPL, bool SL = false>
__device__ __forceinline__ void move_energy(const GDev &P, Shared &sh, const double *__restrict__ pos, int N,
                                            int skip, bool hasA, bool hasB, double asign, KReg<KPL> &K,
                                            const double2 *__restrict__ Sg, double2 *__restrict__ dSg,
                                            void *exy, double out[6], const double *__restrict__ ctrg,
                                            const double *cA, const double *cB) {
    const int tid = threadIdx.x;
    const int ns = P.ns;
    if (P.comcut && tid >= NT - 6) {  // read after the barrier below
        const int set = (tid - (NT - 6)) / 3, i = (tid - (NT - 6)) % 3;
        if (set == 0 && hasA) sh.ctr[0][i] = cA ? cA[i] : mol_centre(sh.rpA, ns, i);
        if (set == 1 && hasB) sh.ctr[1][i] = cB[i];
    }
    // pairs: per species s2 the items (molecule j of s2, its site b, trial site a),
    // item w of lane first + k stride
    double eA = 0.0, eB = 0.0, failA = 0.0;
    float eAf = 0.0f, eBf = 0.0f;
    const float kap32 = (float)P.kappa, rc32 = (float)P.rc2;
    auto pairs = [&](int first, int stride) {
    for (int s2 = 0; s2 < P.nspc; s2++) {
        const bool own = s2 == sh.spc;
        const SpcDev &O = (P.nspc == 1) ? static_cast<const SpcDev &>(P) : P.sp[s2];
        const int ns2 = O.ns;
        const int n2 = own ? N : sh.N[s2];
        const int nss = ns2 * ns;
        const int items = n2 * nss;
        const double inv_nss = own ? P.inv_nss : 1.0 / (double)nss;
        const double *pos2 = pos + (long long)s2 * P.maxN * P.nsmax * 3;
        const double *ctr2 = ctrg + (long long)s2 * P.maxN * 3;
        const double ccl = P.cc_lj[sh.spc][s2], ccc = P.cc_cl[sh.spc][s2];
        for (int w = first; w < items; w += stride) {
            const int j = (int)(((double)w + 0.5) * inv_nss);
            if (own && j == skip) continue;
            const int r = w - j * nss;
            const int b = (int)(((double)r + 0.5) * P.inv_ns);
            const int a = r - b * ns;
            const double *pj = pos2 + ((long long)j * P.nsmax + b) * 3;
            const double xb = pj[0], yb = pj[1], zb = pj[2];
            const int t = P.st[a] * P.nst + O.st[b];
            const double qb = O.q[b];
            const double pb[3] = {xb, yb, zb};
            // ssbasic_mmint's centre-to-centre test per set (ssbasic.F90:
            // 745-756): beyond the radius the pair's term is skipped
            bool ljA = true, clA = true, ljB = true, clB = true;
            if (P.comcut) {
                const double *cj = ctr2 + 3LL * j;
                for (int set = 0; set < 2; set++) {
                    if (!(set == 0 ? hasA : hasB)) continue;
                    double sc[3];
                    min_image(P, sh.ctr[set], cj, sc, SL);
                    const double d2 =
                        __dadd_rn(__dadd_rn(__dmul_rn(sc[0], sc[0]), __dmul_rn(sc[1], sc[1])), __dmul_rn(sc[2], sc[2]));
                    const bool lo = !(ccl > 0.0 && d2 > ccl), co = !(ccc > 0.0 && d2 > ccc);
                    if (set == 0) {
                        ljA = lo;
                        clA = co;
                    } else {
                        ljB = lo;
                        clB = co;
                    }
                }
            }
            if (PREC == MH_PREC_FP64) {
                if (hasA) {
                    double sp[3];
                    min_image(P, &sh.posA[3 * a], pb, sp, SL);
                    if (!pair_energy64(P, t, P.q[a], qb, sp[0], sp[1], sp[2], eA, ljA, clA)) failA = 1.0;
                }
                if (hasB) {
                    double sp[3];
                    min_image(P, &sh.posB[3 * a], pb, sp, SL);
                    pair_energy64(P, t, P.q[a], qb, sp[0], sp[1], sp[2], eB, ljB, clB);
                }
            } else {
                const float qq = P.ewald ? (float)(P.q[a] * qb * kTokcalG) : 0.0f;
                if (hasA) {
                    double d[3];
                    min_image(P, &sh.posA[3 * a], pb, d, SL);
                    if (!pair_energy32(P, t, qq, kap32, rc32, (float)d[0], (float)d[1], (float)d[2], eAf, ljA))
                        failA = 1.0;
                    if (P.coul &&
                        !coul_energy64(P, t, P.q[a], qb, d[0] * d[0] + d[1] * d[1] + d[2] * d[2], eA, clA))
                        failA = 1.0;
                }
                if (hasB) {
                    double d[3];
                    min_image(P, &sh.posB[3 * a], pb, d, SL);
                    pair_energy32(P, t, qq, kap32, rc32, (float)d[0], (float)d[1], (float)d[2], eBf, ljB);
                    if (P.coul) coul_energy64(P, t, P.q[a], qb, d[0] * d[0] + d[1] * d[1] + d[2] * d[2], eB, clB);
                }
            }
        }
    }
    };
    // maps: item l = (set * ns + site) * 8 + corner; the 8 corner lanes of a
    // site are an aligned lane group
    for (int l = tid; l < 16 * ns; l += NT) {
        const int set = l / (8 * ns), a = (l >> 3) % ns, corner = l & 7;
        const bool present = set == 0 ? hasA : hasB;
        const double *r = set == 0 ? &sh.posA[3 * a] : &sh.posB[3 * a];
        double pp = 0.0, pe = 0.0;
        int ok = 1;
        if (present) {
            const SlantCell *sc = SL ? &P.sc : nullptr;
            if (P.has_p[a] && !corner_term<PREC>(P.pm[a], r, corner, pp, sc)) ok = 0;
            if (P.has_e[a] && !corner_term<PREC>(P.em[a], r, corner, pe, sc)) ok = 0;
        }
        for (int off = 1; off < 8; off <<= 1) {
            pp += __shfl_xor(pp, off);
            pe += __shfl_xor(pe, off);
            ok &= __shfl_xor(ok, off);
        }
        if (corner == 0 && present) {
            double e = 0.0;
            if (P.has_p[a]) e = __dadd_rn(e, pp);
            if (P.has_e[a]) e = __dadd_rn(e, __dmul_rn(pe, P.q[a]));
            sh.mapv[set][a] = e;
            sh.mapok[set][a] = ok;
        }
    }
    // e^{i m 2 pi x / L} tables: one lane per (set, site, axis), recurrence in m
    if (P.ewald) {
        const int e0 = NT >= 256 ? 128 : (NT == 128 ? 64 : 0);
        const int l = tid - e0;
        if (l >= 0 && l < 6 * ns) {
            const int set = l / (3 * ns), a = (l / 3) % ns, ax = l % 3;
            const bool present = set == 0 ? hasA : hasB;
            if (present && P.q[a] != 0.0) {
                const double *pp = set == 0 ? sh.posA : sh.posB;
                const double th = (1.0 / P.full[ax]) * kTwoPiR * pp[3 * a + ax];
                double s1, c1;
                sincos(th, &s1, &c1);
                double cr = 1.0, ci = 0.0;
                for (int m = 0; m <= P.kmax; m++) {
                    sh.E[set][a][ax][m] = make_double2(cr, ci);
                    sh.Ef[set][a][ax][m] = make_float2((float)cr, (float)ci);
                    const double nr = cr * c1 - ci * s1, ni = cr * s1 + ci * c1;
                    cr = nr;
                    ci = ni;
                }
            }
        }
    }
    // without centre cutoffs, and when the pair items fit the lanes above the
    // first wave (which holds the map lanes) in one pass, those lanes take them
    // before the barrier, while the map reads are in flight
    int nitems = 0;
    for (int s2 = 0; s2 < P.nspc; s2++)
        nitems += (s2 == sh.spc ? N : sh.N[s2]) * ((P.nspc == 1) ? P.ns : P.sp[s2].ns) * ns;
    const bool early = NT >= 128 && !P.comcut && nitems <= NT - 64;
    if (early && tid >= 64) pairs(tid - 64, NT - 64);
    __syncthreads();
    if (KPL > 0 && P.ewald && P.use_exy) {  // the table lives in the persistent kernel's LDS
        const int W = 2 * P.kmax + 1, per = (P.kmax + 1) * W;
        for (int i = tid; i < 2 * ns * per; i += NT) {
            const int sa = i / per, r = i - sa * per;
            const int set = sa / ns, a = sa - set * ns;
            const int mx = r / W, my = r - mx * W - P.kmax;
            if (!(set == 0 ? hasA : hasB) || P.q[a] == 0.0) continue;
            const double2 ex = sh.E[set][a][0][mx];
            double2 ey = sh.E[set][a][1][my < 0 ? -my : my];
            if (my < 0) ey.y = -ey.y;
            const double2 o = make_double2(ex.x * ey.x - ex.y * ey.y, ex.x * ey.y + ex.y * ey.x);
            if (PREC == MH_PREC_FP64) reinterpret_cast<double2 *>(exy)[i] = o;
            else reinterpret_cast<float2 *>(exy)[i] = make_float2((float)o.x, (float)o.y);
        }
        __syncthreads();
    }
    if (!early) pairs(tid, NT);

    if (PREC != MH_PREC_FP64) {  // FP64 plain-Coulomb sums already in eA, eB
        eA += (double)eAf;
        eB += (double)eBf;
    }
    // reciprocal-space delta over the lane's k vectors
    double eK = 0.0;
    if (P.ewald) {
        if constexpr (KPL > 0) {
          if (P.use_exy) {
            // site-major: the table reads of all KPL k vectors issue together
            const int W = 2 * P.kmax + 1, per = (P.kmax + 1) * W;
            if (PREC == MH_PREC_FP64) {
                double2 dd[KPL];
#pragma unroll
                for (int j = 0; j < KPL; j++) dd[j] = make_double2(0.0, 0.0);
                for (int a = 0; a < ns; a++) {
                    if (P.q[a] == 0.0) continue;
                    for (int set = 0; set < 2; set++) {
                        if (!(set == 0 ? hasA : hasB)) continue;
                        const double sc = set == 0 ? asign * P.q[a] : -P.q[a];
                        const double2 *xy = reinterpret_cast<const double2 *>(exy) + (set * ns + a) * per;
                        const double2 *ez = sh.E[set][a][2];
#pragma unroll
                        for (int j = 0; j < KPL; j++) {
                            const int pk = K.m[j];
                            const double2 u = xy[pk & 1023];
                            double2 v = ez[(pk >> 10) & 31];
                            if (pk & 0x8000) v.y = -v.y;
                            const double ex = u.x * v.x - u.y * v.y, ey = u.x * v.y + u.y * v.x;
                            dd[j].x += sc * ex;
                            dd[j].y += sc * ey;
                        }
                    }
                }
#pragma unroll
                for (int j = 0; j < KPL; j++) {
                    const double2 s = K.S[j], d = dd[j];
                    const double nre = s.x + d.x, nim = s.y + d.y;
                    eK += K.w[j] * ((nre * nre + nim * nim) - (s.x * s.x + s.y * s.y));
                    K.d[j] = d;
                }
            } else {
#pragma clang fp contract(fast)
                float2 dd[KPL];
#pragma unroll
                for (int j = 0; j < KPL; j++) dd[j] = make_float2(0.0f, 0.0f);
                for (int a = 0; a < ns; a++) {
                    if (P.q[a] == 0.0) continue;
                    for (int set = 0; set < 2; set++) {
                        if (!(set == 0 ? hasA : hasB)) continue;
                        const float sc = (float)(set == 0 ? asign * P.q[a] : -P.q[a]);
                        const float2 *xy = reinterpret_cast<const float2 *>(exy) + (set * ns + a) * per;
                        const float2 *ez = sh.Ef[set][a][2];
#pragma unroll
                        for (int j = 0; j < KPL; j++) {
                            const int pk = K.m[j];
                            const float2 u = xy[pk & 1023];
                            float2 v = ez[(pk >> 10) & 31];
                            if (pk & 0x8000) v.y = -v.y;
                            dd[j].x += sc * (u.x * v.x - u.y * v.y);
                            dd[j].y += sc * (u.x * v.y + u.y * v.x);
                        }
                    }
                }
#pragma unroll
                for (int j = 0; j < KPL; j++) {
                    const double2 s = K.S[j];
                    const double dx = dd[j].x, dy = dd[j].y;
                    eK += K.w[j] * (2.0 * (s.x * dx + s.y * dy) + (dx * dx + dy * dy));
                    K.d[j] = make_double2(dx, dy);
                }
            }
          } else {
#pragma unroll
            for (int j = 0; j < KPL; j++) {
                const int k = tid + j * NT;
                if (k < P.nk) {
                    const int pk = K.m[j];
                    const int mx = (int8_t)(pk & 0xff), my = (int8_t)((pk >> 8) & 0xff),
                              mz = (int8_t)((pk >> 16) & 0xff);
                    const double2 d = k_delta<PREC>(P, sh, mx, my, mz, hasA, hasB, asign);
                    const double2 s = K.S[j];
                    if (PREC == MH_PREC_FP64) {
                        const double nre = s.x + d.x, nim = s.y + d.y;
                        eK += K.w[j] * ((nre * nre + nim * nim) - (s.x * s.x + s.y * s.y));
                    } else {
                        eK += K.w[j] * (2.0 * (s.x * d.x + s.y * d.y) + (d.x * d.x + d.y * d.y));
                    }
                    K.d[j] = d;
                }
            }
          }
        } else {
            for (int k = tid; k < P.nk; k += NT) {
                const int8_t *ki = P.kidx + 4 * k;
                const double2 d = k_delta<PREC>(P, sh, ki[0], ki[1], ki[2], hasA, hasB, asign);
                const double2 s = Sg[k];
                const double nre = s.x + d.x, nim = s.y + d.y;
                eK += P.wk[k] * ((nre * nre + nim * nim) - (s.x * s.x + s.y * s.y));
                dSg[k] = d;
            }
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

// S(k) += dS(k) for the lane's k vectors after an accepted move
template <int NT, int KPL>
__device__ __forceinline__ void sk_accept(const GDev &P, KReg<KPL> &K, double2 *__restrict__ Sg,
                                          const double2 *__restrict__ dSg) {
    const int tid = threadIdx.x;
    if constexpr (KPL > 0) {
#pragma unroll
        for (int j = 0; j < KPL; j++) {
            if (tid + j * NT < P.nk) {
                K.S[j].x += K.d[j].x;
                K.S[j].y += K.d[j].y;
            }
        }
    } else {
        for (int k = tid; k < P.nk; k += NT) {
            double2 s = Sg[k], d = dSg[k];
            Sg[k] = make_double2(s.x + d.x, s.y + d.y);
        }
    }
}

// The persistent GCMC loop.  LDSMOL: molecules held in dynamic LDS
// ([maxN][ns][3] sites, [maxN][3] com, [maxN][3] angles, FP64).
// KPLX >= 0: k vectors per lane in registers (orthorhombic cell); KPLX < 0:
// slanted cell (no sorbate Ewald sum).
// rigidmoves_comFromBmap (rigidmoves.F90:1938-1969): a cubelet by weight, a
// point in it (bmap_getbiaspt), a unit cell of the simulation cell; biasf =
// 1 / (weight ncubelets)
__device__ __forceinline__ void bmap_com(const GDev &P, Rng &rng, double c[3], double &biasf) {
    const BiasRef &Bm = P.bias;
    const long long t = bias_index(Bm, rranf(rng));
    const long long cb = Bm.cube[t];
    const long long ci[3] = {cb / ((long long)Bm.ncub[1] * Bm.ncub[2]), (cb / Bm.ncub[2]) % Bm.ncub[1],
                             cb % Bm.ncub[2]};
    biasf = __ddiv_rn(__ddiv_rn(1.0, Bm.wn[t]), Bm.ncubelets);
    for (int i = 0; i < 3; i++) c[i] = __dmul_rn((double)ci[i], Bm.step[i]);
    for (int i = 0; i < 3; i++) c[i] = __dadd_rn(c[i], __dmul_rn(rranf(rng), Bm.step[i]));
    for (int i = 0; i < 3; i++)
        if (c[i] >= Bm.eff[i]) c[i] = __dmul_rn((double)0.9999990f, Bm.eff[i]);
    for (int i = 0; i < 3; i++) {
        const int icell = (int)(rranf(rng) * (double)P.ncell[i]);
        c[i] = __dadd_rn(c[i], __dmul_rn(P.fund[i], (double)icell));
    }
}

// cavitylist_getIndexVec (cavitylist.F90:536-546): Int(r / d) per axis of a
// folded position, z fastest; a coordinate rounding onto the upper face
// stays in the last cube
__device__ __forceinline__ long long cav_cube(const GDev &P, const double *r) {
    long long c = 0;
    for (int i = 0; i < 3; i++) {
        int k = (int)__ddiv_rn(r[i], P.cav_d[i]);
        k = min(max(k, 0), P.cav_n[i] - 1);
        c = c * P.cav_n[i] + k;
    }
    return c;
}

// rigidmoves_getConfigFromCavBias (rigidmoves.F90:1987-2126): cav_T trial
// centres (uniform, or from the bias map) with random angles, drawn in
// order by every lane; lane t checks trial t (every site in a tabulated
// cube holding no site of a listed species); cav_prob(totmol) updated with
// the number found (stats_update, by lane 0, which alone reads and writes
// the table); one of the successes at random.  Returns 0 with the trial in
// com / ang and biasf = its map bias x cav_prob / cav_T, 1 when no trial
// found a cavity (the insertion fails), 5 when the total loading exceeds
// the cav_prob table.  Requires cav_T <= NT.
template <int NT>
__device__ int cavity_trials(const GDev &P, Shared &sh, const Rep &R, int rep, int s, Rng &rng, double com[3],
                             double ang[3], double &biasf) {
    const int tid = threadIdx.x, T = P.cav_T, ns = P.ns;
    double my[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 1.0};
    double mycos = 1.0;
    for (int t = 0; t < T; t++) {
        double c[3], a[3] = {0.0, 0.0, 0.0}, bf = 1.0, ct = 1.0;
        if (P.cav_map) bmap_com(P, rng, c, bf);
        else
            for (int i = 0; i < 3; i++) c[i] = __dmul_rn(P.full[i], rranf(rng));
        if (ns != 1) {  // rigidmoves_reorient with newxyz
            ct = __dsub_rn(1.0, __dmul_rn(rranf(rng), 2.0));
            a[0] = __dsub_rn(3.1415927, __dmul_rn(rranf(rng), kTwoPiR));
            a[2] = __dsub_rn(3.1415927, __dmul_rn(rranf(rng), kTwoPiR));
        }
        if (t == tid) {
            for (int i = 0; i < 3; i++) {
                my[i] = c[i];
                my[3 + i] = a[i];
            }
            my[6] = bf;
            mycos = ct;
        }
    }
    bool found = false;
    if (tid < T) {
        if (ns != 1) my[4] = acos(mycos);
        double sc[3][2];
        for (int k = 0; k < 3; k++) sincos(my[3 + k], &sc[k][0], &sc[k][1]);
        const int *cube = R.cube + (long long)rep * R.cube_stride;
        found = true;
        for (int a = 0; a < ns && found; a++) {
            double r[3];
            for (int i = 0; i < 3; i++) {
                if (ns == 1) {
                    r[i] = wrap_full(my[i], P.full[i]);
                } else {
                    double m[3];
                    euler_row(i, sc[0][0], sc[0][1], sc[1][0], sc[1][1], sc[2][0], sc[2][1], m);
                    r[i] = wrap_full(site_coord_unf(P, a, i, my, m), P.full[i]);
                }
            }
            const long long cb = cav_cube(P, r);
            found = (!P.cav_tab || P.cav_tab[cb]) &&
                    __hip_atomic_load(cube + cb, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0;
        }
    }
    const unsigned long long bal = __ballot(found);
    if ((tid & 63) == 0) sh.cavbits[tid >> 6] = bal;
    __syncthreads();
    int nsucc = 0;
    for (int w = 0; w < NT / 64; w++) nsucc += __popcll(sh.cavbits[w]);
    // config_getTotMoles with the slot of the new molecule opened
    int totmol = 1;
    for (int i = 0; i < P.nspc; i++) totmol += sh.N[i];
    if (totmol >= R.cav_cap) return 5;
    if (tid == 0) {
        const long long o = ((long long)rep * GSPC + s) * R.cav_cap + totmol;
        const long long n1 = R.cav_cnt[o] + 1;
        const double avg = __ddiv_rn(__dadd_rn(__dmul_rn(R.cav_avg[o], __dsub_rn((double)n1, 1.0)), (double)nsucc),
                                     __dmul_rn((double)n1, 1.0));
        R.cav_avg[o] = avg;
        R.cav_cnt[o] = n1;
        sh.cavpc = avg;
    }
    __syncthreads();
    if (nsucc == 0) return 1;
    // the pick-th success in trial order
    const int pick = (int)(rranf(rng) * (double)nsucc);
    int tsel = 0, left = pick;
    for (int w = 0; w < NT / 64; w++) {
        unsigned long long b = sh.cavbits[w];
        const int c = __popcll(b);
        if (left < c) {
            for (int j = 0; j < left; j++) b &= b - 1;
            tsel = 64 * w + __builtin_ctzll(b);
            break;
        }
        left -= c;
    }
    if (tid == tsel)
        for (int i = 0; i < 7; i++) sh.cavsel[i] = my[i];
    __syncthreads();
    for (int i = 0; i < 3; i++) {
        com[i] = sh.cavsel[i];
        ang[i] = sh.cavsel[3 + i];
    }
    biasf = __dmul_rn(sh.cavsel[6], __ddiv_rn(sh.cavpc, (double)T));
    return 0;
}

// molecules entering (sign 1) or leaving (-1) the cavity list: its folded
// sites r[3 ns] (cavitylist_updateMolecInfo, cavitylist.F90:351-513, as
// site counts per cube)
__device__ __forceinline__ void cav_sites(const GDev &P, const Rep &R, int rep, int s, const double *r, int sign) {
    const int tid = threadIdx.x;
    if (P.cav_on && (P.cav_mask >> s & 1) && tid < P.ns)
        atomicAdd(R.cube + (long long)rep * R.cube_stride + cav_cube(P, r + 3 * tid), sign);
}

// CAV: the instantiation with cavity-biased moves (S(k) and molecules in
// HBM); the others carry none of its code or registers
template <int NT, int PREC, int KPLX, bool LDSMOL, bool CAV = false>
__global__ void __launch_bounds__(NT) k_gcmc(const GDev *__restrict__ Ptab, Rep R, int nsteps) {
    constexpr int KPL = KPLX > 0 ? KPLX : 0;
    constexpr bool SL = KPLX < 0;
    extern __shared__ double dyn[];
    __shared__ Shared sh;
    __shared__ GDev P;
    const int rep = blockIdx.x;
    const int tid = threadIdx.x;
    int nrun = nsteps;
    if (R.sys) {
        if (R.sys[rep] < 0) return;  // free slot
        nrun = min(nsteps, R.budget[rep]);
        if (nrun <= 0) return;
    }
    const GDev *Pg = Ptab + (R.sys ? R.sys[rep] : 0);
    for (int i = tid; i < (int)(sizeof(GDev) / 4); i += NT)
        reinterpret_cast<int *>(&P)[i] = reinterpret_cast<const int *>(Pg)[i];
    __syncthreads();
    if (R.err[rep]) return;
    // per-species chain state in LDS (the species of a step: a copy in
    // registers, written back at the end of the step)
    __shared__ double sB[GSPC], sDel[GSPC][2], sSum[GSPC], sSum2[GSPC];
    __shared__ long long sAtt[GSPC][4], sAcc[GSPC][4];
    const int nspc = P.nspc;
    if (tid < GSPC) {
        const int o = GSPC * rep + tid;
        const bool live = tid < nspc;
        sh.N[tid] = live ? R.N[o] : 0;
        sB[tid] = live ? R.B[o] : 0.0;
        sSum[tid] = live ? R.sumN[o] : 0.0;
        sSum2[tid] = live ? R.sumN2[o] : 0.0;
        sDel[tid][0] = live ? R.delta[2 * o] : 0.0;
        sDel[tid][1] = live ? R.delta[2 * o + 1] : 0.0;
        for (int i = 0; i < 4; i++) {
            sAtt[tid][i] = live ? R.att[4 * o + i] : 0;
            sAcc[tid][i] = live ? R.acc[4 * o + i] : 0;
        }
    }
    if (tid == 0) sh.spc = 0;  // the uploaded model is species 0's
    int ns = P.ns, cur = 0;
    Rng rng{R.ix[rep], R.iy[rep], R.first[rep], R.iseed[rep]};
    double energy = R.energy[rep];
    long long nst = R.nsteps[rep];
    const long long slots = (long long)nspc * P.maxN, ms3 = 3LL * P.nsmax;
    double *gcom = R.com + (long long)rep * slots * 3;
    double *gang = R.ang + (long long)rep * slots * 3;
    double *gpos = R.pos + (long long)rep * slots * ms3;
    double *gctr = R.ctr + (long long)rep * slots * 3;  // centres stay in HBM
    double2 *S = R.S + (long long)rep * R.kstride;
    double2 *dS = R.dS + (long long)rep * R.kstride;
    double *pos, *com, *ang;
    void *exy = reinterpret_cast<char *>(dyn) + P.exy_off;
    __syncthreads();
    if constexpr (LDSMOL) {
        pos = dyn;
        com = dyn + slots * ms3;
        ang = com + slots * 3;
        for (int s2 = 0; s2 < nspc; s2++) {
            const long long n2 = sh.N[s2], o3 = (long long)s2 * P.maxN * 3, op = (long long)s2 * P.maxN * ms3;
            for (long long i = tid; i < n2 * ms3; i += NT) pos[op + i] = gpos[op + i];
            for (long long i = tid; i < n2 * 3; i += NT) {
                com[o3 + i] = gcom[o3 + i];
                ang[o3 + i] = gang[o3 + i];
            }
        }
    } else {
        pos = gpos;
        com = gcom;
        ang = gang;
    }
    KReg<KPL> K;
    if constexpr (KPL > 0) {
#pragma unroll
        for (int j = 0; j < KPL; j++) {
            const int k = tid + j * NT;
            if (k < P.nk) {
                K.S[j] = S[k];
                K.w[j] = P.wk[k];
                K.m[j] = P.use_exy ? kpack_xy(P.kidx + 4 * k, P.kmax) : kpack(P.kidx + 4 * k);
            } else {
                // padding k: weight 0, table entry (0, 0, 0)
                K.S[j] = make_double2(0.0, 0.0);
                K.w[j] = 0.0;
                K.m[j] = P.use_exy ? P.kmax : 0;
            }
            K.d[j] = make_double2(0.0, 0.0);
        }
    }
    __syncthreads();
    int err = 0, done = 0;
    for (int it = 0; it < nrun; it++) {
        // gcmc_dosim: a species with equal probability (gcmc.F90:240-242)
        const int s = (int)(rranf(rng) * (double)nspc);
        if (s != cur) {  // its model into the LDS view
            __syncthreads();
            for (int i = tid; i < (int)(sizeof(SpcDev) / 4); i += NT)
                reinterpret_cast<int *>(static_cast<SpcDev *>(&P))[i] = reinterpret_cast<const int *>(&P.sp[s])[i];
            if (tid == 0) sh.spc = s;
            __syncthreads();
            cur = s;
            ns = P.ns;
        }
        int N = sh.N[s];
        long long att[4], acc[4];
        for (int i = 0; i < 4; i++) {
            att[i] = sAtt[s][i];
            acc[i] = sAcc[s][i];
        }
        double dtr = sDel[s][0], drt = sDel[s][1];
        const double B = sB[s];
        double *comS = com + (long long)s * P.maxN * 3, *angS = ang + (long long)s * P.maxN * 3;
        double *posS = pos + (long long)s * P.maxN * ms3;
        double *ctrS = gctr + (long long)s * P.maxN * 3;
        const double prob = rranf(rng);
        int mv = 0;
        for (int i = 0; i < P.nmoves; i++) {
            if (prob > P.cumwt[i]) mv++;
            else break;
        }
        if (mv >= P.nmoves) mv = P.nmoves - 1;
        const int mtype = P.mtype[mv];
        const bool biased = mtype >= MH_MV_BINSERT;
        const int mbase = biased ? mtype - MH_MV_BINSERT : mtype;  // INSERT / DELETE for the biased pair
        double biasf = 1.0;  // biasfactor of rigidmoves_binsert / rigidmoves_bdelete
        int accepted = 0;
        double de = 0.0;
        bool doit = true;
        int mol = -1;
        // ---- trial generation (identical in every lane)
        double tcom[3], tang[3] = {0.0, 0.0, 0.0};
        if (CAV && mbase == MH_MV_INSERT && biased && P.cav_T > 0) {
            att[0]++;
            if (N >= P.maxN) {
                err = 2;  // capacity
                break;
            }
            const int c = cavity_trials<NT>(P, sh, R, rep, s, rng, tcom, tang, biasf);
            if (c == 5) {
                err = 5;
                break;
            }
            doit = c == 0;
        } else if (mbase == MH_MV_INSERT) {
            att[0]++;
            if (N >= P.maxN) {
                err = 2;  // capacity
                break;
            }
            if (!biased) {
                if constexpr (SL) {
                    // simcell_uniformRandPt (simcell.F90:274-295): one draw
                    // per lattice vector, summed in order
                    double v[3][3];
                    for (int k = 0; k < 3; k++) {
                        const double rk = rranf(rng);
                        for (int i = 0; i < 3; i++) v[k][i] = __dmul_rn(P.lvec[k][i], rk);
                    }
                    for (int i = 0; i < 3; i++) tcom[i] = __dadd_rn(__dadd_rn(v[0][i], v[1][i]), v[2][i]);
                } else {
                    for (int i = 0; i < 3; i++) tcom[i] = __dmul_rn(P.full[i], rranf(rng));
                }
            } else {
                bmap_com(P, rng, tcom, biasf);
            }
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
                att[mbase]++;
                for (int i = 0; i < 3; i++) {
                    tcom[i] = comS[3 * mol + i];
                    tang[i] = angS[3 * mol + i];
                }
                if (mbase == MH_MV_DELETE && biased && !(P.cav_T > 0 && !P.cav_map)) {
                    // rigidmoves_bdelete: bmap_getcellwt of the centre of mass
                    // mapped into the unit cell (simcell_maptouc)
                    const BiasRef &Bm = P.bias;
                    long long ci[3];
                    for (int i = 0; i < 3; i++) {
                        ci[i] = (long long)(maptouc(tcom[i], P.fund[i]) / Bm.step[i]);
                        if (ci[i] > Bm.ncub[i] - 1) ci[i] = Bm.ncub[i] - 1;
                    }
                    const int t = Bm.cindex[(ci[0] * Bm.ncub[1] + ci[1]) * Bm.ncub[2] + ci[2]];
                    if (t == 0) {
                        err = 4;  // the reference stops: no bias map index (bmap.F90:469-477)
                        break;
                    }
                    biasf = __dmul_rn(Bm.ncubelets, Bm.wn[t - 1]);
                }
                if (CAV && mbase == MH_MV_DELETE && biased && P.cav_T > 0) {
                    // rigidmoves_bdelete (rigidmoves.F90:1632-1643): divided
                    // by cav_prob(totmol) / cav_T at the current loading; the
                    // table is lane 0's
                    int totmol = 0;
                    for (int i = 0; i < nspc; i++) totmol += sh.N[i];
                    if (totmol >= R.cav_cap) {
                        err = 5;
                        break;
                    }
                    if (tid == 0) sh.cavpc = R.cav_avg[((long long)rep * GSPC + s) * R.cav_cap + totmol];
                    __syncthreads();
                    biasf = __ddiv_rn(biasf, __ddiv_rn(sh.cavpc, (double)P.cav_T));
                    __syncthreads();
                }
                if (mtype == MH_MV_TRANSLATE) {
                    for (int i = 0; i < 3; i++) {
                        double dv = __dmul_rn(__dsub_rn(__dmul_rn(rranf(rng), 2.0), 1.0), dtr);
                        tcom[i] = __dadd_rn(tcom[i], dv);
                    }
                } else if (mtype == MH_MV_ROTATE) {
                    double dphi = __dmul_rn(__dsub_rn(__dmul_rn(rranf(rng), 2.0), 1.0), drt);
                    double dpsi = __dmul_rn(__dsub_rn(__dmul_rn(rranf(rng), 2.0), 1.0), drt);
                    double dcos = __dmul_rn(__dsub_rn(__dmul_rn(rranf(rng), 2.0), 1.0), drt);
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
            const bool hasA = (mbase != MH_MV_DELETE);
            const bool hasB = (mbase != MH_MV_INSERT);
            // ---- site coordinates: set B = the molecule as stored; set A =
            // the trial (or, for a deletion, the stored molecule)
            const int n3 = 3 * ns;
            if (hasB && tid >= 64 - n3 && tid < 64) sh.posB[tid - (64 - n3)] = posS[mol * ms3 + tid - (64 - n3)];
            if (mbase == MH_MV_DELETE) {
                if (tid < n3) sh.posA[tid] = posS[mol * ms3 + tid];
            } else if (SL) {
                if (tid == 0) site_coords(P, tcom, tang, sh.posA, sh.rpA);
            } else if (ns == 1) {
                if (tid < 3) {
                    sh.posA[tid] = wrap_full(tcom[tid], P.full[tid]);
                    sh.rpA[tid] = tcom[tid];
                }
            } else {
                if (tid < 3) {
                    double s_, c_;
                    sincos(tang[tid], &s_, &c_);
                    sh.sc[tid][0] = s_;
                    sh.sc[tid][1] = c_;
                }
                __syncthreads();
                if (tid < n3) {
                    const int a = tid / 3, i = tid - 3 * a;
                    double m[3];
                    euler_row(i, sh.sc[0][0], sh.sc[0][1], sh.sc[1][0], sh.sc[1][1], sh.sc[2][0], sh.sc[2][1], m);
                    const double u = site_coord_unf(P, a, i, tcom, m);
                    sh.posA[tid] = wrap_full(u, P.full[i]);
                    sh.rpA[tid] = u;
                }
            }
            __syncthreads();
            double o[6];
            if (mbase == MH_MV_DELETE) {
                // the molecule's energy in the current system: its sites as set A
                move_energy<NT, PREC, KPL, SL>(P, sh, pos, N, mol, true, false, -1.0, K, S, dS, exy, o, gctr,
                                               ctrS + 3 * mol, nullptr);
            } else {
                move_energy<NT, PREC, KPL, SL>(P, sh, pos, N, mol, hasA, hasB, 1.0, K, S, dS, exy, o, gctr, nullptr,
                                               hasB ? ctrS + 3 * mol : nullptr);
            }
            const double mA = o[0], pA = o[1], mB = o[2], pB = o[3], eK = o[4];
            bool ok = o[5] != 0.0;
            if (mbase == MH_MV_INSERT) {
                double U = __dadd_rn(mA, pA);
                if (P.ewald) U = U + eK + P.molconst;
                if (ok && U > P.maxnrg) ok = false;
                if (ok) {
                    double bias = biasf / P.insdel;
                    double u = -P.rti * U;
                    u = u + log(bias);
                    u = u + B;
                    u = fmin(fmax(u, -30.0), 30.0);
                    double boltz = exp(u) * (1.0 / (double)(float)(N + 1));
                    accepted = boltz > rranf(rng);
                    if (accepted) {
                        if (tid < 3) {
                            comS[3 * N + tid] = tcom[tid];
                            angS[3 * N + tid] = tang[tid];
                        }
                        if (tid < n3) posS[N * ms3 + tid] = sh.posA[tid];
                        if (P.comcut && tid < 3) ctrS[3 * N + tid] = sh.ctr[0][tid];
                        if constexpr (CAV) cav_sites(P, R, rep, s, sh.posA, 1);
                        if (P.ewald) sk_accept<NT, KPL>(P, K, S, dS);
                        N++;
                        energy += U;
                        acc[0]++;
                        de = U;
                    }
                }
            } else if (mbase == MH_MV_DELETE) {
                if (!ok) {
                    err = 3;  // an existing molecule failed to evaluate (mcmoves.F90:599-603)
                    break;
                }
                double U = __dadd_rn(mA, pA);
                if (P.ewald) U = U - eK + P.molconst;
                double bias = biasf * P.insdel;
                double u = -P.rti * (-U);
                u = u + log(bias);
                u = u + (-B);
                u = fmin(fmax(u, -30.0), 30.0);
                double boltz = exp(u) * (1.0 * N);
                accepted = boltz > rranf(rng);
                if (accepted) {
                    const int last = N - 1;
                    // every lane finished reading the molecules (block_sum barrier)
                    if (tid < 3) {
                        comS[3 * mol + tid] = comS[3 * last + tid];
                        angS[3 * mol + tid] = angS[3 * last + tid];
                    }
                    if (tid < n3) posS[mol * ms3 + tid] = posS[last * ms3 + tid];
                    if (P.comcut && tid < 3) ctrS[3 * mol + tid] = ctrS[3 * last + tid];
                    if constexpr (CAV) cav_sites(P, R, rep, s, sh.posA, -1);
                    if (P.ewald) sk_accept<NT, KPL>(P, K, S, dS);
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
                        if (tid < 3) {
                            comS[3 * mol + tid] = tcom[tid];
                            angS[3 * mol + tid] = tang[tid];
                        }
                        if (tid < n3) posS[mol * ms3 + tid] = sh.posA[tid];
                        if (P.comcut && tid < 3) ctrS[3 * mol + tid] = sh.ctr[0][tid];
                        if constexpr (CAV) {
                            cav_sites(P, R, rep, s, sh.posB, -1);
                            cav_sites(P, R, rep, s, sh.posA, 1);
                        }
                        if (P.ewald) sk_accept<NT, KPL>(P, K, S, dS);
                        energy += dU;
                        acc[mtype]++;
                        de = dU;
                    }
                }
                // moves_adjustParams after every attempt (mcmoves.F90:460-461)
                if (mtype == MH_MV_TRANSLATE && P.adj_trans)
                    dtr = adjust_delta(dtr, (double)acc[2] / (double)att[2]);
                else if (mtype == MH_MV_ROTATE && P.adj_rot)
                    drt = adjust_delta(drt, (double)acc[3] / (double)att[3]);
            }
        }
        // cube counts changed by atomics at the device level: complete before
        // the next step's trials read them
        if constexpr (CAV)
            if (P.cav_on && accepted) __threadfence();
        if (tid == 0 && R.tr_cap > 0) {
            long long t = nst - R.tr_base;
            if (t >= 0 && t < R.tr_cap) {
                long long o = (long long)rep * R.tr_cap + t;
                R.tr_move[o] = 16 * s + mtype;
                R.tr_acc[o] = accepted;
                R.tr_N[o] = N;
                R.tr_dE[o] = de;
            }
        }
        // the species' counters back to LDS; gcmcmoves_updatenmolecs: every
        // species' statistics after every step
        if (tid == 0) {
            sh.N[s] = N;
            for (int i = 0; i < 4; i++) {
                sAtt[s][i] = att[i];
                sAcc[s][i] = acc[i];
            }
            sDel[s][0] = dtr;
            sDel[s][1] = drt;
            for (int i = 0; i < nspc; i++) {
                const double n = (double)sh.N[i];
                sSum[i] += n;
                sSum2[i] += n * n;
            }
        }
        nst++;
        done++;
        __syncthreads();
    }
    // write the replica back (a step that stopped on an error leaves its
    // species' counters as they were before it)
    __syncthreads();
    if constexpr (LDSMOL) {
        for (int s2 = 0; s2 < nspc; s2++) {
            const long long n2 = sh.N[s2], o3 = (long long)s2 * P.maxN * 3, op = (long long)s2 * P.maxN * ms3;
            for (long long i = tid; i < n2 * ms3; i += NT) gpos[op + i] = pos[op + i];
            for (long long i = tid; i < n2 * 3; i += NT) {
                gcom[o3 + i] = com[o3 + i];
                gang[o3 + i] = ang[o3 + i];
            }
        }
    }
    if (tid < nspc) {
        const int o = GSPC * rep + tid;
        R.N[o] = sh.N[tid];
        R.sumN[o] = sSum[tid];
        R.sumN2[o] = sSum2[tid];
        R.delta[2 * o] = sDel[tid][0];
        R.delta[2 * o + 1] = sDel[tid][1];
        for (int i = 0; i < 4; i++) {
            R.att[4 * o + i] = sAtt[tid][i];
            R.acc[4 * o + i] = sAcc[tid][i];
        }
    }
    if constexpr (KPL > 0) {
#pragma unroll
        for (int j = 0; j < KPL; j++) {
            const int k = tid + j * NT;
            if (k < P.nk) S[k] = K.S[j];
        }
    }
    if (tid == 0) {
        if (R.budget) R.budget[rep] = err ? 0 : R.budget[rep] - done;
        R.err[rep] = err;
        R.ix[rep] = rng.ix;
        R.iy[rep] = rng.iy;
        R.first[rep] = rng.first;
        R.iseed[rep] = rng.iseed;
        R.energy[rep] = energy;
        R.nsteps[rep] = nst;
    }
}

// Insertion energies of trial molecules (one workgroup per trial)
template <int NT, int PREC>
__global__ void __launch_bounds__(NT) k_trial(const GDev *__restrict__ Ptab, Rep R, int rep, int spc, int n,
                                               const double *__restrict__ tcom, const double *__restrict__ tang,
                                               double2 *__restrict__ dSbuf, double *__restrict__ e,
                                               uint8_t *__restrict__ okout) {
    __shared__ Shared sh;
    __shared__ GDev P;
    const int tid = threadIdx.x;
    const GDev *Pg = Ptab + (R.sys ? R.sys[rep] : 0);
    for (int i = tid; i < (int)(sizeof(GDev) / 4); i += NT)
        reinterpret_cast<int *>(&P)[i] = reinterpret_cast<const int *>(Pg)[i];
    __syncthreads();
    // the trial molecule's species as the view
    for (int i = tid; i < (int)(sizeof(SpcDev) / 4); i += NT)
        reinterpret_cast<int *>(static_cast<SpcDev *>(&P))[i] = reinterpret_cast<const int *>(&Pg->sp[spc])[i];
    if (tid < GSPC) sh.N[tid] = tid < P.nspc ? R.N[GSPC * rep + tid] : 0;
    if (tid == 0) sh.spc = spc;
    __syncthreads();
    const int t = blockIdx.x;
    if (t >= n) return;
    const int N = sh.N[spc];
    const double *pos = R.pos + (long long)rep * P.nspc * P.maxN * P.nsmax * 3;
    const double2 *S = R.S + (long long)rep * R.kstride;
    const double *ctrg = R.ctr + (long long)rep * P.nspc * P.maxN * 3;
    if (tid == 0) site_coords(P, tcom + 3 * t, tang + 3 * t, sh.posA, sh.rpA);
    __syncthreads();
    double o[6];
    KReg<0> K;
    if (P.slanted)
        move_energy<NT, PREC, 0, true>(P, sh, pos, N, -1, true, false, 1.0, K, S, dSbuf + (long long)t * P.nk,
                                       nullptr, o, ctrg, nullptr, nullptr);
    else
        move_energy<NT, PREC, 0>(P, sh, pos, N, -1, true, false, 1.0, K, S, dSbuf + (long long)t * P.nk, nullptr, o,
                                 ctrg, nullptr, nullptr);
    if (tid == 0) {
        double U = __dadd_rn(o[0], o[1]);
        if (P.ewald) U = U + o[4] + P.molconst;
        bool ok = o[5] != 0.0 && !(U > P.maxnrg);
        e[t] = U;
        okout[t] = ok ? 1 : 0;
    }
}

// S(k) of every replica from the coordinates (FP64), and total energies
__global__ void k_sk_recompute(const GDev *__restrict__ Ptab, Rep R) {
    const int rep = blockIdx.y;
    if (R.sys && R.sys[rep] < 0) return;
    const GDev &P = Ptab[R.sys ? R.sys[rep] : 0];
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= P.nk) return;
    const double *pos = R.pos + (long long)rep * P.nspc * P.maxN * P.nsmax * 3;
    const int8_t *ki = P.kidx + 4 * k;
    double kv[3] = {((double)ki[0] / P.full[0]) * kTwoPiR, ((double)ki[1] / P.full[1]) * kTwoPiR,
                    ((double)ki[2] / P.full[2]) * kTwoPiR};
    double re = 0.0, im = 0.0;
    for (int s = 0; s < P.nspc; s++) {
        const SpcDev &Q = P.sp[s];
        const int N = R.N[GSPC * rep + s];
        const double *ps = pos + (long long)s * P.maxN * P.nsmax * 3;
        for (int m = 0; m < N; m++)
            for (int a = 0; a < Q.ns; a++) {
                if (Q.q[a] == 0.0) continue;
                const double *r = ps + ((long long)m * P.nsmax + a) * 3;
                double arg = kv[0] * r[0] + kv[1] * r[1] + kv[2] * r[2];
                double s_, c_;
                sincos(arg, &s_, &c_);
                re += Q.q[a] * c_;
                im += Q.q[a] * s_;
            }
    }
    R.S[(long long)rep * R.kstride + k] = make_double2(re, im);
}

// total energy per replica: sum over molecules of (maps + pairs with the
// molecules of earlier species and the earlier molecules of its species)
// + reciprocal + sum of N_s molconst_s; one workgroup per replica
template <int NT>
__global__ void __launch_bounds__(NT) k_total(const GDev *__restrict__ Ptab, Rep R, double *__restrict__ out) {
    __shared__ GDev P;
    __shared__ double red[NT / 64 * 2];
    const int tid = threadIdx.x;
    if (R.sys && R.sys[blockIdx.x] < 0) {
        if (tid == 0) out[blockIdx.x] = 0.0;
        return;
    }
    const GDev *Pg = Ptab + (R.sys ? R.sys[blockIdx.x] : 0);
    for (int i = tid; i < (int)(sizeof(GDev) / 4); i += NT)
        reinterpret_cast<int *>(&P)[i] = reinterpret_cast<const int *>(Pg)[i];
    __syncthreads();
    const int rep = blockIdx.x;
    const long long ms3 = 3LL * P.nsmax, per = (long long)P.maxN * ms3;
    const double *pos = R.pos + (long long)rep * P.nspc * per;
    const double *ctr = R.ctr + (long long)rep * P.nspc * P.maxN * 3;
    int Ns[GSPC], ntot = 0;
    for (int s = 0; s < GSPC; s++) {
        Ns[s] = s < P.nspc ? R.N[GSPC * rep + s] : 0;
        ntot += Ns[s];
    }
    double e = 0.0, bad = 0.0, nconst = 0.0;
    for (int idx = tid; idx < ntot; idx += NT) {
        int s = 0, m = idx;
        while (m >= Ns[s]) m -= Ns[s++];
        const SpcDev &Q = P.sp[s];
        const double *pm = pos + s * per + m * ms3;
        for (int a = 0; a < Q.ns; a++) {
            double em;
            if (!site_map_energy(P, Q, a, pm + 3 * a, em)) bad = 1.0;
            e += em;
        }
        for (int s2 = 0; s2 <= s; s2++) {
            const SpcDev &O = P.sp[s2];
            const int nj = s2 < s ? Ns[s2] : m;
            for (int j = 0; j < nj; j++) {
                const double *pj = pos + s2 * per + j * ms3;
                bool ljo = true, clo = true;
                if (P.comcut) {
                    double sc[3];
                    min_image(P, ctr + 3 * ((long long)s * P.maxN + m), ctr + 3 * ((long long)s2 * P.maxN + j), sc,
                              P.slanted != 0);
                    const double d2 =
                        __dadd_rn(__dadd_rn(__dmul_rn(sc[0], sc[0]), __dmul_rn(sc[1], sc[1])), __dmul_rn(sc[2], sc[2]));
                    ljo = !(P.cc_lj[s][s2] > 0.0 && d2 > P.cc_lj[s][s2]);
                    clo = !(P.cc_cl[s][s2] > 0.0 && d2 > P.cc_cl[s][s2]);
                }
                for (int a = 0; a < Q.ns; a++)
                    for (int b = 0; b < O.ns; b++) {
                        double sp[3];
                        min_image(P, pm + 3 * a, pj + 3 * b, sp, P.slanted != 0);
                        if (!pair_energy64(P, Q.st[a] * P.nst + O.st[b], Q.q[a], O.q[b], sp[0], sp[1], sp[2], e, ljo,
                                           clo))
                            bad = 1.0;
                    }
            }
        }
        nconst += Q.molconst;
    }
    if (P.ewald) {
        e += nconst;
        for (int k = tid; k < P.nk; k += NT) {
            double2 s = R.S[(long long)rep * R.kstride + k];
            e += P.wk[k] * (s.x * s.x + s.y * s.y);
        }
    }
    double v[2] = {e, bad};
    block_sum<NT, 2>(v, red);
    if (tid == 0) out[rep] = (v[1] != 0.0) ? NAN : v[0];
}

// uncompacted copy of a map: out[n][8] = vals[ptr[n] - 1][8], V = NaN where
// ptr[n] == 0
__global__ void k_densify(const int32_t *__restrict__ ptr, const float *__restrict__ vals, long long nodes,
                          float *__restrict__ out) {
    const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nodes) return;
    const int p = ptr[n];
    float4 a = make_float4(NAN, 0.0f, 0.0f, 0.0f), b = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (p > 0) {
        const float4 *r = reinterpret_cast<const float4 *>(vals + 8 * (long long)(p - 1));
        a = r[0];
        b = r[1];
    }
    float4 *o = reinterpret_cast<float4 *>(out + 8 * n);
    o[0] = a;
    o[1] = b;
}

}  // namespace mh

using namespace mh;


// One molecular system of a batch: sorbate model, maps, cell, k vectors.
// A plain batch has one system; a slot pool holds one per admitted request.
struct SysHost {
    bool used = false;
    GDev H;
    DBuf wk, kidx;
    double volume = 0.0;
    std::vector<int> slots;  // replicas (slots) of this system, admit order
};

// cavity list of the stored configurations (cavitylist_initSorbInfo,
// cavitylist.F90:282-340): cleared, then every site of the listed species
// counted in its cube (two launches: the clear completes first)
__global__ void k_cav_clear(Rep R) {
    int *c = R.cube + (long long)blockIdx.y * R.cube_stride;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < R.cube_stride;
         i += (long long)gridDim.x * blockDim.x)
        c[i] = 0;
}

__global__ void k_cav_fill(const GDev *__restrict__ Ptab, Rep R) {
    const GDev &P = Ptab[0];
    const int rep = blockIdx.x;
    const long long slots = (long long)P.nspc * P.maxN, ms3 = 3LL * P.nsmax;
    const double *gpos = R.pos + (long long)rep * slots * ms3;
    for (int s = 0; s < P.nspc; s++) {
        if (!(P.cav_mask >> s & 1)) continue;
        const int ns = P.sp[s].ns, n = R.N[GSPC * rep + s];
        for (int i = threadIdx.x; i < n * ns; i += blockDim.x) {
            const int m = i / ns, a = i - m * ns;
            const double *r = gpos + ((long long)s * P.maxN + m) * ms3 + 3 * a;
            atomicAdd(R.cube + (long long)rep * R.cube_stride + cav_cube(P, r), 1);
        }
    }
}

struct mh_gcmc {
    mh_ctx *ctx = nullptr;
    int nrep = 0, maxN = 0, prec = 0, threads = 256;
    int nspc = 1;           // species of the (single-system) batch
    int sel = 0;            // species the per-species calls address
    int ns = 1;             // max sites over the systems (LDS layout)
    int nk = 0;             // max k vectors over the systems
    long long kstride = 0;  // S(k) entries per replica
    int kpl = 0;            // k vectors per lane held in registers (0: S(k) in HBM)
    bool ldsmol = false;    // molecules held in LDS during a launch
    bool slanted = false;   // single system in a slanted cell
    size_t lds_bytes = 0;   // dynamic LDS of the launch
    bool exy_off = false;   // autotuner: e^{i(kx x + ky y)} tables disabled
    bool pool = false;      // slot pool (per-replica system and step budget)
    std::vector<SysHost> sys;
    std::vector<int> slot_sys;  // host mirror of sysidx (pool)
    DBuf dP;                    // GDev[sys.size()]
    DBuf sysidx, budget;        // pool only
    DBuf N, err, ix, iy, first, iseed, B, energy, sumN, sumN2, nsteps, att, acc;
    DBuf com, ang, pos, ctr, S, dS, delta;
    DBuf tr_move, tr_acc, tr_N, tr_dE;
    // cavity bias (mh_gcmc_set_cavity): cube counts, cav_prob tables, the
    // tabulation; launches use at least cav_tmax threads (one trial per lane)
    DBuf cube, cav_avg, cav_cnt, cav_tab;
    long long ncube = 0;
    int cav_cap = 0, cav_tmax = 0;
    int tr_cap = 0;
    long long tr_base = 0;
    // hipGraph of one cycle's launch (mh_gcmc_run_cycles) and what it baked in
    hipGraphExec_t gexec = nullptr;
    struct GraphKey {
        Rep r;
        const void *P;
        int nsteps, threads, prec, kpl;
        bool ldsmol, exy_off;
        size_t lds_bytes;
    } gkey;
    ~mh_gcmc() {
        if (gexec) (void)hipGraphExecDestroy(gexec);
    }
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
        r.sumN2 = sumN2.as<double>();
        r.nsteps = nsteps.as<long long>();
        r.att = att.as<long long>();
        r.acc = acc.as<long long>();
        r.com = com.as<double>();
        r.ang = ang.as<double>();
        r.pos = pos.as<double>();
        r.ctr = ctr.as<double>();
        r.delta = delta.as<double>();
        r.S = S.as<double2>();
        r.dS = dS.as<double2>();
        r.tr_move = tr_move.as<int>();
        r.tr_acc = tr_acc.as<int>();
        r.tr_N = tr_N.as<int>();
        r.tr_dE = tr_dE.as<double>();
        r.tr_cap = tr_cap;
        r.tr_base = tr_base;
        r.sys = pool ? sysidx.as<int>() : nullptr;
        r.budget = pool ? budget.as<int>() : nullptr;
        r.kstride = kstride;
        r.cube = cube.as<int>();
        r.cube_stride = ncube;
        r.cav_avg = cav_avg.as<double>();
        r.cav_cnt = cav_cnt.as<long long>();
        r.cav_cap = cav_cap;
        return r;
    }
    const GDev &sys_of(int rep) const { return sys[pool ? slot_sys[rep] : 0].H; }
};

static double erfc_as_host(double x) {
    const double a1 = (double)0.254829592f, a2 = (double)-0.284496736f, a3 = (double)1.421413741f,
                 a4 = (double)-1.453152027f, a5 = (double)1.061405429f, p = (double)0.3275911f;
    double t = 1.0 / (1.0 + p * x);
    return t * (a1 + t * (a2 + t * (a3 + t * (a4 + t * a5)))) * std::exp(-x * x);
}

// register tiles of k vectors per lane compiled for each workgroup size
// (larger tiles spill at that size); 0 = S(k) kept in HBM
static bool kpl_allowed(int threads, int k) {
    switch (threads) {
    case 64: return k == 8 || k == 16;
    case 128: return k == 4 || k == 8 || k == 16;
    case 256: return k == 2 || k == 4 || k == 8 || k == 16;
    case 512: return k == 2 || k == 4;
    default: return false;
    }
}

// smallest register tile of k vectors per lane that holds nk at NT lanes
static int pick_kpl(int nk, int threads) {
    if (nk == 0) return 0;
    const int need = (nk + threads - 1) / threads;
    for (int k : {2, 4, 8, 16})
        if (need <= k && kpl_allowed(threads, k)) return k;
    return 0;
}

template <int PREC, int NT, int KPL>
static void launch_lds(mh_gcmc *g, const GDev *P, const Rep &r, int nsteps) {
    (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&k_gcmc<NT, PREC, KPL, true>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)g->lds_bytes);
    hipLaunchKernelGGL((k_gcmc<NT, PREC, KPL, true>), dim3(g->nrep), dim3(NT), g->lds_bytes, g->ctx->stream, P, r,
                       nsteps);
}

template <int PREC, int NT>
static void launch_nt(mh_gcmc *g, const GDev *P, const Rep &r, int nsteps) {
    if (g->ncube) {  // cavity bias: its own instantiation (plan() keeps S(k) and molecules in HBM)
        hipLaunchKernelGGL((k_gcmc<NT, PREC, 0, false, true>), dim3(g->nrep), dim3(NT), 0, g->ctx->stream, P, r,
                           nsteps);
        return;
    }
    if (g->slanted) {  // slanted cell: its own instantiation (no k vectors)
        if (g->ldsmol) launch_lds<PREC, NT, -1>(g, P, r, nsteps);
        else hipLaunchKernelGGL((k_gcmc<NT, PREC, -1, false>), dim3(g->nrep), dim3(NT), 0, g->ctx->stream, P, r, nsteps);
        return;
    }
    if (!g->ldsmol) {
        hipLaunchKernelGGL((k_gcmc<NT, PREC, 0, false>), dim3(g->nrep), dim3(NT), 0, g->ctx->stream, P, r, nsteps);
        return;
    }
    // only the tiles kpl_allowed() admits for NT are instantiated
    constexpr bool k2 = NT == 256 || NT == 512, k4 = NT >= 128 && NT <= 512, k8 = NT <= 256, k16 = NT <= 256;
    switch (g->kpl) {
    case 2: if constexpr (k2) launch_lds<PREC, NT, 2>(g, P, r, nsteps); break;
    case 4: if constexpr (k4) launch_lds<PREC, NT, 4>(g, P, r, nsteps); break;
    case 8: if constexpr (k8) launch_lds<PREC, NT, 8>(g, P, r, nsteps); break;
    case 16: if constexpr (k16) launch_lds<PREC, NT, 16>(g, P, r, nsteps); break;
    default: launch_lds<PREC, NT, 0>(g, P, r, nsteps); break;
    }
}

template <int PREC>
static void launch_gcmc(mh_gcmc *g, int nsteps) {
    Rep r = g->rep();
    const GDev *P = g->dP.as<GDev>();
    switch (g->threads) {
    case 64: launch_nt<PREC, 64>(g, P, r, nsteps); break;
    case 128: launch_nt<PREC, 128>(g, P, r, nsteps); break;
    case 512: launch_nt<PREC, 512>(g, P, r, nsteps); break;
    case 1024: launch_nt<PREC, 1024>(g, P, r, nsteps); break;
    default: launch_nt<PREC, 256>(g, P, r, nsteps); break;
    }
}

static void launch_any(mh_gcmc *g, int nsteps) {
    if (g->prec == MH_PREC_FP32) launch_gcmc<MH_PREC_FP32>(g, nsteps);
    else launch_gcmc<MH_PREC_FP64>(g, nsteps);
}

// dense copy of a map for the GCMC lookups, built once per map
int mh::map_dense(mh_ctx *ctx, mh_map *m, const float **out) {
    const long long nodes = (long long)m->npnts[0] * m->npnts[1] * m->npnts[2];
    if (!m->dense.p) {
        int rc;
        if ((rc = m->dense.alloc(sizeof(float) * 8 * nodes))) return rc;
        hipLaunchKernelGGL(k_densify, dim3((unsigned)((nodes + 255) / 256)), dim3(256), 0, ctx->stream,
                           m->ptr.as<int32_t>(), m->vals.as<float>(), nodes, m->dense.as<float>());
        MH_HIP(hipGetLastError());
        MH_HIP(hipStreamSynchronize(ctx->stream));
    }
    *out = m->dense.as<float>();
    return MH_OK;
}

// Launch plan over the live systems: molecules in dynamic LDS when they fit
// (gfx950: 160 KB per workgroup, one replica per CU), the k-vector register
// tile, and per system the e^{i(kx x + ky y)} tables behind the molecules;
// uploads the GDev table.
static int plan(mh_gcmc *g) {
    int dev_max = 65536;
    (void)hipDeviceGetAttribute(&dev_max, hipDeviceAttributeMaxSharedMemoryPerBlock, g->ctx->device);
    const long long budget = (long long)dev_max - (long long)sizeof(Shared) - (long long)sizeof(GDev) - 1024;
    int ns = 1, nk = 0;
    for (auto &s : g->sys)
        if (s.used) {
            ns = std::max(ns, s.H.nsmax);
            nk = std::max(nk, s.H.nk);
        }
    g->ns = std::max(g->ns, ns);
    g->nk = nk;
    const long long need = (long long)g->nspc * g->maxN * (3 * g->ns + 6) * (long long)sizeof(double);
    g->ldsmol = !g->ncube && need <= budget;
    g->lds_bytes = g->ldsmol ? (size_t)need : 0;
    g->kpl = g->ldsmol ? pick_kpl(nk, g->threads) : 0;
    const long long off = (g->lds_bytes + 255) / 256 * 256, ent = (g->prec == MH_PREC_FP64) ? 16 : 8;
    long long extra = 0;
    std::vector<GDev> tab(g->sys.size());
    for (size_t i = 0; i < g->sys.size(); i++) {
        GDev &H = g->sys[i].H;
        if (g->sys[i].used) {
            const long long K = H.kmax;
            const long long exy = H.ewald ? 2LL * H.nsmax * (K + 1) * (2 * K + 1) * ent : 0;
            H.use_exy = (!g->exy_off && g->kpl > 0 && exy > 0 && off + exy <= budget) ? 1 : 0;
            H.exy_off = H.use_exy ? (int)off : 0;
            if (H.use_exy) extra = std::max(extra, exy);
            tab[i] = H;
            tab[i].nsmax = g->ns;  // slot layout of the batch
        } else {
            std::memset(&tab[i], 0, sizeof(GDev));
        }
    }
    if (extra) g->lds_bytes = (size_t)(off + extra);
    MH_HIP(hipMemcpy(g->dP.p, tab.data(), sizeof(GDev) * tab.size(), hipMemcpyHostToDevice));
    return MH_OK;
}

// BINSERT / BDELETE draw from the species' bias map unless they are
// cavity biased without one (CAVITY_BIAS, set by mh_gcmc_set_cavity)
static int moves_ready(const GDev &H, int nspc, const char *who) {
    for (int s = 0; s < nspc; s++) {
        const SpcDev &Q = H.sp[s];
        for (int k = 0; k < Q.nmoves; k++) {
            if (Q.mtype[k] < MH_MV_BINSERT || Q.bias.cube || (Q.cav_T > 0 && !Q.cav_map)) continue;
            set_error(std::string(who) + ": BINSERT / BDELETE need a bias map (or CAVITY_BIAS)");
            return MH_EINVAL;
        }
    }
    return MH_OK;
}

// Sorbate model, maps, cell and k vectors of one system.
static int build_system(mh_ctx *ctx, const mh_gcmc_params *p, int maxN, const double fund[3], const double full[3],
                        SysHost &S, const char *who, bool ctx_cell) {
    std::string w(who);
    MH_CHECK(p->nsites >= 1 && p->nsites <= GS, MH_EINVAL, w + ": nsites out of range [1,8]");
    MH_CHECK(p->ntypes >= 1 && p->ntypes <= 8, MH_EINVAL, w + ": ntypes out of range [1,8]");
    MH_CHECK(p->nmoves >= 1 && p->nmoves <= 4, MH_EINVAL, w + ": nmoves out of range [1,4]");
    MH_CHECK(p->temperature > 0.0, MH_EINVAL, w + ": temperature must be > 0");
    MH_CHECK(fund[0] > 0.0 && fund[1] > 0.0 && fund[2] > 0.0 && full[0] > 0.0 && full[1] > 0.0 && full[2] > 0.0,
             MH_ESTATE, w + ": cell edges must be > 0 (upload the framework first)");
    if (p->ewald) {
        MH_CHECK(p->ewald_kmax >= 1 && p->ewald_kmax <= GKMAX, MH_EINVAL, w + ": ewald_kmax out of range [1,15]");
        MH_CHECK(p->ewald_kappa > 0.0, MH_EINVAL, w + ": ewald_kappa must be > 0");
    }
    if (p->coul) {
        MH_CHECK(!p->ewald, MH_EINVAL, w + ": plain Coulomb and Ewald are exclusive");
        MH_CHECK(p->coul_hicut >= 0.0 && p->coul_locut >= 0.0, MH_EINVAL, w + ": Coulomb cutoffs must be >= 0");
    }
    GDev &H = S.H;
    std::memset(&H, 0, sizeof H);
    H.ns = p->nsites;
    for (int a = 0; a < H.ns; a++) {
        for (int i = 0; i < 3; i++) H.ref[a][i] = p->ref[a][i];
        H.st[a] = p->site_type[a];
        H.q[a] = p->site_q[a];
        MH_CHECK(p->site_type[a] >= 0 && p->site_type[a] < p->ntypes, MH_EINVAL, w + ": site_type out of range");
        const mh_map *maps[2] = {p->site_pmap[a], p->site_emap[a]};
        for (int k = 0; k < 2; k++) {
            if (!maps[k]) continue;
            MapRef &m = k == 0 ? H.pm[a] : H.em[a];
            (k == 0 ? H.has_p : H.has_e)[a] = 1;
            int rc = map_dense(ctx, const_cast<mh_map *>(maps[k]), &m.dense);
            if (rc) return rc;
            for (int i = 0; i < 3; i++) {
                m.n[i] = maps[k]->npnts[i];
                m.step[i] = maps[k]->step[i];
                m.ell[i] = fund[i];
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
        H.full[i] = full[i];
        H.fund[i] = fund[i];
    }
    H.coul = p->coul ? 1 : 0;
    if (H.coul) {
        for (int t = 0; t < p->ntypes * p->ntypes; t++) H.con[t] = p->coul_on[t] ? 1 : 0;
        H.cfac = 331.5 * p->coul_shield;
        H.ch2 = p->coul_hicut * p->coul_hicut;
        H.cshift = p->coul_hicut > 0.0 ? 1.0 / p->coul_hicut : 0.0;
        H.cl2 = p->coul_locut * p->coul_locut;
    }
    // centre-of-mass cutoffs of species 0 towards each species (the rows of
    // the other species of a mixture are filled by mh_gcmc_create_mix)
    for (int s2 = 0; s2 < GSPC; s2++) {
        MH_CHECK(p->lj_comcut[s2] >= 0.0 && p->coul_comcut[s2] >= 0.0, MH_EINVAL,
                 w + ": centre-of-mass cutoffs must be >= 0");
        H.cc_lj[0][s2] = p->lj_comcut[s2] * p->lj_comcut[s2];
        H.cc_cl[0][s2] = p->coul_comcut[s2] * p->coul_comcut[s2];
        if (p->lj_comcut[s2] > 0.0 || p->coul_comcut[s2] > 0.0) H.comcut = 1;
    }
    H.ewald = p->ewald ? 1 : 0;
    H.kmax = p->ewald ? p->ewald_kmax : 0;
    H.kappa = p->ewald_kappa;
    H.rc2 = p->ewald_rcut * p->ewald_rcut;
    // k vectors of the half space (mx > 0, or mx = 0 & my > 0, or mx = my = 0 & mz > 0)
    std::vector<double> wv;
    std::vector<int8_t> ki;
    double V = H.full[0] * H.full[1] * H.full[2];
    if (ctx_cell && ctx->slanted) {
        // the simulation cell's lattice vectors (fundcell_genlatticevects,
        // fundcell.F90:174-208, double trigonometry of the angles) and its
        // volume (fundcell_genvolume: (l1 x l2) . l3, vector.F90:405-429)
        MH_CHECK(!p->ewald, MH_EINVAL, w + ": sorbate Ewald sums are not offered in slanted cells");
        MH_CHECK(!p->bias, MH_EINVAL, w + ": biased insertion is not offered in slanted cells");
        H.slanted = 1;
        for (int i = 0; i < 9; i++) {
            H.sc.U[i] = ctx->U[i];
            H.sc.S[i] = ctx->S[i];
        }
        const double degtorad = 3.1415927 / 180.0;
        const double abc = ctx->ang_deg[0] * degtorad, aac = ctx->ang_deg[1] * degtorad,
                     aab = ctx->ang_deg[2] * degtorad;
        double ell[3];
        for (int i = 0; i < 3; i++) ell[i] = fund[i] * (double)std::lround(full[i] / fund[i]);
        double (*L)[3] = H.lvec;
        L[0][0] = ell[0];
        L[0][1] = 0.0;
        L[0][2] = 0.0;
        L[1][0] = ell[1] * std::cos(aab);
        L[1][1] = ell[1] * std::sin(aab);
        L[1][2] = 0.0;
        L[2][0] = ell[2] * std::cos(aac);
        L[2][1] = ell[2] * (std::cos(abc) - std::cos(aac) * std::cos(aab)) / std::sin(aab);
        double m2 = 0.0;
        m2 = m2 + L[2][0] * L[2][0];
        m2 = m2 + L[2][1] * L[2][1];
        m2 = m2 + 0.0 * 0.0;
        L[2][2] = std::sqrt(ell[2] * ell[2] - m2);
        const double cx = L[0][1] * L[1][2] - L[0][2] * L[1][1];
        const double cy = -(L[0][0] * L[1][2] - L[0][2] * L[1][0]);
        const double cz = L[0][0] * L[1][1] - L[0][1] * L[1][0];
        V = cx * L[2][0] + cy * L[2][1] + cz * L[2][2];
    }
    S.volume = V;
    if (H.ewald) {
        int K = H.kmax;
        for (int mx = 0; mx <= K; mx++)
            for (int my = -K; my <= K; my++)
                for (int mz = -K; mz <= K; mz++) {
                    if (mx == 0 && (my < 0 || (my == 0 && mz <= 0))) continue;
                    double kx = ((double)mx / H.full[0]) * kTwoPiR, ky = ((double)my / H.full[1]) * kTwoPiR,
                           kz = ((double)mz / H.full[2]) * kTwoPiR;
                    double k2 = kx * kx + ky * ky + kz * kz;
                    wv.push_back((4.0 * M_PI / V) * std::exp(-k2 / (4.0 * H.kappa * H.kappa)) / k2 * kTokcalG);
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
    H.nk = (int)wv.size();
    H.nmoves = p->nmoves;
    double wsum = 0.0, wi = 0.0, wd = 0.0;
    for (int i = 0; i < p->nmoves; i++) wsum += p->move_weight[i];
    double c = 0.0;
    for (int i = 0; i < p->nmoves; i++) {
        const int mt = p->move_type[i];
        MH_CHECK(mt >= 0 && mt <= MH_MV_BDELETE, MH_EINVAL, w + ": bad move type");
        H.mtype[i] = mt;
        double wn = p->move_weight[i] / wsum;
        c = c + wn;
        H.cumwt[i] = c;
        // mcmoveset_setInsDelRatio: the basic INSERT / DELETE tags
        if (mt == MH_MV_INSERT || mt == MH_MV_BINSERT) {
            MH_CHECK(wi == 0.0, MH_EINVAL, w + ": one insertion move per move set");
            wi = wn;
        }
        if (mt == MH_MV_DELETE || mt == MH_MV_BDELETE) {
            MH_CHECK(wd == 0.0, MH_EINVAL, w + ": one deletion move per move set");
            wd = wn;
        }
    }
    for (int i = 0; i < 3; i++) H.ncell[i] = (int)std::lround(full[i] / fund[i]);
    if (p->bias) {
        const mh_bmap *bm = p->bias;
        MH_CHECK(bm->ctx == ctx, MH_EINVAL, w + ": the bias map belongs to another context");
        for (int i = 0; i < 3; i++)
            MH_CHECK(std::fabs(bm->eff[i] - fund[i]) <= 1e-9 * fund[i], MH_EINVAL,
                     w + ": the bias map's cell differs from the fundamental cell");
        H.bias = bias_ref(bm);
    }
    H.insdel = (wi > 0.0 && wd > 0.0) ? wi / wd : 1.0;  // mcmoveset.F90:424-430
    H.dtrans = p->delta_trans;
    H.drot = p->delta_rot;
    H.adj_trans = p->adjust_trans != 0;
    H.adj_rot = p->adjust_rot != 0;
    H.maxnrg = p->max_nrg;
    H.rti = 1.0 / ((8.314e-3 / 4.184) * p->temperature);
    H.prec = p->precision;
    H.maxN = maxN;
    H.inv_ns = 1.0 / H.ns;
    H.inv_nss = 1.0 / (H.ns * H.ns);
    int rc;
    if ((rc = S.wk.alloc(sizeof(double) * (H.nk + 1))) || (rc = S.kidx.alloc(4 * (H.nk + 1)))) return rc;
    if (H.nk) {
        MH_HIP(hipMemcpy(S.wk.p, wv.data(), sizeof(double) * H.nk, hipMemcpyHostToDevice));
        MH_HIP(hipMemcpy(S.kidx.p, ki.data(), 4 * H.nk, hipMemcpyHostToDevice));
    }
    H.wk = S.wk.as<double>();
    H.kidx = S.kidx.as<int8_t>();
    H.nspc = 1;
    H.nsmax = H.ns;
    H.sp[0] = static_cast<const SpcDev &>(H);
    S.used = true;
    return MH_OK;
}

// replica state buffers for nrep replicas
static int alloc_replicas(mh_gcmc *g) {
    const long long R = g->nrep, slots = (long long)g->nspc * g->maxN, Q = (long long)GSPC * g->nrep;
    int rc;
    if ((rc = g->N.alloc(4 * Q)) || (rc = g->err.alloc(4 * R)) || (rc = g->ix.alloc(4 * R)) ||
        (rc = g->iy.alloc(4 * R)) || (rc = g->first.alloc(4 * R)) || (rc = g->iseed.alloc(4 * R)) ||
        (rc = g->B.alloc(8 * Q)) || (rc = g->energy.alloc(8 * R)) || (rc = g->sumN.alloc(8 * Q)) ||
        (rc = g->sumN2.alloc(8 * Q)) || (rc = g->nsteps.alloc(8 * R)) || (rc = g->att.alloc(32 * Q)) ||
        (rc = g->acc.alloc(32 * Q)) || (rc = g->delta.alloc(16 * Q)) ||
        (rc = g->com.alloc(sizeof(double) * 3 * R * slots)) || (rc = g->ang.alloc(sizeof(double) * 3 * R * slots)) ||
        (rc = g->pos.alloc(sizeof(double) * 3 * R * slots * g->ns)) ||
        (rc = g->ctr.alloc(sizeof(double) * 3 * R * slots)) ||
        (rc = g->S.alloc(sizeof(double2) * (R * g->kstride + 1))) ||
        (rc = g->dS.alloc(sizeof(double2) * (R * g->kstride + 1))) || (rc = g->tr_move.alloc(4)) ||
        (rc = g->tr_acc.alloc(4)) || (rc = g->tr_N.alloc(4)) || (rc = g->tr_dE.alloc(8)))
        return rc;
    MH_HIP(hipMemset(g->S.p, 0, g->S.n));
    MH_HIP(hipMemset(g->err.p, 0, g->err.n));
    MH_HIP(hipMemset(g->N.p, 0, g->N.n));
    return MH_OK;
}

// fresh chain state of replicas reps: no molecules, RNG from seeds, B per
// species from its fugacity fug[i][s] and the system volume
// (gcmc_genfuglist, gcmc.F90:303-319), the species' displacements
static int init_replicas(mh_gcmc *g, const std::vector<int> &reps, const double *fug, const int32_t *seeds,
                         double volume, double T, const GDev &H) {
    const double sivolume = volume * 1.0e-30 * 6.0221367e23;
    const int nspc = H.nspc;
    for (size_t i = 0; i < reps.size(); i++) {
        for (int sp = 0; sp < nspc; sp++)
            MH_CHECK(fug[i * nspc + sp] > 0.0, MH_EINVAL, "gcmc: fugacity must be > 0");
        const int r = reps[i];
        const int32_t zero = 0, one = 1;
        const double zd = 0.0;
        const long long zl = 0;
        int zN[GSPC] = {0, 0, 0, 0};
        double hB[GSPC] = {0, 0, 0, 0}, zD[GSPC] = {0, 0, 0, 0}, d2[2 * GSPC];
        long long zl4[4 * GSPC];
        for (int k = 0; k < 4 * GSPC; k++) zl4[k] = 0;
        for (int sp = 0; sp < GSPC; sp++) {
            if (sp < nspc)
                hB[sp] = std::log(fug[i * nspc + sp] * 1.0e3 * sivolume / (8.314 * T)) - std::log(1.0);
            d2[2 * sp] = sp < nspc ? H.sp[sp].dtrans : 0.0;
            d2[2 * sp + 1] = sp < nspc ? H.sp[sp].drot : 0.0;
        }
        MH_HIP(hipMemcpy(g->N.as<int>() + GSPC * r, zN, 4 * GSPC, hipMemcpyHostToDevice));
        MH_HIP(hipMemcpy(g->err.as<int>() + r, &zero, 4, hipMemcpyHostToDevice));
        MH_HIP(hipMemcpy(g->ix.as<int32_t>() + r, &zero, 4, hipMemcpyHostToDevice));
        MH_HIP(hipMemcpy(g->iy.as<int32_t>() + r, &zero, 4, hipMemcpyHostToDevice));
        MH_HIP(hipMemcpy(g->first.as<int32_t>() + r, &one, 4, hipMemcpyHostToDevice));
        MH_HIP(hipMemcpy(g->iseed.as<int32_t>() + r, &seeds[i], 4, hipMemcpyHostToDevice));
        MH_HIP(hipMemcpy(g->B.as<double>() + GSPC * r, hB, 8 * GSPC, hipMemcpyHostToDevice));
        MH_HIP(hipMemcpy(g->energy.as<double>() + r, &zd, 8, hipMemcpyHostToDevice));
        MH_HIP(hipMemcpy(g->sumN.as<double>() + GSPC * r, zD, 8 * GSPC, hipMemcpyHostToDevice));
        MH_HIP(hipMemcpy(g->sumN2.as<double>() + GSPC * r, zD, 8 * GSPC, hipMemcpyHostToDevice));
        MH_HIP(hipMemcpy(g->nsteps.as<long long>() + r, &zl, 8, hipMemcpyHostToDevice));
        MH_HIP(hipMemcpy(g->att.as<long long>() + 4 * GSPC * r, zl4, 32 * GSPC, hipMemcpyHostToDevice));
        MH_HIP(hipMemcpy(g->acc.as<long long>() + 4 * GSPC * r, zl4, 32 * GSPC, hipMemcpyHostToDevice));
        MH_HIP(hipMemcpy(g->delta.as<double>() + 2 * GSPC * r, d2, 16 * GSPC, hipMemcpyHostToDevice));
        if (g->kstride)
            MH_HIP(hipMemset(g->S.as<double2>() + (long long)r * g->kstride, 0, sizeof(double2) * g->kstride));
    }
    return MH_OK;
}

extern "C" {

int mh_gcmc_create_mix(mh_ctx *ctx, const mh_gcmc_params *spc, int nspc, int nrep, const double *fugacity_kPa,
                       const int32_t *seeds, int maxN, mh_gcmc **out) {
    MH_CHECK(ctx && spc && fugacity_kPa && seeds && out, MH_EINVAL, "mh_gcmc_create: bad arguments");
    MH_CHECK(nspc >= 1 && nspc <= GSPC, MH_EINVAL, "mh_gcmc_create: nspc out of range [1,4]");
    MH_CHECK(nrep >= 1 && nrep <= 65535, MH_EINVAL, "mh_gcmc_create: nrep out of range");
    MH_CHECK(maxN >= 1 && maxN <= (1 << 20), MH_EINVAL, "mh_gcmc_create: maxN out of range");
    const mh_gcmc_params *p = &spc[0];
    MH_CHECK(p->precision == MH_PREC_FP64 || p->precision == MH_PREC_FP32, MH_EINVAL,
             "mh_gcmc_create: bad precision");
    for (int sp = 0; sp < nspc; sp++)
        for (int a = 0; a < spc[sp].nsites && a < MH_MAX_SITES; a++)
            MH_CHECK((!spc[sp].site_pmap[a] || spc[sp].site_pmap[a]->slanted == ctx->slanted) &&
                         (!spc[sp].site_emap[a] || spc[sp].site_emap[a]->slanted == ctx->slanted),
                     MH_EINVAL, "mh_gcmc_create: the maps were built for another cell shape than the context's");
    MH_HIP(hipSetDevice(ctx->device));
    std::unique_ptr<mh_gcmc> g(new mh_gcmc());
    g->ctx = ctx;
    g->nrep = nrep;
    g->maxN = maxN;
    g->nspc = nspc;
    g->prec = p->precision;
    g->sys.resize(1);
    int rc;
    if ((rc = build_system(ctx, p, maxN, ctx->fund_ell, ctx->full_ell, g->sys[0], "mh_gcmc_create", true))) return rc;
    GDev &H = g->sys[0].H;
    // species 1..: their model and move set with species 0's shared fields
    // (site types index species 0's pair tables)
    for (int sp = 1; sp < nspc; sp++) {
        mh_gcmc_params q = *p;
        const mh_gcmc_params &o = spc[sp];
        q.nsites = o.nsites;
        std::memcpy(q.ref, o.ref, sizeof q.ref);
        std::memcpy(q.site_type, o.site_type, sizeof q.site_type);
        std::memcpy(q.site_q, o.site_q, sizeof q.site_q);
        std::memcpy(q.site_pmap, o.site_pmap, sizeof q.site_pmap);
        std::memcpy(q.site_emap, o.site_emap, sizeof q.site_emap);
        q.nmoves = o.nmoves;
        std::memcpy(q.move_type, o.move_type, sizeof q.move_type);
        std::memcpy(q.move_weight, o.move_weight, sizeof q.move_weight);
        q.delta_trans = o.delta_trans;
        q.delta_rot = o.delta_rot;
        q.max_nrg = o.max_nrg;
        q.bias = o.bias;
        q.adjust_trans = o.adjust_trans;
        q.adjust_rot = o.adjust_rot;
        SysHost T;
        if ((rc = build_system(ctx, &q, maxN, ctx->fund_ell, ctx->full_ell, T, "mh_gcmc_create", true))) return rc;
        H.sp[sp] = static_cast<const SpcDev &>(T.H);
        H.nsmax = std::max(H.nsmax, T.H.ns);
        for (int s2 = 0; s2 < GSPC; s2++) {
            MH_CHECK(o.lj_comcut[s2] >= 0.0 && o.coul_comcut[s2] >= 0.0, MH_EINVAL,
                     "mh_gcmc_create: centre-of-mass cutoffs must be >= 0");
            H.cc_lj[sp][s2] = o.lj_comcut[s2] * o.lj_comcut[s2];
            H.cc_cl[sp][s2] = o.coul_comcut[s2] * o.coul_comcut[s2];
            if (o.lj_comcut[s2] > 0.0 || o.coul_comcut[s2] > 0.0) H.comcut = 1;
        }
    }
    for (int a = 0; a < nspc; a++)
        for (int b = 0; b < nspc; b++)
            MH_CHECK(H.cc_lj[a][b] == H.cc_lj[b][a] && H.cc_cl[a][b] == H.cc_cl[b][a], MH_EINVAL,
                     "mh_gcmc_create: centre-of-mass cutoffs of a species pair differ between its two species");
    H.nspc = nspc;
    g->ns = H.nsmax;
    g->slanted = H.slanted != 0;
    g->nk = H.nk;
    g->kstride = g->nk;
    if ((rc = g->dP.alloc(sizeof(GDev)))) return rc;
    if ((rc = alloc_replicas(g.get()))) return rc;
    std::vector<int> reps(nrep);
    for (int i = 0; i < nrep; i++) reps[i] = i;
    if ((rc = init_replicas(g.get(), reps, fugacity_kPa, seeds, g->sys[0].volume, p->temperature, H))) return rc;
    if ((rc = plan(g.get()))) return rc;
    ctx_retain(ctx);
    *out = g.release();
    return MH_OK;
}

int mh_gcmc_create(mh_ctx *ctx, const mh_gcmc_params *p, int nrep, const double *fugacity_kPa,
                   const int32_t *seeds, int maxN, mh_gcmc **out) {
    return mh_gcmc_create_mix(ctx, p, 1, nrep, fugacity_kPa, seeds, maxN, out);
}

int mh_gcmc_species(mh_gcmc *g, int spc) {
    MH_CHECK(g && spc >= 0 && spc < g->nspc, MH_EINVAL, "mh_gcmc_species: no such species");
    g->sel = spc;
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
    MH_CHECK(threads >= g->cav_tmax, MH_EINVAL, "mh_gcmc_set_threads: fewer threads than cavity trials");
    MH_HIP(hipSetDevice(g->ctx->device));
    MH_HIP(hipStreamSynchronize(g->ctx->stream));
    g->threads = threads;
    return plan(g);
}

static int cav_rebuild(mh_gcmc *g) {
    Rep r = g->rep();
    const int bx = (int)std::min<long long>((g->ncube + 255) / 256, 1024);
    hipLaunchKernelGGL(k_cav_clear, dim3(bx, g->nrep), dim3(256), 0, g->ctx->stream, r);
    hipLaunchKernelGGL(k_cav_fill, dim3(g->nrep), dim3(256), 0, g->ctx->stream, g->dP.as<GDev>(), r);
    MH_HIP(hipGetLastError());
    MH_HIP(hipStreamSynchronize(g->ctx->stream));
    return MH_OK;
}

/* ---------------------------------------------------------- slot pools */

int mh_pool_create(mh_ctx *ctx, int nslots, int nsys, int maxN, int kmax_cap, int max_sites, int precision,
                   mh_gcmc **out) {
    MH_CHECK(ctx && out, MH_EINVAL, "mh_pool_create: bad arguments");
    MH_CHECK(nslots >= 1 && nslots <= 65535 && nsys >= 1 && nsys <= nslots, MH_EINVAL,
             "mh_pool_create: need 1 <= nsys <= nslots <= 65535");
    MH_CHECK(maxN >= 1 && maxN <= (1 << 20), MH_EINVAL, "mh_pool_create: maxN out of range");
    MH_CHECK(kmax_cap >= 0 && kmax_cap <= GKMAX, MH_EINVAL, "mh_pool_create: kmax_cap out of range [0,15]");
    MH_CHECK(max_sites >= 1 && max_sites <= GS, MH_EINVAL, "mh_pool_create: max_sites out of range [1,8]");
    MH_CHECK(precision == MH_PREC_FP64 || precision == MH_PREC_FP32, MH_EINVAL, "mh_pool_create: bad precision");
    MH_HIP(hipSetDevice(ctx->device));
    std::unique_ptr<mh_gcmc> g(new mh_gcmc());
    g->ctx = ctx;
    g->pool = true;
    g->nrep = nslots;
    g->maxN = maxN;
    g->prec = precision;
    g->ns = max_sites;
    const long long K = kmax_cap;
    g->kstride = K ? ((2 * K + 1) * (2 * K + 1) * (2 * K + 1) - 1) / 2 : 0;
    g->sys.resize(nsys);
    g->slot_sys.assign(nslots, -1);
    int rc;
    if ((rc = g->dP.alloc(sizeof(GDev) * nsys))) return rc;
    if ((rc = alloc_replicas(g.get()))) return rc;
    if ((rc = g->sysidx.alloc(4 * (size_t)nslots)) || (rc = g->budget.alloc(4 * (size_t)nslots))) return rc;
    std::vector<int> neg(nslots, -1);
    MH_HIP(hipMemcpy(g->sysidx.p, neg.data(), 4 * (size_t)nslots, hipMemcpyHostToDevice));
    MH_HIP(hipMemset(g->budget.p, 0, 4 * (size_t)nslots));
    if ((rc = plan(g.get()))) return rc;
    ctx_retain(ctx);
    *out = g.release();
    return MH_OK;
}

int mh_pool_admit(mh_gcmc *g, const mh_gcmc_params *p, const double fund_ell[3], const double full_ell[3],
                  int nrep, const double *fugacity_kPa, const int32_t *seeds, int nsteps, int *ticket) {
    MH_CHECK(g && g->pool && p && fund_ell && full_ell && fugacity_kPa && seeds && ticket, MH_EINVAL,
             "mh_pool_admit: bad arguments");
    MH_CHECK(nrep >= 1 && nsteps >= 0, MH_EINVAL, "mh_pool_admit: bad nrep / nsteps");
    MH_CHECK(p->nsites <= g->ns, MH_EINVAL, "mh_pool_admit: more sites than the pool's max_sites");
    for (int a = 0; a < p->nsites && a < MH_MAX_SITES; a++)
        MH_CHECK((!p->site_pmap[a] || !p->site_pmap[a]->slanted) && (!p->site_emap[a] || !p->site_emap[a]->slanted),
                 MH_EINVAL, "mh_pool_admit: maps of slanted cells are not offered for GCMC");
    MH_HIP(hipSetDevice(g->ctx->device));
    MH_HIP(hipStreamSynchronize(g->ctx->stream));  // no launch in flight while slots change
    int s = -1;
    for (size_t i = 0; i < g->sys.size(); i++)
        if (!g->sys[i].used) {
            s = (int)i;
            break;
        }
    std::vector<int> reps;
    for (int i = 0; i < g->nrep && (int)reps.size() < nrep; i++)
        if (g->slot_sys[i] < 0) reps.push_back(i);
    if (s < 0 || (int)reps.size() < nrep) {
        set_error("mh_pool_admit: no free system or slots (retry after a release)");
        return MH_ECAPACITY;
    }
    SysHost &S = g->sys[s];
    mh_gcmc_params q = *p;
    q.precision = g->prec;
    int rc = build_system(g->ctx, &q, g->maxN, fund_ell, full_ell, S, "mh_pool_admit", false);
    if (!rc) rc = moves_ready(S.H, 1, "mh_pool_admit");
    if (!rc && S.H.nk > g->kstride) {
        set_error("mh_pool_admit: more k vectors than the pool's kmax_cap");
        rc = MH_ECAPACITY;
    }
    if (rc) {
        S = SysHost();
        return rc;
    }
    if ((rc = init_replicas(g, reps, fugacity_kPa, seeds, S.volume, p->temperature, S.H))) {
        S = SysHost();
        return rc;
    }
    S.slots = reps;
    std::vector<int> bud(nrep, nsteps);
    for (int i = 0; i < nrep; i++) {
        g->slot_sys[reps[i]] = s;
        MH_HIP(hipMemcpy(g->sysidx.as<int>() + reps[i], &s, 4, hipMemcpyHostToDevice));
        MH_HIP(hipMemcpy(g->budget.as<int>() + reps[i], &bud[i], 4, hipMemcpyHostToDevice));
    }
    if ((rc = plan(g))) return rc;
    *ticket = s;
    return MH_OK;
}

int mh_pool_launch(mh_gcmc *g, int nsteps) {
    MH_CHECK(g && g->pool && nsteps > 0, MH_EINVAL, "mh_pool_launch: bad arguments");
    MH_HIP(hipSetDevice(g->ctx->device));
    launch_any(g, nsteps);
    MH_HIP(hipGetLastError());
    return MH_OK;
}

int mh_pool_wait(mh_gcmc *g) {
    MH_CHECK(g && g->pool, MH_EINVAL, "mh_pool_wait: bad arguments");
    MH_HIP(hipSetDevice(g->ctx->device));
    MH_HIP(hipStreamSynchronize(g->ctx->stream));
    return MH_OK;
}

int mh_pool_status(mh_gcmc *g, int ticket, int *nrep, int *slots, int *remaining, int *err, int32_t *N,
                   double *energy, double *sumN, int64_t *nsteps, int64_t *att, int64_t *acc) {
    MH_CHECK(g && g->pool && ticket >= 0 && ticket < (int)g->sys.size() && g->sys[ticket].used && nrep,
             MH_EINVAL, "mh_pool_status: bad ticket");
    MH_HIP(hipSetDevice(g->ctx->device));
    const std::vector<int> &reps = g->sys[ticket].slots;
    *nrep = (int)reps.size();
    for (size_t i = 0; i < reps.size(); i++) {
        const int r = reps[i];
        if (slots) slots[i] = r;
        if (remaining) MH_HIP(hipMemcpy(remaining + i, g->budget.as<int>() + r, 4, hipMemcpyDeviceToHost));
        if (err) MH_HIP(hipMemcpy(err + i, g->err.as<int>() + r, 4, hipMemcpyDeviceToHost));
        if (N) MH_HIP(hipMemcpy(N + i, g->N.as<int>() + GSPC * r, 4, hipMemcpyDeviceToHost));
        if (energy) MH_HIP(hipMemcpy(energy + i, g->energy.as<double>() + r, 8, hipMemcpyDeviceToHost));
        if (sumN) MH_HIP(hipMemcpy(sumN + i, g->sumN.as<double>() + GSPC * r, 8, hipMemcpyDeviceToHost));
        if (nsteps) MH_HIP(hipMemcpy(nsteps + i, g->nsteps.as<long long>() + r, 8, hipMemcpyDeviceToHost));
        if (att) MH_HIP(hipMemcpy(att + 4 * i, g->att.as<long long>() + 4 * GSPC * r, 32, hipMemcpyDeviceToHost));
        if (acc) MH_HIP(hipMemcpy(acc + 4 * i, g->acc.as<long long>() + 4 * GSPC * r, 32, hipMemcpyDeviceToHost));
    }
    return MH_OK;
}

int mh_pool_release(mh_gcmc *g, int ticket) {
    MH_CHECK(g && g->pool && ticket >= 0 && ticket < (int)g->sys.size() && g->sys[ticket].used, MH_EINVAL,
             "mh_pool_release: bad ticket");
    MH_HIP(hipSetDevice(g->ctx->device));
    MH_HIP(hipStreamSynchronize(g->ctx->stream));
    const int neg = -1, zero = 0;
    for (int r : g->sys[ticket].slots) {
        g->slot_sys[r] = -1;
        MH_HIP(hipMemcpy(g->sysidx.as<int>() + r, &neg, 4, hipMemcpyHostToDevice));
        MH_HIP(hipMemcpy(g->budget.as<int>() + r, &zero, 4, hipMemcpyHostToDevice));
    }
    g->sys[ticket] = SysHost();
    return plan(g);
}

int mh_gcmc_set_config(mh_gcmc *g, int rep, int n, const double *com, const double *ang) {
    MH_CHECK(g && rep >= 0 && rep < g->nrep, MH_EINVAL, "mh_gcmc_set_config: bad replica");
    MH_CHECK(n >= 0 && n <= g->maxN && (n == 0 || com), MH_EINVAL, "mh_gcmc_set_config: bad n");
    MH_HIP(hipSetDevice(g->ctx->device));
    MH_CHECK(!g->pool || g->slot_sys[rep] >= 0, MH_EINVAL, "mh_gcmc_set_config: free slot");
    const GDev &G0 = g->sys_of(rep);
    const SpcDev &Q = G0.sp[g->sel];
    const int ns = Q.ns;
    std::vector<double> hc(3 * (size_t)n), ha(3 * (size_t)n, 0.0), hp(3 * (size_t)n * ns), hctr(3 * (size_t)n);
    for (int m = 0; m < n; m++) {
        for (int i = 0; i < 3; i++) {
            hc[3 * m + i] = com[3 * m + i];
            if (ang) ha[3 * m + i] = ang[3 * m + i];
            hctr[3 * m + i] = hc[3 * m + i];  // one site: its unfolded position
        }
        // host restatement of site_coords (rigidcoords_toxyz + simcell_pbc)
        auto slant_pbc_h = [&](double *r) {
            double v[3];
            for (int i = 0; i < 3; i++) {
                double t = 0.0;
                for (int c = 0; c < 3; c++) t = t + G0.sc.S[3 * i + c] * r[c];
                v[i] = t - std::floor(t / G0.full[i]) * G0.full[i];
            }
            for (int i = 0; i < 3; i++) {
                double t = 0.0;
                for (int c = 0; c < 3; c++) t = t + G0.sc.U[3 * i + c] * v[c];
                r[i] = t;
            }
        };
        if (ns == 1 && G0.slanted) {
            for (int i = 0; i < 3; i++) hp[3 * m + i] = hc[3 * m + i];
            slant_pbc_h(&hp[3 * m]);
            continue;
        }
        if (ns == 1) {
            for (int i = 0; i < 3; i++) {
                double c = hc[3 * m + i];
                hp[3 * m + i] = c - std::floor(c / G0.full[i]) * G0.full[i];
            }
            continue;
        }
        double sp = std::sin(ha[3 * m]), cp = std::cos(ha[3 * m]), st = std::sin(ha[3 * m + 1]),
               ct = std::cos(ha[3 * m + 1]), ss = std::sin(ha[3 * m + 2]), cs = std::cos(ha[3 * m + 2]);
        double M[3][3] = {{cs * cp - ct * sp * ss, -ss * cp - ct * sp * cs, st * sp},
                          {cs * sp + ct * cp * ss, -ss * sp + ct * cp * cs, -st * cp},
                          {ss * st, cs * st, ct}};
        double csum[3] = {0.0, 0.0, 0.0};
        for (int a = 0; a < ns; a++) {
            for (int i = 0; i < 3; i++) {
                double d = M[i][0] * Q.ref[a][0] + M[i][1] * Q.ref[a][1] + M[i][2] * Q.ref[a][2];
                double rp = hc[3 * m + i] + d;
                csum[i] = csum[i] + rp;
                hp[(3 * m * ns) + 3 * a + i] = G0.slanted ? rp : rp - std::floor(rp / G0.full[i]) * G0.full[i];
            }
            if (G0.slanted) slant_pbc_h(&hp[(3 * m * ns) + 3 * a]);
        }
        // config_getMolecCTR: the mean of the unfolded site positions
        for (int i = 0; i < 3; i++) hctr[3 * m + i] = csum[i] / (double)ns;
    }
    // species g->sel's slots: [sel maxN, (sel + 1) maxN) of the replica, g->ns sites each
    const long long slot0 = (long long)rep * g->nspc * g->maxN + (long long)g->sel * g->maxN;
    const long long off3 = slot0 * 3, offp = slot0 * g->ns * 3;
    if (n) {
        MH_HIP(hipMemcpy(g->com.as<double>() + off3, hc.data(), sizeof(double) * 3 * n, hipMemcpyHostToDevice));
        MH_HIP(hipMemcpy(g->ang.as<double>() + off3, ha.data(), sizeof(double) * 3 * n, hipMemcpyHostToDevice));
        MH_HIP(hipMemcpy(g->ctr.as<double>() + off3, hctr.data(), sizeof(double) * 3 * n, hipMemcpyHostToDevice));
        MH_HIP(hipMemcpy2D(g->pos.as<double>() + offp, sizeof(double) * 3 * g->ns, hp.data(),
                           sizeof(double) * 3 * ns, sizeof(double) * 3 * ns, n, hipMemcpyHostToDevice));
    }
    MH_HIP(hipMemcpy(g->N.as<int>() + GSPC * rep + g->sel, &n, sizeof(int), hipMemcpyHostToDevice));
    if (g->ncube) {
        int rc = cav_rebuild(g);
        if (rc) return rc;
    }
    // structure factor and energy of every replica from their coordinates
    std::vector<double> e(g->nrep);
    int rc = mh_gcmc_total_energy(g, e.data(), 1);
    return rc;
}

/* ---------------------------------------------------------- cavity bias */

int mh_gcmc_set_cavity(mh_gcmc *g, const mh_cavity *c) {
    MH_CHECK(g && c, MH_EINVAL, "mh_gcmc_set_cavity: bad arguments");
    MH_CHECK(!g->pool, MH_EINVAL, "mh_gcmc_set_cavity: not for slot pools");
    MH_CHECK(!g->slanted, MH_EINVAL, "mh_gcmc_set_cavity: the cube grid needs an orthorhombic cell");
    MH_CHECK(!g->ncube, MH_EINVAL, "mh_gcmc_set_cavity: the batch already has a cavity list");
    GDev &H = g->sys[0].H;
    long long nc = 1;
    for (int i = 0; i < 3; i++) {
        MH_CHECK(c->n[i] >= 1 && c->n[i] <= 4096, MH_EINVAL, "mh_gcmc_set_cavity: cubes per axis out of [1, 4096]");
        nc *= c->n[i];
    }
    MH_CHECK(nc <= (1LL << 28), MH_EINVAL, "mh_gcmc_set_cavity: more than 2^28 cubes");
    MH_CHECK(c->species_mask >= 0 && c->species_mask < (1 << g->nspc), MH_EINVAL,
             "mh_gcmc_set_cavity: species_mask names a species the batch does not have");
    int tmax = 0;
    for (int s = 0; s < g->nspc; s++) {
        const int T = c->ntrials[s];
        MH_CHECK(T >= 0 && T <= 1024, MH_EINVAL, "mh_gcmc_set_cavity: ntrials out of [0, 1024]");
        if (T == 0) continue;
        bool bins = false;
        for (int k = 0; k < H.sp[s].nmoves; k++) bins = bins || H.sp[s].mtype[k] == MH_MV_BINSERT;
        MH_CHECK(bins, MH_EINVAL, "mh_gcmc_set_cavity: a cavity-biased species has no BINSERT move");
        MH_CHECK(!c->map_bias[s] || H.sp[s].bias.cube, MH_EINVAL,
                 "mh_gcmc_set_cavity: MAP_CAVITY_BIAS needs the species' bias map");
        tmax = std::max(tmax, T);
    }
    MH_CHECK(tmax > 0, MH_EINVAL, "mh_gcmc_set_cavity: no species is cavity biased");
    MH_HIP(hipSetDevice(g->ctx->device));
    MH_HIP(hipStreamSynchronize(g->ctx->stream));
    const int cap = g->nspc * g->maxN + 2;
    const long long nt = (long long)g->nrep * GSPC * cap;
    int rc;
    if ((rc = g->cube.alloc(sizeof(int) * nc * g->nrep)) || (rc = g->cav_avg.alloc(sizeof(double) * nt)) ||
        (rc = g->cav_cnt.alloc(sizeof(long long) * nt)) || (c->tab && (rc = g->cav_tab.alloc(nc))))
        return rc;
    if (c->tab) MH_HIP(hipMemcpy(g->cav_tab.p, c->tab, nc, hipMemcpyHostToDevice));
    // rigidmoves_equateInsBias (rigidmoves.F90:2748-2760): every entry
    // starts with one update of ntrials / 2
    std::vector<double> avg(nt, 0.0);
    std::vector<long long> cnt(nt, 1);
    for (int r = 0; r < g->nrep; r++)
        for (int s = 0; s < g->nspc; s++)
            for (int t = 0; t < cap; t++) avg[((size_t)r * GSPC + s) * cap + t] = c->ntrials[s] / 2.0;
    MH_HIP(hipMemcpy(g->cav_avg.p, avg.data(), sizeof(double) * nt, hipMemcpyHostToDevice));
    MH_HIP(hipMemcpy(g->cav_cnt.p, cnt.data(), sizeof(long long) * nt, hipMemcpyHostToDevice));
    H.cav_on = 1;
    H.cav_mask = c->species_mask;
    H.cav_tab = c->tab ? g->cav_tab.as<uint8_t>() : nullptr;
    for (int i = 0; i < 3; i++) {
        H.cav_n[i] = c->n[i];
        H.cav_d[i] = H.full[i] / (double)c->n[i];  // cavitylist_init: ell / n
    }
    for (int s = 0; s < g->nspc; s++) {
        H.sp[s].cav_T = c->ntrials[s];
        H.sp[s].cav_map = c->map_bias[s] ? 1 : 0;
    }
    H.cav_T = H.sp[0].cav_T;  // the base view is species 0's
    H.cav_map = H.sp[0].cav_map;
    g->ncube = nc;
    g->cav_cap = cap;
    g->cav_tmax = tmax;
    if (g->threads < tmax) g->threads = tmax <= 64 ? 64 : tmax <= 128 ? 128 : tmax <= 256 ? 256 : tmax <= 512 ? 512 : 1024;
    if ((rc = plan(g))) return rc;
    return cav_rebuild(g);
}

int mh_gcmc_get_cavity(mh_gcmc *g, int rep, int cap, double *avg, int64_t *n, int32_t *cubes) {
    MH_CHECK(g && rep >= 0 && rep < g->nrep && cap >= 0, MH_EINVAL, "mh_gcmc_get_cavity: bad arguments");
    MH_CHECK(g->ncube, MH_EINVAL, "mh_gcmc_get_cavity: the batch has no cavity list");
    MH_HIP(hipSetDevice(g->ctx->device));
    MH_HIP(hipStreamSynchronize(g->ctx->stream));
    const int k = std::min(cap, g->cav_cap);
    const size_t o = ((size_t)rep * GSPC + g->sel) * g->cav_cap;
    if (avg && k) MH_HIP(hipMemcpy(avg, g->cav_avg.as<double>() + o, sizeof(double) * k, hipMemcpyDeviceToHost));
    if (n && k) MH_HIP(hipMemcpy(n, g->cav_cnt.as<long long>() + o, sizeof(long long) * k, hipMemcpyDeviceToHost));
    if (cubes)
        MH_HIP(hipMemcpy(cubes, g->cube.as<int>() + (size_t)rep * g->ncube, sizeof(int) * g->ncube,
                         hipMemcpyDeviceToHost));
    return MH_OK;
}

int mh_cavity_tab(const mh_map *pmap, double ucutoff, const int n[3], const int ncell[3], const double full_ell[3],
                  uint8_t *tab) {
    MH_CHECK(pmap && n && ncell && full_ell && tab, MH_EINVAL, "mh_cavity_tab: bad arguments");
    int np[3];
    int64_t nf = 0, counts[4];
    double step[3];
    int rc = mh_map_info(pmap, np, &nf, counts, step);
    if (rc) return rc;
    for (int i = 0; i < 3; i++) {
        MH_CHECK(n[i] >= 1 && ncell[i] >= 1 && n[i] % ncell[i] == 0, MH_EINVAL,
                 "mh_cavity_tab: cubes per axis must be a multiple of the unit cells (cavitylist.F90:233-241)");
        MH_CHECK(n[i] <= np[i], MH_EINVAL, "mh_cavity_tab: more cubes than map nodes on an axis");
    }
    const size_t nodes = (size_t)np[0] * np[1] * np[2];
    std::vector<int32_t> ptr(nodes);
    std::vector<float> v8(8 * (size_t)nf);
    if ((rc = mh_map_get(pmap, ptr.data(), v8.data()))) return rc;
    std::memset(tab, 0, (size_t)n[0] * n[1] * n[2]);
    double d[3];
    int per[3];
    for (int i = 0; i < 3; i++) {
        d[i] = full_ell[i] / n[i];
        per[i] = n[i] / ncell[i];
    }
    // cavitylist_initArrPmap (cavitylist.F90:790-978): the cube of each map
    // cubelet centre whose lower-corner node is tabulated below ucutoff,
    // in every unit cell of the simulation cell
    for (int i = 0; i < np[0] - 1; i++)
        for (int j = 0; j < np[1] - 1; j++)
            for (int k = 0; k < np[2] - 1; k++) {
                const int32_t q = ptr[((size_t)i * np[1] + j) * np[2] + k];
                if (q == 0 || !((double)v8[q - 1] < ucutoff)) continue;
                const double x[3] = {i * step[0] + step[0] / 2, j * step[1] + step[1] / 2, k * step[2] + step[2] / 2};
                int ic[3];
                for (int a = 0; a < 3; a++) {
                    ic[a] = (int)(x[a] / d[a]);
                    MH_CHECK(ic[a] < per[a], MH_EINVAL, "mh_cavity_tab: a map cubelet lies beyond the unit cell");
                }
                for (int u = 0; u < ncell[0]; u++)
                    for (int v = 0; v < ncell[1]; v++)
                        for (int w = 0; w < ncell[2]; w++)
                            tab[((size_t)(ic[0] + per[0] * u) * n[1] + (ic[1] + per[1] * v)) * n[2] +
                                (ic[2] + per[2] * w)] = 1;
            }
    return MH_OK;
}

int mh_gcmc_run(mh_gcmc *g, int nsteps) {
    MH_CHECK(g && nsteps >= 0, MH_EINVAL, "mh_gcmc_run: bad arguments");
    if (nsteps == 0) return MH_OK;
    if (!g->pool)
        if (int rc = moves_ready(g->sys[0].H, g->nspc, "mh_gcmc_run")) return rc;
    MH_HIP(hipSetDevice(g->ctx->device));
    KTimer t(g->ctx, g->prec == MH_PREC_FP32 ? "gcmc32" : "gcmc64");
    launch_any(g, nsteps);
    MH_HIP(hipGetLastError());
    MH_HIP(hipStreamSynchronize(g->ctx->stream));
    t.stop();
    std::vector<int> err(g->nrep);
    MH_HIP(hipMemcpy(err.data(), g->err.p, 4 * g->nrep, hipMemcpyDeviceToHost));
    for (int i = 0; i < g->nrep; i++) {
        MH_CHECK(err[i] != 2, MH_ECAPACITY, "mh_gcmc_run: a replica reached maxN molecules");
        MH_CHECK(err[i] != 4, MH_ESTATE,
                 "mh_gcmc_run: BDELETE picked a molecule whose cubelet is not in the bias map");
        MH_CHECK(err[i] != 5, MH_ECAPACITY, "mh_gcmc_run: the total loading reached the cavity-bias table size");
        MH_CHECK(err[i] == 0, MH_ESTATE, "mh_gcmc_run: an existing molecule failed to evaluate");
    }
    g->tr_base = g->tr_base;  // trace window stays anchored at the step set_trace saw
    return MH_OK;
}

int mh_gcmc_run_cycles(mh_gcmc *g, int nsteps, int ncycles) {
    MH_CHECK(g && nsteps >= 0 && ncycles >= 0, MH_EINVAL, "mh_gcmc_run_cycles: bad arguments");
    if (nsteps == 0 || ncycles == 0) return MH_OK;
    if (!g->pool)
        if (int rc = moves_ready(g->sys[0].H, g->nspc, "mh_gcmc_run_cycles")) return rc;
    MH_HIP(hipSetDevice(g->ctx->device));
    mh_gcmc::GraphKey k;
    std::memset(&k, 0, sizeof k);
    k.r = g->rep();
    k.P = g->dP.p;
    k.nsteps = nsteps;
    k.threads = g->threads;
    k.prec = g->prec;
    k.kpl = g->kpl;
    k.ldsmol = g->ldsmol;
    k.exy_off = g->exy_off;
    k.lds_bytes = g->lds_bytes;
    if (!g->gexec || std::memcmp(&k, &g->gkey, sizeof k) != 0) {
        if (g->gexec) {
            MH_HIP(hipGraphExecDestroy(g->gexec));
            g->gexec = nullptr;
        }
        hipStream_t st = g->ctx->stream;
        MH_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        launch_any(g, nsteps);
        hipGraph_t graph = nullptr;
        const hipError_t ec = hipStreamEndCapture(st, &graph);
        MH_HIP(ec);
        const hipError_t ei = hipGraphInstantiate(&g->gexec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        MH_HIP(ei);
        g->gkey = k;
    }
    KTimer t(g->ctx, g->prec == MH_PREC_FP32 ? "gcmc32" : "gcmc64");
    for (int c = 0; c < ncycles; c++) MH_HIP(hipGraphLaunch(g->gexec, g->ctx->stream));
    MH_HIP(hipStreamSynchronize(g->ctx->stream));
    t.stop();
    std::vector<int> err(g->nrep);
    MH_HIP(hipMemcpy(err.data(), g->err.p, 4 * g->nrep, hipMemcpyDeviceToHost));
    for (int i = 0; i < g->nrep; i++) {
        MH_CHECK(err[i] != 2, MH_ECAPACITY, "mh_gcmc_run_cycles: a replica reached maxN molecules");
        MH_CHECK(err[i] != 4, MH_ESTATE,
                 "mh_gcmc_run_cycles: BDELETE picked a molecule whose cubelet is not in the bias map");
        MH_CHECK(err[i] != 5, MH_ECAPACITY, "mh_gcmc_run_cycles: the total loading reached the cavity-bias table size");
        MH_CHECK(err[i] == 0, MH_ESTATE, "mh_gcmc_run_cycles: an existing molecule failed to evaluate");
    }
    return MH_OK;
}

// the selected species' entries (elem bytes each, per species) of a
// per-species device array [rep][GSPC][...]
static int spc_gather(mh_gcmc *g, const DBuf &b, size_t elem, void *out) {
    const size_t R = g->nrep;
    std::vector<char> h(R * GSPC * elem);
    MH_HIP(hipMemcpy(h.data(), b.p, h.size(), hipMemcpyDeviceToHost));
    for (size_t r = 0; r < R; r++)
        std::memcpy((char *)out + r * elem, h.data() + (r * GSPC + g->sel) * elem, elem);
    return MH_OK;
}

int mh_gcmc_get_state(mh_gcmc *g, int32_t *N, double *energy, double *sumN, int64_t *nsteps, int64_t *att,
                      int64_t *acc, int32_t *rng_ix, int32_t *rng_iy) {
    MH_CHECK(g, MH_EINVAL, "mh_gcmc_get_state: g is NULL");
    MH_HIP(hipSetDevice(g->ctx->device));
    const size_t R = g->nrep;
    int rc;
    if (N && (rc = spc_gather(g, g->N, 4, N))) return rc;
    if (energy) MH_HIP(hipMemcpy(energy, g->energy.p, 8 * R, hipMemcpyDeviceToHost));
    if (sumN && (rc = spc_gather(g, g->sumN, 8, sumN))) return rc;
    if (nsteps) MH_HIP(hipMemcpy(nsteps, g->nsteps.p, 8 * R, hipMemcpyDeviceToHost));
    if (att && (rc = spc_gather(g, g->att, 32, att))) return rc;
    if (acc && (rc = spc_gather(g, g->acc, 32, acc))) return rc;
    if (rng_ix) MH_HIP(hipMemcpy(rng_ix, g->ix.p, 4 * R, hipMemcpyDeviceToHost));
    if (rng_iy) MH_HIP(hipMemcpy(rng_iy, g->iy.p, 4 * R, hipMemcpyDeviceToHost));
    return MH_OK;
}

int mh_gcmc_get_moments(mh_gcmc *g, double *sumN, double *sumN2) {
    MH_CHECK(g, MH_EINVAL, "mh_gcmc_get_moments: g is NULL");
    MH_HIP(hipSetDevice(g->ctx->device));
    int rc;
    if (sumN && (rc = spc_gather(g, g->sumN, 8, sumN))) return rc;
    if (sumN2 && (rc = spc_gather(g, g->sumN2, 8, sumN2))) return rc;
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

int mh_gcmc_get_deltas(mh_gcmc *g, double *delta) {
    MH_CHECK(g && delta, MH_EINVAL, "mh_gcmc_get_deltas: bad arguments");
    MH_HIP(hipSetDevice(g->ctx->device));
    MH_HIP(hipStreamSynchronize(g->ctx->stream));
    return spc_gather(g, g->delta, 16, delta);
}

int mh_gcmc_set_deltas(mh_gcmc *g, int rep, double delta_trans, double delta_rot) {
    MH_CHECK(g && rep >= 0 && rep < g->nrep, MH_EINVAL, "mh_gcmc_set_deltas: bad replica");
    MH_CHECK(delta_trans >= 0.0 && delta_rot >= 0.0, MH_EINVAL, "mh_gcmc_set_deltas: negative displacement");
    MH_HIP(hipSetDevice(g->ctx->device));
    MH_HIP(hipStreamSynchronize(g->ctx->stream));
    const double d2[2] = {delta_trans, delta_rot};
    MH_HIP(hipMemcpy(g->delta.as<double>() + 2 * (GSPC * rep + g->sel), d2, 16, hipMemcpyHostToDevice));
    return MH_OK;
}

int mh_gcmc_set_counts(mh_gcmc *g, int rep, const int64_t *att, const int64_t *acc) {
    MH_CHECK(g && rep >= 0 && rep < g->nrep && att && acc, MH_EINVAL, "mh_gcmc_set_counts: bad arguments");
    for (int i = 0; i < 4; ++i)
        MH_CHECK(att[i] >= 0 && acc[i] >= 0 && acc[i] <= att[i], MH_EINVAL,
                 "mh_gcmc_set_counts: counts must satisfy 0 <= acc <= att");
    MH_HIP(hipSetDevice(g->ctx->device));
    MH_HIP(hipStreamSynchronize(g->ctx->stream));
    long long a[4], c[4];
    for (int i = 0; i < 4; ++i) a[i] = att[i], c[i] = acc[i];
    MH_HIP(hipMemcpy(g->att.as<long long>() + 4 * (GSPC * rep + g->sel), a, 32, hipMemcpyHostToDevice));
    MH_HIP(hipMemcpy(g->acc.as<long long>() + 4 * (GSPC * rep + g->sel), c, 32, hipMemcpyHostToDevice));
    return MH_OK;
}

int mh_gcmc_get_config(mh_gcmc *g, int rep, int *n, double *com, double *ang, double *r) {
    MH_CHECK(g && rep >= 0 && rep < g->nrep && n, MH_EINVAL, "mh_gcmc_get_config: bad arguments");
    MH_HIP(hipSetDevice(g->ctx->device));
    int hn = 0;
    MH_HIP(hipMemcpy(&hn, g->N.as<int>() + GSPC * rep + g->sel, 4, hipMemcpyDeviceToHost));
    *n = hn;
    const long long slot0 = (long long)rep * g->nspc * g->maxN + (long long)g->sel * g->maxN;
    const long long off3 = slot0 * 3, offp = slot0 * g->ns * 3;
    const int ns = g->sys_of(rep).sp[g->sel].ns;
    if (hn && com) MH_HIP(hipMemcpy(com, g->com.as<double>() + off3, sizeof(double) * 3 * hn, hipMemcpyDeviceToHost));
    if (hn && ang) MH_HIP(hipMemcpy(ang, g->ang.as<double>() + off3, sizeof(double) * 3 * hn, hipMemcpyDeviceToHost));
    if (hn && r)
        MH_HIP(hipMemcpy2D(r, sizeof(double) * 3 * ns, g->pos.as<double>() + offp, sizeof(double) * 3 * g->ns,
                           sizeof(double) * 3 * ns, hn, hipMemcpyDeviceToHost));
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
        if ((rc = save.alloc(sizeof(double2) * (R * g->kstride + 1)))) return rc;
        MH_HIP(hipMemcpyAsync(save.p, g->S.p, sizeof(double2) * R * g->kstride, hipMemcpyDeviceToDevice,
                              g->ctx->stream));
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
        MH_HIP(hipMemcpyAsync(g->S.p, save.p, sizeof(double2) * R * g->kstride, hipMemcpyDeviceToDevice,
                              g->ctx->stream));
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
                           rep, g->sel, n, dc.as<double>(), da.as<double>(), dsb.as<double2>(), de.as<double>(),
                           dok.as<uint8_t>());
    else
        hipLaunchKernelGGL((k_trial<256, MH_PREC_FP64>), dim3(n), dim3(256), 0, g->ctx->stream, g->dP.as<GDev>(), r,
                           rep, g->sel, n, dc.as<double>(), da.as<double>(), dsb.as<double2>(), de.as<double>(),
                           dok.as<uint8_t>());
    MH_HIP(hipGetLastError());
    MH_HIP(hipStreamSynchronize(g->ctx->stream));
    MH_HIP(hipMemcpy(e, de.p, sizeof(double) * n, hipMemcpyDeviceToHost));
    MH_HIP(hipMemcpy(ok, dok.p, n, hipMemcpyDeviceToHost));
    return MH_OK;
}

int mh_gcmc_autotune(mh_gcmc *g, int probe_steps, int *threads_out, int *exy_out) {
    MH_CHECK(g && probe_steps > 0, MH_EINVAL, "mh_gcmc_autotune: bad arguments");
    if (!g->pool)
        if (int rc = moves_ready(g->sys[0].H, g->nspc, "mh_gcmc_autotune")) return rc;
    MH_HIP(hipSetDevice(g->ctx->device));
    hipStream_t st = g->ctx->stream;
    // snapshot of every buffer a run modifies
    std::vector<DBuf *> bufs = {&g->N, &g->err, &g->ix, &g->iy, &g->first, &g->iseed, &g->energy, &g->sumN,
                                &g->sumN2, &g->nsteps, &g->att, &g->acc, &g->com, &g->ang, &g->pos, &g->ctr, &g->S,
                                &g->delta};
    if (g->pool) bufs.push_back(&g->budget);
    if (g->ncube) {
        bufs.push_back(&g->cube);
        bufs.push_back(&g->cav_avg);
        bufs.push_back(&g->cav_cnt);
    }
    const int nb = (int)bufs.size();
    std::vector<DBuf> snap(nb);
    int rc;
    for (int i = 0; i < nb; i++) {
        if ((rc = snap[i].alloc(bufs[i]->n))) return rc;
        MH_HIP(hipMemcpyAsync(snap[i].p, bufs[i]->p, bufs[i]->n, hipMemcpyDeviceToDevice, st));
    }
    const int tr_cap = g->tr_cap;
    g->tr_cap = 0;  // no trace writes while probing
    auto restore = [&]() -> int {
        for (int i = 0; i < nb; i++)
            MH_HIP(hipMemcpyAsync(bufs[i]->p, snap[i].p, bufs[i]->n, hipMemcpyDeviceToDevice, st));
        MH_HIP(hipStreamSynchronize(st));
        return MH_OK;
    };
    hipEvent_t e0, e1;
    MH_HIP(hipEventCreate(&e0));
    MH_HIP(hipEventCreate(&e1));
    float best = 1e30f;
    int best_t = g->threads;
    bool best_exy_off = g->exy_off;
    for (int t : {64, 128, 256, 512}) {
        if (t < g->cav_tmax) continue;  // one cavity trial per lane
        for (int xo = 0; xo < 2; xo++) {
            g->threads = t;
            g->exy_off = xo != 0;
            if ((rc = plan(g))) break;
            if (xo && g->nk == 0) continue;
            if ((rc = restore())) break;
            for (int rep = 0; rep < 2; rep++) {  // first launch warms the code object
                (void)hipEventRecord(e0, st);
                launch_any(g, probe_steps);
                (void)hipEventRecord(e1, st);
                if (hipGetLastError() != hipSuccess || hipEventSynchronize(e1) != hipSuccess) {
                    rc = MH_EHIP;
                    break;
                }
                float ms = 0.0f;
                (void)hipEventElapsedTime(&ms, e0, e1);
                if (rep == 1 && ms < best) {
                    best = ms;
                    best_t = t;
                    best_exy_off = g->exy_off;
                }
            }
            if (rc) break;
        }
        if (rc) break;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    g->tr_cap = tr_cap;
    g->threads = best_t;
    g->exy_off = best_exy_off;
    (void)plan(g);
    int rc2 = restore();
    if (rc) {
        set_error("mh_gcmc_autotune: a probe launch failed");
        return rc;
    }
    if (rc2) return rc2;
    if (threads_out) *threads_out = best_t;
    if (exy_out) {
        *exy_out = 0;
        for (auto &s : g->sys)
            if (s.used && s.H.use_exy) *exy_out = 1;
    }
    return MH_OK;
}

}  // extern "C"
