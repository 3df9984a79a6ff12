/*
 * music_oracle.c -- CPU FP64 restatement of the Music GCMC energy hot path
 * (map build, compaction, map file, Hermite interpolation).
 *
 * TEST INFRASTRUCTURE ONLY (parity checker and CPU baseline); the product
 * path never links it.  Built with -ffp-contract=off so that every
 * expression is evaluated in the reference's left-to-right order.
 *
 * Pinning (see tests/test_oracle_*.py):
 *  - mo_rranf, mo_erfc, mo_lj_pair_hot, mo_lj_node, mo_ewald_sfactor,
 *    mo_ewald_node, mo_compact, mo_map_write, mo_pothermite:
 *      compared against oracle/_ref (the reference's own Fortran modules
 *      random/utils/lj/ewald/pmap/interpolate compiled with flang, driven by
 *      oracle/ref_harness.f90) on committed fixtures in tests/golden/.
 *  - mo_lj_node/mo_ewald_node/mo_compact: test012/test013 golden rows and
 *      compaction counts (tests/test012-pmapgen/output_reference:284-305,
 *      tests/test013-emapgen/output_reference:122,241-262).
 *  - mo_buck_node: closed form (finite differences of the potential,
 *      reduction to LJ-like inverse powers when A = 0) -- the reference's
 *      buck_snglint stops before evaluating (buck.F90:303-307).
 *  - mo_pothermite/mo_forcehermite: node values reproduced exactly at nodes,
 *      exact reproduction of tricubic polynomials, finite differences.
 */
#include "music_oracle.h"
#include "music_gcmc_oracle.h"
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ RNG */
/* random.F90:51-54 (random_init) */
void mo_rng_init(mo_rng *r, int32_t iseed)
{
    r->iseed = iseed;
    r->first = 1;
    r->ix = -1;
    r->iy = -1;
}

/* am = Nearest(1.0,-1.0)/CONST_IM evaluated in single precision
 * (random.F90:130): (1 - 2^-24) / 2^31, exact in float. */
static double mo_rng_am(void)
{
    float one_m = nextafterf(1.0f, -1.0f);
    float im = (float)2147483647; /* rounds to 2^31 */
    return (double)(one_m / im);
}

/* random.F90:121-152 (rranf_NR) */
double mo_rranf(mo_rng *r)
{
    if (r->first) {
        int32_t a = r->iseed < 0 ? -r->iseed : r->iseed;
        r->first = 0;
        r->iy = (int32_t)((888889999 ^ a) | 1);
        r->ix = (int32_t)(777755555 ^ a);
        r->iseed = a + 1;
    }
    uint32_t x = (uint32_t)r->ix;
    x ^= x << 13;
    x ^= x >> 17;
    x ^= x << 5;
    r->ix = (int32_t)x;
    int32_t k = r->iy / 127773;
    r->iy = 16807 * (r->iy - k * 127773) - 2836 * k;
    if (r->iy < 0) r->iy += 2147483647;
    int32_t v = (int32_t)((2147483647 & (r->ix ^ r->iy)) | 1);
    return mo_rng_am() * (double)v;
}

void mo_rranf_n(int32_t iseed, int n, double *out, int32_t *ix_out, int32_t *iy_out)
{
    mo_rng r;
    mo_rng_init(&r, iseed);
    for (int i = 0; i < n; i++) out[i] = mo_rranf(&r);
    if (ix_out) *ix_out = r.ix;
    if (iy_out) *iy_out = r.iy;
}

/* ----------------------------------------------------------- math */
/* utils.F90:2090-2101: A&S 7.1.26, coefficients are default-real literals
 * (single precision) assigned to double parameters. */
double mo_erfc(double x)
{
    const double a1 = (double)0.254829592f, a2 = (double)-0.284496736f,
                 a3 = (double)1.421413741f, a4 = (double)-1.453152027f,
                 a5 = (double)1.061405429f, p = (double)0.3275911f;
    double t = 1.0 / (1.0 + p * x);
    double xsq = x * x;
    double tp = t * (a1 + t * (a2 + t * (a3 + t * (a4 + t * a5))));
    return tp * exp(-xsq);
}

/* fundcell.F90:640-709, orthorhombic branch: single fold, sep = a - b */
void mo_minimage(const double ell[3], const double a[3], const double b[3], double sep[3])
{
    for (int i = 0; i < 3; i++) {
        double h = 0.5 * ell[i];
        double d = a[i] - b[i];
        if (d > h) d = d - ell[i];
        else if (d < -h) d = d + ell[i];
        sep[i] = d;
    }
}

/* ------------------------------------------------------------ cells */
/* matrix_inverse (matrix.F90:195-229): LU decomposition with implicit
 * scaling and partial pivoting, then one back substitution per column
 * (matrixops_ludcmpReal / matrixops_lubksbReal, matrixops.F90:31-161) */
static void lu_inverse3(const double m[3][3], double inv[3][3])
{
    double a[3][3], vv[3];
    int indx[3];
    memcpy(a, m, sizeof a);
    for (int i = 0; i < 3; i++) {
        double aamax = 0.0;
        for (int j = 0; j < 3; j++)
            if (fabs(a[i][j]) > aamax) aamax = fabs(a[i][j]);
        vv[i] = 1.0 / aamax;
    }
    for (int j = 0; j < 3; j++) {
        for (int i = 1; i < j; i++) {
            double sum = a[i][j];
            for (int k = 0; k < i; k++) sum = sum - a[i][k] * a[k][j];
            a[i][j] = sum;
        }
        double aamax = 0.0;
        int imax = j;
        for (int i = j; i < 3; i++) {
            double sum = a[i][j];
            if (j > 0) {
                for (int k = 0; k < j; k++) sum = sum - a[i][k] * a[k][j];
                a[i][j] = sum;
            }
            double dum = vv[i] * fabs(sum);
            if (dum >= aamax) {
                imax = i;
                aamax = dum;
            }
        }
        if (j != imax) {
            for (int k = 0; k < 3; k++) {
                double dum = a[imax][k];
                a[imax][k] = a[j][k];
                a[j][k] = dum;
            }
            vv[imax] = vv[j];
        }
        indx[j] = imax;
        if (j != 2) {
            if (a[j][j] == 0.0) a[j][j] = (double)1.0e-20f;
            double dum = 1.0 / a[j][j];
            for (int i = j + 1; i < 3; i++) a[i][j] = a[i][j] * dum;
        }
    }
    if (a[2][2] == 0.0) a[2][2] = (double)1.0e-20f;
    for (int col = 0; col < 3; col++) {
        double b[3] = {0.0, 0.0, 0.0};
        b[col] = 1.0;
        int ii = -1;
        for (int i = 0; i < 3; i++) {
            int ll = indx[i];
            double sum = b[ll];
            b[ll] = b[i];
            if (ii >= 0) {
                for (int j = ii; j < i; j++) sum = sum - a[i][j] * b[j];
            } else if (sum != 0.0) {
                ii = i;
            }
            b[i] = sum;
        }
        for (int i = 2; i >= 0; i--) {
            double sum = b[i];
            for (int j = i + 1; j < 3; j++) sum = sum - a[i][j] * b[j];
            b[i] = sum / a[i][i];
        }
        for (int i = 0; i < 3; i++) inv[i][col] = b[i];
    }
}

/* fundcell_init (fundcell.F90:60-112) + fundcell_initxform (:378-412) +
 * fundcell_construct (:345-371) + fundcell_genlatticevects (:174-208).
 * The unslant matrix is formed from the angles rounded to default real
 * (float32) with single-precision cosines, as the reference declares them. */
void mo_cell_init(mo_cell *c, const double fell[3], const double ang_deg[3], const double origin[3],
                  const double eff[3], const int ncell[3])
{
    const double degtorad = MO_PI / 180.0;
    double err = fabs(ang_deg[0] - 90.0) + fabs(ang_deg[1] - 90.0) + fabs(ang_deg[2] - 90.0);
    c->ortho = err > 0.001 ? 0 : 1;
    double abc = ang_deg[0] * degtorad, aac = ang_deg[1] * degtorad, aab = ang_deg[2] * degtorad;
    for (int i = 0; i < 3; i++) {
        c->fell[i] = fell[i];
        c->finv[i] = 1.0 / fell[i];
        c->origin[i] = origin[i];
        c->eff[i] = eff[i];
        double rs = (double)ncell[i];
        c->ell[i] = fell[i] * rs;
        c->half[i] = (0.5 * fell[i]) * rs;
    }
    float alpha = (float)abc, beta = (float)aac, gamma = (float)aab;
    float cg = cosf(gamma), cb = cosf(beta), ca = cosf(alpha);
    memset(c->U, 0, sizeof c->U);
    c->U[0][0] = 1.0;
    c->U[0][1] = (double)cg;
    c->U[0][2] = (double)cb;
    float cg2 = cg * cg;
    c->U[1][1] = sqrt(1.0 - (double)cg2);
    float num = ca - cg * cb;
    c->U[1][2] = (double)num / c->U[1][1];
    c->U[2][2] = sqrt(1.0 - c->U[0][2] * c->U[0][2] - c->U[1][2] * c->U[1][2]);
    lu_inverse3(c->U, c->S);
    c->lvec[0][0] = c->ell[0];
    c->lvec[0][1] = 0.0;
    c->lvec[0][2] = 0.0;
    c->lvec[1][0] = c->ell[1] * cos(aab);
    c->lvec[1][1] = c->ell[1] * sin(aab);
    c->lvec[1][2] = 0.0;
    c->lvec[2][0] = c->ell[2] * cos(aac);
    c->lvec[2][1] = c->ell[2] * (cos(abc) - cos(aac) * cos(aab)) / sin(aab);
    double m2 = 0.0;
    m2 = m2 + c->lvec[2][0] * c->lvec[2][0];
    m2 = m2 + c->lvec[2][1] * c->lvec[2][1];
    m2 = m2 + 0.0 * 0.0;
    c->lvec[2][2] = sqrt(c->ell[2] * c->ell[2] - m2);
}

/* matrix_m_mult_v (matrix.F90:137-150) */
static void mat_vec(const double m[3][3], const double v[3], double out[3])
{
    for (int r = 0; r < 3; r++) {
        double s = 0.0;
        s = s + m[r][0] * v[0];
        s = s + m[r][1] * v[1];
        s = s + m[r][2] * v[2];
        out[r] = s;
    }
}

void mo_cell_minimage(const mo_cell *c, const double a[3], const double b[3], double sep[3])
{
    if (c->ortho) {
        mo_minimage(c->ell, a, b, sep);
        return;
    }
    double va[3], vb[3], d[3];
    mat_vec(c->S, a, va);
    mat_vec(c->S, b, vb);
    for (int i = 0; i < 3; i++) {
        double x = va[i] - vb[i];
        if (x > c->half[i]) x = x - c->ell[i];
        else if (x < -c->half[i]) x = x + c->ell[i];
        d[i] = x;
    }
    mat_vec(c->U, d, sep);
}

void mo_cell_maptouc(const mo_cell *c, const double r[3], double out[3])
{
    double v[3], w[3];
    for (int i = 0; i < 3; i++) v[i] = r[i] - c->origin[i];
    if (!c->ortho) {
        mat_vec(c->S, v, w);
        memcpy(v, w, sizeof v);
    }
    for (int i = 0; i < 3; i++) {
        double d = v[i] * c->finv[i];
        if (d >= 0.0) d = fmod(d, 1.0);
        else d = fmod(d, 1.0) + 1.0;
        v[i] = d * c->fell[i];
    }
    if (!c->ortho) {
        mat_vec(c->U, v, w);
        memcpy(v, w, sizeof v);
    }
    for (int i = 0; i < 3; i++) out[i] = v[i] + c->origin[i];
}

void mo_cell_pbc(const mo_cell *c, const double r[3], double out[3])
{
    double v[3];
    if (c->ortho) memcpy(v, r, sizeof v);
    else mat_vec(c->S, r, v);
    for (int i = 0; i < 3; i++) {
        double idx = floor(v[i] / c->ell[i]);
        v[i] = v[i] - idx * c->ell[i];
    }
    if (c->ortho) memcpy(out, v, sizeof v);
    else mat_vec(c->U, v, out);
}

/* minimum image in the given cell, or the orthorhombic full_ell */
static inline void mi(const double full_ell[3], const mo_cell *cell, const double a[3], const double b[3],
                      double sep[3])
{
    if (cell) mo_cell_minimage(cell, a, b, sep);
    else mo_minimage(full_ell, a, b, sep);
}

static inline double normsq(const double v[3])
{
    double s = 0.0;
    s = s + v[0] * v[0];
    s = s + v[1] * v[1];
    s = s + v[2] * v[2];
    return s;
}

/* lj.F90:860-869 */
void mo_lj_AB(double sig, double eps_K, double *A, double *B)
{
    double s6 = pow(sig, 6);
    *B = 4.0 * eps_K * s6 * MO_KCALMOLE_KB;
    *A = (*B) * s6;
}

/* lj.F90:927-941 (SIG@LBMIX, EPS@LBMIX) */
void mo_lj_mix_lb(double sig1, double eps1, double sig2, double eps2, double *sig, double *eps)
{
    *sig = (sig1 + sig2) / 2.0;
    *eps = sqrt(eps1 * eps2);
}

/* lj.F90:652-733 with hot present and nderivs > 0 */
int mo_lj_pair_hot(double A, double B, double hicut2, double locut2,
                   const double sep[3], double *nrg_out, double hot[8])
{
    double r2 = normsq(sep);
    if (r2 > hicut2) return 1;
    if (r2 < locut2) return 0;
    double r2i = 1.0 / r2;
    double r6i = r2i * r2i * r2i;
    double r12i = r6i * r6i;
    double nrg = A * r12i - B * r6i;
    *nrg_out = nrg;
    double fac1 = (12.0 * A * r12i - 6.0 * B * r6i) * r2i;
    double f0 = sep[0] * fac1, f1 = sep[1] * fac1, f2 = sep[2] * fac1;
    double r10i = r12i * r2;
    double r16i = r10i * r6i;
    hot[0] = nrg;
    hot[1] = -f0;
    hot[2] = -f1;
    hot[3] = -f2;
    fac1 = (168.0 * A * r16i - 48.0 * B * r10i);
    hot[4] = fac1 * sep[0] * sep[1];
    hot[5] = fac1 * sep[0] * sep[2];
    hot[6] = fac1 * sep[1] * sep[2];
    fac1 = (-2688.0 * A * r16i + 480.0 * B * r10i) * r2i;
    hot[7] = fac1 * sep[0] * sep[1] * sep[2];
    return 2;
}

/* Node loop: pairmodel.F90:892-909 (hot += hotatm per atom, stop at the
 * first locut violation), node assembly mapmaker.F90:500-547. */
static int lj_node_c(const double p[3], int natoms, const double *xyz, const int *type,
                     const double *A, const double *B, const double *hicut2,
                     const double *locut2, const double full_ell[3], const mo_cell *cell, double out[8])
{
    double hot[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    double pot = 0.0;
    int ok = 1;
    for (int a = 0; a < natoms; a++) {
        double sep[3];
        mi(full_ell, cell, p, xyz + 3 * a, sep);
        double hotatm[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        double nrg = 0.0;
        int t = type[a];
        int rc = mo_lj_pair_hot(A[t], B[t], hicut2[t], locut2[t], sep, &nrg, hotatm);
        if (rc == 2) pot = pot + nrg;
        for (int i = 0; i < 8; i++) hot[i] = hot[i] + hotatm[i];
        if (rc == 0) { ok = 0; break; }
    }
    if (!ok) {
        out[0] = MO_VERY_LRG_NRG;
    } else {
        out[0] = pot;
    }
    for (int i = 1; i < 8; i++) out[i] = hot[i];
    return ok;
}

int mo_lj_node(const double p[3], int natoms, const double *xyz, const int *type,
               const double *A, const double *B, const double *hicut2,
               const double *locut2, const double full_ell[3], double out[8])
{
    return lj_node_c(p, natoms, xyz, type, A, B, hicut2, locut2, full_ell, NULL, out);
}

/* Buckingham, closed form: V = A exp(-r/rho) - C r^-6 - D r^-8 per pair
 * (buck.F90:329).  Derivatives of a radial V(r):
 *   dV/dx = x f1, f1 = V'/r;  d2V/dxdy = x y g2, g2 = f1'/r;
 *   d3V/dxdydz = x y z g3, g3 = g2'/r. */
static int buck_node_c(const double p[3], int natoms, const double *xyz, const int *type,
                       const double *A, const double *rho, const double *C, const double *D,
                       const double *hicut2, const double *locut2, const double full_ell[3],
                       const mo_cell *cell, double out[8])
{
    double hot[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    double pot = 0.0;
    int ok = 1;
    for (int a = 0; a < natoms; a++) {
        double sep[3];
        mi(full_ell, cell, p, xyz + 3 * a, sep);
        int t = type[a];
        double r2 = normsq(sep);
        if (r2 > hicut2[t]) continue;
        if (r2 < locut2[t]) { ok = 0; break; }
        double r = sqrt(r2);
        double ri = 1.0 / r;
        double r2i = ri * ri;
        double r6i = r2i * r2i * r2i;
        double r8i = r6i * r2i;
        double e = A[t] * exp(-r / rho[t]);
        double ir = 1.0 / rho[t];
        double v = e - C[t] * r6i - D[t] * r8i;
        double f1 = -e * ir * ri + 6.0 * C[t] * r8i + 8.0 * D[t] * r8i * r2i;
        double g2 = e * ir * (ir * r2i + r2i * ri) - 48.0 * C[t] * r8i * r2i
                    - 80.0 * D[t] * r8i * r2i * r2i;
        double g3 = -e * ir * (ir * ir * r2i * ri + 3.0 * ir * r2i * r2i + 3.0 * r2i * r2i * ri)
                    + 480.0 * C[t] * r6i * r6i + 960.0 * D[t] * r8i * r6i;
        pot = pot + v;
        hot[0] += v;
        hot[1] += sep[0] * f1;
        hot[2] += sep[1] * f1;
        hot[3] += sep[2] * f1;
        hot[4] += g2 * sep[0] * sep[1];
        hot[5] += g2 * sep[0] * sep[2];
        hot[6] += g2 * sep[1] * sep[2];
        hot[7] += g3 * sep[0] * sep[1] * sep[2];
    }
    out[0] = ok ? pot : MO_VERY_LRG_NRG;
    for (int i = 1; i < 8; i++) out[i] = hot[i];
    return ok;
}

int mo_buck_node(const double p[3], int natoms, const double *xyz, const int *type,
                 const double *A, const double *rho, const double *C, const double *D,
                 const double *hicut2, const double *locut2, const double full_ell[3],
                 double out[8])
{
    return buck_node_c(p, natoms, xyz, type, A, rho, C, D, hicut2, locut2, full_ell, NULL, out);
}

/* ------------------------------------------------------------- Ewald */
/* ewald.F90:558-681 (ewald_sfactor + ewald_sfSum) */
int mo_ewald_sfactor(int natoms, const double *xyz, const double *q, const double ell[3],
                     double kappa_param, int kmax, double locut, int nkmax,
                     double *k, double *s, double *si, int *kidx, double *kappa_out)
{
    int nk = 0;
    double ialpha = 1.0 / ell[0], ibeta = 1.0 / ell[1], igamma = 1.0 / ell[2];
    double mn = ell[0];
    if (ell[1] < mn) mn = ell[1];
    if (ell[2] < mn) mn = ell[2];
    double kappa = kappa_param / mn;
    if (kappa_out) *kappa_out = kappa;
    double ikap = 1.0 / kappa;
    double pk = MO_PI * ikap;
    double ikapPiSq = pk * pk;
    double iVolPi = ialpha * ibeta * igamma / MO_PI;
    for (int x = -kmax; x <= kmax; x++) {
        double kalpha = (double)x * ialpha;
        double ka2pi = kalpha * MO_TWOPI;
        double ksqx = kalpha * kalpha;
        for (int y = -kmax; y <= kmax; y++) {
            double kbeta = (double)y * ibeta;
            double kb2pi = kbeta * MO_TWOPI;
            double ksqy = kbeta * kbeta;
            for (int z = -kmax; z <= kmax; z++) {
                if (x == 0 && y == 0 && z == 0) continue;
                double kgamma = (double)z * igamma;
                double kg2pi = kgamma * MO_TWOPI;
                double ksqz = kgamma * kgamma;
                double ksq = ksqx + ksqy + ksqz;
                double factor = iVolPi * exp(-ikapPiSq * ksq) / ksq;
                if (!(factor > locut)) continue;
                double sum = 0.0, sumi = 0.0;
                for (int i = 0; i < natoms; i++) {
                    const double *r = xyz + 3 * i;
                    double rdotk = ka2pi * r[0] + kb2pi * r[1] + kg2pi * r[2];
                    sum = sum + cos(rdotk) * q[i];
                    sumi = sumi + sin(rdotk) * q[i];
                }
                double smag = factor * factor * (sum * sum + sumi * sumi);
                if (smag <= locut * 1.0e-3) continue;
                if (nk >= nkmax) return -1;
                s[nk] = sum * factor;
                si[nk] = sumi * factor;
                k[3 * nk + 0] = ka2pi;
                k[3 * nk + 1] = kb2pi;
                k[3 * nk + 2] = kg2pi;
                if (kidx) {
                    kidx[3 * nk + 0] = x;
                    kidx[3 * nk + 1] = y;
                    kidx[3 * nk + 2] = z;
                }
                nk++;
            }
        }
    }
    return nk;
}

/* ewald.F90:334-398 (getInteractionHOT), :856-918 (rwaldHOT),
 * :966-1005 (skwaldHOT); node assembly mapmaker.F90:536-543. */
static int ewald_node_c(const double p[3], int natoms, const double *xyz, const double *q,
                        double charge1, double kappa, int nk, const double *k,
                        const double *s, const double *si, const double full_ell[3],
                        const mo_cell *cell, double out[8])
{
    const double risqrtpi = 0.5641896;
    double hot[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    double lowcut = MO_LOW_DIST_CUTOFF;
    double lowcut2 = lowcut * lowcut;
    /* pass 1: low-cut check over all atoms (ewald.F90:373-375) */
    for (int i = 0; i < natoms; i++) {
        double sep[3];
        mi(full_ell, cell, p, xyz + 3 * i, sep);
        if (normsq(sep) < lowcut2) {
            out[0] = MO_VERY_LRG_NRG;
            for (int j = 1; j < 8; j++) out[j] = 0.0;
            return 0;
        }
    }
    double rpot = 0.0;
    double kappa3 = kappa * kappa * kappa;
    for (int i = 0; i < natoms; i++) {
        double sep[3];
        mi(full_ell, cell, p, xyz + 3 * i, sep);
        double d2 = normsq(sep);
        double dist = sqrt(d2);
        double kr = kappa * dist;
        double erfkr = mo_erfc(kr);
        double ri1 = 1.0 / dist;
        double q2 = q[i];
        rpot = rpot + charge1 * q2 * erfkr * ri1;
        double ri2 = ri1 * ri1;
        double ri3 = ri2 * ri1;
        double ek2r2 = exp(-(kr * kr));
        double temp = -(ri3 * erfkr) - 2.0 * kappa * ri2 * risqrtpi * ek2r2;
        hot[1] = hot[1] + sep[0] * (q2 * temp) * charge1;
        hot[2] = hot[2] + sep[1] * (q2 * temp) * charge1;
        hot[3] = hot[3] + sep[2] * (q2 * temp) * charge1;
        double ri4 = ri3 * ri1;
        double ri5 = ri4 * ri1;
        temp = 3.0 * ri5 * erfkr + 6.0 * kappa * risqrtpi * ri4 * ek2r2
               + 4.0 * kappa3 * risqrtpi * ri2 * ek2r2;
        hot[4] = hot[4] + charge1 * q2 * sep[0] * sep[1] * temp;
        hot[5] = hot[5] + charge1 * q2 * sep[0] * sep[2] * temp;
        hot[6] = hot[6] + charge1 * q2 * sep[1] * sep[2] * temp;
        double ri6 = ri5 * ri1;
        double ri7 = ri6 * ri1;
        temp = -(15.0 * ri7 * erfkr) - 30.0 * kappa * risqrtpi * ri6 * ek2r2
               - 20.0 * kappa3 * risqrtpi * ri4 * ek2r2
               - 8.0 * kappa3 * kappa * kappa * risqrtpi * ri2 * ek2r2;
        hot[7] = hot[7] + charge1 * q2 * sep[0] * sep[1] * sep[2] * temp;
    }
    double kpot = 0.0;
    for (int i = 0; i < nk; i++) {
        const double *kv = k + 3 * i;
        double rdotk = 0.0;
        rdotk = rdotk + p[0] * kv[0];
        rdotk = rdotk + p[1] * kv[1];
        rdotk = rdotk + p[2] * kv[2];
        double c = cos(rdotk), sn = sin(rdotk);
        double realPart = s[i] * c - si[i] * sn;
        double imagPart = s[i] * sn - si[i] * c;
        kpot = kpot + realPart * charge1;
        hot[1] = hot[1] - kv[0] * imagPart * charge1;
        hot[2] = hot[2] - kv[1] * imagPart * charge1;
        hot[3] = hot[3] - kv[2] * imagPart * charge1;
        hot[4] = hot[4] - kv[0] * kv[1] * realPart * charge1;
        hot[5] = hot[5] - kv[0] * kv[2] * realPart * charge1;
        hot[6] = hot[6] - kv[1] * kv[2] * realPart * charge1;
        hot[7] = hot[7] - kv[0] * kv[1] * kv[2] * imagPart * charge1;
    }
    double pot = 0.0 + rpot + kpot;
    pot = pot * MO_TOKCAL;
    out[0] = 0.0 + pot;
    for (int j = 1; j < 8; j++) out[j] = hot[j] * MO_TOKCAL;
    return 1;
}

int mo_ewald_node(const double p[3], int natoms, const double *xyz, const double *q,
                  double charge1, double kappa, int nk, const double *k,
                  const double *s, const double *si, const double full_ell[3],
                  double out[8])
{
    return ewald_node_c(p, natoms, xyz, q, charge1, kappa, nk, k, s, si, full_ell, NULL, out);
}

/* -------------------------------------------------------------- grid */
/* config.F90:762-786 */
void mo_grid_dims(const double eff[3], double gridspace, int npnts[3], double step[3])
{
    for (int i = 0; i < 3; i++) {
        npnts[i] = (int)(eff[i] / gridspace) + 1;
        step[i] = eff[i] / (double)(npnts[i] - 1);
    }
}

void mo_build_nodes(int kind, const mo_framework *fw, const int npnts[3], const double step[3],
                    long long m0, long long m1, double *vals)
{
    long long nx = npnts[0], ny = npnts[1];
    for (long long m = m0; m < m1; m++) {
        long long x = m % nx, y = (m / nx) % ny, z = m / (nx * ny);
        double p[3] = {(double)x * step[0], (double)y * step[1], (double)z * step[2]};
        double *o = vals + 8 * (m - m0);
        if (kind == 1)
            lj_node_c(p, fw->natoms, fw->xyz, fw->type, fw->A, fw->B, fw->hicut2,
                      fw->locut2, fw->full_ell, fw->cell, o);
        else if (kind == 2)
            ewald_node_c(p, fw->natoms, fw->xyz, fw->q, fw->charge1, fw->kappa, fw->nk,
                         fw->k, fw->s, fw->si, fw->full_ell, fw->cell, o);
        else
            buck_node_c(p, fw->natoms, fw->xyz, fw->type, fw->bA, fw->brho, fw->bC, fw->bD,
                        fw->hicut2, fw->locut2, fw->full_ell, fw->cell, o);
    }
}

/* one node at an arbitrary point p (same kinds as mo_build_nodes) */
int mo_node_at(int kind, const mo_framework *fw, const double p[3], double out[8])
{
    if (kind == 1)
        return lj_node_c(p, fw->natoms, fw->xyz, fw->type, fw->A, fw->B, fw->hicut2, fw->locut2, fw->full_ell,
                         fw->cell, out);
    if (kind == 2)
        return ewald_node_c(p, fw->natoms, fw->xyz, fw->q, fw->charge1, fw->kappa, fw->nk, fw->k, fw->s, fw->si,
                            fw->full_ell, fw->cell, out);
    return buck_node_c(p, fw->natoms, fw->xyz, fw->type, fw->bA, fw->brho, fw->bC, fw->bD, fw->hicut2, fw->locut2,
                       fw->full_ell, fw->cell, out);
}

/* simcell_uniformRandPt (simcell.F90:268-295): a uniform point of the
 * simulation cell, ell_i r_i when orthorhombic, else (lvec1 r1 + lvec2 r2)
 * + lvec3 r3 */
void mo_cell_uniform(const mo_cell *c, mo_rng *rng, double out[3])
{
    if (c->ortho) {
        for (int i = 0; i < 3; i++) out[i] = c->ell[i] * mo_rranf(rng);
        return;
    }
    double v[3][3];
    for (int k = 0; k < 3; k++) {
        double r = mo_rranf(rng);
        for (int i = 0; i < 3; i++) v[k][i] = c->lvec[k][i] * r;
    }
    for (int i = 0; i < 3; i++) out[i] = (v[0][i] + v[1][i]) + v[2][i];
}

/* simcell_getvolume -> fundcell_genvolume: vector_tripledotprod (vector.F90:
 * 405-429), the cross product l1 x l2 then its dot product with l3.
 * Pinned: the reference build's simcell_getvolume on the FCELL cells. */
double mo_cell_volume(const mo_cell *c)
{
    const double *l1 = c->lvec[0], *l2 = c->lvec[1], *l3 = c->lvec[2];
    double x = l1[1] * l2[2] - l1[2] * l2[1];
    double y = -(l1[0] * l2[2] - l1[2] * l2[0]);
    double z = l1[0] * l2[1] - l1[1] * l2[0];
    double v = 0.0;
    v = v + x * l3[0];
    v = v + y * l3[1];
    v = v + z * l3[2];
    return v;
}

/* -------------------------------------------------------- compaction */
/* pmap.F90:509-656.  Grid node (i,j,k) (1-based x,y,z) has source index
 * m = (i-1) + nx*((j-1) + ny*(k-1)) (pmap_addpoint assigns ptr = running
 * count in config_makegrid order); ptr is stored (k,j,i), z fastest. */
long long mo_compact(const int npnts[3], const float *r, double cutoff,
                     int32_t *ptr, long long *src_out, long long counts[4])
{
    long long nx = npnts[0], ny = npnts[1], nz = npnts[2];
    float cut = (float)cutoff;
    long long naccept = 0, reject = 0, acceptCut = 0, acceptNeighbor = 0;
#define PIDX(i, j, k) ((((long long)(i) - 1) * ny + ((j) - 1)) * nz + ((k) - 1))
#define MIDX(i, j, k) (((long long)(i) - 1) + nx * (((long long)(j) - 1) + ny * ((long long)(k) - 1)))
    memset(ptr, 0, sizeof(int32_t) * nx * ny * nz);
    for (long long i = 1; i <= nx; i++)
        for (long long j = 1; j <= ny; j++)
            for (long long k = 1; k <= nz; k++) {
                float v = r[MIDX(i, j, k)];
                if (v < cut) {
                    if (ptr[PIDX(i, j, k)] == 0) {
                        acceptCut++;
                        naccept++;
                        ptr[PIDX(i, j, k)] = (int32_t)naccept;
                    }
                    for (int ii = -1; ii <= 1; ii++)
                        for (int jj = -1; jj <= 1; jj++)
                            for (int kk = -1; kk <= 1; kk++) {
                                if (ii == 0 && jj == 0 && kk == 0) continue;
                                long long a = i + ii, b = j + jj, c = k + kk;
                                if (a > nx) a = 1;
                                if (b > ny) b = 1;
                                if (c > nz) c = 1;
                                if (a < 1) a = nx;
                                if (b < 1) b = ny;
                                if (c < 1) c = nz;
                                if (ptr[PIDX(a, b, c)] == 0 && r[MIDX(a, b, c)] >= cut) {
                                    acceptNeighbor++;
                                    naccept++;
                                    ptr[PIDX(a, b, c)] = (int32_t)naccept;
                                }
                            }
                } else {
                    if (ptr[PIDX(i, j, k)] != 0) continue;
                    reject++;
                }
            }
    if (src_out) {
        for (long long i = 1; i <= nx; i++)
            for (long long j = 1; j <= ny; j++)
                for (long long k = 1; k <= nz; k++) {
                    int32_t n = ptr[PIDX(i, j, k)];
                    if (n) src_out[n - 1] = MIDX(i, j, k);
                }
    }
#undef PIDX
#undef MIDX
    counts[0] = acceptCut;
    counts[1] = acceptNeighbor;
    counts[2] = naccept;
    counts[3] = reject;
    return naccept;
}

/* ---------------------------------------------------------- map file */
/* Fortran unformatted sequential records with 4-byte length markers. */
static int rec_write(FILE *f, const void *buf, size_t n)
{
    uint32_t m = (uint32_t)n;
    if (n > 0x7fffffffULL) return -1;
    if (fwrite(&m, 4, 1, f) != 1) return -1;
    if (n && fwrite(buf, 1, n, f) != n) return -1;
    if (fwrite(&m, 4, 1, f) != 1) return -1;
    return 0;
}

static void padcpy(char *dst, const char *src, size_t n)
{
    size_t l = strlen(src);
    memset(dst, ' ', n);
    memcpy(dst, src, l < n ? l : n);
}

/* mapheader.F90:178-232 then interpolate.F90:922-938 */
int mo_map_write(const char *path, int nzatoms, const mo_zatom *zatoms, float pcut_off,
                 const int npnts[3], long long nfsize, const int32_t *ptr, const float *vals8)
{
    FILE *f = fopen(path, "wb");
    if (!f) return -1;
    int rc = 0;
    char s40[40], s60[60], s25[25], s27[27];
    float rj = 0.0f;
    int32_t ij = 0;
    int32_t zt[11] = {0};
    padcpy(s40, "file contains header", 40);
    padcpy(s60, "Old type header 2003", 60);
    padcpy(s25, "probe", 25);
    rc |= rec_write(f, s40, 40);
    rc |= rec_write(f, s60, 60);
    rc |= rec_write(f, s25, 25);
    rc |= rec_write(f, &rj, 4);
    int32_t nz = nzatoms;
    rc |= rec_write(f, &nz, 4);
    rc |= rec_write(f, zt, sizeof zt);
    for (int i = 0; i < nzatoms; i++) {
        padcpy(s27, zatoms[i].name, 25);
        padcpy(s27 + 25, zatoms[i].symbol, 2);
        rc |= rec_write(f, s27, 27);
        rc |= rec_write(f, &zatoms[i].charge, 4);
        float abcd[4] = {zatoms[i].A12, zatoms[i].B6, zatoms[i].C2, zatoms[i].D0};
        rc |= rec_write(f, abcd, 16);
        float cut[2] = {zatoms[i].rcuthi, zatoms[i].rcutlo};
        rc |= rec_write(f, cut, 8);
    }
    rc |= rec_write(f, &ij, 4);
    rc |= rec_write(f, &pcut_off, 4);
    rc |= rec_write(f, &rj, 4);
    rc |= rec_write(f, &rj, 4);
    rc |= rec_write(f, &ij, 4);
    rc |= rec_write(f, &rj, 4);
    int32_t nb[3] = {npnts[0], npnts[1], npnts[2]};
    rc |= rec_write(f, nb, 12);
    int32_t nf = (int32_t)nfsize;
    rc |= rec_write(f, &nf, 4);
    size_t nn = (size_t)npnts[0] * npnts[1] * npnts[2];
    rc |= rec_write(f, ptr, 4 * nn);
    for (int a = 0; a < 8; a++) rc |= rec_write(f, vals8 + (size_t)a * nfsize, 4 * (size_t)nfsize);
    if (fclose(f)) rc = -1;
    return rc ? -1 : 0;
}

static int rec_read(FILE *f, void *buf, size_t cap, size_t *nout)
{
    uint32_t m1, m2;
    if (fread(&m1, 4, 1, f) != 1) return -1;
    if (m1 > cap) return -2;
    if (m1 && fread(buf, 1, m1, f) != m1) return -1;
    if (fread(&m2, 4, 1, f) != 1 || m1 != m2) return -1;
    if (nout) *nout = m1;
    return 0;
}

static int map_read_header(FILE *f, int npnts[3], long long *nfsize, float *pcut, int *nzat)
{
    char buf[128];
    size_t n;
    if (rec_read(f, buf, sizeof buf, &n) || n != 40 || strncmp(buf, "file contains header", 20)) return -1;
    if (rec_read(f, buf, sizeof buf, &n)) return -1;
    if (rec_read(f, buf, sizeof buf, &n)) return -1;
    if (rec_read(f, buf, sizeof buf, &n)) return -1;
    int32_t nz;
    if (rec_read(f, &nz, 4, &n)) return -1;
    if (nzat) *nzat = nz;
    if (rec_read(f, buf, sizeof buf, &n)) return -1;
    for (int i = 0; i < nz; i++)
        for (int r = 0; r < 4; r++)
            if (rec_read(f, buf, sizeof buf, &n)) return -1;
    if (rec_read(f, buf, sizeof buf, &n)) return -1;
    float pc;
    if (rec_read(f, &pc, 4, &n)) return -1;
    if (pcut) *pcut = pc;
    for (int r = 0; r < 4; r++)
        if (rec_read(f, buf, sizeof buf, &n)) return -1;
    int32_t nb[3];
    if (rec_read(f, nb, 12, &n) || n != 12) return -1;
    int32_t nf;
    if (rec_read(f, &nf, 4, &n)) return -1;
    npnts[0] = nb[0];
    npnts[1] = nb[1];
    npnts[2] = nb[2];
    *nfsize = nf;
    return 0;
}

int mo_map_read_sizes(const char *path, int npnts[3], long long *nfsize, float *pcut_off, int *nzatoms)
{
    FILE *f = fopen(path, "rb");
    if (!f) return -1;
    int rc = map_read_header(f, npnts, nfsize, pcut_off, nzatoms);
    fclose(f);
    return rc;
}

int mo_map_read(const char *path, int32_t *ptr, float *vals8)
{
    FILE *f = fopen(path, "rb");
    if (!f) return -1;
    int np[3];
    long long nf;
    int rc = map_read_header(f, np, &nf, NULL, NULL);
    size_t nn = (size_t)np[0] * np[1] * np[2];
    size_t n;
    if (!rc) rc = rec_read(f, ptr, 4 * nn, &n);
    for (int a = 0; a < 8 && !rc; a++) rc = rec_read(f, vals8 + (size_t)a * nf, 4 * (size_t)nf, &n);
    fclose(f);
    return rc;
}

/* ---------------------------------------------------- interpolation */
/* simcell.F90:634-668, orthorhombic, origin 0 */
void mo_maptouc(const double fund_ell[3], const double r[3], double out[3])
{
    for (int i = 0; i < 3; i++) {
        double ell_inv = 1.0 / fund_ell[i];
        double d = r[i] * ell_inv;
        if (d >= 0.0) d = fmod(d, 1.0);
        else d = fmod(d, 1.0) + 1.0;
        out[i] = d * fund_ell[i];
    }
}

static inline double h0f(double x) { return x * x * (3.0 - 2.0 * x); }
static inline double h1f(double x, double ds) { return ds * x * x * (x - 1.0); }
static inline double hd0f(double x) { return 6.0 * x * (1.0 - x); }
static inline double hd1f(double x, double ds) { return ds * x * (3.0 * x - 2.0); }

/* shared setup of pothermite/forcehermite (interpolate.F90:641-723) */
static int herm_setup(const int npnts[3], const double step[3], const int32_t *ptr,
                      const float *vals8, long long nfsize, const double coord[3],
                      double ft[64], double tuv[3], int wrap)
{
    int idx[3], idxp[3];
    for (int d = 0; d < 3; d++) {
        idx[d] = (int)(coord[d] / step[d] + 1.0);
        tuv[d] = (coord[d] - (double)(idx[d] - 1) * step[d]) / step[d];
        idxp[d] = idx[d] + 1;
        /* the wrap of the upper corner holds for orthorhombic cells only
         * (interpolate.F90:675-679); without it a point in the last
         * cubelet has no upper corner and fails */
        if (idx[d] == npnts[d]) {
            if (!wrap) return 0;
            idxp[d] = 1;
        }
        if (idx[d] < 1 || idx[d] > npnts[d]) return 0; /* outside the grid */
    }
    long long ny = npnts[1], nz = npnts[2];
#define P(i, j, k) ptr[(((long long)(i) - 1) * ny + ((j) - 1)) * nz + ((k) - 1)]
    int ii = idx[0], jj = idx[1], kk = idx[2], ip = idxp[0], jp = idxp[1], kp = idxp[2];
    int32_t mi[8] = {P(ii, jj, kk), P(ip, jj, kk), P(ii, jp, kk), P(ii, jj, kp),
                     P(ip, jp, kk), P(ip, jj, kp), P(ii, jp, kp), P(ip, jp, kp)};
#undef P
    for (int i = 0; i < 8; i++)
        if (mi[i] == 0) return 0;
    for (int i = 0; i < 8; i++)
        for (int a = 0; a < 8; a++) ft[8 * i + a] = (double)vals8[(size_t)a * nfsize + (mi[i] - 1)];
    return 1;
}

/* corner pattern (interpolate.F90:725-835): for corner i, which of the two
 * 1-D basis functions (lower/upper) applies per axis */
static const int CX[8] = {0, 1, 0, 0, 1, 1, 0, 1};
static const int CY[8] = {0, 0, 1, 0, 1, 0, 1, 1};
static const int CZ[8] = {0, 0, 0, 1, 0, 1, 1, 1};
/* value slot a of a corner uses the derivative basis (h1) along x/y/z */
static const int DX[8] = {0, 1, 0, 0, 1, 1, 0, 1};
static const int DY[8] = {0, 0, 1, 0, 1, 0, 1, 1};
static const int DZ[8] = {0, 0, 0, 1, 0, 1, 1, 1};

static void basis(double t, double step, double h[2][2])
{
    /* h[lower/upper][value/deriv basis] */
    h[0][0] = h0f(1 - t);
    h[1][0] = h0f(t);
    h[0][1] = h1f(1 - t, -step);
    h[1][1] = h1f(t, step);
}

static void dbasis(double t, double step, double h[2][2])
{
    h[0][0] = -hd0f(1 - t);
    h[1][0] = hd0f(t);
    h[0][1] = -hd1f(1 - t, -step);
    h[1][1] = hd1f(t, step);
}

/* interpolate.F90:603-854 */
static int pothermite_w(const int npnts[3], const double step[3], const int32_t *ptr,
                        const float *vals8, long long nfsize, const double coord[3], double *p, int wrap)
{
    double ft[64], tuv[3];
    if (!herm_setup(npnts, step, ptr, vals8, nfsize, coord, ft, tuv, wrap)) {
        *p = 0.0;
        return 0;
    }
    double hx[2][2], hy[2][2], hz[2][2];
    basis(tuv[0], step[0], hx);
    basis(tuv[1], step[1], hy);
    basis(tuv[2], step[2], hz);
    double sum1 = 0.0;
    for (int i = 0; i < 8; i++)
        for (int a = 0; a < 8; a++) {
            double hi = hx[CX[i]][DX[a]], hj = hy[CY[i]][DY[a]], hk = hz[CZ[i]][DZ[a]];
            sum1 = sum1 + ft[8 * i + a] * hi * hj * hk;
        }
    *p = sum1;
    return 1;
}

int mo_pothermite(const int npnts[3], const double step[3], const int32_t *ptr,
                  const float *vals8, long long nfsize, const double coord[3], double *p)
{
    return pothermite_w(npnts, step, ptr, vals8, nfsize, coord, p, 1);
}

int mo_pothermite_nowrap(const int npnts[3], const double step[3], const int32_t *ptr, const float *vals8,
                         long long nfsize, const double coord[3], double *p)
{
    return pothermite_w(npnts, step, ptr, vals8, nfsize, coord, p, 0);
}

/* interpolate.F90:189-584 */
int mo_forcehermite(const int npnts[3], const double step[3], const int32_t *ptr,
                    const float *vals8, long long nfsize, const double coord[3], double p[4])
{
    double ft[64], tuv[3];
    if (!herm_setup(npnts, step, ptr, vals8, nfsize, coord, ft, tuv, 1)) {
        p[0] = p[1] = p[2] = p[3] = 0.0;
        return 0;
    }
    double hx[2][2], hy[2][2], hz[2][2], dx[2][2], dy[2][2], dz[2][2];
    basis(tuv[0], step[0], hx);
    basis(tuv[1], step[1], hy);
    basis(tuv[2], step[2], hz);
    dbasis(tuv[0], step[0], dx);
    dbasis(tuv[1], step[1], dy);
    dbasis(tuv[2], step[2], dz);
    double s1 = 0, s2 = 0, s3 = 0, s4 = 0;
    for (int i = 0; i < 8; i++)
        for (int a = 0; a < 8; a++) {
            double f = ft[8 * i + a];
            double hi = hx[CX[i]][DX[a]], hj = hy[CY[i]][DY[a]], hk = hz[CZ[i]][DZ[a]];
            double di = dx[CX[i]][DX[a]], dj = dy[CY[i]][DY[a]], dk = dz[CZ[i]][DZ[a]];
            s1 = s1 + f * hi * hj * hk;
            s2 = s2 + f * di * hj * hk;
            s3 = s3 + f * hi * dj * hk;
            s4 = s4 + f * hi * hj * dk;
        }
    p[0] = s1;
    p[1] = -s2 / step[0];
    p[2] = -s3 / step[1];
    p[3] = -s4 / step[2];
    return 1;
}
