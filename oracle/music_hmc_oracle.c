/*
 * music_hmc_oracle.c -- CPU FP64 restatement of Music's HMC burst for
 * flexible sorbates.  TEST INFRASTRUCTURE ONLY (see music_hmc_oracle.h).
 *
 * Pinning (tests/test_hmc_oracle.py): forces against central finite
 * differences of the energy for every term (closed form), energy
 * conservation of the NVE integrator as dt -> 0 (second-order drift), the
 * Box-Muller draws against their closed form from the rranf stream, and the
 * torsion against the textbook dihedral (polymer convention: cos = -cos of
 * the IUPAC angle, cosexpansion.F90:142).
 */
#include "music_hmc_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define KJMOLE_KB 8.314e-3  /* defaults.F90 */
#define SCALEKE 0.01         /* (A/ps)^2 amu -> kJ/mol, defaults.F90:76 */
#define SCALEF 418.4         /* kcal/mol/A / amu -> A/ps^2, defaults.F90:73 */
#define CALTOJ 4.184

static double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

/* harmonic_getinteraction (harmonic.F90:130-160): vr = r2 - r1 */
static double bond_term(double k, double req, const double *r1, const double *r2, double *f1, double *f2)
{
    double vr[3] = {r2[0] - r1[0], r2[1] - r1[1], r2[2] - r1[2]};
    double r = sqrt(dot3(vr, vr));
    double vru[3] = {0, 0, 0};
    if (r != 0.0)
        for (int i = 0; i < 3; i++) vru[i] = vr[i] / r;
    double u = 0.5 * k * (r - req) * (r - req);
    double f = -k * (r - req);
    if (f1)
        for (int i = 0; i < 3; i++) {
            f2[i] += vru[i] * f;
            f1[i] += vru[i] * f * (-1.0);
        }
    return u;
}

/* harmonica (harmonica.F90:100-160): angle 1-2-3 */
static double angle_term(double k, double t0, const double *r1, const double *r2, const double *r3, double *f1,
                         double *f2, double *f3)
{
    double v1[3] = {r1[0] - r2[0], r1[1] - r2[1], r1[2] - r2[2]};
    double v2[3] = {r3[0] - r2[0], r3[1] - r2[1], r3[2] - r2[2]};
    double r12 = dot3(v1, v1), r22 = dot3(v2, v2);
    double r1n = sqrt(r12), r2n = sqrt(r22);
    double adotb = dot3(v2, v1);
    double ct = adotb / (r1n * r2n);
    double th = acos(ct);
    double u = 0.5 * k * (th - t0) * (th - t0);
    if (f1) {
        double st = sqrt(1.0 - ct * ct);
        double at = (fabs(st) < 1.0e-10) ? 0.0 : k * (th - t0) / (r1n * r2n * st);
        double a1[3], a3[3];
        for (int i = 0; i < 3; i++) {
            a1[i] = (v2[i] - v1[i] * (adotb / r12)) * at;
            a3[i] = (v1[i] - v2[i] * (adotb / r22)) * at;
            f1[i] += a1[i];
            f3[i] += a3[i];
            f2[i] += (a1[i] + a3[i]) * (-1.0);
        }
    }
    return u;
}

/* cosexpansion_getinteraction (cosexpansion.F90:88-196), ifvec = .False. */
static double torsion_term(const double *cn, const double *r1, const double *r2, const double *r3, const double *r4,
                           double *f1, double *f2, double *f3, double *f4)
{
    double a[3], b[3], c[3];
    for (int i = 0; i < 3; i++) {
        a[i] = r2[i] - r1[i];
        b[i] = r3[i] - r2[i];
        c[i] = r4[i] - r3[i];
    }
    double aa = dot3(a, a), ab = dot3(a, b), ac = dot3(a, c), bb = dot3(b, b), bc = dot3(b, c), cc = dot3(c, c);
    double T1 = ab * bc - ac * bb, T2 = aa * bb - ab * ab, T3 = bb * cc - bc * bc;
    if (fabs(T2 * T3) < 1.0e-11) return 0.0;
    double inv = 1.0 / sqrt(T2 * T3), T1T2 = T1 / T2, T1T3 = T1 / T3;
    double cs = -(T1 * inv);
    double term = 1.0, sum = 0.0;
    for (int i = 0; i < 6; i++) {
        sum = sum + cn[i] * term;
        term = term * cs;
    }
    if (f1) {
        double da[3], db[3], dc[3];
        for (int i = 0; i < 3; i++) {
            da[i] = ((a[i] * (T1T2 * bb)) + (c[i] * bb) - (b[i] * (bc + ab * T1T2))) * inv;
            db[i] = ((b[i] * (2 * ac + T1T2 * aa + T1T3 * cc)) - (a[i] * (bc + ab * T1T2)) - (c[i] * (ab + T1T3 * bc))) * inv;
            dc[i] = ((a[i] * bb) + (c[i] * (T1T3 * bb)) - (b[i] * (ab + bc * T1T3))) * inv;
        }
        term = 1.0;
        double s = 0.0;
        for (int i = 1; i < 6; i++) {
            s = s + (double)i * cn[i] * term;
            term = term * cs;
        }
        double ffac = -s;
        for (int i = 0; i < 3; i++) {
            f1[i] += (da[i] * (-1.0)) * ffac;
            f2[i] += (da[i] - db[i]) * ffac;
            f3[i] += (db[i] - dc[i]) * ffac;
            f4[i] += dc[i] * ffac;
        }
    }
    return sum;
}

/* One bonded term on its own (for checks against the reference's
 * harmonic / harmonica / cosexpansion): kind 1 bond (par k, req; 2 atoms),
 * 2 angle (par k, theta0; 3 atoms), 3 torsion (par c0..c5; 4 atoms).
 * r[natoms][3] -> energy, forces f[natoms][3] (overwritten). */
double mo_bonded_term(int kind, const double *par, const double *r, double *f)
{
    for (int i = 0; i < 12; i++) f[i] = 0.0;
    if (kind == 1) return bond_term(par[0], par[1], r, r + 3, f, f + 3);
    if (kind == 2) return angle_term(par[0], par[1], r, r + 3, r + 6, f, f + 3, f + 6);
    return torsion_term(par, r, r + 3, r + 6, r + 9, f, f + 3, f + 6, f + 9);
}

static double energy_forces(const mo_hmc_sys *s, int N, const double *rp, const double *r, double *F, int *ok,
                            double *parts);

double mo_hmc_energy_forces(const mo_hmc_sys *s, int N, const double *rp, const double *r, double *F, int *ok)
{
    return energy_forces(s, N, rp, r, F, ok, NULL);
}

double mo_hmc_energy_parts(const mo_hmc_sys *s, int N, const double *rp, const double *r, double *F, int *ok,
                           double *parts)
{
    for (int k = 0; k < MO_HMC_NPARTS; k++) parts[k] = 0.0;
    return energy_forces(s, N, rp, r, F, ok, parts);
}

/* the total e is summed term by term in the order below; parts (may be NULL)
 * collects the same terms by kind: map, LJ, stretch, bending, torsion */
static double energy_forces(const mo_hmc_sys *s, int N, const double *rp, const double *r, double *F, int *ok,
                            double *parts)
{
#define PART(k, x) \
    do {           \
        double t_ = (x); \
        e += t_;   \
        if (parts) parts[k] += t_; \
    } while (0)
    const int ns = s->ns;
    *ok = 1;
    if (F) memset(F, 0, sizeof(double) * 3 * (size_t)N * ns);
    double e = 0.0;
    /* framework maps (forcehermite at the folded site position) */
    for (int m = 0; m < N; m++)
        for (int a = 0; a < ns; a++) {
            int k = s->site_map[a];
            if (k < 0) continue;
            const double *ri = r + ((size_t)m * ns + a) * 3;
            double c[3], p[4];
            mo_maptouc(s->fund_ell, ri, c);
            if (!mo_forcehermite(s->map_npnts[k], s->map_step[k], s->map_ptr[k], s->map_vals[k], s->map_nf[k], c, p)) {
                *ok = 0;
                return 0.0;
            }
            PART(0, p[0]);
            if (F)
                for (int i = 0; i < 3; i++) F[((size_t)m * ns + a) * 3 + i] += p[1 + i];
        }
    /* sorbate-sorbate LJ, molecules m < n */
    for (int m = 0; m < N; m++)
        for (int n = m + 1; n < N; n++)
            for (int a = 0; a < ns; a++)
                for (int b = 0; b < ns; b++) {
                    const double *ri = r + ((size_t)m * ns + a) * 3, *rj = r + ((size_t)n * ns + b) * 3;
                    double sep[3];
                    mo_minimage(s->full_ell, ri, rj, sep);
                    int t = s->stype[a] * s->nst + s->stype[b];
                    if (!s->lj_on[t]) continue;
                    double r2 = dot3(sep, sep);
                    if (r2 > s->lj_h2[t]) continue;
                    if (r2 < s->lj_l2[t]) {
                        *ok = 0;
                        return 0.0;
                    }
                    double r2i = 1.0 / r2, r6i = r2i * r2i * r2i, r12i = r6i * r6i;
                    PART(1, s->lj_A[t] * r12i - s->lj_B[t] * r6i);
                    if (F) {
                        double f = (12.0 * s->lj_A[t] * r12i - 6.0 * s->lj_B[t] * r6i) * r2i;
                        for (int i = 0; i < 3; i++) {
                            F[((size_t)m * ns + a) * 3 + i] += sep[i] * f;
                            F[((size_t)n * ns + b) * 3 + i] -= sep[i] * f;
                        }
                    }
                }
    /* intramolecular terms on the principal (unwrapped) coordinates */
    for (int m = 0; m < N; m++) {
        const double *R = rp + (size_t)m * ns * 3;
        double *G = F ? F + (size_t)m * ns * 3 : NULL;
#define S3(i) (G ? G + 3 * (i) : NULL)
        for (int q = 0; q < s->nb; q++)
            PART(2, bond_term(s->bk[q], s->br0[q], R + 3 * s->bi[q][0], R + 3 * s->bi[q][1], S3(s->bi[q][0]),
                              S3(s->bi[q][1])));
        for (int q = 0; q < s->na; q++)
            PART(3, angle_term(s->ak[q], s->at0[q], R + 3 * s->ai[q][0], R + 3 * s->ai[q][1], R + 3 * s->ai[q][2],
                               S3(s->ai[q][0]), S3(s->ai[q][1]), S3(s->ai[q][2])));
        for (int q = 0; q < s->nt; q++)
            PART(4, torsion_term(s->tc[q], R + 3 * s->ti[q][0], R + 3 * s->ti[q][1], R + 3 * s->ti[q][2],
                                 R + 3 * s->ti[q][3], S3(s->ti[q][0]), S3(s->ti[q][1]), S3(s->ti[q][2]),
                                 S3(s->ti[q][3])));
#undef S3
    }
#undef PART
    return e;
}

/* random_gaussian (random.F90:197-217), Box-Muller with the saved value */
double mo_hmc_gaussian(mo_hmc_state *st, double ave, double sigma)
{
    if (!st->g_have) {
        double r1 = mo_rranf(&st->rng), r2 = mo_rranf(&st->rng);
        if (r1 < 1.0e-10) r1 = 1.0e-10;
        if (r2 < 1.0e-10) r2 = 1.0e-10;
        const double twopi = 2.0 * 3.1415927;
        double z1 = sqrt(-2.0 * log(r1)) * cos(twopi * r2);
        st->g_z2 = sqrt(-2.0 * log(r1)) * sin(twopi * r2);
        st->g_have = 1;
        return ave + sigma * z1;
    }
    st->g_have = 0;
    return ave + sigma * st->g_z2;
}

static double kinetic(const mo_hmc_sys *s, int N, const double *v)
{
    double ke = 0.0;
    for (int m = 0; m < N; m++)
        for (int a = 0; a < s->ns; a++) {
            const double *vv = v + ((size_t)m * s->ns + a) * 3;
            ke = ke + dot3(vv, vv) * s->mass[a];
        }
    return 0.5 * ke * SCALEKE; /* config_kineticEnergy, config.F90:2474-2519 */
}

static void accel(const mo_hmc_sys *s, int N, const double *F, double *acc)
{
    for (int m = 0; m < N; m++)
        for (int a = 0; a < s->ns; a++)
            for (int i = 0; i < 3; i++) {
                size_t o = ((size_t)m * s->ns + a) * 3 + i;
                acc[o] = F[o] * (1.0 / s->mass[a]) * SCALEF; /* config_conv2accel, config.F90:1951-1968 */
            }
}

/* velverlet_nveIntegrate (fast branch), one step */
static int vv_step(const mo_hmc_sys *s, int N, double *rp, double *r, double *v, double *acc, double *F, double *U)
{
    const size_t n3 = (size_t)N * s->ns * 3;
    const double dt = s->dt, dt2 = dt * dt * 0.5;
    for (size_t o = 0; o < n3; o++) {
        rp[o] = rp[o] + v[o] * dt + acc[o] * dt2;
        r[o] = r[o] + v[o] * dt + acc[o] * dt2;
        v[o] = v[o] + acc[o] * 0.5 * dt;
    }
    int ok;
    *U = mo_hmc_energy_forces(s, N, rp, r, F, &ok);
    if (!ok || *U > s->max_nrg_md) return 0;
    accel(s, N, F, acc);
    for (size_t o = 0; o < n3; o++) v[o] = v[o] + acc[o] * 0.5 * dt;
    return 1;
}

int mo_hmc_nve(const mo_hmc_sys *s, int N, double *rp, double *r, double *v, int nsteps, double *etot)
{
    const size_t n3 = (size_t)N * s->ns * 3;
    double *F = malloc(sizeof(double) * n3), *acc = malloc(sizeof(double) * n3);
    int ok;
    double U = mo_hmc_energy_forces(s, N, rp, r, F, &ok);
    int good = ok;
    if (good) {
        accel(s, N, F, acc);
        etot[0] = U + kinetic(s, N, v) / CALTOJ;
        for (int k = 0; k < nsteps && good; k++) {
            good = vv_step(s, N, rp, r, v, acc, F, &U);
            if (good) etot[k + 1] = U + kinetic(s, N, v) / CALTOJ;
        }
    }
    free(F);
    free(acc);
    return good;
}

/* md_doSim with velverlet_nveIntegrate (md.F90:592-623, velverlet.F90:148-288):
 * nsteps steps from rp, r, v; after each step trace[step][MO_MD_NQ] = energy
 * parts, sum m v^2, sum_m |sum m v|^2 / M^2 (config_kineticTemp,
 * config.F90:2528-2615), total momentum.  Returns the number of steps done
 * (fewer on a failed evaluation). */
int mo_hmc_md(const mo_hmc_sys *s, int N, double *rp, double *r, double *v, int nsteps, double *trace)
{
    const int ns = s->ns;
    const size_t n3 = (size_t)N * ns * 3;
    double *F = malloc(sizeof(double) * n3), *acc = malloc(sizeof(double) * n3);
    const double dt = s->dt, dt2 = dt * dt * 0.5;
    double mmass = 0.0, parts[MO_HMC_NPARTS];
    for (int a = 0; a < ns; a++) mmass += s->mass[a];
    const double mmassi2 = 1.0 / mmass / mmass;
    int ok, done = 0;
    energy_forces(s, N, rp, r, F, &ok, NULL);
    if (ok) {
        accel(s, N, F, acc);
        for (int k = 0; k < nsteps; k++) {
            for (size_t o = 0; o < n3; o++) {
                rp[o] = rp[o] + v[o] * dt + acc[o] * dt2;
                r[o] = r[o] + v[o] * dt + acc[o] * dt2;
                v[o] = v[o] + acc[o] * 0.5 * dt;
            }
            double U = mo_hmc_energy_parts(s, N, rp, r, F, &ok, parts);
            if (!ok || U > s->max_nrg_md) break;
            accel(s, N, F, acc);
            for (size_t o = 0; o < n3; o++) v[o] = v[o] + acc[o] * 0.5 * dt;
            double *t = trace + (size_t)k * MO_MD_NQ;
            for (int q = 0; q < MO_HMC_NPARTS; q++) t[q] = parts[q];
            double k2 = 0.0, km = 0.0, P[3] = {0.0, 0.0, 0.0};
            for (int m = 0; m < N; m++) {
                double q[3] = {0.0, 0.0, 0.0};
                for (int a = 0; a < ns; a++) {
                    const double *vv = v + ((size_t)m * ns + a) * 3;
                    k2 = k2 + dot3(vv, vv) * s->mass[a];
                    for (int i = 0; i < 3; i++) {
                        q[i] = vv[i] * s->mass[a] + q[i];
                        P[i] += s->mass[a] * vv[i];
                    }
                }
                km = dot3(q, q) * mmassi2 + km;
            }
            t[5] = k2;
            t[6] = km;
            for (int i = 0; i < 3; i++) t[7 + i] = P[i];
            done++;
        }
    }
    free(F);
    free(acc);
    return done;
}

int mo_hmc_move(const mo_hmc_sys *s, mo_hmc_state *st)
{
    const int N = st->N, ns = s->ns;
    const size_t n3 = (size_t)N * ns * 3;
    st->att++;
    /* hybridmoves_boltzVel: per molecule, per atom, three Gaussians */
    for (int m = 0; m < N; m++)
        for (int a = 0; a < ns; a++) {
            double sigma = sqrt((KJMOLE_KB * s->T) / (SCALEKE * s->mass[a]));
            for (int i = 0; i < 3; i++) st->v[((size_t)m * ns + a) * 3 + i] = mo_hmc_gaussian(st, 0.0, sigma);
        }
    double *F = malloc(sizeof(double) * (n3 + 1)), *acc = malloc(sizeof(double) * (n3 + 1));
    double *rp0 = malloc(sizeof(double) * (n3 + 1)), *r0 = malloc(sizeof(double) * (n3 + 1));
    int ok;
    double U0 = mo_hmc_energy_forces(s, N, st->rp, st->r, F, &ok);
    if (!ok || U0 > s->max_nrg0) {
        free(F), free(acc), free(rp0), free(r0);
        return -1;
    }
    double ke0 = kinetic(s, N, st->v);
    memcpy(rp0, st->rp, sizeof(double) * n3);
    memcpy(r0, st->r, sizeof(double) * n3);
    accel(s, N, F, acc);
    double U1 = U0;
    int good = 1;
    for (int k = 0; k < s->nmd && good; k++) good = vv_step(s, N, st->rp, st->r, st->v, acc, F, &U1);
    int accepted = 0;
    if (good) {
        double ke1 = kinetic(s, N, st->v);
        double dU = (U1 - U0) + (ke1 - ke0) / CALTOJ;
        st->last_dH = dU;
        double rti = 1.0 / (MO_KCALMOLE_KB * s->T);
        double u = -rti * dU;
        u = u + log(1.0);
        if (u > MO_MAX_EXP) u = MO_MAX_EXP;
        if (u < -MO_MAX_EXP) u = -MO_MAX_EXP;
        double boltz = exp(u);
        accepted = boltz > mo_rranf(&st->rng);
    } else {
        st->fail++;
        st->last_dH = NAN;
    }
    if (accepted) {
        st->acc++;
        st->energy = U1;
        for (size_t o = 0; o < n3; o++) { /* simcell_pbc of the simulation-cell coordinates */
            int i = (int)(o % 3);
            st->r[o] = st->r[o] - floor(st->r[o] / s->full_ell[i]) * s->full_ell[i];
        }
    } else {
        st->energy = U0;
        memcpy(st->rp, rp0, sizeof(double) * n3);
        memcpy(st->r, r0, sizeof(double) * n3);
    }
    free(F), free(acc), free(rp0), free(r0);
    return accepted;
}
