/*
 * music_gcmc_oracle.c -- CPU FP64 restatement of the GCMC move loop for
 * rigid sorbates (RINSERT / RDELETE / RTRANSLATE / RROTATE) with energies
 * from potential maps, sorbate-sorbate LJ with minimum image and (a new
 * capability the reference lacks) Ewald electrostatics with structure-factor
 * deltas.  TEST INFRASTRUCTURE ONLY: checker for the HIP GCMC kernel.
 *
 * Step structure and RNG draw order follow:
 *   gcmc_dosim            gcmc.F90:240-250  (species draw every step)
 *   gcmcmoves_move        gcmcmoves.F90:270-364 (move draw, molecule draw)
 *   mcmoveset_pickmove    mcmoveset.F90:314-335
 *   mcmoves_insert        mcmoves.F90:476-566 (no draw when the energy fails)
 *   mcmoves_delete        mcmoves.F90:578-647
 *   mcmoves_perturb       mcmoves.F90:352-463
 *   mcmoves_attempt       mcmoves.F90:290-335 (clamp +-30, boltz > rranf())
 *   rigidmoves_rinsert    rigidmoves.F90:1313-1367 (3 draws, +3 if multi-site)
 *   rigidmoves_randomtranslate / randomrotate  rigidmoves.F90:691-864
 *   rigidmoves_rotperturb rigidmoves.F90:1675-1720
 *   rigidcoords_toxyz     rigidcoords.F90:254-276, matrix_b2feuler matrix.F90:383-403
 *   simcell_pbc           simcell.F90:521-547
 *   config_delmol         config.F90:1401-1413 (last molecule moves into the slot)
 *   rigidmoves_binsert    rigidmoves.F90:1386-1442, rigidmoves_comFromBmap :1938-1969
 *                         (index, 3 in-cubelet, 3 unit-cell draws, +3 if multi-site)
 *   rigidmoves_bdelete    rigidmoves.F90:1578-1654 (bias of the molecule's cubelet)
 *   bmap_cubeinit / bmap_getbiasindex / bmap_getbiaspt / bmap_getcellwt
 *                         bmap.F90:181-636
 *   rigidmoves_getConfigFromCavBias  rigidmoves.F90:1987-2126 (cavity bias:
 *                         ntrials x (3 or 7 centre draws, +3 if multi-site),
 *                         then one selection draw when a trial succeeded)
 *   cavitylist_getIndexVec / _checkcavityfromIndex / _updateMolecInfo
 *                         cavitylist.F90:351-546, 1137-1157; stats_update
 *                         stats.F90:152-195
 *   Cavity bias is pinned by what the mathematics fixes (occupancy recounts,
 *   the running average's closed form, the ideal-gas mean in a tabulated
 *   quarter cell: tests/test_cavity_oracle.py); no reference output covers
 *   it.
 * Molecule energy = map terms (interpolate_int per site, factor q for the
 * emap) + LJ pairs with every other molecule (lj_snglint, nderivs = 0) +
 * Ewald terms.  Failure (LJ locut, untabulated map corner, U > max_nrg)
 * rejects the move without an RNG draw.
 *
 * Ewald (no reference implementation; ssdriver.F90:489-498 stops):
 *   real space  sum q_a q_b erfc(kappa r)/r, r < rc (A&S erfc, utils.F90)
 *   reciprocal  E = sum_{k in half space} w_k |S_k|^2,
 *               w_k = (4 pi / V) exp(-k^2 / 4 kappa^2) / k^2,
 *               k = twopi (mx/Lx, my/Ly, mz/Lz), |m_i| <= kmax
 *   self        -kappa/sqrt(pi) sum q_a^2 per molecule
 *   intra       -sum_{a<b} q_a q_b erf(kappa r_ab)/r_ab per molecule
 * all times 332.064 (ewald.F90:60).  Pinned by closed forms in
 * tests/test_gcmc_oracle.py (delta = full recompute difference; total
 * Ewald of a small neutral system against a direct lattice sum).
 */
#include "music_gcmc_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define TWOPI_R (2.0 * 3.1415927)

/* rigidcoords_toxyz (rigidcoords.F90:254-276) with matrix_b2feuler
 * (matrix.F90:383-403): rp = com + R(phi, theta, psi) ref, the product in
 * matrix_m_mult_v's order (matrix.F90:137-150); a single site sits at com */
void mo_euler_xyz(int ns, const double *ref, const double com[3], const double ang[3], double *rp)
{
    if (ns == 1) {
        rp[0] = com[0];
        rp[1] = com[1];
        rp[2] = com[2];
        return;
    }
    double phi = ang[0], theta = ang[1], psi = ang[2];
    double sp = sin(phi), cp = cos(phi), st = sin(theta), ct = cos(theta), ss = sin(psi), cs = cos(psi);
    double m[3][3];
    m[0][0] = cs * cp - ct * sp * ss;
    m[0][1] = -ss * cp - ct * sp * cs;
    m[0][2] = st * sp;
    m[1][0] = cs * sp + ct * cp * ss;
    m[1][1] = -ss * sp + ct * cp * cs;
    m[1][2] = -st * cp;
    m[2][0] = ss * st;
    m[2][1] = cs * st;
    m[2][2] = ct;
    for (int a = 0; a < ns; a++) {
        const double *v = ref + 3 * a;
        for (int i = 0; i < 3; i++) {
            double d = 0.0;
            d = d + m[i][0] * v[0];
            d = d + m[i][1] * v[1];
            d = d + m[i][2] * v[2];
            rp[3 * a + i] = com[i] + d;
        }
    }
}

static void toxyz(const mo_gcmc_sys *s, const double com[3], const double ang[3], double *rp)
{
    mo_euler_xyz(s->ns, &s->ref[0][0], com, ang, rp);
}

/* coulomb_snglint (coulomb.F90:258-327) for one site pair with charges
 * from the molecules (charges_fromAA_FILE false): skipped beyond hicut when
 * hicut > 0, failure (return 0) below locut, else
 * 331.5 shield q1 q2 (1/r - 1/hicut), or 1/r when hicut = 0. */
int mo_coul_pair(double hicut, double locut, double shield, double r2, double q1, double q2, double *e)
{
    *e = 0.0;
    if (hicut > 0.0 && r2 > hicut * hicut) return 1;
    if (r2 < locut * locut) return 0;
    double rr = sqrt(r2);
    if (hicut > 0.0)
        *e = (331.5 * shield * q1 * q2) * (1.0 / rr - 1.0 / hicut);
    else
        *e = (331.5 * shield * q1 * q2) * (1.0 / rr);
    return 1;
}

static inline int slanted(const mo_gcmc_sys *s)
{
    return s->cell && !s->cell->ortho;
}

/* simcell_pbcarray (simcell.F90:521-547) of ns sites in the cell of s */
static void pbc(const mo_gcmc_sys *s, int ns, const double *rp, double *r)
{
    if (slanted(s)) {
        for (int a = 0; a < ns; a++) mo_cell_pbc(s->cell, rp + 3 * a, r + 3 * a);
        return;
    }
    for (int a = 0; a < ns; a++)
        for (int i = 0; i < 3; i++) {
            double c = rp[3 * a + i];
            double idx = floor(c / s->full_ell[i]);
            r[3 * a + i] = c - idx * s->full_ell[i];
        }
}

static double erfc_as(double x)
{
    return mo_erfc(x);
}

/* ---------------------------------------------------------------- Ewald */
void mo_gcmc_kvecs(const mo_gcmc_sys *s, int *nk, double *kv, double *w)
{
    int K = s->kmax, n = 0;
    double V = s->full_ell[0] * s->full_ell[1] * s->full_ell[2];
    for (int mx = 0; mx <= K; mx++)
        for (int my = -K; my <= K; my++)
            for (int mz = -K; mz <= K; mz++) {
                if (mx == 0 && (my < 0 || (my == 0 && mz <= 0))) continue;
                double kx = ((double)mx / s->full_ell[0]) * TWOPI_R;
                double ky = ((double)my / s->full_ell[1]) * TWOPI_R;
                double kz = ((double)mz / s->full_ell[2]) * TWOPI_R;
                double k2 = kx * kx + ky * ky + kz * kz;
                if (kv) {
                    kv[3 * n] = kx;
                    kv[3 * n + 1] = ky;
                    kv[3 * n + 2] = kz;
                }
                if (w) w[n] = (4.0 * M_PI / V) * exp(-k2 / (4.0 * s->kappa * s->kappa)) / k2 * MO_TOKCAL;
                n++;
            }
    *nk = n;
}

/* structure factor contribution of one molecule of species s (k vectors of
 * the global state st): sre/sim[k] = sum_a q_a e^{ik.r_a} */
static void mol_sk(const mo_gcmc_sys *s, const mo_gcmc_state *st, const double *r, double *sre, double *sim)
{
    for (int k = 0; k < st->nk; k++) {
        double re = 0.0, im = 0.0;
        const double *kv = st->kv + 3 * k;
        for (int a = 0; a < s->ns; a++) {
            if (s->sq[a] == 0.0) continue;
            double arg = kv[0] * r[3 * a] + kv[1] * r[3 * a + 1] + kv[2] * r[3 * a + 2];
            re += s->sq[a] * cos(arg);
            im += s->sq[a] * sin(arg);
        }
        sre[k] = re;
        sim[k] = im;
    }
}

/* E_recip(S + sgn*add) - E_recip(S) (+ optional second change) */
static double drecip(const mo_gcmc_state *st, const double *are, const double *aim, double asgn,
                     const double *bre, const double *bim, double bsgn)
{
    double d = 0.0;
    for (int k = 0; k < st->nk; k++) {
        double re = st->Sre[k], im = st->Sim[k];
        double nre = re, nim = im;
        if (are) {
            nre += asgn * are[k];
            nim += asgn * aim[k];
        }
        if (bre) {
            nre += bsgn * bre[k];
            nim += bsgn * bim[k];
        }
        d += st->wk[k] * ((nre * nre + nim * nim) - (re * re + im * im));
    }
    return d;
}

double mo_gcmc_mol_const(const mo_gcmc_sys *s)
{
    /* self + intramolecular exclusion of one rigid molecule */
    if (!s->ewald) return 0.0;
    double e = 0.0;
    for (int a = 0; a < s->ns; a++) e -= s->kappa / sqrt(M_PI) * s->sq[a] * s->sq[a];
    for (int a = 0; a < s->ns; a++)
        for (int b = a + 1; b < s->ns; b++) {
            double d[3];
            for (int i = 0; i < 3; i++) d[i] = s->ref[a][i] - s->ref[b][i];
            double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
            e -= s->sq[a] * s->sq[b] * (1.0 - erfc_as(s->kappa * r)) / r;
        }
    return e * MO_TOKCAL;
}

/* ---------------------------------------------------- molecule energy */
/* config_getMolecCTR (config.F90:1788-1802): the sum of the ns unfolded site
 * positions in site order, divided by ns */
static void mol_centre(int ns, const double *rp, double c[3])
{
    for (int i = 0; i < 3; i++) {
        double t = 0.0;
        for (int a = 0; a < ns; a++) t = t + rp[3 * a + i];
        c[i] = t / (double)ns;
    }
}

/* Energy of a molecule of species sp (sys[sp]) with wrapped site coordinates
 * r (unfolded rp) against the framework maps of its species and every
 * molecule of every species except molecule 'skip' of species sp.  The pair
 * tables, cell and Ewald / Coulomb settings are those of sys[0] (shared by
 * the species; site types index the common table); the centre-of-mass
 * cutoffs are sys[sp]'s towards each species (ssbasic_mmint,
 * ssbasic.F90:745-756: the molecule pair's LJ or Coulomb term is skipped
 * when the centres are further apart).  Returns 0 on failure. */
static int mol_energy_mix(const mo_gcmc_sys *const *sys, int nspc, mo_gcmc_state *const *sts, int sp,
                          const double *r, const double *rp, int skip, double *umap, double *upair)
{
    const mo_gcmc_sys *s = sys[sp];
    const mo_gcmc_sys *g = sys[0];
    double um = 0.0;
    const int sl = slanted(g);
    /* pmap_int / interpolate_int (interpolate.F90:663-679): slanted cells
     * interpolate without the periodic wrap of the upper corner */
    int (*herm)(const int *, const double *, const int32_t *, const float *, long long, const double *,
                double *) = sl ? mo_pothermite_nowrap : mo_pothermite;
    for (int a = 0; a < s->ns; a++) {
        double c[3];
        if (sl) mo_cell_maptouc(g->cell, r + 3 * a, c);
        else mo_maptouc(g->fund_ell, r + 3 * a, c);
        int pm = s->site_pmap[a];
        if (pm >= 0) {
            double p;
            if (!herm(s->map_npnts[pm], s->map_step[pm], s->map_ptr[pm], s->map_vals[pm], s->map_nf[pm], c, &p))
                return 0;
            um = um + p;
        }
        int em = s->site_emap[a];
        if (em >= 0) {
            double p;
            if (!herm(s->map_npnts[em], s->map_step[em], s->map_ptr[em], s->map_vals[em], s->map_nf[em], c, &p))
                return 0;
            um = um + p * s->sq[a];
        }
    }
    double up = 0.0;
    double ctr[3];
    mol_centre(s->ns, rp, ctr);
    for (int s2 = 0; s2 < nspc; s2++) {
        const mo_gcmc_sys *o = sys[s2];
        const mo_gcmc_state *st = sts[s2];
        const double cl = s->lj_comcut2[s2], cc = s->coul_comcut2[s2];
        for (int m = 0; m < st->N; m++) {
            if (s2 == sp && m == skip) continue;
            const double *ro = st->r + 3 * o->ns * m;
            int lj_pair = 1, coul_pair = 1;
            if (cl > 0.0 || cc > 0.0) {
                double c2[3], sep[3];
                mol_centre(o->ns, st->rp + 3 * o->ns * m, c2);
                if (sl) mo_cell_minimage(g->cell, ctr, c2, sep);
                else mo_minimage(g->full_ell, ctr, c2, sep);
                double d2 = 0.0;
                d2 = d2 + sep[0] * sep[0];
                d2 = d2 + sep[1] * sep[1];
                d2 = d2 + sep[2] * sep[2];
                if (cl > 0.0 && d2 > cl) lj_pair = 0;
                if (cc > 0.0 && d2 > cc) coul_pair = 0;
            }
            for (int a = 0; a < s->ns; a++)
                for (int b = 0; b < o->ns; b++) {
                    double sep[3];
                    if (sl) mo_cell_minimage(g->cell, r + 3 * a, ro + 3 * b, sep);
                    else mo_minimage(g->full_ell, r + 3 * a, ro + 3 * b, sep);
                    double r2 = 0.0;
                    r2 = r2 + sep[0] * sep[0];
                    r2 = r2 + sep[1] * sep[1];
                    r2 = r2 + sep[2] * sep[2];
                    int ta = s->stype[a], tb = o->stype[b];
                    int t = ta * g->nst + tb;
                    if (lj_pair && g->lj_on[t]) {
                        if (r2 <= g->lj_h2[t]) {
                            if (r2 < g->lj_l2[t]) return 0;
                            double r2i = 1.0 / r2;
                            double r6i = r2i * r2i * r2i;
                            double r12i = r6i * r6i;
                            up = up + (g->lj_A[t] * r12i - g->lj_B[t] * r6i);
                        }
                    }
                    if (g->ewald && r2 < g->rc2_real && s->sq[a] != 0.0 && o->sq[b] != 0.0) {
                        double rr = sqrt(r2);
                        up = up + s->sq[a] * o->sq[b] * erfc_as(g->kappa * rr) / rr * MO_TOKCAL;
                    }
                    if (coul_pair && g->coul && g->coul_on[t]) {
                        double e;
                        if (!mo_coul_pair(g->coul_hicut, g->coul_locut, g->coul_shield, r2, s->sq[a], o->sq[b], &e))
                            return 0;
                        up = up + e;
                    }
                }
        }
    }
    *umap = um;
    *upair = up;
    return 1;
}

/* trace of step 'step' in the global state g: move = 16 species + move
 * type, N = molecules of the moved species after the step */
static void record(mo_gcmc_state *g, int step, int spc, int mtype, int acc, int N, double de)
{
    if (!g->trace || step >= g->trace_cap) return;
    mo_gcmc_trace *t = g->trace + step;
    t->move = 16 * spc + mtype;
    t->accepted = acc;
    t->N = N;
    t->dE = de;
}

/* bmap_cubeinit (bmap.F90:181-403): weights exp(-u / kT_bias), exponent
 * clamped to +-30, over the cubelets whose lower-corner node (ptr(k, j, i))
 * is tabulated with u <= 30 kcal/mol (BMAP_DEFAULT_CUTOFF), loop x outer, y,
 * z inner; when min(wt) / sum(wt) <= 1e-6 the energies above
 * -log(1e-6 sum) / rti are lowered to it first.  (When the ratio is larger
 * the reference compares against an unset local, maxnrgval; here nothing is
 * clamped.)  Cumulative probabilities are running sums over cum_wt, both
 * summed in that order.  ptr is [nx][ny][nz] (z fastest, 1-based, 0 =
 * untabulated); V is the map's V record [nfsize]. */
int mo_bmap_build(mo_bmap *b, const int npnts[3], const double step[3], const double eff[3], const int32_t *ptr,
                  const float *V, double T_bias)
{
    memset(b, 0, sizeof *b);
    const int nx1 = npnts[0] - 1, ny1 = npnts[1] - 1, nz1 = npnts[2] - 1;
    b->ncub[0] = nx1;
    b->ncub[1] = ny1;
    b->ncub[2] = nz1;
    b->ncubelets = (long long)nx1 * ny1 * nz1;
    for (int i = 0; i < 3; i++) {
        b->step[i] = step[i];
        b->eff[i] = eff[i];
    }
    const double rti = 1.0 / (T_bias * MO_KCALMOLE_KB);
    const double ucut = 30.0;
    double sum = 0.0, mn = 1.0e20;
    long long numtab = 0;
    for (int i = 0; i < nx1; i++)
        for (int j = 0; j < ny1; j++)
            for (int k = 0; k < nz1; k++) {
                int p = ptr[((long long)i * npnts[1] + j) * npnts[2] + k];
                if (p == 0) continue;
                double u = (double)V[p - 1];
                if (!(u <= ucut)) continue;
                double e = -rti * u;
                if (e > MO_MAX_EXP) e = MO_MAX_EXP;
                else if (e < -MO_MAX_EXP) e = -MO_MAX_EXP;
                double wt = exp(e);
                if (wt < mn) mn = wt;
                sum = sum + wt;
                numtab++;
            }
    if (numtab == 0) return -1;
    b->clamped = !(mn / sum > 1.0e-6);
    b->maxnrgval = b->clamped ? -log(sum * 1.0e-6) / rti : INFINITY;
    b->cube = malloc(sizeof(int32_t) * numtab);
    b->wt = malloc(sizeof(double) * numtab);
    b->cum_prob = malloc(sizeof(double) * numtab);
    b->cube_index = calloc((size_t)b->ncubelets, sizeof(int32_t));
    long long t = 0;
    b->cum_wt = 0.0;
    for (int i = 0; i < nx1; i++)
        for (int j = 0; j < ny1; j++)
            for (int k = 0; k < nz1; k++) {
                int p = ptr[((long long)i * npnts[1] + j) * npnts[2] + k];
                if (p == 0) continue;
                double u = (double)V[p - 1];
                if (!(u <= ucut)) continue;
                if (u > b->maxnrgval) u = b->maxnrgval;
                double e = -rti * u;
                if (e > MO_MAX_EXP) e = MO_MAX_EXP;
                else if (e < -MO_MAX_EXP) e = -MO_MAX_EXP;
                long long c = ((long long)i * ny1 + j) * nz1 + k;
                b->cube[t] = (int32_t)c;
                b->wt[t] = exp(e);
                b->cube_index[c] = (int32_t)(t + 1);
                b->cum_wt = b->cum_wt + b->wt[t];
                t++;
            }
    b->numtab = numtab;
    double run = 0.0;
    for (t = 0; t < numtab; t++) {
        run = run + b->wt[t];
        b->cum_prob[t] = run / b->cum_wt;
    }
    return 0;
}

void mo_bmap_free(mo_bmap *b)
{
    free(b->cube);
    free(b->wt);
    free(b->cum_prob);
    free(b->cube_index);
    memset(b, 0, sizeof *b);
}

/* bmap_getbiasindex (bmap.F90:575-636), the reference's bisection on the
 * cumulative probabilities; 0-based result */
long long mo_bmap_index(const mo_bmap *b, double key)
{
    long long low = 1, high = b->numtab, mid = 1;
    for (;;) {
        if (low > high) break;
        mid = (low + high) / 2;
        double cp = b->cum_prob[mid - 1];
        if (key < cp) {
            high = mid - 1;
        } else if (key > cp) {
            low = mid + 1;
            mid = low;
        } else {
            break;
        }
    }
    return mid - 1;
}

/* mcmoves_attempt (mcmoves.F90:290-335) */
static int attempt(const mo_gcmc_sys *s, mo_gcmc_state *st, double deltaE, double bias, int has_x,
                   double addxtra, int has_p, double prefactor)
{
    double rti = 1.0 / (MO_KCALMOLE_KB * s->T);
    double u = -rti * deltaE;
    u = u + log(bias);
    if (has_x) u = u + addxtra;
    if (u > MO_MAX_EXP) u = MO_MAX_EXP;
    if (u < -MO_MAX_EXP) u = -MO_MAX_EXP;
    double boltz = exp(u);
    if (has_p) boltz = boltz * prefactor;
    return boltz > mo_rranf(&st->rng);
}

/* rigidmoves_adjustdeltatrans / _adjustdeltarot (rigidmoves.F90:2586-2617):
 * 0.95 below a cumulative acceptance ratio of 0.49, 1.05 above 0.51, the
 * displacement bounded by MIN/MAX_TRANS_DELTA = 0.01, 2.0 (:216-217).
 * 0.49 and 0.51 are default-real literals and the ratio succ/(att*one) is
 * REAL(8) (moves.F90:553-560): the comparison widens the float32 values. */
static double adjust_delta(double d, double ratio)
{
    if (ratio < (double)0.49f) return fmax(d * 0.95, 0.01);
    if (ratio > (double)0.51f) return fmin(d * 1.05, 2.0);
    return d;
}

/* rigidmoves_comFromBmap (rigidmoves.F90:1938-1969): a cubelet by its
 * weight, a point in it (bmap_getbiaspt), a unit cell of the simulation
 * cell; *bias = 1 / (weight ncubelets) */
static void bmap_com(const mo_gcmc_sys *s, const mo_gcmc_sys *g, mo_rng *rng, double com[3], double *bias)
{
    const mo_bmap *b = s->bias;
    long long t = mo_bmap_index(b, mo_rranf(rng));
    long long c = b->cube[t];
    long long ci[3] = {c / ((long long)b->ncub[1] * b->ncub[2]), (c / b->ncub[2]) % b->ncub[1], c % b->ncub[2]};
    double wt = b->wt[t] / b->cum_wt;
    *bias = 1.0 / wt / (double)b->ncubelets;
    for (int i = 0; i < 3; i++) com[i] = (double)ci[i] * b->step[i];
    for (int i = 0; i < 3; i++) com[i] = com[i] + mo_rranf(rng) * b->step[i];
    for (int i = 0; i < 3; i++)
        if (com[i] >= b->eff[i]) com[i] = (double)0.9999990f * b->eff[i];
    for (int i = 0; i < 3; i++) {
        int nc = (int)lround(g->full_ell[i] / g->fund_ell[i]);
        int icell = (int)(mo_rranf(rng) * nc);
        com[i] = com[i] + g->fund_ell[i] * icell;
    }
}

/* rigidmoves_reorient with newxyz (random Euler angles; none for one site) */
static void random_angles(int ns, mo_rng *rng, double ang[3])
{
    ang[0] = ang[1] = ang[2] = 0.0;
    if (ns != 1) {
        double costheta = 1.0 - mo_rranf(rng) * 2.0;
        ang[1] = acos(costheta);
        ang[0] = MO_PI - mo_rranf(rng) * MO_TWOPI;
        ang[2] = MO_PI - mo_rranf(rng) * MO_TWOPI;
    }
}

/* cavitylist_getIndexVec (cavitylist.F90:536-546): Int(r / d) per axis of a
 * folded position, cubes numbered with z fastest.  A coordinate that rounds
 * onto the upper face is kept in the last cube. */
static long long cav_cube(const mo_cavity *cv, const double *r)
{
    long long c = 0;
    for (int i = 0; i < 3; i++) {
        int k = (int)(r[i] / cv->d[i]);
        if (k > cv->n[i] - 1) k = cv->n[i] - 1;
        if (k < 0) k = 0;
        c = c * cv->n[i] + k;
    }
    return c;
}

/* cavitylist_updateMolecInfo (cavitylist.F90:351-513) for the sites r of one
 * molecule of species sp entering (sign 1) or leaving (-1) the cavity list.
 * A cube holds a count of sites; it is free when the count is zero, which
 * is the reference's test of its per-species molecule lists. */
static void cav_move(const mo_gcmc_sys *const *sys, mo_gcmc_state *const *sts, int sp, const double *r, int sign)
{
    const mo_cavity *cv = sys[0]->cav;
    if (!cv || !(cv->spc_mask >> sp & 1)) return;
    for (int a = 0; a < sys[sp]->ns; a++) sts[0]->cube_cnt[cav_cube(cv, r + 3 * a)] += sign;
}

/* the cavity list of the current configuration (cavitylist_initSorbInfo,
 * cavitylist.F90:282-340) */
void mo_gcmc_cav_rebuild(const mo_gcmc_sys *const *sys, int nspc, mo_gcmc_state *const *sts)
{
    const mo_cavity *cv = sys[0]->cav;
    if (!cv) return;
    long long nc = (long long)cv->n[0] * cv->n[1] * cv->n[2];
    for (long long c = 0; c < nc; c++) sts[0]->cube_cnt[c] = 0;
    for (int sp = 0; sp < nspc; sp++)
        for (int m = 0; m < sts[sp]->N; m++) cav_move(sys, sts, sp, sts[sp]->r + 3 * sys[sp]->ns * m, 1);
}

/* stats_update (stats.F90:152-195) of cav_prob(t) with the cavities found */
static void cav_stats(mo_gcmc_state *st, int t, double v)
{
    st->cav_n[t]++;
    st->cav_avg[t] = (st->cav_avg[t] * ((double)st->cav_n[t] - 1.0) + v) / ((double)st->cav_n[t] * 1.0);
}

int mo_gcmc_run_mix(const mo_gcmc_sys *const *sys, int nspc, mo_gcmc_state *const *sts, long long nsteps)
{
    const mo_gcmc_sys *g = sys[0];
    mo_gcmc_state *G = sts[0];
    double rnew[3 * MO_MAXS], rpnew[3 * MO_MAXS];
    double *ak = NULL, *bk = NULL, *ck = NULL, *dk = NULL;
    if (g->ewald) {
        ak = malloc(sizeof(double) * G->nk);
        bk = malloc(sizeof(double) * G->nk);
        ck = malloc(sizeof(double) * G->nk);
        dk = malloc(sizeof(double) * G->nk);
    }
    double mconst_s[MO_MAXSPC];
    for (int i = 0; i < nspc; i++) mconst_s[i] = mo_gcmc_mol_const(sys[i]);
    /* cavity-biased trials: centre, angles and bias per successful trial */
    int tmax = 0;
    for (int i = 0; i < nspc; i++)
        if (sys[i]->cav_ntrials > tmax) tmax = sys[i]->cav_ntrials;
    double *ctr_trial = tmax ? malloc(sizeof(double) * 7 * tmax) : NULL;
    for (long long it = 0; it < nsteps; it++) {
        int step = (int)G->nsteps;
        /* gcmc_dosim: a species with equal probability (gcmc.F90:240-242) */
        const int sp = (int)(mo_rranf(&G->rng) * (double)nspc);
        const mo_gcmc_sys *s = sys[sp];
        mo_gcmc_state *st = sts[sp];
        const int ns = s->ns;
        const double mconst = mconst_s[sp];
        double prob = mo_rranf(&G->rng);
        int mv = 0;
        for (int i = 0; i < s->nmoves; i++) {
            if (prob > s->cumwt[i]) mv++;
            else break;
        }
        if (mv >= s->nmoves) mv = s->nmoves - 1;
        int mtype = s->move_type[mv];
        int N = st->N;
        if (mtype == MO_MV_INSERT || mtype == MO_MV_BINSERT) {
            st->att[0]++;
            int accepted = 0;
            double de = 0.0;
            if (N >= st->maxN) return -3; /* capacity of the molecule arrays */
            {
                double com[3], ang[3] = {0.0, 0.0, 0.0};
                double biasfactor = 1.0;
                int found = 1;
                if (mtype == MO_MV_BINSERT && s->cav_ntrials > 0) {
                    /* rigidmoves_getConfigFromCavBias (rigidmoves.F90:1987-2126):
                     * ntrials centres (uniform, or from the bias map), each
                     * with random angles; the trials whose sites all lie in
                     * cavities; cav_prob(totmol) updated with their number;
                     * one of them at random, bias x cav_prob / ntrials.
                     * totmol counts the molecule being inserted
                     * (config_getTotMoles after the slot is opened). */
                    const mo_cavity *cv = g->cav;
                    const int T = s->cav_ntrials;
                    int nsucc = 0;
                    for (int t = 0; t < T; t++) {
                        double c[3], a[3], bf = 1.0;
                        if (s->cav_map) bmap_com(s, g, &G->rng, c, &bf);
                        else
                            for (int i = 0; i < 3; i++) c[i] = g->full_ell[i] * mo_rranf(&G->rng);
                        random_angles(ns, &G->rng, a);
                        toxyz(s, c, a, rpnew);
                        pbc(g, ns, rpnew, rnew);
                        int cav = 1;
                        for (int q = 0; q < ns && cav; q++) {
                            long long cb = cav_cube(cv, rnew + 3 * q);
                            cav = (!cv->tab || cv->tab[cb]) && G->cube_cnt[cb] == 0;
                        }
                        if (cav) {
                            double *o = ctr_trial + 7 * nsucc++;
                            memcpy(o, c, sizeof c);
                            memcpy(o + 3, a, sizeof a);
                            o[6] = bf;
                        }
                    }
                    int totmol = 1;
                    for (int i = 0; i < nspc; i++) totmol += sts[i]->N;
                    if (totmol >= st->cav_cap) return -5; /* capacity of the cav_prob table */
                    cav_stats(st, totmol, (double)nsucc);
                    if (nsucc == 0) {
                        found = 0;
                    } else {
                        int pick = (int)(mo_rranf(&G->rng) * nsucc);
                        const double *o = ctr_trial + 7 * pick;
                        memcpy(com, o, sizeof com);
                        memcpy(ang, o + 3, sizeof ang);
                        biasfactor = o[6] * (st->cav_avg[totmol] / (double)T);
                    }
                } else if (mtype == MO_MV_INSERT) {
                    /* simcell_uniformRandPt (simcell.F90:274-295) */
                    if (slanted(g)) mo_cell_uniform(g->cell, &G->rng, com);
                    else
                        for (int i = 0; i < 3; i++) com[i] = g->full_ell[i] * mo_rranf(&G->rng);
                    random_angles(ns, &G->rng, ang);
                } else {
                    bmap_com(s, g, &G->rng, com, &biasfactor);
                    random_angles(ns, &G->rng, ang);
                }
                toxyz(s, com, ang, rpnew);
                pbc(g, ns, rpnew, rnew);
                double um = 0.0, up = 0.0;
                int ok = found && mol_energy_mix(sys, nspc, sts, sp, rnew, rpnew, -1, &um, &up);
                double U = 0.0;
                if (ok) {
                    U = um + up;
                    if (g->ewald) {
                        mol_sk(s, G, rnew, ak, bk);
                        U = U + drecip(G, ak, bk, 1.0, NULL, NULL, 0.0) + mconst;
                    }
                    if (U > s->max_nrg) ok = 0;
                }
                if (ok) {
                    double bias = biasfactor / s->insdelratio;
                    accepted = attempt(s, G, U, bias, 1, st->B, 1, 1.0 / (double)(float)(N + 1));
                    if (accepted) {
                        memcpy(st->com + 3 * N, com, sizeof com);
                        memcpy(st->ang + 3 * N, ang, sizeof ang);
                        memcpy(st->rp + 3 * ns * N, rpnew, sizeof(double) * 3 * ns);
                        memcpy(st->r + 3 * ns * N, rnew, sizeof(double) * 3 * ns);
                        if (g->ewald)
                            for (int k = 0; k < G->nk; k++) {
                                G->Sre[k] += ak[k];
                                G->Sim[k] += bk[k];
                            }
                        st->N = N + 1;
                        cav_move(sys, sts, sp, rnew, 1);
                        G->energy += U;
                        st->acc[0]++;
                        de = U;
                    }
                }
            }
            record(G, step, sp, mtype, accepted, st->N, de);
        } else if (mtype == MO_MV_DELETE || mtype == MO_MV_BDELETE) {
            int accepted = 0;
            double de = 0.0;
            if (N != 0) {
                int mol = (int)(mo_rranf(&G->rng) * N);
                st->att[1]++;
                double biasfactor = 1.0;
                if (mtype == MO_MV_BDELETE && !(s->cav_ntrials > 0 && !s->cav_map)) {
                    /* rigidmoves_bdelete: bmap_getcellwt of the centre of mass
                     * mapped into the unit cell (simcell_maptouc); a cubelet
                     * outside the bias map stops the reference */
                    const mo_bmap *b = s->bias;
                    double uc[3];
                    mo_maptouc(g->fund_ell, st->com + 3 * mol, uc);
                    long long ci[3];
                    for (int i = 0; i < 3; i++) {
                        ci[i] = (long long)(uc[i] / b->step[i]);
                        if (ci[i] > b->ncub[i] - 1) ci[i] = b->ncub[i] - 1;
                    }
                    long long t = b->cube_index[(ci[0] * b->ncub[1] + ci[1]) * b->ncub[2] + ci[2]];
                    if (t == 0) return -4;
                    double wt = b->wt[t - 1] / b->cum_wt;
                    biasfactor = (double)b->ncubelets * wt;
                }
                if (mtype == MO_MV_BDELETE && s->cav_ntrials > 0) {
                    /* rigidmoves_bdelete (rigidmoves.F90:1632-1643): divided by
                     * cav_prob(totmol) / ntrials at the current loading */
                    int totmol = 0;
                    for (int i = 0; i < nspc; i++) totmol += sts[i]->N;
                    if (totmol >= st->cav_cap) return -5;
                    biasfactor = biasfactor / (st->cav_avg[totmol] / (double)s->cav_ntrials);
                }
                const double *r = st->r + 3 * ns * mol;
                double um, up;
                int ok = mol_energy_mix(sys, nspc, sts, sp, r, st->rp + 3 * ns * mol, mol, &um, &up);
                double U = um + up;
                if (ok && g->ewald) {
                    mol_sk(s, G, r, ak, bk);
                    U = U - drecip(G, ak, bk, -1.0, NULL, NULL, 0.0) + mconst;
                }
                if (!ok) return -1; /* mcmoves.F90:599-603: an existing molecule must evaluate */
                double bias = biasfactor * s->insdelratio;
                accepted = attempt(s, G, -U, bias, 1, -st->B, 1, 1.0 * N);
                if (accepted) {
                    int last = N - 1;
                    cav_move(sys, sts, sp, st->r + 3 * ns * mol, -1);
                    memmove(st->com + 3 * mol, st->com + 3 * last, sizeof(double) * 3);
                    memmove(st->ang + 3 * mol, st->ang + 3 * last, sizeof(double) * 3);
                    memmove(st->rp + 3 * ns * mol, st->rp + 3 * ns * last, sizeof(double) * 3 * ns);
                    memmove(st->r + 3 * ns * mol, st->r + 3 * ns * last, sizeof(double) * 3 * ns);
                    if (g->ewald)
                        for (int k = 0; k < G->nk; k++) {
                            G->Sre[k] -= ak[k];
                            G->Sim[k] -= bk[k];
                        }
                    st->N = N - 1;
                    G->energy -= U;
                    st->acc[1]++;
                    de = -U;
                }
            }
            record(G, step, sp, mtype, accepted, st->N, de);
        } else {
            /* TRANSLATE / ROTATE */
            int accepted = 0;
            double de = 0.0;
            int idx = (mtype == MO_MV_TRANSLATE) ? 2 : 3;
            if (N != 0) {
                int mol = (int)(mo_rranf(&G->rng) * N);
                st->att[idx]++;
                const double *rold = st->r + 3 * ns * mol;
                double um0, up0;
                if (!mol_energy_mix(sys, nspc, sts, sp, rold, st->rp + 3 * ns * mol, mol, &um0, &up0)) return -2;
                double com[3], ang[3];
                memcpy(com, st->com + 3 * mol, sizeof com);
                memcpy(ang, st->ang + 3 * mol, sizeof ang);
                if (mtype == MO_MV_TRANSLATE) {
                    double dv[3];
                    const double dtr = s->adjust_trans ? st->delta[0] : s->dtrans;
                    for (int i = 0; i < 3; i++) dv[i] = ((mo_rranf(&G->rng) * 2.0) - 1.0) * dtr;
                    for (int i = 0; i < 3; i++) com[i] = com[i] + dv[i];
                } else {
                    const double drt = s->adjust_rot ? st->delta[1] : s->drot;
                    double dphi = ((mo_rranf(&G->rng) * 2.0) - 1.0) * drt;
                    double dpsi = ((mo_rranf(&G->rng) * 2.0) - 1.0) * drt;
                    double dcos = ((mo_rranf(&G->rng) * 2.0) - 1.0) * drt;
                    if (ns != 1) {
                        double phi = ang[0] + dphi;
                        double psi = ang[2] + dpsi;
                        double costheta = cos(ang[1]) + dcos;
                        ang[0] = phi - round(phi / MO_TWOPI) * MO_TWOPI;
                        ang[2] = psi - round(psi / MO_TWOPI) * MO_TWOPI;
                        costheta = costheta - round(costheta / 2.0) * 2.0;
                        ang[1] = acos(costheta);
                    }
                }
                toxyz(s, com, ang, rpnew);
                pbc(g, ns, rpnew, rnew);
                double um1, up1;
                int ok = mol_energy_mix(sys, nspc, sts, sp, rnew, rpnew, mol, &um1, &up1);
                double dU = 0.0;
                if (ok) {
                    double U1 = um1 + up1, U0 = um0 + up0;
                    if (U1 > s->max_nrg) ok = 0;
                    dU = U1 - U0;
                    if (ok && g->ewald) {
                        mol_sk(s, G, rold, ck, dk);
                        mol_sk(s, G, rnew, ak, bk);
                        dU = dU + drecip(G, ak, bk, 1.0, ck, dk, -1.0);
                    }
                }
                if (ok) {
                    accepted = attempt(s, G, dU, 1.0, 0, 0.0, 0, 1.0);
                    if (accepted) {
                        cav_move(sys, sts, sp, st->r + 3 * ns * mol, -1);
                        cav_move(sys, sts, sp, rnew, 1);
                        memcpy(st->com + 3 * mol, com, sizeof com);
                        memcpy(st->ang + 3 * mol, ang, sizeof ang);
                        memcpy(st->rp + 3 * ns * mol, rpnew, sizeof(double) * 3 * ns);
                        memcpy(st->r + 3 * ns * mol, rnew, sizeof(double) * 3 * ns);
                        if (g->ewald)
                            for (int k = 0; k < G->nk; k++) {
                                G->Sre[k] += ak[k] - ck[k];
                                G->Sim[k] += bk[k] - dk[k];
                            }
                        G->energy += dU;
                        st->acc[idx]++;
                        de = dU;
                    }
                }
                /* moves_adjustParams after every attempt (mcmoves.F90:460-461) */
                if (mtype == MO_MV_TRANSLATE && s->adjust_trans)
                    st->delta[0] = adjust_delta(st->delta[0], (double)st->acc[2] / (double)st->att[2]);
                else if (mtype != MO_MV_TRANSLATE && s->adjust_rot)
                    st->delta[1] = adjust_delta(st->delta[1], (double)st->acc[3] / (double)st->att[3]);
            }
            record(G, step, sp, mtype, accepted, st->N, de);
        }
        /* gcmcmoves_updatenmolecs: every species after every step */
        for (int i = 0; i < nspc; i++) sts[i]->sumN += sts[i]->N;
        G->nsteps++;
    }
    free(ak);
    free(bk);
    free(ck);
    free(dk);
    free(ctr_trial);
    return 0;
}

/* full-system energy (checks the running total, interact_checknrgs): map
 * and pair terms of every molecule, each pair once (molecules of earlier
 * species and earlier molecules of the same species), reciprocal sum and
 * the per-molecule Ewald constants */
double mo_gcmc_total_energy_mix(const mo_gcmc_sys *const *sys, int nspc, mo_gcmc_state *const *sts)
{
    double e = 0.0;
    int keepN[MO_MAXSPC];
    mo_gcmc_state tmp[MO_MAXSPC];
    mo_gcmc_state *tp[MO_MAXSPC];
    for (int i = 0; i < nspc; i++) {
        keepN[i] = sts[i]->N;
        tmp[i] = *sts[i];
        tp[i] = &tmp[i];
    }
    for (int sp = 0; sp < nspc; sp++) {
        for (int i = 0; i < nspc; i++) tmp[i].N = i < sp ? keepN[i] : 0;
        for (int m = 0; m < keepN[sp]; m++) {
            tmp[sp].N = m; /* molecules 0..m-1 of this species */
            double um, up;
            if (!mol_energy_mix(sys, nspc, tp, sp, sts[sp]->r + 3 * sys[sp]->ns * m, sts[sp]->rp + 3 * sys[sp]->ns * m,
                                -1, &um, &up))
                return NAN; /* a failed evaluation */
            e += um + up;
        }
    }
    if (sys[0]->ewald) {
        const mo_gcmc_state *G = sts[0];
        double er = 0.0;
        for (int k = 0; k < G->nk; k++) er += G->wk[k] * (G->Sre[k] * G->Sre[k] + G->Sim[k] * G->Sim[k]);
        e += er;
        for (int i = 0; i < nspc; i++) e += sts[i]->N * mo_gcmc_mol_const(sys[i]);
    }
    return e;
}

void mo_gcmc_recompute_sk_mix(const mo_gcmc_sys *const *sys, int nspc, mo_gcmc_state *const *sts)
{
    mo_gcmc_state *G = sts[0];
    for (int k = 0; k < G->nk; k++) G->Sre[k] = G->Sim[k] = 0.0;
    if (!sys[0]->ewald) return;
    double *a = malloc(sizeof(double) * G->nk), *b = malloc(sizeof(double) * G->nk);
    for (int sp = 0; sp < nspc; sp++)
        for (int m = 0; m < sts[sp]->N; m++) {
            mol_sk(sys[sp], G, sts[sp]->r + 3 * sys[sp]->ns * m, a, b);
            for (int k = 0; k < G->nk; k++) {
                G->Sre[k] += a[k];
                G->Sim[k] += b[k];
            }
        }
    free(a);
    free(b);
}

/* place molecule m of species s (sys[s], sts[s]) at com/ang (seeded
 * loadings); the cell is sys[0]'s */
void mo_gcmc_set_mol_mix(const mo_gcmc_sys *const *sys, mo_gcmc_state *const *sts, int sp, int m,
                         const double com[3], const double ang[3])
{
    const mo_gcmc_sys *s = sys[sp];
    mo_gcmc_state *st = sts[sp];
    memcpy(st->com + 3 * m, com, sizeof(double) * 3);
    memcpy(st->ang + 3 * m, ang, sizeof(double) * 3);
    toxyz(s, com, ang, st->rp + 3 * s->ns * m);
    pbc(sys[0], s->ns, st->rp + 3 * s->ns * m, st->r + 3 * s->ns * m);
}

/* one species: the mixture routines with nspc = 1 */
int mo_gcmc_run(const mo_gcmc_sys *s, mo_gcmc_state *st, long long nsteps)
{
    return mo_gcmc_run_mix(&s, 1, &st, nsteps);
}

double mo_gcmc_total_energy(const mo_gcmc_sys *s, const mo_gcmc_state *st)
{
    mo_gcmc_state *p = (mo_gcmc_state *)st;
    return mo_gcmc_total_energy_mix(&s, 1, &p);
}

void mo_gcmc_recompute_sk(const mo_gcmc_sys *s, mo_gcmc_state *st)
{
    mo_gcmc_recompute_sk_mix(&s, 1, &st);
}

void mo_gcmc_set_mol(const mo_gcmc_sys *s, mo_gcmc_state *st, int m, const double com[3], const double ang[3])
{
    mo_gcmc_set_mol_mix(&s, &st, 0, m, com, ang);
}

/* gcmc_genfuglist (gcmc.F90:303-319): B = ln(f 1e3 V 1e-30 N_A / (R T)) - ln(Zig) */
double mo_adams_B(double fug_kPa, double volume_A3, double T, double Zig)
{
    double sivolume = volume_A3 * 1.0e-30 * MO_NAV;
    double B = log(fug_kPa * 1.0e3 * sivolume / (MO_RGAS * T));
    return B - log(Zig);
}
