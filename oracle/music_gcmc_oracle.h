/*
 * music_gcmc_oracle.h -- GCMC move-loop oracle (test infrastructure only).
 */
#ifndef MUSIC_GCMC_ORACLE_H
#define MUSIC_GCMC_ORACLE_H
#include "music_oracle.h"
#ifdef __cplusplus
extern "C" {
#endif

enum { MO_MV_INSERT = 0, MO_MV_DELETE = 1, MO_MV_TRANSLATE = 2, MO_MV_ROTATE = 3, MO_MV_BINSERT = 4,
       MO_MV_BDELETE = 5 };

/* Bias map of biased insertion / deletion (bmap_init + bmap_cubeinit,
 * bmap.F90:69-403): the cubelets of a pmap whose lower-corner node is
 * tabulated with V <= 30 kcal/mol, in the reference's loop order (x outer, z
 * inner), Boltzmann weights at the bias temperature and their cumulative
 * probabilities. */
typedef struct {
    int ncub[3];           /* cubelets per axis (npnts - 1) */
    long long ncubelets;   /* all cubelets, header%ncubelets */
    long long numtab;      /* biased cubelets */
    double step[3], eff[3];
    double cum_wt;
    int clamped;           /* energies above maxnrgval were clamped */
    double maxnrgval;
    int32_t *cube;         /* [numtab] cubelet id (i * ny1 + j) * nz1 + k, 0-based */
    double *wt;            /* [numtab] weight */
    double *cum_prob;      /* [numtab] */
    int32_t *cube_index;   /* [ncubelets] 1-based position in the list, 0 = not biased */
} mo_bmap;
int mo_bmap_build(mo_bmap *b, const int npnts[3], const double step[3], const double eff[3], const int32_t *ptr,
                  const float *V, double T_bias);
void mo_bmap_free(mo_bmap *b);
long long mo_bmap_index(const mo_bmap *b, double key);
#define MO_MAXS 8
#define MO_MAXMAPS 8
#define MO_MAXSPC 4

/* Cavity list of cavity-biased insertion (cavitylist.F90): a grid of
 * n[0] x n[1] x n[2] cubes over the simulation cell (cavitylist_init,
 * :155-275); a cube is a cavity when it is tabulated (tab, NULL = all:
 * cavitylist_initArrUniform / _initArrPmap, :744-978) and no site of a
 * molecule of the species in spc_mask lies in it
 * (cavitylist_checkcavityfromIndex, :1137-1157). */
typedef struct {
    int n[3];
    double d[3];          /* cube edges: full_ell / n */
    const int8_t *tab;    /* [n0][n1][n2] (z fastest) 1 = tabulated, or NULL */
    int spc_mask;         /* bit s: molecules of species s occupy cubes */
} mo_cavity;

typedef struct {
    double full_ell[3], fund_ell[3];
    int ns;
    double ref[MO_MAXS][3];
    int stype[MO_MAXS];
    double sq[MO_MAXS];
    int site_pmap[MO_MAXS], site_emap[MO_MAXS];
    int nmaps;
    const int32_t *map_ptr[MO_MAXMAPS];
    const float *map_vals[MO_MAXMAPS]; /* SoA [8][nf] as in the map file */
    int map_npnts[MO_MAXMAPS][3];
    double map_step[MO_MAXMAPS][3];
    long long map_nf[MO_MAXMAPS];
    int nst;                                  /* sorbate site types */
    double lj_A[64], lj_B[64], lj_h2[64], lj_l2[64];
    int lj_on[64];                            /* [nst][nst] */
    int ewald, kmax;
    double kappa, rc2_real;
    double T;
    int nmoves;
    int move_type[4];
    double cumwt[4];
    double insdelratio, dtrans, drot, max_nrg;
    const mo_bmap *bias;                      /* BINSERT / BDELETE, or NULL */
    /* plain Coulomb (coulomb_snglint, coulomb.F90:258-327) for the type pairs
     * with coul_on: 331.5 shield q q (1/r - 1/hicut) within hicut (hicut 0:
     * no cutoff); r < locut fails */
    int coul;
    int coul_on[64];
    double coul_hicut, coul_locut, coul_shield;
    /* adjustable displacements (rigidmoves.F90:2586-2617): the state's delta
     * is used and updated after every attempt of the move */
    int adjust_trans, adjust_rot;
    /* slanted simulation cell (NULL or orthorhombic: full_ell / fund_ell):
     * insertion by simcell_uniformRandPt, pbc, minimum image and maptouc in
     * the slanted basis, Hermite without the corner wrap */
    const mo_cell *cell;
    /* centre-of-mass cutoffs of the sorbate pair terms (ssbasic COMCUT@ /
     * COMTRUNC@, ssbasic.F90:139-160, 745-756), squared, of this species'
     * molecules with those of species [s2]: a pair whose geometric centres
     * (config_getMolecCTR, config.F90:1788-1802: the mean of the unfolded
     * site positions) are further apart (simcell_minimage) has no LJ
     * (lj_comcut2) or plain-Coulomb (coul_comcut2) term; 0 = no cutoff */
    double lj_comcut2[MO_MAXSPC], coul_comcut2[MO_MAXSPC];
    /* cavity bias (rigidmoves_getConfigFromCavBias, rigidmoves.F90:1987-
     * 2126): the cavity list (sys[0]'s is the mixture's); this species'
     * BINSERT / BDELETE are cavity biased when cav_ntrials > 0, with centres
     * uniform in the cell (CAVITY_BIAS) or from its bias map when cav_map
     * (MAP_CAVITY_BIAS) */
    const mo_cavity *cav;
    int cav_ntrials, cav_map;
} mo_gcmc_sys;

typedef struct {
    int move, accepted, N;
    double dE;
} mo_gcmc_trace;

typedef struct {
    int N, maxN;
    double *com, *ang, *rp, *r;  /* [maxN][3], [maxN][3], [maxN][ns][3] x2 */
    int nk;
    double *kv, *wk, *Sre, *Sim; /* [nk][3], [nk], [nk], [nk] */
    double energy, B;
    mo_rng rng;
    long long att[4], acc[4];
    double sumN;
    long long nsteps;
    mo_gcmc_trace *trace;
    int trace_cap;
    double delta[2]; /* translate / rotate displacement of adjustable moves */
    /* cavity bias: sites per cube of the cavity list (sts[0]'s), and this
     * species' running average of the cavities found per insertion at total
     * loading t (cav_prob(t), stats_update, stats.F90:152-195; each entry
     * started with one update of ntrials / 2, rigidmoves.F90:2748-2760),
     * t < cav_cap */
    int *cube_cnt;
    double *cav_avg;
    long long *cav_n;
    int cav_cap;
} mo_gcmc_state;

/* Mixtures (gcmc_dosim, gcmc.F90:224-258): sys[i] / sts[i] for species i;
 * the cell, pair tables (site types index a common table), Coulomb / Ewald
 * settings and temperature are sys[0]'s; sts[0] also holds the RNG, the
 * energy, S(k), the step count and the trace (move = 16 species + type, N =
 * molecules of the moved species); sts[i] holds species i's molecules,
 * counters, displacements and B. */
int mo_gcmc_run_mix(const mo_gcmc_sys *const *sys, int nspc, mo_gcmc_state *const *sts, long long nsteps);
double mo_gcmc_total_energy_mix(const mo_gcmc_sys *const *sys, int nspc, mo_gcmc_state *const *sts);
void mo_gcmc_recompute_sk_mix(const mo_gcmc_sys *const *sys, int nspc, mo_gcmc_state *const *sts);
void mo_gcmc_set_mol_mix(const mo_gcmc_sys *const *sys, mo_gcmc_state *const *sts, int sp, int m,
                         const double com[3], const double ang[3]);
void mo_gcmc_cav_rebuild(const mo_gcmc_sys *const *sys, int nspc, mo_gcmc_state *const *sts);
void mo_gcmc_kvecs(const mo_gcmc_sys *s, int *nk, double *kv, double *w);
double mo_gcmc_mol_const(const mo_gcmc_sys *s);
int mo_gcmc_run(const mo_gcmc_sys *s, mo_gcmc_state *st, long long nsteps);
double mo_gcmc_total_energy(const mo_gcmc_sys *s, const mo_gcmc_state *st);
void mo_gcmc_recompute_sk(const mo_gcmc_sys *s, mo_gcmc_state *st);
void mo_gcmc_set_mol(const mo_gcmc_sys *s, mo_gcmc_state *st, int m, const double com[3], const double ang[3]);
double mo_adams_B(double fug_kPa, double volume_A3, double T, double Zig);
void mo_euler_xyz(int ns, const double *ref, const double com[3], const double ang[3], double *rp);
int mo_coul_pair(double hicut, double locut, double shield, double r2, double q1, double q2, double *e);

#ifdef __cplusplus
}
#endif
#endif
