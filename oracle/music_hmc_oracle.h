/*
 * music_hmc_oracle.h -- CPU FP64 restatement of Music's hybrid MC burst for
 * flexible sorbates (TEST INFRASTRUCTURE ONLY: the parity checker of the
 * HIP HMC kernel; only tests/ and bench.py's CPU baseline may load it).
 *
 * One HMC move of the whole sorbate subsystem (hybridmoves_integrate_boltz,
 * hybridmoves.F90:186-241): Gaussian velocities per atom
 * (hybridmoves_boltzVel, hybridmoves.F90:247-307; random_gaussian,
 * random.F90:197-217), forces and energy of the start configuration, n NVE
 * velocity-Verlet steps (velverlet_nveIntegrate, velverlet.F90:148-288,
 * integratePos/halfIntegrateVel :608-665), Metropolis on dU + dKE
 * (mcmoves_perturb, mcmoves.F90:371-426; mcmoves_attempt :290-335).
 * Energy model: framework maps with forces (forcehermite,
 * interpolate.F90:189-584), sorbate-sorbate LJ with minimum image
 * (lj_snglint, lj.F90:652-733), harmonic bonds (harmonic.F90:130-160),
 * harmonic angles (harmonica.F90:100-160), cosine-series torsions
 * (cosexpansion.F90:88-196).  Units: A, ps, amu, kcal/mol (energies),
 * kJ/mol (kinetic energy, as the reference).
 */
#ifndef MUSIC_HMC_ORACLE_H
#define MUSIC_HMC_ORACLE_H
#include "music_oracle.h"
#ifdef __cplusplus
extern "C" {
#endif

#define MO_HMC_MAXS 8
#define MO_HMC_MAXT 16

typedef struct {
    int ns;
    int stype[MO_HMC_MAXS];
    double mass[MO_HMC_MAXS];
    int site_map[MO_HMC_MAXS];          /* index into maps, -1 = none */
    int nmaps;
    const int32_t *map_ptr[MO_HMC_MAXS];
    const float *map_vals[MO_HMC_MAXS]; /* SoA [8][nf] */
    int map_npnts[MO_HMC_MAXS][3];
    double map_step[MO_HMC_MAXS][3];
    long long map_nf[MO_HMC_MAXS];
    double full_ell[3], fund_ell[3];
    int nst;
    double lj_A[64], lj_B[64], lj_h2[64], lj_l2[64];
    int lj_on[64];
    int nb;
    int bi[MO_HMC_MAXT][2];
    double bk[MO_HMC_MAXT], br0[MO_HMC_MAXT];
    int na;
    int ai[MO_HMC_MAXT][3];
    double ak[MO_HMC_MAXT], at0[MO_HMC_MAXT];
    int nt;
    int ti[MO_HMC_MAXT][4];
    double tc[MO_HMC_MAXT][6];
    double T, dt;
    int nmd;
    double max_nrg_md, max_nrg0;
} mo_hmc_sys;

typedef struct {
    int N, maxN;
    double *rp, *r, *v, *a; /* [maxN][ns][3] each */
    mo_rng rng;
    int g_have;             /* a second Box-Muller value is pending */
    double g_z2;
    double energy;          /* potential energy of the current configuration */
    long long att, acc, fail;
    double last_dH;
} mo_hmc_state;

/* Energy (kcal/mol) and forces F[N][ns][3] (may be NULL) of a configuration;
 * *ok = 0 on a map / locut failure. */
double mo_hmc_energy_forces(const mo_hmc_sys *s, int N, const double *rp, const double *r, double *F, int *ok);
/* random_gaussian (random.F90:197-217) */
double mo_hmc_gaussian(mo_hmc_state *st, double ave, double sigma);
/* One HMC move; returns 1 accepted, 0 rejected, -1 initial energy failure. */
int mo_hmc_move(const mo_hmc_sys *s, mo_hmc_state *st);
/* NVE velocity-Verlet trajectory from given velocities (no acceptance):
 * nsteps steps, writes total energy (U + KE/4.184, kcal/mol) after every
 * step into etot[nsteps + 1]; returns 0 on failure. */
int mo_hmc_nve(const mo_hmc_sys *s, int N, double *rp, double *r, double *v, int nsteps, double *etot);
/* energy parts map, LJ, stretch, bending, torsion (kcal/mol) */
#define MO_HMC_NPARTS 5
double mo_hmc_energy_parts(const mo_hmc_sys *s, int N, const double *rp, const double *r, double *F, int *ok,
                           double *parts);
/* plain NVE MD: per step the energy parts, sum m v^2, sum_m |sum m v|^2 / M^2
 * and the total momentum (MO_MD_NQ values); returns the steps done */
#define MO_MD_NQ 10
int mo_hmc_md(const mo_hmc_sys *s, int N, double *rp, double *r, double *v, int nsteps, double *trace);

double mo_bonded_term(int kind, const double *par, const double *r, double *f);

#ifdef __cplusplus
}
#endif
#endif
