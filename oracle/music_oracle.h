/*
 * music_oracle.h -- CPU FP64 restatement of the Music GCMC energy hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * library (libmusic_hip.so) and the timed CPU baseline in bench.py.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * it.  The product path never calls into it.
 *
 * Every function follows the reference arithmetic order of the Fortran
 * routine it cites (file:line relative to the Music-code source tree).
 * Pinning of each function (golden vectors, oracle/_ref built from the
 * reference's own Fortran sources, closed forms) is listed in
 * music_oracle.c next to the function.
 *
 * Conventions: Angstrom, kcal/mol, elementary charge.  Arrays are
 * C-contiguous; xyz arrays are [n][3].  Cells are orthorhombic with origin
 * 0 unless a mo_cell (slanted cells, fundcell.F90) is given.
 */
#ifndef MUSIC_ORACLE_H
#define MUSIC_ORACLE_H
#include <stdint.h>
#include <stddef.h>
#ifdef __cplusplus
extern "C" {
#endif

/* ---- constants (defaults.F90:33-145, ewald.F90:60, coulomb.F90:292) ---- */
#define MO_PI 3.1415927
#define MO_TWOPI (2.0 * MO_PI)
#define MO_KCALMOLE_KB (8.314e-3 / 4.184)
#define MO_CALTOJ 4.184
#define MO_NAV 6.0221367e23
#define MO_RGAS 8.314
#define MO_MAX_EXP 30.0
#define MO_LOW_DIST_CUTOFF 1.0e-3
#define MO_HIGH_DIST_CUTOFF 13.0
#define MO_VERY_LRG_NRG 1.0e10
#define MO_TOKCAL 332.064
#define MO_COUL_CONST 331.5

/* ---- RNG: random.F90:121-152 ---- */
typedef struct {
    int32_t ix, iy;
    int32_t iseed;
    int32_t first;
} mo_rng;
void mo_rng_init(mo_rng *r, int32_t iseed);
double mo_rranf(mo_rng *r);
/* fill n draws (test helper) */
void mo_rranf_n(int32_t iseed, int n, double *out, int32_t *ix_out, int32_t *iy_out);

/* ---- cells: fundcell.F90 / simcell.F90 ----
 * A fundamental cell (edges fell, angles alpha = bc, beta = ac, gamma = ab
 * in degrees, origin, effective box eff) and the simulation cell made of
 * ncell copies per axis.  U maps slanted to Cartesian coordinates, S = U^-1
 * (fundcell_initxform, fundcell.F90:378-412; matrix_inverse, matrix.F90:
 * 195-229); lvec[i] are the simulation cell's lattice vectors
 * (fundcell_genlatticevects, fundcell.F90:174-208). */
typedef struct {
    int ortho;                /* orthoflag: every angle within 0.001 degree of 90 */
    double fell[3], finv[3];  /* fundamental cell edges and their inverses */
    double ell[3], half[3];   /* simulation cell edges and halves */
    double origin[3];         /* fundamental cell origin */
    double eff[3];            /* fundamental cell effective (map) box */
    double U[3][3], S[3][3];  /* unslant, slant [row][column] */
    double lvec[3][3];        /* simulation cell lattice vectors [vector][xyz] */
} mo_cell;
void mo_cell_init(mo_cell *c, const double fell[3], const double ang_deg[3], const double origin[3],
                  const double eff[3], const int ncell[3]);
/* simcell_minimage (fundcell_snglMinImage, fundcell.F90:582-623): sep = a - b */
void mo_cell_minimage(const mo_cell *c, const double a[3], const double b[3], double sep[3]);
/* simcell_maptouc (simcell.F90:634-668) */
void mo_cell_maptouc(const mo_cell *c, const double r[3], double out[3]);
/* simcell_pbcpt (simcell.F90:460-484) */
void mo_cell_pbc(const mo_cell *c, const double r[3], double out[3]);

/* ---- math helpers ---- */
double mo_erfc(double x);                                   /* utils.F90:2090-2101 */
void mo_minimage(const double ell[3], const double a[3], const double b[3], double sep[3]);
void mo_lj_AB(double sig, double eps_K, double *A, double *B); /* lj.F90:860-869 */
void mo_lj_mix_lb(double sig1, double eps1, double sig2, double eps2, double *sig, double *eps);

/* single LJ pair with higher order terms (lj.F90:652-733).  Returns 2 when
 * the pair was evaluated, 1 when beyond hicut (nrg/hot untouched), 0 when
 * r2 < locut2 (failure). */
int mo_lj_pair_hot(double A, double B, double hicut2, double locut2,
                   const double sep[3], double *nrg, double hot[8]);

/* ---- map build nodes ---- */
/* LJ node: probe at p vs framework atoms (pairmodel.F90:856-911 atom loop,
 * mapmaker.F90:483-555 node assembly).  type[] indexes the per-type arrays;
 * on[t]==0 means no LJ for that type (reference would Stop; callers must not
 * pass it).  out[0] = V (or 1e10 on failure), out[1..7] = hot(2..8).
 * Returns 1 if the node succeeded. */
int mo_lj_node(const double p[3], int natoms, const double *xyz, const int *type,
               const double *A, const double *B, const double *hicut2,
               const double *locut2, const double full_ell[3], double out[8]);

/* Buckingham node (closed form; buck.F90:329 potential, analytic derivatives).
 * V = A exp(-r/rho) - C/r^6 - D/r^8 per pair. */
int mo_buck_node(const double p[3], int natoms, const double *xyz, const int *type,
                 const double *A, const double *rho, const double *C, const double *D,
                 const double *hicut2, const double *locut2, const double full_ell[3],
                 double out[8]);

/* Ewald structure factor (ewald.F90:558-681).  Returns nk (or -1 if more
 * than nkmax).  k: [nk][3], s, si: [nk].  kappa_out = kappaParam/min(ell).
 * kidx (optional, may be NULL): [nk][3] integer (x,y,z) of each k. */
int mo_ewald_sfactor(int natoms, const double *xyz, const double *q, const double ell[3],
                     double kappa_param, int kmax, double locut, int nkmax,
                     double *k, double *s, double *si, int *kidx, double *kappa_out);

/* Ewald emap node (ewald.F90:334-398, :856-918, :966-1005). */
int mo_ewald_node(const double p[3], int natoms, const double *xyz, const double *q,
                  double charge1, double kappa, int nk, const double *k,
                  const double *s, const double *si, const double full_ell[3],
                  double out[8]);

/* ---- grid (config.F90:735-808, mapmaker.F90:360-365) ---- */
void mo_grid_dims(const double eff[3], double gridspace, int npnts[3], double step[3]);

/* Build the full (uncompacted) map: vals[m][8] FP64 with node m in
 * config_makegrid order (x fastest).  kind: 1 = LJ, 2 = Ewald, 3 = Buck.
 * Per-type arrays as above; for Ewald the sfactor arrays are passed. */
typedef struct {
    int natoms;
    const double *xyz;
    const int *type;
    const double *q;
    const double *A, *B, *hicut2, *locut2;   /* LJ (per type) */
    const double *bA, *brho, *bC, *bD;        /* Buckingham (per type) */
    double charge1;
    double kappa;
    int nk;
    const double *k, *s, *si;
    double full_ell[3];
    const mo_cell *cell;  /* slanted cell, or NULL: orthorhombic full_ell */
} mo_framework;
void mo_build_nodes(int kind, const mo_framework *fw, const int npnts[3], const double step[3],
                    long long m0, long long m1, double *vals /* [(m1-m0)][8] */);
/* one node at an arbitrary point */
int mo_node_at(int kind, const mo_framework *fw, const double p[3], double out[8]);
/* simcell_uniformRandPt (simcell.F90:268-295) */
void mo_cell_uniform(const mo_cell *c, mo_rng *rng, double out[3]);
/* simcell_getvolume (simcell.F90:300-304, fundcell_genvolume fundcell.F90:
 * 159-165): (l1 x l2) . l3 of the simulation cell's lattice vectors */
double mo_cell_volume(const mo_cell *c);

/* ---- compaction (pmap.F90:509-656) ----
 * r: FP32 potential per node in grid order (x fastest), cutoff in kcal/mol.
 * ptr_out: [nx][ny][nz] flattened with z fastest (Fortran ptr(k,j,i)),
 * 1-based compact indices, 0 = untabulated.  src_out (optional): for each
 * compact index n (0-based), the source node m (0-based grid order).
 * counts: acceptCut, acceptNeighbor, naccept, reject. */
long long mo_compact(const int npnts[3], const float *r, double cutoff,
                     int32_t *ptr_out, long long *src_out, long long counts[4]);

/* ---- map file (mapheader.F90:178-232, interpolate.F90:922-938) ---- */
typedef struct {
    char name[26];
    char symbol[3];
    float charge, A12, B6, C2, D0, rcuthi, rcutlo;
} mo_zatom;
int mo_map_write(const char *path, int nzatoms, const mo_zatom *zatoms, float pcut_off,
                 const int npnts[3], long long nfsize, const int32_t *ptr, const float *vals8 /* [8][nfsize] SoA */);
/* read header sizes; returns 0 on success */
int mo_map_read_sizes(const char *path, int npnts[3], long long *nfsize, float *pcut_off, int *nzatoms);
int mo_map_read(const char *path, int32_t *ptr, float *vals8);

/* ---- interpolation (interpolate.F90:104-158, :603-874; simcell.F90:634-668) ---- */
void mo_maptouc(const double fund_ell[3], const double r[3], double out[3]);
/* returns 1 if all 8 corners tabulated, p = potential (periodic corner
 * wrap as for orthorhombic cells; mo_pothermite_nowrap without it) */
int mo_pothermite_nowrap(const int npnts[3], const double step[3], const int32_t *ptr, const float *vals8,
                         long long nfsize, const double coord[3], double *p);
int mo_pothermite(const int npnts[3], const double step[3], const int32_t *ptr,
                  const float *vals8, long long nfsize, const double coord[3], double *p);
/* forcehermite: p[0]=V, p[1..3] = gradient terms (interpolate.F90:189-584) */
int mo_forcehermite(const int npnts[3], const double step[3], const int32_t *ptr,
                    const float *vals8, long long nfsize, const double coord[3], double p[4]);

#ifdef __cplusplus
}
#endif
#endif
