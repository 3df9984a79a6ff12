/*
 * music_hip.h -- C-ABI of libmusic_hip.so, the MI355X-native GCMC energy
 * hot path of Music (map build, map interpolation, sorbate-sorbate energies,
 * Ewald structure-factor deltas, device-resident GCMC).
 *
 * Plain C types only (pointers + sizes).  Every function returns an int
 * status: MH_OK (0) or a positive MH_E* code; mh_last_error() gives the
 * message of the last failure on the calling thread.  Nothing here aborts
 * the process.  One context per HIP device; a context is used by one host
 * thread at a time.
 *
 * Units: Angstrom, kcal/mol (maps and energies), elementary charge, K.
 * Cells have their origin at 0 and are orthorhombic (the reference's fast
 * path, fundcell.F90:640-709) unless mh_framework_set_cell gives a slanted
 * one; the calls that do not handle slanted cells reject them with MH_EINVAL.
 * Indices passed in are 0-based; Fortran callers subtract 1.
 */
#ifndef MUSIC_HIP_H
#define MUSIC_HIP_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

enum {
    MH_OK = 0,
    MH_EINVAL = 1,    /* bad argument / unsupported input */
    MH_ENOMEM = 2,    /* device allocation failed */
    MH_EHIP = 3,      /* HIP runtime error */
    MH_EIO = 4,       /* map file I/O */
    MH_ESTATE = 5,    /* call order (e.g. map build before framework) */
    MH_ECAPACITY = 6  /* a fixed capacity (k vectors, molecules) exceeded */
};

enum { MH_MAP_LJ = 1, MH_MAP_EWALD = 2, MH_MAP_BUCK = 3 };
/* MH_PREC_FP64: FP64 everywhere, follows the reference arithmetic.
 * MH_PREC_FP32: FP32 pair arithmetic, FP32 partial sums per atom tile,
 *               FP64 Kahan accumulation across tiles; Ewald reciprocal part
 *               stays FP64. */
enum { MH_PREC_FP64 = 0, MH_PREC_FP32 = 1 };

typedef struct mh_ctx mh_ctx;
typedef struct mh_map mh_map;

const char *mh_last_error(void);
const char *mh_version(void);

/* Replaces the process-wide module state the reference initialises in
 * interact_init / forcefield_init (interact.F90:121-287). */
int mh_ctx_create(int device, mh_ctx **out);
int mh_ctx_destroy(mh_ctx *ctx);
/* Per-kernel timing with HIP events on the context's stream (no
 * counterpart in the reference, whose timers are wtime.F90 CPU clocks).
 * mh_ctx_get_timing writes a JSON object {"kernel": [ms_total, launches]}. */
int mh_ctx_set_timing(mh_ctx *ctx, int on);
int mh_ctx_get_timing(mh_ctx *ctx, char *buf, int len);

/* The fixed framework species (config_init "Fixed", config.F90:179-) after
 * replication to the simulation cell (simcell.F90:200-216).
 * xyz[natoms][3], type[natoms] (atom-type index), q[natoms].
 * fund_ell: fundamental cell edges (grid and map folding, simcell_geteff
 * with fund=.True.); full_ell: simulation cell edges (minimum image,
 * Ewald, insertion volume). */
int mh_framework_upload(mh_ctx *ctx, int natoms, const double *xyz, const int *type,
                        const double *q, const double fund_ell[3], const double full_ell[3]);

/* LJ parameters of the probe atom against each framework atom type
 * (lj_init / lj_calc_AB, lj.F90:88-262, 860-869): A, B in kcal/mol A^12/A^6,
 * squared cutoffs.  on[t] = 0 marks an OFF pair (the reference's map build
 * stops on it, pairmodel.F90:903-907; here such atoms contribute nothing). */
int mh_ff_set_lj(mh_ctx *ctx, int ntypes, const double *A, const double *B,
                 const double *hicut2, const double *locut2, const uint8_t *on);

/* Buckingham A exp(-r/rho) - C r^-6 - D r^-8 per framework type (the
 * reference's buck_snglint, buck.F90:292-358, stops before evaluating). */
int mh_ff_set_buck(mh_ctx *ctx, int ntypes, const double *A, const double *rho,
                   const double *C, const double *D, const double *hicut2,
                   const double *locut2);

/* Slanted fundamental cell of the framework just uploaded (fundcell_init,
 * fundcell_initxform, fundcell.F90:60-112, 378-412): angles alpha (b^c),
 * beta (a^c), gamma (a^b) in degrees and the effective (orthorhombic) box of
 * the maps, eff.  Within 0.001 degree of 90 the cell is orthorhombic and
 * only eff is taken.  A slanted cell folds minimum images in its slanted
 * basis (fundcell_multMinImage, :640-709) in the map build -- the FP64
 * node kernel serves both precisions there -- and map interpolation folds
 * points into the slanted unit cell without the periodic wrap of the upper
 * corner (simcell_maptouc, simcell.F90:634-668; interpolate.F90:675-679).
 * GCMC batches (mh_gcmc_create) in a slanted cell insert along its lattice
 * vectors (simcell_uniformRandPt, simcell.F90:274-295), fold sites and pair
 * separations in the slanted basis and take the cell volume as (l1 x l2).l3
 * for the Adams B; they refuse sorbate Ewald sums and biased insertion
 * there.  Slot pools and HMC batches refuse slanted cells.
 * mh_framework_upload resets the cell to orthorhombic with eff = fund_ell. */
int mh_framework_set_cell(mh_ctx *ctx, const double angles_deg[3], const double eff[3]);

/* Ewald parameters and framework structure factor (ewald_init,
 * ewald_initCalc -> ewald_sfactor, ewald.F90:99-216, 558-681).  kappa =
 * kappa_param / min(full_ell).  Returns the number of accepted k vectors.
 * sign_mode selects the reciprocal-space part of Ewald maps:
 * MH_EWALD_PARITY reproduces ewald_skwaldHOT (ewald.F90:966-1005), whose
 * potential and second derivatives use s cos - si sin (the potential of the
 * framework inverted through the origin) while the first and third
 * derivatives use s sin - si cos; MH_EWALD_PHYSICS gives the potential of
 * the framework itself, s cos + si sin, with all seven derivatives its own.
 * The two agree when the framework is centrosymmetric about the origin. */
enum { MH_EWALD_PARITY = 0, MH_EWALD_PHYSICS = 1 };
int mh_ewald_set(mh_ctx *ctx, double kappa_param, int kmax, double locut, int sign_mode, int *nk_out);
/* copy out k[nk][3], s[nk], si[nk] (any pointer may be NULL) */
int mh_ewald_get(mh_ctx *ctx, double *k, double *s, double *si, double *kappa);

/* Map build: mapmaker_makemap (mapmaker.F90:422-598) = grid (config_makegrid,
 * config.F90:735-808) + per-node V and 7 derivatives + FP32 store
 * (pmap_addpoint, pmap.F90:449-500) + compaction (pmap_compact,
 * pmap.F90:509-656) with cutoff hipotcut_kJ / 4.184.
 * probe_q is the probe charge for Ewald maps (charge1 of
 * ewald_getInteractionHOT).  keep_full != 0 keeps the uncompacted FP32 and
 * FP64 (FP64 precision only) node values for mh_map_get_full. */
int mh_map_build(mh_ctx *ctx, int kind, int precision, double probe_q, double gridspace,
                 double hipotcut_kJ, int keep_full, mh_map **out);
int mh_map_destroy(mh_map *map);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): mh_map_build for LJ and Buckingham maps that calls no
 * hipMalloc / hipFree / allocating library routine.  The node values and
 * compaction arrays live in a file-scope device array (about 100 MB per
 * device, held until the process ends) of at most 2^21 grid nodes; larger
 * grids return MH_ECAPACITY.  The compacted map goes to the caller's device
 * buffers: ptr_dev int32[nx*ny*nz] (the mh_map_get layout) and vals_dev
 * float[vals_capacity][8] (record-major: r, rx, ry, rz, rxy, rxz, ryz, rxyz
 * per tabulated point).  ptr_dev == NULL only fills npnts_out (size query).
 * nfsize_out receives the tabulated count even when it exceeds
 * vals_capacity (then MH_ECAPACITY, nothing written to vals_dev).  Calls
 * on one device share the scratch array: calls that overlap on one device
 * race. */
int mh_map_build_static(mh_ctx *ctx, int kind, int precision, double gridspace, double hipotcut_kJ,
                        int32_t *ptr_dev, float *vals_dev, int64_t vals_capacity, int npnts_out[3],
                        int64_t *nfsize_out, int64_t counts_out[4]);

/* npnts[3] (nodes per axis), nfsize (tabulated points), counts[4] =
 * accepted below cutoff, accepted neighbours, total, rejected (pmap.F90:
 * 651-654), step[3] (grid spacing). */
int mh_map_info(const mh_map *map, int npnts[3], int64_t *nfsize, int64_t counts[4],
                double step[3]);
/* Compacted map in the file layout: ptr[nx][ny][nz] (z fastest, Fortran
 * ptr(k,j,i)), 1-based, 0 = untabulated; vals8[8][nfsize] SoA order
 * r, rx, ry, rz, rxy, rxz, ryz, rxyz (interpolate.F90:30-36). */
int mh_map_get(const mh_map *map, int32_t *ptr, float *vals8);
/* Uncompacted node values in config_makegrid order (x fastest), [m][8]
 * (requires keep_full).  f64 may be NULL; for FP32 builds it receives the
 * FP64 accumulators. */
int mh_map_get_full(const mh_map *map, float *f32, double *f64);

/* Map file (mapheader_makeOldHeader + interpolate_writecore,
 * mapheader.F90:178-232, interpolate.F90:922-938; Fortran unformatted
 * sequential, 4-byte record markers).  zatoms: nzatoms records of
 * name[25] symbol[2] (space padded, not NUL terminated) and 7 floats
 * charge, A12, B6, C2, D0, rcuthi, rcutlo.  pcut_kJ goes to pcut_off. */
int mh_map_write(const mh_map *map, const char *path, int nzatoms, const char *names25,
                 const char *symbols2, const float *zparams7, float pcut_kJ);
/* pmap_readmap (pmap.F90:321-358): read a map file onto the device.
 * fund_ell gives the grid spacing eff/(nbr-1) (mapheader.F90:140-146): the
 * map's effective box, which is the fundamental cell unless the context
 * holds a slanted cell (mh_framework_set_cell), whose points then fold into
 * that cell as in the map build. */
int mh_map_load(mh_ctx *ctx, const char *path, const double fund_ell[3], mh_map **out);
/* Upload a map held by the caller (same layout as mh_map_get). */
int mh_map_from_host(mh_ctx *ctx, const int npnts[3], int64_t nfsize, const int32_t *ptr,
                     const float *vals8, const double fund_ell[3], mh_map **out);

/* Map interpolation at n points (maps_int -> pmap_int -> interpolate_int ->
 * simcell_maptouc + pothermite, maps.F90:87-110, interpolate.F90:104-158,
 * 603-874).  xyz[n][3] host arrays; e[n] = V * factor; ok[n] = 0 where a
 * corner is untabulated (the move fails, mapflag = .False.). */
int mh_map_interp(mh_ctx *ctx, const mh_map *map, int n, const double *xyz, double factor,
                  double *e, uint8_t *ok);

/* ------------------------------------------------------------------ GCMC
 * Device-resident grand-canonical MC of one rigid sorbate species in a
 * fixed framework (gcmc_dosim, gcmc.F90:224-258), many independent replicas
 * (state points) per batch, one workgroup per replica, one persistent kernel
 * launch per mh_gcmc_run call.  Moves RINSERT / RDELETE / RTRANSLATE /
 * RROTATE (rigidmoves.F90:691-1570) with the reference's RNG (rranf_NR,
 * random.F90:121-152, one stream per replica) and draw order, acceptance
 * mcmoves_attempt (mcmoves.F90:290-335).  Molecule energies: framework maps
 * (interpolate_int, factor q for emaps) + sorbate-sorbate LJ with minimum
 * image (ssbasic_msint, lj_snglint) + optional Ewald electrostatics with
 * structure-factor deltas (real space + reciprocal + self + intramolecular
 * exclusion; no reference counterpart: the reference stops,
 * ssdriver.F90:489-498).  Failures (LJ locut, untabulated map corner,
 * U > max_nrg) reject the move without an RNG draw. */
/* ------------------------------------------------------------- bias map
 * Bias map of biased insertion / deletion (bmap_init + bmap_cubeinit,
 * bmap.F90:69-403), built on the device from a pmap at bias_temperature
 * (K): the cubelets whose lower-corner node is tabulated with V <= 30
 * kcal/mol, in the reference's order (x outer, z inner), Boltzmann weights
 * exp(-V / kT_bias) (exponent clamped to +-30; energies above
 * -log(1e-6 sum) / rti lowered to it when min / sum <= 1e-6) and their
 * cumulative probabilities, summed in that order. */
typedef struct mh_bmap mh_bmap;
int mh_bmap_create(mh_ctx *ctx, const mh_map *pmap, double bias_temperature, mh_bmap **out);
int mh_bmap_destroy(mh_bmap *b);
/* numtab: biased cubelets; ncubelets: all cubelets; cum_wt: the weight total;
 * clamp_energy: the energy above which weights were clamped (+inf: none).
 * Any pointer may be NULL. */
int mh_bmap_info(const mh_bmap *b, int64_t *numtab, int64_t *ncubelets, double *cum_wt, double *clamp_energy);
/* Per biased cubelet in list order: id (i * (ny - 1) + j) * (nz - 1) + k
 * (0-based), weight, cumulative probability.  Any pointer may be NULL. */
int mh_bmap_get(const mh_bmap *b, int32_t *cube, double *wt, double *cum_prob);
/* The list position (0-based) the GCMC kernel's search returns for each key
 * in [0, 1): that of bmap_getbiasindex (bmap.F90:575-636). */
int mh_bmap_search(const mh_bmap *b, int n, const double *keys, int64_t *index);

/* BINSERT / BDELETE: biased insertion and deletion (rigidmoves_binsert,
 * rigidmoves.F90:1386-1442 with rigidmoves_comFromBmap :1938-1969;
 * rigidmoves_bdelete :1578-1654) with the bias map params->bias.  A
 * biased insertion draws a cubelet by weight, a point in it and a unit cell
 * of the simulation cell, and enters the acceptance with 1 / (w ncubelets);
 * a biased deletion enters with w ncubelets of the molecule's cubelet (w =
 * weight / total).  Attempts and acceptances count as insert (index 0) and
 * delete (index 1).  Deleting a molecule whose cubelet is not in the bias
 * map stops the reference; here the replica stops with error 4 and
 * mh_gcmc_run returns MH_ESTATE. */
enum { MH_MV_INSERT = 0, MH_MV_DELETE = 1, MH_MV_TRANSLATE = 2, MH_MV_ROTATE = 3, MH_MV_BINSERT = 4,
       MH_MV_BDELETE = 5 };
#define MH_MAX_SITES 8
#define MH_MAX_SPECIES 4

typedef struct {
    int nsites;                          /* 1..8 */
    double ref[MH_MAX_SITES][3];         /* body-frame site coordinates */
    int site_type[MH_MAX_SITES];         /* index into the sorbate LJ table */
    double site_q[MH_MAX_SITES];         /* e */
    const mh_map *site_pmap[MH_MAX_SITES]; /* LJ map of each site or NULL */
    const mh_map *site_emap[MH_MAX_SITES]; /* electrostatic map or NULL (factor q) */
    int ntypes;                          /* sorbate site types, 1..8 */
    double lj_A[64], lj_B[64], lj_hicut2[64], lj_locut2[64]; /* [ntypes][ntypes] */
    uint8_t lj_on[64];
    int ewald;                           /* 1: Ewald sorbate-sorbate electrostatics */
    double ewald_kappa;                  /* 1/A */
    int ewald_kmax;                      /* |m_i| <= kmax, half space */
    double ewald_rcut;                   /* real-space cutoff, A */
    double temperature;                  /* K */
    int nmoves;
    int move_type[4];                    /* MH_MV_* */
    double move_weight[4];
    double delta_trans, delta_rot;       /* fixed displacements (A, rad) */
    double max_nrg;                      /* kcal/mol (1e10 in the reference) */
    int precision;                       /* MH_PREC_FP64 or MH_PREC_FP32 */
    const mh_bmap *bias;                 /* bias map of BINSERT / BDELETE, or NULL;
                                            kept alive by the caller while the batch lives */
    /* plain Coulomb between sorbate sites (coulomb_snglint, coulomb.F90:258-327;
     * atom_atom_file COUL lines under an spc_spc "COUL BASIC"): for the type
     * pairs with coul_on, 331.5 shield q_a q_b (1/r - 1/hicut) kcal/mol within
     * hicut (1/r for all minimum-image pairs when hicut = 0); r < locut fails
     * the evaluation.  Exclusive with ewald. */
    int coul;
    uint8_t coul_on[64];                 /* [ntypes][ntypes] */
    double coul_hicut, coul_locut, coul_shield;
    /* adjustable displacements ("adjust delta" = 1 on the RTRANSLATE /
     * RROTATE lines): after every attempt of the move its displacement is
     * scaled by 0.95 while the move's cumulative acceptance ratio is below
     * 0.49 and by 1.05 while above 0.51, within [0.01, 2.0]
     * (rigidmoves_adjustdeltatrans / _adjustdeltarot, rigidmoves.F90:2586-2617,
     * called from mcmoves_perturb, mcmoves.F90:460-461).  Per replica. */
    int adjust_trans, adjust_rot;
    /* Centre-of-mass cutoffs of the sorbate pair terms (the species-species
     * options COMCUT@r / COMTRUNC@r of ssbasic_init, ssbasic.F90:139-160,
     * applied in ssbasic_mmint, ssbasic.F90:745-756): radius (A) for the pairs
     * of this species' molecules with those of species s2; a molecule pair
     * whose geometric centres (the mean of the unfolded site positions,
     * config_getMolecCTR, config.F90:1788-1802) are further apart, minimum
     * image, has no LJ (lj_comcut) or plain-Coulomb (coul_comcut) term; the
     * Ewald real-space term is not affected.  0 = no cutoff.  The radii of a
     * mixture must be symmetric (species s to s2 = s2 to s). */
    double lj_comcut[MH_MAX_SPECIES], coul_comcut[MH_MAX_SPECIES];
} mh_gcmc_params;

typedef struct mh_gcmc mh_gcmc;

/* nrep replicas; fugacity per replica (kPa, ideal gas: Adams B from
 * gcmc_genfuglist, gcmc.F90:303-319); seeds per replica (iseed of
 * random_init); maxN = molecule capacity per replica. */
int mh_gcmc_create(mh_ctx *ctx, const mh_gcmc_params *p, int nrep, const double *fugacity_kPa,
                   const int32_t *seeds, int maxN, mh_gcmc **out);
/* A mixture of nspc (1..MH_MAX_SPECIES) sorbates (gcmc_dosim, gcmc.F90:
 * 224-258: every step draws a species with equal probability, then a move
 * of its move set; its molecules interact with those of every species).
 * spc[s] gives species s's model, maps, move set, displacements, bias map
 * and max_nrg; the pair tables (site types of all species index them),
 * Coulomb / Ewald settings, temperature and precision are spc[0]'s.
 * fugacity_kPa[nrep][nspc]; maxN = molecule capacity per species and
 * replica.  mh_gcmc_create is this call with nspc = 1.  Traces record
 * move = 16 species + move type and N of the moved species. */
int mh_gcmc_create_mix(mh_ctx *ctx, const mh_gcmc_params *spc, int nspc, int nrep, const double *fugacity_kPa,
                       const int32_t *seeds, int maxN, mh_gcmc **out);

/* Cavity-biased insertion and deletion (BINSERT line "CAVITY" = CAVITY_BIAS
 * or "bmap, CAVITY" = MAP_CAVITY_BIAS; rigidmoves_initInsBias,
 * rigidmoves.F90:620-670).  The cavity list (cavitylist.F90, control file
 * section "Cavity Bias Info", cavitylist_readCtrlFile :1022-1082) is a grid
 * of n[0] x n[1] x n[2] cubes over the simulation cell (edge ell / n); a
 * cube is a cavity when it is tabulated (tab[n0][n1][n2], z fastest, 1 =
 * tabulated; NULL = every cube, cavitylist_initArrUniform) and no site of a
 * molecule of the species in species_mask (bit s) lies in it.  A cavity-
 * biased insertion of species s (rigidmoves_getConfigFromCavBias, :1987-
 * 2126) draws ntrials[s] centres, uniform in the cell or, when
 * map_bias[s], from the species' bias map (params->bias), each with random
 * angles; the trials whose sites all lie in cavities are counted into the
 * running average cav_prob(totmol) (totmol = molecules of all species with
 * the new one; every entry starts with one update of ntrials / 2,
 * rigidmoves_equateInsBias :2748-2760); with none the insertion fails;
 * otherwise one of them at random enters the acceptance with its map bias x
 * cav_prob / ntrials.  The species' BDELETE divides its bias by cav_prob /
 * ntrials at the current total (rigidmoves_bdelete :1632-1643).
 * ntrials[s] = 0: species s moves without cavity bias.  ntrials <= 1024;
 * launches use at least ntrials threads.  Orthorhombic cells, not for slot
 * pools.  Cube counts follow every accepted move and mh_gcmc_set_config. */
typedef struct {
    int n[3];
    const uint8_t *tab;
    int species_mask;
    int ntrials[MH_MAX_SPECIES];
    int map_bias[MH_MAX_SPECIES];
} mh_cavity;
int mh_gcmc_set_cavity(mh_gcmc *g, const mh_cavity *c);
/* cav_prob(t) of the selected species for t < cap (running average, update
 * count) and the site count of every cube, of replica rep (any pointer may
 * be NULL). */
int mh_gcmc_get_cavity(mh_gcmc *g, int rep, int cap, double *avg, int64_t *n, int32_t *cubes);
/* cavitylist_initArrPmap (cavitylist.F90:790-978): the tabulation from a
 * pmap: the cube (edge full_ell / n) holding the centre of each map cubelet
 * whose lower-corner node is tabulated with V < ucutoff (kcal/mol), marked
 * in each of the ncell unit cells of the simulation cell. */
int mh_cavity_tab(const mh_map *pmap, double ucutoff, const int n[3], const int ncell[3], const double full_ell[3],
                  uint8_t *tab);
/* Species addressed by the per-species calls that follow (mh_gcmc_get_state's
 * N / sumN / att / acc, mh_gcmc_get_moments, mh_gcmc_get_deltas /
 * set_deltas, mh_gcmc_set_counts, mh_gcmc_set_config / get_config,
 * mh_trial_energy); 0 after creation.  Energies, step counts and RNG states
 * are per replica. */
int mh_gcmc_species(mh_gcmc *g, int spc);
int mh_gcmc_destroy(mh_gcmc *g);
/* Threads per replica (64..1024, power of 2); chosen by the autotuner. */
int mh_gcmc_set_threads(mh_gcmc *g, int threads);
/* Autotuner (no reference counterpart: the reference is serial): times
 * probe_steps on every replica for each candidate workgroup size (64, 128,
 * 256, 512) and reciprocal-space table mode from a snapshot of the batch,
 * restores the snapshot and keeps the fastest.  threads_out / exy_out may be
 * NULL.  Results are identical for every candidate (FP64: bit-identical
 * moves; only the schedule differs). */
int mh_gcmc_autotune(mh_gcmc *g, int probe_steps, int *threads_out, int *exy_out);
/* Place n molecules (com[n][3], Euler angles phi, theta, psi [n][3]) in a
 * replica; recomputes its structure factor and energy. */
int mh_gcmc_set_config(mh_gcmc *g, int rep, int n, const double *com, const double *ang);
/* nsteps MC steps on every replica in one kernel launch. */
int mh_gcmc_run(mh_gcmc *g, int nsteps);
/* ncycles MC cycles of nsteps steps (gcmc_dosim per cycle, the driver loop
 * music_gcmc.F90:203-237): the persistent launch of one cycle is captured
 * once in a hipGraph and replayed per cycle, with one host synchronization at
 * the end; re-captured when the batch's launch configuration or buffers
 * change.  Same results as ncycles calls of mh_gcmc_run(g, nsteps). */
int mh_gcmc_run_cycles(mh_gcmc *g, int nsteps, int ncycles);
/* Per replica: N, running energy (kcal/mol), sum of N over steps, steps,
 * att/acc [nrep][4] by move type, RNG state (for checkpoints).  Any
 * pointer may be NULL. */
int mh_gcmc_get_state(mh_gcmc *g, int32_t *N, double *energy, double *sumN, int64_t *nsteps,
                      int64_t *att, int64_t *acc, int32_t *rng_ix, int32_t *rng_iy);
/* Per replica: sums over the steps run of N and N^2, from which a host forms
 * the running averages of stats_update (stats.F90:152-195): cumulative mean
 * sumN / steps, standard deviation sqrt(sumN2 / steps - mean^2), and the
 * block mean from the difference of two readings.  Either pointer may be
 * NULL. */
int mh_gcmc_get_moments(mh_gcmc *g, double *sumN, double *sumN2);
/* Restore a checkpoint: N and RNG state per replica (the configuration via
 * mh_gcmc_set_config).  The reference does not save the RNG state. */
int mh_gcmc_set_rng(mh_gcmc *g, int rep, int32_t ix, int32_t iy);
/* Displacements in use per replica, delta[nrep][2] = (translate A, rotate
 * rad): the params' values, changed by adjustable moves. */
int mh_gcmc_get_deltas(mh_gcmc *g, double *delta);
/* Restore a replica's displacements (a checkpoint of adjustable moves). */
int mh_gcmc_set_deltas(mh_gcmc *g, int rep, double delta_trans, double delta_rot);
/* Restore a replica's attempt / acceptance counts att[4], acc[4] by move
 * type (a checkpoint): adjustable displacements follow the cumulative ratio
 * acc / att of their move (moves_adjustParams, moves.F90:553-560), so an
 * exact resume needs the counts as well as the displacements. */
int mh_gcmc_set_counts(mh_gcmc *g, int rep, const int64_t *att, const int64_t *acc);
int mh_gcmc_get_config(mh_gcmc *g, int rep, int *n, double *com, double *ang, double *r);
/* Full FP64 recompute of every replica's energy (interact_checknrgs); with
 * refresh != 0 the structure factors and running energies are reset to it. */
int mh_gcmc_total_energy(mh_gcmc *g, double *e, int refresh);
/* Per-step trace (move, accepted, N after, energy change) of every replica,
 * capacity steps from the next run on. */
int mh_gcmc_set_trace(mh_gcmc *g, int capacity);
int mh_gcmc_get_trace(mh_gcmc *g, int rep, int *move, int *accepted, int *N, double *dE, int *ntrace);
/* Insertion energies of n trial molecules (com, Euler angles) into the
 * current configuration of replica rep (subinteract_int of a new molecule),
 * in the batch precision; ok = 0 where the energy evaluation fails. */
int mh_trial_energy(mh_gcmc *g, int rep, int n, const double *com, const double *ang, double *e,
                    uint8_t *ok);

/* ------------------------------------------------------------------- HMC
 * Hybrid MC bursts of flexible sorbates in a fixed framework
 * (hybridmoves_integrate_boltz, hybridmoves.F90:186-241), many replicas per
 * batch, one workgroup per replica: Gaussian velocities per atom with the
 * reference RNG and Box-Muller order (hybridmoves_boltzVel,
 * hybridmoves.F90:247-307; random_gaussian, random.F90:197-217), md_steps
 * NVE velocity-Verlet steps (velverlet_nveIntegrate, velverlet.F90:148-288),
 * Metropolis on dU + dKE (mcmoves_perturb, mcmoves.F90:371-426).  Energy:
 * framework maps with forces (forcehermite, interpolate.F90:189-584),
 * sorbate-sorbate LJ with minimum image (lj_snglint), harmonic bonds
 * (harmonic.F90:130-160), harmonic angles (harmonica.F90:100-160),
 * cosine-series torsions (cosexpansion.F90:88-196, polymer convention).
 * Units: A, ps, amu; energies kcal/mol; bond k kcal/mol/A^2, angle k
 * kcal/mol/rad^2 (U = k/2 (x - x0)^2), torsion U = sum_n c[n] cos^n. */
typedef struct {
    int nsites;                          /* 1..8 sites per molecule */
    int site_type[MH_MAX_SITES];         /* index into the LJ table */
    double site_mass[MH_MAX_SITES];      /* amu */
    const mh_map *site_map[MH_MAX_SITES];/* LJ map per site or NULL */
    int ntypes;
    double lj_A[64], lj_B[64], lj_hicut2[64], lj_locut2[64]; /* [ntypes][ntypes] */
    uint8_t lj_on[64];
    int nbonds;
    int bond[16][2];
    double bond_k[16], bond_r0[16];
    int nangles;
    int angle[16][3];                    /* i-j-k, j the apex */
    double angle_k[16], angle_theta0[16];/* rad */
    int ntorsions;
    int torsion[16][4];
    double torsion_c[16][6];
    double temperature;                  /* K */
    double dt_ps;                        /* time step, ps */
    int md_steps;                        /* VV steps per move */
    double max_nrg_md;                   /* kcal/mol, energy check per step (1000 in the reference) */
    double max_nrg_start;                /* kcal/mol, start configuration (1e6) */
    int precision;                       /* MH_PREC_FP64 or MH_PREC_FP32 (FP32 pair and map arithmetic) */
} mh_hmc_params;

typedef struct mh_hmc mh_hmc;

int mh_hmc_create(mh_ctx *ctx, const mh_hmc_params *p, int nrep, const int32_t *seeds, int maxN, mh_hmc **out);
int mh_hmc_destroy(mh_hmc *h);
/* Threads per replica: 128, 256, 512 or 1024 (default 512; 1024 in FP64 when
   maxN * nsites > 512). */
int mh_hmc_set_threads(mh_hmc *h, int threads);
/* n molecules of a replica, principal (unwrapped) site coordinates
 * rp[n][nsites][3]; the simulation-cell coordinates are their pbc images. */
int mh_hmc_set_config(mh_hmc *h, int rep, int n, const double *rp);
/* nmoves HMC moves on every replica in one launch. */
int mh_hmc_run(mh_hmc *h, int nmoves);
/* Per replica: N, potential energy of the current configuration, attempted,
 * accepted and failed (energy evaluation) moves, dH of the last move
 * (NaN if it failed), RNG state.  Any pointer may be NULL. */
int mh_hmc_get_state(mh_hmc *h, int32_t *N, double *energy, int64_t *att, int64_t *acc, int64_t *fail,
                     double *last_dH, int32_t *rng_ix, int32_t *rng_iy);
int mh_hmc_get_config(mh_hmc *h, int rep, int *n, double *rp, double *r);
/* Energy and forces F[n][nsites][3] (kcal/mol/A) of a replica's current
 * configuration in the batch precision; ok = 0 on a failed evaluation. */
int mh_hmc_energy_forces(mh_hmc *h, int rep, double *e, double *F, int *ok);

/* ---- plain molecular dynamics on the same replicas (md_doSim with the NVE
 * velocity-Verlet integrator, md.F90:592-623, velverlet.F90:148-288; the
 * reference's MD driver music.F90 with the "MD Information" ctrlfile
 * section).  Velocities v[n][nsites][3] in A/ps (the restart file's
 * _Velocities_ section, config.F90:2114-2430). */
int mh_hmc_set_velocities(mh_hmc *h, int rep, int n, const double *v);
int mh_hmc_get_velocities(mh_hmc *h, int rep, int *n, double *v);
/* Quantities per replica and step in the trace of mh_hmc_md: energy parts
 * (kcal/mol) framework map, sorbate-sorbate LJ, stretch, bending, torsion;
 * sum m v^2 over sites and sum over molecules |sum m v|^2 / M^2
 * (config_kineticTemp, config.F90:2528-2615; amu A^2/ps^2 and A^2/ps^2);
 * total momentum sum m v (amu A/ps). */
#define MH_MD_NQ 10
/* nsteps NVE velocity-Verlet steps (dt of the params) on every replica from
 * the current positions and velocities; trace[nrep][nsteps][MH_MD_NQ]
 * (host memory) receives the values after each step.  A failed energy
 * evaluation (overlap below locut, a site off its map) stops the replica
 * and returns MH_ESTATE (the reference stops the run). */
int mh_hmc_md(mh_hmc *h, int nsteps, double *trace);

/* ------------------------------------------------------------ slot pools
 * Continuous batching of GCMC requests (no counterpart in the reference,
 * which runs one simulation per process).  A pool is an mh_gcmc handle
 * with nslots replica slots and up to nsys systems (sorbate model + maps +
 * cell), each slot carrying its own system and step budget.  One
 * mh_pool_launch advances every occupied slot by min(nsteps, its remaining
 * budget) in one persistent kernel; requests are admitted into free slots
 * and released between launches, so work of different requests shares the
 * device.  A slot's trajectory depends only on its own system, seed and
 * budget: results are identical whether a request runs alone or with
 * others.  mh_gcmc_get_config / set_config / total_energy / trial_energy
 * take slot indices. */
int mh_pool_create(mh_ctx *ctx, int nslots, int nsys, int maxN, int kmax_cap, int max_sites, int precision,
                   mh_gcmc **out);
/* Admit a request: nrep replicas of one system (params, its cell edges;
 * params->precision is ignored, the pool's applies), each running nsteps
 * steps.  ticket receives the request's id.  MH_ECAPACITY when no system
 * entry or not enough free slots are left (admit again after a release). */
int mh_pool_admit(mh_gcmc *g, const mh_gcmc_params *p, const double fund_ell[3], const double full_ell[3],
                  int nrep, const double *fugacity_kPa, const int32_t *seeds, int nsteps, int *ticket);
/* Enqueue one launch of up to nsteps steps per occupied slot (asynchronous). */
int mh_pool_launch(mh_gcmc *g, int nsteps);
/* Wait for the launches enqueued so far. */
int mh_pool_wait(mh_gcmc *g);
/* Per replica of a ticket (admit order): slot index, remaining steps,
 * error flag (2 = reached maxN, 3 = evaluation failure of an existing
 * molecule), N, energy, sum of N over steps, steps done, att/acc [nrep][4].
 * nrep receives the replica count; any other pointer may be NULL. */
int mh_pool_status(mh_gcmc *g, int ticket, int *nrep, int *slots, int *remaining, int *err, int32_t *N,
                   double *energy, double *sumN, int64_t *nsteps, int64_t *att, int64_t *acc);
/* Free the ticket's slots and system entry. */
int mh_pool_release(mh_gcmc *g, int ticket);

#ifdef __cplusplus
}
#endif
#endif
