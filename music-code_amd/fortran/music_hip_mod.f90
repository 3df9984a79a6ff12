! music_hip_mod -- Fortran 2003 iso_c_binding interface to libmusic_hip.so
! (include/music_hip.h).  This is the layer the reference's drivers
! (music_mapmaker.F90, music_gcmc.F90) call in place of mapmaker_makemap
! (mapmaker.F90:422-598), maps_int (maps.F90:87-110) and gcmc_dosim
! (gcmc.F90:224-258).
!
! Conventions: arrays are passed as contiguous Fortran arrays whose memory
! order matches the C layout (xyz(3, natoms) in Fortran is xyz[natoms][3] in
! C); indices passed to the library are 0-based; every function returns an
! integer(c_int) status (MH_OK = 0) and mh_error_message() returns the
! library's last error text.
module music_hip_mod
  use, intrinsic :: iso_c_binding
  implicit none
  private

  integer(c_int), parameter, public :: MH_OK = 0, MH_EINVAL = 1, MH_ENOMEM = 2, MH_EHIP = 3, &
       MH_EIO = 4, MH_ESTATE = 5, MH_ECAPACITY = 6
  integer(c_int), parameter, public :: MH_MAP_LJ = 1, MH_MAP_EWALD = 2, MH_MAP_BUCK = 3
  integer(c_int), parameter, public :: MH_PREC_FP64 = 0, MH_PREC_FP32 = 1
  integer(c_int), parameter, public :: MH_EWALD_PARITY = 0, MH_EWALD_PHYSICS = 1
  integer(c_int), parameter, public :: MH_MV_INSERT = 0, MH_MV_DELETE = 1, MH_MV_TRANSLATE = 2, &
       MH_MV_ROTATE = 3, MH_MV_BINSERT = 4, MH_MV_BDELETE = 5
  integer, parameter, public :: MH_MAX_SITES = 8, MH_MAX_SPECIES = 4

  ! mh_gcmc_params (field order and types as in the C struct)
  type, bind(C), public :: mh_gcmc_params
     integer(c_int) :: nsites
     real(c_double) :: ref(3, MH_MAX_SITES)
     integer(c_int) :: site_type(MH_MAX_SITES)
     real(c_double) :: site_q(MH_MAX_SITES)
     type(c_ptr) :: site_pmap(MH_MAX_SITES)
     type(c_ptr) :: site_emap(MH_MAX_SITES)
     integer(c_int) :: ntypes
     real(c_double) :: lj_A(64), lj_B(64), lj_hicut2(64), lj_locut2(64)
     integer(c_int8_t) :: lj_on(64)
     integer(c_int) :: ewald
     real(c_double) :: ewald_kappa
     integer(c_int) :: ewald_kmax
     real(c_double) :: ewald_rcut
     real(c_double) :: temperature
     integer(c_int) :: nmoves
     integer(c_int) :: move_type(4)
     real(c_double) :: move_weight(4)
     real(c_double) :: delta_trans, delta_rot
     real(c_double) :: max_nrg
     integer(c_int) :: precision
     type(c_ptr) :: bias = c_null_ptr
     integer(c_int) :: coul = 0
     integer(c_int8_t) :: coul_on(64) = 0_c_int8_t
     real(c_double) :: coul_hicut = 0.0d0, coul_locut = 1.0d-3, coul_shield = 1.0d0
     integer(c_int) :: adjust_trans = 0, adjust_rot = 0
     ! centre-of-mass cutoffs (A) of the LJ / plain-Coulomb pair terms with
     ! each partner species (COMCUT@r), 0 = none
     real(c_double) :: lj_comcut(MH_MAX_SPECIES) = 0.0d0, coul_comcut(MH_MAX_SPECIES) = 0.0d0
  end type mh_gcmc_params

  ! cavity-biased insertion / deletion (mh_gcmc_set_cavity): cube grid,
  ! tabulation (uint8 [n1][n2][n3], z fastest, c_null_ptr = every cube),
  ! species occupying cubes (bit s), trials and map bias per species
  type, bind(C), public :: mh_cavity
     integer(c_int) :: n(3) = 1
     type(c_ptr) :: tab = c_null_ptr
     integer(c_int) :: species_mask = 0
     integer(c_int) :: ntrials(MH_MAX_SPECIES) = 0
     integer(c_int) :: map_bias(MH_MAX_SPECIES) = 0
  end type mh_cavity

  public :: mh_error_message
  public :: mh_ctx_create, mh_ctx_destroy, mh_framework_upload, mh_framework_set_cell, mh_ff_set_lj, mh_ff_set_buck
  public :: mh_ewald_set, mh_map_build, mh_map_destroy, mh_map_info, mh_map_get, mh_map_write
  public :: mh_map_load, mh_map_interp
  public :: mh_gcmc_create, mh_gcmc_create_mix, mh_gcmc_species, mh_gcmc_destroy, mh_gcmc_set_threads, &
       mh_gcmc_autotune
  public :: mh_gcmc_set_config, mh_gcmc_run, mh_gcmc_run_cycles, mh_gcmc_get_state, mh_gcmc_set_rng, &
            mh_gcmc_get_moments, mh_gcmc_get_deltas, mh_gcmc_set_deltas, mh_gcmc_set_counts
  public :: mh_gcmc_get_config, mh_gcmc_total_energy, mh_trial_energy
  public :: mh_bmap_create, mh_bmap_destroy, mh_bmap_info
  public :: mh_gcmc_set_cavity, mh_cavity_tab

  interface
     function c_mh_last_error() bind(C, name="mh_last_error") result(p)
       import :: c_ptr
       type(c_ptr) :: p
     end function c_mh_last_error

     function c_strlen(s) bind(C, name="strlen") result(n)
       import :: c_ptr, c_size_t
       type(c_ptr), value :: s
       integer(c_size_t) :: n
     end function c_strlen

     function mh_ctx_create(device, ctx) bind(C, name="mh_ctx_create") result(rc)
       import :: c_int, c_ptr
       integer(c_int), value :: device
       type(c_ptr), intent(out) :: ctx
       integer(c_int) :: rc
     end function mh_ctx_create

     function mh_ctx_destroy(ctx) bind(C, name="mh_ctx_destroy") result(rc)
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int) :: rc
     end function mh_ctx_destroy

     function mh_framework_upload(ctx, natoms, xyz, types, q, fund_ell, full_ell) &
          bind(C, name="mh_framework_upload") result(rc)
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: natoms
       real(c_double), intent(in) :: xyz(3, *), q(*), fund_ell(3), full_ell(3)
       integer(c_int), intent(in) :: types(*)
       integer(c_int) :: rc
     end function mh_framework_upload

     function mh_framework_set_cell(ctx, angles_deg, eff) bind(C, name="mh_framework_set_cell") result(rc)
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: angles_deg(3), eff(3)
       integer(c_int) :: rc
     end function mh_framework_set_cell

     function mh_ff_set_lj(ctx, ntypes, A, B, hicut2, locut2, on) bind(C, name="mh_ff_set_lj") result(rc)
       import :: c_int, c_ptr, c_double, c_int8_t
       type(c_ptr), value :: ctx
       integer(c_int), value :: ntypes
       real(c_double), intent(in) :: A(*), B(*), hicut2(*), locut2(*)
       integer(c_int8_t), intent(in) :: on(*)
       integer(c_int) :: rc
     end function mh_ff_set_lj

     function mh_ff_set_buck(ctx, ntypes, A, rho, C, D, hicut2, locut2) &
          bind(C, name="mh_ff_set_buck") result(rc)
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: ntypes
       real(c_double), intent(in) :: A(*), rho(*), C(*), D(*), hicut2(*), locut2(*)
       integer(c_int) :: rc
     end function mh_ff_set_buck

     function mh_ewald_set(ctx, kappa_param, kmax, locut, sign_mode, nk) bind(C, name="mh_ewald_set") result(rc)
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), value :: kappa_param, locut
       integer(c_int), value :: kmax, sign_mode
       integer(c_int), intent(out) :: nk
       integer(c_int) :: rc
     end function mh_ewald_set

     function mh_map_build(ctx, kind, precision, probe_q, gridspace, hipotcut_kJ, keep_full, map) &
          bind(C, name="mh_map_build") result(rc)
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: kind, precision, keep_full
       real(c_double), value :: probe_q, gridspace, hipotcut_kJ
       type(c_ptr), intent(out) :: map
       integer(c_int) :: rc
     end function mh_map_build

     function mh_map_destroy(map) bind(C, name="mh_map_destroy") result(rc)
       import :: c_int, c_ptr
       type(c_ptr), value :: map
       integer(c_int) :: rc
     end function mh_map_destroy

     function mh_map_info(map, npnts, nfsize, counts, step) bind(C, name="mh_map_info") result(rc)
       import :: c_int, c_ptr, c_double, c_int64_t
       type(c_ptr), value :: map
       integer(c_int), intent(out) :: npnts(3)
       integer(c_int64_t), intent(out) :: nfsize, counts(4)
       real(c_double), intent(out) :: step(3)
       integer(c_int) :: rc
     end function mh_map_info

     function mh_map_get(map, ptr, vals8) bind(C, name="mh_map_get") result(rc)
       import :: c_int, c_ptr, c_int32_t, c_float
       type(c_ptr), value :: map
       integer(c_int32_t), intent(out) :: ptr(*)
       real(c_float), intent(out) :: vals8(*)
       integer(c_int) :: rc
     end function mh_map_get

     function mh_map_write(map, path, nzatoms, names25, symbols2, zparams7, pcut_kJ) &
          bind(C, name="mh_map_write") result(rc)
       import :: c_int, c_ptr, c_char, c_float
       type(c_ptr), value :: map
       character(kind=c_char), intent(in) :: path(*), names25(*), symbols2(*)
       integer(c_int), value :: nzatoms
       real(c_float), intent(in) :: zparams7(*)
       real(c_float), value :: pcut_kJ
       integer(c_int) :: rc
     end function mh_map_write

     function mh_map_load(ctx, path, fund_ell, map) bind(C, name="mh_map_load") result(rc)
       import :: c_int, c_ptr, c_char, c_double
       type(c_ptr), value :: ctx
       character(kind=c_char), intent(in) :: path(*)
       real(c_double), intent(in) :: fund_ell(3)
       type(c_ptr), intent(out) :: map
       integer(c_int) :: rc
     end function mh_map_load

     function mh_map_interp(ctx, map, n, xyz, factor, e, ok) bind(C, name="mh_map_interp") result(rc)
       import :: c_int, c_ptr, c_double, c_int8_t
       type(c_ptr), value :: ctx, map
       integer(c_int), value :: n
       real(c_double), intent(in) :: xyz(3, *)
       real(c_double), value :: factor
       real(c_double), intent(out) :: e(*)
       integer(c_int8_t), intent(out) :: ok(*)
       integer(c_int) :: rc
     end function mh_map_interp

     function mh_gcmc_create(ctx, params, nrep, fugacity_kPa, seeds, maxN, g) &
          bind(C, name="mh_gcmc_create") result(rc)
       import :: c_int, c_ptr, c_double, c_int32_t, mh_gcmc_params
       type(c_ptr), value :: ctx
       type(mh_gcmc_params), intent(in) :: params
       integer(c_int), value :: nrep, maxN
       real(c_double), intent(in) :: fugacity_kPa(*)
       integer(c_int32_t), intent(in) :: seeds(*)
       type(c_ptr), intent(out) :: g
       integer(c_int) :: rc
     end function mh_gcmc_create

     function mh_gcmc_create_mix(ctx, params, nspc, nrep, fugacity_kPa, seeds, maxN, g) &
          bind(C, name="mh_gcmc_create_mix") result(rc)
       import :: c_int, c_ptr, c_double, c_int32_t, mh_gcmc_params
       type(c_ptr), value :: ctx
       type(mh_gcmc_params), intent(in) :: params(*)
       integer(c_int), value :: nspc, nrep, maxN
       real(c_double), intent(in) :: fugacity_kPa(*)
       integer(c_int32_t), intent(in) :: seeds(*)
       type(c_ptr), intent(out) :: g
       integer(c_int) :: rc
     end function mh_gcmc_create_mix

     function mh_gcmc_species(g, spc) bind(C, name="mh_gcmc_species") result(rc)
       import :: c_int, c_ptr
       type(c_ptr), value :: g
       integer(c_int), value :: spc
       integer(c_int) :: rc
     end function mh_gcmc_species

     function mh_gcmc_destroy(g) bind(C, name="mh_gcmc_destroy") result(rc)
       import :: c_int, c_ptr
       type(c_ptr), value :: g
       integer(c_int) :: rc
     end function mh_gcmc_destroy

     function mh_gcmc_set_threads(g, threads) bind(C, name="mh_gcmc_set_threads") result(rc)
       import :: c_int, c_ptr
       type(c_ptr), value :: g
       integer(c_int), value :: threads
       integer(c_int) :: rc
     end function mh_gcmc_set_threads

     function mh_gcmc_autotune(g, probe_steps, threads, exy) bind(C, name="mh_gcmc_autotune") result(rc)
       import :: c_int, c_ptr
       type(c_ptr), value :: g
       integer(c_int), value :: probe_steps
       integer(c_int), intent(out) :: threads, exy
       integer(c_int) :: rc
     end function mh_gcmc_autotune

     function mh_gcmc_set_config(g, rep, n, com, ang) bind(C, name="mh_gcmc_set_config") result(rc)
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: g
       integer(c_int), value :: rep, n
       real(c_double), intent(in) :: com(3, *), ang(3, *)
       integer(c_int) :: rc
     end function mh_gcmc_set_config

     function mh_gcmc_run(g, nsteps) bind(C, name="mh_gcmc_run") result(rc)
       import :: c_int, c_ptr
       type(c_ptr), value :: g
       integer(c_int), value :: nsteps
       integer(c_int) :: rc
     end function mh_gcmc_run

     function mh_gcmc_run_cycles(g, nsteps, ncycles) bind(C, name="mh_gcmc_run_cycles") result(rc)
       import :: c_int, c_ptr
       type(c_ptr), value :: g
       integer(c_int), value :: nsteps, ncycles
       integer(c_int) :: rc
     end function mh_gcmc_run_cycles

     function mh_gcmc_get_state(g, N, energy, sumN, nsteps, att, acc, ix, iy) &
          bind(C, name="mh_gcmc_get_state") result(rc)
       import :: c_int, c_ptr, c_double, c_int32_t, c_int64_t
       type(c_ptr), value :: g
       integer(c_int32_t), intent(out) :: N(*), ix(*), iy(*)
       real(c_double), intent(out) :: energy(*), sumN(*)
       integer(c_int64_t), intent(out) :: nsteps(*), att(4, *), acc(4, *)
       integer(c_int) :: rc
     end function mh_gcmc_get_state

     function mh_gcmc_get_moments(g, sumN, sumN2) bind(C, name="mh_gcmc_get_moments") result(rc)
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: g
       real(c_double), intent(out) :: sumN(*), sumN2(*)
       integer(c_int) :: rc
     end function mh_gcmc_get_moments

     function mh_gcmc_set_rng(g, rep, ix, iy) bind(C, name="mh_gcmc_set_rng") result(rc)
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: g
       integer(c_int), value :: rep
       integer(c_int32_t), value :: ix, iy
       integer(c_int) :: rc
     end function mh_gcmc_set_rng

     function mh_gcmc_get_deltas(g, delta) bind(C, name="mh_gcmc_get_deltas") result(rc)
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: g
       real(c_double) :: delta(*)
       integer(c_int) :: rc
     end function mh_gcmc_get_deltas

     function mh_gcmc_set_deltas(g, rep, dtrans, drot) bind(C, name="mh_gcmc_set_deltas") result(rc)
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: g
       integer(c_int), value :: rep
       real(c_double), value :: dtrans, drot
       integer(c_int) :: rc
     end function mh_gcmc_set_deltas

     function mh_gcmc_set_counts(g, rep, att, acc) bind(C, name="mh_gcmc_set_counts") result(rc)
       import :: c_int, c_ptr, c_int64_t
       type(c_ptr), value :: g
       integer(c_int), value :: rep
       integer(c_int64_t), intent(in) :: att(4), acc(4)
       integer(c_int) :: rc
     end function mh_gcmc_set_counts

     function mh_gcmc_get_config(g, rep, n, com, ang, r) bind(C, name="mh_gcmc_get_config") result(rc)
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: g
       integer(c_int), value :: rep
       integer(c_int), intent(out) :: n
       real(c_double), intent(out) :: com(3, *), ang(3, *), r(3, *)
       integer(c_int) :: rc
     end function mh_gcmc_get_config

     function mh_gcmc_total_energy(g, e, refresh) bind(C, name="mh_gcmc_total_energy") result(rc)
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: g
       real(c_double), intent(out) :: e(*)
       integer(c_int), value :: refresh
       integer(c_int) :: rc
     end function mh_gcmc_total_energy

     function mh_gcmc_set_cavity(g, c) bind(C, name="mh_gcmc_set_cavity") result(rc)
       import :: c_int, c_ptr, mh_cavity
       type(c_ptr), value :: g
       type(mh_cavity), intent(in) :: c
       integer(c_int) :: rc
     end function mh_gcmc_set_cavity

     function mh_cavity_tab(pmap, ucutoff, n, ncell, full_ell, tab) bind(C, name="mh_cavity_tab") result(rc)
       import :: c_int, c_ptr, c_double, c_int8_t
       type(c_ptr), value :: pmap
       real(c_double), value :: ucutoff
       integer(c_int), intent(in) :: n(3), ncell(3)
       real(c_double), intent(in) :: full_ell(3)
       integer(c_int8_t), intent(out) :: tab(*)
       integer(c_int) :: rc
     end function mh_cavity_tab

     function mh_bmap_create(ctx, pmap, bias_temperature, b) bind(C, name="mh_bmap_create") result(rc)
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx, pmap
       real(c_double), value :: bias_temperature
       type(c_ptr), intent(out) :: b
       integer(c_int) :: rc
     end function mh_bmap_create

     function mh_bmap_destroy(b) bind(C, name="mh_bmap_destroy") result(rc)
       import :: c_int, c_ptr
       type(c_ptr), value :: b
       integer(c_int) :: rc
     end function mh_bmap_destroy

     function mh_bmap_info(b, numtab, ncubelets, cum_wt, clamp_energy) bind(C, name="mh_bmap_info") result(rc)
       import :: c_int, c_ptr, c_double, c_int64_t
       type(c_ptr), value :: b
       integer(c_int64_t), intent(out) :: numtab, ncubelets
       real(c_double), intent(out) :: cum_wt, clamp_energy
       integer(c_int) :: rc
     end function mh_bmap_info

     function mh_trial_energy(g, rep, n, com, ang, e, ok) bind(C, name="mh_trial_energy") result(rc)
       import :: c_int, c_ptr, c_double, c_int8_t
       type(c_ptr), value :: g
       integer(c_int), value :: rep, n
       real(c_double), intent(in) :: com(3, *), ang(3, *)
       real(c_double), intent(out) :: e(*)
       integer(c_int8_t), intent(out) :: ok(*)
       integer(c_int) :: rc
     end function mh_trial_energy
  end interface

contains

  ! text of the library's last error (mh_last_error)
  function mh_error_message() result(msg)
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: s(:)
    integer :: n, i
    p = c_mh_last_error()
    if (.not. c_associated(p)) then
       msg = ''
       return
    end if
    n = int(c_strlen(p))
    call c_f_pointer(p, s, [n])
    allocate(character(len=n) :: msg)
    do i = 1, n
       msg(i:i) = s(i)
    end do
  end function mh_error_message

end module music_hip_mod
