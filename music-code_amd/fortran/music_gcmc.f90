! music_gcmc -- GCMC driver over libmusic_hip (usage: music_gcmc ctrlfile
! [fp32] [chain] [maxn=N] [check]), the drop-in for the reference's
! drivers/music_gcmc.F90:79-257 for rigid species (one, or a mixture of up
! to four: every step draws a species, gcmc.F90:240-242) with the rigid
! insert / delete / translate / rotate moves and the biased insert / delete
! (BINSERT with a bias pmap and temperature, BDELETE; rigidmoves_initInsBias,
! rigidmoves.F90:593-675).  It reads the ctrlfile's
! 'General Information' (iterations, print interval, seed; general.F90),
! atom and molecule types, simulation cell, forcefield and 'GCMC
! Information' sections (gcmc_init, gcmc.F90:100-200; gcmcmoves_init and
! gcmcmoves_getpressures, gcmcmoves.F90:96-260, 430-646), and from the
! species-species file the species-framework maps (NCOUL MAP ...
! Atom@PMAP@file, COUL MAP ... all@EMAP@file; ssmap.F90:217-348), the
! species-species LJ (NCOUL BASIC LJ, atom-atom pairs with LBMIX /
! JORGMIX mixing) and optional Ewald electrostatics (COUL SUM EWALD
! KAPPA@k KMAX@n [RCUT@r]; the reference stops there, ssdriver.F90:489-498).
!
! State points: by default every pressure of the list runs at once, one
! replica per point in one device batch, each an independent chain that
! starts empty with seed iseed + p - 1 (point 1 is the reference's run).
! With 'chain' the points run one after another as in the reference
! (music_gcmc.F90:169-245): the configuration and the random-number stream
! carry over from one point to the next and the statistics restart
! (gcmcmoves_beginsim, gcmcmoves.F90:706-718).
!
! Output follows the reference's: every iprint iterations the acceptance
! counts per move and the number of molecules (instantaneous, block
! average, cumulative average, standard deviation: stats_update,
! stats.F90:152-195, formed from the device's sums of N and N^2), and at the
! end of each point the recomputed and stored energies (kJ/mol), their
! deviation (interact_checknrgs) and the loading per unit cell.  With
! 'check' it prints the parsed parameters and stops before any device call.
!
! Restart files: every icrash iterations the configuration of each state
! point goes to <restart file>.<simno + simstart - 1> in the reference's
! layout (config_writerestartfile, config.F90:2357-2430), and its
! random-number state and full-precision configuration to the same name +
! '.state' (the reference keeps neither).  A 'RESTARTFILE name' source in
! 'Configuration Initialization' (config_readrestartfile, config.F90:
! 2114-2247) starts every state point from that configuration; with
! 'resume' the configuration and random-number state come from name.state
! instead, so a run split at a restart gives the same trajectory as an
! unbroken one.
program music_gcmc
  use, intrinsic :: iso_c_binding
  use music_hip_mod
  use music_inputs
  use music_datafile
  implicit none
  type(sim_system) :: sy
  type(mh_gcmc_params), target :: P
  character(len=LL) :: ctrl, arg, line
  character(len=NL) :: spname, mvname(4)
  integer :: i, j, k, a, niter, iprint, iseed, nmicro, npts, bsize, nspc, isp, nmv, nsite, nst, np
  integer :: maxn, nunit, icrash, simstart, iconfig
  character(len=LL) :: confname
  type(df_contents) :: dfc
  logical :: dfon
  integer, allocatable :: dfu(:)
  logical :: chain, check_only, resume
  character(len=LL) :: desc, resname, srcfile
  real(c_double) :: T, wts(4), kappa_param, rcut
  real(c_double), allocatable :: press(:)
  integer :: stypes(MH_MAX_SITES)
  character(len=LL) :: pmapf(MH_MAX_SITES), emapf(MH_MAX_SITES), biasf
  real(c_double) :: bias_T
  type(c_ptr) :: bmaps(MH_MAX_SPECIES)
  type(c_ptr) :: ctx, g
  type(c_ptr), allocatable :: maps(:)
  character(len=LL), allocatable :: mapnames(:)
  integer :: nmaps
  real(c_double) :: ref(3, MH_MAX_SITES)
  integer(8) :: t0, t1, rate
  real(c_double) :: secs
  ! the species of a mixture; the scalars above (spname, P, press, ...) hold
  ! the species 'cur' being read or reported (use_species / save_species)
  integer :: cur
  type(mh_gcmc_params), target :: PS(MH_MAX_SPECIES)
  character(len=NL) :: spnames(MH_MAX_SPECIES), mvnames(4, MH_MAX_SPECIES)
  integer :: isps(MH_MAX_SPECIES), nmvs(MH_MAX_SPECIES), nsites(MH_MAX_SPECIES)
  real(c_double), allocatable :: pressm(:,:)
  character(len=LL) :: pmapfs(MH_MAX_SITES, MH_MAX_SPECIES), emapfs(MH_MAX_SITES, MH_MAX_SPECIES)
  character(len=LL) :: biasfs(MH_MAX_SPECIES), srcfiles(MH_MAX_SPECIES)
  real(c_double) :: bias_Ts(MH_MAX_SPECIES)
  ! cavity bias: trials and map bias per species, the cavity list
  integer :: cav_ntr(MH_MAX_SPECIES) = 0, cav_mb(MH_MAX_SPECIES) = 0
  type(mh_cavity) :: cav
  integer(c_int8_t), allocatable, target :: cavtab(:)
  character(len=LL) :: cavpmap
  real(c_double) :: cavucut

  call get_command_argument(1, ctrl)
  if (len_trim(ctrl) == 0) then
     write(*, '(a)') 'usage: music_gcmc ctrlfile [fp32] [chain] [resume] [maxn=N] [check]'
     stop 1
  end if
  P%precision = MH_PREC_FP64
  chain = .false.
  check_only = .false.
  resume = .false.
  maxn = 512
  do i = 2, command_argument_count()
     call get_command_argument(i, arg)
     if (trim(lower(arg)) == 'fp32') then
        P%precision = MH_PREC_FP32
     else if (trim(lower(arg)) == 'chain') then
        chain = .true.
     else if (trim(lower(arg)) == 'check') then
        check_only = .true.
     else if (trim(lower(arg)) == 'resume') then
        resume = .true.
     else if (index(lower(arg), 'maxn=') == 1) then
        read(arg(6:), *) maxn
     else
        call fail('unknown argument ' // trim(arg))
     end if
  end do
  call read_system(ctrl, sy)

  ! ---- General Information: description, iterations, print interval,
  ! crash and config intervals, first simulation number, seed
  i = find_section(sy%t, 'General Information')
  call need(i, 'General Information')
  desc = sy%t%line(i)
  desc = desc(1:min(len_trim(desc), 96))  ! general.F90:87 reads it into 2 strLen characters
  read(strip(sy%t%line(i + 1)), *) niter
  read(strip(sy%t%line(i + 2)), *) iprint
  read(strip(sy%t%line(i + 3)), *) icrash
  read(strip(sy%t%line(i + 4)), *) iconfig
  read(strip(sy%t%line(i + 5)), *) simstart
  read(strip(sy%t%line(i + 6)), *) iseed
  resname = word(strip(sy%t%line(i + 8)), 1)
  confname = word(strip(sy%t%line(i + 9)), 1)
  ! ---- Main Datafile Information (datafile_initContents): the contents of
  ! the configuration file written every iconfig iterations
  dfon = .false.
  i = find_section(sy%t, 'Main Datafile Information')
  if (i > 0) then
     call df_parse_contents(sy%t%line(i), dfc, dfon)
     if (.not. dfon) call fail('unknown field in the Main Datafile Information line: ' // trim(sy%t%line(i)))
     if (dfc%time) call fail('datafile: the GCMC driver stores no time (the reference stops on "time")')
     if (dfc%intra_nrg) call fail('datafile: intra_energy is not stored (the reference stops on it)')
  end if

  ! ---- GCMC Information
  i = find_section(sy%t, 'GCMC Information')
  call need(i, 'GCMC Information')
  read(strip(sy%t%line(i)), *) nmicro
  read(strip(sy%t%line(i + 1)), *) T
  if (index(lower(sy%t%line(i + 2)), 'ideal') == 0 .and. index(lower(sy%t%line(i + 2)), 'null') == 0) &
       call fail('only the ideal-gas equation of state is offered (fugacity = pressure)')
  read(strip(sy%t%line(i + 3)), *) npts
  read(strip(sy%t%line(i + 4)), *) bsize
  read(strip(sy%t%line(i + 5)), *) nspc
  if (nspc < 1 .or. nspc > MH_MAX_SPECIES) call fail('1 to 4 GCMC species are offered')
  allocate(press(npts), pressm(npts, nspc))
  i = i + 6
  nst = 0
  srcfile = ''
  do cur = 1, nspc
  do while (i <= sy%t%n)
     line = strip(sy%t%line(i))
     if (len_trim(line) > 0 .and. verify(trim(line), '-') /= 0) exit
     i = i + 1
  end do
  spname = word(strip(sy%t%line(i)), 1)
  isp = mol_index(sy, spname)
  do k = 1, cur - 1
     if (trim(spnames(k)) == trim(spname)) call fail('species ' // trim(spname) // ' listed twice')
  end do
  i = i + 1
  call read_pressures(i)
  i = i + 1          ! sitemap line
  read(strip(sy%t%line(i + 1)), *) nmv
  if (nmv < 1 .or. nmv > 4) call fail('1 to 4 move types are offered')
  read(strip(sy%t%line(i + 2)), *) (wts(k), k = 1, nmv)
  i = i + 3
  P%delta_trans = 0.0d0
  P%delta_rot = 0.0d0
  P%adjust_trans = 0
  P%adjust_rot = 0
  biasf = ''
  bias_T = 0.0d0
  do k = 1, nmv
     mvname(k) = upper(word(strip(sy%t%line(i)), 1))
     i = i + 1
     select case (trim(mvname(k)))
     case ('RINSERT', 'INSERT')
        P%move_type(k) = MH_MV_INSERT
     case ('RDELETE', 'DELETE')
        P%move_type(k) = MH_MV_DELETE
     case ('RTRANSLATE', 'TRANSLATE')
        P%move_type(k) = MH_MV_TRANSLATE
        call read_delta(i, P%delta_trans, P%adjust_trans)
        i = i + 1
     case ('RROTATE', 'ROTATE')
        P%move_type(k) = MH_MV_ROTATE
        call read_delta(i, P%delta_rot, P%adjust_rot)
        i = i + 1
     case ('BINSERT')
        ! rigidmoves_initInsBias (rigidmoves.F90:620-670): a bias pmap file
        ! and the bias temperature (MAP_BIAS); 'CAVITY' and the number of
        ! trials (CAVITY_BIAS); 'file, CAVITY', the temperature and the
        ! number of trials (MAP_CAVITY_BIAS)
        P%move_type(k) = MH_MV_BINSERT
        line = strip(sy%t%line(i))
        if (index(line, ',') == 0 .and. trim(upper(word(line, 1))) == 'CAVITY') then
           read(word(strip(sy%t%line(i + 1)), 1), *) cav_ntr(cur)
           cav_mb(cur) = 0
           i = i + 2
        else if (index(line, ',') > 0) then
           if (trim(upper(word(line, 2))) /= 'CAVITY' .or. nwords(line) /= 2) &
                call fail('BINSERT: put the bias map and CAVITY in that order (rigidmoves.F90:649-666)')
           biasf = word(line, 1)
           read(word(strip(sy%t%line(i + 1)), 1), *) bias_T
           read(word(strip(sy%t%line(i + 2)), 1), *) cav_ntr(cur)
           cav_mb(cur) = 1
           i = i + 3
        else
           biasf = word(line, 1)
           read(word(strip(sy%t%line(i + 1)), 1), *) bias_T
           i = i + 2
        end if
        if (index(line, ',') > 0 .or. trim(upper(word(line, 1))) == 'CAVITY') then
           if (cav_ntr(cur) < 1 .or. cav_ntr(cur) > 1024) call fail('BINSERT: 1 to 1024 cavity trials are offered')
        end if
     case ('BDELETE')
        P%move_type(k) = MH_MV_BDELETE
     case default
        call fail('move type ' // trim(mvname(k)) // ' is not offered (rigid moves only)')
     end select
     P%move_weight(k) = wts(k)
  end do
  P%nmoves = nmv
  P%temperature = T
  P%max_nrg = 1.0d10
  if (any(P%move_type(1:nmv) == MH_MV_BINSERT) .neqv. any(P%move_type(1:nmv) == MH_MV_BDELETE)) &
       call fail('BINSERT and BDELETE come as a pair (moves.F90:827-840)')

  ! ---- the species: sites, body-frame coordinates, types (one table over
  ! the site types of every species), charges
  associate (m => sy%mols(isp))
    nsite = m%natoms
    if (nsite > MH_MAX_SITES) call fail('at most 8 sites per molecule')
    call principal_axes(m, sy%atoms, ref(:, 1:nsite))
    P%nsites = nsite
    do a = 1, nsite
       P%ref(:, a) = ref(:, a)
       P%site_q(a) = m%q(a)
       k = 0
       do j = 1, nst
          if (stypes(j) == m%atype(a)) k = j
       end do
       if (k == 0) then
          if (nst == MH_MAX_SITES) call fail('at most 8 site types over the species')
          nst = nst + 1
          stypes(nst) = m%atype(a)
          k = nst
       end if
       P%site_type(a) = k - 1
    end do
  end associate
  do a = 1, MH_MAX_SITES
     P%site_pmap(a) = c_null_ptr
     P%site_emap(a) = c_null_ptr
  end do
  pmapf = ''
  emapf = ''
  call save_species(cur)
  end do
  ! pair tables and Coulomb / Ewald settings (species 1 carries them for
  ! the mixture), species-framework maps per species
  PS(1)%ntypes = nst
  PS(1)%ewald = 0
  PS(1)%coul = 0
  PS(1)%lj_on = 0_c_int8_t
  PS(1)%coul_on = 0_c_int8_t
  PS(1)%lj_A = 0.0d0
  PS(1)%lj_B = 0.0d0
  PS(1)%lj_hicut2 = 0.0d0
  PS(1)%lj_locut2 = 0.0d0
  call read_species_file()
  nunit = sy%nxyz(1) * sy%nxyz(2) * sy%nxyz(3)
  do cur = 1, nspc
     call use_species(cur)
     call read_config_source()
     call save_species(cur)
  end do
  call use_species(1)
  call read_cavity_section()

  write(*, '(a)') 'Welcome to MuSiC (GCMC implementation, libmusic_hip)'
  write(*, '(2a)') 'Reading from control file ', trim(ctrl)
  write(*, '(a, i0, a, i0, a, i0, a, i0, a, i0)') '  iterations ', niter, ' x ', nmicro, ', print every ', iprint, &
       ', block ', bsize, ', seed ', iseed
  write(*, '(a, 3f10.4, a, 3i3)') '  simulation cell (A):', sy%full, ', unit cells', sy%nxyz
  do cur = 1, nspc
  call use_species(cur)
  write(*, '(a, i0, a, f9.3, a)') '  species ' // trim(spname) // ': ', nsite, ' sites, T = ', T, ' K'
  do k = 1, npts
     write(*, '(a, i0, a, es24.16)') '  state point ', k, ': fugacity (kPa) ', press(k)
  end do
  do k = 1, nmv
     write(*, '(a, i0, 2a, f8.4)') '  move ', k, ': ', trim(mvname(k)) // ' weight', P%move_weight(k)
  end do
  write(*, '(a, 2f10.5)') '  delta translate / rotate:', P%delta_trans, P%delta_rot
  do a = 1, nsite
     write(*, '(a, i0, a, 3es24.16, a, i0, a, f8.4, 4a)') '  site ', a, ':', P%ref(:, a), ' type ', &
          P%site_type(a), ' q ', P%site_q(a), ' pmap ', trim(basename(pmapf(a))), ' emap ', trim(basename(emapf(a)))
  end do
  if (len_trim(biasf) > 0) write(*, '(a, a, a, f10.3)') '  bias: ', trim(biasf), ' T ', bias_T
  if (cav_ntr(cur) > 0) write(*, '(a, i0, a, i0)') '  cavity bias: trials ', cav_ntr(cur), ' map ', cav_mb(cur)
  if (len_trim(srcfile) > 0) write(*, '(2a)') '  start configuration: ', trim(srcfile)
  end do
  if (cav%species_mask /= 0) write(*, '(a, 3(1x, i0), a, i0, 3a, es12.4)') '  cavity list: cubes', cav%n, &
       ', species mask ', cav%species_mask, ', tabulation ', trim(cavpmap), ' below ', cavucut
  call use_species(1)
  do k = 1, nst * nst
     if (P%lj_on(k) /= 0) write(*, '(a, i0, a, 4es24.16)') '  lj ', k - 1, ':', P%lj_A(k), P%lj_B(k), &
          P%lj_hicut2(k), P%lj_locut2(k)
  end do
  if (P%ewald /= 0) write(*, '(a, es24.16, a, i0, a, es24.16)') '  ewald: kappa ', P%ewald_kappa, ' kmax ', &
       P%ewald_kmax, ' rcut ', P%ewald_rcut
  if (P%coul /= 0) write(*, '(a, es24.16, a, es24.16, a, es24.16)') '  plain coulomb: hicut ', P%coul_hicut, &
       ' locut ', P%coul_locut, ' shield ', P%coul_shield
  do k = 1, nspc
     do j = k, nspc
        if (PS(k)%lj_comcut(j) > 0.0d0 .or. PS(k)%coul_comcut(j) > 0.0d0) &
             write(*, '(a, 2(1x, a), a, 2es24.16)') '  centre cutoff', trim(spnames(k)), trim(spnames(j)), &
             ': lj / coul', PS(k)%lj_comcut(j), PS(k)%coul_comcut(j)
     end do
  end do
  if (dfon) then
     if (nspc > 1) call fail('datafile: configuration files are written for one GCMC species')
     write(*, '(a, i0, 2a)') '  configuration files every ', iconfig, ' iterations: ', trim(confname)
     if (dfc%pair_nrg .and. (sy%nmol > 1 .or. P%ewald /= 0 .or. P%coul /= 0 .or. any(P%site_q(1:nsite) /= 0.0d0))) &
          call fail('datafile: pair_energy is stored for a single neutral species without framework')
  end if
  write(*, '(a, i0, 2a)') '  restart files every ', icrash, ' iterations: ', trim(resname)
  flush(6)
  if (check_only) go to 900

  ! ---- device: the simulation cell (and framework, for reference), maps
  call check(mh_ctx_create(0_c_int, ctx), 'mh_ctx_create')
  if (sy%ntot > 0) then
     call check(mh_framework_upload(ctx, int(sy%ntot, c_int), sy%xyz, sy%typ, sy%q, sy%fund, sy%full), &
          'mh_framework_upload')
  else
     ! a NULL cell: one placeholder atom defines the cell; no map reads it
     sy%xyz(:, 1) = 0.0d0
     sy%typ(1) = 0
     sy%q(1) = 0.0d0
     call check(mh_framework_upload(ctx, 1_c_int, sy%xyz, sy%typ, sy%q, sy%fund, sy%full), &
          'mh_framework_upload')
  end if
  ! a slanted cell: maps fold into it, insertion follows its lattice vectors
  ! (the library refuses sorbate Ewald sums and biased insertion there)
  if (sy%slanted) call check(mh_framework_set_cell(ctx, sy%angles, sy%eff), 'mh_framework_set_cell')
  allocate(maps(3 * MH_MAX_SITES * MH_MAX_SPECIES), mapnames(3 * MH_MAX_SITES * MH_MAX_SPECIES))
  nmaps = 0
  do cur = 1, nspc
     call use_species(cur)
     do a = 1, nsite
        if (len_trim(pmapf(a)) > 0) P%site_pmap(a) = load_map(pmapf(a))
        if (len_trim(emapf(a)) > 0) P%site_emap(a) = load_map(emapf(a))
     end do
     bmaps(cur) = c_null_ptr
     if (len_trim(biasf) > 0) then
        call check(mh_bmap_create(ctx, load_map(biasf), bias_T, bmaps(cur)), 'mh_bmap_create ' // trim(biasf))
        P%bias = bmaps(cur)
     end if
     call save_species(cur)
  end do
  call use_species(1)
  if (cav%species_mask /= 0 .and. trim(upper(cavpmap)) /= 'NULL') then
     allocate(cavtab(cav%n(1) * cav%n(2) * cav%n(3)))
     call check(mh_cavity_tab(load_map(cavpmap), cavucut, cav%n, sy%nxyz, sy%full, cavtab), &
          'mh_cavity_tab ' // trim(cavpmap))
     cav%tab = c_loc(cavtab)
  end if
  np = 1
  call system_clock(t0, rate)
  if (chain) then
     do np = 1, npts
        call run_points(np, np, np > 1)
     end do
  else
     call run_points(1, npts, .false.)
  end if
  call system_clock(t1)
  secs = real(t1 - t0, c_double) / real(rate, c_double)
  write(*, '(a, f10.3, a, es10.3, a)') 'Main Program   Time: ', secs, ' sec   (', &
       real(npts, c_double) * niter * nmicro / max(secs, 1.0d-9), ' steps/s over all state points)'
  do cur = 1, nspc
     if (c_associated(bmaps(cur))) call check(mh_bmap_destroy(bmaps(cur)), 'mh_bmap_destroy')
  end do
  do k = 1, nmaps
     call check(mh_map_destroy(maps(k)), 'mh_map_destroy')
  end do
  call check(mh_ctx_destroy(ctx), 'mh_ctx_destroy')
900 continue

contains

  ! state points p0..p1 as replicas of one batch; carry: take the previous
  ! point's configuration and random-number state (held in g) first
  subroutine run_points(q0, q1, carry)
    integer, intent(in) :: q0, q1
    logical, intent(in) :: carry
    integer :: nr, r, e, nev, it, n, q
    integer(c_int32_t), allocatable :: seeds(:), Ncur(:), ix(:), iy(:)
    real(c_double), allocatable :: energy(:), sN(:), sN2(:), snap(:,:), snap2(:,:), etot(:), fug(:)
    integer(c_int64_t), allocatable :: nsteps(:), att(:,:), acc(:,:)
    integer(8), allocatable :: ev(:), blk(:)
    integer(8) :: s, sprev, total
    real(c_double), allocatable :: com(:,:), ang(:,:), rr(:,:)
    integer(c_int) :: ncarry
    type(c_ptr) :: gnew
    nr = q1 - q0 + 1
    allocate(seeds(nr), Ncur(nr), ix(nr), iy(nr), energy(nr), sN(nr), sN2(nr), etot(nr), nsteps(nr))
    allocate(att(4, nr), acc(4, nr), fug(nr * nspc))
    do r = 1, nr
       seeds(r) = int(iseed + (q0 + r - 1) - 1, c_int32_t)
       do q = 1, nspc
          fug((r - 1) * nspc + q) = pressm(q0 + r - 1, q)
       end do
    end do
    if (chain) seeds(1) = int(iseed, c_int32_t)
    call check(mh_gcmc_create_mix(ctx, PS, int(nspc, c_int), int(nr, c_int), fug, seeds, int(maxn, c_int), gnew), &
         'mh_gcmc_create')
    if (cav%species_mask /= 0) call check(mh_gcmc_set_cavity(gnew, cav), 'mh_gcmc_set_cavity')
    if (.not. carry) then
       do q = 1, nspc
          call use_species(q)
          call check(mh_gcmc_species(gnew, int(q - 1, c_int)), 'mh_gcmc_species')
          if (len_trim(srcfile) > 0) call start_from_restart(gnew, nr, q)
       end do
    else
       ! every species' configuration, then the replica's random-number state
       do q = 1, nspc
          call use_species(q)
          allocate(com(3, maxn), ang(3, maxn), rr(3, maxn * nsite))
          call check(mh_gcmc_species(g, int(q - 1, c_int)), 'mh_gcmc_species')
          call check(mh_gcmc_species(gnew, int(q - 1, c_int)), 'mh_gcmc_species')
          call check(mh_gcmc_get_config(g, 0_c_int, ncarry, com, ang, rr), 'mh_gcmc_get_config')
          call check(mh_gcmc_set_config(gnew, 0_c_int, ncarry, com, ang), 'mh_gcmc_set_config')
          deallocate(com, ang, rr)
       end do
       call check(mh_gcmc_get_state(g, Ncur, energy, sN, nsteps, att, acc, ix, iy), 'mh_gcmc_get_state')
       call check(mh_gcmc_set_rng(gnew, 0_c_int, ix(1), iy(1)), 'mh_gcmc_set_rng')
       call check(mh_gcmc_destroy(g), 'mh_gcmc_destroy')
    end if
    call use_species(1)
    g = gnew

    do r = 1, nr
       write(*, '(a)') '------------------------------------------------------------'
       write(*, '(a, i0, a, i0, 3x, a, i0, a)') 'Beginning GCMC simulation number ', q0 + r - 1, ' of ', npts, &
            '(', niter, ' steps)'
       write(*, '(a)') '------------------------------------------------------------'
       do q = 1, nspc
          write(*, '(a)') '    Species Name: ' // trim(spnames(q))
          write(*, '(a)') '      Press.(kPa)   Fugacity'
          write(*, '(4x, 2f11.3)') pressm(q0 + r - 1, q), pressm(q0 + r - 1, q)
       end do
    end do
    ! events: every print step and the start of its statistics block
    ! (a block holds the last ((s - 1) mod bsize) + 1 steps)
    nev = 0
    allocate(ev(2 * (niter / max(iprint, 1) + 2) + niter / max(icrash, 1) + niter / max(iconfig, 1) + 2), &
         blk(niter / max(iprint, 1) + 2))
    n = 0
    do it = iprint, niter, iprint
       s = int(it, 8) * nmicro
       n = n + 1
       blk(n) = s - (mod(s - 1, int(bsize, 8)) + 1)
       nev = nev + 1
       ev(nev) = blk(n)
       nev = nev + 1
       ev(nev) = s
    end do
    ! restart files at every icrash iterations (music_gcmc.F90:231-233)
    if (icrash > 0) then
       do it = icrash, niter, icrash
          nev = nev + 1
          ev(nev) = int(it, 8) * nmicro
       end do
    end if
    ! configuration files at every iconfig iterations (music_gcmc.F90:226-228)
    if (dfon .and. iconfig > 0) then
       do it = iconfig, niter, iconfig
          nev = nev + 1
          ev(nev) = int(it, 8) * nmicro
       end do
       allocate(dfu(nr))
       do r = 1, nr
          dfu(r) = open_datafile(q0 + r - 1)
       end do
    end if
    total = int(niter, 8) * nmicro
    allocate(snap(nr, nspc))
    sprev = 0
    it = 0
    e = 1
    do while (.true.)
       ! next event at or after sprev
       s = total
       do k = 1, nev
          if (ev(k) > sprev .and. ev(k) < s) s = ev(k)
       end do
       if (s > sprev) call run_steps(s - sprev)
       sprev = s
       do k = 1, n
          if (int(k * iprint, 8) * nmicro == s) then
             if (blk(k) == 0) snap = 0.0d0
             call show_stats(k * iprint, s, s - blk(k), snap)
          end if
       end do
       if (dfon .and. iconfig > 0 .and. s > 0) then
          if (mod(s, int(iconfig, 8) * nmicro) == 0) then
             do r = 1, nr
                call write_datafile_config(dfu(r), r - 1)
             end do
          end if
       end if
       if (icrash > 0 .and. s > 0) then
          if (mod(s, int(icrash, 8) * nmicro) == 0) then
             do r = 1, nr
                call write_restart(restart_name(q0 + r - 1), s / nmicro, r - 1)
             end do
             call use_species(1)
          end if
       end if
       ! a block start (after the print: the next block may start at this
       ! print step): keep sum N for the block average
       do k = 1, n
          if (blk(k) == s .and. s > 0) then
             do q = 1, nspc
                call check(mh_gcmc_species(g, int(q - 1, c_int)), 'mh_gcmc_species')
                call check(mh_gcmc_get_moments(g, sN, sN2), 'mh_gcmc_get_moments')
                snap(:, q) = sN
             end do
          end if
       end do
       if (s >= total) exit
    end do
    ! end of the points: recomputed against stored energies
    call check(mh_gcmc_total_energy(g, etot, 0_c_int), 'mh_gcmc_total_energy')
    allocate(snap2(nr, nspc))
    do q = 1, nspc
       call check(mh_gcmc_species(g, int(q - 1, c_int)), 'mh_gcmc_species')
       call check(mh_gcmc_get_state(g, Ncur, energy, sN, nsteps, att, acc, ix, iy), 'mh_gcmc_get_state')
       snap2(:, q) = sN
    end do
    do r = 1, nr
       write(*, '(a)') ' ________ ACTUAL ENERGIES    _____________'
       write(*, '(a, f16.3)') '    Energy: ', etot(r) * 4.184d0
       write(*, '(a)') ' ________ STORED ENERGIES    _____________'
       write(*, '(a, f16.3)') '    Energy: ', energy(r) * 4.184d0
       write(*, '(a)') '------------------------------------------------------------'
       write(*, '(a, i0, 4x, a, es10.3)') 'Ending GCMC simulation number ', q0 + r - 1, 'Nrg dev check: ', &
            abs(etot(r) - energy(r)) * 4.184d0
       do q = 1, nspc
          write(*, '(2x, a, a, f10.1, 4x, a, f8.2)') trim(spnames(q)), ':  P (kPa):', pressm(q0 + r - 1, q), &
               'loading (molec/uc):', snap2(r, q) / real(max(nsteps(r), 1_c_int64_t), c_double) / nunit
       end do
       write(*, '(a)') '------------------------------------------------------------'
    end do
    if (allocated(dfu)) then
       do r = 1, nr
          close(dfu(r))
       end do
       deallocate(dfu)
    end if
    if (.not. chain .or. q1 == npts) call check(mh_gcmc_destroy(g), 'mh_gcmc_destroy')
  end subroutine run_points

  ! the configuration file of one simulation (datafile_initout as called by
  ! music_gcmc.F90:171-182): name.<simno + simstart - 1>
  integer function open_datafile(simno) result(u)
    integer, intent(in) :: simno
    character(len=16) :: num
    character(len=NL), allocatable :: names(:)
    logical, allocatable :: fixed(:)
    integer :: k
    write(num, '(i0)') simno + simstart - 1
    allocate(names(sy%nmol), fixed(sy%nmol))
    do k = 1, sy%nmol
       names(k) = sy%mols(k)%name
       fixed(k) = k == sy%ifw
    end do
    u = df_open(trim(confname) // '.' // trim(num), ctrl, simno, df_siminfo(ctrl), dfc, press(simno), T, 'GCMC', &
         names, fixed, niter, iconfig)
  end function open_datafile

  ! one configuration of replica rep: principal site coordinates, the total
  ! energy (kJ/mol) and, for a single neutral species, its LJ pair energy
  subroutine write_datafile_config(u, rep)
    integer, intent(in) :: u, rep
    integer(c_int) :: nmol
    real(c_double), allocatable :: com(:,:), ang(:,:), rr(:,:), rp(:,:), pair(:,:), en(:), sN(:)
    integer(c_int32_t), allocatable :: Nc(:), ix(:), iy(:)
    integer(c_int64_t), allocatable :: nst(:), att(:,:), acc(:,:)
    logical, allocatable :: fixed(:)
    integer :: m, k, npair
    allocate(com(3, maxn), ang(3, maxn), rr(3, maxn * nsite), rp(3, maxn * nsite))
    call check(mh_gcmc_get_config(g, int(rep, c_int), nmol, com, ang, rr), 'mh_gcmc_get_config')
    do m = 1, nmol
       call toxyz(com(:, m), ang(:, m), rp(:, (m - 1) * nsite + 1:m * nsite))
    end do
    allocate(Nc(nrep_now()), ix(nrep_now()), iy(nrep_now()), en(nrep_now()), sN(nrep_now()), nst(nrep_now()))
    allocate(att(4, nrep_now()), acc(4, nrep_now()))
    call check(mh_gcmc_get_state(g, Nc, en, sN, nst, att, acc, ix, iy), 'mh_gcmc_get_state')
    npair = sy%nmol * (sy%nmol + 1) / 2
    allocate(pair(2, npair), fixed(sy%nmol))
    pair = 0.0d0
    pair(1, 1) = en(rep + 1) * 4.184d0
    do k = 1, sy%nmol
       fixed(k) = k == sy%ifw
    end do
    call df_write_config(u, dfc, sy%nmol, fixed, isp, int(nmol), nsite, rp, pair, en(rep + 1) * 4.184d0)
  end subroutine write_datafile_config

  ! nsteps GCMC steps on every replica, in launches of at most 2^30 steps
  subroutine run_steps(ns)
    integer(8), intent(in) :: ns
    integer(8) :: left
    integer(c_int) :: rc
    left = ns
    do while (left > 0)
       rc = mh_gcmc_run(g, int(min(left, 1073741824_8), c_int))
       if (rc == MH_ECAPACITY) call fail('a state point reached maxn molecules; run again with a larger maxn=N')
       call check(rc, 'mh_gcmc_run')
       left = left - min(left, 1073741824_8)
    end do
  end subroutine run_steps

  ! the reference's statistics block (gcmc_displaystats) per replica: per
  ! species its fugacity, the acceptance of its moves and its number of
  ! molecules (gcmcmoves_displaystats)
  subroutine show_stats(it, s, nblk, sblk)
    integer, intent(in) :: it
    integer(8), intent(in) :: s, nblk
    real(c_double), intent(in) :: sblk(:,:)
    integer :: r, nr, k, mt, q
    integer(c_int32_t), allocatable :: Nc(:,:), ix(:), iy(:)
    real(c_double), allocatable :: en(:), sN(:,:), sN2(:,:), dl(:,:,:)
    integer(c_int64_t), allocatable :: nst(:), att(:,:,:), acc(:,:,:)
    real(c_double) :: cum, dev, blkavg, ratio
    character(len=32) :: lab
    nr = size(sblk, 1)
    allocate(Nc(nr, nspc), ix(nr), iy(nr), en(nr), sN(nr, nspc), sN2(nr, nspc), nst(nr))
    allocate(att(4, nr, nspc), acc(4, nr, nspc), dl(2, nr, nspc))
    do q = 1, nspc
       call check(mh_gcmc_species(g, int(q - 1, c_int)), 'mh_gcmc_species')
       call check(mh_gcmc_get_deltas(g, dl(:, :, q)), 'mh_gcmc_get_deltas')
       call check(mh_gcmc_get_state(g, Nc(:, q), en, sN(:, q), nst, att(:, :, q), acc(:, :, q), ix, iy), &
            'mh_gcmc_get_state')
       call check(mh_gcmc_get_moments(g, sN(:, q), sN2(:, q)), 'mh_gcmc_get_moments')
    end do
    do r = 1, nr
       write(*, '(2x, a, i0, a, i0, a, i0)') 'Macro-iteration number ', it, ' (', s, ' micro) of simulation ', &
            merge(np, r, chain)
       write(*, '(2x, a)') '------------------------------------------------------------'
       write(*, '(2x, a)') 'The GCMC Stats:'
       do q = 1, nspc
          call use_species(q)
          write(*, '(4x, a, f10.1, a)') 'Species Name: ' // trim(spname) // '    Pressure:', &
               press(merge(np, r, chain)), ' kPa'
          do k = 1, nmv
             mt = mod(P%move_type(k), 4) + 1  ! BINSERT / BDELETE count as insert / delete
             ratio = 0.0d0
             if (att(mt, r, q) > 0) ratio = real(acc(mt, r, q), c_double) / real(att(mt, r, q), c_double)
             lab = trim(basic_tag(P%move_type(k))) // ' (' // trim(mvname(k)) // ')'
             if (P%move_type(k) == MH_MV_TRANSLATE) then
                write(*, '(6x, a26, a, f5.3, 3x, a, i0, a, i0, a, f6.3)') lab, 'Acc. Ratio: ', ratio, '(', &
                     acc(mt, r, q), ' of ', att(mt, r, q), ')  Max disp: ', dl(1, r, q)
             else if (P%move_type(k) == MH_MV_ROTATE) then
                write(*, '(6x, a26, a, f5.3, 3x, a, i0, a, i0, a, f6.3)') lab, 'Acc. Ratio: ', ratio, '(', &
                     acc(mt, r, q), ' of ', att(mt, r, q), ')  Max disp: ', dl(2, r, q)
             else
                write(*, '(6x, a26, a, f5.3, 3x, a, i0, a, i0, a)') lab, 'Acc. Ratio: ', ratio, '(', &
                     acc(mt, r, q), ' of ', att(mt, r, q), ')'
             end if
          end do
          cum = sN(r, q) / real(s, c_double)
          dev = sN2(r, q) / real(s, c_double) - cum * cum
          blkavg = (sN(r, q) - sblk(r, q)) / real(nblk, c_double)
          write(*, '(6x, a, 4f7.2)') 'Number of Molecules (ins, blk, cum, stdd)', real(Nc(r, q), c_double), &
               blkavg, cum, sqrt(max(dev, 0.0d0))
       end do
       write(*, '(a)') ' '
    end do
    call use_species(1)
  end subroutine show_stats

  function restart_name(simno) result(o)
    integer, intent(in) :: simno
    character(len=LL) :: o
    character(len=16) :: num
    write(num, '(i0)') simno + simstart - 1
    o = trim(resname) // '.' // trim(num)
  end function restart_name

  ! the reference's vector_display(v, "f13.5"): three f13.5 fields, the
  ! first left-adjusted
  function vec13(v) result(o)
    real(c_double), intent(in) :: v(3)
    character(len=64) :: o
    write(o, '(3f13.5)') v
    o = adjustl(o)
  end function vec13

  ! principal (unfolded) site coordinates of a molecule: rigidcoords_toxyz
  ! (rigidcoords.F90:254-276), centre of mass + Euler matrix x body frame
  subroutine toxyz(com, ang, rp)
    real(c_double), intent(in) :: com(3), ang(3)
    real(c_double), intent(out) :: rp(3, nsite)
    real(c_double) :: sp, cp, st, ct, ss, cs, M(3, 3)
    integer :: b
    if (nsite == 1) then
       rp(:, 1) = com
       return
    end if
    sp = sin(ang(1)); cp = cos(ang(1)); st = sin(ang(2)); ct = cos(ang(2)); ss = sin(ang(3)); cs = cos(ang(3))
    M(1, :) = [cs * cp - ct * sp * ss, -ss * cp - ct * sp * cs, st * sp]
    M(2, :) = [cs * sp + ct * cp * ss, -ss * sp + ct * cp * cs, -st * cp]
    M(3, :) = [ss * st, cs * st, ct]
    do b = 1, nsite
       rp(:, b) = com + matmul(M, P%ref(:, b))
    end do
  end subroutine toxyz

  ! config_writerestartfile (config.F90:2357-2430) for replica rep, one
  ! block per GCMC species, and the replica's random-number state beside it
  ! (then per species: molecules, displacements, attempt / acceptance counts)
  subroutine write_restart(path, iter, rep)
    character(len=*), intent(in) :: path
    integer(8), intent(in) :: iter
    integer, intent(in) :: rep
    integer :: u, n, m, b, dof, q
    integer(c_int) :: nmol
    real(c_double), allocatable :: com(:,:), ang(:,:), rr(:,:), rp(:,:), dl(:,:)
    integer(c_int32_t), allocatable :: Nc(:), ix(:), iy(:)
    real(c_double), allocatable :: en(:), sN(:)
    integer(c_int64_t), allocatable :: nst(:), att(:,:), acc(:,:)
    character(len=48) :: nm
    allocate(com(3, maxn), ang(3, maxn), rr(3, maxn * MH_MAX_SITES), rp(3, MH_MAX_SITES))
    allocate(Nc(nrep_now()), ix(nrep_now()), iy(nrep_now()), en(nrep_now()), sN(nrep_now()), nst(nrep_now()))
    allocate(att(4, nrep_now()), acc(4, nrep_now()), dl(2, nrep_now()))
    open(newunit=u, file=trim(path), status='replace', action='write')
    write(u, '(1x, a)') trim(desc)
    write(u, '(i12)') iter
    write(u, '(2x, es22.15e3)') 0.0d0
    do q = 1, nspc
       call use_species(q)
       call check(mh_gcmc_species(g, int(q - 1, c_int)), 'mh_gcmc_species')
       call check(mh_gcmc_get_config(g, int(rep, c_int), nmol, com, ang, rr), 'mh_gcmc_get_config')
       n = nmol
       dof = sy%mols(isp)%dof
       if (dof == 0) dof = merge(3, 6, nsite == 1)
       nm = spname
       write(u, '(2a)') '_MOLECULE_NAME_: ', nm
       write(u, '(i20, a)') n, ' # Nmoles'
       write(u, '(i20, a)') nsite, ' # Natoms'
       write(u, '(a)') ' _Principal Coordinates Only_'
       do m = 1, n
          call toxyz(com(:, m), ang(:, m), rp(:, 1:nsite))
          do b = 1, nsite
             write(u, '(a)') trim(vec13(rp(:, b)))
          end do
       end do
       write(u, '(a)') ' _Velocities_'
       do m = 1, n * nsite
          write(u, '(a)') trim(vec13([0.0d0, 0.0d0, 0.0d0]))
       end do
       write(u, '(a)') ' _Generalized Coordinates_'
       do m = 1, n
          write(u, '(a)') 'FULLY_RIGID'
          write(u, '(a, 3f13.5, i5)') trim(vec13(com(:, m))), ang(2, m), ang(1, m), ang(3, m), dof
       end do
    end do
    close(u)
    open(newunit=u, file=trim(path) // '.state', status='replace', action='write')
    do q = 1, nspc
       call check(mh_gcmc_species(g, int(q - 1, c_int)), 'mh_gcmc_species')
       call check(mh_gcmc_get_config(g, int(rep, c_int), nmol, com, ang, rr), 'mh_gcmc_get_config')
       call check(mh_gcmc_get_state(g, Nc, en, sN, nst, att, acc, ix, iy), 'mh_gcmc_get_state')
       call check(mh_gcmc_get_deltas(g, dl), 'mh_gcmc_get_deltas')
       if (q == 1) write(u, '(2i12)') ix(rep + 1), iy(rep + 1)
       write(u, '(i12)') nmol
       do m = 1, nmol
          write(u, '(6es26.17e3)') com(:, m), ang(:, m)
       end do
       write(u, '(2es26.17e3)') dl(:, rep + 1)
       ! attempt / acceptance counts: adjustable displacements follow their ratio
       write(u, '(8i21)') att(:, rep + 1), acc(:, rep + 1)
    end do
    close(u)
    call use_species(1)
  end subroutine write_restart

  integer function nrep_now()
    nrep_now = merge(1, npts, chain)
  end function nrep_now

  ! 'Configuration Initialization' (config_init, config.F90:220-246): the
  ! source of the GCMC species; GCMC / NULL start empty, RESTARTFILE reads
  ! config_readrestartfile's layout
  subroutine read_config_source()
    integer :: ii
    character(len=LL) :: w1, src
    srcfile = ''
    ii = find_section(sy%t, 'Configuration Initialization')
    if (ii < 0) return
    do while (ii + 1 <= sy%t%n)
       w1 = strip(sy%t%line(ii))
       if (len_trim(w1) == 0 .or. verify(trim(w1), '-') == 0) exit
       if (trim(word(w1, 1)) == trim(spname)) then
          src = strip(sy%t%line(ii + 1))
          select case (trim(upper(word(src, 1))))
          case ('RESTARTFILE')
             srcfile = word(src, 2)
             if (.not. exists(trim(srcfile))) srcfile = trim(sy%base) // '/' // trim(srcfile)
             if (.not. exists(trim(srcfile))) call fail('restart file not found: ' // trim(word(src, 2)))
          case ('GCMC', 'NULL', 'DFLT_SIZE', 'DEFAULT')
          case default
             call fail('configuration source ' // trim(word(src, 1)) // ' is not offered (GCMC or RESTARTFILE)')
          end select
          return
       end if
       ii = ii + 2
    end do
  end subroutine read_config_source

  ! config_readrestartfile (config.F90:2114-2247): the molecules of GCMC
  ! species q (the current species, selected in gb) from their generalized
  ! coordinates (centre of mass, theta, phi, psi) into every replica of batch
  ! gb; with 'resume' the saved random-number state and the species' block
  ! of the state file too
  subroutine start_from_restart(gb, nr, q)
    type(c_ptr), intent(in) :: gb
    integer, intent(in) :: nr, q
    type(text_file) :: rf
    integer :: ii, nm, na, m, r, u, ios, k, mk
    real(c_double), allocatable :: com(:,:), ang(:,:)
    real(c_double) :: v(7), dres(2)
    integer(c_int32_t) :: rx, ry
    integer(c_int64_t) :: cnt(8)
    logical :: have_cnt
    character(len=LL) :: rng
    call read_text(srcfile, rf)
    ii = 1
    do while (ii <= rf%n)
       if (index(rf%line(ii), '_MOLECULE_NAME_:') == 1) then
          if (trim(word(rf%line(ii), 2)) == trim(spname)) exit
       end if
       ii = ii + 1
    end do
    if (ii > rf%n) call fail('species ' // trim(spname) // ' not in the restart file')
    read(strip(rf%line(ii + 1)), *) nm
    read(strip(rf%line(ii + 2)), *) na
    if (na /= nsite) call fail('the restart file has another number of atoms per molecule')
    if (nm > maxn) call fail('the restart file holds more molecules than maxn')
    ! principal coordinates and velocities (3 values per atom, any per line)
    ii = ii + 4 + lines_for(ii + 4, nm * na)
    ii = ii + 1 + lines_for(ii + 1, nm * na)
    ii = ii + 1
    allocate(com(3, max(nm, 1)), ang(3, max(nm, 1)))
    do m = 1, nm
       if (index(upper(rf%line(ii)), 'RIGID') > 0 .or. index(upper(rf%line(ii)), 'FLEXIBLE') > 0) ii = ii + 1
       v = 0.0d0
       read(strip(rf%line(ii)), *, iostat=ios) v(1:6)
       if (ios /= 0) call fail('cannot read generalized coordinates: ' // trim(rf%line(ii)))
       com(:, m) = v(1:3)
       ang(:, m) = [v(5), v(4), v(6)]  ! file order theta, phi, psi
       ii = ii + 1
    end do
    if (resume) then
       rng = trim(srcfile) // '.state'
       if (.not. exists(trim(rng))) call fail('resume: no saved state ' // trim(rng))
       open(newunit=u, file=trim(rng), status='old', action='read')
       read(u, *) rx, ry
       ! the blocks of the species before q, then q's
       do k = 1, q
          read(u, *, iostat=ios) mk
          if (ios /= 0) call fail('resume: the saved state holds fewer species than the ctrlfile')
          if (k == q .and. mk /= nm) call fail('resume: the saved state does not belong to the restart file')
          do m = 1, mk
             if (k == q) then
                read(u, *) com(:, m), ang(:, m)
             else
                read(u, *) v(1:6)
             end if
          end do
          read(u, *, iostat=ios) dres
          if (ios /= 0) dres = [P%delta_trans, P%delta_rot]
          have_cnt = .false.
          if (ios == 0) then
             read(u, *, iostat=ios) cnt
             have_cnt = ios == 0
          end if
          if (k < q .and. .not. have_cnt) call fail('resume: incomplete species block in ' // trim(rng))
       end do
       close(u)
    end if
    do r = 1, nr
       call check(mh_gcmc_set_config(gb, int(r - 1, c_int), int(nm, c_int), com, ang), 'mh_gcmc_set_config')
       if (resume) call check(mh_gcmc_set_rng(gb, int(r - 1, c_int), rx, ry), 'mh_gcmc_set_rng')
       if (resume) call check(mh_gcmc_set_deltas(gb, int(r - 1, c_int), dres(1), dres(2)), 'mh_gcmc_set_deltas')
       if (resume .and. have_cnt) &
            call check(mh_gcmc_set_counts(gb, int(r - 1, c_int), cnt(1:4), cnt(5:8)), 'mh_gcmc_set_counts')
    end do
    write(*, '(a, i0, 2a)') '  started from ', nm, ' molecules of ', trim(srcfile)
  contains
    ! lines holding n 3-vectors from line i0 on
    integer function lines_for(i0, n)
      integer, intent(in) :: i0, n
      integer :: got, l
      got = 0
      l = i0
      do while (got < n)
         got = got + nwords(strip(rf%line(l))) / 3
         l = l + 1
      end do
      lines_for = l - i0
    end function lines_for
  end subroutine start_from_restart

  function basic_tag(mt) result(o)
    integer(c_int), intent(in) :: mt
    character(len=9) :: o
    select case (mt)
    case (MH_MV_INSERT, MH_MV_BINSERT)
       o = 'INSERT'
    case (MH_MV_DELETE, MH_MV_BDELETE)
       o = 'DELETE'
    case (MH_MV_TRANSLATE)
       o = 'TRANSLATE'
    case default
       o = 'ROTATE'
    end select
  end function basic_tag

  ! gcmcmoves_getpressures (gcmcmoves.F90:430-570): 'p' for one point,
  ! 'p1, p2' for a generated list (gcmcmoves_genpressures: geometric steps
  ! when the range spans more than two decades, else linear), or npts
  ! comma-separated values, continued on up to two more lines ending '+'
  subroutine read_pressures(ii)
    integer, intent(inout) :: ii
    integer :: nf, q, nl
    real(c_double) :: pincr
    character(len=LL) :: f(256)
    nf = 0
    nl = 0
    do
       line = strip(sy%t%line(ii))
       call split_commas(line, f, nf)
       nl = nl + 1
       if (trim(adjustl(f(nf))) /= '+' .or. nl == 3) exit
       nf = nf - 1
       ii = ii + 1
    end do
    if (trim(adjustl(f(nf))) == '+') nf = nf - 1
    if (npts == 1 .and. nf > 1) call fail('one state point but more than one pressure given')
    if (nf == 1 .and. npts == 1) then
       read(word(f(1), 1), *, iostat=q) press(1)
       if (q /= 0) call read_pressure_file(word(f(1), 1))
    else if (nf == 1) then
       call read_pressure_file(word(f(1), 1))
    else if (nf == 2) then
       read(word(f(1), 1), *, err=90) press(1)
       read(word(f(2), 1), *, err=90) pincr
       if (abs(log10(pincr / press(1))) > 2.0d0) then
          pincr = exp(log(pincr / press(1)) / (npts - 1.0d0))
          do q = 1, npts - 1
             press(q + 1) = press(1) * pincr**q
          end do
       else
          pincr = (pincr - press(1)) / (npts - 1.0d0)
          do q = 1, npts - 1
             press(q + 1) = press(1) + pincr * q
          end do
       end if
    else if (nf == npts) then
       do q = 1, npts
          read(word(f(q), 1), *, err=90) press(q)
       end do
    else
       call fail('could not parse the pressure line')
    end if
    return
90  call fail('could not parse the pressure line')
  end subroutine read_pressures

  ! gcmcmoves_readpressures (gcmcmoves.F90:575-611): in the named file
  ! (relative to the working directory) the first line holding the species
  ! name, any case, then the number of points, then the pressures (kPa)
  subroutine read_pressure_file(name)
    character(len=*), intent(in) :: name
    type(text_file) :: pf
    integer :: k, j, n, ios
    character(len=:), allocatable :: buf
    if (.not. exists(trim(name))) call fail('pressure file not found: ' // trim(name))
    call read_text(trim(name), pf)
    k = 0
    do j = 1, pf%n
       if (index(lower(pf%line(j)), lower(trim(spname))) > 0) then
          k = j
          exit
       end if
    end do
    if (k == 0 .or. k >= pf%n) call fail('species ' // trim(spname) // ' not found in pressure file ' // trim(name))
    read(pf%line(k + 1), *, iostat=ios) n
    if (ios /= 0 .or. n /= npts) &
         call fail('the number of points in pressure file ' // trim(name) // ' differs from the ctrlfile')
    buf = ''
    do j = k + 2, pf%n
       buf = buf // ' ' // trim(pf%line(j))
    end do
    read(buf, *, iostat=ios) press(1:npts)
    if (ios /= 0) call fail('could not read ' // trim(name) // ': too few pressures')
    write(*, '(3a)') '  pressures read from ', trim(name), ' (gcmcmoves_readpressures)'
  end subroutine read_pressure_file

  ! fields of a line split at commas, appended to f(nf+1:)
  subroutine split_commas(s, f, nf)
    character(len=*), intent(in) :: s
    character(len=LL), intent(inout) :: f(:)
    integer, intent(inout) :: nf
    integer :: st, c
    st = 1
    do c = 1, len_trim(s) + 1
       if (c > len_trim(s) .or. s(min(c, len(s)):min(c, len(s))) == ',') then
          nf = nf + 1
          f(nf) = adjustl(s(st:c - 1))
          st = c + 1
       end if
    end do
  end subroutine split_commas

  ! 'delta, adjust' of a translate / rotate move; adjustable steps are not
  ! offered
  ! displacement and its "adjust delta" flag (0 = fixed, 1 = adjusted
  ! towards a 50 % acceptance after every attempt, rigidmoves.F90:2586-2617)
  subroutine read_delta(ii, d, adj)
    integer, intent(in) :: ii
    real(c_double), intent(out) :: d
    integer(c_int), intent(out) :: adj
    line = strip(sy%t%line(ii))
    read(word(line, 1), *) d
    adj = 0
    if (nwords(line) > 1) read(word(line, 2), *) adj
    if (adj /= 0 .and. adj /= 1) call fail('adjust delta flag must be 0 or 1')
  end subroutine read_delta

  ! species-species file entries of the GCMC species (ssdriver_init,
  ! ssdriver.F90): between two GCMC species (a species with itself, or two of
  ! a mixture) the LJ and Coulomb / Ewald models fill the pair tables over
  ! their site types; between a GCMC species and another molecule, the maps
  ! of its sites.  Missing or OFF lines leave a pair without interaction.
  subroutine read_species_file()
    type(text_file) :: ss
    character(len=LL) :: s, kind, tok
    integer :: l, wv, k1, k2, b, c, s1, s2, ncl
    integer :: ljon(MH_MAX_SPECIES, MH_MAX_SPECIES), ckind(MH_MAX_SPECIES, MH_MAX_SPECIES)
    real(c_double) :: ljcc(MH_MAX_SPECIES, MH_MAX_SPECIES), clcc(MH_MAX_SPECIES, MH_MAX_SPECIES), rcc
    real(c_double) :: ek, er
    integer(c_int) :: kmax
    logical :: ewset
    call read_text(find_file(sy%ssfile, 'NONE', '', sy%base), ss)
    ljon = 0
    ckind = 0
    ljcc = 0.0d0
    clcc = 0.0d0
    ewset = .false.
    call use_species(1)
    do l = 1, ss%n
       s = strip(ss%line(l))
       if (nwords(s) < 3) cycle
       kind = upper(word(s, 3))
       if (trim(kind) /= 'NCOUL' .and. trim(kind) /= 'COUL') cycle
       s1 = species_of(word(s, 1))
       s2 = species_of(word(s, 2))
       if (s1 == 0 .or. s2 == 0) cycle          ! species-framework: maps, below
       if (trim(upper(word(s, 4))) == 'OFF') cycle
       if (trim(kind) == 'NCOUL') then
          if (index(upper(s), 'LJ') == 0) call fail('species-species NCOUL model must be LJ: ' // trim(s))
          ! ssbasic_init options (ssbasic.F90:125-177): FAST is what the
          ! kernel does; COMCUT@r / COMTRUNC@r set the centre-of-mass cutoff;
          ! SMOOTH@ and NEIGHBOR change the energy and are not offered
          do wv = 4, nwords(s)
             tok = upper(word(s, wv))
             if (trim(tok) == 'BASIC' .or. trim(tok) == 'LJ' .or. trim(tok) == 'FAST') cycle
             if (centre_cutoff(tok, rcc)) then
                ljcc(s1, s2) = rcc
                ljcc(s2, s1) = rcc
                cycle
             end if
             call fail('species-species option ' // trim(word(s, wv)) // ' is not offered (BASIC LJ FAST ' // &
                  'COMCUT@ COMTRUNC@): ' // trim(s))
          end do
          ljon(s1, s2) = 1
          ljon(s2, s1) = 1
          do b = 1, nst
             do c = 1, nst
                if (.not. pair_of(s1, s2, b, c)) cycle
                k1 = pair_index(sy, sy%atoms(stypes(b) + 1)%name, sy%atoms(stypes(c) + 1)%name)
                if (k1 == 0) call fail('no atom-atom LJ entry for ' // trim(sy%atoms(stypes(b) + 1)%name) // &
                     ' ' // trim(sy%atoms(stypes(c) + 1)%name))
                k2 = (b - 1) * nst + c
                if (.not. sy%pairs(k1)%ok) cycle
                P%lj_A(k2) = sy%pairs(k1)%A
                P%lj_B(k2) = sy%pairs(k1)%B
                P%lj_hicut2(k2) = sy%pairs(k1)%hicut**2
                P%lj_locut2(k2) = sy%pairs(k1)%locut**2
                P%lj_on(k2) = 1_c_int8_t
             end do
          end do
       else if (trim(upper(word(s, 4))) == 'BASIC' .and. index(upper(s), 'WFCOUL') == 0) then
          ! plain Coulomb from the atom-atom COUL lines (coulomb_snglint);
          ! the ssbasic options other than FAST are refused as for NCOUL
          do wv = 5, nwords(s)
             tok = upper(word(s, wv))
             if (trim(tok) == 'FAST') cycle
             if (centre_cutoff(tok, rcc)) then
                clcc(s1, s2) = rcc
                clcc(s2, s1) = rcc
                cycle
             end if
             call fail('species-species option ' // trim(word(s, wv)) // ' is not offered (BASIC FAST ' // &
                  'COMCUT@ COMTRUNC@): ' // trim(s))
          end do
          ckind(s1, s2) = 1
          ckind(s2, s1) = 1
          call plain_coulomb(s1, s2)
       else
          if (index(upper(s), 'EWALD') == 0) &
               call fail('species-species COUL model must be SUM EWALD, BASIC (plain COUL) or OFF: ' // trim(s))
          kappa_param = 0.0d0
          rcut = 0.0d0
          kmax = 0
          do wv = 4, nwords(s)
             tok = upper(word(s, wv))
             if (index(tok, 'KAPPA@') == 1) read(tok(7:), *) kappa_param
             if (index(tok, 'KMAX@') == 1) read(tok(6:), *) kmax
             if (index(tok, 'RCUT@') == 1) read(tok(6:), *) rcut
          end do
          if (kappa_param <= 0.0d0 .or. kmax <= 0) call fail('Ewald needs KAPPA@ and KMAX@: ' // trim(s))
          ! kappa in 1/A from the reference's kappa parameter (ewald.F90:
          ! kappa = KAPPA / shortest cell edge); without RCUT@ every
          ! minimum-image pair enters the real-space sum, as the ewald
          ! module's default rcut = -1 (ewald.F90:111-114, 812): a cutoff
          ! beyond the longest minimum-image distance
          ek = kappa_param / minval(sy%full)
          if (rcut <= 0.0d0) rcut = 2.0d0 * maxval(sy%full)
          er = rcut
          ! one Ewald sum covers every charged species pair of the mixture
          if (ewset .and. (ek /= P%ewald_kappa .or. kmax /= P%ewald_kmax .or. er /= P%ewald_rcut)) &
               call fail('Ewald: one KAPPA@ / KMAX@ / RCUT@ for all species pairs is offered: ' // trim(s))
          ewset = .true.
          ckind(s1, s2) = 2
          ckind(s2, s1) = 2
          P%ewald = 1
          P%ewald_kappa = ek
          P%ewald_kmax = kmax
          P%ewald_rcut = er
       end if
    end do
    ! the tables are per site-type pair: species pairs with different models
    ! must not share a site-type pair
    do s1 = 1, nspc
       do s2 = 1, nspc
          do b = 1, nst
             do c = 1, nst
                if (.not. pair_of(s1, s2, b, c)) cycle
                k2 = (b - 1) * nst + c
                if ((ljon(s1, s2) == 0 .and. P%lj_on(k2) /= 0) .or. &
                     (ckind(s1, s2) /= 1 .and. P%coul_on(k2) /= 0)) &
                     call fail('species ' // trim(spnames(s1)) // ' and ' // trim(spnames(s2)) // &
                     ' share site types with species pairs of another model')
             end do
          end do
          ! Ewald sums every charge: charged species pairs are all EWALD
          ncl = 0
          if (any(PS(s1)%site_q(1:nsites(s1)) /= 0.0d0) .and. any(PS(s2)%site_q(1:nsites(s2)) /= 0.0d0)) ncl = 1
          if (ewset .and. ncl == 1 .and. ckind(s1, s2) /= 2) &
               call fail('Ewald: every pair of charged species needs the COUL SUM EWALD line (' // &
               trim(spnames(s1)) // ', ' // trim(spnames(s2)) // ')')
       end do
    end do
    call save_species(1)
    do s1 = 1, nspc
       PS(s1)%lj_comcut(1:nspc) = ljcc(s1, 1:nspc)
       PS(s1)%coul_comcut(1:nspc) = clcc(s1, 1:nspc)
    end do
    ! species-framework maps
    do cur = 1, nspc
       call use_species(cur)
       do l = 1, ss%n
          s = strip(ss%line(l))
          if (nwords(s) < 3) cycle
          kind = upper(word(s, 3))
          if (trim(kind) /= 'NCOUL' .and. trim(kind) /= 'COUL') cycle
          s1 = species_of(word(s, 1))
          s2 = species_of(word(s, 2))
          if (s1 /= 0 .and. s2 /= 0) cycle
          if (s1 /= cur .and. s2 /= cur) cycle
          if (trim(upper(word(s, 4))) == 'OFF') cycle
          if (index(upper(s), 'MAP') == 0) &
               call fail('species-framework interactions come from maps (PMAP@ / EMAP@): ' // trim(s))
          do wv = 4, nwords(s)
             tok = word(s, wv)
             k1 = index(upper(tok), '@PMAP@')
             k2 = index(upper(tok), '@EMAP@')
             if (k1 > 0) call assign_map(tok(1:k1 - 1), tok(k1 + 6:), pmapf)
             if (k2 > 0) call assign_map(tok(1:k2 - 1), tok(k2 + 6:), emapf)
          end do
       end do
       call save_species(cur)
    end do
    call use_species(1)
  end subroutine read_species_file

  ! ssbasic_init's centre-of-mass cutoff options (ssbasic.F90:139-160):
  ! COMCUT@r (r at most half the simulation cell's smallest width, else the
  ! reference stops) and COMTRUNC@r (no such check)
  logical function centre_cutoff(tok, r)
    character(len=*), intent(in) :: tok
    real(c_double), intent(out) :: r
    integer :: at, ios
    centre_cutoff = .false.
    r = 0.0d0
    at = index(tok, '@')
    if (at == 0) return
    if (tok(1:at - 1) /= 'COMCUT' .and. tok(1:at - 1) /= 'COMTRUNC') return
    read(tok(at + 1:), *, iostat=ios) r
    if (ios /= 0 .or. r <= 0.0d0) call fail('centre-of-mass cutoff needs a radius > 0: ' // trim(tok))
    if (tok(1:at - 1) == 'COMCUT' .and. r > cell_min_width() / 2.0d0) &
         call fail('mol-mol cut-off ' // trim(tok) // ' must be less than half the minimum width of the ' // &
         'simulation cell (ssbasic.F90:153-160)')
    centre_cutoff = .true.
  end function centre_cutoff

  ! fundcell_genwidths (fundcell.F90:229-247) for the simulation cell (the
  ! fundamental cell's lattice vectors times the unit cells per axis): each
  ! lattice vector projected on the normal of the other two
  real(c_double) function cell_min_width()
    real(c_double) :: l(3, 3), cp(3), w
    integer :: i, j, k
    do i = 1, 3
       l(:, i) = sy%lvec(:, i) * real(sy%nxyz(i), c_double)
       if (.not. sy%slanted .or. all(sy%lvec(:, i) == 0.0d0)) then
          l(:, i) = 0.0d0
          l(i, i) = sy%full(i)
       end if
    end do
    cell_min_width = huge(1.0d0)
    do i = 1, 3
       j = mod(i, 3) + 1
       k = mod(i + 1, 3) + 1
       cp = [l(2, j) * l(3, k) - l(3, j) * l(2, k), l(3, j) * l(1, k) - l(1, j) * l(3, k), &
             l(1, j) * l(2, k) - l(2, j) * l(1, k)]
       w = dot_product(l(:, i), cp) / sqrt(dot_product(cp, cp))
       cell_min_width = min(cell_min_width, w)
    end do
  end function cell_min_width

  ! index (1..nspc) of a GCMC species by name, 0 for another molecule
  integer function species_of(nm)
    character(len=*), intent(in) :: nm
    integer :: q
    species_of = 0
    do q = 1, nspc
       if (trim(spnames(q)) == trim(nm)) species_of = q
    end do
  end function species_of

  ! site-type pair (b, c) of the table joins a site of species s1 with one of
  ! species s2
  logical function pair_of(s1, s2, b, c)
    integer, intent(in) :: s1, s2, b, c
    pair_of = any(PS(s1)%site_type(1:nsites(s1)) == b - 1) .and. any(PS(s2)%site_type(1:nsites(s2)) == c - 1)
    pair_of = pair_of .or. (any(PS(s2)%site_type(1:nsites(s2)) == b - 1) .and. &
         any(PS(s1)%site_type(1:nsites(s1)) == c - 1))
  end function pair_of

  ! plain Coulomb between the site types of species s1 and s2: per type pair
  ! the atom_atom COUL line (on unless OFF); one hicut and shield constant for
  ! all pairs (the kernel's form), locut LOW_DIST_CUTOFF
  subroutine plain_coulomb(s1, s2)
    integer, intent(in) :: s1, s2
    integer :: b, c
    logical :: found, off
    real(c_double) :: hc, sh
    character(len=LL) :: aa
    aa = find_file(sy%aafile, 'NONE', '', sy%base)
    P%coul_locut = LOW_DIST_CUTOFF
    do b = 1, nst
       do c = 1, nst
          if (.not. pair_of(s1, s2, b, c)) cycle
          call read_coul_pair(aa, sy%atoms(stypes(b) + 1)%name, sy%atoms(stypes(c) + 1)%name, found, off, hc, sh)
          if (.not. found) call fail('no atom-atom COUL entry for ' // trim(sy%atoms(stypes(b) + 1)%name) // &
               ' ' // trim(sy%atoms(stypes(c) + 1)%name))
          if (off) cycle
          P%coul_on((b - 1) * nst + c) = 1_c_int8_t
          if (P%coul == 0) then
             P%coul_hicut = hc
             P%coul_shield = sh
             P%coul = 1
          else if (hc /= P%coul_hicut .or. sh /= P%coul_shield) then
             call fail('plain Coulomb: one HICUT and SHIELDCONST for all site pairs is offered')
          end if
       end do
    end do
  end subroutine plain_coulomb

  ! the scalars of the species being read or reported (spname, isp, nsite,
  ! move names, pressures, model, map and bias files, start configuration)
  ! from / into the per-species arrays
  ! the cavity list of cavity-biased insertion: cavitylist_init /
  ! cavitylist_readCtrlFile (cavitylist.F90:155-275, 1022-1082): the number
  ! of sorbates, their names, cubes per axis, the pmap of the tabulation
  ! (NULL: every cube) and, for a pmap, the energy cutoff
  subroutine read_cavity_section()
    integer :: ii, ncs, k2, q, hit
    character(len=NL) :: nm
    if (all(cav_ntr(1:nspc) == 0)) return
    if (sy%slanted) call fail('cavity bias needs an orthorhombic simulation cell')
    ii = find_section(sy%t, 'Cavity Bias Info')
    if (ii < 0) call fail('cavity bias needs the "Cavity Bias Info" section (cavitylist.F90:155-275)')
    read(word(strip(sy%t%line(ii)), 1), *) ncs
    if (ncs < 1) call fail('Cavity Bias Info: no sorbate listed')
    cav%species_mask = 0
    do k2 = 1, ncs
       nm = word(strip(sy%t%line(ii + k2)), 1)
       hit = 0
       do q = 1, nspc
          if (trim(spnames(q)) == trim(nm)) hit = q
       end do
       if (hit == 0) call fail('Cavity Bias Info: ' // trim(nm) // ' is not a species of the simulation')
       if (btest(cav%species_mask, hit - 1)) call fail('Cavity Bias Info: ' // trim(nm) // ' listed twice')
       cav%species_mask = ibset(cav%species_mask, hit - 1)
    end do
    read(strip(sy%t%line(ii + ncs + 1)), *) cav%n
    do q = 1, 3
       if (cav%n(q) < 1 .or. cav%n(q) > 4096) call fail('Cavity Bias Info: 1 to 4096 cubes per axis')
       if (mod(cav%n(q), sy%nxyz(q)) /= 0) &
            call fail('Cavity Bias Info: cubes per axis must be a multiple of the unit cells (cavitylist.F90:233-241)')
    end do
    cavpmap = word(strip(sy%t%line(ii + ncs + 2)), 1)
    cavucut = 0.0d0
    if (trim(upper(cavpmap)) /= 'NULL') read(word(strip(sy%t%line(ii + ncs + 3)), 1), *) cavucut
    do q = 1, nspc
       cav%ntrials(q) = int(cav_ntr(q), c_int)
       cav%map_bias(q) = int(cav_mb(q), c_int)
       if (cav_ntr(q) > 0 .and. .not. btest(cav%species_mask, q - 1)) &
            call fail('Cavity Bias Info: a cavity-biased species must be listed (cavitylist.F90:363-368)')
    end do
  end subroutine read_cavity_section

  subroutine use_species(q)
    integer, intent(in) :: q
    spname = spnames(q)
    isp = isps(q)
    nsite = nsites(q)
    nmv = nmvs(q)
    mvname = mvnames(:, q)
    press = pressm(:, q)
    P = PS(q)
    pmapf = pmapfs(:, q)
    emapf = emapfs(:, q)
    biasf = biasfs(q)
    bias_T = bias_Ts(q)
    srcfile = srcfiles(q)
  end subroutine use_species

  subroutine save_species(q)
    integer, intent(in) :: q
    spnames(q) = spname
    isps(q) = isp
    nsites(q) = nsite
    nmvs(q) = nmv
    mvnames(:, q) = mvname
    pressm(:, q) = press
    PS(q) = P
    pmapfs(:, q) = pmapf
    emapfs(:, q) = emapf
    biasfs(q) = biasf
    bias_Ts(q) = bias_T
    srcfiles(q) = srcfile
  end subroutine save_species

  ! map file of the sites of one atom type ('all': every site)
  subroutine assign_map(aname, fname, dst)
    character(len=*), intent(in) :: aname, fname
    character(len=LL), intent(inout) :: dst(:)
    integer :: b
    logical :: hit
    hit = .false.
    do b = 1, nsite
       if (trim(lower(aname)) == 'all' .or. trim(sy%atoms(sy%mols(isp)%atype(b) + 1)%name) == trim(aname)) then
          dst(b) = fname
          hit = .true.
       end if
    end do
    if (.not. hit) call fail('map entry for an atom type the species does not have: ' // trim(aname))
  end subroutine assign_map

  ! a map file (current directory, $PMAPDIR, the ctrlfile's directory, or
  ! potentials/pmaps), loaded once
  function load_map(fname) result(m)
    character(len=*), intent(in) :: fname
    type(c_ptr) :: m
    character(len=LL) :: path
    integer :: q
    do q = 1, nmaps
       if (trim(mapnames(q)) == trim(fname)) then
          m = maps(q)
          return
       end if
    end do
    if (exists(trim(fname))) then
       path = fname
    else
       path = find_file(fname, 'PMAPDIR', 'pmaps', sy%base)
    end if
    call check(mh_map_load(ctx, trim(path) // c_null_char, sy%fund, m), 'mh_map_load ' // trim(path))
    nmaps = nmaps + 1
    maps(nmaps) = m
    mapnames(nmaps) = fname
  end function load_map

  logical function exists(pth)
    character(len=*), intent(in) :: pth
    inquire(file=trim(pth), exist=exists)
  end function exists

  function basename(pth) result(o)
    character(len=*), intent(in) :: pth
    character(len=LL) :: o
    integer :: q
    q = index(trim(pth), '/', back=.true.)
    o = pth(q + 1:)
    if (len_trim(o) == 0) o = '-'
  end function basename

  subroutine check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    if (rc /= MH_OK) then
       write(*, '(a, a, i3, a, a)') what, ' failed, code ', rc, ': ', mh_error_message()
       stop 1
    end if
  end subroutine check

end program music_gcmc
