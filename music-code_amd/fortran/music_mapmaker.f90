! music_mapmaker -- map maker driver over libmusic_hip (usage:
! music_mapmaker ctrlfile [fp32]), the drop-in for the reference's
! drivers/music_mapmaker.F90:6-153 + mapmaker_makemap (mapmaker.F90:422-598):
! reads the ctrlfile sections 'Atomic Types', 'Molecule Types', 'Simulation
! Cell Information', 'Forcefield Information' and 'Mapmaker Information',
! replicates the fixed species to the simulation cell (simcell.F90:200-216:
! copies along x, then y, then z), builds every requested LJ or Ewald map on
! the device and writes it in the reference's map-file layout under the
! reference's file name (mapmaker_makeName, mapmaker.F90:604-625) in the
! current directory.
program music_mapmaker
  use, intrinsic :: iso_c_binding
  use music_hip_mod
  use music_inputs
  implicit none
  character(len=LL) :: ctrl, fname, arg
  type(sim_system), target :: sy
  type(map_job), allocatable :: jobs(:)
  type(ewald_opts) :: ew
  integer :: i, k, nat, njobs, ifw, iprobe, ptype, nt
  real(c_double), allocatable :: ljA(:), ljB(:), h2(:), l2(:)
  integer(c_int8_t), allocatable :: on(:)
  integer(c_int) :: prec, sign_mode, nk, npnts(3)
  integer(c_int64_t) :: nfsize, counts(4)
  real(c_double) :: step(3)
  type(c_ptr) :: ctx, map
  integer :: nz
  integer, allocatable :: ztypes(:)
  character(kind=c_char), allocatable :: znames(:), zsyms(:)
  real(c_float), allocatable :: zpar(:,:)
  integer(8) :: t0, t1, rate

  call get_command_argument(1, ctrl)
  if (len_trim(ctrl) == 0) then
     write(*, '(a)') 'usage: music_mapmaker ctrlfile [fp32] [physics]'
     stop 1
  end if
  ! fp32: FP32 pair arithmetic with FP64 sums; physics: Ewald maps of the
  ! framework itself (the reference's reciprocal-space signs otherwise)
  prec = MH_PREC_FP64
  sign_mode = MH_EWALD_PARITY
  do i = 2, command_argument_count()
     call get_command_argument(i, arg)
     if (trim(lower(arg)) == 'fp32') prec = MH_PREC_FP32
     if (trim(lower(arg)) == 'physics') sign_mode = MH_EWALD_PHYSICS
  end do
  call read_system(ctrl, sy)
  if (sy%ifw == 0) call fail('the map maker needs a fundamental-cell species')
  nat = sy%nat
  ifw = sy%ifw
  ! ---- map jobs
  i = find_section(sy%t, 'Mapmaker Information')
  call need(i, 'Mapmaker Information')
  read(strip(sy%t%line(i)), *) njobs
  allocate(jobs(njobs))
  i = i + 1
  do k = 1, njobs
     i = next_nonblank(sy%t, i)
     jobs(k)%mapsorb = word(strip(sy%t%line(i)), 1)
     jobs(k)%probe = word(strip(sy%t%line(i + 1)), 1)
     jobs(k)%inter = word(strip(sy%t%line(i + 2)), 1)
     jobs(k)%model = upper(word(strip(sy%t%line(i + 2)), 2))
     read(word(strip(sy%t%line(i + 3)), 1), *) jobs(k)%gridspace
     read(word(strip(sy%t%line(i + 4)), 1), *) jobs(k)%hipotcut
     jobs(k)%fname = word(strip(sy%t%line(i + 5)), 1)
     i = i + 6
  end do

  call check(mh_ctx_create(0_c_int, ctx), 'mh_ctx_create')
  call check(mh_framework_upload(ctx, int(sy%ntot, c_int), sy%xyz, sy%typ, sy%q, sy%fund, sy%full), &
       'mh_framework_upload')
  ! the cell angles and the effective box the grid spans (config_makegrid)
  call check(mh_framework_set_cell(ctx, sy%angles, sy%eff), 'mh_framework_set_cell')
  nt = nat
  allocate(ljA(nt), ljB(nt), h2(nt), l2(nt), on(nt))
  do k = 1, njobs
     if (mol_index(sy, jobs(k)%mapsorb) /= ifw) call fail('the mapped species must be the fundamental cell species')
     iprobe = mol_index(sy, jobs(k)%probe)
     if (sy%mols(iprobe)%natoms /= 1) call fail('the probe must be a single atom (config.F90:754-759)')
     ptype = sy%mols(iprobe)%atype(1)
     call system_clock(t0, rate)
     if (trim(jobs(k)%model) == 'LJ') then
        call probe_table(ptype)
        call check(mh_ff_set_lj(ctx, int(nt, c_int), ljA, ljB, h2, l2, on), 'mh_ff_set_lj')
        call check(mh_map_build(ctx, MH_MAP_LJ, prec, 1.0_c_double, jobs(k)%gridspace, jobs(k)%hipotcut, &
             0_c_int, map), 'mh_map_build')
     else if (trim(jobs(k)%model) == 'EWALD') then
        ew = read_ewald(find_file(sy%ssfile, 'NONE', '', sy%base), jobs(k)%probe, jobs(k)%mapsorb)
        if (.not. ew%found) call fail('no EWALD entry for the probe and map species in the spc_spc file')
        call check(mh_ewald_set(ctx, ew%kappa, int(ew%kmax, c_int), ew%locut, sign_mode, nk), 'mh_ewald_set')
        call check(mh_map_build(ctx, MH_MAP_EWALD, prec, sy%mols(iprobe)%q(1), jobs(k)%gridspace, &
             jobs(k)%hipotcut, 0_c_int, map), 'mh_map_build')
     else
        call fail('unsupported map model ' // trim(jobs(k)%model))
     end if
     call system_clock(t1)
     call check(mh_map_info(map, npnts, nfsize, counts, step), 'mh_map_info')
     fname = jobs(k)%fname
     if (trim(upper(fname)) == 'AUTO') then  ! mapmaker_makeName
        fname = trim(jobs(k)%mapsorb)
        if (trim(upper(jobs(k)%probe)) /= 'PROBE') fname = trim(fname) // '.' // trim(jobs(k)%probe)
        fname = trim(fname) // '.' // trim(jobs(k)%model) // '.map'
     end if
     call header_atoms(trim(jobs(k)%model) == 'LJ', ptype)
     call check(mh_map_write(map, trim(fname) // c_null_char, int(nz, c_int), znames, zsyms, zpar, &
          real(jobs(k)%hipotcut, c_float)), 'mh_map_write')
     write(*, '(a, a, a, 3i6, a, 4i12, a, f10.3, a)') 'map ', trim(fname), ': grid', npnts, ', counts', counts, &
          ', ', real(t1 - t0) / real(rate) * 1.0e3, ' ms'
     call check(mh_map_destroy(map), 'mh_map_destroy')
  end do
  call check(mh_ctx_destroy(ctx), 'mh_ctx_destroy')

contains

  ! LJ parameters of the probe type against every framework atom type
  subroutine probe_table(pt)
    integer, intent(in) :: pt
    integer :: a, j
    ljA = 0.0d0
    ljB = 0.0d0
    h2 = 0.0d0
    l2 = 0.0d0
    on = 0_c_int8_t
    do a = 1, nt
       j = pair_index(sy, sy%atoms(a)%name, sy%atoms(pt + 1)%name)
       if (j == 0) cycle
       if (.not. sy%pairs(j)%ok) cycle
       ljA(a) = sy%pairs(j)%A
       ljB(a) = sy%pairs(j)%B
       h2(a) = sy%pairs(j)%hicut**2
       l2(a) = sy%pairs(j)%locut**2
       on(a) = 1_c_int8_t
    end do
  end subroutine probe_table

  ! map header atom records (mapmaker_hackheader, mapmaker.F90:372-388): one
  ! per atom type of the mapped species, ascending; the charge is that of the
  ! species' i-th atom (the reference's molecules_getcharge(mapSorb, i))
  subroutine header_atoms(lj, pt)
    logical, intent(in) :: lj
    integer, intent(in) :: pt
    integer :: a, j, r
    logical :: seen
    if (allocated(ztypes)) deallocate(ztypes, znames, zsyms, zpar)
    allocate(ztypes(nat))
    nz = 0
    do a = 0, nat - 1
       seen = .false.
       do j = 1, sy%mols(ifw)%natoms
          if (sy%mols(ifw)%atype(j) == a) seen = .true.
       end do
       if (seen) then
          nz = nz + 1
          ztypes(nz) = a
       end if
    end do
    allocate(znames(25 * max(nz, 1)), zsyms(2 * max(nz, 1)), zpar(7, max(nz, 1)))
    znames = ' '
    zsyms = ' '
    zpar = 0.0
    do r = 1, nz
       a = ztypes(r) + 1
       do j = 1, min(25, len_trim(sy%atoms(a)%name))
          znames(25 * (r - 1) + j) = sy%atoms(a)%name(j:j)
       end do
       do j = 1, min(2, len_trim(sy%atoms(a)%symbol))
          zsyms(2 * (r - 1) + j) = sy%atoms(a)%symbol(j:j)
       end do
       if (r <= sy%mols(ifw)%natoms) zpar(1, r) = real(sy%mols(ifw)%q(r), c_float)
       if (lj) then
          j = pair_index(sy, sy%atoms(a)%name, sy%atoms(pt + 1)%name)
          if (j > 0) then
             if (sy%pairs(j)%ok) then
                zpar(2, r) = real(sy%pairs(j)%A, c_float)
                zpar(3, r) = real(sy%pairs(j)%B, c_float)
                zpar(6, r) = real(sy%pairs(j)%hicut, c_float)
                zpar(7, r) = real(sy%pairs(j)%locut, c_float)
             end if
          end if
       end if
    end do
  end subroutine header_atoms

  subroutine check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    if (rc /= MH_OK) then
       write(*, '(a, a, i3, a, a)') what, ' failed, code ', rc, ': ', mh_error_message()
       stop 1
    end if
  end subroutine check

end program music_mapmaker
