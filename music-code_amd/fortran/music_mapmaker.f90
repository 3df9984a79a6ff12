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
  character(len=LL) :: ctrl, base, s, fname, arg
  type(text_file) :: t
  type(atom_type), allocatable :: atoms(:)
  type(molecule), allocatable :: mols(:)
  type(lj_pair), allocatable :: pairs(:)
  type(map_job), allocatable :: jobs(:)
  type(ewald_opts) :: ew
  integer :: i, k, n, nat, nmol, npairs, njobs, ifw, iprobe, ptype, nxyz(3), ax, c, ntot, nt
  character(len=NL) :: fwname, aafile, ssfile
  real(c_double) :: fund(3), full(3), off(3)
  real(c_double), allocatable :: xyz(:,:), q(:), ljA(:), ljB(:), h2(:), l2(:)
  integer(c_int), allocatable :: typ(:)
  integer(c_int8_t), allocatable :: on(:)
  integer(c_int) :: prec, nk, npnts(3)
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
     write(*, '(a)') 'usage: music_mapmaker ctrlfile [fp32]'
     stop 1
  end if
  prec = MH_PREC_FP64
  call get_command_argument(2, arg)
  if (trim(lower(arg)) == 'fp32') prec = MH_PREC_FP32
  base = dirname(ctrl)
  call read_text(ctrl, t)

  ! ---- atom types
  i = find_section(t, 'Atomic Types')
  call need(i, 'Atomic Types')
  read(strip(t%line(i)), *) nat
  allocate(atoms(nat))
  i = i + 1
  do k = 1, nat
     i = next_nonblank(t, i)
     atoms(k)%name = word(strip(t%line(i)), 1)
     s = word(strip(t%line(i + 1)), 1)
     call read_atom(find_file(s, 'ATOMSDIR', 'atoms', base), atoms(k))
     i = i + 2
  end do
  ! ---- molecule types
  i = find_section(t, 'Molecule Types')
  call need(i, 'Molecule Types')
  read(strip(t%line(i)), *) nmol
  allocate(mols(nmol))
  i = i + 1
  do k = 1, nmol
     i = next_nonblank(t, i)
     mols(k)%name = word(strip(t%line(i)), 1)
     s = word(strip(t%line(i + 1)), 1)
     call read_molecule(find_file(s, 'MOLSDIR', 'molecules', base), atoms, mols(k))
     i = i + 2
  end do
  ! ---- simulation cell (fundamental cell from a molecule's Fundcell_Info)
  i = find_section(t, 'Simulation Cell Information')
  call need(i, 'Simulation Cell Information')
  fwname = word(strip(t%line(i)), 1)
  read(strip(t%line(i + 1)), *) nxyz
  ifw = mol_index(fwname)
  if (.not. mols(ifw)%has_fundcell) call fail('fundamental cell molecule has no Fundcell_Info')
  if (any(abs(mols(ifw)%angles - 90.0d0) > 1.0d-9)) call fail('only orthorhombic cells are supported')
  fund = mols(ifw)%ell
  full = fund * nxyz
  ! replicate the fixed species: x, then y, then z
  n = mols(ifw)%natoms
  ntot = n * nxyz(1) * nxyz(2) * nxyz(3)
  allocate(xyz(3, ntot), typ(ntot), q(ntot))
  xyz(:, 1:n) = mols(ifw)%xyz
  typ(1:n) = mols(ifw)%atype
  q(1:n) = mols(ifw)%q
  do ax = 1, 3
     do c = 1, nxyz(ax) - 1
        off = 0.0d0
        off(ax) = fund(ax) * c
        do k = 1, n
           xyz(:, c * n + k) = xyz(:, k) + off
           typ(c * n + k) = typ(k)
           q(c * n + k) = q(k)
        end do
     end do
     n = n * nxyz(ax)
  end do
  ! ---- forcefield files
  i = find_section(t, 'Forcefield Information')
  call need(i, 'Forcefield Information')
  aafile = word(strip(t%line(i + 2)), 1)
  ssfile = word(strip(t%line(i + 3)), 1)
  call read_atom_atom(find_file(aafile, 'NONE', '', base), pairs, npairs)
  ! ---- map jobs
  i = find_section(t, 'Mapmaker Information')
  call need(i, 'Mapmaker Information')
  read(strip(t%line(i)), *) njobs
  allocate(jobs(njobs))
  i = i + 1
  do k = 1, njobs
     i = next_nonblank(t, i)
     jobs(k)%mapsorb = word(strip(t%line(i)), 1)
     jobs(k)%probe = word(strip(t%line(i + 1)), 1)
     jobs(k)%inter = word(strip(t%line(i + 2)), 1)
     jobs(k)%model = upper(word(strip(t%line(i + 2)), 2))
     read(word(strip(t%line(i + 3)), 1), *) jobs(k)%gridspace
     read(word(strip(t%line(i + 4)), 1), *) jobs(k)%hipotcut
     jobs(k)%fname = word(strip(t%line(i + 5)), 1)
     i = i + 6
  end do

  call check(mh_ctx_create(0_c_int, ctx), 'mh_ctx_create')
  call check(mh_framework_upload(ctx, int(ntot, c_int), xyz, typ, q, fund, full), 'mh_framework_upload')
  nt = nat
  allocate(ljA(nt), ljB(nt), h2(nt), l2(nt), on(nt))
  do k = 1, njobs
     if (mol_index(jobs(k)%mapsorb) /= ifw) call fail('the mapped species must be the fundamental cell species')
     iprobe = mol_index(jobs(k)%probe)
     if (mols(iprobe)%natoms /= 1) call fail('the probe must be a single atom (config.F90:754-759)')
     ptype = mols(iprobe)%atype(1)
     call system_clock(t0, rate)
     if (trim(jobs(k)%model) == 'LJ') then
        call probe_table(ptype)
        call check(mh_ff_set_lj(ctx, int(nt, c_int), ljA, ljB, h2, l2, on), 'mh_ff_set_lj')
        call check(mh_map_build(ctx, MH_MAP_LJ, prec, 1.0_c_double, jobs(k)%gridspace, jobs(k)%hipotcut, &
             0_c_int, map), 'mh_map_build')
     else if (trim(jobs(k)%model) == 'EWALD') then
        ew = read_ewald(find_file(ssfile, 'NONE', '', base), jobs(k)%probe, jobs(k)%mapsorb)
        if (.not. ew%found) call fail('no EWALD entry for the probe and map species in the spc_spc file')
        call check(mh_ewald_set(ctx, ew%kappa, int(ew%kmax, c_int), ew%locut, nk), 'mh_ewald_set')
        call check(mh_map_build(ctx, MH_MAP_EWALD, prec, mols(iprobe)%q(1), jobs(k)%gridspace, &
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

  integer function mol_index(name)
    character(len=*), intent(in) :: name
    integer :: j
    do j = 1, nmol
       if (trim(mols(j)%name) == trim(name)) then
          mol_index = j
          return
       end if
    end do
    call fail('unknown species ' // trim(name))
    mol_index = -1
  end function mol_index

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
       j = pair_index(atoms(a)%name, atoms(pt + 1)%name)
       if (j == 0) cycle
       if (.not. pairs(j)%ok) cycle
       ljA(a) = pairs(j)%A
       ljB(a) = pairs(j)%B
       h2(a) = pairs(j)%hicut**2
       l2(a) = pairs(j)%locut**2
       on(a) = 1_c_int8_t
    end do
  end subroutine probe_table

  integer function pair_index(n1, n2)
    character(len=*), intent(in) :: n1, n2
    integer :: j
    pair_index = 0
    do j = 1, npairs
       if ((trim(pairs(j)%a1) == trim(n1) .and. trim(pairs(j)%a2) == trim(n2)) .or. &
           (trim(pairs(j)%a1) == trim(n2) .and. trim(pairs(j)%a2) == trim(n1))) pair_index = j
    end do
  end function pair_index

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
       do j = 1, mols(ifw)%natoms
          if (mols(ifw)%atype(j) == a) seen = .true.
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
       do j = 1, min(25, len_trim(atoms(a)%name))
          znames(25 * (r - 1) + j) = atoms(a)%name(j:j)
       end do
       do j = 1, min(2, len_trim(atoms(a)%symbol))
          zsyms(2 * (r - 1) + j) = atoms(a)%symbol(j:j)
       end do
       if (r <= mols(ifw)%natoms) zpar(1, r) = real(mols(ifw)%q(r), c_float)
       if (lj) then
          j = pair_index(atoms(a)%name, atoms(pt + 1)%name)
          if (j > 0) then
             if (pairs(j)%ok) then
                zpar(2, r) = real(pairs(j)%A, c_float)
                zpar(3, r) = real(pairs(j)%B, c_float)
                zpar(6, r) = real(pairs(j)%hicut, c_float)
                zpar(7, r) = real(pairs(j)%locut, c_float)
             end if
          end if
       end if
    end do
  end subroutine header_atoms

  subroutine need(i, tag)
    integer, intent(in) :: i
    character(len=*), intent(in) :: tag
    if (i < 0) call fail('ctrlfile section not found: ' // tag)
  end subroutine need

  subroutine fail(msg)
    character(len=*), intent(in) :: msg
    write(*, '(a)') trim(msg)
    stop 1
  end subroutine fail

  subroutine check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    if (rc /= MH_OK) then
       write(*, '(a, a, i3, a, a)') what, ' failed, code ', rc, ': ', mh_error_message()
       stop 1
    end if
  end subroutine check

end program music_mapmaker
