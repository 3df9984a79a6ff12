! map_from_list -- minimal Fortran driver of the map build through the
! iso_c_binding layer (music_hip_mod): reads a plain-text framework and LJ
! table, builds one LJ map on the device (mapmaker_makemap,
! mapmaker.F90:422-598) and writes it in the reference's map-file layout.
!
! usage: map_from_list <input> <mapfile>
! input (free format):
!   natoms ntypes
!   Lx Ly Lz  Mx My Mz            (fundamental cell, simulation cell)
!   gridspace hipotcut_kJ precision(0 = FP64, 1 = FP32)
!   natoms lines: x y z type(0-based)
!   ntypes lines: A B hicut2 locut2
! prints: npnts(3), nfsize, counts(4)
program map_from_list
  use, intrinsic :: iso_c_binding
  use music_hip_mod
  implicit none
  character(len=512) :: inpath, mappath
  integer :: natoms, ntypes, i, u, prec
  real(c_double) :: L(3), Lfull(3), gridspace, hipot, step(3)
  real(c_double), allocatable :: xyz(:,:), q(:), A(:), B(:), h2(:), l2(:)
  integer(c_int), allocatable :: typ(:)
  integer(c_int8_t), allocatable :: on(:)
  type(c_ptr) :: ctx, map
  integer(c_int) :: npnts(3)
  integer(c_int64_t) :: nfsize, counts(4)
  character(kind=c_char) :: dummy(1)
  real(c_float) :: zdummy(1)

  call get_command_argument(1, inpath)
  call get_command_argument(2, mappath)
  open(newunit=u, file=trim(inpath), status='old', action='read')
  read(u, *) natoms, ntypes
  read(u, *) L, Lfull
  read(u, *) gridspace, hipot, prec
  allocate(xyz(3, natoms), typ(natoms), q(natoms), A(ntypes), B(ntypes), h2(ntypes), l2(ntypes), on(ntypes))
  do i = 1, natoms
     read(u, *) xyz(:, i), typ(i)
  end do
  do i = 1, ntypes
     read(u, *) A(i), B(i), h2(i), l2(i)
  end do
  close(u)
  q = 0.0_c_double
  on = 1_c_int8_t

  call check(mh_ctx_create(0_c_int, ctx), 'mh_ctx_create')
  call check(mh_framework_upload(ctx, int(natoms, c_int), xyz, typ, q, L, Lfull), 'mh_framework_upload')
  call check(mh_ff_set_lj(ctx, int(ntypes, c_int), A, B, h2, l2, on), 'mh_ff_set_lj')
  call check(mh_map_build(ctx, MH_MAP_LJ, int(prec, c_int), 1.0_c_double, gridspace, hipot, 0_c_int, map), &
       'mh_map_build')
  call check(mh_map_info(map, npnts, nfsize, counts, step), 'mh_map_info')
  call check(mh_map_write(map, trim(mappath)//c_null_char, 0_c_int, dummy, dummy, zdummy, &
       real(hipot, c_float)), 'mh_map_write')
  write(*, '(3i8, i12, 4i12)') npnts, nfsize, counts
  call check(mh_map_destroy(map), 'mh_map_destroy')
  call check(mh_ctx_destroy(ctx), 'mh_ctx_destroy')

contains

  subroutine check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    if (rc /= MH_OK) then
       write(*, '(a, a, i3, a, a)') what, ' failed, code ', rc, ': ', mh_error_message()
       stop 1
    end if
  end subroutine check

end program map_from_list
