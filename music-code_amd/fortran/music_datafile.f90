! The reference's binary datafile ("config file", datafile.F90) written by
! the drivers: Fortran unformatted sequential records.
!   header      FileHeader_Info (datafile.F90:102-108, datafile_writeheader
!               :596-638): ctrl filename, line count, 200 ctrlfile lines,
!               simulation number, run description
!   contents    ContentDetails (:83-87): six logicals
!   extra info  100 lines of 48 characters (datafile_writeExtraInfo
!               :384-454)
!   per configuration (datafile_writeconfig :909-958, datafile_writenrg
!               :1034-1073): pair energies per species pair (if stored),
!               kinetic and intramolecular energy per species, total
!               energy (if stored), then per moving species N and its
!               principal site coordinates and velocities (if stored)
module music_datafile
  use, intrinsic :: iso_c_binding
  implicit none
  private

  integer, parameter :: STRLEN = 48, LSTRLEN = 96, MAX_CTRL_LINES = 200, MAX_EXTRA = 100, MAX_SORBS = 10

  type :: df_header
     sequence
     character(len=STRLEN) :: ctrl_filename
     integer(c_int32_t) :: nlines
     character(len=2 * LSTRLEN) :: ctrl_line(MAX_CTRL_LINES)
     integer(c_int32_t) :: simno
     character(len=10 * STRLEN) :: siminfo
  end type df_header

  type, public :: df_contents
     logical :: tot_nrg = .false., pair_nrg = .false., intra_nrg = .false.
     logical :: posns = .false., vels = .false., time = .false.
  end type df_contents

  public :: df_parse_contents, df_open, df_write_config, df_siminfo

contains

  ! datafile_initContents (datafile.F90:230-285): the comma-separated line
  ! after the "Main Datafile Information" tag; ok = .false. on an unknown
  ! field
  subroutine df_parse_contents(line, c, ok)
    character(len=*), intent(in) :: line
    type(df_contents), intent(out) :: c
    logical, intent(out) :: ok
    character(len=len(line)) :: s
    character(len=64) :: f
    integer :: k, j
    s = line
    k = index(s, '#')
    if (k > 0) s(k:) = ' '
    ok = .true.
    do
       s = adjustl(s)
       if (len_trim(s) == 0) exit
       j = index(s, ',')
       if (j == 0) then
          f = s
          s = ' '
       else
          f = s(1:j - 1)
          s = s(j + 1:)
       end if
       f = adjustl(f)
       call lower_in_place(f)
       select case (trim(f))
       case ('energy')
          c%tot_nrg = .true.
       case ('intra_energy')
          c%intra_nrg = .true.
       case ('pair_energy')
          c%pair_nrg = .true.
       case ('velocity')
          c%vels = .true.
       case ('position')
          c%posns = .true.
       case ('time')
          c%time = .true.
       case default
          ok = .false.
       end select
    end do
  end subroutine df_parse_contents

  subroutine lower_in_place(s)
    character(len=*), intent(inout) :: s
    integer :: i
    do i = 1, len(s)
       if (s(i:i) >= 'A' .and. s(i:i) <= 'Z') s(i:i) = achar(iachar(s(i:i)) + 32)
    end do
  end subroutine lower_in_place

  ! general_init's run description (general.F90:100-106)
  function df_siminfo(ctrl) result(o)
    character(len=*), intent(in) :: ctrl
    character(len=10 * STRLEN) :: o
    character(len=8) :: d
    character(len=10) :: t
    character(len=5) :: z
    call date_and_time(d, t, z)
    o = 'This Simulation was done in   Year : ' // d(1:4) // ',   Month : ' // d(5:6) // ',   on date : ' // &
         d(7:8) // ',    at Time : ' // t(1:2) // ' hrs ' // t(3:4) // ' minutes. ###  The time specified ' // &
         '(local time) varies from GMT by ' // z(1:3) // ' hrs' // z(4:5) // ' minutes.  The control file used was ' &
         // trim(ctrl) // '. '
  end function df_siminfo

  ! datafile_initout (datafile.F90:148-225): open path, write the header,
  ! the contents and the extra-info lines; returns the unit
  integer function df_open(path, ctrl, simno, siminfo, c, pressure, temperature, simtype, names, fixed, &
                           niters, iconfig) result(u)
    character(len=*), intent(in) :: path, ctrl, siminfo, simtype
    integer, intent(in) :: simno, niters, iconfig
    type(df_contents), intent(in) :: c
    real(c_double), intent(in) :: pressure, temperature
    character(len=*), intent(in) :: names(:)
    logical, intent(in) :: fixed(:)
    type(df_header), allocatable :: h
    character(len=STRLEN) :: x(MAX_EXTRA)
    character(len=LSTRLEN) :: buf
    integer :: cu, ios, i, n
    allocate(h)
    h%ctrl_filename = ctrl
    h%ctrl_line = repeat(achar(0), 2 * LSTRLEN)  ! lines past nlines: zero bytes
    n = 0
    open(newunit=cu, file=trim(ctrl), status='old', action='read', iostat=ios)
    if (ios == 0) then
       do
          read(cu, '(a)', iostat=ios) buf
          if (ios /= 0 .or. n == MAX_CTRL_LINES) exit
          n = n + 1
          h%ctrl_line(n) = buf
       end do
       close(cu)
    end if
    h%nlines = n
    h%simno = simno
    h%siminfo = siminfo
    x = ' '
    write(x(1), '(i5, 5x, a)') simno, ' # simno'
    write(x(2), '(f12.5, 5x, a)') pressure, ' # simPressure'
    write(x(3), '(f12.5, 5x, a)') temperature, ' # simTemperature'
    x(4) = trim(simtype) // '     ' // ' # simType'
    write(x(5), '(i5, 5x, a)') size(names), ' # nspcs'
    do i = 1, MAX_SORBS
       if (i <= size(names)) then
          write(x(5 + i), '(a, i5)') names(i)(1:20) // ' # spc', i
          if (fixed(i)) then
             write(x(5 + MAX_SORBS + i), '(a, 5x, a, i5)') 'FIXED', ' # spc', i
          else
             write(x(5 + MAX_SORBS + i), '(a, 5x, a, i5)') 'NOT_FIXED', ' # spc', i
          end if
       else
          write(x(5 + i), '(a, 5x, a)') 'NOSPECIES', ' # spc'
          write(x(5 + MAX_SORBS + i), '(a, 5x, a, i5)') 'NOSPECIES', ' # spc', i
       end if
    end do
    write(x(6 + 2 * MAX_SORBS), '(i12, 5x, a)') niters, ' # niters'
    write(x(7 + 2 * MAX_SORBS), '(i12, 5x, a)') iconfig, ' # iconfig'
    open(newunit=u, file=trim(path), status='replace', action='write', form='unformatted', access='sequential')
    write(u) h
    write(u) c
    do i = 1, MAX_EXTRA
       write(u) x(i)
    end do
  end function df_open

  ! one configuration of a GCMC replica (datafile_writeconfig without a
  ! time): nspc species of which the moving one is isp with nmol molecules
  ! of nsite sites at principal coordinates rp(3, nsite * nmol); pair(2,
  ! npair) the (noncoulombic, coulombic) energies of the species pairs
  ! i <= j in order, total the total energy (kJ/mol); velocities are zero
  ! (Monte Carlo), kinetic and intramolecular energies zero (rigid species)
  subroutine df_write_config(u, c, nspc, fixed, isp, nmol, nsite, rp, pair, total)
    integer, intent(in) :: u, nspc, isp, nmol, nsite
    type(df_contents), intent(in) :: c
    logical, intent(in) :: fixed(:)
    real(c_double), intent(in) :: rp(:,:), pair(:,:), total
    integer :: i, j, k
    real(c_double), allocatable :: v(:,:)
    k = 0
    do i = 1, nspc
       do j = i, nspc
          k = k + 1
          if (c%pair_nrg) write(u) pair(1, k), pair(2, k)
       end do
       write(u) 0.0d0, 0.0d0
    end do
    if (c%tot_nrg) write(u) total
    do i = 1, nspc
       if (fixed(i)) cycle
       if (i /= isp) then
          write(u) 0_c_int32_t
          cycle
       end if
       write(u) int(nmol, c_int32_t)
       if (nmol == 0) cycle
       if (c%posns) write(u) rp(:, 1:nsite * nmol)
       if (c%vels) then
          allocate(v(3, nsite * nmol))
          v = 0.0d0
          write(u) v
       end if
    end do
  end subroutine df_write_config

end module music_datafile
