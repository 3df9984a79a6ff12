! music_inputs -- readers for the reference's input files used by the map
! maker and GCMC drivers: ctrlfile sections (located by tag string as
! general.F90 / filesrchstr do), atom files (atom.F90), molecule files with
! 'Coord_Info: Listed' coordinates and Fundcell_Info (molecule.F90,
! fundcell.F90), the atom-atom forcefield file (lj_init, lj.F90:88-262) and
! the species-species file (ssdriver.F90:81-253, EWALD options of
! ewald_init, ewald.F90:99-180).  Only the subset the GCMC energy path
! needs: LJ and Ewald pair models, orthorhombic cells.
module music_inputs
  use, intrinsic :: iso_c_binding
  implicit none
  private

  integer, parameter, public :: LL = 512, NL = 32
  real(c_double), parameter, public :: LOW_DIST_CUTOFF = 1.0d-3, HIGH_DIST_CUTOFF = 13.0d0
  real(c_double), parameter, public :: KCALMOLE_KB = 8.314d-3 / 4.184d0

  type, public :: text_file
     character(len=LL), allocatable :: line(:)
     integer :: n = 0
  end type text_file

  type, public :: atom_type
     character(len=NL) :: name = '', symbol = ''
     real(c_double) :: mass = 0.0d0, sscharge = 0.0d0
  end type atom_type

  type, public :: molecule
     character(len=NL) :: name = ''
     integer :: natoms = 0
     character(len=NL), allocatable :: aname(:)
     integer(c_int), allocatable :: atype(:)       ! 0-based atom-type index
     real(c_double), allocatable :: xyz(:,:), q(:)  ! xyz(3, natoms)
     logical :: transform = .true.
     logical :: has_fundcell = .false.
     real(c_double) :: ell(3) = 0.0d0, angles(3) = 90.0d0
  end type molecule

  type, public :: lj_pair
     character(len=NL) :: a1 = '', a2 = ''
     logical :: off = .false., ok = .false.
     real(c_double) :: sig = 0.0d0, eps = 0.0d0, A = 0.0d0, B = 0.0d0
     real(c_double) :: hicut = HIGH_DIST_CUTOFF, locut = LOW_DIST_CUTOFF
  end type lj_pair

  type, public :: ewald_opts
     character(len=NL) :: s1 = '', s2 = ''
     logical :: found = .false.
     integer :: kmax = 0
     real(c_double) :: kappa = 0.0d0, locut = 0.0d0
  end type ewald_opts

  type, public :: map_job
     character(len=NL) :: mapsorb = '', probe = '', inter = '', model = '', fname = ''
     real(c_double) :: gridspace = 0.0d0, hipotcut = 0.0d0
  end type map_job

  public :: read_text, strip, lower, upper, find_section, next_nonblank, word, nwords
  public :: find_file, read_atom, read_molecule, read_atom_atom, read_ewald, lj_calc_AB
  public :: dirname

contains

  subroutine read_text(path, t)
    character(len=*), intent(in) :: path
    type(text_file), intent(out) :: t
    integer :: u, ios, n, i
    character(len=LL) :: buf
    open(newunit=u, file=trim(path), status='old', action='read', iostat=ios)
    if (ios /= 0) then
       write(*, '(a, a)') 'cannot open ', trim(path)
       stop 1
    end if
    n = 0
    do
       read(u, '(a)', iostat=ios) buf
       if (ios /= 0) exit
       n = n + 1
    end do
    allocate(t%line(max(n, 1)))
    t%line = ''
    rewind(u)
    do i = 1, n
       read(u, '(a)', iostat=ios) t%line(i)
    end do
    t%n = n
    close(u)
  end subroutine read_text

  ! the line without its '#' comment, left-adjusted
  function strip(s) result(o)
    character(len=*), intent(in) :: s
    character(len=LL) :: o
    integer :: k
    k = index(s, '#')
    if (k > 0) then
       o = adjustl(s(1:k - 1))
    else
       o = adjustl(s)
    end if
  end function strip

  function lower(s) result(o)
    character(len=*), intent(in) :: s
    character(len=len(s)) :: o
    integer :: i, c
    o = s
    do i = 1, len(s)
       c = iachar(s(i:i))
       if (c >= 65 .and. c <= 90) o(i:i) = achar(c + 32)
    end do
  end function lower

  function upper(s) result(o)
    character(len=*), intent(in) :: s
    character(len=len(s)) :: o
    integer :: i, c
    o = s
    do i = 1, len(s)
       c = iachar(s(i:i))
       if (c >= 97 .and. c <= 122) o(i:i) = achar(c - 32)
    end do
  end function upper

  ! index of the line after the first line containing tag (case-insensitive)
  integer function find_section(t, tag) result(i)
    type(text_file), intent(in) :: t
    character(len=*), intent(in) :: tag
    integer :: j
    i = -1
    do j = 1, t%n
       if (index(lower(t%line(j)), lower(tag)) > 0) then
          i = j + 1
          return
       end if
    end do
  end function find_section

  integer function next_nonblank(t, i0) result(i)
    type(text_file), intent(in) :: t
    integer, intent(in) :: i0
    i = i0
    do while (i <= t%n)
       if (len_trim(strip(t%line(i))) > 0) return
       i = i + 1
    end do
  end function next_nonblank

  logical function is_sep(c)
    character, intent(in) :: c
    is_sep = (c == ' ' .or. c == ',' .or. c == achar(9))
  end function is_sep

  integer function nwords(s)
    character(len=*), intent(in) :: s
    integer :: i
    logical :: inw
    nwords = 0
    inw = .false.
    do i = 1, len_trim(s)
       if (is_sep(s(i:i))) then
          inw = .false.
       else if (.not. inw) then
          inw = .true.
          nwords = nwords + 1
       end if
    end do
  end function nwords

  ! k-th word (separators: blank, comma, tab)
  function word(s, k) result(o)
    character(len=*), intent(in) :: s
    integer, intent(in) :: k
    character(len=LL) :: o
    integer :: i, n, st
    logical :: inw
    o = ''
    n = 0
    inw = .false.
    st = 1
    do i = 1, len_trim(s) + 1
       if (i > len_trim(s)) then
          if (inw .and. n == k) o = s(st:i - 1)
          exit
       end if
       if (is_sep(s(i:i))) then
          if (inw .and. n == k) then
             o = s(st:i - 1)
             return
          end if
          inw = .false.
       else if (.not. inw) then
          inw = .true.
          n = n + 1
          st = i
       end if
    end do
  end function word

  function dirname(p) result(o)
    character(len=*), intent(in) :: p
    character(len=LL) :: o
    integer :: k
    k = index(trim(p), '/', back=.true.)
    if (k > 1) then
       o = p(1:k - 1)
    else if (k == 1) then
       o = '/'
    else
       o = '.'
    end if
  end function dirname

  logical function exists(p)
    character(len=*), intent(in) :: p
    inquire(file=trim(p), exist=exists)
  end function exists

  ! a reference input file: $envvar/name, base/name, then the shared
  ! potentials/<sub>/ tree next to base or up to three levels above it
  function find_file(name, envvar, sub, base) result(o)
    character(len=*), intent(in) :: name, envvar, sub, base
    character(len=LL) :: o, d, p
    integer :: st, k
    o = ''
    call get_environment_variable(trim(envvar), d, status=st)
    if (st == 0 .and. len_trim(d) > 0) then
       if (exists(trim(d) // '/' // trim(name))) then
          o = trim(d) // '/' // trim(name)
          return
       end if
    end if
    if (exists(trim(base) // '/' // trim(name))) then
       o = trim(base) // '/' // trim(name)
       return
    end if
    p = base
    do k = 1, 3
       if (len_trim(sub) > 0) then
          if (exists(trim(p) // '/potentials/' // trim(sub) // '/' // trim(name))) then
             o = trim(p) // '/potentials/' // trim(sub) // '/' // trim(name)
             return
          end if
       end if
       p = dirname(p)
    end do
    write(*, '(a, a)') 'input file not found: ', trim(name)
    stop 1
  end function find_file

  subroutine read_atom(path, a)
    character(len=*), intent(in) :: path
    type(atom_type), intent(inout) :: a
    type(text_file) :: t
    character(len=LL) :: s, key, val
    integer :: i, k
    call read_text(path, t)
    do i = 1, t%n
       s = strip(t%line(i))
       k = index(s, ':')
       if (k == 0) cycle
       key = adjustl(s(1:k - 1))
       val = adjustl(s(k + 1:))
       select case (trim(key))
       case ('Atom_Symbol')
          a%symbol = word(val, 1)
       case ('Atom_Mass')
          read(val, *) a%mass
       case ('Atom_SS_Charge')
          read(val, *) a%sscharge
       end select
    end do
  end subroutine read_atom

  ! 'Coord_Info: Listed Cartesian' molecule (index x y z name [charge] ...)
  subroutine read_molecule(path, atoms, m)
    character(len=*), intent(in) :: path
    type(atom_type), intent(in) :: atoms(:)
    type(molecule), intent(inout) :: m
    type(text_file) :: t
    character(len=LL) :: s
    integer :: i, j, n, k, nw
    call read_text(path, t)
    i = 1
    do while (i <= t%n)
       s = strip(t%line(i))
       if (index(s, 'Coord_Info:') == 1) then
          m%transform = index(lower(s), 'notransform') == 0
          i = i + 1
          read(strip(t%line(i)), *) n
          m%natoms = n
          allocate(m%aname(n), m%atype(n), m%xyz(3, n), m%q(n))
          do j = 1, n
             s = strip(t%line(i + j))
             read(word(s, 2), *) m%xyz(1, j)
             read(word(s, 3), *) m%xyz(2, j)
             read(word(s, 4), *) m%xyz(3, j)
             m%aname(j) = word(s, 5)
             nw = nwords(s)
             m%q(j) = 0.0d0
             if (nw >= 6) read(word(s, 6), *) m%q(j)
             m%atype(j) = -1
             do k = 1, size(atoms)
                if (trim(atoms(k)%name) == trim(m%aname(j))) m%atype(j) = k - 1
             end do
             if (m%atype(j) < 0) then
                write(*, '(a, a)') 'atom type not declared: ', trim(m%aname(j))
                stop 1
             end if
          end do
          i = i + n
       else if (index(s, 'Fundcell_Info:') == 1) then
          m%has_fundcell = .true.
          read(strip(t%line(i + 1)), *) m%ell
          read(strip(t%line(i + 2)), *) m%angles
          i = i + 4
       end if
       i = i + 1
    end do
  end subroutine read_molecule

  ! lj_calc_AB (lj.F90:860-869): B = 4 eps sig^6 kB, A = B sig^6
  subroutine lj_calc_AB(sig, eps, A, B)
    real(c_double), intent(in) :: sig, eps
    real(c_double), intent(out) :: A, B
    real(c_double) :: s6
    s6 = sig**6
    B = 4.0d0 * eps * s6 * KCALMOLE_KB
    A = B * s6
  end subroutine lj_calc_AB

  ! atom_atom_file: 'a1 a2 LJ SIG@s EPS@e [HICUT@h] [LOCUT@l]' or '... LJ OFF'
  subroutine read_atom_atom(path, pairs, npairs)
    character(len=*), intent(in) :: path
    type(lj_pair), allocatable, intent(out) :: pairs(:)
    integer, intent(out) :: npairs
    type(text_file) :: t
    character(len=LL) :: s, tok, key, val
    integer :: i, w, k
    call read_text(path, t)
    allocate(pairs(t%n))
    npairs = 0
    do i = 1, t%n
       s = strip(t%line(i))
       if (nwords(s) < 3) cycle
       if (upper(trim(word(s, 3))) /= 'LJ') cycle
       npairs = npairs + 1
       associate (p => pairs(npairs))
         p%a1 = word(s, 1)
         p%a2 = word(s, 2)
         do w = 4, nwords(s)
            tok = word(s, w)
            k = index(tok, '@')
            if (k > 0) then
               key = upper(tok(1:k - 1))
               val = tok(k + 1:)
            else
               key = upper(tok)
               val = ''
            end if
            select case (trim(key))
            case ('OFF')
               p%off = .true.
               p%hicut = 0.0d0
            case ('SIG')
               if (index(upper(val), 'MIX') == 0) read(val, *) p%sig
            case ('EPS')
               if (index(upper(val), 'MIX') == 0) read(val, *) p%eps
            case ('HICUT', 'TRUNC')
               read(val, *) p%hicut
            case ('LOCUT', 'LOWCUT')
               read(val, *) p%locut
            end select
         end do
         p%ok = (.not. p%off) .and. p%sig > 0.0d0
         if (p%ok) call lj_calc_AB(p%sig, p%eps, p%A, p%B)
       end associate
    end do
  end subroutine read_atom_atom

  ! spc_spc_file: 's1 s2 COUL SUM EWALD KAPPA@k KMAX@n LOCUT@l [SFACTOR] ...'
  function read_ewald(path, s1, s2) result(e)
    character(len=*), intent(in) :: path, s1, s2
    type(ewald_opts) :: e
    type(text_file) :: t
    character(len=LL) :: s, tok, key, val
    integer :: i, w, k
    call read_text(path, t)
    do i = 1, t%n
       s = strip(t%line(i))
       if (nwords(s) < 5) cycle
       if (upper(trim(word(s, 3))) /= 'COUL') cycle
       if (.not. ((trim(word(s, 1)) == trim(s1) .and. trim(word(s, 2)) == trim(s2)) .or. &
                  (trim(word(s, 1)) == trim(s2) .and. trim(word(s, 2)) == trim(s1)))) cycle
       if (index(upper(s), 'EWALD') == 0) cycle
       e%found = .true.
       do w = 4, nwords(s)
          tok = word(s, w)
          k = index(tok, '@')
          if (k == 0) cycle
          key = upper(tok(1:k - 1))
          val = tok(k + 1:)
          select case (trim(key))
          case ('KMAX')
             read(val, *) e%kmax
          case ('KAPPA')
             read(val, *) e%kappa
          case ('LOCUT')
             read(val, *) e%locut
          end select
       end do
       return
    end do
  end function read_ewald

end module music_inputs
