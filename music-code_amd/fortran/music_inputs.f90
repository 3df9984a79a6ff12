! music_inputs -- readers for the reference's input files used by the map
! maker and GCMC drivers: ctrlfile sections (located by tag string as
! general.F90 / filesrchstr do), atom files (atom.F90), molecule files with
! 'Coord_Info: Listed' coordinates and Fundcell_Info (molecule.F90,
! fundcell.F90), the atom-atom forcefield file (lj_init, lj.F90:88-262) and
! the species-species file (ssdriver.F90:81-253, EWALD options of
! ewald_init, ewald.F90:99-180).  Only the subset the GCMC energy path
! needs: LJ and Ewald pair models; slanted cells for the map driver.
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
     logical :: rigid = .false., translate = .true., rotate = .true.
     integer :: dof = 0                              ! Molecule_DOF (0: not given)
     logical :: has_fundcell = .false.
     real(c_double) :: ell(3) = 0.0d0, angles(3) = 90.0d0, eff(3) = 0.0d0
  end type molecule

  type, public :: lj_pair
     character(len=NL) :: a1 = '', a2 = ''
     logical :: off = .false., ok = .false.
     integer :: sigmix = 0, epsmix = 0   ! 1: LBMIX, 2: JORGMIX (lj_calc_mix)
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

  ! everything the drivers read from a ctrlfile and the files it names:
  ! atom and molecule types, the simulation cell (a fundamental-cell species
  ! replicated nxyz times, or a NULL cell of given edges) and the atom-atom
  ! pairs of the forcefield
  type, public :: sim_system
     character(len=LL) :: ctrl = '', base = ''
     type(text_file) :: t
     integer :: nat = 0, nmol = 0, ifw = 0, ntot = 0, npairs = 0
     type(atom_type), allocatable :: atoms(:)
     type(molecule), allocatable :: mols(:)
     integer :: nxyz(3) = 1
     real(c_double) :: fund(3) = 0.0d0, full(3) = 0.0d0
     ! cell angles (degrees), effective (map) box, lattice vectors lvec(:, i)
     ! of the fundamental cell; slanted: not orthorhombic within 0.001 degree
     real(c_double) :: angles(3) = 90.0d0, eff(3) = 0.0d0, lvec(3, 3) = 0.0d0
     logical :: slanted = .false.
     real(c_double), allocatable :: xyz(:,:), q(:)   ! replicated fixed species
     integer(c_int), allocatable :: typ(:)
     character(len=NL) :: aafile = '', ssfile = ''
     type(lj_pair), allocatable :: pairs(:)
  end type sim_system

  public :: read_text, strip, lower, upper, find_section, next_nonblank, word, nwords, read_coul_pair
  public :: find_file, read_atom, read_molecule, read_atom_atom, read_ewald, lj_calc_AB
  public :: dirname, fail, need, read_system, mol_index, pair_index, principal_axes

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
          ! molecule.F90:1008-1028: coordsys, gcmodel type, flag
          m%rigid = trim(upper(word(s, 4))) == 'RIGID'
          m%translate = trim(upper(word(s, 5))) /= 'NOTRANSFORM'
          m%rotate = m%translate .and. trim(upper(word(s, 5))) /= 'NOROTATE'
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
       else if (index(s, 'Molecule_DOF:') == 1) then
          read(s(14:), *) m%dof
       else if (index(s, 'Fundcell_Info:') == 1) then
          m%has_fundcell = .true.
          read(strip(t%line(i + 1)), *) m%ell
          read(strip(t%line(i + 2)), *) m%angles
          read(strip(t%line(i + 4)), *) m%eff
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
               if (trim(upper(val)) == 'LBMIX') then
                  p%sigmix = 1
               else if (trim(upper(val)) == 'JORGMIX') then
                  p%sigmix = 2
               else
                  read(val, *) p%sig
               end if
            case ('EPS')
               if (index(upper(val), 'MIX') > 0) then
                  p%epsmix = 1
               else
                  read(val, *) p%eps
               end if
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
    ! mixed pairs from the pure ones (lj_calc_mix, lj.F90:905-979): sigma
    ! arithmetic (LBMIX) or geometric (JORGMIX), epsilon geometric
    do i = 1, npairs
       associate (p => pairs(i))
         if (p%off .or. (p%sigmix == 0 .and. p%epsmix == 0)) cycle
         if (p%sigmix == 0 .or. p%epsmix == 0) call fail('mixing rule needs both SIG and EPS: ' // &
              trim(p%a1) // ' ' // trim(p%a2))
         w = pure_pair(p%a1)
         k = pure_pair(p%a2)
         if (w == 0 .or. k == 0) call fail('mixing rule without the pure pairs: ' // trim(p%a1) // ' ' // &
              trim(p%a2))
         if (p%sigmix == 1) then
            p%sig = (pairs(w)%sig + pairs(k)%sig) / 2.0d0
         else
            p%sig = sqrt(pairs(w)%sig * pairs(k)%sig)
         end if
         p%eps = sqrt(pairs(w)%eps * pairs(k)%eps)
         p%ok = .true.
         call lj_calc_AB(p%sig, p%eps, p%A, p%B)
       end associate
    end do
  contains
    integer function pure_pair(a)
      character(len=*), intent(in) :: a
      integer :: j
      pure_pair = 0
      do j = 1, npairs
         if (trim(pairs(j)%a1) == trim(a) .and. trim(pairs(j)%a2) == trim(a) .and. pairs(j)%sigmix == 0 &
              .and. pairs(j)%sig > 0.0d0) pure_pair = j
      end do
    end function pure_pair
  end subroutine read_atom_atom

  ! atom_atom_file 'a1 a2 COUL [OFF] [HICUT@h] [SHIELDCONST@s]' for one atom
  ! pair (coulomb_init, coulomb.F90:60-231: locut LOW_DIST_CUTOFF, hicut 0 =
  ! no cutoff, shield 1).  found = .false. when the file has no such line;
  ! Q1@/Q2@ (charges from this file), LOCUT@ and TRUNC@ stop the reference
  ! and are refused here.
  subroutine read_coul_pair(path, a1, a2, found, off, hicut, shield)
    character(len=*), intent(in) :: path, a1, a2
    logical, intent(out) :: found, off
    real(c_double), intent(out) :: hicut, shield
    type(text_file) :: t
    character(len=LL) :: s, tok, key, val
    integer :: i, w, k
    found = .false.
    off = .false.
    hicut = 0.0d0
    shield = 1.0d0
    call read_text(path, t)
    do i = 1, t%n
       s = strip(t%line(i))
       if (nwords(s) < 3) cycle
       if (upper(trim(word(s, 3))) /= 'COUL') cycle
       if (.not. ((trim(word(s, 1)) == trim(a1) .and. trim(word(s, 2)) == trim(a2)) .or. &
                  (trim(word(s, 1)) == trim(a2) .and. trim(word(s, 2)) == trim(a1)))) cycle
       found = .true.
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
             off = .true.
          case ('HICUT')
             read(val, *) hicut
          case ('SHIELDCONST')
             read(val, *) shield
          case default
             call fail('atom-atom COUL option not offered: ' // trim(s))
          end select
       end do
       return
    end do
  end subroutine read_coul_pair

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

  subroutine fail(msg)
    character(len=*), intent(in) :: msg
    write(*, '(a)') trim(msg)
    stop 1
  end subroutine fail

  subroutine need(i, tag)
    integer, intent(in) :: i
    character(len=*), intent(in) :: tag
    if (i < 0) call fail('ctrlfile section not found: ' // tag)
  end subroutine need

  integer function mol_index(s, name)
    type(sim_system), intent(in) :: s
    character(len=*), intent(in) :: name
    integer :: j
    do j = 1, s%nmol
       if (trim(s%mols(j)%name) == trim(name)) then
          mol_index = j
          return
       end if
    end do
    call fail('unknown species ' // trim(name))
    mol_index = -1
  end function mol_index

  ! forcefield pair of two atom-type names (either order), 0 if absent
  integer function pair_index(s, n1, n2)
    type(sim_system), intent(in) :: s
    character(len=*), intent(in) :: n1, n2
    integer :: j
    pair_index = 0
    do j = 1, s%npairs
       if ((trim(s%pairs(j)%a1) == trim(n1) .and. trim(s%pairs(j)%a2) == trim(n2)) .or. &
           (trim(s%pairs(j)%a1) == trim(n2) .and. trim(s%pairs(j)%a2) == trim(n1))) pair_index = j
    end do
  end function pair_index

  ! ctrlfile sections 'Atomic Types', 'Molecule Types', 'Simulation Cell
  ! Information' (simcell_init: a fundamental-cell species and nxyz copies,
  ! or NULL followed by the cell edges, angles, origin, effective edges,
  ! nxyz and periodicity) and 'Forcefield Information'; the fixed species is
  ! replicated to the simulation cell as simcell.F90:200-216 does (copies
  ! along x, then y, then z)
  subroutine read_system(ctrl, s)
    character(len=*), intent(in) :: ctrl
    type(sim_system), intent(inout) :: s
    character(len=LL) :: w, fwname
    integer :: i, k, n, ax, c
    real(c_double) :: off(3), ang(3)
    s%ctrl = ctrl
    s%base = dirname(ctrl)
    call read_text(ctrl, s%t)
    i = find_section(s%t, 'Atomic Types')
    call need(i, 'Atomic Types')
    read(strip(s%t%line(i)), *) s%nat
    allocate(s%atoms(s%nat))
    i = i + 1
    do k = 1, s%nat
       i = next_nonblank(s%t, i)
       s%atoms(k)%name = word(strip(s%t%line(i)), 1)
       w = word(strip(s%t%line(i + 1)), 1)
       call read_atom(find_file(w, 'ATOMSDIR', 'atoms', s%base), s%atoms(k))
       i = i + 2
    end do
    i = find_section(s%t, 'Molecule Types')
    call need(i, 'Molecule Types')
    read(strip(s%t%line(i)), *) s%nmol
    allocate(s%mols(s%nmol))
    i = i + 1
    do k = 1, s%nmol
       i = next_nonblank(s%t, i)
       s%mols(k)%name = word(strip(s%t%line(i)), 1)
       w = word(strip(s%t%line(i + 1)), 1)
       call read_molecule(find_file(w, 'MOLSDIR', 'molecules', s%base), s%atoms, s%mols(k))
       i = i + 2
    end do
    i = find_section(s%t, 'Simulation Cell Information')
    call need(i, 'Simulation Cell Information')
    fwname = word(strip(s%t%line(i)), 1)
    if (trim(upper(fwname)) == 'NULL') then
       s%ifw = 0
       read(strip(s%t%line(i + 1)), *) s%fund
       read(strip(s%t%line(i + 2)), *) ang
       read(strip(s%t%line(i + 4)), *) s%eff
       read(strip(s%t%line(i + 5)), *) s%nxyz
       s%angles = ang
       s%full = s%fund * s%nxyz
       s%ntot = 0
       allocate(s%xyz(3, 1), s%typ(1), s%q(1))
       call lattice_vectors(s%fund, s%angles, s%lvec)
    else
       read(strip(s%t%line(i + 1)), *) s%nxyz
       s%ifw = mol_index(s, fwname)
       associate (fw => s%mols(s%ifw))
         if (.not. fw%has_fundcell) call fail('fundamental cell molecule has no Fundcell_Info')
         s%fund = fw%ell
         s%angles = fw%angles
         s%eff = fw%eff
         s%full = s%fund * s%nxyz
         n = fw%natoms
         s%ntot = n * s%nxyz(1) * s%nxyz(2) * s%nxyz(3)
         allocate(s%xyz(3, s%ntot), s%typ(s%ntot), s%q(s%ntot))
         s%xyz(:, 1:n) = fw%xyz
         s%typ(1:n) = fw%atype
         s%q(1:n) = fw%q
       end associate
       ! copies along the lattice vectors a, then b, then c (simcell.F90:
       ! 200-216, molecule_expand, molecule.F90:637-700: shift lvec * c)
       call lattice_vectors(s%fund, s%angles, s%lvec)
       do ax = 1, 3
          do c = 1, s%nxyz(ax) - 1
             off = s%lvec(:, ax) * real(c, c_double)
             do k = 1, n
                s%xyz(:, c * n + k) = s%xyz(:, k) + off
                s%typ(c * n + k) = s%typ(k)
                s%q(c * n + k) = s%q(k)
             end do
          end do
          n = n * s%nxyz(ax)
       end do
    end if
    ! fundcell_init (fundcell.F90:96-102): orthorhombic within 0.001 degree
    s%slanted = sum(abs(s%angles - 90.0d0)) > 0.001d0
    if (all(s%eff <= 0.0d0)) s%eff = s%fund
    i = find_section(s%t, 'Forcefield Information')
    call need(i, 'Forcefield Information')
    s%aafile = word(strip(s%t%line(i + 2)), 1)
    s%ssfile = word(strip(s%t%line(i + 3)), 1)
    call read_atom_atom(find_file(s%aafile, 'NONE', '', s%base), s%pairs, s%npairs)
  end subroutine read_system

  ! fundcell_genlatticevects (fundcell.F90:174-208): a along x, b in the xy
  ! plane, c completing the edge lengths; angles in radians with the
  ! reference's pi = 3.1415927 (defaults.F90:33-35)
  subroutine lattice_vectors(ell, ang, lv)
    real(c_double), intent(in) :: ell(3), ang(3)
    real(c_double), intent(out) :: lv(3, 3)
    real(c_double), parameter :: degtorad = 3.1415927d0 / 180.0d0
    real(c_double) :: abc, aac, aab, m2
    abc = ang(1) * degtorad
    aac = ang(2) * degtorad
    aab = ang(3) * degtorad
    lv(:, 1) = [ell(1), 0.0d0, 0.0d0]
    lv(:, 2) = [ell(2) * cos(aab), ell(2) * sin(aab), 0.0d0]
    lv(1, 3) = ell(3) * cos(aac)
    lv(2, 3) = ell(3) * (cos(abc) - cos(aac) * cos(aab)) / sin(aab)
    m2 = 0.0d0
    m2 = m2 + lv(1, 3) * lv(1, 3)
    m2 = m2 + lv(2, 3) * lv(2, 3)
    lv(3, 3) = sqrt(ell(3) * ell(3) - m2)
  end subroutine lattice_vectors

  ! molecule_genprincipleaxes (molecule.F90:830-926), applied to rigid
  ! molecules of 3 or more atoms (molecule.F90:714-716, 843), in the
  ! reference's arithmetic order: unless NOTRANSFORM, shift by the centre of
  ! mass of the real masses (molecule_getrefcom, molecule.F90:777-798: sum of
  ! r m times the reciprocal of the mass sum); form the inertia tensor with
  ! integer-truncated masses (molecule.F90:833, 872-892); diagonalise it by
  ! Householder tridiagonalisation and implicit QL (tred2 / tql2, restated in
  ! tred2_tql2 below) and, unless NOROTATE or NOTRANSFORM, rotate every atom
  ! by the transposed eigenvector matrix (matrix_m_mult_v, matrix.F90:137-150).
  subroutine principal_axes(m, atoms, ref)
    type(molecule), intent(in) :: m
    type(atom_type), intent(in) :: atoms(:)
    real(c_double), intent(out) :: ref(3, m%natoms)
    real(c_double) :: ms(m%natoms), com(3), r(3, m%natoms), I3(3, 3), z(3, 3), d(3), msum, inv, moi, rsq, v
    integer :: a, i, j, k, mi
    r = m%xyz
    if (.not. m%rigid .or. m%natoms < 3) then
       ref = r
       return
    end if
    do a = 1, m%natoms
       ms(a) = atoms(m%atype(a) + 1)%mass
    end do
    if (m%translate) then
       com = 0.0d0
       msum = 0.0d0
       do a = 1, m%natoms
          msum = msum + ms(a)
          com = com + m%xyz(:, a) * ms(a)
       end do
       inv = 1.0d0 / msum
       com = com * inv
       do a = 1, m%natoms
          r(:, a) = r(:, a) + (-com)
       end do
    end if
    do i = 1, 3
       do j = 1, 3
          moi = 0.0d0
          do a = 1, m%natoms
             mi = int(ms(a))
             if (i == j) then
                rsq = r(1, a) * r(1, a) + r(2, a) * r(2, a) + r(3, a) * r(3, a)
                moi = moi + real(mi, c_double) * (rsq - r(i, a) * r(i, a))
             else
                moi = moi - real(mi, c_double) * r(i, a) * r(j, a)
             end if
          end do
          I3(i, j) = moi
       end do
    end do
    call tred2_tql2(I3, d, z)
    if (.not. m%rotate) then
       ref = r
       return
    end if
    do a = 1, m%natoms
       do k = 1, 3
          v = 0.0d0
          do j = 1, 3
             v = v + z(j, k) * r(j, a)
          end do
          ref(k, a) = v
       end do
    end do
  end subroutine principal_axes

  ! sqrt(a^2 + b^2) by the iteration of EISPACK's pythag (Moler and
  ! Morrison, IBM J. Res. Dev. 27 (1983) 577)
  real(c_double) function pythag(a, b) result(p)
    real(c_double), intent(in) :: a, b
    real(c_double) :: r, s, t, u
    p = max(abs(a), abs(b))
    if (p == 0.0d0) return
    r = (min(abs(a), abs(b)) / p)**2
    do
       t = 4.0d0 + r
       if (t == 4.0d0) exit
       s = r / t
       u = 1.0d0 + 2.0d0 * s
       p = u * p
       r = (s / u)**2 * r
    end do
  end function pythag

  ! Eigenvalues d (ascending) and eigenvectors z (columns) of the symmetric
  ! 3x3 matrix a (lower triangle read): Householder reduction to
  ! tridiagonal form with the transformations accumulated (EISPACK tred2;
  ! Martin, Reinsch and Wilkinson, Numer. Math. 11 (1968) 181), then the
  ! implicit QL method (EISPACK tql2; Bowdler, Martin, Reinsch and
  ! Wilkinson, Numer. Math. 11 (1968) 293).  The steps are those of the
  ! published algorithms, which the reference calls (molecule.F90:896-897),
  ! so the eigenvectors carry the reference's signs, also for equal moments.
  subroutine tred2_tql2(a, d, z)
    real(c_double), intent(in) :: a(3, 3)
    real(c_double), intent(out) :: d(3), z(3, 3)
    integer, parameter :: n = 3
    real(c_double) :: e(3), f, g, h, hh, scale, c, c2, c3, dl1, el1, p, r, s, s2, tst1, tst2
    integer :: i, j, k, l, m, ii, l1, l2, it
    do i = 1, n
       do j = i, n
          z(j, i) = a(j, i)
       end do
       d(i) = a(n, i)
    end do
    do i = n, 2, -1
       l = i - 1
       h = 0.0d0
       scale = 0.0d0
       if (l >= 2) then
          do k = 1, l
             scale = scale + abs(d(k))
          end do
       end if
       if (l < 2 .or. scale == 0.0d0) then
          e(i) = d(l)
          do j = 1, l
             d(j) = z(l, j)
             z(i, j) = 0.0d0
             z(j, i) = 0.0d0
          end do
       else
          do k = 1, l
             d(k) = d(k) / scale
             h = h + d(k) * d(k)
          end do
          f = d(l)
          g = -sign(sqrt(h), f)
          e(i) = scale * g
          h = h - f * g
          d(l) = f - g
          do j = 1, l
             e(j) = 0.0d0
          end do
          do j = 1, l
             f = d(j)
             z(j, i) = f
             g = e(j) + z(j, j) * f
             do k = j + 1, l
                g = g + z(k, j) * d(k)
                e(k) = e(k) + z(k, j) * f
             end do
             e(j) = g
          end do
          f = 0.0d0
          do j = 1, l
             e(j) = e(j) / h
             f = f + e(j) * d(j)
          end do
          hh = f / (h + h)
          do j = 1, l
             e(j) = e(j) - hh * d(j)
          end do
          do j = 1, l
             f = d(j)
             g = e(j)
             do k = j, l
                z(k, j) = z(k, j) - f * e(k) - g * d(k)
             end do
             d(j) = z(l, j)
             z(i, j) = 0.0d0
          end do
       end if
       d(i) = h
    end do
    do i = 2, n
       l = i - 1
       z(n, l) = z(l, l)
       z(l, l) = 1.0d0
       h = d(i)
       if (h /= 0.0d0) then
          do k = 1, l
             d(k) = z(k, i) / h
          end do
          do j = 1, l
             g = 0.0d0
             do k = 1, l
                g = g + z(k, i) * z(k, j)
             end do
             do k = 1, l
                z(k, j) = z(k, j) - g * d(k)
             end do
          end do
       end if
       do k = 1, l
          z(k, i) = 0.0d0
       end do
    end do
    do i = 1, n
       d(i) = z(n, i)
       z(n, i) = 0.0d0
    end do
    z(n, n) = 1.0d0
    e(1) = 0.0d0
    ! implicit QL
    do i = 2, n
       e(i - 1) = e(i)
    end do
    f = 0.0d0
    tst1 = 0.0d0
    e(n) = 0.0d0
    do l = 1, n
       it = 0
       h = abs(d(l)) + abs(e(l))
       if (tst1 < h) tst1 = h
       do m = l, n
          tst2 = tst1 + abs(e(m))
          if (tst2 == tst1) exit
       end do
       if (m /= l) then
          do
             if (it == 30) call fail('principal axes: tql2 did not converge')
             it = it + 1
             l1 = l + 1
             l2 = l1 + 1
             g = d(l)
             p = (d(l1) - g) / (2.0d0 * e(l))
             r = pythag(p, 1.0d0)
             d(l) = e(l) / (p + sign(r, p))
             d(l1) = e(l) * (p + sign(r, p))
             dl1 = d(l1)
             h = g - d(l)
             do i = l2, n
                d(i) = d(i) - h
             end do
             f = f + h
             p = d(m)
             c = 1.0d0
             c2 = c
             el1 = e(l1)
             s = 0.0d0
             s2 = 0.0d0
             c3 = 0.0d0
             do ii = 1, m - l
                c3 = c2
                c2 = c
                s2 = s
                i = m - ii
                g = c * e(i)
                h = c * p
                r = pythag(p, e(i))
                e(i + 1) = s * r
                s = e(i) / r
                c = p / r
                p = c * d(i) - s * g
                d(i + 1) = h + s * (c * g + s * d(i))
                do k = 1, n
                   h = z(k, i + 1)
                   z(k, i + 1) = s * z(k, i) + c * h
                   z(k, i) = c * z(k, i) - s * h
                end do
             end do
             p = -s * s2 * c3 * el1 * e(l) / dl1
             e(l) = s * p
             d(l) = c * p
             tst2 = tst1 + abs(e(l))
             if (.not. tst2 > tst1) exit
          end do
       end if
       d(l) = d(l) + f
    end do
    do ii = 2, n
       i = ii - 1
       k = i
       p = d(i)
       do j = ii, n
          if (d(j) < p) then
             k = j
             p = d(j)
          end if
       end do
       if (k /= i) then
          d(k) = d(i)
          d(i) = p
          do j = 1, n
             p = z(j, i)
             z(j, i) = z(j, k)
             z(j, k) = p
          end do
       end if
    end do
  end subroutine tred2_tql2

end module music_inputs
