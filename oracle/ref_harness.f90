! ref_harness.f90 -- test-only driver around the reference's own Fortran
! modules (random, utils, lj, simcell, ewald, pmap, interpolate), compiled
! from the reference source tree by oracle/Makefile.ref into oracle/_ref/.
!
! It reads one job from stdin and prints results to stdout so that the C
! oracle (oracle/music_oracle.c) can be checked against the reference's own
! arithmetic.  Only the loops that glue the reference routines together are
! written here; every pair term, structure factor, compaction, file record
! and interpolation comes from the reference procedures named below.
!
! Jobs:
!   RNG seed n                           -> n values of rranf()
!   ERFC n x(1..n)                       -> erfc(x)
!   LJNODE  (see read order below)       -> ok, V, hot(2:8) per node
!   EWALD   (see read order below)       -> ok, V, hot(2:8) per node
!   COMPACT (see read order below)       -> counts, writes a map file
!   HERMITE (see read order below)       -> ok, V per point
!   COULOMB (see read order below)       -> ok, V per pair (coulomb_snglint)
!   EULER   (see read order below)       -> site coordinates (matrix_b2feuler)
!   GAUSS seed n                         -> n values of random_gaussian(0, 1)
!   BONDED  (see read order below)       -> energies and forces of harmonic,
!                                           harmonica and cosexpansion terms
!   PAXES ctrlfile name molfile          -> principal moments and body-frame
!                                           coordinates (molecule_init)
!   R4CMP n ratio(1..n)                  -> ratio < 0.49, ratio > 0.51 with
!                                           default-real literals, as in
!                                           rigidmoves_adjustdeltatrans
!   FCELL   (see read order below)       -> slant / unslant matrices, lattice
!                                           vectors, volume, minimum images, maptouc,
!                                           pbc and uniform random points of a
!                                           (possibly slanted) cell
!   LJNODEX / HERMITEX                   -> LJNODE / HERMITE in a cell given
!                                           by a cell block (fundcell_init)
program ref_harness
  Use defaults, Only: RDbl, RSgl
  Use random, Only: random_init, rranf
  Use utils, Only: erfc
  Use vector, Only: VecType
  Use storebase, Only: EnergyPlus, storebase_init, storebase_zero
  Use lj, Only: LJ_Params, lj_snglint
  Use simcell, Only: SimCell_Params, simcell_minimage
  Use ewald, Only: ewald_init, ewald_initCalc, ewald_getInteractionHOT
  Use pmap, Only: Pmap_Params, pmap_newmap, pmap_addpoint, pmap_compact, &
      pmap_write, pmap_pmapinit, pmap_int
  Use coulomb, Only: Coulomb_Params, coulomb_snglint
  Use matrix, Only: MatrixType, matrix_b2feuler, Operator(*)
  Use random, Only: random_gaussian
  Use harmonic, Only: HarmonicParams, harmonic_getinteraction
  Use harmonica, Only: HarmonicAModel, harmonica_getinteraction
  Use cosexpansion, Only: CosExpansionModel, cosexpansion_getinteraction
  Use defaults, Only: d_ctrl_file
  Use file, Only: file_settag
  Use atom, Only: atom_init
  Use molecule, Only: MolecularParams, molecule_init, molecule_getatomcoords, &
      molecule_getmoi
  Use fundcell, Only: Fundamental_Cell, fundcell_init, fundcell_construct
  Use vector, Only: IntVecType
  Use simcell, Only: simcell_maptouc, simcell_pbc, simcell_uniformRandPt, simcell_getvolume
  Implicit None

  Character(len=32) :: job
  Read(*,*) job
  Select Case (Trim(job))
  Case ('RNG')
    Call job_rng()
  Case ('ERFC')
    Call job_erfc()
  Case ('LJNODE')
    Call job_ljnode(.False.)
  Case ('EWALD')
    Call job_ewald()
  Case ('COMPACT')
    Call job_compact()
  Case ('HERMITE')
    Call job_hermite(.False.)
  Case ('COULOMB')
    Call job_coulomb()
  Case ('EULER')
    Call job_euler()
  Case ('GAUSS')
    Call job_gauss()
  Case ('BONDED')
    Call job_bonded()
  Case ('PAXES')
    Call job_paxes()
  Case ('R4CMP')
    Call job_r4cmp()
  Case ('FCELL')
    Call job_fcell()
  Case ('LJNODEX')
    Call job_ljnode(.True.)
  Case ('HERMITEX')
    Call job_hermite(.True.)
  Case Default
    Write(0,*) 'unknown job ', Trim(job)
    Stop 2
  End Select

Contains

  Subroutine set_cell(cell, fund, full)
    Type(SimCell_Params), Intent(Out) :: cell
    Real(RDbl), Intent(In) :: fund(3), full(3)
    cell%fundC%orthoflag = .True.
    cell%fundC%ell = fund
    cell%fundC%half_ell = 0.5_RDbl*fund
    cell%fundC%ell_inv = 1.0_RDbl/fund
    cell%fundC%eff = fund
    cell%fundC%origin%comp = 0.0_RDbl
    cell%fullC%orthoflag = .True.
    cell%fullC%ell = full
    cell%fullC%half_ell = 0.5_RDbl*full
    cell%fullC%ell_inv = 1.0_RDbl/full
    cell%fullC%eff = full
    cell%fullC%origin%comp = 0.0_RDbl
  End Subroutine set_cell

  ! cell block: fund edges a b c; angles alpha beta gamma (degrees); origin;
  ! eff; cells per axis.  fundcell_init reads the first four lines from a
  ! unit (fundcell.F90:60-112); the simulation cell is fundcell_construct of
  ! it (:345-371).
  Subroutine read_cell(cell)
    Type(SimCell_Params), Intent(Out) :: cell
    Real(RDbl) :: v(12)
    Integer :: nc(3), u
    Type(IntVecType) :: sz
    Read(*,*) v(1:3)
    Read(*,*) v(4:6)
    Read(*,*) v(7:9)
    Read(*,*) v(10:12)
    Read(*,*) nc
    Open(newunit=u, status='SCRATCH', form='FORMATTED')
    Write(u,'(3ES26.17E3)') v(1:3)
    Write(u,'(3ES26.17E3)') v(4:6)
    Write(u,'(3ES26.17E3)') v(7:9)
    Write(u,'(3ES26.17E3)') v(10:12)
    Rewind(u)
    Call fundcell_init(cell%fundC, u)
    Close(u)
    sz%comp = nc
    Call fundcell_construct(cell%fundC, sz, cell%fullC)
  End Subroutine read_cell

  Subroutine job_rng()
    Integer :: seed, n, i
    Read(*,*) seed, n
    Call random_init(seed)
    Do i = 1, n
      Write(*,'(ES26.17E3)') rranf()
    End Do
  End Subroutine job_rng

  Subroutine job_erfc()
    Integer :: n, i
    Real(RDbl) :: x
    Read(*,*) n
    Do i = 1, n
      Read(*,*) x
      Write(*,'(ES26.17E3)') erfc(x)
    End Do
  End Subroutine job_erfc

  ! LJNODE input: full(3); natoms; natoms x (x y z type); ntypes;
  ! ntypes x (A B hicut2 locut2); nnodes; nnodes x (x y z)
  ! Loop structure of pairmodel_mmint5 (pairmodel.F90:892-909) and node
  ! assembly of mapmaker_makemap (mapmaker.F90:500-547).
  Subroutine job_ljnode(xcell)
    Logical, Intent(In) :: xcell
    Type(SimCell_Params) :: cell
    Real(RDbl) :: full(3), p(3), pot
    Integer :: natoms, ntypes, nnodes, a, t, nd
    Type(VecType), Allocatable :: atv(:)
    Type(VecType) :: pv(1)
    Type(VecType), Allocatable :: sep(:,:)
    Integer, Allocatable :: typ(:)
    Type(LJ_Params), Allocatable :: par(:)
    Type(EnergyPlus) :: ffout, opp
    Real(RDbl) :: hot(8), hotatm(8)
    Logical :: ok
    If (xcell) Then
      Call read_cell(cell)
    Else
      Read(*,*) full
      Call set_cell(cell, full, full)
    End If
    Read(*,*) natoms
    Allocate(atv(natoms), typ(natoms), sep(1,natoms))
    Do a = 1, natoms
      Read(*,*) atv(a)%comp(1), atv(a)%comp(2), atv(a)%comp(3), typ(a)
    End Do
    Read(*,*) ntypes
    Allocate(par(ntypes))
    Do t = 1, ntypes
      Read(*,*) par(t)%A, par(t)%B, par(t)%hicut2, par(t)%locut2
    End Do
    Call storebase_init(ffout, 2)
    Call storebase_init(opp, 2)
    Read(*,*) nnodes
    Do nd = 1, nnodes
      Read(*,*) p
      pv(1)%comp = p
      Call storebase_zero(ffout)
      hot = 0.0_RDbl
      Call simcell_minimage(cell, pv, atv, sep)
      ok = .True.
      Do a = 1, natoms
        hotatm = 0.0_RDbl
        ok = lj_snglint(par(typ(a)+1), sep(1,a), ffout, opp, hotatm)
        hot = hot + hotatm
        If (.Not. ok) Exit
      End Do
      If (ok) Then
        pot = ffout%nrg
      Else
        pot = 1.0e10_RDbl
      End If
      Write(*,'(L2,8ES26.17E3)') ok, pot, hot(2:8)
    End Do
  End Subroutine job_ljnode

  ! EWALD input: ewald line; full(3); natoms; natoms x (x y z q);
  ! charge1; nnodes; nnodes x (x y z)
  Subroutine job_ewald()
    Character(len=255) :: line
    Type(SimCell_Params) :: cell
    Real(RDbl) :: full(3), p(3), charge1, pot
    Integer :: natoms, nnodes, a, nd
    Type(VecType), Allocatable :: atv(:)
    Type(VecType) :: pv
    Real(RDbl), Allocatable :: q(:)
    Real(RDbl) :: hot(8)
    Logical :: ok
    Read(*,'(a)') line
    Read(*,*) full
    Call set_cell(cell, full, full)
    Read(*,*) natoms
    Allocate(atv(natoms), q(natoms))
    Do a = 1, natoms
      Read(*,*) atv(a)%comp(1), atv(a)%comp(2), atv(a)%comp(3), q(a)
    End Do
    Read(*,*) charge1
    Call ewald_init(line)
    Call ewald_initCalc(atv, q, full)
    Read(*,*) nnodes
    Do nd = 1, nnodes
      Read(*,*) p
      pv%comp = p
      ok = .False.
      Call ewald_getInteractionHOT(cell, pv, atv, charge1, q, pot, hot, ok)
      If (.Not. ok) pot = 1.0e10_RDbl
      Write(*,'(L2,8ES26.17E3)') ok, pot, hot(2:8)
    End Do
  End Subroutine job_ewald

  ! COMPACT input: filename; nx ny nz; cutoff(kcal); pcut(kJ);
  ! nzatoms; nzatoms x (name symbol / charge A12 B6 C2 D0 rcuthi rcutlo);
  ! then nx*ny*nz lines of 8 values in config_makegrid order (x fastest).
  ! Uses pmap_newmap/pmap_addpoint/pmap_compact/pmap_write as
  ! mapmaker_makemap does (mapmaker.F90:466-592).
  Subroutine job_compact()
    Character(len=255) :: fname
    Type(Pmap_Params) :: big, lil
    Integer :: np(3), nzat, i, x, y, z, u
    Real(RDbl) :: cutoff, vals(8), hot(8)
    Real(RSgl) :: pcut
    Type(VecType) :: r
    Character(len=25) :: nm
    Character(len=2) :: sy
    Read(*,'(a)') fname
    Read(*,*) np
    Read(*,*) cutoff
    Read(*,*) pcut
    Call pmap_newmap(big, np(1)*np(2)*np(3), np)
    Call pmap_newmap(lil, np(1)*np(2)*np(3), np)
    big%header%stepx = 1.0_RDbl
    big%header%stepy = 1.0_RDbl
    big%header%stepz = 1.0_RDbl
    Read(*,*) nzat
    lil%header%nzatoms = nzat
    Allocate(lil%header%zatom(nzat))
    Do i = 1, nzat
      Read(*,'(a25,1x,a2)') nm, sy
      lil%header%zatom(i)%atom_name = nm
      lil%header%zatom(i)%symbol = sy
      Read(*,*) lil%header%zatom(i)%charge, lil%header%zatom(i)%A12, &
          lil%header%zatom(i)%B6, lil%header%zatom(i)%C2, &
          lil%header%zatom(i)%D0, lil%header%zatom(i)%rcuthi, &
          lil%header%zatom(i)%rcutlo
    End Do
    lil%header%pcut_off = pcut
    Do z = 1, np(3)
      Do y = 1, np(2)
        Do x = 1, np(1)
          Read(*,*) vals
          r%comp = (/ Real(x-1,RDbl), Real(y-1,RDbl), Real(z-1,RDbl) /)
          hot = vals
          Call pmap_addpoint(big, r, vals(1), hot, .True.)
        End Do
      End Do
    End Do
    Call pmap_compact(big, lil, cutoff)
    ! pmap_write opens with status OLD (file.F90:208-210): create it first
    Open(newunit=u, file=Trim(fname), form='UNFORMATTED', status='REPLACE')
    Close(u)
    Call pmap_write(lil, Trim(fname))
    Write(*,'(I12)') lil%npoints
  End Subroutine job_compact

  ! HERMITE input: map file name (in directory $PMAPDIR); fund(3); npts;
  ! npts x (x y z)
  Subroutine job_hermite(xcell)
    Logical, Intent(In) :: xcell
    Character(len=255) :: fname
    Type(Pmap_Params), Pointer :: map
    Type(SimCell_Params) :: cell
    Type(EnergyPlus) :: out
    Type(VecType) :: pt
    Real(RDbl) :: fund(3), p(3)
    Integer :: n, i
    Logical :: ok
    Read(*,'(a)') fname
    Read(*,*) fund
    If (xcell) Then
      Call read_cell(cell)
    Else
      Call set_cell(cell, fund, fund)
    End If
    Nullify(map)
    Call pmap_pmapinit(map, fund, Trim(fname))
    Call storebase_init(out, 0)
    Read(*,*) n
    Do i = 1, n
      Read(*,*) p
      pt%comp = p
      Call storebase_zero(out)
      ok = pmap_int(map, cell, pt, out)
      Write(*,'(L2,ES26.17E3)') ok, out%nrg
    End Do
  End Subroutine job_hermite

  ! COULOMB input: hicut locut shield; npairs; npairs x (sx sy sz q1 q2)
  ! (hicut <= 0: no cutoff, as coulomb_init leaves it without HICUT)
  Subroutine job_coulomb()
    Type(Coulomb_Params) :: par
    Type(EnergyPlus) :: res, res2
    Type(VecType) :: sep
    Real(RDbl) :: hic, loc, sh, v(5)
    Integer :: n, i
    Logical :: ok
    Read(*,*) hic, loc, sh
    par%off = .False.
    par%initialized = .True.
    par%hicut = hic
    par%hicut2 = hic*hic
    par%locut = loc
    par%locut2 = loc*loc
    par%shield_const = sh
    par%charges_fromAA_FILE = .False.
    par%Q1 = 0.0_RDbl
    par%Q2 = 0.0_RDbl
    Call storebase_init(res, 0)
    Call storebase_init(res2, 0)
    Read(*,*) n
    Do i = 1, n
      Read(*,*) v
      sep%comp = v(1:3)
      Call storebase_zero(res)
      ok = coulomb_snglint(par, sep, v(4), v(5), res, res2)
      Write(*,'(L2,ES26.17E3)') ok, res%nrg
    End Do
  End Subroutine job_coulomb

  ! EULER input: natoms; natoms x (x y z) body frame; n;
  ! n x (comx comy comz phi theta psi) -> natoms rows of xyz per molecule,
  ! com + matrix_b2feuler(phi, theta, psi) * ref as rigidcoords_toxyz
  ! (rigidcoords.F90:254-276)
  Subroutine job_euler()
    Type(MatrixType) :: m
    Type(VecType), Allocatable :: ref(:)
    Type(VecType) :: x
    Real(RDbl) :: v(6)
    Integer :: na, n, i, a
    Read(*,*) na
    Allocate(ref(na))
    Do a = 1, na
      Read(*,*) ref(a)%comp
    End Do
    Read(*,*) n
    Do i = 1, n
      Read(*,*) v
      m = matrix_b2feuler(v(4), v(5), v(6))
      Do a = 1, na
        x = m*ref(a)
        Write(*,'(3ES26.17E3)') v(1:3) + x%comp
      End Do
    End Do
  End Subroutine job_euler

  Subroutine job_gauss()
    Integer :: seed, n, i
    Read(*,*) seed, n
    Call random_init(seed)
    Do i = 1, n
      Write(*,'(ES26.17E3)') random_gaussian(0.0_RDbl, 1.0_RDbl)
    End Do
  End Subroutine job_gauss

  ! BONDED input: kind (1 harmonic bond k req, 2 harmonica angle k theta0,
  ! 3 cosexpansion c0..c5), its parameters; n; n x (coordinates of the 2, 3
  ! or 4 atoms) -> energy and the forces on each atom
  Subroutine job_bonded()
    Type(HarmonicParams) :: hb
    Type(HarmonicAModel), Pointer :: ha
    Type(CosExpansionModel) :: cx
    Type(VecType) :: r(4), f(4)
    Real(RDbl) :: p(6), c(12), u
    Integer :: kind, n, i, na, a
    Read(*,*) kind
    Select Case (kind)
    Case (1)
      Read(*,*) p(1:2)
      hb%k = p(1)
      hb%req = p(2)
      na = 2
    Case (2)
      Read(*,*) p(1:2)
      Allocate(ha)
      ha%ktheta = p(1)
      ha%thetaeq = p(2)
      na = 3
    Case Default
      Read(*,*) p(1:6)
      cx%cn = p(1:6)
      na = 4
    End Select
    Read(*,*) n
    Do i = 1, n
      Read(*,*) c(1:3*na)
      Do a = 1, na
        r(a)%comp = c(3*a-2:3*a)
      End Do
      Select Case (kind)
      Case (1)
        Call harmonic_getinteraction(hb, r(1:2), .False., u, f(1:2))
      Case (2)
        Call harmonica_getinteraction(ha, r(1:3), .False., u, f(1:3))
      Case Default
        Call cosexpansion_getinteraction(cx, r(1:4), .False., u, f(1:4))
      End Select
      Write(*,'(13ES26.17E3)') u, (f(a)%comp, a = 1, na)
    End Do
  End Subroutine job_bonded

  ! PAXES input: control file (its atom types section), molecule name, file
  ! name in $MOLSDIR -> natoms, principal moments, body-frame coordinates
  ! after molecule_init (molecule_coordinit -> molecule_genprincipleaxes,
  ! molecule.F90:982-1112, 830-926)
  Subroutine job_paxes()
    Character(len=255) :: ctrl, nm, fn
    Type(MolecularParams) :: mol
    Type(VecType) :: x
    Real(RDbl) :: moi(3)
    Integer :: a
    Read(*,'(a)') ctrl
    Read(*,'(a)') nm
    Read(*,'(a)') fn
    Call file_settag(Trim(ctrl), d_ctrl_file)
    Call atom_init()
    Call molecule_init(mol, Trim(nm), Trim(fn))
    moi = molecule_getmoi(mol)
    Write(*,'(A,I6)') 'NATOMS', mol%natoms
    Write(*,'(A,3ES26.17E3)') 'MOI', moi
    Do a = 1, mol%natoms
      x = molecule_getatomcoords(mol, a)
      Write(*,'(A,3ES26.17E3)') 'XYZ', x%comp
    End Do
  End Subroutine job_paxes

  ! R4CMP: the comparison of rigidmoves_adjustdeltatrans
  ! (rigidmoves.F90:2592-2597): a REAL(8) ratio against the literals 0.49
  ! and 0.51 written without a kind, exactly as there
  Subroutine job_r4cmp()
    Integer :: n, i
    Real(RDbl) :: ratio
    Read(*,*) n
    Do i = 1, n
      Read(*,*) ratio
      Write(*,'(2L2)') ratio < 0.49, ratio > 0.51
    End Do
  End Subroutine job_r4cmp

  ! FCELL input: cell block; npairs; npairs x (x1 y1 z1 x2 y2 z2); npts;
  ! npts x (x y z); seed nrand.  Output: ORTHO, U (unslant, row-major), S
  ! (slant), L (full-cell lattice vectors), per pair SEP (simcell_minimage,
  ! fundcell.F90:582-623) and dist2, per point UC (simcell_maptouc) and PBC
  ! (simcell_pbc), then RAND (simcell_uniformRandPt) after random_init(seed)
  Subroutine job_fcell()
    Type(SimCell_Params) :: cell
    Type(VecType) :: a, b, sv, pt, o
    Real(RDbl) :: v(6), d2
    Integer :: n, i, j, seed
    Call read_cell(cell)
    Write(*,'(A,L2)') 'ORTHO', cell%fullC%orthoflag
    Write(*,'(A,9ES26.17E3)') 'U', ((cell%fullC%unslantmatrix%comp(i,j), j = 1, 3), i = 1, 3)
    Write(*,'(A,9ES26.17E3)') 'S', ((cell%fullC%slantmatrix%comp(i,j), j = 1, 3), i = 1, 3)
    Write(*,'(A,9ES26.17E3)') 'L', (cell%fullC%lvec(i)%comp, i = 1, 3)
    Write(*,'(A,ES26.17E3)') 'VOL', simcell_getvolume(cell)
    Read(*,*) n
    Do i = 1, n
      Read(*,*) v
      a%comp = v(1:3)
      b%comp = v(4:6)
      Call simcell_minimage(cell, a, b, sv, d2)
      Write(*,'(A,4ES26.17E3)') 'SEP', sv%comp, d2
    End Do
    Read(*,*) n
    Do i = 1, n
      Read(*,*) v(1:3)
      pt%comp = v(1:3)
      o = simcell_maptouc(cell, pt)
      Write(*,'(A,3ES26.17E3)') 'UC', o%comp
      Call simcell_pbc(cell, pt, o)
      Write(*,'(A,3ES26.17E3)') 'PBC', o%comp
    End Do
    Read(*,*) seed, n
    Call random_init(seed)
    Do i = 1, n
      o = simcell_uniformRandPt(cell)
      Write(*,'(A,3ES26.17E3)') 'RAND', o%comp
    End Do
  End Subroutine job_fcell

End Program ref_harness
