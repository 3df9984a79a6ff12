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
  Implicit None

  Character(len=32) :: job
  Read(*,*) job
  Select Case (Trim(job))
  Case ('RNG')
    Call job_rng()
  Case ('ERFC')
    Call job_erfc()
  Case ('LJNODE')
    Call job_ljnode()
  Case ('EWALD')
    Call job_ewald()
  Case ('COMPACT')
    Call job_compact()
  Case ('HERMITE')
    Call job_hermite()
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
  Subroutine job_ljnode()
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
    Read(*,*) full
    Call set_cell(cell, full, full)
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
  Subroutine job_hermite()
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
    Call set_cell(cell, fund, fund)
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

End Program ref_harness
