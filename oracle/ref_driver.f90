!-------------------------------------------------------------------------------
! Golden-vector driver for the reference MOCCa modules.
!
! Test infrastructure only: this program is linked against the reference's
! own Fortran modules (compiled in place by oracle/ref.mk) and writes their
! outputs for seeded inputs, so that the C restatement under oracle/ and the
! HIP library can be compared with what the reference itself computes.
!
! Usage:   mocca_ref_driver <input.bin> <outprefix> <task> < namelists.in
!   (run in a directory holding forces.param)
!
! <input.bin> (stream, little endian), written by tests/golden/make_golden.py:
!   int32 nwt
!   int32 qn(4,nwt)        isospin, parity, signature, timesimplex
!   real8 occ(nwt)
!   real8 grid(nx,ny,nz,4,nwt)
!   real8 U(n,2) B(n,2) NablaB(n,3,2) W(n,3,3,2) DivW(n,3,2)
!   real8 S(n,3,2) A(n,3,2) DivA(n,2)
! Tasks (comma separated): derive, hpsi, dens, am (UpdateAM after dens: the
! angular-momentum and spin expectation values, <prefix>_am.bin = per state
! J(3) JTR(3) JTI(3) J2(3) Sp(3) SpTR(3) SpTI(3), then TotalAngMom(3) JTR(3)
! JTI(3) J2Total(3) TotalAngMom_dens(3) AMIsoblock(2,2,2,3)), step, force (synthetic mean
! fields from the input file), pot (densities -> moments -> Coulomb ->
! potentials -> energy, reference routines), scf (the Evolve loop of
! Main.f90:255-328 for HF, with AlternateStep while alternating constraints
! are on; iteration count = niter from the 4th argument).
! With ParameterEstimation = 1 the scf task reads maxspwf's start grid from
! seed.bin (real8 grid(nx,ny,nz,4)) in place of IterativeEstimation's
! random_number start, and writes dt, momentum after each GradDesc to
! <prefix>_params.bin.  An optional energies.bin (real8 e(nwt)) gives the
! single-particle energies of the start states, fermi.bin (real8 (2)) the
! start Fermi energies, as a wavefunction file would provide them.
! With 'conv' in the tasks the scf task also writes, after the start-up's and
! every iteration's CompEnergy, the convergence tests of ConvergenceCheck
! (Main.f90:368-397) at several precisions (<prefix>_conv.bin: int32
! flags(13) = ConverEnergy at 1e-9, 1e-6, 1e-4, 1e-3, 1e-2, ConverMultipoleAll
! at 1e-4, 1e-2, 1e-1, 0.5, ConverFermi at 1e-4, 1, 10, 100).  With 'final'
! it runs the steps of FinalIteration (Main.f90:482-529) after the loop:
! Lagrange derivatives of the density basis, UpdateDensities(0),
! CompEnergy (<prefix>_final.bin: the energy row as in _scf.bin), then the
! finite-difference derivatives and UpdateDensities(0) again (the density
! rho(n,2) follows the row).
! With 'wv1' the scf task ends by writing a wavefunction file as Evolve does
! at a checkpoint (Main.f90:318-325), after the iteration's UpdateAM (:301):
! WriteMOCCa_v1 to <prefix>.v1.  With 'rv1' the states, densities, energies,
! occupations, Fermi energies (and, with HFB, the HFB matrices) come from the
! wavefunction file restart.v1 (ReadMOCCa_v1, then the HFB matrices as Input
! takes them, InOut.f90:171-183) instead of the input file's seeds, and the
! scf start-up recomputes the densities only when Recalc is set
! (Main.f90:206-215).
!-------------------------------------------------------------------------------
program ref_driver
  use CompilationInfo
  use GenInfo
  use Force
  use Derivatives
  use Spinors
  use WaveFunctions
  use SpwfStorage
  use Densities
  use MeanFields
  use ImaginaryTime
  use InOutput, only : ReadRunInfo, WriteMOCCa_v1, ReadMOCCa_v1, OutputFileName,        &
  &                    InputFileName, InputKappa, InputRho, InputU, InputV, inPC, inTRC,   &
  &                    inIC, filenwt, FileHFBColumns, FileBlocksizes
  use Transform, only : TransformHFBMatrices
  use Moments
  use Coulomb
  use Energy
  use PotentialMixing
  use Pairing, only : SolvePairing, Fermi, PairingEnergy, PairingType, Delta, ConverFermi,   &
  &                   PairingIter, FermiPrec
  use PairingInteraction, only : PairingStrength, Alpha, DDexp, PairingCut, PairingMu,   &
  &                              RhoSat, CutType, DensityFactor, PCutoffs
  use HFB, only : HFBMix, HFBIter, RhoHFB, KappaHFB, blocksizes, QuasiEnergies
  use DensityMixing, only : MixDensities
  use Cranking, only : UpdateAM, TotalAngMom, TotalAngMom_dens
  implicit none

  character(len=512) :: infile, prefix, tasks, arg4
  integer            :: nwtf, i, m, it, uu, n, niter, up, upp, uh, uc
  logical            :: altcheck, havee
  real(KIND=dp), allocatable :: estart(:)
  integer, allocatable :: qn(:,:)
  real(KIND=dp), allocatable :: occ(:), grids(:,:,:,:,:)
  type(Spinor)       :: sp, h

  call get_command_argument(1, infile)
  call get_command_argument(2, prefix)
  call get_command_argument(3, tasks)
  niter = 0
  call get_command_argument(4, arg4)
  if(len_trim(arg4).gt.0) read(arg4,*) niter

  call ReadRunInfo()

  open(newunit=uu, file=trim(infile), access='stream', form='unformatted',      &
  &    status='old')
  read(uu) nwtf
  if(nwtf .ne. nwt) stop 'nwt mismatch between namelist and input file'
  allocate(qn(4,nwt), occ(nwt), grids(nx,ny,nz,4,nwt))
  read(uu) qn
  read(uu) occ
  read(uu) grids

  allocate(HFBasis(nwt))
  do i=1,nwt
    sp = NewSpinor()
    sp%Grid(:,:,:,:,1) = grids(:,:,:,:,i)
    HFBasis(i) = NewWaveFunction(sp, qn(1,i), qn(4,i), qn(2,i), qn(3,i), 1)
    call HFBasis(i)%SetOcc(occ(i))
  enddo
  DensityBasis => HFBasis

  if(index(tasks,'rv1').gt.0) then
    InputFileName = 'restart.v1'
    call ReadMOCCa_v1(62)
    if(allocated(InputKappa)) then
      call TransformHFBMatrices(InputU, InputV, InputRho, InputKappa, inPC, inTRC, inIC, &
      &                         filenwt, FileHFBColumns, FileBlocksizes)
    endif
    DensityBasis => HFBasis
  endif

  if(index(tasks,'hpsi').gt.0 .or. index(tasks,'step').gt.0) then
    ! Mean-field arrays (allocated the way ConstructPotentials does,
    ! PotentialMixing.f90:32-54) and filled with the synthetic fields.
    allocate(BPot(nx,ny,nz,2), NablaBPot(nx,ny,nz,3,2), UPot(nx,ny,nz,2))
    allocate(WPot(nx,ny,nz,3,3,2), DivWPot(nx,ny,nz,3,2))
    allocate(SPot(nx,ny,nz,3,2), APot(nx,ny,nz,3,2), DivAPot(nx,ny,nz,2))
    read(uu) UPot, BPot, NablaBPot, WPot, DivWPot, SPot, APot, DivAPot
    ActionOfB => ActionOfBNew
  endif
  close(uu)

  n = nx*ny*nz

  if(index(tasks,'force').gt.0) then
    open(newunit=uu, file=trim(prefix)//'_force.bin', access='stream',         &
    &    form='unformatted', status='replace')
    write(uu) B1,B2,B3,B4,B5,B6,B7a,B7b,B8a,B8b,Byt3a,Byt3b,B9,B9q,B10,B11,     &
    &        B12a,B12b,B13a,B13b,B14,B15,B16,B17,B18,B19,B20,B21
    write(uu) hbm, nucleonmass, e2, hbar, dv, Neutrons, Protons
    write(uu) COM1Body
    write(uu) PairingType, CutType, HFBIter
    write(uu) PairingStrength, Alpha, DDexp, PairingCut, PairingMu, RhoSat, HFBMix
    write(uu) dble(PairingIter), FermiPrec
    close(uu)
  endif

  call DeriveAll()

  if(index(tasks,'derive').gt.0) then
    open(newunit=uu, file=trim(prefix)//'_derive.bin', access='stream',        &
    &    form='unformatted', status='replace')
    do i=1,nwt
      do m=1,3
        write(uu) HFBasis(i)%Der(m)%Grid(:,:,:,:,1)
      enddo
      write(uu) HFBasis(i)%Lap%Grid(:,:,:,:,1)
    enddo
    close(uu)
  endif

  if(index(tasks,'hpsi').gt.0) then
    open(newunit=uu, file=trim(prefix)//'_hpsi.bin', access='stream',          &
    &    form='unformatted', status='replace')
    do i=1,nwt
      h = hPsi(HFBasis(i))
      write(uu) h%Grid(:,:,:,:,1)
    enddo
    close(uu)
  endif

  if(index(tasks,'dens').gt.0) then
    call ComputeDensity(Density, .false.)
    open(newunit=uu, file=trim(prefix)//'_dens.bin', access='stream',          &
    &    form='unformatted', status='replace')
    write(uu) Density%Rho, Density%Tau, Density%NablaJ, Density%DerRho,         &
    &        Density%LapRho
    if(.not. TRC) then
      write(uu) Density%Vecj, Density%Vecs, Density%RotVecJ, Density%RotS,      &
      &        Density%DerS
    endif

  if(index(tasks,'am').gt.0) then
    ! UpdateAM (Cranking.f90:135-232) after the densities of 'dens': DiagSpin
    ! and DiagAng of every state (Spwf.f90:977-1489), the sums over the
    ! density basis and the integrals of the spin and current densities
    call UpdateAM()
    open(newunit=uu, file=trim(prefix)//'_am.bin', access='stream',            &
    &    form='unformatted', status='replace')
    do i=1,nwt
      write(uu) HFBasis(i)%J, HFBasis(i)%JTR, HFBasis(i)%JTI, HFBasis(i)%J2,     &
      &         HFBasis(i)%Sp, HFBasis(i)%SpTR, HFBasis(i)%SpTI
    enddo
    write(uu) TotalAngMom, JTR, JTI, J2Total, TotalAngMom_dens, AMIsoblock
    close(uu)
  endif
    close(uu)
  endif

  if(index(tasks,'step').gt.0) then
    ! Two heavy-ball steps with fixed dt/momentum (ParameterEstimation=0),
    ! each followed by Gram-Schmidt, as GradDesc does (ImaginaryTime.f90:439).
    open(newunit=uu, file=trim(prefix)//'_step.bin', access='stream',          &
    &    form='unformatted', status='replace')
    do it=1,2
      call GradDesc(it)
      do i=1,nwt
        write(uu) HFBasis(i)%Value%Grid(:,:,:,:,1)
      enddo
      do i=1,nwt
        write(uu) HFBasis(i)%GetEnergy(), HFBasis(i)%GetDispersion()
      enddo
      call DeriveAll()
    enddo
    close(uu)
  endif

  if(index(tasks,'pot').gt.0) then
    ! Densities of the given states (occupations as given), then the
    ! reference's own chain: moments -> Coulomb -> potentials -> energy.
    call ComputeDensity(Density, .false.)
    call CalculateAllMoments(1)
    open(newunit=uu, file=trim(prefix)//'_cutoff.bin', access='stream',        &
    &    form='unformatted', status='replace')
    write(uu) Cutoff
    close(uu)
    call ConstructPotentials
    call CompEnergy
    call dump_moments(trim(prefix)//'_moments.bin')
    call dump_fields(trim(prefix)//'_pot.bin')
    call dump_energy(trim(prefix)//'_energy.bin')
  endif

  if(index(tasks,'scf').gt.0) then
    ! Start-up and loop of Evolve (Main.f90:172-225, 255-328) for HF without
    ! constraints or cranking, ParameterEstimation = 0.
    EvolveSpwf => GradDesc
    ! Optional single-particle energies of the start states (energies.bin,
    ! real8 e(nwt)), as a wavefunction file would provide them.
    inquire(file='energies.bin', exist=havee)
    if(havee) then
      allocate(estart(nwt))
      open(newunit=up, file='energies.bin', access='stream', form='unformatted', status='old')
      read(up) estart
      close(up)
      do i=1,nwt
        call HFBasis(i)%SetEnergy(estart(i))
      enddo
    endif
    ! Optional start Fermi energies (fermi.bin, real8 Fermi(2)).
    inquire(file='fermi.bin', exist=havee)
    if(havee) then
      open(newunit=up, file='fermi.bin', access='stream', form='unformatted', status='old')
      read(up) Fermi
      close(up)
    endif
    if(ParameterEstimation.eq.1) then
      ! maxspwf as the first call of IterativeEstimation makes it
      ! (ImaginaryTime.f90:58-70), from the seed grid instead of random_number;
      ! the loop below then passes iteration numbers >= 2 so that it is kept.
      maxspwf = CopyWaveFunction(HFBasis(nwt))
      open(newunit=up, file='seed.bin', access='stream', form='unformatted',   &
      &    status='old')
      read(up) maxspwf%value%grid
      close(up)
      call maxspwf%compnorm()
      maxspwf%value = 1.0/sqrt(maxspwf%norm) * maxspwf%value
      call maxspwf%compder()
      open(newunit=up, file=trim(prefix)//'_params.bin', access='stream',      &
      &    form='unformatted', status='replace')
    endif
    call SolvePairing
    if(PairingType.eq.2) then
      ! the start-up pairing step: gaps, kappa, rho, density factor, cutoffs
      open(newunit=uh, file=trim(prefix)//'_hfbstart.bin', access='stream',    &
      &    form='unformatted', status='replace')
      write(uh) dble(Delta), dble(KappaHFB), dble(RhoHFB), DensityFactor, PCutoffs
      close(uh)
    endif
    if(PairingType.eq.1) then
      ! the start-up BCS step: gaps, cutoffs, quasiparticle energies,
      ! occupations, Fermi energies
      open(newunit=uh, file=trim(prefix)//'_bcsstart.bin', access='stream',    &
      &    form='unformatted', status='replace')
      write(uh) dble(Delta(:,1,1,1)), PCutoffs, (HFBasis(i)%GetEqp(), i=1,nwt),    &
      &         (HFBasis(i)%GetOcc(), i=1,nwt), Fermi
      close(uh)
    endif
    call DeriveAll
    if(index(tasks,'rv1').eq.0 .or. Recalc) call UpdateDensities(0)
    call CalculateAllMoments(1)
    call CalculateAllMoments(1)
    call CalculateAllMoments(1)
    call ConstructPotentials
    call CompEnergy
    open(newunit=uu, file=trim(prefix)//'_scf.bin', access='stream',            &
    &    form='unformatted', status='replace')
    write(uu) TotalEnergy, sum(Kinetic), CoulombEnergy, CoulombExchange,        &
    &        sum(CoMCorrection), Skyrmeterms
    open(newunit=upp, file=trim(prefix)//'_pairing.bin', access='stream',       &
    &    form='unformatted', status='replace')
    write(upp) Fermi, PairingEnergy
    if(index(tasks,'conv').gt.0) then
      open(newunit=uc, file=trim(prefix)//'_conv.bin', access='stream',         &
      &    form='unformatted', status='replace')
      call dump_conv(uc)
    endif
    altcheck = CheckForAlternateConstraints()
    do it=1,niter
      if(altcheck) call AlternateStep()
      call EvolveSpwf(max(it, 2))
      if(ParameterEstimation.eq.1) write(up) dt, momentum
      call SolvePairing
      altcheck = CheckForAlternateConstraints()
      call DeriveAll
      call UpdateDensities(0)
      call MixDensities(it)
      call CalculateAllMoments(1)
      call ReadjustAllMoments()
      call TurnOffConstraints(it)
      call ConstructPotentials
      call CompEnergy
      write(uu) TotalEnergy, sum(Kinetic), CoulombEnergy, CoulombExchange,      &
      &        sum(CoMCorrection), Skyrmeterms
      write(upp) Fermi, PairingEnergy
      if(index(tasks,'conv').gt.0) call dump_conv(uc)
    enddo
    close(uu)
    close(upp)
    if(index(tasks,'conv').gt.0) close(uc)
    if(ParameterEstimation.eq.1) close(up)
    open(newunit=uu, file=trim(prefix)//'_scfstates.bin', access='stream',      &
    &    form='unformatted', status='replace')
    do i=1,nwt
      write(uu) HFBasis(i)%Value%Grid(:,:,:,:,1)
    enddo
    do i=1,nwt
      write(uu) HFBasis(i)%GetEnergy(), HFBasis(i)%GetOcc()
    enddo
    close(uu)
    call dump_fields(trim(prefix)//'_scfpot.bin')
    if(index(tasks,'wv1').gt.0) then
      call UpdateAM()
      OutputFileName = trim(prefix)//'.v1'
      call WriteMOCCa_v1(61)
    endif
    if(index(tasks,'final').gt.0) then
      ! FinalIteration (Main.f90:482-529), without the printing
      call AssignFDCoefs(-1, -1, CoulombLapOrder)
      do i=1,nwt
        call DensityBasis(i)%CompDer()
      enddo
      call UpdateDensities(0)
      call CompEnergy
      open(newunit=uu, file=trim(prefix)//'_final.bin', access='stream',        &
      &    form='unformatted', status='replace')
      write(uu) TotalEnergy, sum(Kinetic), CoulombEnergy, CoulombExchange,      &
      &        sum(CoMCorrection), Skyrmeterms
      call AssignFDCoefs(MaxFDOrder, MaxFDLapOrder, CoulombLapOrder)
      do i=1,nwt
        call DensityBasis(i)%CompDer()
      enddo
      call UpdateDensities(0)
      write(uu) Density%Rho
      close(uu)
    endif
  endif

contains

  subroutine dump_conv(u)
    ! the tests of ConvergenceCheck (Main.f90:368-397) at several precisions
    integer, intent(in) :: u
    real(KIND=dp), parameter :: ep(5) = (/1d-9, 1d-6, 1d-4, 1d-3, 1d-2/)
    real(KIND=dp), parameter :: mp(4) = (/1d-4, 1d-2, 1d-1, 5d-1/)
    real(KIND=dp), parameter :: pp(4) = (/1d-4, 1d0, 1d1, 1d2/)
    integer :: fl(13), k
    do k=1,5
      fl(k) = merge(1, 0, ConverEnergy(ep(k)))
    enddo
    do k=1,4
      fl(5+k) = merge(1, 0, ConverMultipoleAll(mp(k)))
    enddo
    do k=1,4
      fl(9+k) = merge(1, 0, ConverFermi(pp(k)))
    enddo
    write(u) fl
  end subroutine dump_conv

  subroutine dump_moments(fn)
    character(len=*), intent(in) :: fn
    type(Moment), pointer :: Cur
    integer :: v, cnt
    cnt = 0
    Cur => Root
    do while(associated(Cur))
      if(Cur%l .lt. 0) exit
      cnt = cnt + 1
      Cur => Cur%Next
    enddo
    open(newunit=v, file=fn, access='stream', form='unformatted',              &
    &    status='replace')
    write(v) cnt
    Cur => Root
    do while(associated(Cur))
      if(Cur%l .lt. 0) exit
      write(v) Cur%l, Cur%m, merge(1, 0, Cur%ImPart), Cur%Value
      Cur => Cur%Next
    enddo
    close(v)
  end subroutine dump_moments

  subroutine dump_fields(fn)
    character(len=*), intent(in) :: fn
    integer :: v
    open(newunit=v, file=fn, access='stream', form='unformatted',              &
    &    status='replace')
    write(v) Density%Rho, Density%Tau, Density%NablaJ, Density%DerRho,         &
    &        Density%LapRho
    if(.not. TRC) then
      write(v) Density%Vecj, Density%Vecs, Density%RotVecJ, Density%RotS
    endif
    write(v) CoulombPotential, CoulExchange
    write(v) size(CPotComplete,1), size(CPotComplete,2), size(CPotComplete,3)
    write(v) CPotComplete
    write(v) UPot, BPot, NablaBPot, WPot, DivWPot
    if(.not. TRC) then
      write(v) SPot, APot, DivAPot
    endif
    close(v)
  end subroutine dump_fields

  subroutine dump_energy(fn)
    character(len=*), intent(in) :: fn
    integer :: v
    open(newunit=v, file=fn, access='stream', form='unformatted',              &
    &    status='replace')
    write(v) TotalEnergy, Kinetic, CoulombEnergy, CoulombExchange,             &
    &        CoMCorrection, Skyrmeterms
    close(v)
  end subroutine dump_energy

end program ref_driver
