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
! Tasks (comma separated): derive, hpsi, dens, step, force, scf
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
  use InOutput, only : ReadRunInfo
  implicit none

  character(len=512) :: infile, prefix, tasks
  integer            :: nwtf, i, m, it, u, n
  integer, allocatable :: qn(:,:)
  real(KIND=dp), allocatable :: occ(:), grids(:,:,:,:,:)
  type(Spinor)       :: sp, h

  call get_command_argument(1, infile)
  call get_command_argument(2, prefix)
  call get_command_argument(3, tasks)

  call ReadRunInfo()

  open(newunit=u, file=trim(infile), access='stream', form='unformatted',      &
  &    status='old')
  read(u) nwtf
  if(nwtf .ne. nwt) stop 'nwt mismatch between namelist and input file'
  allocate(qn(4,nwt), occ(nwt), grids(nx,ny,nz,4,nwt))
  read(u) qn
  read(u) occ
  read(u) grids

  allocate(HFBasis(nwt))
  do i=1,nwt
    sp = NewSpinor()
    sp%Grid(:,:,:,:,1) = grids(:,:,:,:,i)
    HFBasis(i) = NewWaveFunction(sp, qn(1,i), qn(4,i), qn(2,i), qn(3,i), 1)
    call HFBasis(i)%SetOcc(occ(i))
  enddo
  DensityBasis => HFBasis

  ! Mean-field arrays (allocated the way ConstructPotentials does,
  ! PotentialMixing.f90:32-54) and filled with the synthetic fields.
  allocate(BPot(nx,ny,nz,2), NablaBPot(nx,ny,nz,3,2), UPot(nx,ny,nz,2))
  allocate(WPot(nx,ny,nz,3,3,2), DivWPot(nx,ny,nz,3,2))
  allocate(SPot(nx,ny,nz,3,2), APot(nx,ny,nz,3,2), DivAPot(nx,ny,nz,2))
  read(u) UPot, BPot, NablaBPot, WPot, DivWPot, SPot, APot, DivAPot
  close(u)
  ActionOfB => ActionOfBNew

  n = nx*ny*nz

  if(index(tasks,'force').gt.0) then
    open(newunit=u, file=trim(prefix)//'_force.bin', access='stream',         &
    &    form='unformatted', status='replace')
    write(u) B1,B2,B3,B4,B5,B6,B7a,B7b,B8a,B8b,Byt3a,Byt3b,B9,B9q,B10,B11,     &
    &        B12a,B12b,B13a,B13b,B14,B15,B16,B17,B18,B19,B20,B21
    write(u) hbm, nucleonmass, e2, hbar, dv, Neutrons, Protons
    write(u) COM1Body
    close(u)
  endif

  call DeriveAll()

  if(index(tasks,'derive').gt.0) then
    open(newunit=u, file=trim(prefix)//'_derive.bin', access='stream',        &
    &    form='unformatted', status='replace')
    do i=1,nwt
      do m=1,3
        write(u) HFBasis(i)%Der(m)%Grid(:,:,:,:,1)
      enddo
      write(u) HFBasis(i)%Lap%Grid(:,:,:,:,1)
    enddo
    close(u)
  endif

  if(index(tasks,'hpsi').gt.0) then
    open(newunit=u, file=trim(prefix)//'_hpsi.bin', access='stream',          &
    &    form='unformatted', status='replace')
    do i=1,nwt
      h = hPsi(HFBasis(i))
      write(u) h%Grid(:,:,:,:,1)
    enddo
    close(u)
  endif

  if(index(tasks,'dens').gt.0) then
    call ComputeDensity(Density, .false.)
    open(newunit=u, file=trim(prefix)//'_dens.bin', access='stream',          &
    &    form='unformatted', status='replace')
    write(u) Density%Rho, Density%Tau, Density%NablaJ, Density%DerRho,         &
    &        Density%LapRho
    if(.not. TRC) then
      write(u) Density%Vecj, Density%Vecs, Density%RotVecJ, Density%RotS,      &
      &        Density%DerS
    endif
    close(u)
  endif

  if(index(tasks,'step').gt.0) then
    ! Two heavy-ball steps with fixed dt/momentum (ParameterEstimation=0),
    ! each followed by Gram-Schmidt, as GradDesc does (ImaginaryTime.f90:439).
    open(newunit=u, file=trim(prefix)//'_step.bin', access='stream',          &
    &    form='unformatted', status='replace')
    do it=1,2
      call GradDesc(it)
      do i=1,nwt
        write(u) HFBasis(i)%Value%Grid(:,:,:,:,1)
      enddo
      do i=1,nwt
        write(u) HFBasis(i)%GetEnergy(), HFBasis(i)%GetDispersion()
      enddo
      call DeriveAll()
    enddo
    close(u)
  endif
end program ref_driver
