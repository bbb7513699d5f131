! Reference-side driver of the gpkde reconstruction: test infrastructure only.
!
! Calls the reference's GridProjectedKDEType%Initialize and %ComputeDensity
! (GridProjectedKDE.F90:354, :3598) the way MPathRW.F90:556-574 and
! :1626-1642 do, on points and weights read from a binary file, and writes
! the density grid and the histogram counts.
!
! usage: gpkde_ref <input> <output>
!   input  (stream): int32 hdr(14) = n, nbins(3), nOptLoops, database,
!          isotropic, weighted, unitVolume, initialSmoothingSelection,
!          exportOptimizationVariables (files gpkde.out<loop> in the working
!          directory), skipErrorConvergence, boundKernelSizeFormat,
!          effectiveWeightFormat, exportFormat (-1 none, 0 text, 1 binary
!          stream: the density rows go to file 'export.out' in the working
!          directory), outputColumnFormat, outputDataId, particleGroupId;
!          real64 origin(3), binSize(3), scalingFactor, smoothingFactor,
!          minHOverDelta, deltaHOverDelta, maxHOverDelta (passed when
!          minHOverDelta > 0), outputDataIdVal;
!          real64 x(n), y(n), z(n), w(n)
!   output (stream): real64 density(nx*ny*nz), counts(nx*ny*nz),
!          initialSmoothing(3)
program gpkde_ref
  use iso_fortran_env, only: int32, real64
  use GridProjectedKDEModule, only: GridProjectedKDEType
  implicit none
  integer(int32) :: hdr(18)
  real(real64) :: origin(3), binSize(3), domainSize(3), sf, smf, hod(3), tval
  real(real64), allocatable :: pts(:, :), w(:)
  character(len=1024) :: fin, fout
  type(GridProjectedKDEType), allocatable :: kde
  integer :: n, nb(3)

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(10, file=trim(fin), access='stream', form='unformatted', status='old')
  read(10) hdr, origin, binSize, sf, smf, hod, tval
  n = hdr(1)
  nb = hdr(2:4)
  allocate(pts(n, 3), w(n))
  read(10) pts(:, 1), pts(:, 2), pts(:, 3), w
  close(10)
  domainSize = nb * binSize

  allocate(kde)
  if (hod(1) > 0d0) then
    call kde%Initialize(domainSize, binSize, domainOrigin=origin,             &
      adaptGridToCoords=.false., nOptimizationLoops=int(hdr(5)),               &
      databaseOptimization=hdr(6) .ne. 0,                                      &
      minHOverDelta=hod(1), deltaHOverDelta=hod(2), maxHOverDelta=hod(3),      &
      initialSmoothingSelection=int(hdr(10)), initialSmoothingFactor=smf,      &
      effectiveWeightFormat=int(hdr(14)),                                      &
      boundKernelSizeFormat=int(hdr(13)), outFileName='/dev/null')
  else
    call kde%Initialize(domainSize, binSize, domainOrigin=origin,             &
      adaptGridToCoords=.false., nOptimizationLoops=int(hdr(5)),               &
      databaseOptimization=hdr(6) .ne. 0,                                      &
      initialSmoothingSelection=int(hdr(10)), initialSmoothingFactor=smf,      &
      effectiveWeightFormat=int(hdr(14)),                                      &
      boundKernelSizeFormat=int(hdr(13)), outFileName='/dev/null')
  end if
  if (hdr(15) .ge. 0) then
    if (hdr(15) .eq. 0) then
      open(12, file='export.out', status='replace', form='formatted',         &
        access='sequential')
    else
      open(12, file='export.out', status='replace', form='unformatted',       &
        access='stream')
    end if
    if (hdr(8) .ne. 0) then
      call kde%ComputeDensity(pts, unitVolume=hdr(9) .ne. 0,                  &
        weightedHistogram=.true., weights=w, scalingFactor=sf,                 &
        histogramScalingFactor=sf, isotropic=hdr(7) .ne. 0,                    &
        skipErrorConvergence=hdr(12) .ne. 0, relativeErrorConvergence=0.02d0,  &
        outputFileUnit=12, outputDataId=int(hdr(17)), outputDataIdVal=tval,    &
        particleGroupId=int(hdr(18)), outputColumnFormat=int(hdr(16)),         &
        outputDataFormat=int(hdr(15)))
    else
      call kde%ComputeDensity(pts, unitVolume=hdr(9) .ne. 0, scalingFactor=sf,&
        histogramScalingFactor=sf, isotropic=hdr(7) .ne. 0,                    &
        skipErrorConvergence=hdr(12) .ne. 0, relativeErrorConvergence=0.02d0,  &
        outputFileUnit=12, outputDataId=int(hdr(17)), outputDataIdVal=tval,    &
        particleGroupId=int(hdr(18)), outputColumnFormat=int(hdr(16)),         &
        outputDataFormat=int(hdr(15)))
    end if
    close(12)
  else if (hdr(8) .ne. 0) then
    call kde%ComputeDensity(pts, unitVolume=hdr(9) .ne. 0,                    &
      weightedHistogram=.true., weights=w, scalingFactor=sf,                   &
      histogramScalingFactor=sf, isotropic=hdr(7) .ne. 0,                      &
      exportOptimizationVariables=hdr(11) .ne. 0,                              &
      skipErrorConvergence=hdr(12) .ne. 0, relativeErrorConvergence=0.02d0)
  else
    call kde%ComputeDensity(pts, unitVolume=hdr(9) .ne. 0, scalingFactor=sf,  &
      histogramScalingFactor=sf, isotropic=hdr(7) .ne. 0,                      &
      exportOptimizationVariables=hdr(11) .ne. 0,                              &
      skipErrorConvergence=hdr(12) .ne. 0, relativeErrorConvergence=0.02d0)
  end if
  open(11, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(11) kde%densityEstimateGrid, kde%histogram%counts, kde%initialSmoothing
  close(11)
end program gpkde_ref
