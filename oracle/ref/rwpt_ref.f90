! Reference-side driver of the random-walk path: test infrastructure only.
!
! Builds the reference's own objects (RectangularGridDisType, FlowModelData,
! TransportModelData, ParticleTrackingOptions, ParticleTrackingEngine) from a
! binary input file instead of MODFLOW files, then runs the particle loop of
! MPathRW.F90:1239-1403 (release checks, dry-cell check, TrackPath, status
! mapping) for each output time.  Every particle is tracked by its own
! OpenMP thread with schedule(static,1), so rng_par_zig stream k (the k-th
! 2^64 jump of the base seed, set up by TrackSubCell.F90:609-620 for
! omp_get_max_threads() streams) belongs to particle k+1 -- the per-particle
! streams the device path uses.
!
! usage: rwpt_ref <input> <output>
!   input  (stream, native endian): int32 hdr(20); real64 tsp(2);
!          real64 delr(ncol), delc(nrow), top, bot, frf, fff, flf (ncell),
!          bflow(6*ncell), porosity, retardation, head, source, sink (ncell);
!          int32 ibound, icbound, zone (ncell);
!          real64 aL, aT, aLV, aTV, dmeff (nd each);
!          int32 cell(np); real64 lx, ly, lz, t (np); int32 status(np);
!          real64 tout(nout)
!   output (stream): per output time: int32 cell(np), status(np);
!          real64 lx, ly, lz, t (np)
program rwpt_ref
  use iso_fortran_env, only: int32, real64, error_unit
  use omp_lib
  use ModflowRectangularGridModule, only: ModflowRectangularGridType
  use RectangularGridDisModule, only: RectangularGridDisType
  use FlowModelDataModule, only: FlowModelDataType
  use TransportModelDataModule, only: TransportModelDataType
  use ModpathSimulationDataModule, only: ModpathSimulationDataType
  use DispersionDataModule, only: DispersionDataType
  use ParticleTrackingOptionsModule, only: ParticleTrackingOptionsType
  use ParticleTrackingEngineModule, only: ParticleTrackingEngineType
  use TrackPathResultModule, only: TrackPathResultType
  use ParticleLocationModule, only: ParticleLocationType
  use ParticleCoordinateModule, only: ParticleCoordinateType
  use ModpathCellDataModule, only: ModpathCellDataType
  implicit none

  integer(int32) :: hdr(20)
  integer :: ncol, nrow, nlay, ncell, np, nout, nd, n, i, j, k, p, kt, nlc, dcount
  real(real64) :: tsp(2)
  real(real64), allocatable :: delr(:), delc(:), xf(:), yf(:), tout(:)
  real(real64), allocatable :: lx(:), ly(:), lz(:), tt(:)
  integer(int32), allocatable :: cell(:), status(:), ibuf(:)
  real(real64) :: time, maxTime, tPoint(1)
  character(len=1024) :: fin, fout, targ
  logical :: teamOk
  integer :: nteam
  real(real64) :: wall0

  class(ModflowRectangularGridType), pointer :: grid
  type(RectangularGridDisType), pointer :: dis
  type(FlowModelDataType), target :: flow
  type(TransportModelDataType), target :: transport
  type(ModpathSimulationDataType), target :: sim
  type(DispersionDataType) :: dispersion
  type(ParticleTrackingOptionsType) :: options
  type(ParticleTrackingEngineType), target :: engine
  type(TrackPathResultType), target :: result
  type(ParticleLocationType) :: loc
  type(ParticleCoordinateType), pointer :: last
  type(ModpathCellDataType) :: cbuf
  integer :: plCount, st

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  ! optional third argument: timing mode with that many OpenMP threads,
  ! one rng stream per thread as MPathRW runs it (draws differ from the
  ! per-particle streams of the default mode); the elapsed tracking time
  ! is printed as "seconds <s>"
  nteam = 0
  if (command_argument_count() >= 3) then
    call get_command_argument(3, targ)
    read(targ, *) nteam
  end if
  open(10, file=trim(fin), access='stream', form='unformatted', status='old')
  read(10) hdr
  ncol = hdr(1); nrow = hdr(2); nlay = hdr(3); np = hdr(4); nout = hdr(5)
  nd = hdr(12)
  ncell = ncol * nrow * nlay
  nlc = ncol * nrow
  read(10) tsp

  ! grid: the arrays RectangularGridDis.f90 derives from a DIS file
  allocate(dis)
  allocate(delr(ncol), delc(nrow), xf(ncol + 1), yf(nrow + 1))
  read(10) delr, delc
  dis%GridType = 1
  dis%LayerCount = nlay
  dis%RowCount = nrow
  dis%ColumnCount = ncol
  dis%CellCount = ncell
  dis%LayerCellCount = nlc
  allocate(dis%Top(ncell), dis%Bottom(ncell), dis%SaturatedTop(ncell))
  allocate(dis%CellX(ncell), dis%CellY(ncell), dis%DelX(ncell), dis%DelY(ncell))
  allocate(dis%CellType(ncell), dis%LayerOffsets(nlay + 1))
  do k = 0, nlay
    dis%LayerOffsets(k + 1) = k * nlc
  end do
  dis%OriginX = 0d0
  dis%OriginY = 0d0
  dis%RotationAngle = 0d0
  xf(1) = 0d0
  do j = 1, ncol
    xf(j + 1) = xf(j) + delr(j)
  end do
  yf(1) = 0d0
  do i = 1, nrow
    yf(1) = yf(1) + delc(i)
  end do
  do i = 1, nrow
    yf(i + 1) = yf(i) - delc(i)
  end do
  read(10) dis%Top, dis%Bottom
  dis%SaturatedTop = dis%Top
  dis%CellType = 0
  n = 0
  do k = 1, nlay
    do i = 1, nrow
      do j = 1, ncol
        n = n + 1
        dis%DelX(n) = delr(j)
        dis%DelY(n) = delc(i)
        dis%CellX(n) = (xf(j) + xf(j + 1)) / 2.0
        dis%CellY(n) = (yf(i) + yf(i + 1)) / 2.0
      end do
    end do
  end do
  call dis%ComputeFaceAssignments()
  grid => dis

  ! flow model data: the arrays FlowModelData.f90 loads from budget files
  allocate(flow%FlowsRightFace(ncell), flow%FlowsFrontFace(ncell), flow%FlowsLowerFace(ncell))
  allocate(flow%BoundaryFlows(6 * ncell), flow%Porosity(ncell), flow%Retardation(ncell))
  allocate(flow%Heads(ncell), flow%SourceFlows(ncell), flow%SinkFlows(ncell))
  allocate(flow%StorageFlows(ncell), flow%IBound(ncell), flow%IBoundTS(ncell), flow%Zones(ncell))
  read(10) flow%FlowsRightFace, flow%FlowsFrontFace, flow%FlowsLowerFace, flow%BoundaryFlows
  read(10) flow%Porosity, flow%Retardation, flow%Heads, flow%SourceFlows, flow%SinkFlows
  flow%StorageFlows = 0d0
  allocate(ibuf(ncell))
  read(10) ibuf
  flow%IBound = ibuf
  flow%IBoundTS = ibuf
  flow%SteadyState = hdr(20) .ne. 0
  allocate(transport%ICBoundTS(ncell))
  read(10) ibuf
  transport%ICBoundTS = ibuf
  read(10) flow%Zones
  transport%defaultICBound = hdr(13)

  ! dispersion parameters (TransportModelData.f90 pointers)
  allocate(transport%AlphaL(nd), transport%AlphaT(nd), transport%AlphaLV(nd))
  allocate(transport%AlphaTV(nd), transport%DMEff(nd))
  read(10) transport%AlphaL, transport%AlphaT, transport%AlphaLV, transport%AlphaTV, transport%DMEff
  transport%AlphaLH => transport%AlphaL
  transport%AlphaTH => transport%AlphaT
  sim%SolutesOption = 1
  transport%simulationData => sim
  dispersion%modelKind = hdr(11)
  dispersion%uniformParameters = nd .eq. 1

  ! tracking options (ModpathSimulationData.f90:1770-2021)
  options%RandomWalkParticleTracking = .true.
  options%advectionKind = hdr(6)
  options%timeStepKind = hdr(7)
  options%timeStepParameters = tsp
  options%dimensionMask = hdr(8:10)
  options%nDim = sum(options%dimensionMask)
  allocate(options%dimensions(options%nDim))
  dcount = 0
  do k = 1, 3
    if (options%dimensionMask(k) .eq. 0) cycle
    dcount = dcount + 1
    options%dimensions(dcount) = k
  end do
  options%randomGenFunction = 2
  options%BackwardTracking = hdr(14) .ne. 0
  options%isUniformPorosity = hdr(15) .ne. 0
  options%StopAtWeakSinks = hdr(16) .ne. 0
  options%StopAtWeakSources = hdr(17) .ne. 0
  options%StopZone = hdr(18)

  ! particles and output times
  allocate(cell(np), status(np), lx(np), ly(np), lz(np), tt(np), tout(nout))
  read(10) cell, lx, ly, lz, tt, status, tout
  close(10)

  ! one rng stream per particle: omp_get_max_threads() streams are made by
  ! TrackSubCell%InitializeRandomWalk during Initialize
  if (nteam < 1) nteam = np
  call omp_set_dynamic(.false.)
  call omp_set_num_threads(nteam)
  call engine%Initialize(grid, options, flow, transport)
  call engine%UpdateDispersionFunction(dispersion)

  open(11, file=trim(fout), access='stream', form='unformatted', status='replace')
  time = 0d0
  teamOk = .true.
  wall0 = omp_get_wtime()
  do kt = 1, nout
    maxTime = tout(kt)
    tPoint(1) = maxTime
    !$omp parallel do num_threads(nteam) schedule(static, 1) default(shared) &
    !$omp private(p, loc, result, last, plCount, st, cbuf) firstprivate(engine)
    do p = 1, np
      if (omp_get_num_threads() .ne. nteam) teamOk = .false.
      if (status(p) .eq. 0) then
        if (tt(p) .lt. time) then
          status(p) = 8
        else if (tt(p) .le. maxTime) then
          status(p) = 1
          if (engine%FlowModelData%IBoundTS(cell(p)) .eq. 0) status(p) = 7
        end if
      end if
      if (status(p) .eq. 7) then
        call engine%FillCellBuffer(cell(p), cbuf)
        call cbuf%VerifyDryCell()
        if (cbuf%partiallyDry) status(p) = 1
      end if
      if (status(p) .eq. 1) then
        loc%CellNumber = cell(p)
        loc%Layer = (cell(p) - 1) / nlc + 1
        loc%LocalX = lx(p)
        loc%LocalY = ly(p)
        loc%LocalZ = lz(p)
        loc%TrackingTime = tt(p)
        call engine%TrackPath(result, .false., 0, 1, p, p, loc, maxTime, tPoint, 1)
        plCount = result%ParticlePath%Pathline%GetItemCount()
        last => result%ParticlePath%Pathline%Items(plCount)
        cell(p) = last%CellNumber
        lx(p) = last%LocalX
        ly(p) = last%LocalY
        lz(p) = last%LocalZ
        tt(p) = last%TrackingTime
        st = result%Status
        if (st .eq. result%Status_ReachedBoundaryFace()) then
          status(p) = 2
        else if (st .eq. result%Status_StopAtWeakSink()) then
          status(p) = 3
        else if (st .eq. result%Status_StopAtWeakSource()) then
          status(p) = 4
        else if (st .eq. result%Status_NoExitPossible()) then
          status(p) = 5
        else if (st .eq. result%Status_StopZoneCell()) then
          status(p) = 6
        else if (st .eq. result%Status_InactiveCell()) then
          status(p) = 7
        else if (st .eq. result%Status_Undefined()) then
          status(p) = 9
        end if
      end if
    end do
    !$omp end parallel do
    if (.not. teamOk) then
      write(error_unit, *) 'rwpt_ref: OpenMP team smaller than the particle count'
      stop 3
    end if
    write(11) cell, status, lx, ly, lz, tt
    time = maxTime
  end do
  close(11)
  write(*, '(A, ES16.8)') 'seconds ', omp_get_wtime() - wall0
end program rwpt_ref
