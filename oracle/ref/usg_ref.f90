! Reference-side driver of the random walk on MODFLOW 6 DISV grids of
! rectangular cells: test infrastructure only (a checker, never part of the
! product).
!
! The grid is the reference's own RectangularGridDisvMf6 read from a binary
! grid file (ReadGRB, CheckRectangular, ComputeFaceAssignments); the flow
! and transport arrays FlowModelData / TransportModelData hold are set from
! a binary input file instead of head and budget files.
!
! Mode "tables": for every cell, FillCellBuffer + FillNeighborCellData
! (pr_FillNeighborCellDataUnstructured, ParticleTrackingEngine.f90:1257-2005)
! and, for every sub-cell, FillMassSubCellDataBuffer,
! TrackSubCell%ComputeCornerVariables and ComputeRandomWalkTimeStep; the
! sub-cell data and the corner variables are written out.
!
! Mode "walk": the particle loop of MPathRW.F90:1239-1403 with one OpenMP
! thread, and so one rng stream, per particle (as oracle/ref/rwpt_ref.f90),
! the threads taking turns,
! and pr_RWPTrackPath's cell loop (ParticleTrackingEngine.f90:1051-1254)
! written out here around the reference's own TrackCell, neighbour fill and
! ConvertFromNeighbor, with one change: every one of the 2 x 18 neighbour
! buffers is Reset before each neighbour fill.  The reference's own reset
! loop (:1269-1273) skips buffers whose `initialized` flag is false, which
! is the case of every buffer filled from a sub-cell or from a refined cell,
! so that on refined grids its buffers carry values from cells visited
! earlier by the same thread.
!
! usage: usg_ref GRB IN OUT walk|tables
!   IN (stream): int32 hdr(20): np nout advection tsk dimmask(3) model nd
!     defaultICBound backward uniformPorosity stopWeakSinks stopWeakSources
!     stopZone steady (the rest 0); real64 tsp(2);
!     real64 flowja(nja), bflow(6 ncell), storage, source, sink, porosity,
!     retardation, head (ncell); int32 ibound, icbound, zone (ncell);
!     real64 aL aT aLV aTV dmeff (nd each);
!     int32 cell(np); real64 lx ly lz t (np); int32 status(np); real64 tout(nout)
!   OUT walk (stream): per output time int32 cell(np), status(np);
!     real64 lx ly lz t (np)
!   OUT tables (stream): int32 ncell; per cell int32 subcount (1 or 4), then
!     per sub-cell (row-major) int32 row col conn(6) massbnd(6) hasexit
!     cellhasexit dry partiallydry, real64 VX1 VX2 VY1 VY2 VZ1 VZ2 DX DY DZ
!     offx(2) offy(2) porosity retardation dtcell source sink,
!     qcorner(4,8), porosity corners(8), Dxxx Dyyy Dzzz Dxyx Dxzx Dyzy (8 each)
program usg_ref
  use iso_fortran_env, only: int32, real64, error_unit
  use omp_lib
  use ModflowRectangularGridModule, only: ModflowRectangularGridType
  use RectangularGridDisvMf6Module, only: RectangularGridDisvMf6Type
  use FlowModelDataModule, only: FlowModelDataType
  use TransportModelDataModule, only: TransportModelDataType
  use ModpathSimulationDataModule, only: ModpathSimulationDataType
  use DispersionDataModule, only: DispersionDataType
  use ParticleTrackingOptionsModule, only: ParticleTrackingOptionsType
  use ParticleTrackingEngineModule, only: ParticleTrackingEngineType
  use TrackCellResultModule, only: TrackCellResultType
  use ParticleLocationModule, only: ParticleLocationType
  use ModpathCellDataModule, only: ModpathCellDataType
  implicit none

  integer(int32) :: hdr(20)
  integer :: ncell, nja, np, nout, nd, k, p, kt, dcount, n, r, c, m
  real(real64) :: tsp(2), time, maxTime, dtcell
  real(real64), allocatable :: tout(:), lx(:), ly(:), lz(:), tt(:)
  integer(int32), allocatable :: cell(:), status(:), ibuf(:)
  character(len=1024) :: fgrb, fin, fout, mode
  logical :: teamOk

  class(ModflowRectangularGridType), pointer :: grid
  type(RectangularGridDisvMf6Type), pointer :: disv
  type(FlowModelDataType), target :: flow
  type(TransportModelDataType), target :: transport
  type(ModpathSimulationDataType), target :: sim
  type(DispersionDataType) :: dispersion
  type(ParticleTrackingOptionsType) :: options
  type(ParticleTrackingEngineType), target :: engine
  type(ModpathCellDataType) :: cbuf
  integer :: nsub, hx, chx, dr, pdr

  call get_command_argument(1, fgrb)
  call get_command_argument(2, fin)
  call get_command_argument(3, fout)
  call get_command_argument(4, mode)

  allocate(disv)
  open(unit=20, file=trim(fgrb), form='unformatted', access='stream', status='old', &
       action='read')
  call disv%ReadData(20, 0, 6)
  close(20)
  grid => disv
  ncell = grid%CellCount
  nja = grid%JaCount

  open(10, file=trim(fin), access='stream', form='unformatted', status='old')
  read(10) hdr
  np = hdr(1); nout = hdr(2); nd = hdr(9)
  read(10) tsp
  allocate(flow%FlowsJA(nja), flow%BoundaryFlows(6 * ncell), flow%StorageFlows(ncell))
  allocate(flow%SourceFlows(ncell), flow%SinkFlows(ncell), flow%Porosity(ncell))
  allocate(flow%Retardation(ncell), flow%Heads(ncell), flow%IBound(ncell))
  allocate(flow%IBoundTS(ncell), flow%Zones(ncell), ibuf(ncell))
  read(10) flow%FlowsJA, flow%BoundaryFlows, flow%StorageFlows, flow%SourceFlows
  read(10) flow%SinkFlows, flow%Porosity, flow%Retardation, flow%Heads
  read(10) ibuf
  flow%IBound = ibuf
  flow%IBoundTS = ibuf
  allocate(transport%ICBoundTS(ncell))
  read(10) ibuf
  transport%ICBoundTS = ibuf
  read(10) flow%Zones
  flow%SteadyState = hdr(16) .ne. 0
  transport%defaultICBound = hdr(10)
  allocate(transport%AlphaL(nd), transport%AlphaT(nd), transport%AlphaLV(nd))
  allocate(transport%AlphaTV(nd), transport%DMEff(nd))
  read(10) transport%AlphaL, transport%AlphaT, transport%AlphaLV, transport%AlphaTV, &
       transport%DMEff
  transport%AlphaLH => transport%AlphaL
  transport%AlphaTH => transport%AlphaT
  sim%SolutesOption = 1
  transport%simulationData => sim
  dispersion%modelKind = hdr(8)
  dispersion%uniformParameters = nd .eq. 1

  options%RandomWalkParticleTracking = .true.
  options%advectionKind = hdr(3)
  options%timeStepKind = hdr(4)
  options%timeStepParameters = tsp
  options%dimensionMask = hdr(5:7)
  options%nDim = sum(options%dimensionMask)
  allocate(options%dimensions(options%nDim))
  dcount = 0
  do k = 1, 3
    if (options%dimensionMask(k) .eq. 0) cycle
    dcount = dcount + 1
    options%dimensions(dcount) = k
  end do
  options%randomGenFunction = 2
  options%BackwardTracking = hdr(11) .ne. 0
  options%isUniformPorosity = hdr(12) .ne. 0
  options%StopAtWeakSinks = hdr(13) .ne. 0
  options%StopAtWeakSources = hdr(14) .ne. 0
  options%StopZone = hdr(15)

  allocate(cell(np), status(np), lx(np), ly(np), lz(np), tt(np), tout(nout))
  read(10) cell, lx, ly, lz, tt, status, tout
  close(10)

  call omp_set_dynamic(.false.)
  call omp_set_num_threads(max(np, 1))
  call engine%Initialize(grid, options, flow, transport)
  call engine%UpdateDispersionFunction(dispersion)
  engine%TrackCell%SteadyState = flow%SteadyState
  engine%TrackCell%TrackingOptions = engine%TrackingOptions

  open(11, file=trim(fout), access='stream', form='unformatted', status='replace')

  if (trim(mode) .eq. 'tables') then
    write(11) int(ncell, int32)
    do n = 1, ncell
      call engine%FillCellBuffer(n, engine%TrackCell%CellData)
      call reset_neighbours(engine)
      call engine%FillNeighborCellData(engine%NeighborCellData)
      nsub = engine%TrackCell%CellData%GetSubCellCount()
      write(11) int(nsub, int32)
      chx = 0
      if (engine%TrackCell%CellData%HasExitFace(options%BackwardTracking)) chx = 1
      do r = 1, engine%TrackCell%CellData%GetSubCellRowCount()
        do c = 1, engine%TrackCell%CellData%GetSubCellColumnCount()
          call engine%TrackCell%CellData%FillMassSubCellDataBuffer( &
               engine%TrackCell%TrackSubCell%SubCellData, r, c, options%BackwardTracking)
          call engine%TrackCell%TrackSubCell%ComputeCornerVariables( &
               engine%TrackCell%CellData, engine%NeighborCellData)
          call engine%TrackCell%TrackSubCell%ComputeRandomWalkTimeStep(options, dtcell)
          associate (s => engine%TrackCell%TrackSubCell%SubCellData, &
                     t => engine%TrackCell%TrackSubCell)
            hx = 0
            if (s%HasExitFace()) hx = 1
            dr = 0
            if (s%dry) dr = 1
            pdr = 0
            if (s%partiallyDry) pdr = 1
            write(11) int([r, c], int32), int(s%Connection, int32), &
                 int(s%MassBoundary, int32), int([hx, chx, dr, pdr], int32)
            write(11) s%VX1, s%VX2, s%VY1, s%VY2, s%VZ1, s%VZ2, s%DX, s%DY, s%DZ, &
                 s%OffsetX, s%OffsetY, s%Porosity, s%Retardation, dtcell, &
                 engine%TrackCell%CellData%SourceFlow, engine%TrackCell%CellData%SinkFlow
            write(11) t%qCorner000, t%qCorner100, t%qCorner010, t%qCorner110, &
                 t%qCorner001, t%qCorner101, t%qCorner011, t%qCorner111
            write(11) t%porosity000, t%porosity100, t%porosity010, t%porosity110, &
                 t%porosity001, t%porosity101, t%porosity011, t%porosity111
            write(11) t%Dxxx, t%Dyyy, t%Dzzz, t%Dxyx, t%Dxzx, t%Dyzy
          end associate
        end do
      end do
    end do
    close(11)
    stop
  end if

  time = 0d0
  teamOk = .true.
  do kt = 1, nout
    maxTime = tout(kt)
    !$omp parallel do num_threads(np) schedule(static, 1) default(shared) &
    !$omp private(p, cbuf) firstprivate(engine)
    do p = 1, np
      if (omp_get_num_threads() .ne. np) teamOk = .false.
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
      ! one particle at a time: the reference keeps implicitly saved
      ! locals that every thread shares (skipSubCells in
      ! pr_FillNeighborSubCellVariables, TrackSubCell.F90:3873, and
      ! sumFlowRate in pr_GetInterpolatedCornerDischarge, :4093), a race
      ! once threads track refined and unrefined cells at the same time
      !$omp critical (usg_track)
      if (status(p) .eq. 1) then
        call track_path(engine, cell(p), lx(p), ly(p), lz(p), tt(p), maxTime, status(p))
      end if
      !$omp end critical (usg_track)
    end do
    !$omp end parallel do
    if (.not. teamOk) then
      write(error_unit, *) 'usg_ref: OpenMP team smaller than the particle count'
      stop 3
    end if
    write(11) cell, status, lx, ly, lz, tt
    time = maxTime
  end do
  close(11)

contains

  subroutine reset_neighbours(e)
    type(ParticleTrackingEngineType), intent(inout) :: e
    integer :: i, j
    do j = 1, 18
      do i = 1, 2
        call e%NeighborCellData(i, j)%Reset()
      end do
    end do
  end subroutine reset_neighbours

  ! pr_RWPTrackPath's loop (ParticleTrackingEngine.f90:1051-1254) for one
  ! stopping time without time-series points, with the neighbour buffers
  ! reset before each fill; writes back the last pathline point and the
  ! particle status of MPathRW.F90:1337-1403.
  subroutine track_path(e, pc, px, py, pz, pt, tmax, st)
    type(ParticleTrackingEngineType), target, intent(inout) :: e
    integer(int32), intent(inout) :: pc, st
    real(real64), intent(inout) :: px, py, pz, pt
    real(real64), intent(in) :: tmax
    type(ParticleLocationType) :: loc, last
    type(TrackCellResultType) :: res
    integer :: cnt, nextCell
    real(real64) :: fx, fy, fz

    loc%CellNumber = pc
    loc%Layer = e%Grid%GetLayer(pc)
    loc%LocalX = px
    loc%LocalY = py
    loc%LocalZ = pz
    loc%TrackingTime = pt
    last = loc
    e%TrackCell%SteadyState = e%FlowModelData%SteadyState
    e%TrackCell%TrackingOptions = e%TrackingOptions
    call e%FillCellBuffer(loc%CellNumber, e%TrackCell%CellData)
    call reset_neighbours(e)
    call e%FillNeighborCellData(e%NeighborCellData)
    do
      if (loc%CellNumber .ne. e%TrackCell%CellData%CellNumber) then
        call e%FillCellBuffer(loc%CellNumber, e%TrackCell%CellData)
        call reset_neighbours(e)
        call e%FillNeighborCellData(e%NeighborCellData)
      end if
      call e%TrackCell%ExecuteRandomWalkParticleTracking(loc, tmax, res, e%NeighborCellData)
      if (res%Status .eq. res%Status_Undefined()) then
        st = 9
        exit
      else if (res%Status .eq. res%Status_ExitAtCellFace()) then
        cnt = res%TrackingPoints%GetItemCount()
        if (cnt .gt. 1) last = res%TrackingPoints%Items(cnt)
        nextCell = res%NextCellNumber
        if (nextCell .gt. 0) then
          if (e%FlowModelData%IBoundTS(nextCell) .ne. 0) then
            fx = res%TrackingPoints%Items(cnt)%LocalX
            fy = res%TrackingPoints%Items(cnt)%LocalY
            fz = res%TrackingPoints%Items(cnt)%LocalZ
            call e%Grid%ConvertFromNeighbor(nextCell, res%CellNumber, fx, fy, fz, loc)
            loc%TrackingTime = res%TrackingPoints%Items(cnt)%TrackingTime
            ! a failed conversion leaves no cell: the oracle's status 9
            if (loc%CellNumber .le. 0) then
              st = 9
              exit
            end if
          else
            st = 2
            exit
          end if
        else
          st = 2
          exit
        end if
      else if (res%Status .eq. res%Status_ReachedStoppingTime()) then
        cnt = res%TrackingPoints%GetItemCount()
        if (cnt .gt. 1) last = res%TrackingPoints%Items(cnt)
        st = 1
        exit
      else
        if (res%Status .eq. res%Status_NoExitPossible()) then
          st = 5
        else if (res%Status .eq. res%Status_StopZoneCell()) then
          st = 6
        else if (res%Status .eq. res%Status_StopAtWeakSink()) then
          st = 3
        else if (res%Status .eq. res%Status_StopAtWeakSource()) then
          st = 4
        else if (res%Status .eq. res%Status_InactiveCell()) then
          st = 7
        else
          st = 9
        end if
        if (st .ne. 9) then
          if (res%TrackingPoints%GetItemCount() .ge. 1) then
            last = res%TrackingPoints%Items(1)
          else
            last = loc
          end if
        end if
        exit
      end if
    end do
    pc = last%CellNumber
    px = last%LocalX
    py = last%LocalY
    pz = last%LocalZ
    pt = last%TrackingTime
  end subroutine track_path

end program usg_ref
