! Test infrastructure (a checker, never part of the product): reads the
! binary grid, head and budget files of a MODFLOW 6 simulation with a DIS
! grid (or, with a sixth argument "disv", a DISV grid of rectangular cells:
! RectangularGridDisvMf6 ReadGRB and CheckRectangular) through the
! reference's own readers (RectangularGridDisMf6 ReadGRB,
! HeadReader, BudgetReader) and FlowModelData%LoadTimeStep, then assembles
! every cell with ModpathCellData%SetMassTransportDataUnstructured, the
! path ParticleTrackingEngine takes for MODFLOW 6 grids
! (ParticleTrackingEngine.f90:711-748).
!
! usage: mf6_ref GRB HDS CBC IN OUT [disv]
!   IN (stream): kper kstp (i4), hnoflow hdry (f8), nif (i4); ibound(ncell)
!   (i4), porosity(ncell), retardation(ncell) (f8); nif labels (16 chars)
!   and nif IFACE values (i4).
!   OUT (stream): ncell nja (i4), steady (i4); IBoundTS (i4), Heads,
!   SourceFlows, SinkFlows, StorageFlows, BoundaryFlows(6 ncell), FlowsJA,
!   Top, Bottom (f8); then per cell Q1..Q6 of the first sub-face,
!   SourceFlow and SinkFlow of the assembled cell, and the face pore
!   velocities VX1 VX2 VY1 VY2 VZ1 VZ2 and DX DY DZ of its sub-cell buffer
!   (FillMassSubCellDataBuffer of sub-cell (1,1), forward tracking) (f8);
!   with disv, then CellX, CellY, DelX, DelY (f8, ncell each); last the
!   grid's SaturatedTop after LoadTimeStep (f8, ncell).
program mf6_ref
  use RectangularGridDisMf6Module, only: RectangularGridDisMf6Type
  use RectangularGridDisvMf6Module, only: RectangularGridDisvMf6Type
  use HeadReaderModule, only: HeadReaderType
  use BudgetReaderModule, only: BudgetReaderType
  use FlowModelDataModule, only: FlowModelDataType
  use ModpathCellDataModule, only: ModpathCellDataType
  use ModflowRectangularGridModule, only: ModflowRectangularGridType
  use ModpathSubCellDataModule, only: ModpathSubCellDataType
  implicit none
  character(len=1024) :: fgrb, fhds, fcbc, fin, fout, kind
  type(RectangularGridDisMf6Type), pointer :: grid
  type(RectangularGridDisvMf6Type), pointer :: gridv
  class(ModflowRectangularGridType), pointer :: gp
  type(HeadReaderType), pointer :: hr
  type(BudgetReaderType), pointer :: br
  type(FlowModelDataType) :: flow
  type(ModpathCellDataType) :: cbuf
  type(ModpathSubCellDataType) :: sc
  integer :: n, ncell, kper, kstp, nif, k, steady
  double precision :: hnoflow, hdry
  integer, allocatable, target :: ibound(:), zones(:), icb(:), ifacev(:)
  double precision, allocatable, target :: por(:), ret(:)
  character(len=16), allocatable :: iflab(:)
  double precision :: bf(6)

  call get_command_argument(1, fgrb)
  call get_command_argument(2, fhds)
  call get_command_argument(3, fcbc)
  call get_command_argument(4, fin)
  call get_command_argument(5, fout)
  kind = 'dis'
  if (command_argument_count() >= 6) call get_command_argument(6, kind)

  allocate(hr, br)
  open(unit=20, file=trim(fgrb), form='unformatted', access='stream', status='old', &
       action='read')
  if (trim(kind) .eq. 'disv') then
    allocate(gridv)
    call gridv%ReadData(20, 0, 6)
    gp => gridv
  else
    allocate(grid)
    call grid%ReadData(20, 0, 6)
    gp => grid
  end if
  close(20)
  ncell = gp%CellCount
  call hr%OpenFile(trim(fhds), 21, 6)
  call br%OpenBudgetFile(trim(fcbc), 22, 6)

  open(unit=10, file=trim(fin), form='unformatted', access='stream', status='old', &
       action='read')
  read(10) kper, kstp, hnoflow, hdry, nif
  allocate(ibound(ncell), zones(ncell), icb(ncell), por(ncell), ret(ncell))
  allocate(iflab(max(nif, 1)), ifacev(max(nif, 1)))
  read(10) ibound, por, ret
  if (nif > 0) then
    read(10) (iflab(k), k = 1, nif)
    read(10) (ifacev(k), k = 1, nif)
  end if
  close(10)
  zones = 1
  icb = 0

  call flow%Initialize(hr, br, gp, hnoflow, hdry)
  if (.not. flow%Initialized) then
    write(*, '(A)') 'mf6_ref: flow model data not initialised'
    stop 2
  end if
  call flow%SetIBound(ibound, ncell)
  call flow%SetPorosity(por, ncell)
  call flow%SetZones(zones, ncell)
  call flow%SetRetardation(ret, ncell)
  call flow%SetDefaultIface(iflab, ifacev, nif)
  call flow%LoadTimeStep(kper, kstp)

  open(unit=30, file=trim(fout), form='unformatted', access='stream', status='replace')
  steady = 0
  if (flow%SteadyState) steady = 1
  write(30) ncell, gp%JaCount, steady
  write(30) flow%IBoundTS, flow%Heads, flow%SourceFlows, flow%SinkFlows, flow%StorageFlows
  write(30) flow%BoundaryFlows, flow%FlowsJA, gp%Top, gp%Bottom
  do n = 1, ncell
    bf = flow%BoundaryFlows(6 * (n - 1) + 1:6 * n)
    call cbuf%SetMassTransportDataUnstructured(n, ncell, gp%JaCount, gp, ibound, &
         flow%IBoundTS, por(n), ret(n), flow%StorageFlows(n), flow%SourceFlows(n), &
         flow%SinkFlows(n), flow%FlowsJA, bf, flow%Heads(n), gp%CellType(n), zones(n), &
         icb, 0)
    call cbuf%FillMassSubCellDataBuffer(sc, 1, 1, .false.)
    write(30) cbuf%Q1(1), cbuf%Q2(1), cbuf%Q3(1), cbuf%Q4(1), cbuf%Q5(1), cbuf%Q6(1), &
         cbuf%SourceFlow, cbuf%SinkFlow, sc%VX1, sc%VX2, sc%VY1, sc%VY2, sc%VZ1, sc%VZ2, &
         sc%DX, sc%DY, sc%DZ
  end do
  if (trim(kind) .eq. 'disv') write(30) gp%CellX, gp%CellY, gp%DelX, gp%DelY
  write(30) gp%SaturatedTop
  close(30)
end program mf6_ref
