! Reference-side driver of MODPATH-RW's timeseries and endpoint writers:
! test infrastructure only.
!
! Calls the reference's WriteTimeseriesHeader / WriteTimeseriesRecord and
! WriteEndpoints / WriteEndpointsMassParticles (ParticleManager.F90:51-250)
! on particles read from a binary file, with a RectangularGridDisType built
! from arrays, as MPathRW.F90:615-712 and :1454, :1743 use them.
!
! usage: writers_ref <input>   (writes ts.out, ep.out, epm.out and the
!                               observation records obsres.out/.bin,
!                               obssink.out/.bin in the working directory)
!   input (stream): int32 hdr(8) = ncol, nrow, nlay, np, ngroups,
!         trackingDirection, timePointIndex, timeStep;
!         real64 referenceTime, originX, originY, rotation, time;
!         real64 delr(ncol), delc(nrow), top(ncell), bot(ncell);
!         int32 zones(ncell);
!         int32 group, sequence, id, status, cell, initialCell,
!         initialFace (np each); real64 lx, ly, lz, t, ilx, ily, ilz, it,
!         mass (np each); int32 solute(ngroups); character(20) names(ngroups);
!         int32 isobs(ncell); real64 rfac(ncell), wvol(ncell), flowRate;
!         optionally real64 saturatedTop(ncell) (Grid%SaturatedTop, which
!         ConvertToModelXYZ uses as the cell top; the tops when absent)
!
! Observation records (WriteResidentObsRecord[Binary], WriteSinkObsRecord
! [Binary], Observation.f90:114-221) for the particles in cells with
! isobs /= 0: resident for status 1 (at the time point, as MPathRW.F90:1470),
! sink for status 3 (MPathRW.F90:1354).
program writers_ref
  use iso_fortran_env, only: int32, real64
  use ModflowRectangularGridModule, only: ModflowRectangularGridType
  use RectangularGridDisModule, only: RectangularGridDisType
  use ModpathSimulationDataModule, only: ModpathSimulationDataType
  use ParticleCoordinateModule, only: ParticleCoordinateType
  use GeoReferenceModule, only: GeoReferenceType
  use ParticleManagerModule
  use ParticleModule, only: ParticleType
  use ObservationModule, only: ObservationType, WriteResidentObsRecord, &
    WriteResidentObsRecordBinary, WriteSinkObsRecord, WriteSinkObsRecordBinary
  implicit none
  integer(int32) :: hdr(8)
  integer :: ncol, nrow, nlay, ncell, np, ng, nlc, i, j, k, n, p, g
  real(real64) :: rtime, ox, oy, rot, time
  real(real64), allocatable :: delr(:), delc(:), xf(:), yf(:)
  integer(int32), allocatable :: igrp(:), iseq(:), iid(:), ist(:), icell(:), icell0(:), iface0(:)
  integer(int32), allocatable :: solute(:)
  real(real64), allocatable :: lx(:), ly(:), lz(:), tt(:), ilx(:), ily(:), ilz(:), it(:), mass(:)
  character(len=20), allocatable :: names(:)
  integer, allocatable :: cnt(:), pos(:)
  character(len=1024) :: fin
  class(ModflowRectangularGridType), pointer :: grid
  type(RectangularGridDisType), pointer :: dis
  type(ModpathSimulationDataType), target :: sim
  type(ParticleCoordinateType) :: pc
  type(GeoReferenceType) :: geo
  integer(int32), allocatable :: isobs(:)
  real(real64), allocatable :: rfac(:), wvol(:)
  real(real64) :: flowRate
  real(real64), allocatable :: satTop(:)
  integer :: ios
  type(ObservationType) :: obs
  type(ParticleType) :: q1
  integer :: c

  call get_command_argument(1, fin)
  open(10, file=trim(fin), access='stream', form='unformatted', status='old')
  read(10) hdr
  ncol = hdr(1); nrow = hdr(2); nlay = hdr(3); np = hdr(4); ng = hdr(5)
  ncell = ncol * nrow * nlay
  nlc = ncol * nrow
  read(10) rtime, ox, oy, rot, time
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
  dis%OriginX = ox
  dis%OriginY = oy
  dis%RotationAngle = rot
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

  allocate(sim%Zones(ncell))
  read(10) sim%Zones
  allocate(igrp(np), iseq(np), iid(np), ist(np), icell(np), icell0(np), iface0(np))
  allocate(lx(np), ly(np), lz(np), tt(np), ilx(np), ily(np), ilz(np), it(np), mass(np))
  allocate(solute(ng), names(ng))
  read(10) igrp, iseq, iid, ist, icell, icell0, iface0
  read(10) lx, ly, lz, tt, ilx, ily, ilz, it, mass
  read(10) solute, names
  allocate(isobs(ncell), rfac(ncell), wvol(ncell))
  read(10) isobs, rfac, wvol, flowRate
  allocate(satTop(ncell))
  read(10, iostat=ios) satTop
  if (ios == 0) dis%SaturatedTop = satTop
  close(10)

  ! timeseries: header, then the active particles in order
  open(11, file='ts.out', status='replace', form='formatted', access='sequential')
  call WriteTimeseriesHeader(11, int(hdr(6)), rtime, ox, oy, rot)
  do p = 1, np
    if (ist(p) .ne. 1) cycle
    pc%CellNumber = icell(p)
    pc%Layer = (icell(p) - 1) / nlc + 1
    pc%LocalX = lx(p)
    pc%LocalY = ly(p)
    pc%LocalZ = lz(p)
    pc%TrackingTime = time
    call grid%ConvertToModelXYZ(pc%CellNumber, pc%LocalX, pc%LocalY, pc%LocalZ, &
      pc%GlobalX, pc%GlobalY, pc%GlobalZ)
    call WriteTimeseriesRecord(int(iseq(p)), int(iid(p)), int(igrp(p)), int(hdr(8)), &
      int(hdr(7)), pc, geo, 11)
  end do
  close(11)

  ! endpoints: particles grouped as the file lists them (group-major)
  sim%TrackingDirection = hdr(6)
  sim%ReferenceTime = rtime
  sim%ParticleGroupCount = ng
  allocate(sim%ParticleGroups(ng), cnt(ng), pos(ng))
  cnt = 0
  do p = 1, np
    cnt(igrp(p)) = cnt(igrp(p)) + 1
  end do
  do g = 1, ng
    sim%ParticleGroups(g)%Group = g
    sim%ParticleGroups(g)%Name = names(g)
    sim%ParticleGroups(g)%Solute = solute(g)
    sim%ParticleGroups(g)%TotalParticleCount = cnt(g)
    allocate(sim%ParticleGroups(g)%Particles(cnt(g)))
  end do
  pos = 0
  do p = 1, np
    g = igrp(p)
    pos(g) = pos(g) + 1
    associate (q => sim%ParticleGroups(g)%Particles(pos(g)))
      q%Group = g
      q%ID = iid(p)
      q%SequenceNumber = iseq(p)
      q%Status = ist(p)
      q%CellNumber = icell(p)
      q%Layer = (icell(p) - 1) / nlc + 1
      q%LocalX = lx(p)
      q%LocalY = ly(p)
      q%LocalZ = lz(p)
      q%TrackingTime = tt(p)
      q%InitialCellNumber = icell0(p)
      q%InitialLayer = (icell0(p) - 1) / nlc + 1
      q%InitialFace = iface0(p)
      q%InitialLocalX = ilx(p)
      q%InitialLocalY = ily(p)
      q%InitialLocalZ = ilz(p)
      q%InitialTrackingTime = it(p)
      q%Mass = mass(p)
    end associate
  end do
  open(12, file='ep.out', status='replace', form='formatted', access='sequential')
  call WriteEndpoints(sim, grid, geo, 12)
  close(12)
  open(13, file='epm.out', status='replace', form='formatted', access='sequential')
  call WriteEndpointsMassParticles(sim, grid, geo, 13)
  close(13)

  ! observation records, particles in file order
  open(14, file='obsres.out', status='replace', form='formatted', access='sequential')
  open(15, file='obsres.bin', status='replace', form='unformatted', access='stream')
  open(16, file='obssink.out', status='replace', form='formatted', access='sequential')
  open(17, file='obssink.bin', status='replace', form='unformatted', access='stream')
  do p = 1, np
    c = icell(p)
    if (isobs(c) .eq. 0) cycle
    q1%ID = iid(p)
    q1%Mass = mass(p)
    q1%Group = igrp(p)
    q1%CellNumber = c
    q1%Layer = (c - 1) / nlc + 1
    q1%TrackingTime = tt(p)
    if (ist(p) .eq. 1) then
      pc%CellNumber = c
      pc%Layer = (c - 1) / nlc + 1
      pc%LocalX = lx(p)
      pc%LocalY = ly(p)
      pc%LocalZ = lz(p)
      pc%TrackingTime = tt(p)
      call grid%ConvertToModelXYZ(pc%CellNumber, pc%LocalX, pc%LocalY, pc%LocalZ, &
        pc%GlobalX, pc%GlobalY, pc%GlobalZ)
      call WriteResidentObsRecord(obs, int(hdr(8)), int(hdr(7)), q1, pc, &
        int(solute(igrp(p))), rfac(c), wvol(c), 14)
      call WriteResidentObsRecordBinary(obs, int(hdr(8)), int(hdr(7)), q1, pc, &
        int(solute(igrp(p))), rfac(c), wvol(c), 15)
    else if (ist(p) .eq. 3) then
      call WriteSinkObsRecord(obs, int(hdr(8)), int(hdr(7)), q1, int(solute(igrp(p))), &
        flowRate, 16)
      call WriteSinkObsRecordBinary(obs, int(hdr(8)), int(hdr(7)), q1, &
        int(solute(igrp(p))), flowRate, 17)
    end if
  end do
  close(14)
  close(15)
  close(16)
  close(17)
end program writers_ref
