! Drives the engine from Fortran through the bind(C) module of
! include/rwpt.f90, as the reference's driver would in place of its
! particle loop (MPathRW.F90:1239-1403) and gpkde calls (:1626-1642):
! create, set_flow, set_dispersion, load_particles, advance through the
! output times, get_particles, kde.  The grid, flows and particles come
! from a stream file written by the calling test.
! usage: drive_engine <input> <output>
!   input:  int32 ncol nrow nlay np nout mask(3) advection model;
!           real64 dt, delr(ncol), delc(nrow), top, bot, frf, fff, flf (ncell),
!           bflow(6 ncell), porosity, retardation (ncell);
!           int32 ibound, icbound (ncell), default_icbound;
!           real64 aL, aT, dmeff (one each);
!           int32 cell(np); real64 lx, ly, lz (np), tout(nout);
!           real64 origin(3), bin(3); int32 nbins(3)
!   output: int32 rc, cell(np), status(np); real64 lx, ly, lz, t (np),
!           density, histogram (nbins product)
program drive_engine
  use iso_c_binding
  use RwptGpuModule
  implicit none
  integer(c_int32_t) :: hdr(10), dicb, nb(3)
  integer :: ncol, nrow, nlay, ncell, np, nout, k, nk
  real(c_double) :: dt, aL(1), aT(1), dm(1), orig(3), bins(3), zero(1)
  real(c_double), allocatable, target :: delr(:), delc(:), top(:), bot(:)
  real(c_double), allocatable :: frf(:), fff(:), flf(:), bflow(:), por(:), ret(:)
  real(c_double), allocatable :: lx(:), ly(:), lz(:), t(:), tout(:), dens(:), hist(:)
  integer(c_int32_t), allocatable :: ibound(:), icbound(:), cell(:), status(:)
  type(rwpt_grid) :: g
  type(rwpt_options) :: o
  type(rwpt_kde_options) :: ko
  type(c_ptr) :: e
  integer(c_int) :: rc
  integer(c_int64_t) :: nact, nsteps
  character(len=1024) :: fin, fout

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(10, file=trim(fin), access='stream', form='unformatted', status='old')
  read(10) hdr
  ncol = hdr(1); nrow = hdr(2); nlay = hdr(3); np = hdr(4); nout = hdr(5)
  ncell = ncol * nrow * nlay
  allocate(delr(ncol), delc(nrow), top(ncell), bot(ncell), frf(ncell), fff(ncell))
  allocate(flf(ncell), bflow(6 * ncell), por(ncell), ret(ncell), ibound(ncell))
  allocate(icbound(ncell), cell(np), status(np), lx(np), ly(np), lz(np), t(np), tout(nout))
  read(10) dt, delr, delc, top, bot, frf, fff, flf, bflow, por, ret
  read(10) ibound, icbound, dicb
  read(10) aL, aT, dm
  read(10) cell, lx, ly, lz, tout
  read(10) orig, bins, nb
  close(10)
  nk = nb(1) * nb(2) * nb(3)
  allocate(dens(nk), hist(nk))

  g%ncol = ncol; g%nrow = nrow; g%nlay = nlay
  g%delr = c_loc(delr); g%delc = c_loc(delc); g%top = c_loc(top); g%bot = c_loc(bot)
  call rwpt_default_options(o)
  o%dimension_mask = hdr(6:8)
  o%advection_kind = hdr(9)
  o%time_step_kind = 4
  o%time_step_params = [dt, 0d0]
  rc = rwpt_create(g, o, e)
  if (rc == 0) rc = rwpt_set_flow(e, frf, fff, flf, bflow, por, ret, ibound, icbound, dicb)
  zero = 0d0
  if (rc == 0) rc = rwpt_set_dispersion(e, hdr(10), 1, aL, aT, zero, zero, dm)
  t = 0d0
  status = 1
  if (rc == 0) rc = rwpt_load_particles(e, int(np, c_int64_t), cell, lx, ly, lz, t, &
                                        status, c_null_ptr, 0_c_int64_t)
  do k = 1, nout
    if (rc == 0) rc = rwpt_advance(e, tout(k), nact, nsteps)
  end do
  if (rc == 0) rc = rwpt_get_particles(e, cell, lx, ly, lz, t, status, c_null_ptr)
  ko%origin = orig
  ko%bin_size = bins
  ko%nbins = nb
  ko%mode = 1
  ko%weighted = 0
  ko%scaling_factor = 1d0
  ko%unit_volume = 0
  ko%n_opt_loops = 0
  ko%isotropic = 0
  ko%skip_error_convergence = 0
  ko%rel_error_convergence = 0d0
  ko%initial_smoothing = 0
  ko%initial_smoothing_factor = 1d0
  ko%bound_kernel_size = 0
  ko%min_h_over_delta = 0d0
  ko%max_h_over_delta = 0d0
  ko%effective_weight_format = 0
  ko%kernel_database = 0
  ko%delta_h_over_delta = 0d0
  ko%member = -1
  ko%initial = 0
  if (rc == 0) rc = rwpt_kde(e, ko, c_null_ptr, dens, hist)
  if (rc /= 0) print '(a,i4)', 'engine error ', rc
  if (c_associated(e)) then
    if (rwpt_destroy(e) /= 0) rc = -100
  end if
  open(11, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(11) rc, cell, status, lx, ly, lz, t, dens, hist
  close(11)
end program drive_engine
