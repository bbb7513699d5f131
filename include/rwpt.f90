! Fortran binding of the C ABI of include/rwpt.h for MODPATH-RW
! (ISO_C_BINDING): derived types with the field order and kinds of the C
! structs, one interface per entry point.  The reference's driver uses it in
! place of the OpenMP particle loop of MPathRW.F90:1239-1403 and of the gpkde
! calls of MPathRW.F90:556-574 and :1626-1642.
module RwptGpuModule
  use iso_c_binding
  implicit none
  private
  public :: rwpt_grid, rwpt_options, rwpt_kde_options
  public :: rwpt_default_options, rwpt_create, rwpt_destroy, rwpt_last_error
  public :: rwpt_set_flow, rwpt_set_cell_extras, rwpt_set_dispersion, rwpt_set_saturated_top
  public :: rwpt_set_graph
  public :: rwpt_load_particles, rwpt_get_particles, rwpt_particle_count
  public :: rwpt_advance, rwpt_set_time, rwpt_set_steady_state, rwpt_global_xyz
  public :: rwpt_kde, rwpt_kde_device, rwpt_kde_info
  public :: rwpt_kde_write, rwpt_get_substeps, rwpt_set_tuning, rwpt_last_launches
  public :: rwpt_kernel_time, rwpt_reset_kernel_time, rwpt_get_cell_tables
  public :: rwpt_create_ensemble, rwpt_autotune, rwpt_write_timeseries_header
  public :: rwpt_write_timeseries, rwpt_write_endpoints, rwpt_adjoint
  public :: rwpt_write_observation, rwpt_adjoint_options, rwpt_adjoint_stats
  public :: rwpt_adjoint_opts
  public :: rwpt_usg_grid, rwpt_create_usg, rwpt_set_flow_usg, rwpt_usg_unit_count
  public :: rwpt_get_usg_tables

  type, bind(C) :: rwpt_adjoint_opts
    real(c_double)     :: graze_cut
    integer(c_int32_t) :: reverse_windows, tangent_windows, frozen_noise, diagnostics
  end type
  type, bind(C) :: rwpt_grid
    integer(c_int32_t) :: ncol, nrow, nlay
    type(c_ptr) :: delr, delc, top, bot          ! c_loc of real(c_double) arrays
  end type
  type, bind(C) :: rwpt_usg_grid
    integer(c_int32_t) :: ncell, nlay
    integer(c_int64_t) :: nja
    type(c_ptr) :: ia, ja, layer                 ! c_loc of integer(c_int32_t) arrays
    type(c_ptr) :: cell_x, cell_y, del_x, del_y, top, bot  ! real(c_double) arrays
    type(c_ptr) :: cell_type                     ! integer(c_int32_t) array or c_null_ptr
  end type
  type, bind(C) :: rwpt_options
    integer(c_int32_t) :: advection_kind, time_step_kind
    real(c_double)     :: time_step_params(2)
    integer(c_int32_t) :: dimension_mask(3), backward_tracking, stop_zone
    integer(c_int32_t) :: stop_at_weak_sinks, stop_at_weak_sources, steady_state
    integer(c_int32_t) :: precision
    integer(c_int64_t) :: base_seed(2)
  end type
  type, bind(C) :: rwpt_kde_options
    real(c_double)     :: origin(3), bin_size(3)
    integer(c_int32_t) :: nbins(3), mode, weighted
    real(c_double)     :: scaling_factor
    integer(c_int32_t) :: unit_volume, n_opt_loops, isotropic, skip_error_convergence
    real(c_double)     :: rel_error_convergence
    integer(c_int32_t) :: initial_smoothing
    real(c_double)     :: initial_smoothing_factor
    integer(c_int32_t) :: bound_kernel_size
    real(c_double)     :: min_h_over_delta, max_h_over_delta
    integer(c_int32_t) :: effective_weight_format, kernel_database
    real(c_double)     :: delta_h_over_delta
    integer(c_int32_t) :: member, initial
  end type

  interface
    subroutine rwpt_default_options(opt) bind(C, name='rwpt_default_options')
      import :: rwpt_options
      type(rwpt_options), intent(out) :: opt
    end subroutine
    integer(c_int) function rwpt_create(grid, opt, e) bind(C, name='rwpt_create')
      import :: c_int, c_ptr, rwpt_grid, rwpt_options
      type(rwpt_grid), intent(in) :: grid
      type(rwpt_options), intent(in) :: opt
      type(c_ptr), intent(out) :: e
    end function
    integer(c_int) function rwpt_create_usg(grid, opt, e) bind(C, name='rwpt_create_usg')
      import :: c_int, c_ptr, rwpt_usg_grid, rwpt_options
      type(rwpt_usg_grid), intent(in) :: grid
      type(rwpt_options), intent(in) :: opt
      type(c_ptr), intent(out) :: e
    end function
    ! bflow, storage, retardation, ibound and icbound may be c_null_ptr
    integer(c_int) function rwpt_set_flow_usg(e, flowja, bflow, storage, porosity, &
        retardation, ibound, icbound, default_icbound) bind(C, name='rwpt_set_flow_usg')
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: e
      real(c_double), intent(in) :: flowja(*), porosity(*)
      type(c_ptr), value :: bflow, storage, retardation, ibound, icbound
      integer(c_int32_t), value :: default_icbound
    end function
    integer(c_int64_t) function rwpt_usg_unit_count(e) bind(C, name='rwpt_usg_unit_count')
      import :: c_int64_t, c_ptr
      type(c_ptr), value :: e
    end function
    integer(c_int) function rwpt_get_usg_tables(e, out) bind(C, name='rwpt_get_usg_tables')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: e
      real(c_double), intent(out) :: out(*)
    end function
    integer(c_int) function rwpt_destroy(e) bind(C, name='rwpt_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: e
    end function
    type(c_ptr) function rwpt_last_error(e) bind(C, name='rwpt_last_error')
      import :: c_ptr
      type(c_ptr), value :: e
    end function
    integer(c_int) function rwpt_set_flow(e, frf, fff, flf, bflow, porosity, &
        retardation, ibound, icbound, default_icbound) bind(C, name='rwpt_set_flow')
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: e
      real(c_double), intent(in) :: frf(*), fff(*), flf(*), bflow(*), porosity(*), retardation(*)
      integer(c_int32_t), intent(in) :: ibound(*), icbound(*)
      integer(c_int32_t), value :: default_icbound
    end function
    integer(c_int) function rwpt_set_cell_extras(e, head, layer_type, zone, &
        source_flow, sink_flow) bind(C, name='rwpt_set_cell_extras')
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: e
      real(c_double), intent(in) :: head(*), source_flow(*), sink_flow(*)
      integer(c_int32_t), intent(in) :: layer_type(*), zone(*)
    end function
    integer(c_int) function rwpt_set_graph(e, launches, max_list) &
        bind(C, name='rwpt_set_graph')
      import :: c_int, c_ptr, c_int32_t, c_int64_t
      type(c_ptr), value :: e
      integer(c_int32_t), value :: launches
      integer(c_int64_t), value :: max_list
    end function
    ! saturated_top: c_loc(Grid%SaturatedTop), or c_null_ptr for the cell tops
    integer(c_int) function rwpt_set_saturated_top(e, saturated_top) &
        bind(C, name='rwpt_set_saturated_top')
      import :: c_int, c_ptr
      type(c_ptr), value :: e, saturated_top
    end function
    integer(c_int) function rwpt_set_dispersion(e, model, uniform, aL, aT, aLV, aTV, &
        dmeff) bind(C, name='rwpt_set_dispersion')
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: e
      integer(c_int32_t), value :: model, uniform
      real(c_double), intent(in) :: aL(*), aT(*), aLV(*), aTV(*), dmeff(*)
    end function
    integer(c_int) function rwpt_load_particles(e, n, cell, lx, ly, lz, t, status, &
        rng, stream_offset) bind(C, name='rwpt_load_particles')
      import :: c_int, c_ptr, c_double, c_int32_t, c_int64_t
      type(c_ptr), value :: e
      integer(c_int64_t), value :: n, stream_offset
      integer(c_int32_t), intent(in) :: cell(*), status(*)
      real(c_double), intent(in) :: lx(*), ly(*), lz(*), t(*)
      type(c_ptr), value :: rng                  ! c_null_ptr: per-particle streams
    end function
    integer(c_int) function rwpt_get_particles(e, cell, lx, ly, lz, t, status, rng) &
        bind(C, name='rwpt_get_particles')
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: e
      integer(c_int32_t), intent(out) :: cell(*), status(*)
      real(c_double), intent(out) :: lx(*), ly(*), lz(*), t(*)
      type(c_ptr), value :: rng
    end function
    integer(c_int64_t) function rwpt_particle_count(e) bind(C, name='rwpt_particle_count')
      import :: c_int64_t, c_ptr
      type(c_ptr), value :: e
    end function
    integer(c_int) function rwpt_set_time(e, t) bind(C, name='rwpt_set_time')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: e
      real(c_double), value :: t
    end function
    integer(c_int) function rwpt_set_steady_state(e, steady_state) &
        bind(C, name='rwpt_set_steady_state')
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: e
      integer(c_int32_t), value :: steady_state
    end function
    integer(c_int) function rwpt_advance(e, t_stop, n_active, n_substeps) &
        bind(C, name='rwpt_advance')
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: e
      real(c_double), value :: t_stop
      integer(c_int64_t), intent(out) :: n_active, n_substeps
    end function
    integer(c_int) function rwpt_global_xyz(e, gx, gy, gz) bind(C, name='rwpt_global_xyz')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: e
      real(c_double), intent(out) :: gx(*), gy(*), gz(*)
    end function
    integer(c_int) function rwpt_kde(e, opt, mass, density, histogram) bind(C, name='rwpt_kde')
      import :: c_int, c_ptr, c_double, rwpt_kde_options
      type(c_ptr), value :: e
      type(rwpt_kde_options), intent(in) :: opt
      type(c_ptr), value :: mass                 ! c_null_ptr: unit masses
      real(c_double), intent(out) :: density(*), histogram(*)
    end function
    integer(c_int) function rwpt_kde_device(e, opt, density_dev, histogram_dev) &
        bind(C, name='rwpt_kde_device')
      import :: c_int, c_ptr, rwpt_kde_options
      type(c_ptr), value :: e
      type(rwpt_kde_options), intent(in) :: opt
      type(c_ptr), value :: density_dev, histogram_dev   ! device addresses
    end function
    integer(c_int) function rwpt_kde_info(e, bandwidth, points, npoints, eff_mass, &
        loops) bind(C, name='rwpt_kde_info')
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: e
      real(c_double), intent(out) :: bandwidth, points, npoints, eff_mass
      integer(c_int32_t), intent(out) :: loops
    end function
    integer(c_int) function rwpt_kde_write(path, append, format, column_format, &
        time_id, time, group_id, opt, density, histogram, histogram_scale) &
        bind(C, name='rwpt_kde_write')
      import :: c_int, c_char, c_double, c_int32_t, rwpt_kde_options
      character(kind=c_char), intent(in) :: path(*)   ! trim(name)//c_null_char
      integer(c_int32_t), value :: append, format, column_format, time_id, group_id
      real(c_double), value :: time, histogram_scale
      type(rwpt_kde_options), intent(in) :: opt
      real(c_double), intent(in) :: density(*), histogram(*)
    end function
    integer(c_int) function rwpt_get_substeps(e, substeps) bind(C, name='rwpt_get_substeps')
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: e
      integer(c_int64_t), intent(out) :: substeps(*)
    end function
    integer(c_int) function rwpt_set_tuning(e, budget, block, sort, min_waves) &
        bind(C, name='rwpt_set_tuning')
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: e
      integer(c_int32_t), value :: budget, block, sort, min_waves
    end function
    integer(c_int) function rwpt_last_launches(e) bind(C, name='rwpt_last_launches')
      import :: c_int, c_ptr
      type(c_ptr), value :: e
    end function
    integer(c_int) function rwpt_kernel_time(e, which, ms, launches) &
        bind(C, name='rwpt_kernel_time')
      import :: c_int, c_ptr, c_int32_t, c_double, c_int64_t
      type(c_ptr), value :: e
      integer(c_int32_t), value :: which
      real(c_double), intent(out) :: ms
      integer(c_int64_t), intent(out) :: launches
    end function
    integer(c_int) function rwpt_reset_kernel_time(e) bind(C, name='rwpt_reset_kernel_time')
      import :: c_int, c_ptr
      type(c_ptr), value :: e
    end function
    integer(c_int) function rwpt_get_cell_tables(e, out) bind(C, name='rwpt_get_cell_tables')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: e
      real(c_double), intent(out) :: out(*)
    end function
    integer(c_int) function rwpt_create_ensemble(grid, opt, members, e) &
        bind(C, name='rwpt_create_ensemble')
      import :: c_int, c_ptr, c_int32_t, rwpt_grid, rwpt_options
      type(rwpt_grid), intent(in) :: grid
      type(rwpt_options), intent(in) :: opt
      integer(c_int32_t), value :: members
      type(c_ptr), intent(out) :: e
    end function
    integer(c_int) function rwpt_autotune(e, t_probe, budget, block, sort, min_waves, &
        ms_best, trial_substeps) bind(C, name='rwpt_autotune')
      import :: c_int, c_ptr, c_double, c_int32_t, c_int64_t
      type(c_ptr), value :: e
      real(c_double), value :: t_probe
      integer(c_int32_t), intent(out) :: budget, block, sort, min_waves
      real(c_double), intent(out) :: ms_best
      integer(c_int64_t), intent(out) :: trial_substeps
    end function
    integer(c_int) function rwpt_write_timeseries_header(path, tracking_direction, &
        reference_time, origin_x, origin_y, rotation) &
        bind(C, name='rwpt_write_timeseries_header')
      import :: c_int, c_char, c_double, c_int32_t
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int32_t), value :: tracking_direction
      real(c_double), value :: reference_time, origin_x, origin_y, rotation
    end function
    integer(c_int) function rwpt_write_timeseries(e, path, time_point, time_step, time, &
        sequence, id, group, select) bind(C, name='rwpt_write_timeseries')
      import :: c_int, c_ptr, c_char, c_double, c_int32_t
      type(c_ptr), value :: e
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int32_t), value :: time_point, time_step
      real(c_double), value :: time
      integer(c_int32_t), intent(in) :: sequence(*), id(*), group(*)
      type(c_ptr), value :: select                  ! c_loc of integer(c_int8_t) or c_null_ptr
    end function
    integer(c_int) function rwpt_write_endpoints(e, path, format, tracking_direction, &
        reference_time, origin_x, origin_y, rotation, n_groups, group_names, &
        group_solute, group, sequence, id, initial_cell, initial_lx, initial_ly, &
        initial_lz, initial_time, initial_face, mass) bind(C, name='rwpt_write_endpoints')
      import :: c_int, c_ptr, c_char, c_double, c_int32_t
      type(c_ptr), value :: e
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int32_t), value :: format, tracking_direction, n_groups
      real(c_double), value :: reference_time, origin_x, origin_y, rotation
      type(c_ptr), intent(in) :: group_names(*)     ! c_loc of null-terminated names
      integer(c_int32_t), intent(in) :: group_solute(*), group(*), sequence(*), id(*)
      integer(c_int32_t), intent(in) :: initial_cell(*), initial_face(*)
      real(c_double), intent(in) :: initial_lx(*), initial_ly(*), initial_lz(*)
      real(c_double), intent(in) :: initial_time(*), mass(*)
    end function
    integer(c_int) function rwpt_write_observation(e, path, append, style, format, n_cells, &
        cells, time_point, time_step, id, group, group_solute, mass, flow_rate, select) &
        bind(C, name='rwpt_write_observation')
      import :: c_int, c_ptr, c_char, c_double, c_int32_t
      type(c_ptr), value :: e
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int32_t), value :: append, style, format, n_cells, time_point, time_step
      integer(c_int32_t), intent(in) :: cells(*), id(*), group(*), group_solute(*)
      type(c_ptr), value :: mass                    ! c_loc of the masses or c_null_ptr
      real(c_double), value :: flow_rate
      type(c_ptr), value :: select                  ! c_loc of integer(c_int8_t) or c_null_ptr
    end function
    integer(c_int) function rwpt_adjoint(e, n_windows, t_windows, n_obs, obs_xyz, target, &
        bandwidth, mass, misfit, conc, grad_al, grad_at, grad_dm, grad_porosity_scale, ms) &
        bind(C, name='rwpt_adjoint')
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: e
      integer(c_int32_t), value :: n_windows, n_obs
      real(c_double), intent(in) :: t_windows(*), obs_xyz(*), target(*), bandwidth(3)
      type(c_ptr), value :: mass                    ! c_loc of the masses or c_null_ptr
      real(c_double), intent(out) :: misfit, conc(*)
      type(c_ptr), value :: grad_al, grad_at, grad_dm, grad_porosity_scale, ms
    end function
    integer(c_int) function rwpt_adjoint_options(e, opt) bind(C, name='rwpt_adjoint_options')
      import :: c_int, c_ptr, rwpt_adjoint_opts
      type(c_ptr), value :: e
      type(rwpt_adjoint_opts), intent(in) :: opt
    end function
    integer(c_int) function rwpt_adjoint_stats(e, out) bind(C, name='rwpt_adjoint_stats')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: e
      real(c_double), intent(out) :: out(33)
    end function
  end interface
end module RwptGpuModule
