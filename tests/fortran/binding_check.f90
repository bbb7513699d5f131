! Exercises include/rwpt.f90 without a GPU: prints the default options as
! the C side fills them, and writes density rows through rwpt_kde_write.
! usage: binding_check <output file>
program binding_check
  use iso_c_binding
  use RwptGpuModule
  implicit none
  type(rwpt_options) :: o
  type(rwpt_kde_options) :: k
  real(c_double) :: dens(6), hist(6)
  character(len=512) :: path
  integer(c_int) :: rc

  call rwpt_default_options(o)
  print '(a,2i4,2es24.16)', 'kinds ', o%advection_kind, o%time_step_kind, o%time_step_params
  print '(a,3i4,6i4)', 'mask ', o%dimension_mask, o%backward_tracking, o%stop_zone, &
    o%stop_at_weak_sinks, o%stop_at_weak_sources, o%steady_state, o%precision
  print '(a,2z17)', 'seed ', o%base_seed

  call get_command_argument(1, path)
  k%origin = [0d0, 0d0, 0d0]
  k%bin_size = [1d0, 1d0, 0d0]
  k%nbins = [3, 2, 1]
  dens = [0d0, 1.5d0, -2d0, 3.25d-7, 4d0, 0d0]
  hist = [1d0, 2d0, 0d0, 4d0, 5d0, 6d0]
  rc = rwpt_kde_write(trim(path)//c_null_char, 0, 0, 1, 1, 2.5d0, 1, k, dens, hist, 2d0)
  print '(a,i4)', 'rc ', rc
end program binding_check
