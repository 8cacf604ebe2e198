!------------------------------------------------------------------------------
! Thin Fortran host of the MI355X hot path.
!
! Holds the particle state in module arrays laid out as the reference's
! (x(ndim,npart), vel(3,npart), ... ; src/variablesND.f90:161-215) and
! provides the reference's argument-less entry points
!   derivs          (src/derivs.f90:27)
!   step            (src/stepND_leapfrog_mhd.f90:31)
!   set_linklist    (src/linkND.f90:45)
!   iterate_density (src/iterate_density.f90:43)
!   get_rates       (src/ratesND_mhd.f90:29)
! each a call through the C-ABI into libndsp_hip.so.  The state stays in HBM
! between calls; ndsp_host_sync copies it back into the module arrays (the
! reference's own host copy is touched only at output/evwrite time).
! A non-zero status is fatal, like the reference's quit
! (src/ndspmhd.f90:369): the message is printed and the program stops.
!------------------------------------------------------------------------------
module ndsp_host
 use iso_c_binding
 use ndsp_capi
 implicit none
 type(c_ptr), save :: ctx = c_null_ptr
 integer, save :: npart = 0, ndim = 0
 integer(c_int32_t), save :: iopt(32)
 real(c_double), save :: dopt(32)
 real(c_double), save :: time = 0.0_c_double, dt = 0.0_c_double
 ! evolved state (reference layout)
 real(c_double), allocatable, save :: x(:,:), vel(:,:), Bevol(:,:), alpha(:,:)
 real(c_double), allocatable, save :: pmass(:), hh(:), en(:), psi(:), dustevol(:)
 integer(c_int32_t), allocatable, save :: itype(:)
 ! derived and rates (filled by ndsp_host_sync)
 real(c_double), allocatable, save :: rho(:), force(:,:), dBevoldt(:,:), dudt(:)
 type(ndsp_stats_t), save :: stats
 type(ndsp_dt_t), save :: dtinfo

contains

 subroutine ndsp_host_check(ierr, what)
    integer(c_int), intent(in) :: ierr
    character(len=*), intent(in) :: what
    if (ierr /= 0) then
       write(*,'(a,a,a,i3,a,a)') ' ndsp: ', what, ' failed, status ', ierr, ': ', trim(ndsp_error_text())
       error stop 2
    endif
 end subroutine ndsp_host_check

 ! allocate the host arrays for n particles in ndim dimensions
 subroutine ndsp_host_alloc(n, nd)
    integer, intent(in) :: n, nd
    npart = n
    ndim = nd
    allocate(x(nd,n), vel(3,n), Bevol(3,n), alpha(3,n), pmass(n), hh(n), en(n), psi(n), dustevol(n), itype(n))
    allocate(rho(n), force(3,n), dBevoldt(3,n), dudt(n))
    itype = 0
    psi = 0.
    dustevol = 0.
 end subroutine ndsp_host_alloc

 ! create the device context with the options and copy the state to HBM
 ! (replaces the module set-up of src/initialiseND_mhd.f90:140-300)
 subroutine ndsp_host_init(precision)
    integer, intent(in) :: precision
    call ndsp_host_check(ndsp_create(iopt, dopt, int(npart, c_int64_t), int(precision, c_int32_t), ctx), &
                         'ndsp_create')
    call ndsp_host_check(ndsp_upload(ctx, x, vel, pmass, hh, en, Bevol, alpha, psi, dustevol, itype), &
                         'ndsp_upload')
    dt = dopt(24)
    time = 0.
    call ndsp_host_check(ndsp_set_dt(ctx, dt, time), 'ndsp_set_dt')
    ! steps replay recorded HIP graphs (same results, fewer launches)
    call ndsp_host_check(ndsp_set_graph(ctx, 1_c_int32_t), 'ndsp_set_graph')
 end subroutine ndsp_host_init

 subroutine get_field(name, a, n)
    character(len=*), intent(in) :: name
    integer, intent(in) :: n
    real(c_double), intent(out) :: a(n)
    integer(c_int32_t) :: w
    call ndsp_host_check(ndsp_download(ctx, trim(name)//c_null_char, a, w), 'ndsp_download '//name)
 end subroutine get_field

 ! HBM -> module arrays
 subroutine ndsp_host_sync()
    call get_field('x', x, ndim*npart)
    call get_field('vel', vel, 3*npart)
    call get_field('Bevol', Bevol, 3*npart)
    call get_field('alpha', alpha, 3*npart)
    call get_field('hh', hh, npart)
    call get_field('en', en, npart)
    call get_field('psi', psi, npart)
    call get_field('dustevol', dustevol, npart)
    call get_field('rho', rho, npart)
    call get_field('force', force, 3*npart)
    call get_field('dBevoldt', dBevoldt, 3*npart)
    call get_field('dudt', dudt, npart)
 end subroutine ndsp_host_sync

 subroutine ndsp_host_finish()
    integer(c_int) :: ierr
    if (c_associated(ctx)) ierr = ndsp_destroy(ctx)
    ctx = c_null_ptr
 end subroutine ndsp_host_finish

end module ndsp_host

!--- the reference's argument-less hot-path entry points ----------------------

subroutine derivs
 use ndsp_host
 implicit none
 call ndsp_host_check(ndsp_derivs(ctx, stats, dtinfo), 'derivs')
end subroutine derivs

subroutine step
 use ndsp_host
 implicit none
 call ndsp_host_check(ndsp_step(ctx, stats, dtinfo), 'step')
 dt = dtinfo%dt
end subroutine step

subroutine set_linklist
 use ndsp_host
 implicit none
 call ndsp_host_check(ndsp_set_linklist(ctx), 'set_linklist')
end subroutine set_linklist

subroutine iterate_density
 use ndsp_host
 implicit none
 call ndsp_host_check(ndsp_iterate_density(ctx, stats), 'iterate_density')
end subroutine iterate_density

subroutine get_rates
 use ndsp_host
 implicit none
 call ndsp_host_check(ndsp_get_rates(ctx, dtinfo), 'get_rates')
end subroutine get_rates
