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
 integer(c_int32_t), allocatable, save :: itype(:), ireal(:)   ! ireal 0-based, -1 none
 ! derived and rates (filled by ndsp_host_sync)
 real(c_double), allocatable, save :: rho(:), force(:,:), dBevoldt(:,:), dudt(:)
 type(ndsp_stats_t), save :: stats
 type(ndsp_dt_t), save :: dtinfo
 ! output units, as the reference's loguns (src/logun.f90:37-40)
 integer, save :: ievfile = 13, idatfile = 12, ireadf = 15
 ! further fields for dumps (filled by ndsp_host_sync_output)
 real(c_double), allocatable, save :: dens(:), uu(:), pr(:), drhodt(:), divB(:), gradh(:)
 real(c_double), allocatable, save :: curlB(:,:), graddivv(:,:), dustfrac(:), deltav(:,:), Bfield(:,:)
 real(c_double), allocatable, save :: rhogas(:), rhodust(:)

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
    allocate(ireal(n))
    ireal = -1
    allocate(rho(n), force(3,n), dBevoldt(3,n), dudt(n))
    allocate(dens(n), uu(n), pr(n), drhodt(n), divB(n), gradh(n), curlB(3,n), graddivv(3,n), Bfield(3,n))
    allocate(dustfrac(n), deltav(3,n), rhogas(n), rhodust(n))
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
    if (any(itype == 1)) call ndsp_host_check(ndsp_set_ireal(ctx, ireal), 'ndsp_set_ireal')
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

 ! HBM -> module arrays, everything a dump holds
 subroutine ndsp_host_sync_output()
    call ndsp_host_sync()
    call get_field('pmass', pmass, npart)
    call get_field('dens', dens, npart)
    call get_field('uu', uu, npart)
    call get_field('pr', pr, npart)
    call get_field('drhodt', drhodt, npart)
    call get_field('divB', divB, npart)
    call get_field('gradh', gradh, npart)
    call get_field('curlB', curlB, 3*npart)
    call get_field('Bfield', Bfield, 3*npart)
    call get_field('graddivv', graddivv, 3*npart)
    call get_field('dustfrac', dustfrac, npart)
    call get_field('deltav', deltav, 3*npart)
    call get_field('rhogas', rhogas, npart)
    call get_field('rhodust', rhodust, npart)
 end subroutine ndsp_host_sync_output

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

!--- output: the .ev line and the dump files ---------------------------------

! One line of the .ev file (src/evwrite_mhd.f90:27): the numbers come from
! the device's FP64 reductions (ndsp_ev_columns); header and line formats are
! the reference's (src/evwrite_mhd.f90:274-308), header on the first call.
subroutine evwrite(t, etot, momtot)
 use ndsp_host
 implicit none
 real(c_double), intent(in)  :: t
 real(c_double), intent(out) :: etot, momtot
 real(c_double) :: cols(25)
 integer(c_int32_t) :: nc
 integer :: imhd, idust
 logical, save :: init = .false.
 call ndsp_host_check(ndsp_ev_columns(ctx, t, dt, cols, nc), 'evwrite')
 etot = cols(6)
 momtot = cols(7)
 imhd = iopt(3)
 idust = iopt(10)
 if (imhd /= 0) then
    if (.not.init) then
       write(ievfile,"('#',24(a18,1x))") 't','ekin','etherm','emag','epot','etot','momtot',&
          'angtot','rhomax','rhomean','dt','emagp','crosshel','betamhdmin','betamhdav',&
          'divBav','divBmax','divBtot','fdotBav','fdotBmax','force_err_av','force_err_max',&
          'omegamhdav','omegamhdmax','fracdivBok'
       init = .true.
    endif
    write(ievfile,"(24(1pe18.10,1x),1pe8.2)") cols(1:25)
 elseif (idust == 1 .or. idust == 3 .or. idust == 4) then
    if (.not.init) then
       write(ievfile,"('#',24(a18,1x))") 't','ekin','etherm','emag','epot','etot','momtot',&
          'angtot','rhomax','rhomean','dt','ekiny','totmassgas','totmassdust'
       init = .true.
    endif
    write(ievfile,"(24(1pe18.10,1x))") cols(1:14)
 else
    if (.not.init) then
       write(ievfile,"('#',24(a18,1x))") 't','ekin','etherm','emag','epot','etot','momtot',&
          'angtot','rhomax','rhomean','dt','ekiny','dmomtot'
       init = .true.
    endif
    write(ievfile,"(24(1pe18.10,1x))") cols(1:13)
 endif
 flush(ievfile)
end subroutine evwrite

! Dump files in the reference's layout (src/readwrite_dumps.f90:32-200):
! Fortran unformatted sequential records, a header record then one record
! per column, SPLASH-readable.  The real particles are written (the
! reference's idumpghost = 0 layout): image particles are not materialised
! on the device.  Geometry is cartesian; del2v (physical viscosity, off on
! this path) is written as zeros, as the reference allocates and zeroes it.
module ndsp_dumpfiles
 use iso_c_binding
 use ndsp_host
 implicit none
 integer, parameter :: ndimV = 3
contains

 subroutine dump_layout(ncolumns, iformat)
    integer, intent(out) :: ncolumns, iformat
    integer :: imhd, idust
    imhd = iopt(3)
    idust = iopt(10)
    ncolumns = ndim + 2*ndimV + 4
    if (imhd /= 0) then
       ncolumns = ncolumns + 8 + 2*ndimV
       iformat = 2
    else
       ncolumns = ncolumns + 5 + 2*ndimV   ! + del2v and graddivv
       iformat = 1
    endif
    if (idust == 1 .or. idust == 3 .or. idust == 4) then
       iformat = 5
       ncolumns = ncolumns + ndimV + 2 + 1
    endif
 end subroutine dump_layout

 subroutine write_dump(t, dumpfile)
    real(c_double), intent(in) :: t
    character(len=*), intent(in) :: dumpfile
    integer :: i, nprint, ncolumns, iformat, ierr, ib(3)
    character(len=12) :: geom
    logical :: dusty
    real(c_double), allocatable :: zeros(:)
    call ndsp_host_sync_output()
    call dump_layout(ncolumns, iformat)
    dusty = iformat == 5
    geom = 'cartesian'
    nprint = npart
    ib(1:3) = iopt(12:14)
    open(unit=idatfile, file=dumpfile, status='replace', form='unformatted', iostat=ierr)
    if (ierr /= 0) then
       write(*,*) 'error: can''t create new dumpfile ', trim(dumpfile)
       error stop 2
    endif
    write(idatfile) t, npart, nprint, dopt(3), dopt(1), ndim, ndimV, ncolumns, iformat, &
                    ib(1:ndim), dopt(14:13+ndim), dopt(17:16+ndim), len(geom), geom
    do i = 1, ndim
       write(idatfile) x(i,1:nprint)
    enddo
    do i = 1, ndimV
       write(idatfile) vel(i,1:nprint)
    enddo
    write(idatfile) hh(1:nprint)
    if (dusty) then
       write(idatfile) rho(1:nprint)
    else
       write(idatfile) dens(1:nprint)
    endif
    write(idatfile) uu(1:nprint)
    write(idatfile) pmass(1:nprint)
    if (iopt(3) /= 0) then
       do i = 1, 3
          write(idatfile) alpha(i,1:nprint)
       enddo
       do i = 1, ndimV
          write(idatfile) Bfield(i,1:nprint)
       enddo
       write(idatfile) psi(1:nprint)
       write(idatfile) pr(1:nprint)
       write(idatfile) -drhodt(1:nprint)/rho(1:nprint)
       write(idatfile) divB(1:nprint)
       do i = 1, ndimV
          write(idatfile) curlB(i,1:nprint)
       enddo
       write(idatfile) gradh(1:nprint)
       do i = 1, ndimV
          write(idatfile) force(i,1:nprint)
       enddo
    else
       do i = 1, 2
          write(idatfile) alpha(i,1:nprint)
       enddo
       write(idatfile) pr(1:nprint)
       write(idatfile) -drhodt(1:nprint)/rho(1:nprint)
       write(idatfile) gradh(1:nprint)
       do i = 1, ndimV
          write(idatfile) force(i,1:nprint)
       enddo
       allocate(zeros(nprint))
       zeros = 0.
       do i = 1, ndimV
          write(idatfile) zeros
       enddo
       do i = 1, ndimV
          write(idatfile) graddivv(i,1:nprint)
       enddo
    endif
    if (dusty) then
       write(idatfile) dustfrac(1:nprint)
       do i = 1, ndimV
          write(idatfile) deltav(i,1:nprint)
       enddo
       write(idatfile) rhogas(1:nprint)
       write(idatfile) rhodust(1:nprint)
    endif
    write(idatfile) itype(1:nprint)
    close(unit=idatfile)
 end subroutine write_dump

 ! Restart state from a dump (src/readwrite_dumps.f90:209-440): time, box
 ! and walls from the header, then the columns a run needs (x, vel, hh, uu,
 ! pmass, alpha, B, psi, dust fraction, itype); the information-only columns
 ! are skipped.  The dust columns are read where write_dump puts them (after
 ! the information columns).  Call before ndsp_host_init.
 subroutine read_dump(dumpfile, tfile)
    character(len=*), intent(in) :: dumpfile
    real(c_double), intent(out) :: tfile
    integer :: i, ndimfile, ndimvfile, npartfile, nprintfile, ncolumns, iformat, lengeom, ierr, nskip
    integer :: ib(3)
    real(c_double) :: gammafile, hfactfile, xminf(3), xmaxf(3), dummy
    character(len=12) :: geomfile
    open(unit=ireadf, file=dumpfile, status='old', form='unformatted', iostat=ierr)
    if (ierr /= 0) then
       write(*,*) 'error opening dump file: ', trim(dumpfile)
       error stop 2
    endif
    read(ireadf, iostat=ierr) tfile, npartfile, nprintfile, gammafile, hfactfile, ndimfile, ndimvfile, &
                              ncolumns, iformat, ib(1:ndim), xminf(1:ndim), xmaxf(1:ndim), lengeom, &
                              geomfile(1:12)
    if (ierr /= 0) error stop 'error reading header from dump file'
    if (ndimfile /= ndim) error stop 'x dimensions not equal between dump file and code'
    if (ndimvfile /= ndimV) error stop 'v dimensions not equal between dump file and code'
    if (npartfile /= npart) error stop 'particle number of the dump differs from the set-up'
    if (abs(gammafile - dopt(3)) > 1.e-3) write(*,"('warning: gamma changed from original setup')")
    if (abs(hfactfile - dopt(1)) > 1.e-3) write(*,"('warning: hfact changed from original setup')")
    iopt(12:11+ndim) = ib(1:ndim)
    dopt(14:13+ndim) = xminf(1:ndim)
    dopt(17:16+ndim) = xmaxf(1:ndim)
    do i = 1, ndim
       read(ireadf, iostat=ierr) x(i,1:npart)
    enddo
    do i = 1, ndimV
       read(ireadf, iostat=ierr) vel(i,1:npart)
    enddo
    read(ireadf, iostat=ierr) hh(1:npart)
    read(ireadf, iostat=ierr) dens(1:npart)
    read(ireadf, iostat=ierr) uu(1:npart)
    read(ireadf, iostat=ierr) pmass(1:npart)
    if (ierr /= 0) error stop 'error reading dumpfile'
    alpha = 0.
    Bevol = 0.
    psi = 0.
    if (iformat == 2 .or. iformat == 4) then
       do i = 1, 3
          read(ireadf, iostat=ierr) alpha(i,1:npart)
       enddo
       do i = 1, ndimV
          read(ireadf, iostat=ierr) Bevol(i,1:npart)
       enddo
       read(ireadf, iostat=ierr) psi(1:npart)
       nskip = 4 + 2*ndimV
    else
       do i = 1, 2
          read(ireadf, iostat=ierr) alpha(i,1:npart)
       enddo
       nskip = 3 + 3*ndimV
    endif
    if (ierr /= 0) error stop 'error reading dumpfile'
    do i = 1, nskip
       read(ireadf, iostat=ierr) dummy
    enddo
    dustevol = 0.
    if (iformat == 5) then
       read(ireadf, iostat=ierr) dustevol(1:npart)
       do i = 1, ndimV + 2
          read(ireadf, iostat=ierr) dummy
       enddo
    endif
    read(ireadf, iostat=ierr) itype(1:npart)
    if (ierr /= 0) write(*,*) 'WARNING: itype array not present in file'
    close(unit=ireadf)
    en = uu
 end subroutine read_dump

end module ndsp_dumpfiles
