!------------------------------------------------------------------------------
! Reference-side binding of the MI355X hot path.
!
! Supplies the reference's argument-less hot-path entry points
!   set_linklist    (src/linkND.f90:45)
!   iterate_density (src/iterate_density.f90:43)
!   get_rates       (src/ratesND_mhd.f90:29)
!   derivs          (src/derivs.f90:27)
!   step            (src/stepND_leapfrog_mhd.f90:31)
! written against the reference's own module globals (src/variablesND.f90:
! part, rates, hterms, linklist, derivB, timestep, options, artvi, bound,
! setup_params, dimen_mhd; eos in src/eos.f90:34-36).  A reference build that
! links this file and ndsp_capi.f90 in place of linkND.f90,
! iterate_density.f90, ratesND_mhd.f90, derivs.f90 and stepND_*.f90 (the
! reference swaps implementations by file selection at link time,
! src/Makefile:227,263) runs its hot path on the GPU with every call site
! unchanged.  Compile with the reference's flags (-fdefault-real-8
! -fdefault-double-8, so that real is c_double).
!
! Ownership and synchronisation:
! - the device context is created on the first call of any entry (in the
!   reference: set_linklist from primitive2conservative,
!   src/conservative2primitive.f90:557) from the option globals as they are
!   then; options the device path does not implement stop the run;
! - set_linklist and derivs take the host's state, because their callers set
!   it just before (primitive2conservative sets h, initialise sets the
!   conserved variables); iterate_density, get_rates and step continue from
!   the state held in HBM;
! - after each entry the results the reference routine would have written go
!   back into the module arrays: iterate_density its h, rho, gradh,
!   gradgradh, drho/dt, dh/dt, psi, numneigh and itsdensity; get_rates the
!   rates, divB, curlB, gradpsi and the timestep criteria; derivs both plus
!   the primitive variables of conservative2primitive; step the whole state,
!   because evolve's evwrite and output read it every step
!   (src/evolve.f90:106-161).  ndsp_ref_set_sync(.false.) keeps the state in
!   HBM between steps; the host then calls ndsp_ref_sync before reading it;
! - step uses timestep:dt for the step and leaves the next timestep in it
!   (src/stepND_leapfrog_mhd.f90:252); evolve advances time itself
!   (src/evolve.f90:93).
! Precision: NDSP_PRECISION=fp64 in the environment selects the plain FP64
! kernels; the default is the mixed-precision path.  A non-zero status stops
! the run through the reference's quit (src/ndspmhd.f90:369).
! All state of this module is private.
!------------------------------------------------------------------------------
module ndsp_refbind
 use iso_c_binding
 use ndsp_capi
 implicit none
 private
 public :: ndsp_ref_set_linklist, ndsp_ref_iterate_density, ndsp_ref_get_rates
 public :: ndsp_ref_derivs, ndsp_ref_step, ndsp_ref_sync, ndsp_ref_set_sync, ndsp_ref_release

 type(c_ptr), save :: ctx = c_null_ptr
 integer, save :: nctx = 0, ndimctx = 0
 logical, save :: sync_every_step = .true.
 type(ndsp_stats_t), save :: stats
 type(ndsp_dt_t), save :: dtout
 real(c_double), allocatable, save :: buf(:)

contains

 subroutine check(ierr, what)
    use loguns, only:iprint
    integer(c_int), intent(in) :: ierr
    character(len=*), intent(in) :: what
    if (ierr /= 0) then
       write(iprint,'(a,a,a,i3,a,a)') ' ndsp: ', what, ' failed, status ', ierr, ': ', &
            trim(ndsp_error_text())
       call quit
    endif
 end subroutine check

 subroutine refuse(what)
    use loguns, only:iprint
    character(len=*), intent(in) :: what
    write(iprint,'(a,a)') ' ndsp: not implemented on the GPU path: ', what
    call quit
 end subroutine refuse

 ! the options of the reference's globals, as the C-ABI's option vectors
 ! (include/ndsp.h); the switches whose other branches the device path does
 ! not implement must have the values it assumes (src/defaults.f90:29-120)
 subroutine options_vectors(iopt, dopt)
    use dimen_mhd,    only:ndim
    use options,      only:ikernel,ikernelalt,iener,imhd,imagforce,idivbzero,iav,iavlim,idust, &
                           idrag_nature,ibound,maxdensits,idustevol,psidecayfact,tolh,Kdrag,damp, &
                           icty,ikernav,ihvar,iprterm,ivisc,iresist,igravity,iexternal_force, &
                           ixsph,iquantum,ind_timesteps,iambipolar,idiffuse,ibiascorrection, &
                           icompute_d2v,usenumdens,geom
    use artvi,        only:alphamin,alphaumin,alphabmin,beta,avdecayconst
    use eos,          only:gamma,polyk
    use setup_params, only:hfact,psep
    use timestep,     only:C_cour,C_force,dt
    use bound,        only:xmin,xmax
    use part,         only:Bconst
    integer(c_int32_t), intent(out) :: iopt(32)
    real(c_double),     intent(out) :: dopt(32)
    integer :: d
    if (icty /= 0) call refuse('icty /= 0')
    if (ikernav /= 3) call refuse('ikernav /= 3')
    if (ihvar /= 2) call refuse('ihvar /= 2')
    if (ikernelalt /= ikernel) call refuse('ikernelalt /= ikernel')
    if (iprterm /= 0) call refuse('iprterm /= 0')
    if (ivisc /= 0) call refuse('ivisc /= 0')
    if (iresist /= 0) call refuse('iresist /= 0')
    if (igravity /= 0) call refuse('igravity /= 0')
    if (iexternal_force /= 0) call refuse('iexternal_force /= 0')
    if (ixsph /= 0) call refuse('ixsph /= 0')
    if (iquantum /= 0) call refuse('iquantum /= 0')
    if (ind_timesteps /= 0) call refuse('ind_timesteps /= 0')
    if (iambipolar /= 0) call refuse('iambipolar /= 0')
    if (idiffuse /= 0) call refuse('idiffuse /= 0')
    if (ibiascorrection /= 0) call refuse('ibiascorrection /= 0')
    if (icompute_d2v /= 0) call refuse('icompute_d2v /= 0')
    if (usenumdens) call refuse('usenumdens')
    if (trim(geom) /= 'cartesian' .and. len_trim(geom) > 0) call refuse('geom '//trim(geom))
    iopt = 0
    dopt = 0.
    iopt(1) = ikernel
    iopt(2) = iener
    iopt(3) = imhd
    iopt(4) = imagforce
    iopt(5) = idivbzero
    iopt(6) = iav
    iopt(7:9) = iavlim(1:3)
    iopt(10) = idust
    iopt(11) = idrag_nature
    do d = 1, ndim
       iopt(11+d) = ibound(d)
    enddo
    iopt(15) = maxdensits
    iopt(16) = idustevol
    iopt(17) = ndim
    dopt(1) = hfact
    dopt(2) = tolh
    dopt(3) = gamma
    dopt(4) = polyk
    dopt(5) = Kdrag
    dopt(6) = alphamin
    dopt(7) = alphaumin
    dopt(8) = alphabmin
    dopt(9) = beta
    dopt(10) = avdecayconst
    dopt(11) = psidecayfact
    dopt(12) = C_cour
    dopt(13) = C_force
    do d = 1, ndim
       dopt(13+d) = xmin(d)
       dopt(16+d) = xmax(d)
    enddo
    dopt(20) = psep
    do d = 1, min(3, size(Bconst))
       dopt(20+d) = Bconst(d)
    enddo
    dopt(24) = dt
    dopt(25) = damp
 end subroutine options_vectors

 ! create the context on first use (or again if npart changed)
 subroutine ensure_context()
    use dimen_mhd, only:ndim
    use part,      only:npart
    integer(c_int32_t) :: iopt(32), prec
    real(c_double)     :: dopt(32)
    character(len=16)  :: env
    integer :: ln, st
    if (c_associated(ctx) .and. nctx == npart) return
    call ndsp_ref_release()
    call options_vectors(iopt, dopt)
    prec = NDSP_PREC_MIXED
    call get_environment_variable('NDSP_PRECISION', env, ln, st)
    if (st == 0 .and. (env(1:4) == 'fp64' .or. env(1:4) == 'FP64')) prec = NDSP_PREC_FP64
    call check(ndsp_create(iopt, dopt, int(npart, c_int64_t), prec, ctx), 'ndsp_create')
    nctx = npart
    ndimctx = ndim
    allocate(buf(3*npart))
 end subroutine ensure_context

 ! host state (real particles) -> HBM, in the C-ABI's layout
 subroutine upload_state()
    use dimen_mhd, only:ndim
    use part,      only:npart,x,vel,pmass,hh,en,Bevol,alpha,psi,dustevol,itype,itypebnd
    use bound,     only:ireal
    use timestep,  only:dt,time
    real(c_double), allocatable :: xs(:,:), vs(:,:), bs(:,:), as(:,:), ds(:)
    integer(c_int32_t), allocatable :: it(:), ir(:)
    integer :: i, n
    n = npart
    allocate(xs(ndim,n), vs(3,n), bs(3,n), as(3,n), ds(n), it(n))
    xs(:,1:n) = x(1:ndim,1:n)
    vs = 0.
    vs(1:min(3,size(vel,1)),1:n) = vel(1:min(3,size(vel,1)),1:n)
    bs = 0.
    bs(1:min(3,size(Bevol,1)),1:n) = Bevol(1:min(3,size(Bevol,1)),1:n)
    as = 0.
    as(1:min(3,size(alpha,1)),1:n) = alpha(1:min(3,size(alpha,1)),1:n)
    ds = 0.
    if (allocated(dustevol)) ds(1:n) = dustevol(1,1:n)
    it(1:n) = itype(1:n)
    call check(ndsp_upload(ctx, xs, vs, pmass, hh, en, bs, as, psi, ds, it), 'ndsp_upload')
    if (any(itype(1:n) == itypebnd)) then
       allocate(ir(n))
       do i = 1, n
          ir(i) = ireal(i) - 1   ! 0-based, -1 none
       enddo
       call check(ndsp_set_ireal(ctx, ir), 'ndsp_set_ireal')
    endif
    call check(ndsp_set_dt(ctx, dt, time), 'ndsp_set_dt')
 end subroutine upload_state

 ! one field of the real particles from HBM; w components per particle
 subroutine fetch(name, w)
    character(len=*), intent(in) :: name
    integer, intent(in) :: w
    integer(c_int32_t) :: wd
    call check(ndsp_download(ctx, trim(name)//c_null_char, buf, wd), 'ndsp_download '//name)
    if (wd /= w) call refuse('width of '//name)
 end subroutine fetch

 subroutine get1(name, a)
    use part, only:npart
    character(len=*), intent(in) :: name
    real, intent(inout) :: a(:)
    call fetch(name, 1)
    a(1:npart) = buf(1:npart)
 end subroutine get1

 subroutine get3(name, a)
    use part, only:npart
    character(len=*), intent(in) :: name
    real, intent(inout) :: a(:,:)
    integer :: i, c
    call fetch(name, 3)
    do i = 1, npart
       do c = 1, min(3, size(a,1))
          a(c,i) = buf(3*(i-1)+c)
       enddo
    enddo
 end subroutine get3

 ! what iterate_density writes (src/iterate_density.f90:95-344)
 subroutine fetch_density()
    use part,     only:npart,hh,rho,psi
    use hterms,   only:gradh,gradgradh,itsdensity
    use rates,    only:drhodt,dhdt
    use linklist, only:numneigh
    use bound,    only:hhmax
    call get1('hh', hh)
    call get1('rho', rho)
    call get1('gradh', gradh)
    call get1('gradgradh', gradgradh)
    call get1('drhodt', drhodt)
    call get1('dhdt', dhdt)
    call get1('psi', psi)
    call fetch('numneigh', 1)
    numneigh(1:npart) = nint(buf(1:npart))
    itsdensity = stats%itsdensity
    hhmax = maxval(hh(1:npart))
 end subroutine fetch_density

 ! what conservative2primitive writes (src/conservative2primitive.f90:42-470)
 subroutine fetch_primitives()
    use part,    only:dens,pr,spsound,uu,Bfield,dustfrac,rhogas,rhodust
    use options, only:imhd,onef_dust
    call get1('dens', dens)
    call get1('pr', pr)
    call get1('spsound', spsound)
    call get1('uu', uu)
    if (imhd /= 0) call get3('Bfield', Bfield)
    if (onef_dust) then
       if (allocated(dustfrac)) call get1('dustfrac', dustfrac(1,:))
       if (allocated(rhogas)) call get1('rhogas', rhogas)
       if (allocated(rhodust)) call get1('rhodust', rhodust(1,:))
    endif
 end subroutine fetch_primitives

 ! what get_rates writes (src/ratesND_mhd.f90:29-979)
 subroutine fetch_rates()
    use part,     only:deltav,graddivv
    use rates,    only:force,dudt,dendt,dBevoldt,daldt,dpsidt,gradpsi,ddustevoldt
    use derivB,   only:divB,curlB
    use options,  only:onef_dust
    use timestep, only:dtcourant,dtforce,dtdrag,vsig2max
    call get3('force', force)
    call get1('dudt', dudt)
    call get1('dendt', dendt)
    call get3('dBevoldt', dBevoldt)
    call get3('daldt', daldt)
    call get1('dpsidt', dpsidt)
    call get1('divB', divB)
    call get3('curlB', curlB)
    call get3('gradpsi', gradpsi)
    if (allocated(graddivv)) call get3('graddivv', graddivv)
    if (onef_dust) then
       if (allocated(ddustevoldt)) call get1('ddustevoldt', ddustevoldt(1,:))
       if (allocated(deltav)) call get3('deltav', deltav)
    endif
    dtcourant = dtout%dtcourant
    dtforce = dtout%dtforce
    dtdrag = dtout%dtdrag
    vsig2max = dtout%vsig2max
 end subroutine fetch_rates

 ! the evolved state
 subroutine fetch_state()
    use dimen_mhd, only:ndim
    use part,      only:npart,x,vel,hh,en,Bevol,alpha,psi,dustevol
    use options,   only:onef_dust,imhd
    integer :: i, d
    call fetch('x', ndim)
    do i = 1, npart
       do d = 1, ndim
          x(d,i) = buf(ndim*(i-1)+d)
       enddo
    enddo
    call get3('vel', vel)
    call get1('hh', hh)
    call get1('en', en)
    if (imhd /= 0) call get3('Bevol', Bevol)
    call get3('alpha', alpha)
    call get1('psi', psi)
    if (onef_dust .and. allocated(dustevol)) call get1('dustevol', dustevol(1,:))
 end subroutine fetch_state

 subroutine ndsp_ref_set_linklist()
    call ensure_context()
    call upload_state()
    call check(ndsp_set_linklist(ctx), 'set_linklist')
 end subroutine ndsp_ref_set_linklist

 subroutine ndsp_ref_iterate_density()
    if (.not. c_associated(ctx)) call ndsp_ref_set_linklist()
    call check(ndsp_iterate_density(ctx, stats), 'iterate_density')
    call fetch_density()
 end subroutine ndsp_ref_iterate_density

 subroutine ndsp_ref_get_rates()
    if (.not. c_associated(ctx)) call refuse('get_rates before any density evaluation')
    call check(ndsp_get_rates(ctx, dtout), 'get_rates')
    call fetch_rates()
 end subroutine ndsp_ref_get_rates

 subroutine ndsp_ref_derivs()
    call ensure_context()
    call upload_state()
    call check(ndsp_derivs(ctx, stats, dtout), 'derivs')
    call fetch_density()
    call fetch_primitives()
    call fetch_rates()
 end subroutine ndsp_ref_derivs

 subroutine ndsp_ref_step()
    use timestep, only:dt,time
    if (.not. c_associated(ctx)) call ndsp_ref_derivs()
    call check(ndsp_set_dt(ctx, dt, time - dt), 'ndsp_set_dt')
    call check(ndsp_step(ctx, stats, dtout), 'step')
    dt = dtout%dt
    if (sync_every_step) call ndsp_ref_sync()
 end subroutine ndsp_ref_step

 ! HBM -> the module arrays (state, density and primitive fields, rates)
 subroutine ndsp_ref_sync()
    if (.not. c_associated(ctx)) return
    call fetch_state()
    call fetch_density()
    call fetch_primitives()
    call fetch_rates()
 end subroutine ndsp_ref_sync

 subroutine ndsp_ref_set_sync(every_step)
    logical, intent(in) :: every_step
    sync_every_step = every_step
 end subroutine ndsp_ref_set_sync

 subroutine ndsp_ref_release()
    integer(c_int) :: ierr
    if (c_associated(ctx)) ierr = ndsp_destroy(ctx)
    ctx = c_null_ptr
    nctx = 0
    if (allocated(buf)) deallocate(buf)
 end subroutine ndsp_ref_release

end module ndsp_refbind

!--- the reference's argument-less entry points --------------------------------

subroutine set_linklist
 use ndsp_refbind, only:ndsp_ref_set_linklist
 implicit none
 call ndsp_ref_set_linklist()
end subroutine set_linklist

subroutine iterate_density
 use ndsp_refbind, only:ndsp_ref_iterate_density
 implicit none
 call ndsp_ref_iterate_density()
end subroutine iterate_density

subroutine get_rates
 use ndsp_refbind, only:ndsp_ref_get_rates
 implicit none
 call ndsp_ref_get_rates()
end subroutine get_rates

subroutine derivs
 use ndsp_refbind, only:ndsp_ref_derivs
 implicit none
 call ndsp_ref_derivs()
end subroutine derivs

subroutine step
 use ndsp_refbind, only:ndsp_ref_step
 implicit none
 call ndsp_ref_step()
end subroutine step
