!------------------------------------------------------------------------------
! Test-only main program for the reference-side binding of the GPU hot path
! (ndspmhd_amd/fortran/ndsp_refbind.f90).  Built by oracle/ref/Makefile from
! the unmodified reference sources with the hot-path files (linkND.f90,
! iterate_density.f90, ratesND_mhd.f90, derivs.f90) left out and the binding
! linked in their place, as the reference swaps implementations by file
! selection (src/Makefile:227,263).  Everything else it runs is the
! reference's own code: its module globals, and its evwrite
! (src/evwrite_mhd.f90:27), which writes the .ev file from the module arrays
! the binding fills.
!
! It fills the globals as harness.f90 does (shared include files), runs the
! reference's primitive2conservative as initialise does, then the loop of
! the reference's evolve (src/evolve.f90:53-106): evwrite at the
! start, then nsteps times time = time + dt, step, evwrite, with dt = 0 for
! the first step.  It writes the final state in harness.f90's output layout
! followed by the step count, time and dt (tests/golden/refio.py, mode 3).
!
!   ndsp_bind3d <input> <output> <nsteps>      (.ev text to <output>.ev)
!------------------------------------------------------------------------------
program ndsp_bind_harness
 use cons2prim, only:primitive2conservative
 include 'harness_decl.inc'
 integer :: nrun, istep
 character(len=32) :: arg

 include 'harness_setup.inc'
 call get_command_argument(3,arg)
 read(arg,*) nrun
!
!--initialise's last part (src/initialiseND_mhd.f90:281-295): the set-up
!  gives primitive variables (the density that the input h implies, u, B)
!  and the reference's primitive2conservative
!  (src/conservative2primitive.f90:478-810) sets h from them, calls
!  set_linklist and iterate_density (the binding: the device context is
!  created there), forms the conserved variables and calls derivs
!
 dens(1:np) = pmass(1:np)*(hfact/hh(1:np))**ndim
 uu(1:np) = en(1:np)
 if (imhd /= 0) Bfield(:,1:np) = Bevol(:,1:np)
 call primitive2conservative
 open(unit=ievfile,file=trim(outfile)//'.ev',status='replace',form='formatted')
 dt = 0.
 call evwrite(time,etot,momtot)
 do istep = 1,nrun
    time = time + dt
    nsteps = nsteps + 1
    call step
    call evwrite(time,etot,momtot)
 enddo
 close(ievfile)
!
!--the final state, as harness.f90 writes its results
!
 open(unit=22,file=trim(outfile),access='stream',form='unformatted',status='replace')
 write(22) npart,ntotal,itsdensity,0
 write(22) dtcourant,dtforce,dtdrag,dtav,vsig2max,hhmax
 write(22) x(1:ndim,1:np)
 write(22) hh(1:np)
 write(22) rho(1:np)
 write(22) gradh(1:np)
 write(22) gradgradh(1:np)
 write(22) drhodt(1:np)
 write(22) dhdt(1:np)
 write(22) dens(1:np)
 write(22) pr(1:np)
 write(22) spsound(1:np)
 write(22) uu(1:np)
 write(22) psi(1:np)
 write(22) force(1:3,1:np)
 write(22) dudt(1:np)
 write(22) dendt(1:np)
 write(22) dBevoldt(1:3,1:np)
 write(22) daldt(1:3,1:np)
 write(22) dpsidt(1:np)
 write(22) divB(1:np)
 write(22) curlB(1:3,1:np)
 write(22) gradpsi(1:3,1:np)
 if (onef_dust) then
    write(22) dustfrac(1,1:np)
    write(22) rhogas(1:np)
    write(22) rhodust(1,1:np)
    write(22) ddustevoldt(1,1:np)
    write(22) deltav(1:3,1:np)
 else
    write(22) zeros,zeros,zeros,zeros,zeros3
 endif
 write(22) graddivv(1:3,1:np)
 write(22) numneigh(1:np)
 write(22) nsteps
 write(22) time,dt
 write(22) vel(1:3,1:np)
 write(22) en(1:np)
 write(22) Bevol(1:3,1:np)
 write(22) alpha(1:3,1:np)
 close(22)

end program ndsp_bind_harness

subroutine quit
 use loguns, only:iprint
 write(iprint,*) 'ndsp_bind_harness: quit called'
 stop 3
end subroutine quit
