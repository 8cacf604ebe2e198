!------------------------------------------------------------------------------
! Fortran driver of the MI355X hot path: the time loop of the reference's
! evolve (src/evolve.f90:93-175) around the argument-less derivs/step entry
! points of ndsp_host, on one GPU.
!
!   ndsp_run <input.bin> <output.bin> <nsteps> [tmax] [precision]
!
! input.bin has the layout of the reference harness input (options iopt/dopt
! and the particle state; tests/golden/refio.py).  Each step prints an .ev
! style line (time, ekin, etherm, emag, etot, momtot; src/evwrite_mhd.f90).
! output.bin: int32 npart, ndim, nsteps, 0; float64 time, dt; then x, vel,
! Bevol, hh, en, alpha, psi, dustevol, rho in the reference layout.
!------------------------------------------------------------------------------
program ndsp_run
 use iso_c_binding
 use ndsp_capi
 use ndsp_host
 implicit none
 character(len=256) :: infile, outfile, arg
 integer :: magic, mode, np, nd, ierr, nmax, nsteps, prec
 real(c_double) :: tmax, cons(17)
 integer(c_int) :: rc

 call get_command_argument(1, infile)
 call get_command_argument(2, outfile)
 call get_command_argument(3, arg)
 read(arg, *) nmax
 tmax = huge(tmax)
 call get_command_argument(4, arg)
 if (len_trim(arg) > 0) read(arg, *) tmax
 prec = NDSP_PREC_MIXED
 call get_command_argument(5, arg)
 if (len_trim(arg) > 0) read(arg, *) prec

 open(unit=21, file=trim(infile), access='stream', form='unformatted', status='old', iostat=ierr)
 if (ierr /= 0) error stop 'ndsp_run: cannot open input'
 read(21) magic, mode, np, nd
 if (magic /= 1313100880) error stop 'ndsp_run: bad magic'
 read(21) iopt
 read(21) dopt
 call ndsp_host_alloc(np, nd)
 read(21) x, vel, pmass, hh, en, Bevol, alpha, psi, dustevol, itype
 close(21)

 call ndsp_host_init(prec)
!--initial derivative evaluation (as primitive2conservative,
!  src/conservative2primitive.f90:799)
 call derivs
 nsteps = 0
 rc = ndsp_conserved(ctx, cons)
 write(*, '(a,6(1x,es18.10))') ' ev', time, cons(1), cons(2), cons(3), cons(4), cons(8)
!--time loop, src/evolve.f90:93-175
 do while (time < tmax .and. nsteps < nmax)
    time = time + dt
    nsteps = nsteps + 1
    call step
    rc = ndsp_conserved(ctx, cons)
    write(*, '(a,6(1x,es18.10))') ' ev', time, cons(1), cons(2), cons(3), cons(4), cons(8)
 enddo
 call ndsp_host_sync()
 open(unit=22, file=trim(outfile), access='stream', form='unformatted', status='replace')
 write(22) int(npart, c_int32_t), int(ndim, c_int32_t), int(nsteps, c_int32_t), 0_c_int32_t
 write(22) time, dt
 write(22) x, vel, Bevol, hh, en, alpha, psi, dustevol, rho
 close(22)
 call ndsp_host_finish()
end program ndsp_run
