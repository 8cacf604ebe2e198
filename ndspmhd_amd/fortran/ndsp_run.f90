!------------------------------------------------------------------------------
! Fortran host program of the MI355X hot path: the time loop of the reference's
! evolve (src/evolve.f90:52-190) around the argument-less derivs/step entry
! points of ndsp_host, on one GPU, with the reference's outputs: the .ev
! file (one line per step, evwrite) and dump files (write_dump), and
! restarts from a dump (read_dump).
!
!   ndsp_run <input.bin> <output.bin> <nsteps> [tmax] [precision] [rootname]
!            [nout] [restart.dat]
!
! input.bin has the layout of the reference harness input (options iopt/dopt,
! the particle state and the partners of fixed particles;
! tests/golden/refio.py).  With a rootname the run writes rootname.ev and the dumps rootname_00000.dat (initial state),
! then one every nout steps (nout > 0) and one after the last step, named as
! the reference's output does (src/outputND_mhd.f90:54).  restart.dat
! replaces the state and time of input.bin by those of a dump.  Every step
! also prints an ' ev' line (time, ekin, etherm, emag, etot, momtot).
! output.bin: int32 npart, ndim, nsteps, 0; float64 time, dt; then x, vel,
! Bevol, hh, en, alpha, psi, dustevol, rho in the reference layout.
!------------------------------------------------------------------------------
program ndsp_run
 use iso_c_binding
 use ndsp_capi
 use ndsp_host
 use ndsp_dumpfiles
 implicit none
 character(len=256) :: infile, outfile, arg, rootname, restart
 integer :: magic, mode, np, nd, ierr, nmax, nsteps, prec, nout, ifile
 real(c_double) :: tmax, tfile, etot, momtot, etotin, momtotin, detot, dmomtot, cons(17)
 integer(c_int) :: rc
 logical :: dumped

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
 call get_command_argument(6, rootname)
 nout = 0
 call get_command_argument(7, arg)
 if (len_trim(arg) > 0) read(arg, *) nout
 call get_command_argument(8, restart)

 open(unit=21, file=trim(infile), access='stream', form='unformatted', status='old', iostat=ierr)
 if (ierr /= 0) error stop 'ndsp_run: cannot open input'
 read(21) magic, mode, np, nd
 if (magic /= 1313100880) error stop 'ndsp_run: bad magic'
 read(21) iopt
 read(21) dopt
 call ndsp_host_alloc(np, nd)
 read(21) x, vel, pmass, hh, en, Bevol, alpha, psi, dustevol, itype, ireal
 close(21)
 tfile = 0.
 if (len_trim(restart) > 0) call read_dump(trim(restart), tfile)

 call ndsp_host_init(prec)
 if (tfile /= 0.) then
    time = tfile
    call ndsp_host_check(ndsp_set_dt(ctx, dt, time), 'ndsp_set_dt')
 endif
!--initial derivative evaluation (as primitive2conservative,
!  src/conservative2primitive.f90:799)
 call derivs
 nsteps = 0
 ifile = -1
 detot = 0.
 dmomtot = 0.
 if (len_trim(rootname) > 0) then
    open(unit=ievfile, file=trim(rootname)//'.ev', status='replace', form='formatted')
    call evwrite(time, etotin, momtotin)
    write(*,"(' Total energy = ',1pe10.2,2x,'Linear momentum = ',1pe10.2)") etotin, momtotin
    call output_dump()
 endif
 rc = ndsp_conserved(ctx, cons)
 write(*, '(a,6(1x,es18.10))') ' ev', time, cons(1), cons(2), cons(3), cons(4), cons(8)
!--time loop, src/evolve.f90:93-175
 do while (time < tmax .and. nsteps < nmax)
    time = time + dt
    nsteps = nsteps + 1
    call step
    dumped = .false.
    if (len_trim(rootname) > 0) then
       if ((nout > 0 .and. mod(nsteps, nout) == 0) .or. nsteps >= nmax .or. time >= tmax) then
          call output_dump()
          dumped = .true.
       endif
       call evwrite(time, etot, momtot)
       detot = max(detot, abs(etot - etotin))
       dmomtot = max(dmomtot, abs(momtot - momtotin))
    endif
    rc = ndsp_conserved(ctx, cons)
    write(*, '(a,6(1x,es18.10))') ' ev', time, cons(1), cons(2), cons(3), cons(4), cons(8)
 enddo
 if (len_trim(rootname) > 0) then
    if (.not. dumped .and. nsteps > 0) call output_dump()
    close(ievfile)
!--conservation summary, src/evolve.f90:183-189
    if (abs(momtotin) < tiny(momtotin)) momtotin = 1.0
    if (abs(etotin) < tiny(etotin)) etotin = 1.0
    write(*,"(/,' Max energy error   : ',1pe10.2,3x,' Max momentum error : ',1pe10.2)") &
          detot/abs(etotin), dmomtot/abs(momtotin)
 endif
 call ndsp_host_sync()
 open(unit=22, file=trim(outfile), access='stream', form='unformatted', status='replace')
 write(22) int(npart, c_int32_t), int(ndim, c_int32_t), int(nsteps, c_int32_t), 0_c_int32_t
 write(22) time, dt
 write(22) x, vel, Bevol, hh, en, alpha, psi, dustevol, rho
 close(22)
 call ndsp_host_finish()

contains

 ! src/outputND_mhd.f90:51-66: next file number, name rootname_NNNNN.dat
 subroutine output_dump()
    character(len=300) :: dumpfile
    ifile = ifile + 1
    write(dumpfile, "(a,'_',i5.5,'.dat')") trim(rootname), ifile
    write(*, "('| t = ',g10.3,' | step = ',i8,' | npart = ',i7,' | nghost = ',i5,' |',a)") &
          time, nsteps, npart, 0, trim(dumpfile)
    call write_dump(time, trim(dumpfile))
 end subroutine output_dump

end program ndsp_run
