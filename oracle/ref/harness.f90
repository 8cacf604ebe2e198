!------------------------------------------------------------------------------
! Test-only main program for the compiled reference (ndspmhd src/, built unmodified
! from the reference tree by oracle/ref/Makefile).  It is the oracle's
! checker: it writes golden vectors for tests/golden, nothing else uses it.
!
! It plays the part of the reference's main program: it fills the module
! globals the way initialise (src/initialiseND_mhd.f90:140-300) does for the
! options read from an input file, calls the reference's own derivs
! (src/derivs.f90:27) once, and writes the results.  Mode 2 additionally
! enumerates neighbour sets with the reference's link list
! (src/linkND.f90:45, src/get_neighbour_lists.f90:176) and the distance tests
! of density_partial (src/density_sums.f90:499-528) and get_rates
! (src/ratesND_mhd.f90:415).
!
! The reference's quit (src/ndspmhd.f90:369) lives beside its main program,
! which this program replaces, so the program supplies its own quit that stops
! with a non-zero status.
!
! Input / output: little-endian stream files, layout documented in
! tests/golden/refio.py; the .ev text of the state goes to <output>.ev and
! the options, as write_infile writes an input file, to <output>.in.
!------------------------------------------------------------------------------
program ndsp_ref_harness
 include 'harness_decl.inc'

 include 'harness_setup.inc'
!
!--the hot path: one derivative evaluation (wall time printed for the CPU
!  baseline)
!
 call system_clock(count0,count_rate)
 call derivs
 call system_clock(count1)
 write(*,"(a,es16.8)") ' NDSP_DERIVS_SECONDS ',real(count1-count0)/real(count_rate)
!
!--results
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

 if (mode == 2) then
!
!--neighbour sets of the real particles, found through the reference's own
!  cells (full stencil) with the gather test of density_partial and the
!  symmetric test of get_rates, at the h the iteration finished with
!
    allocate(listneigh(ntotal),pg(3,400*np+64),ps(3,400*np+64))
    npg = 0
    nps = 0
    do i=1,npart
       call get_neighbour_list_partial(iamincell(i),listneigh,nneigh)
       hi1 = 1./hh(i)
       hi21 = hi1*hi1
       do n=1,nneigh
          j = listneigh(n)
          dx(:) = x(:,i) - x(:,j)
          rij2 = dot_product(dx,dx)
          q2i = rij2*hi21
          hj1 = 1./hh(j)
          q2j = rij2*hj1*hj1
          ip = j
          icode = 0
          if (j > npart) then
             ip = ireal(j)
             do k=ndim,1,-1
                icode = 3*icode
                if (x(k,j) > x(k,ip)) icode = icode + 1
                if (x(k,j) < x(k,ip)) icode = icode + 2
             enddo
          endif
          if (q2i < radkern2) then
             npg = npg + 1
             if (npg > size(pg,2)) error stop 'pair buffer overflow'
             pg(:,npg) = (/i,ip,icode/)
          endif
          if (q2i < radkern2 .or. q2j < radkern2) then
             nps = nps + 1
             if (nps > size(ps,2)) error stop 'pair buffer overflow'
             ps(:,nps) = (/i,ip,icode/)
          endif
       enddo
    enddo
    write(22) npg
    write(22) pg(1:3,1:npg)
    write(22) nps
    write(22) ps(1:3,1:nps)
 endif
 close(22)
!
!--the .ev line of this state, written by the reference's own evwrite
!  (src/evwrite_mhd.f90:27) at time 0 with the input dt
!
 open(unit=ievfile,file=trim(outfile)//'.ev',status='replace',form='formatted')
 call evwrite(time,etot,momtot)
 close(ievfile)
!
!--the options as the reference's own write_infile writes them
!  (src/readwrite_infile.f90:35)
!
 call write_infile(trim(outfile)//'.in')
!--and values that take the writer's other branches (es10.3 below 1e-3 and
!  from 1e4, g16.9 from 10 with all zeros stripped, long and negative values)
 psep = 36.87
 tmax = 0.075
 tout = 1.23456789012
 polyk = 2.5e-4
 C_cour = -0.3
 hsoft = 12345.678
 etamhd = 1.0e-12
 avdecayconst = 0.0999999
 beta = 10.0
 rhocrit = 99.5
 Kdrag = 20.
 alphamin = 0.123456789
 call write_infile(trim(outfile)//'.in2')

end program ndsp_ref_harness

subroutine quit
 use loguns, only:iprint
 write(iprint,*) 'ndsp_ref_harness: reference called quit'
 stop 3
end subroutine quit
