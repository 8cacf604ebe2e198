!------------------------------------------------------------------------------
! iso_c_binding interfaces to the C-ABI of libndsp_hip.so (include/ndsp.h).
! Every function returns an ndsp_status (0 = ok); ndsp_last_error gives the
! text of the last failure.
!------------------------------------------------------------------------------
module ndsp_capi
 use iso_c_binding
 implicit none

 integer(c_int32_t), parameter :: NDSP_PREC_MIXED = 0, NDSP_PREC_FP64 = 1

 type, bind(c) :: ndsp_stats_t
    integer(c_int32_t) :: itsdensity, nmin, nmax, pad
    real(c_double)     :: ncalc_mean
    integer(c_int64_t) :: pairs_density, pairs_force
 end type ndsp_stats_t

 type, bind(c) :: ndsp_dt_t
    real(c_double) :: dt, dtcourant, dtforce, dtdrag, vsig2max, hhmax, stressmax, time
 end type ndsp_dt_t

 interface
    integer(c_int) function ndsp_create(iopt, dopt, npart, precision, ctx) bind(c, name='ndsp_create')
       import :: c_int, c_int32_t, c_int64_t, c_double, c_ptr
       integer(c_int32_t), intent(in) :: iopt(32)
       real(c_double),     intent(in) :: dopt(32)
       integer(c_int64_t), value      :: npart
       integer(c_int32_t), value      :: precision
       type(c_ptr),        intent(out):: ctx
    end function ndsp_create

    integer(c_int) function ndsp_destroy(ctx) bind(c, name='ndsp_destroy')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
    end function ndsp_destroy

    integer(c_int) function ndsp_upload(ctx, x, vel, pmass, hh, en, Bevol, alpha, psi, dustevol, itype) &
                   bind(c, name='ndsp_upload')
       import :: c_int, c_ptr, c_double, c_int32_t
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: x(*), vel(*), pmass(*), hh(*), en(*), Bevol(*), alpha(*), psi(*), dustevol(*)
       integer(c_int32_t), intent(in) :: itype(*)
    end function ndsp_upload

    integer(c_int) function ndsp_download(ctx, name, out, width) bind(c, name='ndsp_download')
       import :: c_int, c_ptr, c_char, c_double, c_int32_t
       type(c_ptr), value :: ctx
       character(kind=c_char), intent(in) :: name(*)
       real(c_double), intent(out) :: out(*)
       integer(c_int32_t), intent(out) :: width
    end function ndsp_download

    integer(c_int) function ndsp_set_linklist(ctx) bind(c, name='ndsp_set_linklist')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
    end function ndsp_set_linklist

    integer(c_int) function ndsp_iterate_density(ctx, stats) bind(c, name='ndsp_iterate_density')
       import :: c_int, c_ptr, ndsp_stats_t
       type(c_ptr), value :: ctx
       type(ndsp_stats_t), intent(out) :: stats
    end function ndsp_iterate_density

    integer(c_int) function ndsp_get_rates(ctx, dt) bind(c, name='ndsp_get_rates')
       import :: c_int, c_ptr, ndsp_dt_t
       type(c_ptr), value :: ctx
       type(ndsp_dt_t), intent(out) :: dt
    end function ndsp_get_rates

    integer(c_int) function ndsp_derivs(ctx, stats, dt) bind(c, name='ndsp_derivs')
       import :: c_int, c_ptr, ndsp_stats_t, ndsp_dt_t
       type(c_ptr), value :: ctx
       type(ndsp_stats_t), intent(out) :: stats
       type(ndsp_dt_t), intent(out) :: dt
    end function ndsp_derivs

    integer(c_int) function ndsp_step(ctx, stats, dt) bind(c, name='ndsp_step')
       import :: c_int, c_ptr, ndsp_stats_t, ndsp_dt_t
       type(c_ptr), value :: ctx
       type(ndsp_stats_t), intent(out) :: stats
       type(ndsp_dt_t), intent(out) :: dt
    end function ndsp_step

    integer(c_int) function ndsp_set_dt(ctx, dt, time) bind(c, name='ndsp_set_dt')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), value :: dt, time
    end function ndsp_set_dt

    integer(c_int) function ndsp_set_graph(ctx, on) bind(c, name='ndsp_set_graph')
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: on
    end function ndsp_set_graph

    integer(c_int) function ndsp_conserved(ctx, out) bind(c, name='ndsp_conserved')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: out(17)
    end function ndsp_conserved

    integer(c_int) function ndsp_ev_columns(ctx, t, dt, cols, ncols) bind(c, name='ndsp_ev_columns')
       import :: c_int, c_ptr, c_double, c_int32_t
       type(c_ptr), value :: ctx
       real(c_double), value :: t, dt
       real(c_double), intent(out) :: cols(25)
       integer(c_int32_t), intent(out) :: ncols
    end function ndsp_ev_columns

    integer(c_int) function ndsp_set_ireal(ctx, ireal) bind(c, name='ndsp_set_ireal')
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), intent(in) :: ireal(*)
    end function ndsp_set_ireal

    type(c_ptr) function ndsp_last_error() bind(c, name='ndsp_last_error')
       import :: c_ptr
    end function ndsp_last_error
 end interface

contains

 ! the text of the last C-ABI error as a Fortran string
 function ndsp_error_text() result(msg)
    character(len=256) :: msg
    character(kind=c_char), pointer :: p(:)
    type(c_ptr) :: cp
    integer :: i
    msg = ' '
    cp = ndsp_last_error()
    if (.not. c_associated(cp)) return
    call c_f_pointer(cp, p, [256])
    do i = 1, 256
       if (p(i) == c_null_char) exit
       msg(i:i) = p(i)
    enddo
 end function ndsp_error_text

end module ndsp_capi
