! iso_c_binding interface to liblbx (include/lbx.h), used by the Fortran
! replacements of laboetie's module_collision, equilibration and
! moment_propagation.  One device context per process holds the lattice; the
! equilibration creates it and the moment propagation reuses it.
module lbx_bind
    use iso_c_binding
    implicit none

    integer(c_int), parameter :: LBX_OK = 0, LBX_ERR_NEG_POP = 1, LBX_ERR_RESTPART = 2, LBX_ERR_TAU = 3, &
                                 LBX_ERR_NOMEM = 4, LBX_ERR_ARG = 5, LBX_ERR_DEVICE = 6, LBX_ERR_STATE = 7
    integer(c_int32_t), parameter :: LBX_PREC_FP64 = 0, LBX_PREC_FP32_SHIFT = 1, LBX_PREC_FP32_STORE = 2
    integer(c_int32_t), parameter :: LBX_EXIT_CONVERGED = 1, LBX_EXIT_CAP = 2

    type(c_ptr), save :: lbx_lattice = c_null_ptr

    interface
        function lbx_status_string(status) bind(c, name='lbx_status_string') result(msg)
            import :: c_int, c_ptr
            integer(c_int), value :: status
            type(c_ptr) :: msg
        end function

        integer(c_int) function lbx_create(lx, ly, lz, nature, tau, precision, out) bind(c, name='lbx_create')
            import :: c_int, c_int32_t, c_int8_t, c_double, c_ptr
            integer(c_int32_t), value :: lx, ly, lz
            integer(c_int8_t), intent(in) :: nature(*)
            real(c_double), value :: tau
            integer(c_int32_t), value :: precision
            type(c_ptr), intent(out) :: out
        end function

        integer(c_int) function lbx_destroy(ctx) bind(c, name='lbx_destroy')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function

        integer(c_int) function lbx_set_state(ctx, n, rho, jx, jy, jz) bind(c, name='lbx_set_state')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(in) :: n(*), rho(*), jx(*), jy(*), jz(*)
        end function

        integer(c_int) function lbx_equilibrate(ctx, target_error, fext, max_steps, l2err_trace, trace_cap, &
                                                steps_done, exit_reason) bind(c, name='lbx_equilibrate')
            import :: c_int, c_ptr, c_double, c_int64_t, c_int32_t
            type(c_ptr), value :: ctx
            real(c_double), value :: target_error
            real(c_double), intent(in) :: fext(3)
            integer(c_int64_t), value :: max_steps
            type(c_ptr), value :: l2err_trace
            integer(c_int64_t), value :: trace_cap
            integer(c_int64_t), intent(out) :: steps_done
            integer(c_int32_t), intent(out) :: exit_reason
        end function

        integer(c_int) function lbx_equilibrate_begin(ctx, target_error, fext, max_steps, flux_sums) &
                bind(c, name='lbx_equilibrate_begin')
            import :: c_int, c_ptr, c_double, c_int64_t, c_int32_t
            type(c_ptr), value :: ctx
            real(c_double), value :: target_error
            real(c_double), intent(in) :: fext(3)
            integer(c_int64_t), value :: max_steps
            integer(c_int32_t), value :: flux_sums
        end function

        integer(c_int) function lbx_equilibrate_advance(ctx, until_step, steps_done, exit_reason) &
                bind(c, name='lbx_equilibrate_advance')
            import :: c_int, c_ptr, c_int64_t, c_int32_t
            type(c_ptr), value :: ctx
            integer(c_int64_t), value :: until_step
            integer(c_int64_t), intent(out) :: steps_done
            integer(c_int32_t), intent(out) :: exit_reason
        end function

        integer(c_int) function lbx_flux_sums(ctx, first_step, count, out, available) bind(c, name='lbx_flux_sums')
            import :: c_int, c_ptr, c_double, c_int64_t
            type(c_ptr), value :: ctx
            integer(c_int64_t), value :: first_step, count
            real(c_double), intent(inout) :: out(*)
            integer(c_int64_t), intent(out) :: available
        end function

        integer(c_int) function lbx_flow(ctx, rho, jx, jy, jz) bind(c, name='lbx_flow')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(out) :: rho(*), jx(*), jy(*), jz(*)
        end function

        integer(c_int) function lbx_l2err_trace(ctx, first_step, count, out, available) &
                bind(c, name='lbx_l2err_trace')
            import :: c_int, c_ptr, c_double, c_int64_t
            type(c_ptr), value :: ctx
            integer(c_int64_t), value :: first_step, count
            real(c_double), intent(out) :: out(*)
            integer(c_int64_t), intent(out) :: available
        end function

        integer(c_int) function lbx_get_state(ctx, n, rho, jx, jy, jz) bind(c, name='lbx_get_state')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(out) :: n(*), rho(*), jx(*), jy(*), jz(*)
        end function

        integer(c_int) function lbx_mp_init_from_flow(ctx, rho, jx, jy, jz, fext, Db, ka, kd, vacf0) &
                bind(c, name='lbx_mp_init_from_flow')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(in) :: rho(*), jx(*), jy(*), jz(*), fext(3)
            real(c_double), value :: Db, ka, kd
            real(c_double), intent(out) :: vacf0(3)
        end function

        integer(c_int) function lbx_mp_run(ctx, nsteps, stop_on_convergence, vacf_trace, steps_done, converged) &
                bind(c, name='lbx_mp_run')
            import :: c_int, c_ptr, c_double, c_int64_t, c_int32_t
            type(c_ptr), value :: ctx
            integer(c_int64_t), value :: nsteps
            integer(c_int32_t), value :: stop_on_convergence
            real(c_double), intent(out) :: vacf_trace(*)
            integer(c_int64_t), intent(out) :: steps_done
            integer(c_int32_t), intent(out) :: converged
        end function

        integer(c_int) function lbx_mp_run_start(ctx, nsteps, stop_on_convergence) bind(c, name='lbx_mp_run_start')
            import :: c_int, c_ptr, c_int64_t, c_int32_t
            type(c_ptr), value :: ctx
            integer(c_int64_t), value :: nsteps
            integer(c_int32_t), value :: stop_on_convergence
        end function

        integer(c_int) function lbx_mp_run_wait(ctx, vacf_trace, steps_done, converged) bind(c, name='lbx_mp_run_wait')
            import :: c_int, c_ptr, c_double, c_int64_t, c_int32_t
            type(c_ptr), value :: ctx
            real(c_double), intent(out) :: vacf_trace(*)
            integer(c_int64_t), intent(out) :: steps_done
            integer(c_int32_t), intent(out) :: converged
        end function

        integer(c_int) function lbx_mp_free(ctx) bind(c, name='lbx_mp_free')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function

        integer(c_int) function lbx_collide(lx, ly, lz, nature, tau, n, density, jx, jy, jz, fx, fy, fz) &
                bind(c, name='lbx_collide')
            import :: c_int, c_int32_t, c_int8_t, c_double
            integer(c_int32_t), value :: lx, ly, lz
            integer(c_int8_t), intent(in) :: nature(*)
            real(c_double), value :: tau
            real(c_double), intent(inout) :: n(*)
            real(c_double), intent(in) :: density(*), jx(*), jy(*), jz(*), fx(*), fy(*), fz(*)
        end function
    end interface

contains

    ! The reference's fatal stops, by status (equilibration.f90:248,
    ! module_moment_propagation.f90:257, module_collision.f90:39); anything else
    ! is a device or argument error of the lbx path itself.
    subroutine lbx_check(status, where)
        integer(c_int), intent(in) :: status
        character(*), intent(in) :: where
        character(kind=c_char), pointer :: msg(:)
        integer :: k
        if (status == LBX_OK) return
        select case (status)
        case (LBX_ERR_NEG_POP)
            error stop "In equilibration, the population n(x,y,z,vel) < 0"
        case (LBX_ERR_RESTPART)
            stop 'somewhere restpart is negative'
        case (LBX_ERR_TAU)
            error stop "relaxation_time must be > 0.5"
        case default
            call c_f_pointer(lbx_status_string(status), msg, [200])
            k = 0
            do while (k < 200)
                if (msg(k + 1) == c_null_char) exit
                k = k + 1
            end do
            print *, "lbx error in ", where, ": ", (msg(1:k))
            error stop "lbx: device call failed"
        end select
    end subroutine lbx_check

    ! node%nature as a contiguous int8 array (x fastest)
    subroutine lbx_nature(nature)
        use system, only: node
        integer(c_int8_t), allocatable, intent(out) :: nature(:)
        allocate (nature(size(node)))
        nature = int(reshape(node%nature, [size(node)]), c_int8_t)
    end subroutine lbx_nature

    ! the process-wide lattice context, created on first use
    ! The lattice, created on first use.  One lb.in key beyond the
    ! reference's selects the precision form: gpu_precision = fp64 (default,
    ! the reference's arithmetic), fp32_shift or fp32_store.
    subroutine lbx_ensure_lattice(tau)
        use system, only: supercell
        use module_input, only: getinput
        real(c_double), intent(in) :: tau
        integer(c_int8_t), allocatable :: nature(:)
        integer :: lx, ly, lz
        integer(c_int32_t) :: prec
        character(len=:), allocatable :: pname
        if (c_associated(lbx_lattice)) return
        lx = supercell%geometry%dimensions%indiceMax(1)
        ly = supercell%geometry%dimensions%indiceMax(2)
        lz = supercell%geometry%dimensions%indiceMax(3)
        pname = trim(adjustl(getinput%char("gpu_precision", defaultvalue="fp64")))
        select case (pname)
        case ("fp64")
            prec = LBX_PREC_FP64
        case ("fp32_shift")
            prec = LBX_PREC_FP32_SHIFT
        case ("fp32_store")
            prec = LBX_PREC_FP32_STORE
        case default
            print *, "gpu_precision must be fp64, fp32_shift or fp32_store, found: ", pname
            stop
        end select
        call lbx_nature(nature)
        call lbx_check(lbx_create(int(lx, c_int32_t), int(ly, c_int32_t), int(lz, c_int32_t), nature, tau, &
                                  prec, lbx_lattice), "lbx_create")
    end subroutine lbx_ensure_lattice

end module lbx_bind
