! Drop-in replacement of laboetie's module moment_propagation
! (module_moment_propagation.f90): same module name, same public procedures
! (init, propagate, deallocate_propagated_quantity, :11) with the same
! interfaces, same lb.in keys, same stdout and output/vacf.dat; the
! propagated quantities live on the GPU (lbx_mp_*).
module moment_propagation

    use precision_kinds
    use iso_c_binding, only: c_double, c_int64_t
    implicit none

    private
    public :: init, propagate, deallocate_propagated_quantity

    real(dp), parameter :: eps = epsilon(1._dp)
    logical, save :: write_vacf = .true.

    ! propagate() runs the propagation ahead in chunks of up to mp_chunk steps
    ! (each stopping at the reference's convergence step) and hands back one
    ! VACF per call from the buffer: VACF(buf_first + k - 1) is buf(:, k),
    ! k = 1..buf_n; buf_conv: the last step held passed the convergence test.
    ! The next chunk is begun (lbx_mp_run_start) as soon as a chunk is in the
    ! buffer, so the GPU steps it while this one's lines are written; it starts
    ! at step pend_first.
    integer, parameter :: mp_chunk = 256
    real(c_double), save :: buf(3, mp_chunk)
    integer(c_int64_t), save :: buf_first = 0, buf_n = 0, pend_first = 0
    logical, save :: buf_conv = .false., pending = .false.
    integer, save :: max_steps = 0  ! maximum_moment_propagation_steps (< 0: until converged)

contains

    ! moment_propagation::init (module_moment_propagation.f90:30-160); the
    ! tracer population that drop_tracers rebuilt into n is rebuilt on the
    ! device from the same flow (node%solventDensity, node%solventFlux) and
    ! f_ext, with the same expression (drop_tracers.f90:97-100).
    subroutine init
        use iso_c_binding, only: c_double
        use system, only: node, f_ext
        use module_input, only: getinput
        use lbx_bind
        implicit none
        real(dp) :: ka, kd, Db, Ds, zt, tau
        real(c_double) :: vacf0(3)
        real(c_double), allocatable :: rho(:), jx(:), jy(:), jz(:)

        ka = getinput%dp('tracer_ka', 0._dp)
        kd = getinput%dp('tracer_kd', 0._dp)
        if (ka < -eps) error stop 'I detected tracer%ka to be <0 in module moment_propagation. STOP.'
        if (kd < -eps) error stop 'I detected tracer%kd to be <0 in module moment_propagation. STOP.'
        Db = getinput%dp('tracer_Db', 0._dp)
        Ds = getinput%dp('tracer_Ds', 0._dp)
        if (abs(Ds) > eps) then
            print *, "Tracers you defined have non-zero surface diffusion coefficient"
            print *, "This is not implemented yet"
            error stop
        end if
        zt = getinput%dp('tracer_z', 0._dp)
        if (abs(zt) > eps) then
            print *, "charged tracers are not implemented"
            error stop "line 85 of module_moment_propagation.f90"
        end if
        tau = getinput%dp('relaxation_time', defaultvalue=1._dp, assert=">0")
        call lbx_ensure_lattice(real(tau, c_double))

        allocate (rho(size(node)), jx(size(node)), jy(size(node)), jz(size(node)))
        rho = reshape(node%solventDensity, [size(node)])
        jx = reshape(node%solventFlux(1), [size(node)])
        jy = reshape(node%solventFlux(2), [size(node)])
        jz = reshape(node%solventFlux(3), [size(node)])
        call lbx_check(lbx_mp_init_from_flow(lbx_lattice, rho, jx, jy, jz, real(f_ext, c_double), &
                                             real(Db, c_double), real(ka, c_double), real(kd, c_double), vacf0), &
                       "lbx_mp_init_from_flow")

        print *, 0, real(vacf0, sp)
        max_steps = getinput%int("maximum_moment_propagation_steps", 0)
        buf_first = 0
        buf_n = 0
        buf_conv = .false.
        pending = .false.
        write_vacf = getinput%log("print_vacf", .true.)
        if (write_vacf) then
            open (99, file='output/vacf.dat')
            write (99, *) '# time t, VACF_x(t), VACF_y(t), VACF_z(t)'
            write (99, *) 0, vacf0
        end if
    end subroutine init

    ! one moment-propagation step (module_moment_propagation.f90:164-289):
    ! VACF(it) to stdout every 10000 steps and to output/vacf.dat, and the
    ! convergence verdict of :284-288.  The steps run on the device a chunk
    ! at a time (drop_tracers calls this with it = 1, 2, ... until the verdict
    ! or maximum_moment_propagation_steps); a chunk never runs past either.
    subroutine propagate(it, is_converged)
        use iso_c_binding, only: c_int32_t
        use lbx_bind
        implicit none
        integer(i2b), intent(in) :: it
        logical, intent(out) :: is_converged
        real(c_double) :: vacf(3)
        integer(c_int64_t) :: k
        integer(c_int32_t) :: converged

        if (it < buf_first .or. it >= buf_first + buf_n) then
            if (.not. pending) call start_chunk(int(it, c_int64_t))
            ! drop_tracers calls it = 1, 2, ...: the begun chunk starts at this step
            if (.not. pending .or. pend_first /= it) error stop "moment propagation called out of step order"
            call lbx_check(lbx_mp_run_wait(lbx_lattice, buf, buf_n, converged), "lbx_mp_run_wait")
            pending = .false.
            if (buf_n < 1) error stop "moment propagation called past its convergence"
            buf_first = it
            buf_conv = converged /= 0
            if (.not. buf_conv) call start_chunk(buf_first + buf_n)
        end if
        k = it - buf_first + 1
        vacf = buf(:, k)
        if (modulo(it, 10000) == 0) print *, it, real(vacf, sp)
        if (write_vacf) write (99, *) it, vacf
        is_converged = buf_conv .and. k == buf_n
    end subroutine propagate

    ! begin the chunk of steps first.. (up to mp_chunk, not past
    ! maximum_moment_propagation_steps); nothing when no step remains
    subroutine start_chunk(first)
        use iso_c_binding, only: c_int32_t
        use lbx_bind
        integer(c_int64_t), intent(in) :: first
        integer(c_int64_t) :: n
        n = mp_chunk
        if (max_steps > 0) n = min(n, int(max_steps, c_int64_t) - first + 1)
        if (n < 1) return
        call lbx_check(lbx_mp_run_start(lbx_lattice, n, 1_c_int32_t), "lbx_mp_run_start")
        pending = .true.
        pend_first = first
    end subroutine start_chunk

    subroutine deallocate_propagated_quantity
        use lbx_bind
        implicit none
        if (c_associated(lbx_lattice)) call lbx_check(lbx_mp_free(lbx_lattice), "lbx_mp_free")
    end subroutine deallocate_propagated_quantity

end module moment_propagation
