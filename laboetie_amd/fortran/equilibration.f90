! Drop-in replacement of laboetie's equilibration (equilibration.f90): same
! argument-less interface, same lb.in keys, same files under output/, same
! module state on exit (n, node%solventDensity, node%solventFlux).  The time
! loop (collide, bounce-back, streaming, moments, convergence and the force
! switch) runs on the GPU (lbx_equilibrate_begin / _advance) in runs of up to
! 256 steps; between them the shim writes the outputs of the steps taken: the
! stdout step lines and l2err.dat, total_mass_flux.dat, and the profiles every
! print_files_frequency steps.
subroutine equilibration

    use iso_c_binding, only: c_double, c_int64_t, c_int32_t, c_null_ptr
    use, intrinsic :: ieee_arithmetic, only: ieee_is_nan
    use precision_kinds, only: dp
    use system, only: node, n
    use module_input, only: getinput
    use constants, only: x, y, z, zerodp
    use lbx_bind

    implicit none
    integer :: n1, n2, n3, print_frequency, print_files_frequency, u13
    integer(c_int64_t) :: steps, done, until, next_prof, t, max_steps
    integer(c_int64_t), parameter :: run_steps = 256
    integer(c_int32_t) :: reason, flux_sums
    integer(c_int) :: status
    logical :: write_total_mass_flux
    real(dp) :: target_error, tau, f_ext_loc(3), l2prev
    real(dp), allocatable :: density(:,:,:), jx(:,:,:), jy(:,:,:), jz(:,:,:)
    real(dp), parameter :: eps = epsilon(1._dp)

    ! laboetie's path is uncharged only (equilibration.f90:28-33)
    if (abs(getinput%dp('sigma', zerodp)) > eps) then
        print *, "ERROR: laboetie can only consider uncharged systems."
        error stop
    end if
    if (getinput%log("compensate_f_ext", .false.)) &
        error stop "compensate_f_ext = T is not available on the lbx path"
    write_total_mass_flux = getinput%log("write_total_mass_flux", .false.)

    n1 = getinput%int("lx", assert=">0")
    n2 = getinput%int("ly", assert=">0")
    n3 = getinput%int("lz", assert=">0")
    print_frequency = getinput%int('print_frequency', defaultvalue=max(int(50000/(n1*n2*n3)), 1), assert=">0")
    print_files_frequency = getinput%int("print_files_frequency", huge(1))
    target_error = getinput%dp("target_error", 1.D-10)
    tau = getinput%dp('relaxation_time', defaultvalue=1._dp, assert=">0")
    if (tau < 0.5_dp) error stop "relaxation_time must be > 0.5"
    f_ext_loc = getinput%dp3("f_ext", [0._dp, 0._dp, 0._dp])
    ! extension key (absent from the reference): cap on the number of steps, 0 = none
    max_steps = int(getinput%int("maximum_equilibration_steps", 0), c_int64_t)

    allocate (density(n1, n2, n3), jx(n1, n2, n3), jy(n1, n2, n3), jz(n1, n2, n3))
    density = node%solventdensity
    jx = 0
    jy = 0
    jz = 0

    open (66, file="output/mass-flux_profile_along_z.dat")
    open (67, file="output/mass-flux_profile_along_y.dat")
    open (68, file="output/mass-flux_profile_along_x.dat")
    write (66, *) "# z, <ρ.v_x>_{x,y}, <ρ.v_y>_{x,y}, <ρ.v_z>_{x,y}"
    write (67, *) "# y, <ρ.v_x>_{x,z}, <ρ.v_y>_{x,z}, <ρ.v_z>_{x,z}"
    write (68, *) "# x, <ρ.v_x>_{y,z}, <ρ.v_y>_{y,z}, <ρ.v_z>_{y,z}"
    open (56, file="output/mean-density_profile_along_z.dat")
    open (57, file="output/mean-density_profile_along_y.dat")
    open (58, file="output/mean-density_profile_along_x.dat")
    if (write_total_mass_flux) open (65, file="output/total_mass_flux.dat")

    print *
    print *, 'Lattice Boltzmann'
    print *, '================='
    print *, '       step'
    print *, '       ----'

    ! profiles of the state that opens step 1 (equilibration.f90:154-176)
    call write_profiles("# timestep", 1, .true.)

    call lbx_ensure_lattice(real(tau, c_double))
    call lbx_check(lbx_set_state(lbx_lattice, n, density, jx, jy, jz), "lbx_set_state")
    flux_sums = 0
    if (write_total_mass_flux) flux_sums = 1
    call lbx_check(lbx_equilibrate_begin(lbx_lattice, real(target_error, c_double), real(f_ext_loc, c_double), &
                                         max_steps, flux_sums), "lbx_equilibrate_begin")

    open (newunit=u13, file="./output/l2err.dat")
    l2prev = 0
    done = 0
    ! the next step that opens with a profile (the one of t = 1 is written)
    next_prof = huge(1_c_int64_t)
    if (print_files_frequency < huge(1)) next_prof = max(print_files_frequency, 2)
    if (next_prof < huge(1_c_int64_t) .and. modulo(next_prof, int(print_files_frequency, c_int64_t)) /= 0) &
        next_prof = next_prof + print_files_frequency - modulo(next_prof, int(print_files_frequency, c_int64_t))
    do
        until = done + run_steps
        if (next_prof < huge(1_c_int64_t)) until = min(until, next_prof - 1)
        status = lbx_equilibrate_advance(lbx_lattice, until, steps, reason)
        call write_step_outputs(done + 1, steps)
        done = steps
        if (status /= LBX_OK .or. reason /= 0) exit
        ! the state that opens step next_prof (equilibration.f90:154-176)
        if (done == next_prof - 1) then
            call lbx_check(lbx_flow(lbx_lattice, density, jx, jy, jz), "lbx_flow")
            call write_profiles("# timestep", int(next_prof), .true.)
            next_prof = next_prof + print_files_frequency
        end if
    end do
    close (u13)
    if (write_total_mass_flux) close (65)
    call lbx_check(status, "lbx_equilibrate")

    call lbx_check(lbx_get_state(lbx_lattice, n, density, jx, jy, jz), "lbx_get_state")

    ! steady-state profiles and the 2-D flux field (equilibration.f90:499-533)
    call write_profiles("# Steady state with convergence criteria", 0, .false.)
    close (66)
    close (67)
    close (68)
    close (56)
    close (57)
    close (58)
    call write_field_x1()

    node%solventdensity = density
    node%solventflux(x) = jx
    node%solventflux(y) = jy
    node%solventflux(z) = jz

contains

    ! the outputs of steps first..last of the run (equilibration.f90:149,
    ! 257-261,339-344): the stdout line that opens step t (every
    ! print_frequency steps, with the l2err of the step before), the
    ! total_mass_flux.dat line (the flux sum of the step before, written once
    ! step t has no negative population) and the l2err.dat line.  The FP32
    ! storage forms measure l2err every 64 steps in stage 2: the other steps
    ! (NaN in the trace) have no l2err.dat line.
    subroutine write_step_outputs(first, last)
        integer(c_int64_t), intent(in) :: first, last
        integer(c_int64_t) :: held, fheld, t, n
        real(dp) :: tr(max(last - first + 1, 1_c_int64_t)), js(3, max(last - first + 1, 1_c_int64_t))
        if (last < first) return
        ! l2err of steps first..min(last, held): a step that stopped on a
        ! negative population has none
        call lbx_check(lbx_l2err_trace(lbx_lattice, 1_c_int64_t, 0_c_int64_t, tr, held), "lbx_l2err_trace")
        n = min(last, held) - first + 1
        if (n > 0) call lbx_check(lbx_l2err_trace(lbx_lattice, first, n, tr, held), "lbx_l2err_trace")
        ! flux sums of steps first-1..last-1 (the one of step 0 is 0)
        js = 0
        if (write_total_mass_flux .and. n > 0) then
            call lbx_check(lbx_flux_sums(lbx_lattice, 1_c_int64_t, 0_c_int64_t, js, fheld), "lbx_flux_sums")
            if (first == 1) then
                if (n > 1) call lbx_check(lbx_flux_sums(lbx_lattice, 1_c_int64_t, n - 1, js(:, 2:), fheld), &
                                          "lbx_flux_sums")
            else
                call lbx_check(lbx_flux_sums(lbx_lattice, first - 1, n, js, fheld), "lbx_flux_sums")
            end if
        end if
        do t = first, last
            if (modulo(t, int(print_frequency, c_int64_t)) == 0) &
                print *, int(t), real(l2prev), "(target", real(target_error, 4), ")"
            if (t > held) exit
            if (write_total_mass_flux) write (65, *) int(t), real(js(:, t - first + 1))
            if (ieee_is_nan(tr(t - first + 1))) cycle
            write (u13, *) int(t), tr(t - first + 1)
            l2prev = tr(t - first + 1)
        end do
    end subroutine write_step_outputs

    ! the six profile files: a header line, then per plane of each axis the
    ! mean density of its non-empty nodes and the summed flux components
    subroutine write_profiles(header, step, with_step)
        character(*), intent(in) :: header
        integer, intent(in) :: step
        logical, intent(in) :: with_step
        integer :: k
        if (with_step) then
            write (66, *) header, step
            write (67, *) header, step
            write (68, *) header, step
            write (56, *) header, step
            write (57, *) header, step
            write (58, *) header, step
        else
            write (66, *) header, real(target_error)
            write (67, *) header, real(target_error)
            write (68, *) header, real(target_error)
            write (56, *) header, real(target_error)
            write (57, *) header, real(target_error)
            write (58, *) header, real(target_error)
        end if
        if (with_step) then
            do k = 1, n3
                write (56, *) k, sum(density(:, :, k))/max(count(density(:, :, k) > eps), 1)
                write (66, *) k, sum(jx(:, :, k)), sum(jy(:, :, k)), sum(jz(:, :, k))
            end do
            do k = 1, n2
                write (57, *) k, sum(density(:, k, :))/max(count(density(:, k, :) > eps), 1)
                write (67, *) k, sum(jx(:, k, :)), sum(jy(:, k, :)), sum(jz(:, k, :))
            end do
            do k = 1, n1
                write (58, *) k, sum(density(k, :, :))/max(count(density(k, :, :) > eps), 1)
                write (68, *) k, sum(jx(k, :, :)), sum(jy(k, :, :)), sum(jz(k, :, :))
            end do
            write (66, *)
            write (67, *)
            write (68, *)
        else
            do k = 1, n3
                write (66, *) k, sum(jx(:, :, k)), sum(jy(:, :, k)), sum(jz(:, :, k))
                write (56, *) k, sum(density(:, :, k))/max(count(density(:, :, k) > eps), 1)
            end do
            do k = 1, n2
                write (67, *) k, sum(jx(:, k, :)), sum(jy(:, k, :)), sum(jz(:, k, :))
                write (57, *) k, sum(density(:, k, :))/max(count(density(:, k, :) > eps), 1)
            end do
            do k = 1, n1
                write (68, *) k, sum(jx(k, :, :)), sum(jy(k, :, :)), sum(jz(k, :, :))
                write (58, *) k, sum(density(k, :, :))/max(count(density(k, :, :) > eps), 1)
            end do
        end if
    end subroutine write_profiles

    subroutine write_field_x1()
        integer :: j, k, u
        open (newunit=u, file="output/mass-flux_field_2d_at_x.eq.1.dat")
        do j = 1, n2
            do k = 1, n3
                write (u, *) j, k, jy(1, j, k), jz(1, j, k)
            end do
        end do
        close (u)
    end subroutine write_field_x1

end subroutine equilibration
