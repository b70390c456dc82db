! Drop-in replacement of laboetie's module_collision (module_collision.f90):
! same module name, same public procedures and interfaces; collide() runs on
! the GPU through lbx_collide.
module module_collision

    use precision_kinds, only: dp
    implicit none

    private
    public :: collide, check_population

contains

    ! collide(n, density, jx, jy, jz, f_ext_x, f_ext_y, f_ext_z),
    ! module_collision.f90:15-130: BGK collision with the Guo force term at
    ! every fluid node; the relaxation time is read from lb.in on the first
    ! call (:37-41).
    subroutine collide(n, density, jx, jy, jz, f_ext_x, f_ext_y, f_ext_z)
        use iso_c_binding, only: c_int8_t, c_int32_t, c_double
        use system, only: node
        use module_input, only: getinput
        use lbx_bind, only: lbx_collide, lbx_check, lbx_nature
        implicit none
        real(dp), intent(in) :: density(:,:,:), jx(:,:,:), jy(:,:,:), jz(:,:,:)
        real(dp), intent(in) :: f_ext_x(:,:,:), f_ext_y(:,:,:), f_ext_z(:,:,:)
        real(dp), intent(inout) :: n(:,:,:,:)
        logical, save :: have_tau = .false.
        real(dp), save :: tau
        integer(c_int8_t), allocatable :: nature(:)

        if (.not. have_tau) then
            tau = getinput%dp('relaxation_time', defaultvalue=1._dp, assert=">0")
            if (tau < 0.5_dp) error stop "relaxation_time must be > 0.5"
            if (getinput%log('first_order_only', defaultvalue=.false.)) &
                error stop "first_order_only = T is not available on the lbx path"
            have_tau = .true.
        end if
        call lbx_nature(nature)
        call lbx_check(lbx_collide(int(size(n, 1), c_int32_t), int(size(n, 2), c_int32_t), int(size(n, 3), c_int32_t), &
                                   nature, real(tau, c_double), n, density, jx, jy, jz, f_ext_x, f_ext_y, f_ext_z), &
                       "collide")
    end subroutine collide

    ! population sanity check, module_collision.f90:143-149
    subroutine check_population(arrayin)
        implicit none
        real(dp), dimension(:,:,:,:), intent(in) :: arrayin
        real(dp), parameter :: eps = epsilon(1.0_dp)
        if (any(abs(arrayin) > 1 + eps)) stop "Critical. Population n_i(r) > 1 somewhere"
        if (any(arrayin < -eps)) stop 'Critical. Population n_i(r) < 0 somewhere.'
    end subroutine check_population

end module module_collision
