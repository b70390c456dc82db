! Test infrastructure, not product code.
!
! norm2_dump (ours): evaluates the Fortran NORM2 intrinsic, as the compiler
! that builds the reference evaluates it, on fixed grids of the vectors the
! reference's geometry builders pass to it (module_geometry.f90:229-237, a
! node minus a corner or the cell centre; :270, a node minus the cylinder
! axis): integer and half-integer components.  tests/test_oracle_pins.py
! regenerates the same grids and checks lbref_norm2 bit for bit against the
! values written here.
!
! Output (stream, native endianness), in this order:
!   dense 3-D grid  components k/2, k = -64..64, first component fastest
!   sparse 3-D grid components k/2, k = -1200..1200 step 25
!   2-D grid        components k/2, k = -600..600
program norm2_dump
    implicit none
    integer, parameter :: dp = kind(1.0d0)
    integer :: a, b, c, u
    character(len=512) :: path
    real(dp) :: v3(3), v2(2)
    call get_command_argument(1, path)
    open(newunit=u, file=trim(path), access='stream', form='unformatted', status='replace')
    do c = -64, 64
        do b = -64, 64
            do a = -64, 64
                v3 = [real(a, dp), real(b, dp), real(c, dp)] / 2.0_dp
                write(u) norm2(v3)
            end do
        end do
    end do
    do c = -1200, 1200, 25
        do b = -1200, 1200, 25
            do a = -1200, 1200, 25
                v3 = [real(a, dp), real(b, dp), real(c, dp)] / 2.0_dp
                write(u) norm2(v3)
            end do
        end do
    end do
    do b = -600, 600
        do a = -600, 600
            v2 = [real(a, dp), real(b, dp)] / 2.0_dp
            write(u) norm2(v2)
        end do
    end do
    close(u)
end program norm2_dump
