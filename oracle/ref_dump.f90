! Test infrastructure, not product code.
!
! A main program of our own that drives the reference's unchanged modules and
! subroutines (maxlevesque/laboetie, src/) and dumps their module arrays as raw
! little-endian binaries, so that the golden fixtures pin every node and
! population, not only what the reference's text outputs print.
!
! Usage (cwd holds lb.in, and geom.in when geometryLabel = 0):
!   ref_dump full      main.f90:12-34 sequence; after equilibration writes
!                      output/eq_state.bin, after drop_tracers output/tracer_n.bin
!   ref_dump geom      init_simu only; writes output/geom.bin
!   ref_dump collide   init_simu, then one module_collision::collide call on the
!                      arrays read from collide_in.bin; writes collide_out.bin
!   ref_dump tables    init_simu only; writes output/tables.bin (lbm%vel)
!
! Binary layouts (all Fortran column-major, x fastest):
!   geom.bin      int32 lx,ly,lz; int8 nature(lx,ly,lz); int8 interfacial(lx,ly,lz)
!   eq_state.bin  int32 lx,ly,lz; f64 n(lx,ly,lz,19); f64 rho(lx,ly,lz);
!                 f64 jx, jy, jz (lx,ly,lz)           (equilibration.f90:551-554)
!   tracer_n.bin  f64 n(19,lx,ly,lz)                   (drop_tracers.f90:85-105)
!   collide_in.bin / collide_out.bin
!                 f64 n(lx,ly,lz,19) [in: then density, jx, jy, jz, fx, fy, fz]
!   tables.bin    int32 coo(3,19); int32 inv(19); f64 a0(19), a1(19), a2(19)
program ref_dump
    use system, only: n, node, supercell
    use module_collision, only: collide
    use mod_lbmodel, only: lbm
    implicit none
    character(len=16) :: mode
    integer :: lx, ly, lz, u, l
    real(8), allocatable :: d(:,:,:), jx(:,:,:), jy(:,:,:), jz(:,:,:), fx(:,:,:), fy(:,:,:), fz(:,:,:)

    mode = 'full'
    if (command_argument_count() >= 1) call get_command_argument(1, mode)

    call init_simu
    lx = supercell%geometry%dimensions%indiceMax(1)
    ly = supercell%geometry%dimensions%indiceMax(2)
    lz = supercell%geometry%dimensions%indiceMax(3)

    select case (trim(mode))
    case ('geom')
        open(newunit=u, file='output/geom.bin', access='stream', form='unformatted', status='replace')
        write(u) lx, ly, lz
        write(u) int(node%nature, 1)
        write(u) merge(1_1, 0_1, node%isInterfacial)
        close(u)

    case ('tables')
        open(newunit=u, file='output/tables.bin', access='stream', form='unformatted', status='replace')
        write(u) (int(lbm%vel(l)%coo(1:3), 4), l = 1, 19)
        write(u) (int(lbm%vel(l)%inv, 4), l = 1, 19)
        write(u) (lbm%vel(l)%a0, l = 1, 19)
        write(u) (lbm%vel(l)%a1, l = 1, 19)
        write(u) (lbm%vel(l)%a2, l = 1, 19)
        close(u)

    case ('collide')
        allocate(d(lx,ly,lz), jx(lx,ly,lz), jy(lx,ly,lz), jz(lx,ly,lz), fx(lx,ly,lz), fy(lx,ly,lz), fz(lx,ly,lz))
        open(newunit=u, file='collide_in.bin', access='stream', form='unformatted', status='old')
        read(u) n
        read(u) d, jx, jy, jz, fx, fy, fz
        close(u)
        call collide(n, d, jx, jy, jz, fx, fy, fz)
        open(newunit=u, file='collide_out.bin', access='stream', form='unformatted', status='replace')
        write(u) n
        close(u)

    case default
        call poisson_nernst_planck
        call equilibration
        open(newunit=u, file='output/eq_state.bin', access='stream', form='unformatted', status='replace')
        write(u) lx, ly, lz
        write(u) n
        write(u) node%solventDensity
        write(u) node%solventFlux(1)
        write(u) node%solventFlux(2)
        write(u) node%solventFlux(3)
        close(u)
        call drop_tracers
        if (allocated(n)) then
            if (size(n, 1) == 19) then
                open(newunit=u, file='output/tracer_n.bin', access='stream', form='unformatted', status='replace')
                write(u) n
                close(u)
            end if
        end if
    end select
end program ref_dump
