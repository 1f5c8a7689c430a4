! bs2e_rkh_multipoles.f90 -- the RKH program's multipole step
! (main_RKH.f90 -> setup_shift_Coul_Ham, RKH_hamiltonian.f90:276-292:
! init_g_l, setup_r_k_int) over the drop-in module bs2e_rkh.
!
! usage: bs2e_rkh_multipoles k m Z h_max r_max k_GL max_k r_0 out_file
!
! out_file (stream): n_b, max_k, r_k_int(0:max_k, n_b, n_b) (8-byte integers
! and doubles).
program bs2e_rkh_multipoles
    use bs2e_mat_els, only: b_spline, generate_grid, bs2e_shutdown
    use bs2e_rkh
    implicit none
    integer, parameter :: dp = kind(1.d0)
    integer :: k, m, Z, k_GL, u
    real(dp) :: h_max, r_max
    real(dp), allocatable :: grid(:)
    type(b_spline) :: splines
    character(len=4096) :: arg, out_file

    call get_command_argument(1, arg); read(arg, *) k
    call get_command_argument(2, arg); read(arg, *) m
    call get_command_argument(3, arg); read(arg, *) Z
    call get_command_argument(4, arg); read(arg, *) h_max
    call get_command_argument(5, arg); read(arg, *) r_max
    call get_command_argument(6, arg); read(arg, *) k_GL
    call get_command_argument(7, arg); read(arg, *) max_k
    call get_command_argument(8, arg); read(arg, *) r_0
    call get_command_argument(9, out_file)

    call generate_grid(k, m, Z, h_max, r_max, grid)
    call splines%init(k, grid)
    call init_g_l(splines, k_GL)
    call setup_r_k_int(splines)
    write(6, '(a,i0,a,i0)') "r_k_int: n_b ", splines%n_b, ", max_k ", max_k

    open(newunit=u, file=trim(out_file), access='stream', form='unformatted', status='replace', action='write')
    write(u) splines%n_b, max_k, r_k_int
    close(u)
    call bs2e_shutdown()
end program bs2e_rkh_multipoles
