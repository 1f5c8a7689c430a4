! bs2e_basis_setup.f90 -- the two-electron branch of the reference's
! basis_setup (main_basis_setup.f90:47-118) written against the drop-in
! modules: knots, B-splines, primitive containers, the R^k tensor (from the
! containers and, a second time, straight from the knots) and the H/S blocks
! of one symmetry.
!
! usage: bs2e_basis_setup k m Z h_max r_max k_GL max_k max_l L pi full key_stride hs_file out_file
!                         [max_L res_dir [gauge]]
!
! With max_L and res_dir, it also builds the blocks of every symmetry of the
! z-polarised two-electron basis (L = 0..max_L) and the dipole blocks
! between every pair of them for q = -1, 0, 1 in the given gauge ('l' by
! default; main_basis_setup.f90:122-145), and stores what the reference's
! basis_setup stores for them (main_basis_setup.f90:154-161):
! res_dir/H_diag.dat, S_diag.dat, D_-1.dat, D_0.dat, D_1.dat, basis.dat,
! basis_info.txt, splines.dat, basis_input.dat.
!
! hs_file (stream): S then H(l) for l = 0..max_l, each n_b x n_b complex(dp)
! column-major (the one-electron matrices, mat_els.f90:85-118 and :294-321);
! hs_file "device": they are computed by setup_S / setup_H_one_particle on
! the device for the hydrogenic potential of charge Z (no CAP).
! out_file (stream, 8-byte integers and doubles) receives every result:
!   n_knots, knots
!   nnz4, K, r_k%data, r_m_k%data, r_k%iv, r_k%i, r_k%j
!   nnz6, r_d_k%data, r_d_k%iv, %i, %i_p, %j, %j_p
!   n_keys, keys(4, n_keys), values(K, n_keys) from compute_R_k_map,
!   values(K, n_keys) from compute_R_k_direct
!   n_config, configs(5, n_config)            (n1, n2, l1, l2, eqv)
!   H_sp: nnz, index_ptr, indices, data;  S_sp: the same
program bs2e_basis_setup
    use bs2e_mat_els
    use bs2e_basis_input, only: set_basis_defaults, write_basis_input, get_basis_input, in_dir => basis_output_dir, &
        in_k => k, in_m => m, in_Z => Z, in_h_max => h_max, in_r_max => r_max, in_k_GL => k_GL, in_max_k => max_k, &
        in_max_L => max_L, in_max_l_1p => max_l_1p, in_max_l2 => max_l2, in_cap_order => CAP_order, &
        in_cap_r_0 => CAP_r_0, in_cap_eta => CAP_eta, in_gauge => gauge, in_full => full, in_z_pol => z_pol, &
        in_two_el => two_el
    implicit none
    integer, parameter :: dp = kind(1.d0)
    integer :: k, m, Z, k_GL, max_k, max_l, L, ipi, ifull, stride
    real(dp) :: h_max, r_max
    character(len=4096) :: hs_file, out_file, arg
    real(dp), allocatable :: grid(:), vals(:,:), vals_direct(:,:)
    type(b_spline) :: splines
    type(sparse_4d) :: r_k, r_m_k
    type(sparse_6d) :: r_d_k
    class(Nd_DOK), allocatable :: R_p
    type(Nd_DOK) :: R_direct
    type(block), allocatable :: H_vec(:)
    complex(dp), allocatable :: S(:,:)
    complex(dp), allocatable :: eigs_v(:,:)   ! main_basis_setup.f90:88 passes one-particle energies to bas%init
    type(sym) :: term
    type(CSR_matrix) :: H_sp, S_sp
    type(hydrogenic) :: pot
    type(CAP) :: CAP_c
    integer, allocatable :: keys(:,:)
    integer :: u, i, a, b, c, d, n_keys, p, q, n_pairs, max_L_all
    character(len=4096) :: res_dir
    type(basis) :: bas
    type(block_diag_CS) :: H_diag, S_diag
    type(block_CS) :: dipoles(-1:1)
    type(radial_dipole) :: radial_dip
    type(sym) :: dsyms(2)
    character :: gauge
    integer :: iq, j
    character(len=2) :: qs

    call get_command_argument(1, arg); read(arg, *) k
    call get_command_argument(2, arg); read(arg, *) m
    call get_command_argument(3, arg); read(arg, *) Z
    call get_command_argument(4, arg); read(arg, *) h_max
    call get_command_argument(5, arg); read(arg, *) r_max
    call get_command_argument(6, arg); read(arg, *) k_GL
    call get_command_argument(7, arg); read(arg, *) max_k
    call get_command_argument(8, arg); read(arg, *) max_l
    call get_command_argument(9, arg); read(arg, *) L
    call get_command_argument(10, arg); read(arg, *) ipi
    call get_command_argument(11, arg); read(arg, *) ifull
    call get_command_argument(12, arg); read(arg, *) stride
    call get_command_argument(13, hs_file)
    call get_command_argument(14, out_file)

    call generate_grid(k, m, Z, h_max, r_max, grid)
    call splines%init(k, grid)
    write(6, *) "n_b: ", splines%n_b

    call r_k%init(max_k, splines)
    call r_m_k%init(max_k, splines)
    call r_d_k%init(max_k, splines)
    call setup_Slater_integrals(splines, max_k, k_GL, r_k, r_m_k, r_d_k)
    allocate(Nd_DOK :: R_p)
    call compute_R_k_map(r_d_k, r_k, r_m_k, splines, max_k, R_p)
    call compute_R_k_direct(splines, max_k, k_GL, R_direct)

    ! a strided sample of the banded keys [a,b,c,d] (|a-c| < k, |b-d| < k)
    n_pairs = 0
    do c = 1, splines%n_b
        do a = max(1, c - k + 1), min(splines%n_b, c + k - 1)
            n_pairs = n_pairs + 1
        end do
    end do
    n_keys = ((n_pairs + stride - 1) / stride) ** 2
    allocate(keys(4, n_keys), vals(0:max_k, n_keys), vals_direct(0:max_k, n_keys))
    n_keys = 0
    p = 0
    do c = 1, splines%n_b
        do a = max(1, c - k + 1), min(splines%n_b, c + k - 1)
            p = p + 1
            if (mod(p - 1, stride) /= 0) cycle
            q = 0
            do d = 1, splines%n_b
                do b = max(1, d - k + 1), min(splines%n_b, d + k - 1)
                    q = q + 1
                    if (mod(q - 1, stride) /= 0) cycle
                    n_keys = n_keys + 1
                    keys(:, n_keys) = [a, b, c, d]
                end do
            end do
        end do
    end do
    do i = 1, n_keys
        call R_p%get_val(keys(:, i), vals(:, i))
        call R_direct%get_val(keys(:, i), vals_direct(:, i))
    end do

    allocate(S(splines%n_b, splines%n_b), H_vec(0:max_l))
    if (trim(hs_file) == "device") then
        call pot%init(Z)
        call CAP_c%init(2, 2.0d0 * r_max, (0.0d0, 0.0d0))
        call setup_S(splines, k_GL, S)
        do i = 0, max_l
            call H_vec(i)%init(splines%n_b, splines%n_b)
            call setup_H_one_particle(pot, CAP_c, i, splines, k_GL, H_vec(i)%data)
        end do
    else
        open(newunit=u, file=trim(hs_file), access='stream', form='unformatted', status='old', action='read')
        read(u) S
        do i = 0, max_l
            call H_vec(i)%init(splines%n_b, splines%n_b)
            read(u) H_vec(i)%data
        end do
        close(u)
    end if

    term%l = L
    term%m = 0
    term%pi = ipi /= 0
    term%configs = count_configs(term, max_l, splines%n_b, splines%k, splines%n_b, splines%n_b, max_l)
    term%n_config = size(term%configs)
    call construct_block_tensor(H_vec, S, splines, term, max_k, R_p, H_sp, S_sp, ifull /= 0)
    write(6, '(a,i0,a,i0,a,i0)') "block L=", L, ": n_config ", term%n_config, ", nnz(H) ", H_sp%nnz

    open(newunit=u, file=trim(out_file), access='stream', form='unformatted', status='replace', action='write')
    write(u) size(grid), grid
    write(u) r_k%nnz, max_k + 1, r_k%data, r_m_k%data, r_k%iv, r_k%i, r_k%j
    write(u) r_d_k%nnz, r_d_k%data, r_d_k%iv, r_d_k%i, r_d_k%i_p, r_d_k%j, r_d_k%j_p
    write(u) n_keys, keys(:, 1:n_keys), vals(:, 1:n_keys), vals_direct(:, 1:n_keys)
    write(u) term%n_config
    do i = 1, term%n_config
        write(u) term%configs(i)%n, term%configs(i)%l, merge(1, 0, term%configs(i)%eqv)
    end do
    write(u) H_sp%nnz, H_sp%index_ptr, H_sp%indices, H_sp%data
    write(u) S_sp%nnz, S_sp%index_ptr, S_sp%indices, S_sp%data
    close(u)

    if (command_argument_count() >= 16) then
        call get_command_argument(15, arg); read(arg, *) max_L_all
        call get_command_argument(16, res_dir)
        allocate(eigs_v(splines%n_b, 0:max_l))
        eigs_v = (0.0_dp, 0.0_dp)
        call bas%init(max_L_all, max_l, splines%n_b, splines%k, splines%n_b, splines%n_b, max_l, .true., .true., &
                      eigs_v)
        call H_diag%init(bas%n_sym)
        call S_diag%init(bas%n_sym)
        call construct_blocks_tensor(H_vec, S, splines, bas%syms, max_k, R_p, H_diag, S_diag, ifull /= 0)
        gauge = 'l'
        if (command_argument_count() >= 17) then
            call get_command_argument(17, arg)
            gauge = arg(1:1)
        end if
        call setup_radial_dip(splines, k_GL, radial_dip, gauge)
        do iq = -1, 1
            call dipoles(iq)%init([bas%n_sym, bas%n_sym])
            do j = 1, bas%n_sym
                dsyms(2) = bas%syms(j)
                do i = 1, bas%n_sym
                    dsyms(1) = bas%syms(i)
                    call construct_dip_block_tensor(dsyms, iq, splines, S, radial_dip, dipoles(iq)%blocks(i, j), &
                                                    ifull /= 0 .or. i <= j)
                end do
            end do
        end do
        call H_diag%store(trim(res_dir) // "H_diag.dat")
        call S_diag%store(trim(res_dir) // "S_diag.dat")
        do iq = -1, 1
            write(qs, '(i0)') iq
            call dipoles(iq)%store(trim(res_dir) // "D_" // trim(qs) // ".dat")
        end do
        call bas%store(trim(res_dir))
        call splines%store(trim(res_dir))
        ! the input record of this run (input_tools.f90:765-773); the CAP is
        ! the one the device one-electron matrices use (order 2 at 2 r_max,
        ! eta = 0)
        call set_basis_defaults()
        in_dir = trim(res_dir)
        in_k = k
        in_m = m
        in_Z = Z
        in_h_max = h_max
        in_r_max = r_max
        in_k_GL = k_GL
        in_max_k = max_k
        in_max_L = max_L_all
        in_max_l_1p = max_l
        in_max_l2 = max_l
        in_cap_order = 2
        in_cap_r_0 = 2.0d0 * r_max
        in_cap_eta = (0.0d0, 0.0d0)
        in_gauge = gauge
        in_full = ifull /= 0
        in_z_pol = .true.
        in_two_el = .true.
        call write_basis_input(trim(res_dir))
        call get_basis_input(trim(res_dir) // "basis_input.dat")   ! what the other programs read back
        if (in_k /= k .or. in_max_L /= max_L_all .or. in_gauge /= gauge) error stop "basis_input.dat does not read back"
        write(6, '(a,i0,a)') "stored ", bas%n_sym, " symmetry blocks in " // trim(res_dir)
    end if

    call R_p%release()
    call R_direct%release()
    call bs2e_shutdown()
end program bs2e_basis_setup
