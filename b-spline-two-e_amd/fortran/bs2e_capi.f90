! bs2e_capi.f90 -- ISO_C_BINDING interfaces to include/bs2e.h (libbs2e.so).
!
! Every integer of the C ABI is int64_t; arrays are passed by reference as
! assumed-size dummies, so Fortran allocatable arrays (contiguous,
! column-major) are handed over without copies.
module bs2e_capi
    use, intrinsic :: iso_c_binding
    implicit none
    private

    integer(c_int64_t), parameter, public :: BS2E_OK = 0, BS2E_EINVAL = -1, BS2E_EDEVICE = -2, &
        BS2E_ENOMEM = -3, BS2E_ERANGE = -4, BS2E_ESTATE = -5, BS2E_PREC_FP64 = 0, &
        BS2E_PREC_MIXED = 1

    public :: bs2e_create, bs2e_destroy, bs2e_last_error, bs2e_generate_grid, bs2e_count_nnz
    public :: bs2e_slater_primitives, bs2e_rk_create, bs2e_rk_from_primitives, bs2e_rk_compute
    public :: bs2e_rk_synchronize, bs2e_rk_info, bs2e_rk_copy_to_host, bs2e_rk_get_val, bs2e_rk_release
    public :: bs2e_ang_k_LS, bs2e_count_configs, bs2e_block_create, bs2e_block_sizes, bs2e_block_fill
    public :: bs2e_block_release, c_string, bs2e_setup_S, bs2e_setup_H_one_particle
    public :: bs2e_setup_radial_dip, bs2e_dip_block_create, bs2e_dip_block_sizes, bs2e_dip_block_fill
    public :: bs2e_dip_block_release, bs2e_setup_r_k_int
    public :: bs2e_blockset_create, bs2e_blockset_sizes, bs2e_blockset_fill, bs2e_blockset_get, bs2e_blockset_release
    public :: bs2e_blockset_fill_rows, bs2e_blockset_row_counts, bs2e_blockset_get_rows, bs2e_blockset_reserve

    interface
        function bs2e_create(ctx, device) bind(C, name="bs2e_create") result(rc)
            import :: c_ptr, c_int64_t
            type(c_ptr), intent(out) :: ctx
            integer(c_int64_t), value :: device
            integer(c_int64_t) :: rc
        end function

        function bs2e_destroy(ctx) bind(C, name="bs2e_destroy") result(rc)
            import :: c_ptr, c_int64_t
            type(c_ptr), value :: ctx
            integer(c_int64_t) :: rc
        end function

        function bs2e_last_error(ctx) bind(C, name="bs2e_last_error") result(msg)
            import :: c_ptr
            type(c_ptr), value :: ctx
            type(c_ptr) :: msg
        end function

        function bs2e_generate_grid(k, m, Z, h_max, r_max, knots, cap) bind(C, name="bs2e_generate_grid") result(n)
            import :: c_ptr, c_int64_t, c_double
            integer(c_int64_t), value :: k, m, Z
            real(c_double), value :: h_max, r_max
            type(c_ptr), value :: knots           ! NULL: count only
            integer(c_int64_t), value :: cap
            integer(c_int64_t) :: n
        end function

        function bs2e_count_nnz(k, n_knots, knots, nnz4, nnz6, n_pairs) bind(C, name="bs2e_count_nnz") result(rc)
            import :: c_int64_t, c_double
            integer(c_int64_t), value :: k, n_knots
            real(c_double), intent(in) :: knots(*)
            integer(c_int64_t), intent(out) :: nnz4, nnz6, n_pairs
            integer(c_int64_t) :: rc
        end function

        function bs2e_slater_primitives(ctx, k, n_knots, knots, k_gl, max_k, nnz4, rk, rmk, iv4, i4, j4, &
                                        nnz6, rdk, iv6, i6, ip6, j6, jp6) &
                bind(C, name="bs2e_slater_primitives") result(rc)
            import :: c_ptr, c_int64_t, c_double
            type(c_ptr), value :: ctx
            integer(c_int64_t), value :: k, n_knots, k_gl, max_k, nnz4, nnz6
            real(c_double), intent(in) :: knots(*)
            real(c_double), intent(out) :: rk(*), rmk(*), rdk(*)
            integer(c_int64_t), intent(out) :: iv4(*), i4(*), j4(*), iv6(*), i6(*), ip6(*), j6(*), jp6(*)
            integer(c_int64_t) :: rc
        end function

        function bs2e_setup_S(ctx, k, n_knots, knots, k_gl, S) bind(C, name="bs2e_setup_S") result(rc)
            import :: c_ptr, c_int64_t, c_double, c_double_complex
            type(c_ptr), value :: ctx
            integer(c_int64_t), value :: k, n_knots, k_gl
            real(c_double), intent(in) :: knots(*)
            complex(c_double_complex), intent(out) :: S(*)
            integer(c_int64_t) :: rc
        end function

        function bs2e_setup_H_one_particle(ctx, k, n_knots, knots, k_gl, Z, l_min, l_max, cap_order, cap_r0, &
                                           eta_re, eta_im, H) bind(C, name="bs2e_setup_H_one_particle") result(rc)
            import :: c_ptr, c_int64_t, c_double, c_double_complex
            type(c_ptr), value :: ctx
            integer(c_int64_t), value :: k, n_knots, k_gl, Z, l_min, l_max, cap_order
            real(c_double), value :: cap_r0, eta_re, eta_im
            real(c_double), intent(in) :: knots(*)
            complex(c_double_complex), intent(out) :: H(*)
            integer(c_int64_t) :: rc
        end function

        function bs2e_rk_create(ctx, k, n_knots, knots, k_gl, max_k, precision, rk) &
                bind(C, name="bs2e_rk_create") result(rc)
            import :: c_ptr, c_int64_t, c_double
            type(c_ptr), value :: ctx
            integer(c_int64_t), value :: k, n_knots, k_gl, max_k, precision
            real(c_double), intent(in) :: knots(*)
            type(c_ptr), intent(out) :: rk
            integer(c_int64_t) :: rc
        end function

        function bs2e_rk_from_primitives(ctx, k, n_knots, knots, max_k, nnz4, rk4, rmk, iv4, i4, j4, &
                                         nnz6, rdk, iv6, i6, ip6, j6, jp6, rk) &
                bind(C, name="bs2e_rk_from_primitives") result(rc)
            import :: c_ptr, c_int64_t, c_double
            type(c_ptr), value :: ctx
            integer(c_int64_t), value :: k, n_knots, max_k, nnz4, nnz6
            real(c_double), intent(in) :: knots(*), rk4(*), rmk(*), rdk(*)
            integer(c_int64_t), intent(in) :: iv4(*), i4(*), j4(*), iv6(*), i6(*), ip6(*), j6(*), jp6(*)
            type(c_ptr), intent(out) :: rk
            integer(c_int64_t) :: rc
        end function

        function bs2e_rk_compute(rk) bind(C, name="bs2e_rk_compute") result(rc)
            import :: c_ptr, c_int64_t
            type(c_ptr), value :: rk
            integer(c_int64_t) :: rc
        end function

        function bs2e_rk_synchronize(rk) bind(C, name="bs2e_rk_synchronize") result(rc)
            import :: c_ptr, c_int64_t
            type(c_ptr), value :: rk
            integer(c_int64_t) :: rc
        end function

        function bs2e_rk_info(rk, info, n) bind(C, name="bs2e_rk_info") result(rc)
            import :: c_ptr, c_int64_t
            type(c_ptr), value :: rk
            integer(c_int64_t), intent(out) :: info(*)
            integer(c_int64_t), value :: n
            integer(c_int64_t) :: rc
        end function

        function bs2e_rk_copy_to_host(rk, dst, n_elems) bind(C, name="bs2e_rk_copy_to_host") result(rc)
            import :: c_ptr, c_int64_t
            type(c_ptr), value :: rk, dst
            integer(c_int64_t), value :: n_elems
            integer(c_int64_t) :: rc
        end function

        function bs2e_rk_get_val(rk, indices, values, n_values) bind(C, name="bs2e_rk_get_val") result(rc)
            import :: c_ptr, c_int64_t, c_double
            type(c_ptr), value :: rk
            integer(c_int64_t), intent(in) :: indices(4)
            real(c_double), intent(out) :: values(*)
            integer(c_int64_t), value :: n_values
            integer(c_int64_t) :: rc
        end function

        function bs2e_rk_release(rk) bind(C, name="bs2e_rk_release") result(rc)
            import :: c_ptr, c_int64_t
            type(c_ptr), value :: rk
            integer(c_int64_t) :: rc
        end function

        function bs2e_ang_k_LS(k, l1a, l1b, l2a, l2b, L) bind(C, name="bs2e_ang_k_LS") result(v)
            import :: c_int64_t, c_double
            integer(c_int64_t), value :: k, l1a, l1b, l2a, l2b, L
            real(c_double) :: v
        end function

        function bs2e_count_configs(L, term_pi, max_l_1p, n_b, k_spline, max_n_b, n_all_l, l_2_max, configs, cap) &
                bind(C, name="bs2e_count_configs") result(n)
            import :: c_int64_t
            import :: c_ptr
            integer(c_int64_t), value :: L, term_pi, max_l_1p, n_b, k_spline, max_n_b, n_all_l, l_2_max, cap
            type(c_ptr), value :: configs         ! NULL: count only
            integer(c_int64_t) :: n
        end function

        function bs2e_block_create(rk, L, n_config, configs, full, blk) bind(C, name="bs2e_block_create") result(rc)
            import :: c_ptr, c_int64_t
            type(c_ptr), value :: rk
            integer(c_int64_t), value :: L, n_config, full
            integer(c_int64_t), intent(in) :: configs(*)
            type(c_ptr), intent(out) :: blk
            integer(c_int64_t) :: rc
        end function

        function bs2e_block_sizes(blk, sizes, n) bind(C, name="bs2e_block_sizes") result(rc)
            import :: c_ptr, c_int64_t
            type(c_ptr), value :: blk
            integer(c_int64_t), intent(out) :: sizes(*)
            integer(c_int64_t), value :: n
            integer(c_int64_t) :: rc
        end function

        function bs2e_block_fill(blk, max_l_1p, Hl, S, ip_H, ix_H, d_H, ip_S, ix_S, d_S) &
                bind(C, name="bs2e_block_fill") result(rc)
            import :: c_ptr, c_int64_t, c_double_complex
            type(c_ptr), value :: blk
            integer(c_int64_t), value :: max_l_1p
            complex(c_double_complex), intent(in) :: Hl(*), S(*)
            integer(c_int64_t), intent(out) :: ip_H(*), ix_H(*), ip_S(*), ix_S(*)
            complex(c_double_complex), intent(out) :: d_H(*), d_S(*)
            integer(c_int64_t) :: rc
        end function

        function bs2e_block_release(blk) bind(C, name="bs2e_block_release") result(rc)
            import :: c_ptr, c_int64_t
            type(c_ptr), value :: blk
            integer(c_int64_t) :: rc
        end function

        function bs2e_blockset_create(rk, n_sym, L, n_config, configs, full, out) &
                bind(C, name="bs2e_blockset_create") result(rc)
            import :: c_ptr, c_int64_t
            type(c_ptr), value :: rk
            integer(c_int64_t), value :: n_sym, full
            integer(c_int64_t), intent(in) :: L(*), n_config(*), configs(*)
            type(c_ptr), intent(out) :: out
            integer(c_int64_t) :: rc
        end function

        function bs2e_blockset_sizes(set, sizes, n) bind(C, name="bs2e_blockset_sizes") result(rc)
            import :: c_ptr, c_int64_t
            type(c_ptr), value :: set
            integer(c_int64_t), intent(out) :: sizes(*)
            integer(c_int64_t), value :: n
            integer(c_int64_t) :: rc
        end function

        function bs2e_blockset_fill(set, max_l_1p, Hl, S) bind(C, name="bs2e_blockset_fill") result(rc)
            import :: c_ptr, c_int64_t, c_double_complex
            type(c_ptr), value :: set
            integer(c_int64_t), value :: max_l_1p
            complex(c_double_complex), intent(in) :: Hl(*), S(*)
            integer(c_int64_t) :: rc
        end function

        function bs2e_blockset_get(set, sym, ipH, ixH, dH, ipS, ixS, dS) bind(C, name="bs2e_blockset_get") result(rc)
            import :: c_ptr, c_int64_t
            type(c_ptr), value :: set
            integer(c_int64_t), value :: sym
            type(c_ptr), value :: ipH, ixH, dH, ipS, ixS, dS
            integer(c_int64_t) :: rc
        end function

        ! A window of global rows (0-based); Hl and S as c_loc of the arrays, or
        ! c_null_ptr for both to reuse the matrices of the set's last fill.
        function bs2e_blockset_fill_rows(set, max_l_1p, Hl, S, row_lo, n_rows) &
                bind(C, name="bs2e_blockset_fill_rows") result(rc)
            import :: c_ptr, c_int64_t
            type(c_ptr), value :: set, Hl, S
            integer(c_int64_t), value :: max_l_1p, row_lo, n_rows
            integer(c_int64_t) :: rc
        end function

        function bs2e_blockset_reserve(set, nnz_H, nnz_S) bind(C, name="bs2e_blockset_reserve") result(rc)
            import :: c_ptr, c_int64_t
            type(c_ptr), value :: set
            integer(c_int64_t), value :: nnz_H, nnz_S
            integer(c_int64_t) :: rc
        end function

        function bs2e_blockset_row_counts(set, row_lo, n_rows, cnt_H, cnt_S) &
                bind(C, name="bs2e_blockset_row_counts") result(rc)
            import :: c_ptr, c_int64_t
            type(c_ptr), value :: set
            integer(c_int64_t), value :: row_lo, n_rows
            integer(c_int64_t), intent(out) :: cnt_H(*), cnt_S(*)
            integer(c_int64_t) :: rc
        end function

        function bs2e_blockset_get_rows(set, sym, row_lo, n_rows, ipH, ixH, dH, ipS, ixS, dS) &
                bind(C, name="bs2e_blockset_get_rows") result(rc)
            import :: c_ptr, c_int64_t
            type(c_ptr), value :: set
            integer(c_int64_t), value :: sym, row_lo, n_rows
            type(c_ptr), value :: ipH, ixH, dH, ipS, ixS, dS
            integer(c_int64_t) :: rc
        end function

        function bs2e_blockset_release(set) bind(C, name="bs2e_blockset_release") result(rc)
            import :: c_ptr, c_int64_t
            type(c_ptr), value :: set
            integer(c_int64_t) :: rc
        end function

        function bs2e_setup_r_k_int(ctx, k, n_knots, knots, k_gl, max_k, r_0, out) &
                bind(C, name="bs2e_setup_r_k_int") result(rc)
            import :: c_ptr, c_int64_t, c_double
            type(c_ptr), value :: ctx
            integer(c_int64_t), value :: k, n_knots, k_gl, max_k
            real(c_double), value :: r_0
            real(c_double), intent(in) :: knots(*)
            real(c_double), intent(out) :: out(*)
            integer(c_int64_t) :: rc
        end function

        function bs2e_setup_radial_dip(ctx, k, n_knots, knots, k_gl, gauge, m1, m2) &
                bind(C, name="bs2e_setup_radial_dip") result(rc)
            import :: c_ptr, c_int64_t, c_double
            type(c_ptr), value :: ctx
            integer(c_int64_t), value :: k, n_knots, k_gl, gauge
            real(c_double), intent(in) :: knots(*)
            type(c_ptr), value :: m1, m2
            integer(c_int64_t) :: rc
        end function

        function bs2e_dip_block_create(ctx, k, n_b, L1, M1, pi1, n1, cf1, L2, M2, pi2, n2, cf2, q, compute, out) &
                bind(C, name="bs2e_dip_block_create") result(rc)
            import :: c_ptr, c_int64_t
            type(c_ptr), value :: ctx
            integer(c_int64_t), value :: k, n_b, L1, M1, pi1, n1, L2, M2, pi2, n2, q, compute
            integer(c_int64_t), intent(in) :: cf1(*), cf2(*)
            type(c_ptr), intent(out) :: out
            integer(c_int64_t) :: rc
        end function

        function bs2e_dip_block_sizes(d, sizes, n) bind(C, name="bs2e_dip_block_sizes") result(rc)
            import :: c_ptr, c_int64_t
            type(c_ptr), value :: d
            integer(c_int64_t), intent(out) :: sizes(*)
            integer(c_int64_t), value :: n
            integer(c_int64_t) :: rc
        end function

        function bs2e_dip_block_fill(d, gauge, S, rad1, rad2, ip, ix, data) bind(C, name="bs2e_dip_block_fill") &
                result(rc)
            import :: c_ptr, c_int64_t
            type(c_ptr), value :: d
            integer(c_int64_t), value :: gauge
            type(c_ptr), value :: S, rad1, rad2, ip, ix, data
            integer(c_int64_t) :: rc
        end function

        function bs2e_dip_block_release(d) bind(C, name="bs2e_dip_block_release") result(rc)
            import :: c_ptr, c_int64_t
            type(c_ptr), value :: d
            integer(c_int64_t) :: rc
        end function

        function c_strlen(s) bind(C, name="strlen") result(n)
            import :: c_ptr, c_size_t
            type(c_ptr), value :: s
            integer(c_size_t) :: n
        end function
    end interface

contains

    ! Fortran copy of a NUL-terminated C string (the library's error messages).
    function c_string(p) result(s)
        type(c_ptr), intent(in) :: p
        character(len=:), allocatable :: s
        character(kind=c_char), pointer :: chars(:)
        integer(c_size_t) :: n, i
        if (.not. c_associated(p)) then
            s = ""
            return
        end if
        n = c_strlen(p)
        call c_f_pointer(p, chars, [n])
        allocate(character(len=n) :: s)
        do i = 1, n
            s(i:i) = chars(i)
        end do
    end function c_string

end module bs2e_capi
