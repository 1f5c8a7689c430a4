! ref_capi.f90 -- bind(C) exports of the reference's own dependency-free
! Fortran modules, compiled from /root/reference by oracle/Makefile into
! oracle/_ref/libbs2e_ref.so.  TEST INFRASTRUCTURE ONLY: it is the checker
! used to pin the oracle's restatement bit for bit (tests and the fixture
! generator load it; the product never does).
!
! Wrapped routines (all from the reference, unchanged):
!   grid_tools.f90:6-55        generate_grid
!   bspline_tools.f90:27-56    b_spline%init / init_support
!   bspline_tools.f90:151-224  BVALUE_D (through eval_d / d_eval_d)
!   bspline_tools.f90:364-373  find_max_n_b
!   orbital_tools.f90:119-216  count_configs
module ref_capi
    use iso_c_binding, only: c_int64_t, c_double
    use grid_tools, only: generate_grid
    use bspline_tools, only: b_spline
    use orbital_tools, only: sym, config, count_configs
    implicit none
contains

    function ref_generate_grid(k, m, z, h_max, r_max, out, cap) result(n) bind(C, name="ref_generate_grid")
        integer(c_int64_t), value :: k, m, z, cap
        real(c_double), value :: h_max, r_max
        real(c_double), intent(out) :: out(cap)
        integer(c_int64_t) :: n
        double precision, allocatable :: grid(:)
        integer :: kk, mm, zz

        kk = k; mm = m; zz = z
        call generate_grid(kk, mm, zz, h_max, r_max, grid)
        n = size(grid)
        if (n <= cap) out(1:n) = grid
    end function ref_generate_grid

    ! Value (or i_deriv-th derivative) of spline s (1-based) at x on interval iv,
    ! with the unit coefficient vector the reference builds in mat_els.f90.
    function ref_bvalue_unit(t, n_knots, k, s, x, i_deriv, iv) result(v) bind(C, name="ref_bvalue_unit")
        integer(c_int64_t), value :: n_knots, k, s, i_deriv, iv
        real(c_double), intent(in) :: t(n_knots)
        real(c_double), value :: x
        real(c_double) :: v
        type(b_spline) :: spl
        double precision, allocatable :: c(:)
        integer :: kk, ivv, nd

        kk = k; ivv = iv; nd = i_deriv
        call spl%init(kk, t)
        allocate(c(spl%n), source=0.d0)
        c(s) = 1.d0
        if (nd == 0) then
            v = spl%eval_d(x, c, ivv)
        else
            v = spl%d_eval_d(x, c, nd, ivv)
        end if
    end function ref_bvalue_unit

    ! Batched form: values of every spline s_list(j) at x_list(j) on iv_list(j).
    subroutine ref_bvalue_batch(t, n_knots, k, npts, s_list, x_list, iv_list, i_deriv, out) &
            bind(C, name="ref_bvalue_batch")
        integer(c_int64_t), value :: n_knots, k, npts, i_deriv
        real(c_double), intent(in) :: t(n_knots)
        integer(c_int64_t), intent(in) :: s_list(npts), iv_list(npts)
        real(c_double), intent(in) :: x_list(npts)
        real(c_double), intent(out) :: out(npts)
        type(b_spline) :: spl
        double precision, allocatable :: c(:)
        integer :: kk, j, ivv, nd

        kk = k; nd = i_deriv
        call spl%init(kk, t)
        allocate(c(spl%n), source=0.d0)
        do j = 1, npts
            c(s_list(j)) = 1.d0
            ivv = iv_list(j)
            if (nd == 0) then
                out(j) = spl%eval_d(x_list(j), c, ivv)
            else
                out(j) = spl%d_eval_d(x_list(j), c, nd, ivv)
            end if
            c(s_list(j)) = 0.d0
        end do
    end subroutine ref_bvalue_batch

    ! n, n_b, number of breakpoints, and the support table (n_iv x n, column-major, 0/1).
    subroutine ref_spline_info(t, n_knots, k, n, n_b, n_bp, support, cap) bind(C, name="ref_spline_info")
        integer(c_int64_t), value :: n_knots, k, cap
        real(c_double), intent(in) :: t(n_knots)
        integer(c_int64_t), intent(out) :: n, n_b, n_bp
        integer(c_int64_t), intent(out) :: support(cap)
        type(b_spline) :: spl
        integer :: kk, i, j, nv

        kk = k
        call spl%init(kk, t)
        n = spl%n
        n_b = spl%n_b
        n_bp = size(spl%breakpoints)
        nv = n_bp - 1
        if (nv*spl%n <= cap) then
            do j = 1, spl%n
                do i = 1, nv
                    if (spl%support(i, j)) then
                        support((j-1)*nv + i) = 1
                    else
                        support((j-1)*nv + i) = 0
                    end if
                end do
            end do
        end if
    end subroutine ref_spline_info

    function ref_find_max_n_b(t, n_knots, k, x) result(res) bind(C, name="ref_find_max_n_b")
        integer(c_int64_t), value :: n_knots, k
        real(c_double), intent(in) :: t(n_knots)
        real(c_double), value :: x
        integer(c_int64_t) :: res
        type(b_spline) :: spl
        integer :: kk

        kk = k
        call spl%init(kk, t)
        res = spl%find_max_n_b(x)
    end function ref_find_max_n_b

    ! count_configs for the symmetry (L, M=0, parity term_pi); writes
    ! n1,n2,l1,l2,eqv per config into out(5*cap).  Orbital energies are only
    ! used for an unused local in the reference, so zeros are passed.
    function ref_count_configs(L, term_pi, max_l_1p, n_b, k_spline, max_n_b, n_all_l, l_2_max, out, cap) &
            result(cnt) bind(C, name="ref_count_configs")
        integer(c_int64_t), value :: L, term_pi, max_l_1p, n_b, k_spline, max_n_b, n_all_l, l_2_max, cap
        integer(c_int64_t), intent(out) :: out(5*cap)
        integer(c_int64_t) :: cnt
        type(sym) :: term
        type(config), allocatable :: confs(:)
        double complex, allocatable :: eigs(:,:)
        integer :: j

        term%l = L
        term%m = 0
        term%pi = (term_pi /= 0)
        allocate(eigs(n_b, 0:max_l_1p), source=dcmplx(0.d0, 0.d0))
        confs = count_configs(term, int(max_l_1p), int(n_b), int(k_spline), int(max_n_b), int(n_all_l), &
                              int(l_2_max), eigs)
        cnt = size(confs)
        if (cnt <= cap) then
            do j = 1, int(cnt)
                out(5*(j-1)+1) = confs(j)%n(1)
                out(5*(j-1)+2) = confs(j)%n(2)
                out(5*(j-1)+3) = confs(j)%l(1)
                out(5*(j-1)+4) = confs(j)%l(2)
                if (confs(j)%eqv) then
                    out(5*(j-1)+5) = 1
                else
                    out(5*(j-1)+5) = 0
                end if
            end do
        end if
    end function ref_count_configs

end module ref_capi
