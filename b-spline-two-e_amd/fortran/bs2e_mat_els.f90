! bs2e_mat_els.f90 -- Fortran drop-in for the two-electron radial path of
! edvinolo/b-spline-two-e, running on the MI355X through the C ABI.
!
! The public names and argument lists are the reference's (trailing arguments
! the path does not read, such as basis%init's eigs, are accepted), so that a
! caller written against them (main_basis_setup.f90:75-118) keeps its calls;
! it changes its `use` lines to this module (which holds all of them, where the
! reference spreads them over mat_els, bspline_tools, sparse_array_tools,
! hamiltonian, orbital_tools, grid_tools, potentials and CAP_tools) and
! declares its tensor class(Nd_DOK) (the reference: class(sparse_Nd) with an
! Nd_DOK allocated into it, main_basis_setup.f90:27,84):
!
!   generate_grid(k, m, Z, h_max, r_max, grid)           grid_tools.f90:6-55
!   b_spline%init(k, t), %find_max_n_b(x)                bspline_tools.f90:27-39, 364-373
!   sparse_4d / sparse_6d %init(max_k, b_splines)        sparse_array_tools.f90:9-31, 159-178
!   setup_Slater_integrals(b_splines, max_k, k_GL,
!                          r_k, r_m_k, r_d_k)            mat_els.f90:172-182
!   Nd_DOK (%init, %set_val, %get_val)                   sparse_array_tools.f90:49-61, 495-555
!   compute_R_k_map(r_d_k, r_k, r_m_k, b_splines,
!                   max_k, R)                            sparse_array_tools.f90:452-493
!   config, sym, count_configs(term, max_l_1p, n_b,
!       k_spline, max_n_b, n_all_l, l_2_max, eigs)       orbital_tools.f90:10-19, 119-216
!   block, CSR_matrix, construct_block_tensor(H, S,
!       b_splines, term, max_k, R_k, H_sp, S_sp, full)   hamiltonian.f90:106-283
!   hydrogenic (%init(Z)), CAP (%init(order, r_0, eta))  potentials.f90, CAP_tools.f90
!   setup_S(b_splines, k_GL, S)                          mat_els.f90:85-118
!   setup_H_one_particle(potential, CAP_c, l,
!                        b_splines, k_GL, H)             mat_els.f90:47-83
!   basis (%init, %compute_size, %store)                 orbital_tools.f90:21-34, 263-397
!   block_diag_CS (%init, %compute_shape, %store,
!                  %load), CSR_matrix%store              block_tools.f90:41-151, 458-524,
!                                                        sparse_array_tools.f90:695-742
!   b_spline%store(dir)                                  bspline_tools.f90:375-387
!
! plus compute_R_k_direct(b_splines, max_k, k_GL, R), which builds the tensor
! on the device straight from the knots without the host containers.
!
! Errors: a negative status from the library becomes `error stop` with the
! library's message, as the reference stops on its own errors.  The device
! context is created on first use (device 0, or bs2e_use_device).
module bs2e_mat_els
    use, intrinsic :: iso_c_binding
    use bs2e_capi
    implicit none
    private

    integer, parameter :: dp = kind(1.d0)

    public :: b_spline, sparse_4d, sparse_6d, Nd_DOK, config, sym, block, CSR_matrix
    public :: generate_grid, setup_Slater_integrals, compute_R_k_map, compute_R_k_direct
    public :: count_configs, construct_block_tensor, ang_k_LS, bs2e_use_device, bs2e_shutdown
    public :: hydrogenic, CAP, setup_S, setup_H_one_particle, basis, block_diag_CS
    public :: block_CS, radial_dipole, setup_radial_dip, construct_dip_block_tensor, rkh_multipoles
    public :: construct_blocks_tensor, BS2E_PREC_FP64, BS2E_PREC_MIXED

    type(c_ptr), save :: g_ctx = c_null_ptr
    integer, save :: g_device = 0

    type :: b_spline
        integer :: k = 0
        integer :: n = 0
        integer :: n_b = 0
        real(dp), allocatable :: knots(:)
        real(dp), allocatable :: breakpoints(:)
    contains
        procedure :: init => b_spline_init
        procedure :: find_max_n_b => b_spline_find_max_n_b
        procedure :: store => b_spline_store
    end type b_spline

    type :: sparse_4d
        integer :: nnz = 0
        real(dp), allocatable :: data(:,:)
        integer, allocatable :: iv(:), i(:), j(:)
    contains
        procedure :: init => init_4d
    end type sparse_4d

    type :: sparse_6d
        integer :: nnz = 0
        real(dp), allocatable :: data(:,:)
        integer, allocatable :: iv(:), i(:), j(:), i_p(:), j_p(:)
    contains
        procedure :: init => init_6d
    end type sparse_6d

    ! The banded R^k tensor, device-resident with a host mirror; the key
    ! [a,b,c,d] is located by arithmetic instead of a hash map.
    type :: Nd_DOK
        integer :: N = 4
        integer :: N_val = 0
        integer :: N_elements = 0
        type(c_ptr) :: handle = c_null_ptr
    contains
        procedure :: init => init_Nd_DOK
        procedure :: set_val => set_val_Nd_DOK
        procedure :: get_val => get_val_Nd_DOK
        procedure :: release => release_Nd_DOK
    end type Nd_DOK

    type :: config
        integer :: n(2) = 0
        integer :: l(2) = 0
        logical :: eqv = .false.
    end type config

    type :: sym
        integer :: l = 0
        integer :: m = 0
        logical :: pi = .false.
        integer :: n_config = 0
        type(config), allocatable :: configs(:)
    end type sym

    type :: block
        complex(dp), allocatable :: data(:,:)
    contains
        procedure :: init => block_init
    end type block

    type :: hydrogenic
        integer :: Z = 1
    contains
        procedure :: init => init_hydrogenic
    end type hydrogenic

    type :: CAP
        integer :: order = 2
        real(dp) :: r_0 = 0.0_dp
        complex(dp) :: eta = (0.0_dp, 0.0_dp)
    contains
        procedure :: init => init_CAP
    end type CAP

    type :: CSR_matrix
        integer :: shape(2) = 0
        integer :: nnz = 0
        integer, allocatable :: index_ptr(:)
        integer, allocatable :: indices(:)
        complex(dp), allocatable :: data(:)
    contains
        procedure :: store => CSR_store
    end type CSR_matrix

    type :: block_diag_CS
        integer :: block_shape(2) = 0
        integer :: shape(2) = 0
        type(CSR_matrix), allocatable :: blocks(:)
    contains
        procedure :: init => init_block_diag_CS
        procedure :: compute_shape => compute_shape_block_diag_CS
        procedure :: store => CS_block_diag_store
        procedure :: load => CS_block_diag_load
    end type block_diag_CS

    ! block_tools.f90:24-39: a matrix of CSR blocks (the dipole operator
    ! between every pair of symmetries)
    type :: block_CS
        integer :: block_shape(2) = 0
        integer :: shape(2) = 0
        type(CSR_matrix), allocatable :: blocks(:,:)
    contains
        procedure :: init => init_block_CS
        procedure :: compute_shape => compute_shape_block_CS
        procedure :: store => CS_block_store
        procedure :: load => CS_block_load
    end type block_CS

    ! mat_els.f90:14-19: radial dipole matrices of one gauge
    type :: radial_dipole
        complex(dp), allocatable :: r_mat(:,:)
        complex(dp), allocatable :: dr_mat(:,:)
        complex(dp), allocatable :: r_inv_mat(:,:)
    end type radial_dipole

    type :: basis
        integer :: max_l_1p = 0
        integer :: max_L = 0
        logical :: two_el = .true.
        integer :: n_sym = 0
        integer :: n_states = 0
        type(sym), allocatable :: syms(:)
        integer, allocatable :: sym_ptr(:)
    contains
        procedure :: init => init_basis
        procedure :: compute_size => compute_basis_size
        procedure :: store => store_basis
    end type basis

contains

    ! ---------------------------------------------------------------- context
    subroutine bs2e_use_device(device)
        integer, intent(in) :: device
        if (c_associated(g_ctx)) call bs2e_shutdown()
        g_device = device
    end subroutine bs2e_use_device

    subroutine bs2e_shutdown()
        integer(c_int64_t) :: rc
        if (c_associated(g_ctx)) rc = bs2e_destroy(g_ctx)
        g_ctx = c_null_ptr
    end subroutine bs2e_shutdown

    function ctx() result(c)
        type(c_ptr) :: c
        if (.not. c_associated(g_ctx)) call check(bs2e_create(g_ctx, int(g_device, c_int64_t)), "bs2e_create")
        c = g_ctx
    end function ctx

    subroutine check(rc, where)
        integer(c_int64_t), intent(in) :: rc
        character(len=*), intent(in) :: where
        if (rc < 0) then
            write(0, '(a,a,i0,a,a)') where, ": bs2e status ", rc, ": ", c_string(bs2e_last_error(c_null_ptr))
            error stop "bs2e: device library call failed"
        end if
    end subroutine check

    ! ---------------------------------------------------------------- basis
    subroutine generate_grid(k, m, Z, h_max, r_max, grid)
        integer, intent(in) :: k, m, Z
        real(dp), intent(in) :: h_max, r_max
        real(dp), allocatable, target, intent(out) :: grid(:)
        integer(c_int64_t) :: n
        n = bs2e_generate_grid(int(k, c_int64_t), int(m, c_int64_t), int(Z, c_int64_t), h_max, r_max, c_null_ptr, &
                               0_c_int64_t)
        call check(n, "generate_grid")
        allocate(grid(n))
        call check(bs2e_generate_grid(int(k, c_int64_t), int(m, c_int64_t), int(Z, c_int64_t), h_max, r_max, &
                                      c_loc(grid), n), "generate_grid")
    end subroutine generate_grid

    subroutine b_spline_init(this, k, t)
        class(b_spline), intent(inout) :: this
        integer, intent(in) :: k
        real(dp), intent(in) :: t(:)
        this%k = k
        this%n = size(t) - k
        this%n_b = this%n - 2
        this%knots = t
        this%breakpoints = t(k:size(t) - k + 1)
    end subroutine b_spline_init

    ! index of the last interval whose left breakpoint is below x (bspline_tools.f90:364-373)
    function b_spline_find_max_n_b(this, x) result(res)
        class(b_spline), intent(in) :: this
        real(dp), intent(in) :: x
        integer :: res, i
        res = -1
        do i = 1, size(this%breakpoints)
            if (this%breakpoints(i) >= x) then
                res = i - 1
                return
            end if
        end do
    end function b_spline_find_max_n_b

    subroutine block_init(this, n, m)
        class(block), intent(inout) :: this
        integer, intent(in) :: n, m
        if (allocated(this%data)) deallocate(this%data)
        allocate(this%data(n, m))
        this%data = (0.0_dp, 0.0_dp)
    end subroutine block_init

    ! ---------------------------------------------------------------- files
    ! Fortran unformatted sequential records, as the reference writes them.
    subroutine b_spline_store(this, dir)
        class(b_spline), intent(in) :: this
        character(len=*), intent(in) :: dir
        integer :: u
        open(file=dir // "splines.dat", newunit=u, action='write', form='unformatted')
        write(u) this%k
        write(u) size(this%knots)
        write(u) this%knots
        close(u)
    end subroutine b_spline_store

    subroutine CSR_store(this, loc)
        class(CSR_matrix), intent(in) :: this
        character(len=*), intent(in) :: loc
        integer :: u
        open(file=loc, newunit=u, action='write', form='unformatted')
        write(u) 'CSR'
        write(u) this%shape
        write(u) this%nnz
        write(u) this%index_ptr
        write(u) this%indices
        write(u) this%data
        close(u)
    end subroutine CSR_store

    subroutine init_block_diag_CS(this, n)
        class(block_diag_CS), intent(inout) :: this
        integer, intent(in) :: n
        this%block_shape = [n, n]
        if (allocated(this%blocks)) deallocate(this%blocks)
        allocate(this%blocks(n))
    end subroutine init_block_diag_CS

    subroutine compute_shape_block_diag_CS(this)
        class(block_diag_CS), intent(inout) :: this
        integer :: i
        this%shape = 0
        do i = 1, this%block_shape(1)
            this%shape = this%shape + this%blocks(i)%shape
        end do
    end subroutine compute_shape_block_diag_CS

    subroutine CS_block_diag_store(this, loc)
        class(block_diag_CS), intent(inout) :: this
        character(len=*), intent(in) :: loc
        integer :: i, u
        call this%compute_shape()
        open(file=loc, newunit=u, action='write', form='unformatted')
        write(u) 'CSR'
        write(u) this%block_shape
        write(u) this%shape
        do i = 1, this%block_shape(1)
            write(u) this%blocks(i)%shape
            write(u) this%blocks(i)%nnz
            if (this%blocks(i)%nnz > 0) then
                write(u) this%blocks(i)%index_ptr
                write(u) this%blocks(i)%indices
                write(u) this%blocks(i)%data
            end if
        end do
        close(u)
    end subroutine CS_block_diag_store

    subroutine CS_block_diag_load(this, loc)
        class(block_diag_CS), intent(inout) :: this
        character(len=*), intent(in) :: loc
        integer :: i, u, nnz, shape(2), block_shape(2)
        character(len=3) :: kind
        open(file=loc, newunit=u, action='read', form='unformatted', status='old')
        read(u) kind
        if (kind /= 'CSR') error stop "block_diag_CS%load: blocks must be CSR"
        read(u) block_shape
        call this%init(block_shape(1))
        read(u) this%shape
        do i = 1, block_shape(1)
            read(u) shape
            read(u) nnz
            this%blocks(i)%shape = shape
            this%blocks(i)%nnz = nnz
            allocate(this%blocks(i)%index_ptr(shape(1) + 1), this%blocks(i)%indices(nnz), this%blocks(i)%data(nnz))
            this%blocks(i)%index_ptr = 1
            if (nnz > 0) then
                read(u) this%blocks(i)%index_ptr
                read(u) this%blocks(i)%indices
                read(u) this%blocks(i)%data
            end if
        end do
        close(u)
    end subroutine CS_block_diag_load

    ! orbital_tools.f90:263-343 init_basis, two-electron symmetries: (0,0,even),
    ! then per L either (L, 0, parity of L) (z_pol) or every (L, M, parity)
    ! with |M| even for even parity and odd for odd parity.
    ! eigs (the one-particle energies the reference passes on to count_configs,
    ! which computes an energy from them and filters nothing with it,
    ! orbital_tools.f90:169-182) is accepted and not read.
    subroutine init_basis(this, max_L, max_l_1p, n_b, k_spline, max_n_b, n_all_l, l_2_max, z_pol, two_el, eigs)
        class(basis), intent(inout) :: this
        integer, intent(in) :: max_L, max_l_1p, n_b, k_spline, max_n_b, n_all_l, l_2_max
        logical, intent(in) :: z_pol, two_el
        complex(dp), intent(in), optional :: eigs(:,:)
        integer :: l, m, p, ptr
        if (.not. two_el) error stop "basis%init: only the two-electron basis is built here"
        this%max_l_1p = max_l_1p
        this%max_L = max_L
        this%two_el = two_el
        if (z_pol) then
            this%n_sym = max_L + 1
        else
            this%n_sym = (max_L + 1)**2
        end if
        if (allocated(this%syms)) deallocate(this%syms)
        allocate(this%syms(this%n_sym))
        this%syms(1)%l = 0
        this%syms(1)%m = 0
        this%syms(1)%pi = .false.
        ptr = 2
        do l = 1, max_L
            if (z_pol) then
                this%syms(ptr)%l = l
                this%syms(ptr)%m = 0
                this%syms(ptr)%pi = mod(l, 2) /= 0
                ptr = ptr + 1
            else
                do p = 0, 1
                    do m = -l, l
                        if (abs(mod(m, 2)) /= p) cycle
                        this%syms(ptr)%l = l
                        this%syms(ptr)%m = m
                        this%syms(ptr)%pi = p == 1
                        ptr = ptr + 1
                    end do
                end do
            end if
        end do
        do ptr = 1, this%n_sym
            this%syms(ptr)%configs = count_configs(this%syms(ptr), max_l_1p, n_b, k_spline, max_n_b, n_all_l, l_2_max)
            this%syms(ptr)%n_config = size(this%syms(ptr)%configs)
        end do
    end subroutine init_basis

    subroutine compute_basis_size(this)
        class(basis), intent(inout) :: this
        integer :: i, ptr
        this%n_states = 0
        do i = 1, this%n_sym
            this%n_states = this%n_states + this%syms(i)%n_config
        end do
        if (allocated(this%sym_ptr)) deallocate(this%sym_ptr)
        allocate(this%sym_ptr(this%n_sym + 1))
        ptr = 1
        do i = 1, this%n_sym
            this%sym_ptr(i) = ptr
            ptr = ptr + this%syms(i)%n_config
        end do
        this%sym_ptr(this%n_sym + 1) = ptr
    end subroutine compute_basis_size

    subroutine store_basis(this, loc)
        class(basis), intent(inout) :: this
        character(len=*), intent(in) :: loc
        integer :: i, u
        call this%compute_size()
        open(file=loc // "basis.dat", newunit=u, action='write', form='unformatted')
        write(u) this%max_l_1p
        write(u) this%max_L
        write(u) this%two_el
        write(u) this%n_sym
        write(u) this%n_states
        write(u) this%sym_ptr
        do i = 1, this%n_sym
            write(u) this%syms(i)%l
            write(u) this%syms(i)%m
            write(u) this%syms(i)%pi
            write(u) this%syms(i)%n_config
            write(u) this%syms(i)%configs
        end do
        close(u)
        open(file=loc // "basis_info.txt", newunit=u, action='write')
        write(u, "(a5,a5,a5,a5,a9)") "# Sym", "L", "M", "Pi", "states"
        do i = 1, this%n_sym
            write(u, "(i5,i5,i5,l5,i9)") i, this%syms(i)%l, this%syms(i)%m, this%syms(i)%pi, this%syms(i)%n_config
        end do
        close(u)
    end subroutine store_basis

    ! ---------------------------------------------------------------- one-electron
    subroutine init_hydrogenic(this, Z)
        class(hydrogenic), intent(inout) :: this
        integer, intent(in) :: Z
        this%Z = Z
    end subroutine init_hydrogenic

    subroutine init_CAP(this, order, r_0, eta)
        class(CAP), intent(inout) :: this
        integer, intent(in) :: order
        real(dp), intent(in) :: r_0
        complex(dp), intent(in) :: eta
        this%order = order
        this%r_0 = r_0
        this%eta = eta
    end subroutine init_CAP

    ! block_tools.f90:97-104
    subroutine init_block_CS(this, shape_in)
        class(block_CS), intent(inout) :: this
        integer, intent(in) :: shape_in(2)
        this%block_shape = shape_in
        if (allocated(this%blocks)) deallocate(this%blocks)
        allocate(this%blocks(shape_in(1), shape_in(2)))
    end subroutine init_block_CS

    ! block_tools.f90:127-140: the shape from the diagonal blocks
    subroutine compute_shape_block_CS(this)
        class(block_CS), intent(inout) :: this
        integer :: i
        this%shape = 0
        do i = 1, this%block_shape(1)
            this%shape(1) = this%shape(1) + this%blocks(i, i)%shape(1)
        end do
        do i = 1, this%block_shape(2)
            this%shape(2) = this%shape(2) + this%blocks(i, i)%shape(2)
        end do
    end subroutine compute_shape_block_CS

    ! block_tools.f90:386-415: 'CSR', block_shape, shape, then per block
    ! (j outer, i inner) its shape, nnz and, when nnz > 0, its arrays
    subroutine CS_block_store(this, loc)
        class(block_CS), intent(inout) :: this
        character(len=*), intent(in) :: loc
        integer :: i, j, u
        call this%compute_shape()
        open(file=loc, newunit=u, action='write', form='unformatted')
        write(u) 'CSR'
        write(u) this%block_shape
        write(u) this%shape
        do j = 1, this%block_shape(2)
            do i = 1, this%block_shape(1)
                write(u) this%blocks(i, j)%shape
                write(u) this%blocks(i, j)%nnz
                if (this%blocks(i, j)%nnz > 0) then
                    write(u) this%blocks(i, j)%index_ptr
                    write(u) this%blocks(i, j)%indices
                    write(u) this%blocks(i, j)%data
                end if
            end do
        end do
        close(u)
    end subroutine CS_block_store

    ! block_tools.f90:417-456
    subroutine CS_block_load(this, loc)
        class(block_CS), intent(inout) :: this
        character(len=*), intent(in) :: loc
        integer :: i, j, u, nnz, shape(2), block_shape(2)
        character(len=3) :: kind
        open(file=loc, newunit=u, action='read', form='unformatted', status='old')
        read(u) kind
        if (kind /= 'CSR') error stop "block_CS%load: only CSR blocks are implemented"
        read(u) block_shape
        call this%init(block_shape)
        read(u) this%shape
        do j = 1, block_shape(2)
            do i = 1, block_shape(1)
                read(u) shape
                read(u) nnz
                this%blocks(i, j)%shape = shape
                this%blocks(i, j)%nnz = nnz
                if (nnz > 0) then
                    allocate(this%blocks(i, j)%index_ptr(shape(1) + 1), this%blocks(i, j)%indices(nnz), &
                             this%blocks(i, j)%data(nnz))
                    read(u) this%blocks(i, j)%index_ptr
                    read(u) this%blocks(i, j)%indices
                    read(u) this%blocks(i, j)%data
                end if
            end do
        end do
        close(u)
    end subroutine CS_block_load

    ! mat_els.f90:120-170 setup_radial_dip on the device: gauge 'l' fills
    ! r_mat, 'v' dr_mat and r_inv_mat (:348-390)
    subroutine setup_radial_dip(b_splines, k_GL, radial_dip, gauge)
        type(b_spline), intent(in) :: b_splines
        integer, intent(in) :: k_GL
        type(radial_dipole), intent(out), target :: radial_dip
        character, intent(in) :: gauge
        integer :: nb
        nb = b_splines%n_b
        if (gauge == 'l') then
            allocate(radial_dip%r_mat(nb, nb))
            call check(bs2e_setup_radial_dip(ctx(), int(b_splines%k, c_int64_t), int(size(b_splines%knots), c_int64_t), &
                       b_splines%knots, int(k_GL, c_int64_t), 0_c_int64_t, c_loc(radial_dip%r_mat), c_null_ptr), &
                       "setup_radial_dip")
        else if (gauge == 'v') then
            allocate(radial_dip%dr_mat(nb, nb), radial_dip%r_inv_mat(nb, nb))
            call check(bs2e_setup_radial_dip(ctx(), int(b_splines%k, c_int64_t), int(size(b_splines%knots), c_int64_t), &
                       b_splines%knots, int(k_GL, c_int64_t), 1_c_int64_t, c_loc(radial_dip%dr_mat), &
                       c_loc(radial_dip%r_inv_mat)), "setup_radial_dip")
        else
            write(6, *) "Error! Unrecognized gauge option in for radial dipole: ", gauge
            stop
        end if
    end subroutine setup_radial_dip

    ! dipole.f90:8-47 construct_dip_block_tensor: the pattern (init_dip_block,
    ! :95-150) and the values ang * dip_mat_neq on the device.  The gauge is
    ! the one radial_dip was set up with.
    subroutine construct_dip_block_tensor(syms, q, b_splines, S, radial_dip, dip_block, compute)
        type(sym), intent(in) :: syms(2)
        integer, intent(in) :: q
        type(b_spline), intent(in) :: b_splines
        complex(dp), intent(in), target, contiguous :: S(:,:)
        type(radial_dipole), intent(in), target :: radial_dip
        type(CSR_matrix), intent(out) :: dip_block
        logical, intent(in) :: compute
        integer(c_int64_t), allocatable :: cf1(:), cf2(:)
        integer(c_int64_t), allocatable, target :: ip(:), ix(:)
        complex(dp), allocatable, target :: d(:)
        integer(c_int64_t) :: sizes(3), gauge
        type(c_ptr) :: blk, r1, r2
        if (allocated(radial_dip%r_mat)) then
            gauge = 0
            r1 = c_loc(radial_dip%r_mat)
            r2 = c_null_ptr
        else if (allocated(radial_dip%dr_mat) .and. allocated(radial_dip%r_inv_mat)) then
            gauge = 1
            r1 = c_loc(radial_dip%dr_mat)
            r2 = c_loc(radial_dip%r_inv_mat)
        else
            error stop "construct_dip_block_tensor: radial_dip has not been set up"
        end if
        call pack_configs(syms(1), cf1)
        call pack_configs(syms(2), cf2)
        call check(bs2e_dip_block_create(ctx(), int(b_splines%k, c_int64_t), int(b_splines%n_b, c_int64_t), &
                   int(syms(1)%l, c_int64_t), int(syms(1)%m, c_int64_t), merge(1_c_int64_t, 0_c_int64_t, syms(1)%pi), &
                   int(syms(1)%n_config, c_int64_t), cf1, int(syms(2)%l, c_int64_t), int(syms(2)%m, c_int64_t), &
                   merge(1_c_int64_t, 0_c_int64_t, syms(2)%pi), int(syms(2)%n_config, c_int64_t), cf2, &
                   int(q, c_int64_t), merge(1_c_int64_t, 0_c_int64_t, compute), blk), "construct_dip_block_tensor")
        call check(bs2e_dip_block_sizes(blk, sizes, 3_c_int64_t), "construct_dip_block_tensor")
        dip_block%shape = [syms(1)%n_config, syms(2)%n_config]
        dip_block%nnz = int(sizes(1))
        if (sizes(1) > 0) then
            allocate(ip(syms(1)%n_config + 1), ix(sizes(1)), d(sizes(1)))
            call check(bs2e_dip_block_fill(blk, gauge, c_loc(S), r1, r2, c_loc(ip), c_loc(ix), c_loc(d)), &
                       "construct_dip_block_tensor")
            dip_block%index_ptr = int(ip)
            dip_block%indices = int(ix)
            dip_block%data = d
        end if
        call check(bs2e_dip_block_release(blk), "construct_dip_block_tensor")
    end subroutine construct_dip_block_tensor

    subroutine pack_configs(term, cf)
        type(sym), intent(in) :: term
        integer(c_int64_t), allocatable, intent(out) :: cf(:)
        integer :: c
        allocate(cf(5 * max(term%n_config, 1)))
        cf = 0
        do c = 1, term%n_config
            cf(5 * (c - 1) + 1:5 * (c - 1) + 2) = int(term%configs(c)%n, c_int64_t)
            cf(5 * (c - 1) + 3:5 * (c - 1) + 4) = int(term%configs(c)%l, c_int64_t)
            cf(5 * c) = merge(1_c_int64_t, 0_c_int64_t, term%configs(c)%eqv)
        end do
    end subroutine pack_configs

    ! RKH_hamiltonian.f90:91-159 setup_r_k_int on the device, with the
    ! reference's module variables (k_GL of init_g_l, max_k, r_0) as arguments
    subroutine rkh_multipoles(b_splines, k_GL, max_k, r_0, r_k_int)
        type(b_spline), intent(in) :: b_splines
        integer, intent(in) :: k_GL, max_k
        real(dp), intent(in) :: r_0
        real(dp), intent(out) :: r_k_int(0:, :, :)
        if (size(r_k_int, 1) /= max_k + 1 .or. size(r_k_int, 2) /= b_splines%n_b .or. &
            size(r_k_int, 3) /= b_splines%n_b) error stop "rkh_multipoles: r_k_int must be (0:max_k, n_b, n_b)"
        call check(bs2e_setup_r_k_int(ctx(), int(b_splines%k, c_int64_t), int(size(b_splines%knots), c_int64_t), &
                   b_splines%knots, int(k_GL, c_int64_t), int(max_k, c_int64_t), r_0, r_k_int), "setup_r_k_int")
    end subroutine rkh_multipoles

    ! mat_els.f90:85-118 on the device
    subroutine setup_S(b_splines, k_GL, S)
        type(b_spline), intent(in) :: b_splines
        integer, intent(in) :: k_GL
        complex(dp), intent(out) :: S(:,:)
        call check(bs2e_setup_S(ctx(), int(b_splines%k, c_int64_t), int(size(b_splines%knots), c_int64_t), &
                                b_splines%knots, int(k_GL, c_int64_t), S), "setup_S")
    end subroutine setup_S

    ! mat_els.f90:47-83 on the device (hydrogenic potential, CAP)
    subroutine setup_H_one_particle(potential, CAP_c, l, b_splines, k_GL, H)
        type(hydrogenic), intent(in) :: potential
        type(CAP), intent(in) :: CAP_c
        integer, intent(in) :: l
        type(b_spline), intent(in) :: b_splines
        integer, intent(in) :: k_GL
        complex(dp), intent(out) :: H(:,:)
        call check(bs2e_setup_H_one_particle(ctx(), int(b_splines%k, c_int64_t), int(size(b_splines%knots), c_int64_t), &
                   b_splines%knots, int(k_GL, c_int64_t), int(potential%Z, c_int64_t), int(l, c_int64_t), &
                   int(l, c_int64_t), int(CAP_c%order, c_int64_t), CAP_c%r_0, real(CAP_c%eta, dp), aimag(CAP_c%eta), H), &
                   "setup_H_one_particle")
    end subroutine setup_H_one_particle

    ! ---------------------------------------------------------------- containers
    subroutine sizes_of(b_splines, nnz4, nnz6, n_pairs)
        type(b_spline), intent(in) :: b_splines
        integer(c_int64_t), intent(out) :: nnz4, nnz6, n_pairs
        call check(bs2e_count_nnz(int(b_splines%k, c_int64_t), int(size(b_splines%knots), c_int64_t), b_splines%knots, &
                                  nnz4, nnz6, n_pairs), "count_nnz")
    end subroutine sizes_of

    subroutine init_4d(this, max_k, b_splines)
        class(sparse_4d), intent(inout) :: this
        integer, intent(in) :: max_k
        type(b_spline), intent(in) :: b_splines
        integer(c_int64_t) :: nnz4, nnz6, n_pairs
        call sizes_of(b_splines, nnz4, nnz6, n_pairs)
        this%nnz = int(nnz4)
        if (allocated(this%data)) deallocate(this%data, this%iv, this%i, this%j)
        allocate(this%data(this%nnz, 0:max_k), this%iv(this%nnz), this%i(this%nnz), this%j(this%nnz))
        this%data = 0.0_dp
        this%iv = 0
        this%i = 0
        this%j = 0
    end subroutine init_4d

    subroutine init_6d(this, max_k, b_splines)
        class(sparse_6d), intent(inout) :: this
        integer, intent(in) :: max_k
        type(b_spline), intent(in) :: b_splines
        integer(c_int64_t) :: nnz4, nnz6, n_pairs
        call sizes_of(b_splines, nnz4, nnz6, n_pairs)
        this%nnz = int(nnz6)
        if (allocated(this%data)) deallocate(this%data, this%iv, this%i, this%j, this%i_p, this%j_p)
        allocate(this%data(this%nnz, 0:max_k), this%iv(this%nnz), this%i(this%nnz), this%j(this%nnz), &
                 this%i_p(this%nnz), this%j_p(this%nnz))
        this%data = 0.0_dp
        this%iv = 0
        this%i = 0
        this%j = 0
        this%i_p = 0
        this%j_p = 0
    end subroutine init_6d

    ! mat_els.f90:172-182: fills the primitive containers (initialised by the
    ! caller) on the device, in the reference's order.
    subroutine setup_Slater_integrals(b_splines, max_k, k_GL, r_k, r_m_k, r_d_k)
        type(b_spline), intent(in) :: b_splines
        integer, intent(in) :: max_k, k_GL
        type(sparse_4d), intent(inout) :: r_k, r_m_k
        type(sparse_6d), intent(inout) :: r_d_k
        integer(c_int64_t), allocatable :: iv4(:), i4(:), j4(:), iv6(:), i6(:), ip6(:), j6(:), jp6(:)
        allocate(iv4(r_k%nnz), i4(r_k%nnz), j4(r_k%nnz))
        allocate(iv6(r_d_k%nnz), i6(r_d_k%nnz), ip6(r_d_k%nnz), j6(r_d_k%nnz), jp6(r_d_k%nnz))
        call check(bs2e_slater_primitives(ctx(), int(b_splines%k, c_int64_t), int(size(b_splines%knots), c_int64_t), &
                   b_splines%knots, int(k_GL, c_int64_t), int(max_k, c_int64_t), int(r_k%nnz, c_int64_t), &
                   r_k%data, r_m_k%data, iv4, i4, j4, int(r_d_k%nnz, c_int64_t), r_d_k%data, &
                   iv6, i6, ip6, j6, jp6), "setup_Slater_integrals")
        r_k%iv = int(iv4)
        r_k%i = int(i4)
        r_k%j = int(j4)
        r_m_k%iv = int(iv4)
        r_m_k%i = int(i4)
        r_m_k%j = int(j4)
        r_d_k%iv = int(iv6)
        r_d_k%i = int(i6)
        r_d_k%i_p = int(ip6)
        r_d_k%j = int(j6)
        r_d_k%j_p = int(jp6)
    end subroutine setup_Slater_integrals

    ! ---------------------------------------------------------------- Nd_DOK
    subroutine init_Nd_DOK(this, N, N_val, nnz)
        class(Nd_DOK), intent(inout) :: this
        integer, intent(in) :: N, N_val, nnz
        call this%release()
        this%N = N
        this%N_val = N_val
        this%N_elements = nnz
    end subroutine init_Nd_DOK

    subroutine set_val_Nd_DOK(this, indices, val)
        class(Nd_DOK), intent(inout) :: this
        integer, intent(in) :: indices(:)
        real(dp), intent(in) :: val(:)
        error stop "Nd_DOK%set_val: the tensor is built on the device by compute_R_k_map; it is read-only"
    end subroutine set_val_Nd_DOK

    ! sparse_array_tools.f90:536-555; reentrant (reads the host mirror).
    subroutine get_val_Nd_DOK(this, indices, values)
        class(Nd_DOK), intent(inout) :: this
        integer, intent(in) :: indices(:)
        real(dp), intent(out) :: values(:)
        integer(c_int64_t) :: idx(4)
        if (.not. c_associated(this%handle)) error stop "Nd_DOK%get_val: no tensor (call compute_R_k_map first)"
        idx = int(indices(1:4), c_int64_t)
        call check(bs2e_rk_get_val(this%handle, idx, values, int(size(values), c_int64_t)), "Nd_DOK%get_val")
    end subroutine get_val_Nd_DOK

    subroutine release_Nd_DOK(this)
        class(Nd_DOK), intent(inout) :: this
        integer(c_int64_t) :: rc
        if (c_associated(this%handle)) rc = bs2e_rk_release(this%handle)
        this%handle = c_null_ptr
    end subroutine release_Nd_DOK

    subroutine finish_tensor(R, max_k)
        class(Nd_DOK), intent(inout) :: R
        integer, intent(in) :: max_k
        integer(c_int64_t) :: info(2)
        call check(bs2e_rk_compute(R%handle), "compute_R_k_map")
        call check(bs2e_rk_copy_to_host(R%handle, c_null_ptr, 0_c_int64_t), "compute_R_k_map")
        call check(bs2e_rk_info(R%handle, info, 2_c_int64_t), "compute_R_k_map")
        R%N = 4
        R%N_val = max_k + 1
        R%N_elements = int(info(2) * info(2))
    end subroutine finish_tensor

    ! sparse_array_tools.f90:452-493: the tensor from the caller's containers.
    subroutine compute_R_k_map(r_d_k, r_k, r_m_k, b_splines, max_k, R)
        type(sparse_6d), intent(in) :: r_d_k
        type(sparse_4d), intent(in) :: r_k, r_m_k
        type(b_spline), intent(in) :: b_splines
        integer, intent(in) :: max_k
        class(Nd_DOK), intent(inout) :: R
        call R%release()
        call check(bs2e_rk_from_primitives(ctx(), int(b_splines%k, c_int64_t), int(size(b_splines%knots), c_int64_t), &
                   b_splines%knots, int(max_k, c_int64_t), int(r_k%nnz, c_int64_t), r_k%data, r_m_k%data, &
                   int(r_k%iv, c_int64_t), int(r_k%i, c_int64_t), int(r_k%j, c_int64_t), int(r_d_k%nnz, c_int64_t), &
                   r_d_k%data, int(r_d_k%iv, c_int64_t), int(r_d_k%i, c_int64_t), int(r_d_k%i_p, c_int64_t), &
                   int(r_d_k%j, c_int64_t), int(r_d_k%j_p, c_int64_t), R%handle), "compute_R_k_map")
        call finish_tensor(R, max_k)
    end subroutine compute_R_k_map

    ! The whole build on the device from the knots (no host containers).
    ! precision: BS2E_PREC_FP64 (default) or BS2E_PREC_MIXED (the off-diagonal
    ! moments on the FP32 matrix cores from exact integer slices).
    subroutine compute_R_k_direct(b_splines, max_k, k_GL, R, precision)
        type(b_spline), intent(in) :: b_splines
        integer, intent(in) :: max_k, k_GL
        class(Nd_DOK), intent(inout) :: R
        integer, intent(in), optional :: precision
        integer(c_int64_t) :: prec
        prec = BS2E_PREC_FP64
        if (present(precision)) prec = int(precision, c_int64_t)
        call R%release()
        call check(bs2e_rk_create(ctx(), int(b_splines%k, c_int64_t), int(size(b_splines%knots), c_int64_t), &
                   b_splines%knots, int(k_GL, c_int64_t), int(max_k, c_int64_t), prec, R%handle), &
                   "compute_R_k_direct")
        call finish_tensor(R, max_k)
    end subroutine compute_R_k_direct

    ! ---------------------------------------------------------------- blocks
    function ang_k_LS(k, confs, L) result(res)
        integer, intent(in) :: k, L
        type(config), intent(in) :: confs(2)
        real(dp) :: res
        res = bs2e_ang_k_LS(int(k, c_int64_t), int(confs(1)%l(1), c_int64_t), int(confs(1)%l(2), c_int64_t), &
                            int(confs(2)%l(1), c_int64_t), int(confs(2)%l(2), c_int64_t), int(L, c_int64_t))
    end function ang_k_LS

    ! orbital_tools.f90:119-216.  eigs is accepted for the signature; the
    ! reference only forms an energy from it that no filter uses.
    function count_configs(term, max_l_1p, n_b, k_spline, max_n_b, n_all_l, l_2_max, eigs) result(res)
        type(sym), intent(in) :: term
        integer, intent(in) :: max_l_1p, n_b, k_spline, max_n_b, n_all_l, l_2_max
        complex(dp), intent(in), optional :: eigs(:,:)
        type(config), allocatable :: res(:)
        integer(c_int64_t), allocatable, target :: buf(:)
        integer(c_int64_t) :: n, pi, args(8)
        integer :: c
        pi = merge(1_c_int64_t, 0_c_int64_t, term%pi)
        args = [int(term%l, c_int64_t), pi, int(max_l_1p, c_int64_t), int(n_b, c_int64_t), &
                int(k_spline, c_int64_t), int(max_n_b, c_int64_t), int(n_all_l, c_int64_t), int(l_2_max, c_int64_t)]
        n = bs2e_count_configs(args(1), args(2), args(3), args(4), args(5), args(6), args(7), args(8), c_null_ptr, &
                               0_c_int64_t)
        call check(n, "count_configs")
        allocate(buf(5 * max(n, 1_c_int64_t)), res(n))
        call check(bs2e_count_configs(args(1), args(2), args(3), args(4), args(5), args(6), args(7), args(8), &
                                      c_loc(buf), n), "count_configs")
        do c = 1, int(n)
            res(c)%n = int(buf(5 * (c - 1) + 1:5 * (c - 1) + 2))
            res(c)%l = int(buf(5 * (c - 1) + 3:5 * (c - 1) + 4))
            res(c)%eqv = buf(5 * c) /= 0
        end do
    end function count_configs

    ! hamiltonian.f90:106-283: CSR H and S of one symmetry block, with the
    ! R^k-dependent values summed over angular channels on the device.
    ! H(l)%data, l = 0..max_l_1p, are the one-electron blocks.
    subroutine construct_block_tensor(H, S, b_splines, term, max_k, R_k, H_sp, S_sp, full)
        type(block), allocatable, intent(in) :: H(:)
        complex(dp), intent(in) :: S(:,:)
        type(b_spline), intent(in) :: b_splines
        type(sym), intent(in) :: term
        integer, intent(in) :: max_k
        class(Nd_DOK), intent(inout) :: R_k
        type(CSR_matrix), intent(out) :: H_sp, S_sp
        logical, intent(in) :: full
        integer(c_int64_t), allocatable :: cf(:), ipH(:), ixH(:), ipS(:), ixS(:)
        complex(dp), allocatable :: Hl(:,:,:), dH(:), dS(:)
        integer(c_int64_t) :: sizes(5)
        type(c_ptr) :: blk
        integer :: c, l, n, nb
        if (.not. c_associated(R_k%handle)) error stop "construct_block_tensor: R_k holds no tensor"
        if (R_k%N_val /= max_k + 1) error stop "construct_block_tensor: max_k differs from the tensor's"
        if (lbound(H, 1) /= 0) error stop "construct_block_tensor: H must be indexed from l = 0"
        n = term%n_config
        nb = b_splines%n_b
        allocate(cf(5 * max(n, 1)))
        do c = 1, n
            cf(5 * (c - 1) + 1:5 * (c - 1) + 2) = int(term%configs(c)%n, c_int64_t)
            cf(5 * (c - 1) + 3:5 * (c - 1) + 4) = int(term%configs(c)%l, c_int64_t)
            cf(5 * c) = merge(1_c_int64_t, 0_c_int64_t, term%configs(c)%eqv)
        end do
        allocate(Hl(nb, nb, 0:ubound(H, 1)))
        do l = 0, ubound(H, 1)
            Hl(:, :, l) = H(l)%data
        end do
        call check(bs2e_block_create(R_k%handle, int(term%l, c_int64_t), int(n, c_int64_t), cf, &
                                     merge(1_c_int64_t, 0_c_int64_t, full), blk), "construct_block_tensor")
        call check(bs2e_block_sizes(blk, sizes, 5_c_int64_t), "construct_block_tensor")
        allocate(ipH(n + 1), ipS(n + 1), ixH(max(sizes(1), 1_c_int64_t)), ixS(max(sizes(2), 1_c_int64_t)))
        allocate(dH(max(sizes(1), 1_c_int64_t)), dS(max(sizes(2), 1_c_int64_t)))
        call check(bs2e_block_fill(blk, int(ubound(H, 1), c_int64_t), Hl, S, ipH, ixH, dH, ipS, ixS, dS), &
                   "construct_block_tensor")
        call check(bs2e_block_release(blk), "construct_block_tensor")
        H_sp%shape = [n, n]
        S_sp%shape = [n, n]
        H_sp%nnz = int(sizes(1))
        S_sp%nnz = int(sizes(2))
        H_sp%index_ptr = int(ipH)
        S_sp%index_ptr = int(ipS)
        H_sp%indices = int(ixH(1:sizes(1)))
        S_sp%indices = int(ixS(1:sizes(2)))
        H_sp%data = dH(1:sizes(1))
        S_sp%data = dS(1:sizes(2))
    end subroutine construct_block_tensor

    ! main_basis_setup.f90:101-118: construct_block_tensor for every
    ! symmetry of syms, with one count and one fill launch over the
    ! configuration rows of all of them; H_diag%blocks(i), S_diag%blocks(i)
    ! receive symmetry i.  H_diag and S_diag must be init'ed with size(syms).
    subroutine construct_blocks_tensor(H, S, b_splines, syms, max_k, R_k, H_diag, S_diag, full)
        type(block), allocatable, intent(in) :: H(:)
        complex(dp), intent(in) :: S(:,:)
        type(b_spline), intent(in) :: b_splines
        type(sym), intent(in) :: syms(:)
        integer, intent(in) :: max_k
        class(Nd_DOK), intent(inout) :: R_k
        type(block_diag_CS), intent(inout) :: H_diag, S_diag
        logical, intent(in) :: full
        integer(c_int64_t), allocatable :: cf(:), Ls(:), ns(:), sizes(:)
        integer(c_int64_t), allocatable, target :: ipH(:), ixH(:), ipS(:), ixS(:)
        complex(dp), allocatable, target :: dH(:), dS(:)
        complex(dp), allocatable :: Hl(:,:,:)
        type(c_ptr) :: set
        integer :: i, c, l, n, nb, ns_tot, off
        if (.not. c_associated(R_k%handle)) error stop "construct_blocks_tensor: R_k holds no tensor"
        if (R_k%N_val /= max_k + 1) error stop "construct_blocks_tensor: max_k differs from the tensor's"
        if (lbound(H, 1) /= 0) error stop "construct_blocks_tensor: H must be indexed from l = 0"
        if (size(H_diag%blocks) /= size(syms) .or. size(S_diag%blocks) /= size(syms)) &
            error stop "construct_blocks_tensor: H_diag and S_diag must hold one block per symmetry"
        nb = b_splines%n_b
        ns_tot = sum(syms(:)%n_config)
        allocate(cf(5 * max(ns_tot, 1)), Ls(size(syms)), ns(size(syms)), sizes(5 * size(syms)))
        off = 0
        do i = 1, size(syms)
            Ls(i) = syms(i)%l
            ns(i) = syms(i)%n_config
            do c = 1, syms(i)%n_config
                cf(off + 5 * (c - 1) + 1:off + 5 * (c - 1) + 2) = int(syms(i)%configs(c)%n, c_int64_t)
                cf(off + 5 * (c - 1) + 3:off + 5 * (c - 1) + 4) = int(syms(i)%configs(c)%l, c_int64_t)
                cf(off + 5 * c) = merge(1_c_int64_t, 0_c_int64_t, syms(i)%configs(c)%eqv)
            end do
            off = off + 5 * syms(i)%n_config
        end do
        allocate(Hl(nb, nb, 0:ubound(H, 1)))
        do l = 0, ubound(H, 1)
            Hl(:, :, l) = H(l)%data
        end do
        call check(bs2e_blockset_create(R_k%handle, int(size(syms), c_int64_t), Ls, ns, cf, &
                                        merge(1_c_int64_t, 0_c_int64_t, full), set), "construct_blocks_tensor")
        call check(bs2e_blockset_sizes(set, sizes, int(size(sizes), c_int64_t)), "construct_blocks_tensor")
        call check(bs2e_blockset_fill(set, int(ubound(H, 1), c_int64_t), Hl, S), "construct_blocks_tensor")
        do i = 1, size(syms)
            n = syms(i)%n_config
            allocate(ipH(n + 1), ipS(n + 1), ixH(max(sizes(5 * i - 4), 1_c_int64_t)), &
                     ixS(max(sizes(5 * i - 3), 1_c_int64_t)))
            allocate(dH(max(sizes(5 * i - 4), 1_c_int64_t)), dS(max(sizes(5 * i - 3), 1_c_int64_t)))
            call check(bs2e_blockset_get(set, int(i, c_int64_t), c_loc(ipH), c_loc(ixH), c_loc(dH), c_loc(ipS), &
                                         c_loc(ixS), c_loc(dS)), "construct_blocks_tensor")
            H_diag%blocks(i)%shape = [n, n]
            S_diag%blocks(i)%shape = [n, n]
            H_diag%blocks(i)%nnz = int(sizes(5 * i - 4))
            S_diag%blocks(i)%nnz = int(sizes(5 * i - 3))
            H_diag%blocks(i)%index_ptr = int(ipH)
            S_diag%blocks(i)%index_ptr = int(ipS)
            H_diag%blocks(i)%indices = int(ixH(1:sizes(5 * i - 4)))
            S_diag%blocks(i)%indices = int(ixS(1:sizes(5 * i - 3)))
            H_diag%blocks(i)%data = dH(1:sizes(5 * i - 4))
            S_diag%blocks(i)%data = dS(1:sizes(5 * i - 3))
            deallocate(ipH, ipS, ixH, ixS, dH, dS)
        end do
        call check(bs2e_blockset_release(set), "construct_blocks_tensor")
    end subroutine construct_blocks_tensor

end module bs2e_mat_els


! The RKH frame's one-electron multipoles (RKH_hamiltonian.f90:14-28, 91-159)
! with the module state the reference keeps: r_0 and max_k (set by
! adiabatic_potentials.f90:127,138 setup_U_r), the GL order of init_g_l and
! the result r_k_int(0:max_k, n_b, n_b).
module bs2e_rkh
    use bs2e_mat_els, only: b_spline, rkh_multipoles
    implicit none
    private
    integer, parameter :: dp = kind(1.d0)
    public :: r_0, max_k, r_k_int, init_g_l, setup_r_k_int

    real(dp) :: r_0 = 0.0_dp
    integer :: max_k = 0
    real(dp), allocatable :: r_k_int(:,:,:)
    integer :: k_GL_rkh = 0

contains

    subroutine init_g_l(b_splines, k_GL)
        type(b_spline), intent(in) :: b_splines
        integer, intent(in) :: k_GL
        if (b_splines%n_b < 1) error stop "init_g_l: b_splines is not initialised"
        k_GL_rkh = k_GL
    end subroutine init_g_l

    subroutine setup_r_k_int(b_splines)
        type(b_spline), intent(in) :: b_splines
        if (k_GL_rkh < 1) error stop "setup_r_k_int: call init_g_l first"
        if (allocated(r_k_int)) deallocate(r_k_int)
        allocate(r_k_int(0:max_k, b_splines%n_b, b_splines%n_b))
        call rkh_multipoles(b_splines, k_GL_rkh, max_k, r_0, r_k_int)
    end subroutine setup_r_k_int

end module bs2e_rkh


! The basis_setup input record the other programs read back from the result
! directory (input_tools.f90:7-43 variables and namelist /basis_input/,
! :254-327 get_basis_input, :765-773 write_basis_input, :825-844
! set_basis_defaults): same group name, variable names and types.
module bs2e_basis_input
    implicit none
    private
    integer, parameter :: dp = kind(1.d0)
    public :: basis_output_dir, k, m, Z, max_k, k_GL, max_L, max_l_1p, max_l2, h_max, r_max, r_2_max, r_all_l
    public :: CAP_order, CAP_r_0, CAP_eta, gauge, z_pol, full, two_el
    public :: set_basis_defaults, write_basis_input, get_basis_input

    character(len=64) :: basis_output_dir = ''
    integer :: k, m, Z, max_k, k_GL, max_L, max_l_1p, max_l2
    real(dp) :: h_max, r_max, r_2_max, r_all_l
    integer :: CAP_order
    real(dp) :: CAP_r_0
    complex(dp) :: CAP_eta
    character :: gauge = 'l'
    logical :: z_pol, full, two_el

    namelist /basis_input/ basis_output_dir, k, m, Z, h_max, r_max, r_2_max, r_all_l, k_GL, CAP_order, CAP_r_0, &
        CAP_eta, max_L, max_l_1p, max_l2, max_k, gauge, z_pol, full, two_el

contains

    subroutine set_basis_defaults()
        k = 6
        m = 3
        Z = 2
        h_max = 0.5_dp
        r_max = 15.0_dp
        r_2_max = -1.0_dp
        r_all_l = -1.0_dp
        k_GL = k + 6
        CAP_order = 2
        CAP_r_0 = 10.0_dp
        CAP_eta = cmplx(1.d-3, 0.d0, kind=dp)
        max_L = 2
        max_l_1p = 5
        max_l2 = 5
        max_k = 4
        z_pol = .true.
        full = .true.
        two_el = .true.
    end subroutine set_basis_defaults

    ! Strings are written between double quotes (delim='quote', what
    ! gfortran's namelist output uses), so that any compiler's namelist
    ! input reads them back.
    subroutine write_basis_input(res_dir)
        character(len=*), intent(in) :: res_dir
        integer :: unit
        open(file=res_dir // 'basis_input.dat', newunit=unit, action='write', delim='quote')
        write(unit, nml=basis_input)
        close(unit)
    end subroutine write_basis_input

    subroutine get_basis_input(input_file)
        character(len=*), intent(in) :: input_file
        integer :: unit
        open(file=input_file, newunit=unit, action='read', status='old')
        read(unit, nml=basis_input)
        close(unit)
        if ((gauge /= 'l') .and. (gauge /= 'v')) error stop "get_basis_input: gauge must be l (length) or v (velocity)"
    end subroutine get_basis_input

end module bs2e_basis_input
