! Fortran side of the libmoed C ABI (include/moed.h).
!
!   module moed_capi : bind(C) interfaces of the moed_* entry points and the moed_params type
!   module moed_ed   : drop-in procedures with the reference's names and argument lists
!                      (ed_solve, ed_calc_green_ftn, lanczos_iteration, multiply_H, apply_c,
!                      generate_basis, calc_dev, minimization) and the module variables the
!                      reference's callers read (eigval, nev_calc, Gr, omega, nwloc, D_ev).
!
! The reference's DMFT driver (main.f90) switches to the GPU solver by replacing
! `use ed_hamiltonian / ed_solver / ed_green` with `use moed_ed` and calling moed_ed_init once
! after ed_read_options, passing its ed_config globals.  Errors stop the program with the
! library's message, as the reference's die() does (sys.F:5-41).
module moed_capi
    use iso_c_binding
    implicit none

    integer(c_int32_t), parameter :: MOED_F_EXACT_CMPLX = 1, MOED_F_BATCHED = 8, MOED_F_FIT_TREE = 16
    integer(c_int32_t), parameter :: MOED_F_FULL_REORTH = 32, MOED_F_NO_DEGENERACY_CHECK = 64

    type, bind(C) :: moed_params
        integer(c_int32_t) :: norb, nsite
        real(c_double) :: U, J, beta
        integer(c_int32_t) :: nstep, nev, nw
        type(c_ptr) :: ek, vk, tij
        integer(c_int32_t) :: nsector
        type(c_ptr) :: nelec, nup
        integer(c_int32_t) :: flags
        real(c_double) :: gs_tol
        integer(c_int32_t) :: gs_maxvec, device
    end type moed_params

    interface
        integer(c_int) function moed_create(p, h) bind(C, name="moed_create")
            import :: moed_params, c_ptr, c_int
            type(moed_params), intent(in) :: p
            type(c_ptr), intent(out) :: h
        end function
        integer(c_int) function moed_destroy(h) bind(C, name="moed_destroy")
            import :: c_ptr, c_int
            type(c_ptr), value :: h
        end function
        integer(c_int) function moed_last_error(h, buf, n) bind(C, name="moed_last_error")
            import :: c_ptr, c_int, c_char, c_int32_t
            type(c_ptr), value :: h
            character(kind=c_char), intent(out) :: buf(*)
            integer(c_int32_t), value :: n
        end function
        integer(c_int) function moed_set_bath(h, ek, vk) bind(C, name="moed_set_bath")
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: h
            real(c_double), intent(in) :: ek(*), vk(*)
        end function
        integer(c_int) function moed_sector_dim(h, nu, nd, dim) bind(C, name="moed_sector_dim")
            import :: c_ptr, c_int, c_int32_t, c_int64_t
            type(c_ptr), value :: h
            integer(c_int32_t), value :: nu, nd
            integer(c_int64_t), intent(out) :: dim
        end function
        integer(c_int) function moed_basis(h, nu, nd, up, dn) bind(C, name="moed_basis")
            import :: c_ptr, c_int, c_int32_t
            type(c_ptr), value :: h
            integer(c_int32_t), value :: nu, nd
            integer(c_int32_t), intent(out) :: up(*), dn(*)
        end function
        integer(c_int) function moed_hv(h, nu, nd, x, y) bind(C, name="moed_hv")
            import :: c_ptr, c_int, c_int32_t, c_double
            type(c_ptr), value :: h
            integer(c_int32_t), value :: nu, nd
            real(c_double), intent(in) :: x(*)
            real(c_double), intent(out) :: y(*)
        end function
        integer(c_int) function moed_apply_c(h, pm, iorb, ispin, nu, nd, x, y) bind(C, name="moed_apply_c")
            import :: c_ptr, c_int, c_int32_t, c_double
            type(c_ptr), value :: h
            integer(c_int32_t), value :: pm, iorb, ispin, nu, nd
            real(c_double), intent(in) :: x(*)
            real(c_double), intent(out) :: y(*)
        end function
        integer(c_int) function moed_lanczos(h, nu, nd, vec, nstep, nsc, a, b) bind(C, name="moed_lanczos")
            import :: c_ptr, c_int, c_int32_t, c_double
            type(c_ptr), value :: h
            integer(c_int32_t), value :: nu, nd, nstep
            real(c_double), intent(in) :: vec(*)
            integer(c_int32_t), intent(out) :: nsc
            real(c_double), intent(out) :: a(*), b(*)
        end function
        integer(c_int) function moed_solve(h, nev_calc, val, prob, sector, level) bind(C, name="moed_solve")
            import :: c_ptr, c_int, c_int32_t, c_double
            type(c_ptr), value :: h
            integer(c_int32_t), intent(out) :: nev_calc
            real(c_double), intent(out) :: val(*), prob(*)
            integer(c_int32_t), intent(out) :: sector(*), level(*)
        end function
        integer(c_int) function moed_green(h, nw, omega, gr) bind(C, name="moed_green")
            import :: c_ptr, c_int, c_int32_t, c_double, c_double_complex
            type(c_ptr), value :: h
            integer(c_int32_t), value :: nw
            real(c_double), intent(in) :: omega(*)
            complex(c_double_complex), intent(out) :: gr(*)
        end function
        integer(c_int) function moed_calc_dev(h, ek, vk, ef, nw, omega, dev) bind(C, name="moed_calc_dev")
            import :: c_ptr, c_int, c_int32_t, c_double, c_double_complex
            type(c_ptr), value :: h
            real(c_double), intent(in) :: ek(*), vk(*), ef(*), omega(*)
            integer(c_int32_t), value :: nw
            complex(c_double_complex), intent(out) :: dev(*)
        end function
        integer(c_int) function moed_chi2_grad(h, nfit, nx, x, nw, omega, dev, f, g) bind(C, name="moed_chi2_grad")
            import :: c_ptr, c_int, c_int32_t, c_double, c_double_complex
            type(c_ptr), value :: h
            integer(c_int32_t), value :: nfit, nx, nw
            real(c_double), intent(in) :: x(*), omega(*)
            complex(c_double_complex), intent(in) :: dev(*)
            real(c_double), intent(out) :: f(*), g(*)
        end function
        integer(c_int) function moed_minimize(h, nfit, nx, x, nw, omega, dev, ftol, fret, iter, status) &
                bind(C, name="moed_minimize")
            import :: c_ptr, c_int, c_int32_t, c_double, c_double_complex
            type(c_ptr), value :: h
            integer(c_int32_t), value :: nfit, nx, nw
            real(c_double), intent(inout) :: x(*)
            real(c_double), intent(in) :: omega(*)
            complex(c_double_complex), intent(in) :: dev(*)
            real(c_double), value :: ftol
            real(c_double), intent(out) :: fret(*)
            integer(c_int32_t), intent(out) :: iter(*), status(*)
        end function
        integer(c_int) function moed_omega_grid(nw, beta, omega) bind(C, name="moed_omega_grid")
            import :: c_int, c_int32_t, c_double
            integer(c_int32_t), value :: nw
            real(c_double), value :: beta
            real(c_double), intent(out) :: omega(*)
        end function
        integer(c_int) function moed_set_fit_launch(h, iterations) bind(C, name="moed_set_fit_launch")
            import :: c_ptr, c_int, c_int32_t
            type(c_ptr), value :: h
            integer(c_int32_t), value :: iterations
        end function
        integer(c_int) function moed_set_eig_chunk(h, vectors) bind(C, name="moed_set_eig_chunk")
            import :: c_ptr, c_int, c_int32_t
            type(c_ptr), value :: h
            integer(c_int32_t), value :: vectors
        end function
        integer(c_int) function moed_set_hv_shape(h, threads, strip) bind(C, name="moed_set_hv_shape")
            import :: c_ptr, c_int, c_int32_t
            type(c_ptr), value :: h
            integer(c_int32_t), value :: threads, strip
        end function
        integer(c_int) function moed_set_persistent(h, max_dim) bind(C, name="moed_set_persistent")
            import :: c_ptr, c_int, c_int64_t
            type(c_ptr), value :: h
            integer(c_int64_t), value :: max_dim
        end function
        integer(c_int) function moed_fit_nfev(h, nfit, nfev) bind(C, name="moed_fit_nfev")
            import :: c_ptr, c_int, c_int32_t
            type(c_ptr), value :: h
            integer(c_int32_t), value :: nfit
            integer(c_int32_t), intent(out) :: nfev(*)
        end function
        integer(c_int) function moed_band_tb(norb, nq, hk) bind(C, name="moed_band_tb")
            import :: c_int, c_int32_t, c_double
            integer(c_int32_t), value :: norb, nq
            real(c_double), intent(out) :: hk(*)
        end function
        integer(c_int) function moed_delta_new(h, nw, nq, rmu, hk, gr, dev, gksum, nocc) bind(C, name="moed_delta_new")
            import :: c_ptr, c_int, c_int32_t, c_double, c_double_complex
            type(c_ptr), value :: h
            integer(c_int32_t), value :: nw, nq
            real(c_double), value :: rmu
            real(c_double), intent(in) :: hk(*)
            complex(c_double_complex), intent(in) :: gr(*)
            complex(c_double_complex), intent(inout) :: dev(*)
            complex(c_double_complex), intent(out) :: gksum(*)
            real(c_double), intent(out) :: nocc(*)
        end function
        integer(c_int) function moed_dos_diag(h, nw, omega_real, small, ev, nsp, nsn, ap, bp, an, bn, aff) &
                bind(C, name="moed_dos_diag")
            import :: c_ptr, c_int, c_int32_t, c_double
            type(c_ptr), value :: h
            integer(c_int32_t), value :: nw, nsp, nsn
            real(c_double), intent(in) :: omega_real(*), ap(*), bp(*), an(*), bn(*)
            real(c_double), value :: small, ev
            real(c_double), intent(out) :: aff(*)
        end function
        integer(c_int) function moed_export_green(h, path) bind(C, name="moed_export_green")
            import :: c_ptr, c_int, c_char
            type(c_ptr), value :: h
            character(kind=c_char), intent(in) :: path(*)
        end function
        integer(c_int) function moed_import_green(h, path) bind(C, name="moed_import_green")
            import :: c_ptr, c_int, c_char
            type(c_ptr), value :: h
            character(kind=c_char), intent(in) :: path(*)
        end function
        integer(c_int) function moed_export_eigvec(dir, isector, iev, node, nloc, vec) bind(C, name="moed_export_eigvec")
            import :: c_int, c_int32_t, c_int64_t, c_double, c_char
            character(kind=c_char), intent(in) :: dir(*)
            integer(c_int32_t), value :: isector, iev, node
            integer(c_int64_t), value :: nloc
            real(c_double), intent(in) :: vec(*)
        end function
        integer(c_int) function moed_dos(h, nw, omega_real, small, weighted, aff) bind(C, name="moed_dos")
            import :: c_ptr, c_int, c_int32_t, c_double
            type(c_ptr), value :: h
            integer(c_int32_t), value :: nw, weighted
            real(c_double), intent(in) :: omega_real(*)
            real(c_double), value :: small
            real(c_double), intent(out) :: aff(*)
        end function
    end interface
end module moed_capi


module moed_ed
    use iso_c_binding
    use moed_capi
    implicit none
    private
    public :: basis_t, eigval_t, moed_ed_init, moed_ed_finalize, moed_ed_set_bath, generate_basis, multiply_H, &
              apply_c, lanczos_iteration, ed_solve, ed_calc_green_ftn, calc_dev, minimization, &
              ed_delta_new_gpu, dos, eigval, nev_calc, Gr, omega, nwloc, D_ev, moed_h, dos_small

    ! ed_basis.F90:10-30 (one process owns the whole sector: nloc = ntot)
    type basis_t
        integer :: nloc, ntot, nup, ndown, ne_up, ne_down
        integer, allocatable :: up(:), down(:)
    end type basis_t

    ! ed_solver.F90:15-20
    type eigval_t
        double precision :: val, prob
        integer :: sector, level
    end type eigval_t

    type(c_ptr), save :: moed_h = c_null_ptr
    type(eigval_t), allocatable, save :: eigval(:)
    integer, save :: nev_calc = 0
    integer, save :: nwloc = 0
    double precision, allocatable, save :: omega(:)
    double complex, allocatable, save :: Gr(:,:), D_ev(:,:)
    integer, save :: m_norb = 0, m_nsite = 0, m_nev = 0, m_nstep = 0
    ! DMFT.small of the reference (ed_config.f90:116), the broadening dos uses
    double precision, save :: dos_small = 0.05d0

    ! storage the handle points into must outlive moed_create
    double precision, allocatable, target, save :: s_ek(:), s_vk(:)
    integer(c_int32_t), allocatable, target, save :: s_nelec(:), s_nup(:)

contains

    subroutine check(rc, where)
        integer(c_int), intent(in) :: rc
        character(len=*), intent(in) :: where
        character(kind=c_char) :: buf(512)
        character(len=512) :: msg
        integer :: i, ic
        if (rc == 0) return
        msg = ' '
        if (c_associated(moed_h)) then
            ic = moed_last_error(moed_h, buf, 512_c_int32_t)
            do i = 1, 512
                if (buf(i) == c_null_char) exit
                msg(i:i) = buf(i)
            enddo
        endif
        write(6, '(a,a,a,i0,a,a)') 'moed: ', where, ' failed (', rc, '): ', trim(msg)
        error stop 1
    end subroutine check

    ! Replaces the solver part of ed_read_options + ed_green_init (ed_config.f90:43, ed_green.F90:25):
    ! ek(1:Nsite), vk(1:Norb, Norb+1:Nsite) and the sectors as in ed_config.f90.
    subroutine moed_ed_init(Norb, Nbath, U, Jex, beta, Nstep, Nev, Nw, ek, vk, Nsector, nelec, nup, flags)
        integer, intent(in) :: Norb, Nbath, Nstep, Nev, Nw, Nsector
        double precision, intent(in) :: U, Jex, beta, ek(:), vk(:,:)
        integer, intent(in) :: nelec(:), nup(:)
        integer, intent(in), optional :: flags
        type(moed_params) :: p
        integer(c_int) :: rc
        m_norb = Norb; m_nsite = Norb + Nbath; m_nev = Nev; m_nstep = Nstep
        allocate(s_ek(m_nsite), s_vk(Norb*Nbath), s_nelec(Nsector), s_nup(Nsector))
        s_ek = ek(1:m_nsite)
        s_vk = reshape(vk(1:Norb, 1:Nbath), [Norb*Nbath])
        s_nelec = nelec(1:Nsector)
        s_nup = nup(1:Nsector)
        p%norb = Norb; p%nsite = m_nsite; p%U = U; p%J = Jex; p%beta = beta
        p%nstep = Nstep; p%nev = Nev; p%nw = Nw
        p%ek = c_loc(s_ek); p%vk = c_loc(s_vk); p%tij = c_null_ptr
        p%nsector = Nsector; p%nelec = c_loc(s_nelec); p%nup = c_loc(s_nup)
        p%flags = 0
        if (present(flags)) p%flags = flags
        p%gs_tol = 0.0d0; p%gs_maxvec = 0; p%device = 0
        rc = moed_create(p, moed_h)
        call check(rc, 'moed_create')
        nwloc = Nw
        allocate(omega(Nw), Gr(Norb, Nw), D_ev(Norb, Nw))
        rc = moed_omega_grid(int(Nw, c_int32_t), beta, omega)
        call check(rc, 'moed_omega_grid')
        Gr = (0.0d0, 0.0d0)
    end subroutine moed_ed_init

    subroutine moed_ed_finalize()
        integer(c_int) :: rc
        if (c_associated(moed_h)) rc = moed_destroy(moed_h)
        moed_h = c_null_ptr
        if (allocated(s_ek)) deallocate(s_ek, s_vk, s_nelec, s_nup)
        if (allocated(omega)) deallocate(omega, Gr, D_ev)
        if (allocated(eigval)) deallocate(eigval)
    end subroutine moed_ed_finalize

    ! the DMFT loop's broadcast of the fitted bath (main.f90:103-106)
    subroutine moed_ed_set_bath(ek, vk)
        double precision, intent(in) :: ek(:), vk(:,:)
        s_ek = ek(1:m_nsite)
        s_vk = reshape(vk, [size(s_vk)])
        call check(moed_set_bath(moed_h, s_ek, s_vk), 'moed_set_bath')
    end subroutine moed_ed_set_bath

    ! generate_basis (ed_basis.F90:35)
    type(basis_t) function generate_basis(ne_up, ne_down) result(b)
        integer, intent(in) :: ne_up, ne_down
        integer(c_int64_t) :: d1, d2
        call check(moed_sector_dim(moed_h, int(ne_up, c_int32_t), 0_c_int32_t, d1), 'generate_basis')
        call check(moed_sector_dim(moed_h, int(ne_down, c_int32_t), 0_c_int32_t, d2), 'generate_basis')
        b%ne_up = ne_up; b%ne_down = ne_down
        b%nup = int(d1); b%ndown = int(d2); b%ntot = b%nup * b%ndown; b%nloc = b%ntot
        allocate(b%up(b%nup), b%down(b%ndown))
        call check(moed_basis(moed_h, int(ne_up, c_int32_t), int(ne_down, c_int32_t), b%up, b%down), 'generate_basis')
    end function generate_basis

    ! multiply_H (ed_hamiltonian.F90:88)
    subroutine multiply_H(basis, A, B)
        type(basis_t), intent(in) :: basis
        double precision, intent(in) :: A(basis%nloc)
        double precision, intent(out) :: B(basis%nloc)
        call check(moed_hv(moed_h, int(basis%ne_up, c_int32_t), int(basis%ne_down, c_int32_t), A, B), 'multiply_H')
    end subroutine multiply_H

    ! apply_c (ed_operators.F90:10)
    subroutine apply_c(basis, vec, pm, iorb, ispin, basis_out, vec_out)
        type(basis_t), intent(in) :: basis
        double precision, intent(in) :: vec(basis%nloc)
        integer, intent(in) :: pm, iorb, ispin
        type(basis_t), intent(out) :: basis_out
        double precision, allocatable, intent(out) :: vec_out(:)
        integer :: d
        d = merge(1, -1, pm == 1)
        if (ispin == 1) then
            basis_out = generate_basis(basis%ne_up + d, basis%ne_down)
        else
            basis_out = generate_basis(basis%ne_up, basis%ne_down + d)
        endif
        allocate(vec_out(basis_out%nloc))
        call check(moed_apply_c(moed_h, int(pm, c_int32_t), int(iorb, c_int32_t), int(ispin, c_int32_t), &
                                int(basis%ne_up, c_int32_t), int(basis%ne_down, c_int32_t), vec, vec_out), 'apply_c')
    end subroutine apply_c

    ! lanczos_iteration (ed_lanczos.F90:14); a, b dimensioned 0:Nstep as in the reference
    subroutine lanczos_iteration(basis, vec, nstep_calc, a, b)
        type(basis_t), intent(in) :: basis
        double precision, intent(in) :: vec(basis%nloc)
        integer, intent(out) :: nstep_calc
        double precision, intent(out) :: a(0:m_nstep), b(0:m_nstep)
        integer(c_int32_t) :: nsc
        call check(moed_lanczos(moed_h, int(basis%ne_up, c_int32_t), int(basis%ne_down, c_int32_t), vec, &
                                int(m_nstep, c_int32_t), nsc, a, b), 'lanczos_iteration')
        nstep_calc = nsc
    end subroutine lanczos_iteration

    ! ed_solve (ed_solver.F90:28): fills eigval(:) and nev_calc; eigenvectors stay on the device
    subroutine ed_solve()
        double precision :: val(m_nev), prob(m_nev)
        integer(c_int32_t) :: sec(m_nev), lev(m_nev), nc
        integer :: i
        call check(moed_solve(moed_h, nc, val, prob, sec, lev), 'ed_solve')
        nev_calc = nc
        if (allocated(eigval)) deallocate(eigval)
        allocate(eigval(nev_calc))
        do i = 1, nev_calc
            eigval(i)%val = val(i); eigval(i)%prob = prob(i)
            eigval(i)%sector = sec(i); eigval(i)%level = lev(i)
        enddo
    end subroutine ed_solve

    ! ed_calc_green_ftn (ed_green.F90:60): Gr(Norb, nwloc) at omega(:)
    subroutine ed_calc_green_ftn(nev_in)
        integer, intent(in) :: nev_in
        if (nev_in /= nev_calc) call check(-1_c_int, 'ed_calc_green_ftn: nev_calc differs from ed_solve')
        call check(moed_green(moed_h, int(nwloc, c_int32_t), omega, Gr), 'ed_calc_green_ftn')
    end subroutine ed_calc_green_ftn

    ! calc_dev (ed_green.F90:260) from the given bath and impurity levels ef
    subroutine calc_dev(ek, vk, ef)
        double precision, intent(in) :: ek(:), vk(:,:), ef(:)
        call check(moed_calc_dev(moed_h, ek, reshape(vk, [size(vk)]), ef, int(nwloc, c_int32_t), omega, D_ev), &
                   'calc_dev')
    end subroutine calc_dev

    ! minimization (func.f90:144) with the reference's argument list; comm is unused
    subroutine minimization(x, nwl, Ns, nxsize, om, Dev, Nbath, Norb, comm, dab)
        integer, intent(in) :: nwl, Ns, nxsize, Nbath, Norb, comm
        double precision, intent(inout) :: x(nxsize)
        double precision, intent(in) :: om(nwl)
        double complex, intent(in) :: Dev(Norb, nwl)
        double precision, intent(out) :: dab
        double precision :: fret(1)
        integer(c_int32_t) :: it(1), st(1)
        call check(moed_minimize(moed_h, 1_c_int32_t, int(nxsize, c_int32_t), x, int(nwl, c_int32_t), om, Dev, &
                                 1.0d-7, fret, it, st), 'minimization')
        if (st(1) /= 0) call check(int(st(1), c_int), 'minimization: frprmn maximum iterations exceeded')
        dab = fret(1)
    end subroutine minimization

    ! ed_delta_new + n_from_gksum (ed_green.F90:187-253) on the device, for the caller's Gr, D_ev
    ! (updated in place) and Gksum; hkd(o + Norb*(q-1)) = dble(Hk(o,o,q)) (moed_band_tb gives the
    ! reference's tight-binding band)
    subroutine ed_delta_new_gpu(Grl, Devl, Gksuml, rMu, Nq, hkd, nocc)
        double complex, intent(in) :: Grl(:,:)
        double complex, intent(inout) :: Devl(:,:)
        double complex, intent(out) :: Gksuml(:,:)
        double precision, intent(in) :: rMu, hkd(:)
        integer, intent(in) :: Nq
        double precision, intent(out) :: nocc(:)
        call check(moed_delta_new(moed_h, int(size(Grl, 2), c_int32_t), int(Nq, c_int32_t), rMu, hkd, Grl, Devl, &
                                  Gksuml, nocc), 'ed_delta_new')
    end subroutine ed_delta_new_gpu

    ! dos (dos.f90:1-63) with the reference's argument list, from the records of the last
    ! ed_calc_green_ftn; each record overwrites its orbital's row as in the reference
    subroutine dos(omega_real, Aff)
        double precision, intent(in) :: omega_real(:)
        double precision, intent(out) :: Aff(:,:)
        call check(moed_dos(moed_h, int(size(omega_real), c_int32_t), omega_real, dos_small, 0_c_int32_t, Aff), &
                   'dos')
    end subroutine dos

end module moed_ed
