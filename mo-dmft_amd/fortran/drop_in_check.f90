! A program in the shape of the reference's main.f90 (main.f90:59-127) that reaches the device only
! through the drop-in procedures of module moed_ed: the single-sector calls (generate_basis,
! multiply_H, apply_c, lanczos_iteration), then one solve and Green's function (ed_solve,
! ed_calc_green_ftn), calc_dev, the lattice step (ed_delta_new), one bath fit (minimization) and the
! real-axis spectral function (dos).  It reads the model from a stream file and writes every
! result to another, so that the results can be compared with the same calls made through the C
! ABI.
!
!   drop_in_check <input> <output>
!
! input (little-endian stream): int32 norb, nbath, nstep, nev, nw, nsector, nu, nd, nq, nwr;
!   real64 U, J, beta, rMu; real64 ek(nsite), vk(norb, nbath); int32 nelec(nsector), nup(nsector);
!   real64 x(dim of sector (nu, nd)), omega_real(nwr)
! output: H x, c+(orbital 1, up) x, nstep_calc (int32), a(0:nstep), b(0:nstep), nev_calc (int32),
!   (val, prob, sector, level) per kept state, Gr(norb, nw), D_ev(norb, nw) after the lattice step,
!   Gksum(norb, nw), nocc(norb), the fitted x(nx), chi^2, Aff(norb, nwr)
program drop_in_check
    use iso_c_binding
    use moed_capi, only: moed_band_tb
    use moed_ed
    implicit none
    character(len=512) :: fin, fout
    integer :: norb, nbath, nstep, nev, nw, nsector, nu, nd, nq, nwr, nsite, nx, i, nsc
    double precision :: U, Jex, beta, rmu, dab
    double precision, allocatable :: ek(:), vk(:,:), x(:), y(:), vout(:), a(:), b(:), hkd(:), nocc(:), xb(:)
    double precision, allocatable :: omr(:), aff(:,:), ef(:)
    integer, allocatable :: nelec(:), nup(:)
    double complex, allocatable :: gks(:,:)
    type(basis_t) :: bas, bout

    call get_command_argument(1, fin)
    call get_command_argument(2, fout)
    open(10, file=trim(fin), access='stream', form='unformatted', status='old')
    read(10) norb, nbath, nstep, nev, nw, nsector, nu, nd, nq, nwr
    read(10) U, Jex, beta, rmu
    nsite = norb + nbath
    allocate(ek(nsite), vk(norb, nbath), nelec(nsector), nup(nsector))
    read(10) ek, vk, nelec, nup
    call moed_ed_init(norb, nbath, U, Jex, beta, nstep, nev, nw, ek, vk, nsector, nelec, nup)
    bas = generate_basis(nu, nd)
    allocate(x(bas%nloc), y(bas%nloc), omr(nwr), aff(norb, nwr))
    read(10) x, omr
    close(10)

    call multiply_H(bas, x, y)
    call apply_c(bas, x, 1, 1, 1, bout, vout)
    allocate(a(0:nstep), b(0:nstep))
    call lanczos_iteration(bas, x, nsc, a, b)

    call ed_solve()
    call ed_calc_green_ftn(nev_calc)
    allocate(ef(norb))
    ef = ek(1:norb)
    call calc_dev(ek, vk, ef)
    allocate(hkd(norb*nq), nocc(norb), gks(norb, nw))
    if (moed_band_tb(int(norb, c_int32_t), int(nq, c_int32_t), hkd) /= 0) error stop 'moed_band_tb'
    call ed_delta_new_gpu(Gr, D_ev, gks, rmu, nq, hkd, nocc)
    ! ev_to_x (func.f90:165-191): bath energies, V orbital by orbital, then ef
    nx = nbath*(1 + norb) + norb
    allocate(xb(nx))
    xb(1:nbath) = ek(norb+1:nsite)
    do i = 1, norb
        xb(nbath*i+1:nbath*(i+1)) = vk(i, :)
    enddo
    xb(nbath*(norb+1)+1:nx) = ef
    call minimization(xb, nwloc, nsite, nx, omega, D_ev, nbath, norb, 0, dab)
    call dos(omr, aff)

    open(11, file=trim(fout), access='stream', form='unformatted', status='replace')
    write(11) y, vout, nsc, a, b, nev_calc
    do i = 1, nev_calc
        write(11) eigval(i)%val, eigval(i)%prob, eigval(i)%sector, eigval(i)%level
    enddo
    write(11) Gr, D_ev, gks, nocc, xb, dab, aff
    close(11)
    call moed_ed_finalize()
end program drop_in_check
