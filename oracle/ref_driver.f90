! Driver for the reference build (oracle/_ref/ref_driver).
!
! TEST INFRASTRUCTURE ONLY: it links the reference's own modules, compiled from the
! reference sources by oracle/Makefile, and runs one of their routines on inputs
! read from a binary stream file, writing the outputs to another.  It is used in
! the build container to pin the C oracle and to write golden fixtures.
!
! Input file (stream, little endian):
!   int32  task, norb, nbath, ne_up, ne_dn, i1, i2, i3
!   real64 U, J, ek(nsite), vk(norb*nbath) column-major
!   payload (task dependent)
! Tasks:
!   1 multiply_H  (ed_hamiltonian.F90:88)   payload x(D)            -> y(D)
!   2 lanczos     (ed_lanczos.F90:14)       i1=nstep, payload v(D)  -> nstep_calc, a(0:nstep), b(0:nstep)
!   3 apply_c     (ed_operators.F90:10)     i1=pm,i2=iorb,i3=ispin, payload v(D) -> D_out, v_out
!   4 FUNC+dfunc  (func.f90:1,47)           i1=nw, payload x(nx), omega(nw), Dev(2*norb*nw) -> f, g(nx)
!   5 minimization(func.f90:144)            same payload -> fret, x(nx)
!   6 SORT        (func.f90:306)            i1=n, payload a(n) -> indx(n)
!   7 H columns   (multiply_H on unit vectors) -> for each column: nnz, rows(nnz), vals(nnz)
!   8 timing      i1 = calls; every MPI rank owns its row block -> seconds per call, ranks, y(1)
!   9 ed_delta_new (ed_green.F90:187) after ed_set_band_structure (ed_hamiltonian.F90:29), then
!                 n_from_gksum (ed_green.F90:226)
!                 i1 = nw, i2 = Nq; payload rMu, beta, Gr(2*norb*nw), D_ev(2*norb*nw)
!                 -> Hk(i,i,q) real parts (norb*Nq), Gksum(2*norb*nw), D_ev(2*norb*nw), nocc(norb)
!  10 dos_diag     (dos.f90:65)            i1 = nw, i2 = nstep_calcp, i3 = nstep_calcn; payload small, ev,
!                 int32 even, omega_real(nw), ap, bp (0:nsp), an, bn (0:nsn) -> Aff(1, 1:nw) for iorb = 1
!  11 export_eigvec (ed_io.F90:13)          i1 = isector, i2 = iev, i3 = nloc; payload eigvec(nloc)
!                 -> file name (150 characters), size in bytes (int32), the file's bytes
!  12 apbpanbn.dat records as ed_calc_green_ftn writes them (ed_green.F90:75, 112-114): i1 = number
!                 of records; per record payload int32 nsp, nsn, iev, iorb, even; real64 val, prob,
!                 ap, bp (0:nsp), an, bn (0:nsn) -> size in bytes (int32), the file's bytes
!  13 dos (dos.f90:1-63) reading such records: i1 = number of records, i2 = nw, i3 = nev_calc;
!                 payload as task 12, then real64 (small, Nstep), omega_real(nw), and (val, prob) of
!                 the nev_calc states -> Aff(norb, nw); Nstep bounds every record's step counts
!  14 timing of lanczos_iteration, Nstep = i1, rows split over the MPI ranks -> seconds per step,
!                 ranks, a(0)
!  15 timing of one minimization with the frequencies split over the ranks as ed_green_init splits
!                 them; payload as task 5 -> seconds, ranks, fret, x(nx)
program ref_driver
    use parallel_params
    use ed_config
    use ed_basis
    use ed_operators
    use ed_hamiltonian
    use ed_lanczos
    use ed_green, only: nwloc, Gr, D_ev, Gksum, ed_delta_new, n_from_gksum, nlocals_w, offsets_w, nocc
    use ed_io, only: export_eigvec
    use ed_solver, only: eigval, nev_calc
    implicit none
    character(len=512) :: fin, fout
    integer :: hdr(8), task, ne_up, ne_dn, i1, i2, i3, i, j, nnz, nstep_calc, nx, nwl
    double precision :: uj(2), fret, t0, t1
    double precision, allocatable :: x(:), y(:), a(:), b(:), om(:), dv(:), g(:)
    double complex, allocatable :: dev(:,:)
    integer, allocatable :: indx(:), rows(:)
    integer :: ieven, nbytes, rhdr(5)
    logical :: even
    character(len=150) :: fname
    integer(kind=1), allocatable :: bytes(:)
    double precision, allocatable :: ap(:), bp(:), an(:), bn(:), aff(:,:)
    type(basis_t) :: bas, bout
    double precision, external :: func

    call MPI_Init(ierr)
    call MPI_Comm_size(comm, Nodes, ierr)
    call MPI_Comm_rank(comm, Node, ierr)
    call get_command_argument(1, fin)
    call get_command_argument(2, fout)
    open(10, file=trim(fin), access='stream', form='unformatted', status='old')
    read(10) hdr
    task = hdr(1); Norb = hdr(2); Nbath = hdr(3); ne_up = hdr(4); ne_dn = hdr(5)
    i1 = hdr(6); i2 = hdr(7); i3 = hdr(8)
    Nsite = Norb + Nbath
    read(10) uj
    U = uj(1); Jex = uj(2)
    allocate(ek(nsite), vk(norb, norb+1:nsite))
    read(10) ek
    read(10) vk
    if (node .eq. 0) open(11, file=trim(fout), access='stream', form='unformatted', status='replace')

    select case (task)
    case (1)
        bas = generate_basis(ne_up, ne_dn)
        allocate(x(bas%ntot), y(bas%ntot))
        read(10) x
        call multiply_H(bas, x, y)
        write(11) y
    case (2)
        Nstep = i1
        bas = generate_basis(ne_up, ne_dn)
        allocate(x(bas%ntot), a(0:nstep), b(0:nstep))
        read(10) x
        call lanczos_iteration(bas, x, nstep_calc, a, b)
        write(11) nstep_calc
        write(11) a
        write(11) b
    case (3)
        bas = generate_basis(ne_up, ne_dn)
        allocate(x(bas%ntot))
        read(10) x
        call apply_c(bas, x, i1, i2, i3, bout, y)
        write(11) bout%ntot
        write(11) y
    case (4, 5)
        nwl = i1
        nx = Nbath + Nbath*Norb + Norb
        allocate(x(nx), g(nx), om(nwl), dv(2*Norb*nwl), dev(Norb, nwl))
        read(10) x
        read(10) om
        read(10) dv
        do i = 1, nwl
            do j = 1, Norb
                dev(j, i) = dcmplx(dv(2*(j-1+Norb*(i-1))+1), dv(2*(j-1+Norb*(i-1))+2))
            enddo
        enddo
        if (task .eq. 4) then
            fret = func(x, Nsite, nwl, nx, om, dev, Nbath, Norb, comm)
            call dfunc(x, g, nwl, Nsite, nx, om, dev, Nbath, Norb, comm)
            write(11) fret
            write(11) g
        else
            call minimization(x, nwl, Nsite, nx, om, dev, Nbath, Norb, comm, fret)
            write(11) fret
            write(11) x
        endif
    case (6)
        allocate(a(i1), indx(i1))
        read(10) a
        call sort(i1, a, indx)
        write(11) indx
    case (8)
        ! time i1 calls of multiply_H after one untimed call; every rank owns nloc rows of the
        ! sector and gathers the full vector each call, as in the reference run
        bas = generate_basis(ne_up, ne_dn)
        allocate(x(bas%ntot), y(bas%ntot))
        do i = 1, bas%nloc
            x(i) = 1.0d0 / dble(i + bas%offsets(node))
        enddo
        call multiply_H(bas, x, y)
        call mpi_barrier(comm, ierr)
        t0 = mpi_wtime()
        do j = 1, i1
            call multiply_H(bas, x, y)
        enddo
        call mpi_barrier(comm, ierr)
        t1 = mpi_wtime()
        if (node .eq. 0) then
            write(11) (t1 - t0) / dble(i1)
            write(11) nodes
            write(11) y(1)
        endif
    case (7)
        bas = generate_basis(ne_up, ne_dn)
        allocate(x(bas%ntot), y(bas%ntot), rows(bas%ntot))
        x = 0.0d0
        do j = 1, bas%ntot
            x(j) = 1.0d0
            call multiply_H(bas, x, y)
            x(j) = 0.0d0
            nnz = 0
            do i = 1, bas%ntot
                if (y(i) .ne. 0.0d0) then
                    nnz = nnz + 1
                    rows(nnz) = i
                    y(nnz) = y(i)
                endif
            enddo
            write(11) nnz
            write(11) rows(1:nnz)
            write(11) y(1:nnz)
        enddo
    case (9)
        nwloc = i1
        Nq = i2
        read(10) uj
        rMu = uj(1)
        beta = uj(2)
        call ed_hamiltonian_init
        call ed_set_band_structure
        allocate(Gr(norb, nwloc), D_ev(norb, nwloc), Gksum(norb, nwloc))
        read(10) Gr
        read(10) D_ev
        call ed_delta_new
        allocate(nlocals_w(0:0), offsets_w(0:0), nocc(norb))
        nlocals_w(0) = nwloc
        offsets_w(0) = 0
        call n_from_gksum
        write(11) ((dble(Hk(i, i, j)), i = 1, norb), j = 1, Nq)
        write(11) Gksum
        write(11) D_ev
        write(11) nocc
    case (10)
        nw = i1
        read(10) uj
        small = uj(1)
        read(10) ieven
        even = ieven .ne. 0
        allocate(om(nw), ap(0:i2), bp(0:i2), an(0:i3), bn(0:i3), aff(norb, nw))
        read(10) om
        read(10) ap
        read(10) bp
        read(10) an
        read(10) bn
        aff = 0.0d0
        call dos_diag(1, aff, om, 1, uj(2), 1.0d0, even, i2, i3, ap, bp, an, bn)
        write(11) aff(1, :)
    case (11)
        allocate(x(i3))
        read(10) x
        call export_eigvec(i1, i2, 0, i3, 1, x)
        write(fname, "(A7,I2.2,A1,I3.3,A1,I4.4)") "eigvec-", i1, "-", i2, "-", 0
        call copy_file(trim(fname))
    case (12)
        call write_records(i1)
        fname = "apbpanbn.dat"
        call copy_file("apbpanbn.dat")
    case (13)
        call write_records(i1)
        nw = i2
        read(10) uj
        small = uj(1)
        Nstep = nint(uj(2))
        allocate(om(nw), aff(norb, nw))
        read(10) om
        nev_calc = i3
        allocate(eigval(nev_calc))
        do j = 1, nev_calc
            read(10) uj
            eigval(j)%val = uj(1)
            eigval(j)%prob = uj(2)
        enddo
        call dos(om, aff)
        write(11) aff
    case (14)
        ! time lanczos_iteration with Nstep = i1 (i1 + 1 products) on the sector; every rank owns
        ! nloc rows, as in the reference run -> seconds per Lanczos step (per product), ranks, a(0)
        Nstep = i1
        bas = generate_basis(ne_up, ne_dn)
        allocate(x(bas%nloc), a(0:nstep), b(0:nstep))
        do i = 1, bas%nloc
            x(i) = 1.0d0 / dble(i + bas%offsets(node))
        enddo
        call mpi_barrier(comm, ierr)
        t0 = mpi_wtime()
        call lanczos_iteration(bas, x, nstep_calc, a, b)
        call mpi_barrier(comm, ierr)
        t1 = mpi_wtime()
        if (node .eq. 0) then
            write(11) (t1 - t0) / dble(nstep + 1)
            write(11) nodes
            write(11) a(0)
        endif
    case (15)
        ! time one minimization (func.f90:144) with the frequencies split over the ranks as
        ! ed_green_init splits them (ed_green.F90:29-51); payload as task 5 (all nw frequencies)
        ! -> seconds, ranks, fret, x
        nwl = i1
        nx = Nbath + Nbath*Norb + Norb
        allocate(x(nx), g(nx), om(nwl), dv(2*Norb*nwl))
        read(10) x
        read(10) om
        read(10) dv
        nwloc = nwl / nodes
        j = mod(nwl, nodes)
        if (node .lt. j) then
            nwloc = nwloc + 1
            i2 = node * nwloc
        else
            i2 = j + node * nwloc
        endif
        allocate(dev(Norb, nwloc))
        do i = 1, nwloc
            do j = 1, Norb
                dev(j, i) = dcmplx(dv(2*(j-1+Norb*(i2+i-1))+1), dv(2*(j-1+Norb*(i2+i-1))+2))
            enddo
        enddo
        call mpi_barrier(comm, ierr)
        t0 = mpi_wtime()
        call minimization(x, nwloc, Nsite, nx, om(i2+1:i2+nwloc), dev, Nbath, Norb, comm, fret)
        call mpi_barrier(comm, ierr)
        t1 = mpi_wtime()
        if (node .eq. 0) then
            write(11) t1 - t0
            write(11) nodes
            write(11) fret
            write(11) x
        endif
    end select
    if (node .eq. 0) close(11)
    close(10)
    call MPI_Finalize(ierr)
contains
    ! apbpanbn.dat with ed_calc_green_ftn's record statements (ed_green.F90:75, 112-114)
    subroutine write_records(nrec)
        integer, intent(in) :: nrec
        integer :: r
        open(89, file="apbpanbn.dat", form="unformatted", status="new")
        do r = 1, nrec
            read(10) rhdr
            even = rhdr(5) .ne. 0
            allocate(a(2), ap(0:rhdr(1)), bp(0:rhdr(1)), an(0:rhdr(2)), bn(0:rhdr(2)))
            read(10) a
            read(10) ap
            read(10) bp
            read(10) an
            read(10) bn
            write(89) rhdr(1), rhdr(2)
            write(89) rhdr(3), rhdr(4), a(1), a(2), even, (ap(i), i = 0, rhdr(1)), (bp(i), i = 0, rhdr(1)), &
                      (an(i), i = 0, rhdr(2)), (bn(i), i = 0, rhdr(2))
            deallocate(a, ap, bp, an, bn)
        enddo
        close(89)
    end subroutine write_records

    ! the bytes of a file written by the reference's I/O, appended to the output
    subroutine copy_file(name)
        character(len=*), intent(in) :: name
        inquire(file=name, size=nbytes)
        allocate(bytes(nbytes))
        open(12, file=name, access='stream', form='unformatted', status='old')
        read(12) bytes
        close(12)
        write(11) fname
        write(11) nbytes
        write(11) bytes
        deallocate(bytes)
    end subroutine copy_file
end program ref_driver
