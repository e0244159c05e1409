! edgpu_dropin.f90 -- drop-in replacements, with the reference's argument
! lists, of the ED solver entry points of yw-choi/DMFT-ED-ZT, computed on the
! GPU through libedgpu.so.  Compiled INSIDE the reference's build (it uses
! the reference's modules dmft_params, dmft_grid, ed_basis, eigpair, io_units,
! mpi, utils); see INTEGRATION.md for the Makefile lines.
!
!   edgpu_ground_state      replaces ground_state       (ground_state.f90:1)
!   edgpu_lanczos_iteration replaces lanczos_iteration  (lanczos_solver.f90:92)
!   edgpu_cluster_green_ftn replaces cluster_green_ftn  (dmft_green.f90:47)
!   edgpu_hx                replaces hx                 (impurity_hamiltonian.f90:77)
!
! The context is created lazily from the dmft_params module globals; errors
! abort through die() as the reference does (fdf/utils.F90:411-459).  The
! solver runs on one GPU; with several MPI ranks every rank calls it
! identically and rank-local vector chunks are gathered/scattered here.
module edgpu_dropin
    use iso_c_binding
    use edgpu_c
    implicit none
    private
    public :: edgpu_ground_state, edgpu_lanczos_iteration, edgpu_cluster_green_ftn, edgpu_hx, &
              edgpu_finalize, edgpu_observables, edgpu_self_energy

    type(c_ptr), save :: ctx = c_null_ptr

    ! Options a driver may set before its first call; neither has a
    ! counterpart in the reference.
    ! edgpu_precision: storage of the Lanczos vectors (EDG_FP64, the
    !   reference's, or EDG_FP32; dot products, alpha/beta, the tridiagonal
    !   solve and the continued fraction stay FP64).
    ! edgpu_gs_to_host: .true. (the reference's interface) fills gs%vec with
    !   this rank's chunk of the ground state; .false. leaves the vector on
    !   the device and gs%vec empty (size 0), for a driver that takes the
    !   occupancies and the self-energy from edgpu_observables and
    !   edgpu_self_energy instead of post_process.f90's loops over gs%vec.
    integer(c_int), public, save :: edgpu_precision = EDG_FP64
    logical, public, save :: edgpu_gs_to_host = .true.

contains

    subroutine check(rc, where)
        use utils, only: die
        integer(c_int), intent(in) :: rc
        character(len=*), intent(in) :: where
        if (rc /= EDG_OK) call die(where, trim(edgpu_error_message(ctx)))
    end subroutine check

    subroutine ensure_ctx()
        use dmft_params, only: norb, nbath, nspin, U, Jex, Up, Jp, mu, maxnstep, diag_solver
        use utils, only: die
        type(edg_params) :: p
        if (c_associated(ctx)) return
        call edg_default_params(p)
        p%norb = norb
        p%nbath = nbath
        p%nspin = nspin
        p%U = U
        p%J = Jex
        p%Up = Up
        p%Jp = Jp
        p%mu = mu
        p%maxnstep = maxnstep
        p%diag_solver = diag_solver
        p%precision = edgpu_precision
        if (edg_create(p, 0_c_int, ctx) /= EDG_OK) call die("edgpu", "edg_create failed")
    end subroutine ensure_ctx

    subroutine edgpu_finalize()
        integer(c_int) :: rc
        if (c_associated(ctx)) rc = edg_destroy(ctx)
        ctx = c_null_ptr
    end subroutine edgpu_finalize

    ! ground_state(em, ek, vmk, gs) of ground_state.f90:1-81
    subroutine edgpu_ground_state(em, ek, vmk, gs)
        use mpi
        use dmft_params, only: norb, nbath, nsector, sectors
        use eigpair, only: eigpair_t
        use ed_basis, only: basis_t, generate_basis, dealloc_basis
        double precision, intent(in) :: em(norb,2), ek(nbath,2), vmk(norb,nbath,2)
        type(eigpair_t), intent(out) :: gs
        integer(c_int) :: sec(2*nsector), isector, nu, nd, ns
        integer(c_long_long) :: dim
        real(c_double) :: e0, er, res
        double precision, allocatable :: full(:)
        type(basis_t) :: basis
        integer :: i
        call ensure_ctx()
        do i = 1, nsector
            sec(2*i-1) = sectors(i,1)
            sec(2*i) = sectors(i,2)
        enddo
        call check(edg_ground_state(ctx, em, ek, vmk, int(nsector, c_int), sec, isector, e0), &
                   "ground_state")
        call check(edg_gs_info(ctx, nu, nd, dim, er, e0, res, ns), "ground_state")
        gs%isector = isector
        gs%val = e0
        if (.not. edgpu_gs_to_host) then
            allocate(gs%vec(0))
            return
        endif
        allocate(full(dim))
        call check(edg_get_gs_vector(ctx, full, dim), "ground_state")
        call generate_basis(int(nu), int(nd), basis)
        allocate(gs%vec(basis%nloc))
        gs%vec = full(basis%offsets(taskid)+1 : basis%offsets(taskid)+basis%nloc)
        call dealloc_basis(basis)
        deallocate(full)
    end subroutine edgpu_ground_state

    ! lanczos_iteration(em, ek, vmk, basis, v_init, nstep, a, b) of
    ! lanczos_solver.f90:92-173 (v_init is this rank's chunk)
    subroutine edgpu_lanczos_iteration(em, ek, vmk, basis, v_init, nstep, a, b)
        use mpi
        use dmft_params, only: norb, nbath, maxnstep
        use ed_basis, only: basis_t
        double precision, intent(in) :: em(norb,2), ek(nbath,2), vmk(norb,nbath,2)
        type(basis_t), intent(in) :: basis
        double precision, intent(in) :: v_init(basis%nloc)
        integer, intent(out) :: nstep
        double precision, intent(out) :: a(maxnstep), b(maxnstep)
        double precision, allocatable :: full(:)
        integer(c_int) :: ns
        call ensure_ctx()
        allocate(full(basis%ntot))
        call mpi_allgatherv(v_init, basis%nloc, mpi_double_precision, full, basis%nlocals, &
                            basis%offsets, mpi_double_precision, comm, mpierr)
        call check(edg_lanczos_iteration(ctx, em, ek, vmk, int(basis%ne_up, c_int), &
                   int(basis%ne_down, c_int), full, ns, a, b), "lanczos_iteration")
        nstep = ns
        deallocate(full)
    end subroutine edgpu_lanczos_iteration

    ! cluster_green_ftn(em, ek, vmk, gs, G_cl) of dmft_green.f90:47-126.
    ! The ground state computed by edgpu_ground_state is still on the device;
    ! when the held one is not gs (another sector or energy, or none), gs%vec
    ! is gathered and installed first.  The master rank writes g_coeffs.dat
    ! as the reference does (:66-69,89-109): a header "maxnstep norb", then
    ! per orbital and spin a line "iorb ispin E0" and maxnstep rows
    ! "ap bp am bm" (spin-1 coefficients repeated for spin 2 when the
    ! `even` shortcut skips its chains, :76-80,102-109).
    subroutine edgpu_cluster_green_ftn(em, ek, vmk, gs, G_cl)
        use mpi
        use utils, only: die
        use io_units, only: IO_G_COEFFS, FN_G_COEFFS
        use dmft_params, only: norb, nbath, nspin, maxnstep, sectors
        use dmft_grid, only: nwloc, omega
        use eigpair, only: eigpair_t
        use ed_basis, only: basis_t, generate_basis, dealloc_basis
        double precision, intent(in) :: em(norb,2), ek(nbath,2), vmk(norb,nbath,2)
        type(eigpair_t), intent(in) :: gs
        double complex, intent(out) :: G_cl(nwloc,norb,nspin)
        double precision, allocatable :: coeffs(:,:,:,:), full(:)
        integer(c_int) :: nu, nd, ns, rc
        integer(c_long_long) :: dim
        real(c_double) :: e0, er, res
        type(basis_t) :: basis
        integer :: iorb, ispin, i
        call ensure_ctx()
        rc = edg_gs_info(ctx, nu, nd, dim, er, e0, res, ns)
        if (rc /= EDG_OK .or. nu /= sectors(gs%isector,1) .or. nd /= sectors(gs%isector,2) .or. &
            e0 /= gs%val) then
            if (size(gs%vec) == 0) call die("cluster_green_ftn", &
                "gs has no vector (edgpu_gs_to_host = .false.) and is not the ground state held on the device")
            call generate_basis(sectors(gs%isector,1), sectors(gs%isector,2), basis)
            allocate(full(basis%ntot))
            call mpi_allgatherv(gs%vec, basis%nloc, mpi_double_precision, full, basis%nlocals, &
                                basis%offsets, mpi_double_precision, comm, mpierr)
            call check(edg_set_gs(ctx, int(basis%ne_up, c_int), int(basis%ne_down, c_int), gs%val, full, &
                                  int(basis%ntot, c_long_long)), "cluster_green_ftn")
            deallocate(full)
            call dealloc_basis(basis)
        endif
        allocate(coeffs(maxnstep,4,2,norb))
        call check(edg_cluster_green(ctx, em, ek, vmk, int(nwloc, c_int), omega, G_cl, coeffs), &
                   "cluster_green_ftn")
        if (master) then
            open(unit=IO_G_COEFFS, file=FN_G_COEFFS, status="replace")
            write(IO_G_COEFFS,*) maxnstep, norb
            do iorb = 1, norb
                do ispin = 1, 2
                    write(IO_G_COEFFS,"(2I4,1x,ES24.16)") iorb, ispin, gs%val
                    do i = 1, maxnstep
                        write(IO_G_COEFFS,"(4(ES24.16,1x))") coeffs(i,1,ispin,iorb), coeffs(i,2,ispin,iorb), &
                            coeffs(i,3,ispin,iorb), coeffs(i,4,ispin,iorb)
                    enddo
                enddo
            enddo
            close(IO_G_COEFFS)
        endif
        deallocate(coeffs)
    end subroutine edgpu_cluster_green_ftn

    ! Occupancies occ(iorb,ispin) and double occupancies docc(iorb) of the
    ! ground state held on the device: the loops of post_process.f90:63-112
    ! over gs%vec, with the vector left on the device.
    subroutine edgpu_observables(occ, docc)
        use dmft_params, only: norb
        double precision, intent(out) :: occ(norb,2), docc(norb)
        call ensure_ctx()
        call check(edg_observables(ctx, occ, docc), "observables")
    end subroutine edgpu_observables

    ! Self-energy Sigma(iw,iorb,ispin) = i w + mu - em - Delta - 1/G_cl and the
    ! quasiparticle weight Z(iorb,ispin) of post_process.f90:40-58,121-185 on
    ! the device (every orbital's Sigma kept).
    subroutine edgpu_self_energy(em, ek, vmk, G_cl, Sigma, Z)
        use dmft_params, only: norb, nbath, nspin
        use dmft_grid, only: nwloc, omega
        double precision, intent(in) :: em(norb,2), ek(nbath,2), vmk(norb,nbath,2)
        double complex, intent(in) :: G_cl(nwloc,norb,nspin)
        double complex, intent(out) :: Sigma(nwloc,norb,nspin)
        double precision, intent(out) :: Z(norb,nspin)
        call ensure_ctx()
        call check(edg_self_energy(ctx, em, ek, vmk, int(nwloc, c_int), omega, G_cl, Sigma, Z), &
                   "self_energy")
    end subroutine edgpu_self_energy

    ! hx(em, ek, vmk, basis, x, y, x_all, py) of impurity_hamiltonian.f90:77-247
    subroutine edgpu_hx(em, ek, vmk, basis, x, y, x_all, py)
        use mpi
        use dmft_params, only: norb, nbath
        use ed_basis, only: basis_t
        type(basis_t), intent(in) :: basis
        double precision, intent(in) :: x(basis%nloc), em(norb,2), ek(nbath,2), vmk(norb,nbath,2)
        logical, intent(in) :: py
        double precision, intent(inout) :: y(basis%nloc)
        double precision, intent(out) :: x_all(basis%ntot)
        double precision, allocatable :: y_all(:)
        integer(c_int) :: ipy
        call ensure_ctx()
        call mpi_allgatherv(x, basis%nloc, mpi_double_precision, x_all, basis%nlocals, &
                            basis%offsets, mpi_double_precision, comm, mpierr)
        allocate(y_all(basis%ntot))
        y_all = 0.d0
        ipy = 0
        if (py) then
            y_all(basis%offsets(taskid)+1 : basis%offsets(taskid)+basis%nloc) = y
            ipy = 1
        endif
        call check(edg_hx(ctx, em, ek, vmk, int(basis%ne_up, c_int), int(basis%ne_down, c_int), &
                          x_all, y_all, ipy), "hx")
        y = y_all(basis%offsets(taskid)+1 : basis%offsets(taskid)+basis%nloc)
        deallocate(y_all)
    end subroutine edgpu_hx

end module edgpu_dropin
