! example_batch.f90 -- a Fortran caller of the batched solver through
! iso_c_binding.  Reads nprob single-orbital problems with 5 bath sites
! from standard input (per problem: U, then 5 levels, then 5
! hybridizations; half filling mu = U/2, sector (3,3)), solves them in one
! batch (FP64 vectors, beta = 100 grid) and prints per problem its E0 and
! G(i w_0).
program example_batch
    use iso_c_binding
    use edgpu_c
    implicit none
    integer, parameter :: nbath = 5, nw = 64, nprob = 3
    type(c_ptr) :: b
    type(edg_problem) :: p
    real(c_double), target :: ek(nbath, 2, nprob), vmk(1, nbath, 2, nprob)
    real(c_double) :: omega(nw), uu(nprob), e0, coeffs(80 * 4 * 2)
    complex(c_double_complex) :: G(nw)
    integer(c_long_long) :: id, ids(nprob)
    integer(c_int) :: rc, pending, nstep
    integer :: i, k
    double precision, parameter :: pi = 3.14159265358979323846d0

    do k = 1, nprob
        read(*,*) uu(k)
        read(*,*) ek(:, 1, k)
        read(*,*) vmk(1, :, 1, k)
        ek(:, 2, k) = ek(:, 1, k)
        vmk(1, :, 2, k) = vmk(1, :, 1, k)
    enddo
    do i = 1, nw
        omega(i) = (2.d0*(i-1) + 1.d0)*pi/100.d0
    enddo
    rc = edg_batch_create(0_c_int, EDG_FP64, 1_c_int, int(nw, c_int), omega, 0_c_long_long, b)
    if (rc /= EDG_OK) stop 1
    do k = 1, nprob
        p%nbath = nbath
        p%U = uu(k)
        p%mu = uu(k) / 2.d0
        p%em = 0.d0
        p%ek = c_loc(ek(1, 1, k))
        p%vmk = c_loc(vmk(1, 1, 1, k))
        p%nup = 3
        p%ndn = 3
        p%maxnstep = 80
        p%maxnstep_green = 0
        p%both_spins = 0
        p%seed = 20 + k
        p%gs_only = 0
        rc = edg_batch_submit(b, p, ids(k))
        if (rc /= EDG_OK) stop 2
    enddo
    rc = edg_batch_run(b)
    if (rc /= EDG_OK) stop 3
    do while (edg_batch_poll(b, id, e0, G, nstep, coeffs) == 1)
        write(*,'(a,i3,es24.16,2es24.16)') 'P ', id, e0, real(G(1)), aimag(G(1))
    enddo
    rc = edg_batch_destroy(b)
end program example_batch
