! example_solve.f90 -- a Fortran caller of libedgpu.so through iso_c_binding.
! Solves the single-orbital impurity with 5 bath sites (Ns=6), sector (3,3),
! for the bath read from standard input (5 levels, then 5 hybridizations),
! U=2, mu=1, and prints E0, n_up, docc and G(i w_0) on the beta=100 grid.
program example_solve
    use iso_c_binding
    use edgpu_c
    implicit none
    integer, parameter :: nbath = 5, nw = 2048
    type(edg_params) :: p
    type(c_ptr) :: ctx
    real(c_double) :: em(1,2), ek(nbath,2), vmk(1,nbath,2), e0, omega(nw), occ(2), docc(1)
    real(c_double) :: coeffs(200*4*2)
    complex(c_double_complex) :: G(nw)
    integer(c_int) :: sectors(2), isector, rc
    integer :: i
    double precision, parameter :: pi = 3.14159265358979323846d0

    read(*,*) ek(:,1)
    read(*,*) vmk(1,:,1)
    ek(:,2) = ek(:,1)
    vmk(1,:,2) = vmk(1,:,1)
    em = 0.d0
    do i = 1, nw
        omega(i) = (2.d0*(i-1) + 1.d0)*pi/100.d0
    enddo

    call edg_default_params(p)
    p%norb = 1
    p%nbath = nbath
    p%U = 2.d0
    p%J = 0.3d0
    p%mu = 1.d0
    p%maxnstep = 200
    p%seed = 9
    rc = edg_create(p, 0_c_int, ctx)
    if (rc /= EDG_OK) stop 1
    sectors = [3, 3]
    rc = edg_ground_state(ctx, em, ek, vmk, 1_c_int, sectors, isector, e0)
    if (rc /= EDG_OK) then
        print *, trim(edgpu_error_message(ctx))
        stop 2
    endif
    rc = edg_observables(ctx, occ, docc)
    if (rc /= EDG_OK) stop 3
    rc = edg_cluster_green(ctx, em, ek, vmk, int(nw, c_int), omega, G, coeffs)
    if (rc /= EDG_OK) stop 4
    write(*,'(a,es24.16)') 'E0 ', e0
    write(*,'(a,es24.16)') 'n_up ', occ(1)
    write(*,'(a,es24.16)') 'docc ', docc(1)
    write(*,'(a,2es24.16)') 'G0 ', real(G(1)), aimag(G(1))
    rc = edg_destroy(ctx)
end program example_solve
