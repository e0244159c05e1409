! Test driver of the Fortran drop-in (dmft-ed-zt_amd/fortran/edgpu_dropin.f90)
! built against the stand-in reference modules of refmods.f90: it sets the
! module globals as dmft_params_read would, then calls the drop-ins with the
! reference's argument lists, as dmft.f90:76-78 calls ground_state and
! cluster_green_ftn.
!
! stdin: U, mu, beta, nw, then ek(1:5) and vmk(1:5) of a spin-symmetric
! single-orbital bath (Nbath = 5, sector (3,3)), then a second bath
! ek(1:5), vmk(1:5).  stdout: lines "key values..." read by
! tests/test_fortran.py; g_coeffs.dat is written in the working directory.
program dropin_driver
    use dmft_params
    use dmft_grid
    use eigpair
    use ed_basis
    use edgpu_dropin
    implicit none
    double precision :: em(1,2), ek(5,2), vmk(1,5,2), ek2(5,2), vmk2(1,5,2), b_(200), a_(200)
    double precision, allocatable :: x(:), y(:), x_all(:)
    double complex, allocatable :: G(:,:,:), G2(:,:,:), G3(:,:,:), Sig(:,:,:)
    double precision :: occ(1,2), dd(1), Z(1,1)
    type(eigpair_t) :: gs, gs2, gs3
    type(basis_t) :: basis
    double precision :: n_up, docc, w
    integer :: i, iup, idn, r, nstep

    read(*,*) U, Mu, beta, nw
    read(*,*) ek(:,1)
    read(*,*) vmk(1,:,1)
    read(*,*) ek2(:,1)
    read(*,*) vmk2(1,:,1)
    ek(:,2) = ek(:,1)
    vmk(1,:,2) = vmk(1,:,1)
    ek2(:,2) = ek2(:,1)
    vmk2(1,:,2) = vmk2(1,:,1)
    em = 0.d0
    norb = 1
    nbath = 5
    nsite = 6
    nspin = 1
    Jex = 0.3d0
    maxnstep = 200
    nsector = 1
    allocate(sectors(1,3))
    sectors(1,1) = 3
    sectors(1,2) = 3
    sectors(1,3) = 400
    nwloc = nw
    allocate(omega(nwloc))
    do i = 1, nwloc
        omega(i) = (2*i - 1) * acos(-1.d0) / beta   ! dmft_grid.f90:42-53
    enddo

    ! ground_state (dmft.f90:76)
    call edgpu_ground_state(em, ek, vmk, gs)
    write(*,'(a,i4,es25.16)') 'E0 ', gs%isector, gs%val
    ! occupancies from gs%vec in the reference's row order (post_process.f90:63-112)
    call generate_basis(3, 3, basis)
    n_up = 0.d0
    docc = 0.d0
    do idn = 1, basis%ndown
        do iup = 1, basis%nup
            r = (idn - 1) * basis%nup + iup
            w = gs%vec(r)**2
            if (btest(basis%up(iup), 0)) n_up = n_up + w
            if (btest(basis%up(iup), 0) .and. btest(basis%down(idn), 0)) docc = docc + w
        enddo
    enddo
    write(*,'(a,es25.16)') 'n_up ', n_up
    write(*,'(a,es25.16)') 'docc ', docc

    ! cluster_green_ftn (dmft.f90:78)
    allocate(G(nwloc,1,1), G2(nwloc,1,1))
    call edgpu_cluster_green_ftn(em, ek, vmk, gs, G)
    do i = 1, nwloc, 97
        write(*,'(a,i6,2es25.16)') 'G ', i, dble(G(i,1,1)), aimag(G(i,1,1))
    enddo

    ! hx and lanczos_iteration on x(r) = cos(0.1 r)
    allocate(x(basis%ntot), y(basis%ntot), x_all(basis%ntot))
    do r = 1, basis%ntot
        x(r) = cos(0.1d0 * r)
    enddo
    y = 0.d0
    call edgpu_hx(em, ek, vmk, basis, x, y, x_all, .false.)
    do r = 1, basis%ntot, 37
        write(*,'(a,i6,es25.16)') 'HX ', r, y(r)
    enddo
    call edgpu_lanczos_iteration(em, ek, vmk, basis, x, nstep, a_, b_)
    write(*,'(a,i6)') 'NSTEP ', nstep
    do i = 1, 8
        write(*,'(a,i4,2es25.16)') 'AB ', i, a_(i), b_(i)
    enddo

    ! a second bath's ground state now held on the device; G of the first
    ! ground state must come from gs%vec, not from the held vector
    call edgpu_ground_state(em, ek2, vmk2, gs2)
    write(*,'(a,es25.16)') 'E0B ', gs2%val
    call edgpu_cluster_green_ftn(em, ek, vmk, gs, G2)
    write(*,'(a,es25.16)') 'GDIFF ', maxval(abs(G2 - G))

    ! the ground state left on the device (edgpu_gs_to_host = .false.):
    ! gs%vec empty, occupancies and self-energy from the device
    edgpu_gs_to_host = .false.
    call edgpu_ground_state(em, ek, vmk, gs3)
    write(*,'(a,i6,es25.16)') 'E0C ', size(gs3%vec), gs3%val
    call edgpu_observables(occ, dd)
    write(*,'(a,3es25.16)') 'OBS ', occ(1,1), occ(1,2), dd(1)
    allocate(G3(nwloc,1,1), Sig(nwloc,1,1))
    call edgpu_cluster_green_ftn(em, ek, vmk, gs3, G3)
    write(*,'(a,es25.16)') 'GDIFFC ', maxval(abs(G3 - G))
    call edgpu_self_energy(em, ek, vmk, G3, Sig, Z)
    do i = 1, nwloc, 97
        write(*,'(a,i6,2es25.16)') 'SIG ', i, dble(Sig(i,1,1)), aimag(Sig(i,1,1))
    enddo
    write(*,'(a,es25.16)') 'Z ', Z(1,1)
    call dealloc_basis(basis)
    call edgpu_finalize()
end program dropin_driver
