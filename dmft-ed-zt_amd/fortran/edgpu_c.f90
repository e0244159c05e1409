! edgpu_c.f90 -- iso_c_binding interfaces to libedgpu.so (include/edgpu.h).
!
! Self-contained: compiles without any other module.  The drop-in wrappers
! with the reference's argument lists live in edgpu_dropin.f90.
module edgpu_c
    use iso_c_binding
    implicit none

    integer(c_int), parameter :: EDG_OK = 0, EDG_ERR_ADMISSION = 2, EDG_ERR_ARG = 3, &
                                 EDG_ERR_SOLVER = 4, EDG_ERR_STATE = 5, EDG_ERR_HIP = -1, &
                                 EDG_ERR_NOMEM = -2
    integer(c_int), parameter :: EDG_FP64 = 0, EDG_FP32 = 1

    ! mirrors struct edg_params field by field
    type, bind(C) :: edg_params
        integer(c_int) :: norb, nbath, nspin
        real(c_double) :: U, J, Up, Jp, mu
        integer(c_int) :: maxnstep, maxnstep_green, diag_solver, precision, both_spins
        integer(c_long_long) :: seed
    end type edg_params

    ! mirrors struct edg_problem (batched solver): ek, vmk are C pointers
    ! (c_loc of the caller's ek(nbath,2), vmk(1,nbath,2))
    type, bind(C) :: edg_problem
        integer(c_int) :: nbath
        real(c_double) :: U, mu
        real(c_double) :: em(2)
        type(c_ptr) :: ek, vmk
        integer(c_int) :: nup, ndn, maxnstep, maxnstep_green, both_spins
        integer(c_long_long) :: seed
        integer(c_int) :: gs_only
    end type edg_problem

    interface
        subroutine edg_default_params(p) bind(C, name="edg_default_params")
            import :: edg_params
            type(edg_params), intent(out) :: p
        end subroutine

        integer(c_int) function edg_create(p, device, ctx) bind(C, name="edg_create")
            import :: c_int, c_ptr, edg_params
            type(edg_params), intent(in) :: p
            integer(c_int), value :: device
            type(c_ptr), intent(out) :: ctx
        end function

        integer(c_int) function edg_destroy(ctx) bind(C, name="edg_destroy")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function

        type(c_ptr) function edg_last_error(ctx) bind(C, name="edg_last_error")
            import :: c_ptr
            type(c_ptr), value :: ctx
        end function

        integer(c_int) function edg_hx(ctx, em, ek, vmk, ne_up, ne_dn, x, y, py) &
                bind(C, name="edg_hx")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(in) :: em(*), ek(*), vmk(*), x(*)
            real(c_double), intent(inout) :: y(*)
            integer(c_int), value :: ne_up, ne_dn, py
        end function

        integer(c_int) function edg_lanczos_iteration(ctx, em, ek, vmk, ne_up, ne_dn, v_init, &
                nstep, a, b) bind(C, name="edg_lanczos_iteration")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(in) :: em(*), ek(*), vmk(*), v_init(*)
            integer(c_int), value :: ne_up, ne_dn
            integer(c_int), intent(out) :: nstep
            real(c_double), intent(out) :: a(*), b(*)
        end function

        integer(c_int) function edg_ground_state(ctx, em, ek, vmk, nsector, sectors, isector, &
                e0) bind(C, name="edg_ground_state")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(in) :: em(*), ek(*), vmk(*)
            integer(c_int), value :: nsector
            integer(c_int), intent(in) :: sectors(*)
            integer(c_int), intent(out) :: isector
            real(c_double), intent(out) :: e0
        end function

        integer(c_int) function edg_gs_info(ctx, ne_up, ne_dn, dim, e0_ritz, e0, residual, &
                nstep) bind(C, name="edg_gs_info")
            import :: c_int, c_ptr, c_double, c_long_long
            type(c_ptr), value :: ctx
            integer(c_int), intent(out) :: ne_up, ne_dn, nstep
            integer(c_long_long), intent(out) :: dim
            real(c_double), intent(out) :: e0_ritz, e0, residual
        end function

        integer(c_int) function edg_get_gs_vector(ctx, vec, dim) bind(C, name="edg_get_gs_vector")
            import :: c_int, c_ptr, c_double, c_long_long
            type(c_ptr), value :: ctx
            real(c_double), intent(out) :: vec(*)
            integer(c_long_long), value :: dim
        end function

        integer(c_int) function edg_set_gs(ctx, ne_up, ne_dn, e0, vec, dim) bind(C, name="edg_set_gs")
            import :: c_int, c_ptr, c_double, c_long_long
            type(c_ptr), value :: ctx
            integer(c_int), value :: ne_up, ne_dn
            real(c_double), value :: e0
            real(c_double), intent(in) :: vec(*)
            integer(c_long_long), value :: dim
        end function

        integer(c_int) function edg_cluster_green(ctx, em, ek, vmk, nw, omega, G_cl, coeffs) &
                bind(C, name="edg_cluster_green")
            import :: c_int, c_ptr, c_double, c_double_complex
            type(c_ptr), value :: ctx
            real(c_double), intent(in) :: em(*), ek(*), vmk(*), omega(*)
            integer(c_int), value :: nw
            complex(c_double_complex), intent(out) :: G_cl(*)
            real(c_double), intent(out) :: coeffs(*)
        end function

        integer(c_int) function edg_observables(ctx, occ, docc) bind(C, name="edg_observables")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(out) :: occ(*), docc(*)
        end function

        integer(c_int) function edg_self_energy(ctx, em, ek, vmk, nw, omega, G_cl, Sigma, Z) &
                bind(C, name="edg_self_energy")
            import :: c_int, c_ptr, c_double, c_double_complex
            type(c_ptr), value :: ctx
            real(c_double), intent(in) :: em(*), ek(*), vmk(*), omega(*)
            integer(c_int), value :: nw
            complex(c_double_complex), intent(in) :: G_cl(*)
            complex(c_double_complex), intent(out) :: Sigma(*)
            real(c_double), intent(out) :: Z(*)
        end function

        integer(c_int) function edg_set_reorthogonalization(ctx, mode) &
                bind(C, name="edg_set_reorthogonalization")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), value :: mode
        end function

        integer(c_int) function edg_apply_c(ctx, ne_up, ne_dn, vec, pm, iorb, ispin, vec_out) &
                bind(C, name="edg_apply_c")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: ne_up, ne_dn, pm, iorb, ispin
            real(c_double), intent(in) :: vec(*)
            real(c_double), intent(out) :: vec_out(*)
        end function

        ! ---- batched solver and request loop (edg_batch_* of include/edgpu.h)
        integer(c_int) function edg_batch_create(device, precision, nspin, nw, omega, budget, b) &
                bind(C, name="edg_batch_create")
            import :: c_int, c_ptr, c_double, c_long_long
            integer(c_int), value :: device, precision, nspin, nw
            real(c_double), intent(in) :: omega(*)
            integer(c_long_long), value :: budget
            type(c_ptr), intent(out) :: b
        end function

        integer(c_int) function edg_batch_destroy(b) bind(C, name="edg_batch_destroy")
            import :: c_int, c_ptr
            type(c_ptr), value :: b
        end function

        integer(c_int) function edg_batch_release_memory(device) bind(C, name="edg_batch_release_memory")
            import :: c_int
            integer(c_int), value :: device
        end function

        integer(c_int) function edg_batch_submit(b, p, id) bind(C, name="edg_batch_submit")
            import :: c_int, c_ptr, c_long_long, edg_problem
            type(c_ptr), value :: b
            type(edg_problem), intent(in) :: p
            integer(c_long_long), intent(out) :: id
        end function

        integer(c_int) function edg_batch_step(b, max_steps, pending) bind(C, name="edg_batch_step")
            import :: c_int, c_ptr
            type(c_ptr), value :: b
            integer(c_int), value :: max_steps
            integer(c_int), intent(out) :: pending
        end function

        integer(c_int) function edg_batch_run(b) bind(C, name="edg_batch_run")
            import :: c_int, c_ptr
            type(c_ptr), value :: b
        end function

        ! returns 1 when a finished problem was popped, 0 when none is ready
        integer(c_int) function edg_batch_poll(b, id, e0, G, nstep, coeffs) bind(C, name="edg_batch_poll")
            import :: c_int, c_ptr, c_long_long, c_double, c_double_complex
            type(c_ptr), value :: b
            integer(c_long_long), intent(out) :: id
            real(c_double), intent(out) :: e0
            complex(c_double_complex), intent(out) :: G(*)
            integer(c_int), intent(out) :: nstep
            real(c_double), intent(out) :: coeffs(*)
        end function
    end interface

contains

    ! Message of the last failing call, as a Fortran string.
    function edgpu_error_message(ctx) result(msg)
        type(c_ptr), intent(in) :: ctx
        character(len=256) :: msg
        character(kind=c_char), pointer :: s(:)
        integer :: i
        msg = ' '
        call c_f_pointer(edg_last_error(ctx), s, [256])
        do i = 1, 256
            if (s(i) == c_null_char) exit
            msg(i:i) = s(i)
        enddo
    end function edgpu_error_message

end module edgpu_c
