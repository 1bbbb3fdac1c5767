! spring_gpu_m -- Fortran side of the libspring_gpu.so boundary (iso_c_binding).
!
! The derived types mirror the C structs of include/spring_gpu.h byte for byte; the
! wrapper routines keep the reference's routine names and argument meaning for the hot
! path and its failure behaviour (print the message, ERROR STOP):
!   ScatterProbAllq1_gpu   <- ThirdOrder::ScatterProbAllq1 (src/Kappa/Src/ThirdOrder_m.f90:41-166),
!                             all q1 of one q0; fills Sctrq1(Ndof,Ndof,Ndof,Nq)
!   CalculateLinewidth_gpu <- ThirdOrder::CalculateLinewidth (ThirdOrder_m.f90:442-543) with the
!                             scattering matrix computed on the fly
!   LinewidthAllq0_gpu     <- the q0 loop of src/Kappa/Src/main.f90:552-715 (three-phonon part)
!   FourPhononLw_gpu       <- ScatterProbq1q2_per + TetrahedronIntegration4th (Gaussian form),
!                             FourthOrder_m.f90:42-283, Tetrahedron4th_sep.f90:37-171
!   kappa_noIter_gpu       <- kappa_noIter / kappa_noIter_symm (ThermalConductivity.f90:36-293)
!   TetrahedronW_gpu, TetrahedronIter_gpu, kappa_iter_gpu
!                          <- the iterative solver (src/Kappa/Src_iterative): 1/tau, one F update,
!                             the self-consistent loop of main.f90:243-361
!   get_phonon_gpu         <- DynaMat::get_phonon (src/Kappa/Src/DynaMat_mod.f90:486-709), non-polar:
!                             Evec, freq, OnebyFreq, nBE, grp_vel of every q, freq_min/freq_max
!   PolarizationStrtEnd    <- Helper::PolarizationStrtEnd (src/Kappa/Src/Helper_m.f90:310-353); the
!                             Polarization3rd/4th it returns are the optional last arguments of
!                             the linewidth and kappa wrappers, as of the routines they replace
!   CalculateLwOffMesh_gpu <- LineWidth.x's ScatterProbAllq1 + CalculateLwTetrahedron
!                             (src/LineWidth/Src/Linewidth_m.f90:48-417) for a q0 off the mesh
!   MakeMeshIrrQ_gpu       <- q_tetrahedron%make_tetrahedron + FindIrrQ (TetrahedronQ_m.f90:56-331)
!                             and the rotations of set_SpaceGrSPGlib, host code of the library
! Index arrays arrive 1-based as in the reference (atmIndx(4,:,:), pos, the irreducible
! grid index) and are shifted to the C ABI's 0-based convention here.
module spring_gpu_m

    use iso_c_binding
    implicit none
    private

    integer, parameter, public          :: dp = c_double
    integer(c_int), parameter, public   :: SPR_PREC_FP64 = 0, SPR_PREC_FP32 = 1, SPR_PREC_BF16 = 2

    type, bind(C) :: spr_crystal
        integer(c_int)  :: nb
        type(c_ptr)     :: mass
        real(c_double)  :: vol
    end type spr_crystal

    type, bind(C) :: spr_ifc
        integer(c_int)  :: order, natm
        type(c_ptr)     :: atm_num, atm_indx, fc
    end type spr_ifc

    type, bind(C) :: spr_mesh
        integer(c_int)  :: mesh(3), nq
        type(c_ptr)     :: q_int
        integer(c_int)  :: nq0
        type(c_ptr)     :: irr_grid, pos, multiplicity
        integer(c_int)  :: nsym
        type(c_ptr)     :: rot
        type(c_ptr)     :: tetra
    end type spr_mesh

    type, bind(C) :: spr_phonons
        integer(c_int)  :: nq, ndof
        type(c_ptr)     :: evec, omega, inv_omega, nbe, grp_vel
    end type spr_phonons

    type, public :: gpu_ctx
        type(c_ptr)     :: h = c_null_ptr
        integer         :: Ndof = 0, Nq = 0, Nq0 = 0
    end type gpu_ctx

    interface
        integer(c_int) function spr_ctx_create(cr, fc3, fc4, mesh, ph, device, out) &
                & bind(C, name="spr_ctx_create")
            import :: c_int, c_ptr, spr_crystal, spr_ifc, spr_mesh, spr_phonons
            type(spr_crystal), intent(in)   :: cr
            type(spr_ifc), intent(in)       :: fc3
            type(c_ptr), value              :: fc4
            type(spr_mesh), intent(in)      :: mesh
            type(spr_phonons), intent(in)   :: ph
            integer(c_int), value           :: device
            type(c_ptr), intent(out)        :: out
        end function spr_ctx_create

        integer(c_int) function spr_ctx_destroy(h) bind(C, name="spr_ctx_destroy")
            import :: c_int, c_ptr
            type(c_ptr), value :: h
        end function spr_ctx_destroy

        type(c_ptr) function spr_last_error() bind(C, name="spr_last_error")
            import :: c_ptr
        end function spr_last_error

        integer(c_int) function spr_v3_w(h, i0, W) bind(C, name="spr_v3_w")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value      :: h
            integer(c_int), value   :: i0
            real(c_double)          :: W(*)
        end function spr_v3_w

        integer(c_int) function spr_lw3(h, i0, sigma, prec, lw) bind(C, name="spr_lw3")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value      :: h
            integer(c_int), value   :: i0, prec
            real(c_double), value   :: sigma
            real(c_double)          :: lw(*)
        end function spr_lw3

        integer(c_int) function spr_lw3_all(h, sigma, prec, lw) bind(C, name="spr_lw3_all")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value      :: h
            integer(c_int), value   :: prec
            real(c_double), value   :: sigma
            real(c_double)          :: lw(*)
        end function spr_lw3_all

        integer(c_int) function spr_lw3_tetra(h, i0, lw, p3) bind(C, name="spr_lw3_tetra")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value      :: h
            integer(c_int), value   :: i0
            real(c_double)          :: lw(*), p3(*)
        end function spr_lw3_tetra

        integer(c_int) function spr_lw3_tetra_all(h, lw, p3) bind(C, name="spr_lw3_tetra_all")
            import :: c_int, c_ptr
            import :: c_double
            type(c_ptr), value      :: h
            real(c_double)          :: lw(*), p3(*)
        end function spr_lw3_tetra_all

        integer(c_int) function spr_lw4_tetra(h, i0, lw) bind(C, name="spr_lw4_tetra")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value      :: h
            integer(c_int), value   :: i0
            real(c_double)          :: lw(*)
        end function spr_lw4_tetra

        integer(c_int) function spr_lw4_tetra_all(h, lw) bind(C, name="spr_lw4_tetra_all")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value      :: h
            real(c_double)          :: lw(*)
        end function spr_lw4_tetra_all

        integer(c_int) function spr_lw4_all(h, sigma, prec, lw) bind(C, name="spr_lw4_all")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value      :: h
            integer(c_int), value   :: prec
            real(c_double), value   :: sigma
            real(c_double)          :: lw(*)
        end function spr_lw4_all

        integer(c_int) function spr_kappa_rta(h, lw, T, symm, kappa, mode_k) bind(C, name="spr_kappa_rta")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value      :: h
            real(c_double)          :: lw(*)
            real(c_double), value   :: T
            integer(c_int), value   :: symm
            real(c_double)          :: kappa(9), mode_k(*)
        end function spr_kappa_rta

        integer(c_int) function spr_bte_irr(h, nirr, irr_grid, deg) bind(C, name="spr_bte_irr")
            import :: c_int, c_ptr
            type(c_ptr), value      :: h
            integer(c_int)          :: nirr
            type(c_ptr), value      :: irr_grid, deg
        end function spr_bte_irr

        integer(c_int) function spr_bte_tau_inv(h, inv_tau) bind(C, name="spr_bte_tau_inv")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value      :: h
            real(c_double)          :: inv_tau(*)
        end function spr_bte_tau_inv

        integer(c_int) function spr_bte_delta_f(h, F, dF) bind(C, name="spr_bte_delta_f")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value      :: h
            real(c_double)          :: F(*), dF(*)
        end function spr_bte_delta_f

        integer(c_int) function spr_bte_solve(h, T, eps_k, max_iter, kappa, kappa0, iterations, F) &
                & bind(C, name="spr_bte_solve")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value      :: h
            real(c_double), value   :: T, eps_k
            integer(c_int), value   :: max_iter
            real(c_double)          :: kappa(9), kappa0(9)
            integer(c_int)          :: iterations
            type(c_ptr), value      :: F
        end function spr_bte_solve

        integer(c_int) function spr_get_phonon(cr, latvec, fc2, nq, q_cart, is_gamma, T, device, evec, &
                & omega, inv_omega, nbe, grp_vel) bind(C, name="spr_get_phonon")
            import :: c_int, c_double, spr_crystal, spr_ifc
            type(spr_crystal), intent(in)   :: cr
            real(c_double), intent(in)      :: latvec(9)
            type(spr_ifc), intent(in)       :: fc2
            integer(c_int), value           :: nq
            real(c_double), intent(in)      :: q_cart(*)
            integer(c_int), intent(in)      :: is_gamma(*)
            real(c_double), value           :: T
            integer(c_int), value           :: device
            complex(c_double)               :: evec(*)
            real(c_double)                  :: omega(*), inv_omega(*), nbe(*), grp_vel(*)
        end function spr_get_phonon

        integer(c_int) function spr_set_branch_rules(h, order, pol) bind(C, name="spr_set_branch_rules")
            import :: c_int, c_ptr
            type(c_ptr), value      :: h
            integer(c_int), value   :: order
            integer(c_int)          :: pol(*)
        end function spr_set_branch_rules

        integer(c_int) function spr_kappa_rta_spectral(h, lw, T, symm, kappa, mode_k, spectral) &
                & bind(C, name="spr_kappa_rta_spectral")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value      :: h
            real(c_double)          :: lw(*)
            real(c_double), value   :: T
            integer(c_int), value   :: symm
            real(c_double)          :: kappa(9), mode_k(*), spectral(*)
        end function spr_kappa_rta_spectral

        integer(c_int) function spr_lw3_offmesh(h, q0, evec0, omega0, inv_omega0, cr, latvec, fc2, T, &
                & zero_centered, delta, sigma, lw, W) bind(C, name="spr_lw3_offmesh")
            import :: c_int, c_ptr, c_double, spr_crystal, spr_ifc
            type(c_ptr), value              :: h
            real(c_double), intent(in)      :: q0(3), omega0(*), inv_omega0(*), latvec(9)
            complex(c_double), intent(in)   :: evec0(*)
            type(spr_crystal), intent(in)   :: cr
            type(spr_ifc), intent(in)       :: fc2
            real(c_double), value           :: T, sigma
            integer(c_int), value           :: zero_centered, delta
            real(c_double)                  :: lw(*)
            type(c_ptr), value              :: W
        end function spr_lw3_offmesh

        integer(c_int) function spr_mesh_build(latvec, nb, frac, typ, mesh, time_rev, symmetry, symprec, &
                & nq0, nsym, q_int, q_cart, irr_grid, pos, multiplicity, rot, tetra) bind(C, name="spr_mesh_build")
            import :: c_int, c_ptr, c_double
            real(c_double), intent(in)      :: latvec(9), frac(*)
            integer(c_int), value           :: nb, time_rev, symmetry
            integer(c_int), intent(in)      :: typ(*), mesh(3)
            real(c_double), value           :: symprec
            integer(c_int)                  :: nq0, nsym
            type(c_ptr), value              :: q_int, q_cart, irr_grid, pos, multiplicity, rot, tetra
        end function spr_mesh_build

        integer(c_size_t) function c_strlen(s) bind(C, name="strlen")
            import :: c_size_t, c_ptr
            type(c_ptr), value :: s
        end function c_strlen
    end interface

    public :: gpu_ctx_create, gpu_ctx_destroy, ScatterProbAllq1_gpu, CalculateLinewidth_gpu, &
            & LinewidthAllq0_gpu, FourPhononLw_gpu, kappa_noIter_gpu, spr_error_text, &
            & CalculateLwTetrahedron_gpu, TetrahedronIntegration4th_gpu, FourPhononLwTetra_gpu, IrrQ_tr_gpu, TetrahedronW_gpu, TetrahedronIter_gpu, &
            & kappa_iter_gpu, get_phonon_gpu, PolarizationStrtEnd, CalculateLwOffMesh_gpu, MakeMeshIrrQ_gpu

contains

    function spr_error_text() result(msg)
        character(len=:), allocatable   :: msg
        type(c_ptr)                     :: p
        character(kind=c_char), pointer :: chars(:)
        integer                         :: n, i
        p = spr_last_error()
        n = int(c_strlen(p))
        call c_f_pointer(p, chars, [n])
        allocate(character(len=n) :: msg)
        do i = 1, n
            msg(i:i) = chars(i)
        end do
    end function spr_error_text

    subroutine check(rc, what)
        integer(c_int), intent(in)      :: rc
        character(len=*), intent(in)    :: what
        if ( rc /= 0 ) then
            write(*, '(A, A, I0, A, A)') what, ' failed (', rc, '): ', spr_error_text()
            ERROR STOP
        end if
    end subroutine check

    ! Upload everything once.  atmIndx(4,:,:) and pos/irr_grid are 1-based (reference layout).
    subroutine gpu_ctx_create(ctx, Nbasis, mass, vol, atmNum3, atmIndx3, fC3, &
                            & mesh, q_pnt_int, irr_grid, pos, multiplicity, Rxyz, &
                            & Evec, omega, OnebyOmega, nBE, grp_vel, device, &
                            & atmNum4, atmIndx4, fC4, tetrahedrons)
        type(gpu_ctx), intent(out)                          :: ctx
        integer, intent(in)                                 :: Nbasis, device
        real(dp), intent(in), target                        :: mass(:), vol
        integer, intent(in)                                 :: atmNum3(:), atmIndx3(:,:,:)
        real(dp), intent(in), target                        :: fC3(:,:,:,:,:,:)
        integer, intent(in)                                 :: mesh(3), q_pnt_int(:,:), irr_grid(:), &
                                                             & pos(:), multiplicity(:)
        real(dp), intent(in), target                        :: Rxyz(:,:,:)
        complex(dp), intent(in), target                     :: Evec(:,:,:)
        real(dp), intent(in), target                        :: omega(:,:), OnebyOmega(:,:), nBE(:,:), &
                                                             & grp_vel(:,:,:)
        integer, intent(in), optional                       :: atmNum4(:), atmIndx4(:,:,:)
        real(dp), intent(in), target, optional              :: fC4(:,:,:,:,:,:,:,:)
        integer, intent(in), optional                       :: tetrahedrons(:,:)   ! (4, 6 Nq), 1-based

        integer(c_int), allocatable, target, save   :: num3(:), indx3(:,:,:), num4(:), indx4(:,:,:)
        integer(c_int), allocatable, target, save   :: qint(:,:), irr0(:), pos0(:), mult(:), tet0(:,:)
        real(dp), allocatable, target, save         :: mass_c(:)
        type(spr_crystal)                           :: cr
        type(spr_ifc), target                       :: f3, f4
        type(spr_mesh)                              :: ms
        type(spr_phonons)                           :: ph
        type(c_ptr)                                 :: p4
        integer(c_int)                              :: rc

        num3 = atmNum3
        indx3 = atmIndx3
        indx3(4, :, :) = indx3(4, :, :) - 1
        mass_c = mass
        cr = spr_crystal(int(Nbasis, c_int), c_loc(mass_c), vol)
        f3 = spr_ifc(3, int(size(atmIndx3, 2), c_int), c_loc(num3), c_loc(indx3), c_loc(fC3))
        p4 = c_null_ptr
        if ( present(fC4) ) then
            num4 = atmNum4
            indx4 = atmIndx4
            indx4(4, :, :) = indx4(4, :, :) - 1
            f4 = spr_ifc(4, int(size(atmIndx4, 2), c_int), c_loc(num4), c_loc(indx4), c_loc(fC4))
            p4 = c_loc(f4)
        end if
        qint = q_pnt_int
        irr0 = irr_grid - 1
        pos0 = pos - 1
        mult = multiplicity
        ms = spr_mesh(int(mesh, c_int), int(size(q_pnt_int, 2), c_int), c_loc(qint), &
                    & int(size(irr_grid), c_int), c_loc(irr0), c_loc(pos0), c_loc(mult), &
                    & int(size(Rxyz, 3), c_int), c_loc(Rxyz), c_null_ptr)
        if ( present(tetrahedrons) ) then
            tet0 = tetrahedrons - 1
            ms%tetra = c_loc(tet0)
        end if
        ph = spr_phonons(int(size(omega, 2), c_int), int(size(omega, 1), c_int), c_loc(Evec), &
                       & c_loc(omega), c_loc(OnebyOmega), c_loc(nBE), c_loc(grp_vel))
        rc = spr_ctx_create(cr, f3, p4, ms, ph, int(device, c_int), ctx%h)
        call check(rc, 'spr_ctx_create')
        ctx%Ndof = size(omega, 1)
        ctx%Nq = size(omega, 2)
        ctx%Nq0 = size(irr_grid)
    end subroutine gpu_ctx_create

    subroutine gpu_ctx_destroy(ctx)
        type(gpu_ctx), intent(inout) :: ctx
        integer(c_int) :: rc
        rc = spr_ctx_destroy(ctx%h)
        ctx%h = c_null_ptr
    end subroutine gpu_ctx_destroy

    ! Sctrq1(s2,s1,s0,q1) for irreducible q0 number i0 (1-based), every q1.
    subroutine ScatterProbAllq1_gpu(ctx, i0, Sctrq1)
        type(gpu_ctx), intent(in)   :: ctx
        integer, intent(in)         :: i0
        real(dp), intent(out)       :: Sctrq1(:,:,:,:)
        call check(spr_v3_w(ctx%h, int(i0 - 1, c_int), Sctrq1), 'ScatterProbAllq1_gpu')
    end subroutine ScatterProbAllq1_gpu

    ! Gaussian-delta three-phonon linewidth lw(Ndof) of irreducible q0 number i0 (1-based).
    subroutine CalculateLinewidth_gpu(ctx, i0, Sigma, prec, lw, Polarization3rd)
        type(gpu_ctx), intent(in)   :: ctx
        integer, intent(in)         :: i0, prec
        real(dp), intent(in)        :: Sigma
        real(dp), intent(out)       :: lw(:)
        integer, intent(in), optional :: Polarization3rd(2, 3)
        if ( present(Polarization3rd) ) call set_rules(ctx, 3, Polarization3rd)
        call check(spr_lw3(ctx%h, int(i0 - 1, c_int), Sigma, int(prec, c_int), lw), 'CalculateLinewidth_gpu')
    end subroutine CalculateLinewidth_gpu

    ! LW_Allq0(Ndof, Nq0) of every irreducible q0 in one launch.
    subroutine LinewidthAllq0_gpu(ctx, Sigma, prec, LW_Allq0, Polarization3rd)
        type(gpu_ctx), intent(in)   :: ctx
        integer, intent(in)         :: prec
        real(dp), intent(in)        :: Sigma
        real(dp), intent(out)       :: LW_Allq0(:,:)
        integer, intent(in), optional :: Polarization3rd(2, 3)
        if ( present(Polarization3rd) ) call set_rules(ctx, 3, Polarization3rd)
        call check(spr_lw3_all(ctx%h, Sigma, int(prec, c_int), LW_Allq0), 'LinewidthAllq0_gpu')
    end subroutine LinewidthAllq0_gpu

    ! Tetrahedron-method linewidth lw(Ndof) and P3_q0(Ndof) of irreducible q0 number i0 (1-based),
    ! the outputs of CalculateLwTetrahedron (ThirdOrder_m.f90:169-188); P3 = sum(P3_q0).
    ! Needs the tetrahedra at gpu_ctx_create.  Polarization3rd (from PolarizationStrtEnd) restricts
    ! the branch loops as the reference's argument of the same name (ThirdOrder_m.f90:207-214).
    subroutine CalculateLwTetrahedron_gpu(ctx, i0, lw, P3_q0, P3, Polarization3rd)
        type(gpu_ctx), intent(in)   :: ctx
        integer, intent(in)         :: i0
        real(dp), intent(out)       :: lw(:), P3_q0(:), P3
        integer, intent(in), optional :: Polarization3rd(2, 3)
        if ( present(Polarization3rd) ) call set_rules(ctx, 3, Polarization3rd)
        call check(spr_lw3_tetra(ctx%h, int(i0 - 1, c_int), lw, P3_q0), 'CalculateLwTetrahedron_gpu')
        P3 = sum(P3_q0)
    end subroutine CalculateLwTetrahedron_gpu

    ! Four-phonon tetrahedron-method sum lw(Ndof) of irreducible q0 number i0 (1-based), the output
    ! of TetrahedronIntegration4th (Tetrahedron4th_sep.f90:37-170) summed over every q1 (the
    ! reference sums the images' parts); the caller adds 0.5*lw (main.f90:660).  FP64.  Needs FC4
    ! and the tetrahedra at gpu_ctx_create.
    subroutine TetrahedronIntegration4th_gpu(ctx, i0, lw, Polarization4th)
        type(gpu_ctx), intent(in)   :: ctx
        integer, intent(in)         :: i0
        real(dp), intent(out)       :: lw(:)
        integer, intent(in), optional :: Polarization4th(2, 4)
        if ( present(Polarization4th) ) call set_rules(ctx, 4, Polarization4th)
        call check(spr_lw4_tetra(ctx%h, int(i0 - 1, c_int), lw), 'TetrahedronIntegration4th_gpu')
    end subroutine TetrahedronIntegration4th_gpu

    ! lw4(Ndof, Nq0) of every irreducible q0 with the tetrahedron method (as above).
    subroutine FourPhononLwTetra_gpu(ctx, lw4, Polarization4th)
        type(gpu_ctx), intent(in)   :: ctx
        real(dp), intent(out)       :: lw4(:,:)
        integer, intent(in), optional :: Polarization4th(2, 4)
        if ( present(Polarization4th) ) call set_rules(ctx, 4, Polarization4th)
        call check(spr_lw4_tetra_all(ctx%h, lw4), 'FourPhononLwTetra_gpu')
    end subroutine FourPhononLwTetra_gpu

    ! Raw four-phonon sum lw4(Ndof, Nq0); the caller adds 0.5*lw4 (main.f90:660).
    subroutine FourPhononLw_gpu(ctx, Sigma, prec, lw4, Polarization4th)
        type(gpu_ctx), intent(in)   :: ctx
        integer, intent(in)         :: prec
        real(dp), intent(in)        :: Sigma
        real(dp), intent(out)       :: lw4(:,:)
        integer, intent(in), optional :: Polarization4th(2, 4)
        if ( present(Polarization4th) ) call set_rules(ctx, 4, Polarization4th)
        call check(spr_lw4_all(ctx%h, Sigma, int(prec, c_int), lw4), 'FourPhononLw_gpu')
    end subroutine FourPhononLw_gpu

    ! kappa(9) in W/mK (column-major 3x3) and Mode_k(9, Ndof); Spectral_k(11, Ndof*Nq0) when
    ! present (-Spectral_k: columns (s, omega, nine terms) as main.f90:780-790 completes them);
    ! Polarization3rd / Polarization4th restrict the mode range as in kappa_noIter's arguments.
    subroutine kappa_noIter_gpu(ctx, LW_Allq0, T, Symm, kappa, Mode_k, Spectral_k, Polarization3rd, Polarization4th)
        type(gpu_ctx), intent(in)   :: ctx
        real(dp), intent(in)        :: LW_Allq0(:,:), T
        logical, intent(in)         :: Symm
        real(dp), intent(out)       :: kappa(9), Mode_k(:,:)
        real(dp), intent(out), optional :: Spectral_k(:,:)
        integer, intent(in), optional :: Polarization3rd(2, 3), Polarization4th(2, 4)
        integer(c_int)              :: s
        real(dp), allocatable       :: lw(:,:)
        s = 0
        if ( Symm ) s = 1
        lw = LW_Allq0
        if ( present(Polarization3rd) ) call set_rules(ctx, 3, Polarization3rd)
        if ( present(Polarization4th) ) call set_rules(ctx, 4, Polarization4th)
        if ( present(Spectral_k) ) then
            call check(spr_kappa_rta_spectral(ctx%h, lw, T, s, kappa, Mode_k, Spectral_k), 'kappa_noIter_gpu')
        else
            call check(spr_kappa_rta(ctx%h, lw, T, s, kappa, Mode_k), 'kappa_noIter_gpu')
        end if
    end subroutine kappa_noIter_gpu

    ! PolarizationStrtEnd (Helper_m.f90:310-353): 'A'/'a' -> branches 1..3, 'O'/'o' -> 4..Ndof, any
    ! other letter -> 1..Ndof, per slot; nd_selection is true when some slot is restricted.
    subroutine PolarizationStrtEnd(order, Ndof, character_arr, out_int_arr, nd_selection)
        integer, intent(in)             :: order, Ndof
        character(len=1), intent(in)    :: character_arr(order)
        integer, intent(out)            :: out_int_arr(2, order)
        logical, intent(out)            :: nd_selection
        integer                         :: k
        nd_selection = .false.
        do k = 1, order
            select case (character_arr(k))
            case ('A', 'a')
                out_int_arr(:, k) = [1, 3]
                nd_selection = .true.
            case ('O', 'o')
                out_int_arr(:, k) = [4, Ndof]
                nd_selection = .true.
            case default
                out_int_arr(:, k) = [1, Ndof]
            end select
        end do
    end subroutine PolarizationStrtEnd

    subroutine set_rules(ctx, order, pol)
        type(gpu_ctx), intent(in)   :: ctx
        integer, intent(in)         :: order, pol(:,:)
        integer(c_int)              :: p(2 * order)
        p = int(reshape(pol, [2 * order]), c_int)
        call check(spr_set_branch_rules(ctx%h, int(order, c_int), p), 'spr_set_branch_rules')
    end subroutine set_rules

    ! ---- iterative BTE (src/Kappa/Src_iterative) on the time-reversal irreducible points ----

    ! num_irr_q, the 1-based grid index of each irreducible point and the grid points it stands
    ! for (Qpoints%irr_q_int(4,:), Irr_q_point_m.f90:296-330)
    subroutine IrrQ_tr_gpu(ctx, num_irr_q, irr_grid, deg)
        type(gpu_ctx), intent(in)                       :: ctx
        integer, intent(out)                            :: num_irr_q
        integer, allocatable, intent(out)               :: irr_grid(:), deg(:)
        integer(c_int)                                  :: n
        integer(c_int), allocatable, target             :: g(:), d(:)
        call check(spr_bte_irr(ctx%h, n, c_null_ptr, c_null_ptr), 'IrrQ_tr_gpu')
        allocate( g(n), d(n) )
        call check(spr_bte_irr(ctx%h, n, c_loc(g), c_loc(d)), 'IrrQ_tr_gpu')
        num_irr_q = n
        irr_grid = g + 1
        deg = d
    end subroutine IrrQ_tr_gpu

    ! Inv_tauqs(Ndof, num_irr_q) <- TetrahedronWplus + TetrahedronWminus
    ! (Quasi_momentum_consrv.f90:701-1208)
    subroutine TetrahedronW_gpu(ctx, Inv_tauqs)
        type(gpu_ctx), intent(in)   :: ctx
        real(dp), intent(out)       :: Inv_tauqs(:,:)
        call check(spr_bte_tau_inv(ctx%h, Inv_tauqs), 'TetrahedronW_gpu')
    end subroutine TetrahedronW_gpu

    ! dF_new(3, Ndof, num_irr_q) of F_vecAll, before multiplyTaudFnew <- TetrahedronIterPlus +
    ! TetrahedronIterMinus (TetrahedronDeltaSum.f90:22-620)
    subroutine TetrahedronIter_gpu(ctx, F_vecAll, dF_new)
        type(gpu_ctx), intent(in)   :: ctx
        real(dp), intent(in)        :: F_vecAll(:,:,:)
        real(dp), intent(out)       :: dF_new(:,:,:)
        real(dp), allocatable       :: F(:,:,:)
        F = F_vecAll
        call check(spr_bte_delta_f(ctx%h, F, dF_new), 'TetrahedronIter_gpu')
    end subroutine TetrahedronIter_gpu

    ! The sc_iter loop of main.f90:243-361: k_tensor(9) and the kappa of F0 k_tensor0(9) (W/mK),
    ! iterNum iterations until max|k - k_old| < eps_k or MAX_ITER
    subroutine kappa_iter_gpu(ctx, T, eps_k, max_iter, k_tensor, k_tensor0, iterNum)
        type(gpu_ctx), intent(in)   :: ctx
        real(dp), intent(in)        :: T, eps_k
        integer, intent(in)         :: max_iter
        real(dp), intent(out)       :: k_tensor(9), k_tensor0(9)
        integer, intent(out)        :: iterNum
        integer(c_int)              :: it
        call check(spr_bte_solve(ctx%h, T, eps_k, int(max_iter, c_int), k_tensor, k_tensor0, it, c_null_ptr), &
                 & 'kappa_iter_gpu')
        iterNum = it
    end subroutine kappa_iter_gpu

    ! get_phonon (DynaMat_mod.f90:486-709) for a non-polar crystal on the device: latvec(3,3) with
    ! the lattice vectors as columns (R = matmul(latvec, N)), FC2 as FC2type holds it
    ! (atmNum2, atmIndx2(4,:,:) with a 1-based basis index, fC2(1,3,3,Natm,Nb), only fC2(1,...)
    ! used), qpoints(3,Nq) Cartesian.  A q with all |q_k| < cmp_prec (1e-8) is Gamma: its first
    ! three modes get grp_vel = nBE = OnebyFreq = 0, and freq_min/freq_max are set from it as
    ! in :629-641 (left 0 when no q is Gamma).
    subroutine get_phonon_gpu(Nbasis, mass, latvec, atmNum2, atmIndx2, fC2, qpoints, T, &
                            & Evec, freq, OnebyFreq, nBE, grp_vel, device, freq_min, freq_max)
        integer, intent(in)                     :: Nbasis, device
        real(dp), intent(in)                    :: mass(:), latvec(3,3), fC2(:,:,:,:,:), qpoints(:,:), T
        integer, intent(in)                     :: atmNum2(:), atmIndx2(:,:,:)
        complex(dp), intent(out)                :: Evec(:,:,:)
        real(dp), intent(out)                   :: freq(:,:), OnebyFreq(:,:), nBE(:,:), grp_vel(:,:,:)
        real(dp), intent(out), optional         :: freq_min, freq_max
        real(dp), parameter                     :: cmp_prec = 1.0e-8_dp
        integer(c_int), allocatable, target     :: num2(:), indx2(:,:,:), gam(:)
        real(dp), allocatable, target           :: mass_c(:), fc(:,:,:,:)
        real(c_double)                          :: lat_c(9)
        type(spr_crystal)                       :: cr
        type(spr_ifc)                           :: f2
        integer                                 :: nq, i

        nq = size(qpoints, 2)
        num2 = atmNum2
        indx2 = atmIndx2
        indx2(4, :, :) = indx2(4, :, :) - 1
        fc = fC2(1, :, :, :, :)
        mass_c = mass
        lat_c = reshape(transpose(latvec), [9])       ! row-major latvec[i][j] = latvec(i,j)
        allocate( gam(nq) )
        do i = 1, nq
            gam(i) = merge(1_c_int, 0_c_int, all(dabs(qpoints(:, i)) < cmp_prec))
        end do
        cr = spr_crystal(int(Nbasis, c_int), c_loc(mass_c), 0.0_dp)
        f2 = spr_ifc(2, int(size(atmIndx2, 2), c_int), c_loc(num2), c_loc(indx2), c_loc(fc))
        call check(spr_get_phonon(cr, lat_c, f2, int(nq, c_int), qpoints, gam, T, int(device, c_int), Evec, &
                 & freq, OnebyFreq, nBE, grp_vel), 'get_phonon_gpu')
        if ( present(freq_min) ) freq_min = 0.0_dp
        if ( present(freq_max) ) freq_max = 0.0_dp
        do i = 1, nq
            if ( gam(i) /= 0 ) then
                if ( present(freq_max) ) freq_max = maxval(freq(:, i))
                if ( present(freq_min) ) freq_min = 1.05_dp * dabs(freq(3, i))
            end if
        end do
    end subroutine get_phonon_gpu

    ! LineWidth.x (Linewidth_m.f90:48-417, main.f90:293-379): the three-phonon linewidth lw(Ndof) of
    ! a q0 off the mesh, q0(3) Cartesian with its phonons (Evec_q0(Ndof,Ndof), omega_q0,
    ! OnebyOmega_q0 as set_Phon_mp gives them); the q1 are the context's mesh, ZeroCentered picks
    ! the reference's BZ test and q1 folding; q2 = G - (q0 + q1) as ScatterProbAllq1 finds it,
    ! with its phonons solved on the device from FC2 (FC2type arrays, as get_phonon_gpu).  Sigma
    ! present: the Gaussian sum with that width (rad/ps) instead of the tetrahedron method.
    ! Sctrq1(Ndof,Ndof,Ndof,Nq) optional: the W rows.  FP64.
    subroutine CalculateLwOffMesh_gpu(ctx, q0, Evec_q0, omega_q0, OnebyOmega_q0, Nbasis, mass, latvec, &
                                    & atmNum2, atmIndx2, fC2, T, ZeroCentered, lw, Sctrq1, Sigma)
        type(gpu_ctx), intent(in)                   :: ctx
        real(dp), intent(in)                        :: q0(3), omega_q0(:), OnebyOmega_q0(:), mass(:), latvec(3,3), &
                                                     & fC2(:,:,:,:,:), T
        complex(dp), intent(in)                     :: Evec_q0(:,:)
        integer, intent(in)                         :: Nbasis, atmNum2(:), atmIndx2(:,:,:)
        logical, intent(in)                         :: ZeroCentered
        real(dp), intent(out)                       :: lw(:)
        real(dp), intent(out), optional, target, contiguous :: Sctrq1(:,:,:,:)
        real(dp), intent(in), optional              :: Sigma
        integer(c_int), allocatable, target         :: num2(:), indx2(:,:,:)
        real(dp), allocatable, target               :: mass_c(:), fc(:,:,:,:)
        real(c_double)                              :: lat_c(9), sig
        type(spr_crystal)                           :: cr
        type(spr_ifc)                               :: f2
        type(c_ptr)                                 :: pw
        integer(c_int)                              :: delta
        num2 = atmNum2
        indx2 = atmIndx2
        indx2(4, :, :) = indx2(4, :, :) - 1
        fc = fC2(1, :, :, :, :)
        mass_c = mass
        lat_c = reshape(transpose(latvec), [9])
        cr = spr_crystal(int(Nbasis, c_int), c_loc(mass_c), 0.0_dp)
        f2 = spr_ifc(2, int(size(atmIndx2, 2), c_int), c_loc(num2), c_loc(indx2), c_loc(fc))
        pw = c_null_ptr
        if ( present(Sctrq1) ) pw = c_loc(Sctrq1)
        delta = 1
        sig = 0.0_dp
        if ( present(Sigma) ) then
            delta = 0
            sig = Sigma
        end if
        call check(spr_lw3_offmesh(ctx%h, q0, Evec_q0, omega_q0, OnebyOmega_q0, cr, lat_c, f2, T, &
                 & merge(1_c_int, 0_c_int, ZeroCentered), delta, sig, lw, pw), 'CalculateLwOffMesh_gpu')
    end subroutine CalculateLwOffMesh_gpu

    ! make_tetrahedron + FindIrrQ (TetrahedronQ_m.f90:56-331, zero-centred, shift 0) and the
    ! Cartesian rotations of set_SpaceGrSPGlib, from the crystal: latvec(3,3) (vectors as
    ! columns), basis_frac(3, Nbasis) (atm_pos), typ_indx(Nbasis), qmesh(3), time_rev, symprec
    ! (Angstrom).  Outputs allocated here, indices 1-based as the reference's: q_pnt_int(3, Nq),
    ! q_pnt(3, Nq), irr_grid(Nq0) (grid index of each irreducible point), pos(Nq),
    ! multiplicity(Nq0), Rxyz(3, 3, Nsym), tetrahedrons(4, 6 Nq).
    subroutine MakeMeshIrrQ_gpu(latvec, basis_frac, typ_indx, qmesh, time_rev, symprec, q_pnt_int, q_pnt, &
                              & irr_grid, pos, multiplicity, Rxyz, tetrahedrons)
        real(dp), intent(in)                        :: latvec(3,3), basis_frac(:,:), symprec
        integer, intent(in)                         :: typ_indx(:), qmesh(3), time_rev
        integer, allocatable, intent(out)           :: q_pnt_int(:,:), irr_grid(:), pos(:), multiplicity(:), &
                                                     & tetrahedrons(:,:)
        real(dp), allocatable, intent(out)          :: q_pnt(:,:), Rxyz(:,:,:)
        integer(c_int), allocatable, target         :: qi(:,:), ig(:), ps(:), mu(:), tt(:,:), typ(:)
        real(dp), allocatable, target               :: qc(:,:), rt(:,:,:), frac(:,:)
        real(c_double)                              :: lat_c(9)
        integer(c_int)                              :: nq0, nsym, m(3)
        integer                                     :: nq
        lat_c = reshape(transpose(latvec), [9])
        frac = basis_frac
        typ = typ_indx
        m = qmesh
        call check(spr_mesh_build(lat_c, int(size(typ), c_int), frac, typ, m, int(time_rev, c_int), 1_c_int, &
                 & symprec, nq0, nsym, c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr, &
                 & c_null_ptr), 'MakeMeshIrrQ_gpu')
        nq = product(qmesh)
        allocate( qi(3, nq), qc(3, nq), ig(nq0), ps(nq), mu(nq0), rt(3, 3, nsym), tt(4, 6 * nq) )
        call check(spr_mesh_build(lat_c, int(size(typ), c_int), frac, typ, m, int(time_rev, c_int), 1_c_int, &
                 & symprec, nq0, nsym, c_loc(qi), c_loc(qc), c_loc(ig), c_loc(ps), c_loc(mu), c_loc(rt), &
                 & c_loc(tt)), 'MakeMeshIrrQ_gpu')
        q_pnt_int = qi
        q_pnt = qc
        irr_grid = ig + 1
        pos = ps + 1
        multiplicity = mu
        Rxyz = rt
        tetrahedrons = tt + 1
    end subroutine MakeMeshIrrQ_gpu

end module spring_gpu_m
