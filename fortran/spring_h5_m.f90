! spring_h5_m -- Fortran side of libspring_h5.so (include/spring_h5.h): the force-constant files
! kappa.x reads and the kappa file it writes, with the reference's routine names and argument
! meaning (src/Kappa/Src/hdf5_wrap_mod.f90):
!   r2ndFC_h5 <- r2ndFC (:40-111): atmNum(Nb), atmIndx(4, Natm, Nb), FC2nd(1, 3, 3, Natm, Nb)
!                (the slice FC2(1, :, :, :, :) that DynMatgv_SR uses)
!   r3rdFC_h5 <- r3rdFC (:114-223): atmNum3, atmIndx3, FC_3rd(3, 3, 3, Natm, Natm, Nb)
!   r4thFC_h5 <- r4thFC (:226-337): atmNum4, atmIndx4, FC_4th(3, 3, 3, 3, Natm, Natm, Natm, Nb)
!   w_kappa_h5 <- the writes of main.f90:773-798 (w1dh5_no_grp, w2dh5_no_grp_type2,
!                w2dh5_no_grp): ThermalConductivityTensor, LineWidth, PhaseSpace3rd (optional),
!                ModeResolvedKappa (optional)
!   w_spectral_h5 <- w2dh5_no_grp(Spectral_k, kappa_outFile, "SpectralResolvedKappa") (main.f90:792)
! atmIndx keeps the files' 1-based basis index, as the reference's readers return it.  Failures
! print the message and ERROR STOP, as the reference's HDF5 calls leave a failed read fatal.
module spring_h5_m

    use iso_c_binding
    implicit none
    private

    integer, parameter  :: dp = c_double

    interface
        integer(c_int) function spr_h5_fc_dims(file, order, nb, natm, lead) bind(C, name="spr_h5_fc_dims")
            import :: c_int, c_char
            character(kind=c_char)  :: file(*)
            integer(c_int), value   :: order
            integer(c_int)          :: nb, natm, lead
        end function spr_h5_fc_dims

        integer(c_int) function spr_h5_read_fc(file, order, nb, natm, atm_num, atm_indx, fc) &
                & bind(C, name="spr_h5_read_fc")
            import :: c_int, c_char, c_double
            character(kind=c_char)  :: file(*)
            integer(c_int), value   :: order, nb, natm
            integer(c_int)          :: atm_num(*), atm_indx(*)
            real(c_double)          :: fc(*)
        end function spr_h5_read_fc

        integer(c_int) function spr_h5_write_kappa(file, kappa, ndof, nq0, omega_q0, lw, p3, mode_k) &
                & bind(C, name="spr_h5_write_kappa")
            import :: c_int, c_char, c_double, c_ptr
            character(kind=c_char)  :: file(*)
            real(c_double)          :: kappa(9)
            integer(c_int), value   :: ndof, nq0
            real(c_double)          :: omega_q0(*), lw(*)
            type(c_ptr), value      :: p3, mode_k
        end function spr_h5_write_kappa

        integer(c_int) function spr_h5_write_spectral(file, ndof, nq0, spec) bind(C, name="spr_h5_write_spectral")
            import :: c_int, c_char, c_double
            character(kind=c_char)  :: file(*)
            integer(c_int), value   :: ndof, nq0
            real(c_double)          :: spec(*)
        end function spr_h5_write_spectral

        type(c_ptr) function spr_h5_last_error() bind(C, name="spr_h5_last_error")
            import :: c_ptr
        end function spr_h5_last_error

        integer(c_size_t) function c_strlen(s) bind(C, name="strlen")
            import :: c_size_t, c_ptr
            type(c_ptr), value :: s
        end function c_strlen
    end interface

    public :: r2ndFC_h5, r3rdFC_h5, r4thFC_h5, w_kappa_h5, w_spectral_h5

contains

    subroutine check(rc, what)
        integer(c_int), intent(in)      :: rc
        character(len=*), intent(in)    :: what
        type(c_ptr)                     :: p
        character(kind=c_char), pointer :: chars(:)
        integer                         :: n
        if ( rc /= 0 ) then
            p = spr_h5_last_error()
            n = int(c_strlen(p))
            call c_f_pointer(p, chars, [n])
            write(*, '(A, A, *(A))') what, ' failed: ', chars
            ERROR STOP
        end if
    end subroutine check

    subroutine sizes(filename, order, nb, natm)
        character(len=*), intent(in)    :: filename
        integer, intent(in)             :: order
        integer, intent(out)            :: nb, natm
        integer(c_int)                  :: b, a, l
        call check(spr_h5_fc_dims(trim(filename)//c_null_char, int(order, c_int), b, a, l), 'spr_h5_fc_dims')
        nb = b
        natm = a
    end subroutine sizes

    subroutine r2ndFC_h5(filename, atmNum, atmIndx, FC2nd)
        character(len=*), intent(in)        :: filename
        integer, allocatable, intent(out)   :: atmNum(:), atmIndx(:,:,:)
        real(dp), allocatable, intent(out)  :: FC2nd(:,:,:,:,:)
        integer                             :: nb, natm
        integer(c_int), allocatable         :: num(:), indx(:,:,:)
        call sizes(filename, 2, nb, natm)
        allocate( num(nb), indx(4, natm, nb), FC2nd(1, 3, 3, natm, nb) )
        call check(spr_h5_read_fc(trim(filename)//c_null_char, 2_c_int, int(nb, c_int), int(natm, c_int), &
                 & num, indx, FC2nd), 'r2ndFC_h5')
        atmNum = num
        atmIndx = indx
    end subroutine r2ndFC_h5

    subroutine r3rdFC_h5(filename, atmNum3, atmIndx3, FC_3rd)
        character(len=*), intent(in)        :: filename
        integer, allocatable, intent(out)   :: atmNum3(:), atmIndx3(:,:,:)
        real(dp), allocatable, intent(out)  :: FC_3rd(:,:,:,:,:,:)
        integer                             :: nb, natm
        integer(c_int), allocatable         :: num(:), indx(:,:,:)
        call sizes(filename, 3, nb, natm)
        allocate( num(nb), indx(4, natm, nb), FC_3rd(3, 3, 3, natm, natm, nb) )
        call check(spr_h5_read_fc(trim(filename)//c_null_char, 3_c_int, int(nb, c_int), int(natm, c_int), &
                 & num, indx, FC_3rd), 'r3rdFC_h5')
        atmNum3 = num
        atmIndx3 = indx
    end subroutine r3rdFC_h5

    subroutine r4thFC_h5(filename, atmNum4, atmIndx4, FC_4th)
        character(len=*), intent(in)        :: filename
        integer, allocatable, intent(out)   :: atmNum4(:), atmIndx4(:,:,:)
        real(dp), allocatable, intent(out)  :: FC_4th(:,:,:,:,:,:,:,:)
        integer                             :: nb, natm
        integer(c_int), allocatable         :: num(:), indx(:,:,:)
        call sizes(filename, 4, nb, natm)
        allocate( num(nb), indx(4, natm, nb), FC_4th(3, 3, 3, 3, natm, natm, natm, nb) )
        call check(spr_h5_read_fc(trim(filename)//c_null_char, 4_c_int, int(nb, c_int), int(natm, c_int), &
                 & num, indx, FC_4th), 'r4thFC_h5')
        atmNum4 = num
        atmIndx4 = indx
    end subroutine r4thFC_h5

    ! kappa(9) W/mK, omega_q0 and LW_Allq0 (Ndof, Nq0), P_Allq0 (Ndof, Nq0) and Mode_k (9, Ndof)
    ! optional
    subroutine w_kappa_h5(filename, kappa, omega_q0, LW_Allq0, P_Allq0, Mode_k)
        character(len=*), intent(in)            :: filename
        real(dp), intent(in)                    :: kappa(9), omega_q0(:,:), LW_Allq0(:,:)
        real(dp), intent(in), optional, target, contiguous :: P_Allq0(:,:), Mode_k(:,:)
        real(dp)                                :: k(9)
        type(c_ptr)                             :: pp, pm
        pp = c_null_ptr
        pm = c_null_ptr
        if ( present(P_Allq0) ) pp = c_loc(P_Allq0)
        if ( present(Mode_k) ) pm = c_loc(Mode_k)
        k = kappa
        call check(spr_h5_write_kappa(trim(filename)//c_null_char, k, int(size(omega_q0, 1), c_int), &
                 & int(size(omega_q0, 2), c_int), omega_q0, LW_Allq0, pp, pm), 'w_kappa_h5')
    end subroutine w_kappa_h5

    ! SpectralResolvedKappa (11, Ndof Nq0) into the kappa file written by w_kappa_h5
    subroutine w_spectral_h5(filename, Spectral_k, Ndof, Nq0)
        character(len=*), intent(in)    :: filename
        real(dp), intent(in)            :: Spectral_k(:,:)
        integer, intent(in)             :: Ndof, Nq0
        real(dp), allocatable           :: a(:,:)
        a = Spectral_k
        call check(spr_h5_write_spectral(trim(filename)//c_null_char, int(Ndof, c_int), int(Nq0, c_int), a), &
                 & 'w_spectral_h5')
    end subroutine w_spectral_h5

end module spring_h5_m
