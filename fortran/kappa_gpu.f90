! kappa_gpu -- Fortran host driver of the GPU scattering-rate path.
!
! Follows the three(+four)-phonon part of src/Kappa/Src/main.f90: the irreducible-q0
! loop (:552-715) producing LW_Allq0 and the per-q0 lines of linewidth_IBZ_<prefix><T>K.txt
! (WriteAsciiOutFile, Helper_m.f90:597-667, same header and G20.12 columns), then the RTA
! kappa (kappa_noIter_symm by default, -Symm of parse_cmd_line.f90) printed as in
! main.f90:799-819, and writes kappa_<prefix><T>K.h5 as main.f90:773-798 does.  Inputs come from a
! directory written by spring_amd.bundle.write_bundle (little-endian raw arrays in the
! reference's column-major layout plus a namelist `bundle.nml`); the force constants are read
! from the reference's own files FC2nd_/FC3rd_/FC4th_<prefix><T>K_F.h5 when the directory holds
! them (through libspring_h5, since the HDF5 Fortran module of this image is not readable by
! this compiler).
!
! usage: kappa_gpu [-inp input.nml] [-T K] [-Symm T|F] [-Spectral_k T|F] [-Mode_resolved_k T|F]
!                  [-prec fp64|fp32|bf16] [-delta tetra|gauss]
!   the reference's command line and inputs (input.nml + FC*_F.h5 in the working directory):
!   kappa_nml_m runs it.  With a directory as the first argument instead:
! usage: kappa_gpu <bundle-dir> [fp64|fp32|bf16] [loop|all|iter|tetra] [gpuph]
!   loop: one spr_lw3 call per q0 (the reference's q0 loop); all: one spr_lw3_all launch;
!   tetra: the reference's default delta, the tetrahedron method, per q0 for the three-phonon
!   (CalculateLwTetrahedron, main.f90:585-600) and the four-phonon sum (TetrahedronIntegration4th,
!   main.f90:655-660), FP64;
!   iter: the iterative solution of src/Kappa/Src_iterative/main.f90:243-361 (tetrahedron delta,
!   time-reversal irreducible points) instead, printing the kappa of F0 and the converged kappa;
!   gpuph: the phonons of the mesh are computed on the device from FC2 (get_phonon_gpu, as
!   set_Phon_mp does with get_phonon, main.f90:279-300) instead of read from the bundle
program kappa_gpu

    use spring_gpu_m
    use spring_h5_m
    use kappa_nml_m
    implicit none

    character(len=512)      :: dir, arg
    character(len=64)       :: prefix
    integer                 :: Nbasis, Natm2, Natm3, Natm4, mesh(3), Nq, Nq0, Nsym, FourPhonon, prec, i0, u
    real(dp)                :: T, Sigma, vol
    integer, allocatable    :: atmNum3(:), atmIndx3(:,:,:), atmNum4(:), atmIndx4(:,:,:)
    integer, allocatable    :: q_int(:,:), irr_grid(:), pos(:), mult(:)
    real(dp), allocatable   :: mass(:), fC3(:,:,:,:,:,:), fC4(:,:,:,:,:,:,:,:), Rxyz(:,:,:)
    complex(dp), allocatable :: Evec(:,:,:)
    real(dp), allocatable   :: omega(:,:), OnebyOmega(:,:), nBE(:,:), grp_vel(:,:,:)
    real(dp), allocatable   :: LW_Allq0(:,:), lw4(:,:), Mode_k(:,:), lw(:), P3_q0(:), P_Allq0(:,:)
    real(dp)                :: kappa(9)
    type(gpu_ctx)           :: ctx
    logical                 :: HeaderWrite, loop_mode, iter_mode, tetra_mode, gpu_phonons
    real(dp)                :: P3
    integer, allocatable    :: atmNum2(:), atmIndx2(:,:,:)
    real(dp), allocatable   :: fC2(:,:,:,:,:), q_cart(:,:)
    real(dp)                :: latvec(3,3), freq_min, freq_max
    integer, allocatable    :: tet(:,:)
    real(dp)                :: k_iter(9), k_iter0(9)
    integer                 :: iterNum
    character(len=32)       :: Temp_char, frmt, Ndof_char
    character(len=600)      :: outfile_ascii, outfile_frmt
    integer                 :: Ndof
    character(len=600)      :: fc2file, fc3file, fc4file, kappa_outFile
    logical                 :: have_h5
    integer, allocatable    :: num_h5(:), indx_h5(:,:,:)
    real(dp), allocatable   :: fc3_h5(:,:,:,:,:,:), fc4_h5(:,:,:,:,:,:,:,:), fc2_h5(:,:,:,:,:)

    namelist /bundle/ prefix, Nbasis, Natm2, Natm3, Natm4, mesh, Nq, Nq0, Nsym, T, Sigma, vol, FourPhonon

    call get_command_argument(1, dir)
    if ( len_trim(dir) == 0 .or. dir(1:1) == '-' ) then
        call run_namelist_host()
        stop
    end if
    prec = SPR_PREC_FP64
    call get_command_argument(2, arg)
    if ( trim(arg) == 'fp32' ) prec = SPR_PREC_FP32
    if ( trim(arg) == 'bf16' ) prec = SPR_PREC_BF16
    call get_command_argument(3, arg)
    loop_mode = trim(arg) == 'loop'
    iter_mode = trim(arg) == 'iter'
    tetra_mode = trim(arg) == 'tetra'
    call get_command_argument(4, arg)
    gpu_phonons = trim(arg) == 'gpuph'

    FourPhonon = 0
    Natm2 = 1
    Natm4 = 1
    open(newunit=u, file=trim(dir)//'/bundle.nml', status='old', action='read')
    read(u, nml=bundle)
    close(u)
    Ndof = 3 * Nbasis
    write(Temp_char, '(F6.1)') T
    fc2file = trim(dir)//'/FC2nd_'//trim(prefix)//trim(adjustl(Temp_char))//'K_F.h5'
    fc3file = trim(dir)//'/FC3rd_'//trim(prefix)//trim(adjustl(Temp_char))//'K_F.h5'
    fc4file = trim(dir)//'/FC4th_'//trim(prefix)//trim(adjustl(Temp_char))//'K_F.h5'
    kappa_outFile = 'kappa_'//trim(prefix)//trim(adjustl(Temp_char))//'K.h5'

    allocate( mass(Nbasis), atmNum3(Nbasis), atmIndx3(4, Natm3, Nbasis), fC3(3,3,3,Natm3,Natm3,Nbasis) )
    allocate( q_int(3, Nq), irr_grid(Nq0), pos(Nq), mult(Nq0), Rxyz(3, 3, Nsym) )
    allocate( Evec(Ndof, Ndof, Nq), omega(Ndof, Nq), OnebyOmega(Ndof, Nq), nBE(Ndof, Nq), grp_vel(3, Ndof, Nq) )
    call rd_r(trim(dir)//'/mass.bin', mass)
    call rd_i(trim(dir)//'/atm_num3.bin', atmNum3)
    call rd_i3(trim(dir)//'/atm_indx3.bin', atmIndx3)
    inquire(file=trim(fc3file), exist=have_h5)
    if ( have_h5 ) then
        ! r3rdFC (hdf5_wrap_mod.f90:114-223); the file's Natm and Nb must match the bundle
        call r3rdFC_h5(trim(fc3file), num_h5, indx_h5, fc3_h5)
        if ( size(fc3_h5, 4) /= Natm3 .or. size(fc3_h5, 6) /= Nbasis ) then
            write(*, *) 'FC3 file shape differs from bundle.nml: ', trim(fc3file)
            ERROR STOP
        end if
        atmNum3 = num_h5
        atmIndx3 = indx_h5
        fC3 = fc3_h5
        write(*, '(A, A)') 'Reading Third-Order Force constant file: ', trim(fc3file)
    else
        call rd_r6(trim(dir)//'/fc3.bin', fC3)
    end if
    call rd_i2(trim(dir)//'/q_int.bin', q_int)
    call rd_i(trim(dir)//'/irr_grid.bin', irr_grid)
    call rd_i(trim(dir)//'/pos.bin', pos)
    call rd_i(trim(dir)//'/multiplicity.bin', mult)
    call rd_r3(trim(dir)//'/rot.bin', Rxyz)
    if ( gpu_phonons ) then
        inquire(file=trim(fc2file), exist=have_h5)
        if ( have_h5 ) then
            call r2ndFC_h5(trim(fc2file), atmNum2, atmIndx2, fc2_h5)   ! r2ndFC, hdf5_wrap_mod.f90:40-111
            fC2 = fc2_h5
            allocate( q_cart(3, Nq) )
            write(*, '(A, A)') 'Reading Second-Order Force constant file: ', trim(fc2file)
        else
            allocate( atmNum2(Nbasis), atmIndx2(4, Natm2, Nbasis), fC2(1, 3, 3, Natm2, Nbasis), q_cart(3, Nq) )
            call rd_i(trim(dir)//'/atm_num2.bin', atmNum2)
            call rd_i3(trim(dir)//'/atm_indx2.bin', atmIndx2)
            open(newunit=u, file=trim(dir)//'/fc2.bin', access='stream', form='unformatted', status='old')
            read(u) fC2
            close(u)
        end if
        open(newunit=u, file=trim(dir)//'/latvec.bin', access='stream', form='unformatted', status='old')
        read(u) latvec
        close(u)
        call rd_r2(trim(dir)//'/q_cart.bin', q_cart)
        call get_phonon_gpu(Nbasis, mass, latvec, atmNum2, atmIndx2, fC2, q_cart, T, Evec, omega, &
                          & OnebyOmega, nBE, grp_vel, 0, freq_min, freq_max)
        write(*, '(A, ES16.8, A, ES16.8)') 'Phonons on the device: freq_min ', freq_min, ' freq_max ', freq_max
    else
        call rd_c3(trim(dir)//'/evec.bin', Evec)
        call rd_r2(trim(dir)//'/omega.bin', omega)
        call rd_r2(trim(dir)//'/inv_omega.bin', OnebyOmega)
        call rd_r2(trim(dir)//'/nbe.bin', nBE)
        call rd_r3(trim(dir)//'/grp_vel.bin', grp_vel)
    end if

    if ( iter_mode ) then
        ! ------------------ iterative BTE (Src_iterative/main.f90:243-361), MAX_ITER = 1000 ------------------
        allocate( tet(4, 6 * Nq) )
        call rd_i2(trim(dir)//'/tetra.bin', tet)
        call gpu_ctx_create(ctx, Nbasis, mass, vol, atmNum3, atmIndx3, fC3, mesh, q_int, irr_grid, pos, &
                          & mult, Rxyz, Evec, omega, OnebyOmega, nBE, grp_vel, 0, tetrahedrons=tet)
        call kappa_iter_gpu(ctx, T, 1.0e-6_dp, 1000, k_iter, k_iter0, iterNum)
        write(*, '(A)') 'Thermal conductivity tensor, relaxation time approximation (W/mK):'
        write(*, '(3(ES22.14, 2X))') k_iter0(1:3)
        write(*, '(3(ES22.14, 2X))') k_iter0(4:6)
        write(*, '(3(ES22.14, 2X))') k_iter0(7:9)
        write(*, '(A, I0, A)') 'Iterative solution of BTE: ', iterNum, ' iterations'
        write(*, '(A)') 'Thermal conductivity tensor (W/mK):'
        write(*, '(3(ES22.14, 2X))') k_iter(1:3)
        write(*, '(3(ES22.14, 2X))') k_iter(4:6)
        write(*, '(3(ES22.14, 2X))') k_iter(7:9)
        call gpu_ctx_destroy(ctx)
        stop
    end if

    if ( tetra_mode ) then
        allocate( tet(4, 6 * Nq) )
        call rd_i2(trim(dir)//'/tetra.bin', tet)
    end if
    if ( FourPhonon /= 0 ) then
        inquire(file=trim(fc4file), exist=have_h5)
        if ( have_h5 ) then
            call r4thFC_h5(trim(fc4file), atmNum4, atmIndx4, fc4_h5)   ! r4thFC, hdf5_wrap_mod.f90:226-337
            fC4 = fc4_h5
            write(*, '(A, A)') 'Reading Fourth-Order Force constant file: ', trim(fc4file)
        else
            allocate( atmNum4(Nbasis), atmIndx4(4, Natm4, Nbasis), fC4(3,3,3,3,Natm4,Natm4,Natm4,Nbasis) )
            call rd_i(trim(dir)//'/atm_num4.bin', atmNum4)
            call rd_i3(trim(dir)//'/atm_indx4.bin', atmIndx4)
            open(newunit=u, file=trim(dir)//'/fc4.bin', access='stream', form='unformatted', status='old')
            read(u) fC4
            close(u)
        end if
        if ( tetra_mode ) then
            call gpu_ctx_create(ctx, Nbasis, mass, vol, atmNum3, atmIndx3, fC3, mesh, q_int, irr_grid, pos, &
                              & mult, Rxyz, Evec, omega, OnebyOmega, nBE, grp_vel, 0, atmNum4, atmIndx4, fC4, &
                              & tetrahedrons=tet)
        else
            call gpu_ctx_create(ctx, Nbasis, mass, vol, atmNum3, atmIndx3, fC3, mesh, q_int, irr_grid, pos, &
                              & mult, Rxyz, Evec, omega, OnebyOmega, nBE, grp_vel, 0, atmNum4, atmIndx4, fC4)
        end if
    else if ( tetra_mode ) then
        call gpu_ctx_create(ctx, Nbasis, mass, vol, atmNum3, atmIndx3, fC3, mesh, q_int, irr_grid, pos, &
                          & mult, Rxyz, Evec, omega, OnebyOmega, nBE, grp_vel, 0, tetrahedrons=tet)
    else
        call gpu_ctx_create(ctx, Nbasis, mass, vol, atmNum3, atmIndx3, fC3, mesh, q_int, irr_grid, pos, &
                          & mult, Rxyz, Evec, omega, OnebyOmega, nBE, grp_vel, 0)
    end if

    ! ---------------------------- three-phonon linewidths (main.f90:552-715) ----------------------------
    allocate( LW_Allq0(Ndof, Nq0), lw(Ndof) )
    if ( tetra_mode ) then
        allocate( P3_q0(Ndof), P_Allq0(Ndof, Nq0) )
        do i0 = 1, Nq0
            call CalculateLwTetrahedron_gpu(ctx, i0, lw, P3_q0, P3)
            LW_Allq0(:, i0) = lw
            P_Allq0(:, i0) = P3_q0 * 1000.0_dp / ( 9.0_dp * dble( Ndof**2 ) * dble( Nq ) )   ! fs, main.f90:689
            if ( FourPhonon /= 0 ) then
                call TetrahedronIntegration4th_gpu(ctx, i0, lw)
                LW_Allq0(:, i0) = LW_Allq0(:, i0) + 0.5_dp * lw      ! main.f90:660
            end if
        end do
    else if ( loop_mode ) then
        do i0 = 1, Nq0
            call CalculateLinewidth_gpu(ctx, i0, Sigma, prec, lw)
            LW_Allq0(:, i0) = lw    !** third-order linewidth, no 0.5 factor (main.f90:585) **!
        end do
    else
        call LinewidthAllq0_gpu(ctx, Sigma, prec, LW_Allq0)
    end if
    if ( FourPhonon /= 0 .and. .not. tetra_mode ) then
        allocate( lw4(Ndof, Nq0) )
        call FourPhononLw_gpu(ctx, Sigma, prec, lw4)
        LW_Allq0 = LW_Allq0 + 0.5_dp * lw4      ! main.f90:660
    end if

    outfile_ascii = 'linewidth_IBZ_'//trim(prefix)//trim(adjustl(Temp_char))//'K.txt'
    write(Ndof_char, '(I4)') (2 * Ndof - 1)
    outfile_frmt = "("//trim(adjustl(Ndof_char))//"(G20.12, '  '), G20.12)"
    HeaderWrite = .true.
    do i0 = 1, Nq0
        call WriteAsciiOutFile(trim(outfile_ascii), trim(outfile_frmt), omega(:, irr_grid(i0)), &
                             & LW_Allq0(:, i0), Ndof, HeaderWrite)
    end do

    ! ---------------------------------- kappa (kappa_noIter_symm) ---------------------------------------
    allocate( Mode_k(9, Ndof) )
    call kappa_noIter_gpu(ctx, LW_Allq0, T, .true., kappa, Mode_k)
    write(*, '(A)') 'Thermal conductivity tensor (W/mK):'
    write(*, '(3(ES22.14, 2X))') kappa(1:3)
    write(*, '(3(ES22.14, 2X))') kappa(4:6)
    write(*, '(3(ES22.14, 2X))') kappa(7:9)
    ! kappa_<prefix><T>K.h5 (main.f90:773-798): the tensor, [omega; linewidth] of every q0 and the
    ! mode-resolved kappa
    if ( tetra_mode ) then
        ! PhaseSpace3rd as main.f90:777 writes it in the reference's (tetrahedron) mode
        call w_kappa_h5(trim(kappa_outFile), kappa, omega(:, irr_grid), LW_Allq0, P_Allq0=P_Allq0, Mode_k=Mode_k)
    else
        call w_kappa_h5(trim(kappa_outFile), kappa, omega(:, irr_grid), LW_Allq0, Mode_k=Mode_k)
    end if
    write(*, '(A, A)') 'Wrote ', trim(kappa_outFile)
    call gpu_ctx_destroy(ctx)

contains

    subroutine rd_r(f, a)
        character(len=*), intent(in) :: f
        real(dp), intent(out) :: a(:)
        integer :: v
        open(newunit=v, file=f, access='stream', form='unformatted', status='old'); read(v) a; close(v)
    end subroutine rd_r
    subroutine rd_r2(f, a)
        character(len=*), intent(in) :: f
        real(dp), intent(out) :: a(:,:)
        integer :: v
        open(newunit=v, file=f, access='stream', form='unformatted', status='old'); read(v) a; close(v)
    end subroutine rd_r2
    subroutine rd_r3(f, a)
        character(len=*), intent(in) :: f
        real(dp), intent(out) :: a(:,:,:)
        integer :: v
        open(newunit=v, file=f, access='stream', form='unformatted', status='old'); read(v) a; close(v)
    end subroutine rd_r3
    subroutine rd_r6(f, a)
        character(len=*), intent(in) :: f
        real(dp), intent(out) :: a(:,:,:,:,:,:)
        integer :: v
        open(newunit=v, file=f, access='stream', form='unformatted', status='old'); read(v) a; close(v)
    end subroutine rd_r6
    subroutine rd_c3(f, a)
        character(len=*), intent(in) :: f
        complex(dp), intent(out) :: a(:,:,:)
        integer :: v
        open(newunit=v, file=f, access='stream', form='unformatted', status='old'); read(v) a; close(v)
    end subroutine rd_c3
    subroutine rd_i(f, a)
        character(len=*), intent(in) :: f
        integer, intent(out) :: a(:)
        integer :: v
        open(newunit=v, file=f, access='stream', form='unformatted', status='old'); read(v) a; close(v)
    end subroutine rd_i
    subroutine rd_i2(f, a)
        character(len=*), intent(in) :: f
        integer, intent(out) :: a(:,:)
        integer :: v
        open(newunit=v, file=f, access='stream', form='unformatted', status='old'); read(v) a; close(v)
    end subroutine rd_i2
    subroutine rd_i3(f, a)
        character(len=*), intent(in) :: f
        integer, intent(out) :: a(:,:,:)
        integer :: v
        open(newunit=v, file=f, access='stream', form='unformatted', status='old'); read(v) a; close(v)
    end subroutine rd_i3

end program kappa_gpu
