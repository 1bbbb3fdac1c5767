! kappa_nml_m -- the namelist-driven host of kappa_gpu: what src/Kappa/Src/main.f90 does before,
! around and after its q0 loop, with the scattering-rate work on the GPU.
!
!   command line  get_command_line (parse_cmd_line.f90): -inp [file] (default input.nml), -T [K]
!                 (default 300), -Symm T/F (default T), -Spectral_k T/F, -Mode_resolved_k T/F
!                 (default F); and for this host -prec fp64|fp32|bf16 and -delta tetra|gauss
!                 (tetra, the reference's, is the default; gauss uses Sigma of &AnhInfo)
!   input         &FCInfo, &CrystalInfo, &AtmInfo (cell%init_data, unit_mod.f90:64-214) and &AnhInfo
!                 (ReadNmlKappa, Helper_m.f90:37-132, same defaults), each read from the start of
!                 the file; FC2nd_/FC3rd_[/FC4th_]<prefix><T>K_F.h5 from the working directory
!                 (main.f90:182-198, 235-238) through libspring_h5
!   mesh          make_tetrahedron + FindIrrQ (TetrahedronQ_m.f90:56-331) and the space-group
!                 rotations: MakeMeshIrrQ_gpu (host code of libspring_gpu, no spglib)
!   phonons       set_Phon_mp (main.f90:279-300): get_phonon_gpu on every mesh point
!   q0 loop       CalculateLwTetrahedron with Polarization3rd (main.f90:585-590), P3 and the
!                 phase space (:592, 725), TetrahedronIntegration4th with Polarization4th times
!                 0.5 (:655-660); linewidth_IBZ_<prefix><T>K.txt per q0 (:697-700)
!   kappa         kappa_noIter_symm / kappa_noIter (:731-752) with Spectral_k and Mode_k, written to
!                 kappa_<prefix><T>K.h5 as :773-794 writes it, and printed as :796-816 prints it
! Not supported, with an error: Force_dd (Ewald term), isotope scattering, shift /= 0; restart and
! 4th-order scattering-matrix files are not written.
module kappa_nml_m

    use spring_gpu_m
    use spring_h5_m
    implicit none
    private

    public :: run_namelist_host, WriteAsciiOutFile, read_cell_nml, cell_volume, cross3, fail

contains

    subroutine run_namelist_host()
        character(len=512)          :: infile
        real(dp)                    :: T
        logical                     :: SpectralResolve, ModeResolve, Symm, gauss
        integer                     :: prec

        ! &FCInfo, &CrystalInfo, &AtmInfo
        character(len=64)           :: prefix
        logical                     :: Force_dd
        integer                     :: Nbasis
        real(dp)                    :: Latvec(3, 3)
        real(dp), allocatable       :: atm_pos(:,:), mass(:)
        integer, allocatable        :: typ_indx(:)
        ! &AnhInfo
        character(len=1)            :: SelectionRule3(3), SelectionRule4(4)
        character(len=512)          :: SctrEl4thRDir
        real(dp)                    :: Sigma, symprec, timelimit
        real(dp), allocatable       :: g2iso(:)
        integer                     :: qmesh(3), shift(3), time_rev
        logical                     :: isotope, FourPhonon, WriteRestart, ReadRestart, WriteSctrEl4th, ReadSctrEl4th
        namelist /AnhInfo/ SelectionRule3, SelectionRule4, SctrEl4thRDir, Sigma, symprec, timelimit, g2iso, &
                         & qmesh, shift, time_rev, isotope, FourPhonon, WriteRestart, ReadRestart, &
                         & WriteSctrEl4th, ReadSctrEl4th

        integer                     :: u, err, Ndof, Nq, Nq0, i0, i
        character(len=256)          :: msg
        character(len=32)           :: Temp_char, Ndof_char
        character(len=600)          :: fc2file, fc3file, fc4file, kappa_outFile, outfile_ascii, outfile_frmt
        real(dp)                    :: vol, freq_min, freq_max, P3, PhaseSpace3
        integer, allocatable        :: atmNum2(:), atmIndx2(:,:,:), atmNum3(:), atmIndx3(:,:,:), atmNum4(:), &
                                     & atmIndx4(:,:,:)
        real(dp), allocatable       :: fC2(:,:,:,:,:), fC3(:,:,:,:,:,:), fC4(:,:,:,:,:,:,:,:)
        integer, allocatable        :: q_pnt_int(:,:), irr_grid(:), pos(:), mult(:), tet(:,:)
        real(dp), allocatable       :: q_pnt(:,:), Rxyz(:,:,:)
        complex(dp), allocatable    :: Evec(:,:,:)
        real(dp), allocatable       :: omega(:,:), OnebyOmega(:,:), nBE(:,:), grp_vel(:,:,:)
        real(dp), allocatable       :: LW_Allq0(:,:), P_Allq0(:,:), lw(:), lw4(:,:), P3_q0(:), Mode_k(:,:), &
                                     & Spectral_k(:,:)
        integer                     :: Polarization3rd(2, 3), Polarization4th(2, 4)
        logical                     :: nd_Select_3rd, nd_Select_4th, HeaderWrite
        real(dp)                    :: kappa(9)
        type(gpu_ctx)               :: ctx

        call get_command_line(infile, T, SpectralResolve, ModeResolve, Symm, prec, gauss)

        ! ---------------- input.nml (unit_mod.f90:64-214, Helper_m.f90:37-132) ----------------
        call read_cell_nml(infile, prefix, Force_dd, Nbasis, Latvec, atm_pos, typ_indx, mass)
        ! ReadNmlKappa defaults (Helper_m.f90:66-83)
        SelectionRule3 = 'D'
        SelectionRule4 = 'D'
        SctrEl4thRDir = './ScatterMatEl4th_'//trim(prefix)
        Sigma = 5.1_dp
        symprec = 1.0e-7_dp
        timelimit = 365.0_dp * 24.0_dp * 60.0_dp
        allocate( g2iso(Nbasis) )
        g2iso = 0.0_dp
        qmesh = [7, 7, 7]
        shift = [0, 0, 0]
        time_rev = 0
        isotope = .false.
        FourPhonon = .false.
        WriteRestart = .false.
        ReadRestart = .false.
        WriteSctrEl4th = .false.
        ReadSctrEl4th = .false.
        open(newunit=u, file=trim(infile), status='old', action='read', delim='apostrophe', iostat=err, iomsg=msg)
        if ( err /= 0 ) call fail('opening namelist file '//trim(infile)//': '//trim(msg))
        read(u, nml=AnhInfo, iostat=err, iomsg=msg)
        if ( err /= 0 ) call fail('reading &AnhInfo of '//trim(infile)//': '//trim(msg))
        close(u)

        if ( Force_dd ) call fail('Force_dd = .true. (Ewald dipole-dipole term) is not supported by this library')
        if ( isotope ) call fail('isotope = .true.: isotope scattering is not supported by this library')
        if ( any(shift /= 0) ) call fail('shift /= 0: only the Gamma-centred mesh is supported')
        if ( WriteRestart .or. ReadRestart .or. WriteSctrEl4th .or. ReadSctrEl4th ) &
            & write(*, '(A)') 'WARNING: restart and scattering-matrix files are neither read nor written'

        Ndof = 3 * Nbasis
        vol = cell_volume(Latvec)

        write(Temp_char, '(F6.1)') T
        fc2file = 'FC2nd_'//trim(prefix)//trim(adjustl(Temp_char))//'K_F.h5'
        fc3file = 'FC3rd_'//trim(prefix)//trim(adjustl(Temp_char))//'K_F.h5'
        fc4file = 'FC4th_'//trim(prefix)//trim(adjustl(Temp_char))//'K_F.h5'
        kappa_outFile = 'kappa_'//trim(prefix)//trim(adjustl(Temp_char))//'K.h5'
        outfile_ascii = 'linewidth_IBZ_'//trim(prefix)//trim(adjustl(Temp_char))//'K.txt'
        write(Ndof_char, '(I4)') (2 * Ndof - 1)
        outfile_frmt = "("//trim(adjustl(Ndof_char))//"(G20.12, '  '), G20.12)"

        call PolarizationStrtEnd(3, Ndof, SelectionRule3, Polarization3rd, nd_Select_3rd)
        if ( FourPhonon ) call PolarizationStrtEnd(4, Ndof, SelectionRule4, Polarization4th, nd_Select_4th)

        ! ---------------- force constants (main.f90:182-198, 235-238) ----------------
        call r2ndFC_h5(trim(fc2file), atmNum2, atmIndx2, fC2)
        call r3rdFC_h5(trim(fc3file), atmNum3, atmIndx3, fC3)
        if ( size(atmNum2) /= Nbasis .or. size(atmNum3) /= Nbasis ) &
            & call fail('the FC files do not have Nbasis basis atoms: '//trim(fc2file)//', '//trim(fc3file))
        if ( FourPhonon ) then
            call r4thFC_h5(trim(fc4file), atmNum4, atmIndx4, fC4)
            if ( size(atmNum4) /= Nbasis ) call fail('the FC4 file does not have Nbasis basis atoms: '//trim(fc4file))
        end if

        ! ---------------- mesh, irreducible q0, rotations (main.f90:281-300) ----------------
        call MakeMeshIrrQ_gpu(Latvec, atm_pos, typ_indx, qmesh, time_rev, symprec, q_pnt_int, q_pnt, &
                            & irr_grid, pos, mult, Rxyz, tet)
        Nq = product(qmesh)
        Nq0 = size(irr_grid)
        write(*, '(A, 2(I4, A), I4, A, I6)') 'Number of irreducible q-points for ( ', qmesh(1), ' X', qmesh(2), &
            & ' X', qmesh(3), ' ) mesh is: ', Nq0

        ! ---------------- phonons (set_Phon_mp) ----------------
        allocate( Evec(Ndof, Ndof, Nq), omega(Ndof, Nq), OnebyOmega(Ndof, Nq), nBE(Ndof, Nq), grp_vel(3, Ndof, Nq) )
        call get_phonon_gpu(Nbasis, mass, Latvec, atmNum2, atmIndx2, fC2, q_pnt, T, Evec, omega, OnebyOmega, &
                          & nBE, grp_vel, 0, freq_min, freq_max)

        if ( FourPhonon ) then
            call gpu_ctx_create(ctx, Nbasis, mass, vol, atmNum3, atmIndx3, fC3, qmesh, q_pnt_int, irr_grid, pos, &
                              & mult, Rxyz, Evec, omega, OnebyOmega, nBE, grp_vel, 0, atmNum4, atmIndx4, fC4, &
                              & tetrahedrons=tet)
        else
            call gpu_ctx_create(ctx, Nbasis, mass, vol, atmNum3, atmIndx3, fC3, qmesh, q_pnt_int, irr_grid, pos, &
                              & mult, Rxyz, Evec, omega, OnebyOmega, nBE, grp_vel, 0, tetrahedrons=tet)
        end if

        ! ---------------- q0 loop (main.f90:552-715) ----------------
        allocate( LW_Allq0(Ndof, Nq0), P_Allq0(Ndof, Nq0), lw(Ndof), P3_q0(Ndof) )
        P_Allq0 = 0.0_dp
        PhaseSpace3 = 0.0_dp
        if ( gauss ) then
            call LinewidthAllq0_gpu(ctx, Sigma, prec, LW_Allq0, Polarization3rd)
            if ( FourPhonon ) then
                allocate( lw4(Ndof, Nq0) )
                call FourPhononLw_gpu(ctx, Sigma, prec, lw4, Polarization4th)
                LW_Allq0 = LW_Allq0 + 0.5_dp * lw4
            end if
        else
            do i0 = 1, Nq0
                call CalculateLwTetrahedron_gpu(ctx, i0, lw, P3_q0, P3, Polarization3rd)
                LW_Allq0(:, i0) = lw     ! third-order linewidth, no 0.5 factor (main.f90:590)
                PhaseSpace3 = PhaseSpace3 + dble(mult(i0)) * P3
                P_Allq0(:, i0) = P3_q0 * 1000.0_dp / ( 9.0_dp * dble(Ndof**2) * dble(Nq) )   ! fs (:689)
                if ( FourPhonon ) then
                    call TetrahedronIntegration4th_gpu(ctx, i0, lw, Polarization4th)
                    LW_Allq0(:, i0) = LW_Allq0(:, i0) + 0.5_dp * lw
                end if
            end do
            PhaseSpace3 = PhaseSpace3 * 1000.0_dp / ( 9.0_dp * dble(Ndof**3) * dble(Nq)**2 )   ! fs (:725)
        end if
        HeaderWrite = .true.
        do i0 = 1, Nq0
            call WriteAsciiOutFile(trim(outfile_ascii), trim(outfile_frmt), omega(:, irr_grid(i0)), &
                                 & LW_Allq0(:, i0), Ndof, HeaderWrite)
        end do

        ! ---------------- kappa (main.f90:731-794) ----------------
        allocate( Mode_k(9, Ndof) )
        Mode_k = 0.0_dp
        if ( SpectralResolve ) then
            allocate( Spectral_k(11, Ndof * Nq0) )
            if ( FourPhonon ) then
                call kappa_noIter_gpu(ctx, LW_Allq0, T, Symm, kappa, Mode_k, Spectral_k, Polarization3rd, Polarization4th)
            else
                call kappa_noIter_gpu(ctx, LW_Allq0, T, Symm, kappa, Mode_k, Spectral_k, Polarization3rd)
            end if
        else
            if ( FourPhonon ) then
                call kappa_noIter_gpu(ctx, LW_Allq0, T, Symm, kappa, Mode_k, Polarization3rd=Polarization3rd, &
                                    & Polarization4th=Polarization4th)
            else
                call kappa_noIter_gpu(ctx, LW_Allq0, T, Symm, kappa, Mode_k, Polarization3rd=Polarization3rd)
            end if
        end if
        if ( ModeResolve ) then
            if ( gauss ) then
                call w_kappa_h5(trim(kappa_outFile), kappa, omega(:, irr_grid), LW_Allq0, Mode_k=Mode_k)
            else
                call w_kappa_h5(trim(kappa_outFile), kappa, omega(:, irr_grid), LW_Allq0, P_Allq0=P_Allq0, Mode_k=Mode_k)
            end if
        else
            if ( gauss ) then
                call w_kappa_h5(trim(kappa_outFile), kappa, omega(:, irr_grid), LW_Allq0)
            else
                call w_kappa_h5(trim(kappa_outFile), kappa, omega(:, irr_grid), LW_Allq0, P_Allq0=P_Allq0)
            end if
        end if
        if ( SpectralResolve ) call w_spectral_h5(trim(kappa_outFile), Spectral_k, Ndof, Nq0)

        if ( nd_Select_3rd ) write(*, '(A, 2(A2, 1X), A2, A)') ' ********  You have chosen non-defult selection rule: (', &
            & SelectionRule3, ' ) ******** '
        if ( FourPhonon ) then
            if ( nd_Select_4th ) write(*, '(A, 3(A2, 1X), A2, A)') &
                & ' ******** (FourPhonon) You have chosen non-defult selection rule: (', SelectionRule4, ' ) ******** '
        end if
        write(*, '(A, A)') 'Thermal conductivity information written in file: ', trim(kappa_outFile)
        write(*, '(37X, "|kxx = ", G20.12, ", kxy = ", G20.12, ", kxz = ", G20.12, "|")') kappa(1:3)
        write(*, '("Thermal conductivity tensor (W/m.K): |kyx = ", G20.12, ", kyy = ", G20.12, ", kyz = ", G20.12, "|")') &
            & kappa(4:6)
        write(*, '(37X, "|kzx = ", G20.12, ", kzy = ", G20.12, ", kzz = ", G20.12, "|")') kappa(7:9)
        if ( .not. gauss ) write(*, '(A, G20.12, A)') '3-phonon scattering phase-space: ', PhaseSpace3, ' fs'
        call gpu_ctx_destroy(ctx)
    end subroutine run_namelist_host

    ! cell%init_data (unit_mod.f90:64-214): &FCInfo, &CrystalInfo and &AtmInfo, each read from the
    ! start of the file with the reference's defaults; mass(Nbasis) per atom from the per-type
    ! masses (:181-186), typ_indx 1-based, atm_pos(3, Nbasis) fractional
    subroutine read_cell_nml(infile, prefix, Force_dd, Nbasis, Latvec, atm_pos, typ_indx, mass_atm)
        character(len=*), intent(in)        :: infile
        character(len=64), intent(out)      :: prefix
        logical, intent(out)                :: Force_dd
        integer, intent(out)                :: Nbasis
        real(dp), intent(out)               :: Latvec(3, 3)
        real(dp), allocatable, intent(out)  :: atm_pos(:,:), mass_atm(:)
        integer, allocatable, intent(out)   :: typ_indx(:)
        logical                     :: SecondFC, ThirdFC, FourthFC
        integer                     :: nn_2nd, nn_3rd, nn_4th
        namelist /FCInfo/ prefix, SecondFC, ThirdFC, FourthFC, Force_dd, nn_2nd, nn_3rd, nn_4th
        integer                     :: Ntyp, Sup_Cell(3)
        namelist /CrystalInfo/ Nbasis, Ntyp, Latvec, Sup_Cell
        real(dp), allocatable       :: BornZ(:,:,:), mass(:)
        integer, allocatable        :: atomic_no(:)
        character(len=3), allocatable :: typ_lbl(:)
        real(dp)                    :: dielec(3, 3)
        namelist /AtmInfo/ atm_pos, typ_indx, typ_lbl, atomic_no, mass, dielec, BornZ
        integer                     :: u, err, i
        character(len=256)          :: msg
        prefix = 'spring'
        SecondFC = .true.
        ThirdFC = .true.
        FourthFC = .false.
        Force_dd = .false.
        nn_2nd = 0
        nn_3rd = 0
        nn_4th = 0
        Nbasis = 2
        Ntyp = 1
        Latvec = 0.0_dp
        Sup_Cell = 5
        open(newunit=u, file=trim(infile), status='old', action='read', delim='apostrophe', iostat=err, iomsg=msg)
        if ( err /= 0 ) call fail('opening namelist file '//trim(infile)//': '//trim(msg))
        read(u, nml=FCInfo, iostat=err, iomsg=msg)
        if ( err /= 0 ) call fail('reading &FCInfo of '//trim(infile)//': '//trim(msg))
        rewind(u)
        read(u, nml=CrystalInfo, iostat=err, iomsg=msg)
        if ( err /= 0 ) call fail('reading &CrystalInfo of '//trim(infile)//': '//trim(msg))
        allocate( atm_pos(3, Nbasis), typ_indx(Nbasis), typ_lbl(Ntyp), atomic_no(Ntyp), mass(Ntyp), &
                & BornZ(3, 3, Nbasis) )
        dielec = 1.0_dp
        BornZ = 0.0_dp
        atomic_no = 0
        typ_lbl = '   '
        rewind(u)
        read(u, nml=AtmInfo, iostat=err, iomsg=msg)
        if ( err /= 0 ) call fail('reading &AtmInfo of '//trim(infile)//': '//trim(msg))
        close(u)
        allocate( mass_atm(Nbasis) )
        do i = 1, Nbasis
            mass_atm(i) = mass(typ_indx(i))
        end do
    end subroutine read_cell_nml

    ! |a1 . (a2 x a3)| of the lattice vectors (columns of Latvec)
    pure function cell_volume(Latvec) result(vol)
        real(dp), intent(in)    :: Latvec(3, 3)
        real(dp)                :: vol
        vol = dabs(dot_product(Latvec(:, 1), cross3(Latvec(:, 2), Latvec(:, 3))))
    end function cell_volume

    pure function cross3(a, b) result(c)
        real(dp), intent(in)    :: a(3), b(3)
        real(dp)                :: c(3)
        c = [a(2) * b(3) - a(3) * b(2), a(3) * b(1) - a(1) * b(3), a(1) * b(2) - a(2) * b(1)]
    end function cross3

    ! get_command_line (parse_cmd_line.f90): "-name value" pairs, the reference's defaults, plus
    ! -prec and -delta of this host
    subroutine get_command_line(infile, T, SpectralResolve, ModeResolve, Symm, prec, gauss)
        character(len=*), intent(out)   :: infile
        real(dp), intent(out)           :: T
        logical, intent(out)            :: SpectralResolve, ModeResolve, Symm, gauss
        integer, intent(out)            :: prec
        character(len=512)              :: name, val
        integer                         :: n, k
        infile = 'input.nml'
        T = 300.0_dp
        SpectralResolve = .false.
        ModeResolve = .false.
        Symm = .true.
        prec = SPR_PREC_FP64
        gauss = .false.
        n = command_argument_count()
        k = 1
        do while ( k <= n )
            call get_command_argument(k, name)
            val = ''
            if ( k + 1 <= n ) call get_command_argument(k + 1, val)
            select case (trim(name))
            case ('-inp')
                infile = val
            case ('-T')
                read(val, *) T
            case ('-omegacut')
                continue
            case ('-Spectral_k')
                read(val, *) SpectralResolve
            case ('-Mode_resolved_k')
                read(val, *) ModeResolve
            case ('-Symm')
                read(val, *) Symm
            case ('-prec')
                if ( trim(val) == 'fp32' ) prec = SPR_PREC_FP32
                if ( trim(val) == 'bf16' ) prec = SPR_PREC_BF16
            case ('-delta')
                gauss = trim(val) == 'gauss'
            case default
                write(*, '(A, A)') 'WARNING: Unrecognised command name -> ', trim(name)
            end select
            k = k + 2
        end do
        write(*, '(A, A)') 'Input file to be read: ', trim(infile)
        write(*, '(A, F7.2, A)') 'Temperature: ', T, 'K'
    end subroutine get_command_line

    subroutine fail(msg)
        character(len=*), intent(in) :: msg
        write(*, '(A, A)') 'ERROR: ', msg
        ERROR STOP
    end subroutine fail

    ! WriteAsciiOutFile (Helper_m.f90:597-667): header on the first call, one row per q0
    subroutine WriteAsciiOutFile(filename, write_frmt, freq, LW_q0, Ndof, HeaderWrite)
        character(len=*), intent(in)    :: filename, write_frmt
        real(dp), intent(in)            :: freq(:), LW_q0(:)
        integer, intent(in)             :: Ndof
        logical, intent(inout)          :: HeaderWrite
        integer                         :: ui
        if ( HeaderWrite ) then
            open(newunit=ui, file=filename, status='REPLACE', action='WRITE')
            write(ui, 20) Ndof, Ndof
            HeaderWrite = .false.
        else
            open(newunit=ui, file=filename, status='OLD', action='WRITE', position='APPEND')
        end if
        write(ui, write_frmt) freq, LW_q0
        close(ui)
        20 FORMAT( '#--------- phonon_frequency(rad.THz) ( ', I4, ' columns) ---------      ', &
                 & '---------- linewidth(rad.THz) ( ', I4, ' columns) ----------')
    end subroutine WriteAsciiOutFile

end module kappa_nml_m
