! linewidth_gpu -- the three-phonon linewidth along a dispersion path on the GPU: what
! src/LineWidth/Src/main.f90:21-408 (LineWidth.x) does, with the scattering-rate work in
! libspring_gpu.
!
! usage: linewidth_gpu [-inp input.nml] [-T K] [-omegacut x] [-delta tetra|gauss]
!   command line  parse_cmd_line.f90 of LineWidth: -inp (default input.nml), -T (default 300 K),
!                 -omegacut (accepted, unused, as there); and for this host -delta gauss, the
!                 Gaussian sum with Sigma of &LWInfo (rad/ps), which main.f90:340-351 keeps
!                 switched off; the tetrahedron method (:353-363) is the default
!   input         &FCInfo, &CrystalInfo, &AtmInfo (cell%init_data, unit_mod.f90:64-214), &LWInfo
!                 (Sigma, qmesh, num_points) and &HighSymPathLW (q_high_sym(4, num_points)),
!                 main.f90:135-177; FC2nd_/FC3rd_<prefix><T>K_F.h5 from the working directory
!                 (:117-124) through libspring_h5
!   path          q_points_highsymm (CreateMesh_mod.f90:85-172): count(n+1) points on the
!                 segment n -> n+1, the start node excluded; a segment is zero-centred when an end
!                 has a negative coordinate or all coordinates of an end are within [-1/2, 1/2]
!   q1 mesh       make_tetrahedron of qmesh, shift 0 (main.f90:187-198): MakeMeshIrrQ_gpu; the
!                 zero-centred and the non-zero-centred mesh share its tetrahedra and phonons
!                 (a q1 and q1 + G have the same dynamical matrix), the off-mesh entry folds q1
!                 and applies the BZ test of the segment's kind
!   phonons       set_Phon_mp of the path and of the mesh (:229-235): get_phonon_gpu
!   q0 loop       ScatterProbAllq1 + CalculateLwTetrahedron per path point (:293-379):
!                 CalculateLwOffMesh_gpu; one row per point of linewidth_<prefix><T>K.txt
!                 (WriteAsciiOutFile, Helper_m.f90:167-239: distance, Ndof frequencies, Ndof
!                 linewidths, G20.12, rad/ps)
! Not supported, with an error: Force_dd (Ewald term).  The first node's count is not used (the
! reference allocates rows for it and never fills them); a nonzero one draws a warning.
program linewidth_gpu

    use spring_gpu_m
    use spring_h5_m
    use kappa_nml_m
    implicit none

    character(len=512)          :: infile
    real(dp)                    :: T
    logical                     :: gauss
    character(len=64)           :: prefix
    logical                     :: Force_dd
    integer                     :: Nbasis, Ndof, Nq, Nq0, Npath, i0, i, n, k, u, err
    real(dp)                    :: Latvec(3, 3), G(3, 3), vol
    real(dp), allocatable       :: atm_pos(:,:), mass(:)
    integer, allocatable        :: typ_indx(:)
    ! &LWInfo, &HighSymPathLW
    real(dp)                    :: Sigma
    integer                     :: qmesh(3), num_points
    namelist /LWInfo/ Sigma, qmesh, num_points
    real(dp), allocatable       :: q_high_sym(:,:)
    namelist /HighSymPathLW/ q_high_sym
    character(len=256)          :: msg
    character(len=32)           :: Temp_char, Ndof_char
    character(len=600)          :: fc2file, fc3file, outfile_ascii, outfile_frmt
    integer, allocatable        :: atmNum2(:), atmIndx2(:,:,:), atmNum3(:), atmIndx3(:,:,:)
    real(dp), allocatable       :: fC2(:,:,:,:,:), fC3(:,:,:,:,:,:)
    integer, allocatable        :: q_pnt_int(:,:), irr_grid(:), pos(:), mult(:), tet(:,:)
    real(dp), allocatable       :: q_pnt(:,:), Rxyz(:,:,:)
    real(dp), allocatable       :: q0(:,:), dist_arr(:)
    logical, allocatable        :: ZeroCentered(:)
    complex(dp), allocatable    :: Evec(:,:,:), Evec0(:,:,:)
    real(dp), allocatable       :: omega(:,:), OnebyOmega(:,:), nBE(:,:), grp_vel(:,:,:)
    real(dp), allocatable       :: omega0(:,:), OnebyOmega0(:,:), nBE0(:,:), grp_vel0(:,:,:)
    real(dp), allocatable       :: lw(:)
    logical                     :: HeaderWrite
    integer(8)                  :: t_start, t_end, t_rate
    type(gpu_ctx)               :: ctx

    call system_clock(t_start, t_rate)
    call get_lw_command_line(infile, T, gauss)
    call read_cell_nml(infile, prefix, Force_dd, Nbasis, Latvec, atm_pos, typ_indx, mass)
    if ( Force_dd ) call fail('Force_dd = .true. (Ewald dipole-dipole term) is not supported by this library')
    Ndof = 3 * Nbasis
    vol = cell_volume(Latvec)
    do i = 1, 3   ! unit_mod.f90:190-202
        G(:, i) = cross3(Latvec(:, modulo(i, 3) + 1), Latvec(:, modulo(i + 1, 3) + 1)) * 2.0_dp * acos(-1.0_dp) / vol
    end do

    ! ---------------- &LWInfo, &HighSymPathLW (main.f90:135-177) ----------------
    Sigma = 0.0_dp
    qmesh = 0
    num_points = 0
    open(newunit=u, file=trim(infile), status='old', action='read', delim='apostrophe', iostat=err, iomsg=msg)
    if ( err /= 0 ) call fail('opening namelist file '//trim(infile)//': '//trim(msg))
    read(u, nml=LWInfo, iostat=err, iomsg=msg)
    if ( err /= 0 ) call fail('reading &LWInfo of '//trim(infile)//': '//trim(msg))
    if ( num_points < 2 ) call fail('&LWInfo: num_points must be at least 2 (the path''s nodes)')
    if ( any(qmesh < 1) ) call fail('&LWInfo: qmesh must be three positive integers')
    if ( gauss .and. Sigma <= 0.0_dp ) call fail('-delta gauss needs Sigma > 0 in &LWInfo')
    allocate( q_high_sym(4, num_points) )
    q_high_sym = 0.0_dp
    rewind(u)
    read(u, nml=HighSymPathLW, iostat=err, iomsg=msg)
    if ( err /= 0 ) call fail('reading &HighSymPathLW of '//trim(infile)//': '//trim(msg))
    close(u)
    if ( int(q_high_sym(4, 1)) /= 0 ) write(*, '(A)') &
        & 'WARNING: the point count of the first node of &HighSymPathLW is not used'

    call q_points_highsymm(q_high_sym, G, q0, dist_arr, ZeroCentered)
    Npath = size(dist_arr)
    if ( Npath == 0 ) call fail('&HighSymPathLW gives no path points')

    write(Temp_char, '(F6.1)') T
    fc2file = 'FC2nd_'//trim(prefix)//trim(adjustl(Temp_char))//'K_F.h5'
    fc3file = 'FC3rd_'//trim(prefix)//trim(adjustl(Temp_char))//'K_F.h5'
    outfile_ascii = 'linewidth_'//trim(prefix)//trim(adjustl(Temp_char))//'K.txt'
    write(Ndof_char, '(I4)') (2 * Ndof - 1)
    outfile_frmt = "(G20.12,'   ', "//trim(adjustl(Ndof_char))//"(G20.12, '  '), G20.12)"

    call r2ndFC_h5(trim(fc2file), atmNum2, atmIndx2, fC2)
    call r3rdFC_h5(trim(fc3file), atmNum3, atmIndx3, fC3)
    if ( size(atmNum2) /= Nbasis .or. size(atmNum3) /= Nbasis ) &
        & call fail('the FC files do not have Nbasis basis atoms: '//trim(fc2file)//', '//trim(fc3file))

    ! ---------------- q1 mesh and phonons (main.f90:187-235) ----------------
    call MakeMeshIrrQ_gpu(Latvec, atm_pos, typ_indx, qmesh, 0, 1.0e-5_dp, q_pnt_int, q_pnt, irr_grid, pos, &
                        & mult, Rxyz, tet)
    Nq = product(qmesh)
    Nq0 = size(irr_grid)
    allocate( Evec(Ndof, Ndof, Nq), omega(Ndof, Nq), OnebyOmega(Ndof, Nq), nBE(Ndof, Nq), grp_vel(3, Ndof, Nq) )
    call get_phonon_gpu(Nbasis, mass, Latvec, atmNum2, atmIndx2, fC2, q_pnt, T, Evec, omega, OnebyOmega, &
                      & nBE, grp_vel, 0)
    allocate( Evec0(Ndof, Ndof, Npath), omega0(Ndof, Npath), OnebyOmega0(Ndof, Npath), nBE0(Ndof, Npath), &
            & grp_vel0(3, Ndof, Npath) )
    call get_phonon_gpu(Nbasis, mass, Latvec, atmNum2, atmIndx2, fC2, q0, T, Evec0, omega0, OnebyOmega0, &
                      & nBE0, grp_vel0, 0)
    call gpu_ctx_create(ctx, Nbasis, mass, vol, atmNum3, atmIndx3, fC3, qmesh, q_pnt_int, irr_grid, pos, &
                      & mult, Rxyz, Evec, omega, OnebyOmega, nBE, grp_vel, 0, tetrahedrons=tet)
    write(*, '(A, I6, A, 2(I4, A), I4, A)') 'Linewidth of ', Npath, ' path points over the ( ', qmesh(1), ' X', &
        & qmesh(2), ' X', qmesh(3), ' ) q1 mesh'

    ! ---------------- path loop (main.f90:293-379) ----------------
    allocate( lw(Ndof) )
    HeaderWrite = .true.
    do i0 = 1, Npath
        if ( gauss ) then
            call CalculateLwOffMesh_gpu(ctx, q0(:, i0), Evec0(:, :, i0), omega0(:, i0), OnebyOmega0(:, i0), Nbasis, &
                                      & mass, Latvec, atmNum2, atmIndx2, fC2, T, ZeroCentered(i0), lw, Sigma=Sigma)
        else
            call CalculateLwOffMesh_gpu(ctx, q0(:, i0), Evec0(:, :, i0), omega0(:, i0), OnebyOmega0(:, i0), Nbasis, &
                                      & mass, Latvec, atmNum2, atmIndx2, fC2, T, ZeroCentered(i0), lw)
        end if
        call WritePathRow(trim(outfile_ascii), trim(outfile_frmt), dist_arr(i0), omega0(:, i0), lw, Ndof, HeaderWrite)
    end do
    call gpu_ctx_destroy(ctx)

    call system_clock(t_end)
    write(*, '(A, A)') 'Linewidths written in file: ', trim(outfile_ascii)
    write(*, '(10X, "|", 23X, "      Phonon Linewidth: ", F10.4, " min", 24X, "|")') &
        & dble(t_end - t_start) / dble(t_rate) / 60.0_dp

contains

    ! q_points_highsymm (CreateMesh_mod.f90:85-172); q_high_sym(1:3, n) reduced, q_high_sym(4, n)
    ! the number of points of the segment that ends at node n
    subroutine q_points_highsymm(q_high_sym, G, qpoints, dist_arr, ZeroCenteredMesh)
        real(dp), intent(in)                :: q_high_sym(:,:), G(3, 3)
        real(dp), allocatable, intent(out)  :: qpoints(:,:), dist_arr(:)
        logical, allocatable, intent(out)   :: ZeroCenteredMesh(:)
        real(dp)                            :: s(3), e(3), v_hat(3), v_len, step_len, sd, dist
        integer                             :: nn, n, k, cnt, row
        logical                             :: zc
        nn = size(q_high_sym, 2)
        row = 0
        do n = 2, nn
            row = row + int(q_high_sym(4, n))
        end do
        allocate( qpoints(3, row), dist_arr(row), ZeroCenteredMesh(row) )
        row = 0
        sd = 0.0_dp
        dist = 0.0_dp
        do n = 1, nn - 1
            zc = any(q_high_sym(1:3, n) < 0.0_dp) .or. any(q_high_sym(1:3, n + 1) < 0.0_dp) .or. &
               & all(dabs(q_high_sym(1:3, n)) <= 0.5_dp) .or. all(dabs(q_high_sym(1:3, n + 1)) <= 0.5_dp)
            s = matmul(G, q_high_sym(1:3, n))
            e = matmul(G, q_high_sym(1:3, n + 1))
            cnt = int(q_high_sym(4, n + 1))
            v_len = norm2(e - s)
            step_len = 0.0_dp
            if ( cnt /= 0 ) step_len = v_len / cnt
            if ( cnt /= 0 .and. v_len == 0.0_dp ) call fail('&HighSymPathLW: a segment with points has equal ends')
            if ( cnt /= 0 ) v_hat = (e - s) / v_len
            do k = 1, cnt
                dist = sd + k * step_len
                row = row + 1
                qpoints(:, row) = s + (k * step_len) * v_hat
                dist_arr(row) = dist
                ZeroCenteredMesh(row) = zc
            end do
            sd = dist
        end do
    end subroutine q_points_highsymm

    ! LineWidth's get_command_line (parse_cmd_line.f90): -inp, -T, -omegacut; plus -delta
    subroutine get_lw_command_line(infile, T, gauss)
        character(len=*), intent(out)   :: infile
        real(dp), intent(out)           :: T
        logical, intent(out)            :: gauss
        character(len=512)              :: name, val
        integer                         :: n, k
        infile = 'input.nml'
        T = 300.0_dp
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
            case ('-delta')
                gauss = trim(val) == 'gauss'
            case default
                write(*, '(A, A)') 'WARNING: Unrecognised command name -> ', trim(name)
            end select
            k = k + 2
        end do
        write(*, '(A, A)') 'Input file to be read: ', trim(infile)
        write(*, '(A, F7.2, A)') 'Temperature: ', T, 'K'
    end subroutine get_lw_command_line

    ! WriteAsciiOutFile of LineWidth (Helper_m.f90:167-239): header on the first call, then one
    ! row per path point (distance, frequencies, linewidths)
    subroutine WritePathRow(filename, write_frmt, dist, freq, LW_q0, Ndof, HeaderWrite)
        character(len=*), intent(in)    :: filename, write_frmt
        real(dp), intent(in)            :: dist, freq(:), LW_q0(:)
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
        write(ui, write_frmt) dist, freq, LW_q0
        close(ui)
        20 FORMAT( '#        distance        ', &
                 & '--------- phonon_frequency(rad.THz) ( ', I4, ' columns) ---------      ', &
                 & '--------- linewidth(rad.THz) ( ', I4, ' columns) ---------')
    end subroutine WritePathRow

end program linewidth_gpu
