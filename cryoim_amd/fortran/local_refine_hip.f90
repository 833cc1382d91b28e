! local_refine on the MI355X library: the same key=value command line
! (local_refine/SRC/0input_parameters.f90, 0main.f90), the same local
! ortfile (load_original_ort.f90) and the same output records (format 901).
!   default:       refine_PR.f90 -- symmetry-equivalent pick (sym /= c1)
!                  and Ncycle stencil cycles of central slices against each
!                  sub-particle (cim_lref_refine)
!   Zhigh_df=y:    refine_PR_Zhigh.f90 -- the same under the CTF at
!                  df -/+ local_z0 (sign), always applied
!   particle_df=y: refine_particle_df.f90 -- 401 defoci local_z0 + 4 (i-200) A
!                  scanned at the starting orientation (cim_lref_defocus)
! The model (model3d) is soft-masked and transformed once on the device
! (cim_lref_model); sub-particles are cut from the whole images of the
! particle stacks (or imgstck) around (local_hx0, local_ky0).
program local_refine_hip
    use iso_c_binding
    use cim_iface
    implicit none
    type lpart_t
        character(len=256) :: stck
        integer :: pid, imgid, imgpid, limgid, hx0, ky0
        real :: theta, phi, omega, x, y, pr, df1, df2, astig, lth, lph, lom, lx, ly, lz, lpr
    end type
    character(len=64) :: coeff(22)
    character(len=256) :: arg, key, val, ortfile = '', newort = '', model3d = '', imgstck = '', stck0
    character(len=8) :: sym = 'c1', sgn = ''
    character :: applyctf = 'y', pdf = 'n', zhigh = 'n'
    integer :: nargs, i, eq, first = 1, last = -1, fftsize, fftsize0, total, u, ncycle = 12
    real :: apix = 1.0, maxres = 10.0, minres = 200.0, imgmask = 0.0, vol = 200.0, cs = 2.7, ampwg = 0.1
    real :: refine_angle = 1.0, refine_shift = 1.0, convf = 0.8
    logical :: have_mask = .false., have_last = .false., use_stck = .false., have_ang = .false.
    logical :: have_shift = .false.
    type(lpart_t), allocatable :: part(:)
    type(c_ptr) :: ctx
    type(cim_params) :: p
    integer(c_int) :: rc

    coeff = [character(len=64) :: 'local_ort0', 'local_newort', 'model3d', 'imgstck', 'minRes', 'apix', 'first', &
             'last', 'imgmask', 'applyCTF', 'Ncycle', 'vol', 'Cs', 'refine_angle', 'refine_shift', 'maxRes', &
             'convF', 'particle_df', 'Zhigh_df', 'sym', 'sign', 'ampwg']
    nargs = command_argument_count()
    if (nargs < 5) then
        print *, 'usage: local_refine_hip local_ort0=... local_newort=... model3d=... [key=value ...]'
        print *, '       (the keys of local_refine)'
        stop
    end if
    do i = 1, nargs
        call get_command_argument(i, arg)
        eq = index(arg, '=')
        key = arg(1:max(eq - 1, 0))
        val = arg(eq + 1:)
        if (.not. any(coeff == key)) then
            print *, 'unknown parameter: ', trim(arg)
            stop 2
        end if
        select case (trim(key))
        case ('local_ort0'); ortfile = val
        case ('local_newort'); newort = val
        case ('model3d'); model3d = val
        case ('imgstck'); imgstck = val; use_stck = .true.
        case ('minRes'); read(val, *) minres
        case ('apix'); read(val, *) apix
        case ('first'); read(val, *) first
        case ('last'); read(val, *) last; have_last = .true.
        case ('imgmask'); read(val, *) imgmask; have_mask = .true.
        case ('applyCTF'); applyctf = val(1:1)
        case ('Ncycle'); read(val, *) ncycle
        case ('vol'); read(val, *) vol
        case ('Cs'); read(val, *) cs
        case ('refine_angle'); read(val, *) refine_angle; have_ang = .true.
        case ('refine_shift'); read(val, *) refine_shift; have_shift = .true.
        case ('maxRes'); read(val, *) maxres
        case ('convF'); read(val, *) convf
        case ('particle_df'); pdf = val(1:1)
        case ('Zhigh_df'); zhigh = val(1:1)
        case ('sym'); sym = val(1:8)
        case ('sign'); sgn = val(1:8)
        case ('ampwg'); read(val, *) ampwg
        end select
    end do
    if ((zhigh == 'y' .or. zhigh == 'Y' .or. pdf == 'y' .or. pdf == 'Y') .and. &
        trim(sgn) /= '+' .and. trim(sgn) /= '-') then
        ! refine_PR_Zhigh.f90:124-130 / refine_particle_df.f90:121-127 leave df unset otherwise
        print *, 'Zhigh_df and particle_df need sign=+ or sign=-'
        stop 2
    end if

    ! calc_initialvilues (0calc_initialvalue.f90:9-71)
    call mrc_dims(model3d, fftsize)
    if (use_stck) then
        call mrc_dims(imgstck, fftsize0)
    else
        open(newunit=u, file=trim(ortfile), status='old')
        read(u, *) stck0
        close(u)
        call mrc_dims(stck0, fftsize0)
    end if
    call load_local_ortfile()
    if (have_mask) then
        imgmask = imgmask / apix
    else
        imgmask = real(fftsize / 2 - 2)
    end if
    maxres = apix * real(fftsize) / maxres
    minres = apix * real(fftsize) / minres
    if (.not. have_ang) refine_angle = 360.0 / (maxres * sqrt(5.0) * 3.14159)
    if (.not. have_shift) refine_shift = real(fftsize) / (maxres * sqrt(5.0))

    rc = cim_init(ctx, 0)
    if (rc /= CIM_OK) then
        print *, 'cim_init failed: ', rc
        stop 1
    end if
    call run()
    rc = cim_finalize(ctx)

contains

    subroutine mrc_dims(fname, nx)
        character(len=*), intent(in) :: fname
        integer, intent(out) :: nx
        integer :: v
        open(newunit=v, file=trim(fname), form='unformatted', access='stream', status='old')
        read(v) nx
        close(v)
    end subroutine

    ! load_original_ort.f90:1-20 over the particle count of calc_initialvilues
    subroutine load_local_ortfile()
        integer :: v, ios, j
        character(len=512) :: line
        total = 0
        open(newunit=v, file=trim(ortfile), status='old')
        do
            read(v, '(a)', iostat=ios) line
            if (ios /= 0) exit
            if (len_trim(line) > 0) total = total + 1
        end do
        rewind(v)
        if (.not. have_last .or. last > total) last = total
        allocate(part(first:last))
        do j = 1, first - 1
            read(v, *)
        end do
        do j = first, last
            read(v, *) part(j)%stck, part(j)%pid, part(j)%theta, part(j)%phi, part(j)%omega, part(j)%x, &
                       part(j)%y, part(j)%pr, part(j)%imgid, part(j)%imgpid, part(j)%df1, part(j)%df2, &
                       part(j)%astig, part(j)%limgid, part(j)%hx0, part(j)%ky0, part(j)%lth, part(j)%lph, &
                       part(j)%lom, part(j)%lx, part(j)%ly, part(j)%lz, part(j)%lpr
        end do
        close(v)
    end subroutine

    subroutine run()
        integer, parameter :: B = 256
        real(c_float), allocatable :: model(:), stck(:), imgs(:), rec(:, :)
        character(len=256) :: cur
        integer :: i0, i1, j, kk, v, nimg, pos
        logical :: zh, scan
        zh = zhigh == 'y' .or. zhigh == 'Y'
        scan = .not. zh .and. (pdf == 'y' .or. pdf == 'Y')
        p%box = fftsize
        p%minR = 0
        p%maxR = fftsize / 2 - 2
        ! refine_PR applies the CTF when applyCTF=y; the Zhigh and particle_df variants always do
        p%apply_ctf = merge(1, 0, zh .or. scan .or. applyctf == 'y' .or. applyctf == 'Y')
        p%mask_on = 1                                 ! refine_PR.f90:128: img2d_mask always
        p%conj_out = 0                                ! local_refine's img2d_FFT keeps +Im
        p%apix = apix
        p%imgmask = imgmask
        p%vol = vol
        p%cs = cs
        p%ampwg = ampwg
        p%df_scale = 1.0                              ! defocus in Angstrom
        p%ncycle = ncycle
        p%delta_angle = refine_angle
        p%delta_trans = refine_shift
        ! refine_PR.f90:21-54: header, then FFTsize**3 reals; soft mask, checkerboard, 3D FFT
        allocate(model(int(fftsize, 8)**3))
        open(newunit=v, file=trim(model3d), form='unformatted', access='stream', status='old')
        read(v, pos=1025) model
        close(v)
        rc = cim_lref_model(ctx, p, model)
        call cim_check(ctx, rc, 'cim_lref_model')
        deallocate(model)
        allocate(imgs(int(fftsize0, 8) * fftsize0 * B), rec(CIM_PART_STRIDE, B))
        open(41, file=trim(newort))
        cur = ''
        pos = 0
        do i0 = first, last, B
            i1 = min(last, i0 + B - 1)
            do j = i0, i1
                kk = j - i0 + 1
                if (use_stck) then
                    if (.not. allocated(stck)) then
                        open(newunit=v, file=trim(imgstck), form='unformatted', access='stream', status='old')
                        read(v, pos=9) nimg
                        allocate(stck(int(fftsize0, 8) * fftsize0 * nimg))
                        read(v, pos=1025) stck
                        close(v)
                        pos = first - 1
                    end if
                    imgs(int(kk - 1, 8) * fftsize0 * fftsize0 + 1:int(kk, 8) * fftsize0 * fftsize0) = &
                        stck(int(pos, 8) * fftsize0 * fftsize0 + 1:int(pos + 1, 8) * fftsize0 * fftsize0)
                    pos = pos + 1
                else
                    if (trim(cur) /= trim(part(j)%stck)) then
                        if (allocated(stck)) deallocate(stck)
                        cur = part(j)%stck
                        open(newunit=v, file=trim(cur), form='unformatted', access='stream', status='old')
                        read(v, pos=9) nimg
                        allocate(stck(int(fftsize0, 8) * fftsize0 * nimg))
                        read(v, pos=1025) stck
                        close(v)
                    end if
                    imgs(int(kk - 1, 8) * fftsize0 * fftsize0 + 1:int(kk, 8) * fftsize0 * fftsize0) = &
                        stck(int(part(j)%imgpid - 1, 8) * fftsize0 * fftsize0 + 1: &
                             int(part(j)%imgpid, 8) * fftsize0 * fftsize0)
                end if
                rec(:, kk) = 0.0
                rec(P_THETA, kk) = part(j)%lth
                rec(P_PHI, kk) = part(j)%lph
                rec(P_OMEGA, kk) = part(j)%lom
                rec(P_X, kk) = part(j)%lx
                rec(P_Y, kk) = part(j)%ly
                rec(P_DF1, kk) = part(j)%df1
                rec(P_DF2, kk) = part(j)%df2
                if (zh) then
                    ! refine_PR_Zhigh.f90:124-130: df -/+ local_z0
                    if (trim(sgn) == '+') then
                        rec(P_DF1, kk) = part(j)%df1 + part(j)%lz
                        rec(P_DF2, kk) = part(j)%df2 + part(j)%lz
                    else
                        rec(P_DF1, kk) = part(j)%df1 - part(j)%lz
                        rec(P_DF2, kk) = part(j)%df2 - part(j)%lz
                    end if
                end if
                rec(P_ASTIG, kk) = part(j)%astig
                rec(P_BOXX, kk) = real(part(j)%hx0)
                rec(P_BOXY, kk) = real(part(j)%ky0)
                rec(P_Z, kk) = part(j)%lz
            end do
            if (scan) then
                rc = cim_lref_defocus(ctx, p, i1 - i0 + 1, imgs, fftsize0, rec, merge(1.0, -1.0, trim(sgn) == '+'), &
                                      maxres)
                call cim_check(ctx, rc, 'cim_lref_defocus')
            else
                rc = cim_lref_refine(ctx, p, i1 - i0 + 1, imgs, fftsize0, rec, trim(sym) // c_null_char, convf, &
                                     maxres)
                call cim_check(ctx, rc, 'cim_lref_refine')
            end if
            do j = i0, i1
                kk = j - i0 + 1
                call write_record(part(j), rec(:, kk), zh, scan)
            end do
        end do
        close(41)
    end subroutine

    ! format 901 of refine_PR.f90:249-256, with local_dz and local_dPR for
    ! refine_PR_Zhigh.f90:273-280 and refine_particle_df.f90:165-172;
    ! refine_particle_df writes the particle's own PR in the PR column
    subroutine write_record(pt, r, zh, scan)
        type(lpart_t), intent(in) :: pt
        real(c_float), intent(in) :: r(:)
        logical, intent(in) :: zh, scan
        real :: th, ph, om, x, y, z, pr
        character(len=64) :: s
        th = r(P_THETA); ph = r(P_PHI); om = r(P_OMEGA); x = r(P_X); y = r(P_Y); pr = r(P_PR)
        z = pt%lz
        if (scan) then
            th = pt%lth; ph = pt%lph; om = pt%lom; x = pt%lx; y = pt%ly; z = r(P_Z)
        end if
        s = pt%stck
        if (zh .or. scan) then
            write(41, 902) s, pt%pid, pt%theta, pt%phi, pt%omega, pt%x, pt%y, merge(pt%pr, pr, scan), pt%imgid, &
                pt%imgpid, pt%df1, pt%df2, pt%astig, pt%limgid, pt%hx0, pt%ky0, th, ph, om, x, y, z, pr, &
                th - pt%lth, ph - pt%lph, om - pt%lom, x - pt%lx, y - pt%ly, z - pt%lz, pr - pt%lpr
        else
            write(41, 901) s, pt%pid, pt%theta, pt%phi, pt%omega, pt%x, pt%y, pr, pt%imgid, pt%imgpid, &
                pt%df1, pt%df2, pt%astig, pt%limgid, pt%hx0, pt%ky0, th, ph, om, x, y, z, pr, &
                th - pt%lth, ph - pt%lph, om - pt%lom, x - pt%lx, y - pt%ly
        end if
901     format(a64, i9, 6f10.4, i11, i7, 2f15.4, f10.4, i9, 2i6, 5f10.4, f11.4, f10.4, 5f10.4)
902     format(a64, i9, 6f10.4, i11, i7, 2f15.4, f10.4, i9, 2i6, 5f10.4, f11.4, f10.4, 7f10.4)
    end subroutine

end program
