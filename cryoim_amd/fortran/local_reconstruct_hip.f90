! local_reconstruct on the MI355X library: the same key=value command line
! (local_reconstruct/SRC/0input_parameters.f90, 0main.f90), the same local
! ortfile (load_original_ortcent.f90) and the same outputs.
!   default:   reconstruct.f90 -- local_FFTsize^3 map (MRC, maphead layout)
!   subrec=y:  reconstruct_substep.f90 -- compact tnf sums with the header
!              (FFTsize, maxR, apix, imgmask, 0, 0)
! Each sub-particle is the local_FFTsize box cut around (local_hx0,
! local_ky0) from its full image (img2d_local_box.f90, periodic), inserted at
! its local orientation and shift with local_reconstruct's operators
! (cim_local_reconstruct_sym_ops); defocus in Angstrom, optionally moved by
! -/+ local_z (Zhigh_df=y, sign=-/+).
program local_reconstruct_hip
    use iso_c_binding
    use cim_iface
    implicit none
    type lpart_t
        character(len=256) :: stck
        integer :: pid, imgid, imgpid, limgid, hx0, ky0
        real :: theta, phi, omega, x, y, pr, df1, df2, astig, lth, lph, lom, lx, ly, lz
        logical :: good
    end type
    character(len=64) :: coeff(21)
    character(len=256) :: arg, key, val, ortfile = '', result3d = 'local_map.mrc', imgstck = '', stck0
    character(len=4) :: sym = ''
    character :: applyctf = 'y', subrec = 'n', zhigh = 'n'
    character(len=8) :: sgn = ''
    integer :: nargs, i, eq, first = 1, last = -1, fftsize = 256, fftsize0, maxR, total, u
    real :: apix = 1.0, maxres = 10.0, imgmask = 0.0, vol = 200.0, cs = 2.7, ampwg = 0.1, bfactor = 0.0
    real :: pr_threshold = 90.0, boundx = 100.0, tilt = 90.0
    logical :: have_mask = .false., have_last = .false., use_stck = .false.
    type(lpart_t), allocatable :: part(:)
    type(c_ptr) :: ctx
    type(cim_params) :: p
    integer(c_int) :: rc, nops
    real(c_double), target :: ops(9 * 60)

    coeff = [character(len=64) :: 'local_ort', 'local_sym', 'local_FFTsize', 'apix', 'maxRes', 'imgmask', 'first', &
             'last', 'result3d', 'applyCTF', 'vol', 'Cs', 'Bfactor', 'ampwg', 'PR_threshold', 'subrec', &
             'imgstck', 'boundX', 'Zhigh_df', 'sign', 'tilt']
    nargs = command_argument_count()
    if (nargs < 1) then
        print *, 'usage: local_reconstruct_hip local_ort=... local_FFTsize=... [key=value ...] (local_reconstruct keys)'
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
        case ('local_ort'); ortfile = val
        case ('local_sym'); sym = val(1:4)
        case ('local_FFTsize'); read(val, *) fftsize
        case ('apix'); read(val, *) apix
        case ('maxRes'); read(val, *) maxres
        case ('imgmask'); read(val, *) imgmask; have_mask = imgmask > 0.0
        case ('first'); read(val, *) first
        case ('last'); read(val, *) last; have_last = .true.
        case ('result3d'); result3d = val
        case ('applyCTF'); applyctf = val(1:1)
        case ('vol'); read(val, *) vol
        case ('Cs'); read(val, *) cs
        case ('Bfactor'); read(val, *) bfactor        ! read and unused, as there
        case ('ampwg'); read(val, *) ampwg
        case ('PR_threshold'); read(val, *) pr_threshold
        case ('subrec'); subrec = val(1:1)
        case ('imgstck'); imgstck = val; use_stck = .true.
        case ('boundX'); read(val, *) boundx
        case ('Zhigh_df'); zhigh = val(1:1)
        case ('sign'); sgn = val
        case ('tilt'); read(val, *) tilt
        end select
    end do

    ! calc_initialvalues (0calc_initialvalues.f90:7-62)
    if (use_stck) then
        call mrc_dims(imgstck, fftsize0)
    else
        open(newunit=u, file=trim(ortfile), status='old')
        read(u, *) stck0
        close(u)
        call mrc_dims(stck0, fftsize0)
    end if
    if (have_mask) then
        imgmask = imgmask / apix
        if (imgmask >= real(fftsize) / 2.0) imgmask = real(fftsize) / 2.0 - 1.0
    else
        imgmask = real(fftsize) / 2.0 - 1.0
    end if
    maxres = apix * fftsize / maxres
    maxR = int(maxres)
    if (maxR >= fftsize / 2) maxR = fftsize / 2 - 2
    rc = cim_local_reconstruct_sym_ops(trim(sym) // c_null_char, ops, nops)
    if (rc /= CIM_OK) then
        ! a name 0calc_initialvalues does not know keeps Nsym = 1 and C_matrix(1)
        rc = cim_local_reconstruct_sym_ops('c1' // c_null_char, ops, nops)
    end if

    call load_local_ortfile()
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

    ! load_original_ortcent.f90:1-34: records first..last; quality 'g' when
    ! PR <= PR_threshold, |local_x|, |local_y| < boundX, |local_theta-90| < tilt
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
                       part(j)%lom, part(j)%lx, part(j)%ly, part(j)%lz
            part(j)%good = part(j)%pr <= pr_threshold .and. abs(part(j)%lx) < boundx .and. &
                           abs(part(j)%ly) < boundx .and. abs(part(j)%lth - 90.0) < tilt
        end do
        close(v)
    end subroutine

    ! reconstruct.f90:53-147 / reconstruct_substep.f90: batches of cut boxes
    subroutine run()
        integer, parameter :: B = 256
        real(c_float), allocatable :: full(:), imgs(:), rec(:, :), fa(:), fb(:), wg(:), map(:)
        real(c_float), allocatable :: stck(:)
        character(len=256) :: cur
        integer :: i0, i1, j, kk, v, nt, used, imgid0, pos, nimg
        real :: d1, d2, ast
        logical :: sub
        sub = subrec == 'y'                           ! 0main.f90:111 tests 'y' only
        p%box = fftsize
        p%maxR = maxR
        p%apply_ctf = merge(1, 0, applyctf == 'y' .or. applyctf == 'Y')
        p%mask_on = merge(1, 0, sub)                  ! only the substep masks (reconstruct_substep.f90:94)
        p%norm_on = merge(2, 1, sub)                  ! the substep masks before it normalises
        p%conj_out = 1
        p%apix = apix
        p%imgmask = imgmask
        p%vol = vol
        p%cs = cs
        p%ampwg = ampwg
        p%df_scale = 1.0                              ! defocus in Angstrom (getCTF2d.f90:14)
        rc = cim_rec_begin(ctx, fftsize, maxR, merge(1, 0, sub), real(merge(maxR - 1, maxR, sub), c_float), &
                           merge(0.1_c_float, 0.01_c_float, sub), nops, c_loc(ops))
        call cim_check(ctx, rc, 'cim_rec_begin')
        allocate(full(fftsize0 * fftsize0), imgs(fftsize * fftsize * B), rec(CIM_PART_STRIDE, B))
        cur = ''
        used = 0
        imgid0 = -100
        d1 = 0.0; d2 = 0.0; ast = 0.0
        pos = 0
        do i0 = first, last, B
            i1 = min(last, i0 + B - 1)
            do j = i0, i1
                kk = j - i0 + 1
                if (use_stck) then
                    ! imgstck: the images in order from particle first (reconstruct.f90:36-52)
                    if (.not. allocated(stck)) then
                        open(newunit=v, file=trim(imgstck), form='unformatted', access='stream', status='old')
                        read(v, pos=9) nimg
                        allocate(stck(int(fftsize0, 8) * fftsize0 * nimg))
                        read(v, pos=1025) stck
                        close(v)
                        pos = first - 1
                    end if
                    full = stck(int(pos, 8) * fftsize0 * fftsize0 + 1:int(pos + 1, 8) * fftsize0 * fftsize0)
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
                    full = stck(int(part(j)%imgpid - 1, 8) * fftsize0 * fftsize0 + 1: &
                                int(part(j)%imgpid, 8) * fftsize0 * fftsize0)
                end if
                call local_box(full, part(j)%hx0, part(j)%ky0, imgs((kk - 1) * fftsize * fftsize + 1:))
                rec(:, kk) = 0.0
                if (part(j)%good) then
                    used = used + 1
                    if (zhigh == 'y' .or. zhigh == 'Y') then
                        ! reconstruct.f90:72-83: df -/+ local_z, CTF per sub-particle
                        d1 = part(j)%df1; d2 = part(j)%df2
                        if (trim(sgn) == '-') then
                            d1 = d1 - part(j)%lz; d2 = d2 - part(j)%lz
                        else if (trim(sgn) == '+') then
                            d1 = d1 + part(j)%lz; d2 = d2 + part(j)%lz
                        end if
                        ast = part(j)%astig
                    else if (part(j)%imgid /= imgid0) then
                        d1 = part(j)%df1; d2 = part(j)%df2; ast = part(j)%astig
                        imgid0 = part(j)%imgid
                    end if
                end if
                rec(P_THETA, kk) = part(j)%lth
                rec(P_PHI, kk) = part(j)%lph
                rec(P_OMEGA, kk) = part(j)%lom
                rec(P_X, kk) = part(j)%lx
                rec(P_Y, kk) = part(j)%ly
                rec(P_PR, kk) = part(j)%pr
                rec(P_DF1, kk) = d1
                rec(P_DF2, kk) = d2
                rec(P_ASTIG, kk) = ast
                rec(P_MASKX, kk) = part(j)%lx
                rec(P_MASKY, kk) = part(j)%ly
                rec(P_USE, kk) = merge(1.0, 0.0, part(j)%good)
            end do
            rc = cim_rec_insert(ctx, p, i1 - i0 + 1, imgs, rec)
            call cim_check(ctx, rc, 'cim_rec_insert')
        end do
        print *, 'used sub-particles:', used
        if (sub) then
            nt = (2 * maxR + 1)**2 * (maxR + 1)
            allocate(fa(nt), fb(nt), wg(nt))
            rc = cim_rec_get_tnf(ctx, fa, fb, wg)
            call cim_check(ctx, rc, 'cim_rec_get_tnf')
            open(newunit=v, file=trim(result3d), form='unformatted', access='stream', status='replace')
            write(v) fftsize, maxR, apix, imgmask, 0, 0
            write(v) fa
            write(v) fb
            write(v) wg
            close(v)
        else
            allocate(map(int(fftsize, 8)**3))
            rc = cim_rec_finish(ctx, map)
            call cim_check(ctx, rc, 'cim_rec_finish')
            open(newunit=v, file=trim(result3d), form='unformatted', access='stream', status='replace')
            call write_maphead(v, fftsize, apix)
            write(v) map
            close(v)
        end if
    end subroutine

    ! img2d_local_box.f90:1-19: the n x n box centred on (ix, iy), periodic
    subroutine local_box(img, ix, iy, out)
        real(c_float), intent(in) :: img(:)
        integer, intent(in) :: ix, iy
        real(c_float), intent(inout) :: out(:)
        integer :: q, hx, ky
        do q = 1, fftsize * fftsize
            ky = (q - 1) / fftsize
            hx = mod(q - 1, fftsize)
            hx = hx + ix - fftsize / 2
            ky = ky + iy - fftsize / 2
            if (hx < 0) hx = hx + fftsize0
            if (hx >= fftsize0) hx = hx - fftsize0
            if (ky < 0) ky = ky + fftsize0
            if (ky >= fftsize0) ky = ky - fftsize0
            if (ky * fftsize0 + hx + 1 >= 1 .and. ky * fftsize0 + hx + 1 <= fftsize0 * fftsize0) &
                out(q) = img(ky * fftsize0 + hx + 1)
        end do
    end subroutine

    ! The 1024-byte header of local_reconstruct/SRC/mrchead.f90 for an n^3
    ! float map (single-character arrays hold the first character of each
    ! string assigned to them).
    subroutine write_maphead(v, n, ap)
        integer, intent(in) :: v, n
        real, intent(in) :: ap
        integer :: q
        write(v) n, n, n, 2, -n / 2, -n / 2, -n / 2, n - 1, n - 1, n - 1
        write(v) real(n - 1) * ap, real(n - 1) * ap, real(n - 1) * ap, 90.0, 90.0, 90.0
        write(v) 1, 2, 3, 0.0, 0.0, 0.0, 0, 0, (0, q = 1, 25)
        write(v) -1.0 * n, -1.0 * n, -1.0 * n
        write(v) 'mmmm', 'llll', 0.0, 0
        write(v) repeat('H', 800)
    end subroutine

end program
