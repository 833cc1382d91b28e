! icosprocess_hip: the icosprocess command line (key=value parameters,
! ortfile / MRC / tnf files, icosprocess/SRC/0main.f90) driving the MI355X
! library through ISO_C_BINDING for
!   mode=1            3D reconstruction (reconstruct.f90; subrec=y writes
!                     the tnf partial volume of reconstruct_substep.f90)
!   mode=4            global orientation search (ort_globesearch.f90)
!   mode=3            centre finding: calc_center.f90 against model3d, or
!                     calc_center_selfrot180.f90 without it; output_ort.f90
!   mode=5            local orientation/centre refinement (ort_refinement.f90)
!   mode=6            local reconstruction of loc_box windows around (x0, y0)
!                     with a local symmetry (reconstruction_local.f90;
!                     subrec=y: reconstruction_local_substep.f90)
! Parameter keys, defaults and derived values follow input_parameters,
! initialvalue and calc_initialvalues.  Output files have the reference's
! formats, so the programs chained after icosprocess read them unchanged.
program icosprocess_hip
    use cim_iface
    implicit none
    integer, parameter :: NPAR = 32
    character(len=64) :: coeff(NPAR)
    character(len=256) :: arg, key, val
    character(len=256) :: ortfile = '', newortfile = 'newort.dat', result3d = 'map3d.mrc'
    character(len=256) :: templateort = '', templatestck = '', stck0, model3d = ''
    character(len=4) :: sym = 'icos'
    character :: applyctf = 'y', imgmask_on = 'y', subrec = 'n', realspaceavg = 'n'
    integer :: mode = 1, first = 1, last = 0, ncycle = 6, nargs, i, j, k, eq, fftsize, minR, maxR, T
    integer :: loc_box = -1
    real :: apix = 1.0, minres = 300.0, maxres = 8.0, imgmask = -1.0, vol = 300.0, cs = 2.7, ampwg = 0.1
    real :: pr_threshold = 90.0, searchstep = 2.0, boundx = -1.0, deltaangle, deltatrans, maxcentshift = 100.0
    logical :: have_mask = .false., have_boundx = .false.

    type particle_t
        character(len=54) :: stck
        integer :: pid, imgid, imgpid
        real :: theta, phi, omega, x, y, pr, df1, df2, astig, x0 = 0.0, y0 = 0.0
    end type
    type(particle_t), allocatable :: part(:)
    type(c_ptr) :: ctx
    type(cim_params) :: p
    integer :: nx, ny, nz, rc

    coeff = [character(len=64) :: 'apix', 'minRes', 'maxRes', 'imgmask', 'first', 'last', 'applyCTF', 'Ncycle', &
             'templatestck', 'vol', 'Cs', 'Bfactor', 'ampwg', 'PR_threshold', 'result3d', 'mode', 'newortfile', &
             'model3d', 'templateort', 'searchstep', 'realspaceavg', 'ortfile', 'recISAF', 'subrec', 'FSC', &
             'imgstck', 'sym', 'maxcentshift', 'corrCTF', 'boundX', 'loc_box', 'check']
    nargs = command_argument_count()
    if (nargs < 2) then
        print *, 'usage: icosprocess_hip mode=<1|3|4|5|6> ortfile=... [key=value ...] (icosprocess keys)'
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
        ! the reference compares several keys against a different spelling
        ! when it reads the value (Bfactor, FSC, corrCTF, recISAF), so those
        ! are accepted and ignored, as there (0main.f90:122-174)
        select case (trim(key))
        case ('mode'); read(val, *) mode
        case ('ortfile'); ortfile = val
        case ('templateort'); templateort = val
        case ('templatestck'); templatestck = val
        case ('minRes'); read(val, *) minres
        case ('maxRes'); read(val, *) maxres
        case ('imgmask'); read(val, *) imgmask; have_mask = .true.
        case ('applyCTF'); applyctf = val(1:1)
        case ('apix'); read(val, *) apix
        case ('vol'); read(val, *) vol
        case ('Cs'); read(val, *) cs
        case ('ampwg'); read(val, *) ampwg
        case ('first'); read(val, *) first
        case ('last'); read(val, *) last
        case ('PR_threshold'); read(val, *) pr_threshold
        case ('Ncycle'); read(val, *) ncycle
        case ('result3d'); result3d = val
        case ('newortfile'); newortfile = val
        case ('searchstep'); read(val, *) searchstep
        case ('realspaceavg'); realspaceavg = val(1:1)
        case ('subrec'); subrec = val(1:1)
        case ('sym'); sym = val(1:4)
        case ('boundX'); read(val, *) boundx; have_boundx = .true.
        case ('maxcentshift'); read(val, *) maxcentshift
        case ('loc_box'); read(val, *) loc_box
        case ('model3d'); model3d = val
        end select
    end do
    if (mode /= 1 .and. mode /= 6 .and. trim(sym) /= 'icos') then
        print *, 'icosprocess_hip runs modes 3, 4, 5 with the icosahedral (sym=icos) symmetry only'
        stop 2
    end if

    ! mode 0 (0main.f90:69-86): the template stack projected from model3d (project2d.f90)
    if (mode == 0) then
        call run_project0()
        stop
    end if

    ! initialvalue: box from the first stack, particle count from the ortfile
    open(21, file=trim(ortfile), status='old')
    read(21, *) stck0
    rewind(21)
    k = 0
    do
        read(21, *, iostat=rc) arg
        if (rc /= 0) exit
        k = k + 1
    end do
    close(21)
    if (last <= 0 .or. last > k) last = k
    call mrc_dims(stck0, nx, ny, nz)
    fftsize = nx
    if (.not. have_mask) imgmask = real(fftsize / 2 - 2)
    if (.not. have_boundx) boundx = real(fftsize / 2 - 2)
    if (loc_box <= 0) loc_box = fftsize                ! initialvalue: loc_box = FFTsize
    ! calc_initialvalues (0calc_initialvalues.f90:4-21): mode 6 measures resolution on loc_box
    if (mode /= 6) then
        minres = apix * fftsize / minres
        maxres = apix * fftsize / maxres
    else
        minres = apix * loc_box / minres
        maxres = apix * loc_box / maxres
    end if
    minR = int(minres)
    maxR = int(maxres)
    if (mod(maxR, 2) /= 0) maxR = maxR + 1
    if (imgmask_on == 'y') imgmask = imgmask / apix
    deltaangle = 360.0 / (maxres * sqrt(5.0) * 3.14159)
    deltatrans = real(fftsize) / (maxres * sqrt(5.0))

    call load_particles()

    p%box = fftsize
    p%minR = minR
    p%maxR = maxR
    p%apply_ctf = merge(1, 0, applyctf == 'y' .or. applyctf == 'Y')
    p%mask_on = 1
    p%conj_out = 1
    p%apix = apix
    p%imgmask = imgmask
    p%vol = vol
    p%cs = cs
    p%ampwg = ampwg
    p%df_scale = 1.0e4
    p%searchstep = searchstep
    p%ncycle = ncycle
    p%delta_angle = deltaangle
    p%delta_trans = deltatrans

    rc = cim_init(ctx, 0)
    if (rc /= CIM_OK) then
        print *, 'no HIP device (cim_init returned', rc, ')'
        stop 1
    end if
    select case (mode)
    case (1)
        call run_reconstruct()
    case (3)
        call run_center()
    case (4, 5)
        call run_search()
    case (6)
        call run_local()
    case default
        print *, 'icosprocess_hip implements modes 0, 1, 3, 4, 5 and 6'
        stop 2
    end select
    rc = cim_finalize(ctx)

contains

    subroutine mrc_dims(fname, nx, ny, nz)
        character(len=*), intent(in) :: fname
        integer, intent(out) :: nx, ny, nz
        integer :: u
        open(newunit=u, file=trim(fname), form='unformatted', access='stream', status='old')
        read(u) nx, ny, nz
        close(u)
    end subroutine

    ! load_original_ortcent.f90:12-26 and checkstckfile.f90:6-24
    subroutine load_particles()
        integer :: u, i, n, ios
        allocate(part(first:last))
        open(newunit=u, file=trim(ortfile), status='old')
        do i = 1, first - 1
            read(u, *)
        end do
        do i = first, last
            if (mode /= 6) then
                read(u, *) part(i)%stck, part(i)%pid, part(i)%theta, part(i)%phi, part(i)%omega, part(i)%x, &
                           part(i)%y, part(i)%pr, part(i)%imgid, part(i)%imgpid, part(i)%df1, part(i)%df2, &
                           part(i)%astig
            else
                read(u, *) part(i)%stck, part(i)%pid, part(i)%theta, part(i)%phi, part(i)%omega, part(i)%x, &
                           part(i)%y, part(i)%pr, part(i)%imgid, part(i)%imgpid, part(i)%df1, part(i)%df2, &
                           part(i)%astig, part(i)%x0, part(i)%y0
            end if
        end do
        close(u)
        do i = first, last
            open(newunit=u, file=trim(part(i)%stck), form='unformatted', access='stream', status='old', iostat=ios)
            if (ios /= 0) then
                print *, 'stack file missing: ', trim(part(i)%stck)
                stop 1
            end if
            read(u) n, n, n
            close(u)
            if (part(i)%imgpid > n) then
                print *, 'stack ', trim(part(i)%stck), ' has ', n, ' images; particle asks for ', part(i)%imgpid
                stop 1
            end if
        end do
    end subroutine

    subroutine read_image(pt, img)
        type(particle_t), intent(in) :: pt
        real(c_float), intent(out) :: img(:)
        integer :: u
        integer(8) :: pos
        pos = 1024_8 + 4_8 * int(fftsize, 8)**2 * int(pt%imgpid - 1, 8) + 1_8
        open(newunit=u, file=trim(pt%stck), form='unformatted', access='stream', status='old')
        read(u, pos=pos) img
        close(u)
    end subroutine

    ! Records for a batch.  The CTF is recomputed only when imgID changes in
    ! reconstruct.f90:85-89 and ort_globesearch.f90:61-65, so the defocus of
    ! the first particle of each micrograph run is carried forward there.
    ! Particle records for first..i1 written from i0.  bx >= 0 applies the
    ! reconstruction filter PR <= PR_threshold, |x|, |y| <= bx; mcs >= 0 the
    ! filter PR <= PR_threshold, sqrt(x^2 + y^2) <= mcs of
    ! reconstruct_C1_substep.f90:70-71 (unused records get P_USE = 0).  The
    ! per-micrograph CTF cache follows only particles that pass, as the
    ! reconstruct loops update imgID0 inside that test (reconstruct.f90:77-90,
    ! reconstruction_local.f90:96-108).
    subroutine fill_records(i0, i1, rec, cache_ctf, mask_first, bx, mcs)
        integer, intent(in) :: i0, i1
        real(c_float), intent(out) :: rec(CIM_PART_STRIDE, *)
        logical, intent(in) :: cache_ctf, mask_first
        real, intent(in) :: bx, mcs
        integer :: i, k, imgid0
        real :: d1, d2, ast
        logical :: use
        imgid0 = -100
        d1 = 0.0; d2 = 0.0; ast = 0.0
        do i = first, i1
            use = .true.
            if (bx >= 0.0) use = part(i)%pr <= pr_threshold .and. abs(part(i)%x) <= bx .and. abs(part(i)%y) <= bx
            if (mcs >= 0.0) use = part(i)%pr <= pr_threshold .and. sqrt(part(i)%x**2 + part(i)%y**2) <= mcs
            if (use .and. (.not. cache_ctf .or. part(i)%imgid /= imgid0)) then
                d1 = part(i)%df1; d2 = part(i)%df2; ast = part(i)%astig
                imgid0 = part(i)%imgid
            end if
            if (i < i0) cycle
            k = i - i0 + 1
            rec(:, k) = 0.0
            rec(P_THETA, k) = part(i)%theta
            rec(P_PHI, k) = part(i)%phi
            rec(P_OMEGA, k) = part(i)%omega
            rec(P_X, k) = part(i)%x
            rec(P_Y, k) = part(i)%y
            rec(P_PR, k) = part(i)%pr
            rec(P_DF1, k) = d1
            rec(P_DF2, k) = d2
            rec(P_ASTIG, k) = ast
            if (mask_first) then
                rec(P_MASKX, k) = part(first)%x
                rec(P_MASKY, k) = part(first)%y
            else
                rec(P_MASKX, k) = part(i)%x
                rec(P_MASKY, k) = part(i)%y
            end if
            rec(P_USE, k) = merge(1.0, 0.0, use)
        end do
    end subroutine

    ! icosprocess mode 1.  sym = icos: reconstruct.f90 / reconstruct_substep.f90.
    ! Any other sym (0main.f90:199-205): reconstruct_C1.f90 -- the cyclic
    ! group's operators, mask at each particle's own shift -- or
    ! reconstruct_C1_substep.f90 -- the identity only, rcut maxR-1, the
    ! maxcentshift filter; both cache the CTF per micrograph.
    subroutine run_reconstruct()
        integer, parameter :: B = 256
        real(c_float), allocatable :: imgs(:), rec(:, :), fa(:), fb(:), wg(:), map(:)
        real(c_double), target :: ops(9 * 15)
        integer :: i0, i1, i, k, layout, nt, u, used
        integer(c_int) :: nops
        real :: bx, mcs, rcut
        logical :: sub, c1path, cache, mfirst
        type(c_ptr) :: optr
        sub = subrec == 'y' .or. subrec == 'Y'
        c1path = trim(sym) /= 'icos'
        bx = boundx
        mcs = -1.0
        nops = 60
        optr = c_null_ptr
        layout = merge(1, 0, sub)
        if (.not. c1path) then
            if (sub) bx = fftsize / 4.0               ! reconstruct_substep.f90:22 overrides boundX
            rcut = real(merge(maxR - 2, maxR, sub))
            cache = .not. sub
            mfirst = .true.
        else if (sub) then
            ops = 0.0_c_double
            ops(1) = 1.0_c_double; ops(5) = 1.0_c_double; ops(9) = 1.0_c_double
            nops = 1                                  ! do isym=1,1 (reconstruct_C1_substep.f90:109)
            optr = c_loc(ops)
            bx = -1.0
            mcs = maxcentshift
            rcut = real(maxR - 1)
            cache = .true.
            mfirst = .true.
        else
            rc = cim_icosprocess_sym_ops(trim(sym) // c_null_char, ops, nops)
            if (rc /= CIM_OK) then
                print *, 'mode 1 has no operators for sym=', trim(sym)
                stop 2
            end if
            optr = c_loc(ops)
            rcut = real(maxR)
            cache = .true.
            mfirst = .false.
        end if
        p%norm_on = 1
        p%mask_on = merge(1, 0, imgmask_on == 'y')
        p%maxR = maxR
        rc = cim_rec_begin(ctx, fftsize, maxR, layout, rcut, merge(0.0_c_float, 0.01_c_float, sub), nops, optr)
        call cim_check(ctx, rc, 'cim_rec_begin')
        allocate(imgs(fftsize * fftsize * B), rec(CIM_PART_STRIDE, B))
        used = 0
        do i0 = first, last, B
            i1 = min(last, i0 + B - 1)
            call fill_records(i0, i1, rec, cache, mfirst, bx, mcs)
            do i = i0, i1
                k = i - i0 + 1
                call read_image(part(i), imgs((k - 1) * fftsize * fftsize + 1:k * fftsize * fftsize))
                if (rec(P_USE, k) > 0.0) used = used + 1
            end do
            rc = cim_rec_insert(ctx, p, i1 - i0 + 1, imgs, rec)
            call cim_check(ctx, rc, 'cim_rec_insert')
        end do
        print *, 'inserted particles:', used
        if (sub) then
            nt = (2 * maxR + 1)**2 * (maxR + 1)
            allocate(fa(nt), fb(nt), wg(nt))
            rc = cim_rec_get_tnf(ctx, fa, fb, wg)
            call cim_check(ctx, rc, 'cim_rec_get_tnf')
            open(newunit=u, file=trim(result3d), form='unformatted', access='stream', status='replace')
            write(u) fftsize, maxR, apix, imgmask, 0, merge(1, 0, realspaceavg == 'y' .or. realspaceavg == 'Y')
            write(u) fa
            write(u) fb
            write(u) wg
            close(u)
        else
            allocate(map(int(fftsize, 8)**3))
            rc = cim_rec_finish(ctx, map)
            call cim_check(ctx, rc, 'cim_rec_finish')
            if (realspaceavg == 'y' .or. realspaceavg == 'Y') then
                ! reconstruct.f90:253-255, reconstruct_C1.f90:286-288: img3d_icosavg(fa,
                ! FFTsize, imgmask) over the icosahedral group or the cyclic group of sym
                if (c1path .and. (trim(sym) == 'c1' .or. trim(sym) == 'c7')) then
                    print *, 'img3d_icosavg has no operators for sym=', trim(sym)
                    stop 2
                end if
                rc = cim_icosavg(ctx, map, int(size(map), c_long_long), fftsize, imgmask, CIM_ICOSAVG_FULL, &
                                 merge(nops, 0, c1path), optr)
                call cim_check(ctx, rc, 'cim_icosavg')
            end if
            call write_map(result3d, map, fftsize)
        end if
    end subroutine

    ! icosprocess mode 6 (reconstruction_local.f90:65-176, reconstruction_local_substep.f90):
    ! the loc_box window around (int(x0) + FFTsize/2, int(y0) + FFTsize/2) of each
    ! particle, masked around its own shift, inserted with the local symmetry.
    subroutine run_local()
        integer, parameter :: B = 256
        real(c_float), allocatable :: img(:), imgs(:), rec(:, :), fa(:), fb(:), wg(:), map(:)
        real(c_double), target :: ops(9 * 15)
        integer(c_int) :: nops
        integer :: i0, i1, i, k, layout, nt, u, used, xc, yc, ix, iy, lb2
        logical :: sub
        sub = subrec == 'y'                           ! 0main.f90:254 tests 'y' only
        rc = cim_icosprocess_sym_ops(trim(sym) // c_null_char, ops, nops)
        if (trim(sym) == 'c7') rc = -1                ! reconstruction_local has no c7 branch
        if (rc /= CIM_OK) then
            print *, 'mode 6 has no local symmetry operators for sym=', trim(sym)
            stop 2
        end if
        layout = merge(1, 0, sub)
        p%box = loc_box
        p%maxR = maxR
        p%norm_on = 1
        p%mask_on = merge(1, 0, imgmask_on == 'y')
        rc = cim_rec_begin(ctx, loc_box, maxR, layout, real(merge(maxR - 1, maxR, sub), c_float), 0.01_c_float, &
                           nops, c_loc(ops))
        call cim_check(ctx, rc, 'cim_rec_begin')
        lb2 = loc_box / 2
        allocate(img(fftsize * fftsize), imgs(loc_box * loc_box * B), rec(CIM_PART_STRIDE, B))
        used = 0
        do i0 = first, last, B
            i1 = min(last, i0 + B - 1)
            call fill_records(i0, i1, rec, .true., .false., boundx, -1.0)
            do i = i0, i1
                k = i - i0 + 1
                call read_image(part(i), img)
                xc = int(part(i)%x0) + fftsize / 2
                yc = int(part(i)%y0) + fftsize / 2
                if (xc - lb2 < 0 .or. yc - lb2 < 0 .or. xc + lb2 > fftsize .or. yc + lb2 > fftsize) then
                    print *, 'particle ', i, ': loc_box window around (x0, y0) leaves the image'
                    stop 1
                end if
                do iy = 0, loc_box - 1
                    do ix = 0, loc_box - 1
                        imgs((k - 1) * loc_box * loc_box + iy * loc_box + ix + 1) = &
                            img((yc - lb2 + iy) * fftsize + (xc - lb2 + ix) + 1)
                    end do
                end do
                if (rec(P_USE, k) > 0.0) used = used + 1
            end do
            rc = cim_rec_insert(ctx, p, i1 - i0 + 1, imgs, rec)
            call cim_check(ctx, rc, 'cim_rec_insert')
        end do
        print *, 'usedparticle:', used
        if (sub) then
            nt = (2 * maxR + 1)**2 * (maxR + 1)
            allocate(fa(nt), fb(nt), wg(nt))
            rc = cim_rec_get_tnf(ctx, fa, fb, wg)
            call cim_check(ctx, rc, 'cim_rec_get_tnf')
            open(newunit=u, file=trim(result3d), form='unformatted', access='stream', status='replace')
            write(u) loc_box, maxR, apix, imgmask, 0, 0
            write(u) fa
            write(u) fb
            write(u) wg
            close(u)
        else
            allocate(map(int(loc_box, 8)**3))
            rc = cim_rec_finish(ctx, map)
            call cim_check(ctx, rc, 'cim_rec_finish')
            call write_map(result3d, map, loc_box)
        end if
    end subroutine

    ! MRC header of mrchead.f90:2-58 (maphead) / maphead2.f90 for an n^3 map;
    ! map(4), machinestamp(4) and label(10,80) are single-character arrays, so
    ! each holds the first character of the string assigned to it.
    subroutine write_map(fname, map, n)
        character(len=*), intent(in) :: fname
        real(c_float), intent(in) :: map(:)
        integer, intent(in) :: n
        integer(c_int32_t) :: h(256)
        integer :: u
        h = 0
        h(1:3) = n
        h(4) = 2
        h(5:7) = -n / 2
        h(8:10) = n - 1
        h(11) = transfer(real((n - 1) * apix, c_float), h(1))
        h(12) = h(11)
        h(13) = h(11)
        h(14:16) = transfer(90.0_c_float, h(1))
        h(17) = 1; h(18) = 2; h(19) = 3
        h(50:52) = transfer(real(-n, c_float), h(1))
        h(53) = transfer('mmmm', h(1))
        h(54) = transfer('llll', h(1))
        h(57:256) = transfer('HHHH', h(1))
        open(newunit=u, file=trim(fname), form='unformatted', access='stream', status='replace')
        write(u) h
        write(u) map
        close(u)
    end subroutine

    subroutine run_search()
        integer, parameter :: B = 256
        real(c_float), allocatable :: imgs(:), rec(:, :), timgs(:), tort(:, :), model(:)
        integer :: i0, i1, i, k, u, pid
        real :: ox
        ! template bank (load_template.f90:18-33)
        open(newunit=u, file=trim(templateort), status='old')
        T = 0
        do
            read(u, *, iostat=rc) pid, ox
            if (rc /= 0) exit
            T = T + 1
        end do
        rewind(u)
        allocate(tort(5, T), timgs(fftsize * fftsize * T))
        do i = 1, T
            read(u, *) pid, tort(1, i), tort(2, i), tort(3, i), tort(4, i), tort(5, i)
        end do
        close(u)
        p%mask_on = merge(1, 0, imgmask_on == 'y')
        if (len_trim(templatestck) > 0) then
            open(newunit=u, file=trim(templatestck), form='unformatted', access='stream', status='old')
            read(u, pos=1025) timgs
            close(u)
            rc = cim_search_templates(ctx, p, T, timgs, tort)
            call cim_check(ctx, rc, 'cim_search_templates')
        else
            ! load_template.f90:32-33: no templatestck, project the templates from model3d
            ! (project_template2d.f90:21-26: header, then FFTsize**3 reals)
            allocate(model(int(fftsize, 8)**3))
            open(newunit=u, file=trim(model3d), form='unformatted', access='stream', status='old')
            read(u, pos=1025) model
            close(u)
            rc = cim_search_templates_model(ctx, p, T, model, tort)
            call cim_check(ctx, rc, 'cim_search_templates_model')
            deallocate(model)
        end if
        allocate(imgs(fftsize * fftsize * B), rec(CIM_PART_STRIDE, B))
        open(41, file=trim(newortfile))
        do i0 = first, last, B
            i1 = min(last, i0 + B - 1)
            ! mode 4 caches the CTF per micrograph; mode 5 recomputes it per particle
            call fill_records(i0, i1, rec, mode == 4, .false., -1.0, -1.0)
            do i = i0, i1
                k = i - i0 + 1
                call read_image(part(i), imgs((k - 1) * fftsize * fftsize + 1:k * fftsize * fftsize))
            end do
            if (mode == 4) then
                rc = cim_search_global(ctx, p, i1 - i0 + 1, imgs, rec)
            else
                rc = cim_refine_local(ctx, p, i1 - i0 + 1, imgs, rec)
            end if
            call cim_check(ctx, rc, 'search')
            do i = i0, i1
                k = i - i0 + 1
                call write_record(part(i), rec(:, k), mode == 5)
            end do
        end do
        close(41)
    end subroutine

    ! icosprocess mode 3 (0main.f90:237-241): calc_center when model3d is
    ! given, calc_center_selfrot180 otherwise, then output_ort.f90:1-17
    ! (dtheta, dphi, domega stay 0 from load_original_ortcent.f90:7).
    ! project2d.f90:2-86: the model3d header with nz = the template count, then one
    ! real-space projection per templateort line
    subroutine run_project0()
        integer(c_int32_t) :: h(256)
        real(c_float), allocatable :: model(:), ang(:, :), out(:)
        integer :: u, i, T, pid, n
        real :: th, ph, om
        open(newunit=u, file=trim(model3d), form='unformatted', access='stream', status='old')
        read(u) h
        n = h(1)
        allocate(model(int(n, 8)**3))
        read(u) model
        close(u)
        open(newunit=u, file=trim(templateort), status='old')
        T = 0
        do
            read(u, *, iostat=rc) pid, th, ph, om
            if (rc /= 0) exit
            T = T + 1
        end do
        rewind(u)
        allocate(ang(3, T), out(int(n, 8) * n * T))
        do i = 1, T
            read(u, *) pid, ang(1, i), ang(2, i), ang(3, i)
        end do
        close(u)
        rc = cim_init(ctx, 0)
        if (rc /= CIM_OK) then
            print *, 'no HIP device (cim_init returned', rc, ')'
            stop 1
        end if
        rc = cim_project_templates(ctx, n, T, model, ang, out)
        call cim_check(ctx, rc, 'cim_project_templates')
        rc = cim_finalize(ctx)
        h(3) = T
        open(newunit=u, file=trim(templatestck), form='unformatted', access='stream', status='replace')
        write(u) h
        write(u) out
        close(u)
    end subroutine

    subroutine run_center()
        integer, parameter :: B = 1024
        real(c_float), allocatable :: imgs(:), rec(:, :), vol(:), cc(:)
        integer :: i0, i1, i, k, u
        integer(c_int) :: method
        if (len_trim(model3d) > 0) then
            method = CIM_CTR_MODEL
            allocate(vol(int(fftsize, 8)**3))
            open(newunit=u, file=trim(model3d), form='unformatted', access='stream', status='old')
            read(u, pos=1025) vol                     ! calc_center.f90:21-23: header, then FFTsize**3 reals
            close(u)
            rc = cim_center_model(ctx, p, vol)
            call cim_check(ctx, rc, 'cim_center_model')
            deallocate(vol)
        else
            method = CIM_CTR_SELFROT180
        end if
        allocate(imgs(fftsize * fftsize * B), rec(CIM_PART_STRIDE, B), cc(B))
        open(41, file=trim(newortfile))
        do i0 = first, last, B
            i1 = min(last, i0 + B - 1)
            call fill_records(i0, i1, rec, .true., .false., -1.0, -1.0)
            do i = i0, i1
                k = i - i0 + 1
                call read_image(part(i), imgs((k - 1) * fftsize * fftsize + 1:k * fftsize * fftsize))
            end do
            rc = cim_center(ctx, p, method, i1 - i0 + 1, imgs, rec, cc)
            call cim_check(ctx, rc, 'cim_center')
            do i = i0, i1
                k = i - i0 + 1
                print *, i, rec(P_X, k), rec(P_Y, k), cc(k)
                write(41, 900) part(i)%stck, part(i)%pid, part(i)%theta, part(i)%phi, part(i)%omega, rec(P_X, k), &
                               rec(P_Y, k), rec(P_PR, k), part(i)%imgid, part(i)%imgpid, part(i)%df1, part(i)%df2, &
                               part(i)%astig, 0.0, 0.0, 0.0, rec(P_X, k) - part(i)%x, rec(P_Y, k) - part(i)%y, &
                               rec(P_PR, k) - part(i)%pr
            end do
        end do
        close(41)
900     format(a54,i8,f10.4,f10.4,f10.4,f10.4,f10.4,f10.4,i12,i6,f10.4,f10.4,f10.4,f10.4,f10.4,f10.4,f10.4,f10.4,f10.4)
    end subroutine

    ! ort_globesearch.f90:109-139 / ort_refinement.f90:96-122
    subroutine write_record(pt, r, centres)
        type(particle_t), intent(in) :: pt
        real(c_float), intent(in) :: r(:)
        logical, intent(in) :: centres
        real :: th, ph, om, x, y, pr
        th = r(P_THETA); ph = r(P_PHI); om = r(P_OMEGA); pr = r(P_PR)
        x = pt%x; y = pt%y
        if (centres) then
            x = r(P_X); y = r(P_Y)
        end if
        write(41, 900) pt%stck, pt%pid, th, ph, om, x, y, pr, pt%imgid, pt%imgpid, pt%df1, pt%df2, pt%astig, &
                       th - pt%theta, ph - pt%phi, om - pt%omega, x - pt%x, y - pt%y, pr - pt%pr
900     format(a54,i8,f10.4,f10.4,f10.4,f10.4,f10.4,f10.4,i12,i6,f10.4,f10.4,f10.4,f10.4,f10.4,f10.4,f10.4,f10.4,f10.4)
    end subroutine

end program
