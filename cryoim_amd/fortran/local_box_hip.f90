! local_box on the MI355X library: the same key=value command line
! (local_box/SRC/0input_parameters.f90, 0main.f90), the same icosahedral
! ortfile (load_ort.f90) and the same outputs.
!   without local_imgstck: loc_box_coordinate.f90 -- local_ort gets one line
!       per good particle and operator: the particle's own fields (defocus
!       x 10000), the operator number, the integer corner and fraction of
!       the sub-particle centre in the image, the local orientation, its
!       height, PR;
!   local_imgstck=file: loc_box.f90 -- file gets the sub_FFTsize boxes of
!       every good particle and operator (header of local_maphead.f90 with
!       nz = Nsym * all particles of the ortfile), local_ort the short lines.
! Operators: the matrices of sym_matrix (three lines per matrix, one column
! per line; read_sym_matrix.f90) or cim_local_box_sym_ops(sym).  The key
! Nsym is accepted and has no effect, as there.
program local_box_hip
    use iso_c_binding
    use cim_iface
    implicit none
    type bpart_t
        character(len=64) :: stck
        integer :: pid, imgid, imgpid
        real :: theta, phi, omega, x, y, pr, df1, df2, astig
        logical :: good
    end type
    character(len=64) :: coeff(16)
    character(len=256) :: arg, key, val
    character(len=64) :: icos_ort = '', local_ort = '', local_imgstck = '', icos_imgstck = ''
    character(len=64) :: sym_matrix = 'sym_matrix.txt'
    character(len=4) :: sym = 'i2'
    integer :: nargs, i, eq, first = 1, last = 0, sub = 256, fftsize, total, nsym
    real :: apix = 1.0, pr_threshold = 90.0, boundx = 100.0
    real(c_float) :: x0(3) = 0.0
    logical :: have_last = .false., use_matrix = .false., use_istck = .false., use_lstck = .false.
    type(bpart_t), allocatable :: part(:)
    real(c_double), allocatable :: ops(:)
    type(c_ptr) :: ctx
    integer(c_int) :: rc, nops
    ! batches: particle records, their (particle, operator) records, images, boxes
    real(c_float), allocatable :: rec(:, :), lrec(:, :), imgs(:), boxes(:)
    integer, allocatable :: idx(:)
    integer :: B, nb, nwritten
    integer(8) :: b2, s2

    coeff = [character(len=64) :: 'icos_ort', 'local_ort', 'local_imgstck', 'local_x', 'local_y', 'local_z', &
             'apix', 'PR_threshold', 'sub_FFTsize', 'icos_imgstck', 'Nsym', 'sym_matrix', 'sym', 'first', 'last', &
             'boundX']
    nargs = command_argument_count()
    if (nargs < 4) then
        print *, 'usage: local_box_hip icos_ort=... local_ort=... local_x=... local_y=... local_z=... ', &
                 '[local_imgstck=... sub_FFTsize=... sym=... | sym_matrix=... apix= PR_threshold= boundX= first= last=]'
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
        case ('icos_ort'); read(val, *) icos_ort
        case ('local_ort'); read(val, *) local_ort
        case ('local_imgstck'); read(val, *) local_imgstck; use_lstck = .true.
        case ('local_x'); read(val, *) x0(1)
        case ('local_y'); read(val, *) x0(2)
        case ('local_z'); read(val, *) x0(3)
        case ('apix'); read(val, *) apix
        case ('PR_threshold'); read(val, *) pr_threshold
        case ('sub_FFTsize'); read(val, *) sub
        case ('icos_imgstck'); read(val, *) icos_imgstck; use_istck = .true.
        case ('sym'); read(val, *) sym
        case ('sym_matrix'); read(val, *) sym_matrix; use_matrix = .true.
        case ('first'); read(val, *) first
        case ('last'); read(val, *) last; have_last = .true.
        case ('boundX'); read(val, *) boundx
        end select
    end do

    call load_operators()
    call load_ort()
    rc = cim_init(ctx, 0)
    if (rc /= CIM_OK) then
        print *, 'cim_init failed: ', rc
        stop 1
    end if
    if (use_lstck) then
        call run_stack()
    else
        call run_coordinates()
    end if
    rc = cim_finalize(ctx)

contains

    ! 0main.f90:80-122: the operators of sym, or the matrices of sym_matrix
    ! (Nsym = the number of complete groups of three list-directed reads).
    subroutine load_operators()
        integer :: v, ios, k, q
        real :: a(3), b(3), c(3)
        if (.not. use_matrix) then
            allocate(ops(9 * 60))
            rc = cim_local_box_sym_ops(trim(sym) // c_null_char, ops, nops)
            nsym = nops
            return
        end if
        nsym = 0
        open(newunit=v, file=trim(sym_matrix), status='old')
        do
            read(v, *, iostat=ios) a
            if (ios /= 0) exit
            read(v, *, iostat=ios) b
            if (ios /= 0) exit
            read(v, *, iostat=ios) c
            if (ios /= 0) exit
            nsym = nsym + 1
        end do
        rewind(v)
        allocate(ops(9 * max(nsym, 1)))
        do k = 1, nsym
            do q = 1, 3
                read(v, *) a
                ! line q holds column q: ops row-major (row r, column q)
                ops(9 * (k - 1) + q) = a(1)
                ops(9 * (k - 1) + 3 + q) = a(2)
                ops(9 * (k - 1) + 6 + q) = a(3)
            end do
        end do
        close(v)
        nops = nsym
    end subroutine

    ! load_ort.f90:1-42: the records of icos_ort; status good when
    ! PR <= PR_threshold and |x|, |y| <= boundX; FFTsize from the header of
    ! the first record's stack.
    subroutine load_ort()
        integer :: v, ios, j, ip, ii, iq
        character(len=64) :: s
        real :: t(8)
        open(newunit=v, file=trim(icos_ort), status='old')
        total = 0
        do
            read(v, *, iostat=ios) s, ip, t(1:5), t(6), ii, iq, t(7), t(7), t(8)
            if (ios /= 0) exit
            total = total + 1
        end do
        rewind(v)
        print *, 'totalparticle=', total
        if (.not. have_last) last = total
        if (have_last .and. last > total) then
            print *, 'last should smaller than totalparticle'
            stop
        end if
        allocate(part(max(total, 1)))
        do j = 1, total
            read(v, *) part(j)%stck, part(j)%pid, part(j)%theta, part(j)%phi, part(j)%omega, part(j)%x, part(j)%y, &
                       part(j)%pr, part(j)%imgid, part(j)%imgpid, part(j)%df1, part(j)%df2, part(j)%astig
            part(j)%good = part(j)%pr <= pr_threshold .and. abs(part(j)%x) <= boundx .and. abs(part(j)%y) <= boundx
        end do
        close(v)
        open(newunit=v, file=trim(part(1)%stck), form='unformatted', access='stream', status='old')
        read(v) fftsize
        close(v)
    end subroutine

    subroutine fill_record(j, r)
        integer, intent(in) :: j
        real(c_float), intent(out) :: r(CIM_PART_STRIDE)
        r = 0.0
        r(P_THETA) = part(j)%theta
        r(P_PHI) = part(j)%phi
        r(P_OMEGA) = part(j)%omega
        r(P_X) = part(j)%x
        r(P_Y) = part(j)%y
    end subroutine

    ! loc_box_coordinate.f90:16-41, batches of good particles
    subroutine run_coordinates()
        integer :: v, j
        B = 4096
        allocate(rec(CIM_PART_STRIDE, B), lrec(CIM_LBOX_REC, B * max(nops, 1)), idx(B))
        open(newunit=v, file=trim(local_ort), status='replace')
        nb = 0
        do j = first, last
            if (.not. part(j)%good) cycle
            nb = nb + 1
            idx(nb) = j
            call fill_record(j, rec(:, nb))
            if (nb == B) call flush_coordinates(v)
        end do
        if (nb > 0) call flush_coordinates(v)
        close(v)
    end subroutine

    subroutine flush_coordinates(v)
        integer, intent(in) :: v
        integer :: k, j, ii
        if (nops > 0) then
            rc = cim_local_box_coords(ctx, nb, rec, nops, ops, x0, fftsize, apix, lrec)
            call cim_check(ctx, rc, 'cim_local_box_coords')
        end if
        do k = 1, nb
            j = idx(k)
            do ii = 1, nops
                associate (o => lrec(:, (k - 1) * nops + ii))
                    write(v, 900) part(j)%stck, part(j)%pid, part(j)%theta, part(j)%phi, part(j)%omega, &
                        part(j)%x, part(j)%y, part(j)%pr, part(j)%imgid, part(j)%imgpid, part(j)%df1 * 10000.0, &
                        part(j)%df2 * 10000.0, part(j)%astig, ii, int(o(4)), int(o(5)), o(1), o(2), o(3), o(6), &
                        o(7), o(8), part(j)%pr
                end associate
            end do
        end do
        nb = 0
900     format(a64,i8,f10.4,f10.4,f10.4,f10.4,f10.4,f10.4,i12,i6,f15.4,f15.4,f10.4,i9,i6,i6,f10.4,f10.4,f10.4, &
               f10.4,f10.4,f11.4,f11.4)
    end subroutine

    ! loc_box.f90:11-91: images of first..last (icos_imgstck in order from
    ! particle first, else image imgparticleID of the record's stack), the
    ! good ones in batches through cim_local_box_stack.
    subroutine run_stack()
        integer :: vs, vo, vi, vp, j
        integer(8) :: pos
        b2 = int(fftsize, 8) * fftsize
        s2 = int(sub, 8) * sub
        B = int(max(1_8, min(1024_8, (512_8 * 1024 * 1024) / (4 * (b2 + s2 * max(nops, 1))))))
        allocate(imgs(b2 * B), rec(CIM_PART_STRIDE, B), lrec(CIM_LBOX_REC, B * max(nops, 1)), idx(B))
        allocate(boxes(s2 * max(nops, 1) * B))
        open(newunit=vs, file=trim(local_imgstck), form='unformatted', access='stream', status='replace')
        call write_maphead(vs)
        open(newunit=vo, file=trim(local_ort), status='replace')
        if (use_istck) open(newunit=vi, file=trim(icos_imgstck), form='unformatted', access='stream', status='old')
        nwritten = 0
        nb = 0
        do j = first, last
            if (.not. part(j)%good) cycle
            nb = nb + 1
            idx(nb) = j
            if (use_istck) then
                pos = 1025_8 + 4_8 * b2 * (j - 1)
                read(vi, pos=pos) imgs((nb - 1) * b2 + 1:nb * b2)
            else
                open(newunit=vp, file=trim(part(j)%stck), form='unformatted', access='stream', status='old')
                pos = 1025_8 + 4_8 * b2 * (part(j)%imgpid - 1)
                read(vp, pos=pos) imgs((nb - 1) * b2 + 1:nb * b2)
                close(vp)
            end if
            call fill_record(j, rec(:, nb))
            if (nb == B) call flush_stack(vs, vo)
        end do
        if (nb > 0) call flush_stack(vs, vo)
        if (use_istck) close(vi)
        close(vs)
        close(vo)
        print *, 'sub-particles written:', nwritten
    end subroutine

    subroutine flush_stack(vs, vo)
        integer, intent(in) :: vs, vo
        integer :: k, j, ii
        if (nops > 0) then
            rc = cim_local_box_stack(ctx, nb, imgs, rec, fftsize, nops, ops, x0, sub, apix, boxes, lrec)
            call cim_check(ctx, rc, 'cim_local_box_stack')
            write(vs) boxes(1:s2 * nops * nb)
        end if
        do k = 1, nb
            j = idx(k)
            do ii = 1, nops
                nwritten = nwritten + 1
                associate (o => lrec(:, (k - 1) * nops + ii))
                    write(vo, 900) part(j)%stck, nwritten, o(1), o(2), o(3), o(6), o(7), part(j)%pr, &
                        part(j)%imgid, ii, part(j)%df1, part(j)%df2, part(j)%astig, o(8)
                end associate
            end do
        end do
        nb = 0
900     format(a64,i8,f10.4,f10.4,f10.4,f10.4,f10.4,f10.4,i12,i6,f10.4,f10.4,f10.4,f12.4)
    end subroutine

    ! local_maphead.f90:1-57 with nz = Nsym * totalparticle (loc_box.f90:24);
    ! single-character arrays hold the first character of each string
    ! assigned to them.
    subroutine write_maphead(v)
        integer, intent(in) :: v
        integer :: q
        write(v) sub, sub, nsym * total, 2, -sub / 2, -sub / 2, -sub / 2, sub - 1, sub - 1, sub - 1
        write(v) real(sub - 1) * apix, real(sub - 1) * apix, real(sub - 1) * apix, 90.0, 90.0, 90.0
        write(v) 1, 2, 3, 0.0, 0.0, 0.0, 0, 0, (0, q = 1, 25)
        write(v) -1.0 * sub, -1.0 * sub, -1.0 * sub
        write(v) 'mmmm', 'llll', 0.0, 0
        write(v) repeat('H', 800)
    end subroutine

end program
