! merge_icosprocess_subrec on the MI355X library: the same command line
! (merge_subreconstruct result3d1 result3d2 ... map3dfile, help.f90:1-6),
! the same tnf inputs and the same outputs as
! merge_icosprocess_subrec/SRC/main.f90:1-31.
!   nFSC = 0 in the first tnf header: merge_reconstruct.f90 -- the partial
!     sums of every listed tnf file, folded and inverse-transformed, written
!     as the (FFTsize-8)^3 map (cim_merge_map).
!   nFSC = 1: merge_reconstruct_FSC.f90 -- the <name>oddtnf / <name>evetnf
!     pairs of every listed name summed, FSC.txt written, and the
!     FSC-weighted map (cim_fsc_map, CIM_FSC_MERGE).
! nicosavg = 1 in the header (realspaceavg=y) adds merge's real-space
! icosahedral average (img3d_icosavg.f90): the plain merge takes it, imgmask
! and apix from the last file read (merge_reconstruct.f90:26 re-reads each
! header), the FSC merge from the first (main.f90:21).
program merge_icosprocess_subrec_hip
    use iso_c_binding
    use cim_iface
    implicit none
    character(len=512), allocatable :: par(:)
    integer :: nargs, i, u, ios, fftsize, maxr, nfsc, nicosavg, f2, m2, m, r, ntnf, last_avg
    real :: apix, imgmask, last_apix, last_mask
    real(c_float), allocatable :: a0(:), b0(:), w0(:), a1(:), b1(:), w1(:), a2(:), b2(:), w2(:), map(:), fsc(:)
    type(c_ptr) :: ctx
    integer(c_int) :: rc

    nargs = command_argument_count()
    if (nargs < 2) then
        print *, 'Used to merge data set from icosprocess--subrec to a 3D map'
        print *, 'Usage: merge_subreconstruct result3d1 result3d2 ... map3dfile'
        stop
    end if
    allocate(par(nargs))
    do i = 1, nargs
        call get_command_argument(i, par(i))
    end do

    ! header of the first input: <name>oddtnf if it exists, else <name> (main.f90:18-21)
    open(newunit=u, file=trim(par(1))//'oddtnf', form='unformatted', access='stream', status='old', iostat=ios)
    if (ios /= 0) open(newunit=u, file=trim(par(1)), form='unformatted', access='stream', status='old')
    read(u) fftsize, maxr, apix, imgmask, nfsc, nicosavg
    close(u)
    print *, fftsize, maxr, apix, imgmask, nfsc, nicosavg
    if (nfsc /= 0 .and. nfsc /= 1) stop
    ntnf = (maxr + 1) * (2 * maxr + 1)**2
    m = fftsize - 8
    allocate(a0(ntnf), b0(ntnf), w0(ntnf), map(m**3))
    rc = cim_init(ctx, 0)
    if (rc /= CIM_OK) then
        print *, 'cim_init failed: ', rc
        stop 1
    end if

    if (nfsc == 0) then
        rc = cim_rec_begin(ctx, fftsize, maxr, 1, real(maxr, c_float), 0.01_c_float, 60, c_null_ptr)
        call cim_check(ctx, rc, 'cim_rec_begin')
        do i = 1, nargs - 1
            print *, 'reading data form ', trim(par(i))
            call read_tnf(trim(par(i)))
            rc = cim_rec_add_tnf(ctx, a0, b0, w0)
            call cim_check(ctx, rc, 'cim_rec_add_tnf')
        end do
        apix = last_apix
        rc = cim_merge_map(ctx, last_avg, last_mask, map)
        call cim_check(ctx, rc, 'cim_merge_map')
    else
        allocate(a1(ntnf), b1(ntnf), w1(ntnf), a2(ntnf), b2(ntnf), w2(ntnf), fsc(0:maxr))
        a1 = 0; b1 = 0; w1 = 0; a2 = 0; b2 = 0; w2 = 0
        do i = 1, nargs - 1
            call read_tnf(trim(par(i))//'oddtnf')
            a1 = a1 + a0; b1 = b1 + b0; w1 = w1 + w0
            call read_tnf(trim(par(i))//'evetnf')
            a2 = a2 + a0; b2 = b2 + b0; w2 = w2 + w0
        end do
        print *, 'calculating FSC...'
        rc = cim_fsc_map(ctx, fftsize, maxr, imgmask, CIM_FSC_MERGE, nicosavg, a1, b1, w1, a2, b2, w2, fsc, map)
        call cim_check(ctx, rc, 'cim_fsc_map')
        open(newunit=u, file='FSC.txt')
        do r = 0, maxr
            write(u, *) r, fsc(r), sqrt(2.0 * fsc(r) / (1.0 + fsc(r)))
        end do
        close(u)
    end if
    rc = cim_finalize(ctx)

    print *, 'writing 3D map...'
    open(newunit=u, file=trim(par(nargs)), form='unformatted', access='stream', status='replace')
    call write_maphead(u, m, apix)
    write(u) map
    close(u)

contains

    subroutine read_tnf(name)
        character(len=*), intent(in) :: name
        integer :: v, h1, h2, h3, h4
        real :: r1, r2
        open(newunit=v, file=name, form='unformatted', access='stream', status='old')
        read(v) h1, h2, r1, r2, h3, h4
        last_apix = r1
        last_mask = r2
        last_avg = h4
        if (h1 /= fftsize .or. h2 /= maxr) then
            print *, trim(name), ': FFTsize/maxR ', h1, h2, ' differ from the first file ', fftsize, maxr
            stop 1
        end if
        read(v) a0
        read(v) b0
        read(v) w0
        close(v)
    end subroutine

    ! The 1024-byte MRC header of merge_icosprocess_subrec/SRC/mrchead.f90:2-58
    ! for an n^3 float map (its single-character arrays repeat the first
    ! character of each assigned string).
    subroutine write_maphead(v, n, ap)
        integer, intent(in) :: v, n
        real, intent(in) :: ap
        integer :: k
        write(v) n, n, n, 2, -n / 2, -n / 2, -n / 2, n - 1, n - 1, n - 1
        write(v) real(n - 1) * ap, real(n - 1) * ap, real(n - 1) * ap, 90.0, 90.0, 90.0
        write(v) 1, 2, 3, 0.0, 0.0, 0.0, 0, 0, (0, k = 1, 25)
        write(v) -1.0 * n, -1.0 * n, -1.0 * n
        write(v) 'mmmm', 'llll', 0.0, 0
        write(v) repeat('H', 800)
    end subroutine

end program
