! Unit driver (own code) for single local_refine routines; same protocol
! as icos_units.  Usage: lref_units <case> <in> <out>
program lref_units
    use common_loc_refine
    implicit none
    character(len=256) :: what, fin, fout
    integer :: nx
    real :: th, ph, om, x0, y0, mnR, mxR, pr, msk
    real, allocatable :: fr(:), fi(:), a(:), b(:), a2(:), b2(:), rs(:,:,:), ang(:,:)
    integer :: kind, nn, i, j, k
    character(len=4) :: symin

    call getarg(1, what); call getarg(2, fin); call getarg(3, fout)
    open(11, file=trim(fin), form='unformatted', access='stream', status='old')
    open(12, file=trim(fout), form='unformatted', access='stream', status='replace')
    select case (trim(what))
    case ('slice')
        read(11) nx
        read(11) th, ph, om
        allocate(fr(nx**3), fi(nx**3), a(nx*nx), b(nx*nx))
        read(11) fr; read(11) fi
        call FFT3d_slice(fr, fi, nx, th, ph, om, a, b)
        write(12) a; write(12) b
    case ('prft')
        read(11) nx
        read(11) mnR, mxR
        allocate(a(nx*nx), b(nx*nx), a2(nx*nx), b2(nx*nx))
        read(11) a; read(11) b; read(11) a2; read(11) b2
        call cal_PR_FT(a, b, a2, b2, nx, mnR, mxR, pr)
        write(12) pr
    case ('ftphase')
        read(11) nx
        read(11) x0, y0
        allocate(a(nx*nx), b(nx*nx)); read(11) a; read(11) b
        call img2d_FTphase(a, b, nx, x0, y0)
        write(12) a; write(12) b
    case ('softmask')
        read(11) nx
        read(11) msk
        allocate(fr(nx**3)); read(11) fr
        call img3d_softmask(fr, nx, msk)
        write(12) fr
    case ('fft3')
        read(11) nx
        allocate(fr(nx**3), fi(nx**3)); read(11) fr; read(11) fi
        call img3d_FFT(fr, fi, nx, 1)
        write(12) fr; write(12) fi
    case ('symops')
        read(11) kind, nn
        allocate(rs(3,3,nn))
        if (kind == 0) call icos2f_sym_matc(rs, nn)
        if (kind == 1) call icos2f_sym_matasym(rs, nn)
        if (kind == 2) call icos_matrix_i2(rs, nn)
        if (kind == 3) call icos_matrix_i5(rs, nn)
        write(12) (((rs(i,j,k), j=1,3), i=1,3), k=1,nn)
    case ('equalort')
        read(11) kind, nn
        read(11) th, ph, om
        Nsym = nn
        sym = 'c5'
        if (kind == 2) sym = 'i2'
        if (kind == 3) sym = 'i5'
        allocate(ang(3, nn))
        call icos_equalort(th, ph, om, ang)
        write(12) ang
    case default
        print *, 'unknown case ', trim(what)
        stop 2
    end select
    close(11); close(12)
end program
