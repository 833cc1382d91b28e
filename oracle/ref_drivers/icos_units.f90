! Unit driver (own code): calls single icosprocess routines on inputs read
! from a binary file and writes their outputs to another binary file, so
! that golden vectors can be recorded.  Usage: icos_units <case> <in> <out>
program icos_units
    use common_icosproc
    use FFTW3SPACE
    implicit none
    character(len=256) :: what, fin, fout
    integer :: n, i, j, k, nx, mnR, mxR, T, nr, actf, nom
    real :: r(16), rotm(3,3,60), rm(3,3), th, ph, om, x0, y0, step
    real, allocatable :: a(:), b(:), c(:), al1(:), al2(:), fc(:), pro(:)

    call getarg(1, what); call getarg(2, fin); call getarg(3, fout)
    open(11, file=trim(fin), form='unformatted', access='stream', status='old')
    open(12, file=trim(fout), form='unformatted', access='stream', status='replace')
    select case (trim(what))
    case ('symops')
        call icos2f_sym_mat(rotm)
        write(12) (((rotm(i,j,k), j=1,3), i=1,3), k=1,60)
    case ('euler')
        read(11) n
        do k = 1, n
            read(11) th, ph, om
            call euler_mat(th, ph, om, rm)
            write(12) ((rm(i,j), j=1,3), i=1,3)
        end do
    case ('ctf')
        read(11) nx
        read(11) r(1:7)
        allocate(c(nx*nx))
        call getCTF2d(nx, r(1), r(2), r(3), r(4), r(5), r(6), r(7), c)
        write(12) c
    case ('mask')
        read(11) nx
        read(11) r(1:3)
        allocate(a(nx*nx)); read(11) a
        call img2d_mask(a, nx, r(1), r(2), r(3))
        write(12) a
    case ('norm')
        read(11) nx
        allocate(a(nx*nx)); read(11) a
        call img2d_norm(a, nx)
        write(12) a
    case ('fft2')
        read(11) nx
        allocate(a(nx*nx), b(nx*nx)); read(11) a; read(11) b
        call FFTNEW(4)
        call img2d_FFT(a, b, nx, nx, 1)
        write(12) a; write(12) b
        call img2d_FFT(a, b, nx, nx, -1)
        write(12) a; write(12) b
    case ('centshift')
        read(11) nx
        read(11) x0, y0
        allocate(a(nx*nx), b(nx*nx)); read(11) a; read(11) b
        call img2d_FT_centshift(a, b, nx, x0, y0)
        write(12) a; write(12) b
    case ('expfit')
        read(11) nx, mnR, mxR
        allocate(a(nx*nx), b(nx*nx), fc(mnR:mxR)); read(11) a; read(11) b
        call FT_amp_expfit(a, b, nx, mnR, mxR, fc)
        write(12) fc
    case ('crosslin')
        read(11) n
        allocate(al1(60), al2(60))
        do k = 1, n
            read(11) r(1:6)
            call icos_crosslin_angle(r(1), r(2), r(3), r(4), r(5), r(6), al1, al2)
            write(12) al1; write(12) al2
        end do
    case ('au')
        read(11) n
        do k = 1, n
            read(11) th, ph, om
            call return_icos_asymmetry_unit(th, ph, om)
            write(12) th, ph, om
        end do
    case ('pr_omega', 'pr')
        ! nx minR maxR T applyCTF ; step theta phi omega x y ; particle a b ;
        ! Fcut0 ; templates a b (T) ; template angles (T*3) ; template Fcut (T*nr) ; ctf
        read(11) nx, mnR, mxR, T, actf
        read(11) step, th, ph, om, x0, y0
        FFTsize = nx; minR = mnR; maxR = mxR; totaltemplate = T; Nlin = 60; searchstep = step
        proc2d%applyCTF = 'n'
        if (actf == 1) proc2d%applyCTF = 'y'
        nr = mxR - mnR + 1
        allocate(template(0:T), template2d(nx*nx, 2, 0:T), Fcut(mnR:mxR, 0:T), ctf2d(nx*nx))
        read(11) template2d(:,1,0); read(11) template2d(:,2,0)
        read(11) Fcut(:,0)
        do k = 1, T
            read(11) template2d(:,1,k); read(11) template2d(:,2,k)
        end do
        do k = 1, T
            read(11) template(k)%theta, template(k)%phi, template(k)%omega
        end do
        do k = 1, T
            read(11) Fcut(:,k)
        end do
        read(11) ctf2d
        template(0)%theta = th; template(0)%phi = ph; template(0)%omega = om
        template(0)%x = x0; template(0)%y = y0
        if (trim(what) == 'pr_omega') then
            nom = int(360.0/step)
            allocate(pro(0:nom-1))
            call cal_crosslin_PR_omega(nom, pro)
            write(12) pro
        else
            call cal_crosslin_PR(r(1))
            write(12) r(1)
        end if
    case ('project')
        ! project_template2d.f90 through the module state load_template sets up:
        ! FFTsize, totaltemplate, template(:) angles and the model3d file
        read(11) nx, T
        FFTsize = nx
        totaltemplate = T
        allocate(template(0:T), template2d(nx*nx, 2, 0:T))
        do k = 1, T
            read(11) th, ph, om
            template(k)%theta = th
            template(k)%phi = ph
            template(k)%omega = om
        end do
        allocate(a(nx*nx*nx)); read(11) a
        model3d = trim(fout) // '.model'
        open(13, file=trim(model3d), form='unformatted', access='stream', status='replace')
        write(13) (0, i = 1, 256)
        write(13) a
        close(13)
        call project_template2d
        do k = 1, T
            write(12) template2d(:, 1, k)
        end do
    case default
        print *, 'unknown case ', trim(what)
        stop 2
    end select
    close(11); close(12)
end program
