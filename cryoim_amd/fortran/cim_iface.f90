! ISO_C_BINDING interface to libcryoim_hip.so (include/cryoim.h).
! A Fortran program that uses this module can replace the reference's
! loops (reconstruct.f90, ort_globesearch.f90, ort_refinement.f90,
! refine_PR.f90) with calls into the MI355X library.  Arrays are passed by
! address (first element), exactly the memory order the reference keeps:
! images a(box*box) with x fastest, particle records rec(12, n).
module cim_iface
    use iso_c_binding
    implicit none

    integer(c_int), parameter :: CIM_OK = 0
    integer(c_int), parameter :: CIM_PART_STRIDE = 12
    integer(c_int), parameter :: CIM_FSC_RECONSTRUCT = 0, CIM_FSC_MERGE = 1
    ! CIM_CENTER_MODEL / CIM_CENTER_SELFROT180 of cryoim.h (Fortran names are
    ! case-insensitive, so CIM_CENTER_MODEL would clash with cim_center_model)
    integer(c_int), parameter :: CIM_CTR_MODEL = 0, CIM_CTR_SELFROT180 = 1
    integer(c_int), parameter :: CIM_ICOSAVG_FULL = 0, CIM_ICOSAVG_MERGE = 1
    integer(c_int), parameter :: CIM_LBOX_REC = 8
    ! 1-based positions of the record fields in rec(:, i)
    integer, parameter :: P_THETA = 1, P_PHI = 2, P_OMEGA = 3, P_X = 4, P_Y = 5, P_PR = 6
    integer, parameter :: P_DF1 = 7, P_DF2 = 8, P_ASTIG = 9, P_MASKX = 10, P_MASKY = 11, P_USE = 12
    ! local_refine records: the sub-particle box centre and the height local_z0
    integer, parameter :: P_BOXX = 10, P_BOXY = 11, P_Z = 12

    type, bind(C) :: cim_params
        integer(c_int) :: box = 0, minR = 0, maxR = 0, apply_ctf = 1, mask_on = 1, norm_on = 0
        integer(c_int) :: ncycle = 6, conj_out = 1
        integer(c_int) :: search_method = 0, search_topk = 0, corr_cycles = 0, search_second = 0
        integer(c_int) :: reserved_i(4) = 0
        real(c_float) :: apix = 1.0, imgmask = 0.0, vol = 300.0, cs = 2.7, ampwg = 0.1, df_scale = 1.0e4
        real(c_float) :: searchstep = 2.0, delta_angle = 0.0, delta_trans = 0.0, corr_step = 0.0
        real(c_float) :: reserved_f(6) = 0.0
    end type

    interface
        integer(c_int) function cim_init(ctx, device) bind(C, name='cim_init')
            import :: c_ptr, c_int
            type(c_ptr) :: ctx
            integer(c_int), value :: device
        end function

        integer(c_int) function cim_finalize(ctx) bind(C, name='cim_finalize')
            import :: c_ptr, c_int
            type(c_ptr), value :: ctx
        end function

        type(c_ptr) function cim_last_error(ctx) bind(C, name='cim_last_error')
            import :: c_ptr
            type(c_ptr), value :: ctx
        end function

        integer(c_int) function cim_fft2d(ctx, re, im, n, batch, dir, conj_out) bind(C, name='cim_fft2d')
            import :: c_ptr, c_int, c_float
            type(c_ptr), value :: ctx
            real(c_float) :: re(*), im(*)
            integer(c_int), value :: n, batch, dir, conj_out
        end function

        integer(c_int) function cim_rec_begin(ctx, box, maxR, layout, rcut, wg_init, nops, ops) &
                bind(C, name='cim_rec_begin')
            import :: c_ptr, c_int, c_float
            type(c_ptr), value :: ctx, ops
            integer(c_int), value :: box, maxR, layout, nops
            real(c_float), value :: rcut, wg_init
        end function

        integer(c_int) function cim_rec_insert(ctx, p, n, images, parts) bind(C, name='cim_rec_insert')
            import :: c_ptr, c_int, c_float, cim_params
            type(c_ptr), value :: ctx
            type(cim_params) :: p
            integer(c_int), value :: n
            real(c_float) :: images(*), parts(*)
        end function

        integer(c_int) function cim_rec_get_tnf(ctx, fa, fb, wg) bind(C, name='cim_rec_get_tnf')
            import :: c_ptr, c_int, c_float
            type(c_ptr), value :: ctx
            real(c_float) :: fa(*), fb(*), wg(*)
        end function

        integer(c_int) function cim_rec_finish(ctx, map) bind(C, name='cim_rec_finish')
            import :: c_ptr, c_int, c_float
            type(c_ptr), value :: ctx
            real(c_float) :: map(*)
        end function

        integer(c_int) function cim_rec_add_tnf(ctx, fa, fb, wg) bind(C, name='cim_rec_add_tnf')
            import :: c_ptr, c_int, c_float
            type(c_ptr), value :: ctx
            real(c_float) :: fa(*), fb(*), wg(*)
        end function

        integer(c_int) function cim_merge_map(ctx, nicosavg, imgmask, map) bind(C, name='cim_merge_map')
            import :: c_ptr, c_int, c_float
            type(c_ptr), value :: ctx
            integer(c_int), value :: nicosavg
            real(c_float), value :: imgmask
            real(c_float) :: map(*)
        end function

        integer(c_int) function cim_fsc_map(ctx, box, maxR, imgmask, variant, nicosavg, ofa, ofb, owg, efa, efb, &
                                            ewg, fsc, map) bind(C, name='cim_fsc_map')
            import :: c_ptr, c_int, c_float
            type(c_ptr), value :: ctx
            integer(c_int), value :: box, maxR, variant, nicosavg
            real(c_float), value :: imgmask
            real(c_float) :: ofa(*), ofb(*), owg(*), efa(*), efb(*), ewg(*), fsc(*), map(*)
        end function

        integer(c_int) function cim_icosavg(ctx, map, len, n, r0, variant, nops, ops) bind(C, name='cim_icosavg')
            import :: c_ptr, c_int, c_float, c_long_long
            type(c_ptr), value :: ctx, ops
            real(c_float) :: map(*)
            integer(c_long_long), value :: len
            integer(c_int), value :: n, variant, nops
            real(c_float), value :: r0
        end function

        integer(c_int) function cim_search_templates(ctx, p, T, images, tort) bind(C, name='cim_search_templates')
            import :: c_ptr, c_int, c_float, cim_params
            type(c_ptr), value :: ctx
            type(cim_params) :: p
            integer(c_int), value :: T
            real(c_float) :: images(*), tort(*)
        end function

        integer(c_int) function cim_search_templates_model(ctx, p, T, model, tort) &
                bind(C, name='cim_search_templates_model')
            import :: c_ptr, c_int, c_float, cim_params
            type(c_ptr), value :: ctx
            type(cim_params) :: p
            integer(c_int), value :: T
            real(c_float) :: model(*), tort(*)
        end function

        integer(c_int) function cim_project_templates(ctx, box, T, model, angles, out) &
                bind(C, name='cim_project_templates')
            import :: c_ptr, c_int, c_float
            type(c_ptr), value :: ctx
            integer(c_int), value :: box, T
            real(c_float) :: model(*), angles(*), out(*)
        end function

        integer(c_int) function cim_search_global(ctx, p, n, images, parts) bind(C, name='cim_search_global')
            import :: c_ptr, c_int, c_float, cim_params
            type(c_ptr), value :: ctx
            type(cim_params) :: p
            integer(c_int), value :: n
            real(c_float) :: images(*), parts(*)
        end function

        integer(c_int) function cim_center_model(ctx, p, model) bind(C, name='cim_center_model')
            import :: c_ptr, c_int, c_float, cim_params
            type(c_ptr), value :: ctx
            type(cim_params) :: p
            real(c_float) :: model(*)
        end function

        integer(c_int) function cim_center(ctx, p, method, n, images, parts, cc) bind(C, name='cim_center')
            import :: c_ptr, c_int, c_float, cim_params
            type(c_ptr), value :: ctx
            type(cim_params) :: p
            integer(c_int), value :: method, n
            real(c_float) :: images(*), parts(*), cc(*)
        end function

        integer(c_int) function cim_refine_local(ctx, p, n, images, parts) bind(C, name='cim_refine_local')
            import :: c_ptr, c_int, c_float, cim_params
            type(c_ptr), value :: ctx
            type(cim_params) :: p
            integer(c_int), value :: n
            real(c_float) :: images(*), parts(*)
        end function

        integer(c_int) function cim_lref_model(ctx, p, vol) bind(C, name='cim_lref_model')
            import :: c_ptr, c_int, c_float, cim_params
            type(c_ptr), value :: ctx
            type(cim_params) :: p
            real(c_float) :: vol(*)
        end function

        integer(c_int) function cim_lref_refine(ctx, p, n, images, box0, parts, sym, convF, maxres) &
                bind(C, name='cim_lref_refine')
            import :: c_ptr, c_int, c_float, c_char, cim_params
            type(c_ptr), value :: ctx
            type(cim_params) :: p
            integer(c_int), value :: n, box0
            real(c_float) :: images(*), parts(*)
            character(kind=c_char) :: sym(*)
            real(c_float), value :: convF, maxres
        end function

        integer(c_int) function cim_lref_defocus(ctx, p, n, images, box0, parts, sign, maxres) &
                bind(C, name='cim_lref_defocus')
            import :: c_ptr, c_int, c_float, cim_params
            type(c_ptr), value :: ctx
            type(cim_params) :: p
            integer(c_int), value :: n, box0
            real(c_float) :: images(*), parts(*)
            real(c_float), value :: sign, maxres
        end function

        integer(c_int) function cim_local_reconstruct_sym_ops(sym, ops, nops) &
                bind(C, name='cim_local_reconstruct_sym_ops')
            import :: c_int, c_char, c_double
            character(kind=c_char) :: sym(*)
            real(c_double) :: ops(*)
            integer(c_int) :: nops
        end function

        integer(c_int) function cim_icosprocess_sym_ops(sym, ops, nops) bind(C, name='cim_icosprocess_sym_ops')
            import :: c_int, c_char, c_double
            character(kind=c_char) :: sym(*)
            real(c_double) :: ops(*)
            integer(c_int) :: nops
        end function

        integer(c_int) function cim_local_box_sym_ops(sym, ops, nops) bind(C, name='cim_local_box_sym_ops')
            import :: c_int, c_char, c_double
            character(kind=c_char) :: sym(*)
            real(c_double) :: ops(*)
            integer(c_int) :: nops
        end function

        integer(c_int) function cim_local_box_coords(ctx, n, parts, nops, ops, x0, box, apix, rec) &
                bind(C, name='cim_local_box_coords')
            import :: c_ptr, c_int, c_float, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: n, nops, box
            real(c_float) :: parts(*), x0(3), rec(*)
            real(c_double) :: ops(*)
            real(c_float), value :: apix
        end function

        integer(c_int) function cim_local_box_stack(ctx, n, images, parts, box, nops, ops, x0, sub, apix, boxes, rec) &
                bind(C, name='cim_local_box_stack')
            import :: c_ptr, c_int, c_float, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: n, box, nops, sub
            real(c_float) :: images(*), parts(*), x0(3), boxes(*), rec(*)
            real(c_double) :: ops(*)
            real(c_float), value :: apix
        end function

        function c_strlen(s) bind(C, name='strlen')
            import :: c_ptr, c_size_t
            type(c_ptr), value :: s
            integer(c_size_t) :: c_strlen
        end function
    end interface

contains

    ! Print the library's message for a failed call and stop, as the
    ! reference stops on unusable input (checkstckfile.f90:19-22).
    subroutine cim_check(ctx, rc, what)
        type(c_ptr), intent(in) :: ctx
        integer(c_int), intent(in) :: rc
        character(len=*), intent(in) :: what
        character(kind=c_char), pointer :: msg(:)
        type(c_ptr) :: m
        integer(c_size_t) :: l
        if (rc == CIM_OK) return
        m = cim_last_error(ctx)
        l = c_strlen(m)
        call c_f_pointer(m, msg, [l])
        print *, trim(what), ' failed (', rc, '): ', msg
        stop 1
    end subroutine

end module
