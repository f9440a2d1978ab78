! ISO_C_BINDING interfaces to libsalfort (include/salfort.h) for the Fortran drop-in modules.
! Ids crossing this boundary are 0-based; the drop-in modules convert to the 1-based ids of the
! reference's types.  Every C call returns a status; sal_check turns a failure into a message and
! an error stop.
module salfort_capi
    use iso_c_binding
    implicit none

    interface
        function c_sal_last_error() bind(C, name="sal_last_error")
            import :: c_ptr
            type(c_ptr) :: c_sal_last_error
        end function

        integer(c_int) function c_sal_geom_create(x, y, z, n, thresh, g) bind(C, name="sal_geom_create")
            import :: c_int, c_int64_t, c_double, c_ptr
            real(c_double), intent(in) :: x(*), y(*), z(*)
            integer(c_int64_t), value :: n
            integer(c_int), value :: thresh
            type(c_ptr), intent(out) :: g
        end function

        integer(c_int) function c_sal_geom_destroy(g) bind(C, name="sal_geom_destroy")
            import :: c_int, c_ptr
            type(c_ptr), value :: g
        end function

        integer(c_int) function c_sal_geom_counts(g, np, maxlev, tot) bind(C, name="sal_geom_counts")
            import :: c_int, c_int64_t, c_ptr
            type(c_ptr), value :: g
            integer(c_int), intent(out) :: np, maxlev
            integer(c_int64_t), intent(out) :: tot
        end function

        integer(c_int) function c_sal_geom_export(g, level, is_leaf, parent, child4, face, bounds6, radius, &
                                                  cnt, offsets, pts) bind(C, name="sal_geom_export")
            import :: c_int, c_int64_t, c_double, c_ptr
            type(c_ptr), value :: g
            integer(c_int), intent(out) :: level(*), is_leaf(*), parent(*), child4(*), face(*), cnt(*), pts(*)
            real(c_double), intent(out) :: bounds6(*), radius(*)
            integer(c_int64_t), intent(out) :: offsets(*)
        end function

        integer(c_int) function c_sal_assign_points(g, x, y, z, n, nlev, pp) bind(C, name="sal_assign_points")
            import :: c_int, c_int64_t, c_double, c_ptr
            type(c_ptr), value :: g
            real(c_double), intent(in) :: x(*), y(*), z(*)
            integer(c_int64_t), value :: n
            integer(c_int), value :: nlev
            integer(c_int), intent(out) :: pp(*)
        end function

        integer(c_int) function c_sal_dtt(g, theta, thresh, n) bind(C, name="sal_dtt")
            import :: c_int, c_int64_t, c_double, c_ptr
            type(c_ptr), value :: g
            real(c_double), value :: theta
            integer(c_int), value :: thresh
            integer(c_int64_t), intent(out) :: n
        end function

        integer(c_int) function c_sal_dtt_export(g, it, is, ty) bind(C, name="sal_dtt_export")
            import :: c_int, c_ptr
            type(c_ptr), value :: g
            integer(c_int), intent(out) :: it(*), is(*), ty(*)
        end function

        integer(c_int) function c_sal_dtt_import(g, it, is, ty, n) bind(C, name="sal_dtt_import")
            import :: c_int, c_int64_t, c_ptr
            type(c_ptr), value :: g
            integer(c_int), intent(in) :: it(*), is(*), ty(*)
            integer(c_int64_t), value :: n
        end function

        integer(c_int) function c_sal_stt(g, theta, x, y, z, n, npc, npp) bind(C, name="sal_stt")
            import :: c_int, c_int64_t, c_double, c_ptr
            type(c_ptr), value :: g
            real(c_double), value :: theta
            real(c_double), intent(in) :: x(*), y(*), z(*)
            integer(c_int64_t), value :: n
            integer(c_int64_t), intent(out) :: npc, npp
        end function

        integer(c_int) function c_sal_stt_export(g, pct, pcs, ppt, pps) bind(C, name="sal_stt_export")
            import :: c_int, c_ptr
            type(c_ptr), value :: g
            integer(c_int), intent(out) :: pct(*), pcs(*), ppt(*), pps(*)
        end function

        integer(c_int) function c_sal_fast_sum(g, area, ssh, degree, gx, gy, acc) bind(C, name="sal_fast_sum")
            import :: c_int, c_double, c_ptr
            type(c_ptr), value :: g
            real(c_double), intent(in) :: area(*), ssh(*)
            integer(c_int), value :: degree, acc
            real(c_double), intent(inout) :: gx(*), gy(*)
        end function

        integer(c_int) function c_sal_interact(g, kind, it, is, degree, area, ssh, gx, gy, acc) &
                bind(C, name="sal_interact")
            import :: c_int, c_double, c_ptr
            type(c_ptr), value :: g
            integer(c_int), value :: kind, it, is, degree, acc
            real(c_double), intent(in) :: area(*), ssh(*)
            real(c_double), intent(inout) :: gx(*), gy(*)
        end function

        integer(c_int) function c_sal_mom6_proxy_source(g, area, ssh, degree, w) bind(C, name="sal_mom6_proxy_source")
            import :: c_int, c_double, c_ptr
            type(c_ptr), value :: g
            real(c_double), intent(in) :: area(*), ssh(*)
            integer(c_int), value :: degree
            real(c_double), intent(out) :: w(*)
        end function

        integer(c_int) function c_sal_mom6_pc_compute(g, w, degree, transposed, gx, gy, acc) &
                bind(C, name="sal_mom6_pc_compute")
            import :: c_int, c_double, c_ptr
            type(c_ptr), value :: g
            real(c_double), intent(in) :: w(*)
            integer(c_int), value :: degree, transposed, acc
            real(c_double), intent(inout) :: gx(*), gy(*)
        end function

        integer(c_int) function c_sal_mom6_interact(g, kind, x, y, z, source, degree, transposed, w, area, ssh, &
                                                    gx, gy) bind(C, name="sal_mom6_interact")
            import :: c_int, c_double, c_ptr
            type(c_ptr), value :: g
            integer(c_int), value :: kind, source, degree, transposed
            real(c_double), value :: x, y, z
            real(c_double), intent(in) :: w(*), area(*), ssh(*)
            real(c_double), intent(inout) :: gx, gy
        end function

        integer(c_int) function c_sal_mom6_pp_points(g, x, y, z, ns, sx, sy, sz, ssh, area, gx, gy) &
                bind(C, name="sal_mom6_pp_points")
            import :: c_int, c_double, c_ptr
            type(c_ptr), value :: g
            real(c_double), value :: x, y, z
            integer(c_int), value :: ns
            real(c_double), intent(in) :: sx(*), sy(*), sz(*), ssh(*), area(*)
            real(c_double), intent(inout) :: gx, gy
        end function

        integer(c_int) function c_sal_mom6_pp_compute(g, area, ssh, gx, gy, acc) bind(C, name="sal_mom6_pp_compute")
            import :: c_int, c_double, c_ptr
            type(c_ptr), value :: g
            real(c_double), intent(in) :: area(*), ssh(*)
            integer(c_int), value :: acc
            real(c_double), intent(inout) :: gx(*), gy(*)
        end function
    end interface

contains

    ! Print the library's message for a failed call and stop.
    subroutine sal_check(rc, where)
        integer(c_int), intent(in) :: rc
        character(len=*), intent(in) :: where
        character(kind=c_char), pointer :: msg(:)
        type(c_ptr) :: p
        integer :: k
        character(len=512) :: text
        if (rc == 0) return
        p = c_sal_last_error()
        text = ''
        if (c_associated(p)) then
            call c_f_pointer(p, msg, [512])
            do k = 1, 512
                if (msg(k) == c_null_char) exit
                text(k:k) = msg(k)
            end do
        end if
        print '(a, a, i0, a, a)', where, ': salfort status ', rc, ': ', trim(text)
        error stop 1
    end subroutine
end module salfort_capi
