! Test infrastructure only: a bind(C) bridge onto the reference's own Fortran modules, compiled by
! oracle/Makefile from the sources in the reference tree into oracle/_ref/libsalref.so.  Nothing in
! the product links this; tests and the CPU-baseline leg of bench.py load it through ctypes.
!
! Every routine forwards to the reference procedure named in its comment and only converts between
! the reference's derived types and flat C arrays.  The tree and the interaction lists are held in
! module variables between calls (one grid at a time).  Ids are exported exactly as the reference
! stores them (1-based).
module ref_bridge
    use iso_c_binding
    use cube_sphere_module
    use bli_module
    use tree_module
    use fast_sum_module, only: fs_gfunc => sal_grad_gfunc, fast_sum, fs_pp => pp_interact, &
                               fs_pc => pc_interact, fs_cp => cp_interact, fs_cc => cc_interact
    use mom6_fast_sum_module, only: proxy_source_compute, pc_compute, pp_compute, &
                                    m6_gfunc => sal_grad_gfunc
    implicit none

    type(cube_panel), allocatable, save :: panels(:)
    type(interaction_pair), allocatable, save :: ilist(:), pc_list(:), pp_list(:)

contains

    ! tree_traverse (tree_module.f90:59)
    subroutine ref_tree_build(n, x, y, z, thresh, n_panels) bind(C, name="ref_tree_build")
        integer(c_int), value :: n, thresh
        real(c_double), intent(in) :: x(n), y(n), z(n)
        integer(c_int), intent(out) :: n_panels
        if (allocated(panels)) deallocate(panels)
        call tree_traverse(panels, x, y, z, n, thresh, 0)
        n_panels = size(panels)
    end subroutine

    subroutine ref_tree_total_inside(total) bind(C, name="ref_tree_total_inside")
        integer(c_int64_t), intent(out) :: total
        integer :: i
        total = 0
        do i = 1, size(panels)
            total = total + panels(i)%panel_point_count
        end do
    end subroutine

    ! flat export of type cube_panel (tree_module.f90:5-28); points_inside concatenated in panel order
    subroutine ref_tree_export(level, is_leaf, id, parent, child, face, bounds, radius, cnt, offsets, pts) &
            bind(C, name="ref_tree_export")
        integer(c_int), intent(out) :: level(*), is_leaf(*), id(*), parent(*), child(4, *), face(*), cnt(*)
        real(c_double), intent(out) :: bounds(6, *), radius(*)
        integer(c_int64_t), intent(out) :: offsets(*)
        integer(c_int), intent(out) :: pts(*)
        integer :: i, j
        integer(c_int64_t) :: off
        off = 0
        do i = 1, size(panels)
            level(i) = panels(i)%level
            is_leaf(i) = merge(1, 0, panels(i)%is_leaf)
            id(i) = panels(i)%id
            parent(i) = panels(i)%parent_panel
            child(1, i) = panels(i)%child_panel_1
            child(2, i) = panels(i)%child_panel_2
            child(3, i) = panels(i)%child_panel_3
            child(4, i) = panels(i)%child_panel_4
            face(i) = panels(i)%face
            bounds(1, i) = panels(i)%min_xi
            bounds(2, i) = panels(i)%mid_xi
            bounds(3, i) = panels(i)%max_xi
            bounds(4, i) = panels(i)%min_eta
            bounds(5, i) = panels(i)%mid_eta
            bounds(6, i) = panels(i)%max_eta
            radius(i) = panels(i)%radius
            cnt(i) = panels(i)%panel_point_count
            offsets(i) = off
            do j = 1, panels(i)%panel_point_count
                pts(off + j) = panels(i)%points_inside(j)
            end do
            off = off + panels(i)%panel_point_count
        end do
        offsets(size(panels) + 1) = off
    end subroutine

    ! DUAL_TREE_TRAVERSAL (tree_module.f90:277)
    subroutine ref_dtt(theta, thresh, n_int) bind(C, name="ref_dtt")
        real(c_double), value :: theta
        integer(c_int), value :: thresh
        integer(c_int64_t), intent(out) :: n_int
        if (allocated(ilist)) deallocate(ilist)
        call dual_tree_traversal(ilist, panels, theta, thresh)
        n_int = size(ilist)
    end subroutine

    subroutine ref_dtt_export(it, is, ty) bind(C, name="ref_dtt_export")
        integer(c_int), intent(out) :: it(*), is(*), ty(*)
        integer :: i
        do i = 1, size(ilist)
            it(i) = ilist(i)%index_target
            is(i) = ilist(i)%index_source
            ty(i) = ilist(i)%interact_type
        end do
    end subroutine

    ! fast_sum (fast_sum_module.f90:291); accumulates into gx, gy like the reference
    subroutine ref_fast_sum(n, x, y, z, area, ssh, degree, rank, nproc, gx, gy) bind(C, name="ref_fast_sum")
        integer(c_int), value :: n, degree, rank, nproc
        real(c_double), intent(in) :: x(n), y(n), z(n), area(n), ssh(n)
        real(c_double), intent(inout) :: gx(n), gy(n)
        call fast_sum(gx, gy, ilist, panels, x, y, z, area, ssh, n, degree, rank, nproc)
    end subroutine

    ! fast_sum restricted to the interaction range [first, last] (1-based, inclusive): used to time a
    ! bounded sample of the reference's own evaluation loop
    subroutine ref_fast_sum_range(n, x, y, z, area, ssh, degree, first, last, gx, gy) &
            bind(C, name="ref_fast_sum_range")
        integer(c_int), value :: n, degree
        integer(c_int64_t), value :: first, last
        real(c_double), intent(in) :: x(n), y(n), z(n), area(n), ssh(n)
        real(c_double), intent(inout) :: gx(n), gy(n)
        call fast_sum(gx, gy, ilist(first:last), panels, x, y, z, area, ssh, n, degree, 0, 1)
    end subroutine

    ! one interaction of each kind, exactly as fast_sum dispatches them (fast_sum_module.f90:309-327)
    subroutine ref_interact(kind, i_t, i_s, n, x, y, z, area, ssh, degree, gx, gy) bind(C, name="ref_interact")
        integer(c_int), value :: kind, i_t, i_s, n, degree
        real(c_double), intent(in) :: x(n), y(n), z(n), area(n), ssh(n)
        real(c_double), intent(inout) :: gx(n), gy(n)
        select case (kind)
        case (0)
            call fs_pp(gx, gy, panels(i_t), panels(i_s), x, y, z, area, ssh, degree)
        case (1)
            call fs_pc(gx, gy, panels(i_t), panels(i_s), x, y, z, area, ssh, degree)
        case (2)
            call fs_cp(gx, gy, panels(i_t), panels(i_s), x, y, z, area, ssh, degree)
        case (3)
            call fs_cc(gx, gy, panels(i_t), panels(i_s), x, y, z, area, ssh, degree)
        end select
    end subroutine

    ! assign_points_to_panels (tree_module.f90:249); the caller pre-fills -1 like mom6_code.f90:188-189
    subroutine ref_assign(n, x, y, z, nlev, pp) bind(C, name="ref_assign")
        integer(c_int), value :: n, nlev
        real(c_double), intent(in) :: x(n), y(n), z(n)
        integer(c_int), intent(inout) :: pp(nlev, n)
        call assign_points_to_panels(x, y, z, pp, panels, n)
    end subroutine

    ! SINGLE_TREE_TRAVERSAL (tree_module.f90:401)
    subroutine ref_stt(theta, thresh, n, x, y, z, npc, npp) bind(C, name="ref_stt")
        real(c_double), value :: theta
        integer(c_int), value :: thresh, n
        real(c_double), intent(in) :: x(n), y(n), z(n)
        integer(c_int64_t), intent(out) :: npc, npp
        if (allocated(pc_list)) deallocate(pc_list)
        if (allocated(pp_list)) deallocate(pp_list)
        call single_tree_traversal(pc_list, pp_list, panels, theta, thresh, x, y, z, n)
        npc = size(pc_list)
        npp = size(pp_list)
    end subroutine

    subroutine ref_stt_export(pc_t, pc_s, pp_t, pp_s) bind(C, name="ref_stt_export")
        integer(c_int), intent(out) :: pc_t(*), pc_s(*), pp_t(*), pp_s(*)
        integer :: i
        do i = 1, size(pc_list)
            pc_t(i) = pc_list(i)%index_target
            pc_s(i) = pc_list(i)%index_source
        end do
        do i = 1, size(pp_list)
            pp_t(i) = pp_list(i)%index_target
            pp_s(i) = pp_list(i)%index_source
        end do
    end subroutine

    ! proxy_source_compute (mom6_fast_sum_module.f90:33); zero-filled first, as mom6_code.f90:273 does
    subroutine ref_proxy_source(n, nlev, pp, x, y, z, ssh, area, degree, w) bind(C, name="ref_proxy_source")
        integer(c_int), value :: n, nlev, degree
        integer(c_int), intent(in) :: pp(nlev, n)
        real(c_double), intent(in) :: x(n), y(n), z(n), ssh(n), area(n)
        real(c_double), intent(out) :: w((degree+1)*(degree+1)*size(panels))
        w = 0.0_8
        call proxy_source_compute(w, panels, pp, x, y, z, ssh, area, degree, n)
    end subroutine

    ! pc_compute then pp_compute (mom6_fast_sum_module.f90:153, :177) at one rank: no halo points
    subroutine ref_mom6_sums(n, x, y, z, area, ssh, degree, w, gx, gy) bind(C, name="ref_mom6_sums")
        integer(c_int), value :: n, degree
        real(c_double), intent(in) :: x(n), y(n), z(n), area(n), ssh(n)
        real(c_double), intent(in) :: w((degree+1)*(degree+1)*size(panels))
        real(c_double), intent(inout) :: gx(n), gy(n)
        real(c_double) :: e(0)
        call pc_compute(gx, gy, pc_list, panels, x, y, z, degree, w)
        call pp_compute(gx, gy, pp_list, panels, x, y, z, area, ssh, e, e, e, e, e, n)
    end subroutine

    ! sal_grad_gfunc (fast_sum_module.f90:9); which=1 picks the MOM6 copy (mom6_fast_sum_module.f90:9)
    subroutine ref_gfunc(which, tx, ty, tz, sx, sy, sz, gx, gy) bind(C, name="ref_gfunc")
        integer(c_int), value :: which
        real(c_double), value :: tx, ty, tz, sx, sy, sz
        real(c_double), intent(out) :: gx, gy
        if (which == 1) then
            call m6_gfunc(tx, ty, tz, sx, sy, sz, gx, gy)
        else
            call fs_gfunc(tx, ty, tz, sx, sy, sz, gx, gy)
        end if
    end subroutine

    ! bli_coeffs / bli_interp_points_shift / interp_vals_bli (bli_module.f90:6, :28, :50)
    subroutine ref_bli_coeffs(degree, c) bind(C, name="ref_bli_coeffs")
        integer(c_int), value :: degree
        real(c_double), intent(out) :: c(degree+1)
        real(8), allocatable :: t(:)
        call bli_coeffs(t, degree)
        c = t
    end subroutine

    subroutine ref_bli_points(a, b, degree, p) bind(C, name="ref_bli_points")
        real(c_double), value :: a, b
        integer(c_int), value :: degree
        real(c_double), intent(out) :: p(degree+1)
        real(8), allocatable :: t(:)
        call bli_interp_points_shift(t, a, b, degree)
        p = t
    end subroutine

    subroutine ref_interp_vals(xi, eta, a, b, c, d, degree, v) bind(C, name="ref_interp_vals")
        real(c_double), value :: xi, eta, a, b, c, d
        integer(c_int), value :: degree
        real(c_double), intent(out) :: v(degree+1, degree+1)
        real(8), allocatable :: t(:,:)
        call interp_vals_bli(t, xi, eta, a, b, c, d, degree)
        v = t
    end subroutine

    ! cube_sphere_module.f90:7, :38, :64
    integer(c_int) function ref_face(x, y, z) bind(C, name="ref_face")
        real(c_double), value :: x, y, z
        ref_face = face_from_xyz(x, y, z)
    end function

    subroutine ref_xieta(x, y, z, face, xi, eta) bind(C, name="ref_xieta")
        real(c_double), value :: x, y, z
        integer(c_int), value :: face
        real(c_double), intent(out) :: xi, eta
        if (face < 1) then
            call xieta_from_xyz(x, y, z, xi, eta)
        else
            call xieta_from_xyz(x, y, z, xi, eta, face)
        end if
    end subroutine

    subroutine ref_xyz(xi, eta, face, x, y, z) bind(C, name="ref_xyz")
        real(c_double), value :: xi, eta
        integer(c_int), value :: face
        real(c_double), intent(out) :: x, y, z
        call xyz_from_xieta(x, y, z, xi, eta, face)
    end subroutine

    ! flushes what the reference printed on standard output (its unit 6 is buffered)
    subroutine ref_flush() bind(C, name="ref_flush")
        flush(6)
    end subroutine
end module ref_bridge
