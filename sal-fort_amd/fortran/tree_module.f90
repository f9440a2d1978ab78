! Drop-in replacement of the reference's tree_module (tree_module.f90:1-627).
!
! The public types (cube_panel with its components and defaults, interaction_pair) and procedures
! keep the reference's names, arguments, kinds and intents, so the MPAS and MOM6 drivers compile
! against this module unchanged.  The work runs in libsalfort on the GPU:
!
!   tree_traverse            sal_geom_create + sal_geom_export (tree built on the GPU, then copied
!                            into the Fortran panel array with 1-based ids)
!   assign_points_to_panels  sal_assign_points
!   DUAL_TREE_TRAVERSAL      sal_dtt + sal_dtt_export
!   SINGLE_TREE_TRAVERSAL    sal_stt + sal_stt_export
!
! The device copy of the tree stays resident in this module (sal_tree) so fast_sum and the MOM6
! sums do not upload it again.  unowned_sources and relabel_copy_unowned are the reference's
! multi-rank halo bookkeeping (integer work on host arrays), kept with the same results.
MODULE Tree_Module
    use iso_c_binding
    use cube_sphere_module
    use salfort_capi
    implicit none

    type :: cube_panel ! tree_module.f90:5-28
        integer :: level = 0
        logical :: is_leaf = .true.
        integer :: id
        integer :: parent_panel = -1
        integer :: child_panel_1 = -1
        integer :: child_panel_2 = -1
        integer :: child_panel_3 = -1
        integer :: child_panel_4 = -1
        integer :: face
        real(8) :: min_xi
        real(8) :: mid_xi
        real(8) :: max_xi
        real(8) :: min_eta
        real(8) :: mid_eta
        real(8) :: max_eta
        real(8) :: radius
        integer, allocatable :: points_inside(:)
        integer, allocatable :: relabeled_points_inside(:)
        integer :: panel_point_count = 0
        contains
            procedure :: contains_point
    end type cube_panel

    type :: interaction_pair ! tree_module.f90:30-34
        integer :: index_target
        integer :: index_source
        integer :: interact_type
    end type interaction_pair

    ! device geometry of the last tree_traverse, and a fingerprint of the list last sent to it
    type(c_ptr), save, protected :: sal_tree = c_null_ptr
    integer, save, protected :: sal_tree_n = -1, sal_tree_panels = -1
    integer(c_int64_t), save, private :: list_n = -1, list_hash = -1
    integer(c_int64_t), save, private :: stt_pc_n = -1, stt_pp_n = -1, stt_hash = -1
    ! this rank's share of the grid given to tree_traverse, as relabel_copy_unowned sets it up:
    ! global ids own_first .. own_first + own_count - 1 are its own points, halo_global(c) is the
    ! global id of halo copy c (the e_ arrays); one rank owns everything (own_count = -1: not set)
    integer, save, protected :: sal_own_first = 1, sal_own_count = -1
    integer, allocatable, save, protected :: sal_halo_global(:)

    private :: list_fingerprint, find_in_panel

CONTAINS

    ! point in panel: same face, xi in [min, max), eta in [min, max) (tree_module.f90:39-56)
    logical function contains_point(self, x, y, z) result(contains)
        class(cube_panel), intent(in) :: self
        real(8), intent(in) :: x, y, z
        real(8) :: xi, eta
        contains = .false.
        if (face_from_xyz(x, y, z) /= self%face) return
        call xieta_from_xyz(x, y, z, xi, eta, self%face)
        contains = (xi >= self%min_xi) .and. (xi < self%max_xi) .and. (eta >= self%min_eta) .and. &
                   (eta < self%max_eta)
    end function

    SUBROUTINE tree_traverse(tree_panels, xs, ys, zs, point_count, cluster_thresh, process_rank)
        type(cube_panel), allocatable, intent(out) :: tree_panels(:)
        integer, intent(in) :: point_count, cluster_thresh, process_rank
        real(8), intent(in) :: xs(point_count), ys(point_count), zs(point_count)
        integer(c_int) :: np, maxlev
        integer(c_int64_t) :: tot
        integer(c_int), allocatable :: lev(:), leaf(:), par(:), ch(:), fc(:), cnt(:), pts(:)
        real(c_double), allocatable :: bnd(:), rad(:)
        integer(c_int64_t), allocatable :: off(:)
        integer :: i, k
        if (c_associated(sal_tree)) call sal_check(c_sal_geom_destroy(sal_tree), 'tree_traverse')
        sal_tree = c_null_ptr
        call sal_check(c_sal_geom_create(xs, ys, zs, int(point_count, c_int64_t), int(cluster_thresh, c_int), &
                                         sal_tree), 'tree_traverse')
        call sal_check(c_sal_geom_counts(sal_tree, np, maxlev, tot), 'tree_traverse')
        allocate(lev(np), leaf(np), par(np), ch(4*np), fc(np), cnt(np), bnd(6*np), rad(np), off(np+1))
        allocate(pts(max(tot, 1_c_int64_t)))
        call sal_check(c_sal_geom_export(sal_tree, lev, leaf, par, ch, fc, bnd, rad, cnt, off, pts), 'tree_traverse')
        allocate(tree_panels(np))
        do i = 1, np
            tree_panels(i)%level = lev(i)
            tree_panels(i)%is_leaf = leaf(i) /= 0
            tree_panels(i)%id = i
            tree_panels(i)%parent_panel = merge(par(i) + 1, -1, par(i) >= 0)
            tree_panels(i)%child_panel_1 = merge(ch(4*i-3) + 1, -1, ch(4*i-3) >= 0)
            tree_panels(i)%child_panel_2 = merge(ch(4*i-2) + 1, -1, ch(4*i-2) >= 0)
            tree_panels(i)%child_panel_3 = merge(ch(4*i-1) + 1, -1, ch(4*i-1) >= 0)
            tree_panels(i)%child_panel_4 = merge(ch(4*i) + 1, -1, ch(4*i) >= 0)
            tree_panels(i)%face = fc(i)
            tree_panels(i)%min_xi = bnd(6*i-5)
            tree_panels(i)%mid_xi = bnd(6*i-4)
            tree_panels(i)%max_xi = bnd(6*i-3)
            tree_panels(i)%min_eta = bnd(6*i-2)
            tree_panels(i)%mid_eta = bnd(6*i-1)
            tree_panels(i)%max_eta = bnd(6*i)
            tree_panels(i)%radius = rad(i)
            tree_panels(i)%panel_point_count = cnt(i)
            allocate(tree_panels(i)%points_inside(cnt(i)))
            do k = 1, cnt(i)
                tree_panels(i)%points_inside(k) = pts(off(i) + k) + 1
            end do
            tree_panels(i)%relabeled_points_inside = tree_panels(i)%points_inside
        end do
        sal_tree_n = point_count
        sal_tree_panels = np
        sal_own_first = 1
        sal_own_count = -1
        if (allocated(sal_halo_global)) deallocate(sal_halo_global)
        list_n = -1
        stt_pc_n = -1
    END SUBROUTINE

    ! the device tree behind tree_panels (error stop if these panels were not built here)
    function sal_device_tree(tree_panels, where) result(g)
        type(cube_panel), intent(in) :: tree_panels(:)
        character(len=*), intent(in) :: where
        type(c_ptr) :: g
        if (.not. c_associated(sal_tree) .or. size(tree_panels) /= sal_tree_panels) then
            print '(a, a)', where, ': the panels were not built by this tree_module''s tree_traverse'
            error stop 1
        end if
        g = sal_tree
    end function

    SUBROUTINE assign_points_to_panels(xs, ys, zs, points_panels, tree_panels, own_points)
        real(8), intent(in) :: xs(:), ys(:), zs(:)
        integer, intent(out) :: points_panels(:,:)
        type(cube_panel), intent(in) :: tree_panels(:)
        integer, intent(in) :: own_points
        integer(c_int), allocatable :: pp(:)
        integer :: nlev, i, l
        type(c_ptr) :: g
        g = sal_device_tree(tree_panels, 'assign_points_to_panels')
        nlev = size(points_panels, 1)
        points_panels = -1
        if (own_points <= 0) return
        allocate(pp(nlev * own_points))
        call sal_check(c_sal_assign_points(g, xs, ys, zs, int(own_points, c_int64_t), int(nlev, c_int), pp), &
                       'assign_points_to_panels')
        do i = 1, own_points
            do l = 1, nlev
                if (pp((i-1)*nlev + l) >= 0) points_panels(l, i) = pp((i-1)*nlev + l) + 1
            end do
        end do
    END SUBROUTINE

    ! order-sensitive fingerprint of a list (size + polynomial hash modulo a prime)
    function list_fingerprint(list) result(h)
        type(interaction_pair), intent(in) :: list(:)
        integer(c_int64_t) :: h
        integer :: i
        h = size(list, kind=c_int64_t)
        do i = 1, size(list)
            h = mod(h * 1000003_c_int64_t + list(i)%index_target, 2147483647_c_int64_t)
            h = mod(h * 1000003_c_int64_t + list(i)%index_source, 2147483647_c_int64_t)
            h = mod(h * 1000003_c_int64_t + list(i)%interact_type, 2147483647_c_int64_t)
        end do
    end function

    ! the device tree with `list` as its interaction list (uploaded when it is not the one there)
    function sal_device_list(tree_panels, list, where) result(g)
        type(cube_panel), intent(in) :: tree_panels(:)
        type(interaction_pair), intent(in) :: list(:)
        character(len=*), intent(in) :: where
        type(c_ptr) :: g
        integer(c_int64_t) :: h
        integer(c_int), allocatable :: it(:), is(:), ty(:)
        integer :: i
        g = sal_device_tree(tree_panels, where)
        h = list_fingerprint(list)
        if (size(list, kind=c_int64_t) == list_n .and. h == list_hash) return
        allocate(it(max(size(list), 1)), is(max(size(list), 1)), ty(max(size(list), 1)))
        do i = 1, size(list)
            it(i) = list(i)%index_target - 1
            is(i) = list(i)%index_source - 1
            ty(i) = list(i)%interact_type
        end do
        call sal_check(c_sal_dtt_import(g, it, is, ty, size(list, kind=c_int64_t)), where)
        list_n = size(list, kind=c_int64_t)
        list_hash = h
    end function

    ! the device tree whose single-tree lists are these (error stop if they were not built here)
    function sal_device_stt(tree_panels, pc_list, pp_list, where) result(g)
        type(cube_panel), intent(in) :: tree_panels(:)
        type(interaction_pair), intent(in), optional :: pc_list(:), pp_list(:)
        character(len=*), intent(in) :: where
        type(c_ptr) :: g
        logical :: ok
        g = sal_device_tree(tree_panels, where)
        ok = stt_pc_n >= 0
        if (present(pc_list)) ok = ok .and. size(pc_list, kind=c_int64_t) == stt_pc_n
        if (present(pp_list)) ok = ok .and. size(pp_list, kind=c_int64_t) == stt_pp_n
        if (.not. ok) then
            print '(a, a)', where, ': the lists were not built by this tree_module''s SINGLE_TREE_TRAVERSAL'
            error stop 1
        end if
    end function

    SUBROUTINE DUAL_TREE_TRAVERSAL(interaction_list, tree_panels, param_theta, param_cluster_thresh)
        type(interaction_pair), allocatable, intent(out) :: interaction_list(:)
        type(cube_panel), intent(in) :: tree_panels(:)
        real(8), intent(in) :: param_theta
        integer, intent(in) :: param_cluster_thresh
        integer(c_int64_t) :: n
        integer(c_int), allocatable :: it(:), is(:), ty(:)
        integer :: i
        type(c_ptr) :: g
        g = sal_device_tree(tree_panels, 'DUAL_TREE_TRAVERSAL')
        call sal_check(c_sal_dtt(g, real(param_theta, c_double), int(param_cluster_thresh, c_int), n), &
                       'DUAL_TREE_TRAVERSAL')
        allocate(it(max(n, 1_c_int64_t)), is(max(n, 1_c_int64_t)), ty(max(n, 1_c_int64_t)))
        call sal_check(c_sal_dtt_export(g, it, is, ty), 'DUAL_TREE_TRAVERSAL')
        allocate(interaction_list(n))
        do i = 1, int(n)
            interaction_list(i)%index_target = it(i) + 1
            interaction_list(i)%index_source = is(i) + 1
            interaction_list(i)%interact_type = ty(i)
        end do
        list_n = n
        list_hash = list_fingerprint(interaction_list)
    END SUBROUTINE

    SUBROUTINE SINGLE_TREE_TRAVERSAL(pc_interaction_list, pp_interaction_list, tree_panels, param_theta, &
                                     param_cluster_thresh, xs, ys, zs, own_points)
        type(interaction_pair), allocatable, intent(out) :: pc_interaction_list(:), pp_interaction_list(:)
        type(cube_panel), intent(in) :: tree_panels(:)
        real(8), intent(in) :: param_theta, xs(:), ys(:), zs(:)
        integer, intent(in) :: param_cluster_thresh, own_points
        integer(c_int64_t) :: npc, npp
        integer(c_int), allocatable :: pct(:), pcs(:), ppt(:), pps(:)
        integer :: i
        type(c_ptr) :: g
        g = sal_device_tree(tree_panels, 'SINGLE_TREE_TRAVERSAL')
        call sal_check(c_sal_stt(g, real(param_theta, c_double), xs, ys, zs, int(own_points, c_int64_t), npc, npp), &
                       'SINGLE_TREE_TRAVERSAL')
        allocate(pct(max(npc, 1_c_int64_t)), pcs(max(npc, 1_c_int64_t)))
        allocate(ppt(max(npp, 1_c_int64_t)), pps(max(npp, 1_c_int64_t)))
        call sal_check(c_sal_stt_export(g, pct, pcs, ppt, pps), 'SINGLE_TREE_TRAVERSAL')
        allocate(pc_interaction_list(npc), pp_interaction_list(npp))
        do i = 1, int(npc)
            pc_interaction_list(i) = interaction_pair(pct(i) + 1, pcs(i) + 1, 1)
        end do
        do i = 1, int(npp)
            pp_interaction_list(i) = interaction_pair(ppt(i) + 1, pps(i) + 1, 0)
        end do
        stt_pc_n = npc
        stt_pp_n = npp
    END SUBROUTINE

    ! Non-owned points of the PP source panels, grouped by owning rank, in first-seen order
    ! (tree_module.f90:490-561).  Ranks own contiguous global index ranges starting at
    ! starting_point_proc(k).
    SUBROUTINE unowned_sources(starting_point_proc, pp_interaction_list, tree_panels, P, ID, unowned_points, &
                               points_per_proc, total_points)
        integer, intent(in) :: starting_point_proc(:), P, ID, total_points
        type(interaction_pair), intent(in) :: pp_interaction_list(:)
        type(cube_panel), intent(in) :: tree_panels(:)
        integer, intent(out), allocatable :: unowned_points(:,:), points_per_proc(:)
        integer, allocatable :: tmp(:,:)
        logical, allocatable :: seen(:)
        integer :: lo, hi, i, j, s, pnt, k, owner
        allocate(points_per_proc(P), source=0)
        allocate(tmp(total_points, P), seen(total_points))
        seen = .false.
        lo = starting_point_proc(ID+1)
        hi = merge(total_points, starting_point_proc(min(ID+2, P)) - 1, ID+1 == P)
        do i = 1, size(pp_interaction_list)
            s = pp_interaction_list(i)%index_source
            do j = 1, tree_panels(s)%panel_point_count
                pnt = tree_panels(s)%points_inside(j)
                if (pnt >= lo .and. pnt <= hi) cycle
                if (seen(pnt)) cycle
                seen(pnt) = .true.
                owner = 0
                do k = 1, P
                    if (pnt >= starting_point_proc(k)) owner = k
                end do
                if (owner == 0) cycle
                points_per_proc(owner) = points_per_proc(owner) + 1
                tmp(points_per_proc(owner), owner) = pnt
            end do
        end do
        allocate(unowned_points(maxval([0, points_per_proc]), P), source=-1)
        do k = 1, P
            unowned_points(1:points_per_proc(k), k) = tmp(1:points_per_proc(k), k)
        end do
    END SUBROUTINE

    ! position of global point pnt in panel k's ascending points_inside, 0 if absent
    integer function find_in_panel(panel, pnt) result(pos)
        type(cube_panel), intent(in) :: panel
        integer, intent(in) :: pnt
        integer :: a, b, c
        pos = 0
        a = 1
        b = panel%panel_point_count
        do while (a <= b)
            c = (a + b) / 2
            if (panel%points_inside(c) == pnt) then
                pos = c
                return
            else if (panel%points_inside(c) < pnt) then
                a = c + 1
            else
                b = c - 1
            end if
        end do
    end function

    ! Local numbering of owned points and of the halo copies (tree_module.f90:566-625): owned global
    ! id g becomes g - first + 1; the c-th halo point becomes own_points + c in every panel holding it.
    SUBROUTINE relabel_copy_unowned(e_xs, e_ys, e_zs, e_areas, unowned_points, xs_t, ys_t, zs_t, areas_t, &
                                    tree_panels, P, ID, unowned_points_count, own_points, starting_points)
        real(8), intent(out) :: e_xs(unowned_points_count), e_ys(unowned_points_count), e_zs(unowned_points_count), &
                                e_areas(unowned_points_count)
        integer, intent(in) :: unowned_points(:,:), P, ID, unowned_points_count, own_points, starting_points(:)
        real(8), intent(in) :: xs_t(:), ys_t(:), zs_t(:), areas_t(:)
        type(cube_panel), intent(inout) :: tree_panels(:)
        integer :: lo, i, j, k, pnt, c, pos, nxt
        lo = starting_points(ID+1)
        sal_own_first = lo
        sal_own_count = own_points
        if (allocated(sal_halo_global)) deallocate(sal_halo_global)
        allocate(sal_halo_global(unowned_points_count))
        do i = 1, size(tree_panels)
            do j = 1, tree_panels(i)%panel_point_count
                pnt = tree_panels(i)%points_inside(j)
                if (pnt >= lo .and. pnt < lo + own_points) tree_panels(i)%relabeled_points_inside(j) = pnt - lo + 1
            end do
        end do
        c = 0
        do i = 1, P
            do j = 1, size(unowned_points, 1)
                pnt = unowned_points(j, i)
                if (pnt == -1) cycle
                c = c + 1
                sal_halo_global(c) = pnt
                e_xs(c) = xs_t(pnt)
                e_ys(c) = ys_t(pnt)
                e_zs(c) = zs_t(pnt)
                e_areas(c) = areas_t(pnt)
                k = 1  ! walk the panels holding pnt, root to leaf
                do while (k /= -1 .and. k <= size(tree_panels))
                    pos = find_in_panel(tree_panels(k), pnt)
                    if (pos > 0) then
                        tree_panels(k)%relabeled_points_inside(pos) = own_points + c
                        nxt = tree_panels(k)%child_panel_1
                        k = nxt
                    else
                        k = k + 1
                    end if
                end do
            end do
        end do
    END SUBROUTINE
END MODULE Tree_Module
