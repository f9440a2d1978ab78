! Drop-in replacement of the reference's mom6_fast_sum_module (mom6_fast_sum_module.f90:1-193), the
! MOM6 evaluator: same names and arguments, same accumulate-into-the-caller semantics, and by default
! the reference's eta-major read of the proxy weights in pc_interact (:96-104).  Setting
! mom6_weights_as_stored = .true. reads them the way proxy_source_compute stores them instead
! (the interpolation the algorithm intends).  The work runs in libsalfort on the GPU with the
! single-tree lists tree_module keeps resident.
!
! One rank may hold the whole grid (own points = all points), or several ranks each their own points
! and halo copies, as the MOM6 driver sets them up (tree_traverse on the whole grid's coordinates,
! relabel_copy_unowned for the rank's share): the per-rank sums then place the rank's own and halo
! fields at their global ids for the device, and the driver's Sum_Array adds the proxy weights.
MODULE MOM6_FAST_SUM_MODULE
    use iso_c_binding
    use tree_module
    use bli_module
    use cube_sphere_module
    use salfort_capi
    implicit none
    logical, save :: mom6_weights_as_stored = .false.
    private :: need_all_points, layout_flag

CONTAINS

    ! SAL-gradient kernel (mom6_fast_sum_module.f90:9-31, the same function as fast_sum_module's)
    SUBROUTINE sal_grad_gfunc(tx, ty, tz, sx, sy, sz, sal_x, sal_y)
        real(8), intent(in) :: tx, ty, tz, sx, sy, sz
        real(8), intent(out) :: sal_x, sal_y
        real(8), parameter :: cons = -2.343256857908601e-9_8
        real(8) :: g, c2, c, p1, p2, val, a, cp
        sal_x = 0.0_8
        sal_y = 0.0_8
        if ((abs(tz - 1.0_8) <= 1e-15_8) .or. (abs(tz + 1.0_8) <= 1e-15_8)) return
        g = max(min(tx*sx + ty*sy + tz*sz, 1.0_8), -1.0_8)
        c2 = 2.0_8 - 2.0_8*g
        c = sqrt(c2)
        p1 = (1.0_8 - 6.21196_8) / (c*c2 + 1e-16_8)
        p2 = (2.7_8 + 6.0_8)*(2_8*g + c) / (2.0_8*(g*g - 1.0_8) + 1e-16_8)
        val = (p1 + p2)*cons
        a = tz*tz
        cp = sqrt(1.0_8 - a)
        sal_y = (sz*(1.0_8 - a) - tz*(tx*sx + ty*sy))/cp*val
        sal_x = (tx*sy - ty*sx)/cp*val
    END SUBROUTINE

    subroutine need_all_points(n, where)
        integer, intent(in) :: n
        character(len=*), intent(in) :: where
        if (.not. c_associated(sal_tree) .or. n /= sal_tree_n) then
            print '(a, a)', where, ': needs the whole grid on this rank, as given to tree_traverse'
            error stop 1
        end if
    end subroutine

    ! Grid-length area and ssh arrays for the device sums when this rank holds only its own points
    ! (and the halo copies of others'): own values at their global ids (from relabel_copy_unowned),
    ! halo values at theirs when given, zero elsewhere.  The device tree holds every point's
    ! coordinates (tree_traverse gets the whole grid), so only the fields need placing.
    subroutine rank_fields(areas, sshs, own, af, sf, where, e_areas, e_sshs)
        real(8), intent(in) :: areas(:), sshs(:)
        integer, intent(in) :: own
        real(c_double), allocatable, intent(out) :: af(:), sf(:)
        character(len=*), intent(in) :: where
        real(8), intent(in), optional :: e_areas(:), e_sshs(:)
        integer :: c, h
        if (.not. c_associated(sal_tree) .or. own /= sal_own_count .or. &
            sal_own_first + own - 1 > sal_tree_n) then
            print '(a, a)', where, ': this rank''s points are not those relabel_copy_unowned set up'
            error stop 1
        end if
        allocate(af(sal_tree_n), sf(sal_tree_n), source=0.0_c_double)
        af(sal_own_first:sal_own_first + own - 1) = areas(1:own)
        sf(sal_own_first:sal_own_first + own - 1) = sshs(1:own)
        if (present(e_areas) .and. present(e_sshs)) then
            do c = 1, size(sal_halo_global)
                h = sal_halo_global(c)
                af(h) = e_areas(c)
                sf(h) = e_sshs(c)
            end do
        end if
    end subroutine

    integer(c_int) function layout_flag()
        layout_flag = merge(0_c_int, 1_c_int, mom6_weights_as_stored)
    end function

    ! weights of every panel from the points on its root path, xi-major (mom6_fast_sum_module.f90:33-74).
    ! The reference adds into its intent(out) dummy, which its caller has zeroed (mom6_code.f90:273);
    ! here the dummy is zeroed first and then defined, the same result without reading an
    ! intent(out) argument before it is set.
    SUBROUTINE proxy_source_compute(proxy_source_weights, tree_panels, points_panels, xs, ys, zs, sshs, areas, &
                                    interp_degree, point_count)
        real(8), intent(out) :: proxy_source_weights(:)
        real(8), intent(in) :: xs(:), ys(:), zs(:), sshs(:), areas(:)
        type(cube_panel), intent(in) :: tree_panels(:)
        integer, intent(in) :: points_panels(:,:), interp_degree, point_count
        real(c_double), allocatable :: w(:), af(:), sf(:)
        type(c_ptr) :: g
        integer :: nw
        g = sal_device_tree(tree_panels, 'proxy_source_compute')
        nw = (interp_degree+1)*(interp_degree+1)*size(tree_panels)
        allocate(w(nw))
        if (point_count == sal_tree_n) then
            call sal_check(c_sal_mom6_proxy_source(g, areas, sshs, int(interp_degree, c_int), w), &
                           'proxy_source_compute')
        else  ! one rank of several: its own points' share, which the driver's Sum_Array adds up
            call rank_fields(areas, sshs, point_count, af, sf, 'proxy_source_compute')
            call sal_check(c_sal_mom6_proxy_source(g, af, sf, int(interp_degree, c_int), w), 'proxy_source_compute')
        end if
        proxy_source_weights = 0.0_8
        proxy_source_weights(1:nw) = w
    END SUBROUTINE

    ! one target against one source panel's proxies (mom6_fast_sum_module.f90:76-107)
    SUBROUTINE pc_interact(sal_longrad, sal_latgrad, i_t, source_panel, x, y, z, interp_degree, proxy_weights)
        real(8), intent(inout) :: sal_longrad(:), sal_latgrad(:)
        integer, intent(in) :: i_t, interp_degree
        real(8), intent(in) :: proxy_weights(:), x, y, z
        type(cube_panel), intent(in) :: source_panel
        real(c_double) :: dummy(1)
        if (.not. c_associated(sal_tree)) call need_all_points(-1, 'pc_interact')
        call sal_check(c_sal_mom6_interact(sal_tree, 1_c_int, x, y, z, int(source_panel%id - 1, c_int), &
                                           int(interp_degree, c_int), layout_flag(), proxy_weights, dummy, dummy, &
                                           sal_longrad(i_t), sal_latgrad(i_t)), 'pc_interact')
    END SUBROUTINE

    ! one target against one source panel's points (mom6_fast_sum_module.f90:109-150): the panel's
    ! relabeled points are the rank's own points (ids up to own_points) or halo copies (the rest), read
    ! from the caller's arrays as the reference reads them; the entry's points go to the device, nothing
    ! of the grid's size
    SUBROUTINE pp_interact(sal_longrad, sal_latgrad, i_t, source_panel, xs, ys, zs, as, sshs, e_xs, e_ys, e_zs, &
                           e_as, e_sshs, own_points)
        real(8), intent(inout) :: sal_longrad(:), sal_latgrad(:)
        integer, intent(in) :: i_t, own_points
        real(8), intent(in) :: xs(:), ys(:), zs(:), as(:), sshs(:), e_xs(:), e_ys(:), e_zs(:), e_as(:), e_sshs(:)
        type(cube_panel), intent(in) :: source_panel
        real(c_double), allocatable :: sx(:), sy(:), sz(:), ss(:), sa(:)
        integer :: i, i_s, ns
        if (.not. c_associated(sal_tree)) call need_all_points(-1, 'pp_interact')
        ns = source_panel%panel_point_count
        allocate(sx(max(ns, 1)), sy(max(ns, 1)), sz(max(ns, 1)), ss(max(ns, 1)), sa(max(ns, 1)))
        do i = 1, ns
            i_s = source_panel%relabeled_points_inside(i)
            if (i_s > own_points) then
                i_s = i_s - own_points
                sx(i) = e_xs(i_s); sy(i) = e_ys(i_s); sz(i) = e_zs(i_s); ss(i) = e_sshs(i_s); sa(i) = e_as(i_s)
            else
                sx(i) = xs(i_s); sy(i) = ys(i_s); sz(i) = zs(i_s); ss(i) = sshs(i_s); sa(i) = as(i_s)
            end if
        end do
        call sal_check(c_sal_mom6_pp_points(sal_tree, xs(i_t), ys(i_t), zs(i_t), int(ns, c_int), sx, sy, sz, ss, sa, &
                                            sal_longrad(i_t), sal_latgrad(i_t)), 'pp_interact')
    END SUBROUTINE

    ! all particle-cluster entries (mom6_fast_sum_module.f90:153-174)
    SUBROUTINE pc_compute(sal_longrad, sal_latgrad, pc_interaction_list, cube_panels, xs, ys, zs, interp_degree, &
                          proxy_weights)
        real(8), intent(in) :: xs(:), ys(:), zs(:), proxy_weights(:)
        type(interaction_pair), intent(in) :: pc_interaction_list(:)
        real(8), intent(inout) :: sal_longrad(:), sal_latgrad(:)
        type(cube_panel), intent(in) :: cube_panels(:)
        integer, intent(in) :: interp_degree
        type(c_ptr) :: g
        g = sal_device_stt(cube_panels, pc_list=pc_interaction_list, where='pc_compute')
        call sal_check(c_sal_mom6_pc_compute(g, proxy_weights, int(interp_degree, c_int), layout_flag(), &
                                             sal_longrad, sal_latgrad, 1_c_int), 'pc_compute')
    END SUBROUTINE

    ! all particle-particle entries (mom6_fast_sum_module.f90:177-192)
    SUBROUTINE pp_compute(sal_longrad, sal_latgrad, pp_interaction_list, cube_panels, xs, ys, zs, areas, sshs, e_xs, &
                          e_ys, e_zs, e_areas, e_sshs, own_points)
        real(8), intent(inout) :: sal_longrad(:), sal_latgrad(:)
        type(interaction_pair), intent(in) :: pp_interaction_list(:)
        type(cube_panel), intent(in) :: cube_panels(:)
        real(8), intent(in) :: xs(:), ys(:), zs(:), areas(:), sshs(:), e_xs(:), e_ys(:), e_zs(:), e_areas(:), e_sshs(:)
        integer, intent(in) :: own_points
        type(c_ptr) :: g
        real(c_double), allocatable :: af(:), sf(:)
        g = sal_device_stt(cube_panels, pp_list=pp_interaction_list, where='pp_compute')
        if (own_points == sal_tree_n) then
            call sal_check(c_sal_mom6_pp_compute(g, areas, sshs, sal_longrad, sal_latgrad, 1_c_int), 'pp_compute')
        else  ! sources among this rank's own points and halo copies (relabeled_points_inside)
            call rank_fields(areas, sshs, own_points, af, sf, 'pp_compute', e_areas, e_sshs)
            call sal_check(c_sal_mom6_pp_compute(g, af, sf, sal_longrad, sal_latgrad, 1_c_int), 'pp_compute')
        end if
    END SUBROUTINE
END MODULE
