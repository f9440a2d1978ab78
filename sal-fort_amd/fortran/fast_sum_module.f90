! Drop-in replacement of the reference's fast_sum_module (fast_sum_module.f90:1-333), the MPAS
! evaluator: same names, arguments and accumulate-into-the-caller semantics.  The sums run in
! libsalfort on the GPU against the tree kept resident by tree_module (sal_fast_sum evaluates the
! whole list; the per-interaction routines evaluate one entry with the same kernels).
!
! Ranks: the reference deals interactions round-robin to MPI ranks and sums the partial arrays
! afterwards (fast_sum_module.f90:304, code.f90:99-100).  Here rank 0 evaluates the whole list and
! the other ranks add nothing, which gives the same totals after that sum.
MODULE FAST_SUM_MODULE
    use iso_c_binding
    use tree_module
    use bli_module
    use cube_sphere_module
    use salfort_capi
    implicit none
    private :: one_interaction

CONTAINS

    ! SAL-gradient kernel (east, north) of target t and source s (fast_sum_module.f90:9-31); zero
    ! within 1e-15 of the poles
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

    subroutine one_interaction(kind, sal_longrad, sal_latgrad, target_panel, source_panel, area, ssh, &
                               interp_degree, where)
        integer, intent(in) :: kind, interp_degree
        real(8), intent(inout) :: sal_longrad(:), sal_latgrad(:)
        type(cube_panel), intent(in) :: target_panel, source_panel
        real(8), intent(in) :: area(:), ssh(:)
        character(len=*), intent(in) :: where
        if (.not. c_associated(sal_tree) .or. size(area) /= sal_tree_n) then
            print '(a, a)', where, ': no device tree for these points (call tree_traverse first)'
            error stop 1
        end if
        call sal_check(c_sal_interact(sal_tree, int(kind, c_int), int(target_panel%id - 1, c_int), &
                                      int(source_panel%id - 1, c_int), int(interp_degree, c_int), area, ssh, &
                                      sal_longrad, sal_latgrad, 1_c_int), where)
    end subroutine

    SUBROUTINE pp_interact(sal_longrad, sal_latgrad, target_panel, source_panel, xs, ys, zs, area, ssh, interp_degree)
        real(8), intent(in) :: xs(:), ys(:), zs(:), area(:), ssh(:)
        type(cube_panel), intent(in) :: target_panel, source_panel
        integer, intent(in) :: interp_degree
        real(8), intent(inout) :: sal_longrad(:), sal_latgrad(:)
        call one_interaction(0, sal_longrad, sal_latgrad, target_panel, source_panel, area, ssh, interp_degree, &
                             'pp_interact')
    END SUBROUTINE

    SUBROUTINE pc_interact(sal_longrad, sal_latgrad, target_panel, source_panel, xs, ys, zs, area, ssh, interp_degree)
        real(8), intent(in) :: xs(:), ys(:), zs(:), area(:), ssh(:)
        type(cube_panel), intent(in) :: target_panel, source_panel
        integer, intent(in) :: interp_degree
        real(8), intent(inout) :: sal_longrad(:), sal_latgrad(:)
        call one_interaction(1, sal_longrad, sal_latgrad, target_panel, source_panel, area, ssh, interp_degree, &
                             'pc_interact')
    END SUBROUTINE

    SUBROUTINE cp_interact(sal_longrad, sal_latgrad, target_panel, source_panel, xs, ys, zs, area, ssh, interp_degree)
        real(8), intent(in) :: xs(:), ys(:), zs(:), area(:), ssh(:)
        type(cube_panel), intent(in) :: target_panel, source_panel
        integer, intent(in) :: interp_degree
        real(8), intent(inout) :: sal_longrad(:), sal_latgrad(:)
        call one_interaction(2, sal_longrad, sal_latgrad, target_panel, source_panel, area, ssh, interp_degree, &
                             'cp_interact')
    END SUBROUTINE

    SUBROUTINE cc_interact(sal_longrad, sal_latgrad, target_panel, source_panel, xs, ys, zs, area, ssh, interp_degree)
        real(8), intent(in) :: xs(:), ys(:), zs(:), area(:), ssh(:)
        type(cube_panel), intent(in) :: target_panel, source_panel
        integer, intent(in) :: interp_degree
        real(8), intent(inout) :: sal_longrad(:), sal_latgrad(:)
        call one_interaction(3, sal_longrad, sal_latgrad, target_panel, source_panel, area, ssh, interp_degree, &
                             'cc_interact')
    END SUBROUTINE

    SUBROUTINE fast_sum(sal_longrad, sal_latgrad, interaction_list, cube_panels, xs, ys, zs, area, ssh, point_count, &
                        interp_degree, rank, process_count)
        real(8), intent(in) :: xs(:), ys(:), zs(:), area(:), ssh(:)
        type(cube_panel), intent(in) :: cube_panels(:)
        type(interaction_pair), intent(in) :: interaction_list(:)
        integer, intent(in) :: point_count, interp_degree, rank, process_count
        real(8), intent(inout) :: sal_longrad(:), sal_latgrad(:)
        type(c_ptr) :: g
        integer :: done
        done = 0
        if (rank == 0) then
            if (point_count /= sal_tree_n) then
                print '(a)', 'fast_sum: point_count differs from the tree''s'
                error stop 1
            end if
            g = sal_device_list(cube_panels, interaction_list, 'fast_sum')
            call sal_check(c_sal_fast_sum(g, area, ssh, int(interp_degree, c_int), sal_longrad, sal_latgrad, &
                                          1_c_int), 'fast_sum')
            done = size(interaction_list)
        end if
        print *, 'Rank: ', rank, ' interactions: ', done
    END SUBROUTINE
END MODULE
