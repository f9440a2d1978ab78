! Test infrastructure: the MPAS evaluator's per-interaction routines through the drop-in modules.
!
!     mpas_interact IN.bin OUT.bin
!
! IN.bin (stream): int32 n, thresh, degree; real64 theta; real64 x(n), y(n), z(n), area(n), ssh(n).
! Builds the tree and the list with tree_traverse and DUAL_TREE_TRAVERSAL, then sums every entry by
! calling pp_interact / pc_interact / cp_interact / cc_interact one entry at a time, as the
! reference's fast_sum dispatches them (fast_sum_module.f90:303-329), and once more with fast_sum
! itself.  OUT.bin (stream): int32 entries; real64 gx(n), gy(n) of the per-entry calls, then of fast_sum.
program mpas_interact
    use tree_module
    use fast_sum_module
    implicit none
    character(len=512) :: fin, fout
    integer :: n, thresh, degree, i, it, is, u
    real(8) :: theta
    real(8), allocatable :: xs(:), ys(:), zs(:), area(:), ssh(:), gx(:), gy(:), fx(:), fy(:)
    type(cube_panel), allocatable :: tree_panels(:)
    type(interaction_pair), allocatable :: interaction_list(:)

    call get_command_argument(1, fin)
    call get_command_argument(2, fout)
    open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
    read(u) n, thresh, degree, theta
    allocate(xs(n), ys(n), zs(n), area(n), ssh(n))
    read(u) xs, ys, zs, area, ssh
    close(u)
    call tree_traverse(tree_panels, xs, ys, zs, n, thresh, 0)
    call DUAL_TREE_TRAVERSAL(interaction_list, tree_panels, theta, thresh)
    allocate(gx(n), gy(n), fx(n), fy(n), source=0.0_8)
    do i = 1, size(interaction_list)
        it = interaction_list(i)%index_target
        is = interaction_list(i)%index_source
        select case (interaction_list(i)%interact_type)
        case (0)
            call pp_interact(gx, gy, tree_panels(it), tree_panels(is), xs, ys, zs, area, ssh, degree)
        case (1)
            call pc_interact(gx, gy, tree_panels(it), tree_panels(is), xs, ys, zs, area, ssh, degree)
        case (2)
            call cp_interact(gx, gy, tree_panels(it), tree_panels(is), xs, ys, zs, area, ssh, degree)
        case (3)
            call cc_interact(gx, gy, tree_panels(it), tree_panels(is), xs, ys, zs, area, ssh, degree)
        end select
    end do
    call fast_sum(fx, fy, interaction_list, tree_panels, xs, ys, zs, area, ssh, n, degree, 0, 1)
    open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
    write(u) int(size(interaction_list), 4), gx, gy, fx, fy
    close(u)
end program
