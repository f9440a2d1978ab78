! Test infrastructure: the MOM6 evaluator's per-interaction routines through the drop-in modules.
!
!     mom6_interact IN.bin OUT.bin
!
! IN.bin as for mpas_interact.  One rank holding every point, as the MOM6 driver sets it up
! (mom6_code.f90:177-276): tree_traverse, assign_points_to_panels, SINGLE_TREE_TRAVERSAL,
! proxy_source_compute; then every pc entry through pc_interact with its panel's slice of the
! weights and every pp entry through pp_interact, one entry at a time as pc_compute and pp_compute
! loop them (mom6_fast_sum_module.f90:153-192), and once more through pc_compute and pp_compute.
! OUT.bin (stream): int32 pc entries, pp entries; real64 gx(n), gy(n) of the per-entry calls, then of
! the list routines.
program mom6_interact
    use tree_module
    use mom6_fast_sum_module
    implicit none
    character(len=512) :: fin, fout
    integer :: n, thresh, degree, i, i_s, i_t, m, u, max_level
    real(8) :: theta
    real(8), allocatable :: xs(:), ys(:), zs(:), area(:), ssh(:), gx(:), gy(:), fx(:), fy(:), w(:), e(:)
    integer, allocatable :: points_panels(:,:)
    type(cube_panel), allocatable :: tree_panels(:)
    type(interaction_pair), allocatable :: pc_list(:), pp_list(:)

    call get_command_argument(1, fin)
    call get_command_argument(2, fout)
    open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
    read(u) n, thresh, degree, theta
    allocate(xs(n), ys(n), zs(n), area(n), ssh(n), e(0))
    read(u) xs, ys, zs, area, ssh
    close(u)
    call tree_traverse(tree_panels, xs, ys, zs, n, thresh, 0)
    max_level = tree_panels(size(tree_panels))%level
    allocate(points_panels(max_level + 1, n), source=-1)
    call assign_points_to_panels(xs, ys, zs, points_panels, tree_panels, n)
    call SINGLE_TREE_TRAVERSAL(pc_list, pp_list, tree_panels, theta, thresh, xs, ys, zs, n)
    m = (degree + 1) * (degree + 1)
    allocate(w(m * size(tree_panels)), source=0.0_8)
    call proxy_source_compute(w, tree_panels, points_panels, xs, ys, zs, ssh, area, degree, n)
    allocate(gx(n), gy(n), fx(n), fy(n), source=0.0_8)
    do i = 1, size(pc_list)
        i_s = pc_list(i)%index_source
        i_t = pc_list(i)%index_target
        call pc_interact(gx, gy, i_t, tree_panels(i_s), xs(i_t), ys(i_t), zs(i_t), degree, w(m * (i_s - 1) + 1:m * i_s))
    end do
    do i = 1, size(pp_list)
        call pp_interact(gx, gy, pp_list(i)%index_target, tree_panels(pp_list(i)%index_source), xs, ys, zs, area, ssh, &
                         e, e, e, e, e, n)
    end do
    call pc_compute(fx, fy, pc_list, tree_panels, xs, ys, zs, degree, w)
    call pp_compute(fx, fy, pp_list, tree_panels, xs, ys, zs, area, ssh, e, e, e, e, e, n)
    open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
    write(u) int(size(pc_list), 4), int(size(pp_list), 4), gx, gy, fx, fy
    close(u)
end program
