! Drop-in replacement of the reference's bli_module (bli_module.f90:1-114): barycentric Lagrange
! interpolation on Chebyshev points of the second kind, same names and arguments.  Host routines;
! the GPU library evaluates the same basis on the device for the sums.
MODULE BLI_MODULE
    implicit none
    private :: basis_1d

CONTAINS

    ! barycentric weights [1/2, -1, 1, ..., (-1)^d / 2]; [1] for degree 0 (bli_module.f90:6-26)
    SUBROUTINE bli_coeffs(coeffs, degree)
        real(8), allocatable, intent(out) :: coeffs(:)
        integer, intent(in) :: degree
        integer :: k
        allocate(coeffs(degree+1))
        if (degree == 0) then
            coeffs(1) = 1.0_8
            return
        end if
        do k = 0, degree
            coeffs(k+1) = merge(1.0_8, -1.0_8, mod(k, 2) == 0)
        end do
        coeffs(1) = 0.5_8 * coeffs(1)
        coeffs(degree+1) = 0.5_8 * coeffs(degree+1)
    END SUBROUTINE

    ! Chebyshev-2 nodes cos(k pi / d) mapped onto [min_x, max_x], k = 0..d (descending) (:28-48)
    SUBROUTINE bli_interp_points_shift(interp_points, min_x, max_x, degree)
        real(8), allocatable, intent(out) :: interp_points(:)
        real(8), intent(in) :: min_x, max_x
        integer, intent(in) :: degree
        real(8) :: half, centre, pi
        integer :: k
        pi = 4.D0*DATAN(1.D0)
        allocate(interp_points(degree+1))
        half = 0.5*(max_x - min_x)
        centre = 0.5*(max_x + min_x)
        if (degree == 0) then
            interp_points(1) = centre
            return
        end if
        do k = 0, degree
            interp_points(k+1) = COS(pi/degree*k)*half + centre
        end do
    END SUBROUTINE

    ! one dimension of the basis: l_k(x), exact 0/1 values when x sits on a node (:71-105)
    SUBROUTINE basis_1d(f, x, nodes, w, degree)
        real(8), intent(out) :: f(:)
        real(8), intent(in) :: x, nodes(:), w(:)
        integer, intent(in) :: degree
        real(8) :: s
        integer :: k
        f = 0.0_8
        if (any(abs(x - nodes(1:degree+1)) < 1.0e-16_8)) then
            where (abs(x - nodes(1:degree+1)) < 1.0e-16_8) f(1:degree+1) = 1.0_8
            return
        end if
        s = 0.0_8
        do k = 1, degree+1
            f(k) = w(k) / (x - nodes(k))
            s = s + f(k)
        end do
        f(1:degree+1) = f(1:degree+1) / s
    END SUBROUTINE

    ! tensor basis vals(i, j) = l_i(xi) l_j(eta) on the panel's Chebyshev grid (:50-113)
    SUBROUTINE interp_vals_bli(vals, xi, eta, min_xi, max_xi, min_eta, max_eta, degree)
        real(8), allocatable, intent(out) :: vals(:,:)
        real(8), intent(in) :: xi, eta, min_xi, max_xi, min_eta, max_eta
        integer, intent(in) :: degree
        real(8), allocatable :: nx(:), ne(:), w(:)
        real(8) :: fx(degree+1), fe(degree+1)
        integer :: j
        call bli_interp_points_shift(nx, min_xi, max_xi, degree)
        call bli_interp_points_shift(ne, min_eta, max_eta, degree)
        call bli_coeffs(w, degree)
        call basis_1d(fx, xi, nx, w, degree)
        call basis_1d(fe, eta, ne, w, degree)
        allocate(vals(degree+1, degree+1))
        do j = 1, degree+1
            vals(:, j) = fx * fe(j)
        end do
    END SUBROUTINE
END MODULE
