! Drop-in replacement of the reference's cube_sphere_module (cube_sphere_module.f90:1-96): the
! equiangular gnomonic cubed-sphere maps, same names, arguments and results.  Host scalar
! functions; the GPU library carries its own device copies of the same maps.
MODULE CUBE_SPHERE_MODULE
    implicit none

CONTAINS

    ! face 1..6 of a point: the largest |coordinate| wins, ties go x, then y, then z; the sign picks
    ! the + face (1 x, 2 y, 5 z) or the - face (3, 4, 6) (cube_sphere_module.f90:7-35)
    INTEGER FUNCTION face_from_xyz(x, y, z)
        real(8), intent(in) :: x, y, z
        if ((abs(x) >= abs(y)) .and. (abs(x) >= abs(z))) then
            face_from_xyz = merge(1, 3, x >= 0)
        else if ((abs(y) >= abs(x)) .and. (abs(y) >= abs(z))) then
            face_from_xyz = merge(2, 4, y >= 0)
        else
            face_from_xyz = merge(5, 6, z >= 0)
        end if
    END FUNCTION

    ! face-local angles (xi, eta) of a unit vector; in_face overrides the face (:38-61)
    SUBROUTINE xieta_from_xyz(x, y, z, xi, eta, in_face)
        real(8), intent(in) :: x, y, z
        integer, intent(in), optional :: in_face
        real(8), intent(out) :: xi, eta
        integer :: f
        if (present(in_face)) then
            f = in_face
        else
            f = face_from_xyz(x, y, z)
        end if
        select case (f)
        case (1)
            xi = atan(y / x); eta = atan(z / x)
        case (2)
            xi = atan(-x / y); eta = atan(z / y)
        case (3)
            xi = atan(y / x); eta = atan(-z / x)
        case (4)
            xi = atan(-x / y); eta = atan(-z / y)
        case (5)
            xi = atan(y / z); eta = atan(-x / z)
        case (6)
            xi = atan(-y / z); eta = atan(-x / z)
        end select
    END SUBROUTINE

    ! unit vector of (xi, eta) on a face (:64-95)
    SUBROUTINE xyz_from_xieta(x, y, z, xi, eta, face)
        real(8), intent(in) :: xi, eta
        integer, intent(in) :: face
        real(8), intent(out) :: x, y, z
        real(8) :: tx, te, s
        tx = tan(xi)
        te = tan(eta)
        s = sqrt(1 + tx*tx + te*te)
        select case (face)
        case (1)
            x = 1.0_8 / s; y = tx*x; z = te*x
        case (2)
            y = 1.0_8 / s; x = -tx*y; z = te*y
        case (3)
            x = -1.0_8 / s; y = tx*x; z = -te*x
        case (4)
            y = -1.0_8 / s; x = -tx*y; z = -te*y
        case (5)
            z = 1.0_8 / s; x = -te*z; y = tx*z
        case default
            z = -1.0_8 / s; x = -te*z; y = -tx*z
        end select
    END SUBROUTINE
END MODULE
