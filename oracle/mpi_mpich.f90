! Test infrastructure: the `mpi` module the reference drivers `use`, made from the MPI library
! installed in this image (MPICH's own mpif.h; the library itself is /opt/conda/lib/libmpi*).
module mpi
    include 'mpif.h'
end module mpi
