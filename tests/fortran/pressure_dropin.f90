! A Fortran caller of the pressure backend that declares the backend exactly as the reference's
! src/solver/mod_pressure_solver.f90:23-74 does (the same bind(C) interface block: names, kinds,
! value/intent attributes) and links libdopamine_hip.so in place of the PETSc backend.
!
! Usage: pressure_dropin <input> <output>
! input (native stream): int32 n, n_faces, n_dirichlet_rows, reference_cell, max_iterations;
!   float64 rtol, atol; int32 length + characters of the solver name; int32 length + characters of
!   an option string (length 0: none); int32 local_to_global(n), face_owner(n_faces),
!   face_neighbor_local(n_faces), face_neighbor_global(n_faces), skip_face(n_faces);
!   float64 diagonal(n), rhs(n), face_coeff(n_faces); int32 dirichlet_rows(max(1, n_dir));
!   float64 dirichlet_values(max(1, n_dir)), solution(n) (the initial guess)
! output (native stream): int32 available, initialize, solve return code, iteration_count,
!   finalize; float64 residual_norm, solution(n)
program pressure_dropin
  use iso_c_binding
  implicit none

  interface
    integer(c_int) function fvm_petsc_backend_available() bind(C)
      use iso_c_binding
    end function

    integer(c_int) function fvm_petsc_backend_initialize() bind(C)
      use iso_c_binding
    end function

    integer(c_int) function fvm_petsc_backend_finalize() bind(C)
      use iso_c_binding
    end function

    subroutine fvm_petsc_set_pressure_solver(solver_name) bind(C)
      use iso_c_binding
      character(kind=c_char), intent(in) :: solver_name(*)
    end subroutine

    integer(c_int) function fvm_petsc_device_io_supported() bind(C)
      use iso_c_binding
    end function

    subroutine fvm_petsc_set_io_on_device(on) bind(C)
      use iso_c_binding
      integer(c_int), value :: on
    end subroutine

    subroutine fvm_petsc_insert_option(opt_str) bind(C)
      use iso_c_binding
      character(kind=c_char), intent(in) :: opt_str(*)
    end subroutine

    integer(c_int) function fvm_petsc_solve_pressure(n_cells_local, n_cells_global, n_faces, local_to_global, &
      face_owner, face_neighbor_local, face_neighbor_global, skip_face, diagonal, rhs, face_coeff, reference_cell, &
      max_iterations, rtol, atol, dirichlet_rows, dirichlet_values, n_dirichlet_rows, solution, iteration_count, &
      residual_norm) bind(C)
      use iso_c_binding
      integer(c_int), value :: n_cells_local, n_cells_global, n_faces, reference_cell, max_iterations, n_dirichlet_rows
      integer(c_int), intent(in) :: local_to_global(*), face_owner(*), face_neighbor_local(*), &
                                    face_neighbor_global(*), skip_face(*)
      real(c_double), intent(in) :: diagonal(*), rhs(*), face_coeff(*)
      real(c_double), value :: rtol, atol
      integer(c_int), intent(in) :: dirichlet_rows(*)
      real(c_double), intent(in) :: dirichlet_values(*)
      real(c_double), intent(out) :: solution(*)
      integer(c_int), intent(out) :: iteration_count
      real(c_double), intent(out) :: residual_norm
    end function
  end interface

  integer(c_int) :: n, nf, ndir, ref_cell, maxit, ls, lo, avail, init, rc, its, fin
  real(c_double) :: rtol, atol, resid
  character(len=256) :: fin_name, fout_name
  character(kind=c_char, len=:), allocatable :: solver_name, opt
  integer(c_int), allocatable :: l2g(:), own(:), nbl(:), nbg(:), skip(:), drows(:)
  real(c_double), allocatable :: diag(:), b(:), coeff(:), dvals(:), x(:)
  integer :: u

  if (command_argument_count() /= 2) error stop 'usage: pressure_dropin <input> <output>'
  call get_command_argument(1, fin_name)
  call get_command_argument(2, fout_name)
  open(newunit=u, file=trim(fin_name), access='stream', form='unformatted', status='old', action='read')
  read(u) n, nf, ndir, ref_cell, maxit
  read(u) rtol, atol
  read(u) ls
  allocate(character(kind=c_char, len=ls) :: solver_name)
  if (ls > 0) read(u) solver_name
  read(u) lo
  allocate(character(kind=c_char, len=lo) :: opt)
  if (lo > 0) read(u) opt
  allocate(l2g(n), own(nf), nbl(nf), nbg(nf), skip(nf), diag(n), b(n), coeff(nf), x(n))
  allocate(drows(max(1, ndir)), dvals(max(1, ndir)))
  read(u) l2g, own, nbl, nbg, skip
  read(u) diag, b, coeff
  read(u) drows
  read(u) dvals, x
  close(u)

  avail = fvm_petsc_backend_available()
  init = fvm_petsc_backend_initialize()
  call fvm_petsc_set_pressure_solver(solver_name // c_null_char)
  if (lo > 0) call fvm_petsc_insert_option(opt // c_null_char)
  call fvm_petsc_set_io_on_device(0_c_int)
  its = -1
  resid = -1.0_c_double
  rc = fvm_petsc_solve_pressure(n, n, nf, l2g, own, nbl, nbg, skip, diag, b, coeff, ref_cell, maxit, rtol, atol, &
                                drows, dvals, ndir, x, its, resid)
  fin = fvm_petsc_backend_finalize()

  open(newunit=u, file=trim(fout_name), access='stream', form='unformatted', status='replace', action='write')
  write(u) avail, init, rc, its, fin
  write(u) resid, x
  close(u)
  write(*, '(A,I0,A,I0,A,ES10.3)') 'pressure_dropin: rc ', rc, ' iterations ', its, ' residual ', resid
end program pressure_dropin
