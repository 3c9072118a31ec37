! Fortran drop-in modules over the HIP library (libdopamine_hip.so).
!
! The reference's solver calls its numerics, SGS and wall-model routines from Fortran modules
! (src/numerics/mod_gradient.f90, mod_advection.f90, mod_viscous.f90, src/sgs/mod_vreman.f90,
! mod_amd.f90, mod_wall_model.f90) and steps through the type-bound procedures of FVSolver
! (src/solver/mod_fv_solver.f90:270-285). These modules keep those routine names, argument lists
! and array layouts (assumed-shape FP64 arrays, component fastest: U(3, n_cells),
! grad_U(3, 3, n_cells) with grad(i, j) = du_j/dx_i) and run the routines as HIP kernels through
! the library's C ABI (iso_c_binding). The mesh lives on the GPU inside a dopamine solver handle
! (dopamine_solver_create, include/dopamine_hip.h); type(Mesh) here carries that handle and the
! counts. Each routine copies its array arguments to the GPU, runs one kernel and copies the
! result back; the whole-step path (FVSolver%advance) keeps every field on the GPU across steps.
! Optional boundary-value arguments of the reference routines are required here (U_bnd,
! phi_bnd): the routines stop with a message when they are absent. The wall model's sampling
! height must be the one the solver was created with, and its filter state lives in the solver.

module dopamine_capi
  use iso_c_binding
  implicit none
  public

  interface
    function dopamine_last_error() bind(C, name='dopamine_last_error') result(p)
      import :: c_ptr
      type(c_ptr) :: p
    end function
    function dopamine_op_sizes(h, out) bind(C, name='dopamine_op_sizes') result(rc)
      import :: c_ptr, c_int, c_int64_t
      type(c_ptr), value :: h
      integer(c_int64_t), intent(out) :: out(3)
      integer(c_int) :: rc
    end function
    function dopamine_op_ls_gradient_vector(h, u, ub, g) bind(C, name='dopamine_op_ls_gradient_vector') result(rc)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: u(*), ub(*)
      real(c_double), intent(out) :: g(*)
      integer(c_int) :: rc
    end function
    function dopamine_op_ls_gradient_scalar(h, phi, pb, g) bind(C, name='dopamine_op_ls_gradient_scalar') result(rc)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: phi(*), pb(*)
      real(c_double), intent(out) :: g(*)
      integer(c_int) :: rc
    end function
    function dopamine_op_gg_gradient_vector(h, u, ub, gprev, g) bind(C, name='dopamine_op_gg_gradient_vector') &
        result(rc)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: h, gprev
      real(c_double), intent(in) :: u(*), ub(*)
      real(c_double), intent(out) :: g(*)
      integer(c_int) :: rc
    end function
    function dopamine_op_gg_gradient_scalar(h, phi, pb, gprev, g) bind(C, name='dopamine_op_gg_gradient_scalar') &
        result(rc)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: h, gprev
      real(c_double), intent(in) :: phi(*), pb(*)
      real(c_double), intent(out) :: g(*)
      integer(c_int) :: rc
    end function
    function dopamine_op_sgs_viscosity(h, model, u, g, coef, nu, wall_gradient, dt_stab, nus) &
        bind(C, name='dopamine_op_sgs_viscosity') result(rc)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: model, wall_gradient
      real(c_double), intent(in) :: u(*), g(*)
      real(c_double), value :: coef, nu, dt_stab
      real(c_double), intent(out) :: nus(*)
      integer(c_int) :: rc
    end function
    function dopamine_op_dynamic_smagorinsky(h, u, g, ub, dt_stab, nus) &
        bind(C, name='dopamine_op_dynamic_smagorinsky') result(rc)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: u(*), g(*), ub(*)
      real(c_double), value :: dt_stab
      real(c_double), intent(out) :: nus(*)
      integer(c_int) :: rc
    end function
    function dopamine_op_wall_model(h, u, nu, h_wm, t_filt, dt, tf_fac, tau) bind(C, name='dopamine_op_wall_model') &
        result(rc)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: u(*)
      real(c_double), value :: nu, h_wm, t_filt, dt, tf_fac
      real(c_double), intent(out) :: tau(*)
      integer(c_int) :: rc
    end function
    function dopamine_op_advection_rhs_vector(h, scheme, u, ub, ff, g, rhs) &
        bind(C, name='dopamine_op_advection_rhs_vector') result(rc)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: h, g
      integer(c_int), value :: scheme
      real(c_double), intent(in) :: u(*), ub(*), ff(*)
      real(c_double), intent(out) :: rhs(*)
      integer(c_int) :: rc
    end function
    function dopamine_op_viscous_rhs_vector(h, nu, u, nus, g, ub, rhs) &
        bind(C, name='dopamine_op_viscous_rhs_vector') result(rc)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: h
      real(c_double), value :: nu
      real(c_double), intent(in) :: u(*), nus(*), g(*), ub(*)
      real(c_double), intent(inout) :: rhs(*)
      integer(c_int) :: rc
    end function
    function dopamine_op_advection_rhs_scalar(h, scheme, phi, u, pb, ub, ff, g, rhs) &
        bind(C, name='dopamine_op_advection_rhs_scalar') result(rc)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: h, g
      integer(c_int), value :: scheme
      real(c_double), intent(in) :: phi(*), u(*), pb(*), ub(*), ff(*)
      real(c_double), intent(out) :: rhs(*)
      integer(c_int) :: rc
    end function
    function dopamine_op_viscous_rhs_scalar(h, phi, alpha_mol, nus, pr_t, g, pb, rhs) &
        bind(C, name='dopamine_op_viscous_rhs_scalar') result(rc)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: h
      real(c_double), value :: alpha_mol, pr_t
      real(c_double), intent(in) :: phi(*), nus(*), g(*), pb(*)
      real(c_double), intent(inout) :: rhs(*)
      integer(c_int) :: rc
    end function
    ! the solver handle (FVSolver's type-bound procedures)
    function dopamine_solver_advance(h, dt_req, dt_used) bind(C, name='dopamine_solver_advance') result(rc)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: h
      real(c_double), value :: dt_req
      real(c_double), intent(out) :: dt_used
      integer(c_int) :: rc
    end function
    function dopamine_solver_project_once(h, dt) bind(C, name='dopamine_solver_project_once') result(rc)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: h
      real(c_double), value :: dt
      integer(c_int) :: rc
    end function
    function dopamine_solver_compute_dt(h) bind(C, name='dopamine_solver_compute_dt') result(dt)
      import :: c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double) :: dt
    end function
    function dopamine_solver_state_is_finite(h) bind(C, name='dopamine_solver_state_is_finite') result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: h
      integer(c_int) :: rc
    end function
    function dopamine_solver_time(h) bind(C, name='dopamine_solver_time') result(t)
      import :: c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double) :: t
    end function
    function dopamine_solver_get_field(h, name, host, n) bind(C, name='dopamine_solver_get_field') result(rc)
      import :: c_ptr, c_int, c_char, c_double, c_int64_t
      type(c_ptr), value :: h
      character(kind=c_char), intent(in) :: name(*)
      real(c_double), intent(out) :: host(*)
      integer(c_int64_t), value :: n
      integer(c_int) :: rc
    end function
    function dopamine_solver_set_field(h, name, host) bind(C, name='dopamine_solver_set_field') result(rc)
      import :: c_ptr, c_int, c_char, c_double
      type(c_ptr), value :: h
      character(kind=c_char), intent(in) :: name(*)
      real(c_double), intent(in) :: host(*)
      integer(c_int) :: rc
    end function
    function dopamine_solver_stats(h, out) bind(C, name='dopamine_solver_stats') result(rc)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: out(8)
      integer(c_int) :: rc
    end function
    function c_strlen(p) bind(C, name='strlen') result(n)
      import :: c_ptr, c_size_t
      type(c_ptr), value :: p
      integer(c_size_t) :: n
    end function
  end interface

contains

  !> Text of the library's last error.
  function last_error() result(msg)
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: s(:)
    integer :: i, n
    p = dopamine_last_error()
    n = int(c_strlen(p))
    call c_f_pointer(p, s, [max(n, 1)])
    allocate(character(len=n) :: msg)
    do i = 1, n
      msg(i:i) = s(i)
    end do
  end function

  !> Stops with the routine's name and the library's message on a non-zero return.
  subroutine check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    if (rc /= 0) then
      write(*, '(A,A,A,I0,A,A)') 'dopamine: ', what, ' failed (', rc, '): ', last_error()
      error stop 'dopamine: HIP routine failed'
    end if
  end subroutine

end module dopamine_capi


module mod_mesh
  use iso_c_binding
  use iso_fortran_env, only: real64, int32
  use dopamine_capi
  implicit none
  private
  integer, parameter, public :: rp = real64, ip = int32
  ! src/mod_parameters.f90:54-60
  integer(ip), parameter, public :: SCHEME_VANLEER_TVD = 1, SCHEME_KOREN_TVD = 2, SCHEME_CDS = 3, &
                                    SCHEME_CENTRAL = 3, SCHEME_KEP = 4, SCHEME_SKEW = 4, SCHEME_QUICK = 5

  !> The mesh as the HIP library holds it: a solver handle (include/dopamine_hip.h) whose device
  !> mesh the routines run on, and its counts.
  type, public :: Mesh
    type(c_ptr) :: handle = c_null_ptr
    integer(ip) :: n_cells = 0, n_faces = 0, n_bnd_faces = 0
  contains
    procedure :: attach => mesh_attach
  end type

contains

  subroutine mesh_attach(self, handle)
    class(Mesh), intent(inout) :: self
    type(c_ptr), intent(in) :: handle
    integer(c_int64_t) :: n(3)
    self%handle = handle
    call check(dopamine_op_sizes(handle, n), 'mesh_attach')
    self%n_cells = int(n(1), ip)
    self%n_faces = int(n(2), ip)
    self%n_bnd_faces = int(n(3), ip)
  end subroutine

end module mod_mesh


module mod_gradient
  use iso_c_binding
  use dopamine_capi
  use mod_mesh
  implicit none
  private
  public :: ls_gradient_vector, ls_gradient_scalar, green_gauss_gradient_vector, green_gauss_gradient_scalar

contains

  ! src/numerics/mod_gradient.f90:425
  subroutine ls_gradient_vector(U, msh, grad_U, U_bnd)
    real(rp), intent(in) :: U(:,:)
    type(Mesh), intent(in), target :: msh
    real(rp), intent(inout) :: grad_U(:,:,:)
    real(rp), intent(in), optional, target :: U_bnd(:,:)
    if (.not. present(U_bnd)) error stop 'ls_gradient_vector: U_bnd is required'
    call check(dopamine_op_ls_gradient_vector(msh%handle, U, U_bnd, grad_U), 'ls_gradient_vector')
  end subroutine

  ! src/numerics/mod_gradient.f90:307
  subroutine ls_gradient_scalar(phi, msh, grad_phi, phi_bnd)
    real(rp), intent(in) :: phi(:)
    type(Mesh), intent(in), target :: msh
    real(rp), intent(inout) :: grad_phi(:,:)
    real(rp), intent(in), optional, target :: phi_bnd(:)
    if (.not. present(phi_bnd)) error stop 'ls_gradient_scalar: phi_bnd is required'
    call check(dopamine_op_ls_gradient_scalar(msh%handle, phi, phi_bnd, grad_phi), 'ls_gradient_scalar')
  end subroutine

  ! src/numerics/mod_gradient.f90:172
  subroutine green_gauss_gradient_vector(U, msh, grad_U, U_bnd, grad_U_prev)
    real(rp), intent(in) :: U(:,:)
    type(Mesh), intent(in), target :: msh
    real(rp), intent(inout) :: grad_U(:,:,:)
    real(rp), intent(in), optional, target :: U_bnd(:,:)
    real(rp), intent(in), optional, target :: grad_U_prev(:,:,:)
    real(rp), allocatable, target :: prev(:)
    type(c_ptr) :: pp
    if (.not. present(U_bnd)) error stop 'green_gauss_gradient_vector: U_bnd is required'
    pp = c_null_ptr
    if (present(grad_U_prev)) then
      prev = reshape(grad_U_prev, [size(grad_U_prev)])
      pp = c_loc(prev)
    end if
    call check(dopamine_op_gg_gradient_vector(msh%handle, U, U_bnd, pp, grad_U), 'green_gauss_gradient_vector')
  end subroutine

  ! src/numerics/mod_gradient.f90:38
  subroutine green_gauss_gradient_scalar(phi, msh, grad_phi, phi_bnd, grad_phi_prev)
    real(rp), intent(in) :: phi(:)
    type(Mesh), intent(in), target :: msh
    real(rp), intent(inout) :: grad_phi(:,:)
    real(rp), intent(in), optional, target :: phi_bnd(:)
    real(rp), intent(in), optional, target :: grad_phi_prev(:,:)
    real(rp), allocatable, target :: prev(:)
    type(c_ptr) :: pp
    if (.not. present(phi_bnd)) error stop 'green_gauss_gradient_scalar: phi_bnd is required'
    pp = c_null_ptr
    if (present(grad_phi_prev)) then
      prev = reshape(grad_phi_prev, [size(grad_phi_prev)])
      pp = c_loc(prev)
    end if
    call check(dopamine_op_gg_gradient_scalar(msh%handle, phi, phi_bnd, pp, grad_phi), 'green_gauss_gradient_scalar')
  end subroutine

end module mod_gradient


module mod_advection
  use iso_c_binding
  use dopamine_capi
  use mod_mesh
  implicit none
  private
  public :: advection_rhs_vector, advection_rhs_scalar

contains

  ! src/numerics/mod_advection.f90:679; inv_vol is accepted and unused (the mesh holds 1/V)
  subroutine advection_rhs_vector(U, msh, scheme, rhs, U_bnd, face_flux, inv_vol, grad_U)
    real(rp), intent(in) :: U(:,:)
    type(Mesh), intent(in) :: msh
    integer(ip), intent(in) :: scheme
    real(rp), intent(inout) :: rhs(:,:)
    real(rp), intent(in), optional :: U_bnd(:,:)
    real(rp), intent(in), optional :: face_flux(:)
    real(rp), intent(in), optional :: inv_vol(:)
    real(rp), intent(in), optional :: grad_U(:,:,:)
    real(rp), allocatable, target :: g(:)
    type(c_ptr) :: gp
    if (.not. present(U_bnd) .or. .not. present(face_flux)) &
      error stop 'advection_rhs_vector: U_bnd and face_flux are required'
    if (present(inv_vol)) continue
    gp = c_null_ptr
    if (present(grad_U)) then
      g = reshape(grad_U, [size(grad_U)])
      gp = c_loc(g)
    end if
    call check(dopamine_op_advection_rhs_vector(msh%handle, int(scheme, c_int), U, U_bnd, face_flux, gp, rhs), &
               'advection_rhs_vector')
  end subroutine

  ! src/numerics/mod_advection.f90:56
  subroutine advection_rhs_scalar(phi, U, msh, scheme, rhs, phi_bnd, U_bnd, face_flux, inv_vol, grad_phi)
    real(rp), intent(in) :: phi(:)
    real(rp), intent(in) :: U(:,:)
    type(Mesh), intent(in) :: msh
    integer(ip), intent(in) :: scheme
    real(rp), intent(inout) :: rhs(:)
    real(rp), intent(in), optional :: phi_bnd(:)
    real(rp), intent(in), optional :: U_bnd(:,:)
    real(rp), intent(in), optional :: face_flux(:)
    real(rp), intent(in), optional :: inv_vol(:)
    real(rp), intent(in), optional :: grad_phi(:,:)
    real(rp), allocatable, target :: g(:)
    type(c_ptr) :: gp
    if (.not. present(phi_bnd) .or. .not. present(U_bnd) .or. .not. present(face_flux)) &
      error stop 'advection_rhs_scalar: phi_bnd, U_bnd and face_flux are required'
    if (present(inv_vol)) continue
    gp = c_null_ptr
    if (present(grad_phi)) then
      g = reshape(grad_phi, [size(grad_phi)])
      gp = c_loc(g)
    end if
    call check(dopamine_op_advection_rhs_scalar(msh%handle, int(scheme, c_int), phi, U, phi_bnd, U_bnd, face_flux, &
                                                gp, rhs), 'advection_rhs_scalar')
  end subroutine

end module mod_advection


module mod_viscous
  use iso_c_binding
  use dopamine_capi
  use mod_mesh
  implicit none
  private
  public :: viscous_rhs_vector, viscous_rhs_scalar

contains

  ! src/numerics/mod_viscous.f90:37 (adds the viscous term to rhs)
  subroutine viscous_rhs_vector(U, nu, nu_sgs, grad_U, msh, rhs, U_bnd, inv_vol)
    real(rp), intent(in) :: U(:,:)
    real(rp), intent(in) :: nu
    real(rp), intent(in) :: nu_sgs(:)
    real(rp), intent(in) :: grad_U(:,:,:)
    type(Mesh), intent(in), target :: msh
    real(rp), intent(inout) :: rhs(:,:)
    real(rp), intent(in), optional, target :: U_bnd(:,:)
    real(rp), intent(in), optional, target :: inv_vol(:)
    if (.not. present(U_bnd)) error stop 'viscous_rhs_vector: U_bnd is required'
    if (present(inv_vol)) continue
    call check(dopamine_op_viscous_rhs_vector(msh%handle, nu, U, nu_sgs, grad_U, U_bnd, rhs), 'viscous_rhs_vector')
  end subroutine

  ! src/numerics/mod_viscous.f90:286 (adds (alpha + nu_t / Pr_t) diffusion to rhs)
  subroutine viscous_rhs_scalar(phi, alpha_mol, nu_sgs, Pr_t, grad_phi, msh, rhs, phi_bnd, inv_vol)
    real(rp), intent(in) :: phi(:)
    real(rp), intent(in) :: alpha_mol
    real(rp), intent(in) :: nu_sgs(:)
    real(rp), intent(in) :: Pr_t
    real(rp), intent(in) :: grad_phi(:,:)
    type(Mesh), intent(in), target :: msh
    real(rp), intent(inout) :: rhs(:)
    real(rp), intent(in), optional, target :: phi_bnd(:)
    real(rp), intent(in), optional, target :: inv_vol(:)
    if (.not. present(phi_bnd)) error stop 'viscous_rhs_scalar: phi_bnd is required'
    if (present(inv_vol)) continue
    call check(dopamine_op_viscous_rhs_scalar(msh%handle, phi, alpha_mol, nu_sgs, Pr_t, grad_phi, phi_bnd, rhs), &
               'viscous_rhs_scalar')
  end subroutine

end module mod_viscous


module mod_vreman
  use iso_c_binding
  use dopamine_capi
  use mod_mesh
  use mod_gradient, only: ls_gradient_vector
  implicit none
  private
  real(rp), parameter, public :: C_VREMAN = 0.025_rp  ! src/mod_parameters.f90:19-20
  public :: vreman_sgs_viscosity, smagorinsky_sgs

contains

  ! src/sgs/mod_vreman.f90:50; without grad_U_ext the LS gradient of U is taken first
  subroutine vreman_sgs_viscosity(U, msh, nu_sgs, U_bnd, grad_U_ext, dt_stab)
    real(rp), intent(in) :: U(:,:)
    type(Mesh), intent(inout), target :: msh
    real(rp), intent(inout) :: nu_sgs(:)
    real(rp), intent(in), optional :: U_bnd(:,:)
    real(rp), intent(in), optional, target :: grad_U_ext(:,:,:)
    real(rp), intent(in), optional :: dt_stab
    real(rp), allocatable :: g(:,:,:)
    real(rp) :: dts
    dts = 0.0_rp
    if (present(dt_stab)) dts = dt_stab
    if (present(grad_U_ext)) then
      call check(dopamine_op_sgs_viscosity(msh%handle, 1_c_int, U, grad_U_ext, C_VREMAN, 0.0_rp, 0_c_int, dts, nu_sgs), &
                 'vreman_sgs_viscosity')
    else
      allocate(g(3, 3, msh%n_cells))
      call ls_gradient_vector(U, msh, g, U_bnd)
      call check(dopamine_op_sgs_viscosity(msh%handle, 1_c_int, U, g, C_VREMAN, 0.0_rp, 0_c_int, dts, nu_sgs), &
                 'vreman_sgs_viscosity')
    end if
  end subroutine

  ! src/sgs/mod_vreman.f90:295
  subroutine smagorinsky_sgs(grad_U, msh, nu_sgs, cs_in, dt_stab)
    real(rp), intent(in) :: grad_U(:,:,:)
    type(Mesh), intent(inout) :: msh
    real(rp), intent(inout) :: nu_sgs(:)
    real(rp), intent(in), optional :: cs_in
    real(rp), intent(in), optional :: dt_stab
    real(rp) :: cs, dts
    cs = 0.1_rp
    if (present(cs_in)) cs = cs_in
    dts = 0.0_rp
    if (present(dt_stab)) dts = dt_stab
    call check(dopamine_op_sgs_viscosity(msh%handle, 2_c_int, grad_U, grad_U, cs, 0.0_rp, 0_c_int, dts, nu_sgs), &
               'smagorinsky_sgs')
  end subroutine

end module mod_vreman


module mod_dynamic_smagorinsky
  use iso_c_binding
  use dopamine_capi
  use mod_mesh
  implicit none
  private
  public :: dynamic_smagorinsky_sgs

contains

  ! src/sgs/mod_dynamic_smagorinsky.f90:39 (U_bnd required; without dt_stab no cap)
  subroutine dynamic_smagorinsky_sgs(U, grad_U, msh, nu_sgs, U_bnd, dt_stab)
    real(rp), intent(in) :: U(:,:)
    real(rp), intent(in) :: grad_U(:,:,:)
    type(Mesh), intent(inout) :: msh
    real(rp), intent(inout) :: nu_sgs(:)
    real(rp), intent(in), optional :: U_bnd(:,:)
    real(rp), intent(in), optional :: dt_stab
    real(rp) :: dts
    if (.not. present(U_bnd)) error stop 'dynamic_smagorinsky_sgs: U_bnd is required'
    dts = 0.0_rp
    if (present(dt_stab)) dts = dt_stab
    call check(dopamine_op_dynamic_smagorinsky(msh%handle, U, grad_U, U_bnd, dts, nu_sgs), 'dynamic_smagorinsky_sgs')
  end subroutine

end module mod_dynamic_smagorinsky


module mod_amd
  use iso_c_binding
  use dopamine_capi
  use mod_mesh
  use mod_gradient, only: ls_gradient_vector
  implicit none
  private
  public :: amd_sgs_viscosity

contains

  ! src/sgs/mod_amd.f90:69
  subroutine amd_sgs_viscosity(U, msh, nu_sgs, nu, c_amd, wall_gradient, U_bnd, grad_U_ext, dt_stab)
    real(rp), intent(in) :: U(:,:)
    type(Mesh), intent(inout), target :: msh
    real(rp), intent(inout) :: nu_sgs(:)
    real(rp), intent(in) :: nu
    real(rp), intent(in) :: c_amd
    logical, intent(in) :: wall_gradient
    real(rp), intent(in), optional :: U_bnd(:,:)
    real(rp), intent(in), optional, target :: grad_U_ext(:,:,:)
    real(rp), intent(in), optional :: dt_stab
    real(rp), allocatable :: g(:,:,:)
    real(rp) :: dts
    integer(c_int) :: wg
    dts = 0.0_rp
    if (present(dt_stab)) dts = dt_stab
    wg = 0
    if (wall_gradient) wg = 1
    if (present(grad_U_ext)) then
      call check(dopamine_op_sgs_viscosity(msh%handle, 3_c_int, U, grad_U_ext, c_amd, nu, wg, dts, nu_sgs), &
                 'amd_sgs_viscosity')
    else
      allocate(g(3, 3, msh%n_cells))
      call ls_gradient_vector(U, msh, g, U_bnd)
      call check(dopamine_op_sgs_viscosity(msh%handle, 3_c_int, U, g, c_amd, nu, wg, dts, nu_sgs), 'amd_sgs_viscosity')
    end if
  end subroutine

end module mod_amd


module mod_wall_model
  use iso_c_binding
  use dopamine_capi
  use mod_mesh
  implicit none
  private
  public :: apply_wall_model_bc

contains

  ! src/sgs/mod_wall_model.f90:176; h_wm must equal the solver's wall_model_height (the sampling
  ! cells are found when the solver is created); the filter state persists in the solver
  subroutine apply_wall_model_bc(U, nu, msh, tau_w_bnd, h_wm, t_filt, dt, tf_fac)
    real(rp), intent(in) :: U(:,:)
    real(rp), intent(in) :: nu
    type(Mesh), intent(in), target :: msh
    real(rp), intent(inout) :: tau_w_bnd(:,:)
    real(rp), intent(in), optional :: h_wm
    real(rp), intent(in), optional :: t_filt
    real(rp), intent(in), optional :: dt
    real(rp), intent(in), optional :: tf_fac
    real(rp) :: h_use, t_use, dt_use, tf_use
    h_use = 0.0_rp; if (present(h_wm)) h_use = h_wm
    t_use = 0.0_rp; if (present(t_filt)) t_use = t_filt
    dt_use = 0.0_rp; if (present(dt)) dt_use = dt
    tf_use = 0.0_rp; if (present(tf_fac)) tf_use = tf_fac
    call check(dopamine_op_wall_model(msh%handle, U, nu, h_use, t_use, dt_use, tf_use, tau_w_bnd), &
               'apply_wall_model_bc')
  end subroutine

end module mod_wall_model


module mod_fv_solver
  use iso_c_binding
  use dopamine_capi
  use mod_mesh
  implicit none
  private

  !> The RK3W solver on the GPU (type-bound procedures of src/solver/mod_fv_solver.f90:270-285):
  !> every field stays in HBM; sync_host copies U and p out.
  type, public :: FVSolver
    type(c_ptr) :: handle = c_null_ptr
    type(Mesh) :: mesh
    real(rp) :: time = 0.0_rp
  contains
    procedure :: attach => solver_attach
    procedure :: advance => solver_advance
    procedure :: project_once => solver_project_once
    procedure :: compute_dt => solver_compute_dt
    procedure :: state_is_finite => solver_state_is_finite
    procedure :: sync_host => solver_sync_host
    procedure :: to_device => solver_to_device
  end type

contains

  subroutine solver_attach(self, handle)
    class(FVSolver), intent(inout) :: self
    type(c_ptr), intent(in) :: handle
    self%handle = handle
    call self%mesh%attach(handle)
    self%time = dopamine_solver_time(handle)
  end subroutine

  subroutine solver_advance(self, dt)
    class(FVSolver), intent(inout) :: self
    real(rp), intent(in) :: dt
    real(c_double) :: used
    call check(dopamine_solver_advance(self%handle, dt, used), 'advance')
    self%time = dopamine_solver_time(self%handle)
  end subroutine

  subroutine solver_project_once(self, dt)
    class(FVSolver), intent(inout) :: self
    real(rp), intent(in) :: dt
    call check(dopamine_solver_project_once(self%handle, dt), 'project_once')
  end subroutine

  real(rp) function solver_compute_dt(self) result(dt)
    class(FVSolver), intent(inout) :: self
    dt = dopamine_solver_compute_dt(self%handle)
  end function

  logical function solver_state_is_finite(self) result(ok)
    class(FVSolver), intent(in) :: self
    ok = dopamine_solver_state_is_finite(self%handle) == 1
  end function

  subroutine solver_sync_host(self, U, p)
    class(FVSolver), intent(in) :: self
    real(rp), intent(out) :: U(:,:), p(:)
    call check(dopamine_solver_get_field(self%handle, 'U' // c_null_char, U, int(size(U), c_int64_t)), 'sync_host')
    call check(dopamine_solver_get_field(self%handle, 'p' // c_null_char, p, int(size(p), c_int64_t)), 'sync_host')
  end subroutine

  subroutine solver_to_device(self, U)
    class(FVSolver), intent(inout) :: self
    real(rp), intent(in) :: U(:,:)
    call check(dopamine_solver_set_field(self%handle, 'U' // c_null_char, U), 'to_device')
  end subroutine

end module mod_fv_solver
