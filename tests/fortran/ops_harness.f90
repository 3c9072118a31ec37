! bind(C) harness over the Fortran drop-in modules (dopamine_amd/fortran/dopamine_modules.f90), so
! the tests can call the Fortran routines (mod_gradient, mod_advection, mod_viscous, mod_vreman,
! mod_amd, mod_wall_model, FVSolver) from Python with a solver handle and flat FP64 arrays. Every
! routine builds the Fortran arrays with the reference's shapes (U(3, n_cells), grad_U(3, 3, n_cells),
! U_bnd(3, n_bnd_faces)) and calls the module routine with its reference argument list. A null
! pointer for an optional argument leaves it absent.
module ops_harness
  use iso_c_binding
  use mod_mesh
  use mod_gradient
  use mod_advection
  use mod_viscous
  use mod_vreman
  use mod_amd
  use mod_dynamic_smagorinsky
  use mod_wall_model
  use mod_fv_solver
  implicit none

contains

  function ft_gradient_vector(h, gg, iterations, pu, pub, pg) bind(C, name='ft_gradient_vector') result(rc)
    type(c_ptr), value :: h, pu, pub, pg
    integer(c_int), value :: gg, iterations
    integer(c_int) :: rc
    type(Mesh) :: msh
    real(rp), pointer :: U(:,:), Ub(:,:), g(:,:,:)
    real(rp), allocatable :: prev(:,:,:)
    integer :: it
    call msh%attach(h)
    call c_f_pointer(pu, U, [3, msh%n_cells])
    call c_f_pointer(pub, Ub, [3, max(1, msh%n_bnd_faces)])
    call c_f_pointer(pg, g, [3, 3, msh%n_cells])
    if (gg == 0) then
      call ls_gradient_vector(U, msh, g, Ub)
    else
      call green_gauss_gradient_vector(U, msh, g, Ub)
      do it = 2, iterations
        prev = g
        call green_gauss_gradient_vector(U, msh, g, Ub, prev)
      end do
    end if
    rc = 0
  end function

  function ft_gradient_scalar(h, gg, iterations, pphi, ppb, pg) bind(C, name='ft_gradient_scalar') result(rc)
    type(c_ptr), value :: h, pphi, ppb, pg
    integer(c_int), value :: gg, iterations
    integer(c_int) :: rc
    type(Mesh) :: msh
    real(rp), pointer :: phi(:), pb(:), g(:,:)
    real(rp), allocatable :: prev(:,:)
    integer :: it
    call msh%attach(h)
    call c_f_pointer(pphi, phi, [msh%n_cells])
    call c_f_pointer(ppb, pb, [max(1, msh%n_bnd_faces)])
    call c_f_pointer(pg, g, [3, msh%n_cells])
    if (gg == 0) then
      call ls_gradient_scalar(phi, msh, g, pb)
    else
      call green_gauss_gradient_scalar(phi, msh, g, pb)
      do it = 2, iterations
        prev = g
        call green_gauss_gradient_scalar(phi, msh, g, pb, prev)
      end do
    end if
    rc = 0
  end function

  function ft_advection_vector(h, scheme, pu, pub, pff, pgu, prhs) bind(C, name='ft_advection_vector') result(rc)
    type(c_ptr), value :: h, pu, pub, pff, pgu, prhs
    integer(c_int), value :: scheme
    integer(c_int) :: rc
    type(Mesh) :: msh
    real(rp), pointer :: U(:,:), Ub(:,:), ff(:), gU(:,:,:), rhs(:,:)
    real(rp), allocatable :: inv_vol(:)
    call msh%attach(h)
    call c_f_pointer(pu, U, [3, msh%n_cells])
    call c_f_pointer(pub, Ub, [3, max(1, msh%n_bnd_faces)])
    call c_f_pointer(pff, ff, [msh%n_faces])
    call c_f_pointer(prhs, rhs, [3, msh%n_cells])
    allocate(inv_vol(msh%n_cells))
    inv_vol = 0.0_rp  ! accepted and unused: the device mesh holds 1/V
    if (c_associated(pgu)) then
      call c_f_pointer(pgu, gU, [3, 3, msh%n_cells])
      call advection_rhs_vector(U, msh, int(scheme, ip), rhs, Ub, ff, inv_vol, gU)
    else
      call advection_rhs_vector(U, msh, int(scheme, ip), rhs, Ub, ff, inv_vol)
    end if
    rc = 0
  end function

  function ft_viscous_vector(h, nu, pu, pnus, pgu, pub, prhs) bind(C, name='ft_viscous_vector') result(rc)
    type(c_ptr), value :: h, pu, pnus, pgu, pub, prhs
    real(c_double), value :: nu
    integer(c_int) :: rc
    type(Mesh) :: msh
    real(rp), pointer :: U(:,:), nus(:), gU(:,:,:), Ub(:,:), rhs(:,:)
    call msh%attach(h)
    call c_f_pointer(pu, U, [3, msh%n_cells])
    call c_f_pointer(pnus, nus, [msh%n_cells])
    call c_f_pointer(pgu, gU, [3, 3, msh%n_cells])
    call c_f_pointer(pub, Ub, [3, max(1, msh%n_bnd_faces)])
    call c_f_pointer(prhs, rhs, [3, msh%n_cells])
    call viscous_rhs_vector(U, nu, nus, gU, msh, rhs, Ub)
    rc = 0
  end function

  function ft_advection_scalar(h, scheme, pphi, pu, ppb, pub, pff, pg, prhs) bind(C, name='ft_advection_scalar') &
      result(rc)
    type(c_ptr), value :: h, pphi, pu, ppb, pub, pff, pg, prhs
    integer(c_int), value :: scheme
    integer(c_int) :: rc
    type(Mesh) :: msh
    real(rp), pointer :: phi(:), U(:,:), pb(:), Ub(:,:), ff(:), g(:,:), rhs(:)
    call msh%attach(h)
    call c_f_pointer(pphi, phi, [msh%n_cells])
    call c_f_pointer(pu, U, [3, msh%n_cells])
    call c_f_pointer(ppb, pb, [max(1, msh%n_bnd_faces)])
    call c_f_pointer(pub, Ub, [3, max(1, msh%n_bnd_faces)])
    call c_f_pointer(pff, ff, [msh%n_faces])
    call c_f_pointer(prhs, rhs, [msh%n_cells])
    if (c_associated(pg)) then
      call c_f_pointer(pg, g, [3, msh%n_cells])
      call advection_rhs_scalar(phi, U, msh, int(scheme, ip), rhs, pb, Ub, ff, grad_phi=g)
    else
      call advection_rhs_scalar(phi, U, msh, int(scheme, ip), rhs, pb, Ub, ff)
    end if
    rc = 0
  end function

  function ft_viscous_scalar(h, pphi, alpha, pnus, prt, pg, ppb, prhs) bind(C, name='ft_viscous_scalar') result(rc)
    type(c_ptr), value :: h, pphi, pnus, pg, ppb, prhs
    real(c_double), value :: alpha, prt
    integer(c_int) :: rc
    type(Mesh) :: msh
    real(rp), pointer :: phi(:), nus(:), g(:,:), pb(:), rhs(:)
    call msh%attach(h)
    call c_f_pointer(pphi, phi, [msh%n_cells])
    call c_f_pointer(pnus, nus, [msh%n_cells])
    call c_f_pointer(pg, g, [3, msh%n_cells])
    call c_f_pointer(ppb, pb, [max(1, msh%n_bnd_faces)])
    call c_f_pointer(prhs, rhs, [msh%n_cells])
    call viscous_rhs_scalar(phi, alpha, nus, prt, g, msh, rhs, pb)
    rc = 0
  end function

  !> model 1: vreman_sgs_viscosity, 2: smagorinsky_sgs (cs_in = coef), 3: amd_sgs_viscosity
  !> (c_amd = coef), 4: dynamic_smagorinsky_sgs; a null grad_U takes the LS gradient of U first
  !> (Vreman and AMD)
  function ft_sgs(h, model, pu, pub, pgu, coef, nu, wall_gradient, dt_stab, pnus) bind(C, name='ft_sgs') result(rc)
    type(c_ptr), value :: h, pu, pub, pgu, pnus
    integer(c_int), value :: model, wall_gradient
    real(c_double), value :: coef, nu, dt_stab
    integer(c_int) :: rc
    type(Mesh) :: msh
    real(rp), pointer :: U(:,:), Ub(:,:), gU(:,:,:), nus(:)
    call msh%attach(h)
    call c_f_pointer(pu, U, [3, msh%n_cells])
    call c_f_pointer(pub, Ub, [3, max(1, msh%n_bnd_faces)])
    call c_f_pointer(pnus, nus, [msh%n_cells])
    rc = 0
    if (c_associated(pgu)) call c_f_pointer(pgu, gU, [3, 3, msh%n_cells])
    select case (model)
    case (1)
      if (c_associated(pgu)) then
        call vreman_sgs_viscosity(U, msh, nus, Ub, gU, dt_stab)
      else
        call vreman_sgs_viscosity(U, msh, nus, Ub, dt_stab=dt_stab)
      end if
    case (2)
      call smagorinsky_sgs(gU, msh, nus, coef, dt_stab)
    case (3)
      if (c_associated(pgu)) then
        call amd_sgs_viscosity(U, msh, nus, nu, coef, wall_gradient /= 0, Ub, gU, dt_stab)
      else
        call amd_sgs_viscosity(U, msh, nus, nu, coef, wall_gradient /= 0, Ub, dt_stab=dt_stab)
      end if
    case (4)
      call dynamic_smagorinsky_sgs(U, gU, msh, nus, Ub, dt_stab)
    case default
      rc = -1
    end select
  end function

  !> apply_wall_model_bc with all its optional arguments (a negative value leaves one absent)
  function ft_wall_model(h, pu, nu, h_wm, t_filt, dt, tf_fac, ptau) bind(C, name='ft_wall_model') result(rc)
    type(c_ptr), value :: h, pu, ptau
    real(c_double), value :: nu, h_wm, t_filt, dt, tf_fac
    integer(c_int) :: rc
    type(Mesh) :: msh
    real(rp), pointer :: U(:,:), tau(:,:)
    call msh%attach(h)
    call c_f_pointer(pu, U, [3, msh%n_cells])
    call c_f_pointer(ptau, tau, [3, max(1, msh%n_bnd_faces)])
    if (h_wm < 0.0_rp) then
      call apply_wall_model_bc(U, nu, msh, tau)
    else
      call apply_wall_model_bc(U, nu, msh, tau, h_wm=h_wm, t_filt=t_filt, dt=dt, tf_fac=tf_fac)
    end if
    rc = 0
  end function

  !> n_steps RK3W steps of size dt through FVSolver%advance; U and p copied out by sync_host
  function ft_solver_run(h, n_steps, dt, pu, pp, t_out) bind(C, name='ft_solver_run') result(rc)
    type(c_ptr), value :: h, pu, pp
    integer(c_int), value :: n_steps
    real(c_double), value :: dt
    real(c_double), intent(out) :: t_out
    integer(c_int) :: rc
    type(FVSolver) :: sol
    real(rp), pointer :: U(:,:), p(:)
    integer :: k
    call sol%attach(h)
    do k = 1, n_steps
      call sol%advance(dt)
    end do
    rc = 0
    if (.not. sol%state_is_finite()) rc = -1
    call c_f_pointer(pu, U, [3, sol%mesh%n_cells])
    call c_f_pointer(pp, p, [sol%mesh%n_cells])
    call sol%sync_host(U, p)
    t_out = sol%time
  end function

end module ops_harness
