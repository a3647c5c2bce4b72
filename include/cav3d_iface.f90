! Fortran 2003 interface to libcav3d (include/cav3d.h): the bind(c) module a
! Fortran host (the reference's solve_module / update_module) uses to call
! the HIP library.  Types mirror cav3d_params and cav3d_block field for
! field; every function returns 0 on success and a negative code on error
! (cav3d_last_error gives the message).  Arrays keep the reference's layouts
! (component fastest, column-major), passed by address.
module cav3d_iface
  use, intrinsic :: iso_c_binding
  implicit none
  integer(c_int), parameter :: CAV3D_BC_WALL = 1, CAV3D_BC_INFLOW_SUBSONIC = 2, CAV3D_BC_OUTFLOW_SUBSONIC = 3, &
       CAV3D_BC_FARFIELD = 4, CAV3D_BC_SYMMETRY = 5, CAV3D_BC_INFLOW_SUPERSONIC = 6, CAV3D_BC_OUTFLOW_SUPERSONIC = 7, &
       CAV3D_BC_DEGENERATE = 8, CAV3D_BC_INFLOW = 9, CAV3D_BC_OUTFLOW = 10, CAV3D_BC_SURFACE = 256
  integer(c_int), parameter :: CAV3D_FP64 = 0, CAV3D_MIXED = 1, CAV3D_CLIP_AS_WRITTEN = 0, CAV3D_CLIP_SPECIES = 1

  type, bind(c) :: cav3d_params
    integer(c_int) :: iturb, nscheme, nmuscl, nlim, local, prec, ncav, nsteady, precision, cav_clip
    integer(c_int) :: ides, reserved
    real(c_double) :: cfl, pref, uref, str, dt_phy, c_v, c_c, dp_ref, t_ref, t_crit
    real(c_double) :: bc_pv(9), bc_pressure, emutref
  end type

  type, bind(c) :: cav3d_block
    integer(c_int) :: imax, jmax, kmax, nbc
    type(c_ptr)    :: bc          ! int(7,nbc): kind, istart(3), iend(3)
    integer(c_int) :: ncon
    type(c_ptr)    :: con         ! int(22,ncon): donor, istart(3), iend(3), istart_donor(3), iend_donor(3), transmat(3,3)
  end type

  interface
    integer(c_int) function cav3d_create(p, nb, blocks, ctx) bind(c)
      import; type(cav3d_params), intent(in) :: p; integer(c_int), value :: nb
      type(cav3d_block), intent(in) :: blocks(*); type(c_ptr), intent(out) :: ctx
    end function
    integer(c_int) function cav3d_set_metrics(ctx, b, cx, ex, tx, grd) bind(c)
      import; type(c_ptr), value :: ctx; integer(c_int), value :: b
      real(c_double), intent(in) :: cx(*), ex(*), tx(*), grd(*)
    end function
    integer(c_int) function cav3d_set_grid(ctx, b, x) bind(c)
      import; type(c_ptr), value :: ctx; integer(c_int), value :: b
      real(c_double), intent(in) :: x(*)
    end function
    integer(c_int) function cav3d_build_metrics(ctx) bind(c)
      import; type(c_ptr), value :: ctx
    end function
    integer(c_int) function cav3d_get_metrics(ctx, b, cx, ex, tx, grd) bind(c)
      import; type(c_ptr), value :: ctx; integer(c_int), value :: b
      real(c_double), intent(out) :: cx(*), ex(*), tx(*), grd(*)
    end function
    integer(c_int) function cav3d_put_state(ctx, b, pv, dv, tv, qq) bind(c)
      import; type(c_ptr), value :: ctx; integer(c_int), value :: b
      real(c_double), intent(in) :: pv(*), dv(*), tv(*), qq(*)
    end function
    integer(c_int) function cav3d_get_state(ctx, b, pv, dv, tv, qq) bind(c)
      import; type(c_ptr), value :: ctx; integer(c_int), value :: b
      real(c_double), intent(out) :: pv(*), dv(*), tv(*), qq(*)
    end function
    integer(c_int) function cav3d_timeinteg(ctx, nt_phy, nt, timeprev, time) bind(c)
      import; type(c_ptr), value :: ctx; integer(c_int), value :: nt_phy, nt
      real(c_double), value :: timeprev; real(c_double), intent(out) :: time
    end function
    integer(c_int) function cav3d_setbc(ctx) bind(c)
      import; type(c_ptr), value :: ctx
    end function
    integer(c_int) function cav3d_calrhs(ctx, time, b, res, icav, itt, omega_cut) bind(c)
      import; type(c_ptr), value :: ctx; real(c_double), value :: time; integer(c_int), value :: b
      real(c_double), intent(out) :: res(*), icav(*), itt(*), omega_cut(*)
    end function
    integer(c_int) function cav3d_timestep(ctx, b, dt) bind(c)
      import; type(c_ptr), value :: ctx; integer(c_int), value :: b
      real(c_double), intent(out) :: dt(*)
    end function
    integer(c_int) function cav3d_residual(ctx, l2, ijk_max, resmax) bind(c)
      import; type(c_ptr), value :: ctx
      real(c_double), intent(out) :: l2(*), resmax; integer(c_int), intent(out) :: ijk_max(4)
    end function
    integer(c_int) function cav3d_unsteady_shift(ctx) bind(c)
      import; type(c_ptr), value :: ctx
    end function
    integer(c_int) function cav3d_eos(ctx, n, p, t, y1, y2, dv, tv) bind(c)
      import; type(c_ptr), value :: ctx; integer(c_int), value :: n
      real(c_double), intent(in) :: p(*), t(*), y1(*), y2(*); real(c_double), intent(out) :: dv(*), tv(*)
    end function
    integer(c_int) function cav3d_eos_host(n, p, t, y1, y2, dv, tv) bind(c)
      import; integer(c_int), value :: n
      real(c_double), intent(in) :: p(*), t(*), y1(*), y2(*); real(c_double), intent(out) :: dv(*), tv(*)
    end function
    integer(c_int) function cav3d_integrals(ctx, rhoref, uref, area, aoa, out) bind(c)
      import; type(c_ptr), value :: ctx; real(c_double), value :: rhoref, uref, area, aoa
      real(c_double), intent(out) :: out(10)
    end function
    integer(c_int) function cav3d_integrals_noalloc(ctx, rhoref, uref, area, aoa, out) bind(c)
      import; type(c_ptr), value :: ctx; real(c_double), value :: rhoref, uref, area, aoa
      real(c_double), intent(out) :: out(10)
    end function
    integer(c_int) function cav3d_export(ctx, path, nt_phy, nt, time) bind(c)
      import; type(c_ptr), value :: ctx; character(kind=c_char), intent(in) :: path(*)
      integer(c_int), value :: nt_phy, nt; real(c_double), value :: time
    end function
    integer(c_int) function cav3d_import(ctx, path, nps, nts, timeprev) bind(c)
      import; type(c_ptr), value :: ctx; character(kind=c_char), intent(in) :: path(*)
      integer(c_int), intent(out) :: nps, nts; real(c_double), intent(out) :: timeprev
    end function
    integer(c_int) function cav3d_rhs_vjp(ctx, res_bar, pv_bar, scal_bar) bind(c)
      import; type(c_ptr), value :: ctx; real(c_double), intent(in) :: res_bar(*)
      real(c_double), intent(out) :: pv_bar(*), scal_bar(2)
    end function
    integer(c_int) function cav3d_rhs_vjp_eddy(ctx, res_bar, pv_bar, scal_bar) bind(c)
      import; type(c_ptr), value :: ctx; real(c_double), intent(in) :: res_bar(*)
      real(c_double), intent(out) :: pv_bar(*), scal_bar(2)
    end function
    integer(c_int) function cav3d_eddy(ctx) bind(c)
      import; type(c_ptr), value :: ctx
    end function
    integer(c_int) function cav3d_rhs_vjp_params(ctx, res_bar, par_bar) bind(c)
      import; type(c_ptr), value :: ctx; real(c_double), intent(in) :: res_bar(*)
      real(c_double), intent(out) :: par_bar(14)
    end function
    integer(c_int) function cav3d_objective(ctx, w_vap, w_fx, j, pv_bar, dj_dpref) bind(c)
      import; type(c_ptr), value :: ctx; real(c_double), value :: w_vap, w_fx
      real(c_double), intent(out) :: j, pv_bar(*), dj_dpref
    end function
    integer(c_int) function cav3d_profile(ctx, enable, out) bind(c)
      import; type(c_ptr), value :: ctx; integer(c_int), value :: enable
      real(c_double), intent(out) :: out(10)
    end function
    integer(c_int) function cav3d_synchronize(ctx) bind(c)
      import; type(c_ptr), value :: ctx
    end function
    type(c_ptr) function cav3d_last_error() bind(c)
      import
    end function
    subroutine cav3d_destroy(ctx) bind(c)
      import; type(c_ptr), value :: ctx
    end subroutine
  end interface
end module cav3d_iface
