!==============================================================================!
  module Tflows_Gpu_Mod
!------------------------------------------------------------------------------!
!>  Fortran binding of the tflows_gpu C ABI (include/tflows_gpu.h) for the
!>  T-Flows Process program.  Tflows_Gpu_Solve keeps the argument list of
!>  Nat % Solve_Native(solver, prec, A, x, b, miter, niter, tol, fin_res, norm)
!>  (Sources/Process/Cpu/Native_Mod/Solve_Native.f90:2-29), with the matrix
!>  passed as the handle of its uploaded structure plus its values A % val.
!>  Any failure stops the program, as Message % Error does for the
!>  reference's fatal conditions (Message_Mod/Error.f90:73-74).
!------------------------------------------------------------------------------!
  use iso_c_binding
  implicit none
!==============================================================================!

  ! tfg_options / tfg_report (tflows_gpu.h) of the extended solve
  integer(c_int), parameter :: TFG_FP64 = 0, TFG_FP32_IR = 1
  integer(c_int), parameter :: TFG_CRIT_NRMS = 0, TFG_CRIT_RELRES = 1
  type, bind(C) :: Tflows_Gpu_Options
    integer(c_int) :: precision = 0, criterion = 0, path = 0, inner_max = 0
    real(c_double) :: inner_tol = 0.0_c_double
    integer(c_int) :: timing = 0, true_residual = 0, textbook_bicg = 0,     &
                      deflation = 0
  end type
  type, bind(C) :: Tflows_Gpu_Report
    integer(c_int) :: status, iterations, refinements
    real(c_double) :: true_relres, solve_ms, spmv_ms
    integer(c_int) :: spmv_samples, deflated
  end type

  ! tfg_momentum (tflows_gpu.h): one velocity component's momentum assembly
  ! (Compute_Momentum.f90:119-366).  Device addresses; cell arrays at cell 1,
  ! e.g. c_loc(ui % n(1)) for the (-n_bnd:n) arrays phi_n, gx, gy, gz, uj_i,
  ! uk_i and density.
  type, bind(C) :: Tflows_Gpu_Momentum
    integer(c_int) :: component, n_bnd, adv_scheme, blend_matrix, td_scheme
    real(c_double) :: blend, dt, urf, p_drop
    type(c_ptr)    :: f, fw, fc, v_flux, vis_eff, f_stress
    type(c_ptr)    :: sx, sy, sz, dx, dy, dz, wall
    type(c_ptr)    :: phi_n, phi_o, phi_oo, gx, gy, gz, uj_i, uk_i, density,  &
                      vol, p_i, cell_f, st
  end type

  ! tfg_rhie_chow (tflows_gpu.h): the Rhie-Chow volume fluxes
  ! (Rhie_And_Chow.f90:86-259); device addresses of cell (1..n) and face arrays.
  type, bind(C) :: Tflows_Gpu_Rhie_Chow
    type(c_ptr)    :: f, fc, sx, sy, sz, dx, dy, dz
    type(c_ptr)    :: u, v, w, p, px, py, pz, v_m
    integer(c_int) :: choi
    type(c_ptr)    :: u_o, u_oo, v_o, v_oo, w_o_, w_oo_, density, flux_o, flux_oo
    real(c_double) :: dt, w_o, w_oo
    integer(c_int) :: gu
    type(c_ptr)    :: cell_fx, cell_fy, cell_fz, face_fx, face_fy, face_fz
    integer(c_int) :: vof
    real(c_double) :: sigma
    type(c_ptr)    :: curv, col, col_x, col_y, col_z
  end type

  interface

    integer(c_int) function tfg_matrix_create(ni, nt, nnz, row, col, dia,  &
                                              handle)                      &
                                              bind(C, name='tfg_matrix_create')
      import :: c_int, c_ptr
      integer(c_int), value       :: ni, nt, nnz
      integer(c_int), intent(in)  :: row(*), col(*), dia(*)
      type(c_ptr),    intent(out) :: handle
    end function

    integer(c_int) function tfg_matrix_destroy(handle)                     &
                                              bind(C, name='tfg_matrix_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: handle
    end function

    integer(c_int) function tfg_solve_native(handle, solver, solver_len,   &
                                             prec, prec_len, val, x, b,    &
                                             miter, niter, tol, fin_res,   &
                                             norm)                         &
                                             bind(C, name='tfg_solve_native')
      import :: c_int, c_ptr, c_char, c_double
      type(c_ptr),            value          :: handle
      character(kind=c_char), intent(in)     :: solver(*), prec(*)
      integer(c_int),         value          :: solver_len, prec_len
      real(c_double),         intent(in)     :: val(*)
      real(c_double),         intent(inout)  :: x(*), b(*)
      integer(c_int),         value          :: miter
      integer(c_int),         intent(out)    :: niter
      real(c_double),         value          :: tol
      real(c_double),         intent(out)    :: fin_res
      real(c_double), optional, intent(in)   :: norm
    end function

    ! The same call on device addresses (device_ptrs = 1; c_loc of host
    ! arrays with device_ptrs = 0), with options and a report
    integer(c_int) function tfg_solve_native_ex(handle, solver, solver_len,   &
                                                prec, prec_len, val, x, b,    &
                                                miter, niter, tol, fin_res,   &
                                                norm, device_ptrs, opt, rep)  &
                                           bind(C, name='tfg_solve_native_ex')
      import :: c_int, c_ptr, c_char, c_double, Tflows_Gpu_Options,          &
                Tflows_Gpu_Report
      type(c_ptr),              value         :: handle, val, x, b
      character(kind=c_char),   intent(in)    :: solver(*), prec(*)
      integer(c_int),           value         :: solver_len, prec_len, miter
      integer(c_int),           intent(out)   :: niter
      real(c_double),           value         :: tol
      real(c_double),           intent(out)   :: fin_res
      real(c_double), optional, intent(in)    :: norm
      integer(c_int),           value         :: device_ptrs
      type(Tflows_Gpu_Options), intent(in)    :: opt
      type(Tflows_Gpu_Report),  intent(out)   :: rep
    end function

    ! Device memory of the handle's device (bytes), copies to and from it
    integer(c_int) function tfg_device_alloc(dptr, bytes)                    &
                                             bind(C, name='tfg_device_alloc')
      import :: c_int, c_ptr, c_long_long
      type(c_ptr),          intent(out) :: dptr
      integer(c_long_long), value       :: bytes
    end function
    integer(c_int) function tfg_device_free(dptr)                            &
                                             bind(C, name='tfg_device_free')
      import :: c_int, c_ptr
      type(c_ptr), value :: dptr
    end function
    integer(c_int) function tfg_memcpy_h2d(dst, src, bytes)                  &
                                             bind(C, name='tfg_memcpy_h2d')
      import :: c_int, c_ptr, c_long_long
      type(c_ptr),          value :: dst, src
      integer(c_long_long), value :: bytes
    end function
    integer(c_int) function tfg_memcpy_d2h(dst, src, bytes)                  &
                                             bind(C, name='tfg_memcpy_d2h')
      import :: c_int, c_ptr, c_long_long
      type(c_ptr),          value :: dst, src
      integer(c_long_long), value :: bytes
    end function
    integer(c_int) function tfg_synchronize(handle)                          &
                                             bind(C, name='tfg_synchronize')
      import :: c_int, c_ptr
      type(c_ptr), value :: handle
    end function
    integer(c_int) function tfg_faces_destroy(faces)                         &
                                             bind(C, name='tfg_faces_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: faces
    end function

    type(c_ptr) function tfg_last_error() bind(C, name='tfg_last_error')
      import :: c_ptr
    end function

    ! Multi-rank hooks: Comm_Mod's Global % Sum/Min/Max_Real and
    ! Grid % Exchange_Cells_Real stand-ins, passed with c_funloc of bind(C)
    ! functions of the interfaces below (Tflows_Gpu_Allreduce_If,
    ! Tflows_Gpu_Exchange_If).
    integer(c_int) function tfg_matrix_set_comm(handle, allreduce, exchange,  &
                                                ctx)                        &
                                          bind(C, name='tfg_matrix_set_comm')
      import :: c_int, c_ptr, c_funptr
      type(c_ptr),    value :: handle
      type(c_funptr), value :: allreduce, exchange
      type(c_ptr),    value :: ctx
    end function

    ! Pressure-system assembly on the device from face data
    integer(c_int) function tfg_faces_create(handle, n_faces, c1, c2, pos,    &
                                             diag_bc, b_region, faces)        &
                                             bind(C, name='tfg_faces_create')
      import :: c_int, c_ptr
      type(c_ptr),    value       :: handle
      integer(c_int), value       :: n_faces
      integer(c_int), intent(in)  :: c1(*), c2(*), pos(*), diag_bc(*), b_region(*)
      type(c_ptr),    intent(out) :: faces
    end function

    integer(c_int) function tfg_pressure_assemble(faces, fc, fw, vm, vflux,   &
                                                  val, b, singular)         &
                                         bind(C, name='tfg_pressure_assemble')
      import :: c_int, c_ptr
      type(c_ptr),    value :: faces, fc, fw, vm, vflux, val, b
      integer(c_int), value :: singular
    end function

    ! The same with pp % n for ambient regions: pp = c_loc(pp % n(1))
    integer(c_int) function tfg_pressure_assemble_ex(faces, fc, fw, vm,      &
                                                     vflux, pp, val, b,      &
                                                     singular)               &
                                      bind(C, name='tfg_pressure_assemble_ex')
      import :: c_int, c_ptr
      type(c_ptr),    value :: faces, fc, fw, vm, vflux, pp, val, b
      integer(c_int), value :: singular
    end function

    ! p_min, p_max of Compute_Pressure from device p % n (before Min/Max_Real)
    integer(c_int) function tfg_range(handle, v, vmin, vmax)                 &
                                         bind(C, name='tfg_range')
      import :: c_int, c_ptr, c_double
      type(c_ptr),    value       :: handle, v
      real(c_double), intent(out) :: vmin, vmax
    end function

    ! vel_max of Compute_Momentum from device u, v, w (before Max_Real)
    integer(c_int) function tfg_vel_max(handle, u, v, w, vel_max)            &
                                         bind(C, name='tfg_vel_max')
      import :: c_int, c_ptr, c_double
      type(c_ptr),    value       :: handle, u, v, w
      real(c_double), intent(out) :: vel_max
    end function

    integer(c_int) function tfg_rhie_chow_fluxes(faces, rc, v_flux)          &
                                         bind(C, name='tfg_rhie_chow_fluxes')
      import :: c_int, c_ptr, Tflows_Gpu_Rhie_Chow
      type(c_ptr),                value      :: faces, v_flux
      type(Tflows_Gpu_Rhie_Chow), intent(in) :: rc
    end function

    integer(c_int) function tfg_momentum_assemble(faces, m, val, b, v_m)    &
                                         bind(C, name='tfg_momentum_assemble')
      import :: c_int, c_ptr, Tflows_Gpu_Momentum
      type(c_ptr),               value      :: faces, val, b, v_m
      type(Tflows_Gpu_Momentum), intent(in) :: m
    end function

  end interface

  ! Signatures of the two hooks (tflows_gpu.h: tfg_allreduce_fn,
  ! tfg_exchange_fn); values, ops: count entries; d_x: device address.
  abstract interface
    integer(c_int) function Tflows_Gpu_Allreduce_If(ctx, values, ops,       &
                                                    count) bind(C)
      import :: c_int, c_ptr, c_double
      type(c_ptr),    value         :: ctx
      real(c_double), intent(inout) :: values(*)
      integer(c_int), intent(in)    :: ops(*)
      integer(c_int), value         :: count
    end function
    integer(c_int) function Tflows_Gpu_Exchange_If(ctx, d_x) bind(C)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, d_x
    end function
  end interface

  contains

!==============================================================================!
  function Tflows_Gpu_Error_Text() result(text)
!------------------------------------------------------------------------------!
!>  Message of the last failed call.
!------------------------------------------------------------------------------!
  character(len=256)               :: text
  character(kind=c_char), pointer  :: chars(:)
  integer                          :: i
!==============================================================================!
  text = ' '
  call c_f_pointer(tfg_last_error(), chars, [256])
  do i = 1, 256
    if(chars(i) == c_null_char) exit
    text(i:i) = chars(i)
  end do
  end function

!==============================================================================!
  subroutine Tflows_Gpu_Check(rc, where)
!------------------------------------------------------------------------------!
  integer(c_int), intent(in) :: rc
  character(*),   intent(in) :: where
!==============================================================================!
  if(rc /= 0) then
    print '(a,a,a,i0,a,a)', ' # ', where, ' failed (', rc, '): ',           &
                            trim(Tflows_Gpu_Error_Text())
    stop 1
  end if
  end subroutine

!==============================================================================!
  subroutine Tflows_Gpu_Create(handle, ni, nt, a_row, a_col, a_dia)
!------------------------------------------------------------------------------!
!>  Uploads the CSR structure of Matrix_Type (Matrix_Mod.f90:57-76) once per
!>  grid: a_row(nt+1), a_col(nnz), a_dia(nt), all 1-based as built by
!>  Create_Matrix.f90:87-119.
!------------------------------------------------------------------------------!
  type(c_ptr),    intent(out) :: handle
  integer,        intent(in)  :: ni, nt
  integer(c_int), intent(in)  :: a_row(:), a_col(:), a_dia(:)
!==============================================================================!
  call Tflows_Gpu_Check(tfg_matrix_create(ni, nt, a_row(nt+1) - 1, a_row,    &
                                          a_col, a_dia, handle),             &
                        'tfg_matrix_create')
  end subroutine

!==============================================================================!
  subroutine Tflows_Gpu_Solve(handle, solver, prec, a_val, x, b,            &
                              miter, niter, tol, fin_res, norm)
!------------------------------------------------------------------------------!
!>  Solve_Native's call (Solve_Native.f90:2-46) on the GPU.  x(1:nt) holds
!>  x0 on entry and the solution on exit; a_val and b are left unchanged.
!------------------------------------------------------------------------------!
  type(c_ptr),              intent(in)    :: handle
  character(*),             intent(in)    :: solver, prec
  real(c_double),           intent(in)    :: a_val(:)
  real(c_double),           intent(inout) :: x(:), b(:)
  integer,                  intent(in)    :: miter
  integer,                  intent(out)   :: niter
  real(c_double),           intent(in)    :: tol
  real(c_double),           intent(out)   :: fin_res
  real(c_double), optional, intent(in)    :: norm
!-----------------------------------[Locals]-----------------------------------!
  integer(c_int) :: n_out
!==============================================================================!
  call Tflows_Gpu_Check(tfg_solve_native(handle, solver, len(solver),        &
                                         prec, len(prec), a_val, x, b,       &
                                         miter, n_out, tol, fin_res, norm),  &
                        'tfg_solve_native')
  niter = n_out
  end subroutine

!==============================================================================!
  subroutine Tflows_Gpu_Solve_Device(handle, solver, prec, d_val, d_x, d_b,  &
                                     miter, niter, tol, fin_res, opt, rep,   &
                                     norm)
!------------------------------------------------------------------------------!
!>  Solve_Native's call on device-resident arrays (device addresses of
!>  A % val, x(1:nt) and b, e.g. from tfg_device_alloc or an offload
!>  runtime), with the options of tfg_solve_native_ex (precision, stopping
!>  criterion, ...) and its report.
!------------------------------------------------------------------------------!
  type(c_ptr),              intent(in)  :: handle, d_val, d_x, d_b
  character(*),             intent(in)  :: solver, prec
  integer,                  intent(in)  :: miter
  integer,                  intent(out) :: niter
  real(c_double),           intent(in)  :: tol
  real(c_double),           intent(out) :: fin_res
  type(Tflows_Gpu_Options), intent(in)  :: opt
  type(Tflows_Gpu_Report),  intent(out) :: rep
  real(c_double), optional, intent(in)  :: norm
!-----------------------------------[Locals]-----------------------------------!
  integer(c_int) :: n_out
!==============================================================================!
  call Tflows_Gpu_Check(tfg_solve_native_ex(handle, solver, len(solver),     &
                                            prec, len(prec), d_val, d_x,     &
                                            d_b, miter, n_out, tol, fin_res, &
                                            norm, 1, opt, rep),              &
                        'tfg_solve_native_ex')
  niter = n_out
  end subroutine

!==============================================================================!
  subroutine Tflows_Gpu_Destroy(handle)
!------------------------------------------------------------------------------!
  type(c_ptr), intent(in) :: handle
!==============================================================================!
  call Tflows_Gpu_Check(tfg_matrix_destroy(handle), 'tfg_matrix_destroy')
  end subroutine

  end module
