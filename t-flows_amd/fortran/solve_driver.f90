!==============================================================================!
  program Solve_Driver
!------------------------------------------------------------------------------!
!>  Calls the GPU solver through Tflows_Gpu_Mod the way Solver_Mod % Run calls
!>  Solve_Native (Solver_Mod/Run.f90:23-33).  Reads a system written as raw
!>  stream files in a directory (dims: ni nt nnz; row, col, dia: int32;
!>  val, b, x0: float64), solves it with the solver, preconditioner,
!>  iteration cap, tolerance and optional norm ("-" when absent) given on the
!>  command line, writes x to x.bin and prints "niter fin_res".  Mode "host"
!>  (default) passes host arrays (Tflows_Gpu_Solve); "device" copies them to
!>  device memory and solves there (Tflows_Gpu_Solve_Device, FP64), "fp32"
!>  the same with FP32 sweeps inside FP64 refinement; both also print the
!>  true relative residual.
!>    solve_driver <dir> <solver> <prec> <miter> <tol> [norm|-] [host|device|fp32]
!------------------------------------------------------------------------------!
  use Tflows_Gpu_Mod
  implicit none
!-----------------------------------[Locals]-----------------------------------!
  character(len=512)          :: dir, arg
  character(len=16)           :: solver, prec, mode
  integer                     :: ni, nt, nnz, miter, niter, u
  integer(c_int), allocatable, target :: row(:), col(:), dia(:)
  real(c_double), allocatable, target :: val(:), b(:), x(:)
  real(c_double)              :: tol, fin_res, norm
  logical                     :: has_norm
  type(c_ptr)                 :: h, d_val, d_x, d_b
  type(Tflows_Gpu_Options)    :: opt
  type(Tflows_Gpu_Report)     :: rep
!==============================================================================!

  call get_command_argument(1, dir)
  call get_command_argument(2, solver)
  call get_command_argument(3, prec)
  call get_command_argument(4, arg);  read(arg, *) miter
  call get_command_argument(5, arg);  read(arg, *) tol

  open(newunit=u, file=trim(dir)//'/dims.txt', status='old')
  read(u, *) ni, nt, nnz
  close(u)
  allocate(row(nt+1), col(nnz), dia(nt), val(nnz), b(nt), x(nt))
  call read_int(trim(dir)//'/row.bin', row)
  call read_int(trim(dir)//'/col.bin', col)
  call read_int(trim(dir)//'/dia.bin', dia)
  call read_real(trim(dir)//'/val.bin', val)
  call read_real(trim(dir)//'/b.bin',   b)
  call read_real(trim(dir)//'/x0.bin',  x)

  has_norm = .false.
  if(command_argument_count() >= 6) then
    call get_command_argument(6, arg)
    if(trim(arg) /= '-') then
      read(arg, *) norm
      has_norm = .true.
    end if
  end if
  mode = 'host'
  if(command_argument_count() >= 7) call get_command_argument(7, mode)

  call Tflows_Gpu_Create(h, ni, nt, row, col, dia)
  if(trim(mode) == 'host') then
    if(has_norm) then
      call Tflows_Gpu_Solve(h, trim(solver), trim(prec), val, x, b,         &
                            miter, niter, tol, fin_res, norm)
    else
      call Tflows_Gpu_Solve(h, trim(solver), trim(prec), val, x, b,         &
                            miter, niter, tol, fin_res)
    end if
  else
    call device_copy(d_val, c_loc(val), 8_c_long_long * nnz)
    call device_copy(d_x,   c_loc(x),   8_c_long_long * nt)
    call device_copy(d_b,   c_loc(b),   8_c_long_long * nt)
    opt % true_residual = 1
    if(trim(mode) == 'fp32') opt % precision = TFG_FP32_IR
    if(has_norm) then
      call Tflows_Gpu_Solve_Device(h, trim(solver), trim(prec), d_val, d_x, &
                                   d_b, miter, niter, tol, fin_res, opt,    &
                                   rep, norm)
    else
      call Tflows_Gpu_Solve_Device(h, trim(solver), trim(prec), d_val, d_x, &
                                   d_b, miter, niter, tol, fin_res, opt, rep)
    end if
    call Tflows_Gpu_Check(tfg_memcpy_d2h(c_loc(x), d_x, 8_c_long_long * nt), &
                          'tfg_memcpy_d2h')
    call Tflows_Gpu_Check(tfg_device_free(d_val), 'tfg_device_free')
    call Tflows_Gpu_Check(tfg_device_free(d_x),   'tfg_device_free')
    call Tflows_Gpu_Check(tfg_device_free(d_b),   'tfg_device_free')
  end if
  call Tflows_Gpu_Destroy(h)

  open(newunit=u, file=trim(dir)//'/x.bin', access='stream',                &
       form='unformatted', status='replace')
  write(u) x
  close(u)
  if(trim(mode) == 'host') then
    print '(i0,1x,es24.16)', niter, fin_res
  else
    print '(i0,1x,es24.16,1x,es24.16)', niter, fin_res, rep % true_relres
  end if

  contains

  subroutine device_copy(d, src, bytes)
    type(c_ptr),          intent(out) :: d
    type(c_ptr),          intent(in)  :: src
    integer(c_long_long), intent(in)  :: bytes
    call Tflows_Gpu_Check(tfg_device_alloc(d, bytes), 'tfg_device_alloc')
    call Tflows_Gpu_Check(tfg_memcpy_h2d(d, src, bytes), 'tfg_memcpy_h2d')
  end subroutine

  subroutine read_int(name, a)
    character(*),   intent(in)  :: name
    integer(c_int), intent(out) :: a(:)
    integer :: v
    open(newunit=v, file=name, access='stream', form='unformatted', status='old')
    read(v) a
    close(v)
  end subroutine

  subroutine read_real(name, a)
    character(*),   intent(in)  :: name
    real(c_double), intent(out) :: a(:)
    integer :: v
    open(newunit=v, file=name, access='stream', form='unformatted', status='old')
    read(v) a
    close(v)
  end subroutine

  end program
