!==============================================================================!
  program Solve_Driver
!------------------------------------------------------------------------------!
!>  Calls the GPU solver through Tflows_Gpu_Mod the way Solver_Mod % Run calls
!>  Solve_Native (Solver_Mod/Run.f90:23-33).  Reads a system written as raw
!>  stream files in a directory (dims: ni nt nnz; row, col, dia: int32;
!>  val, b, x0: float64), solves it with the solver, preconditioner,
!>  iteration cap, tolerance and optional norm given on the command line,
!>  writes x to x.bin and prints "niter fin_res".
!>    solve_driver <dir> <solver> <prec> <miter> <tol> [norm]
!------------------------------------------------------------------------------!
  use Tflows_Gpu_Mod
  implicit none
!-----------------------------------[Locals]-----------------------------------!
  character(len=512)          :: dir, arg
  character(len=16)           :: solver, prec
  integer                     :: ni, nt, nnz, miter, niter, u
  integer(c_int), allocatable :: row(:), col(:), dia(:)
  real(c_double), allocatable :: val(:), b(:), x(:)
  real(c_double)              :: tol, fin_res, norm
  type(c_ptr)                 :: h
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

  call Tflows_Gpu_Create(h, ni, nt, row, col, dia)
  if(command_argument_count() >= 6) then
    call get_command_argument(6, arg);  read(arg, *) norm
    call Tflows_Gpu_Solve(h, trim(solver), trim(prec), val, x, b,           &
                          miter, niter, tol, fin_res, norm)
  else
    call Tflows_Gpu_Solve(h, trim(solver), trim(prec), val, x, b,           &
                          miter, niter, tol, fin_res)
  end if
  call Tflows_Gpu_Destroy(h)

  open(newunit=u, file=trim(dir)//'/x.bin', access='stream',                &
       form='unformatted', status='replace')
  write(u) x
  close(u)
  print '(i0,1x,es24.16)', niter, fin_res

  contains

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
