!-------------------------------------------------------------------------------
! iso_c_binding interface to libflusi_hip.so (include/flusi_hip.h).
!
! A Fortran host (FLUSI's time_step / FluidTimestep callers) binds the
! MI355X hot path through this module.  Arrays are passed with the
! reference's explicit-shape layouts: uk(0:nz-1,0:ny-1,0:nx/2,1:3) complex(8),
! u(0:nx-1,0:ny-1,0:nz-1,1:3) real(8), mask_color integer(2).  Every
! function returns 0 or the reference's abort code; hip_check() turns a
! non-zero code into abort(code, msg) like vars.f90:327-336.
!-------------------------------------------------------------------------------
module flusi_hip
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: FLUSI_FP64 = 0, FLUSI_FP32_MIXED = 1
  integer, parameter :: FLUSI_INTEGRALS_NV = 61

  interface
    integer(c_int) function flusi_init(ini_path, precision, ctx) bind(C, name="flusi_init")
      import :: c_char, c_int, c_ptr
      character(kind=c_char), dimension(*), intent(in) :: ini_path
      integer(c_int), value :: precision
      type(c_ptr), intent(out) :: ctx
    end function

    subroutine flusi_free(ctx) bind(C, name="flusi_free")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end subroutine

    type(c_ptr) function flusi_last_error(ctx) bind(C, name="flusi_last_error")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end function

    integer(c_int) function flusi_get_dims(ctx, nx, ny, nz) bind(C, name="flusi_get_dims")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), intent(out) :: nx, ny, nz
    end function

    integer(c_int) function flusi_set_output_dir(ctx, dir) bind(C, name="flusi_set_output_dir")
      import :: c_char, c_int, c_ptr
      type(c_ptr), value :: ctx
      character(kind=c_char), dimension(*), intent(in) :: dir
    end function

    integer(c_int) function flusi_init_fields(ctx) bind(C, name="flusi_init_fields")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    integer(c_int) function flusi_set_uk(ctx, uk) bind(C, name="flusi_set_uk")
      import :: c_int, c_ptr, c_double_complex
      type(c_ptr), value :: ctx
      complex(c_double_complex), dimension(*), intent(in) :: uk
    end function

    integer(c_int) function flusi_get_uk(ctx, uk) bind(C, name="flusi_get_uk")
      import :: c_int, c_ptr, c_double_complex
      type(c_ptr), value :: ctx
      complex(c_double_complex), dimension(*), intent(out) :: uk
    end function

    integer(c_int) function flusi_get_u(ctx, u) bind(C, name="flusi_get_u")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), dimension(*), intent(out) :: u
    end function

    integer(c_int) function flusi_get_vorticity(ctx, w) bind(C, name="flusi_get_vorticity")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), dimension(*), intent(out) :: w
    end function

    ! rhs.f90:2 cal_nlk(time,it,nlk,uk,u,...): nlk and u may be c_null_ptr
    integer(c_int) function flusi_cal_nlk(ctx, time, it, nlk, u) bind(C, name="flusi_cal_nlk")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: time
      integer(c_int), value :: it
      type(c_ptr), value :: nlk, u
    end function

    ! fluid_time_step.f90:34 FluidTimestep(time,it,dt0,dt1,...)
    integer(c_int) function flusi_fluid_timestep(ctx, time, it, dt0, dt1) bind(C, name="flusi_fluid_timestep")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(inout) :: time, dt0
      integer(c_int), intent(inout) :: it
      real(c_double), intent(out) :: dt1
    end function

    integer(c_int) function flusi_step(ctx) bind(C, name="flusi_step")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    integer(c_int) function flusi_run(ctx, nsteps) bind(C, name="flusi_run")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: nsteps
    end function

    integer(c_int) function flusi_advance(ctx, nsteps) bind(C, name="flusi_advance")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: nsteps
    end function

    ! save_fields.f90:2 save_fields
    integer(c_int) function flusi_save_fields(ctx) bind(C, name="flusi_save_fields")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    ! runtime_backuping.f90:3 dump_runtime_backup
    integer(c_int) function flusi_dump_runtime_backup(ctx) bind(C, name="flusi_dump_runtime_backup")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    ! runtime_backuping.f90:198 read_runtime_backup (file: NUL-terminated)
    integer(c_int) function flusi_read_runtime_backup(ctx, file) bind(C, name="flusi_read_runtime_backup")
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: ctx
      character(kind=c_char), dimension(*) :: file
    end function

    integer(c_int) function flusi_get_time(ctx, time, it, dt0, dt1) bind(C, name="flusi_get_time")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: time, dt0, dt1
      integer(c_int), intent(out) :: it
    end function

    ! mask.f90:2 create_mask(time,...)
    integer(c_int) function flusi_create_mask(ctx, time) bind(C, name="flusi_create_mask")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: time
    end function

    integer(c_int) function flusi_get_mask(ctx, mask, us, color) bind(C, name="flusi_get_mask")
      import :: c_int, c_ptr, c_double, c_int16_t
      type(c_ptr), value :: ctx
      real(c_double), dimension(*), intent(out) :: mask, us
      integer(c_int16_t), dimension(*), intent(out) :: color
    end function

    ! integrals.f90:48 write_integrals_fsi (cal_drag + energies)
    integer(c_int) function flusi_write_integrals(ctx, time, out) bind(C, name="flusi_write_integrals")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: time
      real(c_double), dimension(*), intent(out) :: out
    end function

    ! drag.f90:17 cal_drag(time,u,Insect); u may be c_null_ptr (ifft3 of uk)
    integer(c_int) function flusi_cal_drag(ctx, time, u, out) bind(C, name="flusi_cal_drag")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: time
      type(c_ptr), value :: u
      real(c_double), dimension(*), intent(out) :: out
    end function

    integer(c_int) function flusi_energy_k(ctx, e) bind(C, name="flusi_energy_k")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: e
    end function

    ! vis.f90:9 cal_vis(dt,vis)
    integer(c_int) function flusi_cal_vis(ctx, dt, vis) bind(C, name="flusi_cal_vis")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: dt
      real(c_double), dimension(*), intent(out) :: vis
    end function

    ! cof_p3dfft.f90:33 fft(inx,outk) / :65 ifft(ink,outx)
    integer(c_int) function flusi_fft(ctx, f, fk) bind(C, name="flusi_fft")
      import :: c_int, c_ptr, c_double, c_double_complex
      type(c_ptr), value :: ctx
      real(c_double), dimension(*), intent(in) :: f
      complex(c_double_complex), dimension(*), intent(out) :: fk
    end function

    integer(c_int) function flusi_ifft(ctx, fk, f) bind(C, name="flusi_ifft")
      import :: c_int, c_ptr, c_double, c_double_complex
      type(c_ptr), value :: ctx
      complex(c_double_complex), dimension(*), intent(in) :: fk
      real(c_double), dimension(*), intent(out) :: f
    end function

    integer(c_int) function flusi_timers(ctx, enable, ms, names) bind(C, name="flusi_timers")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: enable
      real(c_double), dimension(*), intent(out) :: ms
      type(c_ptr), value :: names
    end function

    integer(c_int) function flusi_get_pressure(ctx, time, p) bind(C, name="flusi_get_pressure")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: time
      real(c_double), dimension(*), intent(out) :: p
    end function

    integer(c_int) function flusi_autotune(ctx, reps, variant, ms) bind(C, name="flusi_autotune")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: reps
      integer(c_int), dimension(6), intent(out) :: variant
      real(c_double), intent(out) :: ms
    end function

    ! discrete adjoint: d(mean lift)/d(wing kinematics Fourier coefficients)
    integer(c_int) function flusi_adjoint_ncontrols(ctx, n) bind(C, name="flusi_adjoint_ncontrols")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), intent(out) :: n
    end function

    integer(c_int) function flusi_adjoint_get_controls(ctx, theta) bind(C, name="flusi_adjoint_get_controls")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), dimension(*), intent(out) :: theta
    end function

    integer(c_int) function flusi_adjoint_forward(ctx, theta, nsteps, every, J) bind(C, name="flusi_adjoint_forward")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), dimension(*), intent(in) :: theta
      integer(c_int), value :: nsteps, every
      real(c_double), intent(out) :: J
    end function

    integer(c_int) function flusi_adjoint_backward(ctx, grad) bind(C, name="flusi_adjoint_backward")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), dimension(*), intent(out) :: grad
    end function

    integer(c_size_t) function c_strlen(s) bind(C, name="strlen")
      import :: c_ptr, c_size_t
      type(c_ptr), value :: s
    end function
  end interface

contains

  ! message of the last failed call, as a Fortran string
  function hip_error_message(ctx) result(msg)
    type(c_ptr), intent(in) :: ctx
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: chars(:)
    integer :: n, i
    p = flusi_last_error(ctx)
    n = int(c_strlen(p))
    call c_f_pointer(p, chars, [n])
    allocate(character(len=n) :: msg)
    do i = 1, n
      msg(i:i) = chars(i)
    end do
  end function

  ! abort(code, msg) on a failed call (vars.f90:327-336 semantics, one rank)
  subroutine hip_check(ctx, rc)
    type(c_ptr), intent(in) :: ctx
    integer(c_int), intent(in) :: rc
    if (rc /= 0) then
      write(*,'("flusi_hip: abort code ",i0,": ",a)') rc, hip_error_message(ctx)
      error stop 1
    endif
  end subroutine

end module flusi_hip
