!-------------------------------------------------------------------------------
! flusi_amd PARAMS.ini [fp64|fp32]
!
! Fortran host for the fluid hot path: reads a FLUSI .ini, initialises the
! fields (init_fields.f90), and runs the time_step.f90 loop on the GPU,
! writing the .t time series (forces.t, energy.t, ekin.t, dt.t, ...) to the
! current directory, as `flusi PARAMS.ini` does for this path.
!-------------------------------------------------------------------------------
program flusi_amd_main
  use iso_c_binding
  use flusi_hip
  implicit none
  type(c_ptr) :: ctx
  character(len=1024) :: ini, mode
  integer(c_int) :: rc, prec, it, nx, ny, nz
  real(c_double) :: time, dt0, dt1

  if (command_argument_count() < 1) then
    write(*,*) "usage: flusi_amd PARAMS.ini [fp64|fp32]"
    stop 1
  endif
  call get_command_argument(1, ini)
  prec = FLUSI_FP64
  if (command_argument_count() >= 2) then
    call get_command_argument(2, mode)
    if (trim(mode) == "fp32") prec = FLUSI_FP32_MIXED
  endif

  rc = flusi_init(trim(ini)//c_null_char, prec, ctx)
  call hip_check(ctx, rc)
  call hip_check(ctx, flusi_get_dims(ctx, nx, ny, nz))
  write(*,'("flusi_amd: ",i0,"x",i0,"x",i0," grid, precision ",i0)') nx, ny, nz, prec
  call hip_check(ctx, flusi_set_output_dir(ctx, "."//c_null_char))
  call hip_check(ctx, flusi_init_fields(ctx))
  call hip_check(ctx, flusi_run(ctx, 0_c_int))
  call hip_check(ctx, flusi_get_time(ctx, time, it, dt0, dt1))
  write(*,'("flusi_amd: done, time=",es15.8," it=",i0," dt=",es12.4)') time, it, dt1
  call flusi_free(ctx)
end program flusi_amd_main
