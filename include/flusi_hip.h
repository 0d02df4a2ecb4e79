/*
 * flusi_hip.h -- C ABI of the MI355X implementation of FLUSI's fluid hot path.
 *
 * Every entry point returns 0 on success or the reference's abort code
 * (vars.f90:327-336 abort(code,msg)); flusi_last_error() returns the message.
 * Arrays crossing the boundary are plain host arrays in the reference's
 * single-rank layouts (see below); the step loop keeps all fields resident on
 * the GPU.  One context per GPU, one HIP stream, single host thread, not
 * re-entrant (as the reference, which is one thread per MPI rank).
 *
 * Layouts (0-based, first index fastest, as the Fortran arrays):
 *   real field       u(ix,iy,iz,comp)   nx*ny*nz doubles per component
 *   spectral field   uk(iz,iy,ix,comp)  complex(8) interleaved re/im,
 *                    nz*ny*(nx/2+1) values per component, ix <-> kx >= 0
 *   mask_color       int16 (ix,iy,iz)   (vars.f90:9)
 */
#ifndef FLUSI_HIP_H
#define FLUSI_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct flusi_ctx flusi_ctx;

#define FLUSI_FP64 0        /* plain FP64 solver (the reference's pr=8) */
#define FLUSI_FP32_MIXED 1  /* FP32 FFTs/storage, FP64 pointwise, IF and reductions */

/* Reads the .ini (params.f90:3-44 get_params, ini_files_parser.f90) and
 * allocates the device state (flusi.f90:112-286 Start_Simulation). */
int flusi_init(const char* ini_path, int precision, flusi_ctx** ctx);
void flusi_free(flusi_ctx* ctx);
const char* flusi_last_error(const flusi_ctx* ctx);
int flusi_get_dims(const flusi_ctx* ctx, int* nx, int* ny, int* nz);

/* Directory for the .t time series (forces.t, energy.t, ekin.t, dt.t,
 * kinematics.t, ...); NULL or "" disables file output. */
int flusi_set_output_dir(flusi_ctx* ctx, const char* dir);

/* inicond/init_fields.f90:2-114 (VortexRing, MeanFlow, quiescent), the
 * insect initialisation (insect_init_clean.f90:1-250) and the first
 * create_mask. */
int flusi_init_fields(flusi_ctx* ctx);

/* State copies in the reference layouts (uk: 3 components). */
int flusi_set_uk(flusi_ctx* ctx, const double* uk);
int flusi_get_uk(flusi_ctx* ctx, double* uk);
/* u = ifft3(uk) (cof_p3dfft.f90:122-145), 3 real components. */
int flusi_get_u(flusi_ctx* ctx, double* u);
/* vorticity = ifft3(curl(uk)) (basic_operators.f90:185-215). */
int flusi_get_vorticity(flusi_ctx* ctx, double* vort);
/* The pressure field save_fields writes (save_fields.f90:98-133): the mask at
 * `time` (iMoving=1), the unprojected right-hand side (rhs.f90:147-330),
 * pressure (rhs.f90:342-379: -i k.nlk/k^2), ifft, minus |u|^2/2.  p: one
 * real field. */
int flusi_get_pressure(flusi_ctx* ctx, double time, double* p);

/* rhs.f90:2-115 cal_nlk: projected right-hand side of the current uk at
 * `time` (mask regenerated when iMoving=1).  nlk: 3 spectral components or
 * NULL; u: physical velocity (3 components) or NULL. */
int flusi_cal_nlk(flusi_ctx* ctx, double time, int it, double* nlk, double* u);

/* fluid_time_step.f90:34-150 FluidTimestep for the configured stepper
 * (AB2 with euler_startup at it=0, RK2, RK4): advances uk from `time` by one
 * step and returns dt1; dt0 is the previous step (as in time_step.f90:61). */
int flusi_fluid_timestep(flusi_ctx* ctx, double* time, int* it, double* dt0, double* dt1);

/* One iteration of time_step.f90:57-167: dt0=dt1, FluidTimestep, swap
 * n0/n1, time+=dt1, it+=1, write_integrals on the tintegral/itdrag cadence,
 * save_fields on the tsave cadence (time_step.f90:97-105; HDF5 snapshots are
 * written only when an output directory is set). */
int flusi_step(flusi_ctx* ctx);
/* time_step.f90:1-180: unless the fields came from a backup, the initial
 * snapshot (init_fields.f90:108-111) and write_integrals; then steps while
 * time<Tmax and it<=nt (nsteps>0 stops after nsteps iterations), a runtime
 * backup every truntime wall-clock hours, the wtimemax stop, and a final
 * backup (iDoBackup=1 and an output directory set). */
int flusi_run(flusi_ctx* ctx, int nsteps);
/* nsteps iterations of flusi_step without the initial write_integrals. */
int flusi_advance(flusi_ctx* ctx, int nsteps);

/* save_fields.f90:46-237 save_fields_fsi at the current time: the iSave*
 * fields of the [Saving] section as HDF5 files "<field>_<tag>.h5" in the
 * output directory (tag: nint(1000 t), or it with naming=timestep), dataset
 * <field>, field_precision single (default) or double, with the attributes
 * of flusi_hdf5_interface.f90:58-63.  Needs libflusi_h5.so (HDF5). */
int flusi_save_fields(flusi_ctx* ctx);
/* runtime_backuping.f90:3-193 dump_runtime_backup: u and the registers
 * nlk(0), nlk(1) in physical space to runtime_backup<n>.h5 (n alternates
 * 0, 1; backup_type one-file-backup) or runtime_backup<n>_<dset>.h5, each
 * dataset with bckp = (time, dt1, dt0, n1, it, nx, ny, nz). */
int flusi_dump_runtime_backup(flusi_ctx* ctx);
/* runtime_backuping.f90:198-361 read_runtime_backup: state, time, dt0, dt1,
 * n1, it from a backup file (relative names in the output directory), the
 * integrating factor of dt1, and the mask at the restored time; 77776 on a
 * resolution mismatch.  inicond=backup::<file> or backup (the newer of
 * runtime_backup0/1.h5) in the .ini does the same inside flusi_init_fields. */
int flusi_read_runtime_backup(flusi_ctx* ctx, const char* file);
int flusi_get_time(const flusi_ctx* ctx, double* time, int* it, double* dt0, double* dt1);

/* mask.f90:2-31 create_mask at `time` (sphere, tethered insect, cavity). */
int flusi_create_mask(flusi_ctx* ctx, double time);
int flusi_get_mask(flusi_ctx* ctx, double* mask, double* us, int16_t* color);

/* integrals.f90:48-238 write_integrals_fsi at `time` from the current uk:
 * cal_drag (drag.f90:17-454), energies, divergence, mean flow, mask volume.
 * out (may be NULL) receives FLUSI_INTEGRALS_NV values:
 *   [0..2] Force (colours 1..5)   [3..5] Torque about (x0,y0,z0)
 *   [6..9] penalisation power (total, x, y, z)
 *   [10..13] E_kin_f (tot,x,y,z) [14..17] diss_f  [18..21] E_kin  [22..25] diss
 *   [26] max div  [27] max div in fluid  [28] mask volume  [29..31] sum mask*us dV
 *   [32..43] per-part force (body, left, right)   [44..52] per-part torque
 *   [53..58] sqrt(u_residual) colours 0..5  [59] aero power [60] inertial power */
#define FLUSI_INTEGRALS_NV 61
int flusi_write_integrals(flusi_ctx* ctx, double time, double* out);

/* drag.f90:17-454 cal_drag(time, u, Insect): per-colour forces, moments about
 * (x0,y0,z0) (the body centre for insects) and about the wing pivots,
 * penalisation power, the residual velocity min|u-us| inside the mask and,
 * for insects, aerodynamic and inertial power, from the physical velocity u
 * (3 real components in the reference layout; NULL: ifft3 of the current uk)
 * with the mask, us and colour as they stand -- a moving mask is redrawn by
 * flusi_create_mask first, as write_integrals_fsi does (integrals.f90:87-100).
 * Appends the forces.t, forces_part<n>.t and u_residual.t rows.  out (may be
 * NULL): the FLUSI_INTEGRALS_NV layout of flusi_write_integrals with slots
 * 0..9 and 32..60 set and the others zero.  Does nothing without
 * penalisation (drag.f90:53). */
int flusi_cal_drag(flusi_ctx* ctx, double time, const double* u, double* out);

/* integrals.f90:522-552 compute_energies_k of the current uk. */
int flusi_energy_k(flusi_ctx* ctx, double* E);
/* flusi_energy_k without any device allocation (no counterpart in the
 * reference).  For callers that cannot allocate device memory: code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc.
 * flusi_energy_k allocates its partial sums on its first call; this variant
 * keeps them (512 doubles), the reduction rows and the result in a
 * file-scope device array of 4096 doubles (32 KiB) that each GPU holds until
 * the process ends, and returns the same value to the bit.  Calls on one GPU
 * share that array: calls that overlap on one GPU (contexts driven from
 * different host threads) race.  900022 if the scratch would exceed its
 * maximum. */
int flusi_energy_k_noalloc(flusi_ctx* ctx, double* E);

/* vis.f90:9-66 cal_vis (real, spectral layout). */
int flusi_cal_vis(flusi_ctx* ctx, double dt, double* vis);

/* cof_p3dfft.f90:33-83 fft (R2C times 1/N) and ifft (C2R, unnormalised) of
 * one field in the reference layouts, through the device passes. */
int flusi_fft(flusi_ctx* ctx, const double* f, double* fk);
int flusi_ifft(flusi_ctx* ctx, const double* fk, double* f);

/* Device time in milliseconds accumulated per pass since the last reset
 * (module_timing.f90 toc slots); names[i] are static strings.  The passes
 * fuse several of the reference's slots: ZINV and YINV are the z and y parts
 * of "RHS (cal_nlk_fsi::ifft3)" and "(cal_nlk_fsi::curl+ifft3)"; XPASS their
 * x part, "(cal_nlk_fsi::nonlinear terms)", "(cal_nlk_fsi::vorticity_sponge)"
 * and the x part of "FFT"; YFWD the y part of "FFT"; ZSTEP its z part,
 * add_grad_pressure and the stepper update of "FluidTimestep (wrapper)";
 * MASK "create_mask (main wrapper)"; DIAG "Statistics (cal_drag)". */
#define FLUSI_TIMER_SLOTS 8
int flusi_timers(flusi_ctx* ctx, int enable, double* ms, const char** names);

/* Kernel-variant autotuner (no counterpart in the reference, which has one
 * FFT library plan per transform, cof_p3dfft.f90:160-180 p3dfft_setup).
 * Runs `reps` right-hand sides of the current state for every variant the
 * grid allows and keeps the fastest for this context; results are cached per
 * grid, precision and batch size for the life of the process.
 * variant[6] (may be NULL): in-wave y/z passes, block-pass prefetch, one-wave
 * x pass, its waves-per-SIMD register budget, in-wave z step, epilogue
 * positions the block z step schedules together (0 = all).  ms: time of one
 * right-hand side with the chosen variant. */
int flusi_autotune(flusi_ctx* ctx, int reps, int* variant, double* ms);

/* Discrete adjoint (no counterpart in the reference).  The objective is the
 * mean lift J = (1/K) sum_{n=1..K} F_z(t_n) over K RK4 steps from the initial
 * condition, F_z the z-force dV/eta sum chi (u_z - us_z) of the insect
 * (drag.f90:17-454 with the mask regenerated at t_n, integrals.f90:87-100).
 * The controls are the Fourier coefficients of the wings' from_file
 * kinematics (wings_motion.f90:98-222), shared by both wings when they read
 * one series: for angle a in (phi, alpha, theta) and n = nfft, control
 * a*2n + i is ai(i) and a*2n + n + i is bi(i), in the file's units.
 * Requires a moving insect mask, RK4, dt_fixed > 0, free mean flow and no
 * sponge.  flusi_adjoint_forward re-initialises the fields, runs K steps with
 * the given controls and keeps a checkpoint every `every` steps;
 * flusi_adjoint_backward then returns dJ/dtheta (the solver is left at the
 * initial state).  From the first adjoint run on, the context redraws the
 * whole insect at every mask update (the reference keeps the body from its
 * first drawing and clears only the previous wings, a history that makes the
 * mask non-differentiable in the kinematics where wing and body overlap). */
int flusi_adjoint_ncontrols(flusi_ctx* ctx, int* n);
int flusi_adjoint_get_controls(flusi_ctx* ctx, double* theta);
int flusi_adjoint_forward(flusi_ctx* ctx, const double* theta, int nsteps, int every, double* J);
int flusi_adjoint_backward(flusi_ctx* ctx, double* grad);

/* Batched ensembles: many independent simulations, each from its own .ini,
 * advanced together (no counterpart in the reference, which runs one
 * simulation per MPI job).  Members whose grids and solver parameters agree
 * form a batch whose passes run as one launch; with mode
 * FLUSI_ENSEMBLE_GRAPH one step of every batch is captured into a hipGraph
 * and replayed.  Members are ordinary contexts owned by the ensemble
 * (flusi_ensemble_member; flusi_free on them does nothing), so every
 * per-context call above works on them between advances.  Batched modes need
 * dt_fixed > 0; the graph mode also needs RK2 or RK4 and no time-dependent
 * forcing (sinusoidal mean flow, startup conditioner). */
typedef struct flusi_ensemble flusi_ensemble;
#define FLUSI_ENSEMBLE_MEMBERS 0  /* every member steps on its own */
#define FLUSI_ENSEMBLE_BATCHED 1  /* one launch per pass and batch */
#define FLUSI_ENSEMBLE_GRAPH 2    /* batched step captured once, replayed */
int flusi_ensemble_create(int n, const char* const* ini_paths, int precision, flusi_ensemble** ens);
void flusi_ensemble_free(flusi_ensemble* ens);
const char* flusi_ensemble_last_error(const flusi_ensemble* ens);
int flusi_ensemble_size(const flusi_ensemble* ens, int* nmembers, int* nbatches);
/* batch b: its member count, grid, and the index of the member that launches
 * for it (its flusi_timers count the batched launches) */
int flusi_ensemble_batch(flusi_ensemble* ens, int b, int* nmem, int* nx, int* ny, int* nz, int* lead);
int flusi_ensemble_member(flusi_ensemble* ens, int i, flusi_ctx** member);
int flusi_ensemble_init_fields(flusi_ensemble* ens);
/* variants: 6 ints per batch, ms: one per batch (either may be NULL) */
int flusi_ensemble_autotune(flusi_ensemble* ens, int reps, int* variants, double* ms);
/* advances every member by nsteps time steps (time_step.f90:57-167 each) */
int flusi_ensemble_advance(flusi_ensemble* ens, int nsteps, int mode);

#ifdef __cplusplus
}
#endif
#endif /* FLUSI_HIP_H */
