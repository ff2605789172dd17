/*
 * oshun_hip.h -- C ABI of the MI355X (gfx950) distribution-function update.
 *
 * One context owns the device-resident state of one OSHUN run (f_lm and the
 * six field components), its coefficient tables and one HIP stream.  Calls on
 * one context are serialised by the caller; separate contexts are independent.
 * No C++ types, no exceptions and no torch types cross this boundary.  Every
 * function returns 0 on success and a negative OSHUN_ERR_* code on failure;
 * oshun_last_error() then describes the failure.
 *
 * The entry points replace, one for one, the reference's interfaces on the
 * hot path (OSHUN-1, file:line):
 *   Runge_Kutta_4::advance_p1            def-RK.cpp:110-145, decl-RK.h:48-79
 *   Actions::Adv                         def-Actions.cpp:73-94, decl-actions.h:47-69
 *   Explicit_EC_FP::loop                 def-FokkerPlanck.cpp:590-625, decl-fokkerplanck.h:140-163
 *   Euler_Backward::advance_1/advance_lm def-FokkerPlanck.cpp:1270-1304, decl-fokkerplanck.h:289-311
 *   main_loop step body                  main.cpp:79-107
 *   Stat storage (upload/download)       decl-state.h:302-334, def-State.cpp:413-427
 *
 * Host layouts.  A harmonic is complex<double>, p fastest, then x, then y,
 * exactly a reference SHarmonic (matrices.h:713-716) when nbc = RKLevel guard
 * cells are included on each side of x and y; nbc = 0 means interior only.
 * Harmonics are in the reference's DFunc order ind(l,m) (def-State.cpp:422-426).
 * Fields are complex<double>[ny+2nbc][nx+2nbc] in the order Ex,Ey,Ez,Bx,By,Bz.
 * The state is real for m = 0 and for the fields; imaginary parts of those
 * are ignored on upload and written as 0 on download.
 */
#ifndef OSHUN_HIP_H
#define OSHUN_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define OSHUN_OK 0
#define OSHUN_ERR_ARG (-1)
#define OSHUN_ERR_HIP (-2)
#define OSHUN_ERR_ALLOC (-3)
#define OSHUN_ERR_STATE (-4)

/* precision of the stored distribution function and of the Vlasov RHS */
#define OSHUN_FP64 0       /* everything in double                                 */
#define OSHUN_FP32_SPLIT 1 /* f_lm + Vlasov RHS in float; fields, collisions double.
                              The float plane of f00 holds f00 - bg(p), bg the FP64
                              x,y-mean of f00 at the last upload, so rounding acts on
                              the deviation; uploads/downloads/moments add bg back. */

/* collision switches (bit mask) */
#define OSHUN_COLL_EE 1  /* e-e on f00            (Explicit_EC_FP)             */
#define OSHUN_COLL_FLM 2 /* implicit l >= 1 step  (Euler_Backward)             */

typedef struct oshun_grid {
  int l0, m0;           /* harmonic truncation, l <= l0, m <= min(l, m0)         */
  int nump, numx, numy; /* interior cells: momentum, x, y (numy = 1 for 1D)      */
  double pmin, pmax;    /* momentum axis p_i = pmin + i (pmax-pmin)/(nump-1)      */
                        /* (matrices.h:860-866; decks use pmin = pmax/(2 nump))  */
  double dx, dy;        /* cell sizes (the reference's x.dx(), y.dx())            */
  int precision;        /* OSHUN_FP64 or OSHUN_FP32_SPLIT                         */
  int bndx, bndy;       /* boundaries: 0 periodic, 1 mirror (bndX, bndY of the
                           deck, def-Parallel.cpp:1121-1307)                     */
  int nbc;              /* with a mirror boundary: the guard cells per side
                           (RKLevel); the device then keeps the whole arrays,
                           guards included, and refreshes the guards at the end
                           of each oshun_step as Neighbor_Communications does;
                           numx, numy >= nbc.  Ignored when both are periodic. */
} oshun_grid;

typedef struct oshun_phys {
  double density_np;  /* reference density [cm^-3] (Coulomb logs, collision rate) */
  double zeta;        /* ion charge Z                                             */
  double small_dt;    /* e-e sub-step target: numh = int(Dt/small_dt)+1          */
  int nb_algorithms;  /* low-velocity cells with the parabolic G (NB)           */
  int collisions;     /* OR of OSHUN_COLL_EE, OSHUN_COLL_FLM                      */
  int ref_quirks;     /* 1: reproduce def-VlasovMax.cpp:918 (y-advection slip)   */
} oshun_phys;

typedef struct oshun_ctx oshun_ctx;

int oshun_create(const oshun_grid* grid, const oshun_phys* phys, oshun_ctx** out);
int oshun_destroy(oshun_ctx* ctx);
/* The kernels a context of this grid would run, decided on the host alone
 * (nothing is allocated; no GPU is needed): *stage = 0 the 1D stage-pair
 * kernels, 1 the 1D single-stage kernel, 2 the compile-time specialised
 * general kernel, 3 the term-list interpreter, 4 the l-streaming general
 * kernel (m0 = 1..4, any l0); *collide = 0 the lane-layout
 * collision kernel, 1 one wavefront per column.  The stage pairs address the
 * whole state with 32-bit byte offsets and the lane-layout kernel one plane,
 * so states of 2 GiB or more (planes of 2 GiB) take the 64-bit kernels.  No
 * counterpart in the reference. */
int oshun_kernel_plan(const oshun_grid* grid, const oshun_phys* phys, int* stage, int* collide);
const char* oshun_last_error(const oshun_ctx* ctx);
int oshun_num_harmonics(const oshun_ctx* ctx);

/* Stat <-> device.  One host pointer per harmonic and per field (the shape
 * of a reference Stat, decl-state.h:302-334); guards are skipped on upload
 * and refilled with periodic images on download.  A mirror-boundary context
 * takes nbc = its guard count (the whole arrays, guards copied as they are)
 * or nbc = 0 (the interior; upload then refreshes the guards). */
int oshun_upload_harmonics(oshun_ctx* ctx, const double* const* sh, const double* const* fld, int nbc);
int oshun_download_harmonics(oshun_ctx* ctx, double* const* sh, double* const* fld, int nbc);
/* Same with all harmonics (resp. all fields) contiguous in one buffer. */
int oshun_upload_state(oshun_ctx* ctx, const double* f, const double* fld, int nbc);
int oshun_download_state(oshun_ctx* ctx, double* f, double* fld, int nbc);

/* Clock shared by the RK4 driver and both collision operators
 * (def-RK.cpp:53, def-FokkerPlanck.cpp:572,1262). */
int oshun_set_time(oshun_ctx* ctx, double t);
int oshun_get_time(const oshun_ctx* ctx, double* t);

/* Actions::Adv of the current state; the slope goes to host buffers laid out
 * like oshun_download_state with nbc = 0. */
int oshun_rhs(oshun_ctx* ctx, double* f_slope, double* fld_slope);
/* Runge_Kutta_4::advance_p1: one classic RK4 step of size h; t += h. */
int oshun_rk4_step(oshun_ctx* ctx, double h);
/* Explicit_EC_FP::loop(tnew): e-e relaxation of f00 over tnew - t_ee. */
int oshun_collide_ee(oshun_ctx* ctx, double tnew);
/* Euler_Backward::advance_1 (which = 1: l = 1) and advance_lm (which = 2:
 * l >= 2, then t_flm = tnew). */
int oshun_collide_flm(oshun_ctx* ctx, double tnew, int which);
/* nsteps x main.cpp:79-107 (RK4, e-e, f1, flm, guard refresh; with periodic
 * boundaries the guards are implicit). */
int oshun_step(oshun_ctx* ctx, double h, int nsteps);
/* Parallel_Environment::Neighbor_Communications of one node
 * (def-Parallel.cpp:3040-3127 with sameNode_periodic_X/Y and
 * sameNode_mirror_X/Y, :1121-1307) on a mirror-boundary context; oshun_step
 * runs it at the end of every step.  OSHUN_ERR_ARG on periodic contexts. */
int oshun_fill_guards(oshun_ctx* ctx);
/* nsteps x main.cpp:79-107 with implicit_E and if_implicit1D (the RKLevel = 2
 * loop): Runge_Kutta_2_Imp::advance_p1 (def-RK.cpp:470-498), e-e,
 * Implicit_E_Field::advance (def-ImplicitE.cpp:472-606, 1D variant),
 * Runge_Kutta_2_Imp::advance_p2 (def-RK.cpp:503-524), Euler_Backward::
 * advance_1/advance_lm.  if_implicitES: Adv1 without Faraday
 * (def-Actions.cpp:118-132).  1D runs with m0 = 0, 3 <= l0 <= 12, FP64;
 * otherwise OSHUN_ERR_ARG. */
int oshun_step_implicit_1d(oshun_ctx* ctx, double h, int nsteps, int if_implicitES);
/* nsteps x the same loop without if_implicit1D, the variant of the production
 * decks (def-Input.cpp:87-95 of each POP deck): Runge_Kutta_2_Imp::advance_p1 (Adv1:
 * advection, and B rotation + Faraday unless if_implicitES), e-e, sources,
 * Implicit_E_Field::advance (def-ImplicitE.cpp:472-606: J0 and the Ex, Ey, Ez
 * probes through Implicit_Ex / Implicit_Eyz, def-VlasovMax.cpp:157-291, each
 * followed by Euler_Backward::advance_1; JN from Ampere or, if_implicitES, the
 * current set below; the 3x3 conductivity solve by Cramer's rule,
 * def-Nmethods.cpp:215-251, retried up to 4 times on zero determinants),
 * advance_p2 (Exyz + CLEAN_LOW_PC), advance_1/advance_lm.  l0 >= 3, m0 >= 2,
 * Np <= 512, FP64; otherwise OSHUN_ERR_ARG.  More than 8 cells with a zero
 * determinant after the last attempt, where the reference stops the run
 * (:597-600), return OSHUN_ERR_STATE after the step's field solve. */
int oshun_step_implicit_2d(oshun_ctx* ctx, double h, int nsteps, int if_implicitES);
/* JN of the electrostatic 2D variant.  With if_implicitES the reference skips
 * Ampere and JN keeps what Current_xyz's constructor gave it: a copy of the Ex
 * field at the solver's construction, in all three components
 * (def-ImplicitE.cpp:56-58, 446).  jn: 3 complex fields (x, y, z) laid out
 * like three of oshun_upload_state's fields, guards of width nbc included;
 * the real parts are used (a mirror-boundary context takes its whole arrays:
 * nbc = its guard count).  Zero until set. */
int oshun_set_implicit_es_current(oshun_ctx* ctx, const double* jn, int nbc);
/* Cells whose conductivity determinant was still zero after the last
 * oshun_step_implicit_2d step's final attempt (their field is set to 0, as
 * the reference does; it prints a warning). */
int oshun_implicit_zero_dets(const oshun_ctx* ctx);

/* ---- Sources (def-PLSource.cpp) -------------------------------------------
 * The deck's laser parameters (Inputdata input_list, decl-input.h:104-109). */
typedef struct oshun_sources {
  int linear_Ivst, polynomial_Ivst; /* PLaserSource on when either is set (decl-plsource.h:47)         */
  int inverse_bremsstrahlung;       /* InverseBremsstrahlung on when 1 (def-PLSource.cpp:772)          */
  int polarization_direction;       /* ExternalEfield: 1 Ex, 2 Ey, 3 Ez (def-PLSource.cpp:356-364)     */
  double I_0, lambda_0;             /* intensity [W/cm^2], wavelength [um]                              */
  double ab_eta, ab_depth;          /* absorbed fraction, absorption depth                              */
  double cntr_x0, cntr_r0, spot_w0; /* profile centre in x, in y, and spot size                         */
  double rise_time, flat_time, fall_time; /* time envelope                                             */
  double smaller_dt;                /* inverse-bremsstrahlung sub-step: numh = int(Dt/smaller_dt)+1     */
} oshun_sources;
/* Sets the sources of the context (the constructors of PLaserSource,
 * InverseBremsstrahlung and ExternalEfield, def-PLSource.cpp:73-133, :308-353,
 * :861-946).  x[nx], y[ny]: positions of the interior cells on the
 * reference's axes (Inputdata::decl_input, def-Input.cpp:313-325) -- of all
 * cells, guards included, on a mirror-boundary context, where the laser
 * source and the external field act on every cell and inverse bremsstrahlung
 * on the interior, as in the reference (def-PLSource.cpp:152-230, 395-410,
 * 838-839); xglob:
 * its global x axis (nxglob points, the laser profile's normalisation).  The
 * source clocks start at the context's time.  oshun_step then runs
 * PLaserSource::Laser_Source and InverseBremsstrahlung::loop after the e-e
 * collisions of every step, where main_loop calls them (main.cpp:90-92), as
 * does oshun_step_implicit_1d. */
int oshun_set_sources(oshun_ctx* ctx, const oshun_sources* src, const double* x, const double* y,
                      const double* xglob, int nxglob);
/* PLaserSource::Laser_Source(tnew) (def-PLSource.cpp:138-252).  A heating
 * rate outside [0, 0.99], where the reference stops the run, returns
 * OSHUN_ERR_STATE (those cells are left unheated); oshun_step checks it too. */
int oshun_laser_source(oshun_ctx* ctx, double tnew);
/* InverseBremsstrahlung::loop(tnew) (def-PLSource.cpp:780-861). */
int oshun_inverse_bremsstrahlung(oshun_ctx* ctx, double tnew);
/* ExternalEfield::Laser_Efield(tnew) (def-PLSource.cpp:368-418): sets the
 * polarisation's field component on every cell; main_loop has this call
 * commented out (main.cpp:91), so no step runs it. */
int oshun_laser_efield(oshun_ctx* ctx, double tnew);

/* Nonrelativistic moments of the current state in double
 * (Parallel_Output::nonrelativistic_moments, def-Parallel.cpp:1677-2253;
 * the reference computes them in float).  out = double[OSHUN_NMOMENTS][ny][nx]
 * in the order: density, Vx, Vy, VxVx, VyVy, VzVz, Qx, VxVy, VxVz, VyVz,
 * Vsq (<v^2> - Vx^2 - Vy^2), T [eV], P [Mbar], N_D, nu, Qy, Kx, Ky.
 * Harmonics a run does not have count as zero.  Kx and Ky difference
 * <v^2>^2 across the interior as the single-rank reference does (end cells
 * take their neighbour's difference); Ky is 0 when ny < 3.  The first call
 * allocates the context's moment buffers. */
#define OSHUN_NMOMENTS 18
int oshun_moments(oshun_ctx* ctx, double* out);
/* The same moments into a DEVICE buffer out_dev (double[OSHUN_NMOMENTS][ny][nx]),
 * enqueued on the context's stream (oshun_synchronize / oshun_get_stream to
 * wait), without allocating: no hipMalloc, hipFree or other allocation in
 * this call, ever, also on its first call.  No counterpart in the reference;
 * it is for callers that cannot allocate (code under HIP graph capture, tight
 * serving loops).  The partial p sums go to one static device array per GPU
 * (8 MiB, kept until the process ends), so nx*ny is limited to 104857
 * (OSHUN_ERR_ARG beyond).  Calls that overlap on one GPU (contexts on
 * different streams of the same device) share that array and race; calls on
 * one context are ordered by its stream.  Both entry points split the p sums
 * into the same chunks (as many as the array holds, at most 32), so their
 * results are bitwise equal. */
int oshun_moments_noalloc(oshun_ctx* ctx, double* out_dev);

/* Stream and device views, for callers that keep data on the GPU. */
int oshun_synchronize(oshun_ctx* ctx);
int oshun_get_stream(const oshun_ctx* ctx, void** hip_stream);
/* Device state: f planes (float or double per precision) and 6 double field
 * planes.  Plane layout [y][p][x] (x fastest); see oshun_plane_info.  In
 * OSHUN_FP32_SPLIT plane 0 is f00 minus the background (see above). */
int oshun_device_state(oshun_ctx* ctx, void** f_dev, double** fld_dev);
/* Copies the device state to (into_ctx = 0) or from (into_ctx = 1) caller
 * device buffers of the same layout (f in the storage precision); either
 * pointer may be NULL.  Synchronises the context's stream. */
int oshun_copy_device_state(oshun_ctx* ctx, void* f_dev, double* fld_dev, int into_ctx);
/* ncomp real planes; comp_l/comp_m/comp_im describe each plane (arrays of
 * length ncomp, may be NULL); plane_elems = ny*np*nx. */
int oshun_plane_info(const oshun_ctx* ctx, int* ncomp, long long* plane_elems, int* comp_l, int* comp_m,
                     int* comp_im);

/* Per-kernel timing with HIP events on the context's stream (off by default).
 * Kernel classes: 0 Vlasov stage, 1 field stage, 2 collisions. */
int oshun_profile_enable(oshun_ctx* ctx, int on);
int oshun_profile_read(oshun_ctx* ctx, double* ms_total, long long* launches, int nclasses);

/* Diagnostics: a plain device copy of nbytes (<= the stored f size) between
 * two scratch buffers of the context, used to calibrate hardware byte
 * counters (FETCH_SIZE / WRITE_SIZE) against a kernel of known traffic. */
int oshun_diag_copy(oshun_ctx* ctx, long long nbytes);
/* Diagnostics: which kernel runs the Vlasov stage of this context
 * (*which = 0 the 1D m0 = 0 kernels, 1 the compile-time specialised general
 * kernel, 2 the term-list interpreter, 3 the l-streaming general kernel).
 * set = 1 forces the interpreter (also for the implicit step's programs),
 * 0 restores the default, -1 only queries.  The general kernels evaluate
 * Actions::Adv (def-Actions.cpp:73-94) with the same per-target sums. */
int oshun_stage_kernel(oshun_ctx* ctx, int set, int* which);
/* p rows of the 1D stage-pair kernels.  Positive: the tile kernels with
 * tiles of 10 rows (6 for l0 <= 8).  Negative: the marched kernels with -rows
 * rows per wavefront (4 .. 32).  By default grids that fill the chip take the
 * marched kernels with the fewest rows that fit one round of wavefronts, small
 * grids the tile kernels (6 rows where 10 would give fewer than two
 * workgroups per CU).  set chooses (6, 10, -4 .. -32), 0 only queries (*rows =
 * 0 on contexts without stage pairs).  The choice groups the p sum of the
 * current differently (results agree to rounding, not bitwise); a batch
 * member equals its run alone bitwise at the same tile height.  No
 * counterpart in the reference. */
int oshun_pair_rows(oshun_ctx* ctx, int set, int* rows);

/* ---- Batched independent runs ------------------------------------------
 * Many 1D runs (m0 = 0, ny = 1, l0 <= 8, one precision for all) of uneven
 * Nx, Np and l0 advanced together: every step is one hipGraph launch whose
 * kernels each cover the tiles of all runs.  No counterpart in the
 * reference, which runs one deck per process; each run here behaves as
 * main.cpp:79-107 (Runge_Kutta_4::advance_p1, Explicit_EC_FP::loop,
 * Euler_Backward::advance_1/advance_lm) with its own step h[i].
 * oshun_batch_run returns the run's context for upload/download/moments;
 * its device buffers are owned by the batch. */
typedef struct oshun_batch oshun_batch;
int oshun_batch_create(const oshun_grid* grids, const oshun_phys* phys, const double* h, int nruns,
                       oshun_batch** out);
int oshun_batch_destroy(oshun_batch* batch);
const char* oshun_batch_last_error(const oshun_batch* batch);
int oshun_batch_size(const oshun_batch* batch);
oshun_ctx* oshun_batch_run(oshun_batch* batch, int i);
int oshun_batch_step(oshun_batch* batch, int nsteps);
int oshun_batch_synchronize(oshun_batch* batch);
/* Times `steps` steps with each candidate stage-pair kernel (tiles of 6, 10
 * and in FP32 14 p rows; marched kernels with 6, 8, 12 rows per wavefront) on
 * the current state (restored afterwards) and keeps the fastest; reports its
 * p rows as oshun_pair_rows does (negative: marched) and ms per step. */
int oshun_batch_autotune(oshun_batch* batch, int steps, int* chosen_rows, double* ms_per_step);
/* The batch's stage-pair kernels as oshun_pair_rows sets a run's (6, 10, 14
 * in FP32, -4 .. -32 marched; 0 only queries); re-records the step graphs.
 * A member equals its run alone bitwise at the same setting.  No counterpart
 * in the reference. */
int oshun_batch_pair_rows(oshun_batch* batch, int set, int* rows);

/* ---- Adjoint of the step -------------------------------------------------
 * Reverse-mode derivative (vector-Jacobian product) of oshun_step: RK4 of the
 * Vlasov-Maxwell right-hand side (Runge_Kutta_4::advance_p1), then the e-e
 * sub-cycles (Explicit_EC_FP::loop) and the implicit l >= 1 solves
 * (Euler_Backward::advance_1/advance_lm), each transposed by hand.  No
 * counterpart in the reference.  1D runs with m0 = 0 and 1 <= l0 <= 8 in
 * OSHUN_FP64 only (other contexts get OSHUN_ERR_ARG).
 * Cotangents are DEVICE buffers in the device-state layout (oshun_device_state):
 * lam_f = ncomp planes of np*nx doubles, lam_fld = 6*nx doubles.
 *
 * oshun_adjoint_step: the context holds the step's input y_n (unchanged on
 * return); on entry lam_* is the cotangent of y_{n+1} = step(y_n, h), on
 * return that of y_n. */
int oshun_adjoint_step(oshun_ctx* ctx, double h, double* lam_f, double* lam_fld);
/* Checkpointed runs: oshun_adjoint_forward advances nsteps steps of size h
 * from the current state, keeping a copy of the state every `every` steps on
 * the device; oshun_adjoint_backward then turns the cotangent of the final
 * state into that of the initial state, recomputing each segment's steps from
 * its checkpoint (the context ends at the final state again).  A context keeps
 * one recorded run: every oshun_adjoint_forward replaces the previous one and
 * returns its generation number (if `generation` is not NULL);
 * oshun_adjoint_backward with generation > 0 fails with OSHUN_ERR_STATE unless
 * that run is still the recorded one (generation <= 0: the recorded run). */
int oshun_adjoint_forward(oshun_ctx* ctx, double h, int nsteps, int every, long long* generation);
int oshun_adjoint_backward(oshun_ctx* ctx, long long generation, double* lam_f, double* lam_fld);
/* frees the adjoint work buffers and checkpoints (also done by oshun_destroy) */
int oshun_adjoint_release(oshun_ctx* ctx);

#ifdef __cplusplus
}
#endif
#endif /* OSHUN_HIP_H */
