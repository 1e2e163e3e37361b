/*
 * smaug_hip.h -- C-ABI of the MI355X-native SMAUG MHD time-step engine.
 *
 * Plain C: pointers, sizes and status codes only.  Every entry point returns
 * SMAUG_OK (0) or a negative SMAUG_E* code; smaug_last_error() returns a
 * human-readable message for the calling thread's last failure.
 *
 * Data contract (kept identical to the reference so states drop in):
 *   - ghost width 2 per side; n[d] counts ghost cells (reference
 *     src/iosmaug.c:47-57);
 *   - field-major SoA, i (dim 0) fastest:
 *       idx = f*N + (k*n1 + j)*n0 + i,   N = n0*n1*n2
 *     (reference src/gradops.cu:89-97, fencode3/encode3);
 *   - field order of the state vector (reference include/iotypes.h:211,215):
 *       2D (NVAR=10): rho mom1 mom2 energy b1 b2 energyb rhob b1b b2b
 *       3D (NVAR=13): rho mom1 mom2 mom3 energy b1 b2 b3 energyb rhob b1b b2b b3b
 *     The first 2+2*ndim fields are perturbations that evolve, the rest are
 *     the static background;
 *   - boundary codes (reference include/iotypes.h:144):
 *       0 periodic, 1 mpi, 2 mpiperiod, 3 cont, 4 contcd4, 5 fixed,
 *       6 symm, 7 asymm, per field and direction (boundtype[f][d][0]).
 *       This build implements 0 (2D), 3, 4, 5, 6 and 7; a fixed ghost keeps
 *       the value of the initial state (or of the state last loaded), and
 *       the directions are applied in the order 0, 1, 2 as the reference's
 *       boundary passes do.
 *
 * Reference interfaces replaced (file:line in mikeg64/smaug):
 *   smaug_create      <- cuinit (src/init.cu:2222) + initgrid (src/init.cu:2502)
 *                        + initial cuboundary pass (src/iosmaug.c:480-488)
 *   smaug_step        <- body of the time loop (src/iosmaug.c:496-1144):
 *                        rkon=0 single stage (src/iosmaug.c:631-835) or the
 *                        intended VAC four-step RK (src/advance.cu:76-99)
 *   smaug_cfl         <- cucomputec + cucomputemaxc and the dt rule
 *                        (src/computedervfields.cu:2234,1864; src/iosmaug.c:541-616)
 *   smaug_fill_ghosts <- cuboundary (src/boundary.cu:231,
 *                        models/boundary.cu.fluxtube1:1137)
 *   smaug_get_state   <- cuupdate device->host copy (src/update.cu:165-236)
 *   smaug_set_state   <- host->device state upload in cuinit (src/init.cu:2222)
 *   smaug_get_grid    <- the wd positions/delx written by initgrid
 *                        (src/init.cu:2926-3010)
 *   smaug_destroy     <- cufinish (src/update.cu:477)
 */
#ifndef SMAUG_HIP_H_
#define SMAUG_HIP_H_

#ifdef __cplusplus
extern "C" {
#endif

#define SMAUG_OK 0
#define SMAUG_E_ARG -1      /* invalid argument or parameter set         */
#define SMAUG_E_HIP -2      /* a HIP runtime call failed                 */
#define SMAUG_E_NOMEM -3    /* allocation failed                         */
#define SMAUG_E_NUMERIC -4  /* non-finite wave speed / state (blow-up)   */
#define SMAUG_E_UNSUPPORTED -5

#define SMAUG_MAXVAR 13
#define SMAUG_MAXDIM 3

/* term-group mask: which parts of the right-hand side a step applies.
 * SMAUG_TERMS_ALL reproduces the reference step; the others exist so each
 * kernel family can be checked on its own. */
#define SMAUG_T_FLUX 0x01      /* centdiff1/centdiff2 flux divergence      */
#define SMAUG_T_ETERM 0x02     /* background energy terms (centdiff2c)     */
#define SMAUG_T_GRAV 0x04      /* gravity (cugrav)                         */
#define SMAUG_T_HRHO 0x08      /* density hyperdiffusion                   */
#define SMAUG_T_HE 0x10        /* energy hyperdiffusion                    */
#define SMAUG_T_HMOM 0x20      /* momentum hyperdiffusion (diag + cross)   */
#define SMAUG_T_HB 0x40        /* magnetic hyperdiffusion (diag + cross)   */
#define SMAUG_T_SOURCE 0x80    /* user source (driver) plug-in             */
#define SMAUG_TERMS_ALL 0xFF

/* user-source plug-ins (reference models/usersource.cu.*), the
 * addsourceterms2 of a model applied to the post-update value with the
 * stage's dt (src/centdiff2.cu:790-872); parameters in drv[] below */
#define SMAUG_DRIVER_NONE 0       /* usersource.cu.default                 */
#define SMAUG_DRIVER_VERTICAL 1   /* Gaussian vertical-velocity driver     */
#define SMAUG_DRIVER_TORSIONAL 2  /* usersource.cu.fluxtube1 (3D)          */
#define SMAUG_DRIVER_SPICULE1 3   /* usersource.cu.spicule1                */
#define SMAUG_NDRV 12

/* arithmetic of the evolved state */
#define SMAUG_PREC_FP64 0
#define SMAUG_PREC_MIXED 1        /* FP32 stencils, FP64 state and update  */

typedef struct smaug_params {
  int ndim;                 /* 2 or 3                                        */
  int n[3];                 /* cells per dim including 2+2 ghosts (n[2]=1 in 2D) */
  double gamma;             /* PGAMMA                                        */
  double mu;                /* PMU                                           */
  double eta;               /* PETA                                          */
  double g[3];              /* PGRAV0..2                                     */
  double dt;                /* DT: fixed step, or the cap when moddton=1     */
  double courant;           /* PCOURANT                                      */
  int rkon;                 /* PRKON: 0 single stage, 1 four-step RK         */
  int moddton;              /* PMODDTON: 1 = CFL step from the wave speeds   */
  int hyperdifmom;          /* PHYPERDIFMOM                                  */
  int divbon;               /* PDIVBON: 1 calls the divB source term         */
  double chyp[SMAUG_MAXVAR];/* PCHYP* per field                              */
  double xmin[3], xmax[3];  /* domain (interior)                             */
  double dx[3];             /* p->dx = (xmax-xmin)/NI, used by the CFL rule  */
  int boundtype[SMAUG_MAXVAR][3][2]; /* [field][dir][0=upper BCU,1=lower BCL] */
  int driver;               /* SMAUG_DRIVER_*                                */
  double drv[SMAUG_NDRV];   /* driver parameters (positions are the cells' wd
                               positions unless stated; dims 0, 1, 2):
     VERTICAL: A [m/s], period P [s], z0, x0, y0 [m], sigma_z, sigma_xy [m];
       v_z = A exp(-dz^2/sz^2 - dx^2/sxy^2 [- dy^2/sxy^2]) sin(2 pi qt / P),
       m_0 += dt v_z rho_t, e += dt v_z^2 rho_t / 2
     TORSIONAL (models/usersource.cu.fluxtube1:3-90): aa [m/s], tdep, xc1,
       xc2, xc3 [m], dz, dx, dy [m] (Gaussian widths along dims 0, 1, 2),
       xxmax, yymax [m]; with zp, xp, yp the offsets from xc1, xc2, xc3 along
       dims 0, 1, 2 and E = exp(-xp^2/dx^2) exp(-yp^2/dy^2) exp(-zp^2/dz^2):
       v_1 = (aa yp / yymax) E tdep, v_2 = -(aa xp / xxmax) E tdep,
       m_1 += dt v_1 rho_t, m_2 += dt v_2 rho_t,
       e += dt (v_1^2 + v_2^2) rho_t / 2                          (3D only)
     SPICULE1 (models/usersource.cu.spicule1:4-136): AA [m/s], tdep, xc1,
       xc2 [m], unused, delta_z, delta_x [m]; only in the step whose counter
       p->it is 1, at the grid positions xmin[d] + i_d dx[d] (i_d counting
       ghost cells): v = AA exp(-(x0-xc1)^2/delta_z^2)
       exp(-(x1-xc2)^2/delta_x^2) tdep, m_0 += dt v rho_t             */
  unsigned terms;           /* SMAUG_T_* mask, SMAUG_TERMS_ALL normally      */
  int precision;            /* SMAUG_PREC_*                                  */
  double qt0;               /* initial value of the nominal clock p->qt      */
  int it0;                  /* initial step counter p->it                    */
  int divbfix;              /* PDIVBFIX: with divbon=1, nonzero applies the
                               divB source terms (src/divb.cu:165-338); the
                               reference's kernel acts in 2D only            */
} smaug_params;

typedef struct smaug_ctx smaug_t;

/* Allocates the device state, uploads the initial state w0 (NVAR*N doubles,
 * ghost cells included) and the cell positions pos (ndim*N doubles, the
 * reference's wd pos1..pos3 slab), derives delx by the reference's grid
 * extrapolation and fills the ghost cells of the perturbation fields. */
int smaug_create(const smaug_params* p, const double* w0, const double* pos,
                 smaug_t** out);

/* Advances nsteps full time steps (rkon=0: one stage each; rkon=1: four). */
int smaug_step(smaug_t* s, int nsteps);

/* Wave-speed maxima of the current state per direction (max over all cells
 * of max(cfast_dir, c_sound)) and the time step the CFL rule would pick. */
int smaug_cfl(smaug_t* s, double* cmax_per_dim, double* dt);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the computation of smaug_cfl on a caller's device state,
 * with no context and no allocation of any kind.
 *   w_device   -- NVAR*N doubles on the device, the layout of smaug_create's
 *                 w0 (evolved fields, then the background), ghost cells
 *                 filled;
 *   out_device -- ndim + 2 doubles on the device: the wave-speed maxima per
 *                 direction, the time step the CFL rule picks (p->dt when
 *                 p->moddton is 0), and 1.0 when a wave speed is not finite
 *                 (blow-up) else 0.0;
 *   hip_stream -- the stream to queue on (NULL: the null stream).
 * Work is only queued (no synchronisation): read out_device after the
 * stream.  Its reduction runs in a fixed SMAUG_NOALLOC_SCRATCH_BYTES-byte
 * scratch array that the library holds once per GPU until the process ends;
 * the scratch does not grow with the grid, which is limited as smaug_create
 * limits it.  Every call on one GPU shares that array: calls that overlap on
 * one GPU (two streams, two threads) race and must be serialised by the
 * caller. */
#define SMAUG_NOALLOC_SCRATCH_BYTES 64
int smaug_cfl_noalloc(const smaug_params* p, const double* w_device, double* out_device, void* hip_stream);

/* Re-applies the boundary conditions to the current perturbation fields. */
int smaug_fill_ghosts(smaug_t* s);

/* Copies the full current state (NVAR*N doubles) to / from host memory. */
int smaug_get_state(const smaug_t* s, double* w_host);
int smaug_set_state(smaug_t* s, const double* w_host);

/* Copies positions (ndim*N) and delx (ndim*N) as used by the solver. */
int smaug_get_grid(const smaug_t* s, double* pos_host, double* delx_host);

/* Current step counter, simulated time (sum of dt), nominal clock qt and
 * the dt of the last step. */
int smaug_get_clock(const smaug_t* s, int* it, double* t, double* qt,
                    double* dt_last);

/* Device pointer of the current state (NVAR*N doubles, FP64) so a caller that
 * owns the stream can read it in place; valid until the next smaug_step. */
int smaug_state_device_ptr(const smaug_t* s, void** dptr);

/* Blocks until all work queued by the context has finished. */
int smaug_sync(smaug_t* s);

int smaug_destroy(smaug_t* s);

const char* smaug_last_error(void);
const char* smaug_version(void);

/* VAC configuration files (host only, no device needed).
 *   smaug_ini_read   <- readasciivacconfig (src/readwrite.c:975-1082): reads
 *                       an ASCII .ini of ndim dims n[] (ghosts included) into
 *                       w (NVAR*N) and pos (ndim*N); the header must match.
 *   smaug_ini_write  <- writeasciivacconfig (src/readwrite.c:1085-1300),
 *                       same header and %lG rows (6 significant digits).
 *   smaug_out_write  <- writevacconfig (src/readwrite.c:160-366): binary
 *                       snapshot, 79-byte name, it, t, ndim, neqpar, nw, n[],
 *                       eqpar[], 79-byte variable names, positions and
 *                       fields as doubles (i fastest), trailing '\n'.
 *   smaug_write_out  <- the call in the time loop (src/iosmaug.c:517): the
 *                       context's current state to <prefix>_<it>.out. */
int smaug_ini_read(const char* path, int ndim, const int* n, double* w, double* pos, int* it, double* t);
int smaug_ini_write(const char* path, const char* name, int ndim, const int* n, const double* w,
                    const double* pos, int it, double t, double gamma, double eta, const double* g);
int smaug_out_write(const char* path, const char* name, int ndim, const int* n, const double* w,
                    const double* pos, int it, double t, double gamma, double eta, const double* g);
int smaug_write_out(smaug_t* s, const char* prefix, const char* name);

/* ---- driver plug-in, streams, checkpoints --------------------------------
 *   smaug_set_driver_map  <- the amplitude of the vertical-velocity driver,
 *                            installed as addsourceterms2 (src/centdiff2.cu:
 *                            790-872).  The driver is a hybrid of two model
 *                            files: the vertical forcing m_0 += dt v_z rho_t
 *                            of models/usersource.cu.spicule1:131, with the
 *                            3D Gaussian shape (pos1..3 from wd) and the
 *                            energy term e += dt v^2 rho_t/2 of
 *                            models/usersource.cu.fluxtube1:40-79,
 *                            given per (j,k) column: n1*n2 values, column
 *                            (j,k) at k*n1+j, ghost columns included.  It
 *                            replaces A exp(-dx^2/sxy^2 [- dy^2/sxy^2]):
 *                            v_z = amap(j,k) exp(-dz^2/sz^2) sin(2 pi qt/P).
 *                            amap NULL restores the scalar form; on_device=1
 *                            when amap is a device pointer.
 *   smaug_set_stream      -- queue all further work on a caller's HIP stream
 *                            (NULL: the context's own stream).
 *   smaug_set_clock       -- step counter and nominal clock p->qt (to restart
 *                            from a checkpoint).
 *   smaug_save_pert / smaug_load_pert -- device-to-device copy of the current
 *                            evolved fields (np*N doubles, np = 2+2*ndim) to /
 *                            from a caller's device buffer; load re-fills the
 *                            ghost cells.
 * Every entry point that takes a caller's device buffer (these, the adjoint
 * set below and smaug_cfl_noalloc) first checks that the buffer lies inside
 * one device allocation large enough for the call and returns SMAUG_E_ARG
 * otherwise. */
int smaug_set_driver_map(smaug_t* s, const double* amap, int on_device);
int smaug_set_stream(smaug_t* s, void* hip_stream);
int smaug_set_clock(smaug_t* s, int it, double qt);
int smaug_save_pert(const smaug_t* s, double* dst_device);
int smaug_load_pert(smaug_t* s, const double* src_device);

/* ---- adjoint (reverse-mode derivative) of the step ------------------------
 * No counterpart in the reference: the tangent/adjoint step entry points of
 * the differentiable binding.  All buffers are device pointers.
 * smaug_set_driver_map with a map, smaug_step_taped and smaug_adjoint_step
 * return SMAUG_E_UNSUPPORTED unless the environment sets SMAUG_ADJOINT=1:
 * their first MI355X run ended in a GPU fault whose cause is not identified
 * (they are checked in a host build of the same sources instead).
 *   smaug_tape_doubles  -- size of one step's tape: the stage inputs of the
 *                          step (4*np*N doubles for rkon=1, np*N for rkon=0)
 *                          and 8 doubles of metadata (dt, qt, driver factor).
 *   smaug_step_taped    -- smaug_step(s, 1), also writing the step's tape.
 *   smaug_adjoint_step  -- given the tape of a step, the state after it
 *                          (wnext, np*N) and in lam (np*N) the cotangent of
 *                          that state, overwrites lam with the cotangent of the
 *                          state before the step and adds d(loss)/d(amap)
 *                          (per column, n1*n2 values; NULL to skip) to gmap.
 *                          The hyperdiffusion coefficients (windowed maxima)
 *                          and wave-speed maxima are held at their forward
 *                          values, and dt is the taped one.  Cotangents live
 *                          on interior cells (ghost entries of lam are zero on
 *                          return; a ghost cotangent on entry is ignored, so
 *                          fold it onto its image first). */
long long smaug_tape_doubles(const smaug_t* s);
int smaug_step_taped(smaug_t* s, double* tape_device);
int smaug_adjoint_step(smaug_t* s, const double* tape_device, const double* wnext_device, double* lam_device,
                       double* gmap_device);

/* ---- tangent-linear (forward-mode derivative) run -------------------------
 * No counterpart in the reference: the FP64 tangent-linear model of the
 * step, the forward-mode derivative of the differentiable binding and the
 * reference its reverse mode is checked against.  The stage kernel runs on
 * (value, tangent) pairs: the values follow smaug_step from the context's
 * current state, the tangents its derivative, with the hyperdiffusion
 * coefficients, the wave-speed maxima and dt held at their values along that
 * trajectory.  The run has its own buffers and clock: the context's state and
 * clock do not change.  Host arrays.  With SMAUG_PREC_MIXED the stencils
 * run on FP32 pairs (the value and tangent shadows, as the mixed step's FP32
 * shadows) over FP64 states, else everything is FP64.  divbon = 1 is
 * rejected.
 *   smaug_tl_begin -- starts a run from the current state in the direction
 *                     dw0 (np*N doubles, field-major; its ghost cells are
 *                     re-filled with the boundary maps), dbg (the static
 *                     background fields, nb*N doubles: e_b, rho_b, b_b per
 *                     cell, ghosts included; NULL = 0) and, when a driver
 *                     amplitude map is set, damap (n1*n2 values, NULL = 0).
 *   smaug_tl_driver_amplitude -- adds to the run's direction the tangent
 *                     d_amplitude of the driver's amplitude drv[0] (vertical,
 *                     torsional and spicule drivers; 0 at smaug_tl_begin)
 *                     for the steps that follow; refused without a driver or
 *                     with a driver amplitude map (the map is the amplitude).
 *   smaug_tl_step  -- advances the run by nsteps steps.
 *   smaug_tl_get   -- copies the run's state (w, np*N) and its tangent (dw,
 *                     np*N); either may be NULL.
 *   smaug_tl_end   -- frees the run (smaug_destroy does too). */
int smaug_tl_begin(smaug_t* s, const double* dw0_host, const double* dbg_host, const double* damap_host);
int smaug_tl_driver_amplitude(smaug_t* s, double d_amplitude);
int smaug_tl_step(smaug_t* s, int nsteps);
int smaug_tl_get(smaug_t* s, double* w_host, double* dw_host);
int smaug_tl_end(smaug_t* s);

/* Kernel timing hooks for the benchmark: launches `reps` RK stages of the
 * configured step on the current state and reports the mean device time of
 * each kernel family measured with HIP events on the context's stream. */
int smaug_time_kernels(smaug_t* s, int nsteps, double* ms_prim,
                       double* ms_rhs, double* ms_step);

#ifdef __cplusplus
}
#endif
#endif /* SMAUG_HIP_H_ */
