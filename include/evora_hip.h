/*
 * evora_hip.h -- C-ABI of the MI355X hydro step of evora.
 *
 * A drop-in for the reference's hydro routines, called once per time step by
 * the unchanged driver loop (source/main.f90:109-155).  The Fortran binding
 * is evora_amd/fortran/evora_hip.f90 (iso_c_binding); Python binds the same
 * symbols with ctypes (evora_amd/hydro.py).  All entry points return an
 * EVH_* status; none throws.  One context owns one HIP stream and the
 * device-resident conserved state; contexts are not thread-safe.
 *
 * Host arrays use the reference layouts (component fastest, Fortran order):
 *   q   (6, nx, ny, nz)                 source/main.f90:72
 *   w   (5, -1:nx+2, -1:ny+2, -1:nz+2)  source/main.f90:73
 *   evo (5, 0:nx+1, 0:ny+1, 0:nz+1)     source/main.f90:74
 *   s   (5, 3, 0:nx+1, 0:ny+1, 0:nz+1)  source/main.f90:75
 */
#ifndef EVORA_HIP_H
#define EVORA_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

/* status codes; 1-7 are the reference's own error stops */
enum {
  EVH_OK = 0,
  EVH_ERR_NEG_PRESSURE = 1,     /* comp.f90:74-83   'negative pressure occured'      (stop)      */
  EVH_ERR_RIEMANN_PRESSURE = 2, /* riemann.f90:61-66,111-116 'unphysical pressure'     (stop)      */
  EVH_ERR_HLL = 3,              /* riemann.f90:87-90  'hll solver'                    (stop)      */
  EVH_ERR_HLLC = 4,             /* riemann.f90:148-151 'hllc solver'                  (stop)      */
  EVH_ERR_DT = 5,               /* timestep.f90:51-55 'hydrostep is negative'         (cleanstop) */
  EVH_ERR_SOLVER = 6,           /* riemann.f90:37-41  'choose proper solver'          (cleanstop) */
  EVH_ERR_LIMITER = 7,          /* muscl.f90:164-167  'wrong slope limiter'           (stop)      */
  EVH_ERR_CONFIG = 8,           /* unsupported grid or parameter combination                       */
  EVH_ERR_HIP = 9,              /* HIP runtime failure (out of memory, no device, launch failure)   */
  EVH_ERR_STATE = 10            /* call out of order (e.g. sweep before hydrostep)                  */
};

enum { EVH_FP64 = 0, EVH_FP32_MIXED = 1 };
enum { EVH_MODE_FUSED = 0, EVH_MODE_STAGED = 1, EVH_MODE_FUSED_EXACT = 2 };

/* Constants of module `global` the step reads (global.f90:44-86), as derived
 * by compparameter (parameter.f90:164-265).  The host passes them; the
 * library never recomputes them. */
typedef struct evh_config {
  int nx, ny, nz;
  int boundtype;    /* 1 zero-gradient (mirror), 2 periodic, 3 zero-gradient + reflecting upper x */
  int solver;       /* 1 HLL, 2 HLLC */
  int limiter;      /* 1 minmod, 2 monotonized central, 3 van Leer */
  int secondorder;  /* 0: first order, compevo skipped (main.f90:151) */
  int dualenergy;
  int onedimension, twodimension;
  int precision;    /* EVH_FP64 or EVH_FP32_MIXED */
  double g, g1, g1inv, gg1, cfl, de_thresh1;
  double dx, dy, dz, dxinv, dyinv, dzinv;
  double volume;
  double mach_fp64; /* FP32_MIXED: pressure from E in FP64 where Mach > mach_fp64 (e.g. 5) */
  /* thermal conduction (parameter.f90:77-78, units.f90:50,63): only
   * simplekappa, without dual energy; 1-D and 2-D grids in EVH_MODE_STAGED */
  int conduction, simplekappa;
  double c_cond, temp0, kappa0;
  /* self-gravity (parameter.f90:235-241,254-263): 3-D, 2-D (nz = 1) and
   * 1-D (ny = nz = 1) grids with nx >= 2, one rank; the potential is
   * periodic whatever the boundtype, as in the reference */
  int gravity, supercomoving;
  double grav, fb, h2, kfx2, kfy2, kfz2;
  int nyqx, nyqy, nyqz;
  double dx2inv, dy2inv, dz2inv;
  /* 1: the namelist's hydro = .false. (parameter.f90:44): evh_step then skips
   * hydrostep and sweep, as main.f90:119,155 do (conduction-only runs) */
  int hydro_off;
} evh_config;

typedef struct evh_ctx evh_ctx;

/* Create a context on the current HIP device; allocates the device state. */
int evh_create(const evh_config *cfg, evh_ctx **out);
int evh_destroy(evh_ctx *h);

/* Select the kernel structure: EVH_MODE_FUSED (one HBM pass per step with
 * refined hardware reciprocals and roots, results within a few ulp of the
 * reference's), EVH_MODE_FUSED_EXACT (the same single pass with the
 * reference's IEEE arithmetic and operation order, FP64 only: bit-identical
 * to the reference except where the entropy takes a power), or
 * EVH_MODE_STAGED (one kernel per reference routine, with w, evo and s
 * materialised in HBM; the path for 1-D/2-D).  3-D grids start in
 * EVH_MODE_FUSED. */
int evh_set_mode(evh_ctx *h, int mode);

/* Host q (reference layout) <-> device state. */
int evh_upload_q(evh_ctx *h, const double *q);
int evh_download_q(evh_ctx *h, double *q);

/* ---- the reference's step-level routines, same order of calls ---- */
int evh_primitive(evh_ctx *h);                   /* primitive(w,q)     muscl.f90:26-43        */
int evh_bound(evh_ctx *h);                       /* bound(w)           boundary.f90:26-113    */
int evh_hydrostep(evh_ctx *h, double *dt);       /* hydrostep(w)       timestep.f90:26-57:
                                                    dt = min(cfl*v, dt), in/out              */
int evh_compevo(evh_ctx *h, double dt);          /* compevo(w,evo,s)   muscl.f90:45-113       */
int evh_sweep(evh_ctx *h, double dt);            /* sweep(w,q,evo,s)   muscl.f90:171-274      */
/* thermal conduction (conduction on in the config): */
int evh_thermalstep(evh_ctx *h, double *dt);     /* thermalstep(q,w)   timestep.f90:72-104:
                                                    dt = min(c_cond dx^2 min(rho/kappa)/temp0, dt) */
int evh_thermalconduction(evh_ctx *h, double dt); /* thermalconduction(q,w) conduction.f90:26-123,
                                                    with the primitive and bound main.f90:171-172
                                                    calls before it; simplekappa only        */

/* self-gravity (gravity on in the config): */
int evh_gravpotential(evh_ctx *h, double cosmo);  /* gravpotential(q,phi) gravitation.f90:26-52: the potential of
                                                    the current q, kept on the device; cosmo is the expansion
                                                    factor of the supercomoving source (ignored otherwise) */
int evh_gravintegration(evh_ctx *h, double dt);   /* gravintegration(q,phi,..) gravitation.f90:172-254 */
int evh_download_phi(evh_ctx *h, double *phi);    /* the potential phi(nx,ny,nz) of the last evh_gravpotential */

/* One pass of main.f90:109-175 on the device (with thermalstep and the
 * conduction of main.f90:121,170-175 when conduction is on, and the
 * gravity of main.f90:157-164 when gravity is on, with the expansion factor
 * of the last evh_gravpotential call, 1 at first), with updatetime
 * (timestep.f90:163-169): t, dt, go in/out.  evh_run repeats it nsteps
 * times (or until go turns 0), syncing only for dt. */
int evh_step(evh_ctx *h, double *t, double tend, double *dt, int *go);
int evh_run(evh_ctx *h, int nsteps, double *t, double tend, double *dt, int *done);

/* ---- one rank's z-slab of a grid decomposed along z (parallel.f90:26-66;
 * the context's nz is the slab's, every other parameter the whole grid's) ----
 * evh_slab: with lo (hi) nonzero the two ghost planes below (above) the slab
 * are read from a halo buffer of the context instead of the z rules of bound
 * (boundary.f90:90-110): the planes communicate (parallel.f90:317-381)
 * receives from the neighbouring slab, as q.  Each buffer holds the 6
 * conserved components of the two planes, [6][2][ny][nx], in the storage
 * type (double; float with EVH_FP32_MIXED); *lo_halo / *hi_halo receive
 * their device addresses (NULL for a side that keeps bound's rules; either
 * pointer may be NULL).  The caller fills them before each evh_sweep, while
 * the context is idle (its sweep kernel reads them: after evh_synchronize or
 * evh_hydrostep, which waits for the context's stream), and
 * drives the step through evh_primitive, evh_hydrostep (dt reduced to the
 * minimum over the slabs, reducestep parallel.f90:155-165), evh_compevo and
 * evh_sweep; evh_step and evh_run refuse a slab.  Fused modes on grids of
 * two planes or more; not with gravity (its Poisson solve spans the grid). */
int evh_slab(evh_ctx *h, int lo, int hi, void **lo_halo, void **hi_halo);
/* The slab's two lowest (lo_dst) and two highest (hi_dst) planes of the
 * current q, [6][2][ny][nx] in the storage type, into device memory (a
 * neighbour's halo buffer, or a buffer a send reads): the planes
 * communicate sends (parallel.f90:334,348).  Either pointer may be NULL;
 * complete when it returns. */
int evh_slab_faces(evh_ctx *h, void *lo_dst, void *hi_dst);
/* With conduction on, the exchange main.f90:170-175 makes before
 * thermalconduction: evh_slab_temp gives the device addresses of the halo
 * buffers for T of the plane next to the slab on each side with a neighbour
 * ([ny][nx], storage type; NULL for a side without), which the caller fills
 * before each evh_thermalconduction while the context is idle;
 * evh_slab_temp_faces copies T of the slab's lowest and highest planes of the
 * current state into device memory (either pointer may be NULL; complete
 * when it returns). */
int evh_slab_temp(evh_ctx *h, void **lo_halo, void **hi_halo);
int evh_slab_temp_faces(evh_ctx *h, void *lo_dst, void *hi_dst);

/* snapshot.f90:169-172 sums of q1, q5, q6 (without the volume factor), FP64
 * compensated on the device. */
int evh_sums(evh_ctx *h, double out[3]);

/* Staged mode only: the last primitive+bound w, and the last compevo evo, s. */
int evh_download_w(evh_ctx *h, double *w);
int evh_download_evo_s(evh_ctx *h, double *evo, double *s);

/* Last error: status code, cell (Fortran indices) and a message. */
int evh_last_error(evh_ctx *h, int *code, int cell[3], char *msg, int msglen);

/* Timing support: events recorded on the context's stream. */
int evh_event_record(evh_ctx *h, int slot);                 /* slot 0..7 */
int evh_event_elapsed(evh_ctx *h, int a, int b, float *ms); /* syncs on slot b */
int evh_synchronize(evh_ctx *h);

/* Launch shape of the fused kernel (no counterpart in the reference; set by
 * the autotuner): tile height 8 or 12 columns, occupancy target (0: the
 * lower, 1: the higher of the two register budgets compiled per precision),
 * z planes per block (rows on 2-D grids; 0, the default: 64 planes or 32
 * rows, halved down to 4 while the grid gives fewer than 1024 blocks).
 * Invalid values are rejected.  A z-slab with halo buffers (evh_slab) runs
 * tile height 12 at the higher budget; its z chunk is the tuning's.  The
 * results do not depend on the launch shape. */
int evh_set_tuning(evh_ctx *h, int tile_ty, int wpe_hi, int zchunk);
int evh_get_tuning(evh_ctx *h, int *tile_ty, int *wpe_hi, int *zchunk);

/* Kernel timing (no counterpart in the reference; for benchmarks): when on,
 * every fused or staged sweep launch, conduction-kernel launch and gravity
 * stage is bracketed by events on the context's stream and its duration is
 * added to a running total at the next sync. */
int evh_set_timing(evh_ctx *h, int on);
int evh_kernel_stats(evh_ctx *h, double *total_ms, long long *launches);
/* the same for the conduction kernel of evh_thermalconduction */
int evh_conduction_stats(evh_ctx *h, double *total_ms, long long *launches);
/* the same for any stage: the sweep, the conduction kernel, the potential
 * (mean, source, both FFTs and the Green's function) or the force update */
enum { EVH_STAGE_SWEEP = 0, EVH_STAGE_CONDUCTION = 1, EVH_STAGE_GRAVPOTENTIAL = 2, EVH_STAGE_GRAVINTEGRATION = 3 };
int evh_stage_stats(evh_ctx *h, int stage, double *total_ms, long long *launches);

/* Bytes of device memory the context holds. */
long long evh_device_bytes(evh_ctx *h);

/* Diagnostics (no counterpart in the reference): out[i] = f(x[i], y[i]) for
 * the device arithmetic of the fast forms, on host arrays of n values, so
 * that it can be held against the host's IEEE operations and libm:
 *   EVH_MATH_POW      x**y for x > 0 (exp(y ln x), the general power)
 *   EVH_MATH_POW_G1   x**g1 with g1 = y (the cube-root form when g1 is 2/3)
 *   EVH_MATH_RPOW_G1  x**(-g1) with g1 = y
 *   EVH_MATH_DIV      x / y from the refined reciprocal of y (IEEE quotient)
 *   EVH_MATH_RCP      1 / x, correctly rounded
 *   EVH_MATH_SQRT     sqrt(x) for x > 0
 *   EVH_MATH_POWF     x**y in single precision (hardware exp2 and log2)
 * Allocates and frees device memory; synchronous. */
enum { EVH_MATH_POW = 0, EVH_MATH_POW_G1 = 1, EVH_MATH_RPOW_G1 = 2, EVH_MATH_DIV = 3, EVH_MATH_RCP = 4,
       EVH_MATH_SQRT = 5, EVH_MATH_POWF = 6 };
int evh_math_eval(int fn, long long n, const double *x, const double *y, double *out);

/* ---- batched form (no counterpart in the reference, whose equivalent is
 * one run per box): many independent 3-D boxes of uneven sizes, each its own
 * evora run (own t, dt from its own CFL limit, own error stop), advanced
 * together.  One step is a device-side dt kernel (hydrostep + updatetime per
 * box, timestep.f90:26-57,163-169) and one persistent sweep kernel over a
 * (box, tile) work list; with use_graph the pair is replayed from a hipGraph
 * and the host does not synchronise between steps.  All boxes share
 * solver, limiter, precision, secondorder and dualenergy; grids, spacing,
 * boundtype and cfl may differ.  The batched form runs the hydro step only:
 * boxes with conduction or gravity on are rejected (EVH_ERR_CONFIG). */
typedef struct evh_batch evh_batch;
int evh_batch_create(const evh_config *cfgs, int nbox, evh_batch **out);
int evh_batch_destroy(evh_batch *b);
int evh_batch_upload_q(evh_batch *b, int box, const double *q);
int evh_batch_download_q(evh_batch *b, int box, double *q);
/* per-box start time and tend (defaults 0 and 1e99); either may be NULL */
int evh_batch_set_time(evh_batch *b, const double *t, const double *tend);
/* advance every running box by nsteps steps; ms = device time of the steps */
int evh_batch_run(evh_batch *b, int nsteps, int use_graph, float *ms);
/* per-box t, last dt, steps done and go flag; any pointer may be NULL */
int evh_batch_times(evh_batch *b, double *t, double *dt, int *steps, int *go);
/* launch shape: tile height 8|12, occupancy target 0|1, z chunk, blocks per CU */
int evh_batch_set_tuning(evh_batch *b, int tile_ty, int wpe_hi, int zchunk, int blocks_per_cu);
int evh_batch_get_tuning(evh_batch *b, int *tile_ty, int *wpe_hi, int *zchunk, int *blocks_per_cu);
/* time every launch shape for trial_steps steps on a snapshot of the state,
 * keep the fastest, restore the state; the timings are written to log */
int evh_batch_autotune(evh_batch *b, int trial_steps, char *log, int loglen);
int evh_batch_last_error(evh_batch *b, int *code, int *box, int cell[3], char *msg, int msglen);

/* Allocation-free batched run.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  The same steps as
 * evh_batch_create + evh_batch_run(nsteps, eager) on the caller's own device
 * buffers q0 and q1, each holding the boxes' SoA states (box i: 6 arrays of
 * nx*ny*nz values, [6][nz][ny][nx], in the storage precision, at element
 * offset 6 * the cells of boxes 0..i-1) in the buffer side[i] names.  Every
 * running box (go[i] != 0) first takes its CFL limit from its current state,
 * as after an upload.  side, t, steps and go are updated in place (tend is
 * read); the return value is EVH_OK or the first stopped box's status, whose
 * message goes to err (may be NULL).  Descriptors, work queues and per-box
 * status live in a fixed 48 KiB device array of the library, one per GPU:
 * at most 64 boxes (more is rejected with EVH_ERR_CONFIG); no memory is
 * allocated or freed.  Calls that overlap on one GPU race on that array.
 * Synchronous (the default stream). */
int evh_batch_run_noalloc(const evh_config *cfgs, int nbox, void *q0, void *q1, int *side, double *t,
                          const double *tend, int *steps, int *go, int nsteps, char *err, int errlen);

#ifdef __cplusplus
}
#endif
#endif
