/*
 * evora_oracle.h -- CPU restatement of evora's hydro step (TEST INFRASTRUCTURE).
 *
 * This is test infrastructure, never the product: only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it.  It restates, in plain C with
 * the reference's evaluation order, the routines of
 *   source/comp.f90:26-85        q2w, pressure
 *   source/muscl.f90:26-290      primitive, compevo, comp_slope, sweep, integration
 *   source/riemann.f90:26-288    riemannsolver, hll, hllc, flux, state, statex
 *   source/boundary.f90:26-113   bound
 *   source/timestep.f90:26-57    hydrostep
 *   source/timestep.f90:147-171  updatetime (non-supercomoving branch)
 *   source/timestep.f90:72-104   thermalstep
 *   source/conduction.f90:26-164 thermalconduction, thermalflux (simplekappa),
 *                                thermalintegration; comp.f90:87-102 press2temp
 *                                (without chemistry)
 *   source/gravitation.f90:26-88,172-254 gravpotential (one rank), multigreen,
 *                                gravintegration, grav_force; fft.f90:58-80
 *                                as the defining DFT sums; comp.f90:241-252 kvalue
 *   source/snapshot.f90:159-193  evolution sums
 *
 * Array layouts are the reference's Fortran layouts (component fastest):
 *   q  (nq, 1:nx, 1:ny, 1:nz)                      nq = 6
 *   w  (nw, -1:nx+2, -1:ny+2, -1:nz+2)              nw = 5
 *   evo(nw, 0:nx+1, 0:ny+1, 0:nz+1)
 *   s  (nw, 3, 0:nx+1, 0:ny+1, 0:nz+1)
 *
 * Errors: where the reference prints and `stop`s, the oracle returns an
 * EV_ERR_* code and records the offending cell in ev_err_cell.
 *
 * Parity pins (tests/test_oracle_*.py): bitwise against the compiled
 * reference (oracle/_ref) on every solver x limiter x boundtype x dimension
 * fixture, Sod against the exact Riemann solution, and invariants
 * (uniform state, mass conservation, mirror and axis-permutation symmetry).
 */
#ifndef EVORA_ORACLE_H
#define EVORA_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum {
  EV_OK = 0,
  EV_ERR_NEG_PRESSURE = 1,     /* comp.f90:74-83 */
  EV_ERR_RIEMANN_PRESSURE = 2, /* riemann.f90:61-66, 111-116 */
  EV_ERR_HLL = 3,              /* riemann.f90:87-90 */
  EV_ERR_HLLC = 4,             /* riemann.f90:148-151 */
  EV_ERR_DT = 5,               /* timestep.f90:51-55 */
  EV_ERR_SOLVER = 6,           /* riemann.f90:37-41 */
  EV_ERR_LIMITER = 7,          /* muscl.f90:164-167 */
  EV_ERR_CONFIG = 8
};

/* Scalars the reference keeps in module `global` (global.f90:45-86). */
typedef struct {
  int nx, ny, nz;
  int boundtype;    /* 1 mirror zero-gradient, 2 periodic, 3 zero-gradient + upper-x reflecting */
  int solver;       /* 1 HLL, 2 HLLC */
  int limiter;      /* 1 minmod, 2 monotonized central, 3 van Leer */
  int secondorder;
  int dualenergy;
  int onedimension, twodimension;
  double g, g1, g1inv, gg1, cfl, de_thresh1;
  double dx, dy, dz, dxinv, dyinv, dzinv;
  double volume;
  /* thermal conduction (parameter.f90:77-78; units.f90:50,63) */
  int conduction, simplekappa;
  double c_cond, temp0, kappa0;
  /* gravity (parameter.f90:235-241,254-263) */
  int gravity, supercomoving;
  double grav, fb, h2, kfx2, kfy2, kfz2;
  int nyqx, nyqy, nyqz;
  double dx2inv, dy2inv, dz2inv;
  /* hydro = .false. (parameter.f90:44): no hydrostep, no sweep (main.f90:119,155) */
  int hydro_off;
} ev_params;

/* cell of the last error: ix, iy, iz (Fortran indices) */
extern int ev_err_cell[3];

long ev_size_q(const ev_params *p);
long ev_size_w(const ev_params *p);
long ev_size_evo(const ev_params *p);
long ev_size_s(const ev_params *p);

int ev_q2w(const ev_params *p, double *q6, double *w5);
int ev_primitive(const ev_params *p, double *w, double *q);
int ev_bound(const ev_params *p, double *w);
int ev_hydrostep(const ev_params *p, const double *w, double *dt);
int ev_comp_slope(const ev_params *p, const double *wl, const double *w, const double *wr, double *s);
int ev_compevo(const ev_params *p, const double *w, double *evo, double *s, double dt);
int ev_riemannsolver(const ev_params *p, const double *wl, const double *wr, double *f, int dir);
int ev_sweep(const ev_params *p, const double *w, double *q, const double *evo, const double *s, double dt);
void ev_flux(const ev_params *p, const double *w, double *f, int dir);
void ev_state(const ev_params *p, const double *w, double *q);
void ev_statex(const ev_params *p, const double *w, double *qx, double s, double sm, int dir);
int ev_thermalstep(const ev_params *p, double *q, const double *w, double *dt);
int ev_thermalconduction(const ev_params *p, double *q, const double *w, double dt);

/* gravity; phi is the real potential (nx, ny, nz), re/im complex fields of that shape */
void ev_dft3(const ev_params *p, double *re, double *im, int sign);
void ev_multigreen(const ev_params *p, double *re, double *im);
int ev_gravpotential(const ev_params *p, const double *q, double *phi, double cosmo);
int ev_gravintegration(const ev_params *p, double *q, const double *phi, double dt);

/* One pass of main.f90:109-175 (hydro, conduction and non-comoving gravity); t, dt, go are in/out. */
int ev_step(const ev_params *p, double *q, double *w, double *evo, double *s,
            double *t, double tend, double *dt, int *go);
/* nsteps passes (stops early when go turns false); returns steps done in *done. */
int ev_run(const ev_params *p, double *q, int nsteps, double *t, double tend,
           double *dt_last, int *done);

/* snapshot.f90:159-193 sums (before the `volume` factor): sum q1, q5, q6 */
void ev_sums(const ev_params *p, const double *q, double out[3]);

#ifdef __cplusplus
}
#endif
#endif
