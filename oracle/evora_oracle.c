/*
 * evora_oracle.c -- CPU restatement of evora's hydro step (TEST INFRASTRUCTURE).
 *
 * The reference the HIP library is tested against, never the product.  Every function keeps
 * the reference's operand order so that, compiled without FP contraction,
 * it reproduces the compiled Fortran bit for bit (pinned against the fixtures
 * oracle/_ref/evref writes, by tests/test_oracle_golden.py).  OpenMP only splits independent cells or
 * pencils; per-cell arithmetic does not depend on the thread count.
 */
#include "evora_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define NQ 6
#define NW 5

int ev_err_cell[3];

/* ---- index helpers (Fortran bounds from parallel.f90:33-64, one rank) ---- */
/* w(nw, -1:nx+2, -1:ny+2, -1:nz+2) */
static inline long widx(const ev_params *p, int i, int j, int k) {
  return (long)NW * ((long)(i + 2 - 1) + (long)(p->nx + 4) * ((long)(j + 1) + (long)(p->ny + 4) * (long)(k + 1)));
}
/* evo(nw, 0:nx+1, ...) and s(nw, 3, 0:nx+1, ...) share the cell index */
static inline long eidx(const ev_params *p, int i, int j, int k) {
  return (long)i + (long)(p->nx + 2) * ((long)j + (long)(p->ny + 2) * (long)k);
}
/* q(nq, 1:nx, 1:ny, 1:nz) */
static inline long qidx(const ev_params *p, int i, int j, int k) {
  return (long)NQ * ((long)(i - 1) + (long)p->nx * ((long)(j - 1) + (long)p->ny * (long)(k - 1)));
}

long ev_size_q(const ev_params *p) { return (long)NQ * p->nx * p->ny * p->nz; }
long ev_size_w(const ev_params *p) { return (long)NW * (p->nx + 4) * (p->ny + 4) * (p->nz + 4); }
long ev_size_evo(const ev_params *p) { return (long)NW * (p->nx + 2) * (p->ny + 2) * (p->nz + 2); }
long ev_size_s(const ev_params *p) { return 3L * NW * (p->nx + 2) * (p->ny + 2) * (p->nz + 2); }

static void set_err_cell(int i, int j, int k) {
  ev_err_cell[0] = i;
  ev_err_cell[1] = j;
  ev_err_cell[2] = k;
}

static inline double fmin2(double a, double b) { return a < b ? a : b; }
static inline double fmax2(double a, double b) { return a > b ? a : b; }

/* ---- comp.f90:45-85  pressure(q, press): may rewrite q(5) or q(6) ---- */
static int pressure(const ev_params *p, double *q, double *press) {
  double m2 = (q[1] * q[1]) + (q[2] * q[2]) + (q[3] * q[3]);
  double kin = 0.5 * m2 / q[0];
  double ratio = kin / q[4];
  double pr;
  if (p->dualenergy) {
    if (ratio > p->de_thresh1) {
      pr = q[5] * pow(q[0], p->g1);     /* comp.f90:63 */
      q[4] = kin + pr * p->g1inv;       /* comp.f90:64 */
    } else {
      pr = p->g1 * (q[4] - kin);        /* comp.f90:67 */
      q[5] = pr / pow(q[0], p->g1);     /* comp.f90:68 */
    }
  } else {
    pr = p->g1 * (q[4] - kin);          /* comp.f90:71 */
  }
  *press = pr;
  if (pr <= 0.0) return EV_ERR_NEG_PRESSURE; /* comp.f90:74-83 */
  return EV_OK;
}

/* ---- comp.f90:26-43  q2w ---- */
int ev_q2w(const ev_params *p, double *q, double *w) {
  double densinv = 1.0 / q[0];
  w[0] = q[0];
  w[1] = q[1] * densinv;
  w[2] = q[2] * densinv;
  w[3] = q[3] * densinv;
  return pressure(p, q, &w[4]);
}

/* ---- muscl.f90:26-43  primitive ---- */
int ev_primitive(const ev_params *p, double *w, double *q) {
  long first_bad = -1;
  int iz;
#pragma omp parallel for schedule(static)
  for (iz = 1; iz <= p->nz; iz++) {
    for (int iy = 1; iy <= p->ny; iy++) {
      for (int ix = 1; ix <= p->nx; ix++) {
        int e = ev_q2w(p, &q[qidx(p, ix, iy, iz)], &w[widx(p, ix, iy, iz)]);
        if (e) {
          long lin = (long)(ix - 1) + (long)p->nx * ((long)(iy - 1) + (long)p->ny * (iz - 1));
#pragma omp critical(ev_err)
          if (first_bad < 0 || lin < first_bad) first_bad = lin;
        }
      }
    }
  }
  if (first_bad >= 0) {
    int ix = (int)(first_bad % p->nx) + 1;
    int iy = (int)((first_bad / p->nx) % p->ny) + 1;
    int izz = (int)(first_bad / ((long)p->nx * p->ny)) + 1;
    set_err_cell(ix, iy, izz);
    return EV_ERR_NEG_PRESSURE;
  }
  return EV_OK;
}

/* ---- boundary.f90:26-113  bound ---- */
static inline void wcopy(double *dst, const double *src) {
  for (int c = 0; c < NW; c++) dst[c] = src[c];
}
static inline void wcopy_reflect(double *dst, const double *src) {
  dst[0] = src[0];
  dst[1] = -src[1];
  dst[2] = -src[2];
  dst[3] = -src[3];
  dst[4] = src[4];
}

int ev_bound(const ev_params *p, double *w) {
  const int nx = p->nx, ny = p->ny, nz = p->nz, bt = p->boundtype;
  /* x boundary over the full y,z extents (boundary.f90:33-66) */
  if (nx != 1) {
    for (int k = -1; k <= nz + 2; k++)
      for (int j = -1; j <= ny + 2; j++) {
        if (bt == 1 || bt == 3) {
          wcopy(&w[widx(p, -1, j, k)], &w[widx(p, 2, j, k)]);
          wcopy(&w[widx(p, 0, j, k)], &w[widx(p, 1, j, k)]);
        } else if (bt == 2) {
          wcopy(&w[widx(p, -1, j, k)], &w[widx(p, nx - 1, j, k)]);
          wcopy(&w[widx(p, 0, j, k)], &w[widx(p, nx, j, k)]);
        }
        if (bt == 1) {
          wcopy(&w[widx(p, nx + 2, j, k)], &w[widx(p, nx - 1, j, k)]);
          wcopy(&w[widx(p, nx + 1, j, k)], &w[widx(p, nx, j, k)]);
        } else if (bt == 2) {
          wcopy(&w[widx(p, nx + 2, j, k)], &w[widx(p, 2, j, k)]);
          wcopy(&w[widx(p, nx + 1, j, k)], &w[widx(p, 1, j, k)]);
        } else if (bt == 3) {
          wcopy_reflect(&w[widx(p, nx + 1, j, k)], &w[widx(p, nx, j, k)]);
          wcopy_reflect(&w[widx(p, nx + 2, j, k)], &w[widx(p, nx - 1, j, k)]);
        }
      }
  }
  /* y boundary over the full x,z extents (boundary.f90:69-87) */
  if (ny != 1) {
    for (int k = -1; k <= nz + 2; k++)
      for (int i = -1; i <= nx + 2; i++) {
        if (bt == 1 || bt == 3) {
          wcopy(&w[widx(p, i, -1, k)], &w[widx(p, i, 2, k)]);
          wcopy(&w[widx(p, i, 0, k)], &w[widx(p, i, 1, k)]);
          wcopy(&w[widx(p, i, ny + 2, k)], &w[widx(p, i, ny - 1, k)]);
          wcopy(&w[widx(p, i, ny + 1, k)], &w[widx(p, i, ny, k)]);
        } else if (bt == 2) {
          wcopy(&w[widx(p, i, -1, k)], &w[widx(p, i, ny - 1, k)]);
          wcopy(&w[widx(p, i, 0, k)], &w[widx(p, i, ny, k)]);
          wcopy(&w[widx(p, i, ny + 2, k)], &w[widx(p, i, 2, k)]);
          wcopy(&w[widx(p, i, ny + 1, k)], &w[widx(p, i, 1, k)]);
        }
      }
  }
  /* z boundary over the full x,y extents (boundary.f90:90-110), one rank */
  if (nz != 1) {
    for (int j = -1; j <= ny + 2; j++)
      for (int i = -1; i <= nx + 2; i++) {
        if (bt == 1 || bt == 3) {
          wcopy(&w[widx(p, i, j, -1)], &w[widx(p, i, j, 2)]);
          wcopy(&w[widx(p, i, j, 0)], &w[widx(p, i, j, 1)]);
          wcopy(&w[widx(p, i, j, nz + 2)], &w[widx(p, i, j, nz - 1)]);
          wcopy(&w[widx(p, i, j, nz + 1)], &w[widx(p, i, j, nz)]);
        } else if (bt == 2) {
          wcopy(&w[widx(p, i, j, -1)], &w[widx(p, i, j, nz - 1)]);
          wcopy(&w[widx(p, i, j, 0)], &w[widx(p, i, j, nz)]);
          wcopy(&w[widx(p, i, j, nz + 2)], &w[widx(p, i, j, 2)]);
          wcopy(&w[widx(p, i, j, nz + 1)], &w[widx(p, i, j, 1)]);
        }
      }
  }
  return EV_OK;
}

/* ---- timestep.f90:26-57  hydrostep: dt = min(cfl * v, dt) ---- */
int ev_hydrostep(const ev_params *p, const double *w, double *dt) {
  double v = 1e6;
  int iz;
#pragma omp parallel for reduction(min : v) schedule(static)
  for (iz = 1; iz <= p->nz; iz++)
    for (int iy = 1; iy <= p->ny; iy++)
      for (int ix = 1; ix <= p->nx; ix++) {
        const double *c = &w[widx(p, ix, iy, iz)];
        double a = sqrt(p->g * c[4] / c[0]);
        double vx = p->dx / (fabs(c[1]) + a);
        double vy = p->dy / (fabs(c[2]) + a);
        double vz = p->dz / (fabs(c[3]) + a);
        v = fmin2(fmin2(fmin2(v, vx), vy), vz);
      }
  *dt = fmin2(p->cfl * v, *dt);
  if (*dt <= 0.0) return EV_ERR_DT;
  return EV_OK;
}

/* ---- muscl.f90:115-169  comp_slope ---- */
int ev_comp_slope(const ev_params *p, const double *wl, const double *w, const double *wr, double *s) {
  double dwl[NW], dwr[NW];
  for (int i = 0; i < NW; i++) {
    dwl[i] = w[i] - wl[i];
    dwr[i] = wr[i] - w[i];
  }
  if (p->limiter == 1) { /* minmod */
    for (int i = 0; i < NW; i++) {
      int lpos = dwl[i] >= 0.0, rpos = dwr[i] >= 0.0;
      if (lpos != rpos) s[i] = 0.0;
      else if (rpos) s[i] = fmin2(dwl[i], dwr[i]);
      else s[i] = fmax2(dwl[i], dwr[i]);
    }
  } else if (p->limiter == 2) { /* monotonized central */
    for (int i = 0; i < NW; i++) {
      double dwc = (wr[i] - wl[i]) * 0.5;
      int lpos = dwl[i] >= 0.0, rpos = dwr[i] >= 0.0;
      if (lpos != rpos) s[i] = 0.0;
      else if (rpos) s[i] = fmin2(fmin2(2.0 * dwl[i], dwc), 2.0 * dwr[i]);
      else s[i] = fmax2(fmax2(2.0 * dwl[i], dwc), 2.0 * dwr[i]);
    }
  } else if (p->limiter == 3) { /* van Leer */
    for (int i = 0; i < NW; i++) {
      double pro = dwl[i] * dwr[i];
      if (pro <= 0.0) s[i] = 0.0;
      else s[i] = 2.0 * pro / (dwr[i] + dwl[i]);
    }
  } else {
    return EV_ERR_LIMITER;
  }
  return EV_OK;
}

/* predictor contribution of direction d (muscl.f90:100-104), a[] += nothing */
static inline void predictor_dir(const ev_params *p, const double *c, const double *sd, int d, double *a) {
  for (int i = 0; i < NW; i++) a[i] = c[d + 1] * sd[i];
  a[4] = a[4] + p->g * c[4] * sd[d + 1];
  a[d + 1] = a[d + 1] + sd[4] / c[0];
}

/* ---- muscl.f90:45-113  compevo ---- */
int ev_compevo(const ev_params *p, const double *w, double *evo, double *s, double dt) {
  if (p->limiter < 1 || p->limiter > 3) return EV_ERR_LIMITER;
  const double dt2 = -0.5 * dt;
  const long sx = NW, sy = (long)NW * (p->nx + 4), sz = (long)NW * (p->nx + 4) * (p->ny + 4);
  if (p->onedimension) {
    for (int ix = 0; ix <= p->nx + 1; ix++) {
      const double *c = &w[widx(p, ix, 1, 1)];
      double *sc = &s[(long)3 * NW * eidx(p, ix, 1, 1)];
      double a[NW];
      ev_comp_slope(p, c - sx, c, c + sx, sc);
      predictor_dir(p, c, sc, 0, a);
      double *e = &evo[(long)NW * eidx(p, ix, 1, 1)];
      for (int i = 0; i < NW; i++) e[i] = dt2 * a[i] * p->dxinv; /* muscl.f90:68 */
    }
  } else if (p->twodimension) {
    for (int iy = 0; iy <= p->ny + 1; iy++)
      for (int ix = 0; ix <= p->nx + 1; ix++) {
        const double *c = &w[widx(p, ix, iy, 1)];
        double *sc = &s[(long)3 * NW * eidx(p, ix, iy, 1)];
        double a1[NW], a2[NW];
        ev_comp_slope(p, c - sx, c, c + sx, sc);
        ev_comp_slope(p, c - sy, c, c + sy, sc + NW);
        predictor_dir(p, c, sc, 0, a1);
        predictor_dir(p, c, sc + NW, 1, a2);
        double *e = &evo[(long)NW * eidx(p, ix, iy, 1)];
        for (int i = 0; i < NW; i++) e[i] = dt2 * (a1[i] * p->dxinv + a2[i] * p->dyinv); /* muscl.f90:86 */
      }
  } else {
    int iz;
#pragma omp parallel for schedule(static)
    for (iz = 0; iz <= p->nz + 1; iz++)
      for (int iy = 0; iy <= p->ny + 1; iy++)
        for (int ix = 0; ix <= p->nx + 1; ix++) {
          const double *c = &w[widx(p, ix, iy, iz)];
          double *sc = &s[(long)3 * NW * eidx(p, ix, iy, iz)];
          double a1[NW], a2[NW], a3[NW];
          ev_comp_slope(p, c - sx, c, c + sx, sc);
          ev_comp_slope(p, c - sy, c, c + sy, sc + NW);
          ev_comp_slope(p, c - sz, c, c + sz, sc + 2 * NW);
          predictor_dir(p, c, sc, 0, a1);
          predictor_dir(p, c, sc + NW, 1, a2);
          predictor_dir(p, c, sc + 2 * NW, 2, a3);
          double *e = &evo[(long)NW * eidx(p, ix, iy, iz)];
          for (int i = 0; i < NW; i++)
            e[i] = dt2 * (a1[i] * p->dxinv + a2[i] * p->dyinv + a3[i] * p->dzinv); /* muscl.f90:106-107 */
        }
  }
  return EV_OK;
}

/* ---- riemann.f90:157-205  flux ---- */
void ev_flux(const ev_params *p, const double *w, double *f, int dir) {
  double u2 = w[1] * w[1] + w[2] * w[2] + w[3] * w[3];
  double ud = w[dir];
  f[0] = w[0] * ud;
  for (int i = 1; i <= 3; i++) {
    if (i == dir) f[i] = w[0] * (ud * ud) + w[4];
    else f[i] = w[0] * ud * w[i];
  }
  f[4] = ud * (0.5 * w[0] * u2 + p->gg1 * w[4]);
  f[5] = ud * w[4] / pow(w[0], p->g1);
}

/* ---- riemann.f90:207-239  state ---- */
void ev_state(const ev_params *p, const double *w, double *q) {
  q[0] = w[0];
  q[1] = w[0] * w[1];
  q[2] = w[0] * w[2];
  q[3] = w[0] * w[3];
  double u2 = (w[1] * w[1]) + (w[2] * w[2]) + (w[3] * w[3]);
  q[4] = 0.5 * w[0] * u2 + w[4] / p->g1;
  q[5] = w[4] / pow(w[0], p->g1);
}

/* ---- riemann.f90:241-288  statex (HLLC star state) ---- */
void ev_statex(const ev_params *p, const double *w, double *qx, double s, double sm, int dir) {
  double ud = w[dir];
  qx[0] = w[0] * (s - ud) / (s - sm);
  for (int i = 1; i <= 3; i++) qx[i] = (i == dir ? sm : w[i]) * qx[0];
  double energy = 0.5 * w[0] * ((w[1] * w[1]) + (w[2] * w[2]) + (w[3] * w[3])) + w[4] / p->g1;
  qx[4] = qx[0] * (energy / w[0] + (sm - ud) * (sm + (w[4] / (w[0] * (s - ud)))));
  qx[5] = qx[0] * w[4] / pow(w[0], p->g);
}

/* ---- riemann.f90:47-92  hll ---- */
static int hll(const ev_params *p, const double *wl, const double *wr, double *f, int dir) {
  if (wl[4] <= 0 || wr[4] <= 0) return EV_ERR_RIEMANN_PRESSURE;
  double al = sqrt(p->g * wl[4] / wl[0]);
  double ar = sqrt(p->g * wr[4] / wr[0]);
  double sl = fmin2(wl[dir] - al, wr[dir] - ar);
  double sr = fmax2(wl[dir] + al, wr[dir] + ar);
  if (0 < sl) {
    ev_flux(p, wl, f, dir);
  } else if (0 > sr) {
    ev_flux(p, wr, f, dir);
  } else if (sl <= 0 && 0 <= sr) {
    double fl[NQ], fr[NQ], ql[NQ], qr[NQ];
    ev_flux(p, wl, fl, dir);
    ev_flux(p, wr, fr, dir);
    ev_state(p, wl, ql);
    ev_state(p, wr, qr);
    for (int i = 0; i < NQ; i++) f[i] = (sr * fl[i] - sl * fr[i] + sl * sr * (qr[i] - ql[i])) / (sr - sl);
  } else {
    return EV_ERR_HLL;
  }
  return EV_OK;
}

/* ---- riemann.f90:96-153  hllc ---- */
static int hllc(const ev_params *p, const double *wl, const double *wr, double *f, int dir) {
  if (wl[4] <= 0 || wr[4] <= 0) return EV_ERR_RIEMANN_PRESSURE;
  double al = sqrt(p->g * wl[4] / wl[0]);
  double ar = sqrt(p->g * wr[4] / wr[0]);
  double sl = fmin2(wl[dir] - al, wr[dir] - ar);
  double sr = fmax2(wl[dir] + al, wr[dir] + ar);
  double facl = wl[0] * (sl - wl[dir]);
  double facr = wr[0] * (sr - wr[dir]);
  double sm = (wr[4] - wl[4] + facl * wl[dir] - facr * wr[dir]) / (facl - facr);
  if (0 < sl) {
    ev_flux(p, wl, f, dir);
  } else if (0 > sr) {
    ev_flux(p, wr, f, dir);
  } else if (sl < 0 && 0 < sm) {
    double fl[NQ], ql[NQ], qx[NQ];
    ev_flux(p, wl, fl, dir);
    ev_state(p, wl, ql);
    ev_statex(p, wl, qx, sl, sm, dir);
    for (int i = 0; i < NQ; i++) f[i] = fl[i] + sl * (qx[i] - ql[i]);
  } else if (sm <= 0 && 0 < sr) {
    double fr[NQ], qr[NQ], qx[NQ];
    ev_flux(p, wr, fr, dir);
    ev_state(p, wr, qr);
    ev_statex(p, wr, qx, sr, sm, dir);
    for (int i = 0; i < NQ; i++) f[i] = fr[i] + sr * (qx[i] - qr[i]);
  } else {
    return EV_ERR_HLLC;
  }
  return EV_OK;
}

/* ---- riemann.f90:26-43  riemannsolver; dir 1..3 ---- */
int ev_riemannsolver(const ev_params *p, const double *wl, const double *wr, double *f, int dir) {
  if (p->solver == 1) return hll(p, wl, wr, f, dir);
  if (p->solver == 2) return hllc(p, wl, wr, f, dir);
  return EV_ERR_SOLVER;
}

/* face states of muscl.f90:207-210 */
static inline void face_states(const ev_params *p, const double *w0, const double *w1, const double *s0,
                               const double *s1, const double *e0, const double *e1, double *wl, double *wr) {
  for (int c = 0; c < NW; c++) {
    wl[c] = w0[c];
    wr[c] = w1[c];
  }
  if (p->secondorder) {
    for (int c = 0; c < NW; c++) {
      wl[c] = wl[c] + 0.5 * s0[c] + e0[c];
      wr[c] = wr[c] - 0.5 * s1[c] + e1[c];
    }
  }
}

/* one pencil of sweep (muscl.f90:191-215 and the y/z copies): cells
 * i = 1..n along `dir`, with w strides ws, evo stride es, s stride es*3*NW. */
static int pencil(const ev_params *p, const double *w, double *q, const double *evo, const double *s,
                  long w0, long wst, long e0, long est, long q0, long qst, int n, int dir, double delta,
                  long *bad_face) {
  double wl[NW], wr[NW], fl[NQ], fr[NQ];
  const int sd = (dir - 1) * NW;
  /* first face: cells 0 and 1 */
  face_states(p, &w[w0], &w[w0 + wst], &s[3 * NW * e0 + sd], &s[3 * NW * (e0 + est) + sd],
              &evo[NW * e0], &evo[NW * (e0 + est)], wl, wr);
  int err = ev_riemannsolver(p, wl, wr, fr, dir);
  if (err) {
    *bad_face = 0;
    return err;
  }
  for (int i = 1; i <= n; i++) {
    for (int c = 0; c < NQ; c++) fl[c] = fr[c];
    long wi = w0 + i * wst, ei = e0 + i * est;
    face_states(p, &w[wi], &w[wi + wst], &s[3 * NW * ei + sd], &s[3 * NW * (ei + est) + sd],
                &evo[NW * ei], &evo[NW * (ei + est)], wl, wr);
    err = ev_riemannsolver(p, wl, wr, fr, dir);
    if (err) {
      *bad_face = i;
      return err;
    }
    double *qc = &q[q0 + (i - 1) * qst];
    for (int c = 0; c < NQ; c++) qc[c] = qc[c] + delta * (fl[c] - fr[c]); /* muscl.f90:287 */
  }
  return EV_OK;
}

/* ---- muscl.f90:171-274  sweep (x, then y, then z) ---- */
int ev_sweep(const ev_params *p, const double *w, double *q, const double *evo, const double *s, double dt) {
  if (p->solver != 1 && p->solver != 2) return EV_ERR_SOLVER;
  const int nx = p->nx, ny = p->ny, nz = p->nz;
  const long wsx = NW, wsy = (long)NW * (nx + 4), wsz = (long)NW * (nx + 4) * (ny + 4);
  const long esx = 1, esy = nx + 2, esz = (long)(nx + 2) * (ny + 2);
  const long qsx = NQ, qsy = (long)NQ * nx, qsz = (long)NQ * nx * ny;
  int err = EV_OK;
  if (nx != 1) {
    double delta = dt * p->dxinv;
    int iz;
#pragma omp parallel for schedule(static)
    for (iz = 1; iz <= nz; iz++)
      for (int iy = 1; iy <= ny; iy++) {
        long bf;
        int e = pencil(p, w, q, evo, s, widx(p, 0, iy, iz), wsx, eidx(p, 0, iy, iz), esx, qidx(p, 1, iy, iz), qsx,
                       nx, 1, delta, &bf);
        if (e) {
#pragma omp critical(ev_err)
          {
            err = e;
            set_err_cell((int)bf, iy, iz);
          }
        }
      }
    if (err) return err;
  }
  if (ny != 1) {
    double delta = dt * p->dyinv;
    int iz;
#pragma omp parallel for schedule(static)
    for (iz = 1; iz <= nz; iz++)
      for (int ix = 1; ix <= nx; ix++) {
        long bf;
        int e = pencil(p, w, q, evo, s, widx(p, ix, 0, iz), wsy, eidx(p, ix, 0, iz), esy, qidx(p, ix, 1, iz), qsy,
                       ny, 2, delta, &bf);
        if (e) {
#pragma omp critical(ev_err)
          {
            err = e;
            set_err_cell(ix, (int)bf, iz);
          }
        }
      }
    if (err) return err;
  }
  if (nz != 1) {
    double delta = dt * p->dzinv;
    int iy;
#pragma omp parallel for schedule(static)
    for (iy = 1; iy <= ny; iy++)
      for (int ix = 1; ix <= nx; ix++) {
        long bf;
        int e = pencil(p, w, q, evo, s, widx(p, ix, iy, 0), wsz, eidx(p, ix, iy, 0), esz, qidx(p, ix, iy, 1), qsz,
                       nz, 3, delta, &bf);
        if (e) {
#pragma omp critical(ev_err)
          {
            err = e;
            set_err_cell(ix, iy, (int)bf);
          }
        }
      }
  }
  return err;
}

/* ---- main.f90:109-155 for pure hydro, with updatetime (timestep.f90:163-169) ---- */
/* ---- timestep.f90:72-104  thermalstep: dt = min(c_cond dx^2 min(rho/kappa) / temp0, dt).
 * pressure(q) is evaluated on q itself, as the reference passes q(:,ix,iy,iz)
 * (it rewrites q only with dual energy, which the conduction path refuses). */
int ev_thermalstep(const ev_params *p, double *q, const double *w, double *dt) {
  if (p->dualenergy) return EV_ERR_CONFIG;
  double v = 1e99;
  int iz;
#pragma omp parallel for reduction(min : v) schedule(static)
  for (iz = 1; iz <= p->nz; iz++)
    for (int iy = 1; iy <= p->ny; iy++)
      for (int ix = 1; ix <= p->nx; ix++) {
        double *qc = &q[qidx(p, ix, iy, iz)];
        double press;
        pressure(p, qc, &press);
        const double temp = press / qc[0] * p->temp0;  /* comp.f90:99 */
        const double kappa = p->kappa0 * pow(temp, 2.5);
        const double r = w[widx(p, ix, iy, iz)] / kappa;
        v = r < v ? r : v;  /* min(w(1)/kappa, v) */
      }
  v = p->c_cond * (p->dx * p->dx) * v / p->temp0;
  if (v < *dt) *dt = v;
  if (*dt <= 0.0) return EV_ERR_DT;
  return EV_OK;
}

/* temperature of a w cell (comp.f90:96-100 without chemistry) */
static inline double wtemp(const ev_params *p, const double *wc) { return wc[4] / wc[0] * p->temp0; }

/* conduction.f90:126-144 with simplekappa: j = -(1/g1) (tr - tl) / delta */
static inline double tflux(const ev_params *p, double tl, double tr, double deltainv) {
  const double kappa_eff = 1.0 / p->g1;
  const double dtempdx = (tr - tl) * deltainv;
  return -kappa_eff * dtempdx;
}

/* conduction.f90:146-162 */
static inline void tintegrate(const ev_params *p, double jl, double jr, double *q, double delta) {
  const double de = delta * (jl - jr);
  q[4] = q[4] + de;
  q[5] = q[5] + p->g1 * de / pow(q[0], p->g1);
}

/* ---- conduction.f90:26-123  thermalconduction: x, y, z sweeps of the
 * temperature flux, all from the same w, accumulated into q(5), q(6).
 * Only simplekappa: the reference's other branch (conduction.f90:135-141)
 * reads dtempdx before assigning it. */
int ev_thermalconduction(const ev_params *p, double *q, const double *w, double dt) {
  if (!p->simplekappa || p->dualenergy) return EV_ERR_CONFIG;
  const long sx = NW, sy = (long)NW * (p->nx + 4), sz = (long)NW * (p->nx + 4) * (p->ny + 4);
  const int nd[3] = {p->nx, p->ny, p->nz};
  const double dinv[3] = {p->dxinv, p->dyinv, p->dzinv};
  const long step[3] = {sx, sy, sz};
  for (int d = 0; d < 3; d++) {
    if (nd[d] == 1) continue;
    const double delta = dt * dinv[d];
    /* pencils along d; the two other indices run over the interior */
    const int a = d == 0 ? 1 : 0, b = d == 2 ? 1 : 2;
    const int na = nd[a], nb = nd[b];
    int ib;
#pragma omp parallel for schedule(static)
    for (ib = 1; ib <= nb; ib++)
      for (int ia = 1; ia <= na; ia++) {
        int idx[3];
        idx[a] = ia;
        idx[b] = ib;
        idx[d] = 0;
        const double *w0 = &w[widx(p, idx[0], idx[1], idx[2])];
        double templ = wtemp(p, w0);
        double tempr = wtemp(p, w0 + step[d]);
        double jr = tflux(p, templ, tempr, dinv[d]);
        for (int i = 1; i <= nd[d]; i++) {
          idx[d] = i;
          const double jl = jr;
          templ = tempr;
          tempr = wtemp(p, &w[widx(p, idx[0], idx[1], idx[2])] + step[d]);
          jr = tflux(p, templ, tempr, dinv[d]);
          tintegrate(p, jl, jr, &q[qidx(p, idx[0], idx[1], idx[2])], delta);
        }
      }
  }
  return EV_OK;
}

/* ---- gravitation.f90:26-88 (gravpotential, multigreen), :172-254 (gravintegration, grav_force) ---- */

/* The transforms of fft.f90:58-80 (FFTW, complex to complex, in place, no
 * normalisation): sign -1 forward, +1 backward, along every axis with
 * n > 1.  Restated as the defining sums a_k = sum_j a_j exp(sign 2 pi i jk/n),
 * evaluated directly along each axis (O(n^2) per line, for test sizes), with
 * the twiddles of the true pi. */
static const double kPi = 3.14159265358979323846;

static void dft_axis(double *re, double *im, int n, long stride, long nlines, const long *line0, int sign) {
  double *c = malloc(sizeof(double) * (size_t)n), *sn = malloc(sizeof(double) * (size_t)n);
  double *br = malloc(sizeof(double) * (size_t)n), *bi = malloc(sizeof(double) * (size_t)n);
  for (int j = 0; j < n; j++) {
    c[j] = cos(2.0 * kPi * j / n);
    sn[j] = sign * sin(2.0 * kPi * j / n);
  }
  for (long l = 0; l < nlines; l++) {
    const long b = line0[l];
    for (int k = 0; k < n; k++) {
      double sr = 0.0, si = 0.0;
      for (int j = 0; j < n; j++) {
        const int m = (int)(((long)j * k) % n);
        const double xr = re[b + stride * j], xi = im[b + stride * j];
        sr += xr * c[m] - xi * sn[m];
        si += xr * sn[m] + xi * c[m];
      }
      br[k] = sr;
      bi[k] = si;
    }
    for (int k = 0; k < n; k++) {
      re[b + stride * k] = br[k];
      im[b + stride * k] = bi[k];
    }
  }
  free(c);
  free(sn);
  free(br);
  free(bi);
}

void ev_dft3(const ev_params *p, double *re, double *im, int sign) {
  const int nx = p->nx, ny = p->ny, nz = p->nz;
  long *l0 = malloc(sizeof(long) * (size_t)((long)nx * ny + (long)nx * nz + (long)ny * nz));
  long m = 0;
  if (nx > 1) {
    m = 0;
    for (int k = 0; k < nz; k++)
      for (int j = 0; j < ny; j++) l0[m++] = (long)nx * (j + (long)ny * k);
    dft_axis(re, im, nx, 1, m, l0, sign);
  }
  if (ny > 1) {
    m = 0;
    for (int k = 0; k < nz; k++)
      for (int i = 0; i < nx; i++) l0[m++] = i + (long)nx * ny * k;
    dft_axis(re, im, ny, nx, m, l0, sign);
  }
  if (nz > 1) {
    m = 0;
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) l0[m++] = i + (long)nx * j;
    dft_axis(re, im, nz, (long)nx * ny, m, l0, sign);
  }
  free(l0);
}

/* comp.f90:241-252 */
static double kvalue(int i, int n, int nyq) { return i <= nyq ? (double)(i - 1) : -(double)(n - (i - 1)); }

/* gravitation.f90:55-89: phi(k) = phi(k) / n3 * green(k), green = -h2 / sum sin^2(kf/2 k), 0 at k = 0.
 * `phi / dble(n3)` divides a complex by a real promoted to complex; the
 * reference's compiler forms the textbook quotient (a c + b d) / (c^2 + d^2),
 * (b c - a d) / (c^2 + d^2) with d = 0, which is not a / c bit for bit when
 * n3 is not a power of two, and that is the form restated here. */
void ev_multigreen(const ev_params *p, double *re, double *im) {
  const double n3 = (double)p->nx * p->ny * p->nz;
  for (int iz = 1; iz <= p->nz; iz++) {
    const double sz = sin(p->kfz2 * kvalue(iz, p->nz, p->nyqz));
    const double sin2z = sz * sz;
    for (int iy = 1; iy <= p->ny; iy++) {
      const double sy = sin(p->kfy2 * kvalue(iy, p->ny, p->nyqy));
      const double sin2y = sy * sy;
      for (int ix = 1; ix <= p->nx; ix++) {
        const double sx = sin(p->kfx2 * kvalue(ix, p->nx, p->nyqx));
        const double sin2x = sx * sx;
        const double green = (ix == 1 && iy == 1 && iz == 1) ? 0.0 : -p->h2 / (sin2x + sin2y + sin2z);
        const long c = (long)(ix - 1) + (long)p->nx * ((long)(iy - 1) + (long)p->ny * (iz - 1));
        const double a = re[c], b = im[c], den = n3 * n3 + 0.0 * 0.0;
        re[c] = (a * n3 + b * 0.0) / den * green;
        im[c] = (b * n3 - a * 0.0) / den * green;
      }
    }
  }
}

/* gravitation.f90:26-52 for q(nq,...): the real part of the potential into
 * phi(1:nx,1:ny,1:nz) (grav_force reads dble(phi) only).  cosmo is the
 * expansion factor of the supercomoving source. */
int ev_gravpotential(const ev_params *p, const double *q, double *phi, double cosmo) {
  const long n = (long)p->nx * p->ny * p->nz;
  double *re = malloc(sizeof(double) * (size_t)n), *im = calloc((size_t)n, sizeof(double));
  if (!re || !im) {
    free(re);
    free(im);
    return EV_ERR_CONFIG;
  }
  if (p->supercomoving) {
    for (long c = 0; c < n; c++) re[c] = cosmo * p->grav * (q[NQ * c] / p->fb - 1.0);
  } else {
    double meanrho = 0.0;
    for (long c = 0; c < n; c++) meanrho += q[NQ * c];
    meanrho = meanrho / (double)n;
    for (long c = 0; c < n; c++) re[c] = p->grav * (q[NQ * c] - meanrho);
  }
  ev_dft3(p, re, im, -1);
  ev_multigreen(p, re, im);
  ev_dft3(p, re, im, +1);
  memcpy(phi, re, sizeof(double) * (size_t)n);
  free(re);
  free(im);
  return EV_OK;
}

/* gravitation.f90:172-254, one rank (periodic in every direction) */
int ev_gravintegration(const ev_params *p, double *q, const double *phi, double dt) {
  const int nx = p->nx, ny = p->ny, nz = p->nz;
#define PHI(i, j, k) phi[(long)((i) - 1) + (long)nx * ((long)((j) - 1) + (long)ny * ((k) - 1))]
  for (int iz = 1; iz <= nz; iz++)
    for (int iy = 1; iy <= ny; iy++)
      for (int ix = 1; ix <= nx; ix++) {
        double f[3];
        if (nx == 1) f[0] = 0;
        else if (ix == 1) f[0] = -(PHI(2, iy, iz) - PHI(nx, iy, iz)) * p->dx2inv;
        else if (ix == nx) f[0] = -(PHI(1, iy, iz) - PHI(nx - 1, iy, iz)) * p->dx2inv;
        else f[0] = -(PHI(ix + 1, iy, iz) - PHI(ix - 1, iy, iz)) * p->dx2inv;
        if (ny == 1) f[1] = 0;
        else if (iy == 1) f[1] = -(PHI(ix, 2, iz) - PHI(ix, ny, iz)) * p->dy2inv;
        else if (iy == ny) f[1] = -(PHI(ix, 1, iz) - PHI(ix, ny - 1, iz)) * p->dy2inv;
        else f[1] = -(PHI(ix, iy + 1, iz) - PHI(ix, iy - 1, iz)) * p->dy2inv;
        if (nz == 1) f[2] = 0;
        else if (iz == 1) f[2] = -(PHI(ix, iy, 2) - PHI(ix, iy, nz)) * p->dz2inv;
        else if (iz == nz) f[2] = -(PHI(ix, iy, 1) - PHI(ix, iy, nz - 1)) * p->dz2inv;
        else f[2] = -(PHI(ix, iy, iz + 1) - PHI(ix, iy, iz - 1)) * p->dz2inv;
        double *qc = &q[qidx(p, ix, iy, iz)];
        qc[4] = qc[4] + dt * (qc[1] * f[0] + qc[2] * f[1] + qc[3] * f[2]);
        for (int d = 0; d < 3; d++) qc[1 + d] = qc[1 + d] + dt * qc[0] * f[d];
      }
#undef PHI
  return EV_OK;
}

int ev_step(const ev_params *p, double *q, double *w, double *evo, double *s, double *t, double tend,
            double *dt, int *go) {
  int err = ev_primitive(p, w, q);
  if (err) return err;
  ev_bound(p, w);
  *dt = 1e99;
  if (!p->hydro_off) { /* main.f90:119 */
    err = ev_hydrostep(p, w, dt);
    if (err) return err;
  }
  if (p->conduction) { /* main.f90:121 */
    err = ev_thermalstep(p, q, w, dt);
    if (err) return err;
  }
  if (tend - *t <= *dt) {
    *dt = tend - *t;
    *go = 0;
  }
  *t = *t + *dt;
  if (p->secondorder) {
    err = ev_compevo(p, w, evo, s, *dt);
    if (err) return err;
  }
  if (!p->hydro_off) { /* main.f90:155 */
    err = ev_sweep(p, w, q, evo, s, *dt);
    if (err) return err;
  }
  if (p->gravity) { /* main.f90:157-164, one rank, not comoving (cosmo = 1) */
    double *phi = malloc(sizeof(double) * (size_t)((long)p->nx * p->ny * p->nz));
    if (!phi) return EV_ERR_CONFIG;
    err = ev_gravpotential(p, q, phi, 1.0);
    if (!err) err = ev_gravintegration(p, q, phi, *dt);
    free(phi);
    if (err) return err;
  }
  if (!p->conduction) return EV_OK;
  /* main.f90:170-175 */
  err = ev_primitive(p, w, q);
  if (err) return err;
  ev_bound(p, w);
  return ev_thermalconduction(p, q, w, *dt);
}

int ev_run(const ev_params *p, double *q, int nsteps, double *t, double tend, double *dt_last, int *done) {
  double *w = calloc((size_t)ev_size_w(p), sizeof(double));
  double *evo = calloc((size_t)ev_size_evo(p), sizeof(double));
  double *s = calloc((size_t)ev_size_s(p), sizeof(double));
  int err = EV_OK, go = 1;
  *done = 0;
  if (!w || !evo || !s) {
    err = EV_ERR_CONFIG;
  } else {
    for (int k = 0; k < nsteps && go; k++) {
      err = ev_step(p, q, w, evo, s, t, tend, dt_last, &go);
      if (err) break;
      *done = k + 1;
    }
  }
  free(w);
  free(evo);
  free(s);
  return err;
}

/* ---- snapshot.f90:169-172 sums, without the `volume` factor ---- */
void ev_sums(const ev_params *p, const double *q, double out[3]) {
  double m = 0, e = 0, s = 0;
  long n = (long)p->nx * p->ny * p->nz;
  for (long i = 0; i < n; i++) {
    m += q[NQ * i + 0];
    e += q[NQ * i + 4];
    s += q[NQ * i + 5];
  }
  out[0] = m;
  out[1] = e;
  out[2] = s;
}
