/*
 * smaug_oracle.c -- CPU restatement of the SMAUG MHD time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity oracle for the HIP
 * engine in smaug_amd/csrc: only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it, and only as the checker.  The
 * product never calls it.
 *
 * It restates, kernel by kernel and on the full grid including ghost cells,
 * what the reference CUDA code computes (mikeg64/smaug, paths relative to
 * the repository's smaug/ directory), with the reference's masks,
 * temporaries and floating-point expression order, compiled with
 * -ffp-contract=off.  Each function cites the reference lines it follows.
 * Known reference quirks that change results are reproduced and named Qn
 * (numbering of the quirk register kept with the design notes):
 *   Q1  boundary-gradient zeroing with C precedence (A || B && C)
 *   Q2  derived/temporary arrays index boundtype[] with their own enum value
 *   Q3  v and p_t evaluated on transverse-interior cells only
 *   Q4  CFL uses the last direction's cmax; dt keeps its previous value when
 *       the CFL step is not below DT
 *   Q7  right-chain hyperdiffusion halo: stray write of q with the unpadded
 *       encoder, resolved as "threads run in increasing linear index"
 *   Q8  left-chain hyperdiffusion halo rule (padded 0 <- 6)
 *   Q10 user source on the post-update state with p->dt and p->qt
 *   Q11 grid extrapolation of ghost positions before delx
 *   Q15 hyperdiffusion sources touch ghost cells (overwritten by the BCs)
 *   Q17 cmax for the hyperdiffusion computed once per dim in the rho loop
 * Periodic boundaries are supported in 2D only (the reference's 3D periodic
 * kernels fill only the k=0 plane, src/boundary.cu:96,121).
 */
#include "smaug_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define BIGD 9.9e12    /* include/iotypes.h:19 */
#define SMALLD 1.0e-12 /* include/iotypes.h:20 */

struct orc {
  smaug_params p;
  int nd, n0, n1, n2;
  long N, NP1, NP2;
  int nvar, nderv;
  int nslot;
  double* w;      /* nslot * nvar * N                         */
  double* wd;     /* nderv * N                                */
  double* wt;     /* 8 * N      (wtemp tmp1..tmp8)            */
  double* wt1;    /* 2 * NP1    (wtemp1 d1, d3)               */
  double* wt2;    /* NP2        (wtemp2 tmpnui)               */
  double* dwn1;   /* nvar * N                                 */
  double dt;      /* p->dt                                    */
  double qt;      /* p->qt                                    */
  double t;       /* state->t                                 */
  int it;
  double cmax_dim[3];
  int cur;        /* slot holding the current state           */
  double* amap;   /* driver amplitude per column (n1*n2), or NULL */
  double* fixed;  /* evolved fields of the initial state: fixed-boundary values */
  int nu_mode;    /* 0 compute nu, 1 compute and record, 2 replay */
  double* nu_tape;
  long nu_len, nu_cap, nu_pos;
};

/* ---- enums (include/iotypes.h:211-242), as functions of ndim ---------- */
#define F_RHO 0
#define F_MOM(d) (1 + (d))
#define F_E (1 + o->nd)
#define F_B(d) (2 + o->nd + (d))
#define F_EB (2 + 2 * o->nd)
#define F_RHOB (3 + 2 * o->nd)
#define F_BB(d) (4 + 2 * o->nd + (d))
#define D_VEL(d) (d)
#define D_FLUX (o->nd)
#define D_HDNUR (o->nd + 1)
#define D_HDNUL (o->nd + 2)
#define D_NUSHK(d) (o->nd + 3 + (d))
#define D_CS (2 * o->nd + 3)
#define D_PT (2 * o->nd + 4)
#define D_PK (2 * o->nd + 5)
#define D_PTB (2 * o->nd + 9)
#define D_PKB (2 * o->nd + 10)
#define D_CF (2 * o->nd + 8)
#define D_POS(d) (2 * o->nd + 11 + (d))
#define D_DELX(d) (3 * o->nd + 11 + (d))
enum { TMP1 = 0, TMP2, TMP3, TMP4, TMP5, TMP6, TMP7, TMP8 };
enum { T1_D1 = 0, T1_D3 = 1 };

static inline long enc(const orc_t* o, int i, int j, int k) {
  return ((long)k * o->n1 + j) * o->n0 + i;
}
static inline long enc1(const orc_t* o, int i, int j, int k) {
  if (o->nd == 2) k = 0;
  return ((long)k * (o->n1 + 1) + j) * (o->n0 + 1) + i;
}
static inline long enc2(const orc_t* o, int i, int j, int k) {
  if (o->nd == 2) k = 0;
  return ((long)k * (o->n1 + 2) + j) * (o->n0 + 2) + i;
}
static inline double* W(orc_t* o, int slot, int f) {
  return o->w + ((long)slot * o->nvar + f) * o->N;
}
static inline double* WD(orc_t* o, int f) { return o->wd + (long)f * o->N; }
static inline double* WT(orc_t* o, int f) { return o->wt + (long)f * o->N; }
static inline int inr(int x, int lo, int hi) { return x > lo && x < hi; }

#define FOR_ALL(o)                              \
  for (int k = 0; k < (o)->n2; k++)             \
    for (int j = 0; j < (o)->n1; j++)           \
      for (int i = 0; i < (o)->n0; i++)

static inline int nof(const orc_t* o, int d) {
  return d == 0 ? o->n0 : (d == 1 ? o->n1 : o->n2);
}
static inline int coord(int i, int j, int k, int d) {
  return d == 0 ? i : (d == 1 ? j : k);
}
static inline long stride(const orc_t* o, int d) {
  return d == 0 ? 1 : (d == 1 ? o->n0 : (long)o->n0 * o->n1);
}
static inline int nonper(const orc_t* o, int bcidx, int d) {
  return o->p.boundtype[bcidx][d][0] != 0;
}

/* ---------------- differential operators (src/gradops.cu) ------------- */

/* grad3dn_MODID, src/gradops.cu:437-602: 4th-order central difference over
 * 2*delx, zeroed at the first/last two interior layers of non-periodic
 * directions with the (A || B && C) precedence of Q1. */
static double g3dn(orc_t* o, const double* a, int bcidx, int i, int j, int k,
                   int dir) {
  const int x = coord(i, j, k, dir), nn = nof(o, dir);
  const long c = enc(o, i, j, k), s = stride(o, dir);
  double g = 0;
  if (x > 1 && x < nn - 2)
    g = ((8 * a[c + s] - 8 * a[c - s] + a[c - 2 * s] - a[c + 2 * s]) / 6.0) /
        (2.0 * WD(o, D_DELX(dir))[c]);
  if (nonper(o, bcidx, dir)) {
    int t;
    if (o->nd == 3) {
      if (dir == 0) t = inr(j, 1, o->n1 - 2) && inr(k, 1, o->n2 - 2);
      else if (dir == 1) t = inr(i, 1, o->n0 - 2) && inr(k, 1, o->n2 - 2);
      else t = inr(i, 1, o->n0 - 2) && inr(j, 1, o->n1 - 2);
    } else {
      t = dir == 0 ? inr(j, 1, o->n1 - 2) : inr(i, 1, o->n0 - 2);
    }
    if (x == nn - 3 || (x == nn - 4 && t)) g = 0;
    else if (x == 2 || (x == 3 && t)) g = 0;
  }
  return g;
}

/* transverse conditions of the first-order operators' zeroing (Q1), with
 * the asymmetric k>1 of grad1l3n dir 0 (src/gradops.cu:627) and j>1 of
 * grad13n dir 2 (src/gradops.cu:816). kind: 0 = grad1l3n, 1 = grad1r3n,
 * 2 = grad13n.  hi selects the condition attached to n-3 (1) or 2 (0). */
static int g1_trans(orc_t* o, int kind, int dir, int hi, int i, int j, int k) {
  if (o->nd == 2) return dir == 0 ? inr(j, 0, o->n1 - 1) : inr(i, 0, o->n0 - 1);
  if (dir == 0) {
    int klo = (kind == 0 && hi) ? 1 : 0;
    return inr(j, 0, o->n1 - 1) && inr(k, klo, o->n2 - 1);
  }
  if (dir == 1) return inr(i, 0, o->n0 - 1) && inr(k, 0, o->n2 - 1);
  {
    int jlo = (kind == 2 && !hi) ? 1 : 0;
    return inr(i, 0, o->n0 - 1) && inr(j, jlo, o->n1 - 1);
  }
}

static double g1_zero(orc_t* o, double g, int kind, int bcidx, int i, int j,
                      int k, int dir) {
  if (!nonper(o, bcidx, dir)) return g;
  const int x = coord(i, j, k, dir), nn = nof(o, dir);
  if (x == nn - 2 || (x == nn - 3 && g1_trans(o, kind, dir, 1, i, j, k))) return 0;
  if (x == 1 || (x == 2 && g1_trans(o, kind, dir, 0, i, j, k))) return 0;
  return g;
}

/* grad1l3n_MODID, src/gradops.cu:608-677 */
static double g1l(orc_t* o, const double* a, int bcidx, int i, int j, int k,
                  int dir) {
  const int x = coord(i, j, k, dir), nn = nof(o, dir);
  if (!(x > 0 && x < nn)) return 0;
  const long c = enc(o, i, j, k), s = stride(o, dir);
  double g = (a[c] - a[c - s]) / (WD(o, D_DELX(dir))[c]);
  return g1_zero(o, g, 0, bcidx, i, j, k, dir);
}

/* grad1r3n_MODID, src/gradops.cu:680-750 */
static double g1r(orc_t* o, const double* a, int bcidx, int i, int j, int k,
                  int dir) {
  const int x = coord(i, j, k, dir), nn = nof(o, dir);
  if (!(x < nn - 1)) return 0;
  const long c = enc(o, i, j, k), s = stride(o, dir);
  double g = (a[c + s] - a[c]) / (WD(o, D_DELX(dir))[c]);
  return g1_zero(o, g, 1, bcidx, i, j, k, dir);
}

/* grad13n_MODID, src/gradops.cu:755-823 */
static double g13(orc_t* o, const double* a, int bcidx, int i, int j, int k,
                  int dir) {
  const int x = coord(i, j, k, dir), nn = nof(o, dir);
  if (!(x > 0 && x < nn - 1)) return 0;
  const long c = enc(o, i, j, k), s = stride(o, dir);
  double g = (a[c + s] - a[c - s]) / (WD(o, D_DELX(dir))[c]) / 2.0;
  return g1_zero(o, g, 2, bcidx, i, j, k, dir);
}

/* ---------------- derived fields (src/dervfields.cu) ------------------ */

static inline double rt(orc_t* o, int slot, long c) {
  return W(o, slot, F_RHO)[c] + W(o, slot, F_RHOB)[c];
}

/* computevel3, src/dervfields.cu:79-98 */
static void vel_cell(orc_t* o, int s, long c) {
  for (int d = 0; d < o->nd; d++)
    WD(o, D_VEL(d))[c] = W(o, s, F_MOM(d))[c] / (W(o, s, F_RHO)[c] + W(o, s, F_RHOB)[c]);
}

/* computept3, src/dervfields.cu:107-153 */
static void pt_cell(orc_t* o, int s, long c) {
  const double gm = o->p.gamma;
  double bbb = 0, bb2 = 0, mm = 0;
  for (int d = 0; d < o->nd; d++) {
    if (d == 0) {
      bbb = W(o, s, F_B(0))[c] * W(o, s, F_BB(0))[c];
      bb2 = W(o, s, F_B(0))[c] * W(o, s, F_B(0))[c];
      mm = W(o, s, F_MOM(0))[c] * W(o, s, F_MOM(0))[c];
    } else {
      bbb = bbb + W(o, s, F_B(d))[c] * W(o, s, F_BB(d))[c];
      bb2 = bb2 + W(o, s, F_B(d))[c] * W(o, s, F_B(d))[c];
      mm = mm + W(o, s, F_MOM(d))[c] * W(o, s, F_MOM(d))[c];
    }
  }
  double v = (gm - 2.0) * (bbb + 0.5 * bb2);
  WD(o, D_PT)[c] = (gm - 1.0) * (W(o, s, F_E)[c] - 0.5 * mm / (W(o, s, F_RHO)[c] + W(o, s, F_RHOB)[c])) - v;
}

/* computepbg3, src/dervfields.cu:157-191 */
static void pbg_cell(orc_t* o, int s, long c) {
  const double gm = o->p.gamma;
  double bb = 0;
  for (int d = 0; d < o->nd; d++) {
    double t = W(o, s, F_BB(d))[c] * W(o, s, F_BB(d))[c];
    bb = d == 0 ? t : bb + t;
  }
  WD(o, D_PTB)[c] = (gm - 1) * W(o, s, F_EB)[c] - 0.5 * (gm - 2) * bb;
}

/* computec3, src/dervfields.cu:241-309 */
static void c_cell(orc_t* o, int s, long c, int dir) {
  const double gm = o->p.gamma;
  double mm = 0, b2 = 0, bbb = 0, bb2 = 0, bbb2 = 0;
  for (int d = 0; d < o->nd; d++) {
    double tm = W(o, s, F_MOM(d))[c] * W(o, s, F_MOM(d))[c];
    double tb = W(o, s, F_B(d))[c] * W(o, s, F_B(d))[c];
    double tbb = W(o, s, F_BB(d))[c] * W(o, s, F_B(d))[c];
    double tb2 = W(o, s, F_BB(d))[c] * W(o, s, F_BB(d))[c];
    mm = d == 0 ? tm : mm + tm;
    b2 = d == 0 ? tb : b2 + tb;
    bbb = d == 0 ? tbb : bbb + tbb;
    bb2 = d == 0 ? tb2 : bb2 + tb2;
  }
  bbb2 = bbb;
  const double rho_t = W(o, s, F_RHO)[c] + W(o, s, F_RHOB)[c];
  double pk = (gm - 1) * (W(o, s, F_E)[c] - 0.5 * (mm / rho_t) - 0.5 * b2 - bbb);
  double cs2 = ((gm) * (pk + ((gm)-1) * (W(o, s, F_EB)[c] - 0.5 * bb2))) / rho_t;
  WD(o, D_CS)[c] = cs2;
  double cf = (b2 + bb2 + 2.0 * bbb2) / rho_t + cs2;
  WD(o, D_CF)[c] = cf;
  double btd = W(o, s, F_BB(dir))[c] + W(o, s, F_B(dir))[c];
  double cfi = 0.5 * (cf + sqrt(cf * cf - 4.0 * cs2 * (btd * btd) / rho_t));
  WD(o, D_CF)[c] = sqrt(cfi) + fabs(W(o, s, F_MOM(dir))[c] / rho_t);
  WD(o, D_CS)[c] = sqrt(WD(o, D_CS)[c]);
}

/* transverse-interior mask of computevels/computept/centdiff flux (Q3),
 * src/computedervfields.cu:82-105 */
static int tint(orc_t* o, int i, int j, int k, int dir) {
  if (o->nd == 3) {
    if (dir == 0) return inr(j, 1, o->n1 - 2) && inr(k, 1, o->n2 - 2);
    if (dir == 1) return inr(i, 1, o->n0 - 2) && inr(k, 1, o->n2 - 2);
    return inr(i, 1, o->n0 - 2) && inr(j, 1, o->n1 - 2);
  }
  return dir == 0 ? inr(j, 1, o->n1 - 2) : inr(i, 1, o->n0 - 2);
}
static int interior(orc_t* o, int i, int j, int k) {
  if (!(inr(i, 1, o->n0 - 2) && inr(j, 1, o->n1 - 2))) return 0;
  return o->nd == 2 || inr(k, 1, o->n2 - 2);
}

/* cucomputevels -> computevels_parallel, src/computedervfields.cu:18-126 */
static void k_vels(orc_t* o, int s, int dir) {
#pragma omp parallel for collapse(2)
  FOR_ALL(o) if (tint(o, i, j, k, dir)) vel_cell(o, s, enc(o, i, j, k));
}

/* cucomputept: computeptzero then computept, src/computedervfields.cu:2266-2299 */
static void k_pt(orc_t* o, int s, int dir) {
#pragma omp parallel for collapse(2)
  FOR_ALL(o) WD(o, D_PT)[enc(o, i, j, k)] = 0.0;
#pragma omp parallel for collapse(2)
  FOR_ALL(o) if (tint(o, i, j, k, dir)) pt_cell(o, s, enc(o, i, j, k));
}

/* cucomputepbg, src/computedervfields.cu:527-661 (all cells) */
static void k_pbg(orc_t* o, int s) {
#pragma omp parallel for collapse(2)
  FOR_ALL(o) pbg_cell(o, s, enc(o, i, j, k));
}

/* cucomputec + cucomputemaxc: max over all cells of cfast then soundspeed,
 * src/computedervfields.cu:1457-1533,1864-2050.  Returns p->cmax. */
static double k_cmax(orc_t* o, int s, int dir) {
#pragma omp parallel for collapse(2)
  FOR_ALL(o) c_cell(o, s, enc(o, i, j, k), dir);
  double cm = 0.0;
  for (long c = 0; c < o->N; c++)
    if (WD(o, D_CF)[c] > cm) cm = WD(o, D_CF)[c];
  for (long c = 0; c < o->N; c++)
    if (WD(o, D_CS)[c] > cm) cm = WD(o, D_CS)[c];
  return cm;
}

/* ---------------- fluxes (src/centdiff1.cu, src/centdiff2.cu) --------- */

/* computeflux, src/centdiff1.cu:51-256 */
static double flux1(orc_t* o, int s, long c, int f, int d) {
  const double rho = W(o, s, F_RHO)[c], rhob = W(o, s, F_RHOB)[c];
  const double md = W(o, s, F_MOM(d))[c];
  if (f == F_RHO) return md * rho / (rho + rhob) + (rhob * md) / (rhob + rho);
  const int cf = f - 1;
  const double tr = md * W(o, s, f)[c] / (rho + rhob);
  const double bf = W(o, s, F_B(cf))[c], bbf = W(o, s, F_BB(cf))[c];
  const double bd = W(o, s, F_B(d))[c], bbd = W(o, s, F_BB(d))[c];
  double fm;
  if (d == cf) fm = WD(o, D_PT)[c] - (bf * bbd + bbf * bd) - bf * bd;
  else fm = -(bf * bbd + bbf * bd) - bf * bd;
  return tr + fm;
}

/* computefluxe / computefluxb*, src/centdiff2.cu:158-356 */
static double flux2(orc_t* o, int s, long c, int f, int d) {
  const double rho = W(o, s, F_RHO)[c], rhob = W(o, s, F_RHOB)[c];
  const double md = W(o, s, F_MOM(d))[c];
  if (f == F_E) {
    double bbm = 0, bm = 0;
    for (int q = 0; q < o->nd; q++) {
      double t1 = W(o, s, F_BB(q))[c] * W(o, s, F_MOM(q))[c];
      double t2 = W(o, s, F_B(q))[c] * W(o, s, F_MOM(q))[c];
      bbm = q == 0 ? t1 : bbm + t1;
      bm = q == 0 ? t2 : bm + t2;
    }
    double ft = md * (WD(o, D_PT)[c]);
    ft -= W(o, s, F_B(d))[c] * bbm;
    ft -= W(o, s, F_BB(d))[c] * bm;
    ft /= (rho + rhob);
    ft += md * W(o, s, F_EB)[c] / (rho + rhob);
    ft -= W(o, s, F_B(d))[c] * bm / (rho + rhob);
    return md * W(o, s, F_E)[c] / (rho + rhob) + ft;
  }
  const int cf = f - F_B(0);
  if (d == cf) return 0.0;
  return md * W(o, s, f)[c] / (rho + rhob) -
         (W(o, s, F_B(d))[c] + W(o, s, F_BB(d))[c]) * W(o, s, F_MOM(cf))[c] / (rho + rhob) +
         (W(o, s, F_BB(cf))[c]) * md / (rho + rhob);
}

/* cucentdiff1, src/centdiff1.cu:822-851: init, flux, divergence, update */
static void k_centdiff1(orc_t* o, int so, int sn, double dt, int f, int dir) {
  double* dw = o->dwn1 + (long)f * o->N;
  double* fl = WD(o, D_FLUX);
#pragma omp parallel for collapse(2)
  FOR_ALL(o) {
    long c = enc(o, i, j, k);
    dw[c] = 0.0;
    fl[c] = 0.0;
  }
#pragma omp parallel for collapse(2)
  FOR_ALL(o) if (tint(o, i, j, k, dir)) {
    long c = enc(o, i, j, k);
    fl[c] = 0.0;
    fl[c] = flux1(o, so, c, f, dir);
  }
#pragma omp parallel for collapse(2)
  FOR_ALL(o) if (interior(o, i, j, k)) {
    long c = enc(o, i, j, k);
    dw[c] += g3dn(o, fl, D_FLUX, i, j, k, dir);
  }
  double* wn = W(o, sn, f);
#pragma omp parallel for collapse(2)
  FOR_ALL(o) if (tint(o, i, j, k, dir)) {
    long c = enc(o, i, j, k);
    wn[c] = wn[c] - dt * dw[c];
  }
}

/* addenergyterms_cd2, src/centdiff2.cu:80-115 */
static void eterms_cell(orc_t* o, int sn, int i, int j, int k, int dir) {
  const long c = enc(o, i, j, k);
  double* e = W(o, sn, F_E);
  const double dt = o->dt;
  e[c] -= +(dt)*WD(o, D_PTB)[c] * g3dn(o, WD(o, D_VEL(dir)), D_VEL(dir), i, j, k, dir);
  for (int idim = 0; idim < o->nd; idim++)
    e[c] += (dt)*W(o, sn, F_BB(idim))[c] * W(o, sn, F_BB(dir))[c] *
            g3dn(o, WD(o, D_VEL(dir)), D_VEL(dir), i, j, k, idim);
}

/* cucentdiff2, src/centdiff2.cu:1117-1171: init, flux, divergence, update,
 * background pressure of the new slot, energy terms */
static void k_centdiff2(orc_t* o, int so, int sn, double dt, int f, int dir) {
  double* dw = o->dwn1 + (long)f * o->N;
  double* fl = WD(o, D_FLUX);
#pragma omp parallel for collapse(2)
  FOR_ALL(o) {
    long c = enc(o, i, j, k);
    dw[c] = 0.0;
    fl[c] = 0.0;
  }
#pragma omp parallel for collapse(2)
  FOR_ALL(o) if (tint(o, i, j, k, dir)) {
    long c = enc(o, i, j, k);
    fl[c] = 0.0;
    fl[c] = flux2(o, so, c, f, dir);
  }
#pragma omp parallel for collapse(2)
  FOR_ALL(o) if (interior(o, i, j, k)) {
    long c = enc(o, i, j, k);
    dw[c] = g3dn(o, fl, D_FLUX, i, j, k, dir);
  }
  double* wn = W(o, sn, f);
#pragma omp parallel for collapse(2)
  FOR_ALL(o) if (tint(o, i, j, k, dir)) {
    long c = enc(o, i, j, k);
    wn[c] = wn[c] - dt * dw[c];
  }
  /* centdiff2ci: p_tb from the output slot, all cells */
  k_pbg(o, sn);
#pragma omp parallel for collapse(2)
  FOR_ALL(o) dw[enc(o, i, j, k)] = 0.0;
  if (f == F_E && (o->p.terms & SMAUG_T_ETERM)) {
    /* centdiff2c region: 3D interior; 2D i<n0-2 && j<n1-2 (src/centdiff2.cu:644-649) */
#pragma omp parallel for collapse(2)
    FOR_ALL(o) {
      int ok = o->nd == 3 ? interior(o, i, j, k) : (i < o->n0 - 2 && j < o->n1 - 2);
      if (ok) eterms_cell(o, sn, i, j, k, dir);
    }
  }
}

/* cugrav -> addgrav_cd2, src/centdiff2.cu:118-154,721-787 */
static void k_grav(orc_t* o, int so, int sn) {
  const double dt = o->dt;
#pragma omp parallel for collapse(2)
  FOR_ALL(o) {
    int ok = o->nd == 3 ? interior(o, i, j, k) : (i < o->n0 - 2 && j < o->n1 - 2);
    if (!ok) continue;
    long c = enc(o, i, j, k);
    for (int d = 0; d < o->nd; d++)
      W(o, sn, F_MOM(d))[c] += (dt) * (o->p.g[d]) * W(o, so, F_RHO)[c];
    for (int d = 0; d < o->nd; d++)
      W(o, sn, F_E)[c] += (dt)*W(o, so, F_RHO)[c] * ((o->p.g[d]) * W(o, so, F_MOM(d))[c]) /
                          (W(o, so, F_RHO)[c] + W(o, so, F_RHOB)[c]);
  }
}

/* ---------------- divB source terms ----------------------------------- */

/* grad3d_MODID, src/gradops.cu:130-260: the 4th-order central difference
 * over 2*p->dx (not delx), zeroed at the first/last two interior layers
 * with the (A || B && C) precedence whatever the boundary types. */
static double g3d(orc_t* o, const double* a, int i, int j, int k, int dir) {
  const int x = coord(i, j, k, dir), nn = nof(o, dir);
  const long c = enc(o, i, j, k), s = stride(o, dir);
  double g = 0;
  if (x > 1 && x < nn - 2)
    g = ((8 * a[c + s] - 8 * a[c - s] + a[c - 2 * s] - a[c + 2 * s]) / 6.0) / (2.0 * o->p.dx[dir]);
  int t;
  if (o->nd == 3) {
    if (dir == 0) t = inr(j, 1, o->n1 - 2) && inr(k, 1, o->n2 - 2);
    else if (dir == 1) t = inr(i, 1, o->n0 - 2) && inr(k, 1, o->n2 - 2);
    else t = inr(i, 1, o->n0 - 2) && inr(j, 1, o->n1 - 2);
  } else {
    t = dir == 0 ? inr(j, 1, o->n1 - 2) : inr(i, 1, o->n0 - 2);
  }
  if (x == nn - 3 || (x == nn - 4 && t)) g = 0;
  else if (x == 2 || (x == 3 && t)) g = 0;
  return g;
}

/* cudivb -> divb_parallel, src/divb.cu:165-338, called after cugrav
 * (src/iosmaug.c:678-679) when divbon == 1; it acts when divbfix != 0.
 * In 2D, on cells 2 < i,j < n-2: divb = d1 b1 + d2 b2 + d1 b1b + d2 b2b
 * (grad3d), dwn1 = 0 then the sources dbsource* (src/divb.cu:17-100):
 * m_d: -divb b_d, b_d: -divb m_d/rho_t (NaN -> 0), e: -divb wd[bdotv]
 * with wd[bdotv] never computed (zeroed by the order-0 pass), rho: 0; then
 * w[ordero] -= dt dwn1 on 1 < i,j < n-2.  The launch covers n0*n1 threads
 * with k = 0 only and its 3D region test needs k > 2, so in 3D the kernel
 * changes nothing. */
static void k_divb(orc_t* o, int so, int sn, double dt) {
  if (o->p.divbon != 1 || !o->p.divbfix || o->nd != 2) return;
  const int nvf = F_B(o->nd - 1) + 1;
  double* divb = WD(o, 2 * o->nd + 7);
  double* bdotv = WD(o, 2 * o->nd + 6);
  FOR_ALL(o) {
    const long c = enc(o, i, j, k);
    for (int f = F_RHO; f < nvf; f++) o->dwn1[(long)f * o->N + c] = 0;
  }
  FOR_ALL(o) {
    if (!(i > 2 && j > 2 && i < o->n0 - 2 && j < o->n1 - 2)) continue;
    const long c = enc(o, i, j, k);
    divb[c] = g3d(o, W(o, so, F_B(0)), i, j, k, 0) + g3d(o, W(o, so, F_B(1)), i, j, k, 1);
    divb[c] += g3d(o, W(o, so, F_BB(0)), i, j, k, 0) + g3d(o, W(o, so, F_BB(1)), i, j, k, 1);
    const double rho_t = W(o, so, F_RHO)[c] + W(o, so, F_RHOB)[c];
    for (int d = 0; d < 2; d++) {
      double sm = -divb[c] * W(o, so, F_B(d))[c];
      double sb = -divb[c] * W(o, so, F_MOM(d))[c] / rho_t;
      o->dwn1[(long)F_MOM(d) * o->N + c] += isnan(sm) ? 0 : sm;
      o->dwn1[(long)F_B(d) * o->N + c] += isnan(sb) ? 0 : sb;
    }
    o->dwn1[(long)F_E * o->N + c] += -divb[c] * bdotv[c];
  }
  FOR_ALL(o) {
    if (!(i > 1 && j > 1 && i < o->n0 - 2 && j < o->n1 - 2)) continue;
    const long c = enc(o, i, j, k);
    for (int f = F_RHO; f < nvf; f++) W(o, sn, f)[c] = W(o, sn, f)[c] - dt * o->dwn1[(long)f * o->N + c];
  }
}

/* ---------------- hyperdiffusion coefficient -------------------------- */

/* q of field (hyperdifvisc1r_parallel phase 2, src/hyperdifvisc1r.cu:1225-1253) */
static double q_cell(orc_t* o, int s, long c, int field) {
  if (field == F_E) {
    double b2 = 0, mm = 0;
    for (int d = 0; d < o->nd; d++) {
      double tb = W(o, s, F_B(d))[c] * W(o, s, F_B(d))[c];
      double tm = W(o, s, F_MOM(d))[c] * W(o, s, F_MOM(d))[c];
      b2 = d == 0 ? tb : b2 + tb;
      mm = d == 0 ? tm : mm + tm;
    }
    return W(o, s, F_E)[c] - 0.5 * (b2 + mm / (W(o, s, F_RHO)[c] + W(o, s, F_RHOB)[c]));
  }
  double v = W(o, s, field)[c];
  if (field >= F_MOM(0) && field <= F_MOM(o->nd - 1))
    v = W(o, s, field)[c] / ((W(o, s, F_RHO)[c] + W(o, s, F_RHOB)[c]));
  return v;
}

/* hyperdifvisc1r_parallel / hyperdifvisc1l_parallel phase 1 for thread
 * (i,j,k): zero the temporaries (src/hyperdifvisc1r.cu:1099-1188) */
static void visc_zero_thread(orc_t* o, int i, int j, int k) {
  const long c = enc(o, i, j, k);
  for (int f = TMP1; f <= TMP8; f++) WT(o, f)[c] = 0;
  for (int f = T1_D1; f <= T1_D3; f++) o->wt1[f * o->NP1 + enc1(o, i, j, k)] = 0;
  o->wt2[enc2(o, i, j, k)] = 0;
  const int li = i == o->n0 - 1, lj = j == o->n1 - 1, lk = o->nd == 3 && k == o->n2 - 1;
  if (li) {
    for (int f = T1_D1; f <= T1_D3; f++) o->wt1[f * o->NP1 + enc1(o, i + 1, j, k)] = 0;
    o->wt2[enc2(o, i + 1, j, k)] = 0;
    o->wt2[enc2(o, i + 2, j, k)] = 0;
  }
  if (lj) {
    for (int f = T1_D1; f <= T1_D3; f++) o->wt1[f * o->NP1 + enc1(o, i, j + 1, k)] = 0;
    o->wt2[enc2(o, i, j + 1, k)] = 0;
    o->wt2[enc2(o, i, j + 2, k)] = 0;
  }
  if (lk) {
    for (int f = T1_D1; f <= T1_D3; f++) o->wt1[f * o->NP1 + enc1(o, i, j, k + 1)] = 0;
    o->wt2[enc2(o, i, j, k + 1)] = 0;
    o->wt2[enc2(o, i, j, k + 2)] = 0;
  }
  if (lj && li) {
    for (int f = T1_D1; f <= T1_D3; f++) o->wt1[f * o->NP1 + enc1(o, i + 1, j + 1, k)] = 0;
    for (int di = 0; di < 2; di++)
      for (int dj = 0; dj < 2; dj++) o->wt2[enc2(o, i + 1 + di, j + 1 + dj, k)] = 0;
  }
  if (li && lk) {
    for (int f = T1_D1; f <= T1_D3; f++) o->wt1[f * o->NP1 + enc1(o, i + 1, j, k + 1)] = 0;
    for (int di = 0; di < 2; di++)
      for (int dk = 0; dk < 2; dk++) o->wt2[enc2(o, i + 1 + di, j, k + 1 + dk)] = 0;
  }
  if (lj && lk) {
    /* the reference zeroes wtemp1 at (i+1,j+1,k) here */
    for (int f = T1_D1; f <= T1_D3; f++) o->wt1[f * o->NP1 + enc1(o, i + 1, j + 1, k)] = 0;
    for (int dk = 0; dk < 2; dk++)
      for (int dj = 0; dj < 2; dj++) o->wt2[enc2(o, i, j + 1 + dj, k + 1 + dk)] = 0;
  }
  if (li && lj && lk) {
    for (int f = T1_D1; f <= T1_D3; f++) o->wt1[f * o->NP1 + enc1(o, i + 1, j + 1, k + 1)] = 0;
    for (int dk = 0; dk < 2; dk++)
      for (int dj = 0; dj < 2; dj++)
        for (int di = 0; di < 2; di++) o->wt2[enc2(o, i + 1 + di, j + 1 + dj, k + 1 + dk)] = 0;
  }
}

/* hyperdifvisc1r_parallel: phase 1 + phase 2 with the stray write of Q7,
 * emulated with threads running in increasing linear index. */
static void k_visc1r(orc_t* o, int s, int field) {
  FOR_ALL(o) {
    long c = enc(o, i, j, k);
    visc_zero_thread(o, i, j, k);
    double q = q_cell(o, s, c, field);
    WT(o, TMP6)[c] = q;
    /* wtemp2[encode3(p,i+1,j+1,k+1,tmpnui)]: unpadded encoder on the
     * padded array (src/hyperdifvisc1r.cu:1235,1253) */
    long L = o->nd == 3 ? ((long)(k + 1) * o->n1 + (j + 1)) * o->n0 + (i + 1)
                        : (long)(j + 1) * o->n0 + (i + 1);
    o->wt2[L] = q;
    WD(o, D_HDNUR)[c] = 0;
  }
}

/* hyperdifvisc1l_parallel: same zeroing, q, no stray write */
static void k_visc1l(orc_t* o, int s, int field) {
  FOR_ALL(o) {
    long c = enc(o, i, j, k);
    visc_zero_thread(o, i, j, k);
    WT(o, TMP6)[c] = q_cell(o, s, c, field);
    WD(o, D_HDNUL)[c] = 0;
  }
}

/* hyperdifvisc1ar / 1al: padded interior <- tmp6 */
static void k_visc1a(orc_t* o) {
  FOR_ALL(o) o->wt2[enc2(o, i + 1, j + 1, k + 1)] = WT(o, TMP6)[enc(o, i, j, k)];
}

/* hyperdifvisc1bcl -> bc_hyperdifl, src/hyperdifvisc1l.cu:17-62 (Q8) */
static void k_visc1bcl(orc_t* o, int dim) {
  const int is = 1;
  FOR_ALL(o) {
    double* t = o->wt2;
    if (o->nd == 2) {
      if (dim == 0 && i == o->n0 - 1) t[enc2(o, i + 2, j + is, k)] = t[enc2(o, o->n0 - 5, j + is, k)];
      else if (dim == 1 && j == o->n1 - 1) t[enc2(o, i + is, j + 2, k)] = t[enc2(o, i + is, o->n1 - 5, k)];
      else if (dim == 0 && i == 0) t[enc2(o, 0, j + is, k)] = t[enc2(o, 6, j + is, k)];
      else if (dim == 1 && j == 0) t[enc2(o, i + is, 0, k)] = t[enc2(o, i + is, 6, k)];
    } else {
      if (dim == 0 && i == o->n0 - 1) t[enc2(o, i + 2, j + is, k + is)] = t[enc2(o, o->n0 - 5, j + is, k + is)];
      else if (dim == 1 && j == o->n1 - 1) t[enc2(o, i + is, j + 2, k + is)] = t[enc2(o, i + is, o->n1 - 5, k + is)];
      else if (dim == 2 && k == o->n2 - 1) t[enc2(o, i + is, j + is, k + 2)] = t[enc2(o, i + is, j + is, o->n2 - 5)];
      else if (dim == 0 && i == 0) t[enc2(o, 0, j + is, k + is)] = t[enc2(o, 6, j + is, k + is)];
      else if (dim == 1 && j == 0) t[enc2(o, i + is, 0, k + is)] = t[enc2(o, i + is, 6, k + is)];
      else if (dim == 2 && k == 0) t[enc2(o, i + is, j + is, 0)] = t[enc2(o, i + is, j + is, 6)];
    }
  }
}

/* hyperdifvisc2r / 2l: |third| and |first| differences on padded arrays,
 * src/hyperdifvisc1r.cu:713-843, src/hyperdifvisc1l.cu:723-853 */
static void k_visc2(orc_t* o, int dim, int right) {
  const int e0 = dim == 0, e1 = dim == 1, e2 = dim == 2;
  double* t2 = o->wt2;
  double* d1 = o->wt1 + T1_D1 * o->NP1;
  double* d3 = o->wt1 + T1_D3 * o->NP1;
  FOR_ALL(o) {
    int ok = i > 1 && j > 1 && (o->nd == 2 || k > 1);
    if (!ok) continue;
    if (right)
      d3[enc1(o, i, j, k)] = fabs(3.0 * (t2[enc2(o, i + e0, j + e1, k + e2)] - t2[enc2(o, i, j, k)]) -
                                  (t2[enc2(o, i + 2 * e0, j + 2 * e1, k + 2 * e2)] -
                                   t2[enc2(o, i - e0, j - e1, k - e2)]));
    else
      d3[enc1(o, i, j, k)] = fabs(3.0 * (t2[enc2(o, i, j, k)] - t2[enc2(o, i - e0, j - e1, k - e2)]) -
                                  (t2[enc2(o, i + e0, j + e1, k + e2)] -
                                   t2[enc2(o, i - 2 * e0, j - 2 * e1, k - 2 * e2)]));
  }
  FOR_ALL(o) {
    if (right)
      d1[enc1(o, i + 1, j + 1, k + 1)] =
          fabs((t2[enc2(o, i + e0 + 1, j + e1 + 1, k + e2 + 1)] - t2[enc2(o, i + 1, j + 1, k + 1)]));
    else
      d1[enc1(o, i + 1, j + 1, k + 1)] =
          fabs(t2[enc2(o, i + 1, j + 1, k + 1)] - (t2[enc2(o, i - e0 + 1, j - e1 + 1, k - e2 + 1)]));
  }
}

/* hyperdifvisc3r / 3l: windowed maxima, src/hyperdifvisc1r.cu:571-708 */
static void k_visc3(orc_t* o, int dim) {
  const int e0 = dim == 0, e1 = dim == 1, e2 = dim == 2;
  double* d1 = o->wt1 + T1_D1 * o->NP1;
  double* d3 = o->wt1 + T1_D3 * o->NP1;
  const int k1 = o->nd == 3 ? 1 : 0;
#pragma omp parallel for collapse(2)
  FOR_ALL(o) {
    if (!interior(o, i, j, k)) continue;
    double m1 = 0;
    for (int s = -1; s <= 1; s++) {
      double v = d3[enc1(o, i + 1 + s * e0, j + 1 + s * e1, k + k1 + s * e2)];
      if (v > m1) m1 = v;
    }
    WT(o, TMP4)[enc(o, i, j, k)] = m1;
    double m2 = 0;
    for (int s = -2; s <= 2; s++) {
      double v = d1[enc1(o, i + 1 + s * e0, j + 1 + s * e1, k + k1 + s * e2)];
      if (v > m2) m2 = v;
    }
    WT(o, TMP5)[enc(o, i, j, k)] = m2;
  }
}

/* hyperdifvisc4r / 4l: nu = delx*cmax*chyp*tmp4/tmp5, src/hyperdifvisc1r.cu:426-565 */
static void k_visc4(orc_t* o, int dim, int field, int out) {
  const double cmax = o->cmax_dim[dim];
#pragma omp parallel for collapse(2)
  FOR_ALL(o) {
    int ok = inr(i, 1, o->n0 - 1) && inr(j, 1, o->n1 - 1) && (o->nd == 2 || inr(k, 1, o->n2 - 1));
    if (!ok) continue;
    long c = enc(o, i, j, k);
    if (WT(o, TMP5)[c] > 0)
      WD(o, out)[c] = (WD(o, D_DELX(dim))[c]) * (cmax) * (o->p.chyp[field]) * WT(o, TMP4)[c] / WT(o, TMP5)[c];
    else
      WD(o, out)[c] = 0;
  }
}

static void k_visc_compute(orc_t* o, int s, int field, int dim);

/* cuhyperdifvisc1ir + cuhyperdifvisc1r + cuhyperdifvisc1l, with the
 * record/replay tape of the frozen-coefficient derivative checks: mode 1
 * appends every (nu_r, nu_l) pair computed, mode 2 returns the recorded
 * pairs in call order instead of computing them. */
static void k_visc(orc_t* o, int s, int field, int dim) {
  const long N = o->N;
  if (o->nu_mode == 2 && o->nu_pos + 2 * N <= o->nu_len) {
    memcpy(WD(o, D_HDNUR), o->nu_tape + o->nu_pos, sizeof(double) * N);
    memcpy(WD(o, D_HDNUL), o->nu_tape + o->nu_pos + N, sizeof(double) * N);
    o->nu_pos += 2 * N;
    return;
  }
  k_visc_compute(o, s, field, dim);
  if (o->nu_mode == 1) {
    if (o->nu_len + 2 * N > o->nu_cap) {
      long cap = o->nu_cap ? 2 * o->nu_cap : 64 * N;
      while (cap < o->nu_len + 2 * N) cap *= 2;
      o->nu_tape = (double*)realloc(o->nu_tape, sizeof(double) * cap);
      o->nu_cap = cap;
    }
    memcpy(o->nu_tape + o->nu_len, WD(o, D_HDNUR), sizeof(double) * N);
    memcpy(o->nu_tape + o->nu_len + N, WD(o, D_HDNUL), sizeof(double) * N);
    o->nu_len += 2 * N;
  }
}

static void k_visc_compute(orc_t* o, int s, int field, int dim) {
  k_visc1r(o, s, field); /* cuhyperdifvisc1ir */
  k_visc1r(o, s, field); /* cuhyperdifvisc1r */
  k_visc1a(o);
  k_visc2(o, dim, 1);
  k_visc3(o, dim);
  k_visc4(o, dim, field, D_HDNUR);
  k_visc1l(o, s, field); /* cuhyperdifvisc1l */
  k_visc1a(o);
  k_visc1bcl(o, dim);
  k_visc2(o, dim, 0);
  k_visc3(o, dim);
  k_visc4(o, dim, field, D_HDNUL);
}

/* ---------------- hyperdiffusion sources ------------------------------ */

static inline double rdx_of(orc_t* o, long c, int dim) { return WD(o, D_DELX(dim))[c]; }
static int core(orc_t* o, int i, int j, int k) { /* 1 .. n-2 in every dim */
  return inr(i, 0, o->n0 - 1) && inr(j, 0, o->n1 - 1) && (o->nd == 2 || inr(k, 0, o->n2 - 1));
}

/* cuhyperdifrhosource1, src/hyperdifrhosource1.cu:15-275 */
static void k_hrho(orc_t* o, int so, int sn, int dim, double dt) {
#pragma omp parallel for collapse(2)
  FOR_ALL(o) {
    long c = enc(o, i, j, k);
    o->dwn1[(long)F_RHO * o->N + c] = 0.0;
    WT(o, TMP1)[c] = 0.0;
    WT(o, TMP2)[c] = 0.0;
  }
#pragma omp parallel for collapse(2)
  FOR_ALL(o) if (core(o, i, j, k)) {
    long c = enc(o, i, j, k);
    WT(o, TMP1)[c] = g1r(o, W(o, so, F_RHO), F_RHO, i, j, k, dim);
    WT(o, TMP2)[c] = g1l(o, W(o, so, F_RHO), F_RHO, i, j, k, dim);
  }
  double* wn = W(o, sn, F_RHO);
  const double* nsh = WD(o, D_NUSHK(dim));
#pragma omp parallel for collapse(2)
  FOR_ALL(o) {
    long c = enc(o, i, j, k);
    wn[c] = wn[c] + dt * ((WD(o, D_HDNUR)[c] + nsh[c]) * WT(o, TMP1)[c] -
                          (WD(o, D_HDNUL)[c] + nsh[c]) * WT(o, TMP2)[c]) / rdx_of(o, c, dim);
  }
}

/* cuhyperdifesource1, src/hyperdifesource1.cu:18-554 */
static void k_he(orc_t* o, int so, int sn, int dim, double dt) {
#pragma omp parallel for collapse(2)
  FOR_ALL(o) {
    long c = enc(o, i, j, k);
    for (int f = TMP1; f <= TMP8; f++) WT(o, f)[c] = 0.0;
    o->dwn1[(long)F_E * o->N + c] = 0.0;
  }
#pragma omp parallel for collapse(2)
  FOR_ALL(o) {
    long c = enc(o, i, j, k);
    WT(o, TMP1)[c] = q_cell(o, so, c, F_E);
  }
#pragma omp parallel for collapse(2)
  FOR_ALL(o) if (core(o, i, j, k)) {
    long c = enc(o, i, j, k);
    WT(o, TMP2)[c] = g1l(o, WT(o, TMP1), TMP1, i, j, k, dim);
    WT(o, TMP3)[c] = g1r(o, WT(o, TMP1), TMP1, i, j, k, dim);
  }
  double* wn = W(o, sn, F_E);
  const double* nsh = WD(o, D_NUSHK(dim));
#pragma omp parallel for collapse(2)
  FOR_ALL(o) {
    long c = enc(o, i, j, k);
    wn[c] = wn[c] + dt * ((WD(o, D_HDNUR)[c] + nsh[c]) * WT(o, TMP3)[c] -
                          (WD(o, D_HDNUL)[c] + nsh[c]) * WT(o, TMP2)[c]) / rdx_of(o, c, dim);
  }
}

/* cuhyperdifmomsource1 (diagonal), src/hyperdifmomsource1.cu:17-535 */
static void k_hmom(orc_t* o, int so, int sn, int field, int dim, int ii0, double dt) {
  const int e0 = dim == 0, e1 = dim == 1, e2 = dim == 2;
#pragma omp parallel for collapse(2)
  FOR_ALL(o) {
    long c = enc(o, i, j, k);
    for (int f = TMP1; f <= TMP8; f++) WT(o, f)[c] = 0.0;
    o->dwn1[(long)F_E * o->N + c] = 0.0;
    o->dwn1[(long)F_MOM(ii0) * o->N + c] = 0.0;
  }
#pragma omp parallel for collapse(2)
  FOR_ALL(o) {
    long c = enc(o, i, j, k);
    WT(o, TMP4)[c] = W(o, so, F_MOM(field))[c] / (W(o, so, F_RHO)[c] + W(o, so, F_RHOB)[c]);
  }
  const double* rho = W(o, so, F_RHO);
  const double* rhob = W(o, so, F_RHOB);
  const double* m0 = W(o, so, F_MOM(ii0));
#pragma omp parallel for collapse(2)
  FOR_ALL(o) if (core(o, i, j, k)) {
    long c = enc(o, i, j, k);
    long cm = enc(o, i - e0, j - e1, k - e2), cp = enc(o, i + e0, j + e1, k + e2);
    WT(o, TMP2)[c] = (rho[c] + rhob[c] + rho[cm] + rhob[cm]) / 2;
    WT(o, TMP3)[c] = (rho[c] + rhob[c] + rho[cp] + rhob[cp]) / 2;
  }
#pragma omp parallel for collapse(2)
  FOR_ALL(o) if (core(o, i, j, k)) {
    long c = enc(o, i, j, k);
    long cm = enc(o, i - e0, j - e1, k - e2), cp = enc(o, i + e0, j + e1, k + e2);
    WT(o, TMP5)[c] = (m0[c] + m0[cm]) / 2;
    WT(o, TMP6)[c] = (m0[c] + m0[cp]) / 2;
  }
#pragma omp parallel for collapse(2)
  FOR_ALL(o) if (core(o, i, j, k)) {
    long c = enc(o, i, j, k);
    WT(o, TMP8)[c] = g1r(o, WT(o, TMP4), TMP4, i, j, k, dim);
    WT(o, TMP7)[c] = g1l(o, WT(o, TMP4), TMP4, i, j, k, dim);
  }
  const double* nur = WD(o, D_HDNUR);
  const double* nul = WD(o, D_HDNUL);
  const double* nsh = WD(o, D_NUSHK(dim));
  double* dwm = o->dwn1 + (long)F_MOM(ii0) * o->N;
  double* dwe = o->dwn1 + (long)F_E * o->N;
#pragma omp parallel for collapse(2)
  FOR_ALL(o) {
    long c = enc(o, i, j, k);
    dwm[c] = (WT(o, TMP3)[c] * (nur[c] + nsh[c]) * WT(o, TMP8)[c] -
              WT(o, TMP2)[c] * (nul[c] + nsh[c]) * WT(o, TMP7)[c]) / (rdx_of(o, c, dim)) / 2;
    W(o, sn, F_MOM(ii0))[c] = W(o, sn, F_MOM(ii0))[c] + dt * dwm[c];
  }
#pragma omp parallel for collapse(2)
  FOR_ALL(o) {
    long c = enc(o, i, j, k);
    dwe[c] = (WT(o, TMP6)[c] * (nur[c] + nsh[c]) * WT(o, TMP8)[c] -
              WT(o, TMP5)[c] * (nul[c] + nsh[c]) * WT(o, TMP7)[c]) / (rdx_of(o, c, dim)) / 2;
    W(o, sn, F_E)[c] = W(o, sn, F_E)[c] + dt * dwe[c];
  }
}

/* cuhyperdifmomsourcene1 (off-diagonal), src/hyperdifmomsourcene1.cu:16-752 */
static void k_hmomne(orc_t* o, int so, int sn, int field, int dim, int ii, int ii0, double dt) {
#pragma omp parallel for collapse(2)
  FOR_ALL(o) {
    long c = enc(o, i, j, k);
    for (int f = TMP1; f <= TMP8; f++) WT(o, f)[c] = 0.0;
    o->dwn1[(long)F_E * o->N + c] = 0.0;
    o->dwn1[(long)F_MOM(ii0) * o->N + c] = 0.0;
  }
#pragma omp parallel for collapse(2)
  FOR_ALL(o) {
    long c = enc(o, i, j, k);
    WT(o, TMP1)[c] = W(o, so, F_RHO)[c] + W(o, so, F_RHOB)[c];
    WT(o, TMP4)[c] = W(o, so, F_MOM(field))[c] / (W(o, so, F_RHO)[c] + W(o, so, F_RHOB)[c]);
  }
#pragma omp parallel for collapse(2)
  FOR_ALL(o) if (core(o, i, j, k)) WT(o, TMP5)[enc(o, i, j, k)] = (g13(o, WT(o, TMP4), TMP4, i, j, k, dim));
  const double* nsh = WD(o, D_NUSHK(dim));
#pragma omp parallel for collapse(2)
  FOR_ALL(o) {
    long c = enc(o, i, j, k);
    WT(o, TMP6)[c] = WT(o, TMP5)[c] * ((WD(o, D_HDNUR)[c] + WD(o, D_HDNUL)[c] + 2.0 * nsh[c])) / 4.0;
  }
#pragma omp parallel for collapse(2)
  FOR_ALL(o) {
    long c = enc(o, i, j, k);
    WT(o, TMP7)[c] = WT(o, TMP1)[c] * WT(o, TMP6)[c];
    WT(o, TMP8)[c] = WT(o, TMP6)[c] * W(o, so, F_MOM(ii0))[c];
  }
  double* dwm = o->dwn1 + (long)F_MOM(ii0) * o->N;
  double* dwe = o->dwn1 + (long)F_E * o->N;
#pragma omp parallel for collapse(2)
  FOR_ALL(o) dwm[enc(o, i, j, k)] = (g13(o, WT(o, TMP7), TMP7, i, j, k, ii));
#pragma omp parallel for collapse(2)
  FOR_ALL(o) {
    long c = enc(o, i, j, k);
    W(o, sn, F_MOM(ii0))[c] = W(o, sn, F_MOM(ii0))[c] + dt * dwm[c];
  }
#pragma omp parallel for collapse(2)
  FOR_ALL(o) if (core(o, i, j, k)) dwe[enc(o, i, j, k)] = (g13(o, WT(o, TMP8), TMP8, i, j, k, ii));
#pragma omp parallel for collapse(2)
  FOR_ALL(o) {
    long c = enc(o, i, j, k);
    W(o, sn, F_E)[c] = W(o, sn, F_E)[c] + dt * dwe[c];
  }
}

/* cuhyperdifbsource1 (diagonal), src/hyperdifbsource1.cu:21-441 */
static void k_hb(orc_t* o, int so, int sn, int field, int dim, int jj, int ii0, double sb, double dt) {
  const int e0 = dim == 0, e1 = dim == 1, e2 = dim == 2;
#pragma omp parallel for collapse(2)
  FOR_ALL(o) {
    long c = enc(o, i, j, k);
    for (int f = TMP1; f <= TMP8; f++) WT(o, f)[c] = 0.0;
    o->dwn1[(long)F_E * o->N + c] = 0.0;
    o->dwn1[(long)F_B(ii0) * o->N + c] = 0.0;
  }
  const double* bj = W(o, so, F_B(jj));
#pragma omp parallel for collapse(2)
  FOR_ALL(o) if (core(o, i, j, k)) {
    long c = enc(o, i, j, k);
    long cm = enc(o, i - e0, j - e1, k - e2), cp = enc(o, i + e0, j + e1, k + e2);
    WT(o, TMP2)[c] = (bj[c] + bj[cm]) / 2;
    WT(o, TMP3)[c] = (bj[c] + bj[cp]) / 2;
    WT(o, TMP1)[c] = W(o, so, F_B(field))[c];
  }
#pragma omp parallel for collapse(2)
  FOR_ALL(o) if (core(o, i, j, k)) {
    long c = enc(o, i, j, k);
    WT(o, TMP4)[c] = g1l(o, WT(o, TMP1), TMP1, i, j, k, dim);
    WT(o, TMP5)[c] = g1r(o, WT(o, TMP1), TMP1, i, j, k, dim);
  }
  const double* nur = WD(o, D_HDNUR);
  const double* nul = WD(o, D_HDNUL);
  double* dwb = o->dwn1 + (long)F_B(ii0) * o->N;
  double* dwe = o->dwn1 + (long)F_E * o->N;
#pragma omp parallel for collapse(2)
  FOR_ALL(o) {
    long c = enc(o, i, j, k);
    double rdx = rdx_of(o, c, dim);
    dwb[c] = sb * (WT(o, TMP5)[c] * nur[c] - WT(o, TMP4)[c] * nul[c]) / rdx;
    dwe[c] = sb * (WT(o, TMP3)[c] * WT(o, TMP5)[c] * nur[c] - WT(o, TMP2)[c] * WT(o, TMP4)[c] * nul[c]) / rdx;
  }
#pragma omp parallel for collapse(2)
  FOR_ALL(o) {
    long c = enc(o, i, j, k);
    W(o, sn, F_B(ii0))[c] = W(o, sn, F_B(ii0))[c] + dt * dwb[c];
    W(o, sn, F_E)[c] = W(o, sn, F_E)[c] + dt * dwe[c];
  }
}

/* cuhyperdifbsourcene1 (off-diagonal), src/hyperdifbsourcene1.cu:18-829 */
static void k_hbne(orc_t* o, int so, int sn, int field, int dim, int jj, int ii0, int mm, double sb,
                   double dt) {
#pragma omp parallel for collapse(2)
  FOR_ALL(o) {
    long c = enc(o, i, j, k);
    for (int f = TMP1; f <= TMP8; f++) WT(o, f)[c] = 0.0;
    o->dwn1[(long)F_E * o->N + c] = 0.0;
    o->dwn1[(long)F_B(ii0) * o->N + c] = 0.0;
  }
#pragma omp parallel for collapse(2)
  FOR_ALL(o) WT(o, TMP1)[enc(o, i, j, k)] = W(o, so, F_B(field))[enc(o, i, j, k)];
#pragma omp parallel for collapse(2)
  FOR_ALL(o) if (core(o, i, j, k)) WT(o, TMP2)[enc(o, i, j, k)] = g13(o, WT(o, TMP1), TMP1, i, j, k, dim);
#pragma omp parallel for collapse(2)
  FOR_ALL(o) {
    long c = enc(o, i, j, k);
    WT(o, TMP3)[c] = WT(o, TMP2)[c] * (WD(o, D_HDNUL)[c] + WD(o, D_HDNUR)[c]) / 2;
  }
#pragma omp parallel for collapse(2)
  FOR_ALL(o) if (core(o, i, j, k)) WT(o, TMP4)[enc(o, i, j, k)] = g13(o, WT(o, TMP3), TMP3, i, j, k, mm);
#pragma omp parallel for collapse(2)
  FOR_ALL(o) {
    long c = enc(o, i, j, k);
    WT(o, TMP5)[c] = WT(o, TMP3)[c] * W(o, so, F_B(jj))[c];
  }
  double* dwb = o->dwn1 + (long)F_B(ii0) * o->N;
  double* dwe = o->dwn1 + (long)F_E * o->N;
#pragma omp parallel for collapse(2)
  FOR_ALL(o) {
    long c = enc(o, i, j, k);
    dwb[c] = sb * WT(o, TMP4)[c];
    W(o, sn, F_B(ii0))[c] = W(o, sn, F_B(ii0))[c] + dt * dwb[c];
  }
  /* hyperdifbsourcene5: k < n2-1 without a lower bound (src/hyperdifbsourcene1.cu:266) */
#pragma omp parallel for collapse(2)
  FOR_ALL(o) {
    int ok = inr(i, 0, o->n0 - 1) && inr(j, 0, o->n1 - 1) && (o->nd == 2 || k < o->n2 - 1);
    if (ok) WT(o, TMP6)[enc(o, i, j, k)] = g13(o, WT(o, TMP5), TMP5, i, j, k, mm);
  }
#pragma omp parallel for collapse(2)
  FOR_ALL(o) {
    long c = enc(o, i, j, k);
    dwe[c] = sb * WT(o, TMP6)[c];
    W(o, sn, F_E)[c] = W(o, sn, F_E)[c] + dt * dwe[c];
  }
}

/* ---------------- user source plug-in --------------------------------- */

/* source_parallel -> addsourceterms2 (src/centdiff2.cu:790-872) with the
 * vertical-velocity driver; the form follows models/usersource.cu.fluxtube1:
 * m += dt*v*rho_t, e += dt*v^2*rho_t/2, positions from wd, time from qt. */
/* models/usersource.cu.fluxtube1:3-90 (3D): torsional velocity about the
 * dim-0 axis through (xc2, xc3), positions from wd, constant tdep;
 * m_1, m_2 and e updated in the reference's order and expression forms */
static void k_source_torsional(orc_t* o, int sn) {
  const double* dv = o->p.drv;
  const double aa = dv[0], tdep = dv[1], xc1 = dv[2], xc2 = dv[3], xc3 = dv[4];
  const double dz = dv[5], dx = dv[6], dy = dv[7], xxmax = dv[8], yymax = dv[9];
  const double dt = o->dt;
#pragma omp parallel for collapse(2)
  FOR_ALL(o) {
    long c = enc(o, i, j, k);
    double xp = WD(o, D_POS(1))[c] - xc2;
    double zp = WD(o, D_POS(0))[c] - xc1;
    double yp = WD(o, D_POS(2))[c] - xc3;
    double exp_z = exp(-zp * zp / (dz * dz));
    double exp_x = exp(-xp * xp / (dx * dx));
    double exp_y = exp(-yp * yp / (dy * dy));
    double exp_xyz = exp_x * exp_y * exp_z;
    double vx = (aa * yp / yymax) * exp_xyz * tdep;
    double vy = -(aa * xp / xxmax) * exp_xyz * tdep;
    W(o, sn, F_MOM(1))[c] += (dt)*vx * (W(o, sn, F_RHO)[c] + W(o, sn, F_RHOB)[c]);
    W(o, sn, F_MOM(2))[c] += (dt)*vy * (W(o, sn, F_RHO)[c] + W(o, sn, F_RHOB)[c]);
    W(o, sn, F_E)[c] += (dt) * (vx * vx + vy * vy) * (W(o, sn, F_RHO)[c] + W(o, sn, F_RHOB)[c]) / 2.0;
  }
}

/* models/usersource.cu.spicule1:4-136: only in the step whose counter
 * p->it is 1 (:67), positions from the grid origin and p->dx with the
 * index counting ghost cells (:69-70), m_0 only (the energy term is
 * commented out, :133) */
static void k_source_spicule1(orc_t* o, int sn) {
  if (o->it != 1) return;
  const double* dv = o->p.drv;
  const double AA = dv[0], tdep = dv[1], xc1 = dv[2], xc2 = dv[3], delta_z = dv[5], delta_x = dv[6];
  const double dt = o->dt;
#pragma omp parallel for collapse(2)
  FOR_ALL(o) {
    long c = enc(o, i, j, k);
    double xp = (o->p.xmin[0]) + (((double)i) * (o->p.dx[0]));
    double yp = (o->p.xmin[1]) + (((double)j) * (o->p.dx[1]));
    double r1 = (xp - xc1) * (xp - xc1);
    double r2 = (yp - xc2) * (yp - xc2);
    double exp_z = exp(-r1 / (delta_z * delta_z));
    double exp_x = exp(-r2 / (delta_x * delta_x));
    double exp_xyz = exp_x * exp_z;
    double vvz = AA * exp_xyz * tdep;
    W(o, sn, F_MOM(0))[c] = W(o, sn, F_MOM(0))[c] + (dt)*vvz * (W(o, sn, F_RHO)[c] + W(o, sn, F_RHOB)[c]);
  }
}

static void k_source(orc_t* o, int sn) {
  if (o->p.driver == SMAUG_DRIVER_TORSIONAL) {
    k_source_torsional(o, sn);
    return;
  }
  if (o->p.driver == SMAUG_DRIVER_SPICULE1) {
    k_source_spicule1(o, sn);
    return;
  }
  if (o->p.driver != SMAUG_DRIVER_VERTICAL) return;
  const double* dv = o->p.drv;
  const double A = dv[0], P = dv[1], z0 = dv[2], x0 = dv[3], y0 = dv[4], sz = dv[5], sxy = dv[6];
  const double tdep = sin(2.0 * M_PI * o->qt / P);
  const double dt = o->dt;
#pragma omp parallel for collapse(2)
  FOR_ALL(o) {
    long c = enc(o, i, j, k);
    double zp = WD(o, D_POS(0))[c] - z0;
    double xp = WD(o, D_POS(1))[c] - x0;
    double ez = exp(-zp * zp / (sz * sz));
    double ex = exp(-xp * xp / (sxy * sxy));
    double exyz = ex * ez;
    if (o->nd == 3) {
      double yp = WD(o, D_POS(2))[c] - y0;
      double ey = exp(-yp * yp / (sxy * sxy));
      exyz = ex * ey * ez;
    }
    double vz = A * exyz * tdep;
    /* amplitude map: one amplitude per (j,k) column in place of
     * A*exp(-dx^2/sxy^2 [- dy^2/sxy^2]) */
    if (o->amap) vz = o->amap[(long)k * o->n1 + j] * ez * tdep;
    double rho_t = W(o, sn, F_RHO)[c] + W(o, sn, F_RHOB)[c];
    W(o, sn, F_MOM(0))[c] += (dt)*vz * rho_t;
    W(o, sn, F_E)[c] += (dt) * (vz * vz) * rho_t / 2.0;
  }
}

/* ---------------- boundaries ------------------------------------------ */

/* One boundary pass for field f along dir, in place, as the reference's
 * per-direction helpers do it: periodic (src/boundary.cu:72-225, 2D),
 * contcd4 (models/boundary.cu.fluxtube1:16-85,168-238,323-381), cont
 * (bc3_cont_dir, src/gradops.cu:947-1060), fixed (bc3_fixed_dir,
 * src/gradops.cu:1234-1268: the values the initialisation pass of
 * src/iosmaug.c:416-425 means to capture, i.e. the initial state's),
 * symm/asymm (bc3_symm_dir/bc3_asymm_dir, src/gradops.cu:1325-1466, whose
 * 2D and 3D maps differ). */
static void k_bound(orc_t* o, int s, int f, int dir) {
  double* a = W(o, s, f);
  const int bt = o->p.boundtype[f][dir][0];
  const int nn = nof(o, dir);
  const long st = stride(o, dir);
  FOR_ALL(o) {
    const int x = coord(i, j, k, dir);
    const long c = enc(o, i, j, k);
    if (x > 1 && x < nn - 2) continue;
    if (bt == 0) {
      if (x == 0 || x == 1) a[c] = a[c + (nn - 4) * st];
      else a[c] = a[c - (nn - 4) * st];
    } else if (bt == 4) {
      if (x == 0 || x == 1) a[c] = a[c + (4 - 2 * x) * st];
      else if (x == nn - 1) a[c] = a[c - 4 * st];
      else a[c] = a[c - 2 * st];
    } else if (bt == 3) {
      if (x == 0 || x == 1) a[c] = a[c + (2 - x) * st];
      else a[c] = a[c - (x - (nn - 3)) * st];
    } else if (bt == 5) {
      a[c] = o->fixed[(long)f * o->N + c];
    } else if (bt == 6 || bt == 7) {
      const double sg = bt == 7 ? -1.0 : 1.0;
      long src;
      if (o->nd == 3) src = (x == 0) ? 4 : (x == 1) ? 3 : (x == nn - 1) ? nn - 5 : nn - 4;
      else src = (x == 0) ? 3 : (x == 1) ? 2 : (x == nn - 1) ? nn - 4 : nn - 3;
      a[c] = sg * a[c + (src - x) * st];
    }
  }
}

static void k_bound_all(orc_t* o, int s) {
  for (int f = 0; f <= F_B(o->nd - 1); f++)
    for (int dir = 0; dir < o->nd; dir++) k_bound(o, s, f, dir);
}

/* ---------------- grid (src/init.cu:2502-3051, Q11) ------------------- */

static void initgrid(orc_t* o, const double* pos) {
  const int nd = o->nd;
  const long NP2 = o->NP2;
  double* tt = (double*)calloc((size_t)3 * NP2, sizeof(double));
  /* padded interior <- positions */
  for (int d = 0; d < nd; d++)
    FOR_ALL(o) tt[d * NP2 + enc2(o, i + 1, j + 1, nd == 3 ? k + 1 : 0)] = pos[(long)d * o->N + enc(o, i, j, k)];
  const int P0 = o->n0 + 2, P1 = o->n1 + 2, P2 = nd == 3 ? o->n2 + 2 : 1;
  for (int dir = 0; dir < nd; dir++) {
    const int nn = nof(o, dir);
    for (int side = 0; side < 2; side++) {
      int ixmin, ixmax, ixe, ixf;
      if (side == 0) {
        ixmax = nn + 1;
        ixmin = nn - 1;
        ixe = ixmin - 1;
        ixf = ixe - 1;
      } else {
        ixmin = 0;
        ixmax = dir == 0 ? 1 : 2;
        ixe = ixmax + 1;
        ixf = ixe + 1;
      }
      for (int d1 = 0; d1 < nd; d1++) {
        double* T = tt + d1 * NP2;
        for (int c2 = 0; c2 < P2; c2++)
          for (int c1 = 0; c1 < P1; c1++)
            for (int c0 = 0; c0 < P0; c0++) {
              int cc[3] = {c0, c1, c2};
              int ix = cc[dir];
              if (ix < ixmin || ix > ixmax) continue;
              int e1[3] = {c0, c1, c2}, f1[3] = {c0, c1, c2};
              e1[dir] = ixe;
              f1[dir] = ixf;
              double a = (double)(1 + abs(ixe - ix)), b = (double)abs(ixe - ix);
              T[enc2(o, c0, c1, c2)] = a * (T[enc2(o, e1[0], e1[1], e1[2])]) - b * (T[enc2(o, f1[0], f1[1], f1[2])]);
            }
      }
    }
  }
  for (int d = 0; d < nd; d++) {
    const int e0 = d == 0, e1 = d == 1, e2 = d == 2;
    const double* T = tt + d * NP2;
    FOR_ALL(o) {
      long c = enc(o, i, j, k);
      int kp = nd == 3 ? k + 1 : 0;
      WD(o, D_POS(d))[c] = T[enc2(o, i + 1, j + 1, kp)];
      WD(o, D_DELX(d))[c] = 0.5 * (T[enc2(o, i + 1 + e0, j + 1 + e1, kp + e2)] - T[enc2(o, i + 1 - e0, j + 1 - e1, kp - e2)]);
    }
  }
  free(tt);
}

/* ---------------- step (src/iosmaug.c:496-1144) ----------------------- */

static void copy_pert(orc_t* o, int dst, int src) {
  for (int f = 0; f <= F_B(o->nd - 1); f++) memcpy(W(o, dst, f), W(o, src, f), sizeof(double) * o->N);
}

/* computedervfields_parallel with order 0: zero wd[vel1..pkb] */
static void zero_derived(orc_t* o) {
  for (int f = D_VEL(0); f <= D_PKB; f++) memset(WD(o, f), 0, sizeof(double) * o->N);
}

/* one evaluation of the loop body: fluxes, gravity, hyperdiffusion,
 * source and boundaries, reading slot so and accumulating into slot sn */
static void stage(orc_t* o, int so, int sn, double dt) {
  const int nd = o->nd;
  const unsigned T = o->p.terms;
  o->dt = dt;
  zero_derived(o);
  for (int dir = 0; dir < nd; dir++) k_vels(o, so, dir);
  for (int dir = 0; dir < nd; dir++) {
    for (int f = F_RHO; f <= F_MOM(nd - 1); f++) {
      if (f == F_MOM(dir)) {
        k_pt(o, so, dir);
        k_pbg(o, so);
      }
      if (T & SMAUG_T_FLUX) k_centdiff1(o, so, sn, dt, f, dir);
    }
    for (int f = F_E; f <= F_B(nd - 1); f++) {
      if (f == F_E) {
        k_vels(o, so, dir);
        k_pbg(o, so);
        k_pt(o, so, dir);
      }
      if (T & (SMAUG_T_FLUX | SMAUG_T_ETERM)) {
        if (T & SMAUG_T_FLUX) {
          k_centdiff2(o, so, sn, dt, f, dir);
        } else {
          /* energy terms without the flux divergence */
          k_pbg(o, sn);
          if (f == F_E)
            FOR_ALL(o) {
              int ok = nd == 3 ? interior(o, i, j, k) : (i < o->n0 - 2 && j < o->n1 - 2);
              if (ok) eterms_cell(o, sn, i, j, k, dir);
            }
        }
      }
    }
  }
  if (T & SMAUG_T_GRAV) k_grav(o, so, sn);
  k_divb(o, so, sn, dt); /* src/iosmaug.c:678-679 */
  if (o->p.hyperdifmom == 1) {
    for (int dim = 0; dim < nd; dim++) {
      o->cmax_dim[dim] = k_cmax(o, so, dim);
      if (T & SMAUG_T_HRHO) {
        k_visc(o, so, F_RHO, dim);
        k_hrho(o, so, sn, dim, dt);
      }
    }
    if (T & SMAUG_T_HE)
      for (int dim = 0; dim < nd; dim++) {
        k_visc(o, so, F_E, dim);
        k_he(o, so, sn, dim, dt);
      }
    if (T & SMAUG_T_HMOM)
      for (int dim = 0; dim < nd; dim++)
        for (int f = 0; f < nd; f++) {
          k_visc(o, so, F_MOM(f), dim);
          for (int ii1 = 0; ii1 <= 1; ii1++) {
            int ii, ii0;
            if (ii1 == 0) {
              ii = dim;
              ii0 = f;
            } else {
              ii = f;
              ii0 = dim;
            }
            if (ii == dim) k_hmom(o, so, sn, f, dim, ii0, dt);
            else k_hmomne(o, so, sn, f, dim, ii, ii0, dt);
          }
        }
    if (T & SMAUG_T_HB)
      for (int dim = 0; dim < nd; dim++)
        for (int f = 0; f < nd; f++) {
          if (f == dim) continue;
          k_visc(o, so, F_B(f), dim);
          for (int ii1 = 0; ii1 <= 1; ii1++) {
            int jj, mm, ii0;
            double sb;
            if (ii1 == 0) {
              jj = dim;
              mm = f;
              sb = -1.0;
              ii0 = dim;
            } else {
              ii0 = f;
              mm = dim;
              sb = 1.0;
              jj = f;
            }
            if (mm == dim) k_hb(o, so, sn, f, dim, jj, ii0, sb, dt);
            else k_hbne(o, so, sn, f, dim, jj, ii0, mm, sb, dt);
          }
        }
  }
  if (T & SMAUG_T_SOURCE) k_source(o, sn);
}

/* moddton block, src/iosmaug.c:541-616 (Q4): the step restarts from DT
 * (dt=DT; p->dt=DT, :549-550) and takes the CFL step when it lies in
 * (SMALLDOUBLE, DT) (:568-569) */
static void cfl(orc_t* o, int s) {
  double cm = 0;
  for (int dim = 0; dim < o->nd; dim++) cm = k_cmax(o, s, dim);
  double courantmax = SMALLD;
  for (int dim = 0; dim < o->nd; dim++) {
    double t = cm / (o->p.dx[dim]);
    if (t > courantmax) courantmax = t;
  }
  o->dt = o->p.dt;
  double dtt = (o->p.courant) / courantmax;
  if (dtt > SMALLD && dtt < o->p.dt) o->dt = dtt;
}

int orc_step(orc_t* o, int nsteps) {
  for (int n = 0; n < nsteps; n++) {
    const int nd = o->nd;
    if (o->p.rkon == 0) {
      copy_pert(o, 0, 1); /* computedervfields order 0 */
      if (o->p.moddton == 1) cfl(o, 0);
      o->qt = o->qt + o->p.dt;
      stage(o, 0, 1, o->dt);
      k_bound_all(o, 1);
      o->cur = 1;
    } else {
      /* intended four-step RK (src/advance.cu:76-99) on top of the stage
       * operator: slots 0=w, 1=w1, 2=w2, 3=w3, 4=new */
      if (o->p.moddton == 1) cfl(o, 0);
      o->qt = o->qt + o->p.dt;
      const double h = o->dt;
      for (int s = 1; s <= 4; s++) memcpy(W(o, s, 0), W(o, 0, 0), sizeof(double) * o->nvar * o->N);
      stage(o, 0, 1, h / 2.0);
      k_bound_all(o, 1);
      stage(o, 1, 2, h / 2.0);
      k_bound_all(o, 2);
      stage(o, 2, 3, h);
      k_bound_all(o, 3);
      stage(o, 3, 4, h / 6.0);
      for (int f = 0; f <= F_B(nd - 1); f++) {
        double *w0 = W(o, 0, f), *w1 = W(o, 1, f), *w2 = W(o, 2, f), *w3 = W(o, 3, f), *w4 = W(o, 4, f);
        for (long c = 0; c < o->N; c++) {
          double comb = (w1[c] + 2.0 * w2[c] + w3[c] - 4.0 * w0[c]) / 3;
          w4[c] = w4[c] + comb;
        }
      }
      k_bound_all(o, 4);
      copy_pert(o, 0, 4);
      o->dt = h;
      o->cur = 0;
    }
    o->it = o->it + 1;
    o->t = o->t + o->dt;
  }
  return 0;
}

/* ---------------- public API ------------------------------------------ */

int orc_create(const smaug_params* p, const double* w0, const double* pos, orc_t** out) {
  if (!p || !w0 || !pos || !out) return -1;
  if (p->ndim != 2 && p->ndim != 3) return -1;
  orc_t* o = (orc_t*)calloc(1, sizeof(orc_t));
  o->p = *p;
  o->nd = p->ndim;
  o->n0 = p->n[0];
  o->n1 = p->n[1];
  o->n2 = p->ndim == 3 ? p->n[2] : 1;
  if (o->n0 < 8 || o->n1 < 8 || (o->nd == 3 && o->n2 < 8)) {
    free(o);
    return -1;
  }
  for (int f = 0; f < SMAUG_MAXVAR; f++)
    for (int d = 0; d < o->nd; d++) {
      int bt = p->boundtype[f][d][0];
      if (bt == 1 || bt == 2 || bt < 0 || bt > 7 || (bt == 0 && o->nd == 3)) {
        if (f < 2 + 2 * o->nd) {
          free(o);
          return -5;
        }
      }
    }
  o->N = (long)o->n0 * o->n1 * o->n2;
  o->NP1 = (long)(o->n0 + 1) * (o->n1 + 1) * (o->nd == 3 ? o->n2 + 1 : 1);
  o->NP2 = (long)(o->n0 + 2) * (o->n1 + 2) * (o->nd == 3 ? o->n2 + 2 : 1);
  o->nvar = 4 + 3 * o->nd;
  o->nderv = 4 * o->nd + 11;
  o->nslot = 5;
  o->w = (double*)calloc((size_t)o->nslot * o->nvar * o->N, sizeof(double));
  o->wd = (double*)calloc((size_t)o->nderv * o->N, sizeof(double));
  o->wt = (double*)calloc((size_t)8 * o->N, sizeof(double));
  o->wt1 = (double*)calloc((size_t)2 * o->NP1, sizeof(double));
  o->wt2 = (double*)calloc((size_t)o->NP2, sizeof(double));
  o->dwn1 = (double*)calloc((size_t)o->nvar * o->N, sizeof(double));
  if (!o->w || !o->wd || !o->wt || !o->wt1 || !o->wt2 || !o->dwn1) {
    orc_destroy(o);
    return -3;
  }
  for (int s = 0; s < o->nslot; s++) memcpy(W(o, s, 0), w0, sizeof(double) * o->nvar * o->N);
  o->fixed = (double*)malloc(sizeof(double) * (2 + 2 * o->nd) * o->N);
  memcpy(o->fixed, w0, sizeof(double) * (2 + 2 * o->nd) * o->N);
  initgrid(o, pos);
  o->dt = p->moddton == 1 ? 1.0e-8 : p->dt; /* src/iosmaug.c:105,333-336 */
  o->qt = p->qt0;
  o->it = p->it0;
  o->t = 0;
  if (p->rkon == 0) {
    k_bound_all(o, 1);
    o->cur = 1;
  } else {
    k_bound_all(o, 0);
    o->cur = 0;
  }
  *out = o;
  return 0;
}

int orc_get_state(orc_t* o, double* w) {
  memcpy(w, W(o, o->cur, 0), sizeof(double) * o->nvar * o->N);
  return 0;
}

int orc_set_state(orc_t* o, const double* w) {
  memcpy(W(o, o->cur, 0), w, sizeof(double) * o->nvar * o->N);
  memcpy(o->fixed, w, sizeof(double) * (2 + 2 * o->nd) * o->N);
  return 0;
}

int orc_get_wd(orc_t* o, double* wd) {
  memcpy(wd, o->wd, sizeof(double) * o->nderv * o->N);
  return 0;
}

int orc_cfl(orc_t* o, double* cmax_per_dim, double* dt) {
  for (int d = 0; d < o->nd; d++) cmax_per_dim[d] = k_cmax(o, o->cur, d);
  double saved = o->dt;
  cfl(o, o->cur);
  if (dt) *dt = o->dt;
  o->dt = saved;
  return 0;
}

int orc_get_clock(orc_t* o, int* it, double* t, double* qt, double* dt) {
  if (it) *it = o->it;
  if (t) *t = o->t;
  if (qt) *qt = o->qt;
  if (dt) *dt = o->dt;
  return 0;
}

/* ---------------- probes for the known-answer tests ------------------- */

/* One differential operator applied to an arbitrary array at every cell:
 * kind 0 grad3dn, 1 grad1l3n, 2 grad1r3n, 3 grad13n; bcidx selects the
 * boundtype row that drives the boundary zeroing. */
int orc_probe_grad(orc_t* o, int kind, const double* a, int bcidx, int dir, double* out) {
  if (kind < 0 || kind > 3 || dir < 0 || dir >= o->nd || bcidx < 0 || bcidx >= SMAUG_MAXVAR) return -1;
  FOR_ALL(o) {
    const long c = enc(o, i, j, k);
    double g;
    if (kind == 0) g = g3dn(o, a, bcidx, i, j, k, dir);
    else if (kind == 1) g = g1l(o, a, bcidx, i, j, k, dir);
    else if (kind == 2) g = g1r(o, a, bcidx, i, j, k, dir);
    else g = g13(o, a, bcidx, i, j, k, dir);
    out[c] = g;
  }
  return 0;
}

/* nu_{i+1/2}, nu_{i-1/2} of one field of the current state along dim, with
 * cmax of that dim, as the hyperdiffusion sources receive them. */
int orc_probe_visc(orc_t* o, int field, int dim, double* nur, double* nul, double* cmax) {
  if (field < 0 || field > F_B(o->nd - 1) || dim < 0 || dim >= o->nd) return -1;
  o->cmax_dim[dim] = k_cmax(o, o->cur, dim);
  k_visc(o, o->cur, field, dim);
  memcpy(nur, WD(o, D_HDNUR), sizeof(double) * o->N);
  memcpy(nul, WD(o, D_HDNUL), sizeof(double) * o->N);
  if (cmax) *cmax = o->cmax_dim[dim];
  return 0;
}

/* restart from state w (nvar*N) at step it and clock qt, as orc_create */
int orc_restart(orc_t* o, const double* w, int it, double qt) {
  for (int s = 0; s < o->nslot; s++) memcpy(W(o, s, 0), w, sizeof(double) * o->nvar * o->N);
  memcpy(o->fixed, w, sizeof(double) * (2 + 2 * o->nd) * o->N);
  o->dt = o->p.moddton == 1 ? 1.0e-8 : o->p.dt;
  o->qt = qt;
  o->it = it;
  o->t = 0;
  o->cur = o->p.rkon == 0 ? 1 : 0;
  k_bound_all(o, o->cur);
  return 0;
}

/* driver amplitude map (n1*n2 values, column (j,k) at k*n1+j), NULL clears */
int orc_set_driver_map(orc_t* o, const double* amap) {
  const long nc = (long)o->n1 * o->n2;
  if (!amap) {
    free(o->amap);
    o->amap = NULL;
    return 0;
  }
  if (!o->amap) o->amap = (double*)malloc(sizeof(double) * nc);
  memcpy(o->amap, amap, sizeof(double) * nc);
  return 0;
}

/* driver parameter drv[q] for the steps that follow (the driver sources
 * read p.drv at every step); differences in the driver amplitude */
int orc_set_driver_param(orc_t* o, int q, double v) {
  if (q < 0 || q >= SMAUG_NDRV) return -1;
  o->p.drv[q] = v;
  return 0;
}

/* nu tape: mode 0 off, 1 record (clears the tape), 2 replay from the start.
 * Returns the number of recorded (nu_r, nu_l) pairs. */
long orc_nu_mode(orc_t* o, int mode) {
  if (mode == 1) o->nu_len = 0;
  o->nu_pos = 0;
  o->nu_mode = mode;
  return o->nu_len / (2 * o->N);
}

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void orc_destroy(orc_t* o) {
  if (!o) return;
  free(o->w);
  free(o->wd);
  free(o->wt);
  free(o->wt1);
  free(o->wt2);
  free(o->dwn1);
  free(o->amap);
  free(o->fixed);
  free(o->nu_tape);
  free(o);
}
