// smaug_adj.hip -- adjoint (reverse-mode derivative) of one SMAUG stage.
//
// The stage kernel k_rhs (smaug_rhs.hip) computes, for every interior cell c,
//   acc(c) = flux divergences + background energy terms + gravity +
//            hyperdiffusion sources of the stage input win,
//   out(c) = base(c) + h acc(c) + driver source(out_rho(c)) [+ RK combination],
// and copies out to the ghost cells that are images of c.  The adjoint of a
// stage is three launches:
//   k_adj_mu   (pointwise): mu = h * (cotangent of acc), the cotangents of
//              base and of the RK combination inputs, and the derivative of
//              the loss with respect to the driver amplitude, per cell;
//   k_adj_rhs  (gather, one thread per cell p): T(p) = (d acc/d win)^T mu at
//              p, the sum over every interior cell c whose stencil reads p of
//              mu(c) times the partial derivative of acc(c) by win(p).  The
//              hyperdiffusion coefficients nu_{i+-1/2} (windowed maxima) and
//              the wave-speed maxima they scale with are held at their values
//              for win: the derivative is that of the stage with those
//              coefficients frozen;
//   k_adj_fold (pointwise): the cotangent of a ghost cell is added to the
//              interior cell it is an image of.
// Every block below transposes one block of k_rhs, which restates
// src/centdiff1.cu, src/centdiff2.cu and the src/hyperdif*source*.cu
// kernels; the forward expression is quoted beside each transpose.
#include <math.h>

#include "../../include/smaug_hip.h"
#include "smaug_dev.h"
#include "smaug_ops.h"

namespace smaug {

namespace {

__device__ inline int axis_off(const Geo& G, int d) { return d == 0 ? 0 : (d == 1 ? G.n[0] : G.n[0] + G.n[1]); }

// 1/delx and delx along d at cell y
__device__ inline double rdx_at(const Geo& G, const AdjArgs& A, int d, const int* y) {
  return A.rdx1 ? A.rdx1[axis_off(G, d) + y[d]] : A.rdelx[(long long)d * G.N + gix(G, y[0], y[1], y[2])];
}
__device__ inline double dx_at(const Geo& G, const AdjArgs& A, int d, const int* y) {
  return A.dx1 ? A.dx1[axis_off(G, d) + y[d]] : A.delx[(long long)d * G.N + gix(G, y[0], y[1], y[2])];
}

template <int ND>
struct Nb;

// nu_{i+-1/2} at the cells x + k e_d, k = -1, 0, 1, that are interior
// (ok[k+1]), from one 9-value window of q along d (the three 7-value
// windows overlap); r = 1/delx along d at those cells
template <int ND>
__device__ inline void nu3_at(const Geo& G, const AdjArgs& A, const double* __restrict__ qa, const int* x, int d,
                              double scale, bool* ok, double* nr, double* nl, double* r);

template <int ND>
struct Nb {  // cell y = x + k e_d (and optionally + k2 e_d2)
  int y[3];
  __device__ Nb(const int* x, int d, int k) {
    y[0] = x[0];
    y[1] = x[1];
    y[2] = x[2];
    y[d] += k;
  }
  __device__ Nb(const int* x, int d, int k, int d2, int k2) {
    y[0] = x[0];
    y[1] = x[1];
    y[2] = x[2];
    y[d] += k;
    y[d2] += k2;
  }
};

template <int ND>
__device__ inline void nu3_at(const Geo& G, const AdjArgs& A, const double* __restrict__ qa, const int* x, int d,
                              double scale, bool* ok, double* nr, double* nl, double* r) {
  double Q[9];
#pragma unroll
  for (int o = 0; o < 9; o++) {
    int z[3] = {x[0], x[1], x[2]};
    z[d] += o - 4;
    Q[o] = qval<ND, double>(G, qa, z);
  }
#pragma unroll
  for (int k = -1; k <= 1; k++) {
    Nb<ND> c(x, d, k);
    ok[k + 1] = interior_c<ND>(G, c.y);
    nr[k + 1] = nl[k + 1] = r[k + 1] = 0.0;
    if (ok[k + 1]) {
      nu7(&Q[k + 1], dx_at(G, A, d, c.y) * scale, nr[k + 1], nl[k + 1]);
      r[k + 1] = rdx_at(G, A, d, c.y);
    }
  }
}

// (dF^d/du)^T Gf at one cell, added to lam.  F^d as in fluxes2:
//   F_rho  = v_d rho + rhob v_d                       (= m_d)
//   F_m_f  = v_d m_f + [f=d] p_t - (b_f bb_d + bb_f b_d) - b_f b_d
//   F_e    = v_d e + v_d p_t - b_d bb.v - bb_d b.v + v_d e_b - b_d b.v
//   F_b_f  = v_d b_f - (b_d + bb_d) v_f + bb_f v_d   (f != d; 0 for f = d)
//   v = m/rho_t,  p_t = (g-1)(e - m.m/2rho_t) - (g-2)(b.bb + b.b/2)
template <int ND>
__device__ inline void flux_vjp(const Geo& G, const double* u, const double* bgc, int d, const double* Gf,
                                double* lam) {
  const double gm = G.gamma;
  const double rho = u[fRHO()], e = u[fE(ND)], eb = bgc[bEB()], rhob = bgc[bRHOB()];
  const double rt = rho + rhob;
  double m[ND], b[ND], bb[ND], v[ND];
#pragma unroll
  for (int q = 0; q < ND; q++) {
    m[q] = u[fMOM(q)];
    b[q] = u[fB(ND, q)];
    bb[q] = bgc[bBB(q)];
    v[q] = m[q] / rt;
  }
  double mm = 0, b2 = 0, bbb = 0, bbv = 0, bv = 0;
#pragma unroll
  for (int q = 0; q < ND; q++) {
    mm += m[q] * m[q];
    b2 += b[q] * b[q];
    bbb += b[q] * bb[q];
    bbv += bb[q] * v[q];
    bv += b[q] * v[q];
  }
  const double pt = (gm - 1.0) * (e - 0.5 * mm / rt) - (gm - 2.0) * (bbb + 0.5 * b2);
  double vb[ND], mb[ND], bbar[ND];
#pragma unroll
  for (int q = 0; q < ND; q++) vb[q] = mb[q] = bbar[q] = 0.0;
  double rtb = 0, ebar = 0, ptb = 0, bbvb = 0, bvb = 0;
  mb[d] += Gf[fRHO()];  // dF_rho/dm_d = 1, dF_rho/drho = 0
#pragma unroll
  for (int f = 0; f < ND; f++) {
    const double g = Gf[fMOM(f)];
    vb[d] += g * m[f];
    mb[f] += g * v[d];
    if (f == d) ptb += g;
    bbar[f] -= g * (bb[d] + b[d]);
    bbar[d] -= g * (bb[f] + b[f]);
  }
  {
    const double g = Gf[fE(ND)];
    vb[d] += g * (e + pt + eb);
    ebar += g * v[d];
    ptb += g * v[d];
    bbar[d] -= g * (bbv + bv);
    bbvb -= g * b[d];
    bvb -= g * (bb[d] + b[d]);
  }
#pragma unroll
  for (int f = 0; f < ND; f++) {
    if (f == d) continue;
    const double g = Gf[fB(ND, f)];
    vb[d] += g * (b[f] + bb[f]);
    bbar[f] += g * v[d];
    bbar[d] -= g * v[f];
    vb[f] -= g * (b[d] + bb[d]);
  }
#pragma unroll
  for (int q = 0; q < ND; q++) {
    vb[q] += bbvb * bb[q] + bvb * b[q];
    bbar[q] += bvb * v[q];
  }
  ebar += ptb * (gm - 1.0);
  rtb += ptb * (gm - 1.0) * 0.5 * mm / (rt * rt);
#pragma unroll
  for (int q = 0; q < ND; q++) {
    mb[q] -= ptb * (gm - 1.0) * m[q] / rt;
    bbar[q] -= ptb * (gm - 2.0) * (bb[q] + b[q]);
    mb[q] += vb[q] / rt;
    rtb -= vb[q] * v[q] / rt;
  }
  lam[fRHO()] += rtb;
  lam[fE(ND)] += ebar;
#pragma unroll
  for (int q = 0; q < ND; q++) {
    lam[fMOM(q)] += mb[q];
    lam[fB(ND, q)] += bbar[q];
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// k_adj_mu: cotangents through the stage update, the driver source and the
// RK combination (pointwise; see AdjMuArgs).
//   out = base + h acc;  out_m0 += h vz (out_rho + rhob);
//   out_e += h vz^2 (out_rho + rhob)/2;  [out += (r1 + 2 r2 + r3 - 4 base)/3]
template <int ND>
__global__ __launch_bounds__(256) void k_adj_mu(Geo G, AdjMuArgs A) {
  constexpr int NP = 2 + 2 * ND;
  const long long c = (long long)blockIdx.x * 256 + threadIdx.x;
  if (c >= G.N) return;
  const long long N = G.N;
  int x[3];
  x[0] = (int)(c % G.n[0]);
  x[1] = (int)((c / G.n[0]) % G.n[1]);
  x[2] = ND == 3 ? (int)(c / ((long long)G.n[0] * G.n[1])) : 0;
  const bool in = interior_c<ND>(G, x);
  double lo[NP], lp[NP];
#pragma unroll
  for (int f = 0; f < NP; f++) lo[f] = in ? A.lam_out[(long long)f * N + c] : 0.0;
#pragma unroll
  for (int f = 0; f < NP; f++) lp[f] = lo[f];
  const double h = A.hfrac * A.meta[0];
  const double tdep = A.meta[2];
  double gc = 0.0;
  if (in && (G.terms & SMAUG_T_SOURCE) && G.driver == SMAUG_DRIVER_VERTICAL) {
    const double* dv = G.drv;
    const double zp = A.pos[c] - dv[2];
    const double ez = exp(-zp * zp / (dv[5] * dv[5]));
    const double xp = A.pos[N + c] - dv[3];
    double exyz = exp(-xp * xp / (dv[6] * dv[6])) * ez;
    if (ND == 3) {
      const double yp = A.pos[2 * N + c] - dv[4];
      exyz = exyz * exp(-yp * yp / (dv[6] * dv[6]));
    }
    // vz as the stage kernel forms it, and d vz / d amplitude
    const double vz = A.amap ? A.amap[x[1] + (ND == 3 ? (long long)G.n[1] * x[2] : 0)] * ez * tdep
                             : dv[0] * exyz * tdep;
    const double shape = A.amap ? ez * tdep : exyz * tdep;
    double orho = A.outr[c];
    if (A.c0) orho -= (A.c1[c] + 2.0 * A.c2[c] + A.c3[c] - 4.0 * A.c0[c]) / 3;
    const double rto = orho + A.bg[(long long)bRHOB() * N + c];
    lp[fRHO()] = lo[fRHO()] + h * vz * lo[fMOM(0)] + h * vz * vz * 0.5 * lo[fE(ND)];
    gc = (lo[fMOM(0)] * h * rto + lo[fE(ND)] * h * vz * rto) * shape;
  }
  // torsional driver: m_1 += h v_1 rho_t, m_2 += h v_2 rho_t,
  // e += h (v_1^2 + v_2^2) rho_t / 2, rho_t of the updated value
  if (ND == 3 && in && (G.terms & SMAUG_T_SOURCE) && G.driver == SMAUG_DRIVER_TORSIONAL) {
    const double* dv = G.drv;
    const double xp = A.pos[N + c] - dv[3], zp = A.pos[c] - dv[2], yp = A.pos[2 * N + c] - dv[4];
    const double exyz = exp(-xp * xp / (dv[6] * dv[6])) * exp(-yp * yp / (dv[7] * dv[7])) * exp(-zp * zp / (dv[5] * dv[5]));
    const double vx = (dv[0] * yp / dv[9]) * exyz * tdep, vy = -(dv[0] * xp / dv[8]) * exyz * tdep;
    lp[fRHO()] = lo[fRHO()] + h * vx * lo[fMOM(1)] + h * vy * lo[fMOM(ND - 1)] + h * (vx * vx + vy * vy) * 0.5 * lo[fE(ND)];
  }
  // spicule1: m_0 += h v rho_t in the step whose counter is 1 (tdep != 0)
  if (in && (G.terms & SMAUG_T_SOURCE) && G.driver == SMAUG_DRIVER_SPICULE1 && tdep != 0.0) {
    const double* dv = G.drv;
    const double xp = G.pxmin[0] + (double)x[0] * G.pdx[0], yp = G.pxmin[1] + (double)x[1] * G.pdx[1];
    const double r1 = (xp - dv[2]) * (xp - dv[2]), r2 = (yp - dv[3]) * (yp - dv[3]);
    const double vvz = dv[0] * (exp(-r2 / (dv[6] * dv[6])) * exp(-r1 / (dv[5] * dv[5]))) * tdep;
    lp[fRHO()] = lo[fRHO()] + h * vvz * lo[fMOM(0)];
  }
#pragma unroll
  for (int f = 0; f < NP; f++) {
    const long long q = (long long)f * N + c;
    A.mu[q] = h * lp[f];
    const double lb = lp[f] + A.cb * lo[f];
    A.lam_base[q] = A.base_set ? lb : A.lam_base[q] + lb;
    if (A.lam_r1) {
      A.lam_r1[q] = lo[f] / 3.0;
      A.lam_r2[q] = 2.0 * lo[f] / 3.0;
      A.lam_r3[q] = lo[f] / 3.0;
    }
  }
  if (A.gcell) A.gcell[c] = A.gcell_set ? gc : A.gcell[c] + gc;
}

// ---------------------------------------------------------------------------
// k_adj_rhs: T(p) = (d acc / d win)^T mu, one thread per cell p, in several
// launches that each keep a modest register footprint:
//   PART 0          fluxes (T = ...)
//   PART 1          energy terms and gravity (T += ...)
//   PART 4, FSEL 0  rho hyperdiffusion, FSEL 1 e hyperdiffusion (T += ...)
//   PART 2, FSEL f  momentum hyperdiffusion of v_f (T += ...)
//   PART 3, FSEL f  magnetic hyperdiffusion of b_f (T += ...)
// Each part chains its own cotangents of v and q_e back to (rho, m, e, b) at
// p, which is exact because the chain is linear.
template <int ND, int PART, int FSEL>
__global__ __launch_bounds__(256) void k_adj_rhs(Geo G, AdjArgs A) {
  constexpr int NP = 2 + 2 * ND;
  const long long p = (long long)blockIdx.x * 256 + threadIdx.x;
  if (p >= G.N) return;
  const long long N = G.N;
  int x[3];
  x[0] = (int)(p % G.n[0]);
  x[1] = (int)((p / G.n[0]) % G.n[1]);
  x[2] = ND == 3 ? (int)(p / ((long long)G.n[0] * G.n[1])) : 0;
  double lam[NP];
#pragma unroll
  for (int f = 0; f < NP; f++) lam[f] = 0.0;
  int ng = 0;
#pragma unroll
  for (int d = 0; d < ND; d++) ng += (x[d] < 2 || x[d] > G.n[d] - 3) ? 1 : 0;
  // cells with two ghost coordinates are read by no stencil of an interior cell
  if (ng > 1) {
    if (PART == 0)
#pragma unroll
      for (int f = 0; f < NP; f++) A.out[(long long)f * N + p] = 0.0;
    return;
  }
  const double* __restrict__ w = A.win;
  const double* __restrict__ bg = A.bg;
  const double* __restrict__ mu = A.mu;
  const unsigned T = G.terms;
  auto at = [&](const int* y) { return gix(G, y[0], y[1], y[2]); };
  auto RT = [&](const int* y) -> double {
    const long long q = at(y);
    return w[q] + bg[(long long)bRHOB() * N + q];
  };
  auto WV = [&](int f, const int* y) -> double { return w[(long long)f * N + at(y)]; };
  double cm[ND];
#pragma unroll
  for (int d = 0; d < ND; d++) cm[d] = __longlong_as_double((long long)A.cmax_bits[d]);

  double vbar[ND], rtbar = 0.0, qebar = 0.0;
#pragma unroll
  for (int d = 0; d < ND; d++) vbar[d] = 0.0;

  // ---------------------------------------------------------- fluxes
  // acc_f(c) -= w_d(c) (8F(c+s) - 8F(c-s) + F(c-2s) - F(c+2s)),
  // w_d(c) = [c interior][no grad3dn zeroing] rdx_d(c)/12
  if (PART == 0 && (T & SMAUG_T_FLUX)) {
    double u[NP], bgc[2 + ND];
#pragma unroll
    for (int f = 0; f < NP; f++) u[f] = w[(long long)f * N + p];
#pragma unroll
    for (int b = 0; b < 2 + ND; b++) bgc[b] = bg[(long long)b * N + p];
    const int ks[4] = {1, -1, -2, 2};
    const double ak[4] = {8.0, -8.0, 1.0, -1.0};
#pragma unroll
    for (int d = 0; d < ND; d++) {
      double Gf[NP];
#pragma unroll
      for (int f = 0; f < NP; f++) Gf[f] = 0.0;
#pragma unroll
      for (int t = 0; t < 4; t++) {
        Nb<ND> c(x, d, -ks[t]);
        if (!interior_c<ND>(G, c.y) || z4(G, zFLUX(ND), d, c.y[d])) continue;
        const double wt = ak[t] * rdx_at(G, A, d, c.y) * (1.0 / 12.0);
#pragma unroll
        for (int f = 0; f < NP; f++) Gf[f] -= wt * mu[(long long)f * N + at(c.y)];
      }
      flux_vjp<ND>(G, u, bgc, d, Gf, lam);
    }
  }

  // ---------------------------------------------------------- background energy terms
  // acc_e(c) += sum_{f,i} k_fi(c) D_i[v_f](c), k_fi = bb_i bb_f - [i=f] p_tb,
  // D_i the masked 4th-order difference (addenergyterms_cd2)
  if (PART == 1 && (T & SMAUG_T_ETERM)) {
    const int ks[4] = {1, -1, -2, 2};
    const double ak[4] = {8.0, -8.0, 1.0, -1.0};
#pragma unroll
    for (int i2 = 0; i2 < ND; i2++) {
#pragma unroll
      for (int t = 0; t < 4; t++) {
        Nb<ND> c(x, i2, -ks[t]);
        if (!interior_c<ND>(G, c.y)) continue;
        const long long cq = at(c.y);
        const double me = mu[(long long)fE(ND) * N + cq];
        if (me == 0.0) continue;
        double bbc[ND], bb2 = 0.0;
#pragma unroll
        for (int q = 0; q < ND; q++) {
          bbc[q] = bg[(long long)bBB(q) * N + cq];
          bb2 += bbc[q] * bbc[q];
        }
        const double ptb = (G.gamma - 1) * bg[(long long)bEB() * N + cq] - 0.5 * (G.gamma - 2) * bb2;
        const double wt = me * ak[t] * rdx_at(G, A, i2, c.y) * (1.0 / 12.0);
#pragma unroll
        for (int f = 0; f < ND; f++) {
          if (z4(G, zVEL(f), i2, c.y[i2])) continue;
          const double k = bbc[i2] * bbc[f] - (i2 == f ? ptb : 0.0);
          vbar[f] += wt * k;
        }
      }
    }
  }

  const bool inp = interior_c<ND>(G, x);
  const double rtp = RT(x);

  // ---------------------------------------------------------- gravity
  // acc_m_d += g_d rho;  acc_e += rho g_d m_d / rho_t
  if (PART == 1 && (T & SMAUG_T_GRAV) && inp) {
    const double me = mu[(long long)fE(ND) * N + p];
    const double rho = w[p], rhob = bg[(long long)bRHOB() * N + p];
#pragma unroll
    for (int d = 0; d < ND; d++) {
      const double md = w[(long long)fMOM(d) * N + p];
      lam[fRHO()] += G.g[d] * mu[(long long)fMOM(d) * N + p] + me * G.g[d] * md * rhob / (rtp * rtp);
      lam[fMOM(d)] += me * rho * G.g[d] / rtp;
    }
  }

  // ---------------------------------------------------------- rho and e hyperdiffusion
  // acc(c) += (nu_r z (q(c+s) - q(c)) - nu_l z (q(c) - q(c-s))) rdx(c)^2
  // gathered at p: mu(p-s) A_r(p-s) - mu(p)(A_r + A_l)(p) + mu(p+s) A_l(p+s)
  if (PART == 4 && G.hyper) {
#pragma unroll
    for (int which = 0; which < 2; which++) {
      if (which != FSEL) continue;
      if (which == 0 && !(T & SMAUG_T_HRHO)) continue;
      if (which == 1 && !(T & SMAUG_T_HE)) continue;
      const int fld = which == 0 ? fRHO() : fE(ND);
      const double* qa = which == 0 ? w : A.prim + (long long)pQE(ND) * N;
      const int zi = which == 0 ? zRHO : zTMP1;
      double acc = 0.0;
#pragma unroll
      for (int d = 0; d < ND; d++) {
        bool ok[3];
        double nr[3], nl[3], r[3];
        nu3_at<ND>(G, A, qa, x, d, cm[d] * G.chyp[fld], ok, nr, nl, r);
#pragma unroll
        for (int k = -1; k <= 1; k++) {
          Nb<ND> c(x, d, k);
          if (!ok[k + 1] || z1(G, zi, d, c.y[d])) continue;
          const double m = mu[(long long)fld * N + at(c.y)];
          const double coef = k == -1 ? nr[0] : (k == 1 ? nl[2] : -(nr[1] + nl[1]));
          acc += m * coef * r[k + 1] * r[k + 1];
        }
      }
      if (which == 0) lam[fRHO()] += acc;
      else qebar += acc;
    }
  }

  // ---------------------------------------------------------- momentum hyperdiffusion (q = v_f)
  if (PART == 2 && G.hyper && (T & SMAUG_T_HMOM)) {
#pragma unroll
    for (int f = 0; f < ND; f++) {
      if (f != FSEL) continue;
      const double* qa = A.prim + (long long)pV(f) * N;
      const double Kc = G.chyp[fMOM(f)];
#pragma unroll
      for (int d = 0; d < ND; d++) {
        const double kap = d == f ? 2.0 : 1.0;
        // nu, rdx and the masks at c = p + k e_d, k = -1, 0, 1
        double nr[3], nl[3], r[3];
        bool ok[3];
        nu3_at<ND>(G, A, qa, x, d, cm[d] * Kc, ok, nr, nl, r);
        // diagonal (src/hyperdifmomsource1.cu): with P = kap nu z rdx^2/2,
        //   acc_m_f(c) += P_r rr (v(c+s) - v(c)) - P_l rl (v(c) - v(c-s))
        //   acc_e(c)   += P_r mr (v(c+s) - v(c)) - P_l ml (v(c) - v(c-s))
        //   rr/rl, mr/ml face averages of rho_t and m_f
        {
          const double vp = qa[p];
          const double mfp = w[(long long)fMOM(f) * N + p];
#pragma unroll
          for (int k = -1; k <= 1; k++) {
            if (!ok[k + 1]) continue;
            Nb<ND> c(x, d, k);
            if (z1(G, zTMP4, d, c.y[d])) continue;
            const long long cq = at(c.y);
            const double mm_ = mu[(long long)fMOM(f) * N + cq], me = mu[(long long)fE(ND) * N + cq];
            const double s2 = kap * r[k + 1] * r[k + 1] * 0.5;
            const double Pr = nr[k + 1] * s2, Pl = nl[k + 1] * s2;
            if (k == -1) {  // p = c + s
              const double rr = (RT(c.y) + rtp) * 0.5, mr = (WV(fMOM(f), c.y) + mfp) * 0.5;
              const double dr = vp - qa[cq];
              vbar[f] += mm_ * Pr * rr + me * Pr * mr;
              rtbar += 0.5 * mm_ * Pr * dr;
              lam[fMOM(f)] += 0.5 * me * Pr * dr;
            } else if (k == 1) {  // p = c - s
              const double rl = (RT(c.y) + rtp) * 0.5, ml = (WV(fMOM(f), c.y) + mfp) * 0.5;
              const double dl = qa[cq] - vp;
              vbar[f] += mm_ * Pl * rl + me * Pl * ml;
              rtbar -= 0.5 * mm_ * Pl * dl;
              lam[fMOM(f)] -= 0.5 * me * Pl * dl;
            } else {  // p = c
              Nb<ND> cp(x, d, 1), cm_(x, d, -1);
              const double rr = (rtp + RT(cp.y)) * 0.5, rl = (rtp + RT(cm_.y)) * 0.5;
              const double mr = (mfp + WV(fMOM(f), cp.y)) * 0.5, ml = (mfp + WV(fMOM(f), cm_.y)) * 0.5;
              const double dr = qa[at(cp.y)] - vp, dl = vp - qa[at(cm_.y)];
              vbar[f] -= mm_ * (Pr * rr + Pl * rl) + me * (Pr * mr + Pl * ml);
              rtbar += 0.5 * mm_ * (Pr * dr - Pl * dl);
              lam[fMOM(f)] += 0.5 * me * (Pr * dr - Pl * dl);
            }
          }
        }
        if (d == f) continue;
        // cross terms (src/hyperdifmomsourcene1.cu), at interior c':
        //   t5 = z (v_f(c'+e_d) - v_f(c'-e_d)) rdx_d/2,  t6 = t5 (nu_r + nu_l)/4
        //   g7 = rho_t t6,  g8 = t6 m_d
        //   acc_m_d(c) += z7 (g7(c+e_f) - g7(c-e_f)) rdx_f(c)/2
        //   acc_e(c)   += z8 (g8(c+e_f) - g8(c-e_f)) rdx_f(c)/2
        // with cotangents gb7(c') = mu_m_d(c'-e_f) w7(c'-e_f) - mu_m_d(c'+e_f) w7(c'+e_f)
#pragma unroll
        for (int k = -1; k <= 1; k++) {
          if (!ok[k + 1]) continue;
          Nb<ND> c(x, d, k);
          double gb7 = 0.0, gb8 = 0.0;
#pragma unroll
          for (int sg = -1; sg <= 1; sg += 2) {
            Nb<ND> cc(c.y, f, sg);  // c' + sg e_f
            if (!interior_c<ND>(G, cc.y)) continue;
            const long long cq = at(cc.y);
            const double wr = rdx_at(G, A, f, cc.y) * 0.5;
            // c = c' - e_f contributes +, c = c' + e_f contributes -
            const double sgn = sg < 0 ? 1.0 : -1.0;
            if (!z1(G, zTMP7, f, cc.y[f])) gb7 += sgn * mu[(long long)fMOM(d) * N + cq] * wr;
            if (!z1(G, zTMP8, f, cc.y[f])) gb8 += sgn * mu[(long long)fE(ND) * N + cq] * wr;
          }
          if (gb7 == 0.0 && gb8 == 0.0) continue;
          const double nsum = (nr[k + 1] + nl[k + 1]) * 0.25;
          const bool zc = z1(G, zTMP4, d, c.y[d]);
          if (k == 0) {
            Nb<ND> cp(x, d, 1), cm_(x, d, -1);
            const double t5 = zc ? 0.0 : (qa[at(cp.y)] - qa[at(cm_.y)]) * r[1] * 0.5;
            const double t6 = t5 * nsum;
            rtbar += gb7 * t6;
            lam[fMOM(d)] += gb8 * t6;
          } else if (!zc) {
            const double t5b = (gb7 * RT(c.y) + gb8 * WV(fMOM(d), c.y)) * nsum;
            // p = c' + e_d (k = -1) gets +, p = c' - e_d (k = +1) gets -
            vbar[f] += (k == -1 ? 1.0 : -1.0) * t5b * r[k + 1] * 0.5;
          }
        }
      }
    }
  }

  // ---------------------------------------------------------- magnetic hyperdiffusion (q = b_f)
  if (PART == 3 && G.hyper && (T & SMAUG_T_HB)) {
#pragma unroll
    for (int f = 0; f < ND; f++) {
      if (f != FSEL) continue;
      const double* qa = w + (long long)fB(ND, f) * N;
      const double Kc = G.chyp[fB(ND, f)];
      const double bfp = qa[p];
#pragma unroll
      for (int d = 0; d < ND; d++) {
        if (d == f) continue;
        double nr[3], nl[3], r[3];
        bool ok[3];
        nu3_at<ND>(G, A, qa, x, d, cm[d] * Kc, ok, nr, nl, r);
        // diagonal (src/hyperdifbsource1.cu), Q = nu z rdx^2:
        //   acc_b_f(c) += Q_r (b(c+s) - b(c)) - Q_l (b(c) - b(c-s))
        //   acc_e(c)   += Q_r (b(c+s)^2 - b(c)^2)/2 - Q_l (b(c)^2 - b(c-s)^2)/2
        {
          double sb = 0.0, se = 0.0;
#pragma unroll
          for (int k = -1; k <= 1; k++) {
            if (!ok[k + 1]) continue;
            Nb<ND> c(x, d, k);
            if (z1(G, zTMP1, d, c.y[d])) continue;
            const long long cq = at(c.y);
            const double r2 = r[k + 1] * r[k + 1];
            const double coef = (k == -1 ? nr[k + 1] : (k == 1 ? nl[k + 1] : -(nr[k + 1] + nl[k + 1]))) * r2;
            sb += mu[(long long)fB(ND, f) * N + cq] * coef;
            se += mu[(long long)fE(ND) * N + cq] * coef;
          }
          lam[fB(ND, f)] += sb + bfp * se;
        }
        // curl-form cross terms (src/hyperdifbsourcene1.cu), interior c':
        //   t2 = z (b_f(c'+e_d) - b_f(c'-e_d)) rdx_d/2,  t3 = t2 (nu_l + nu_r)/2
        //   g3 = t3,  g5 = t3 b_d
        //   acc_b_d(c) -= z3 (g3(c+e_f) - g3(c-e_f)) rdx_f(c)/2
        //   acc_e(c)   -= z5 (g5(c+e_f) - g5(c-e_f)) rdx_f(c)/2
#pragma unroll
        for (int k = -1; k <= 1; k++) {
          if (!ok[k + 1]) continue;
          Nb<ND> c(x, d, k);
          double gb3 = 0.0, gb5 = 0.0;
#pragma unroll
          for (int sg = -1; sg <= 1; sg += 2) {
            Nb<ND> cc(c.y, f, sg);
            if (!interior_c<ND>(G, cc.y)) continue;
            const long long cq = at(cc.y);
            const double wr = rdx_at(G, A, f, cc.y) * 0.5;
            // c = c' - e_f contributes -(+1), c = c' + e_f contributes -(-1)
            const double sgn = sg < 0 ? -1.0 : 1.0;
            if (!z1(G, zTMP3, f, cc.y[f])) gb3 += sgn * mu[(long long)fB(ND, d) * N + cq] * wr;
            if (!z1(G, zTMP5, f, cc.y[f])) gb5 += sgn * mu[(long long)fE(ND) * N + cq] * wr;
          }
          if (gb3 == 0.0 && gb5 == 0.0) continue;
          const double nsum = (nl[k + 1] + nr[k + 1]) * 0.5;
          const bool zc = z1(G, zTMP1, d, c.y[d]);
          if (k == 0) {
            Nb<ND> cp(x, d, 1), cm_(x, d, -1);
            const double t2 = zc ? 0.0 : (qa[at(cp.y)] - qa[at(cm_.y)]) * r[1] * 0.5;
            lam[fB(ND, d)] += gb5 * t2 * nsum;
          } else if (!zc) {
            const double t2b = (gb3 + gb5 * WV(fB(ND, d), c.y)) * nsum;
            lam[fB(ND, f)] += (k == -1 ? 1.0 : -1.0) * t2b * r[k + 1] * 0.5;
          }
        }
      }
    }
  }

  // ---------------------------------------------------------- chain to the state at p
  //   v_f = m_f/rho_t;  q_e = e - (b.b + m.m/rho_t)/2;  rho_t = rho + rhob
  {
    double mm = 0.0;
#pragma unroll
    for (int q = 0; q < ND; q++) {
      const double mq = w[(long long)fMOM(q) * N + p];
      mm += mq * mq;
      lam[fMOM(q)] += vbar[q] / rtp - qebar * mq / rtp;
      rtbar -= vbar[q] * (mq / rtp) / rtp;
      lam[fB(ND, q)] -= qebar * w[(long long)fB(ND, q) * N + p];
    }
    lam[fE(ND)] += qebar;
    rtbar += 0.5 * qebar * mm / (rtp * rtp);
    lam[fRHO()] += rtbar;
  }
#pragma unroll
  for (int f = 0; f < NP; f++) {
    const long long q = (long long)f * N + p;
    A.out[q] = PART == 0 ? lam[f] : A.out[q] + lam[f];
  }
}

// ---------------------------------------------------------------------------
// k_adj_fold: lam(p) += T(p) + sign * T(ghost images of p), interior p, per
// field with that field's boundary map (the transpose of the ghost fill; a
// fixed ghost is a constant and passes nothing on).
template <int ND>
__global__ __launch_bounds__(256) void k_adj_fold(Geo G, const double* __restrict__ Tm, double* __restrict__ lam) {
  const long long c = (long long)blockIdx.x * 256 + threadIdx.x;
  if (c >= G.N) return;
  int x[3];
  x[0] = (int)(c % G.n[0]);
  x[1] = (int)((c / G.n[0]) % G.n[1]);
  x[2] = ND == 3 ? (int)(c / ((long long)G.n[0] * G.n[1])) : 0;
  if (!interior_c<ND>(G, x)) return;
  for (int f = 0; f < G.np; f++) {
    const double* Tf = Tm + (long long)f * G.N;
    double s = Tf[c];
#pragma unroll
    for (int d = 0; d < ND; d++) {
      int im[4];
      const int m = bc_images(G.bct[f][d], x[d], G.n[d], ND, im);
      const double sg = bc_sign(G.bct[f][d]);
      for (int q = 0; q < m; q++) {
        int y[3] = {x[0], x[1], x[2]};
        y[d] = im[q];
        s += sg * Tf[gix(G, y[0], y[1], y[2])];
      }
    }
    lam[(long long)f * G.N + c] += s;
  }
}

// gmap[col] += sum_i gcell(i, col)
__global__ __launch_bounds__(256) void k_colsum(Geo G, const double* __restrict__ gcell, double* __restrict__ gmap) {
  const long long col = (long long)blockIdx.x * 256 + threadIdx.x;
  const long long ncol = G.N / G.n[0];
  if (col >= ncol) return;
  const double* g = gcell + col * G.n[0];
  double s = 0.0;
  for (int i = 0; i < G.n[0]; i++) s += g[i];
  gmap[col] += s;
}

__global__ void k_meta(double* meta, const double* dts, double qt, double tdep) {
  if (threadIdx.x == 0) {
    meta[0] = dts[0];
    meta[1] = qt;
    meta[2] = tdep;
  }
}

static unsigned nblk(long long n) { return (unsigned)((n + 255) / 256); }

void launch_adj_mu(const Geo& G, hipStream_t st, const AdjMuArgs& A) {
  if (G.nd == 3)
    hipLaunchKernelGGL(k_adj_mu<3>, dim3(nblk(G.N)), dim3(256), 0, st, G, A);
  else
    hipLaunchKernelGGL(k_adj_mu<2>, dim3(nblk(G.N)), dim3(256), 0, st, G, A);
}

template <int ND>
static void launch_adj_rhs_t(const Geo& G, hipStream_t st, const AdjArgs& A) {
  const dim3 g(nblk(G.N)), b(256);
  hipLaunchKernelGGL((k_adj_rhs<ND, 0, 0>), g, b, 0, st, G, A);
  hipLaunchKernelGGL((k_adj_rhs<ND, 1, 0>), g, b, 0, st, G, A);
  if (G.hyper && (G.terms & SMAUG_T_HRHO)) hipLaunchKernelGGL((k_adj_rhs<ND, 4, 0>), g, b, 0, st, G, A);
  if (G.hyper && (G.terms & SMAUG_T_HE)) hipLaunchKernelGGL((k_adj_rhs<ND, 4, 1>), g, b, 0, st, G, A);
  if (G.hyper && (G.terms & SMAUG_T_HMOM)) {
    hipLaunchKernelGGL((k_adj_rhs<ND, 2, 0>), g, b, 0, st, G, A);
    hipLaunchKernelGGL((k_adj_rhs<ND, 2, 1>), g, b, 0, st, G, A);
    if (ND == 3) hipLaunchKernelGGL((k_adj_rhs<ND, 2, 2>), g, b, 0, st, G, A);
  }
  if (G.hyper && (G.terms & SMAUG_T_HB)) {
    hipLaunchKernelGGL((k_adj_rhs<ND, 3, 0>), g, b, 0, st, G, A);
    hipLaunchKernelGGL((k_adj_rhs<ND, 3, 1>), g, b, 0, st, G, A);
    if (ND == 3) hipLaunchKernelGGL((k_adj_rhs<ND, 3, 2>), g, b, 0, st, G, A);
  }
}

void launch_adj_rhs(const Geo& G, hipStream_t st, const AdjArgs& A) {
  if (G.nd == 3)
    launch_adj_rhs_t<3>(G, st, A);
  else
    launch_adj_rhs_t<2>(G, st, A);
}

void launch_adj_fold(const Geo& G, hipStream_t st, const double* T, double* lam) {
  if (G.nd == 3)
    hipLaunchKernelGGL(k_adj_fold<3>, dim3(nblk(G.N)), dim3(256), 0, st, G, T, lam);
  else
    hipLaunchKernelGGL(k_adj_fold<2>, dim3(nblk(G.N)), dim3(256), 0, st, G, T, lam);
}

void launch_colsum(const Geo& G, hipStream_t st, const double* gcell, double* gmap) {
  hipLaunchKernelGGL(k_colsum, dim3(nblk(G.N / G.n[0])), dim3(256), 0, st, G, gcell, gmap);
}

void launch_meta(hipStream_t st, double* meta, const double* dts, double qt, double tdep) {
  hipLaunchKernelGGL(k_meta, dim3(1), dim3(64), 0, st, meta, dts, qt, tdep);
}

}  // namespace smaug
