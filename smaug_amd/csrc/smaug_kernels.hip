// smaug_kernels.hip -- MI355X (gfx950) kernels of the SMAUG MHD stage.
//
// One RK stage (or the reference's single Euler stage) is two launches:
//   k_prim  (all cells, pointwise): velocities, total perturbation pressure,
//           the energy hyperdiffusion variable, and the per-direction wave
//           speed maxima reduced through wavefront shuffles, LDS and one
//           64-bit atomicMax per block (the reference copies the whole wd
//           array to the host for this, src/computedervfields.cu:1914).
//   k_rhs   (interior cells): flux divergence (centdiff1/centdiff2),
//           background energy terms, gravity, all hyperdiffusion
//           coefficients and sources, the stage update, the user source on
//           the updated state, the RK combination and the ghost fill.
// The CFL time step is chosen on the device by k_dt so a step needs no
// host round trip.
#include <math.h>

#include "../../include/smaug_hip.h"
#include "smaug_dev.h"

namespace smaug {

__device__ inline double bits2d(unsigned long long b) { return __longlong_as_double((long long)b); }

// ---------------------------------------------------------------------------
// k_prim: pointwise derived fields + wave-speed maxima
//   v_d  = m_d/(rho+rhob)                         src/dervfields.cu:79-98
//   p_t  = (g-1)(e - m^2/2rho_t) - (g-2)(b.bb + b^2/2) src/dervfields.cu:107-153
//   q_e  = e - (b^2 + m^2/rho_t)/2                src/hyperdifesource1.cu:391-395
//   c_s, c_fast,d                                 src/dervfields.cu:241-309
template <int ND>
__global__ __launch_bounds__(256) void k_prim(Geo G, const double* __restrict__ w,
                                              const double* __restrict__ bg, double* __restrict__ prim,
                                              unsigned long long* __restrict__ cmax_bits,
                                              int* __restrict__ flag) {
  const long long N = G.N;
  double mx[3] = {0.0, 0.0, 0.0};
  int bad = 0;
  const long long stride = (long long)gridDim.x * 256;
  for (long long c = (long long)blockIdx.x * 256 + threadIdx.x; c < N; c += stride) {
    const double gm = G.gamma;
    const double rho = w[c], rhob = bg[(long long)bRHOB() * N + c];
    double m[ND], b[ND], bb[ND];
#pragma unroll
    for (int d = 0; d < ND; d++) {
      m[d] = w[(long long)fMOM(d) * N + c];
      b[d] = w[(long long)fB(ND, d) * N + c];
      bb[d] = bg[(long long)bBB(d) * N + c];
    }
    const double e = w[(long long)fE(ND) * N + c];
    const double eb = bg[(long long)bEB() * N + c];
    const double rt = rho + rhob;
    double mm = m[0] * m[0], b2 = b[0] * b[0], bbb = b[0] * bb[0], bb2 = bb[0] * bb[0];
    double bbb_c = bb[0] * b[0];
#pragma unroll
    for (int d = 1; d < ND; d++) {
      mm = mm + m[d] * m[d];
      b2 = b2 + b[d] * b[d];
      bbb = bbb + b[d] * bb[d];
      bb2 = bb2 + bb[d] * bb[d];
      bbb_c = bbb_c + bb[d] * b[d];
    }
#pragma unroll
    for (int d = 0; d < ND; d++) prim[(long long)pV(d) * N + c] = m[d] / rt;
    const double v = (gm - 2.0) * (bbb + 0.5 * b2);
    prim[(long long)pPT(ND) * N + c] = (gm - 1.0) * (e - 0.5 * mm / rt) - v;
    prim[(long long)pQE(ND) * N + c] = e - 0.5 * (b2 + mm / rt);
    // sound and fast speeds
    const double pk = (gm - 1) * (e - 0.5 * (mm / rt) - 0.5 * b2 - bbb_c);
    const double cs2 = (gm * (pk + (gm - 1) * (eb - 0.5 * bb2))) / rt;
    const double cf = (b2 + bb2 + 2.0 * bbb_c) / rt + cs2;
    const double cs = sqrt(cs2);
#pragma unroll
    for (int d = 0; d < ND; d++) {
      const double btd = bb[d] + b[d];
      const double cfi = 0.5 * (cf + sqrt(cf * cf - 4.0 * cs2 * (btd * btd) / rt));
      const double cfd = sqrt(cfi) + fabs(m[d] / rt);
      if (!(cfd <= 1.0e300) || !(cs <= 1.0e300)) bad = 1;  // NaN or inf
      double t = cfd > cs ? cfd : cs;
      if (t > mx[d]) mx[d] = t;
    }
  }
  // wavefront (64 lanes) then block reduction of the three maxima
#pragma unroll
  for (int d = 0; d < ND; d++) {
    double t = mx[d];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      double o = __shfl_xor(t, off, 64);
      t = o > t ? o : t;
    }
    mx[d] = t;
  }
  __shared__ double red[4][3];
  const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0)
    for (int d = 0; d < ND; d++) red[wid][d] = mx[d];
  if (bad) atomicOr(flag, 1);
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int d = 0; d < ND; d++) {
      double t = red[0][d];
      for (int q = 1; q < 4; q++) t = red[q][d] > t ? red[q][d] : t;
      if (t > 0.0) atomicMax(&cmax_bits[d], (unsigned long long)__double_as_longlong(t));
    }
  }
}

__global__ void k_zero3(unsigned long long* p) {
  if (threadIdx.x < 3) p[threadIdx.x] = 0ull;
}

// CFL step (src/iosmaug.c:541-616, quirk Q4): the last direction's maximum
// over every dir's dx; the step keeps its previous value unless the CFL step
// lies in (1e-12, DT).  dts[0] = dt, dts[1] = accumulated time.
__global__ void k_dt(const unsigned long long* cmax_bits, int nd, double dx0, double dx1, double dx2,
                     double courant, double dtcap, int moddton, double* dts) {
  if (threadIdx.x != 0) return;
  double dt = dts[0];
  if (moddton == 1) {
    const double cm = bits2d(cmax_bits[nd - 1]);
    const double dxs[3] = {dx0, dx1, dx2};
    double courantmax = 1.0e-12;
    for (int d = 0; d < nd; d++) {
      double t = cm / dxs[d];
      if (t > courantmax) courantmax = t;
    }
    const double dtt = courant / courantmax;
    if (dtt > 1.0e-12 && dtt < dtcap) dt = dtt;
  }
  dts[0] = dt;
  dts[1] = dts[1] + dt;
}

// ---------------------------------------------------------------------------
// stencil helpers on a field array a (length N) at interior cell c

// 4th-order central difference / (2 delx) along d (grad3dn_MODID,
// src/gradops.cu:437-602); the caller applies the Q1 zeroing.
__device__ inline double d4(const double* __restrict__ a, long long c, long long s, double dx) {
  return ((8 * a[c + s] - 8 * a[c - s] + a[c - 2 * s] - a[c + 2 * s]) / 6.0) / (2.0 * dx);
}

template <int ND>
__device__ inline bool is_interior(const Geo& G, const int* x) {
#pragma unroll
  for (int d = 0; d < ND; d++)
    if (x[d] < 2 || x[d] > G.n[d] - 3) return false;
  return true;
}

// Hyperdiffusion coefficients nu_{i+1/2} (hdnur) and nu_{i-1/2} (hdnul) of
// the variable qa along d at interior cell c (src/hyperdifvisc1r.cu:426-843,
// src/hyperdifvisc1l.cu:17-853): third and first differences at interfaces,
// maxima over 3 and 5 interfaces, nu = delx cmax chyp max3/max5 (0 if the
// denominator is 0).  Past the grid the right chain reads the stray halo of
// quirk Q7 and the left chain reads q at coordinate 5 (quirk Q8).
template <int ND>
__device__ inline void nu_lr(const Geo& G, const double* __restrict__ qa, long long c, const int* x, int d,
                             double dxc, double cmc, double& nur, double& nul) {
  const long long s = G.st[d];
  const int xd = x[d], nn = G.n[d];
  double Q[7];
#pragma unroll
  for (int o = -3; o <= 3; o++) {
    const int xx = xd + o;
    double v;
    if (xx < 0) {
      v = qa[c + (long long)(5 - xd) * s];
    } else if (xx >= nn) {
      int pc[3];
#pragma unroll
      for (int q = 0; q < 3; q++) pc[q] = (q < ND ? x[q] : 0) + 1;
      pc[d] = nn + 1;
      const long long h = halo_r_src(G, pc);
      v = h < 0 ? 0.0 : qa[h];
    } else {
      v = qa[c + (long long)o * s];
    }
    Q[o + 3] = v;
  }
  double D3[6], D1[6];
#pragma unroll
  for (int a = 1; a <= 4; a++) D3[a] = fabs(3.0 * (Q[a + 1] - Q[a]) - (Q[a + 2] - Q[a - 1]));
#pragma unroll
  for (int a = 0; a <= 5; a++) D1[a] = fabs(Q[a + 1] - Q[a]);
  double m3r = 0, m3l = 0, m5r = 0, m5l = 0;
#pragma unroll
  for (int a = 2; a <= 4; a++) m3r = D3[a] > m3r ? D3[a] : m3r;
#pragma unroll
  for (int a = 1; a <= 3; a++) m3l = D3[a] > m3l ? D3[a] : m3l;
#pragma unroll
  for (int a = 1; a <= 5; a++) m5r = D1[a] > m5r ? D1[a] : m5r;
#pragma unroll
  for (int a = 0; a <= 4; a++) m5l = D1[a] > m5l ? D1[a] : m5l;
  nur = m5r > 0 ? dxc * cmc * m3r / m5r : 0.0;
  nul = m5l > 0 ? dxc * cmc * m3l / m5l : 0.0;
}

// fluxes of every evolved field along d at cell cc (src/centdiff1.cu:51-256,
// src/centdiff2.cu:158-356), written with v = m/rho_t from k_prim
template <int ND>
__device__ inline void fluxes(const Geo& G, const double* __restrict__ w, const double* __restrict__ bg,
                              const double* __restrict__ pr, long long cc, int d, double* F) {
  const long long N = G.N;
  double v[ND], b[ND], bb[ND];
#pragma unroll
  for (int q = 0; q < ND; q++) {
    v[q] = pr[(long long)pV(q) * N + cc];
    b[q] = w[(long long)fB(ND, q) * N + cc];
    bb[q] = bg[(long long)bBB(q) * N + cc];
  }
  const double vd = v[d], bd = b[d], bbd = bb[d];
  const double rho = w[cc], rhob = bg[(long long)bRHOB() * N + cc];
  const double pt = pr[(long long)pPT(ND) * N + cc];
  F[fRHO()] = vd * rho + rhob * vd;
#pragma unroll
  for (int f = 0; f < ND; f++) {
    const double mf = w[(long long)fMOM(f) * N + cc];
    double fm = -(b[f] * bbd + bb[f] * bd) - b[f] * bd;
    if (f == d) fm = pt - (b[f] * bbd + bb[f] * bd) - b[f] * bd;
    F[fMOM(f)] = vd * mf + fm;
  }
  double bbv = bb[0] * v[0], bv = b[0] * v[0];
#pragma unroll
  for (int q = 1; q < ND; q++) {
    bbv = bbv + bb[q] * v[q];
    bv = bv + b[q] * v[q];
  }
  const double e = w[(long long)fE(ND) * N + cc], eb = bg[(long long)bEB() * N + cc];
  F[fE(ND)] = vd * e + ((vd * pt - bd * bbv - bbd * bv) + vd * eb - bd * bv);
#pragma unroll
  for (int f = 0; f < ND; f++)
    F[fB(ND, f)] = (f == d) ? 0.0 : vd * b[f] - (bd + bbd) * v[f] + bb[f] * vd;
}

template <int ND>
__global__ __launch_bounds__(256) void k_rhs(Geo G, RhsArgs A) {
  constexpr int NP = 2 + 2 * ND;
  int x[3];
  x[0] = 2 + blockIdx.x * 64 + threadIdx.x;
  x[1] = 2 + blockIdx.y * 4 + threadIdx.y;
  x[2] = ND == 3 ? 2 + blockIdx.z : 0;
  if (A.zero_next && blockIdx.x == 0 && blockIdx.y == 0 && blockIdx.z == 0 && threadIdx.x < 3 &&
      threadIdx.y == 0)
    A.zero_next[threadIdx.x] = 0ull;
  if (x[0] > G.n[0] - 3 || x[1] > G.n[1] - 3) return;
  const long long N = G.N;
  const long long c = ((long long)x[2] * G.n[1] + x[1]) * G.n[0] + x[0];
  const double* __restrict__ w = A.win;
  const double* __restrict__ bg = A.bg;
  const double* __restrict__ pr = A.prim;
  const double* __restrict__ dxa = A.delx;
  const unsigned T = G.terms;
  double acc[NP];
#pragma unroll
  for (int f = 0; f < NP; f++) acc[f] = 0.0;
  double dx[ND];
#pragma unroll
  for (int d = 0; d < ND; d++) dx[d] = dxa[(long long)d * N + c];

  // ---- flux divergence (cucentdiff1, cucentdiff2)
  if (T & SMAUG_T_FLUX) {
#pragma unroll
    for (int d = 0; d < ND; d++) {
      if (z4(G, zFLUX(ND), d, x[d])) continue;
      const long long s = G.st[d];
      double Fm2[NP], Fm1[NP], Fp1[NP], Fp2[NP];
      fluxes<ND>(G, w, bg, pr, c - 2 * s, d, Fm2);
      fluxes<ND>(G, w, bg, pr, c - s, d, Fm1);
      fluxes<ND>(G, w, bg, pr, c + s, d, Fp1);
      fluxes<ND>(G, w, bg, pr, c + 2 * s, d, Fp2);
#pragma unroll
      for (int f = 0; f < NP; f++)
        acc[f] -= ((8 * Fp1[f] - 8 * Fm1[f] + Fm2[f] - Fp2[f]) / 6.0) / (2.0 * dx[d]);
    }
  }
  const double rho_c = w[c], rhob_c = bg[(long long)bRHOB() * N + c];
  const double rt_c = rho_c + rhob_c;
  // ---- background energy terms (addenergyterms_cd2, src/centdiff2.cu:80-115)
  if (T & SMAUG_T_ETERM) {
    double bb[ND], bb2;
#pragma unroll
    for (int d = 0; d < ND; d++) bb[d] = bg[(long long)bBB(d) * N + c];
    bb2 = bb[0] * bb[0];
#pragma unroll
    for (int d = 1; d < ND; d++) bb2 = bb2 + bb[d] * bb[d];
    const double ptb = (G.gamma - 1) * bg[(long long)bEB() * N + c] - 0.5 * (G.gamma - 2) * bb2;
#pragma unroll
    for (int d = 0; d < ND; d++) {
      const double* vd = pr + (long long)pV(d) * N;
      double gdd = z4(G, zVEL(d), d, x[d]) ? 0.0 : d4(vd, c, G.st[d], dx[d]);
      acc[fE(ND)] += -ptb * gdd;
#pragma unroll
      for (int i2 = 0; i2 < ND; i2++) {
        double gi = z4(G, zVEL(d), i2, x[i2]) ? 0.0 : d4(vd, c, G.st[i2], dx[i2]);
        acc[fE(ND)] += bb[i2] * bb[d] * gi;
      }
    }
  }
  // ---- gravity (addgrav_cd2, src/centdiff2.cu:118-154)
  if (T & SMAUG_T_GRAV) {
#pragma unroll
    for (int d = 0; d < ND; d++) {
      acc[fMOM(d)] += G.g[d] * rho_c;
      acc[fE(ND)] += rho_c * (G.g[d] * w[(long long)fMOM(d) * N + c]) / rt_c;
    }
  }
  // ---- hyperdiffusion (src/iosmaug.c:685-812)
  if (G.hyper) {
    double cm[ND];
#pragma unroll
    for (int d = 0; d < ND; d++) cm[d] = bits2d(A.cmax_bits[d]);
    const double* rho = w;
    const double* rhob = bg + (long long)bRHOB() * N;
    if (T & SMAUG_T_HRHO) {  // src/hyperdifrhosource1.cu:15-104
#pragma unroll
      for (int dim = 0; dim < ND; dim++) {
        double nur, nul;
        nu_lr<ND>(G, rho, c, x, dim, dx[dim], cm[dim] * G.chyp[fRHO()], nur, nul);
        const long long s = G.st[dim];
        const bool z = z1(G, zRHO, dim, x[dim]);
        const double gr = z ? 0.0 : (rho[c + s] - rho[c]) / dx[dim];
        const double gl = z ? 0.0 : (rho[c] - rho[c - s]) / dx[dim];
        acc[fRHO()] += (nur * gr - nul * gl) / dx[dim];
      }
    }
    if (T & SMAUG_T_HE) {  // src/hyperdifesource1.cu:18-119
      const double* qe = pr + (long long)pQE(ND) * N;
#pragma unroll
      for (int dim = 0; dim < ND; dim++) {
        double nur, nul;
        nu_lr<ND>(G, qe, c, x, dim, dx[dim], cm[dim] * G.chyp[fE(ND)], nur, nul);
        const long long s = G.st[dim];
        const bool z = z1(G, zTMP1, dim, x[dim]);
        const double gr = z ? 0.0 : (qe[c + s] - qe[c]) / dx[dim];
        const double gl = z ? 0.0 : (qe[c] - qe[c - s]) / dx[dim];
        acc[fE(ND)] += (nur * gr - nul * gl) / dx[dim];
      }
    }
    if (T & SMAUG_T_HMOM) {
#pragma unroll
      for (int dim = 0; dim < ND; dim++) {
        const long long s = G.st[dim];
#pragma unroll
        for (int f = 0; f < ND; f++) {
          const double* vf = pr + (long long)pV(f) * N;
          const double ch = cm[dim] * G.chyp[fMOM(f)];
          double nur, nul;
          nu_lr<ND>(G, vf, c, x, dim, dx[dim], ch, nur, nul);
          // diagonal (hyperdifmomsource1, src/hyperdifmomsource1.cu:17-116,300-480)
          const bool z = z1(G, zTMP4, dim, x[dim]);
          const double gr = z ? 0.0 : (vf[c + s] - vf[c]) / dx[dim];
          const double gl = z ? 0.0 : (vf[c] - vf[c - s]) / dx[dim];
          const double rr = (rho[c] + rhob[c] + rho[c + s] + rhob[c + s]) / 2;
          const double rl = (rho[c] + rhob[c] + rho[c - s] + rhob[c - s]) / 2;
          const double* mf = w + (long long)fMOM(f) * N;
          const double mr = (mf[c] + mf[c + s]) / 2, ml = (mf[c] + mf[c - s]) / 2;
          const double dmv = (rr * nur * gr - rl * nul * gl) / dx[dim] / 2;
          const double dev = (mr * nur * gr - ml * nul * gl) / dx[dim] / 2;
          acc[fMOM(f)] += dmv;
          acc[fE(ND)] += dev;
          if (f == dim) {
            acc[fMOM(f)] += dmv;
            acc[fE(ND)] += dev;
          } else {
            // off-diagonal (hyperdifmomsourcene1, src/hyperdifmomsourcene1.cu:16-752)
            const long long sf = G.st[f];
            const double* md = w + (long long)fMOM(dim) * N;
            double g7[2], g8[2];
#pragma unroll
            for (int sg = 0; sg < 2; sg++) {
              int xq[3] = {x[0], x[1], x[2]};
              xq[f] += sg ? 1 : -1;
              const long long cq = sg ? c + sf : c - sf;
              g7[sg] = 0.0;
              g8[sg] = 0.0;
              if (is_interior<ND>(G, xq)) {
                double nr, nl;
                nu_lr<ND>(G, vf, cq, xq, dim, dxa[(long long)dim * N + cq], ch, nr, nl);
                const double t5 = z1(G, zTMP4, dim, xq[dim])
                                      ? 0.0
                                      : (vf[cq + s] - vf[cq - s]) / dxa[(long long)dim * N + cq] / 2.0;
                const double t6 = t5 * (nr + nl) / 4.0;
                g7[sg] = (rho[cq] + rhob[cq]) * t6;
                g8[sg] = t6 * md[cq];
              }
            }
            const double o7 = z1(G, zTMP7, f, x[f]) ? 0.0 : (g7[1] - g7[0]) / dx[f] / 2.0;
            const double o8 = z1(G, zTMP8, f, x[f]) ? 0.0 : (g8[1] - g8[0]) / dx[f] / 2.0;
            acc[fMOM(dim)] += o7;
            acc[fE(ND)] += o8;
          }
        }
      }
    }
    if (T & SMAUG_T_HB) {
#pragma unroll
      for (int dim = 0; dim < ND; dim++) {
        const long long s = G.st[dim];
#pragma unroll
        for (int f = 0; f < ND; f++) {
          if (f == dim) continue;
          const double* bf = w + (long long)fB(ND, f) * N;
          const double ch = cm[dim] * G.chyp[fB(ND, f)];
          double nur, nul;
          nu_lr<ND>(G, bf, c, x, dim, dx[dim], ch, nur, nul);
          // off-diagonal, sb=-1 (hyperdifbsourcene1, src/hyperdifbsourcene1.cu:18-829)
          {
            const long long sf = G.st[f];
            const double* bd = w + (long long)fB(ND, dim) * N;
            double g3[2], g5[2];
#pragma unroll
            for (int sg = 0; sg < 2; sg++) {
              int xq[3] = {x[0], x[1], x[2]};
              xq[f] += sg ? 1 : -1;
              const long long cq = sg ? c + sf : c - sf;
              g3[sg] = 0.0;
              g5[sg] = 0.0;
              if (is_interior<ND>(G, xq)) {
                double nr, nl;
                nu_lr<ND>(G, bf, cq, xq, dim, dxa[(long long)dim * N + cq], ch, nr, nl);
                const double t2 = z1(G, zTMP1, dim, xq[dim])
                                      ? 0.0
                                      : (bf[cq + s] - bf[cq - s]) / dxa[(long long)dim * N + cq] / 2.0;
                const double t3 = t2 * (nl + nr) / 2;
                g3[sg] = t3;
                g5[sg] = t3 * bd[cq];
              }
            }
            const double o3 = z1(G, zTMP3, f, x[f]) ? 0.0 : (g3[1] - g3[0]) / dx[f] / 2.0;
            const double o5 = z1(G, zTMP5, f, x[f]) ? 0.0 : (g5[1] - g5[0]) / dx[f] / 2.0;
            acc[fB(ND, dim)] += -1.0 * o3;
            acc[fE(ND)] += -1.0 * o5;
          }
          // diagonal, sb=+1 (hyperdifbsource1, src/hyperdifbsource1.cu:21-387)
          {
            const bool z = z1(G, zTMP1, dim, x[dim]);
            const double gr = z ? 0.0 : (bf[c + s] - bf[c]) / dx[dim];
            const double gl = z ? 0.0 : (bf[c] - bf[c - s]) / dx[dim];
            const double br = (bf[c] + bf[c + s]) / 2, bl = (bf[c] + bf[c - s]) / 2;
            acc[fB(ND, f)] += (gr * nur - gl * nul) / dx[dim];
            acc[fE(ND)] += (br * gr * nur - bl * gl * nul) / dx[dim];
          }
        }
      }
    }
  }

  // ---- stage update
  const double h = A.hfrac * A.dts[0];
  double out[NP];
#pragma unroll
  for (int f = 0; f < NP; f++) out[f] = A.base[(long long)f * N + c] + h * acc[f];

  // ---- user source on the updated state (source_parallel, src/centdiff2.cu:790-872)
  if ((T & SMAUG_T_SOURCE) && G.driver == SMAUG_DRIVER_VERTICAL) {
    const double* dv = G.drv;
    const double zp = A.pos[c] - dv[2];
    const double xp = A.pos[N + c] - dv[3];
    const double ez = exp(-zp * zp / (dv[5] * dv[5]));
    const double ex = exp(-xp * xp / (dv[6] * dv[6]));
    double exyz = ex * ez;
    if (ND == 3) {
      const double yp = A.pos[2 * N + c] - dv[4];
      const double ey = exp(-yp * yp / (dv[6] * dv[6]));
      exyz = ex * ey * ez;
    }
    const double vz = dv[0] * exyz * A.tdep;
    const double rto = out[fRHO()] + rhob_c;
    out[fMOM(0)] += h * vz * rto;
    out[fE(ND)] += h * (vz * vz) * rto / 2.0;
  }
  // ---- RK combination (src/advance.cu:89-95)
  if (A.r1) {
#pragma unroll
    for (int f = 0; f < NP; f++) {
      const long long q = (long long)f * N + c;
      out[f] = out[f] + (A.r1[q] + 2.0 * A.r2[q] + A.r3[q] - 4.0 * A.base[q]) / 3;
    }
  }
  // ---- write the cell and its ghost images (boundaries fused, see bc_images)
  double* __restrict__ wo = A.wout;
  if (G.bc_uniform) {
    int li[ND][3], ln[ND];
#pragma unroll
    for (int d = 0; d < ND; d++) {
      li[d][0] = x[d];
      ln[d] = 1 + bc_images(G.bct[0][d], x[d], G.n[d], &li[d][1]);
    }
    if (ND == 2) {
      for (int a = 0; a < ln[0]; a++)
        for (int b2 = 0; b2 < ln[1]; b2++) {
          const long long cc = (long long)li[1][b2] * G.n[0] + li[0][a];
#pragma unroll
          for (int f = 0; f < NP; f++) wo[(long long)f * N + cc] = out[f];
        }
    } else {
      for (int a = 0; a < ln[0]; a++)
        for (int b2 = 0; b2 < ln[1]; b2++)
          for (int c2 = 0; c2 < (ND == 3 ? ln[ND - 1] : 1); c2++) {
            const long long cc = ((long long)li[ND - 1][c2] * G.n[1] + li[1][b2]) * G.n[0] + li[0][a];
#pragma unroll
            for (int f = 0; f < NP; f++) wo[(long long)f * N + cc] = out[f];
          }
    }
  } else {
#pragma unroll
    for (int f = 0; f < NP; f++) {
      int li[3][3], ln[3] = {1, 1, 1};
      for (int d = 0; d < 3; d++) li[d][0] = d < ND ? x[d] : 0;
      for (int d = 0; d < ND; d++) ln[d] = 1 + bc_images(G.bct[f][d], x[d], G.n[d], &li[d][1]);
      for (int a = 0; a < ln[0]; a++)
        for (int b2 = 0; b2 < ln[1]; b2++)
          for (int c2 = 0; c2 < ln[2]; c2++) {
            const long long cc = ((long long)li[2][c2] * G.n[1] + li[1][b2]) * G.n[0] + li[0][a];
            wo[(long long)f * N + cc] = out[f];
          }
    }
  }
}

// Ghost fill of the evolved fields (cuboundary for every field and dir,
// src/iosmaug.c:830-833): after the reference's dir-by-dir passes a ghost
// cell holds the interior cell reached by mapping each coordinate.
template <int ND>
__global__ __launch_bounds__(256) void k_fill(Geo G, double* __restrict__ w) {
  const long long c = (long long)blockIdx.x * 256 + threadIdx.x;
  if (c >= G.N) return;
  int x[3];
  x[0] = (int)(c % G.n[0]);
  x[1] = (int)((c / G.n[0]) % G.n[1]);
  x[2] = ND == 3 ? (int)(c / ((long long)G.n[0] * G.n[1])) : 0;
  bool ghost = false;
  for (int d = 0; d < ND; d++) ghost |= (x[d] < 2 || x[d] > G.n[d] - 3);
  if (!ghost) return;
  for (int f = 0; f < G.np; f++) {
    int y[3] = {0, 0, 0};
    for (int d = 0; d < ND; d++) y[d] = bc_src(G.bct[f][d], x[d], G.n[d]);
    const long long src = ((long long)y[2] * G.n[1] + y[1]) * G.n[0] + y[0];
    w[(long long)f * G.N + c] = w[(long long)f * G.N + src];
  }
}

// ---------------------------------------------------------------------------
// launch wrappers (host)

void launch_prim(const Geo& G, hipStream_t st, const double* w, const double* bg, double* prim,
                 unsigned long long* cmax, int* flag) {
  // grid-stride: at most 2048 blocks so the per-block atomics stay few
  long long nb0 = (G.N + 255) / 256;
  const unsigned nb = (unsigned)(nb0 < 2048 ? nb0 : 2048);
  if (G.nd == 3)
    hipLaunchKernelGGL(k_prim<3>, dim3(nb), dim3(256), 0, st, G, w, bg, prim, cmax, flag);
  else
    hipLaunchKernelGGL(k_prim<2>, dim3(nb), dim3(256), 0, st, G, w, bg, prim, cmax, flag);
}

void launch_zero3(hipStream_t st, unsigned long long* p) {
  hipLaunchKernelGGL(k_zero3, dim3(1), dim3(64), 0, st, p);
}

void launch_dt(hipStream_t st, const unsigned long long* cmax, int nd, const double* dx, double courant,
               double dtcap, int moddton, double* dts) {
  hipLaunchKernelGGL(k_dt, dim3(1), dim3(64), 0, st, cmax, nd, dx[0], dx[1], nd == 3 ? dx[2] : 1.0, courant,
                     dtcap, moddton, dts);
}

void launch_rhs_v1(const Geo& G, hipStream_t st, const RhsArgs& A) {
  const int ni0 = G.n[0] - 4, ni1 = G.n[1] - 4;
  dim3 grid((ni0 + 63) / 64, (ni1 + 3) / 4, G.nd == 3 ? G.n[2] - 4 : 1);
  if (G.nd == 3)
    hipLaunchKernelGGL(k_rhs<3>, grid, dim3(64, 4), 0, st, G, A);
  else
    hipLaunchKernelGGL(k_rhs<2>, grid, dim3(64, 4), 0, st, G, A);
}

void launch_rhs(const Geo& G, hipStream_t st, const RhsArgs& A) {
  if (G.kver == 1)
    launch_rhs_v1(G, st, A);
  else
    launch_rhs_tiled(G, st, A);
}

void launch_fill(const Geo& G, hipStream_t st, double* w) {
  const unsigned nb = (unsigned)((G.N + 255) / 256);
  if (G.nd == 3)
    hipLaunchKernelGGL(k_fill<3>, dim3(nb), dim3(256), 0, st, G, w);
  else
    hipLaunchKernelGGL(k_fill<2>, dim3(nb), dim3(256), 0, st, G, w);
}

}  // namespace smaug
