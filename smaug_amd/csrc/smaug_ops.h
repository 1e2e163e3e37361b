// smaug_ops.h -- device operators shared by the stage kernel (k_rhs) and
// its adjoint (k_adj_rhs): grid index helpers, the hyperdiffusion halo
// substitutions, the nu coefficient and the cell fluxes.
#pragma once
#include <math.h>

#include "smaug_dev.h"

namespace smaug {

__device__ inline long long gix(const Geo& G, int a, int b, int c) {
  return ((long long)c * G.n[1] + b) * G.n[0] + a;
}

// The same cell index as a 32-bit value: a context holds fewer than 2^29
// cells (checked at creation), so every byte offset within one field fits
// 32 bits.  at(base, i) reads element i of a field whose base pointer is
// the same for the whole wavefront, as base + 32-bit byte offset: the form
// of the global load with a scalar base and a 32-bit vector offset (no
// 64-bit address arithmetic per lane); put(base, i, v) writes it.
__device__ inline unsigned cix(const Geo& G, int a, int b, int c) {
  return ((unsigned)c * (unsigned)G.n[1] + (unsigned)b) * (unsigned)G.n[0] + (unsigned)a;
}

template <typename T>
__device__ inline T at(const T* __restrict__ base, unsigned i) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + i * (unsigned)sizeof(T));
}

template <typename T>
__device__ inline void put(T* __restrict__ base, unsigned i, T v) {
  *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + i * (unsigned)sizeof(T)) = v;
}

// |x| and max(a, b) in the arithmetic type of the call (FP64, or FP32 for
// the mixed-precision stencils); max ignores a NaN operand
__device__ inline double tabs(double x) { return fabs(x); }
__device__ inline float tabs(float x) { return fabsf(x); }
__device__ inline double tmax(double a, double b) { return fmax(a, b); }
__device__ inline float tmax(float a, float b) { return fmaxf(a, b); }

// 1/x for a positive normal x: the hardware reciprocal refined by Newton-
// Raphson steps (two in FP64, one in FP32), within an ulp or two of the
// correctly rounded quotient in 5 (FP64) or 3 (FP32) instructions instead
// of the dozen of a full-precision division.
__device__ inline double rcp_pos(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
}
__device__ inline float rcp_pos(float x) {
  const float r = __builtin_amdgcn_rcpf(x);
  const float e = fmaf(-x, r, 1.0f);
  return fmaf(r, e, r);
}

// Dual number (value v, tangent t) of the tangent-linear step: k_rhs
// instantiated with it carries d(state) along with the state.  Sums,
// products and quotients follow the rules of differentiation; comparisons
// and maxima are not defined on it (the one place they occur, the
// hyperdiffusion coefficient, is frozen below).  Trivially constructible,
// so it can fill an LDS array.
struct Dual {
  double v, t;
  Dual() = default;
  __host__ __device__ constexpr Dual(double a) : v(a), t(0.0) {}
  __host__ __device__ constexpr Dual(double a, double b) : v(a), t(b) {}
};
__host__ __device__ inline Dual operator+(Dual a, Dual b) { return Dual(a.v + b.v, a.t + b.t); }
__host__ __device__ inline Dual operator-(Dual a, Dual b) { return Dual(a.v - b.v, a.t - b.t); }
__host__ __device__ inline Dual operator-(Dual a) { return Dual(-a.v, -a.t); }
__host__ __device__ inline Dual operator*(Dual a, Dual b) { return Dual(a.v * b.v, a.t * b.v + a.v * b.t); }
__host__ __device__ inline Dual operator/(Dual a, Dual b) {
  const double q = a.v / b.v;
  return Dual(q, (a.t - q * b.t) / b.v);
}
__host__ __device__ inline Dual operator+(Dual a, double b) { return Dual(a.v + b, a.t); }
__host__ __device__ inline Dual operator+(double a, Dual b) { return Dual(a + b.v, b.t); }
__host__ __device__ inline Dual operator-(Dual a, double b) { return Dual(a.v - b, a.t); }
__host__ __device__ inline Dual operator-(double a, Dual b) { return Dual(a - b.v, -b.t); }
__host__ __device__ inline Dual operator*(Dual a, double b) { return Dual(a.v * b, a.t * b); }
__host__ __device__ inline Dual operator*(double a, Dual b) { return Dual(a * b.v, a * b.t); }
__host__ __device__ inline Dual operator/(Dual a, double b) { return Dual(a.v / b, a.t / b); }
__host__ __device__ inline Dual operator/(double a, Dual b) { return Dual(a) / b; }
__host__ __device__ inline Dual& operator+=(Dual& a, Dual b) { return a = a + b; }
__host__ __device__ inline Dual& operator-=(Dual& a, Dual b) { return a = a - b; }
__host__ __device__ inline Dual& operator*=(Dual& a, Dual b) { return a = a * b; }
__device__ inline Dual rcp_pos(Dual x) {
  const double r = rcp_pos(x.v);
  return Dual(r, -x.t * r * r);
}

// FP32 dual number: the stencils of the mixed-precision tangent-linear step
// (value and tangent rounded to FP32, as the FP32 shadows of the mixed
// forward step), over an FP64 Dual own-cell state
struct DualF {
  float v, t;
  DualF() = default;
  __host__ __device__ constexpr DualF(float a) : v(a), t(0.0f) {}
  __host__ __device__ constexpr DualF(float a, float b) : v(a), t(b) {}
  __host__ __device__ explicit DualF(double a) : v((float)a), t(0.0f) {}
  __host__ __device__ explicit DualF(int a) : v((float)a), t(0.0f) {}
  __host__ __device__ explicit DualF(const Dual& d) : v((float)d.v), t((float)d.t) {}
  __host__ __device__ explicit operator Dual() const { return Dual(v, t); }
};
__host__ __device__ inline DualF operator+(DualF a, DualF b) { return DualF(a.v + b.v, a.t + b.t); }
__host__ __device__ inline DualF operator-(DualF a, DualF b) { return DualF(a.v - b.v, a.t - b.t); }
__host__ __device__ inline DualF operator-(DualF a) { return DualF(-a.v, -a.t); }
__host__ __device__ inline DualF operator*(DualF a, DualF b) { return DualF(a.v * b.v, a.t * b.v + a.v * b.t); }
__host__ __device__ inline DualF operator/(DualF a, DualF b) {
  const float q = a.v / b.v;
  return DualF(q, (a.t - q * b.t) / b.v);
}
__host__ __device__ inline DualF operator*(DualF a, float b) { return DualF(a.v * b, a.t * b); }
__host__ __device__ inline DualF operator*(float a, DualF b) { return DualF(a * b.v, a * b.t); }
__host__ __device__ inline DualF& operator+=(DualF& a, DualF b) { return a = a + b; }
__host__ __device__ inline DualF& operator-=(DualF& a, DualF b) { return a = a - b; }
__device__ inline DualF rcp_pos(DualF x) {
  const float r = rcp_pos(x.v);
  return DualF(r, -x.t * r * r);
}

// own-cell type of the stage kernel for a stencil type: FP64 for the FP64
// and mixed-precision stencils, Dual for the tangent-linear steps
template <typename SH>
struct OwnT {
  using T = double;
};
template <>
struct OwnT<Dual> {
  using T = Dual;
};
template <>
struct OwnT<DualF> {
  using T = Dual;
};
// A value formed from a parameter: with the parameter's tangent t when the
// own-cell type carries tangents (tangent-linear run), else the value alone.
template <typename OT>
__device__ inline OT with_tangent(double v, double t);
template <>
__device__ inline double with_tangent<double>(double v, double) { return v; }
template <>
__device__ inline Dual with_tangent<Dual>(double v, double t) { return Dual(v, t); }
// a staged value in the precision sums with the background take
__device__ inline double widen(float x) { return x; }
__device__ inline double widen(double x) { return x; }
__device__ inline Dual widen(Dual x) { return x; }
__device__ inline Dual widen(DualF x) { return Dual(x.v, x.t); }

template <int ND>
__device__ inline bool in_grid(const Geo& G, const int* x) {
#pragma unroll
  for (int d = 0; d < ND; d++)
    if (x[d] < 0 || x[d] >= G.n[d]) return false;
  return true;
}

template <int ND>
__device__ inline bool interior_c(const Geo& G, const int* x) {
#pragma unroll
  for (int d = 0; d < ND; d++)
    if (x[d] < 2 || x[d] > G.n[d] - 3) return false;
  return true;
}

// Cell whose q stands at grid coordinates x, one step past the grid allowed
// in one direction: coordinate -1 reads coordinate 5 (left-chain halo, quirk
// Q8), coordinate n reads the right-chain stray halo (quirk Q7); -1 = the
// value is 0.  (Cell indices fit an int: a context holds 51 doubles per cell.)
template <int ND>
__device__ inline int qsrc(const Geo& G, const int* x) {
  int out = -1, nout = 0;
#pragma unroll
  for (int d = 0; d < ND; d++)
    if (x[d] < 0 || x[d] >= G.n[d]) {
      out = d;
      nout++;
    }
  if (nout == 0) return (int)gix(G, x[0], x[1], ND == 3 ? x[2] : 0);
  if (nout > 1) return -1;
  int y[3] = {x[0], x[1], ND == 3 ? x[2] : 0};
  if (x[out] == -1) {
    y[out] = 5;
    return (int)gix(G, y[0], y[1], y[2]);
  }
  if (x[out] == G.n[out]) {
    int pc[3];
#pragma unroll
    for (int q = 0; q < 3; q++) pc[q] = y[q] + 1;
    pc[out] = G.n[out] + 1;
    return (int)halo_r_src(G, pc);
  }
  return -1;
}

template <int ND, typename SH>
__device__ inline SH qval(const Geo& G, const SH* __restrict__ qa, const int* x) {
  const int h = qsrc<ND>(G, x);
  return h < 0 ? (SH)0 : qa[h];
}

// nu_{i+1/2}, nu_{i-1/2} from the 7 values q(i-3..i+3) along one direction:
// third differences at 4 interfaces, first differences at 6, maxima over 3
// and 5 interfaces, nu = dxk * max3 / max5 (src/hyperdifvisc1r.cu:426-708,
// src/hyperdifvisc1l.cu:723-853); dxk = delx * cmax * chyp.  T is the
// arithmetic type (double; float for the mixed-precision stencils).
template <typename T>
__device__ inline void nu7(const T* Q, T dxk, T& nur, T& nul) {
  T D3[5], D1[6];
#pragma unroll
  for (int a = 1; a <= 4; a++) D3[a] = tabs(T(3) * (Q[a + 1] - Q[a]) - (Q[a + 2] - Q[a - 1]));
#pragma unroll
  for (int a = 0; a <= 5; a++) D1[a] = tabs(Q[a + 1] - Q[a]);
  // maxima over the right (i+1/2) and left (i-1/2) windows share their
  // middle; max ignores NaN as the reference's "if (v > max)" does
  const T t3 = tmax(D3[2], D3[3]);
  const T t5 = tmax(tmax(D1[1], D1[2]), tmax(D1[3], D1[4]));
  const T m3r = tmax(tmax(t3, D3[4]), T(0)), m3l = tmax(tmax(t3, D3[1]), T(0));
  const T m5r = tmax(tmax(t5, D1[5]), T(0)), m5l = tmax(tmax(t5, D1[0]), T(0));
  // one division for both: r = 1 / (m5r m5l), a zero maximum standing in
  // as 1 (its nu is 0 anyway), then m3r / m5r = (m3r m5l) r, each ratio
  // carrying two roundings more than a quotient.  The ratios are bounded
  // (every third difference of a window is made of first differences of the
  // same window: m3 <= 6 m5), so nothing overflows as long as r does;
  // products outside the normal range take the two divisions.
  const T lo = sizeof(T) == 8 ? T(1e-290) : T(1e-30), hi = sizeof(T) == 8 ? T(1e290) : T(1e30);
  const T a = m5r > T(0) ? m5r : T(1), b = m5l > T(0) ? m5l : T(1);
  const T p = a * b;
  if (p > lo && p < hi) {
    const T r = rcp_pos(p);
    nur = m5r > T(0) ? dxk * ((m3r * b) * r) : T(0);
    nul = m5l > T(0) ? dxk * ((m3l * a) * r) : T(0);
  } else {
    nur = m5r > T(0) ? dxk * m3r / m5r : T(0);
    nul = m5l > T(0) ? dxk * m3l / m5l : T(0);
  }
}

// nu from a window of q staged in LDS (centre index ci, stride s)
template <typename SH, typename T>
__device__ inline void nu_win(const SH* __restrict__ qb, int ci, int s, T dxk, T& nur, T& nul) {
  T Q[7];
#pragma unroll
  for (int o = 0; o < 7; o++) Q[o] = (T)qb[ci + (o - 3) * s];
  nu7<T>(Q, dxk, nur, nul);
}

// the tangent-linear step holds nu at its value for the state (the
// derivative is that of the step with the windowed-maximum coefficients
// frozen, as in the adjoint): nu from the values, tangent 0
template <>
__device__ inline void nu_win<Dual, Dual>(const Dual* __restrict__ qb, int ci, int s, Dual dxk, Dual& nur,
                                          Dual& nul) {
  double Q[7];
#pragma unroll
  for (int o = 0; o < 7; o++) Q[o] = qb[ci + (o - 3) * s].v;
  double r, l;
  nu7<double>(Q, dxk.v, r, l);
  nur = Dual(r);
  nul = Dual(l);
}
template <>
__device__ inline void nu_win<DualF, DualF>(const DualF* __restrict__ qb, int ci, int s, DualF dxk, DualF& nur,
                                            DualF& nul) {
  float Q[7];
#pragma unroll
  for (int o = 0; o < 7; o++) Q[o] = qb[ci + (o - 3) * s].v;
  float r, l;
  nu7<float>(Q, dxk.v, r, l);
  nur = DualF(r);
  nul = DualF(l);
}

// fluxes of every evolved field along d at cell cc (src/centdiff1.cu:51-256,
// src/centdiff2.cu:158-356)
// the state of one cell that its fluxes need, in the arithmetic type T
template <int ND, typename T = double>
struct CellState {
  T m[ND], b[ND], bb[ND], rho, rhob, e, eb;
};

template <int ND, typename SH, typename T = double, typename BT = double>
__device__ inline void load_cell(const Geo& G, const SH* __restrict__ w, const BT* __restrict__ bg, unsigned cc,
                                 CellState<ND, T>& S) {
  const long long N = G.N;
#pragma unroll
  for (int q = 0; q < ND; q++) {
    S.m[q] = (T)at(w + (long long)fMOM(q) * N, cc);
    S.b[q] = (T)at(w + (long long)fB(ND, q) * N, cc);
    S.bb[q] = (T)at(bg + (long long)bBB(q) * N, cc);
  }
  S.rho = (T)at(w, cc);
  S.rhob = (T)at(bg + (long long)bRHOB() * N, cc);
  S.e = (T)at(w + (long long)fE(ND) * N, cc);
  S.eb = (T)at(bg + (long long)bEB() * N, cc);
}

template <int ND, typename T = double>
__device__ inline void fluxes_of(const Geo& G, const CellState<ND, T>& S, int d, T* F) {
  // v = m/rho_t and p_t are formed from the cell's state (k_prim keeps no
  // p_t, and forming v here takes fewer loads per flux cell than reading
  // its copies); p_t follows computept3, src/dervfields.cu:107-153
  const T *m = S.m, *b = S.b, *bb = S.bb;
  const T rho = S.rho, rhob = S.rhob, e = S.e, eb = S.eb;
  const T rt = rho + rhob;
  // one reciprocal for the ND + 1 quotients by rho_t (rho_t > 0; each
  // quotient then carries a rounding or two more than a division)
  const T rrt = rcp_pos(rt);
  T v[ND];
#pragma unroll
  for (int q = 0; q < ND; q++) v[q] = m[q] * rrt;
  T mm = m[0] * m[0], b2 = b[0] * b[0], bbb = b[0] * bb[0];
#pragma unroll
  for (int q = 1; q < ND; q++) {
    mm = mm + m[q] * m[q];
    b2 = b2 + b[q] * b[q];
    bbb = bbb + b[q] * bb[q];
  }
  const T gam = (T)G.gamma;
  const T pt = (gam - T(1)) * (e - T(0.5) * mm * rrt) - (gam - T(2)) * (bbb + T(0.5) * b2);
  T vd = v[0], bd = b[0], bbd = bb[0];
#pragma unroll
  for (int q = 1; q < ND; q++)
    if (q == d) {
      vd = v[q];
      bd = b[q];
      bbd = bb[q];
    }
  F[fRHO()] = vd * rho + rhob * vd;
#pragma unroll
  for (int f = 0; f < ND; f++) {
    const T mag = -(b[f] * bbd + bb[f] * bd) - b[f] * bd;
    F[fMOM(f)] = vd * m[f] + (f == d ? pt + mag : mag);
  }
  T bbv = bb[0] * v[0], bv = b[0] * v[0];
#pragma unroll
  for (int q = 1; q < ND; q++) {
    bbv = bbv + bb[q] * v[q];
    bv = bv + b[q] * v[q];
  }
  F[fE(ND)] = vd * e + ((vd * pt - bd * bbv - bbd * bv) + vd * eb - bd * bv);
#pragma unroll
  for (int f = 0; f < ND; f++)
    F[fB(ND, f)] = (f == d) ? T(0) : vd * b[f] - (bd + bbd) * v[f] + bb[f] * vd;
}

template <int ND, typename SH>
__device__ inline void fluxes2(const Geo& G, const SH* __restrict__ w, const double* __restrict__ bg, long long cc,
                               int d, double* F) {
  CellState<ND> S;
  load_cell<ND, SH>(G, w, bg, (unsigned)cc, S);
  fluxes_of<ND>(G, S, d, F);
}

}  // namespace smaug
