// smaug_kernels.hip -- MI355X (gfx950) kernels of the SMAUG MHD stage.
//
// One RK stage (or the reference's single Euler stage) is two launches:
//   k_prim  (all cells, pointwise): velocities, total perturbation pressure,
//           the energy hyperdiffusion variable, and the per-direction wave
//           speed maxima reduced through wavefront shuffles, LDS and one
//           64-bit atomicMax per block (the reference copies the whole wd
//           array to the host for this, src/computedervfields.cu:1914).
//   k_rhs   (interior cells, smaug_rhs.hip): flux divergence
//           (centdiff1/centdiff2), background energy terms, gravity, all
//           hyperdiffusion coefficients and sources, the stage update, the
//           user source on the updated state, the RK combination and the
//           ghost fill.
// The CFL time step is chosen on the device by k_dt so a step needs no
// host round trip.
#include <math.h>

#include "../../include/smaug_hip.h"
#include "smaug_dev.h"
#include "smaug_ops.h"

namespace smaug {

__device__ inline double bits2d(unsigned long long b) { return __longlong_as_double((long long)b); }

// ---------------------------------------------------------------------------
// k_prim: pointwise derived fields + wave-speed maxima
//   v_d  = m_d/(rho+rhob)                         src/dervfields.cu:79-98
//   q_e  = e - (b^2 + m^2/rho_t)/2                src/hyperdifesource1.cu:391-395
// (p_t, src/dervfields.cu:107-153, is formed by k_rhs's flux function from
// the cell state where it is needed)
//   c_s, c_fast,d                                 src/dervfields.cu:241-309
// With SH = float (mixed precision) prim is stored in FP32 and the FP32
// shadow of the evolved fields (what the stencils of k_rhs read) is written
// from the FP64 state; the wave speeds and their maxima stay FP64.
template <int ND, typename SH>
__global__ __launch_bounds__(256) void k_prim(Geo G, const double* __restrict__ w,
                                              const double* __restrict__ bg, SH* __restrict__ prim,
                                              float* __restrict__ shadow,
                                              unsigned long long* __restrict__ cmax_bits,
                                              int* __restrict__ flag) {
  const long long N = G.N;
  double mx[3] = {0.0, 0.0, 0.0};
  int bad = 0;
  struct In {
    double rho, rhob, e, eb, m[ND], b[ND], bb[ND];
  };
  auto load = [&](long long c, In& q) {
    q.rho = w[c];
    q.rhob = bg[(long long)bRHOB() * N + c];
#pragma unroll
    for (int d = 0; d < ND; d++) {
      q.m[d] = w[(long long)fMOM(d) * N + c];
      q.b[d] = w[(long long)fB(ND, d) * N + c];
      q.bb[d] = bg[(long long)bBB(d) * N + c];
    }
    q.e = w[(long long)fE(ND) * N + c];
    q.eb = bg[(long long)bEB() * N + c];
  };
  auto cell = [&](long long c, const In& q) {
    const double gm = G.gamma;
    const double *m = q.m, *b = q.b, *bb = q.bb;
    const double rho = q.rho, e = q.e, eb = q.eb;
    const double rt = rho + q.rhob;
    double mm = m[0] * m[0], b2 = b[0] * b[0], bb2 = bb[0] * bb[0];
    double bbb_c = bb[0] * b[0];
#pragma unroll
    for (int d = 1; d < ND; d++) {
      mm = mm + m[d] * m[d];
      b2 = b2 + b[d] * b[d];
      bb2 = bb2 + bb[d] * bb[d];
      bbb_c = bbb_c + bb[d] * b[d];
    }
    if (prim) {  // null: wave speeds only (smaug_cfl_noalloc)
#pragma unroll
      for (int d = 0; d < ND; d++) prim[(long long)pV(d) * N + c] = (SH)(m[d] / rt);
      prim[(long long)pQE(ND) * N + c] = (SH)(e - 0.5 * (b2 + mm / rt));
    }
    if (shadow) {
      shadow[c] = (float)rho;
      shadow[(long long)fE(ND) * N + c] = (float)e;
#pragma unroll
      for (int d = 0; d < ND; d++) {
        shadow[(long long)fMOM(d) * N + c] = (float)m[d];
        shadow[(long long)fB(ND, d) * N + c] = (float)b[d];
      }
    }
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
  };
  const long long stride = (long long)gridDim.x * 256;
  for (long long c = (long long)blockIdx.x * 256 + threadIdx.x; c < N; c += stride) {
    In q0;
    load(c, q0);
    cell(c, q0);
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

// k_prim of the tangent-linear step: w and bg hold Dual values (value and
// tangent of every evolved and background field); v and q_e are formed in Dual arithmetic
// (the same expressions as k_prim), the wave speeds and their maxima from
// the values alone (frozen, as the hyperdiffusion coefficients they scale)
// PT = DualF (the mixed-precision run): prim rounded to FP32 duals, and the
// FP32 dual shadow of the state for the stencils
template <int ND, typename PT>
__global__ __launch_bounds__(256) void k_prim_tl(Geo G, const Dual* __restrict__ w, const Dual* __restrict__ bg,
                                                 PT* __restrict__ prim, PT* __restrict__ shadow,
                                                 unsigned long long* __restrict__ cmax_bits, int* __restrict__ flag) {
  const long long N = G.N;
  double mx[3] = {0.0, 0.0, 0.0};
  int bad = 0;
  const long long stride = (long long)gridDim.x * 256;
  for (long long c = (long long)blockIdx.x * 256 + threadIdx.x; c < N; c += stride) {
    const Dual rho = w[c], e = w[(long long)fE(ND) * N + c];
    const Dual rhob = bg[(long long)bRHOB() * N + c], eb = bg[(long long)bEB() * N + c];
    Dual m[ND], b[ND], bb[ND];
#pragma unroll
    for (int d = 0; d < ND; d++) {
      m[d] = w[(long long)fMOM(d) * N + c];
      b[d] = w[(long long)fB(ND, d) * N + c];
      bb[d] = bg[(long long)bBB(d) * N + c];
    }
    const Dual rt = rho + rhob;
    Dual mm = m[0] * m[0], b2 = b[0] * b[0];
#pragma unroll
    for (int d = 1; d < ND; d++) {
      mm = mm + m[d] * m[d];
      b2 = b2 + b[d] * b[d];
    }
#pragma unroll
    for (int d = 0; d < ND; d++) prim[(long long)pV(d) * N + c] = (PT)(m[d] / rt);
    prim[(long long)pQE(ND) * N + c] = (PT)(e - 0.5 * (b2 + mm / rt));
    if (shadow) {
      shadow[c] = (PT)rho;
      shadow[(long long)fE(ND) * N + c] = (PT)e;
#pragma unroll
      for (int d = 0; d < ND; d++) {
        shadow[(long long)fMOM(d) * N + c] = (PT)m[d];
        shadow[(long long)fB(ND, d) * N + c] = (PT)b[d];
      }
    }
    // wave speeds of the values (k_prim's expressions)
    const double gm = G.gamma, rtv = rt.v;
    double mmv = m[0].v * m[0].v, b2v = b[0].v * b[0].v, bb2 = bb[0].v * bb[0].v, bbb = bb[0].v * b[0].v;
#pragma unroll
    for (int d = 1; d < ND; d++) {
      mmv = mmv + m[d].v * m[d].v;
      b2v = b2v + b[d].v * b[d].v;
      bb2 = bb2 + bb[d].v * bb[d].v;
      bbb = bbb + bb[d].v * b[d].v;
    }
    const double pk = (gm - 1) * (e.v - 0.5 * (mmv / rtv) - 0.5 * b2v - bbb);
    const double cs2 = (gm * (pk + (gm - 1) * (eb.v - 0.5 * bb2))) / rtv;
    const double cf = (b2v + bb2 + 2.0 * bbb) / rtv + cs2;
    const double cs = sqrt(cs2);
#pragma unroll
    for (int d = 0; d < ND; d++) {
      const double btd = bb[d].v + b[d].v;
      const double cfi = 0.5 * (cf + sqrt(cf * cf - 4.0 * cs2 * (btd * btd) / rtv));
      const double cfd = sqrt(cfi) + fabs(m[d].v / rtv);
      if (!(cfd <= 1.0e300) || !(cs <= 1.0e300)) bad = 1;
      const double t = cfd > cs ? cfd : cs;
      if (t > mx[d]) mx[d] = t;
    }
  }
#pragma unroll
  for (int d = 0; d < ND; d++) {
    double t = mx[d];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      const double o = __shfl_xor(t, off, 64);
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

// Dual array of n values from value and tangent arrays, and back
__global__ __launch_bounds__(256) void k_dual_pack(long long n, const double* __restrict__ v,
                                                   const double* __restrict__ t, Dual* __restrict__ d) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i < n) d[i] = Dual(v[i], t ? t[i] : 0.0);
}
__global__ __launch_bounds__(256) void k_dual_unpack(long long n, const Dual* __restrict__ d, double* __restrict__ v,
                                                     double* __restrict__ t) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  if (v) v[i] = d[i].v;
  if (t) t[i] = d[i].t;
}

__global__ void k_zero3(unsigned long long* p) {
  if (threadIdx.x < 3) p[threadIdx.x] = 0ull;
}

__global__ void k_clear_flag(int* flag) {
  if (threadIdx.x == 0) *flag = 0;
}

// CFL step (src/iosmaug.c:541-616, quirk Q4): the last direction's maximum
// over every dir's dx.  With moddton=1 each step starts from DT (dt=DT,
// p->dt=DT, :549-550) and takes the CFL step when it lies in (1e-12, DT)
// (:568-569); otherwise the step keeps its value.  dts[0] = dt, dts[1] =
// accumulated time.
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
    dt = dtcap;
    const double dtt = courant / courantmax;
    if (dtt > 1.0e-12 && dtt < dtcap) dt = dtt;
  }
  dts[0] = dt;
  dts[1] = dts[1] + dt;
}

// Ghost fill of the evolved fields (cuboundary for every field and dir,
// src/iosmaug.c:830-833).  The reference applies the directions in order
// 0, 1, 2, each pass overwriting the ghost layers of its direction, so a
// ghost cell ends with the value of the last direction in which it is a
// ghost: a fixed ghost keeps its own value, an image-type ghost takes the
// cell its map names as it stood before that pass (the remaining
// directions), with the sign of asymm maps.
template <int ND>
__global__ __launch_bounds__(256) void k_fill(Geo G, double* __restrict__ w) {
  const long long c = (long long)blockIdx.x * 256 + threadIdx.x;
  if (c >= G.N) return;
  int x[3];
  x[0] = (int)(c % G.n[0]);
  x[1] = (int)((c / G.n[0]) % G.n[1]);
  x[2] = ND == 3 ? (int)(c / ((long long)G.n[0] * G.n[1])) : 0;
  bool ghost = false;
  for (int d = 0; d < ND; d++) ghost |= bc_is_ghost(x[d], G.n[d]);
  if (!ghost) return;
  for (int f = 0; f < G.np; f++) {
    int y[3] = {x[0], x[1], x[2]};
    double sg = 1.0;
    for (int d = ND - 1; d >= 0; d--) {
      if (!bc_is_ghost(y[d], G.n[d])) continue;
      const int t = G.bct[f][d];
      if (t == BC_FIXED) break;  // the value the cell holds
      y[d] = bc_src(t, y[d], G.n[d], ND);
      sg *= bc_sign(t);
    }
    const long long src = ((long long)y[2] * G.n[1] + y[1]) * G.n[0] + y[0];
    if (src != c) w[(long long)f * G.N + c] = sg * w[(long long)f * G.N + src];
  }
}

// ---------------------------------------------------------------------------
// launch wrappers (host)

void launch_prim(const Geo& G, hipStream_t st, const double* w, const double* bg, void* prim, float* shadow,
                 unsigned long long* cmax, int* flag) {
  // grid-stride: at most 2048 blocks so the per-block atomics stay few
  long long nb = (G.N + 255) / 256;
  if (nb > 2048) nb = 2048;
  if (G.mixed) {
    if (G.nd == 3)
      hipLaunchKernelGGL((k_prim<3, float>), dim3((unsigned)nb), dim3(256), 0, st, G, w, bg, (float*)prim, shadow,
                         cmax, flag);
    else
      hipLaunchKernelGGL((k_prim<2, float>), dim3((unsigned)nb), dim3(256), 0, st, G, w, bg, (float*)prim, shadow,
                         cmax, flag);
  } else {
    if (G.nd == 3)
      hipLaunchKernelGGL((k_prim<3, double>), dim3((unsigned)nb), dim3(256), 0, st, G, w, bg, (double*)prim,
                         (float*)nullptr, cmax, flag);
    else
      hipLaunchKernelGGL((k_prim<2, double>), dim3((unsigned)nb), dim3(256), 0, st, G, w, bg, (double*)prim,
                         (float*)nullptr, cmax, flag);
  }
}

void launch_prim_tl(const Geo& G, hipStream_t st, const Dual* w, const Dual* bg, void* prim, void* shadow,
                    unsigned long long* cmax, int* flag) {
  long long nb = (G.N + 255) / 256;
  if (nb > 2048) nb = 2048;
  const dim3 g((unsigned)nb), b(256);
  if (G.mixed) {
    if (G.nd == 3)
      hipLaunchKernelGGL((k_prim_tl<3, DualF>), g, b, 0, st, G, w, bg, (DualF*)prim, (DualF*)shadow, cmax, flag);
    else
      hipLaunchKernelGGL((k_prim_tl<2, DualF>), g, b, 0, st, G, w, bg, (DualF*)prim, (DualF*)shadow, cmax, flag);
  } else {
    if (G.nd == 3)
      hipLaunchKernelGGL((k_prim_tl<3, Dual>), g, b, 0, st, G, w, bg, (Dual*)prim, (Dual*)nullptr, cmax, flag);
    else
      hipLaunchKernelGGL((k_prim_tl<2, Dual>), g, b, 0, st, G, w, bg, (Dual*)prim, (Dual*)nullptr, cmax, flag);
  }
}

void launch_dual_pack(hipStream_t st, long long n, const double* v, const double* t, Dual* d) {
  hipLaunchKernelGGL(k_dual_pack, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, v, t, d);
}

void launch_dual_unpack(hipStream_t st, long long n, const Dual* d, double* v, double* t) {
  hipLaunchKernelGGL(k_dual_unpack, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, d, v, t);
}

void launch_zero3(hipStream_t st, unsigned long long* p) {
  hipLaunchKernelGGL(k_zero3, dim3(1), dim3(64), 0, st, p);
}

void launch_clear_flag(hipStream_t st, int* flag) { hipLaunchKernelGGL(k_clear_flag, dim3(1), dim3(64), 0, st, flag); }

// Scratch of the CFL reduction without a context (NoallocSlots layout):
// maxima and flag to 0, (dt, t) to (dt0, 0).  A kernel, not a memset, so a
// caller may capture it in a HIP graph.
__global__ void k_noalloc_init(unsigned long long* s, double dt0) {
  if (threadIdx.x < kNoallocSlots) {
    unsigned long long v = 0ull;
    if (threadIdx.x == kNaDts) v = (unsigned long long)__double_as_longlong(dt0);
    s[threadIdx.x] = v;
  }
}

// out[0..nd-1] = per-direction maxima, out[nd] = dt, out[nd+1] = 1 when a
// wave speed was not finite, else 0
__global__ void k_noalloc_out(const unsigned long long* s, int nd, double* out) {
  if (threadIdx.x != 0) return;
  for (int d = 0; d < nd; d++) out[d] = bits2d(s[kNaCmax + d]);
  out[nd] = bits2d(s[kNaDts]);
  out[nd + 1] = *reinterpret_cast<const int*>(s + kNaFlag) ? 1.0 : 0.0;
}

void launch_noalloc_init(hipStream_t st, unsigned long long* s, double dt0) {
  hipLaunchKernelGGL(k_noalloc_init, dim3(1), dim3(64), 0, st, s, dt0);
}

void launch_noalloc_out(hipStream_t st, const unsigned long long* s, int nd, double* out) {
  hipLaunchKernelGGL(k_noalloc_out, dim3(1), dim3(64), 0, st, s, nd, out);
}

void launch_dt(hipStream_t st, const unsigned long long* cmax, int nd, const double* dx, double courant,
               double dtcap, int moddton, double* dts) {
  hipLaunchKernelGGL(k_dt, dim3(1), dim3(64), 0, st, cmax, nd, dx[0], dx[1], nd == 3 ? dx[2] : 1.0, courant,
                     dtcap, moddton, dts);
}

void launch_rhs(const Geo& G, hipStream_t st, const RhsArgs& A) { launch_rhs_tiled(G, st, A); }

// per-axis factors of the vertical driver (the Gaussian of
// models/usersource.cu.fluxtube1:59-62):
// the same expressions k_rhs evaluates per cell when the grid is not separable
__global__ __launch_bounds__(256) void k_drvtab(Geo G, const double* __restrict__ pos1, double* __restrict__ dshape) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  const int n0 = G.n[0], n1 = G.n[1], n2 = G.nd == 3 ? G.n[2] : 0;
  if (t >= n0 + n1 + n2) return;
  const double* dv = G.drv;
  if (t < n0) {
    const double zp = pos1[t] - dv[2];
    dshape[t] = exp(-zp * zp / (dv[5] * dv[5]));
  } else if (t < n0 + n1) {
    const double xp = pos1[t] - dv[3];
    dshape[t] = exp(-xp * xp / (dv[6] * dv[6]));
  } else {
    const double yp = pos1[t] - dv[4];
    dshape[t] = exp(-yp * yp / (dv[6] * dv[6]));
  }
}

void launch_drvtab(const Geo& G, hipStream_t st, const double* pos1, double* dshape) {
  const int n = G.n[0] + G.n[1] + (G.nd == 3 ? G.n[2] : 0);
  hipLaunchKernelGGL(k_drvtab, dim3((n + 255) / 256), dim3(256), 0, st, G, pos1, dshape);
}

void launch_fill(const Geo& G, hipStream_t st, double* w) {
  const unsigned nb = (unsigned)((G.N + 255) / 256);
  if (G.nd == 3)
    hipLaunchKernelGGL(k_fill<3>, dim3(nb), dim3(256), 0, st, G, w);
  else
    hipLaunchKernelGGL(k_fill<2>, dim3(nb), dim3(256), 0, st, G, w);
}

}  // namespace smaug
