// smaug_dev.h -- geometry, index maps and stencil operators shared by the
// MI355X kernels (device) and the host set-up code.
//
// Layout (identical to the reference, src/gradops.cu:89-97): field-major,
// dim 0 fastest, ghost width 2.  Internally the evolved perturbation fields
// and the static background are separate arrays:
//   pert  : rho, mom[0..nd), energy, b[0..nd)         (np = 2+2*nd fields)
//   bg    : energyb, rhob, bb[0..nd)                  (nb = 2+nd fields)
//   prim  : v[0..nd), q_e                             (cell-local, per stage)
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace smaug {

constexpr int kMaxPert = 8;
constexpr int kMaxVar = 13;

// run constants passed by value to every kernel
struct Geo {
  int nd;
  int n[3];
  long long N;
  long long st[3];       // strides of dims 0,1,2
  int np;                // evolved fields
  double gamma;
  double g[3];
  double chyp[kMaxPert]; // hyperdiffusion coefficient per evolved field
  unsigned zm[kMaxVar];  // bit d: boundtype[idx][d][0] != 0 (gradient zeroing)
  int bct[kMaxPert][3];  // boundary type per evolved field and dir
  int has_fixed;         // some evolved field has a fixed (5) boundary
  int bc_uniform;        // 1 when every evolved field has the same types
  unsigned terms;        // SMAUG_T_* mask
  int hyper;             // PHYPERDIFMOM
  int driver;
  double drv[SMAUG_NDRV];
  int mixed;             // 1: stencils read FP32 shadows (SMAUG_PREC_MIXED)
  int divb;              // 1: divB source terms (divbon=1, divbfix != 0, 2D)
  double pdx[3];         // p->dx = (xmax-xmin)/NI (the divB gradients' spacing)
  double pxmin[3];       // p->xmin (with pdx, the spicule1 driver's grid positions)
};

// evolved-field indices
__host__ __device__ inline int fRHO() { return 0; }
__host__ __device__ inline int fMOM(int d) { return 1 + d; }
__host__ __device__ inline int fE(int nd) { return 1 + nd; }
__host__ __device__ inline int fB(int nd, int d) { return 2 + nd + d; }
// background indices
__host__ __device__ inline int bEB() { return 0; }
__host__ __device__ inline int bRHOB() { return 1; }
__host__ __device__ inline int bBB(int d) { return 2 + d; }
// prim indices
__host__ __device__ inline int pV(int d) { return d; }
__host__ __device__ inline int pQE(int nd) { return nd; }

// Boundary-table indices the reference's gradient calls pass as "field"
// (quirk Q2: derived and temporary arrays index boundtype[] with their own
// enum value).  Values of the DEV/TEV enums in include/iotypes.h:225-242.
__host__ __device__ inline int zFLUX(int nd) { return nd; }  // DEV flux
__host__ __device__ inline int zVEL(int d) { return d; }     // DEV vel1+d
constexpr int zRHO = 0, zTMP1 = 0, zTMP3 = 2, zTMP4 = 3, zTMP5 = 4, zTMP7 = 6, zTMP8 = 7;

// ---------------------------------------------------------------------------
// boundary index maps, per dim, for ghost width 2:
//   0 periodic  g <- g + (n-4) low, g - (n-4) high      (src/boundary.cu:132-136,209-213)
//   3 cont      0,1 <- 2;  n-2,n-1 <- n-3               (src/gradops.cu:947-1060)
//   4 contcd4   0 <- 4, 1 <- 3, n-1 <- n-5, n-2 <- n-4  (models/boundary.cu.fluxtube1:59-64)
//   5 fixed     the ghost keeps the value the initial state gives it
//               (src/gradops.cu:1201-1268, src/iosmaug.c:416-425)
//   6 symm      mirror: 3D as contcd4; 2D 0 <- 3, 1 <- 2, n-1 <- n-4, n-2 <- n-3
//   7 asymm     the symm map with the sign flipped    (src/gradops.cu:1325-1466)
constexpr int BC_PERIODIC = 0, BC_CONT = 3, BC_CONTCD4 = 4, BC_FIXED = 5, BC_SYMM = 6, BC_ASYMM = 7;

__host__ __device__ inline bool bc_is_ghost(int x, int n) { return x < 2 || x > n - 3; }
__host__ __device__ inline double bc_sign(int type) { return type == BC_ASYMM ? -1.0 : 1.0; }

// interior coordinate the ghost coordinate x copies (x itself when x is
// not a ghost, -1 for a fixed ghost)
__host__ __device__ inline int bc_src(int type, int x, int n, int nd) {
  if (!bc_is_ghost(x, n)) return x;
  switch (type) {
    case BC_PERIODIC:
      return x < 2 ? x + n - 4 : x - (n - 4);
    case BC_CONT:
      return x < 2 ? 2 : n - 3;
    case BC_FIXED:
      return -1;
    case BC_SYMM:
    case BC_ASYMM:
      if (nd == 2) return x < 2 ? 3 - x : 2 * n - 5 - x;  // 0<-3 1<-2 n-2<-n-3 n-1<-n-4
      return x < 2 ? 4 - x : 2 * n - 6 - x;                // 0<-4 1<-3 n-2<-n-4 n-1<-n-5
    default:  // contcd4
      return x < 2 ? 4 - x : 2 * n - 6 - x;
  }
}

// ghost coordinates whose value is copied from interior coordinate y (at
// most mmax of them; with n >= 8 there are at most two); returns how many
// were written to out[].  The inverse of bc_src, written out per type.
__host__ __device__ inline int bc_images(int type, int y, int n, int nd, int* out, int mmax = 4) {
  int m = 0;
  auto put = [&](int g) {
    if (m < mmax) out[m++] = g;
  };
  if (type == BC_FIXED) return 0;
  if (type == BC_PERIODIC) {
    if (y == n - 4) put(0);
    if (y == n - 3) put(1);
    if (y == 2) put(n - 2);
    if (y == 3) put(n - 1);
  } else if (type == BC_CONT) {
    if (y == 2) {
      put(0);
      put(1);
    }
    if (y == n - 3) {
      put(n - 2);
      put(n - 1);
    }
  } else if (nd == 2 && (type == BC_SYMM || type == BC_ASYMM)) {
    if (y == 3) put(0);
    if (y == 2) put(1);
    if (y == n - 4) put(n - 1);
    if (y == n - 3) put(n - 2);
  } else {  // contcd4; symm/asymm in 3D
    if (y == 4) put(0);
    if (y == 3) put(1);
    if (y == n - 5) put(n - 1);
    if (y == n - 4) put(n - 2);
  }
  return m;
}

// The same images as (count, first, second) for n >= 8, where a coordinate
// has at most two: closed forms, no array written at a run-time index.
struct Img2 {
  int m, g0, g1;
};
__host__ __device__ inline Img2 bc_images2(int type, int y, int n, int nd) {
  if (type == BC_FIXED) return Img2{0, 0, 0};
  if (type == BC_PERIODIC) {
    if (y == n - 4) return Img2{1, 0, 0};
    if (y == n - 3) return Img2{1, 1, 0};
    if (y == 2) return Img2{1, n - 2, 0};
    if (y == 3) return Img2{1, n - 1, 0};
    return Img2{0, 0, 0};
  }
  if (type == BC_CONT) {
    if (y == 2) return Img2{2, 0, 1};
    if (y == n - 3) return Img2{2, n - 2, n - 1};
    return Img2{0, 0, 0};
  }
  if (nd == 2 && (type == BC_SYMM || type == BC_ASYMM)) {
    if (y == 3) return Img2{1, 0, 0};
    if (y == 2) return Img2{1, 1, 0};
    if (y == n - 4) return Img2{1, n - 1, 0};
    if (y == n - 3) return Img2{1, n - 2, 0};
    return Img2{0, 0, 0};
  }
  // contcd4; symm/asymm in 3D
  if (y == 4) return Img2{1, 0, 0};
  if (y == 3) return Img2{1, 1, 0};
  if (y == n - 5) return Img2{1, n - 1, 0};
  if (y == n - 4) return Img2{1, n - 2, 0};
  return Img2{0, 0, 0};
}

// zeroing of grad3dn at interior cells: first/last two interior layers
__device__ inline bool z4(const Geo& G, int idx, int d, int x) {
  const int n = G.n[d];
  return ((G.zm[idx] >> d) & 1u) && (x == 2 || x == 3 || x == n - 4 || x == n - 3);
}
// zeroing of grad1l3n/grad1r3n/grad13n at cells whose other coordinates
// are inside 1..n-2: the first and last interior layer
__device__ inline bool z1(const Geo& G, int idx, int d, int x) {
  return ((G.zm[idx] >> d) & 1u) && (x == 2 || x == G.n[d] - 3);
}
// the same tests where the caller knows whether x can lie on a zeroed
// layer at all (EDGE false: x in [5, n-6], never zeroed)
template <bool EDGE>
__device__ inline bool z4e(const Geo& G, int idx, int d, int x) {
  return EDGE && z4(G, idx, d, x);
}
template <bool EDGE>
__device__ inline bool z1e(const Geo& G, int idx, int d, int x) {
  return EDGE && z1(G, idx, d, x);
}


// ---------------------------------------------------------------------------
// Right-chain hyperdiffusion halo (quirk Q7).  The reference writes q of
// cell t into the padded (n+2)^d array at linear index t + C (unpadded
// encoder on a padded array, src/hyperdifvisc1r.cu:1235,1253) in the same
// launch that zeroes that array; with threads ordered by linear index the
// stray value survives at padded cell H iff its writer t = L(H) - C is a
// valid thread not earlier than H's zeroing thread.  Returns the unpadded
// linear index of the source cell, or -1 for zero.
__host__ __device__ inline long long halo_r_src(const Geo& G, const int* pc) {
  const long long P0 = G.n[0] + 2, P1 = G.n[1] + 2;
  long long LH, C, tz;
  int z0 = pc[0] < G.n[0] - 1 ? pc[0] : G.n[0] - 1;
  int z1 = pc[1] < G.n[1] - 1 ? pc[1] : G.n[1] - 1;
  if (G.nd == 3) {
    int z2 = pc[2] < G.n[2] - 1 ? pc[2] : G.n[2] - 1;
    LH = ((long long)pc[2] * P1 + pc[1]) * P0 + pc[0];
    C = (long long)G.n[0] * G.n[1] + G.n[0] + 1;
    tz = ((long long)z2 * G.n[1] + z1) * G.n[0] + z0;
  } else {
    LH = (long long)pc[1] * P0 + pc[0];
    C = G.n[0] + 1;
    tz = (long long)z1 * G.n[0] + z0;
  }
  long long h = LH - C;
  if (h < 0 || h >= G.N || h < tz) return -1;
  return h;
}

struct RhsArgs {
  const double* win;    // stage input, evolved fields (FP64, own-cell values)
  const void* wsh;      // stage input as the stencils read it: win itself
                        // (FP64) or its FP32 shadow written by k_prim (mixed)
  const double* bg;     // background
  const void* prim;     // k_prim output of win: double (FP64) or float (mixed)
  const double* delx;   // nd*N
  const double* rdelx;  // nd*N, 1/delx
  const double* dx1;    // separable grid: delx along d at coordinate x is
  const double* rdx1;   // dx1[off(d) + x], off = 0, n0, n0+n1 (else null)
  const double* pos;    // nd*N
  const double* base;   // stage base (w^n), read at the cell only
  const double* r1;     // RK combination inputs (final stage only, else null)
  const double* r2;
  const double* r3;
  double* wout;
  const unsigned long long* cmax_bits;  // this stage's maxima
  unsigned long long* zero_next;        // maxima slot of the next stage
  const double* dts;                    // device dt
  double hfrac;                         // stage step = hfrac*dt
  double tdep;                          // driver time factor sin(2 pi qt/P)
  double damp;                          // tangent of the driver amplitude drv[0]
                                        // (tangent-linear run; else 0)
  const double* amap;                   // driver amplitude per (j,k) column, or null
  const double* dshape;                 // vertical driver on a separable grid: its
                                        // factors exp(-dz^2/sz^2) over dim 0, then
                                        // exp(-dx^2/sxy^2) over dim 1 and
                                        // exp(-dy^2/sxy^2) over dim 2 (else null)
};

// Adjoint of one stage (smaug_adj.hip).  The stage maps its input win to
//   out = base + h*acc(win) + src(out_rho)   [+ (r1 + 2 r2 + r3 - 4 base)/3]
// on interior cells, with the ghost cells of out images of interior cells.
// acc is differentiated with the hyperdiffusion coefficients nu and the
// wave-speed maxima frozen at their values for win.
struct AdjMuArgs {
  const double* lam_out;  // cotangent of out (np*N, interior cells)
  const double* outr;     // rho of the stage output as written (np*N array, field 0)
  const double* c0;       // combine inputs (base, r1, r2, r3) of the final RK stage, else null
  const double* c1;
  const double* c2;
  const double* c3;
  const double* bg;
  const double* pos;
  const double* amap;
  const double* meta;     // tape metadata of the step: [0] dt, [2] driver time factor
  double hfrac;
  double cb;              // weight of lam_out in the base cotangent (-4/3 on the final RK stage)
  int base_set;           // 1: lam_base = ..., 0: lam_base += ...
  double* mu;             // h * cotangent of acc (np*N, zero off the interior)
  double* lam_base;       // cotangent of base (interior cells)
  double* lam_r1;         // final RK stage: cotangents of r1..r3 (set), else null
  double* lam_r2;
  double* lam_r3;
  double* gcell;          // per-cell d(loss)/d(driver amplitude), or null
  int gcell_set;
};

struct AdjArgs {
  const double* win;      // stage input (FP64, ghosts filled)
  const double* bg;
  const double* prim;     // FP64 k_prim of win
  const double* delx;
  const double* rdelx;
  const double* dx1;
  const double* rdx1;
  const double* mu;
  const unsigned long long* cmax_bits;
  double* out;            // J^T mu at every cell (np*N)
};

void launch_adj_mu(const Geo& G, hipStream_t st, const AdjMuArgs& A);
void launch_adj_rhs(const Geo& G, hipStream_t st, const AdjArgs& A);
// lam += fold(T): T plus T at every ghost image, on interior cells
void launch_adj_fold(const Geo& G, hipStream_t st, const double* T, double* lam);
// gmap[col] += sum over the column of gcell
void launch_colsum(const Geo& G, hipStream_t st, const double* gcell, double* gmap);
// meta[0] = dts[0], meta[1] = qt, meta[2] = tdep
void launch_meta(hipStream_t st, double* meta, const double* dts, double qt, double tdep);

// k_prim: prim (double, or float when mixed) and, when mixed, the FP32
// shadow of the evolved fields of w
void launch_prim(const Geo& G, hipStream_t st, const double* w, const double* bg, void* prim, float* shadow,
                 unsigned long long* cmax, int* flag);
void launch_zero3(hipStream_t st, unsigned long long* p);
// the blow-up flag k_prim raises, back to 0 (a new state replaces the old)
void launch_clear_flag(hipStream_t st, int* flag);
void launch_dt(hipStream_t st, const unsigned long long* cmax, int nd, const double* dx, double courant,
               double dtcap, int moddton, double* dts);
// the scratch of smaug_cfl_noalloc, in 8-byte slots: the three maxima
// (IEEE bits, atomicMax), the blow-up flag (an int in its slot), dt and t
// of the CFL rule
constexpr int kNaCmax = 0, kNaFlag = 3, kNaDts = 4, kNoallocSlots = 6;
void launch_noalloc_init(hipStream_t st, unsigned long long* s, double dt0);
void launch_noalloc_out(hipStream_t st, const unsigned long long* s, int nd, double* out);
void launch_rhs(const Geo& G, hipStream_t st, const RhsArgs& A);
// Tangent-linear step (smaug_tl_*): the stage kernel and k_prim on Dual
// (value, tangent) arrays, with nu and the wave-speed maxima frozen
struct Dual;
void launch_rhs_tl(const Geo& G, hipStream_t st, const RhsArgs& A);
// (G.mixed: prim and shadow hold FP32 duals, the shadow of w for the
// stencils; else prim holds Dual and shadow is unused)
void launch_prim_tl(const Geo& G, hipStream_t st, const Dual* w, const Dual* bg, void* prim, void* shadow,
                    unsigned long long* cmax, int* flag);
// d[i] = (v[i], t[i]) (t null: tangent 0), and back (either output may be null)
void launch_dual_pack(hipStream_t st, long long n, const double* v, const double* t, Dual* d);
void launch_dual_unpack(hipStream_t st, long long n, const Dual* d, double* v, double* t);
void launch_rhs_tiled(const Geo& G, hipStream_t st, const RhsArgs& A);
void launch_fill(const Geo& G, hipStream_t st, double* w);
// the per-axis factors of the vertical driver's shape from per-axis
// positions pos1 (n0 + n1 [+ n2] values) into dshape (same length)
void launch_drvtab(const Geo& G, hipStream_t st, const double* pos1, double* dshape);

}  // namespace smaug
