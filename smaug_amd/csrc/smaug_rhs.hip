// smaug_rhs.hip -- tiled right-hand side of one SMAUG stage (k_rhs).
//
// One workgroup owns a tile of interior cells (8x4x8 in 3D, 16x16 in 2D,
// 256 threads, one cell each) and builds every term of the reference step
// for it from LDS-staged data, so each intermediate is computed once:
//   1. flux phase, per direction d: the eight fluxes F_f^d of the tile
//      extended by two cells along d go to LDS, then each thread takes the
//      4th-order divergence (centdiff1/centdiff2).
//   2. hyperdiffusion phase, per variable q (rho, e_int, v_f, b_f): q is
//      staged as the tile plus a 3-cell halo (with the reference's halo
//      substitutions of quirks Q7/Q8); for every direction the coefficients
//      nu_{i+-1/2} and the cross-term products of the off-diagonal momentum
//      and magnetic sources are formed once per cell of the tile extended by
//      one cell along the cross direction, then each thread applies the
//      diagonal and off-diagonal sources.  The v_f phase also applies the
//      background energy terms, which need gradients of v_f only.
//   3. gravity, the stage update, the user source, the RK combination and
//      the ghost-cell images (boundaries fused).
// The stencils work in the stencil precision SH: FP64 (the state itself),
// or, in mixed precision, FP32 -- the shadows k_prim writes, FP32 LDS boxes
// and FP32 arithmetic, which halves the LDS and most of the registers of a
// workgroup and so raises the number resident per CU.  The own-cell values,
// the background terms, the stage update (base + h*acc), the driver and the
// RK combination are FP64 in both.
#include <math.h>

#include <type_traits>

#include "../../include/smaug_hip.h"
#include "smaug_dev.h"
#include "smaug_ops.h"

namespace smaug {

template <int ND>
struct Tile;
// 3D tile 8 x 4 x 8: dim 0 (contiguous) 8 wide for coalesced rows; the
// workgroups resident on one XCD cover whole x-rows and neighbouring
// y-rows, so the tile is deep in z, the direction whose neighbour tiles are
// least likely to still be in L2 (fewer halo planes re-fetched per cell)
template <>
struct Tile<3> {
  static constexpr int X = 8, Y = 4, Z = 8;
};
template <>
struct Tile<2> {
  static constexpr int X = 16, Y = 16, Z = 1;
};

constexpr int imax(int a, int b) { return a > b ? a : b; }

template <int ND>
struct Lay {
  static constexpr int TX = Tile<ND>::X, TY = Tile<ND>::Y, TZ = Tile<ND>::Z;
  static constexpr int NP = 2 + 2 * ND;
  // q box: tile + 3-cell halo
  static constexpr int QX = TX + 6, QY = TY + 6, QZ = ND == 3 ? TZ + 6 : 1;
  static constexpr int QN = QX * QY * QZ;
  // region: tile extended by one cell along one axis
  static constexpr int RN = ND == 3 ? imax(imax((TX + 2) * TY * TZ, TX * (TY + 2) * TZ), TX * TY * (TZ + 2))
                                    : imax((TX + 2) * TY, TX * (TY + 2));
  // flux region: tile extended by two cells along one axis
  static constexpr int FN = ND == 3 ? imax(imax((TX + 4) * TY * TZ, TX * (TY + 4) * TZ), TX * TY * (TZ + 4))
                                    : imax((TX + 4) * TY, TX * (TY + 4));
  static constexpr int HYP = QN + 4 * RN;
  static constexpr int FLX = NP * FN;
  // cell box: tile + 1-cell halo; box 0 rho_t, boxes 1..ND the
  // components of the phase's vector (m in the momentum phases, b in the
  // field phases); it sits ahead of the area the phases share
  static constexpr int PX = TX + 2, PY = TY + 2, PZ = ND == 3 ? TZ + 2 : 1;
  static constexpr int PN = PX * PY * PZ;
  static constexpr int CB = (ND + 1) * PN;
  static constexpr int SHR = HYP > FLX ? HYP : FLX;  // area the phases share
  // spacing table of a tensor-product grid: delx and 1/delx at the tile's
  // coordinates along each axis (for the cross-term border cells)
  static constexpr int DXT = TX + TY + TZ;
  static constexpr int SMEM = CB + SHR + 2 * DXT;
};

// cell box of the tile plus one cell: rho_t = rho + rho_b (when with_rho)
// and the ND components f0.. of the state, in the stencil precision SH
// (cells outside the grid hold 0 and are never read).  Loads and stores are
// separate steps so the loads can share a round trip with a q box's.
template <int ND, typename SH>
struct CellBox {
  SH* CBX;  // null: no cell box with this q box
  bool with_rho;
  int f0;
  bool shell = false;  // 3D: the tile's own cells are already in the box
                       // (written from the flux phase's cell state); only
                       // the 1-cell shell around the tile is loaded
};

// cell l of the 1-cell shell of the 3D cell box (l < SHN): both z faces,
// then both y faces between them, then both x faces between those
template <int ND>
struct Shell {
  using L = Lay<ND>;
  static constexpr int ZF = L::PX * L::PY, YF = L::PX * (L::PZ - 2), XF = (L::PY - 2) * (L::PZ - 2);
  static constexpr int SHN = 2 * (ZF + YF + XF);
  __device__ static inline void cell(int l, int& a, int& b, int& c) {
    if (l < 2 * ZF) {
      const int side = l >= ZF ? 1 : 0, r = l - side * ZF;
      a = r % L::PX;
      b = r / L::PX;
      c = side ? L::PZ - 1 : 0;
    } else if (l < 2 * (ZF + YF)) {
      const int r0 = l - 2 * ZF, side = r0 >= YF ? 1 : 0, r = r0 - side * YF;
      a = r % L::PX;
      b = side ? L::PY - 1 : 0;
      c = 1 + r / L::PX;
    } else {
      const int r0 = l - 2 * (ZF + YF), side = r0 >= XF ? 1 : 0, r = r0 - side * XF;
      a = side ? L::PX - 1 : 0;
      b = 1 + r % (L::PY - 2);
      c = 1 + r / (L::PY - 2);
    }
  }
};

template <int ND, typename SH>
struct CellLoads {
  static constexpr int NK = (Lay<ND>::PN + 255) / 256;
  SH v[NK][ND + 1];
};

template <int ND, typename SH, bool EDGE>
__device__ inline void cells_load(const Geo& G, const SH* __restrict__ ws, const typename OwnT<SH>::T* __restrict__ bg,
                                  const CellBox<ND, SH>& C, const int* o, CellLoads<ND, SH>& R) {
  using L = Lay<ND>;
  constexpr int NK = CellLoads<ND, SH>::NK;
  int g[NK];  // cell of each slot, -1 outside the grid (or past the box)
#pragma unroll
  for (int k = 0; k < NK; k++) {
    const int l = threadIdx.x + 256 * k;
    int a = l % L::PX, b = (l / L::PX) % L::PY, cz = l / (L::PX * L::PY);
    if (ND == 3 && C.shell && l < Shell<ND>::SHN) Shell<ND>::cell(l, a, b, cz);
    const int x[3] = {o[0] - 1 + a, o[1] - 1 + b, ND == 3 ? o[2] - 1 + cz : 0};
    g[k] = l < (ND == 3 && C.shell ? Shell<ND>::SHN : L::PN) && (!EDGE || in_grid<ND>(G, x)) ? (int)cix(G, x[0], x[1], x[2]) : -1;
  }
#pragma unroll
  for (int k = 0; k < NK; k++) {
    R.v[k][0] = C.with_rho && g[k] >= 0 ? (SH)(widen(at(ws, (unsigned)g[k])) + at(bg + (long long)bRHOB() * G.N, (unsigned)g[k]))
                                        : SH(0);
#pragma unroll
    for (int d = 0; d < ND; d++)
      R.v[k][1 + d] = g[k] >= 0 ? at(ws + (long long)(C.f0 + d) * G.N, (unsigned)g[k]) : SH(0);
  }
}

template <int ND, typename SH>
__device__ inline void cells_store(const CellBox<ND, SH>& C, const CellLoads<ND, SH>& R) {
  using L = Lay<ND>;
#pragma unroll
  for (int k = 0; k < CellLoads<ND, SH>::NK; k++) {
    const int l0 = threadIdx.x + 256 * k;
    int l = l0;
    if (ND == 3 && C.shell) {
      if (l0 >= Shell<ND>::SHN) continue;
      int a, b, cz;
      Shell<ND>::cell(l0, a, b, cz);
      l = a + L::PX * (b + L::PY * cz);
    }
    if (l < L::PN) {
      if (C.with_rho) C.CBX[l] = R.v[k][0];
#pragma unroll
      for (int d = 0; d < ND; d++) C.CBX[(1 + d) * L::PN + l] = R.v[k][1 + d];
    }
  }
}

// q box of a tile whose box crosses the grid edge (with the halo
// substitutions): in batches of NB cells per thread, every source first,
// then the NB loads in flight together, then the stores
template <int ND, typename SH>
__device__ inline void stage_q_edge(const Geo& G, const SH* __restrict__ qa, SH* __restrict__ qb, const int* o) {
  using L = Lay<ND>;
  constexpr int NB = ND == 3 ? 4 : 1;
#pragma unroll 1
  for (int l0 = threadIdx.x; l0 < L::QN; l0 += 256 * NB) {
    int src[NB];
#pragma unroll
    for (int k = 0; k < NB; k++) {
      const int l = l0 + 256 * k;
      src[k] = -1;
      if (l < L::QN) {
        const int a = l % L::QX, b = (l / L::QX) % L::QY, c = l / (L::QX * L::QY);
        const int x[3] = {o[0] - 3 + a, o[1] - 3 + b, ND == 3 ? o[2] - 3 + c : 0};
        src[k] = qsrc<ND>(G, x);
      }
    }
    SH v[NB];
#pragma unroll
    for (int k = 0; k < NB; k++) v[k] = src[k] >= 0 ? at(qa, (unsigned)src[k]) : (SH)0;
#pragma unroll
    for (int k = 0; k < NB; k++)
      if (l0 + 256 * k < L::QN) qb[l0 + 256 * k] = v[k];
  }
}

// q box (tile + 3-cell halo) of field qa into qb, with it the cell box C
// when C.CBX is set and a second q box (qa2 into qb2) when qa2 is set: all
// of their loads are issued before the first store.  EDGE false: the
// caller's tile lies at least five cells inside the grid (the box is inside)
template <int ND, typename SH, bool EDGE>
__device__ inline void stage_q(const Geo& G, const SH* __restrict__ qa, SH* __restrict__ qb, const int* o,
                               const SH* __restrict__ ws = nullptr, const typename OwnT<SH>::T* __restrict__ bg = nullptr,
                               CellBox<ND, SH> C = CellBox<ND, SH>{nullptr, false, 0}, const SH* __restrict__ qa2 = nullptr,
                               SH* __restrict__ qb2 = nullptr) {
  using L = Lay<ND>;
  CellLoads<ND, SH> cl;
  // the wave issues its staging loads at raised priority, so they leave
  // ahead of the other workgroups' stencil arithmetic on the SIMD, and drops
  // back once they are in flight (FP64 0.6 % faster at 256^3; the flux
  // phase's cell loads measured slower with it)
  __builtin_amdgcn_s_setprio(3);
  if (C.CBX) cells_load<ND, SH, EDGE>(G, ws, bg, C, o, cl);
  // fast path: the halo box lies inside the grid, no substitutions needed;
  // each thread owns (a,b) columns of the box and walks their planes
  bool inside = !EDGE || (o[0] - 3 >= 0 && o[1] - 3 >= 0 && o[0] + L::TX + 3 <= G.n[0] && o[1] + L::TY + 3 <= G.n[1]);
  if (EDGE && ND == 3) inside = inside && o[2] - 3 >= 0 && o[2] + L::TZ + 3 <= G.n[2];
  if (inside) {
    constexpr int PL = L::QX * L::QY, NKF = (PL + 255) / 256;
    const unsigned zs = (unsigned)G.n[0] * (unsigned)G.n[1];
    SH v[NKF][L::QZ], v2[NKF][L::QZ];
#pragma unroll
    for (int k = 0; k < NKF; k++) {
      const int l = threadIdx.x + 256 * k;
      if (l < PL) {
        const int a = l % L::QX, b = l / L::QX;
        const unsigned g0 = cix(G, o[0] - 3 + a, o[1] - 3 + b, ND == 3 ? o[2] - 3 : 0);
#pragma unroll
        for (int cz = 0; cz < L::QZ; cz++) v[k][cz] = at(qa, g0 + cz * zs);
        if (qa2) {
#pragma unroll
          for (int cz = 0; cz < L::QZ; cz++) v2[k][cz] = at(qa2, g0 + cz * zs);
        }
      }
    }
    __builtin_amdgcn_s_setprio(0);
    if (C.CBX) cells_store<ND, SH>(C, cl);
#pragma unroll
    for (int k = 0; k < NKF; k++) {
      const int l = threadIdx.x + 256 * k;
      if (l < PL) {
#pragma unroll
        for (int cz = 0; cz < L::QZ; cz++) qb[cz * PL + l] = v[k][cz];
        if (qa2) {
#pragma unroll
          for (int cz = 0; cz < L::QZ; cz++) qb2[cz * PL + l] = v2[k][cz];
        }
      }
    }
    return;
  }
  if constexpr (EDGE) {
    if (C.CBX) cells_store<ND, SH>(C, cl);
    stage_q_edge<ND, SH>(G, qa, qb, o);
    if (qa2) stage_q_edge<ND, SH>(G, qa2, qb2, o);
  }
  __builtin_amdgcn_s_setprio(0);
}

// The terms of one tile whose origin (first interior cell) is o.  EDGE
// false: every cell of the tile lies in [5, n-6] along every direction, so
// no cell of it or of its halos is a ghost or on a zeroed layer, every
// staged box lies inside the grid and no cell has a ghost image; the
// boundary checks are then compiled out (the arithmetic is the same).  The
// own cell's interior test stays a run-time test in both bodies: known true,
// it lets the compiler keep values loaded in one phase live through later
// ones, and the FP64 kernel then spills (336 bytes per lane at gfx950).
template <int ND, typename SH, bool EDGE>
__device__ __forceinline__ void rhs_tile(const Geo& G, const RhsArgs& A, SH* __restrict__ sm, const int* o) {
  using L = Lay<ND>;
  using AT = SH;  // arithmetic type of the stencils
  // own-cell type: FP64 (the own-cell update and the RK combination are
  // FP64 with FP32 stencils too), Dual in the tangent-linear step
  using OT = typename OwnT<SH>::T;
  constexpr int NP = L::NP;
  constexpr int TX = L::TX, TY = L::TY, TZ = L::TZ;
  const int tid = threadIdx.x;
  const int lc[3] = {tid % TX, (tid / TX) % TY, ND == 3 ? tid / (TX * TY) : 0};
  int x[3] = {o[0] + lc[0], o[1] + lc[1], o[2] + lc[2]};
  const bool valid = interior_c<ND>(G, x);
  const long long N = G.N;
  const unsigned c = cix(G, x[0], x[1], x[2]);  // own cell (used by valid threads only)
  const OT* __restrict__ w = (const OT*)A.win;    // own-cell values
  const SH* __restrict__ ws = (const SH*)A.wsh;    // neighbour values (stencil precision)
  const OT* __restrict__ bg = (const OT*)A.bg;  // background (Dual in the tangent-linear step)
  const SH* __restrict__ pr = (const SH*)A.prim;
  const unsigned T = G.terms;
  AT acc[NP];
#pragma unroll
  for (int f = 0; f < NP; f++) acc[f] = AT(0);
  AT dx[ND], rdx[ND];
#pragma unroll
  for (int d = 0; d < ND; d++) {
    if (A.dx1) {  // tensor-product grid: one spacing per coordinate
      const int od = d == 0 ? 0 : (d == 1 ? G.n[0] : G.n[0] + G.n[1]);
      dx[d] = valid ? (AT)A.dx1[od + x[d]] : AT(1);
      rdx[d] = valid ? (AT)A.rdx1[od + x[d]] : AT(1);
    } else {
      dx[d] = valid ? (AT)at(A.delx + (long long)d * N, c) : AT(1);
      rdx[d] = valid ? (AT)at(A.rdelx + (long long)d * N, c) : AT(1);
    }
  }

  AT* const SDX = sm + L::CB + L::SHR;  // spacing table (tensor-product grids)
  if (A.dx1 && tid < L::DXT) {  // visible after the first barrier
    const int d = tid < TX ? 0 : (tid < TX + TY ? 1 : 2);
    const int xx = o[d] + tid - (d == 0 ? 0 : (d == 1 ? TX : TX + TY));
    const int od = d == 0 ? 0 : (d == 1 ? G.n[0] : G.n[0] + G.n[1]);
    const bool ing = d < ND && xx < G.n[d];
    SDX[tid] = ing ? (AT)A.dx1[od + xx] : AT(1);
    SDX[L::DXT + tid] = ing ? (AT)A.rdx1[od + xx] : AT(1);
  }

  AT* const CBX = sm;  // cell box (tile + 1)
  const int pc = (lc[0] + 1) + L::PX * ((lc[1] + 1) + L::PY * (ND == 3 ? lc[2] + 1 : 0));
  const int ps[3] = {1, L::PX, L::PX * L::PY};
  const bool want_mom = G.hyper && (T & SMAUG_T_HMOM);
  const bool want_eterm = (T & SMAUG_T_ETERM) != 0;

  // FP64 3D: the flux phase's own-cell state fills the tile's cells of the
  // momentum phases' cell box (the same rho + rho_b sum as the staging), so
  // that staging loads only the box's shell
  const bool own_cb = sizeof(SH) == 8 && ND == 3 && want_mom && (T & SMAUG_T_FLUX);
  // ------------------------------------------------------------ fluxes
  if (T & SMAUG_T_FLUX) {
    AT* FB = sm + L::CB;
    // the thread's own tile cell: its state is loaded once for every
    // direction (and the flux function's v, p_t, ... of it are shared); a
    // second slot per direction covers the 2 + 2 halo layers along it
    const bool in0 = !EDGE || in_grid<ND>(G, x);
    CellState<ND, AT> S0;
    if (in0) load_cell<ND, SH, AT>(G, ws, bg, c, S0);
    if (own_cb) {  // the tile's cells of the momentum phases' cell box: rho_t, m
      CBX[pc] = in0 ? S0.rho + S0.rhob : AT(0);
#pragma unroll
      for (int q = 0; q < ND; q++) CBX[(1 + q) * L::PN + pc] = in0 ? S0.m[q] : AT(0);
    }
#pragma unroll
    for (int d = 0; d < ND; d++) {
      const int ex[3] = {d == 0 ? 2 : 0, d == 1 ? 2 : 0, d == 2 ? 2 : 0};
      const int rx = TX + 2 * ex[0], ry = TY + 2 * ex[1];
      const int Td = d == 0 ? TX : (d == 1 ? TY : TZ);
      const int pa = d == 0 ? 1 : 0, qa = d == 2 ? 1 : 2;  // the other two axes
      const int Tp = pa == 0 ? TX : TY;
      const int Fd = Tp * (ND == 3 ? (qa == 1 ? TY : TZ) : 1);  // face of the tile across d
      static_assert(ND == 3 ? 4 * TY * TZ <= 256 && 4 * TX * TZ <= 256 && 4 * TX * TY <= 256
                            : 4 * TX <= 256 && 4 * TY <= 256,
                    "the halo layers of a direction fit one slot per thread");
      CellState<ND, AT> S1;
      bool in1 = false;
      int l1 = 0;
      const bool h1 = tid < 4 * Fd;
      if (h1) {
        const int layer = tid / Fd, r = tid - layer * Fd;
        int rc[3] = {0, 0, 0};
        rc[d] = layer < 2 ? layer : Td + layer;  // region coordinates 0, 1 and Td+2, Td+3
        rc[pa] = r % Tp;
        if (ND == 3) rc[qa] = r / Tp;
        const int xr[3] = {o[0] - ex[0] + rc[0], o[1] - ex[1] + rc[1], ND == 3 ? o[2] - ex[2] + rc[2] : 0};
        in1 = !EDGE || in_grid<ND>(G, xr);
        l1 = rc[0] + rx * (rc[1] + ry * rc[2]);
        if (in1) load_cell<ND, SH, AT>(G, ws, bg, cix(G, xr[0], xr[1], xr[2]), S1);
      }
      {
        AT F[NP];
        if (in0) {
          fluxes_of<ND, AT>(G, S0, d, F);
        } else {
#pragma unroll
          for (int f = 0; f < NP; f++) F[f] = AT(0);
        }
        const int l0 = (lc[0] + ex[0]) + rx * ((lc[1] + ex[1]) + ry * (lc[2] + ex[2]));
        // the flux of b_d along d is zero: its slot is neither written nor read
#pragma unroll
        for (int f = 0; f < NP; f++)
          if (f != fB(ND, d)) FB[f * L::FN + l0] = F[f];
      }
      if (h1) {
        AT F[NP];
        if (in1) {
          fluxes_of<ND, AT>(G, S1, d, F);
        } else {
#pragma unroll
          for (int f = 0; f < NP; f++) F[f] = AT(0);
        }
#pragma unroll
        for (int f = 0; f < NP; f++)
          if (f != fB(ND, d)) FB[f * L::FN + l1] = F[f];
      }
      __syncthreads();
      if (valid && !z4e<EDGE>(G, zFLUX(ND), d, x[d])) {
        const int s = d == 0 ? 1 : (d == 1 ? rx : rx * ry);
        const int li = (lc[0] + ex[0]) + rx * ((lc[1] + ex[1]) + ry * (lc[2] + ex[2]));
        const AT inv = rdx[d] * AT(1.0 / 12.0);
#pragma unroll
        for (int f = 0; f < NP; f++) {
          if (f == fB(ND, d)) continue;
          const AT* Fp = FB + f * L::FN + li;
          acc[f] -= ((8 * Fp[s] - 8 * Fp[-s]) + (Fp[-2 * s] - Fp[2 * s])) * inv;
        }
      }
      __syncthreads();
    }
  }

  const OT rho_c = valid ? at(w, c) : OT(0.0);
  const OT rhob_c = valid ? at(bg + (long long)bRHOB() * N, c) : OT(1.0);
  const AT rt_c = (AT)(rho_c + rhob_c);

  // ------------------------------------------------------------ hyperdiffusion + energy terms
  SH* qb = sm + L::CB;                                 // q box
  AT* NR = sm + L::CB + L::QN;                         // region arrays
  AT* RG[4] = {NR, NR + L::RN, NR + 2 * L::RN, NR + 3 * L::RN};  // GA/GB of the two cross directions
  // tile-cell index in the q box and its strides
  const int qc = (lc[0] + 3) + L::QX * ((lc[1] + 3) + L::QY * (ND == 3 ? lc[2] + 3 : 0));
  const int qs[3] = {1, L::QX, L::QX * L::QY};

  AT cm[ND];
#pragma unroll
  for (int d = 0; d < ND; d++) cm[d] = (AT)__longlong_as_double((long long)A.cmax_bits[d]);

  // rho (src/hyperdifrhosource1.cu) and e_int (src/hyperdifesource1.cu):
  // both boxes are staged in one round trip, the e_int box after the rho box
  const bool h_rho = G.hyper && (T & SMAUG_T_HRHO), h_e = G.hyper && (T & SMAUG_T_HE);
  static_assert(2 * L::QN <= L::SHR, "two q boxes in the shared area");
  if (h_rho || h_e) {
    const SH* qe = pr + (long long)pQE(ND) * N;
    SH* qbe = h_rho ? qb + L::QN : qb;
    stage_q<ND, SH, EDGE>(G, h_rho ? ws : qe, qb, o, ws, bg, CellBox<ND, SH>{nullptr, false, 0}, h_rho && h_e ? qe : nullptr,
                    h_rho && h_e ? qbe : nullptr);
    __syncthreads();
    if (valid && h_rho) {
#pragma unroll
      for (int dim = 0; dim < ND; dim++) {
        AT nur, nul;
        nu_win<SH, AT>(qb, qc, qs[dim], dx[dim] * cm[dim] * (AT)G.chyp[fRHO()], nur, nul);
        const bool z = z1e<EDGE>(G, zRHO, dim, x[dim]);
        const AT q0 = qb[qc];
        const AT gr = z ? AT(0) : (qb[qc + qs[dim]] - q0) * rdx[dim];
        const AT gl = z ? AT(0) : (q0 - qb[qc - qs[dim]]) * rdx[dim];
        acc[fRHO()] += (nur * gr - nul * gl) * rdx[dim];
      }
    }
    if (valid && h_e) {
#pragma unroll
      for (int dim = 0; dim < ND; dim++) {
        AT nur, nul;
        nu_win<SH, AT>(qbe, qc, qs[dim], dx[dim] * cm[dim] * (AT)G.chyp[fE(ND)], nur, nul);
        const bool z = z1e<EDGE>(G, zTMP1, dim, x[dim]);
        const AT q0 = qbe[qc];
        const AT gr = z ? AT(0) : (qbe[qc + qs[dim]] - q0) * rdx[dim];
        const AT gl = z ? AT(0) : (q0 - qbe[qc - qs[dim]]) * rdx[dim];
        acc[fE(ND)] += (nur * gr - nul * gl) * rdx[dim];
      }
    }
    __syncthreads();
  }

  AT ptb = AT(0), bbc[ND];  // background terms, formed in FP64
  if (want_eterm && valid) {
    OT bb2 = OT(0.0), bbd[ND];
#pragma unroll
    for (int d = 0; d < ND; d++) {
      bbd[d] = at(bg + (long long)bBB(d) * N, c);
      bbc[d] = (AT)bbd[d];
      bb2 = d == 0 ? bbd[d] * bbd[d] : bb2 + bbd[d] * bbd[d];
    }
    ptb = (AT)((G.gamma - 1) * at(bg + (long long)bEB() * N, c) - 0.5 * (G.gamma - 2) * bb2);
  }

  // momentum phases (q = v_f): energy terms, diagonal and off-diagonal sources
  if (want_mom || want_eterm) {
#pragma unroll
    for (int f = 0; f < ND; f++) {
      // the first box brings rho_t and m on the cell box with it
      stage_q<ND, SH, EDGE>(G, pr + (long long)pV(f) * N, qb, o, ws, bg,
                      CellBox<ND, SH>{f == 0 && want_mom ? CBX : nullptr, true, fMOM(0), own_cb});
      __syncthreads();
      if (want_eterm && valid) {  // addenergyterms_cd2 for dir f
        AT gf[ND];
#pragma unroll
        for (int i2 = 0; i2 < ND; i2++) {
          const int s = qs[i2];
          gf[i2] = z4e<EDGE>(G, zVEL(f), i2, x[i2])
                       ? AT(0)
                       : ((8 * qb[qc + s] - 8 * qb[qc - s]) + (qb[qc - 2 * s] - qb[qc + 2 * s])) * (rdx[i2] * AT(1.0 / 12.0));
        }
        AT t = -ptb * gf[f];
#pragma unroll
        for (int i2 = 0; i2 < ND; i2++) t += bbc[i2] * bbc[f] * gf[i2];
        acc[fE(ND)] += t;
      }
      if (want_mom) {
        // one region pass per f for both cross directions: the tile
        // extended by one layer along f, products rho_t*t6 and t6*m_d for
        // each d != f; the own cell's nu in every direction stays in
        // registers (src/hyperdifmomsource1.cu, src/hyperdifmomsourcene1.cu)
        const int ex[3] = {f == 0 ? 1 : 0, f == 1 ? 1 : 0, f == 2 ? 1 : 0};
        const int rx = TX + 2 * ex[0], ry = TY + 2 * ex[1];
        const int sf = f == 0 ? 1 : (f == 1 ? rx : rx * ry);
        const int ri = (lc[0] + ex[0]) + rx * ((lc[1] + ex[1]) + ry * (lc[2] + ex[2]));
        const AT Kc = (AT)G.chyp[fMOM(f)];
        AT nur[ND], nul[ND];
#pragma unroll
        for (int d = 0; d < ND; d++) {
          nur[d] = nul[d] = AT(0);
          if (valid) nu_win<SH, AT>(qb, qc, qs[d], dx[d] * cm[d] * Kc, nur[d], nul[d]);
          if (d != f) {
            const int k = d < f ? d : d - 1;
            AT g7 = 0, g8 = 0;
            if (valid) {
              const AT t5 =
                  z1e<EDGE>(G, zTMP4, d, x[d]) ? AT(0) : (qb[qc + qs[d]] - qb[qc - qs[d]]) * rdx[d] * AT(0.5);
              const AT t6 = t5 * (nur[d] + nul[d]) * AT(0.25);
              g7 = rt_c * t6;
              g8 = t6 * CBX[(1 + d) * L::PN + pc];
            }
            RG[2 * k][ri] = g7;
            RG[2 * k + 1][ri] = g8;
          }
        }
        {  // the two layers just outside the tile along f
          const int pd = f == 0 ? 1 : 0, qd = f == 2 ? 1 : 2;
          const int Tp = pd == 0 ? TX : TY, Tq = ND == 3 ? (qd == 1 ? TY : TZ) : 1;
          const int Fc = Tp * Tq;
          // FP64: one (border cell, cross direction) task per thread, 2 Fc
          // (ND-1) <= 256 tasks; kt is uniform per wavefront (2 Fc is 64 or
          // 128 in 3D), so each wavefront runs one direction's body (measured
          // 2 % faster at 256^3).  FP32: one border cell per thread for all
          // cross directions (the split measured 1.6 % slower there).
          constexpr int NTK = sizeof(AT) == 8 ? ND - 1 : 1;
          if (tid < 2 * Fc * NTK) {
            const int kt = NTK > 1 ? tid / (2 * Fc) : -1, r2 = tid - (NTK > 1 ? kt * 2 * Fc : 0);
            const int side = r2 >= Fc ? 1 : 0, r = r2 - side * Fc;
            int bc[3] = {0, 0, 0};
            bc[f] = side ? (f == 0 ? TX : (f == 1 ? TY : TZ)) + 1 : 0;
            bc[pd] = r % Tp;
            if (ND == 3) bc[qd] = r / Tp;
            const int l = bc[0] + rx * (bc[1] + ry * bc[2]);
            const int rel[3] = {bc[0] - ex[0], bc[1] - ex[1], ND == 3 ? bc[2] - ex[2] : 0};
            int xr[3] = {o[0] + rel[0], o[1] + rel[1], ND == 3 ? o[2] + rel[2] : 0};
            const bool in = !EDGE || interior_c<ND>(G, xr);
            const unsigned gq = in ? cix(G, xr[0], xr[1], xr[2]) : 0u;
            const int pi = (rel[0] + 1) + L::PX * ((rel[1] + 1) + L::PY * (ND == 3 ? rel[2] + 1 : 0));
            const AT rt = in ? CBX[pi] : AT(0);
            const int qi = (rel[0] + 3) + L::QX * ((rel[1] + 3) + L::QY * (ND == 3 ? rel[2] + 3 : 0));
#pragma unroll
            for (int d = 0; d < ND; d++) {
              if (d == f) continue;
              const int k = d < f ? d : d - 1;
              if (NTK > 1 && k != kt) continue;
              AT g7 = 0, g8 = 0;
              if (in) {
                // d != f: the border cell's coordinate along d is a tile coordinate
                const int ti = (d == 0 ? 0 : (d == 1 ? TX : TX + TY)) + rel[d];
                const AT dxr = A.dx1 ? SDX[ti] : (AT)at(A.delx + (long long)d * N, gq);
                const AT rdxr = A.dx1 ? SDX[L::DXT + ti] : (AT)at(A.rdelx + (long long)d * N, gq);
                AT nr, nl;
                nu_win<SH, AT>(qb, qi, qs[d], dxr * cm[d] * Kc, nr, nl);
                const AT t5 =
                    z1e<EDGE>(G, zTMP4, d, xr[d]) ? AT(0) : (qb[qi + qs[d]] - qb[qi - qs[d]]) * rdxr * AT(0.5);
                const AT t6 = t5 * (nr + nl) * AT(0.25);
                g7 = rt * t6;
                g8 = t6 * CBX[(1 + d) * L::PN + pi];
              }
              RG[2 * k][l] = g7;
              RG[2 * k + 1][l] = g8;
            }
          }
        }
        __syncthreads();
        if (valid) {
          const AT* mf = CBX + (1 + f) * L::PN + pc;
          const AT mc = mf[0];
#pragma unroll
          for (int d = 0; d < ND; d++) {
            const int s = qs[d];
            const int sp = ps[d];
            const bool z = z1e<EDGE>(G, zTMP4, d, x[d]);
            const AT q0 = qb[qc];
            const AT gr = z ? AT(0) : (qb[qc + s] - q0) * rdx[d];
            const AT gl = z ? AT(0) : (q0 - qb[qc - s]) * rdx[d];
            const AT rtp = CBX[pc + sp], rtm = CBX[pc - sp];
            const AT rr = (rt_c + rtp) * AT(0.5), rl = (rt_c + rtm) * AT(0.5);
            const AT mr = (mc + mf[sp]) * AT(0.5), ml = (mc + mf[-sp]) * AT(0.5);
            const AT dmv = (rr * nur[d] * gr - rl * nul[d] * gl) * (rdx[d] * AT(0.5));
            const AT dev = (mr * nur[d] * gr - ml * nul[d] * gl) * (rdx[d] * AT(0.5));
            acc[fMOM(f)] += dmv;
            acc[fE(ND)] += dev;
            if (d == f) {
              acc[fMOM(f)] += dmv;
              acc[fE(ND)] += dev;
            } else {
              const int k = d < f ? d : d - 1;
              const AT* GAk = RG[2 * k];
              const AT* GBk = RG[2 * k + 1];
              const AT o7 = z1e<EDGE>(G, zTMP7, f, x[f]) ? AT(0) : (GAk[ri + sf] - GAk[ri - sf]) * (rdx[f] * AT(0.5));
              const AT o8 = z1e<EDGE>(G, zTMP8, f, x[f]) ? AT(0) : (GBk[ri + sf] - GBk[ri - sf]) * (rdx[f] * AT(0.5));
              acc[fMOM(d)] += o7;
              acc[fE(ND)] += o8;
            }
          }
        }
      }
      __syncthreads();
    }
  }

  // magnetic phases (q = b_f): curl-form cross terms (sb = -1,
  // src/hyperdifbsourcene1.cu) and diagonal terms (sb = +1,
  // src/hyperdifbsource1.cu), one region pass per f for both d != f
  if (G.hyper && (T & SMAUG_T_HB)) {
#pragma unroll
    for (int f = 0; f < ND; f++) {
      // the first box brings b on the cell box with it (the momentum
      // phases are done with the cell box: each ended with a barrier)
      stage_q<ND, SH, EDGE>(G, ws + (long long)fB(ND, f) * N, qb, o, ws, bg,
                      CellBox<ND, SH>{f == 0 ? CBX : nullptr, false, fB(ND, 0)});
      __syncthreads();
      const int ex[3] = {f == 0 ? 1 : 0, f == 1 ? 1 : 0, f == 2 ? 1 : 0};
      const int rx = TX + 2 * ex[0], ry = TY + 2 * ex[1];
      const int sf = f == 0 ? 1 : (f == 1 ? rx : rx * ry);
      const int ri = (lc[0] + ex[0]) + rx * ((lc[1] + ex[1]) + ry * (lc[2] + ex[2]));
      const AT Kc = (AT)G.chyp[fB(ND, f)];
      AT nur[ND], nul[ND];
#pragma unroll
      for (int d = 0; d < ND; d++) {
        nur[d] = nul[d] = AT(0);
        if (d == f) continue;
        const int k = d < f ? d : d - 1;
        AT g3 = 0, g5 = 0;
        if (valid) {
          nu_win<SH, AT>(qb, qc, qs[d], dx[d] * cm[d] * Kc, nur[d], nul[d]);
          const AT t2 =
              z1e<EDGE>(G, zTMP1, d, x[d]) ? AT(0) : (qb[qc + qs[d]] - qb[qc - qs[d]]) * rdx[d] * AT(0.5);
          const AT t3 = t2 * (nul[d] + nur[d]) * AT(0.5);
          g3 = t3;
          g5 = t3 * CBX[(1 + d) * L::PN + pc];
        }
        RG[2 * k][ri] = g3;
        RG[2 * k + 1][ri] = g5;
      }
      {  // the two layers just outside the tile along f
        const int pd = f == 0 ? 1 : 0, qd = f == 2 ? 1 : 2;
        const int Tp = pd == 0 ? TX : TY, Tq = ND == 3 ? (qd == 1 ? TY : TZ) : 1;
        const int Fc = Tp * Tq;
        // FP64: one (border cell, cross direction) task per thread, 2 Fc
        // (ND-1) <= 256 tasks; kt is uniform per wavefront (2 Fc is 64 or
        // 128 in 3D), so each wavefront runs one direction's body (measured
        // 2 % faster at 256^3).  FP32: one border cell per thread for all
        // cross directions (the split measured 1.6 % slower there).
        constexpr int NTK = sizeof(AT) == 8 ? ND - 1 : 1;
        if (tid < 2 * Fc * NTK) {
          const int kt = NTK > 1 ? tid / (2 * Fc) : -1, r2 = tid - (NTK > 1 ? kt * 2 * Fc : 0);
          const int side = r2 >= Fc ? 1 : 0, r = r2 - side * Fc;
          int bc[3] = {0, 0, 0};
          bc[f] = side ? (f == 0 ? TX : (f == 1 ? TY : TZ)) + 1 : 0;
          bc[pd] = r % Tp;
          if (ND == 3) bc[qd] = r / Tp;
          const int l = bc[0] + rx * (bc[1] + ry * bc[2]);
          const int rel[3] = {bc[0] - ex[0], bc[1] - ex[1], ND == 3 ? bc[2] - ex[2] : 0};
          int xr[3] = {o[0] + rel[0], o[1] + rel[1], ND == 3 ? o[2] + rel[2] : 0};
          const bool in = !EDGE || interior_c<ND>(G, xr);
          const unsigned gq = in ? cix(G, xr[0], xr[1], xr[2]) : 0u;
          const int pi = (rel[0] + 1) + L::PX * ((rel[1] + 1) + L::PY * (ND == 3 ? rel[2] + 1 : 0));
          const int qi = (rel[0] + 3) + L::QX * ((rel[1] + 3) + L::QY * (ND == 3 ? rel[2] + 3 : 0));
#pragma unroll
          for (int d = 0; d < ND; d++) {
            if (d == f) continue;
            const int k = d < f ? d : d - 1;
            if (NTK > 1 && k != kt) continue;
            AT g3 = 0, g5 = 0;
            if (in) {
              // d != f: the border cell's coordinate along d is a tile coordinate
              const int ti = (d == 0 ? 0 : (d == 1 ? TX : TX + TY)) + rel[d];
              const AT dxr = A.dx1 ? SDX[ti] : (AT)at(A.delx + (long long)d * N, gq);
              const AT rdxr = A.dx1 ? SDX[L::DXT + ti] : (AT)at(A.rdelx + (long long)d * N, gq);
              AT nr, nl;
              nu_win<SH, AT>(qb, qi, qs[d], dxr * cm[d] * Kc, nr, nl);
              const AT t2 =
                  z1e<EDGE>(G, zTMP1, d, xr[d]) ? AT(0) : (qb[qi + qs[d]] - qb[qi - qs[d]]) * rdxr * AT(0.5);
              const AT t3 = t2 * (nl + nr) * AT(0.5);
              g3 = t3;
              g5 = t3 * CBX[(1 + d) * L::PN + pi];
            }
            RG[2 * k][l] = g3;
            RG[2 * k + 1][l] = g5;
          }
        }
      }
      __syncthreads();
      if (valid) {
#pragma unroll
        for (int d = 0; d < ND; d++) {
          if (d == f) continue;
          const int k = d < f ? d : d - 1;
          const AT* GAk = RG[2 * k];
          const AT* GBk = RG[2 * k + 1];
          const AT o3 = z1e<EDGE>(G, zTMP3, f, x[f]) ? AT(0) : (GAk[ri + sf] - GAk[ri - sf]) * (rdx[f] * AT(0.5));
          const AT o5 = z1e<EDGE>(G, zTMP5, f, x[f]) ? AT(0) : (GBk[ri + sf] - GBk[ri - sf]) * (rdx[f] * AT(0.5));
          acc[fB(ND, d)] += -o3;
          acc[fE(ND)] += -o5;
          const int s = qs[d];
          const bool z = z1e<EDGE>(G, zTMP1, d, x[d]);
          const AT q0 = qb[qc];
          const AT gr = z ? AT(0) : (qb[qc + s] - q0) * rdx[d];
          const AT gl = z ? AT(0) : (q0 - qb[qc - s]) * rdx[d];
          const AT br = (q0 + qb[qc + s]) * AT(0.5), bl = (q0 + qb[qc - s]) * AT(0.5);
          acc[fB(ND, f)] += (gr * nur[d] - gl * nul[d]) * rdx[d];
          acc[fE(ND)] += (br * gr * nur[d] - bl * gl * nul[d]) * rdx[d];
        }
      }
      __syncthreads();
    }
  }

  if (!valid) return;
  // ------------------------------------------------------------ gravity
  if (T & SMAUG_T_GRAV) {
#pragma unroll
    for (int d = 0; d < ND; d++) {
      if (G.g[d] == 0.0) continue;  // no gravity along d: both terms are zero
      acc[fMOM(d)] += (AT)(G.g[d] * rho_c);
      acc[fE(ND)] += (AT)(rho_c * (G.g[d] * at(w + (long long)fMOM(d) * N, c)) / (rho_c + rhob_c));
    }
  }
  // ------------------------------------------------------------ update, source, RK, ghosts
  const double h = A.hfrac * A.dts[0];
  const OT* __restrict__ base = (const OT*)A.base;
  OT out[NP];
#pragma unroll
  for (int f = 0; f < NP; f++) out[f] = at(base + (long long)f * N, c) + h * (OT)acc[f];
  // divB source terms (divb_parallel, src/divb.cu:165-338, 2D), in FP64 on
  // cells 2 < i,j < n-2: divb from the 4th-order differences over 2 p->dx of
  // b + b_b, zeroed on the layers 3, n-4, n-3 (grad3d's mask with interior
  // transverse indices); m_d += h divb b_d, b_d += h divb m_d/rho_t (a NaN
  // source counts 0); the energy source multiplies wd[bdotv], which the
  // reference never fills, so it is 0
  if constexpr (sizeof(OT) == 8) {  // FP64 state only (the tangent-linear step excludes divB)
  if (ND == 2 && G.divb && x[0] > 2 && x[1] > 2 && x[0] < G.n[0] - 2 && x[1] < G.n[1] - 2) {
    auto g3 = [&](const double* a, int d) -> double {
      const int xd = x[d], nn = G.n[d];
      if (xd == 3 || xd == nn - 4 || xd == nn - 3) return 0.0;
      const unsigned s = d == 0 ? 1u : (unsigned)G.n[0];
      return ((8 * at(a, c + s) - 8 * at(a, c - s) + at(a, c - 2 * s) - at(a, c + 2 * s)) / 6.0) / (2.0 * G.pdx[d]);
    };
    const double divb = (g3(w + (long long)fB(ND, 0) * N, 0) + g3(w + (long long)fB(ND, 1) * N, 1)) +
                        (g3(bg + (long long)bBB(0) * N, 0) + g3(bg + (long long)bBB(1) * N, 1));
    const double rt = rho_c + rhob_c;
#pragma unroll
    for (int d = 0; d < ND; d++) {
      const double sm = -divb * at(w + (long long)fB(ND, d) * N, c);
      const double sb = -divb * at(w + (long long)fMOM(d) * N, c) / rt;
      out[fMOM(d)] -= h * (isnan(sm) ? 0.0 : sm);
      out[fB(ND, d)] -= h * (isnan(sb) ? 0.0 : sb);
    }
  }
  }
  if ((T & SMAUG_T_SOURCE) && G.driver == SMAUG_DRIVER_VERTICAL) {
    const double* dv = G.drv;
    double ez, exyz;
    if (A.dshape) {  // separable grid: per-axis factors formed at set-up
      ez = A.dshape[x[0]];
      const double exx = A.dshape[G.n[0] + x[1]];
      exyz = exx * ez;
      if (ND == 3) exyz = exx * A.dshape[G.n[0] + G.n[1] + x[2]] * ez;
    } else {
      const double zp = at(A.pos, c) - dv[2];
      const double xp = at(A.pos + N, c) - dv[3];
      ez = exp(-zp * zp / (dv[5] * dv[5]));
      const double exx = exp(-xp * xp / (dv[6] * dv[6]));
      exyz = exx * ez;
      if (ND == 3) {
        const double yp = at(A.pos + 2 * N, c) - dv[4];
        const double ey = exp(-yp * yp / (dv[6] * dv[6]));
        exyz = exx * ey * ez;
      }
    }
    // amplitude map: one amplitude per (j,k) column in place of
    // A exp(-dx^2/sxy^2 [- dy^2/sxy^2])
    const OT vz = A.amap ? ((const OT*)A.amap)[x[1] + (ND == 3 ? (long long)G.n[1] * x[2] : 0)] * ez * A.tdep
                         : with_tangent<OT>(dv[0] * exyz * A.tdep, A.damp * exyz * A.tdep);
    const OT rto = out[fRHO()] + rhob_c;
    out[fMOM(0)] += h * vz * rto;
    out[fE(ND)] += h * (vz * vz) * rto / 2.0;
  }
  // fluxtube1's torsional driver (models/usersource.cu.fluxtube1:3-90), 3D
  if (ND == 3 && (T & SMAUG_T_SOURCE) && G.driver == SMAUG_DRIVER_TORSIONAL) {
    const double* dv = G.drv;
    const double xp = at(A.pos + N, c) - dv[3];
    const double zp = at(A.pos, c) - dv[2];
    const double yp = at(A.pos + 2 * N, c) - dv[4];
    const double exz = exp(-zp * zp / (dv[5] * dv[5]));
    const double exx = exp(-xp * xp / (dv[6] * dv[6]));
    const double exy = exp(-yp * yp / (dv[7] * dv[7]));
    const double exyz = exx * exy * exz;
    const OT vx = with_tangent<OT>((dv[0] * yp / dv[9]) * exyz * A.tdep, (A.damp * yp / dv[9]) * exyz * A.tdep);
    const OT vy = with_tangent<OT>(-(dv[0] * xp / dv[8]) * exyz * A.tdep, -(A.damp * xp / dv[8]) * exyz * A.tdep);
    const OT rto = out[fRHO()] + rhob_c;
    out[fMOM(1)] += h * vx * rto;
    out[fMOM(2)] += h * vy * rto;
    out[fE(ND)] += h * (vx * vx + vy * vy) * rto / 2.0;
  }
  // spicule1's driver (models/usersource.cu.spicule1:4-136): tdep is 0
  // outside the step whose counter is 1; grid positions xmin + i dx
  if ((T & SMAUG_T_SOURCE) && G.driver == SMAUG_DRIVER_SPICULE1 && A.tdep != 0.0) {
    const double* dv = G.drv;
    const double xp = G.pxmin[0] + (double)x[0] * G.pdx[0];
    const double yp = G.pxmin[1] + (double)x[1] * G.pdx[1];
    const double r1 = (xp - dv[2]) * (xp - dv[2]), r2 = (yp - dv[3]) * (yp - dv[3]);
    const double g = exp(-r2 / (dv[6] * dv[6])) * exp(-r1 / (dv[5] * dv[5]));
    const OT vvz = with_tangent<OT>(dv[0] * g * A.tdep, A.damp * g * A.tdep);
    out[fMOM(0)] = out[fMOM(0)] + h * vvz * (out[fRHO()] + rhob_c);
  }
  if (A.r1) {
#pragma unroll
    for (int f = 0; f < NP; f++) {
      const long long o_f = (long long)f * N;
      out[f] = out[f] + (at((const OT*)A.r1 + o_f, c) + 2.0 * at((const OT*)A.r2 + o_f, c) + at((const OT*)A.r3 + o_f, c) -
                         4.0 * at(base + o_f, c)) / 3;
    }
  }
  OT* __restrict__ wo = (OT*)A.wout;
  // the cell itself, then its ghost images: every combination of its images
  // across the directions (a fixed direction has none; with n >= 8 an
  // interior coordinate feeds at most two ghosts per direction), with the
  // asymm signs.  All indices are compile-time so no register array is
  // indexed at run time; waves away from the domain edge skip the images.
#pragma unroll
  for (int q = 0; q < NP; q++) put(wo + (long long)q * N, c, out[q]);
  if (!EDGE) return;  // no cell of an inner tile has a ghost image
#pragma unroll
  for (int f = 0; f < NP; f++) {
    if (f > 0 && G.bc_uniform) break;
    Img2 I[3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    unsigned neg = 0;  // bit d: direction d flips the sign
#pragma unroll
    for (int d = 0; d < ND; d++) {
      I[d] = bc_images2(G.bct[f][d], x[d], G.n[d], ND);
      if (G.bct[f][d] == BC_ASYMM) neg |= 1u << d;
    }
    if (I[0].m + I[1].m + I[2].m == 0) continue;
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
      for (int b2 = 0; b2 < 3; b2++)
#pragma unroll
        for (int c2 = 0; c2 < (ND == 3 ? 3 : 1); c2++) {
          if (a + b2 + c2 == 0 || a > I[0].m || b2 > I[1].m || c2 > I[2].m) continue;
          const int ga = a == 0 ? x[0] : (a == 1 ? I[0].g0 : I[0].g1);
          const int gb = b2 == 0 ? x[1] : (b2 == 1 ? I[1].g0 : I[1].g1);
          const int gc = c2 == 0 ? x[2] : (c2 == 1 ? I[2].g0 : I[2].g1);
          const unsigned cc = cix(G, ga, gb, gc);
          const unsigned flip = ((a ? 1u : 0u) & neg) ^ ((b2 ? 2u : 0u) & neg) ^ ((c2 ? 4u : 0u) & neg);
          const bool minus = __builtin_popcount(flip) & 1;
          if (G.bc_uniform) {
#pragma unroll
            for (int q = 0; q < NP; q++) put(wo + (long long)q * N, cc, minus ? -out[q] : out[q]);
          } else {
            put(wo + (long long)f * N, cc, minus ? -out[f] : out[f]);
          }
        }
  }
}

// FP32 stencils: at least five wavefronts per SIMD (96 VGPRs; measured
// faster at 256^3 than four without a spill or six with a 72-byte spill).
// FP64: three, the number its 52 KB of LDS per workgroup allows (146
// VGPRs); stating it lets the scheduler spend registers up to that bound
// (measured 0.8 % faster than a bound of one workgroup).  Dual (the
// tangent-linear step): its 105 KB of LDS in 3D leave one workgroup per CU
// whatever the bound, so the compiler may use the accumulation registers
// too (256 + 10 at gfx950, no scratch); DualF (its mixed-precision form, FP32
// pairs in 52 KB of LDS): two wavefronts per SIMD.
template <int ND, typename SH>
__global__ __launch_bounds__(256, sizeof(SH) == 4 ? 5 : (std::is_same<SH, double>::value ? 3 : 2)) void k_rhs_tiled(Geo G, RhsArgs A) {
  using L = Lay<ND>;
  constexpr int TX = L::TX, TY = L::TY, TZ = L::TZ;
  __shared__ SH sm[L::SMEM];
  const int tid = threadIdx.x;
  if (A.zero_next && blockIdx.x == 0 && tid < 3) A.zero_next[tid] = 0ull;
  // XCD-aware tile order: the hardware deals consecutive workgroups to the
  // 8 XCDs in turn, so workgroup b runs on XCD b%8.  In 3D XCD x takes the
  // slab of tile layers [x*gz/8, (x+1)*gz/8) and walks it x fastest, then
  // z, then y, so the z neighbour of a tile runs gx tiles after it (its
  // halo planes still in that XCD's L2) and the y neighbour gx*gz/8 tiles
  // after it (measured 1.7 % faster at 256^3 than x, y, z order, whose z
  // neighbour comes gx*gy tiles later); otherwise XCD x takes the tiles
  // [x*nb/8, (x+1)*nb/8) in x, y, z order.
  const int gx = (G.n[0] - 4 + TX - 1) / TX, gy = (G.n[1] - 4 + TY - 1) / TY;
  const int gz = ND == 3 ? (G.n[2] - 4 + TZ - 1) / TZ : 1;
  const int nb = gx * gy * gz;
  const int b = blockIdx.x;
  int t = b;
  if ((nb & 7) == 0) {
    const int xcd = b & 7, tl = b >> 3;
    if (ND == 3 && (gz & 7) == 0) {
      const int zc = gz >> 3;  // tile layers per XCD
      const int ux = tl % gx, r = tl / gx;
      const int uz = xcd * zc + r % zc, uy = r / zc;
      t = (uz * gy + uy) * gx + ux;
    } else {
      t = xcd * (nb >> 3) + tl;
    }
  }
  const int tbx = t % gx, tby = (t / gx) % gy, tbz = t / (gx * gy);
  const int o[3] = {2 + tbx * TX, 2 + tby * TY, ND == 3 ? 2 + tbz * TZ : 0};
  // inner tile: every cell in [5, n-6] along every direction (the test is
  // uniform per workgroup; the first and last tile layer along an axis are
  // edge tiles)
  bool edge = o[0] < 5 || o[0] + TX - 1 > G.n[0] - 6 || o[1] < 5 || o[1] + TY - 1 > G.n[1] - 6;
  if (ND == 3) edge = edge || o[2] < 5 || o[2] + TZ - 1 > G.n[2] - 6;
  if (edge)
    rhs_tile<ND, SH, true>(G, A, sm, o);
  else
    rhs_tile<ND, SH, false>(G, A, sm, o);
}

template <int ND, typename SH>
static void launch_tiled_t(const Geo& G, hipStream_t st, const RhsArgs& A) {
  using L = Lay<ND>;
  int nb = ((G.n[0] - 4 + L::TX - 1) / L::TX) * ((G.n[1] - 4 + L::TY - 1) / L::TY);
  if (ND == 3) nb *= (G.n[2] - 4 + L::TZ - 1) / L::TZ;
  hipLaunchKernelGGL((k_rhs_tiled<ND, SH>), dim3(nb), dim3(256), 0, st, G, A);
}

// the tangent-linear stage: every state array of A (win, base, r1..r3,
// wout, bg and the driver map) holds Dual values, and wsh and prim hold Dual
// values too, or FP32 DualF values when G.mixed
void launch_rhs_tl(const Geo& G, hipStream_t st, const RhsArgs& A) {
  if (G.mixed) {  // FP32 dual stencils (wsh, prim), FP64 dual own cells
    if (G.nd == 3)
      launch_tiled_t<3, DualF>(G, st, A);
    else
      launch_tiled_t<2, DualF>(G, st, A);
  } else if (G.nd == 3) {
    launch_tiled_t<3, Dual>(G, st, A);
  } else {
    launch_tiled_t<2, Dual>(G, st, A);
  }
}

void launch_rhs_tiled(const Geo& G, hipStream_t st, const RhsArgs& A) {
  if (G.nd == 3) {
    if (G.mixed) launch_tiled_t<3, float>(G, st, A);
    else launch_tiled_t<3, double>(G, st, A);
  } else {
    if (G.mixed) launch_tiled_t<2, float>(G, st, A);
    else launch_tiled_t<2, double>(G, st, A);
  }
}

}  // namespace smaug
