// Field kernels on the world lattice: fdtd2 curl updates, E -= J, binomial
// current filters, periodic ghost copy / fold, FP64 energy partials.
//
// Arithmetic follows the reference expression by expression (the library is
// compiled with -ffp-contract=off, so a*b+c stays two roundings):
//   push_b_fdtd2  src/runko/emf/yee_lattice_fdtd2.c++:19-58
//   push_e_fdtd2  src/runko/emf/yee_lattice_fdtd2.c++:68-111
//   push_b_stencil src/runko/emf/yee_lattice_stencil.c++:17-297
//   add_current   src/runko/emf/yee_lattice.c++:170-179
//   binomial2     src/runko/emf/yee_lattice_current_filter_binomial2.c++:23-76
//   binomial2_unrolled                                       :88-157
//
// All kernels are HBM-bound stencils.  Launch shape: 64 lanes along z (the
// contiguous dimension), 4 rows along y per workgroup, one x plane per
// grid.z, so every wave reads 256 contiguous bytes per component.
#include "rk_internal.h"

#include <cmath>
#include <cstring>

namespace rk {

namespace {

constexpr int BX = 64;
constexpr int BY = 4;

struct Region {
  int r0[3];
  int r1[3];
};

inline dim3 region_grid(const Region& r)
{
  const int ez = r.r1[2] - r.r0[2], ey = r.r1[1] - r.r0[1], ex = r.r1[0] - r.r0[0];
  return dim3((ez + BX - 1) / BX, (ey + BY - 1) / BY, ex);
}

__global__ void __launch_bounds__(BX* BY)
  k_curl_b(Geo g, Region r, float dt, const float* __restrict__ Ex, const float* __restrict__ Ey,
           const float* __restrict__ Ez, float* __restrict__ Bx, float* __restrict__ By,
           float* __restrict__ Bz)
{
  const int k = r.r0[2] + int(blockIdx.x) * BX + int(threadIdx.x);
  const int j = r.r0[1] + int(blockIdx.y) * BY + int(threadIdx.y);
  const int i = r.r0[0] + int(blockIdx.z);
  if(k >= r.r1[2] || j >= r.r1[1]) return;
  const uint32_t c = g.at(i, j, k);
  const uint32_t si = uint32_t(g.L[1]) * uint32_t(g.L[2]), sj = uint32_t(g.L[2]);
  const float ex = Ex[c], ey = Ey[c], ez = Ez[c];
  const float DkEy = Ey[c + 1] - ey;
  const float DjEz = Ez[c + sj] - ez;
  Bx[c] = Bx[c] + dt * (DkEy - DjEz);
  const float DiEz = Ez[c + si] - ez;
  const float DkEx = Ex[c + 1] - ex;
  By[c] = By[c] + dt * (DiEz - DkEx);
  const float DjEx = Ex[c + sj] - ex;
  const float DiEy = Ey[c + si] - ey;
  Bz[c] = Bz[c] + dt * (DjEx - DiEy);
}

// push_b_stencil, src/runko/emf/yee_lattice_stencil.c++:17-297: each
// derivative D*_a F sums, over rows n = 0..2 and then columns c = 0..4, M_a[n][c]
// times the difference of F between axial offsets +n+1 and -n: column 0 on
// the axial line, columns 1 / 3 the pairs at perp1 offsets +-1 / +-2, columns
// 2 / 4 at perp2 offsets +-1 / +-2 (perp1, perp2 cyclic: x -> (y, z),
// y -> (z, x), z -> (x, y)); each pair is (+o) + (-o).
struct StencilM {
  float M[3][3][5];
};

// sum over n, col of M[n][col] * diff(n, col) in the reference's order;
// at(ax, o1, o2) reads the field at offset ax along the derivative's axis a
// and o1, o2 along the cyclic perpendicular axes (a+1)%3, (a+2)%3
template<class At>
__device__ inline float stencil_sum(At&& at, const float (&M)[3][5])
{
  float sum = 0.0f;
#pragma unroll
  for(int n = 0; n < 3; ++n) {
    const int hi = n + 1, lo = -n;
#pragma unroll
    for(int col = 0; col < 5; ++col) {
      float diff;
      if(col == 0) {
        diff = at(hi, 0, 0) - at(lo, 0, 0);
      } else {
        const int o = col <= 2 ? 1 : 2;
        if(col == 1 || col == 3)
          diff = (at(hi, o, 0) + at(hi, -o, 0)) - (at(lo, o, 0) + at(lo, -o, 0));
        else
          diff = (at(hi, 0, o) + at(hi, 0, -o)) - (at(lo, 0, o) + at(lo, 0, -o));
      }
      const float term = M[n][col] * diff;
      sum = (n == 0 && col == 0) ? term : sum + term;
    }
  }
  return sum;
}

// The stencil reads offsets -2..3 along the derivative's axis and -2..2
// across it (324 reads of E per cell).  A block (BX cells along z, BY along
// y) marches through SCH x planes, keeping the last six planes of E over
// y, z in [start-2, end+3) in an LDS ring: each plane is read from memory
// once per block, in rows along z, and every cell takes its six derivatives
// from the ring.
constexpr int SSY = BY + 5, SSZ = BX + 5, SPL = SSY * SSZ, SCH = 32;

// One plane of the ring is 3 * SSY rows of SSZ floats; wave w of the block
// moves rows w, w + 4, ... (lanes along z), through registers so the loads
// of the next plane overlap the current plane's arithmetic.
constexpr int SWAVES = BX * BY / 64, SROWS = (3 * SSY + SWAVES - 1) / SWAVES, SLANES = (SSZ + 63) / 64;
static_assert(BX % 64 == 0, "a wave covers whole rows along z");

__device__ inline void stencil_load(const Geo& g, float (&v)[SROWS][SLANES], int gi, int j0, int k0,
                                    const float* __restrict__ Ex, const float* __restrict__ Ey,
                                    const float* __restrict__ Ez)
{
  const int wave = __builtin_amdgcn_readfirstlane((int(threadIdx.y) * BX + int(threadIdx.x)) / 64);
  const int lane = int(threadIdx.x) % 64;
  const bool xin = gi >= 0 && gi < g.L[0];
#pragma unroll
  for(int q = 0; q < SROWS; ++q) {
    const int row = wave + SWAVES * q, c = row / SSY, y = row % SSY, gj = j0 - 2 + y;
    const float* __restrict__ F = c == 0 ? Ex : (c == 1 ? Ey : Ez);
    const bool rin = row < 3 * SSY && xin && gj >= 0 && gj < g.L[1];
#pragma unroll
    for(int h = 0; h < SLANES; ++h) {
      const int gk = k0 - 2 + lane + 64 * h;
      v[q][h] = (rin && lane + 64 * h < SSZ && gk >= 0 && gk < g.L[2]) ? F[g.at(gi, gj, gk)] : 0.0f;
    }
  }
}

__device__ inline void stencil_store(float (&E)[3][6 * SPL], int slot, const float (&v)[SROWS][SLANES])
{
  const int wave = __builtin_amdgcn_readfirstlane((int(threadIdx.y) * BX + int(threadIdx.x)) / 64);
  const int lane = int(threadIdx.x) % 64;
#pragma unroll
  for(int q = 0; q < SROWS; ++q) {
    const int row = wave + SWAVES * q, c = row / SSY, y = row % SSY;
#pragma unroll
    for(int h = 0; h < SLANES; ++h)
      if(row < 3 * SSY && lane + 64 * h < SSZ) E[c][slot * SPL + y * SSZ + lane + 64 * h] = v[q][h];
  }
}

// derivative along axis A of the staged component Ec at the ring cell `cell`;
// sb[d] is the ring offset of plane x + d - 2
template<int A>
__device__ inline float stencil_lds(const float* Ec, const int (&sb)[6], int cell, const float (&M)[3][5])
{
  return stencil_sum(
    [&](int ax, int o1, int o2) {
      int d[3];
      d[A] = ax;
      d[(A + 1) % 3] = o1;
      d[(A + 2) % 3] = o2;
      return Ec[sb[d[0] + 2] + cell + d[1] * SSZ + d[2]];
    },
    M);
}

__global__ void __launch_bounds__(BX* BY)
  k_curl_b_stencil(Geo g, Region r, float dt, StencilM S, const float* __restrict__ Ex,
                   const float* __restrict__ Ey, const float* __restrict__ Ez, float* __restrict__ Bx,
                   float* __restrict__ By, float* __restrict__ Bz)
{
  __shared__ float E[3][6 * SPL];
  const int k0 = r.r0[2] + int(blockIdx.x) * BX, j0 = r.r0[1] + int(blockIdx.y) * BY;
  const int i0 = r.r0[0] + int(blockIdx.z) * SCH, i1 = min(r.r1[0], i0 + SCH);
  const int k = k0 + int(threadIdx.x), j = j0 + int(threadIdx.y);
  const bool active = k < r.r1[2] && j < r.r1[1];
  const int cell = (2 + int(threadIdx.y)) * SSZ + 2 + int(threadIdx.x);
  float v[SROWS][SLANES];
  for(int p = 0; p < 5; ++p) {
    stencil_load(g, v, i0 - 2 + p, j0, k0, Ex, Ey, Ez);
    stencil_store(E, p, v);
  }
  stencil_load(g, v, i0 + 3, j0, k0, Ex, Ey, Ez);
  int base = 0;  // ring slot of plane i - 2
  for(int i = i0; i < i1; ++i) {
    stencil_store(E, base == 0 ? 5 : base - 1, v);
    __syncthreads();
    if(i + 1 < i1) stencil_load(g, v, i + 4, j0, k0, Ex, Ey, Ez);
    if(active) {
      int sb[6];
#pragma unroll
      for(int d = 0; d < 6; ++d) sb[d] = (base + d < 6 ? base + d : base + d - 6) * SPL;
      const int64_t c = g.at(i, j, k);
      const float DzEy = stencil_lds<2>(E[1], sb, cell, S.M[2]);
      const float DyEz = stencil_lds<1>(E[2], sb, cell, S.M[1]);
      Bx[c] = Bx[c] + dt * (DzEy - DyEz);
      const float DxEz = stencil_lds<0>(E[2], sb, cell, S.M[0]);
      const float DzEx = stencil_lds<2>(E[0], sb, cell, S.M[2]);
      By[c] = By[c] + dt * (DxEz - DzEx);
      const float DyEx = stencil_lds<1>(E[0], sb, cell, S.M[1]);
      const float DxEy = stencil_lds<0>(E[1], sb, cell, S.M[0]);
      Bz[c] = Bz[c] + dt * (DyEx - DxEy);
    }
    __syncthreads();
    base = base == 5 ? 0 : base + 1;
  }
}

template<bool ADDJ>
__global__ void __launch_bounds__(BX* BY)
  k_curl_e(Geo g, Region r, float dt, const float* __restrict__ Bx, const float* __restrict__ By,
           const float* __restrict__ Bz, float* __restrict__ Ex, float* __restrict__ Ey,
           float* __restrict__ Ez, const float* __restrict__ Jx, const float* __restrict__ Jy,
           const float* __restrict__ Jz)
{
  const int k = r.r0[2] + int(blockIdx.x) * BX + int(threadIdx.x);
  const int j = r.r0[1] + int(blockIdx.y) * BY + int(threadIdx.y);
  const int i = r.r0[0] + int(blockIdx.z);
  if(k >= r.r1[2] || j >= r.r1[1]) return;
  const uint32_t c = g.at(i, j, k);
  const uint32_t si = uint32_t(g.L[1]) * uint32_t(g.L[2]), sj = uint32_t(g.L[2]);
  const float bx = Bx[c], by = By[c], bz = Bz[c];
  const float DkBy = By[c - 1] - by;
  const float DjBz = Bz[c - sj] - bz;
  float ex = Ex[c] + dt * (DkBy - DjBz);
  const float DiBz = Bz[c - si] - bz;
  const float DkBx = Bx[c - 1] - bx;
  float ey = Ey[c] + dt * (DiBz - DkBx);
  const float DjBx = Bx[c - sj] - bx;
  const float DiBy = By[c - si] - by;
  float ez = Ez[c] + dt * (DjBx - DiBy);
  if constexpr(ADDJ) {
    ex = ex - Jx[c];
    ey = ey - Jy[c];
    ez = ez - Jz[c];
  }
  Ex[c] = ex;
  Ey[c] = ey;
  Ez[c] = ez;
}

__global__ void __launch_bounds__(BX* BY)
  k_add_current(Geo g, Region r, float* __restrict__ Ex, float* __restrict__ Ey, float* __restrict__ Ez,
                const float* __restrict__ Jx, const float* __restrict__ Jy, const float* __restrict__ Jz)
{
  const int k = r.r0[2] + int(blockIdx.x) * BX + int(threadIdx.x);
  const int j = r.r0[1] + int(blockIdx.y) * BY + int(threadIdx.y);
  const int i = r.r0[0] + int(blockIdx.z);
  if(k >= r.r1[2] || j >= r.r1[1]) return;
  const uint32_t c = g.at(i, j, k);
  Ex[c] = Ex[c] - Jx[c];
  Ey[c] = Ey[c] - Jy[c];
  Ez[c] = Ez[c] - Jz[c];
}

// 27-tap (1,2,1)^3/64, taps in layout_right order, accumulator starting at 0.
// Every coefficient is a power of two (1, 2, 4 or 8 over 64), so the product
// c*P is exact and fma(c, P, acc) rounds exactly as acc + c*P does (unless
// c*P falls below the smallest normal float, 2^-126).
// Same 27-tap filter, each lane sweeping a column of FCH outputs along x: the
// 3x3 (y, z) neighbourhood of three consecutive x planes stays in registers,
// so every input is loaded once per column instead of three times.  The
// three components share the sweep, so a plane's 27 loads are in flight
// together.  Each output still accumulates its 27 taps in layout_right order
// from 0.
constexpr int FCH = 8;

__device__ inline void filter_columns(const Geo& g, const float* const (&S)[3], float* const (&D)[3], int i0, int i1,
                                      int j, int k)
{
  const uint32_t sj = uint32_t(g.L[2]);
  float P[3][3][3][3];  // [component][plane i-1, i, i+1][b][c]
  auto load = [&](int p, int i) {
    const uint32_t c0 = g.at(i, j, k);
#pragma unroll
    for(int q = 0; q < 3; ++q)
#pragma unroll
      for(int b = 0; b < 3; ++b)
#pragma unroll
        for(int cc = 0; cc < 3; ++cc) P[q][p][b][cc] = S[q][c0 + uint32_t(b - 1) * sj + uint32_t(cc - 1)];
  };
  load(0, i0 - 1);
  load(1, i0);
  for(int i = i0; i < i1; ++i) {
    load(2, i + 1);
    const float w[3] = {1.0f, 2.0f, 1.0f};
    const uint32_t c = g.at(i, j, k);
#pragma unroll
    for(int q = 0; q < 3; ++q) {
      float acc = 0.0f;
#pragma unroll
      for(int a = 0; a < 3; ++a)
#pragma unroll
        for(int b = 0; b < 3; ++b)
#pragma unroll
          for(int cc = 0; cc < 3; ++cc) acc = __builtin_fmaf((w[a] * w[b] * w[cc]) * (1.0f / 64.0f), P[q][a][b][cc], acc);
      D[q][c] = acc;
    }
#pragma unroll
    for(int q = 0; q < 3; ++q)
#pragma unroll
      for(int b = 0; b < 3; ++b)
#pragma unroll
        for(int cc = 0; cc < 3; ++cc) {
          P[q][0][b][cc] = P[q][1][b][cc];
          P[q][1][b][cc] = P[q][2][b][cc];
        }
  }
}

__global__ void __launch_bounds__(BX* BY)
  k_filter27c(Geo g, Region r, const float* __restrict__ S0, const float* __restrict__ S1,
              const float* __restrict__ S2, float* __restrict__ D0, float* __restrict__ D1, float* __restrict__ D2)
{
  const int k = r.r0[2] + int(blockIdx.x) * BX + int(threadIdx.x);
  const int j = r.r0[1] + int(blockIdx.y) * BY + int(threadIdx.y);
  const int i0 = r.r0[0] + int(blockIdx.z) * FCH;
  const int i1 = min(i0 + FCH, r.r1[0]);
  if(k >= r.r1[2] || j >= r.r1[1]) return;
  const float* const S[3] = {S0, S1, S2};
  float* const D[3] = {D0, D1, D2};
  filter_columns(g, S, D, i0, i1, j, k);
}

// separable z, y, x passes (0.25, 0.5, 0.25), each output evaluated with the
// reference's intermediate values computed in the same order.  A lane sweeps
// a column of FCH outputs along x; the z-then-y partial of a plane,
//   T(p) = 0.25 Z(p, j-1) + 0.5 Z(p, j) + 0.25 Z(p, j+1),
// depends on the plane alone, so each is computed once and reused by the
// three outputs that take it.
__device__ inline float sep_z(const float* __restrict__ S, uint32_t q)
{
  return 0.25f * S[q - 1] + 0.5f * S[q] + 0.25f * S[q + 1];
}

__device__ inline float sep_zy(const Geo& g, const float* __restrict__ S, int i, int j, int k)
{
  const float y0 = sep_z(S, g.at(i, j - 1, k));
  const float y1 = sep_z(S, g.at(i, j, k));
  const float y2 = sep_z(S, g.at(i, j + 1, k));
  return 0.25f * y0 + 0.5f * y1 + 0.25f * y2;
}

__device__ inline void sep_columns(const Geo& g, const float* const (&S)[3], float* const (&D)[3], int i0, int i1,
                                   int j, int k)
{
  float t0[3], t1[3];
#pragma unroll
  for(int q = 0; q < 3; ++q) {
    t0[q] = sep_zy(g, S[q], i0 - 1, j, k);
    t1[q] = sep_zy(g, S[q], i0, j, k);
  }
  for(int i = i0; i < i1; ++i) {
    float t2[3];
#pragma unroll
    for(int q = 0; q < 3; ++q) t2[q] = sep_zy(g, S[q], i + 1, j, k);
    const uint32_t c = g.at(i, j, k);
#pragma unroll
    for(int q = 0; q < 3; ++q) {
      D[q][c] = 0.25f * t0[q] + 0.5f * t1[q] + 0.25f * t2[q];
      t0[q] = t1[q];
      t1[q] = t2[q];
    }
  }
}

__global__ void __launch_bounds__(BX* BY)
  k_filter_sep(Geo g, Region r, const float* __restrict__ S0, const float* __restrict__ S1,
               const float* __restrict__ S2, float* __restrict__ D0, float* __restrict__ D1,
               float* __restrict__ D2)
{
  const int k = r.r0[2] + int(blockIdx.x) * BX + int(threadIdx.x);
  const int j = r.r0[1] + int(blockIdx.y) * BY + int(threadIdx.y);
  const int i0 = r.r0[0] + int(blockIdx.z) * FCH;
  const int i1 = min(i0 + FCH, r.r1[0]);
  if(k >= r.r1[2] || j >= r.r1[1]) return;
  const float* const S[3] = {S0, S1, S2};
  float* const D[3] = {D0, D1, D2};
  sep_columns(g, S, D, i0, i1, j, k);
}

// Three binomial2 passes in one launch, for a periodic world whose J ghost
// layer is current (3 deep, as after a ghost copy): every interior node of the
// third pass depends on inputs at most 3 nodes away, so the result equals the
// passes run one by one (with or without a ghost copy between them) bit for
// bit.  A workgroup owns an F3Y x F3Z (y, z) column of output nodes and marches
// along x over F3XS planes.  Pass level k (1..3) covers its output's (y, z)
// extent widened by 3-k nodes per side.  When a plane of level k-1 arrives,
// every level-k node reads its 3x3 (y, z) neighbourhood from LDS once and
// feeds the three 27-tap sums that take this plane: it completes the node
// one plane behind (third x offset), continues the node of this plane
// (second) and starts the node one plane ahead (first).  Each sum thus takes
// its taps in the order of filter_columns (x offset, then y, then z, from 0),
// and a level's completed planes reach the next level through LDS (planes
// double-buffered per level, one barrier per step).  The x segment is at
// most F3XS planes and shorter on small worlds, so that the launch still has
// a few workgroups per CU.
constexpr int F3Y = 8, F3Z = 32, F3T = 256, F3XS = 128;
constexpr int F3IY = F3Y + 6, F3IZ = F3Z + 6, F3IN = F3IY * F3IZ, F3IM = (F3IN + F3T - 1) / F3T;

template<int K>
struct F3Level {
  static constexpr int WY = F3Y + 2 * (3 - K), WZ = F3Z + 2 * (3 - K), NP = WY * WZ;
  static constexpr int NM = (NP + F3T - 1) / F3T;  // nodes per thread
  static constexpr int SZ = WZ + 2;                // row length of the level below
};

struct F3Chains {
  float s1, s2;  // sums of the nodes one plane ahead (one x offset taken) and of this plane (two taken)
};

// one level-K position (py, pz): take plane `src` (row length SZ) of the level below
template<int K>
__device__ inline float f3_take(const float* src, int py, int pz, F3Chains& ch)
{
  constexpr int SZ = F3Level<K>::SZ;
  const float w[3] = {1.0f, 2.0f, 1.0f};
  float v[3][3];
#pragma unroll
  for(int b = 0; b < 3; ++b)
#pragma unroll
    for(int c = 0; c < 3; ++c) v[b][c] = src[(py + b) * SZ + pz + c];
  float out = ch.s2, mid = ch.s1, head = 0.0f;
#pragma unroll
  for(int b = 0; b < 3; ++b)
#pragma unroll
    for(int c = 0; c < 3; ++c) {
      out = __builtin_fmaf((w[2] * w[b] * w[c]) * (1.0f / 64.0f), v[b][c], out);
      mid = __builtin_fmaf((w[1] * w[b] * w[c]) * (1.0f / 64.0f), v[b][c], mid);
      head = __builtin_fmaf((w[0] * w[b] * w[c]) * (1.0f / 64.0f), v[b][c], head);
    }
  ch.s2 = mid;
  ch.s1 = head;
  return out;
}

__global__ void __launch_bounds__(F3T)
  k_filter27_x3(Geo g, int xseg, const float* __restrict__ S0, const float* __restrict__ S1,
                const float* __restrict__ S2, float* __restrict__ D0, float* __restrict__ D1, float* __restrict__ D2)
{
  __shared__ float In[2][3][F3IN];
  __shared__ float L1[2][3][F3Level<1>::NP];
  __shared__ float L2[2][3][F3Level<2>::NP];
  const int tid = int(threadIdx.x);
  const int z0 = HALO + int(blockIdx.x) * F3Z, y0 = HALO + int(blockIdx.y) * F3Y;
  const int x0 = HALO + int(blockIdx.z) * xseg;
  const int xs = min(xseg, HALO + g.N[0] - x0);
  const float* const S[3] = {S0, S1, S2};
  float* const D[3] = {D0, D1, D2};
  // input plane q (0 .. xs+5) is haloed plane x0 - 3 + q; rows and columns
  // past the lattice (partial workgroups) read its last node and feed only
  // outputs that are not stored
  uint32_t ioff[F3IM];
  bool iok[F3IM];
#pragma unroll
  for(int m = 0; m < F3IM; ++m) {
    const int e = tid + m * F3T;
    iok[m] = e < F3IN;
    const int r = e / F3IZ, c = e - r * F3IZ;
    const int yy = min(y0 - HALO + r, g.L[1] - 1), zz = min(z0 - HALO + c, g.L[2] - 1);
    ioff[m] = uint32_t(yy) * uint32_t(g.L[2]) + uint32_t(zz);
  }
  const uint32_t plane = uint32_t(g.L[1]) * uint32_t(g.L[2]);
  float pre[3][F3IM];
  auto load = [&](int q) {
    const uint32_t base = uint32_t(x0 - HALO + q) * plane;
#pragma unroll
    for(int comp = 0; comp < 3; ++comp)
#pragma unroll
      for(int m = 0; m < F3IM; ++m)
        if(iok[m]) pre[comp][m] = S[comp][base + ioff[m]];
  };
  F3Chains c1[3][F3Level<1>::NM] = {}, c2[3][F3Level<2>::NM] = {}, c3[3][F3Level<3>::NM] = {};
  const int nin = xs + 6;  // input planes
  load(0);
  for(int s = 0; s < xs + 9; ++s) {
    const int cur = s & 1, prev = cur ^ 1;
    if(s < nin) {
#pragma unroll
      for(int comp = 0; comp < 3; ++comp)
#pragma unroll
        for(int m = 0; m < F3IM; ++m)
          if(iok[m]) In[cur][comp][tid + m * F3T] = pre[comp][m];
      if(s + 1 < nin) load(s + 1);
    }
    // level 1 takes input plane s-1 and completes its node s-2 (input-relative index)
    if(s >= 1 && s <= nin) {
#pragma unroll
      for(int m = 0; m < F3Level<1>::NM; ++m) {
        const int p = tid + m * F3T;
        if(p < F3Level<1>::NP) {
          const int py = p / F3Level<1>::WZ, pz = p - py * F3Level<1>::WZ;
#pragma unroll
          for(int comp = 0; comp < 3; ++comp)
            L1[cur][comp][p] = f3_take<1>(In[prev][comp], py, pz, c1[comp][m]);
        }
      }
    }
    // level 2 takes level-1 node s-3 and completes its node s-4
    if(s >= 3 && s <= nin + 2) {
#pragma unroll
      for(int m = 0; m < F3Level<2>::NM; ++m) {
        const int p = tid + m * F3T;
        if(p < F3Level<2>::NP) {
          const int py = p / F3Level<2>::WZ, pz = p - py * F3Level<2>::WZ;
#pragma unroll
          for(int comp = 0; comp < 3; ++comp)
            L2[cur][comp][p] = f3_take<2>(L1[prev][comp], py, pz, c2[comp][m]);
        }
      }
    }
    // level 3 takes level-2 node s-5 and completes output node s-6 (haloed x0 - 9 + s)
    if(s >= 5) {
#pragma unroll
      for(int m = 0; m < F3Level<3>::NM; ++m) {
        const int p = tid + m * F3T;
        if(p < F3Level<3>::NP) {
          const int py = p / F3Level<3>::WZ, pz = p - py * F3Level<3>::WZ;
          float o[3];
#pragma unroll
          for(int comp = 0; comp < 3; ++comp) o[comp] = f3_take<3>(L2[prev][comp], py, pz, c3[comp][m]);
          const int y = y0 + py, z = z0 + pz;
          if(s >= 9 && y < HALO + g.N[1] && z < HALO + g.N[2]) {
            const uint32_t q = g.at(x0 - 9 + s, y, z);
#pragma unroll
            for(int comp = 0; comp < 3; ++comp) D[comp][q] = o[comp];
          }
        }
      }
    }
    __syncthreads();
  }
}

__device__ inline int wrap_idx(int c, int N) { return c < HALO ? c + N : (c >= N + HALO ? c - N : c); }

// periodic ghost copy: every ghost node takes the value of its interior image
// (sources are interior nodes, destinations ghost nodes, so the copies are
// independent and one launch does them all).  The first nrows workgroups
// copy the whole z rows whose i or j lies in the ghost shell: 2*HALO*L1 rows
// with a ghost i, then N0*2*HALO rows with an interior i and a ghost j, one
// workgroup per row.  The workgroups after them copy the z ghosts of the
// interior rows, one lane per ghost node.
__global__ void __launch_bounds__(64) k_ghost_copy(Geo g, int nrows, float* F0, float* F1, float* F2)
{
  int b = int(blockIdx.x);
  if(b >= nrows) {
    const int64_t t = int64_t(b - nrows) * 64 + threadIdx.x;
    if(t >= int64_t(g.N[0]) * g.N[1] * 2 * HALO) return;
    const int h = int(t % (2 * HALO));
    const int64_t row = t / (2 * HALO);
    const int i = HALO + int(row / g.N[1]), j = HALO + int(row % g.N[1]);
    const int k = h < HALO ? h : g.N[2] + h;
    const uint32_t d = g.at(i, j, k), s = g.at(i, j, wrap_idx(k, g.N[2]));
    F0[d] = F0[s];
    F1[d] = F1[s];
    F2[d] = F2[s];
    return;
  }
  int i, j;
  const int nA = 2 * HALO * g.L[1];
  if(b < nA) {
    const int ii = b / g.L[1];
    j = b - ii * g.L[1];
    i = ii < HALO ? ii : g.N[0] + ii;
  } else {
    b -= nA;
    const int jj = b % (2 * HALO);
    i = HALO + b / (2 * HALO);
    j = jj < HALO ? jj : g.N[1] + jj;
  }
  const int si = wrap_idx(i, g.N[0]), sj = wrap_idx(j, g.N[1]);
  for(int k = int(threadIdx.x); k < g.L[2]; k += int(blockDim.x)) {
    const uint32_t d = g.at(i, j, k), s = g.at(si, sj, wrap_idx(k, g.N[2]));
    F0[d] = F0[s];
    F1[d] = F1[s];
    F2[d] = F2[s];
  }
}

// images of an interior coordinate c in the ghost shell: c-N (c >= N) and c+N (c < 2*HALO)
__device__ inline int images(int c, int N, int out[3])
{
  int m = 0;
  out[m++] = c;
  if(c >= N) out[m++] = c - N;
  if(c < 2 * HALO) out[m++] = c + N;
  return m;
}

// periodic ghost fold (J_exchange): interior node += all of its ghost images
__global__ void __launch_bounds__(256) k_ghost_fold(Geo g, float* F0, float* F1, float* F2)
{
  const int j = HALO + int(blockIdx.x);
  const int i = HALO + int(blockIdx.y);
  int ii[3], jj[3], kk[3];
  const int mi = images(i, g.N[0], ii);
  const int mj = images(j, g.N[1], jj);
  // with N >= 2*HALO per axis a node has at most two images per axis: all of
  // its (up to 7) ghost images are addressed first and loaded together, then
  // added in the order of the general loop below (bit-identical)
  const bool two = g.N[0] >= 2 * HALO && g.N[1] >= 2 * HALO && g.N[2] >= 2 * HALO;
  auto fold = [&](int k) {
    const int mk = images(k, g.N[2], kk);
    if(mi * mj * mk == 1) return;
    const uint32_t c = g.at(i, j, k);
    float s0 = F0[c], s1 = F1[c], s2 = F2[c];
    if(two) {
      float v[8][3];
#pragma unroll
      for(int a = 0; a < 2; ++a)
#pragma unroll
        for(int b = 0; b < 2; ++b)
#pragma unroll
          for(int e = 0; e < 2; ++e) {
            const bool ok = a < mi && b < mj && e < mk && (a | b | e) != 0;
            const uint32_t q = ok ? g.at(ii[a], jj[b], kk[e]) : c;
            const int n = (a * 2 + b) * 2 + e;
            v[n][0] = F0[q];
            v[n][1] = F1[q];
            v[n][2] = F2[q];
          }
#pragma unroll
      for(int n = 1; n < 8; ++n) {
        const int a = n >> 2, b = (n >> 1) & 1, e = n & 1;
        const bool ok = a < mi && b < mj && e < mk;
        const float t0 = s0 + v[n][0], t1 = s1 + v[n][1], t2 = s2 + v[n][2];
        s0 = ok ? t0 : s0;
        s1 = ok ? t1 : s1;
        s2 = ok ? t2 : s2;
      }
      F0[c] = s0;
      F1[c] = s1;
      F2[c] = s2;
      return;
    }
    for(int a = 0; a < mi; ++a)
      for(int b = 0; b < mj; ++b)
        for(int e = 0; e < mk; ++e) {
          if(a == 0 && b == 0 && e == 0) continue;
          const uint32_t q = g.at(ii[a], jj[b], kk[e]);
          s0 = s0 + F0[q];
          s1 = s1 + F1[q];
          s2 = s2 + F2[q];
        }
    F0[c] = s0;
    F1[c] = s1;
    F2[c] = s2;
  };
  if(mi * mj > 1) {
    for(int k = HALO + int(threadIdx.x); k < g.N[2] + HALO; k += int(blockDim.x)) fold(k);
  } else {
    const int t = int(threadIdx.x);
    if(t < 2 * HALO) {
      const int k = t < HALO ? HALO + t : g.N[2] + t - HALO;
      if(k >= HALO && k < g.N[2] + HALO && (t < HALO || k >= HALO + HALO)) fold(k);
    }
  }
}

__global__ void __launch_bounds__(256) k_fill(float* F0, float* F1, float* F2, int64_t n, float v)
{
  for(int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; q < n; q += int64_t(gridDim.x) * blockDim.x) {
    F0[q] = v;
    F1[q] = v;
    F2[q] = v;
  }
}

// FP64 partial sums of (Fx^2 + Fy^2) + Fz^2 (float per node, as the reference)
__global__ void __launch_bounds__(256)
  k_energy(Geo g, Region r, const float* Ex, const float* Ey, const float* Ez, const float* Bx,
           const float* By, const float* Bz, double* partial)
{
  const int i = r.r0[0] + int(blockIdx.y);
  double se = 0.0, sb = 0.0;
  const int ey = r.r1[1] - r.r0[1], ez = r.r1[2] - r.r0[2];
  for(int q = int(threadIdx.x); q < ey * ez; q += int(blockDim.x)) {
    const int j = r.r0[1] + q / ez, k = r.r0[2] + q % ez;
    const uint32_t c = g.at(i, j, k);
    const float e2 = Ex[c] * Ex[c] + Ey[c] * Ey[c] + Ez[c] * Ez[c];
    const float b2 = Bx[c] * Bx[c] + By[c] * By[c] + Bz[c] * Bz[c];
    se += double(e2);
    sb += double(b2);
  }
  __shared__ double red[2][256];
  red[0][threadIdx.x] = se;
  red[1][threadIdx.x] = sb;
  __syncthreads();
  for(int o = 128; o > 0; o >>= 1) {
    if(int(threadIdx.x) < o) {
      red[0][threadIdx.x] += red[0][threadIdx.x + o];
      red[1][threadIdx.x] += red[1][threadIdx.x + o];
    }
    __syncthreads();
  }
  if(threadIdx.x == 0) {
    partial[2 * blockIdx.y] = red[0][0];
    partial[2 * blockIdx.y + 1] = red[1][0];
  }
}

Region make_region(const int r0[3], const int r1[3])
{
  Region r;
  for(int d = 0; d < 3; ++d) {
    r.r0[d] = r0[d];
    r.r1[d] = r1[d];
  }
  return r;
}

bool empty(const Region& r) { return r.r1[0] <= r.r0[0] || r.r1[1] <= r.r0[1] || r.r1[2] <= r.r0[2]; }

}  // namespace

int launch_curl_b(rk_world* w, const int r0[3], const int r1[3], float dt, hipStream_t s)
{
  const Region r = make_region(r0, r1);
  if(empty(r)) return RK_OK;
  hipLaunchKernelGGL(k_curl_b, region_grid(r), dim3(BX, BY), 0, s, w->g, r, dt, w->E[0], w->E[1], w->E[2],
                     w->B[0], w->B[1], w->B[2]);
  RK_HIP(hipGetLastError());
  return RK_OK;
}

int launch_curl_b_stencil(rk_world* w, const int r0[3], const int r1[3], float dt, hipStream_t s)
{
  const Region r = make_region(r0, r1);
  if(empty(r)) return RK_OK;
  StencilM S;
  std::memcpy(S.M, w->stencil, sizeof(S.M));
  dim3 grid = region_grid(r);
  grid.z = unsigned((r.r1[0] - r.r0[0] + SCH - 1) / SCH);
  hipLaunchKernelGGL(k_curl_b_stencil, grid, dim3(BX, BY), 0, s, w->g, r, dt, S, w->E[0], w->E[1],
                     w->E[2], w->B[0], w->B[1], w->B[2]);
  RK_HIP(hipGetLastError());
  return RK_OK;
}

int launch_curl_e(rk_world* w, const int r0[3], const int r1[3], float dt, bool add_j, hipStream_t s)
{
  const Region r = make_region(r0, r1);
  if(empty(r)) return RK_OK;
  if(add_j)
    hipLaunchKernelGGL(k_curl_e<true>, region_grid(r), dim3(BX, BY), 0, s, w->g, r, dt, w->B[0], w->B[1],
                       w->B[2], w->E[0], w->E[1], w->E[2], w->J[0], w->J[1], w->J[2]);
  else
    hipLaunchKernelGGL(k_curl_e<false>, region_grid(r), dim3(BX, BY), 0, s, w->g, r, dt, w->B[0], w->B[1],
                       w->B[2], w->E[0], w->E[1], w->E[2], w->J[0], w->J[1], w->J[2]);
  RK_HIP(hipGetLastError());
  return RK_OK;
}

int launch_add_current(rk_world* w, const int r0[3], const int r1[3], hipStream_t s)
{
  const Region r = make_region(r0, r1);
  if(empty(r)) return RK_OK;
  hipLaunchKernelGGL(k_add_current, region_grid(r), dim3(BX, BY), 0, s, w->g, r, w->E[0], w->E[1], w->E[2],
                     w->J[0], w->J[1], w->J[2]);
  RK_HIP(hipGetLastError());
  return RK_OK;
}

int launch_filter(rk_world* w, const int r0[3], const int r1[3], hipStream_t s)
{
  const Region r = make_region(r0, r1);
  if(empty(r)) return RK_OK;
  if(w->d.current_filter == 0) {
    dim3 grid = region_grid(r);
    grid.z = unsigned((r.r1[0] - r.r0[0] + FCH - 1) / FCH);
    hipLaunchKernelGGL(k_filter27c, grid, dim3(BX, BY), 0, s, w->g, r, w->J[0], w->J[1], w->J[2], w->Jf[0],
                       w->Jf[1], w->Jf[2]);
  }
  else {
    dim3 grid = region_grid(r);
    grid.z = unsigned((r.r1[0] - r.r0[0] + FCH - 1) / FCH);
    hipLaunchKernelGGL(k_filter_sep, grid, dim3(BX, BY), 0, s, w->g, r, w->J[0], w->J[1], w->J[2], w->Jf[0],
                       w->Jf[1], w->Jf[2]);
  }
  RK_HIP(hipGetLastError());
  return RK_OK;
}

int launch_filter_x3(rk_world* w, hipStream_t s)
{
  const Geo& g = w->g;
  const int64_t cols = int64_t((g.N[2] + F3Z - 1) / F3Z) * ((g.N[1] + F3Y - 1) / F3Y);
  int xseg = F3XS;
  while(xseg > 4 && cols * ((g.N[0] + xseg - 1) / xseg) < 512) xseg /= 2;  // 2 workgroups per CU
  const dim3 grid(unsigned((g.N[2] + F3Z - 1) / F3Z), unsigned((g.N[1] + F3Y - 1) / F3Y),
                  unsigned((g.N[0] + xseg - 1) / xseg));
  hipLaunchKernelGGL(k_filter27_x3, grid, dim3(F3T), 0, s, g, xseg, w->J[0], w->J[1], w->J[2], w->Jf[0], w->Jf[1],
                     w->Jf[2]);
  RK_HIP(hipGetLastError());
  return RK_OK;
}

int launch_ghost_copy(rk_world* w, float* const F[3], hipStream_t s)
{
  const Geo& g = w->g;
  const int nrows = 2 * HALO * g.L[1] + g.N[0] * 2 * HALO;
  const int64_t nz = int64_t(g.N[0]) * g.N[1] * 2 * HALO;
  hipLaunchKernelGGL(k_ghost_copy, dim3(unsigned(nrows + (nz + 63) / 64)), dim3(64), 0, s, g, nrows, F[0], F[1], F[2]);
  RK_HIP(hipGetLastError());
  return RK_OK;
}

int launch_ghost_fold(rk_world* w, float* const F[3], hipStream_t s)
{
  hipLaunchKernelGGL(k_ghost_fold, dim3(w->g.N[1], w->g.N[0]), dim3(64), 0, s, w->g, F[0], F[1], F[2]);
  RK_HIP(hipGetLastError());
  return RK_OK;
}

int launch_fill(float* const F[3], int64_t n, float v, hipStream_t s)
{
  const int64_t blocks = std::min<int64_t>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(k_fill, dim3(unsigned(blocks)), dim3(256), 0, s, F[0], F[1], F[2], n, v);
  RK_HIP(hipGetLastError());
  return RK_OK;
}

int launch_field_energy_partials(rk_world* w, const int r0[3], const int r1[3], double* partial, hipStream_t s)
{
  const Region r = make_region(r0, r1);
  if(empty(r)) return RK_OK;
  const int planes = r.r1[0] - r.r0[0];
  hipLaunchKernelGGL(k_energy, dim3(1, planes), dim3(256), 0, s, w->g, r, w->E[0], w->E[1], w->E[2], w->B[0],
                     w->B[1], w->B[2], partial);
  RK_HIP(hipGetLastError());
  return RK_OK;
}

int field_energies(rk_world* w, const int r0[3], const int r1[3], double out[2], hipStream_t s)
{
  const Region r = make_region(r0, r1);
  out[0] = out[1] = 0;
  if(empty(r)) return RK_OK;
  const int planes = r.r1[0] - r.r0[0];
  if(w->red_cap < 2 * planes) return (set_error("reduction scratch too small"), RK_EINVAL);
  if(int rc = launch_field_energy_partials(w, r0, r1, w->red, s)) return rc;
  std::vector<double> h(size_t(2 * planes));
  RK_HIP(hipMemcpyAsync(h.data(), w->red, sizeof(double) * h.size(), hipMemcpyDeviceToHost, s));
  RK_HIP(hipStreamSynchronize(s));
  for(int p = 0; p < planes; ++p) {
    out[0] += h[size_t(2 * p)];
    out[1] += h[size_t(2 * p + 1)];
  }
  out[0] /= 2.0;
  out[1] /= 2.0;
  return RK_OK;
}

}  // namespace rk
