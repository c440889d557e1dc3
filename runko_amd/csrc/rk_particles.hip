// Particle kernels: interpolation + Boris push, zigzag current deposit,
// tile/brick classification with compaction, migration append with the
// periodic wrap, in-bin cell sort, injection, readback, FP64 kinetic energy,
// and the fused batched particle step.
//
// Reference semantics followed (hel-astro-lab/runko v5.0.0):
//   interpolation  src/runko/emf/yee_lattice_interpolate_linear_1st.h:29-138
//   Boris          src/runko/pic/particle_boris.h:15-62
//   Higuera-Cary   src/runko/pic/particle_higuera_cary.h:15-77
//   Faraday        src/runko/pic/particle_faraday.h:37-109
//   zigzag         src/runko/pic/particle_current_zigzag_1st.c++:225-339
//   divide         src/runko/pic/particle.c++:198-348
//   append/wrap    src/runko/pic/particle.h:449-572
//   sort key       src/runko/pic/tile.c++:417-438
//   kinetic energy src/runko/pic/particle.c++:351-377
//
// Execution model of the per-tile methods: one workgroup (256 lanes, 4
// waves) per particle bin (brick).  The brick's node-centred E/B means (brick
// + 1 node) are staged in LDS once; every particle of the brick gathers its 8
// corners from LDS.  The deposit accumulates the brick's current in an FP64
// LDS window over brick + (1,2) nodes and adds it to J once.  A particle
// whose stencil leaves the staged window (possible only on the plain per-tile
// path, e.g. two pushes without a pack) takes the same arithmetic through
// global memory.  Leavers are written to SoA migration records reserved with
// one atomic per wave (ballot + popcount) and their slots are refilled from
// the bin tail (swap-remove) so bins stay compact.  The fused step (k_fused)
// does push, walls, classification, deposit and in-order compaction of a
// brick in one workgroup.
#include "rk_internal.h"
#include "rk_reflect.h"

#include <algorithm>

namespace rk {

namespace {

constexpr int WG = 256;
constexpr int NWAVE = WG / 64;

struct SpParams {
  float* pos[RK_MAX_SPECIES][3];
  float* vel[RK_MAX_SPECIES][3];
  uint64_t* id[RK_MAX_SPECIES];
  const int64_t* start[RK_MAX_SPECIES];
  const uint32_t* cap[RK_MAX_SPECIES];
  uint32_t* count[RK_MAX_SPECIES];
  float qm[RK_MAX_SPECIES];
  float charge[RK_MAX_SPECIES];
  int nsp;
};

struct FieldPtrs {
  const float* E[3];
  const float* B[3];
  float* J[3];
};

// record sets: 0 tile leavers, 1 brick changes, 2 per-brick mailboxes
struct MigPtrs {
  float* pos[3][3];
  float* vel[3][3];
  uint64_t* id[3];
  uint32_t* meta[3];
  int64_t cap[3];
  int64_t per_brick;   // mailbox records per brick (set 2)
  uint32_t* mcount;    // mailbox fill per brick
};

SpParams make_sp(const rk_world* w)
{
  SpParams p {};
  p.nsp = w->d.n_species;
  for(int s = 0; s < p.nsp; ++s) {
    for(int d = 0; d < 3; ++d) {
      p.pos[s][d] = w->sp[s].pos[d];
      p.vel[s][d] = w->sp[s].vel[d];
    }
    p.id[s] = w->sp[s].id;
    p.start[s] = w->sp[s].bin_start;
    p.cap[s] = w->sp[s].bin_cap;
    p.count[s] = w->sp[s].bin_count;
    p.qm[s] = w->sp[s].qm;
    p.charge[s] = w->sp[s].charge;
  }
  return p;
}

FieldPtrs make_fields(const rk_world* w)
{
  FieldPtrs f {};
  for(int d = 0; d < 3; ++d) {
    f.E[d] = w->E[d];
    f.B[d] = w->B[d];
    f.J[d] = w->J[d];
  }
  return f;
}

MigPtrs make_mig(const rk_world* w)
{
  MigPtrs m {};
  for(int s = 0; s < 3; ++s) {
    const MigSet& src = s < 2 ? w->mig[s] : w->mbox;
    for(int d = 0; d < 3; ++d) {
      m.pos[s][d] = src.pos[d];
      m.vel[s][d] = src.vel[d];
    }
    m.id[s] = src.id;
    m.meta[s] = src.meta;
    m.cap[s] = src.capacity;
  }
  m.per_brick = w->mbox_per_brick;
  m.mcount = w->mbox_count;
  return m;
}

// ---------------------------------------------------------------------------
// brick geometry
// ---------------------------------------------------------------------------
struct Brick {
  int bin;      // brick index (bin index inside a species pool)
  int tile;
  int tc[3];    // tile coords in the world
  int lb[3];    // brick coords in the tile
  int toffh[3]; // world haloed index of the tile's haloed index 0
  int c0[3];    // world haloed index of the brick's first cell
  float tmin[3], tmax[3], origo[3];
};

__device__ inline Brick brick_of(const Geo& g, int b)
{
  Brick B;
  B.bin = b;
  g.brick_coords(b, B.tile, B.lb);
  g.tile_coords(B.tile, B.tc);
#pragma unroll
  for(int d = 0; d < 3; ++d) {
    B.toffh[d] = B.tc[d] * g.n[d];
    B.c0[d] = B.toffh[d] + HALO + B.lb[d] * g.br[d];
    B.tmin[d] = g.tile_min(B.tc[d], d);
    B.tmax[d] = float((g.toff[d] + B.tc[d] + 1) * g.n[d]);
    // static_cast<float>(mins) - halo_size, src/runko/pic/tile.c++:329-332
    B.origo[d] = B.tmin[d] - 3.0f;
  }
  return B;
}

// brick (inside tile with coords tc) of a position; cells clamped to the tile
__device__ inline int brick_in_tile(const Geo& g, int tile, const float origo[3], const float p[3])
{
  int lb[3];
#pragma unroll
  for(int d = 0; d < 3; ++d) {
    int c = int(floorf(p[d] - origo[d])) - HALO;
    c = c < 0 ? 0 : (c >= g.n[d] ? g.n[d] - 1 : c);
    lb[d] = g.div_br(c, d);
  }
  return tile * g.bpt + (lb[0] * g.nb[1] + lb[1]) * g.nb[2] + lb[2];
}

// XCD-aware block -> brick map.  Blocks are dealt round-robin over the 8 XCDs;
// XCD x takes the groups x, x + 8, x + 16, ... of `span` consecutive bricks, so
// bricks that share halo lines in L2 stay on one XCD while every XCD gets an
// even share of a world whose density varies along x (span = whole tiles, an
// eighth of a tile column).  span = ceil(n/8) gives each XCD one contiguous
// eighth.  The grid is xcd_grid(n, span); blocks mapped past n exit.
__device__ inline int xcd_brick(int blk, int span)
{
  const int x = blk % 8, j = blk / 8;
  return ((j / span) * 8 + x) * span + j % span;
}

// ---------------------------------------------------------------------------
// node-centred field means, interpolation_linear_1st :91-123
// mean(a,b) = (a+b)/2 ; mean(a,b,c,d) = (a+(b+(c+d)))/4  (right fold)
// ---------------------------------------------------------------------------
struct Means {
  float ex, ey, ez, bx, by, bz;
};

__device__ inline Means node_means(const Geo& g, const FieldPtrs& f, int I, int J, int K)
{
  const uint32_t c = g.at(I, J, K);
  const uint32_t si = uint32_t(g.L[1]) * uint32_t(g.L[2]), sj = uint32_t(g.L[2]);
  Means m;
  m.ex = (f.E[0][c - si] + f.E[0][c]) * 0.5f;
  m.ey = (f.E[1][c - sj] + f.E[1][c]) * 0.5f;
  m.ez = (f.E[2][c - 1] + f.E[2][c]) * 0.5f;
  m.bx = (f.B[0][c] + (f.B[0][c - sj] + (f.B[0][c - 1] + f.B[0][c - sj - 1]))) * 0.25f;
  m.by = (f.B[1][c] + (f.B[1][c - si] + (f.B[1][c - 1] + f.B[1][c - si - 1]))) * 0.25f;
  m.bz = (f.B[2][c] + (f.B[2][c - si] + (f.B[2][c - sj] + f.B[2][c - si - sj]))) * 0.25f;
  return m;
}

__device__ inline float lerpf(float x, float A, float B) { return (1.0f - x) * A + x * B; }

// lerp3D: x first, then y, then z; v[ic*4 + jc*2 + kc]
__device__ inline float lerp3(const float d[3], const float v[8])
{
  const float c00 = lerpf(d[0], v[0], v[4]);
  const float c01 = lerpf(d[0], v[1], v[5]);
  const float c10 = lerpf(d[0], v[2], v[6]);
  const float c11 = lerpf(d[0], v[3], v[7]);
  const float c0 = lerpf(d[1], c00, c10);
  const float c1 = lerpf(d[1], c01, c11);
  return lerpf(d[2], c0, c1);
}

// LDS window of means: M[d] nodes starting at world haloed index m0[d]
struct MeanWin {
  int m0[3];
  int M[3];
  __device__ int vol() const { return M[0] * M[1] * M[2]; }
};

__device__ inline void stage_means(const Geo& g, const FieldPtrs& f, const MeanWin& W, float* lds)
{
  const int V = W.vol();
  for(int q = int(threadIdx.x); q < V; q += WG) {
    const int lk = q % W.M[2], lj = (q / W.M[2]) % W.M[1], li = q / (W.M[2] * W.M[1]);
    const Means m = node_means(g, f, W.m0[0] + li, W.m0[1] + lj, W.m0[2] + lk);
    lds[0 * V + q] = m.ex;
    lds[1 * V + q] = m.ey;
    lds[2 * V + q] = m.ez;
    lds[3 * V + q] = m.bx;
    lds[4 * V + q] = m.by;
    lds[5 * V + q] = m.bz;
  }
}

// Interpolated E and B at a particle.  I,J,K: world haloed index of the
// particle's cell; d: offsets inside the cell.
__device__ inline void gather(const Geo& g, const FieldPtrs& f, const MeanWin& W, const float* lds, int I, int J,
                              int K, const float d[3], float E[3], float B[3], uint32_t* err)
{
  const int li = I - W.m0[0], lj = J - W.m0[1], lk = K - W.m0[2];
  float v[6][8];
  if(li >= 0 && lj >= 0 && lk >= 0 && li + 1 < W.M[0] && lj + 1 < W.M[1] && lk + 1 < W.M[2]) {
    const int V = W.vol();
#pragma unroll
    for(int c = 0; c < 8; ++c) {
      const int q = ((li + (c >> 2)) * W.M[1] + (lj + ((c >> 1) & 1))) * W.M[2] + (lk + (c & 1));
#pragma unroll
      for(int comp = 0; comp < 6; ++comp) v[comp][c] = lds[comp * V + q];
    }
  } else {
    // outside the staged window: same arithmetic from HBM (clamped to the lattice)
    int I0 = I, J0 = J, K0 = K;
    if(I0 < 1 || J0 < 1 || K0 < 1 || I0 + 1 > g.L[0] - 1 || J0 + 1 > g.L[1] - 1 || K0 + 1 > g.L[2] - 1) {
      atomicAdd(err, 1u);
      I0 = min(max(I0, 1), g.L[0] - 2);
      J0 = min(max(J0, 1), g.L[1] - 2);
      K0 = min(max(K0, 1), g.L[2] - 2);
    }
#pragma unroll
    for(int c = 0; c < 8; ++c) {
      const Means m = node_means(g, f, I0 + (c >> 2), J0 + ((c >> 1) & 1), K0 + (c & 1));
      v[0][c] = m.ex;
      v[1][c] = m.ey;
      v[2][c] = m.ez;
      v[3][c] = m.bx;
      v[4][c] = m.by;
      v[5][c] = m.bz;
    }
  }
  E[0] = lerp3(d, v[0]);
  E[1] = lerp3(d, v[1]);
  E[2] = lerp3(d, v[2]);
  B[0] = lerp3(d, v[3]);
  B[1] = lerp3(d, v[4]);
  B[2] = lerp3(d, v[5]);
}

__device__ inline float dot3(const float a[3], const float b[3])
{
  float r = a[0] * b[0];
  r = r + a[1] * b[1];
  r = r + a[2] * b[2];
  return r;
}

// x / cfl, correctly rounded.  For 2^-100 <= |x| <= 2^100 the world may use
// one multiply by RN(1/cfl) and one FMA remainder correction (Markstein's
// scheme) -- only when rk_world_create verified it bit-identical to the IEEE
// quotient for every float in that range; other x (zero, subnormal-range,
// huge, NaN) take the division itself.
__device__ inline bool fast_div_range(float x)
{
  const float ax = fabsf(x);
  return ax >= 0x1p-100f && ax <= 0x1p100f;
}

// The IEEE operations behind the fast forms below, kept out of line: with the
// fast forms enabled they run only for inputs outside the verified ranges,
// and inlined they would hold registers across the particle loops.
__device__ __attribute__((noinline)) float ieee_div(float a, float b) { return a / b; }
__device__ __attribute__((noinline)) float ieee_div_sqrt(float a, float x) { return a / sqrtf(x); }

// FAST: 0 = read the world's flags (fast_div, fast_sqrt); -1 = IEEE forms;
// 1 = reciprocal x/cfl only; 2 = reciprocal x/cfl and the sqrt/rcp forms
template<int FAST = 0>
__device__ inline float div_cfl(const Geo& g, float x)
{
  if((FAST > 0 || (FAST == 0 && g.fast_div)) && fast_div_range(x)) {
    const float q = x * g.rcp_cfl;
    const float r = fmaf(-q, g.cfl, x);
    return fmaf(r, g.rcp_cfl, q);
  }
  return FAST > 0 ? ieee_div(x, g.cfl) : x / g.cfl;
}

// Correctly rounded sqrt for 2^-60 <= x <= 2^60 from the hardware estimate
// and a one-ulp residual test on each side, and c / s from the hardware
// reciprocal (one Newton step) and one FMA remainder correction.  Used only
// where the world verified the composite forms below bit-identical to the
// IEEE operations for every float of their range; other inputs take sqrtf
// and '/' themselves.
__device__ inline float sqrt_near(float x)
{
  const float s = __builtin_amdgcn_sqrtf(x);
  const float sm = __uint_as_float(__float_as_uint(s) - 1u);
  const float sp = __uint_as_float(__float_as_uint(s) + 1u);
  const float rm = fmaf(-sm, s, x);
  const float rp = fmaf(-sp, s, x);
  const float r = rm <= 0.0f ? sm : s;
  return rp > 0.0f ? sp : r;
}

__device__ inline float div_near(float c, float s)
{
  float r = __builtin_amdgcn_rcpf(s);
  r = fmaf(fmaf(-s, r, 1.0f), r, r);
  const float q = c * r;
  const float e = fmaf(-q, s, c);
  return fmaf(e, r, q);
}

// cfl / sqrtf(X), X = cfl^2 + u.u (particle_boris.h:43,56)
template<int FAST = 0>
__device__ inline float cfl_over_sqrt(const Geo& g, float X)
{
  if((FAST > 1 || (FAST == 0 && g.fast_sqrt)) && X >= g.cfl * g.cfl && X <= 0x1p60f)
    return div_near(g.cfl, sqrt_near(X));
  return FAST > 1 ? ieee_div_sqrt(g.cfl, X) : g.cfl / sqrtf(X);
}

// c / t for c in {1, 2} and t = 1 + v.v (particle_boris.h:47, zigzag 1/sqrt)
template<int FAST = 0>
__device__ inline float two_over(const Geo& g, float t)
{
  if((FAST > 1 || (FAST == 0 && g.fast_sqrt)) && t >= 1.0f && t <= 0x1p60f) return div_near(2.0f, t);
  return FAST > 1 ? ieee_div(2.0f, t) : 2.0f / t;
}

// 1 / t for t = 1 + b.b (particle_faraday.h)
template<int FAST = 0>
__device__ inline float one_over(const Geo& g, float t)
{
  if((FAST > 1 || (FAST == 0 && g.fast_sqrt)) && t >= 1.0f && t <= 0x1p60f) return div_near(1.0f, t);
  return FAST > 1 ? ieee_div(1.0f, t) : 1.0f / t;
}

template<int FAST = 0>
__device__ inline float inv_sqrt(const Geo& g, float t)
{
  if((FAST > 1 || (FAST == 0 && g.fast_sqrt)) && t >= 1.0f && t <= 0x1p60f) return div_near(1.0f, sqrt_near(t));
  return FAST > 1 ? ieee_div_sqrt(1.0f, t) : 1.0f / sqrtf(t);
}

// particle_boris.h:39-59, expression order as written there
template<int FAST = 0>
__device__ inline void boris(const Geo& g, float qm, const float E[3], const float B[3], float pos[3], float vel[3])
{
  const float cfl = g.cfl;
  const float hq = 0.5f * qm;
  float u0[3], E0[3];
#pragma unroll
  for(int c = 0; c < 3; ++c) {
    E0[c] = E[c] * hq;
    u0[c] = vel[c] * cfl + E0[c];
  }
  const float ginv = cfl_over_sqrt<FAST>(g, cfl * cfl + dot3(u0, u0));
  const float bfac = hq * ginv;
  float B0[3];
#pragma unroll
  for(int c = 0; c < 3; ++c) B0[c] = div_cfl<FAST>(g, B[c] * bfac);
  const float f = two_over<FAST>(g, 1.0f + dot3(B0, B0));
  float u1[3];
  u1[0] = (u0[0] + (u0[1] * B0[2] - u0[2] * B0[1])) * f;
  u1[1] = (u0[1] + (-u0[0] * B0[2] + u0[2] * B0[0])) * f;
  u1[2] = (u0[2] + (u0[0] * B0[1] - u0[1] * B0[0])) * f;
  float u2[3];
  u2[0] = (u0[0] + (u1[1] * B0[2] - u1[2] * B0[1])) + E0[0];
  u2[1] = (u0[1] + (-u1[0] * B0[2] + u1[2] * B0[0])) + E0[1];
  u2[2] = (u0[2] + (u1[0] * B0[1] - u1[1] * B0[0])) + E0[2];
#pragma unroll
  for(int c = 0; c < 3; ++c) vel[c] = div_cfl<FAST>(g, u2[c]);
  const float ginv2 = cfl_over_sqrt<FAST>(g, cfl * cfl + dot3(u2, u2));
#pragma unroll
  for(int c = 0; c < 3; ++c) pos[c] = pos[c] + (vel[c] * ginv2) * cfl;
}

// ---------------------------------------------------------------------------
// zigzag deposit.  x2: particle position relative to the lattice origo of the
// tile that owns it after migration.  Store(i, j, k, mask, jx, jy, jz) receives
// tile-haloed node indices; mask bit c set when component c is structurally
// non-zero.  particle_current_zigzag_1st.c++:249-336
// ---------------------------------------------------------------------------
struct Zig {
  int i1[3], i2[3];
  float Fx1, Fy1, Fz1, Fx2, Fy2, Fz2;
  float Wx1, Wy1, Wz1, Wx2, Wy2, Wz2;
};

template<int FAST = -1>
__device__ inline Zig zigzag_setup(const Geo& g, const float x2[3], const float u[3], float charge, float cfl)
{
  Zig z;
  const float invgam = inv_sqrt<FAST>(g, 1.0f + dot3(u, u));
  const float s = cfl * invgam;
  float x1[3], fi1[3], fi2[3], xr[3], F1[3], F2[3], W1[3], W2[3];
#pragma unroll
  for(int d = 0; d < 3; ++d) {
    x1[d] = x2[d] - u[d] * s;
    fi1[d] = floorf(x1[d]);
    fi2[d] = floorf(x2[d]);
    const float a = (fi1[d] < fi2[d] ? fi1[d] : fi2[d]) + 1.0f;
    const float b1 = (fi1[d] > fi2[d] ? fi1[d] : fi2[d]);
    const float b2 = 0.5f * (x1[d] + x2[d]);
    const float b = (b1 > b2 ? b1 : b2);
    xr[d] = (a < b ? a : b);
    F1[d] = (xr[d] - x1[d]) * charge;
    F2[d] = (x2[d] - xr[d]) * charge;
    W1[d] = (x1[d] + xr[d]) * 0.5f - fi1[d];
    W2[d] = (x2[d] + xr[d]) * 0.5f - fi2[d];
    z.i1[d] = int(fi1[d]);
    z.i2[d] = int(fi2[d]);
  }
  z.Fx1 = F1[0]; z.Fy1 = F1[1]; z.Fz1 = F1[2];
  z.Fx2 = F2[0]; z.Fy2 = F2[1]; z.Fz2 = F2[2];
  z.Wx1 = W1[0]; z.Wy1 = W1[1]; z.Wz1 = W1[2];
  z.Wx2 = W2[0]; z.Wy2 = W2[1]; z.Wz2 = W2[2];
  return z;
}

template<class Store>
__device__ inline void zigzag_emit(const Zig& z, Store&& st)
{
  const float one = 1.0f;
  const int* a = z.i1;
  const int* b = z.i2;
  st(a[0], a[1], a[2], 7, z.Fx1 * (one - z.Wy1) * (one - z.Wz1), z.Fy1 * (one - z.Wx1) * (one - z.Wz1),
     z.Fz1 * (one - z.Wx1) * (one - z.Wy1));
  st(b[0], b[1], b[2], 7, z.Fx2 * (one - z.Wy2) * (one - z.Wz2), z.Fy2 * (one - z.Wx2) * (one - z.Wz2),
     z.Fz2 * (one - z.Wx2) * (one - z.Wy2));
  st(a[0] + 1, a[1], a[2], 6, 0.0f, z.Fy1 * z.Wx1 * (one - z.Wz1), z.Fz1 * z.Wx1 * (one - z.Wy1));
  st(b[0] + 1, b[1], b[2], 6, 0.0f, z.Fy2 * z.Wx2 * (one - z.Wz2), z.Fz2 * z.Wx2 * (one - z.Wy2));
  st(a[0], a[1] + 1, a[2], 5, z.Fx1 * z.Wy1 * (one - z.Wz1), 0.0f, z.Fz1 * (one - z.Wx1) * z.Wy1);
  st(b[0], b[1] + 1, b[2], 5, z.Fx2 * z.Wy2 * (one - z.Wz2), 0.0f, z.Fz2 * (one - z.Wx2) * z.Wy2);
  st(a[0], a[1], a[2] + 1, 3, z.Fx1 * (one - z.Wy1) * z.Wz1, z.Fy1 * (one - z.Wx1) * z.Wz1, 0.0f);
  st(b[0], b[1], b[2] + 1, 3, z.Fx2 * (one - z.Wy2) * z.Wz2, z.Fy2 * (one - z.Wx2) * z.Wz2, 0.0f);
  st(a[0], a[1] + 1, a[2] + 1, 1, z.Fx1 * z.Wy1 * z.Wz1, 0.0f, 0.0f);
  st(b[0], b[1] + 1, b[2] + 1, 1, z.Fx2 * z.Wy2 * z.Wz2, 0.0f, 0.0f);
  st(a[0] + 1, a[1], a[2] + 1, 2, 0.0f, z.Fy1 * z.Wx1 * z.Wz1, 0.0f);
  st(b[0] + 1, b[1], b[2] + 1, 2, 0.0f, z.Fy2 * z.Wx2 * z.Wz2, 0.0f);
  st(a[0] + 1, a[1] + 1, a[2], 4, 0.0f, 0.0f, z.Fz1 * z.Wx1 * z.Wy1);
  st(b[0] + 1, b[1] + 1, b[2], 4, 0.0f, 0.0f, z.Fz2 * z.Wx2 * z.Wy2);
}

// LDS current window: D[d] nodes starting at world haloed index d0[d]
struct CurWin {
  int d0[3];
  int D[3];
  __device__ int vol() const { return D[0] * D[1] * D[2]; }
};

// The brick's current accumulates in FP64 in LDS: on gfx950 ds_add_f64 runs
// about 9x faster than ds_add_f32, and the FP64 sums are exact to far below
// float32 rounding; each node is rounded to float once, when it leaves LDS.
__device__ inline void lds_add(double* p, float v)
{
  __hip_atomic_fetch_add(p, double(v), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ inline void lds_add_d(double* p, double v)
{
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Deposit one particle.  toffh: world haloed index of the owning tile's
// haloed 0; shift: periodic image shift (world cells) applied to fit the window.
__device__ inline void deposit_particle(const Geo& g, const FieldPtrs& f, const CurWin& C, double* ldsJ, const Zig& z,
                                        const int toffh[3], uint32_t* err)
{
  int base1[3], base2[3];
  bool in_win = true;
#pragma unroll
  for(int d = 0; d < 3; ++d) {
    int l1 = toffh[d] + z.i1[d] - C.d0[d];
    int l2 = toffh[d] + z.i2[d] - C.d0[d];
    if(g.periodic) {
      // a particle that wrapped around the world is deposited at the periodic
      // image nearest to this brick (its ghost image is folded later)
      if(l2 < -1) {
        l1 += g.N[d];
        l2 += g.N[d];
      } else if(l2 > C.D[d]) {
        l1 -= g.N[d];
        l2 -= g.N[d];
      }
    }
    base1[d] = l1;
    base2[d] = l2;
    in_win = in_win && l1 >= 0 && l2 >= 0 && l1 + 1 < C.D[d] && l2 + 1 < C.D[d];
  }
  if(in_win) {
    const int V = C.vol();
    // translate tile-haloed node indices into LDS offsets
    const int o1[3] = {base1[0] - z.i1[0], base1[1] - z.i1[1], base1[2] - z.i1[2]};
    (void)base2;
    zigzag_emit(z, [&](int i, int j, int k, int mask, float jx, float jy, float jz) {
      const int q = ((i + o1[0]) * C.D[1] + (j + o1[1])) * C.D[2] + (k + o1[2]);
      if(mask & 1) lds_add(&ldsJ[q], jx);
      if(mask & 2) lds_add(&ldsJ[V + q], jy);
      if(mask & 4) lds_add(&ldsJ[2 * V + q], jz);
    });
  } else {
    // straight to HBM with global atomics (same contributions)
    zigzag_emit(z, [&](int i, int j, int k, int mask, float jx, float jy, float jz) {
      int I = toffh[0] + i, J = toffh[1] + j, K = toffh[2] + k;
      if(I < 0 || J < 0 || K < 0 || I >= g.L[0] || J >= g.L[1] || K >= g.L[2]) {
        atomicAdd(err, 1u);
        return;
      }
      const uint32_t q = g.at(I, J, K);
      if(mask & 1) unsafeAtomicAdd(&f.J[0][q], jx);
      if(mask & 2) unsafeAtomicAdd(&f.J[1][q], jy);
      if(mask & 4) unsafeAtomicAdd(&f.J[2][q], jz);
    });
  }
}

// plain per-tile deposit: add the brick's window into J with global atomics
__device__ inline void flush_current(const Geo& g, const FieldPtrs& f, const CurWin& C, const double* ldsJ)
{
  const int V = C.vol();
  for(int q = int(threadIdx.x); q < V; q += int(blockDim.x)) {
    const int lk = q % C.D[2], lj = (q / C.D[2]) % C.D[1], li = q / (C.D[2] * C.D[1]);
    const uint32_t c = g.at(C.d0[0] + li, C.d0[1] + lj, C.d0[2] + lk);
#pragma unroll
    for(int comp = 0; comp < 3; ++comp) {
      const float v = float(ldsJ[comp * V + q]);
      if(v != 0.0f) unsafeAtomicAdd(&f.J[comp][c], v);
    }
  }
}

// fused step: store the brick's window (rounded to float) to its staging slot
__device__ inline void stage_current(const CurWin& C, const double* ldsJ, float* stage)
{
  const int V3 = 3 * C.vol();
  for(int q = int(threadIdx.x); q < V3; q += int(blockDim.x)) stage[q] = float(ldsJ[q]);
}

// ---------------------------------------------------------------------------
// block-wide helpers
// ---------------------------------------------------------------------------
__device__ inline int lane_id() { return int(threadIdx.x) & 63; }

__device__ inline int64_t readfirstlane64(int64_t v)
{
  const uint32_t lo = __builtin_amdgcn_readfirstlane(uint32_t(uint64_t(v)));
  const uint32_t hi = __builtin_amdgcn_readfirstlane(uint32_t(uint64_t(v) >> 32));
  return int64_t((uint64_t(hi) << 32) | lo);
}
__device__ inline int wave_id() { return int(threadIdx.x) >> 6; }

__device__ inline uint64_t lanemask_lt() { return (uint64_t(1) << lane_id()) - 1; }

// exclusive block scan of a 0/1 flag over NT lanes; returns rank, writes total.
// The per-wave counts are scanned across lanes 0..NW-1 with DPP row shifts and
// picked with readlane at the (scalar) wave index: no per-wave select masks.
template<int NT = WG>
__device__ inline int block_rank(bool flag, int* wcnt, int& total)
{
  constexpr int NW = NT / 64;
  static_assert(NW <= 16, "one DPP row");
  const uint64_t bal = __ballot(flag);
  const int wid = __builtin_amdgcn_readfirstlane(int(threadIdx.x) >> 6);
  if(lane_id() == 0) wcnt[wid] = __popcll(bal);
  __syncthreads();
  const int l = lane_id();
  const int c = l < NW ? wcnt[l] : 0;
  int x = c;
  x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, true);  // row_shr:1
  x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, true);  // row_shr:2
  x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, true);  // row_shr:4
  if(NW > 8) x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, true);  // row_shr:8
  const int off = __builtin_amdgcn_readlane(x - c, wid);
  total = __builtin_amdgcn_readlane(x, NW - 1);
  __syncthreads();
  return off + __popcll(bal & lanemask_lt());
}

// Same scan with one barrier: the per-wave counts go to one of two buffers
// that consecutive calls alternate (wcnt holds 2*NT/64 ints, par is flipped
// here).  Call k+2 overwrites call k's buffer only after every wave passed
// call k+1's barrier, i.e. after every read of call k.
template<int NT>
__device__ inline int block_rank1(bool flag, int* wcnt, int& par, int& total)
{
  constexpr int NW = NT / 64;
  int* buf = wcnt + par * NW;
  par ^= 1;
  const uint64_t bal = __ballot(flag);
  const int wid = __builtin_amdgcn_readfirstlane(int(threadIdx.x) >> 6);
  if(lane_id() == 0) buf[wid] = __popcll(bal);
  __syncthreads();
  const int l = lane_id();
  const int c = l < NW ? buf[l] : 0;
  int x = c;
  x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, true);  // row_shr:1
  x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, true);  // row_shr:2
  x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, true);  // row_shr:4
  if(NW > 8) x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, true);  // row_shr:8
  const int off = __builtin_amdgcn_readlane(x - c, wid);
  total = __builtin_amdgcn_readlane(x, NW - 1);
  return off + __popcll(bal & lanemask_lt());
}

// reserve one slot per flagged lane in a counter with a single atomic per wave
__device__ inline uint32_t wave_reserve(bool flag, uint32_t* counter)
{
  const uint64_t bal = __ballot(flag);
  if(bal == 0) return 0;
  const int leader = __ffsll((unsigned long long)bal) - 1;
  uint32_t base = 0;
  if(lane_id() == leader) base = atomicAdd(counter, uint32_t(__popcll(bal)));
  base = __shfl(base, leader);
  return base + uint32_t(__popcll(bal & lanemask_lt()));
}

__device__ inline void write_record(const MigPtrs& m, int set, uint32_t slot, const float p[3], const float v[3],
                                    uint64_t id, uint32_t meta)
{
#pragma unroll
  for(int d = 0; d < 3; ++d) {
    m.pos[set][d][slot] = p[d];
    m.vel[set][d][slot] = v[d];
  }
  m.id[set][slot] = id;
  m.meta[set][slot] = meta;
}

// LDS for bin compaction
template<int NT = WG>
struct CompactLdsT {
  uint32_t holes[MAX_HOLES];
  uint32_t tailbits[MAX_HOLES / 32];
  int wcnt[2 * (NT / 64)];  // block_rank uses the first half, block_rank1 both
  int nholes;
};
using CompactLds = CompactLdsT<WG>;

// Swap-remove compaction of one bin after the leavers' slots were recorded in
// L.holes (ascending) and their ids set to DEAD_ID.
template<int NT = WG>
__device__ void compact_bin(const SpParams& sp, int s, int64_t start, int count, int nholes, CompactLdsT<NT>& L,
                            uint32_t* counters)
{
  const int newc = count - nholes;
  if(nholes == 0) return;
  float* P[3] = {sp.pos[s][0], sp.pos[s][1], sp.pos[s][2]};
  float* U[3] = {sp.vel[s][0], sp.vel[s][1], sp.vel[s][2]};
  uint64_t* ID = sp.id[s];
  if(nholes <= MAX_HOLES) {
    // tail bitmap of holes at index >= newc
    for(int q = int(threadIdx.x); q < (nholes + 31) / 32; q += NT) L.tailbits[q] = 0;
    __syncthreads();
    for(int q = int(threadIdx.x); q < nholes; q += NT) {
      const uint32_t h = L.holes[q];
      if(int(h) >= newc) atomicOr(&L.tailbits[(int(h) - newc) >> 5], 1u << ((int(h) - newc) & 31));
    }
    __syncthreads();
    int moved = 0;
    for(int base = 0; base < nholes; base += NT) {
      const int t = base + int(threadIdx.x);
      bool stay = false;
      if(t < nholes) stay = !((L.tailbits[t >> 5] >> (t & 31)) & 1u);
      int tot;
      const int r = block_rank<NT>(stay, L.wcnt, tot);
      if(stay) {
        const int64_t src = start + newc + t;
        const int64_t dst = start + int64_t(L.holes[moved + r]);
#pragma unroll
        for(int d = 0; d < 3; ++d) {
          P[d][dst] = P[d][src];
          U[d][dst] = U[d][src];
        }
        ID[dst] = ID[src];
      }
      moved += tot;
    }
  } else {
    // slow path: order-preserving in-place compaction over DEAD_ID marks
    if(threadIdx.x == 0) atomicAdd(&counters[C_SLOWCOMPACT], 1u);
    int wr = 0;
    for(int base = 0; base < count; base += NT) {
      const int t = base + int(threadIdx.x);
      bool alive = false;
      float p[3], v[3];
      uint64_t id = 0;
      if(t < count) {
        id = ID[start + t];
        alive = id != DEAD_ID;
        if(alive) {
#pragma unroll
          for(int d = 0; d < 3; ++d) {
            p[d] = P[d][start + t];
            v[d] = U[d][start + t];
          }
        }
      }
      int tot;
      const int r = block_rank<NT>(alive, L.wcnt, tot);
      if(alive) {
        const int64_t dst = start + wr + r;
#pragma unroll
        for(int d = 0; d < 3; ++d) {
          P[d][dst] = p[d];
          U[d][dst] = v[d];
        }
        ID[dst] = id;
      }
      wr += tot;
      __syncthreads();
    }
  }
}

__device__ inline int tile_dir(const float p[3], const float lo[3], const float hi[3])
{
  int d[3];
#pragma unroll
  for(int a = 0; a < 3; ++a) d[a] = int(p[a] >= lo[a]) - int(p[a] < hi[a]);
  return 9 * (d[0] + 1) + 3 * (d[1] + 1) + (d[2] + 1);
}

__device__ inline float wrap_coord(float x, float a, float b)
{
  // (x < 0 ? max : min) + fmod(x, L), particle.h:523-549
  const float Ld = b - a;
  return (x < 0 ? b : a) + fmodf(x, Ld);
}

__device__ inline int bin_total(const SpParams& sp, int b)
{
  int t = 0;
  for(int s = 0; s < sp.nsp; ++s) t += int(sp.count[s][b]);
  return t;
}

__device__ inline float dot3p(const float a[3], const float b[3])
{
  float r = a[0] * b[0];
  r = r + a[1] * b[1];
  r = r + a[2] * b[2];
  return r;
}

__device__ inline void cross3p(const float a[3], const float b[3], float r[3])
{
  r[0] = a[1] * b[2] - a[2] * b[1];
  r[1] = -a[0] * b[2] + a[2] * b[0];
  r[2] = a[0] * b[1] - a[1] * b[0];
}

// push_particles_higuera_cary, src/runko/pic/particle_higuera_cary.h:34-73,
// expression order as written there (scalar * vector is v(i) * s)
template<int FAST = 0>
__device__ inline void higuera_cary(const Geo& g, float qm, const float E[3], const float B[3], float pos[3],
                                    float vel[3])
{
  const float cfl = g.cfl;
  const float hqm = 0.5f * qm;
  const float cfl2 = cfl * cfl;
  const float cinv = 1.0f / cfl;
  const float cinv2 = cinv * cinv;
  float E0[3], u0[3], Bt[3];
#pragma unroll
  for(int c = 0; c < 3; ++c) {
    E0[c] = E[c] * hqm;
    u0[c] = vel[c] * cfl + E0[c];
    Bt[c] = B[c] * hqm;
  }
  const float u0sq = dot3p(u0, u0);
  const float b2 = dot3p(Bt, Bt);
  const float bdotu = dot3p(Bt, u0);
  const float gmb = 1.0f + u0sq * cinv2 - b2 * cinv2;
  const float disc = gmb * gmb + 4.0f * (b2 * cinv2 + bdotu * bdotu * cinv2);
  const float ginv = inv_sqrt<FAST>(g, 0.5f * (gmb + sqrtf(disc)));
  const float gc = ginv * cinv;
  float B0[3];
#pragma unroll
  for(int c = 0; c < 3; ++c) B0[c] = Bt[c] * gc;
  const float f = two_over<FAST>(g, 1.0f + gc * gc * b2);
  float cr[3], u1[3], u2[3];
  cross3p(u0, B0, cr);
#pragma unroll
  for(int c = 0; c < 3; ++c) u1[c] = (u0[c] + cr[c]) * f;
  cross3p(u1, B0, cr);
#pragma unroll
  for(int c = 0; c < 3; ++c) u2[c] = (u0[c] + cr[c]) + E0[c];
  const float ginv2 = cfl_over_sqrt<FAST>(g, cfl2 + dot3p(u2, u2));
#pragma unroll
  for(int c = 0; c < 3; ++c) {
    vel[c] = u2[c] * cinv;
    pos[c] = pos[c] + u2[c] * ginv2;
  }
}

// push_particles_faraday, src/runko/pic/particle_faraday.h:54-104 (Cayley
// transform of the field bivector, Sherman-Morrison reduction)
template<int FAST = 0>
__device__ inline void faraday(const Geo& g, float qm, const float E[3], const float B[3], float pos[3],
                               float vel[3])
{
  const float cfl = g.cfl;
  float v0[3];
#pragma unroll
  for(int c = 0; c < 3; ++c) v0[c] = vel[c] * cfl;
  const float gcfl = sqrtf(cfl * cfl + dot3p(v0, v0));
  const float hq = 0.5f * qm;
  float u0[3];
#pragma unroll
  for(int c = 0; c < 3; ++c) u0[c] = v0[c] + E[c] * hq;
  const float geff_cfl = sqrtf(cfl * cfl + dot3p(u0, u0));
  const float kappa = hq / geff_cfl;
  float eps[3], beta[3];
#pragma unroll
  for(int c = 0; c < 3; ++c) {
    eps[c] = E[c] * kappa;
    beta[c] = B[c] * kappa;
  }
  const float w0 = gcfl + dot3p(eps, v0);
  float cr[3], W[3];
  cross3p(v0, beta, cr);
#pragma unroll
  for(int c = 0; c < 3; ++c) W[c] = ((v0[c] + eps[c] * gcfl) + cr[c]) + eps[c] * w0;
  const float b2 = dot3p(beta, beta);
  const float f = one_over<FAST>(g, 1.0f + b2);
  cross3p(beta, W, cr);
  const float bW = dot3p(beta, W);
  float Wr[3];
#pragma unroll
  for(int c = 0; c < 3; ++c) Wr[c] = ((W[c] - cr[c]) + beta[c] * bW) * f;
  const float bde = dot3p(beta, eps);
  cross3p(beta, eps, cr);
  float er[3];
#pragma unroll
  for(int c = 0; c < 3; ++c) er[c] = ((eps[c] - cr[c]) + beta[c] * bde) * f;
  const float D = 1.0f - f * (dot3p(eps, eps) + bde * bde);
  const float sc = dot3p(eps, Wr) / D;
  float u2[3];
#pragma unroll
  for(int c = 0; c < 3; ++c) u2[c] = Wr[c] + er[c] * sc;
#pragma unroll
  for(int c = 0; c < 3; ++c) vel[c] = div_cfl<FAST>(g, u2[c]);
  const float ginv2 = cfl_over_sqrt<FAST>(g, cfl * cfl + dot3p(u2, u2));
#pragma unroll
  for(int c = 0; c < 3; ++c) pos[c] = pos[c] + vel[c] * ginv2 * cfl;
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
template<int PUSHER>
__global__ void __launch_bounds__(WG)
  k_push(Geo g, SpParams sp, FieldPtrs f, int bin0, int nbins, int span, uint32_t* counters)
{
  extern __shared__ float lds[];
  const int b = bin0 + xcd_brick(int(blockIdx.x), span);
  if(b >= bin0 + nbins) return;
  if(bin_total(sp, b) == 0) return;
  const Brick Bk = brick_of(g, b);
  MeanWin W;
#pragma unroll
  for(int d = 0; d < 3; ++d) {
    W.m0[d] = Bk.c0[d];
    W.M[d] = g.br[d] + 2;
  }
  stage_means(g, f, W, lds);
  __syncthreads();
  for(int s = 0; s < sp.nsp; ++s) {
    const int64_t start = sp.start[s][b];
    const int count = int(sp.count[s][b]);
    const float qm = sp.qm[s];
    for(int t = int(threadIdx.x); t < count; t += WG) {
      const int64_t q = start + t;
      float pos[3] = {sp.pos[s][0][q], sp.pos[s][1][q], sp.pos[s][2][q]};
      float vel[3] = {sp.vel[s][0][q], sp.vel[s][1][q], sp.vel[s][2][q]};
      float pr[3], dd[3];
      int ci[3];
#pragma unroll
      for(int d = 0; d < 3; ++d) {
        pr[d] = pos[d] - Bk.origo[d];
        ci[d] = int(uint32_t(pr[d]));
        dd[d] = pr[d] - float(uint32_t(ci[d]));
      }
      float E[3], Bf[3];
      gather(g, f, W, lds, Bk.toffh[0] + ci[0], Bk.toffh[1] + ci[1], Bk.toffh[2] + ci[2], dd, E, Bf,
             &counters[C_STENCIL]);
      if constexpr(PUSHER == 1)
        higuera_cary(g, qm, E, Bf, pos, vel);
      else if constexpr(PUSHER == 2)
        faraday(g, qm, E, Bf, pos, vel);
      else
        boris(g, qm, E, Bf, pos, vel);
#pragma unroll
      for(int d = 0; d < 3; ++d) {
        sp.pos[s][d][q] = pos[d];
        sp.vel[s][d][q] = vel[d];
      }
    }
  }
}

__global__ void __launch_bounds__(WG)
  k_deposit(Geo g, SpParams sp, FieldPtrs f, int bin0, int nbins, int span, uint32_t* counters)
{
  extern __shared__ double ldsd[];
  double* lds = ldsd;
  const int b = bin0 + xcd_brick(int(blockIdx.x), span);
  if(b >= bin0 + nbins) return;
  if(bin_total(sp, b) == 0) return;
  const Brick Bk = brick_of(g, b);
  CurWin C;
#pragma unroll
  for(int d = 0; d < 3; ++d) {
    C.d0[d] = Bk.c0[d] - 1;
    C.D[d] = g.br[d] + 3;
  }
  const int V = C.vol();
  for(int q = int(threadIdx.x); q < 3 * V; q += WG) lds[q] = 0.0;
  __syncthreads();
  for(int s = 0; s < sp.nsp; ++s) {
    const int64_t start = sp.start[s][b];
    const int count = int(sp.count[s][b]);
    for(int t = int(threadIdx.x); t < count; t += WG) {
      const int64_t q = start + t;
      const float u[3] = {sp.vel[s][0][q], sp.vel[s][1][q], sp.vel[s][2][q]};
      const float x2[3] = {sp.pos[s][0][q] - Bk.origo[0], sp.pos[s][1][q] - Bk.origo[1],
                           sp.pos[s][2][q] - Bk.origo[2]};
      const Zig z = zigzag_setup<0>(g, x2, u, sp.charge[s], g.cfl);
      deposit_particle(g, f, C, lds, z, Bk.toffh, &counters[C_STENCIL]);
    }
  }
  __syncthreads();
  flush_current(g, f, C, lds);
}

// pack_outgoing: tile leavers -> record set 0, brick changes -> set 1
__global__ void __launch_bounds__(WG)
  k_pack(Geo g, SpParams sp, MigPtrs m, int bin0, int nbins, uint32_t* counters)
{
  __shared__ CompactLds L;
  const int b = bin0 + int(blockIdx.x);
  if(b >= bin0 + nbins) return;
  const Brick Bk = brick_of(g, b);
  for(int s = 0; s < sp.nsp; ++s) {
    const int64_t start = sp.start[s][b];
    const int count = int(sp.count[s][b]);
    if(count == 0) continue;
    int nholes = 0;
    for(int base = 0; base < count; base += WG) {
      const int t = base + int(threadIdx.x);
      bool lv0 = false, lv1 = false;
      float p[3], v[3];
      uint64_t id = 0;
      int dir = 13;
      if(t < count) {
        const int64_t q = start + t;
#pragma unroll
        for(int d = 0; d < 3; ++d) {
          p[d] = sp.pos[s][d][q];
          v[d] = sp.vel[s][d][q];
        }
        dir = tile_dir(p, Bk.tmin, Bk.tmax);
        if(dir != 13) lv0 = true;
        else lv1 = brick_in_tile(g, Bk.tile, Bk.origo, p) != b;
        if(lv0 || lv1) id = sp.id[s][q];
      }
      uint32_t slot0 = wave_reserve(lv0, &counters[C_MIG0]);
      uint32_t slot1 = wave_reserve(lv1, &counters[C_MIG1]);
      if(lv0 && int64_t(slot0) >= m.cap[0]) {
        lv0 = false;
        atomicAdd(&counters[C_MIGLOST], 1u);
      }
      if(lv1 && int64_t(slot1) >= m.cap[1]) {
        lv1 = false;
        atomicAdd(&counters[C_MIGLOST], 1u);
      }
      if(lv0) write_record(m, 0, slot0, p, v, id, meta_pack(s, 0, dir, uint32_t(b)));
      if(lv1) write_record(m, 1, slot1, p, v, id, meta_pack(s, 1, 13, uint32_t(b)));
      const bool hole = lv0 || lv1;
      if(hole) sp.id[s][start + t] = DEAD_ID;
      int tot;
      const int r = block_rank(hole, L.wcnt, tot);
      if(hole && nholes + r < MAX_HOLES) L.holes[nholes + r] = uint32_t(t);
      nholes += tot;
    }
    __syncthreads();
    compact_bin(sp, s, start, count, nholes, L, counters);
    __syncthreads();
    if(threadIdx.x == 0) sp.count[s][b] = uint32_t(count - nholes);
    __syncthreads();
  }
}

// Destination of one migration record.  Tile leavers (kind 0) go to the
// neighbouring tile in their direction with every coordinate wrapped by the
// global box; brick changes (kind 1) stay in the source tile.
struct RecordDest {
  int s;     // species
  int db;    // destination bin
  float p[3];  // position (wrapped for tile leavers)
};

__device__ inline RecordDest record_dest(const Geo& g, const MigPtrs& m, int st, uint32_t r)
{
  RecordDest R;
  const uint32_t meta = m.meta[st][r];
  R.s = meta_species(meta);
  const int kind = meta_kind(meta);
  const int srcb = int(meta_bin(meta));
  int tile = srcb / g.bpt;
#pragma unroll
  for(int d = 0; d < 3; ++d) R.p[d] = m.pos[st][d][r];
  if(kind == 0) {
    int tc[3];
    g.tile_coords(tile, tc);
    const int dir = meta_dir(meta);
    const int dd[3] = {dir / 9 - 1, (dir / 3) % 3 - 1, dir % 3 - 1};
#pragma unroll
    for(int d = 0; d < 3; ++d) {
      tc[d] = (tc[d] + dd[d] + g.nt[d]) % g.nt[d];
      R.p[d] = wrap_coord(R.p[d], g.wmin[d], g.wmax[d]);
    }
    tile = g.tile_lin(tc[0], tc[1], tc[2]);
  }
  int tc[3];
  g.tile_coords(tile, tc);
  float origo[3];
#pragma unroll
  for(int d = 0; d < 3; ++d) origo[d] = g.tile_min(tc[d], d) - 3.0f;
  R.db = brick_in_tile(g, tile, origo, R.p);
  return R;
}

// Store a record at slot `slot` of its destination bin, or, past the bin's
// capacity, keep it (position wrapped, re-tagged as a brick change into db)
// on the overflow list for the retry after a repack.
__device__ inline void place_record(const SpParams& sp, const MigPtrs& m, int st, uint32_t r, const RecordDest& R,
                                    uint32_t slot, uint32_t* counters, uint32_t* ovf, int64_t ovf_cap)
{
  if(slot < sp.cap[R.s][R.db]) {
    const int64_t q = sp.start[R.s][R.db] + slot;
#pragma unroll
    for(int d = 0; d < 3; ++d) {
      sp.pos[R.s][d][q] = R.p[d];
      sp.vel[R.s][d][q] = m.vel[st][d][r];
    }
    sp.id[R.s][q] = m.id[st][r];
  } else {
#pragma unroll
    for(int d = 0; d < 3; ++d) m.pos[st][d][r] = R.p[d];
    m.meta[st][r] = meta_pack(R.s, 1, 13, uint32_t(R.db));
    const uint32_t o = atomicAdd(&counters[C_OVF], 1u);
    if(int64_t(o) < ovf_cap) ovf[o] = (uint32_t(st) << 30) | r;
    else atomicAdd(&counters[C_MIGLOST], 1u);
  }
}

__device__ inline void append_record(const Geo& g, const SpParams& sp, const MigPtrs& m, int st, uint32_t r,
                                     uint32_t* counters, uint32_t* ovf, int64_t ovf_cap)
{
  const RecordDest R = record_dest(g, m, st, r);
  place_record(sp, m, st, r, R, atomicAdd(&sp.count[R.s][R.db], 1u), counters, ovf, ovf_cap);
}

// append the records of set 0 or 1 (list == nullptr) or of an overflow list
__global__ void __launch_bounds__(WG)
  k_append(Geo g, SpParams sp, MigPtrs m, int set, uint32_t* counters, uint32_t* ovf, int64_t ovf_cap,
           const uint32_t* list, int list_n)
{
  // the counter can exceed the buffer when reservations were refused (those
  // particles stayed in their bins); never read past the record buffer
  const int n = list ? list_n : int(min(int64_t(counters[set == 0 ? C_MIG0 : C_MIG1]), m.cap[set]));
  for(int e = int(blockIdx.x * blockDim.x + threadIdx.x); e < n; e += int(gridDim.x * blockDim.x)) {
    int st = set;
    uint32_t r = uint32_t(e);
    if(list) {
      st = int(list[e] >> 30);
      r = list[e] & 0x3fffffffu;
      if(st > 2 || int64_t(r) >= m.cap[st]) {  // not a record of the current buffers
        atomicAdd(&counters[C_MIGLOST], 1u);
        continue;
      }
    }
    append_record(g, sp, m, st, r, counters, ovf, ovf_cap);
  }
}

// Append the records of sets 0 and 1 (blockIdx.y is the set), then clear
// both record counters: every workgroup reads its set's count before it
// signs off on C_APPDONE, and the last workgroup to sign off clears C_MIG0,
// C_MIG1 and C_APPDONE, so the counters are zero for the next step without
// separate launches.
__global__ void __launch_bounds__(WG)
  k_append_records(Geo g, SpParams sp, MigPtrs m, uint32_t* counters, uint32_t* ovf, int64_t ovf_cap)
{
  const int set = int(blockIdx.y);
  const int n = int(min(int64_t(counters[set == 0 ? C_MIG0 : C_MIG1]), m.cap[set]));
  for(int e = int(blockIdx.x * blockDim.x + threadIdx.x); e < n; e += int(gridDim.x * blockDim.x))
    append_record(g, sp, m, set, uint32_t(e), counters, ovf, ovf_cap);
  __syncthreads();
  if(threadIdx.x == 0) {
    __threadfence();
    const uint32_t done = atomicAdd(&counters[C_APPDONE], 1u);
    if(done == gridDim.x * gridDim.y - 1) {
      counters[C_MIG0] = 0;
      counters[C_MIG1] = 0;
      counters[C_APPDONE] = 0;
    }
  }
}

// Global brick coordinate of bin b along axis d.
__device__ inline int global_brick(const Geo& g, int b, int d)
{
  int tile, lb[3], tc[3];
  g.brick_coords(b, tile, lb);
  g.tile_coords(tile, tc);
  return tc[d] * g.nb[d] + lb[d];
}

// Append every brick's mailbox, then empty it.  Records of one brick go to
// the brick itself or its 26 periodic neighbours: per chunk of WG records
// the slots of each (species, neighbour) pair are ranked with LDS atomics and
// reserved with one global atomic per pair.
__global__ void __launch_bounds__(WG)
  k_append_mailbox(Geo g, SpParams sp, MigPtrs m, int nbins, uint32_t* counters, uint32_t* ovf, int64_t ovf_cap, int span)
{
  constexpr int NK = RK_MAX_SPECIES * 27;
  __shared__ uint32_t cnt[NK];
  __shared__ uint32_t kbase[NK];
  __shared__ int kdb[NK];
  // one block per brick, with the fused step's XCD-aware brick map: a brick's
  // neighbours mostly run on the same XCD, so the appends to one destination
  // bin's tail mostly go through one L2
  {
    const int b = xcd_brick(int(blockIdx.x), span);
    if(b >= nbins) return;  // uniform over the block
    const int n = int(m.mcount[b]);
    if(n == 0) return;  // uniform over the block
    const uint32_t base = uint32_t(int64_t(b) * m.per_brick);
    int gsrc[3];
#pragma unroll
    for(int d = 0; d < 3; ++d) gsrc[d] = global_brick(g, b, d);
    for(int c0 = 0; c0 < n; c0 += WG) {
      for(int k = int(threadIdx.x); k < NK; k += WG) cnt[k] = 0;
      __syncthreads();
      const int e = c0 + int(threadIdx.x);
      RecordDest R;
      int key = -1;
      uint32_t rank = 0;
      // the record's payload is loaded with its destination, so its latency
      // overlaps the ranking and the reservation
      float rv[3] = {0.0f, 0.0f, 0.0f};
      uint64_t rid = 0;
      if(e < n) {
        const uint32_t r = base + uint32_t(e);
#pragma unroll
        for(int d = 0; d < 3; ++d) rv[d] = m.vel[2][d][r];
        rid = m.id[2][r];
        R = record_dest(g, m, 2, r);
        int off = 0;
        bool near = true;
#pragma unroll
        for(int d = 0; d < 3; ++d) {
          const int NB = g.nt[d] * g.nb[d];
          int df = global_brick(g, R.db, d) - gsrc[d];
          df = df < 0 ? df + NB : df;
          const int o = df == 0 ? 0 : (df == 1 ? 1 : (df == NB - 1 ? -1 : 2));
          near = near && o != 2;
          off = off * 3 + (o + 1);
        }
        if(near) {
          key = R.s * 27 + off;
          rank = atomicAdd(&cnt[key], 1u);
          kdb[key] = R.db;  // every writer of a key stores the same bin
        }
      }
      __syncthreads();
      for(int k = int(threadIdx.x); k < NK; k += WG)
        if(cnt[k]) kbase[k] = atomicAdd(&sp.count[k / 27][kdb[k]], cnt[k]);
      __syncthreads();
      if(e < n) {
        const uint32_t slot = key >= 0 ? kbase[key] + rank : atomicAdd(&sp.count[R.s][R.db], 1u);
        const uint32_t cap = sp.cap[R.s][R.db];
        const int64_t q = sp.start[R.s][R.db] + slot;
        if(slot < cap) {
#pragma unroll
          for(int d = 0; d < 3; ++d) {
            sp.pos[R.s][d][q] = R.p[d];
            sp.vel[R.s][d][q] = rv[d];
          }
          sp.id[R.s][q] = rid;
        } else {
          place_record(sp, m, 2, base + uint32_t(e), R, slot, counters, ovf, ovf_cap);  // overflow path
        }
      }
      __syncthreads();
    }
    if(threadIdx.x == 0) m.mcount[b] = 0;
  }
}

__global__ void k_reset(uint32_t* counters, int which)
{
  if(threadIdx.x == 0) counters[which] = 0;
}

// ---------------------------------------------------------------------------
// fused batched step
// ---------------------------------------------------------------------------
// Brick shape known at compile time (BR > 0: BR^3 bricks) or read from Geo
// (BR == 0).  With a compile-time shape every LDS address of the gather and
// of the deposit is one per-particle base plus immediate offsets.
template<int BR>
struct Shape {
  int b[3];
  __device__ explicit Shape(const Geo& g)
  {
#pragma unroll
    for(int d = 0; d < 3; ++d) b[d] = BR > 0 ? BR : g.br[d];
  }
  __device__ int M(int d) const { return b[d] + 2; }  // node means staged per axis
  __device__ int D(int d) const { return b[d] + 3; }  // current nodes per axis
  __device__ int MV() const { return M(0) * M(1) * M(2); }
  __device__ int DV() const { return D(0) * D(1) * D(2); }
  __device__ int div(int c, const Geo& g, int d) const { return BR > 0 ? c / BR : g.div_br(c, d); }
};

// node means of the brick window, one array per component (ex ey ez bx by bz)
template<int BR, int NT>
__device__ inline void stage_means6(const Geo& g, const FieldPtrs& f, const Shape<BR>& S, const int m0[3], float* lds)
{
  const int V = S.MV();
  for(int q = int(threadIdx.x); q < V; q += NT) {
    const int lk = q % S.M(2), lj = (q / S.M(2)) % S.M(1), li = q / (S.M(2) * S.M(1));
    const Means m = node_means(g, f, m0[0] + li, m0[1] + lj, m0[2] + lk);
    lds[q] = m.ex;
    lds[V + q] = m.ey;
    lds[2 * V + q] = m.ez;
    lds[3 * V + q] = m.bx;
    lds[4 * V + q] = m.by;
    lds[5 * V + q] = m.bz;
  }
}

// gather from the staged means; false when a corner lies outside the window
template<int BR>
__device__ inline bool gather6(const Shape<BR>& S, const float* lds, int li, int lj, int lk, const float d[3],
                               float E[3], float B[3])
{
  if(!(li >= 0 && lj >= 0 && lk >= 0 && li + 1 < S.M(0) && lj + 1 < S.M(1) && lk + 1 < S.M(2))) return false;
  const int V = S.MV();
  const int p = (li * S.M(1) + lj) * S.M(2) + lk;
  float v[6][8];
#pragma unroll
  for(int comp = 0; comp < 6; ++comp) {
    // the component's base index stays one register: its 8 corners (at most
    // M1*M2 + M2 + 1 floats apart) are then immediate offsets of LDS reads
    const __attribute__((address_space(3))) float* pc =
      (const __attribute__((address_space(3))) float*)(lds + p + comp * V);
    asm("" : "+v"(pc));
#pragma unroll
    for(int c = 0; c < 8; ++c) {
      const int o = ((c >> 2) * S.M(1) + ((c >> 1) & 1)) * S.M(2) + (c & 1);
      v[comp][c] = pc[o];
    }
  }
  E[0] = lerp3(d, v[0]);
  E[1] = lerp3(d, v[1]);
  E[2] = lerp3(d, v[2]);
  B[0] = lerp3(d, v[3]);
  B[1] = lerp3(d, v[4]);
  B[2] = lerp3(d, v[5]);
  return true;
}

// the 12 non-zero contributions of a zigzag segment, in the order of
// particle_current_zigzag_1st.c++:300-336: node offsets (x, y, z bits) and
// component of each
struct SegVals {
  float v[12];
};

__device__ inline SegVals segment_values(float Fx, float Fy, float Fz, float Wx, float Wy, float Wz)
{
  const float one = 1.0f;
  SegVals r;
  r.v[0] = Fx * (one - Wy) * (one - Wz);   // x at 000
  r.v[1] = Fy * (one - Wx) * (one - Wz);   // y at 000
  r.v[2] = Fz * (one - Wx) * (one - Wy);   // z at 000
  r.v[3] = Fy * Wx * (one - Wz);           // y at 100
  r.v[4] = Fz * Wx * (one - Wy);           // z at 100
  r.v[5] = Fx * Wy * (one - Wz);           // x at 010
  r.v[6] = Fz * (one - Wx) * Wy;           // z at 010
  r.v[7] = Fx * (one - Wy) * Wz;           // x at 001
  r.v[8] = Fy * (one - Wx) * Wz;           // y at 001
  r.v[9] = Fx * Wy * Wz;                   // x at 011
  r.v[10] = Fy * Wx * Wz;                  // y at 101
  r.v[11] = Fz * Wx * Wy;                  // z at 110
  return r;
}

// The fused step accumulates each brick's current in LDS as 64-bit fixed
// point.  The deposit runs with the charge scaled by jscale (a power of two
// with max|q| * jscale <= 2^30), so every contribution is value * jscale
// exactly and lies in [-2^30, 2^30]; it is rounded to an integer (half-way
// cases up) by v_cvt_rpi_i32_f32 and sign-extended.  Integer adds are exact
// and order-free (the window is bit-reproducible), and ds_add_u64 runs about
// twice as fast as ds_add_f64 on gfx950.  The rounding unit is 2^-30 max|q|,
// far below the float rounding of a node's current.
__device__ inline int fixed_of(float v)
{
  int r;
  asm("v_cvt_rpi_i32_f32 %0, %1" : "=v"(r) : "v"(v));
  return r;
}

__device__ inline void lds_add_u64(uint64_t* p, uint64_t v)
{
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// add 12 fixed-point contributions (the node order of segment_values) at
// window node base p
template<int BR>
__device__ inline void emit_fixed(const Shape<BR>& S, uint64_t* p, const int (&r)[12])
{
  const int sx = S.D(1) * S.D(2), sy = S.D(2), V = S.DV();
  uint64_t* px = p;
  uint64_t* py = p + V;
  uint64_t* pz = p + 2 * V;
  auto x = [](int v) { return uint64_t(int64_t(v)); };
  lds_add_u64(px, x(r[0]));
  lds_add_u64(py, x(r[1]));
  lds_add_u64(pz, x(r[2]));
  lds_add_u64(py + sx, x(r[3]));
  lds_add_u64(pz + sx, x(r[4]));
  lds_add_u64(px + sy, x(r[5]));
  lds_add_u64(pz + sy, x(r[6]));
  lds_add_u64(px + 1, x(r[7]));
  lds_add_u64(py + 1, x(r[8]));
  lds_add_u64(px + sy + 1, x(r[9]));
  lds_add_u64(py + sx + 1, x(r[10]));
  lds_add_u64(pz + sx + sy, x(r[11]));
}

// Deposit into the brick's current window; z carries the scaled charge.
// off[d] maps a tile-haloed node index of the owning tile to a window index
// (periodic image included).  False when a node falls outside the window
// (the caller then uses HBM).  Each segment's 12 contributions are rounded to
// fixed point as the reference's atomic deposit adds them, one segment after
// the other.  A particle that stays in its cell adds both segments at the
// same 12 nodes: their integer sums go in with one set of atomics, which
// leaves every window sum unchanged (integer adds are exact and order-free).
// The sums cannot overflow: both segments then move the same way along each
// axis, so |F1| + |F2| = |q| |x2 - x1| < |q| on every axis, every weight lies
// in [0, 1], and each contribution is at most 2^30 (max|q| * jscale) before
// rounding to an integer.
template<int BR>
__device__ inline bool deposit_win(const Shape<BR>& S, uint64_t* ldsJ, const Zig& z, const int off[3])
{
  int a[3], b[3];
  bool in = true;
#pragma unroll
  for(int d = 0; d < 3; ++d) {
    a[d] = z.i1[d] + off[d];
    b[d] = z.i2[d] + off[d];
    in = in && a[d] >= 0 && b[d] >= 0 && a[d] + 1 < S.D(d) && b[d] + 1 < S.D(d);
  }
  if(!in) return false;
  const SegVals v1 = segment_values(z.Fx1, z.Fy1, z.Fz1, z.Wx1, z.Wy1, z.Wz1);
  const SegVals v2 = segment_values(z.Fx2, z.Fy2, z.Fz2, z.Wx2, z.Wy2, z.Wz2);
  int r1[12], r2[12];
#pragma unroll
  for(int n = 0; n < 12; ++n) {
    r1[n] = fixed_of(v1.v[n]);
    r2[n] = fixed_of(v2.v[n]);
  }
  const bool same = a[0] == b[0] && a[1] == b[1] && a[2] == b[2];
  if(same) {
#pragma unroll
    for(int n = 0; n < 12; ++n) r1[n] += r2[n];
  }
  emit_fixed(S, ldsJ + (a[0] * S.D(1) + a[1]) * S.D(2) + a[2], r1);
  if(!same) emit_fixed(S, ldsJ + (b[0] * S.D(1) + b[1]) * S.D(2) + b[2], r2);
  return true;
}

// the segment weights of a deposit set up with the charge scaled by a power
// of two, scaled back by its inverse (exact)
__device__ inline Zig unscaled(Zig z, float inv)
{
  z.Fx1 *= inv;
  z.Fy1 *= inv;
  z.Fz1 *= inv;
  z.Fx2 *= inv;
  z.Fy2 *= inv;
  z.Fz2 *= inv;
  return z;
}

// same contributions straight to HBM (global atomics), for nodes outside the window
__device__ inline void deposit_global(const Geo& g, const FieldPtrs& f, const Zig& z, const int toffh[3], uint32_t* err)
{
  zigzag_emit(z, [&](int i, int j, int k, int mask, float jx, float jy, float jz) {
    const int I = toffh[0] + i, J = toffh[1] + j, K = toffh[2] + k;
    if(I < 0 || J < 0 || K < 0 || I >= g.L[0] || J >= g.L[1] || K >= g.L[2]) {
      atomicAdd(err, 1u);
      return;
    }
    const uint32_t q = g.at(I, J, K);
    if(mask & 1) unsafeAtomicAdd(&f.J[0][q], jx);
    if(mask & 2) unsafeAtomicAdd(&f.J[1][q], jy);
    if(mask & 4) unsafeAtomicAdd(&f.J[2][q], jz);
  });
}

// fmodf(x, L) for L > 0; exact shortcut for |x| < 2L (Sterbenz: x -+ L is exact)
__device__ inline float fmod_near(float x, float L)
{
  const float ax = fabsf(x);
  if(ax < L) return x;
  if(ax < 2.0f * L) return x - copysignf(L, x);
  return fmodf(x, L);
}

// Element q of a bin array through a 32-bit byte offset from the bin's
// (uniform) base, so the access is one scalar base plus one lane offset.
template<class T>
__device__ inline T& at32(T* base, uint32_t q)
{
  return *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + q * uint32_t(sizeof(T)));
}

// Fused batched step of one brick: gather + Boris + classification + zigzag
// deposit (in the post-migration frame) + mailbox record + compaction, then
// the brick's current window goes to its staging slot (no global atomics).
constexpr int FWG = 512;  // 8 waves per brick

// Per-brick constants of the tile-leaver path, indexed by axis and by the
// step dv + 1 in {0, 1, 2} towards the destination tile (kept in LDS so they
// occupy no scalar registers in the particle loop).
struct NbrTab {
  float dorig[3][3];  // destination tile lattice origo: tile mins - 3
  int dtoffh[3][3];   // destination tile's world haloed offset
  int off[3][3];      // destination tile-haloed node -> window index (unwrapped image)
  float wlo[3], whi[3], wlen[3];  // periodic box
};

// WALL: reflecting walls between the push and the classification
// (reflector_wall.c++:253-289 inside the step): the walls that lie in the
// brick's tile reflect or park each pushed particle, reflections add their
// correction current to the world buffer Jc, parked particles leave the bin
// without a record and deposit nothing.
// PUSHER: 0 Boris, 1 Higuera-Cary, 2 Faraday (the same device functions as
// the per-tile push, so the fused step moves particles bit-identically).
// a leaver's record in a shared record set, when its brick's mailbox is full
// (rare; out of line for the same reason as the two paths below)
__device__ __attribute__((noinline)) void fused_spill_record(const MigPtrs* mtab, int set, uint32_t slot, float p0,
                                                            float p1, float p2, float v0, float v1, float v2,
                                                            uint64_t id, uint32_t meta)
{
  const float p[3] = {p0, p1, p2}, v[3] = {v0, v1, v2};
  write_record(*mtab, set, slot, p, v, id, meta);
}

// The fused step's two rare paths, for a particle whose stencil leaves the
// brick's LDS windows (a particle kept in a bin it has left): interpolation
// from the world lattice and a deposit with global atomics.  They are kept out
// of line, with scalar arguments and results by value: inlined, their
// registers counted against the particle loop's and pushed its bin pointers
// out of SGPRs into VGPR lanes (read back on every particle load and store).
struct Fields6 {
  float e0, e1, e2, b0, b1, b2;
};

__device__ __attribute__((noinline)) Fields6 fused_gather_global(const float* E0, const float* E1, const float* E2,
                                                                const float* B0, const float* B1, const float* B2,
                                                                int L0, int L1, int L2, int I, int J, int K, float d0,
                                                                float d1, float d2, uint32_t* err)
{
  Geo g {};
  g.L[0] = L0;
  g.L[1] = L1;
  g.L[2] = L2;
  FieldPtrs f {{E0, E1, E2}, {B0, B1, B2}, {nullptr, nullptr, nullptr}};
  const float d[3] = {d0, d1, d2};
  float E[3], B[3];
  gather(g, f, MeanWin {{0, 0, 0}, {0, 0, 0}}, nullptr, I, J, K, d, E, B, err);
  return Fields6 {E[0], E[1], E[2], B[0], B[1], B[2]};
}

__device__ __attribute__((noinline)) void fused_deposit_global(float* J0, float* J1, float* J2, int L0, int L1, int L2,
                                                              Zig z, int t0, int t1, int t2, uint32_t* err)
{
  Geo g {};
  g.L[0] = L0;
  g.L[1] = L1;
  g.L[2] = L2;
  FieldPtrs f {{nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}, {J0, J1, J2}};
  const int toffh[3] = {t0, t1, t2};
  deposit_global(g, f, z, toffh, err);
}

template<int BR, int FAST, int WALL, int PUSHER, int WGS>
__device__ __forceinline__ void fused_brick(const Geo& g, const SpParams& sp, const FieldPtrs& f, const MigPtrs& m,
                                            const MigPtrs* mtab, int b, uint32_t* counters, float* jstage,
                                            int64_t jstage_stride, const WallSet& walls, float* Jc0, float* Jc1,
                                            float* Jc2, double* lds_dyn, int* wcnt, NbrTab& NT)
{
  const Shape<BR> S(g);
  const Brick Bk = brick_of(g, b);
  uint64_t* ldsJ = reinterpret_cast<uint64_t*>(lds_dyn);
  float* ldsM = reinterpret_cast<float*>(lds_dyn + 3 * S.DV());
  for(int q = int(threadIdx.x); q < 3 * S.DV(); q += WGS) ldsJ[q] = 0;
  const bool any = bin_total(sp, b) != 0;
  if(any) stage_means6<BR, WGS>(g, f, S, Bk.c0, ldsM);
  __syncthreads();
  // window index of a particle's cell: tile-haloed cell - 3 - lb*br (means),
  // tile-haloed node - 2 - lb*br (current)
  int moff[3], coff[3];
#pragma unroll
  for(int d = 0; d < 3; ++d) {
    moff[d] = -HALO - Bk.lb[d] * S.b[d];
    coff[d] = moff[d] + 1;
  }
  if(threadIdx.x < 9) {
    const int d = int(threadIdx.x) / 3, k = int(threadIdx.x) % 3;
    int tc = Bk.tc[d] + k - 1;
    tc = tc < 0 ? tc + g.nt[d] : (tc >= g.nt[d] ? tc - g.nt[d] : tc);
    NT.dorig[d][k] = g.tile_min(tc, d) - 3.0f;
    NT.dtoffh[d][k] = tc * g.n[d];
    // the destination's nodes sit next to this tile in the unwrapped frame:
    // the image nearest to this brick
    NT.off[d][k] = (k - 1) * g.n[d] - HALO + 1 - Bk.lb[d] * S.b[d];
    if(k == 0) {
      NT.wlo[d] = g.wmin[d];
      NT.whi[d] = g.wmax[d];
      NT.wlen[d] = g.wmax[d] - g.wmin[d];
    }
  }
  __syncthreads();
  // walls in this brick's tile (uniform over the block)
  const unsigned wmask = WALL ? walls_in_tile(walls, Bk.tmin[0], Bk.tmax[0], g.cfl) : 0u;
  int nrec = 0;  // records in this brick's mailbox (uniform over the block)
  for(int s = 0; any && s < sp.nsp; ++s) {
    // bin start and count are uniform: keep them (and the bin's array bases)
    // scalar so particle addresses are a 32-bit lane offset from a scalar base
    const int64_t start = readfirstlane64(sp.start[s][b]);
    const int count = __builtin_amdgcn_readfirstlane(int(sp.count[s][b]));
    if(count == 0) continue;
    const float qm = sp.qm[s];
    const float charge = sp.charge[s];
    // the deposit's charge in the window's fixed-point unit (exact: a power of two)
    const float qscaled = charge * float(g.jscale), jinv_f = float(g.jinv);
    float* const P0 = sp.pos[s][0] + start;
    float* const P1 = sp.pos[s][1] + start;
    float* const P2 = sp.pos[s][2] + start;
    float* const U0 = sp.vel[s][0] + start;
    float* const U1 = sp.vel[s][1] + start;
    float* const U2 = sp.vel[s][2] + start;
    uint64_t* const ID = sp.id[s] + start;
    // In-order compaction: the survivors of every chunk are written at the
    // running offset wr (wr <= the chunk's first slot, and every lane holds
    // its particle in registers before the scan's barrier), so the bin ends
    // compact with whole-line writes and no separate compaction pass.
    int wr = 0;
    int par = 0;
    // the next chunk's particles are loaded one iteration ahead, so their
    // latency overlaps this chunk's arithmetic
    float npos[3] = {0.0f, 0.0f, 0.0f}, nvel[3] = {0.0f, 0.0f, 0.0f};
    uint64_t nid = 0;
    if(int(threadIdx.x) < count) {
      const uint32_t q0 = threadIdx.x;
      npos[0] = at32(P0, q0);
      npos[1] = at32(P1, q0);
      npos[2] = at32(P2, q0);
      nvel[0] = at32(U0, q0);
      nvel[1] = at32(U1, q0);
      nvel[2] = at32(U2, q0);
      nid = at32(ID, q0);
    }
    for(int base = 0; base < count; base += WGS) {
      const int t = base + int(threadIdx.x);
      bool lv0 = false, lv1 = false, dead = false;
      float pos[3] = {npos[0], npos[1], npos[2]}, vel[3] = {nvel[0], nvel[1], nvel[2]};
      const uint64_t id = nid;
      int dir = 13;
      if(t + WGS < count) {
        const uint32_t qn = uint32_t(t + WGS);
        npos[0] = at32(P0, qn);
        npos[1] = at32(P1, qn);
        npos[2] = at32(P2, qn);
        nvel[0] = at32(U0, qn);
        nvel[1] = at32(U1, qn);
        nvel[2] = at32(U2, qn);
        nid = at32(ID, qn);
      }
      if(t < count) {
        float pr[3], dd[3];
        int ci[3];
#pragma unroll
        for(int d = 0; d < 3; ++d) {
          pr[d] = pos[d] - Bk.origo[d];
          ci[d] = int(uint32_t(pr[d]));
          dd[d] = pr[d] - float(uint32_t(ci[d]));
        }
        float E[3], Bf[3];
        if(!gather6(S, ldsM, ci[0] + moff[0], ci[1] + moff[1], ci[2] + moff[2], dd, E, Bf)) {
          const Fields6 eb = fused_gather_global(f.E[0], f.E[1], f.E[2], f.B[0], f.B[1], f.B[2], g.L[0], g.L[1], g.L[2],
                                                 Bk.toffh[0] + ci[0], Bk.toffh[1] + ci[1], Bk.toffh[2] + ci[2], dd[0],
                                                 dd[1], dd[2], &counters[C_STENCIL]);
          E[0] = eb.e0;
          E[1] = eb.e1;
          E[2] = eb.e2;
          Bf[0] = eb.b0;
          Bf[1] = eb.b1;
          Bf[2] = eb.b2;
        }
        if constexpr(PUSHER == 1) higuera_cary<FAST>(g, qm, E, Bf, pos, vel);
        else if constexpr(PUSHER == 2) faraday<FAST>(g, qm, E, Bf, pos, vel);
        else boris<FAST>(g, qm, E, Bf, pos, vel);
        if(WALL && wmask) {
          float* const Jc[3] = {Jc0, Jc1, Jc2};
          dead = !reflect_particle(g, walls, wmask, g.cfl, Bk.origo, Bk.toffh, Jc, charge, pos, vel,
                                   &counters[C_STENCIL]);
        }
      }
      if(t < count && !dead) {
        // classification against the tile (pack_outgoing) and the brick
        dir = tile_dir(pos, Bk.tmin, Bk.tmax);
        float x2[3];
        int off[3], dtoffh[3];
        if(dir != 13) {
          lv0 = true;
          const int k3[3] = {dir / 9, (dir / 3) % 3, dir % 3};
#pragma unroll
          for(int d = 0; d < 3; ++d) {
            // every coordinate wrapped by the global box, particle.h:523-549
            const float pw = (pos[d] < 0 ? NT.whi[d] : NT.wlo[d]) + fmod_near(pos[d], NT.wlen[d]);
            x2[d] = pw - NT.dorig[d][k3[d]];
            dtoffh[d] = NT.dtoffh[d][k3[d]];
            off[d] = NT.off[d][k3[d]];
          }
        } else {
          int lb[3];
#pragma unroll
          for(int d = 0; d < 3; ++d) {
            x2[d] = pos[d] - Bk.origo[d];
            const int c = min(max(int(floorf(x2[d])) - HALO, 0), g.n[d] - 1);
            lb[d] = BR > 0 ? int(uint32_t(c) / uint32_t(BR)) : S.div(c, g, d);
            off[d] = coff[d];
            dtoffh[d] = Bk.toffh[d];
          }
          lv1 = lb[0] != Bk.lb[0] || lb[1] != Bk.lb[1] || lb[2] != Bk.lb[2];
        }
        // zigzag deposit in the frame of the tile that owns the particle after migration
        const Zig z = zigzag_setup<FAST>(g, x2, vel, qscaled, g.cfl);
        if(!deposit_win(S, ldsJ, z, off))
          fused_deposit_global(f.J[0], f.J[1], f.J[2], g.L[0], g.L[1], g.L[2], unscaled(z, jinv_f), dtoffh[0],
                               dtoffh[1], dtoffh[2], &counters[C_STENCIL]);
      }
      // leavers take consecutive slots of this brick's mailbox (rank from a
      // block scan: no global atomics); past its capacity they spill into
      // the shared record sets
      const bool hole = lv0 || lv1;
      int tot;
      const int r = block_rank1<WGS>(hole, wcnt, par, tot);
      const int64_t local = int64_t(nrec) + r;
      bool spill = false, kept = false;
      uint32_t slot0 = 0, slot1 = 0;
      if(int64_t(nrec) + tot > m.per_brick) {
        // uniform: this chunk fills the mailbox
        spill = hole && local >= m.per_brick;
        slot0 = wave_reserve(spill && lv0, &counters[C_MIG0]);
        slot1 = wave_reserve(spill && lv1, &counters[C_MIG1]);
        if(spill && ((lv0 && int64_t(slot0) >= mtab->cap[0]) || (lv1 && int64_t(slot1) >= mtab->cap[1]))) {
          // no room anywhere: the particle stays in this bin (pushed and
          // deposited); it is classified again next step
          kept = true;
          atomicAdd(&counters[C_MIGLOST], 1u);
        }
      }
      const bool stays = t < count && !dead && (!hole || kept);
      // survivor's slot: wr + lanes before it in the chunk - leavers before
      // it that departed (r counts all leavers, rk the kept ones)
      int kept_tot = 0;
      int rk = 0;
      if(int64_t(nrec) + tot > m.per_brick) rk = block_rank1<WGS>(hole && kept, wcnt, par, kept_tot);
      // parked particles before this lane (they are not leavers)
      int dead_tot = 0, rd = 0;
      if(WALL && wmask) rd = block_rank1<WGS>(dead, wcnt, par, dead_tot);
      const int dst = wr + int(threadIdx.x) - (r - rk) - rd;
      if(t < count && hole && !kept) {
        // records keep the unwrapped position; the append applies the wrap
        const uint32_t meta = lv0 ? meta_pack(s, 0, dir, uint32_t(b)) : meta_pack(s, 1, 13, uint32_t(b));
        if(!spill) write_record(m, 2, uint32_t(int64_t(b) * m.per_brick + local), pos, vel, id, meta);
        else
          fused_spill_record(mtab, lv0 ? 0 : 1, lv0 ? slot0 : slot1, pos[0], pos[1], pos[2], vel[0], vel[1], vel[2], id,
                             meta);
      }
      if(stays) {
        const uint32_t d = uint32_t(dst);
        at32(P0, d) = pos[0];
        at32(P1, d) = pos[1];
        at32(P2, d) = pos[2];
        at32(U0, d) = vel[0];
        at32(U1, d) = vel[1];
        at32(U2, d) = vel[2];
        if(d != uint32_t(t)) at32(ID, d) = id;
      }
      wr += min(WGS, count - base) - (tot - kept_tot) - dead_tot;
      nrec += (int64_t(nrec) + tot <= m.per_brick) ? tot : int(m.per_brick - nrec);
    }
    __syncthreads();
    if(threadIdx.x == 0) sp.count[s][b] = uint32_t(wr);
    __syncthreads();
  }
  if(threadIdx.x == 0) m.mcount[b] = uint32_t(nrec);
  __syncthreads();
  const int V3 = 3 * S.DV();
  float* stage = jstage + int64_t(b) * jstage_stride;
  for(int q = int(threadIdx.x); q < V3; q += WGS) stage[q] = float(double(int64_t(ldsJ[q])) * g.jinv);
}

// WGS: threads per workgroup (a multiple of 64).
// QUEUE 0: one workgroup per brick, the XCD-aware map of blockIdx.  QUEUE 1:
// persistent workgroups (as many as stay resident); the workgroups of XCD x
// take bricks from that XCD's queue, an atomic counter over the same brick
// order the map gives XCD x, until it runs past the last brick.  The queue
// counters are cleared before the launch.
template<int BR, int FAST, int WALL = 0, int PUSHER = 0, int QUEUE = 0, int WGS = FWG>
__global__ void __launch_bounds__(WGS) __attribute__((amdgpu_waves_per_eu(4)))
  k_fused(Geo g, SpParams sp, FieldPtrs f, MigPtrs m, const MigPtrs* mtab, int nbins, int span, uint32_t* counters,
          float* jstage, int64_t jstage_stride, WallSet walls, const rk_reflector_wall* wdev, float* Jc0, float* Jc1,
          float* Jc2)
{
  extern __shared__ double lds_dyn[];
  __shared__ int wcnt[2 * (WGS / 64)];
  __shared__ NbrTab NT;
  // the walls (from the launch arguments, or from device memory when wdev
  // is given: walls that move on the device), staged in LDS
  __shared__ WallSet sw;
  if constexpr(WALL) {
    if(threadIdx.x == 0) sw.n = walls.n;
    for(int i = 0; i < walls.n; ++i)
      if(int(threadIdx.x) == i) sw.w[i] = wdev ? wdev[i] : walls.w[i];
    __syncthreads();
  }
  if constexpr(QUEUE == 0) {
    const int b = xcd_brick(int(blockIdx.x), span);
    if(b >= nbins) return;
    fused_brick<BR, FAST, WALL, PUSHER, WGS>(g, sp, f, m, mtab, b, counters, jstage, jstage_stride, sw, Jc0, Jc1, Jc2,
                                        lds_dyn, wcnt, NT);
  } else {
    __shared__ int next;
    const int x = int(blockIdx.x) % 8;
    for(;;) {
      if(threadIdx.x == 0) next = xcd_brick(x + 8 * int(atomicAdd(&counters[C_QUEUE + x], 1u)), span);
      __syncthreads();
      const int b = next;
      __syncthreads();
      if(b >= nbins) break;  // uniform: every later pop of this queue is past the end too
      fused_brick<BR, FAST, WALL, PUSHER, WGS>(g, sp, f, m, mtab, b, counters, jstage, jstage_stride, sw, Jc0, Jc1,
                                          Jc2, lds_dyn, wcnt, NT);
      __syncthreads();  // the brick's LDS is read out before the next brick fills it
    }
  }
}

__global__ void k_reset_queues(uint32_t* counters)
{
  if(threadIdx.x < 8) counters[C_QUEUE + threadIdx.x] = 0;
}

// J on every lattice node += sum of the staged windows of the bricks that
// cover it, in a fixed order.  The window of global brick gb along one axis
// covers haloed nodes [2 + gb*br, 4 + (gb+1)*br].
__global__ void __launch_bounds__(128)
  k_current_gather(Geo g, const float* __restrict__ stage, int64_t stride, float* J0, float* J1, float* J2)
{
  // Workgroups are dealt to the 8 XCDs round-robin by linear index; XCD x
  // takes the x-th contiguous share of the rows instead, so the rows that
  // share a window's cache lines (neighbouring Jn) go through one L2.
  const int total = int(gridDim.x * gridDim.y);
  const int lin = int(blockIdx.y * gridDim.x + blockIdx.x);
  const int xcd = lin % 8, q8 = total / 8, r8 = total % 8;
  const int row = xcd * q8 + min(xcd, r8) + lin / 8;
  const int Jn = row % int(gridDim.x), In = row / int(gridDim.x);
  const int D[3] = {g.br[0] + 3, g.br[1] + 3, g.br[2] + 3};
  const int V = D[0] * D[1] * D[2];
  const int Nb[3] = {g.nt[0] * g.nb[0], g.nt[1] * g.nb[1], g.nt[2] * g.nb[2]};
  // candidate bricks along x and y for this row
  int bx[4], ox[4], nx = 0, by[4], oy[4], ny = 0;
  for(int gb = min((In - 2) / g.br[0], Nb[0] - 1); In >= 2 && gb >= 0 && nx < 4; --gb) {
    const int o = In - (2 + gb * g.br[0]);
    if(o >= D[0]) break;
    bx[nx] = gb;
    ox[nx++] = o;
  }
  for(int gb = min((Jn - 2) / g.br[1], Nb[1] - 1); Jn >= 2 && gb >= 0 && ny < 4; --gb) {
    const int o = Jn - (2 + gb * g.br[1]);
    if(o >= D[1]) break;
    by[ny] = gb;
    oy[ny++] = o;
  }
  if(nx <= 2 && ny <= 2 && D[2] <= 2 * g.br[2]) {
    // every node lies in at most two windows per axis: the (up to 8) windows
    // of a node are addressed first and their 24 values loaded together,
    // then summed in the order of the general loop below (bit-identical)
    int64_t wxy[2][2];
#pragma unroll
    for(int a = 0; a < 2; ++a)
#pragma unroll
      for(int c = 0; c < 2; ++c) {
        const int ga = a < nx ? bx[a] : 0, gc = c < ny ? by[c] : 0;
        const int tx = ga / g.nb[0], lx = ga - tx * g.nb[0];
        const int ty = gc / g.nb[1], ly = gc - ty * g.nb[1];
        const int64_t bin0 = int64_t(g.tile_lin(tx, ty, 0)) * g.bpt + (lx * g.nb[1] + ly) * g.nb[2];
        wxy[a][c] = bin0 * stride + ((a < nx ? ox[a] : 0) * D[1] + (c < ny ? oy[c] : 0)) * D[2];
      }
    for(int K = int(threadIdx.x); K < g.L[2]; K += int(blockDim.x)) {
      int64_t wz[2];
      bool vz[2];
      const int gz1 = K >= 2 ? min((K - 2) / g.br[2], Nb[2] - 1) : -1;
#pragma unroll
      for(int zi = 0; zi < 2; ++zi) {
        const int gz = gz1 - zi;
        const int oz = K - (2 + gz * g.br[2]);
        vz[zi] = gz >= 0 && oz < D[2];
        const int tz = vz[zi] ? gz / g.nb[2] : 0, lz = vz[zi] ? gz - tz * g.nb[2] : 0;
        wz[zi] = (int64_t(tz) * g.bpt + lz) * stride + oz;
      }
      vz[1] = vz[1] && vz[0];
      float v[2][2][2][3];
#pragma unroll
      for(int zi = 0; zi < 2; ++zi)
#pragma unroll
        for(int a = 0; a < 2; ++a)
#pragma unroll
          for(int c = 0; c < 2; ++c) {
            // unused windows read the first staged value (in bounds) and are
            // not added: the loads stay unconditional, so all 24 are in
            // flight together
            const bool ok = vz[zi] && a < nx && c < ny;
            const float* w = stage + (ok ? wxy[a][c] + wz[zi] : 0);
            v[zi][a][c][0] = w[0];
            v[zi][a][c][1] = w[V];
            v[zi][a][c][2] = w[2 * V];
          }
      float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f;
#pragma unroll
      for(int zi = 0; zi < 2; ++zi)
#pragma unroll
        for(int a = 0; a < 2; ++a)
#pragma unroll
          for(int c = 0; c < 2; ++c) {
            const bool ok = vz[zi] && a < nx && c < ny;
            const float s0 = a0 + v[zi][a][c][0], s1 = a1 + v[zi][a][c][1], s2 = a2 + v[zi][a][c][2];
            a0 = ok ? s0 : a0;
            a1 = ok ? s1 : a1;
            a2 = ok ? s2 : a2;
          }
      const uint32_t q = g.at(In, Jn, K);
      J0[q] = J0[q] + a0;
      J1[q] = J1[q] + a1;
      J2[q] = J2[q] + a2;
    }
    return;
  }
  for(int K = int(threadIdx.x); K < g.L[2]; K += int(blockDim.x)) {
    float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f;
    for(int gz = min((K - 2) / g.br[2], Nb[2] - 1); K >= 2 && gz >= 0; --gz) {
      const int oz = K - (2 + gz * g.br[2]);
      if(oz >= D[2]) break;
      const int tz = gz / g.nb[2], lz = gz - tz * g.nb[2];
      for(int a = 0; a < nx; ++a) {
        const int tx = bx[a] / g.nb[0], lx = bx[a] - tx * g.nb[0];
        for(int c = 0; c < ny; ++c) {
          const int ty = by[c] / g.nb[1], ly = by[c] - ty * g.nb[1];
          const int bin = g.tile_lin(tx, ty, tz) * g.bpt + (lx * g.nb[1] + ly) * g.nb[2] + lz;
          const float* w = stage + int64_t(bin) * stride + (ox[a] * D[1] + oy[c]) * D[2] + oz;
          a0 = a0 + w[0];
          a1 = a1 + w[V];
          a2 = a2 + w[2 * V];
        }
      }
    }
    // J was cleared before the particle step; it holds only the rare
    // contributions that missed their brick's window (global atomics)
    const uint32_t q = g.at(In, Jn, K);
    J0[q] = J0[q] + a0;
    J1[q] = J1[q] + a1;
    J2[q] = J2[q] + a2;
  }
}

// in-bin counting sort by brick-local cell (layout_right), out-of-brick last
__global__ void __launch_bounds__(WG)
  k_sort(Geo g, SpParams sp, int bin0, int nbins, float* stage, int64_t stage_stride)
{
  extern __shared__ uint32_t cnt[];  // nbuck + 1
  const int cells = g.br[0] * g.br[1] * g.br[2];
  const int nbuck = cells + 1;
  float* P = stage + int64_t(blockIdx.x) * stage_stride * 8;
  for(int bi = int(blockIdx.x); bi < nbins; bi += int(gridDim.x)) {
    const int b = bin0 + bi;
    const Brick Bk = brick_of(g, b);
    for(int s = 0; s < sp.nsp; ++s) {
      const int64_t start = sp.start[s][b];
      const int count = int(sp.count[s][b]);
      if(count < 2) continue;
      for(int q = int(threadIdx.x); q <= nbuck; q += WG) cnt[q] = 0;
      __syncthreads();
      auto key = [&](int64_t q) {
        int c[3];
        bool in = true;
#pragma unroll
        for(int d = 0; d < 3; ++d) {
          c[d] = int(floorf(sp.pos[s][d][q] - Bk.origo[d])) - HALO - Bk.lb[d] * g.br[d];
          in = in && c[d] >= 0 && c[d] < g.br[d];
        }
        return in ? (c[0] * g.br[1] + c[1]) * g.br[2] + c[2] : cells;
      };
      for(int t = int(threadIdx.x); t < count; t += WG) atomicAdd(&cnt[key(start + t)], 1u);
      __syncthreads();
      if(threadIdx.x == 0) {
        uint32_t acc = 0;
        for(int q = 0; q < nbuck; ++q) {
          const uint32_t c = cnt[q];
          cnt[q] = acc;
          acc += c;
        }
      }
      __syncthreads();
      for(int t = int(threadIdx.x); t < count; t += WG) {
        const int64_t q = start + t;
        const uint32_t dst = atomicAdd(&cnt[key(q)], 1u);
#pragma unroll
        for(int d = 0; d < 3; ++d) {
          P[int64_t(d) * stage_stride + dst] = sp.pos[s][d][q];
          P[int64_t(3 + d) * stage_stride + dst] = sp.vel[s][d][q];
        }
        reinterpret_cast<uint64_t*>(P + 6 * stage_stride)[dst] = sp.id[s][q];
      }
      __syncthreads();
      for(int t = int(threadIdx.x); t < count; t += WG) {
        const int64_t q = start + t;
#pragma unroll
        for(int d = 0; d < 3; ++d) {
          sp.pos[s][d][q] = P[int64_t(d) * stage_stride + t];
          sp.vel[s][d][q] = P[int64_t(3 + d) * stage_stride + t];
        }
        sp.id[s][q] = reinterpret_cast<const uint64_t*>(P + 6 * stage_stride)[t];
      }
      __syncthreads();
    }
  }
}

// injection: input particles -> record set 1 (kind 1: in-tile, brick by position)
__global__ void __launch_bounds__(WG)
  k_inject_records(MigPtrs m, int64_t n, const float* px, const float* py, const float* pz, const float* vx,
                   const float* vy, const float* vz, const uint64_t* id, uint32_t meta, uint32_t* counters)
{
  for(int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < n; e += int64_t(gridDim.x) * blockDim.x) {
    const int64_t r = e;
    m.pos[1][0][r] = px[e];
    m.pos[1][1][r] = py[e];
    m.pos[1][2][r] = pz[e];
    m.vel[1][0][r] = vx[e];
    m.vel[1][1][r] = vy[e];
    m.vel[1][2][r] = vz[e];
    m.id[1][r] = id[e];
    m.meta[1][r] = meta;
  }
  if(blockIdx.x == 0 && threadIdx.x == 0) counters[C_MIG1] = uint32_t(n);
}

__global__ void __launch_bounds__(WG)
  k_repack(const float* const* __restrict__ src, float* const* __restrict__ dst, const uint64_t* sid, uint64_t* did,
           const int64_t* ostart, const uint32_t* ocap, uint32_t* count, const int64_t* nstart, int nbins)
{
  for(int b = int(blockIdx.x); b < nbins; b += int(gridDim.x)) {
    const uint32_t c = min(count[b], ocap[b]);
    const int64_t so = ostart[b], no = nstart[b];
    for(uint32_t t = threadIdx.x; t < c; t += WG) {
#pragma unroll
      for(int d = 0; d < 6; ++d) dst[d][no + t] = src[d][so + t];
      did[no + t] = sid[so + t];
    }
    __syncthreads();
    if(threadIdx.x == 0) count[b] = c;
  }
}

// The seven arrays of a particle store (pool or staging buffer).
struct PoolPtrs {
  float* f[6];  // x y z ux uy uz
  uint64_t* id;
};

// Segment copies: segment i moves len[i] particles from src at
// src_off[i] - src_base to dst at dst_off[i] - dst_base.  One workgroup per
// segment.  The source and destination ranges of one launch are disjoint.
__global__ void __launch_bounds__(WG)
  k_copy_segments(PoolPtrs src, PoolPtrs dst, const int64_t* __restrict__ src_off, int64_t src_base,
                  const int64_t* __restrict__ dst_off, int64_t dst_base, const uint32_t* __restrict__ len, int nseg)
{
  for(int i = int(blockIdx.x); i < nseg; i += int(gridDim.x)) {
    const int64_t so = src_off[i] - src_base, d0 = dst_off[i] - dst_base;
    const uint32_t n = len[i];
    for(uint32_t t = threadIdx.x; t < n; t += WG) {
#pragma unroll
      for(int d = 0; d < 6; ++d) dst.f[d][d0 + t] = src.f[d][so + t];
      dst.id[d0 + t] = src.id[so + t];
    }
  }
}

// Moves the listed bins of one species to new places (disjoint from every
// bin: the pool's spare arena).  A full bin keeps min(count, cap) particles;
// its pending records are re-appended afterwards.
__global__ void __launch_bounds__(WG)
  k_move_bins(PoolPtrs p, int64_t* bin_start, uint32_t* bin_cap, uint32_t* bin_count, const int32_t* __restrict__ bins,
              const int64_t* __restrict__ new_start, const uint32_t* __restrict__ new_cap, int n)
{
  for(int i = int(blockIdx.x); i < n; i += int(gridDim.x)) {
    const int b = bins[i];
    const int64_t so = bin_start[b], no = new_start[i];
    const uint32_t c = min(bin_count[b], bin_cap[b]);
    for(uint32_t t = threadIdx.x; t < c; t += WG) {
#pragma unroll
      for(int d = 0; d < 6; ++d) p.f[d][no + t] = p.f[d][so + t];
      p.id[no + t] = p.id[so + t];
    }
    __syncthreads();
    if(threadIdx.x == 0) {
      bin_start[b] = no;
      bin_cap[b] = new_cap[i];
      bin_count[b] = c;
    }
  }
}

__global__ void __launch_bounds__(WG)
  k_gather_tile(SpParams sp, int s, int bin0, int nbins, const int64_t* offs, float* ox, float* oy, float* oz,
                float* ux, float* uy, float* uz, uint64_t* oid)
{
  const int bi = int(blockIdx.x);
  if(bi >= nbins) return;
  const int b = bin0 + bi;
  const int64_t start = sp.start[s][b];
  const int count = int(sp.count[s][b]);
  const int64_t o = offs[bi];
  for(int t = int(threadIdx.x); t < count; t += WG) {
    ox[o + t] = sp.pos[s][0][start + t];
    oy[o + t] = sp.pos[s][1][start + t];
    oz[o + t] = sp.pos[s][2][start + t];
    ux[o + t] = sp.vel[s][0][start + t];
    uy[o + t] = sp.vel[s][1][start + t];
    uz[o + t] = sp.vel[s][2][start + t];
    oid[o + t] = sp.id[s][start + t];
  }
}

// sum over alive particles of (sqrt(1 + u.u) - 1) in float, accumulated in FP64
__global__ void __launch_bounds__(WG) k_kinetic(SpParams sp, int s, int bin0, int nbins, double* partial)
{
  double acc = 0.0;
  for(int bi = int(blockIdx.x); bi < nbins; bi += int(gridDim.x)) {
    const int b = bin0 + bi;
    const int64_t start = sp.start[s][b];
    const int count = int(sp.count[s][b]);
    for(int t = int(threadIdx.x); t < count; t += WG) {
      const float u[3] = {sp.vel[s][0][start + t], sp.vel[s][1][start + t], sp.vel[s][2][start + t]};
      const float e = sqrtf(1.0f + dot3(u, u)) - 1.0f;
      acc += double(e);
    }
  }
  __shared__ double red[WG];
  red[threadIdx.x] = acc;
  __syncthreads();
  for(int o = WG / 2; o > 0; o >>= 1) {
    if(int(threadIdx.x) < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if(threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

}  // namespace

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
namespace {
void bin_range(const rk_world* w, int tile, int& b0, int& nb)
{
  if(tile < 0) {
    b0 = 0;
    nb = w->g.nbricks;
  } else {
    b0 = tile * w->g.bpt;
    nb = w->g.bpt;
  }
}

// span of xcd_brick() for a launch over bricks [b0, b0 + n): groups of whole
// tiles, an eighth of a tile column (tiles sharing an x index) per group, when
// the launch covers the world; else one contiguous eighth per XCD
int xcd_span(const Geo& g, int n)
{
  const int eighth = (n + 7) / 8;
  if(n != g.nbricks) return std::max(eighth, 1);
  const int tiles = std::max(1, g.nt[1] * g.nt[2] / 8);
  return std::max(1, std::min(tiles * g.bpt, eighth));
}

// grid for xcd_brick(): every group of every XCD, so the map covers [0, n)
unsigned xcd_grid(int n, int span)
{
  const int groups = (n + span - 1) / span;
  return unsigned((groups + 7) / 8 * 8 * span);
}

// Compare the fast forms of x/cfl, cfl/sqrt(X), 2/t, 1/t and 1/sqrt(t) with the
// IEEE operations over every float of their fast ranges; mismatches summed.
__global__ void __launch_bounds__(256) k_verify_fast(float cfl, float rcp, uint32_t* mismatches)
{
  const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
  uint32_t bad = 0, bad2 = 0;
  for(uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < (uint64_t(1) << 32); i += stride) {
    const float x = __uint_as_float(uint32_t(i));
    if(fast_div_range(x)) {
      const float q = x * rcp;
      const float r = fmaf(-q, cfl, x);
      bad += __float_as_uint(fmaf(r, rcp, q)) != __float_as_uint(x / cfl) ? 1u : 0u;
    }
    if(x >= cfl * cfl && x <= 0x1p60f)
      bad2 += __float_as_uint(div_near(cfl, sqrt_near(x))) != __float_as_uint(cfl / sqrtf(x)) ? 1u : 0u;
    if(x >= 1.0f && x <= 0x1p60f) {
      bad2 += __float_as_uint(div_near(2.0f, x)) != __float_as_uint(2.0f / x) ? 1u : 0u;
      bad2 += __float_as_uint(div_near(1.0f, x)) != __float_as_uint(1.0f / x) ? 1u : 0u;
      bad2 += __float_as_uint(div_near(1.0f, sqrt_near(x))) != __float_as_uint(1.0f / sqrtf(x)) ? 1u : 0u;
    }
  }
  if(__ballot(bad != 0 || bad2 != 0) == 0) return;
  for(int o = 32; o > 0; o >>= 1) {
    bad += __shfl_down(bad, o);
    bad2 += __shfl_down(bad2, o);
  }
  if(lane_id() == 0) {
    atomicAdd(&mismatches[0], bad);
    atomicAdd(&mismatches[1], bad2);
  }
}

size_t means_lds(const Geo& g) { return sizeof(float) * 6 * size_t(g.br[0] + 2) * (g.br[1] + 2) * (g.br[2] + 2); }
size_t cur_lds(const Geo& g) { return sizeof(double) * 3 * size_t(g.br[0] + 3) * (g.br[1] + 3) * (g.br[2] + 3); }
}  // namespace

int verify_fast_division(float cfl, float rcp, uint64_t mismatches[2])
{
  uint32_t* d = nullptr;
  RK_HIP(hipMalloc(reinterpret_cast<void**>(&d), 2 * sizeof(uint32_t)));
  hipError_t e = hipMemset(d, 0, 2 * sizeof(uint32_t));
  if(e == hipSuccess) {
    hipLaunchKernelGGL(k_verify_fast, dim3(8192), dim3(256), 0, 0, cfl, rcp, d);
    e = hipGetLastError();
  }
  uint32_t h[2] = {0, 0};
  if(e == hipSuccess) e = hipMemcpy(h, d, 2 * sizeof(uint32_t), hipMemcpyDeviceToHost);
  (void)hipFree(d);
  RK_HIP(e);
  mismatches[0] = h[0];
  mismatches[1] = h[1];
  return RK_OK;
}

int launch_push(rk_world* w, int tile, hipStream_t s)
{
  int b0, nb;
  bin_range(w, tile, b0, nb);
  auto* kern = w->g.pusher == 1 ? k_push<1> : (w->g.pusher == 2 ? k_push<2> : k_push<0>);
  const int span = xcd_span(w->g, nb);
  hipLaunchKernelGGL(kern, dim3(xcd_grid(nb, span)), dim3(WG), means_lds(w->g), s, w->g, make_sp(w), make_fields(w),
                     b0, nb, span, w->counters);
  RK_HIP(hipGetLastError());
  return RK_OK;
}

int launch_deposit(rk_world* w, int tile, hipStream_t s)
{
  int b0, nb;
  bin_range(w, tile, b0, nb);
  const int span = xcd_span(w->g, nb);
  hipLaunchKernelGGL(k_deposit, dim3(xcd_grid(nb, span)), dim3(WG), cur_lds(w->g), s, w->g, make_sp(w),
                     make_fields(w), b0, nb, span, w->counters);
  RK_HIP(hipGetLastError());
  return RK_OK;
}

int launch_pack(rk_world* w, int tile, hipStream_t s)
{
  int b0, nb;
  bin_range(w, tile, b0, nb);
  hipLaunchKernelGGL(k_pack, dim3(unsigned(nb)), dim3(WG), 0, s, w->g, make_sp(w), make_mig(w), b0, nb, w->counters);
  RK_HIP(hipGetLastError());
  return RK_OK;
}

int launch_append(rk_world* w, int set, hipStream_t s)
{
  const uint32_t* none = nullptr;
  hipLaunchKernelGGL(k_append, dim3(1024), dim3(WG), 0, s, w->g, make_sp(w), make_mig(w), set, w->counters,
                     w->ovf_list, w->ovf_cap, none, 0);
  RK_HIP(hipGetLastError());
  return RK_OK;
}

int launch_append_mailbox(rk_world* w, hipStream_t s)
{
  // one block per brick (measured faster than a capped grid)
  const int span = xcd_span(w->g, w->g.nbricks);
  hipLaunchKernelGGL(k_append_mailbox, dim3(xcd_grid(w->g.nbricks, span)), dim3(WG), 0, s, w->g, make_sp(w),
                     make_mig(w), w->g.nbricks, w->counters, w->ovf_list, w->ovf_cap, span);
  RK_HIP(hipGetLastError());
  return RK_OK;
}

int launch_append_records(rk_world* w, hipStream_t s)
{
  // the sets hold only what spilled past full mailboxes (usually nothing):
  // a small grid, since every workgroup signs off with a fence and an atomic
  hipLaunchKernelGGL(k_append_records, dim3(64, 2), dim3(WG), 0, s, w->g, make_sp(w), make_mig(w), w->counters,
                     w->ovf_list, w->ovf_cap);
  RK_HIP(hipGetLastError());
  return RK_OK;
}

int launch_reset_counter(rk_world* w, int which, hipStream_t s)
{
  hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, s, w->counters, which);
  RK_HIP(hipGetLastError());
  return RK_OK;
}

int launch_fused(rk_world* w, hipStream_t s, const WallSet* walls, const rk_reflector_wall* wdev)
{
  const Geo& g = w->g;
  const size_t dyn = means_lds(g) + cur_lds(g);
  const int span = xcd_span(g, g.nbricks);
  const dim3 grid(xcd_grid(g.nbricks, span));
  WallSet ws {};
  if(walls) ws = *walls;
  auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, grid, dim3(FWG), dyn, s, g, make_sp(w), make_fields(w), make_mig(w),
                       static_cast<const MigPtrs*>(w->mig_table), g.nbricks, span, w->counters, w->jstage,
                       w->jstage_stride, ws, wdev, w->corrJ[0], w->corrJ[1], w->corrJ[2]);
  };
  const bool cube8 = g.br[0] == 8 && g.br[1] == 8 && g.br[2] == 8;
  const bool cube4 = g.br[0] == 4 && g.br[1] == 4 && g.br[2] == 4;
  const int lvl = g.fast_div ? (g.fast_sqrt ? 2 : 1) : -1;
  if(ws.n > 0 && !w->corrJ[0]) {
    set_error("wall correction buffer not set");
    return RK_EINVAL;
  }
  if(g.pusher != 0) {
    // Higuera-Cary or Faraday, 8^3 bricks or the generic shape; on 8^3 bricks
    // with both fast forms verified, the pusher, the deposit and the
    // classification take the verified sqrt/rcp forms (the IEEE results),
    // otherwise the IEEE operations
    const bool hc = g.pusher == 1;
    if(ws.n > 0) {
      if(cube8 && lvl == 2) hc ? go(k_fused<8, 2, 1, 1>) : go(k_fused<8, 2, 1, 2>);
      else if(cube8) hc ? go(k_fused<8, -1, 1, 1>) : go(k_fused<8, -1, 1, 2>);
      else hc ? go(k_fused<0, -1, 1, 1>) : go(k_fused<0, -1, 1, 2>);
    } else if(cube8 && lvl == 2) hc ? go(k_fused<8, 2, 0, 1>) : go(k_fused<8, 2, 0, 2>);
    else if(cube8) hc ? go(k_fused<8, -1, 0, 1>) : go(k_fused<8, -1, 0, 2>);
    else hc ? go(k_fused<0, -1, 0, 1>) : go(k_fused<0, -1, 0, 2>);
  } else if(w->fused_wg == 256 && cube8 && !w->fused_queue) {
    auto go256 = [&](auto kern) {
      hipLaunchKernelGGL(kern, grid, dim3(256), dyn, s, g, make_sp(w), make_fields(w), make_mig(w),
                         static_cast<const MigPtrs*>(w->mig_table), g.nbricks, span, w->counters, w->jstage,
                         w->jstage_stride, ws, wdev, w->corrJ[0], w->corrJ[1], w->corrJ[2]);
    };
    if(ws.n > 0)
      lvl == 2 ? go256(k_fused<8, 2, 1, 0, 0, 256>)
               : (lvl == 1 ? go256(k_fused<8, 1, 1, 0, 0, 256>) : go256(k_fused<8, -1, 1, 0, 0, 256>));
    else
      lvl == 2 ? go256(k_fused<8, 2, 0, 0, 0, 256>)
               : (lvl == 1 ? go256(k_fused<8, 1, 0, 0, 0, 256>) : go256(k_fused<8, -1, 0, 0, 0, 256>));
  } else if(w->fused_queue && cube8) {
    // persistent queue: as many workgroups as stay resident, a multiple of 8
    auto goq = [&](auto kern) -> int {
      int per_cu = 0, cus = 0, dev = 0;
      RK_HIP(hipGetDevice(&dev));
      RK_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
      RK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, FWG, dyn));
      const int blocks = std::max(8, (std::max(per_cu, 1) * cus + 7) / 8 * 8);
      hipLaunchKernelGGL(k_reset_queues, dim3(1), dim3(64), 0, s, w->counters);
      hipLaunchKernelGGL(kern, dim3(unsigned(blocks)), dim3(FWG), dyn, s, g, make_sp(w), make_fields(w), make_mig(w),
                         static_cast<const MigPtrs*>(w->mig_table), g.nbricks, span, w->counters, w->jstage,
                         w->jstage_stride, ws, wdev, w->corrJ[0], w->corrJ[1], w->corrJ[2]);
      return RK_OK;
    };
    int e = RK_OK;
    if(ws.n > 0)
      e = lvl == 2 ? goq(k_fused<8, 2, 1, 0, 1>) : (lvl == 1 ? goq(k_fused<8, 1, 1, 0, 1>) : goq(k_fused<8, -1, 1, 0, 1>));
    else
      e = lvl == 2 ? goq(k_fused<8, 2, 0, 0, 1>) : (lvl == 1 ? goq(k_fused<8, 1, 0, 0, 1>) : goq(k_fused<8, -1, 0, 0, 1>));
    if(e) return e;
  } else if(ws.n > 0) {
    if(cube8) lvl == 2 ? go(k_fused<8, 2, 1>) : (lvl == 1 ? go(k_fused<8, 1, 1>) : go(k_fused<8, -1, 1>));
    else if(cube4) lvl == 2 ? go(k_fused<4, 2, 1>) : (lvl == 1 ? go(k_fused<4, 1, 1>) : go(k_fused<4, -1, 1>));
    else go(k_fused<0, 0, 1>);
  } else if(cube8) lvl == 2 ? go(k_fused<8, 2>) : (lvl == 1 ? go(k_fused<8, 1>) : go(k_fused<8, -1>));
  else if(cube4) lvl == 2 ? go(k_fused<4, 2>) : (lvl == 1 ? go(k_fused<4, 1>) : go(k_fused<4, -1>));
  else go(k_fused<0, 0>);
  RK_HIP(hipGetLastError());
  return RK_OK;
}

int launch_current_gather(rk_world* w, hipStream_t s)
{
  hipLaunchKernelGGL(k_current_gather, dim3(unsigned(w->g.L[1]), unsigned(w->g.L[0])), dim3(128), 0, s, w->g,
                     w->jstage, w->jstage_stride, w->J[0], w->J[1], w->J[2]);
  RK_HIP(hipGetLastError());
  return RK_OK;
}

int launch_sort(rk_world* w, int tile, hipStream_t s)
{
  int b0, nb;
  bin_range(w, tile, b0, nb);
  uint32_t maxcap = 0;
  {
    std::vector<uint32_t> caps(size_t(w->g.nbricks));
    for(int sp = 0; sp < w->d.n_species; ++sp) {
      RK_HIP(hipMemcpyAsync(caps.data(), w->sp[sp].bin_cap, sizeof(uint32_t) * caps.size(), hipMemcpyDeviceToHost, s));
      RK_HIP(hipStreamSynchronize(s));
      for(int b = b0; b < b0 + nb; ++b) maxcap = std::max(maxcap, caps[size_t(b)]);
    }
  }
  if(maxcap == 0) return RK_OK;
  const int grid = std::min(nb, 512);
  const int64_t stride = (int64_t(maxcap) + 63) / 64 * 64;
  float* stage = nullptr;
  RK_HIP(hipMallocAsync(reinterpret_cast<void**>(&stage), sizeof(float) * 8 * size_t(stride) * size_t(grid), s));
  const int cells = w->g.br[0] * w->g.br[1] * w->g.br[2];
  hipLaunchKernelGGL(k_sort, dim3(unsigned(grid)), dim3(WG), sizeof(uint32_t) * size_t(cells + 2), s, w->g, make_sp(w),
                     b0, nb, stage, stride);
  RK_HIP(hipGetLastError());
  RK_HIP(hipFreeAsync(stage, s));
  return RK_OK;
}

int launch_inject(rk_world* w, int tile, int s, int64_t n, const float* const pos[3], const float* const vel[3],
                  const uint64_t* id, hipStream_t st)
{
  if(n <= 0) return RK_OK;
  if(n > w->mig[1].capacity) return (set_error("inject: record set 1 smaller than the batch"), RK_EOVERFLOW);
  const unsigned blocks = unsigned(std::min<int64_t>((n + WG - 1) / WG, 4096));
  hipLaunchKernelGGL(k_inject_records, dim3(blocks), dim3(WG), 0, st, make_mig(w), n, pos[0], pos[1], pos[2], vel[0],
                     vel[1], vel[2], id, meta_pack(s, 1, 13, uint32_t(tile * w->g.bpt)), w->counters);
  RK_HIP(hipGetLastError());
  int rc = launch_append(w, 1, st);
  if(rc) return rc;
  return launch_reset_counter(w, C_MIG1, st);
}

int launch_retry(rk_world* w, hipStream_t st)
{
  // move the overflow list aside and re-append its records
  uint32_t n = 0;
  RK_HIP(hipMemcpyAsync(&n, &w->counters[C_OVF], sizeof(uint32_t), hipMemcpyDeviceToHost, st));
  RK_HIP(hipStreamSynchronize(st));
  if(n == 0) return RK_OK;
  uint32_t* list = nullptr;
  RK_HIP(hipMallocAsync(reinterpret_cast<void**>(&list), sizeof(uint32_t) * n, st));
  RK_HIP(hipMemcpyAsync(list, w->ovf_list, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, st));
  int rc = launch_reset_counter(w, C_OVF, st);
  if(rc) return rc;
  hipLaunchKernelGGL(k_append, dim3(1024), dim3(WG), 0, st, w->g, make_sp(w), make_mig(w), 0, w->counters, w->ovf_list,
                     w->ovf_cap, list, int(n));
  RK_HIP(hipGetLastError());
  RK_HIP(hipFreeAsync(list, st));
  return RK_OK;
}

int launch_repack(rk_world* w, int s, const Species& nsp, hipStream_t st)
{
  const Species& o = w->sp[s];
  float* src[6] = {o.pos[0], o.pos[1], o.pos[2], o.vel[0], o.vel[1], o.vel[2]};
  float* dst[6] = {nsp.pos[0], nsp.pos[1], nsp.pos[2], nsp.vel[0], nsp.vel[1], nsp.vel[2]};
  float** dsrc = nullptr;
  float** ddst = nullptr;
  RK_HIP(hipMallocAsync(reinterpret_cast<void**>(&dsrc), sizeof(src), st));
  RK_HIP(hipMallocAsync(reinterpret_cast<void**>(&ddst), sizeof(dst), st));
  RK_HIP(hipMemcpyAsync(dsrc, src, sizeof(src), hipMemcpyHostToDevice, st));
  RK_HIP(hipMemcpyAsync(ddst, dst, sizeof(dst), hipMemcpyHostToDevice, st));
  const int grid = std::min(w->g.nbricks, 4096);
  hipLaunchKernelGGL(k_repack, dim3(unsigned(grid)), dim3(WG), 0, st, dsrc, ddst, o.id, nsp.id, o.bin_start, o.bin_cap,
                     o.bin_count, nsp.bin_start, w->g.nbricks);
  RK_HIP(hipGetLastError());
  RK_HIP(hipStreamSynchronize(st));
  RK_HIP(hipFreeAsync(dsrc, st));
  RK_HIP(hipFreeAsync(ddst, st));
  return RK_OK;
}

namespace {

PoolPtrs pool_ptrs(const Species& p)
{
  PoolPtrs q {};
  for(int d = 0; d < 3; ++d) {
    q.f[d] = p.pos[d];
    q.f[3 + d] = p.vel[d];
  }
  q.id = p.id;
  return q;
}

}  // namespace

int launch_move_bins(rk_world* w, int s, int n, const int32_t* bins, const int64_t* new_start, const uint32_t* new_cap,
                     hipStream_t st)
{
  if(n <= 0) return RK_OK;
  const Species& p = w->sp[s];
  hipLaunchKernelGGL(k_move_bins, dim3(unsigned(std::min(n, 8192))), dim3(WG), 0, st, pool_ptrs(p), p.bin_start,
                     p.bin_cap, p.bin_count, bins, new_start, new_cap, n);
  RK_HIP(hipGetLastError());
  return RK_OK;
}

// In-place relayout of species s (synchronous).  Bins are taken in the order
// of their current starts s_i; bin i goes to f_i = sum of the new capacities
// before it, with its n_i stored particles.  One pass moves each bin once:
// first the bins with f_i < s_i in ascending order, then those with f_i > s_i
// in descending order (a bin with f_i = s_i stays).  No bin lands on the old
// place of a bin not yet moved: a bin moving towards the front ends below
// s_i + n_i <= s_j for every later bin j, and lies above s_j + n_j < f_j + n_j
// <= f_i for every earlier bin j that moves towards the end; a bin moving
// towards the end starts above s_i >= s_j + n_j for every earlier bin j; the
// new places are disjoint.  Batches of whole bins go through the staging
// buffer (gathered, then scattered), so a bin's old and new places may
// overlap.
int relayout_species(rk_world* w, int s, const uint32_t* new_cap, void* tmp, int64_t tmp_bytes, int64_t* new_start,
                     hipStream_t st)
{
  Species& p = w->sp[s];
  const int nb = w->g.nbricks;
  const size_t nbs = static_cast<size_t>(nb);
  std::vector<int64_t> start(nbs, 0);
  std::vector<uint32_t> cap(nbs, 0), cnt(nbs, 0);
  RK_HIP(hipMemcpyAsync(start.data(), p.bin_start, sizeof(int64_t) * size_t(nb), hipMemcpyDeviceToHost, st));
  RK_HIP(hipMemcpyAsync(cap.data(), p.bin_cap, sizeof(uint32_t) * size_t(nb), hipMemcpyDeviceToHost, st));
  RK_HIP(hipMemcpyAsync(cnt.data(), p.bin_count, sizeof(uint32_t) * size_t(nb), hipMemcpyDeviceToHost, st));
  RK_HIP(hipStreamSynchronize(st));
  std::vector<int32_t> order(nbs);
  for(int b = 0; b < nb; ++b) order[size_t(b)] = b;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return start[size_t(a)] < start[size_t(b)]; });
  // sorted tables: old start, final start, stored count
  std::vector<int64_t> so(nbs), fo(nbs);
  std::vector<uint32_t> n(nbs);
  int64_t f = 0, end = 0;
  uint32_t nmax = 0;
  for(int i = 0; i < nb; ++i) {
    const int b = order[size_t(i)];
    const uint32_t stored = std::min(cnt[size_t(b)], cap[size_t(b)]);
    if(cap[size_t(b)] > 0 && start[size_t(b)] < end) {
      set_error("relayout: bins of species " + std::to_string(s) + " overlap");
      return RK_ESTATE;
    }
    if(new_cap[b] < stored) {
      set_error("relayout: new capacity of bin " + std::to_string(b) + " below its " + std::to_string(stored) +
                " stored particles");
      return RK_EINVAL;
    }
    end = std::max(end, start[size_t(b)] + int64_t(cap[size_t(b)]));
    so[size_t(i)] = start[size_t(b)];
    fo[size_t(i)] = f;
    n[size_t(i)] = stored;
    f += new_cap[b];
    nmax = std::max(nmax, stored);
  }
  if(end > p.capacity || f > p.capacity) {
    set_error("relayout: species " + std::to_string(s) + " needs " + std::to_string(f) + " slots, pool holds " +
              std::to_string(p.capacity));
    return RK_EOVERFLOW;
  }
  // the moves in processing order: old start, new start, staging offset, count
  std::vector<int64_t> mso, mfo, mst;
  std::vector<uint32_t> mn;
  int64_t st_off = 0;
  auto add_move = [&](int i) {
    if(n[size_t(i)] == 0 || so[size_t(i)] == fo[size_t(i)]) return;
    mso.push_back(so[size_t(i)]);
    mfo.push_back(fo[size_t(i)]);
    mst.push_back(st_off);
    mn.push_back(n[size_t(i)]);
    st_off += n[size_t(i)];
  };
  for(int i = 0; i < nb; ++i)
    if(fo[size_t(i)] < so[size_t(i)]) add_move(i);
  for(int i = nb - 1; i >= 0; --i)
    if(fo[size_t(i)] > so[size_t(i)]) add_move(i);
  const int nm = int(mn.size());
  // staging: the three offset tables, the count table, then [7][W] particles
  const int64_t table = int64_t(nb) * (3 * sizeof(int64_t) + sizeof(uint32_t));
  const int64_t tab_al = (table + 255) / 256 * 256;
  const int64_t W = tmp_bytes > tab_al ? (tmp_bytes - tab_al) / 32 : 0;
  if(!tmp || W < int64_t(nmax) || W < 1) {
    set_error("relayout: staging buffer of " + std::to_string(tmp_bytes) + " bytes is too small");
    return RK_EINVAL;
  }
  char* base = static_cast<char*>(tmp);
  int64_t* d_so = reinterpret_cast<int64_t*>(base);
  int64_t* d_fo = d_so + nb;
  int64_t* d_st = d_fo + nb;
  uint32_t* d_n = reinterpret_cast<uint32_t*>(d_st + nb);
  if(nm > 0) {
    RK_HIP(hipMemcpyAsync(d_so, mso.data(), sizeof(int64_t) * size_t(nm), hipMemcpyHostToDevice, st));
    RK_HIP(hipMemcpyAsync(d_fo, mfo.data(), sizeof(int64_t) * size_t(nm), hipMemcpyHostToDevice, st));
    RK_HIP(hipMemcpyAsync(d_st, mst.data(), sizeof(int64_t) * size_t(nm), hipMemcpyHostToDevice, st));
    RK_HIP(hipMemcpyAsync(d_n, mn.data(), sizeof(uint32_t) * size_t(nm), hipMemcpyHostToDevice, st));
  }
  PoolPtrs stage {};
  float* fbase = reinterpret_cast<float*>(base + tab_al);
  for(int d = 0; d < 6; ++d) stage.f[d] = fbase + int64_t(d) * W;
  stage.id = reinterpret_cast<uint64_t*>(fbase + 6 * W);
  const PoolPtrs pool = pool_ptrs(p);
  // batches of consecutive moves whose particles fit the staging buffer
  int k0 = 0;
  while(k0 < nm) {
    int k1 = k0;
    int64_t T = 0;
    while(k1 < nm && T + mn[size_t(k1)] <= W) T += mn[size_t(k1++)];
    const int ns = k1 - k0;
    const unsigned grid = unsigned(std::min(ns, 8192));
    hipLaunchKernelGGL(k_copy_segments, dim3(grid), dim3(WG), 0, st, pool, stage, d_so + k0, int64_t(0), d_st + k0,
                       mst[size_t(k0)], d_n + k0, ns);
    RK_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_copy_segments, dim3(grid), dim3(WG), 0, st, stage, pool, d_st + k0, mst[size_t(k0)],
                       d_fo + k0, int64_t(0), d_n + k0, ns);
    RK_HIP(hipGetLastError());
    k0 = k1;
  }
  for(int i = 0; i < nb; ++i) {
    const int b = order[size_t(i)];
    new_start[b] = fo[size_t(i)];
    cnt[size_t(b)] = n[size_t(i)];
  }
  RK_HIP(hipMemcpyAsync(p.bin_start, new_start, sizeof(int64_t) * size_t(nb), hipMemcpyHostToDevice, st));
  RK_HIP(hipMemcpyAsync(p.bin_cap, new_cap, sizeof(uint32_t) * size_t(nb), hipMemcpyHostToDevice, st));
  RK_HIP(hipMemcpyAsync(p.bin_count, cnt.data(), sizeof(uint32_t) * size_t(nb), hipMemcpyHostToDevice, st));
  RK_HIP(hipStreamSynchronize(st));
  return RK_OK;
}

int launch_gather_tile(rk_world* w, int tile, int s, float* const pos[3], float* const vel[3], uint64_t* id,
                       const int64_t* offs, hipStream_t st)
{
  const int b0 = tile * w->g.bpt, nb = w->g.bpt;
  hipLaunchKernelGGL(k_gather_tile, dim3(unsigned(nb)), dim3(WG), 0, st, make_sp(w), s, b0, nb, offs, pos[0], pos[1],
                     pos[2], vel[0], vel[1], vel[2], id);
  RK_HIP(hipGetLastError());
  return RK_OK;
}

int kinetic_partial_count(rk_world* w, int tile)
{
  int b0, nb;
  bin_range(w, tile, b0, nb);
  return std::min(nb, 1024);
}

int launch_kinetic_partials(rk_world* w, int tile, int s, double* partial, hipStream_t st)
{
  int b0, nb;
  bin_range(w, tile, b0, nb);
  const int grid = std::min(nb, 1024);
  hipLaunchKernelGGL(k_kinetic, dim3(unsigned(grid)), dim3(WG), 0, st, make_sp(w), s, b0, nb, partial);
  RK_HIP(hipGetLastError());
  return RK_OK;
}

int kinetic_energies(rk_world* w, int tile, double* out, hipStream_t st)
{
  const int grid = kinetic_partial_count(w, tile);
  if(w->red_cap < grid) return (set_error("reduction scratch too small"), RK_EINVAL);
  std::vector<double> h(static_cast<size_t>(grid));
  for(int s = 0; s < w->d.n_species; ++s) {
    if(int rc = launch_kinetic_partials(w, tile, s, w->red, st)) return rc;
    RK_HIP(hipMemcpyAsync(h.data(), w->red, sizeof(double) * h.size(), hipMemcpyDeviceToHost, st));
    RK_HIP(hipStreamSynchronize(st));
    double acc = 0;
    for(double v : h) acc += v;
    out[s] = acc;
  }
  return RK_OK;
}

}  // namespace rk

namespace rk {
// device copy of the record-set pointers (read by rarely taken kernel paths)
int upload_mig_table(rk_world* w)
{
  if(!w->mig_table) RK_HIP(hipMalloc(&w->mig_table, sizeof(MigPtrs)));
  const MigPtrs m = make_mig(w);
  RK_HIP(hipDeviceSynchronize());
  RK_HIP(hipMemcpy(w->mig_table, &m, sizeof(MigPtrs), hipMemcpyHostToDevice));
  return RK_OK;
}
}  // namespace rk
