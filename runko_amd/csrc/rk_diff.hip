// Differentiable periodic box: one PIC step and its adjoint (reverse mode).
//
// Forward.  The discrete step of one lap of projects/pic-turbulence/pic.py:187-221
// on a single periodic box, with the expression order of the reference
// kernels (and of the tile path), compiled with -ffp-contract=off:
//   B1 = B + (cfl/2) curlE(E)                       emf/yee_lattice_fdtd2.c++:19-58
//   per particle: E, B at x (node means + lerp3D)   emf/yee_lattice_interpolate_linear_1st.h:29-138
//                 Boris: u', x' = x + cfl u'/gamma'  pic/particle_boris.h:15-62
//                 x'' = periodic wrap of x'          pic/tile_communication.c++:98-117
//                 zigzag deposit of (x'', u') -> J   pic/particle_current_zigzag_1st.c++:225-339
//   Jf = binomial2^P(J) (periodic, 27 taps)         emf/yee_lattice_current_filter_binomial2.c++:23-76
//   B2 = B1 + (cfl/2) curlE(E)
//   E' = E + cfl curlB(B2) - Jf                     emf/yee_lattice_fdtd2.c++:68-111, yee_lattice.c++:170-179
// On a periodic box the tile halos, J_exchange and the copies between filter
// passes reduce to periodic indexing, so a step gives the tile path's values
// (J up to the order of its float sums).
//
// Adjoint.  Vector-Jacobian products of each stage, in reverse order.  The
// discrete curls are negative transposes of each other (curlE^T = -curlB),
// the periodic binomial stencil is symmetric (F^T = F), so the field part
// reuses the forward stencils with negated time steps.  The particle part is
// one kernel per particle that recomputes its forward step and then applies,
// by hand-derived chain rule:
//   the deposit adjoint  (gather of lambda_J at the 24 zigzag nodes),
//   the position update and Boris Jacobian-transposes,
//   the interpolation adjoint (scatter of lambda_E, lambda_B to node-mean
//   adjoints, 48 atomics, then one grid pass maps node means back to the
//   staggered components) and the gradient of the weights in x.
// Atomic targets are interleaved per node ([cell][4] current, [cell][8]
// node-mean adjoints), so one particle's atomics at a node share a cache line.
// Floors, the zigzag relay point's branch and the wrap are piecewise
// constant / identity, so their derivatives are taken on the branch the
// forward step took.
#include "rk_internal.h"

#include <hipcub/hipcub.hpp>

#include <cmath>

namespace rk {

namespace {

constexpr int DWG = 256;

template<class T>
struct DBox {
  int n[3];
  int nsp;
  int64_t off[RK_MAX_SPECIES + 1];  // species s holds particles [off[s], off[s+1])
  T q[RK_MAX_SPECIES];
  T qm[RK_MAX_SPECIES];
  T cfl;
};

template<class T>
struct G3 {
  T* c[3];
};

// one box of an ensemble: its geometry and where its cells and particles
// start in the concatenated arrays (a single box is an ensemble of one)
template<class T>
struct DEnt {
  DBox<T> b;
  int64_t c0, p0;
  // particle and cell kernels run on flat grids: the box's blocks are
  // [pb0, pb0 + pnb) and [cb0, cb0 + cnb), a per-block map gives the box
  int64_t pb0, pnb, cb0, cnb;
};

template<class T>
__device__ inline G3<T> sh(G3<T> g, int64_t o)
{
#pragma unroll
  for(int c = 0; c < 3; ++c)
    if(g.c[c]) g.c[c] += o;
  return g;
}

__host__ __device__ inline int64_t cidx(const int n[3], int i, int j, int k)
{
  return (int64_t(i) * n[1] + j) * n[2] + k;
}

__device__ inline int wrap1(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

template<class T>
__device__ inline T dot3d(const T a[3], const T b[3])
{
  T r = a[0] * b[0];
  r = r + a[1] * b[1];
  r = r + a[2] * b[2];
  return r;
}

template<class T>
__device__ inline void cross3d(const T a[3], const T b[3], T r[3])
{
  r[0] = a[1] * b[2] - a[2] * b[1];
  r[1] = -a[0] * b[2] + a[2] * b[0];
  r[2] = a[0] * b[1] - a[1] * b[0];
}

template<class T>
__device__ inline T lerp_d(T x, T A, T B)
{
  return (T(1) - x) * A + x * B;
}

template<class T>
__device__ inline T lerp3_d(const T d[3], const T m[2][2][2])
{
  const T c00 = lerp_d(d[0], m[0][0][0], m[1][0][0]);
  const T c01 = lerp_d(d[0], m[0][0][1], m[1][0][1]);
  const T c10 = lerp_d(d[0], m[0][1][0], m[1][1][0]);
  const T c11 = lerp_d(d[0], m[0][1][1], m[1][1][1]);
  const T c0 = lerp_d(d[1], c00, c10);
  const T c1 = lerp_d(d[1], c01, c11);
  return lerp_d(d[2], c0, c1);
}

template<class T>
__device__ inline void atomic_add(T* p, T v)
{
  unsafeAtomicAdd(p, v);
}

// ---------------------------------------------------------------- field stencils
// out = in + dt * curlE(E)   (push_b_fdtd2 order; out may alias in)
template<class T>
__global__ void __launch_bounds__(DWG)
  k_dcurl_b(const DEnt<T>* __restrict__ tab, G3<T> out, G3<T> in, G3<T> E, T dt)
{
  const DEnt<T> en = tab[blockIdx.z];
  const DBox<T>& b = en.b;
  if(int(blockIdx.x) >= b.n[1] || int(blockIdx.y) >= b.n[0]) return;
  out = sh(out, en.c0);
  in = sh(in, en.c0);
  E = sh(E, en.c0);
  const int j = int(blockIdx.x), i = int(blockIdx.y);
  const int n0 = b.n[0], n1 = b.n[1], n2 = b.n[2];
  const int ip = wrap1(i + 1, n0), jp = wrap1(j + 1, n1);
  for(int k = int(threadIdx.x); k < n2; k += int(blockDim.x)) {
    const int kp = wrap1(k + 1, n2);
    const int64_t c = cidx(b.n, i, j, k);
    const T DkEy = E.c[1][cidx(b.n, i, j, kp)] - E.c[1][c];
    const T DjEz = E.c[2][cidx(b.n, i, jp, k)] - E.c[2][c];
    const T DiEz = E.c[2][cidx(b.n, ip, j, k)] - E.c[2][c];
    const T DkEx = E.c[0][cidx(b.n, i, j, kp)] - E.c[0][c];
    const T DjEx = E.c[0][cidx(b.n, i, jp, k)] - E.c[0][c];
    const T DiEy = E.c[1][cidx(b.n, ip, j, k)] - E.c[1][c];
    out.c[0][c] = in.c[0][c] + dt * (DkEy - DjEz);
    out.c[1][c] = in.c[1][c] + dt * (DiEz - DkEx);
    out.c[2][c] = in.c[2][c] + dt * (DjEx - DiEy);
  }
}

// out = in + dt * curlB(B), then out -= J when J is given (push_e_fdtd2 +
// add_current order; out may alias in, not B)
template<class T>
__global__ void __launch_bounds__(DWG)
  k_dcurl_e(const DEnt<T>* __restrict__ tab, G3<T> out, G3<T> in, G3<T> B, T dt, G3<T> J)
{
  const DEnt<T> en = tab[blockIdx.z];
  const DBox<T>& b = en.b;
  if(int(blockIdx.x) >= b.n[1] || int(blockIdx.y) >= b.n[0]) return;
  out = sh(out, en.c0);
  in = sh(in, en.c0);
  B = sh(B, en.c0);
  J = sh(J, en.c0);
  const int j = int(blockIdx.x), i = int(blockIdx.y);
  const int n0 = b.n[0], n1 = b.n[1], n2 = b.n[2];
  const int im = wrap1(i - 1, n0), jm = wrap1(j - 1, n1);
  for(int k = int(threadIdx.x); k < n2; k += int(blockDim.x)) {
    const int km = wrap1(k - 1, n2);
    const int64_t c = cidx(b.n, i, j, k);
    const T DkBy = B.c[1][cidx(b.n, i, j, km)] - B.c[1][c];
    const T DjBz = B.c[2][cidx(b.n, i, jm, k)] - B.c[2][c];
    const T DiBz = B.c[2][cidx(b.n, im, j, k)] - B.c[2][c];
    const T DkBx = B.c[0][cidx(b.n, i, j, km)] - B.c[0][c];
    const T DjBx = B.c[0][cidx(b.n, i, jm, k)] - B.c[0][c];
    const T DiBy = B.c[1][cidx(b.n, im, j, k)] - B.c[1][c];
    T e0 = in.c[0][c] + dt * (DkBy - DjBz);
    T e1 = in.c[1][c] + dt * (DiBz - DkBx);
    T e2 = in.c[2][c] + dt * (DjBx - DiBy);
    if(J.c[0]) {
      e0 = e0 - J.c[0][c];
      e1 = e1 - J.c[1][c];
      e2 = e2 - J.c[2][c];
    }
    out.c[0][c] = e0;
    out.c[1][c] = e1;
    out.c[2][c] = e2;
  }
}

// binomial2, 27 taps in layout_right order (yee_lattice_current_filter_binomial2.c++:23-76)
template<class T>
__global__ void __launch_bounds__(DWG) k_dfilter(const DEnt<T>* __restrict__ tab, G3<T> out, G3<T> in)
{
  const DEnt<T> en = tab[blockIdx.z];
  const DBox<T>& b = en.b;
  if(int(blockIdx.x) >= b.n[1] || int(blockIdx.y) >= b.n[0]) return;
  out = sh(out, en.c0);
  in = sh(in, en.c0);
  const int j = int(blockIdx.x), i = int(blockIdx.y);
  const int n0 = b.n[0], n1 = b.n[1], n2 = b.n[2];
  const T w1 = T(1. / 64.), w2 = T(2. / 64.), w4 = T(4. / 64.), w8 = T(8. / 64.);
  const int I[3] = {wrap1(i - 1, n0), i, wrap1(i + 1, n0)};
  const int Jn[3] = {wrap1(j - 1, n1), j, wrap1(j + 1, n1)};
  for(int k = int(threadIdx.x); k < n2; k += int(blockDim.x)) {
    const int K[3] = {wrap1(k - 1, n2), k, wrap1(k + 1, n2)};
#pragma unroll
    for(int c = 0; c < 3; ++c) {
      const T* src = in.c[c];
      T acc = T(0);
#pragma unroll
      for(int a = 0; a < 3; ++a)
#pragma unroll
        for(int bb = 0; bb < 3; ++bb)
#pragma unroll
          for(int cc = 0; cc < 3; ++cc) {
            const int o = (a == 1) + (bb == 1) + (cc == 1);
            const T w = o == 3 ? w8 : (o == 2 ? w4 : (o == 1 ? w2 : w1));
            acc = acc + w * src[cidx(b.n, I[a], Jn[bb], K[cc])];
          }
      out.c[c][cidx(b.n, i, j, k)] = acc;
    }
  }
}

template<class T>
__global__ void __launch_bounds__(DWG) k_dfill(T* p, int64_t n, T v)
{
  for(int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < n; e += int64_t(gridDim.x) * blockDim.x)
    p[e] = v;
}

// lambda_E += NM_E^T(ME), lambda_B += NM_B^T(MB): node means back to the
// staggered components.  E_x(node N) = (Ex[N-x] + Ex[N]) / 2, so Ex[n]
// receives (ME_x[n] + ME_x[n+x]) / 2; B_x(node N) = (Bx[N] + Bx[N-y] + Bx[N-z]
// + Bx[N-y-z]) / 4, so Bx[n] receives the mean of MB_x over n, n+y, n+z, n+y+z.
// M8 holds per node ME_x, ME_y, ME_z, MB_x, MB_y, MB_z (and two pad slots).
template<class T>
__global__ void __launch_bounds__(DWG)
  k_nodemean_adj(const DEnt<T>* __restrict__ tab, G3<T> LE, G3<T> LB, const T* __restrict__ M8)
{
  const DEnt<T> en = tab[blockIdx.z];
  const DBox<T>& b = en.b;
  if(int(blockIdx.x) >= b.n[1] || int(blockIdx.y) >= b.n[0]) return;
  LE = sh(LE, en.c0);
  LB = sh(LB, en.c0);
  M8 += 8 * en.c0;
  const int j = int(blockIdx.x), i = int(blockIdx.y);
  const int ip = wrap1(i + 1, b.n[0]), jp = wrap1(j + 1, b.n[1]);
  for(int k = int(threadIdx.x); k < b.n[2]; k += int(blockDim.x)) {
    const int kp = wrap1(k + 1, b.n[2]);
    const int64_t c = cidx(b.n, i, j, k);
    const int64_t cx = cidx(b.n, ip, j, k), cy = cidx(b.n, i, jp, k), cz = cidx(b.n, i, j, kp);
    const int64_t cxy = cidx(b.n, ip, jp, k), cxz = cidx(b.n, ip, j, kp), cyz = cidx(b.n, i, jp, kp);
    auto M = [&](int64_t q, int comp) { return M8[q * 8 + comp]; };
    LE.c[0][c] += (M(c, 0) + M(cx, 0)) * T(0.5);
    LE.c[1][c] += (M(c, 1) + M(cy, 1)) * T(0.5);
    LE.c[2][c] += (M(c, 2) + M(cz, 2)) * T(0.5);
    LB.c[0][c] += ((M(c, 3) + M(cy, 3)) + (M(cz, 3) + M(cyz, 3))) * T(0.25);
    LB.c[1][c] += ((M(c, 4) + M(cx, 4)) + (M(cz, 4) + M(cxz, 4))) * T(0.25);
    LB.c[2][c] += ((M(c, 5) + M(cx, 5)) + (M(cy, 5) + M(cxy, 5))) * T(0.25);
  }
}

// interleaved [cell][4] current <-> planar components
template<class T>
__global__ void __launch_bounds__(DWG) k_j4_split(const T* __restrict__ J4, G3<T> J, int64_t n)
{
  for(int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; q < n; q += int64_t(gridDim.x) * blockDim.x) {
    J.c[0][q] = J4[q * 4 + 0];
    J.c[1][q] = J4[q * 4 + 1];
    J.c[2][q] = J4[q * 4 + 2];
  }
}

template<class T>
__global__ void __launch_bounds__(DWG) k_j4_pack(G3<T> J, T* __restrict__ J4, int64_t n)
{
  for(int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; q < n; q += int64_t(gridDim.x) * blockDim.x) {
    J4[q * 4 + 0] = J.c[0][q];
    J4[q * 4 + 1] = J.c[1][q];
    J4[q * 4 + 2] = J.c[2][q];
    J4[q * 4 + 3] = T(0);
  }
}

// ---------------------------------------------------------------- particles
template<class T>
struct PState {
  T x[3], u[3];
};

// forward quantities of one particle's step that its adjoint reuses
template<class T>
struct Fwd {
  int I[3][3];        // wrapped box indices of base cell -1, 0, +1 per axis
  T d[3];             // intra-cell offsets of the interpolation
  T mE[3][2][2][2];   // node means at the 8 corners
  T mB[3][2][2][2];
  T EBE[3], EBB[3];   // interpolated fields
  T xn[3], un[3];     // position after the push (before the wrap), momentum
  T xw[3];            // wrapped position
};

template<class T>
__device__ inline bool in_box(const DBox<T>& b, const T x[3])
{
  // the stencils below address cells -1 .. n+1 around the box
  return x[0] >= T(-1) && x[0] <= T(b.n[0] + 1) && x[1] >= T(-1) && x[1] <= T(b.n[1] + 1) && x[2] >= T(-1) &&
         x[2] <= T(b.n[2] + 1);
}

// Node means of E and B (interpolation_EB_linear_1st's corner values) at
// every node, interleaved [node][8]: E_x, E_y, E_z, B_x, B_y, B_z, 0, 0.
// E_x(N) = (Ex[N-x] + Ex[N]) / 2, B_x(N) = (Bx[N] + (Bx[N-y] + (Bx[N-z] +
// Bx[N-y-z]))) / 4: the expressions of the per-corner computation, so the
// particles read the same values they would compute.
template<class T>
__global__ void __launch_bounds__(DWG) k_node_means(const DEnt<T>* __restrict__ tab, G3<T> E, G3<T> B, T* __restrict__ NM8)
{
  const DEnt<T> en = tab[blockIdx.z];
  const DBox<T>& b = en.b;
  if(int(blockIdx.x) >= b.n[1] || int(blockIdx.y) >= b.n[0]) return;
  E = sh(E, en.c0);
  B = sh(B, en.c0);
  NM8 += 8 * en.c0;
  const int j = int(blockIdx.x), i = int(blockIdx.y);
  const int im = wrap1(i - 1, b.n[0]), jm = wrap1(j - 1, b.n[1]);
  for(int k = int(threadIdx.x); k < b.n[2]; k += int(blockDim.x)) {
    const int km = wrap1(k - 1, b.n[2]);
    const int64_t c = cidx(b.n, i, j, k);
    T* o = NM8 + c * 8;
    o[0] = (E.c[0][cidx(b.n, im, j, k)] + E.c[0][c]) / T(2);
    o[1] = (E.c[1][cidx(b.n, i, jm, k)] + E.c[1][c]) / T(2);
    o[2] = (E.c[2][cidx(b.n, i, j, km)] + E.c[2][c]) / T(2);
    o[3] = (B.c[0][c] + (B.c[0][cidx(b.n, i, jm, k)] + (B.c[0][cidx(b.n, i, j, km)] + B.c[0][cidx(b.n, i, jm, km)]))) /
           T(4);
    o[4] = (B.c[1][c] + (B.c[1][cidx(b.n, im, j, k)] + (B.c[1][cidx(b.n, i, j, km)] + B.c[1][cidx(b.n, im, j, km)]))) /
           T(4);
    o[5] = (B.c[2][c] + (B.c[2][cidx(b.n, im, j, k)] + (B.c[2][cidx(b.n, i, jm, k)] + B.c[2][cidx(b.n, im, jm, k)]))) /
           T(4);
    o[6] = T(0);
    o[7] = T(0);
  }
}

template<class T>
__device__ inline void load_node(const T* p, T v[6]);

template<>
__device__ inline void load_node<float>(const float* p, float v[6])
{
  const float4 a = *reinterpret_cast<const float4*>(p);
  const float2 c = *reinterpret_cast<const float2*>(p + 4);
  v[0] = a.x;
  v[1] = a.y;
  v[2] = a.z;
  v[3] = a.w;
  v[4] = c.x;
  v[5] = c.y;
}

template<>
__device__ inline void load_node<double>(const double* p, double v[6])
{
  const double2 a = *reinterpret_cast<const double2*>(p);
  const double2 c = *reinterpret_cast<const double2*>(p + 2);
  const double2 d = *reinterpret_cast<const double2*>(p + 4);
  v[0] = a.x;
  v[1] = a.y;
  v[2] = c.x;
  v[3] = c.y;
  v[4] = d.x;
  v[5] = d.y;
}

template<class T>
__device__ inline void gather_fields(const DBox<T>& b, const T* NM8, Fwd<T>& f, const T x[3])
{
  T p[3];
  uint32_t ii[3];
#pragma unroll
  for(int a = 0; a < 3; ++a) {
    p[a] = x[a] + T(3);  // x - origo, origo = mins - halo = -3
    ii[a] = static_cast<uint32_t>(p[a]);
    const int base = int(ii[a]) - 3;
#pragma unroll
    for(int o = 0; o < 3; ++o) f.I[a][o] = wrap1(wrap1(base + o - 1, b.n[a]), b.n[a]);
    f.d[a] = p[a] - static_cast<T>(ii[a]);
  }
#pragma unroll
  for(int ic = 0; ic < 2; ++ic)
#pragma unroll
    for(int jc = 0; jc < 2; ++jc)
#pragma unroll
      for(int kc = 0; kc < 2; ++kc) {
        T v[6];
        load_node(NM8 + cidx(b.n, f.I[0][ic + 1], f.I[1][jc + 1], f.I[2][kc + 1]) * 8, v);
#pragma unroll
        for(int c = 0; c < 3; ++c) {
          f.mE[c][ic][jc][kc] = v[c];
          f.mB[c][ic][jc][kc] = v[3 + c];
        }
      }
#pragma unroll
  for(int c = 0; c < 3; ++c) {
    f.EBE[c] = lerp3_d(f.d, f.mE[c]);
    f.EBB[c] = lerp3_d(f.d, f.mB[c]);
  }
}

template<class T>
struct BorisT {
  T v0[3], E0[3], u0[3], g, bfac, B0[3], f, w[3], u1[3], u2[3];
};

template<class T>
__device__ inline void boris_fwd(const T cfl, const T qm, const T EBE[3], const T EBB[3], const T x[3],
                                 const T vel[3], BorisT<T>& s, T xn[3], T un[3])
{
  const T hq = T(0.5) * qm;
#pragma unroll
  for(int c = 0; c < 3; ++c) s.v0[c] = vel[c] * cfl;
#pragma unroll
  for(int c = 0; c < 3; ++c) s.E0[c] = EBE[c] * hq;
#pragma unroll
  for(int c = 0; c < 3; ++c) s.u0[c] = s.v0[c] + s.E0[c];
  s.g = cfl / sqrt(cfl * cfl + dot3d(s.u0, s.u0));
  s.bfac = hq * s.g;
#pragma unroll
  for(int c = 0; c < 3; ++c) s.B0[c] = (EBB[c] * s.bfac) / cfl;
  s.f = T(2) / (T(1) + dot3d(s.B0, s.B0));
  T cr[3];
  cross3d(s.u0, s.B0, cr);
#pragma unroll
  for(int c = 0; c < 3; ++c) s.w[c] = s.u0[c] + cr[c];
#pragma unroll
  for(int c = 0; c < 3; ++c) s.u1[c] = s.w[c] * s.f;
  cross3d(s.u1, s.B0, cr);
#pragma unroll
  for(int c = 0; c < 3; ++c) s.u2[c] = (s.u0[c] + cr[c]) + s.E0[c];
#pragma unroll
  for(int c = 0; c < 3; ++c) un[c] = s.u2[c] / cfl;
  const T ginv2 = cfl / sqrt(cfl * cfl + dot3d(s.u2, s.u2));
#pragma unroll
  for(int c = 0; c < 3; ++c) xn[c] = x[c] + (un[c] * ginv2) * cfl;
}

template<class T>
__device__ inline void wrap_pos(const DBox<T>& b, const T xn[3], T xw[3])
{
  // leavers get (x < 0 ? L : 0) + fmod(x, L); |x| < 2L here, so fmod is x or x - L
#pragma unroll
  for(int a = 0; a < 3; ++a) {
    const T L = T(b.n[a]);
    const T v = xn[a];
    xw[a] = v < T(0) ? L + v : (v >= L ? v - L : v);
  }
}

// zigzag geometry of one particle (particle_current_zigzag_1st.c++:225-339)
template<class T>
struct Zig {
  T x1[3], x2[3], fi1[3], fi2[3], xr[3], F1[3], F2[3], W1[3], W2[3];
  bool mid[3];   // relay point on its (x1+x2)/2 branch
  int c1[3][2], c2[3][2];  // wrapped box indices of the segment cells and +1
  T s, invgam;
};

template<class T>
__device__ inline void zig_setup(const DBox<T>& b, const T xw[3], const T u[3], T charge, Zig<T>& z)
{
  const T one = T(1);
  z.invgam = one / sqrt(one + dot3d(u, u));
  z.s = b.cfl * z.invgam;
#pragma unroll
  for(int d = 0; d < 3; ++d) {
    z.x2[d] = xw[d] + T(3);
    z.x1[d] = z.x2[d] - u[d] * z.s;
    z.fi1[d] = floor(z.x1[d]);
    z.fi2[d] = floor(z.x2[d]);
    const T a = (z.fi1[d] < z.fi2[d] ? z.fi1[d] : z.fi2[d]) + one;
    const T b1 = (z.fi1[d] > z.fi2[d] ? z.fi1[d] : z.fi2[d]);
    const T b2 = T(0.5) * (z.x1[d] + z.x2[d]);
    const T bb = (b1 > b2 ? b1 : b2);
    z.xr[d] = (a < bb ? a : bb);
    z.mid[d] = !(b1 > b2) && !(a < bb);
    z.F1[d] = (z.xr[d] - z.x1[d]) * charge;
    z.F2[d] = (z.x2[d] - z.xr[d]) * charge;
    z.W1[d] = (z.x1[d] + z.xr[d]) * T(0.5) - z.fi1[d];
    z.W2[d] = (z.x2[d] + z.xr[d]) * T(0.5) - z.fi2[d];
    const int a1 = int(static_cast<uint32_t>(z.fi1[d])) - 3, a2 = int(static_cast<uint32_t>(z.fi2[d])) - 3;
    z.c1[d][0] = wrap1(wrap1(a1, b.n[d]), b.n[d]);
    z.c1[d][1] = wrap1(wrap1(a1 + 1, b.n[d]), b.n[d]);
    z.c2[d][0] = wrap1(wrap1(a2, b.n[d]), b.n[d]);
    z.c2[d][1] = wrap1(wrap1(a2 + 1, b.n[d]), b.n[d]);
  }
}

// the 7 (node, component) contributions of one segment, in store order:
// v[m][c] goes to node (c[0][o0], c[1][o1], c[2][o2]) with offsets ZOFF[m]
__constant__ const int ZOFF[7][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {0, 1, 1}, {1, 0, 1}, {1, 1, 0}};

template<class T>
__device__ inline void zig_values(const T F[3], const T W[3], T v[7][3])
{
  const T one = T(1);
  v[0][0] = F[0] * (one - W[1]) * (one - W[2]);
  v[0][1] = F[1] * (one - W[0]) * (one - W[2]);
  v[0][2] = F[2] * (one - W[0]) * (one - W[1]);
  v[1][0] = T(0);
  v[1][1] = F[1] * W[0] * (one - W[2]);
  v[1][2] = F[2] * W[0] * (one - W[1]);
  v[2][0] = F[0] * W[1] * (one - W[2]);
  v[2][1] = T(0);
  v[2][2] = F[2] * (one - W[0]) * W[1];
  v[3][0] = F[0] * (one - W[1]) * W[2];
  v[3][1] = F[1] * (one - W[0]) * W[2];
  v[3][2] = T(0);
  v[4][0] = F[0] * W[1] * W[2];
  v[4][1] = T(0);
  v[4][2] = T(0);
  v[5][0] = T(0);
  v[5][1] = F[1] * W[0] * W[2];
  v[5][2] = T(0);
  v[6][0] = T(0);
  v[6][1] = T(0);
  v[6][2] = F[2] * W[0] * W[1];
}

template<class T>
__device__ inline int species_of(const DBox<T>& b, int64_t p)
{
  int s = 0;
  while(s + 1 < b.nsp && p >= b.off[s + 1]) ++s;
  return s;
}

template<class T>
__global__ void __launch_bounds__(DWG)
  k_dpush(const DEnt<T>* __restrict__ tab, const uint16_t* __restrict__ bmap, const T* __restrict__ NM8, G3<T> xin, G3<T> uin, G3<T> xout, G3<T> uout,
          uint32_t* __restrict__ count, uint32_t* __restrict__ slot2, unsigned long long* err)
{
  const DEnt<T> en = tab[bmap[blockIdx.x]];
  const DBox<T>& b = en.b;
  const int64_t blk = int64_t(blockIdx.x) - en.pb0, nblk = en.pnb;
  NM8 += 8 * en.c0;
  xin = sh(xin, en.p0);
  uin = sh(uin, en.p0);
  xout = sh(xout, en.p0);
  uout = sh(uout, en.p0);
  count += en.c0;
  slot2 += 2 * en.p0;
  const int64_t np = b.off[b.nsp];
  for(int64_t p = blk * blockDim.x + threadIdx.x; p < np; p += nblk * blockDim.x) {
    const int s = species_of(b, p);
    T x[3], u[3];
#pragma unroll
    for(int a = 0; a < 3; ++a) {
      x[a] = xin.c[a][p];
      u[a] = uin.c[a][p];
    }
    if(!in_box(b, x)) {
      atomicAdd(err, 1ull);
#pragma unroll
      for(int a = 0; a < 3; ++a) {
        xout.c[a][p] = x[a];
        uout.c[a][p] = u[a];
      }
      slot2[2 * p] = UINT32_MAX;
      slot2[2 * p + 1] = UINT32_MAX;
      continue;
    }
    Fwd<T> f;
    gather_fields(b, NM8, f, x);
    BorisT<T> bs;
    boris_fwd(b.cfl, b.qm[s], f.EBE, f.EBB, x, u, bs, f.xn, f.un);
    wrap_pos(b, f.xn, f.xw);
    Zig<T> z;
    zig_setup(b, f.xw, f.un, b.q[s], z);
    // a slot in the bucket of each segment's cell (one when both share it)
    const bool same = z.fi1[0] == z.fi2[0] && z.fi1[1] == z.fi2[1] && z.fi1[2] == z.fi2[2];
    slot2[2 * p] = atomicAdd(&count[cidx(b.n, z.c1[0][0], z.c1[1][0], z.c1[2][0])], 1u);
    slot2[2 * p + 1] = same ? UINT32_MAX : atomicAdd(&count[cidx(b.n, z.c2[0][0], z.c2[1][0], z.c2[2][0])], 1u);
#pragma unroll
    for(int a = 0; a < 3; ++a) {
      xout.c[a][p] = f.xw[a];
      uout.c[a][p] = f.un[a];
    }
  }
}

// the 12 structurally non-zero (node, component) values of zig_values, in
// ZOFF order: node 0 x,y,z; node 1 y,z; node 2 x,z; node 3 x,y; node 4 x;
// node 5 y; node 6 z
template<class T>
__device__ inline void pack12(const T v[7][3], T* o)
{
  o[0] = v[0][0];
  o[1] = v[0][1];
  o[2] = v[0][2];
  o[3] = v[1][1];
  o[4] = v[1][2];
  o[5] = v[2][0];
  o[6] = v[2][2];
  o[7] = v[3][0];
  o[8] = v[3][1];
  o[9] = v[4][0];
  o[10] = v[5][1];
  o[11] = v[6][2];
}

__constant__ const int P12[12][2] = {{0, 0}, {0, 1}, {0, 2}, {1, 1}, {1, 2}, {2, 0},
                                     {2, 2}, {3, 0}, {3, 1}, {4, 0}, {5, 1}, {6, 2}};

// current records: each segment's 12 values in its cell's bucket, from the
// pushed state (the zigzag geometry is recomputed from x'', u')
template<class T>
__global__ void __launch_bounds__(DWG)
  k_dep_records(const DEnt<T>* __restrict__ tab, const uint16_t* __restrict__ bmap, G3<T> xout, G3<T> uout, const uint32_t* __restrict__ cstart,
                const uint32_t* __restrict__ slot2, T* __restrict__ rec)
{
  const DEnt<T> en = tab[bmap[blockIdx.x]];
  const DBox<T>& b = en.b;
  const int64_t blk = int64_t(blockIdx.x) - en.pb0, nblk = en.pnb;
  xout = sh(xout, en.p0);
  uout = sh(uout, en.p0);
  cstart += en.c0;
  slot2 += 2 * en.p0;
  const int64_t np = b.off[b.nsp];
  for(int64_t p = blk * blockDim.x + threadIdx.x; p < np; p += nblk * blockDim.x) {
    const uint32_t s1 = slot2[2 * p], s2 = slot2[2 * p + 1];
    if(s1 == UINT32_MAX) continue;
    const int s = species_of(b, p);
    const T xw[3] = {xout.c[0][p], xout.c[1][p], xout.c[2][p]};
    const T un[3] = {uout.c[0][p], uout.c[1][p], uout.c[2][p]};
    Zig<T> z;
    zig_setup(b, xw, un, b.q[s], z);
    T v1[7][3], v2[7][3];
    zig_values(z.F1, z.W1, v1);
    zig_values(z.F2, z.W2, v2);
    if(s2 == UINT32_MAX) {
#pragma unroll
      for(int m = 0; m < 7; ++m)
#pragma unroll
        for(int c = 0; c < 3; ++c) v1[m][c] = v1[m][c] + v2[m][c];
    } else {
      pack12(v2, rec + (int64_t(cstart[cidx(b.n, z.c2[0][0], z.c2[1][0], z.c2[2][0])]) + s2) * 12);
    }
    pack12(v1, rec + (int64_t(cstart[cidx(b.n, z.c1[0][0], z.c1[1][0], z.c1[2][0])]) + s1) * 12);
  }
}

// one thread per cell: sum its segment records, 12 atomics into the
// interleaved current at the cell's 7 zigzag nodes
template<class T>
__global__ void __launch_bounds__(DWG)
  k_cell_deposit(const DEnt<T>* __restrict__ tab, const uint16_t* __restrict__ bmap, const T* __restrict__ rec, const uint32_t* __restrict__ cstart,
                 const uint32_t* __restrict__ count, T* J4)
{
  const DEnt<T> en = tab[bmap[blockIdx.x]];
  const DBox<T>& b = en.b;
  const int64_t blk = int64_t(blockIdx.x) - en.cb0, nblk = en.cnb;
  const int64_t ncell = int64_t(b.n[0]) * b.n[1] * b.n[2];
  cstart += en.c0;
  count += en.c0;
  J4 += 4 * en.c0;
  for(int64_t c = blk * blockDim.x + threadIdx.x; c < ncell; c += nblk * blockDim.x) {
    const uint32_t cnt = count[c];
    if(cnt == 0) continue;
    const int64_t r0 = cstart[c];
    T acc[12];
#pragma unroll
    for(int q = 0; q < 12; ++q) acc[q] = T(0);
    for(uint32_t r = 0; r < cnt; ++r) {
      const T* o = rec + (r0 + r) * 12;
#pragma unroll
      for(int q = 0; q < 12; ++q) acc[q] += o[q];
    }
    const int k = int(c % b.n[2]), j = int((c / b.n[2]) % b.n[1]), i = int(c / (int64_t(b.n[1]) * b.n[2]));
#pragma unroll
    for(int q = 0; q < 12; ++q) {
      const int m = P12[q][0];
      const int64_t node = cidx(b.n, wrap1(i + ZOFF[m][0], b.n[0]), wrap1(j + ZOFF[m][1], b.n[1]),
                                wrap1(k + ZOFF[m][2], b.n[2]));
      atomic_add(&J4[node * 4 + P12[q][1]], acc[q]);
    }
  }
}

// gathered adjoint of one zigzag segment: dL/dF and dL/dW (per axis)
template<class T>
__device__ inline void zig_seg_adj(const DBox<T>& b, const T* L4, T jsign, const int c[3][2], const T F[3],
                                   const T W[3], T gF[3], T gW[3])
{
  const T one = T(1);
  gW[0] = gW[1] = gW[2] = T(0);
#pragma unroll
  for(int d = 0; d < 3; ++d) {
    const int e = d == 0 ? 1 : 0, g = d == 2 ? 1 : 2;  // the other two axes, e < g
    T v[2][2];
#pragma unroll
    for(int de = 0; de < 2; ++de)
#pragma unroll
      for(int dg = 0; dg < 2; ++dg) {
        int o[3] = {0, 0, 0};
        o[e] = de;
        o[g] = dg;
        v[de][dg] = jsign * L4[cidx(b.n, c[0][o[0]], c[1][o[1]], c[2][o[2]]) * 4 + d];
      }
    const T we = W[e], wg = W[g];
    const T G = (v[0][0] * (one - we) + v[1][0] * we) * (one - wg) + (v[0][1] * (one - we) + v[1][1] * we) * wg;
    const T dGe = (v[1][0] - v[0][0]) * (one - wg) + (v[1][1] - v[0][1]) * wg;
    const T dGg = (v[0][1] - v[0][0]) * (one - we) + (v[1][1] - v[1][0]) * we;
    gF[d] = G;
    gW[e] += F[d] * dGe;
    gW[g] += F[d] * dGg;
  }
}

template<class T>
__device__ inline void cross_adj(const T a[3], const T b[3], const T gc[3], T ga[3], T gb[3])
{
  // c = a x b: dL/da = b x gc, dL/db = gc x a (accumulated)
  T t[3];
  cross3d(b, gc, t);
#pragma unroll
  for(int k = 0; k < 3; ++k) ga[k] += t[k];
  cross3d(gc, a, t);
#pragma unroll
  for(int k = 0; k < 3; ++k) gb[k] += t[k];
}

template<class T>
__global__ void __launch_bounds__(DWG)
  k_dpush_adj(const DEnt<T>* __restrict__ tab, const uint16_t* __restrict__ bmap, const T* __restrict__ NM8, G3<T> xin, G3<T> uin, const T* L4, T jsign,
              G3<T> lx_out, G3<T> lu_out, G3<T> lx_in, G3<T> lu_in, const uint32_t* __restrict__ cstart,
              const uint32_t* __restrict__ pslot, T* __restrict__ rec)
{
  const DEnt<T> en = tab[bmap[blockIdx.x]];
  const DBox<T>& b = en.b;
  const int64_t blk = int64_t(blockIdx.x) - en.pb0, nblk = en.pnb;
  NM8 += 8 * en.c0;
  xin = sh(xin, en.p0);
  uin = sh(uin, en.p0);
  L4 += 4 * en.c0;
  lx_out = sh(lx_out, en.p0);
  lu_out = sh(lu_out, en.p0);
  lx_in = sh(lx_in, en.p0);
  lu_in = sh(lu_in, en.p0);
  cstart += en.c0;
  pslot += en.p0;
  const int64_t np = b.off[b.nsp];
  const T cfl = b.cfl;
  // every lane of a wave runs every iteration (the scatter below reduces
  // across lanes); lanes past the end or outside the box compute on a dummy
  // particle and neither scatter nor store
  for(int64_t base = blk * blockDim.x; base < np; base += nblk * blockDim.x) {
    const int64_t p = base + int64_t(threadIdx.x);
    const bool live = p < np;
    const int64_t pl = live ? p : np - 1;
    const int s = species_of(b, pl);
    T x[3], u[3], gxo[3], guo[3];
#pragma unroll
    for(int a = 0; a < 3; ++a) {
      x[a] = xin.c[a][pl];
      u[a] = uin.c[a][pl];
      gxo[a] = lx_out.c[a][pl];
      guo[a] = lu_out.c[a][pl];
    }
    const bool ok = live && in_box(b, x);
    if(live && !ok) {  // counted by the forward step; passes its adjoint through
#pragma unroll
      for(int a = 0; a < 3; ++a) {
        lx_in.c[a][p] = gxo[a];
        lu_in.c[a][p] = guo[a];
      }
    }
    if(!ok) {
#pragma unroll
      for(int a = 0; a < 3; ++a) {
        x[a] = T(0.5);
        u[a] = T(0);
        gxo[a] = T(0);
        guo[a] = T(0);
      }
    }
    // ---- recompute the forward step
    Fwd<T> f;
    gather_fields(b, NM8, f, x);
    BorisT<T> bs;
    boris_fwd(cfl, b.qm[s], f.EBE, f.EBB, x, u, bs, f.xn, f.un);
    wrap_pos(b, f.xn, f.xw);
    Zig<T> z;
    const T q = b.q[s];
    zig_setup(b, f.xw, f.un, q, z);

    // ---- deposit adjoint: J(F, W) -> x1, x2, xr -> x'' and u'
    T gF1[3], gW1[3], gF2[3], gW2[3];
    zig_seg_adj(b, L4, jsign, z.c1, z.F1, z.W1, gF1, gW1);
    zig_seg_adj(b, L4, jsign, z.c2, z.F2, z.W2, gF2, gW2);
    T gx1[3], gx2[3];
#pragma unroll
    for(int d = 0; d < 3; ++d) {
      const T gxr = q * gF1[d] - q * gF2[d] + T(0.5) * (gW1[d] + gW2[d]);
      gx1[d] = -q * gF1[d] + T(0.5) * gW1[d];
      gx2[d] = q * gF2[d] + T(0.5) * gW2[d];
      if(z.mid[d]) {
        gx1[d] += T(0.5) * gxr;
        gx2[d] += T(0.5) * gxr;
      }
    }
    // x1 = x2 - u' s(u'), s = cfl / sqrt(1 + u'.u')
    T gun[3], gxw[3];
    const T ig3 = z.invgam * z.invgam * z.invgam;
    const T g1u = dot3d(gx1, f.un);
#pragma unroll
    for(int d = 0; d < 3; ++d) {
      gxw[d] = gx2[d] + gx1[d] + gxo[d];
      gun[d] = guo[d] - z.s * gx1[d] + cfl * ig3 * g1u * f.un[d];
    }
    // ---- position update: x' = x + cfl u' / sqrt(1 + u'.u'); the wrap is the identity
    T gx[3];
    const T gxu = dot3d(gxw, f.un);
#pragma unroll
    for(int d = 0; d < 3; ++d) {
      gx[d] = gxw[d];
      gun[d] += cfl * (z.invgam * gxw[d] - ig3 * gxu * f.un[d]);
    }
    // ---- Boris Jacobian-transpose
    T gu2[3], gu0[3], gE0[3], gB0[3] = {T(0), T(0), T(0)}, gu1[3] = {T(0), T(0), T(0)};
#pragma unroll
    for(int d = 0; d < 3; ++d) {
      gu2[d] = gun[d] / cfl;
      gu0[d] = gu2[d];
      gE0[d] = gu2[d];
    }
    cross_adj(bs.u1, bs.B0, gu2, gu1, gB0);  // u2 = u0 + u1 x B0 + E0
    T gw[3];
#pragma unroll
    for(int d = 0; d < 3; ++d) gw[d] = gu1[d] * bs.f;  // u1 = w f
    const T gf = dot3d(gu1, bs.w);
#pragma unroll
    for(int d = 0; d < 3; ++d) gu0[d] += gw[d];  // w = u0 + u0 x B0
    cross_adj(bs.u0, bs.B0, gw, gu0, gB0);
#pragma unroll
    for(int d = 0; d < 3; ++d) gB0[d] += -(bs.f * bs.f) * gf * bs.B0[d];  // f = 2 / (1 + B0.B0)
    const T hq = T(0.5) * b.qm[s];
    T gEB[3], gEE[3];
#pragma unroll
    for(int d = 0; d < 3; ++d) gEB[d] = gB0[d] * (bs.bfac / cfl);  // B0 = EBB bfac / cfl
    const T gbfac = dot3d(gB0, f.EBB) / cfl;
    const T gg = hq * gbfac;  // bfac = hq g
    const T g3 = bs.g * bs.g * bs.g / (cfl * cfl);
#pragma unroll
    for(int d = 0; d < 3; ++d) gu0[d] += -gg * g3 * bs.u0[d];  // g = cfl / sqrt(cfl^2 + u0.u0)
    T gvel[3];
#pragma unroll
    for(int d = 0; d < 3; ++d) {
      gE0[d] += gu0[d];          // u0 = v0 + E0
      gvel[d] = gu0[d] * cfl;    // v0 = vel cfl
      gEE[d] = gE0[d] * hq;      // E0 = EBE hq
    }
    // ---- interpolation adjoint.  The particle's node-mean adjoints
    // (dL/dE, dL/dB at the particle times the trilinear weights of its 8
    // corners) are left as one record in its base cell's bucket; a pass over
    // cells sums them (k_cell_scatter).  Gradient of the weights in x here,
    // with the corner values loaded again (cache hits) instead of kept live.
    const T wx[2] = {T(1) - f.d[0], f.d[0]}, wy[2] = {T(1) - f.d[1], f.d[1]}, wz[2] = {T(1) - f.d[2], f.d[2]};
    T gd[3] = {T(0), T(0), T(0)};
#pragma unroll
    for(int ic = 0; ic < 2; ++ic)
#pragma unroll
      for(int jc = 0; jc < 2; ++jc)
#pragma unroll
        for(int kc = 0; kc < 2; ++kc) {
          const int64_t node = cidx(b.n, f.I[0][ic + 1], f.I[1][jc + 1], f.I[2][kc + 1]);
          T v[6];
          load_node(NM8 + node * 8, v);
          T sum = T(0);
#pragma unroll
          for(int c = 0; c < 3; ++c) sum += gEE[c] * v[c] + gEB[c] * v[3 + c];
          // d(trilinear weight)/dd: the weight with its own axis factor replaced by -1 / +1
          gd[0] += (ic ? sum : -sum) * (wy[jc] * wz[kc]);
          gd[1] += (jc ? sum : -sum) * (wx[ic] * wz[kc]);
          gd[2] += (kc ? sum : -sum) * (wx[ic] * wy[jc]);
        }
    if(ok) {
      const int64_t key = cidx(b.n, f.I[0][1], f.I[1][1], f.I[2][1]);
      T* o = rec + (int64_t(cstart[key]) + pslot[p]) * 12;
#pragma unroll
      for(int c = 0; c < 3; ++c) {
        o[c] = gEE[c];
        o[3 + c] = gEB[c];
        o[6 + c] = f.d[c];
      }
    }
    if(!ok) continue;
#pragma unroll
    for(int a = 0; a < 3; ++a) {
      lx_in.c[a][p] = gx[a] + gd[a];
      lu_in.c[a][p] = gvel[a];
    }
  }
}

// bucket slots: particle p of base cell c gets slot pslot[p] in c's bucket
template<class T>
__global__ void __launch_bounds__(DWG)
  k_cell_count(const DEnt<T>* __restrict__ tab, const uint16_t* __restrict__ bmap, G3<T> xin, uint32_t* __restrict__ count, uint32_t* __restrict__ pslot)
{
  const DEnt<T> en = tab[bmap[blockIdx.x]];
  const DBox<T>& b = en.b;
  const int64_t blk = int64_t(blockIdx.x) - en.pb0, nblk = en.pnb;
  xin = sh(xin, en.p0);
  count += en.c0;
  pslot += en.p0;
  const int64_t np = b.off[b.nsp];
  for(int64_t p = blk * blockDim.x + threadIdx.x; p < np; p += nblk * blockDim.x) {
    const T x[3] = {xin.c[0][p], xin.c[1][p], xin.c[2][p]};
    if(!in_box(b, x)) {
      pslot[p] = UINT32_MAX;
      continue;
    }
    int I[3];
#pragma unroll
    for(int a = 0; a < 3; ++a) {
      const uint32_t ii = static_cast<uint32_t>(x[a] + T(3));
      I[a] = wrap1(wrap1(int(ii) - 3, b.n[a]), b.n[a]);
    }
    pslot[p] = atomicAdd(&count[cidx(b.n, I[0], I[1], I[2])], 1u);
  }
}

// one thread per cell: sum its particles' records into the 8 corner nodes'
// node-mean adjoints (48 atomics per occupied cell instead of per particle)
template<class T>
__global__ void __launch_bounds__(DWG)
  k_cell_scatter(const DEnt<T>* __restrict__ tab, const uint16_t* __restrict__ bmap, const T* __restrict__ rec, const uint32_t* __restrict__ cstart,
                 const uint32_t* __restrict__ count, T* M8)
{
  const DEnt<T> en = tab[bmap[blockIdx.x]];
  const DBox<T>& b = en.b;
  const int64_t blk = int64_t(blockIdx.x) - en.cb0, nblk = en.cnb;
  const int64_t ncell = int64_t(b.n[0]) * b.n[1] * b.n[2];
  cstart += en.c0;
  count += en.c0;
  M8 += 8 * en.c0;
  for(int64_t c = blk * blockDim.x + threadIdx.x; c < ncell; c += nblk * blockDim.x) {
    const uint32_t cnt = count[c];
    if(cnt == 0) continue;
    const int64_t r0 = cstart[c];
    T acc[8][6];
#pragma unroll
    for(int m = 0; m < 8; ++m)
#pragma unroll
      for(int q = 0; q < 6; ++q) acc[m][q] = T(0);
    for(uint32_t r = 0; r < cnt; ++r) {
      const T* o = rec + (r0 + r) * 12;
      T g[6], d[3];
#pragma unroll
      for(int q = 0; q < 6; ++q) g[q] = o[q];
#pragma unroll
      for(int a = 0; a < 3; ++a) d[a] = o[6 + a];
      const T wx[2] = {T(1) - d[0], d[0]}, wy[2] = {T(1) - d[1], d[1]}, wz[2] = {T(1) - d[2], d[2]};
#pragma unroll
      for(int ic = 0; ic < 2; ++ic)
#pragma unroll
        for(int jc = 0; jc < 2; ++jc)
#pragma unroll
          for(int kc = 0; kc < 2; ++kc) {
            const T w = wx[ic] * wy[jc] * wz[kc];
#pragma unroll
            for(int q = 0; q < 6; ++q) acc[(ic * 2 + jc) * 2 + kc][q] += g[q] * w;
          }
    }
    const int k = int(c % b.n[2]), j = int((c / b.n[2]) % b.n[1]), i = int(c / (int64_t(b.n[1]) * b.n[2]));
#pragma unroll
    for(int ic = 0; ic < 2; ++ic)
#pragma unroll
      for(int jc = 0; jc < 2; ++jc)
#pragma unroll
        for(int kc = 0; kc < 2; ++kc) {
          const int64_t node = cidx(b.n, wrap1(i + ic, b.n[0]), wrap1(j + jc, b.n[1]), wrap1(k + kc, b.n[2]));
#pragma unroll
          for(int q = 0; q < 6; ++q) atomic_add(&M8[node * 8 + q], acc[(ic * 2 + jc) * 2 + kc][q]);
        }
  }
}

}  // namespace

// ---------------------------------------------------------------- host side
struct DiffImpl {
  rk_diff_desc d;                    // the first box (precision, cfl, filter passes are common)
  std::vector<rk_diff_desc> boxes;   // every box of the ensemble
  std::vector<int64_t> c0, p0;       // first cell / particle of each box in the concatenated arrays
  int nmax[3] = {0, 0, 0};           // largest side per axis (the field kernels' grid)
  int64_t npmax = 0, ncmax = 0;      // largest box's particles and cells
  void* tab = nullptr;               // device DEnt<T>[boxes]
  std::vector<int64_t> pb0, pnb, cb0, cnb;  // flat-grid block ranges per box (particles, cells)
  int64_t pblocks = 0, cblocks_box = 0;     // flat-grid sizes
  uint16_t* pmap = nullptr;                 // particle-kernel block -> box
  uint16_t* cmap = nullptr;                 // cell-kernel block -> box
  int64_t ncell = 0, np = 0;         // totals
  size_t vb = 4;
  void* scratch[5][3] = {};  // B1, J, Jt, LB2, LE (planar)
  void* j4 = nullptr;        // [cell][4] current (deposit) / its adjoint (gather)
  void* m8 = nullptr;        // [cell][8] node-mean adjoints of E and B
  void* nm8 = nullptr;       // [cell][8] node means of E and B1
  uint32_t* ccount = nullptr;  // particles per base cell
  uint32_t* cstart = nullptr;  // bucket offsets (exclusive sum of ccount)
  uint32_t* pslot = nullptr;   // bucket slots: per particle (adjoint) or per zigzag segment (forward)
  void* rec = nullptr;         // [record][12] bucketed by cell: current segments or interpolation adjoints
  void* scan_tmp = nullptr;
  size_t scan_bytes = 0;
  unsigned long long* err = nullptr;
};

namespace {

enum { S_B1 = 0, S_J, S_JT, S_LB2, S_LE };

template<class T>
DBox<T> make_box(const rk_diff_desc& d)
{
  DBox<T> b {};
  for(int a = 0; a < 3; ++a) b.n[a] = d.n[a];
  b.nsp = d.n_species;
  b.off[0] = 0;
  for(int s = 0; s < b.nsp; ++s) {
    b.off[s + 1] = b.off[s] + d.count[s];
    b.q[s] = static_cast<T>(d.charge[s]);
    const double sg = (0.0 < d.charge[s]) - (d.charge[s] < 0.0);
    b.qm[s] = static_cast<T>(sg / d.mass[s]);
  }
  b.cfl = static_cast<T>(d.cfl);
  return b;
}

template<class T>
hipError_t upload_table(DiffImpl& h)
{
  std::vector<DEnt<T>> t(h.boxes.size());
  for(size_t i = 0; i < t.size(); ++i)
    t[i] = DEnt<T> {make_box<T>(h.boxes[i]), h.c0[i], h.p0[i], h.pb0[i], h.pnb[i], h.cb0[i], h.cnb[i]};
  hipError_t e = hipMalloc(&h.tab, sizeof(DEnt<T>) * t.size());
  if(e == hipSuccess) e = hipMemcpy(h.tab, t.data(), sizeof(DEnt<T>) * t.size(), hipMemcpyHostToDevice);
  // block -> box maps of the flat grids
  std::vector<uint16_t> pm(size_t(std::max<int64_t>(h.pblocks, 1))), cm(size_t(std::max<int64_t>(h.cblocks_box, 1)));
  for(size_t i = 0; i < t.size(); ++i) {
    for(int64_t k = 0; k < h.pnb[i]; ++k) pm[size_t(h.pb0[i] + k)] = uint16_t(i);
    for(int64_t k = 0; k < h.cnb[i]; ++k) cm[size_t(h.cb0[i] + k)] = uint16_t(i);
  }
  if(e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&h.pmap), sizeof(uint16_t) * pm.size());
  if(e == hipSuccess) e = hipMemcpy(h.pmap, pm.data(), sizeof(uint16_t) * pm.size(), hipMemcpyHostToDevice);
  if(e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&h.cmap), sizeof(uint16_t) * cm.size());
  if(e == hipSuccess) e = hipMemcpy(h.cmap, cm.data(), sizeof(uint16_t) * cm.size(), hipMemcpyHostToDevice);
  return e;
}

template<class T>
const DEnt<T>* table(const DiffImpl& h)
{
  return static_cast<const DEnt<T>*>(h.tab);
}

template<class T>
G3<T> g3(void* const p[3])
{
  return G3<T> {{static_cast<T*>(p[0]), static_cast<T*>(p[1]), static_cast<T*>(p[2])}};
}

template<class T>
G3<T> g3s(const DiffImpl& h, int which)
{
  return G3<T> {{static_cast<T*>(h.scratch[which][0]), static_cast<T*>(h.scratch[which][1]),
                 static_cast<T*>(h.scratch[which][2])}};
}

int dfail(int code, const char* msg)
{
  set_error(msg);
  return code;
}

template<class T>
int fill_n(T* p, int64_t n, hipStream_t st)
{
  const unsigned blocks = unsigned(std::max<int64_t>(1, std::min<int64_t>((n + DWG - 1) / DWG, 4096)));
  hipLaunchKernelGGL(k_dfill<T>, dim3(blocks), dim3(DWG), 0, st, p, n, T(0));
  RK_HIP(hipGetLastError());
  return RK_OK;
}

// flat passes over every cell of every box
inline unsigned cblocks(const DiffImpl& h)
{
  return unsigned(std::max<int64_t>(1, std::min<int64_t>((h.ncell + DWG - 1) / DWG, 4096)));
}

// per-box launches: blockIdx.z is the box; x (and y) are sized for the largest box
inline dim3 grid_of(const DiffImpl& h)
{
  return dim3(unsigned(h.nmax[1]), unsigned(h.nmax[0]), unsigned(h.boxes.size()));
}
inline dim3 block_of(const DiffImpl& h) { return dim3(unsigned(std::min(h.nmax[2], DWG))); }
// flat grids over every box's particle / cell blocks (block -> box by map)
inline dim3 pgrid(const DiffImpl& h) { return dim3(unsigned(std::max<int64_t>(h.pblocks, 1))); }
inline dim3 cgrid(const DiffImpl& h) { return dim3(unsigned(std::max<int64_t>(h.cblocks_box, 1))); }

// Jf = F^P(J) in the J / Jt pair; returns the scratch slot holding the result
template<class T>
int filter_passes(const DiffImpl& h, const DEnt<T>* b, int src, hipStream_t st, int& result)
{
  int a = src, o = src == S_J ? S_JT : S_J;
  for(int pass = 0; pass < h.d.filter_passes; ++pass) {
    hipLaunchKernelGGL(k_dfilter<T>, grid_of(h), block_of(h), 0, st, b, g3s<T>(h, o), g3s<T>(h, a));
    RK_HIP(hipGetLastError());
    std::swap(a, o);
  }
  result = a;
  return RK_OK;
}

template<class T>
int step_t(DiffImpl& h, const rk_diff_state* in, rk_diff_state* out, hipStream_t st)
{
  const DEnt<T>* b = table<T>(h);
  const T half = static_cast<T>(h.d.cfl / 2), full = static_cast<T>(h.d.cfl);
  const G3<T> none {{nullptr, nullptr, nullptr}};
  const G3<T> E = g3<T>(in->E), B = g3<T>(in->B);
  // B1 = B + cfl/2 curlE(E)
  hipLaunchKernelGGL(k_dcurl_b<T>, grid_of(h), block_of(h), 0, st, b, g3s<T>(h, S_B1), B, E, half);
  RK_HIP(hipGetLastError());
  T* J4 = static_cast<T*>(h.j4);
  T* NM8 = static_cast<T*>(h.nm8);
  if(int rc = fill_n<T>(J4, 4 * h.ncell, st)) return rc;
  hipLaunchKernelGGL(k_node_means<T>, grid_of(h), block_of(h), 0, st, b, E, g3s<T>(h, S_B1), NM8);
  RK_HIP(hipGetLastError());
  if(h.np > 0) {
    // push, then the current through cell buckets: slots, offsets, records, per-cell sums
    RK_HIP(hipMemsetAsync(h.ccount, 0, sizeof(uint32_t) * size_t(h.ncell), st));
    hipLaunchKernelGGL(k_dpush<T>, pgrid(h), dim3(DWG), 0, st, b, h.pmap, NM8, g3<T>(in->x), g3<T>(in->u),
                       g3<T>(out->x), g3<T>(out->u), h.ccount, h.pslot, h.err);
    RK_HIP(hipGetLastError());
    size_t bytes = h.scan_bytes;
    RK_HIP(hipcub::DeviceScan::ExclusiveSum(h.scan_tmp, bytes, h.ccount, h.cstart, int(h.ncell), st));
    T* rec = static_cast<T*>(h.rec);
    hipLaunchKernelGGL(k_dep_records<T>, pgrid(h), dim3(DWG), 0, st, b, h.pmap, g3<T>(out->x), g3<T>(out->u),
                       h.cstart, h.pslot, rec);
    RK_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_cell_deposit<T>, cgrid(h), dim3(DWG), 0, st, b, h.cmap, rec, h.cstart, h.ccount, J4);
    RK_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(k_j4_split<T>, dim3(cblocks(h)), dim3(DWG), 0, st, J4, g3s<T>(h, S_J), h.ncell);
  RK_HIP(hipGetLastError());
  int jf = S_J;
  if(int rc = filter_passes<T>(h, b, S_J, st, jf)) return rc;
  // B2 = B1 + cfl/2 curlE(E), written to out.B (E is still the input E)
  hipLaunchKernelGGL(k_dcurl_b<T>, grid_of(h), block_of(h), 0, st, b, g3<T>(out->B), g3s<T>(h, S_B1), E, half);
  RK_HIP(hipGetLastError());
  // E' = E + cfl curlB(B2) - Jf
  hipLaunchKernelGGL(k_dcurl_e<T>, grid_of(h), block_of(h), 0, st, b, g3<T>(out->E), E, g3<T>(out->B), full,
                     g3s<T>(h, jf));
  RK_HIP(hipGetLastError());
  (void)none;
  return RK_OK;
}

template<class T>
int adjoint_t(DiffImpl& h, const rk_diff_state* in, const rk_diff_state* ao, rk_diff_state* ai, hipStream_t st)
{
  const DEnt<T>* b = table<T>(h);
  const T half = static_cast<T>(h.d.cfl / 2), full = static_cast<T>(h.d.cfl);
  const G3<T> none {{nullptr, nullptr, nullptr}};
  const G3<T> E = g3<T>(in->E), B = g3<T>(in->B);
  const G3<T> LEo = g3<T>(ao->E), LBo = g3<T>(ao->B);
  // forward B1
  hipLaunchKernelGGL(k_dcurl_b<T>, grid_of(h), block_of(h), 0, st, b, g3s<T>(h, S_B1), B, E, half);
  RK_HIP(hipGetLastError());
  // E' = E + cfl curlB(B2) - Jf:  lambda_B2 = LB' + cfl curlB^T(LE') = LB' - cfl curlE(LE')
  hipLaunchKernelGGL(k_dcurl_b<T>, grid_of(h), block_of(h), 0, st, b, g3s<T>(h, S_LB2), LBo, LEo, T(-full));
  RK_HIP(hipGetLastError());
  // B2 = B1 + cfl/2 curlE(E):  lambda_E = LE' + cfl/2 curlE^T(LB2) = LE' - cfl/2 curlB(LB2)
  hipLaunchKernelGGL(k_dcurl_e<T>, grid_of(h), block_of(h), 0, st, b, g3s<T>(h, S_LE), LEo, g3s<T>(h, S_LB2),
                     T(-half), none);
  RK_HIP(hipGetLastError());
  // lambda_J = -F^P(LE'): the filter passes on a copy of LE' in J, sign at the gather
  for(int c = 0; c < 3; ++c)
    RK_HIP(hipMemcpyAsync(h.scratch[S_J][c], ao->E[c], h.vb * size_t(h.ncell), hipMemcpyDeviceToDevice, st));
  int lj = S_J;
  if(int rc = filter_passes<T>(h, b, S_J, st, lj)) return rc;
  T* L4 = static_cast<T*>(h.j4);
  T* M8 = static_cast<T*>(h.m8);
  hipLaunchKernelGGL(k_j4_pack<T>, dim3(cblocks(h)), dim3(DWG), 0, st, g3s<T>(h, lj), L4, h.ncell);
  RK_HIP(hipGetLastError());
  if(int rc = fill_n<T>(M8, 8 * h.ncell, st)) return rc;
  T* NM8 = static_cast<T*>(h.nm8);
  hipLaunchKernelGGL(k_node_means<T>, grid_of(h), block_of(h), 0, st, b, E, g3s<T>(h, S_B1), NM8);
  RK_HIP(hipGetLastError());
  if(h.np > 0) {
    // bucket the particles by base cell: counts, offsets, then the records
    RK_HIP(hipMemsetAsync(h.ccount, 0, sizeof(uint32_t) * size_t(h.ncell), st));
    hipLaunchKernelGGL(k_cell_count<T>, pgrid(h), dim3(DWG), 0, st, b, h.pmap, g3<T>(in->x), h.ccount, h.pslot);
    RK_HIP(hipGetLastError());
    size_t bytes = h.scan_bytes;
    RK_HIP(hipcub::DeviceScan::ExclusiveSum(h.scan_tmp, bytes, h.ccount, h.cstart, int(h.ncell), st));
    T* rec = static_cast<T*>(h.rec);
    hipLaunchKernelGGL(k_dpush_adj<T>, pgrid(h), dim3(DWG), 0, st, b, h.pmap, NM8, g3<T>(in->x),
                       g3<T>(in->u), L4, T(-1), g3<T>(ao->x), g3<T>(ao->u), g3<T>(ai->x), g3<T>(ai->u), h.cstart,
                       h.pslot, rec);
    RK_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_cell_scatter<T>, cgrid(h), dim3(DWG), 0, st, b, h.cmap, rec, h.cstart, h.ccount, M8);
    RK_HIP(hipGetLastError());
  }
  // node-mean adjoints: lambda_E += NM_E^T(ME), lambda_B1 = LB2 + NM_B^T(MB)
  hipLaunchKernelGGL(k_nodemean_adj<T>, grid_of(h), block_of(h), 0, st, b, g3s<T>(h, S_LE), g3s<T>(h, S_LB2), M8);
  RK_HIP(hipGetLastError());
  // B1 = B + cfl/2 curlE(E): lambda_B = lambda_B1, lambda_E -= cfl/2 curlB(lambda_B1)
  for(int c = 0; c < 3; ++c)
    RK_HIP(hipMemcpyAsync(ai->B[c], h.scratch[S_LB2][c], h.vb * size_t(h.ncell), hipMemcpyDeviceToDevice, st));
  hipLaunchKernelGGL(k_dcurl_e<T>, grid_of(h), block_of(h), 0, st, b, g3<T>(ai->E), g3s<T>(h, S_LE),
                     g3s<T>(h, S_LB2), T(-half), none);
  RK_HIP(hipGetLastError());
  return RK_OK;
}

bool state_ok(const rk_diff_state* s, bool particles)
{
  if(!s) return false;
  for(int c = 0; c < 3; ++c) {
    if(!s->E[c] || !s->B[c]) return false;
    if(particles && (!s->x[c] || !s->u[c])) return false;
  }
  return true;
}

}  // namespace

}  // namespace rk

using namespace rk;

struct rk_diff : rk::DiffImpl {};

extern "C" {

int rk_diff_create_ensemble(rk_diff** out, const rk_diff_desc* descs, int32_t n_boxes)
{
  if(!out || !descs) return dfail(RK_EINVAL, "null argument");
  *out = nullptr;
  if(n_boxes < 1 || n_boxes > 65535) return dfail(RK_EINVAL, "diff ensemble: 1..65535 boxes");
  const rk_diff_desc& d0 = descs[0];
  if(d0.precision != 4 && d0.precision != 8) return dfail(RK_EINVAL, "diff box: precision must be 4 or 8");
  if(!(d0.cfl > 0.0) || d0.filter_passes < 0) return dfail(RK_EINVAL, "diff box: invalid cfl or filter passes");
  auto* h = new rk_diff();
  h->d = d0;
  h->vb = size_t(d0.precision);
  for(int i = 0; i < n_boxes; ++i) {
    const rk_diff_desc& d = descs[i];
    const char* bad = nullptr;
    for(int a = 0; a < 3; ++a)
      if(d.n[a] < 2 || d.n[a] > 4096) bad = "diff box: every side must be 2..4096 cells";
    if(d.n_species < 0 || d.n_species > RK_MAX_SPECIES) bad = "diff box: invalid species count";
    if(d.precision != d0.precision || d.cfl != d0.cfl || d.filter_passes != d0.filter_passes)
      bad = "diff ensemble: every box needs the same precision, cfl and filter passes";
    int64_t np = 0;
    for(int sp = 0; sp < d.n_species && !bad; ++sp) {
      if(d.count[sp] < 0 || !(d.mass[sp] > 0.0)) bad = "diff box: invalid particle count or mass";
      np += d.count[sp];
    }
    if(bad) {
      delete h;
      return dfail(RK_EINVAL, bad);
    }
    const int64_t nc = int64_t(d.n[0]) * d.n[1] * d.n[2];
    h->boxes.push_back(d);
    h->c0.push_back(h->ncell);
    h->p0.push_back(h->np);
    h->ncell += nc;
    h->np += np;
    h->ncmax = std::max(h->ncmax, nc);
    h->npmax = std::max(h->npmax, np);
    for(int a = 0; a < 3; ++a) h->nmax[a] = std::max(h->nmax[a], d.n[a]);
  }
  if(h->ncell >= (int64_t(1) << 31) || h->np >= (int64_t(1) << 31)) {
    delete h;
    return dfail(RK_EINVAL, "diff ensemble: cells and particles must each total below 2^31");
  }
  for(size_t i = 0; i < h->boxes.size(); ++i) {
    const int64_t np = (i + 1 < h->boxes.size() ? h->p0[i + 1] : h->np) - h->p0[i];
    const int64_t nc = (i + 1 < h->boxes.size() ? h->c0[i + 1] : h->ncell) - h->c0[i];
    h->pb0.push_back(h->pblocks);
    h->pnb.push_back(std::min<int64_t>((np + DWG - 1) / DWG, int64_t(1) << 20));
    h->pblocks += h->pnb.back();
    h->cb0.push_back(h->cblocks_box);
    h->cnb.push_back(std::max<int64_t>(1, std::min<int64_t>((nc + DWG - 1) / DWG, 4096)));
    h->cblocks_box += h->cnb.back();
  }
  hipError_t e = h->vb == 4 ? upload_table<float>(*h) : upload_table<double>(*h);
  for(int i = 0; i < 5 && e == hipSuccess; ++i)
    for(int c = 0; c < 3 && e == hipSuccess; ++c) e = hipMalloc(&h->scratch[i][c], h->vb * size_t(h->ncell));
  if(e == hipSuccess) e = hipMalloc(&h->j4, 4 * h->vb * size_t(h->ncell));
  if(e == hipSuccess) e = hipMalloc(&h->m8, 8 * h->vb * size_t(h->ncell));
  if(e == hipSuccess) e = hipMalloc(&h->nm8, 8 * h->vb * size_t(h->ncell));
  if(e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&h->ccount), sizeof(uint32_t) * size_t(h->ncell));
  if(e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&h->cstart), sizeof(uint32_t) * size_t(h->ncell));
  if(e == hipSuccess)
    e = hipMalloc(reinterpret_cast<void**>(&h->pslot), 2 * sizeof(uint32_t) * size_t(std::max<int64_t>(h->np, 1)));
  if(e == hipSuccess) e = hipMalloc(&h->rec, 2 * 12 * h->vb * size_t(std::max<int64_t>(h->np, 1)));
  if(e == hipSuccess)
    e = hipcub::DeviceScan::ExclusiveSum(nullptr, h->scan_bytes, h->ccount, h->cstart, int(h->ncell));
  if(e == hipSuccess) e = hipMalloc(&h->scan_tmp, std::max<size_t>(h->scan_bytes, 16));
  if(e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&h->err), sizeof(unsigned long long));
  if(e == hipSuccess) e = hipMemset(h->err, 0, sizeof(unsigned long long));
  if(e != hipSuccess) {
    rk_diff_destroy(h);
    return hip_fail(e, "rk_diff_create");
  }
  *out = h;
  return RK_OK;
}

int rk_diff_create(rk_diff** out, const rk_diff_desc* desc)
{
  if(!desc) return dfail(RK_EINVAL, "null argument");
  return rk_diff_create_ensemble(out, desc, 1);
}

int rk_diff_destroy(rk_diff* h)
{
  if(!h) return RK_OK;
  for(auto& s : h->scratch)
    for(auto& p : s)
      if(p) (void)hipFree(p);
  if(h->j4) (void)hipFree(h->j4);
  if(h->m8) (void)hipFree(h->m8);
  if(h->nm8) (void)hipFree(h->nm8);
  for(void* q : {static_cast<void*>(h->ccount), static_cast<void*>(h->cstart), static_cast<void*>(h->pslot), h->rec,
                 h->scan_tmp})
    if(q) (void)hipFree(q);
  if(h->err) (void)hipFree(h->err);
  if(h->tab) (void)hipFree(h->tab);
  if(h->pmap) (void)hipFree(h->pmap);
  if(h->cmap) (void)hipFree(h->cmap);
  delete h;
  return RK_OK;
}

int rk_diff_step(rk_diff* h, const rk_diff_state* in, rk_diff_state* out, void* stream)
{
  if(!h || !state_ok(in, h->np > 0) || !state_ok(out, h->np > 0)) return dfail(RK_EINVAL, "null state buffer");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  return h->vb == 4 ? step_t<float>(*h, in, out, st) : step_t<double>(*h, in, out, st);
}

int rk_diff_step_adjoint(rk_diff* h, const rk_diff_state* in, const rk_diff_state* adj_out, rk_diff_state* adj_in,
                         void* stream)
{
  if(!h || !state_ok(in, h->np > 0) || !state_ok(adj_out, h->np > 0) || !state_ok(adj_in, h->np > 0))
    return dfail(RK_EINVAL, "null state buffer");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  return h->vb == 4 ? adjoint_t<float>(*h, in, adj_out, adj_in, st) : adjoint_t<double>(*h, in, adj_out, adj_in, st);
}

int rk_diff_error_count(rk_diff* h, int64_t* out, void* stream)
{
  if(!h || !out) return dfail(RK_EINVAL, "null argument");
  unsigned long long v = 0;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  RK_HIP(hipMemcpyAsync(&v, h->err, sizeof(v), hipMemcpyDeviceToHost, st));
  RK_HIP(hipStreamSynchronize(st));
  *out = int64_t(v);
  return RK_OK;
}

}  // extern "C"
