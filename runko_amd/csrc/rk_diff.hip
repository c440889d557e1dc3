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
// Floors, the zigzag relay point's branch and the wrap are piecewise
// constant / identity, so their derivatives are taken on the branch the
// forward step took.
#include "rk_internal.h"

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
  k_dcurl_b(DBox<T> b, G3<T> out, G3<T> in, G3<T> E, T dt)
{
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
  k_dcurl_e(DBox<T> b, G3<T> out, G3<T> in, G3<T> B, T dt, G3<T> J)
{
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
__global__ void __launch_bounds__(DWG) k_dfilter(DBox<T> b, G3<T> out, G3<T> in)
{
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
template<class T>
__global__ void __launch_bounds__(DWG)
  k_nodemean_adj(DBox<T> b, G3<T> LE, G3<T> LB, G3<T> ME, G3<T> MB)
{
  const int j = int(blockIdx.x), i = int(blockIdx.y);
  const int ip = wrap1(i + 1, b.n[0]), jp = wrap1(j + 1, b.n[1]);
  for(int k = int(threadIdx.x); k < b.n[2]; k += int(blockDim.x)) {
    const int kp = wrap1(k + 1, b.n[2]);
    const int64_t c = cidx(b.n, i, j, k);
    const int64_t cx = cidx(b.n, ip, j, k), cy = cidx(b.n, i, jp, k), cz = cidx(b.n, i, j, kp);
    const int64_t cxy = cidx(b.n, ip, jp, k), cxz = cidx(b.n, ip, j, kp), cyz = cidx(b.n, i, jp, kp);
    LE.c[0][c] += (ME.c[0][c] + ME.c[0][cx]) * T(0.5);
    LE.c[1][c] += (ME.c[1][c] + ME.c[1][cy]) * T(0.5);
    LE.c[2][c] += (ME.c[2][c] + ME.c[2][cz]) * T(0.5);
    LB.c[0][c] += ((MB.c[0][c] + MB.c[0][cy]) + (MB.c[0][cz] + MB.c[0][cyz])) * T(0.25);
    LB.c[1][c] += ((MB.c[1][c] + MB.c[1][cx]) + (MB.c[1][cz] + MB.c[1][cxz])) * T(0.25);
    LB.c[2][c] += ((MB.c[2][c] + MB.c[2][cx]) + (MB.c[2][cy] + MB.c[2][cxy])) * T(0.25);
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

template<class T>
__device__ inline void gather_fields(const DBox<T>& b, const G3<T>& E, const G3<T>& B, Fwd<T>& f, const T x[3])
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
  auto at = [&](int a, int bb, int c) { return cidx(b.n, f.I[0][a], f.I[1][bb], f.I[2][c]); };
#pragma unroll
  for(int ic = 0; ic < 2; ++ic)
#pragma unroll
    for(int jc = 0; jc < 2; ++jc)
#pragma unroll
      for(int kc = 0; kc < 2; ++kc) {
        f.mE[0][ic][jc][kc] = (E.c[0][at(ic, jc + 1, kc + 1)] + E.c[0][at(ic + 1, jc + 1, kc + 1)]) / T(2);
        f.mE[1][ic][jc][kc] = (E.c[1][at(ic + 1, jc, kc + 1)] + E.c[1][at(ic + 1, jc + 1, kc + 1)]) / T(2);
        f.mE[2][ic][jc][kc] = (E.c[2][at(ic + 1, jc + 1, kc)] + E.c[2][at(ic + 1, jc + 1, kc + 1)]) / T(2);
        f.mB[0][ic][jc][kc] =
          (B.c[0][at(ic + 1, jc + 1, kc + 1)] +
           (B.c[0][at(ic + 1, jc, kc + 1)] + (B.c[0][at(ic + 1, jc + 1, kc)] + B.c[0][at(ic + 1, jc, kc)]))) /
          T(4);
        f.mB[1][ic][jc][kc] =
          (B.c[1][at(ic + 1, jc + 1, kc + 1)] +
           (B.c[1][at(ic, jc + 1, kc + 1)] + (B.c[1][at(ic + 1, jc + 1, kc)] + B.c[1][at(ic, jc + 1, kc)]))) /
          T(4);
        f.mB[2][ic][jc][kc] =
          (B.c[2][at(ic + 1, jc + 1, kc + 1)] +
           (B.c[2][at(ic, jc + 1, kc + 1)] + (B.c[2][at(ic + 1, jc, kc + 1)] + B.c[2][at(ic, jc, kc + 1)]))) /
          T(4);
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

template<class T>
__device__ inline void zig_store(const DBox<T>& b, const G3<T>& J, const int c[3][2], const T F[3], const T W[3])
{
  const T one = T(1);
  auto st = [&](int di, int dj, int dk, T jx, T jy, T jz) {
    const int64_t q = cidx(b.n, c[0][di], c[1][dj], c[2][dk]);
    if(jx != T(0)) atomic_add(&J.c[0][q], jx);
    if(jy != T(0)) atomic_add(&J.c[1][q], jy);
    if(jz != T(0)) atomic_add(&J.c[2][q], jz);
  };
  st(0, 0, 0, F[0] * (one - W[1]) * (one - W[2]), F[1] * (one - W[0]) * (one - W[2]),
     F[2] * (one - W[0]) * (one - W[1]));
  st(1, 0, 0, T(0), F[1] * W[0] * (one - W[2]), F[2] * W[0] * (one - W[1]));
  st(0, 1, 0, F[0] * W[1] * (one - W[2]), T(0), F[2] * (one - W[0]) * W[1]);
  st(0, 0, 1, F[0] * (one - W[1]) * W[2], F[1] * (one - W[0]) * W[2], T(0));
  st(0, 1, 1, F[0] * W[1] * W[2], T(0), T(0));
  st(1, 0, 1, T(0), F[1] * W[0] * W[2], T(0));
  st(1, 1, 0, T(0), T(0), F[2] * W[0] * W[1]);
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
  k_dpush(DBox<T> b, G3<T> E, G3<T> B1, G3<T> xin, G3<T> uin, G3<T> xout, G3<T> uout, G3<T> J,
          unsigned long long* err)
{
  const int64_t np = b.off[b.nsp];
  for(int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; p < np; p += int64_t(gridDim.x) * blockDim.x) {
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
      continue;
    }
    Fwd<T> f;
    gather_fields(b, E, B1, f, x);
    BorisT<T> bs;
    boris_fwd(b.cfl, b.qm[s], f.EBE, f.EBB, x, u, bs, f.xn, f.un);
    wrap_pos(b, f.xn, f.xw);
    Zig<T> z;
    zig_setup(b, f.xw, f.un, b.q[s], z);
    zig_store(b, J, z.c1, z.F1, z.W1);
    zig_store(b, J, z.c2, z.F2, z.W2);
#pragma unroll
    for(int a = 0; a < 3; ++a) {
      xout.c[a][p] = f.xw[a];
      uout.c[a][p] = f.un[a];
    }
  }
}

// gathered adjoint of one zigzag segment: dL/dF and dL/dW (per axis)
template<class T>
__device__ inline void zig_seg_adj(const DBox<T>& b, const G3<T>& LJ, T jsign, const int c[3][2], const T F[3],
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
        v[de][dg] = jsign * LJ.c[d][cidx(b.n, c[0][o[0]], c[1][o[1]], c[2][o[2]])];
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
  k_dpush_adj(DBox<T> b, G3<T> E, G3<T> B1, G3<T> xin, G3<T> uin, G3<T> LJ, T jsign, G3<T> lx_out,
              G3<T> lu_out, G3<T> lx_in, G3<T> lu_in, G3<T> ME, G3<T> MB)
{
  const int64_t np = b.off[b.nsp];
  const T cfl = b.cfl;
  for(int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; p < np; p += int64_t(gridDim.x) * blockDim.x) {
    const int s = species_of(b, p);
    T x[3], u[3], gxo[3], guo[3];
#pragma unroll
    for(int a = 0; a < 3; ++a) {
      x[a] = xin.c[a][p];
      u[a] = uin.c[a][p];
      gxo[a] = lx_out.c[a][p];
      guo[a] = lu_out.c[a][p];
    }
    if(!in_box(b, x)) {  // counted by the forward step; passes its adjoint through
#pragma unroll
      for(int a = 0; a < 3; ++a) {
        lx_in.c[a][p] = gxo[a];
        lu_in.c[a][p] = guo[a];
      }
      continue;
    }
    // ---- recompute the forward step
    Fwd<T> f;
    gather_fields(b, E, B1, f, x);
    BorisT<T> bs;
    boris_fwd(cfl, b.qm[s], f.EBE, f.EBB, x, u, bs, f.xn, f.un);
    wrap_pos(b, f.xn, f.xw);
    Zig<T> z;
    const T q = b.q[s];
    zig_setup(b, f.xw, f.un, q, z);

    // ---- deposit adjoint: J(F, W) -> x1, x2, xr -> x'' and u'
    T gF1[3], gW1[3], gF2[3], gW2[3];
    zig_seg_adj(b, LJ, jsign, z.c1, z.F1, z.W1, gF1, gW1);
    zig_seg_adj(b, LJ, jsign, z.c2, z.F2, z.W2, gF2, gW2);
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
    // ---- interpolation adjoint: scatter to node means, gradient in x
    const T wx[2] = {T(1) - f.d[0], f.d[0]}, wy[2] = {T(1) - f.d[1], f.d[1]}, wz[2] = {T(1) - f.d[2], f.d[2]};
    T gd[3] = {T(0), T(0), T(0)};
#pragma unroll
    for(int ic = 0; ic < 2; ++ic)
#pragma unroll
      for(int jc = 0; jc < 2; ++jc)
#pragma unroll
        for(int kc = 0; kc < 2; ++kc) {
          const T w = wx[ic] * wy[jc] * wz[kc];
          const int64_t node = cidx(b.n, f.I[0][ic + 1], f.I[1][jc + 1], f.I[2][kc + 1]);
#pragma unroll
          for(int c = 0; c < 3; ++c) {
            atomic_add(&ME.c[c][node], gEE[c] * w);
            atomic_add(&MB.c[c][node], gEB[c] * w);
          }
        }
#pragma unroll
    for(int c = 0; c < 3; ++c) {
#pragma unroll
      for(int jc = 0; jc < 2; ++jc)
#pragma unroll
        for(int kc = 0; kc < 2; ++kc) {
          const T w = wy[jc] * wz[kc];
          gd[0] += (gEE[c] * (f.mE[c][1][jc][kc] - f.mE[c][0][jc][kc]) +
                    gEB[c] * (f.mB[c][1][jc][kc] - f.mB[c][0][jc][kc])) * w;
        }
#pragma unroll
      for(int ic = 0; ic < 2; ++ic)
#pragma unroll
        for(int kc = 0; kc < 2; ++kc) {
          const T w = wx[ic] * wz[kc];
          gd[1] += (gEE[c] * (f.mE[c][ic][1][kc] - f.mE[c][ic][0][kc]) +
                    gEB[c] * (f.mB[c][ic][1][kc] - f.mB[c][ic][0][kc])) * w;
        }
#pragma unroll
      for(int ic = 0; ic < 2; ++ic)
#pragma unroll
        for(int jc = 0; jc < 2; ++jc) {
          const T w = wx[ic] * wy[jc];
          gd[2] += (gEE[c] * (f.mE[c][ic][jc][1] - f.mE[c][ic][jc][0]) +
                    gEB[c] * (f.mB[c][ic][jc][1] - f.mB[c][ic][jc][0])) * w;
        }
    }
#pragma unroll
    for(int a = 0; a < 3; ++a) {
      lx_in.c[a][p] = gx[a] + gd[a];
      lu_in.c[a][p] = gvel[a];
    }
  }
}

}  // namespace

// ---------------------------------------------------------------- host side
struct DiffImpl {
  rk_diff_desc d;
  int64_t ncell = 0, np = 0;
  size_t vb = 4;
  void* scratch[7][3] = {};  // B1, J, Jt, LB2, LE, ME, MB
  unsigned long long* err = nullptr;
};

namespace {

enum { S_B1 = 0, S_J, S_JT, S_LB2, S_LE, S_ME, S_MB };

template<class T>
DBox<T> make_box(const DiffImpl& h)
{
  DBox<T> b {};
  for(int a = 0; a < 3; ++a) b.n[a] = h.d.n[a];
  b.nsp = h.d.n_species;
  b.off[0] = 0;
  for(int s = 0; s < b.nsp; ++s) {
    b.off[s + 1] = b.off[s] + h.d.count[s];
    b.q[s] = static_cast<T>(h.d.charge[s]);
    const double sg = (0.0 < h.d.charge[s]) - (h.d.charge[s] < 0.0);
    b.qm[s] = static_cast<T>(sg / h.d.mass[s]);
  }
  b.cfl = static_cast<T>(h.d.cfl);
  return b;
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
int fill3(const DiffImpl& h, G3<T> g, hipStream_t st)
{
  const unsigned blocks = unsigned(std::min<int64_t>((h.ncell + DWG - 1) / DWG, 4096));
  for(int c = 0; c < 3; ++c) {
    hipLaunchKernelGGL(k_dfill<T>, dim3(blocks), dim3(DWG), 0, st, g.c[c], h.ncell, T(0));
    RK_HIP(hipGetLastError());
  }
  return RK_OK;
}

inline dim3 grid_of(const DiffImpl& h) { return dim3(unsigned(h.d.n[1]), unsigned(h.d.n[0])); }
inline dim3 block_of(const DiffImpl& h) { return dim3(unsigned(std::min(h.d.n[2], DWG))); }
inline unsigned pblocks(const DiffImpl& h)
{
  return unsigned(std::max<int64_t>(1, std::min<int64_t>((h.np + DWG - 1) / DWG, int64_t(1) << 20)));
}

// Jf = F^P(J) in the J / Jt pair; returns the scratch slot holding the result
template<class T>
int filter_passes(const DiffImpl& h, const DBox<T>& b, int src, hipStream_t st, int& result)
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
  const DBox<T> b = make_box<T>(h);
  const T half = static_cast<T>(h.d.cfl / 2), full = static_cast<T>(h.d.cfl);
  const G3<T> none {{nullptr, nullptr, nullptr}};
  const G3<T> E = g3<T>(in->E), B = g3<T>(in->B);
  // B1 = B + cfl/2 curlE(E)
  hipLaunchKernelGGL(k_dcurl_b<T>, grid_of(h), block_of(h), 0, st, b, g3s<T>(h, S_B1), B, E, half);
  RK_HIP(hipGetLastError());
  if(int rc = fill3<T>(h, g3s<T>(h, S_J), st)) return rc;
  if(h.np > 0) {
    hipLaunchKernelGGL(k_dpush<T>, dim3(pblocks(h)), dim3(DWG), 0, st, b, E, g3s<T>(h, S_B1), g3<T>(in->x),
                       g3<T>(in->u), g3<T>(out->x), g3<T>(out->u), g3s<T>(h, S_J), h.err);
    RK_HIP(hipGetLastError());
  }
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
  const DBox<T> b = make_box<T>(h);
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
  if(int rc = fill3<T>(h, g3s<T>(h, S_ME), st)) return rc;
  if(int rc = fill3<T>(h, g3s<T>(h, S_MB), st)) return rc;
  if(h.np > 0) {
    hipLaunchKernelGGL(k_dpush_adj<T>, dim3(pblocks(h)), dim3(DWG), 0, st, b, E, g3s<T>(h, S_B1), g3<T>(in->x),
                       g3<T>(in->u), g3s<T>(h, lj), T(-1), g3<T>(ao->x), g3<T>(ao->u), g3<T>(ai->x), g3<T>(ai->u),
                       g3s<T>(h, S_ME), g3s<T>(h, S_MB));
    RK_HIP(hipGetLastError());
  }
  // node-mean adjoints: lambda_E += NM_E^T(ME), lambda_B1 = LB2 + NM_B^T(MB)
  hipLaunchKernelGGL(k_nodemean_adj<T>, grid_of(h), block_of(h), 0, st, b, g3s<T>(h, S_LE), g3s<T>(h, S_LB2),
                     g3s<T>(h, S_ME), g3s<T>(h, S_MB));
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

int rk_diff_create(rk_diff** out, const rk_diff_desc* desc)
{
  if(!out || !desc) return dfail(RK_EINVAL, "null argument");
  *out = nullptr;
  const rk_diff_desc& d = *desc;
  for(int a = 0; a < 3; ++a)
    if(d.n[a] < 2 || d.n[a] > 4096) return dfail(RK_EINVAL, "diff box: every side must be 2..4096 cells");
  if(d.n_species < 0 || d.n_species > RK_MAX_SPECIES) return dfail(RK_EINVAL, "diff box: invalid species count");
  if(d.precision != 4 && d.precision != 8) return dfail(RK_EINVAL, "diff box: precision must be 4 or 8");
  if(!(d.cfl > 0.0) || d.filter_passes < 0) return dfail(RK_EINVAL, "diff box: invalid cfl or filter passes");
  auto* h = new rk_diff();
  h->d = d;
  h->vb = size_t(d.precision);
  h->ncell = int64_t(d.n[0]) * d.n[1] * d.n[2];
  for(int s = 0; s < d.n_species; ++s) {
    if(d.count[s] < 0 || !(d.mass[s] > 0.0)) {
      delete h;
      return dfail(RK_EINVAL, "diff box: invalid particle count or mass");
    }
    h->np += d.count[s];
  }
  hipError_t e = hipSuccess;
  for(int i = 0; i < 7 && e == hipSuccess; ++i)
    for(int c = 0; c < 3 && e == hipSuccess; ++c) e = hipMalloc(&h->scratch[i][c], h->vb * size_t(h->ncell));
  if(e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&h->err), sizeof(unsigned long long));
  if(e == hipSuccess) e = hipMemset(h->err, 0, sizeof(unsigned long long));
  if(e != hipSuccess) {
    rk_diff_destroy(h);
    return hip_fail(e, "rk_diff_create");
  }
  *out = h;
  return RK_OK;
}

int rk_diff_destroy(rk_diff* h)
{
  if(!h) return RK_OK;
  for(auto& s : h->scratch)
    for(auto& p : s)
      if(p) (void)hipFree(p);
  if(h->err) (void)hipFree(h->err);
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
