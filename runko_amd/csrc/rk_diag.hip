// FP64 Gauss-law diagnostic: residual of div E - F(rho), where rho is the
// first-order (cloud-in-cell) charge on the lattice nodes and F is the same
// number of (1,2,1)^3 binomial passes that filter the current every step.
//
// Why this is the conserved quantity: the zigzag deposit is charge
// conserving, rho^{n+1} - rho^n = -div J with Jx[i] the flux from node i to
// node i+1 (src/runko/pic/particle_current_zigzag_1st.c++:225-339), and the
// field update subtracts the filtered current, E -= F(J)
// (src/runko/emf/yee_lattice.c++:170-179).  The filter is linear and
// shift-invariant, so div F(J) = F(div J) and div E - F(rho) stays at its
// initial value (0 for a neutral start with E = 0).  div E at node i is
// (Ex[i]-Ex[i-1]) + (Ey[j]-Ey[j-1]) + (Ez[k]-Ez[k-1]).
#include "rk_internal.h"

#include <algorithm>
#include <cmath>

namespace rk {

namespace {

constexpr int WG = 256;

__global__ void __launch_bounds__(WG)
  k_rho(Geo g, const float* px, const float* py, const float* pz, const int64_t* start, const uint32_t* count,
        double q, double* rho)
{
  const int b = int(blockIdx.x);
  int tile, lb[3], tc[3];
  g.brick_coords(b, tile, lb);
  g.tile_coords(tile, tc);
  float origo[3];
  int toffh[3];
  for(int d = 0; d < 3; ++d) {
    origo[d] = g.tile_min(tc[d], d) - 3.0f;
    toffh[d] = tc[d] * g.n[d];
  }
  const int64_t s0 = start[b];
  const int n = int(count[b]);
  for(int t = int(threadIdx.x); t < n; t += WG) {
    const float p[3] = {px[s0 + t] - origo[0], py[s0 + t] - origo[1], pz[s0 + t] - origo[2]};
    int I[3];
    double w1[3];
    bool ok = true;
    for(int d = 0; d < 3; ++d) {
      const float f = floorf(p[d]);
      I[d] = toffh[d] + int(f);
      w1[d] = double(p[d]) - double(f);
      ok = ok && I[d] >= 0 && I[d] + 1 < g.L[d];
    }
    if(!ok) continue;
    for(int c = 0; c < 8; ++c) {
      const int a = c >> 2, bb = (c >> 1) & 1, e = c & 1;
      const double wgt = (a ? w1[0] : 1.0 - w1[0]) * (bb ? w1[1] : 1.0 - w1[1]) * (e ? w1[2] : 1.0 - w1[2]);
      atomicAdd(&rho[g.at(I[0] + a, I[1] + bb, I[2] + e)], q * wgt);
    }
  }
}

__device__ inline int wrapc(int c, int N) { return c < HALO ? c + N : (c >= N + HALO ? c - N : c); }

__global__ void k_ghost_fold_d(Geo g, double* R)
{
  // interior nodes += images in the ghost shell
  const int j = HALO + int(blockIdx.x), i = HALO + int(blockIdx.y);
  for(int k = HALO + int(threadIdx.x); k < g.N[2] + HALO; k += int(blockDim.x)) {
    int ii[3], jj[3], kk[3];
    int mi = 0, mj = 0, mk = 0;
    ii[mi++] = i;
    if(i >= g.N[0]) ii[mi++] = i - g.N[0];
    if(i < 2 * HALO) ii[mi++] = i + g.N[0];
    jj[mj++] = j;
    if(j >= g.N[1]) jj[mj++] = j - g.N[1];
    if(j < 2 * HALO) jj[mj++] = j + g.N[1];
    kk[mk++] = k;
    if(k >= g.N[2]) kk[mk++] = k - g.N[2];
    if(k < 2 * HALO) kk[mk++] = k + g.N[2];
    if(mi * mj * mk == 1) continue;
    double s = R[g.at(i, j, k)];
    for(int a = 0; a < mi; ++a)
      for(int b = 0; b < mj; ++b)
        for(int e = 0; e < mk; ++e)
          if(a || b || e) s += R[g.at(ii[a], jj[b], kk[e])];
    R[g.at(i, j, k)] = s;
  }
}

__global__ void k_ghost_copy_d(Geo g, double* R)
{
  const int j = int(blockIdx.x), i = int(blockIdx.y);
  for(int k = int(threadIdx.x); k < g.L[2]; k += int(blockDim.x)) {
    const bool ghost = i < HALO || j < HALO || k < HALO || i >= g.N[0] + HALO || j >= g.N[1] + HALO ||
                       k >= g.N[2] + HALO;
    if(ghost) R[g.at(i, j, k)] = R[g.at(wrapc(i, g.N[0]), wrapc(j, g.N[1]), wrapc(k, g.N[2]))];
  }
}

__global__ void k_filter_d(Geo g, const double* S, double* D)
{
  const int j = HALO + int(blockIdx.x), i = HALO + int(blockIdx.y);
  const double w[3] = {1.0, 2.0, 1.0};
  for(int k = HALO + int(threadIdx.x); k < g.N[2] + HALO; k += int(blockDim.x)) {
    double acc = 0.0;
    for(int a = 0; a < 3; ++a)
      for(int b = 0; b < 3; ++b)
        for(int c = 0; c < 3; ++c) acc += w[a] * w[b] * w[c] / 64.0 * S[g.at(i + a - 1, j + b - 1, k + c - 1)];
    D[g.at(i, j, k)] = acc;
  }
}

__global__ void k_residual(Geo g, const float* Ex, const float* Ey, const float* Ez, const double* R, double* part)
{
  const int j = HALO + int(blockIdx.x), i = HALO + int(blockIdx.y);
  double mx = 0.0, r2 = 0.0, d2 = 0.0;
  const uint32_t si = uint32_t(g.L[1]) * uint32_t(g.L[2]), sj = uint32_t(g.L[2]);
  for(int k = HALO + int(threadIdx.x); k < g.N[2] + HALO; k += int(blockDim.x)) {
    const uint32_t c = g.at(i, j, k);
    const double div = (double(Ex[c]) - double(Ex[c - si])) + (double(Ey[c]) - double(Ey[c - sj])) +
                       (double(Ez[c]) - double(Ez[c - 1]));
    const double rho = R[c];
    mx = fmax(mx, fabs(div - rho));
    r2 += rho * rho;
    d2 += div * div;
  }
  __shared__ double sm[3][64];
  sm[0][threadIdx.x] = mx;
  sm[1][threadIdx.x] = r2;
  sm[2][threadIdx.x] = d2;
  __syncthreads();
  for(int o = 32; o > 0; o >>= 1) {
    if(int(threadIdx.x) < o) {
      sm[0][threadIdx.x] = fmax(sm[0][threadIdx.x], sm[0][threadIdx.x + o]);
      sm[1][threadIdx.x] += sm[1][threadIdx.x + o];
      sm[2][threadIdx.x] += sm[2][threadIdx.x + o];
    }
    __syncthreads();
  }
  if(threadIdx.x == 0) {
    const int64_t r = int64_t(blockIdx.y) * gridDim.x + blockIdx.x;
    part[3 * r] = sm[0][0];
    part[3 * r + 1] = sm[1][0];
    part[3 * r + 2] = sm[2][0];
  }
}

}  // namespace

int gauss_residual(rk_world* w, int passes, double* out, hipStream_t s)
{
  const Geo& g = w->g;
  const size_t nodes = size_t(g.nodes());
  double *rho = nullptr, *tmp = nullptr, *part = nullptr;
  const size_t rows = size_t(g.N[0]) * size_t(g.N[1]);
  RK_HIP(hipMallocAsync(reinterpret_cast<void**>(&rho), sizeof(double) * nodes, s));
  RK_HIP(hipMallocAsync(reinterpret_cast<void**>(&tmp), sizeof(double) * nodes, s));
  RK_HIP(hipMallocAsync(reinterpret_cast<void**>(&part), sizeof(double) * 3 * rows, s));
  RK_HIP(hipMemsetAsync(rho, 0, sizeof(double) * nodes, s));
  RK_HIP(hipMemsetAsync(tmp, 0, sizeof(double) * nodes, s));
  for(int sp = 0; sp < w->d.n_species; ++sp) {
    const Species& P = w->sp[sp];
    hipLaunchKernelGGL(k_rho, dim3(unsigned(g.nbricks)), dim3(WG), 0, s, g, P.pos[0], P.pos[1], P.pos[2], P.bin_start,
                       P.bin_count, w->d.charge[sp], rho);
    RK_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(k_ghost_fold_d, dim3(unsigned(g.N[1]), unsigned(g.N[0])), dim3(64), 0, s, g, rho);
  for(int p = 0; p < passes; ++p) {
    hipLaunchKernelGGL(k_ghost_copy_d, dim3(unsigned(g.L[1]), unsigned(g.L[0])), dim3(64), 0, s, g, rho);
    hipLaunchKernelGGL(k_filter_d, dim3(unsigned(g.N[1]), unsigned(g.N[0])), dim3(64), 0, s, g, rho, tmp);
    std::swap(rho, tmp);
  }
  hipLaunchKernelGGL(k_residual, dim3(unsigned(g.N[1]), unsigned(g.N[0])), dim3(64), 0, s, g, w->E[0], w->E[1],
                     w->E[2], rho, part);
  RK_HIP(hipGetLastError());
  std::vector<double> h(3 * rows);
  RK_HIP(hipMemcpyAsync(h.data(), part, sizeof(double) * h.size(), hipMemcpyDeviceToHost, s));
  RK_HIP(hipStreamSynchronize(s));
  double mx = 0, r2 = 0, d2 = 0;
  for(size_t r = 0; r < rows; ++r) {
    mx = std::max(mx, h[3 * r]);
    r2 += h[3 * r + 1];
    d2 += h[3 * r + 2];
  }
  const double ncell = double(g.N[0]) * g.N[1] * g.N[2];
  out[0] = mx;
  out[1] = std::sqrt(r2 / ncell);
  out[2] = std::sqrt(d2 / ncell);
  RK_HIP(hipFreeAsync(rho, s));
  RK_HIP(hipFreeAsync(tmp, s));
  RK_HIP(hipFreeAsync(part, s));
  RK_HIP(hipStreamSynchronize(s));
  return RK_OK;
}

}  // namespace rk
