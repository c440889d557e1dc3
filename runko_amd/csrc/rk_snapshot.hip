// Snapshot packing for the "RNKO" v3 output files (the reference's MPI-IO
// writers, src/runko/io/snapshots/): the device side that turns the world's
// fields and particles into the files' arrays.  The host side writes the
// 512-byte header and the payload.
//
// Reference semantics followed (hel-astro-lab/runko v5.0.0):
//   fields    mpiio_fields.c++:218-321   E, B sampled at stride points, J summed
//                                        over stride^3 cells (z, y, x loop order),
//                                        particle counts per coarse cell
//   particles mpiio_particles.c++:196-259 E, B interpolated at sampled particles
//                                        (interpolate_EB_linear_1st)
//   spectra   mpiio_spectra.c++:130-209  log10(|u|) and beta_{x,y,z} histograms per
//                                        coarse x bin, out-of-range values in the
//                                        edge bins
// Output arrays are float and laid out as the files store them:
//   fields  [field][oz][oy][ox]            (z outermost, x innermost)
//   spectra [field][tile z][tile y][x bin][bin]
// Particle counts and histogram entries are sums of 1.0f (exact below 2^24).
#include "rk_internal.h"

#include <cmath>

namespace rk {

namespace {

constexpr int SWG = 256;

int sfail(int code, const char* msg)
{
  set_error(msg);
  return code;
}

struct SnapDims {
  int stride;
  int nt[3];   // output points per tile per axis: max(1, n / stride)
  int no[3];   // global output points per axis
};

// E, B sampled at (ix, iy, iz) * stride of each tile; J summed over the
// stride^3 cells from there, z outermost, x innermost; counts zeroed.
__global__ void __launch_bounds__(SWG)
  k_snap_fields(Geo g, SnapDims s, const float* E0, const float* E1, const float* E2, const float* B0,
                const float* B1, const float* B2, const float* J0, const float* J1, const float* J2, float* out,
                int nfields)
{
  const int oy = int(blockIdx.x), oz = int(blockIdx.y);
  const int64_t plane = int64_t(s.no[0]) * s.no[1] * s.no[2];
  const int ty = oy / s.nt[1], iy = oy % s.nt[1], tz = oz / s.nt[2], iz = oz % s.nt[2];
  for(int ox = int(threadIdx.x); ox < s.no[0]; ox += int(blockDim.x)) {
    const int tx = ox / s.nt[0], ix = ox % s.nt[0];
    // world haloed lattice index of the sample point
    const int si = HALO + tx * g.n[0] + ix * s.stride;
    const int sj = HALO + ty * g.n[1] + iy * s.stride;
    const int sk = HALO + tz * g.n[2] + iz * s.stride;
    const uint32_t q = g.at(si, sj, sk);
    const int64_t o = (int64_t(oz) * s.no[1] + oy) * s.no[0] + ox;
    out[0 * plane + o] = E0[q];
    out[1 * plane + o] = E1[q];
    out[2 * plane + o] = E2[q];
    out[3 * plane + o] = B0[q];
    out[4 * plane + o] = B1[q];
    out[5 * plane + o] = B2[q];
    float sx = 0.0f, sy = 0.0f, sz = 0.0f;
    for(int kk = 0; kk < s.stride; ++kk)
      for(int jj = 0; jj < s.stride; ++jj)
        for(int ii = 0; ii < s.stride; ++ii) {
          const uint32_t r = g.at(si + ii, sj + jj, sk + kk);
          sx += J0[r];
          sy += J1[r];
          sz += J2[r];
        }
    out[6 * plane + o] = sx;
    out[7 * plane + o] = sy;
    out[8 * plane + o] = sz;
    for(int f = 9; f < nfields; ++f) out[f * plane + o] = 0.0f;
  }
}

// particle counts of one species at the coarse cell floor((x - mins) / stride)
// (x inv_stride in float); particles whose coarse cell is outside the
// tile's output block are not counted
__global__ void __launch_bounds__(SWG)
  k_snap_density(Geo g, SnapDims s, const float* px, const float* py, const float* pz, const int64_t* start,
                 const uint32_t* count, float* dens, float inv_stride)
{
  const int b = int(blockIdx.x);
  int t, lb[3];
  g.brick_coords(b, t, lb);
  int tc[3];
  g.tile_coords(t, tc);
  const float mx = g.tile_min(tc[0], 0), my = g.tile_min(tc[1], 1), mz = g.tile_min(tc[2], 2);
  const int64_t st = start[b];
  const int cnt = int(count[b]);
  for(int i = int(threadIdx.x); i < cnt; i += int(blockDim.x)) {
    const float x = px[st + i] - mx, y = py[st + i] - my, z = pz[st + i] - mz;
    const float fx = floorf(x * inv_stride), fy = floorf(y * inv_stride), fz = floorf(z * inv_stride);
    if(!(fx >= 0.0f && fy >= 0.0f && fz >= 0.0f)) continue;
    if(fx >= float(s.nt[0]) || fy >= float(s.nt[1]) || fz >= float(s.nt[2])) continue;
    const int ox = tc[0] * s.nt[0] + int(fx), oy = tc[1] * s.nt[1] + int(fy), oz = tc[2] * s.nt[2] + int(fz);
    atomicAdd(&dens[(int64_t(oz) * s.no[1] + oy) * s.no[0] + ox], 1.0f);
  }
}

struct SpecParams {
  int nxt, nxg, nbins, ntile_y, ntile_z;
  float inv_stride, log_umin, inv_dlog, inv_dbeta, last_bin;
};

__device__ inline int spec_bin(float raw, float last)
{
  // clamp to [0, last] (out-of-range values land in the edge bins)
  const float c = fminf(fmaxf(raw, 0.0f), last);
  return int(c);
}

// four histograms per species: log10|u|, beta_x, beta_y, beta_z, per coarse x bin of each tile
__global__ void __launch_bounds__(SWG)
  k_snap_spectra(Geo g, SpecParams p, const float* px, const float* ux_, const float* uy_, const float* uz_,
                 const int64_t* start, const uint32_t* count, float* out, int base_field)
{
  const int b = int(blockIdx.x);
  int t, lb[3];
  g.brick_coords(b, t, lb);
  int tc[3];
  g.tile_coords(t, tc);
  const float mx = g.tile_min(tc[0], 0);
  const int64_t st = start[b];
  const int cnt = int(count[b]);
  const int64_t field = int64_t(p.ntile_z) * p.ntile_y * p.nxg * p.nbins;
  for(int i = int(threadIdx.x); i < cnt; i += int(blockDim.x)) {
    const float fx = floorf((px[st + i] - mx) * p.inv_stride);
    if(!(fx >= 0.0f) || fx >= float(p.nxt)) continue;
    const float ux = ux_[st + i], uy = uy_[st + i], uz = uz_[st + i];
    const float u2 = ux * ux + uy * uy + uz * uz;
    const float u_mag = sqrtf(u2);
    const float gamma = sqrtf(1.0f + u2);
    const float inv_gamma = 1.0f / gamma;
    // log10 through double and one rounding to float: the correctly rounded value
    const float lu = float(log10(double(u_mag)));
    const int ib[4] = {spec_bin(floorf((lu - p.log_umin) * p.inv_dlog), p.last_bin),
                       spec_bin(floorf((ux * inv_gamma + 1.0f) * p.inv_dbeta), p.last_bin),
                       spec_bin(floorf((uy * inv_gamma + 1.0f) * p.inv_dbeta), p.last_bin),
                       spec_bin(floorf((uz * inv_gamma + 1.0f) * p.inv_dbeta), p.last_bin)};
    const int64_t row = ((int64_t(tc[2]) * p.ntile_y + tc[1]) * p.nxg + tc[0] * p.nxt + int(fx)) * p.nbins;
#pragma unroll
    for(int f = 0; f < 4; ++f) atomicAdd(&out[(base_field + f) * field + row + ib[f]], 1.0f);
  }
}

// E, B at arbitrary positions of one tile: interpolate_EB_linear_1st on the
// tile's haloed window (same expressions as the push)
__device__ inline float lerp_s(float x, float A, float B) { return (1.0f - x) * A + x * B; }

__device__ inline float lerp3_s(const float d[3], const float m[2][2][2])
{
  const float c00 = lerp_s(d[0], m[0][0][0], m[1][0][0]);
  const float c01 = lerp_s(d[0], m[0][0][1], m[1][0][1]);
  const float c10 = lerp_s(d[0], m[0][1][0], m[1][1][0]);
  const float c11 = lerp_s(d[0], m[0][1][1], m[1][1][1]);
  const float c0 = lerp_s(d[1], c00, c10);
  const float c1 = lerp_s(d[1], c01, c11);
  return lerp_s(d[2], c0, c1);
}

__global__ void __launch_bounds__(SWG)
  k_interp_at(Geo g, int tile, int64_t n, const float* px, const float* py, const float* pz, const float* E0,
              const float* E1, const float* E2, const float* B0, const float* B1, const float* B2, float* oe0,
              float* oe1, float* oe2, float* ob0, float* ob1, float* ob2, uint32_t* err)
{
  int tc[3];
  g.tile_coords(tile, tc);
  const float origo[3] = {g.tile_min(tc[0], 0) - 3.0f, g.tile_min(tc[1], 1) - 3.0f, g.tile_min(tc[2], 2) - 3.0f};
  const int woff[3] = {tc[0] * g.n[0], tc[1] * g.n[1], tc[2] * g.n[2]};  // tile window origin in the world lattice
  for(int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < n; e += int64_t(gridDim.x) * blockDim.x) {
    const float p[3] = {px[e] - origo[0], py[e] - origo[1], pz[e] - origo[2]};
    // the stencil reads haloed indices i-1 .. i+1 of the tile window
    bool ok = true;
#pragma unroll
    for(int a = 0; a < 3; ++a) ok = ok && p[a] >= 1.0f && p[a] < float(g.n[a] + 2 * HALO - 1);
    if(!ok) {
      atomicAdd(err, 1u);
      oe0[e] = oe1[e] = oe2[e] = ob0[e] = ob1[e] = ob2[e] = 0.0f;
      continue;
    }
    const int i = int(static_cast<uint32_t>(p[0])), j = int(static_cast<uint32_t>(p[1])),
              k = int(static_cast<uint32_t>(p[2]));
    auto at = [&](int a, int b, int c) { return g.at(woff[0] + a, woff[1] + b, woff[2] + c); };
    float ex[2][2][2], ey[2][2][2], ez[2][2][2], bx[2][2][2], by[2][2][2], bz[2][2][2];
#pragma unroll
    for(int ic = 0; ic < 2; ++ic)
#pragma unroll
      for(int jc = 0; jc < 2; ++jc)
#pragma unroll
        for(int kc = 0; kc < 2; ++kc) {
          const int ii = i + ic, jj = j + jc, kk = k + kc;
          ex[ic][jc][kc] = (E0[at(ii - 1, jj, kk)] + E0[at(ii, jj, kk)]) / 2.0f;
          ey[ic][jc][kc] = (E1[at(ii, jj - 1, kk)] + E1[at(ii, jj, kk)]) / 2.0f;
          ez[ic][jc][kc] = (E2[at(ii, jj, kk - 1)] + E2[at(ii, jj, kk)]) / 2.0f;
          bx[ic][jc][kc] =
            (B0[at(ii, jj, kk)] + (B0[at(ii, jj - 1, kk)] + (B0[at(ii, jj, kk - 1)] + B0[at(ii, jj - 1, kk - 1)]))) /
            4.0f;
          by[ic][jc][kc] =
            (B1[at(ii, jj, kk)] + (B1[at(ii - 1, jj, kk)] + (B1[at(ii, jj, kk - 1)] + B1[at(ii - 1, jj, kk - 1)]))) /
            4.0f;
          bz[ic][jc][kc] =
            (B2[at(ii, jj, kk)] + (B2[at(ii - 1, jj, kk)] + (B2[at(ii, jj - 1, kk)] + B2[at(ii - 1, jj - 1, kk)]))) /
            4.0f;
        }
    const float d[3] = {p[0] - float(i), p[1] - float(j), p[2] - float(k)};
    oe0[e] = lerp3_s(d, ex);
    oe1[e] = lerp3_s(d, ey);
    oe2[e] = lerp3_s(d, ez);
    ob0[e] = lerp3_s(d, bx);
    ob1[e] = lerp3_s(d, by);
    ob2[e] = lerp3_s(d, bz);
  }
}

bool dims_of(const Geo& g, int stride, SnapDims& s)
{
  s.stride = stride;
  for(int a = 0; a < 3; ++a) {
    if(stride < 1 || stride > g.n[a]) return false;
    s.nt[a] = std::max(1, g.n[a] / stride);
    s.no[a] = g.nt[a] * s.nt[a];
  }
  return true;
}

}  // namespace

}  // namespace rk

using namespace rk;

extern "C" {

int rk_snapshot_fields(rk_world* w, int32_t stride, int32_t nspecies, float* out, void* stream)
{
  if(!w || !out) return sfail(RK_EINVAL, "null argument");
  const Geo& g = w->g;
  SnapDims s {};
  if(!dims_of(g, stride, s)) return sfail(RK_EINVAL, "snapshot stride must be 1 .. cells per tile");
  if(nspecies < 0) return sfail(RK_EINVAL, "invalid species count");
  if(!w->E[0] || !w->J[0]) return sfail(RK_EINVAL, "fields not set");
  const int ns = std::min<int>(nspecies, 5);
  const int nfields = 9 + ns;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(k_snap_fields, dim3(unsigned(s.no[1]), unsigned(s.no[2])), dim3(SWG), 0, st, g, s, w->E[0],
                     w->E[1], w->E[2], w->B[0], w->B[1], w->B[2], w->J[0], w->J[1], w->J[2], out, nfields);
  RK_HIP(hipGetLastError());
  const int64_t plane = int64_t(s.no[0]) * s.no[1] * s.no[2];
  const float inv_stride = 1.0f / float(stride);
  for(int sp = 0; sp < std::min(ns, w->d.n_species); ++sp) {
    if(!w->sp[sp].id) return sfail(RK_EINVAL, "particle pools not set");
    hipLaunchKernelGGL(k_snap_density, dim3(unsigned(g.nbricks)), dim3(SWG), 0, st, g, s, w->sp[sp].pos[0],
                       w->sp[sp].pos[1], w->sp[sp].pos[2], w->sp[sp].bin_start, w->sp[sp].bin_count,
                       out + (9 + sp) * plane, inv_stride);
    RK_HIP(hipGetLastError());
  }
  return RK_OK;
}

int rk_snapshot_spectra(rk_world* w, int32_t stride, int32_t nbins, float umin, float umax, int32_t nspecies,
                        float* out, void* stream)
{
  if(!w || !out) return sfail(RK_EINVAL, "null argument");
  const Geo& g = w->g;
  if(stride < 1 || stride > g.n[0]) return sfail(RK_EINVAL, "spectra stride must be 1 .. cells per tile along x");
  if(nbins < 1) return sfail(RK_EINVAL, "spectra need at least one bin");
  if(!(umin > 0.0f) || !(umax > umin)) return sfail(RK_EINVAL, "spectra need 0 < umin < umax");
  SpecParams p {};
  p.nxt = std::max(1, g.n[0] / stride);
  p.nxg = g.nt[0] * p.nxt;
  p.nbins = nbins;
  p.ntile_y = g.nt[1];
  p.ntile_z = g.nt[2];
  p.inv_stride = 1.0f / float(stride);
  // SpectraWriter constructor, mpiio_spectra.c++:111-116; the logarithms are
  // the correctly rounded float values (through double), the rest float
  p.log_umin = float(std::log10(double(umin)));
  const float log_umax = float(std::log10(double(umax)));
  p.inv_dlog = float(nbins) / (log_umax - p.log_umin);
  p.inv_dbeta = float(nbins) / 2.0f;
  p.last_bin = float(nbins - 1);
  const int ns = std::min<int>(std::max(nspecies, 0), 3);
  const int64_t field = int64_t(p.ntile_z) * p.ntile_y * p.nxg * p.nbins;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  RK_HIP(hipMemsetAsync(out, 0, sizeof(float) * size_t(field) * size_t(4 * ns), st));
  for(int sp = 0; sp < std::min(ns, w->d.n_species); ++sp) {
    if(!w->sp[sp].id) return sfail(RK_EINVAL, "particle pools not set");
    hipLaunchKernelGGL(k_snap_spectra, dim3(unsigned(g.nbricks)), dim3(SWG), 0, st, g, p, w->sp[sp].pos[0],
                       w->sp[sp].vel[0], w->sp[sp].vel[1], w->sp[sp].vel[2], w->sp[sp].bin_start,
                       w->sp[sp].bin_count, out, 4 * sp);
    RK_HIP(hipGetLastError());
  }
  return RK_OK;
}

int rk_interpolate_at(rk_world* w, int32_t tile, int64_t n, const float* const pos[3], float* const E[3],
                      float* const B[3], void* stream)
{
  if(!w || !pos || !E || !B) return sfail(RK_EINVAL, "null argument");
  const Geo& g = w->g;
  if(tile < 0 || tile >= g.ntiles()) return sfail(RK_EINVAL, "tile index out of range");
  if(n <= 0) return RK_OK;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  const unsigned blocks = unsigned(std::min<int64_t>((n + SWG - 1) / SWG, 65536));
  hipLaunchKernelGGL(k_interp_at, dim3(blocks), dim3(SWG), 0, st, g, tile, n, pos[0], pos[1], pos[2], w->E[0],
                     w->E[1], w->E[2], w->B[0], w->B[1], w->B[2], E[0], E[1], E[2], B[0], B[1], B[2],
                     &w->counters[C_STENCIL]);
  RK_HIP(hipGetLastError());
  return RK_OK;
}

}  // extern "C"
