// Reflecting wall (pic::reflector_wall) and edge boundary conditions
// (emf::edge_bc), the boundary tools of the reference's shock setup
// (projects/pic-shock/pic.py).
//
// Reference semantics followed (hel-astro-lab/runko v5.0.0):
//   reflect_at_wall     src/runko/pic/reflector_wall.c++:125-237
//   reflect_particles   src/runko/pic/reflector_wall.c++:257-298
//   correction current  src/runko/pic/reflector_wall.c++:36-118, pic/tile.c++:411-414
//   edge bc             src/runko/emf/yee_lattice.c++:263-306, emf/tile.c++:808-847
//
// Every expression of the reflection is evaluated in the order the reference
// writes it (-ffp-contract=off), so reflected positions and velocities are
// bit-identical to the CPU restatement.  The correction currents go to a
// world-lattice buffer with float atomics (the reference also adds them with
// atomics); deposit_current of each affected tile adds its window of that
// buffer to J and clears it.
#include "rk_internal.h"
#include "rk_reflect.h"

#include <algorithm>
#include <vector>

namespace rk {

namespace {

constexpr int WWG = 256;

struct SpeciesPtrs {
  float* pos[RK_MAX_SPECIES][3];
  float* vel[RK_MAX_SPECIES][3];
  uint64_t* id[RK_MAX_SPECIES];
  const int64_t* start[RK_MAX_SPECIES];
  uint32_t* count[RK_MAX_SPECIES];
  float charge[RK_MAX_SPECIES];
  int nsp;
};

SpeciesPtrs species_ptrs(const rk_world* w)
{
  SpeciesPtrs p {};
  p.nsp = w->d.n_species;
  for(int s = 0; s < p.nsp; ++s) {
    for(int d = 0; d < 3; ++d) {
      p.pos[s][d] = w->sp[s].pos[d];
      p.vel[s][d] = w->sp[s].vel[d];
    }
    p.id[s] = w->sp[s].id;
    p.start[s] = w->sp[s].bin_start;
    p.count[s] = w->sp[s].bin_count;
    p.charge[s] = w->sp[s].charge;
  }
  return p;
}

// One block per bin of the tile; every particle through the tile's walls in
// registration order.
__global__ void __launch_bounds__(WWG)
  k_reflect(Geo g, SpeciesPtrs sp, WallSet walls, int b0, float cfl, float origo0, float origo1, float origo2,
            int toffh0, int toffh1, int toffh2, float* Jc0, float* Jc1, float* Jc2, uint32_t* counters)
{
  const int b = b0 + int(blockIdx.x);
  const float origo[3] = {origo0, origo1, origo2};
  const int toffh[3] = {toffh0, toffh1, toffh2};
  float* const Jc[3] = {Jc0, Jc1, Jc2};
  const unsigned all = (1u << walls.n) - 1u;
  for(int s = 0; s < sp.nsp; ++s) {
    const int64_t start = sp.start[s][b];
    const int count = int(sp.count[s][b]);
    const float charge = sp.charge[s];
    for(int t = int(threadIdx.x); t < count; t += WWG) {
      const int64_t q = start + t;
      if(sp.id[s][q] == DEAD_ID) continue;
      float pos[3] = {sp.pos[s][0][q], sp.pos[s][1][q], sp.pos[s][2][q]};
      float u[3] = {sp.vel[s][0][q], sp.vel[s][1][q], sp.vel[s][2][q]};
      if(pos[0] >= walls_max_x(walls)) continue;  // in front of every wall
      const bool alive = reflect_particle(g, walls, all, cfl, origo, toffh, Jc, charge, pos, u, &counters[C_STENCIL]);
#pragma unroll
      for(int d = 0; d < 3; ++d) sp.pos[s][d][q] = pos[d];
      sp.vel[s][0][q] = u[0];
      if(!alive) sp.id[s][q] = DEAD_ID;
    }
  }
}

// In-order removal of dead slots from every bin of a tile (one block per bin).
__global__ void __launch_bounds__(WWG) k_compact_dead(SpeciesPtrs sp, int b0)
{
  __shared__ int wc[WWG / 64];
  const int b = b0 + int(blockIdx.x);
  const int lane = int(threadIdx.x) & 63, wv = int(threadIdx.x) >> 6;
  for(int s = 0; s < sp.nsp; ++s) {
    const int64_t start = sp.start[s][b];
    const int count = int(sp.count[s][b]);
    int wr = 0;
    for(int base = 0; base < count; base += WWG) {
      const int t = base + int(threadIdx.x);
      bool alive = false;
      float p[3] = {0, 0, 0}, v[3] = {0, 0, 0};
      uint64_t id = 0;
      if(t < count) {
        id = sp.id[s][start + t];
        alive = id != DEAD_ID;
#pragma unroll
        for(int d = 0; d < 3; ++d) {
          p[d] = sp.pos[s][d][start + t];
          v[d] = sp.vel[s][d][start + t];
        }
      }
      const uint64_t bal = __ballot(alive);
      if(lane == 0) wc[wv] = __popcll(bal);
      __syncthreads();
      int off = 0, tot = 0;
#pragma unroll
      for(int k = 0; k < WWG / 64; ++k) {
        off += k < wv ? wc[k] : 0;
        tot += wc[k];
      }
      __syncthreads();
      if(alive) {
        const int64_t dst = start + wr + off + __popcll(bal & ((uint64_t(1) << lane) - 1));
#pragma unroll
        for(int d = 0; d < 3; ++d) {
          sp.pos[s][d][dst] = p[d];
          sp.vel[s][d][dst] = v[d];
        }
        sp.id[s][dst] = id;
      }
      wr += tot;
    }
    __syncthreads();
    if(threadIdx.x == 0) sp.count[s][b] = uint32_t(wr);
    __syncthreads();
  }
}

// J += Jc, Jc = 0 over one haloed tile window
__global__ void k_add_corr(Geo g, int i0, int j0, int k0, int ni, int nj, int nk, float* J0, float* J1, float* J2,
                           float* C0, float* C1, float* C2)
{
  const int j = j0 + int(blockIdx.x), i = i0 + int(blockIdx.y);
  if(j >= j0 + nj || i >= i0 + ni) return;
  for(int k = k0 + int(threadIdx.x); k < k0 + nk; k += int(blockDim.x)) {
    const uint32_t q = g.at(i, j, k);
    J0[q] = J0[q] + C0[q];
    J1[q] = J1[q] + C1[q];
    J2[q] = J2[q] + C2[q];
    C0[q] = 0.0f;
    C1[q] = 0.0f;
    C2[q] = 0.0f;
  }
}

struct Box {
  int b0[3], bn[3];  // launch box: the tile's region with its halo
  int r0[3], r1[3];  // the tile's own cells of the region
};

// Sets the masked components on the launch box where a cell lies in the
// tile's own region or in the world's outer halo (the cells that stand in
// for a boundary tile's private halo copies).
__global__ void k_set_region(Geo g, Box b, float* F0, float* F1, float* F2, float v0, float v1, float v2, int mask)
{
  const int j = b.b0[1] + int(blockIdx.x), i = b.b0[0] + int(blockIdx.y);
  if(j >= b.b0[1] + b.bn[1] || i >= b.b0[0] + b.bn[0]) return;
  const bool ij_in = i >= b.r0[0] && i < b.r1[0] && j >= b.r0[1] && j < b.r1[1];
  const bool ij_out = i < HALO || i >= HALO + g.N[0] || j < HALO || j >= HALO + g.N[1];
  for(int k = b.b0[2] + int(threadIdx.x); k < b.b0[2] + b.bn[2]; k += int(blockDim.x)) {
    const bool in = ij_in && k >= b.r0[2] && k < b.r1[2];
    const bool out = ij_out || k < HALO || k >= HALO + g.N[2];
    if(!in && !out) continue;
    const uint32_t q = g.at(i, j, k);
    if(mask & 1) F0[q] = v0;
    if(mask & 2) F1[q] = v1;
    if(mask & 4) F2[q] = v2;
  }
}

int fail_w(int code, const char* msg)
{
  set_error(msg);
  return code;
}

}  // namespace

// the world buffer of wall correction currents, allocated (zeroed) on first use
int ensure_corrections(rk_world* w)
{
  if(w->corrJ[0]) return RK_OK;
  const size_t nb = sizeof(float) * size_t(w->g.nodes());
  for(int c = 0; c < 3; ++c) {
    RK_HIP(hipMalloc(reinterpret_cast<void**>(&w->corrJ[c]), nb));
    RK_HIP(hipMemset(w->corrJ[c], 0, nb));
  }
  w->corr_pending.assign(size_t(w->g.ntiles()), 0);
  return RK_OK;
}

// tiles that a wall of ws lies in get their corrections added by the next
// apply_corrections (the fused step's reflect stage writes them)
int mark_wall_tiles(rk_world* w, const WallSet& ws)
{
  if(int rc = ensure_corrections(w)) return rc;
  const Geo& g = w->g;
  for(int t = 0; t < g.ntiles(); ++t) {
    int tc[3];
    g.tile_coords(t, tc);
    const float tmin = g.tile_min(tc[0], 0);
    const float tmax = float((g.toff[0] + tc[0] + 1) * g.n[0]);
    if(walls_in_tile(ws, tmin, tmax, g.cfl)) {
      w->corr_pending[size_t(t)] = 1;
      w->corr_any = 1;
    }
  }
  return RK_OK;
}

// J += corrections and clear them on every node of the world lattice (walls
// whose tiles the host does not track: they move on the device)
__global__ void k_add_corr_world(int64_t n, float* J0, float* J1, float* J2, float* C0, float* C1, float* C2)
{
  for(int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; q < n; q += int64_t(gridDim.x) * blockDim.x) {
    J0[q] = J0[q] + C0[q];
    J1[q] = J1[q] + C1[q];
    J2[q] = J2[q] + C2[q];
    C0[q] = 0.0f;
    C1[q] = 0.0f;
    C2[q] = 0.0f;
  }
}

// advance_reflector_walls on the device: walloc += betawall * cfl in float32
// (reflector_wall.c++:293-300)
__global__ void k_advance_walls(rk_reflector_wall* wl, int n, float cfl)
{
  const int i = int(threadIdx.x);
  if(i < n) wl[i].walloc = wl[i].walloc + wl[i].betawall * cfl;
}

int add_corrections_world(rk_world* w, hipStream_t s)
{
  const int64_t n = int64_t(w->g.nodes());
  const unsigned blocks = unsigned(std::min<int64_t>((n + 255) / 256, 8192));
  hipLaunchKernelGGL(k_add_corr_world, dim3(blocks), dim3(256), 0, s, n, w->J[0], w->J[1], w->J[2], w->corrJ[0],
                     w->corrJ[1], w->corrJ[2]);
  RK_HIP(hipGetLastError());
  return RK_OK;
}

int apply_corrections(rk_world* w, int tile, hipStream_t s)
{
  if(!w->corr_any) return RK_OK;
  const Geo& g = w->g;
  const int t0 = tile < 0 ? 0 : tile, t1 = tile < 0 ? g.ntiles() : tile + 1;
  for(int t = t0; t < t1; ++t) {
    if(!w->corr_pending[size_t(t)]) continue;
    int tc[3];
    g.tile_coords(t, tc);
    const int n0 = g.n[0] + 2 * HALO, n1 = g.n[1] + 2 * HALO, n2 = g.n[2] + 2 * HALO;
    hipLaunchKernelGGL(k_add_corr, dim3(unsigned(n1), unsigned(n0)), dim3(64), 0, s, g, tc[0] * g.n[0],
                       tc[1] * g.n[1], tc[2] * g.n[2], n0, n1, n2, w->J[0], w->J[1], w->J[2], w->corrJ[0],
                       w->corrJ[1], w->corrJ[2]);
    RK_HIP(hipGetLastError());
    w->corr_pending[size_t(t)] = 0;
  }
  w->corr_any = std::any_of(w->corr_pending.begin(), w->corr_pending.end(), [](char c) { return c != 0; });
  return RK_OK;
}

}  // namespace rk

using namespace rk;

extern "C" {

int rk_reflect_particles(rk_world* w, int32_t tile, const rk_reflector_wall* walls, int32_t n_walls, void* stream)
{
  if(!w) return fail_w(RK_EINVAL, "null world");
  const Geo& g = w->g;
  if(tile < -1 || tile >= g.ntiles()) return fail_w(RK_EINVAL, "tile index out of range");
  if(n_walls < 0 || n_walls > MAX_WALLS) return fail_w(RK_EINVAL, "at most 8 reflector walls per tile");
  if(n_walls == 0) return RK_OK;
  if(!walls) return fail_w(RK_EINVAL, "null wall array");
  for(int s = 0; s < w->d.n_species; ++s)
    if(!w->sp[s].id) return fail_w(RK_EINVAL, "particle pools not set");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  if(int rc = ensure_corrections(w)) return rc;
  const float cfl = g.cfl;
  const SpeciesPtrs sp = species_ptrs(w);
  const int t0 = tile < 0 ? 0 : tile, t1 = tile < 0 ? g.ntiles() : tile + 1;
  for(int t = t0; t < t1; ++t) {
    int tc[3];
    g.tile_coords(t, tc);
    const float tmin = g.tile_min(tc[0], 0);
    const float tmax = float((g.toff[0] + tc[0] + 1) * g.n[0]);
    // wall_is_in_tile, reflector_wall.c++:261-264
    WallSet ws {};
    for(int i = 0; i < n_walls; ++i)
      if(walls[i].walloc >= tmin - cfl && walls[i].walloc <= tmax) ws.w[ws.n++] = walls[i];
    if(ws.n == 0) continue;
    w->corr_pending[size_t(t)] = 1;
    w->corr_any = 1;
    const int b0 = t * g.bpt;
    ProfScope ps(w, 8, st);
    hipLaunchKernelGGL(k_reflect, dim3(unsigned(g.bpt)), dim3(WWG), 0, st, g, sp, ws, b0, cfl,
                       g.tile_min(tc[0], 0) - 3.0f, g.tile_min(tc[1], 1) - 3.0f, g.tile_min(tc[2], 2) - 3.0f,
                       tc[0] * g.n[0], tc[1] * g.n[1], tc[2] * g.n[2], w->corrJ[0], w->corrJ[1], w->corrJ[2],
                       w->counters);
    RK_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_compact_dead, dim3(unsigned(g.bpt)), dim3(WWG), 0, st, sp, b0);
    RK_HIP(hipGetLastError());
  }
  return RK_OK;
}

int rk_world_set_walls(rk_world* w, const rk_reflector_wall* walls, int32_t n_walls)
{
  if(!w) return fail_w(RK_EINVAL, "null world");
  if(n_walls < 0 || n_walls > MAX_WALLS) return fail_w(RK_EINVAL, "at most 8 reflector walls");
  if(n_walls > 0 && !walls) return fail_w(RK_EINVAL, "null wall array");
  if(!w->walls_dev)
    RK_HIP(hipMalloc(reinterpret_cast<void**>(&w->walls_dev), sizeof(rk_reflector_wall) * MAX_WALLS));
  if(n_walls > 0) {
    if(int rc = ensure_corrections(w)) return rc;
    RK_HIP(hipMemcpy(w->walls_dev, walls, sizeof(rk_reflector_wall) * size_t(n_walls), hipMemcpyHostToDevice));
  }
  w->walls_dev_n = n_walls;
  return RK_OK;
}

int rk_advance_walls(rk_world* w, void* stream)
{
  if(!w) return fail_w(RK_EINVAL, "null world");
  if(w->walls_dev_n == 0) return RK_OK;
  hipLaunchKernelGGL(k_advance_walls, dim3(1), dim3(64), 0, reinterpret_cast<hipStream_t>(stream), w->walls_dev,
                     w->walls_dev_n, w->g.cfl);
  RK_HIP(hipGetLastError());
  return RK_OK;
}

int rk_read_walls(rk_world* w, rk_reflector_wall* out, int32_t n_walls, void* stream)
{
  if(!w || !out) return fail_w(RK_EINVAL, "null argument");
  if(n_walls != w->walls_dev_n) return fail_w(RK_EINVAL, "read_walls: n_walls differs from the walls set");
  if(n_walls == 0) return RK_OK;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  RK_HIP(hipMemcpyAsync(out, w->walls_dev, sizeof(rk_reflector_wall) * size_t(n_walls), hipMemcpyDeviceToHost, st));
  RK_HIP(hipStreamSynchronize(st));
  return RK_OK;
}

int rk_apply_edge_bc(rk_world* w, int32_t tile, const rk_edge_bc* bc, int32_t mode, void* stream)
{
  if(!w || !bc) return fail_w(RK_EINVAL, "null argument");
  const Geo& g = w->g;
  if(tile < -1 || tile >= g.ntiles()) return fail_w(RK_EINVAL, "tile index out of range");
  if(bc->direction < 0 || bc->direction > 2 || bc->side < 0 || bc->side > 1)
    return fail_w(RK_EINVAL, "edge_bc: direction must be 0..2 and side 0 or 1");
  float* const* F = nullptr;
  const float* v = nullptr;
  int mask = 0;
  switch(mode) {
    case RK_COMM_EMF_E: F = w->E; v = bc->E; mask = bc->E_components; break;
    case RK_COMM_EMF_B: F = w->B; v = bc->B; mask = bc->B_components; break;
    case RK_COMM_EMF_J: F = w->J; v = bc->J; mask = bc->J_components; break;
    default: return fail_w(RK_EINVAL, "YeeLattice::apply_edge_bc does not support given communication mode");
  }
  if(!F[0]) return fail_w(RK_EINVAL, "fields not set");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  const int d = bc->direction;
  // tile -1: one launch per slab of tiles along d (tile index tc[d]); every
  // tile of a slab has the same width, and the union of their regions and
  // world-halo cells is the slab's region over the whole world across d
  const bool all = tile < 0;
  const int t0 = all ? 0 : tile, t1 = all ? g.nt[d] : tile + 1;
  for(int t = t0; t < t1; ++t) {
    int tc[3] = {0, 0, 0};
    if(all) tc[d] = t;
    else g.tile_coords(t, tc);
    // Tile::edge_bc_width, emf/tile.c++:810-828
    const float tmin = g.tile_min(tc[d], d), tmax = float((g.toff[d] + tc[d] + 1) * g.n[d]);
    const int Nd = g.n[d];
    size_t width = 0;
    if(bc->side == 0) {
      if(bc->position <= tmin) continue;
      width = bc->position >= tmax ? size_t(Nd) : static_cast<size_t>(bc->position - tmin) + 1;
    } else {
      if(bc->position >= tmax) continue;
      width = bc->position <= tmin ? size_t(Nd) : size_t(Nd) - static_cast<size_t>(bc->position - tmin);
    }
    if(width == 0) continue;
    const int wdt = int(std::min(width, size_t(Nd)));
    // YeeLattice::apply_edge_bc sets the region on the tile's haloed
    // lattice: the halo in the other two directions and the halo on the
    // region's side of direction d.  Halo cells inside the world belong to
    // neighbour tiles, which lie in the region by the same width rule and
    // set those cells themselves; halo cells outside the world are set here.
    Box b;
    for(int a = 0; a < 3; ++a) {
      const bool whole = all && a != d;
      b.r0[a] = HALO + (whole ? 0 : tc[a] * g.n[a]);
      b.r1[a] = b.r0[a] + (whole ? g.N[a] : g.n[a]);
      b.b0[a] = b.r0[a] - HALO;
      b.bn[a] = (whole ? g.N[a] : g.n[a]) + 2 * HALO;
    }
    if(bc->side == 0) {
      b.r1[d] = b.r0[d] + wdt;
      b.bn[d] = HALO + wdt;
    } else {
      b.r0[d] = b.r1[d] - wdt;
      b.b0[d] = b.r0[d];
      b.bn[d] = wdt + HALO;
    }
    hipLaunchKernelGGL(k_set_region, dim3(unsigned(b.bn[1]), unsigned(b.bn[0])), dim3(64), 0, st, g, b, F[0], F[1],
                       F[2], v[0], v[1], v[2], mask);
    RK_HIP(hipGetLastError());
  }
  return RK_OK;
}

}  // extern "C"
