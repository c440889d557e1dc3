// C-ABI entry points (include/runko_amd.h): world lifecycle, the per-tile /
// batched methods with the call-sequence rules that make one world lattice
// behave like the reference's tiles, communication, overflow handling and
// readback.
#include "rk_internal.h"
#include "rk_reflect.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <string>

namespace rk {

thread_local std::string g_err;

void set_error(const std::string& msg) { g_err = msg; }

int hip_fail(hipError_t e, const char* where)
{
  g_err = std::string("HIP error ") + hipGetErrorString(e) + " at " + where;
  return RK_EHIP;
}

}  // namespace rk

using namespace rk;

namespace {

hipStream_t S(void* s) { return reinterpret_cast<hipStream_t>(s); }

int fail(int code, const std::string& msg)
{
  set_error(msg);
  return code;
}

int sign_of(double v) { return (0.0 < v) - (v < 0.0); }

// interior region of one tile (or the world interior for tile < 0)
void interior(const rk_world* w, int tile, int r0[3], int r1[3])
{
  const Geo& g = w->g;
  if(tile < 0) {
    for(int d = 0; d < 3; ++d) {
      r0[d] = HALO;
      r1[d] = HALO + g.N[d];
    }
    return;
  }
  int tc[3];
  g.tile_coords(tile, tc);
  for(int d = 0; d < 3; ++d) {
    r0[d] = HALO + tc[d] * g.n[d];
    r1[d] = r0[d] + g.n[d];
  }
}

// filter range: interior, extended to [1, L-1) on the world boundary (the
// reference filters [1, L-1) of every tile's haloed lattice)
void filter_region(const rk_world* w, int tile, int r0[3], int r1[3])
{
  const Geo& g = w->g;
  if(tile < 0) {
    for(int d = 0; d < 3; ++d) {
      r0[d] = 1;
      r1[d] = g.L[d] - 1;
    }
    return;
  }
  int tc[3];
  g.tile_coords(tile, tc);
  interior(w, tile, r0, r1);
  for(int d = 0; d < 3; ++d) {
    if(tc[d] == 0) r0[d] = 1;
    if(tc[d] == g.nt[d] - 1) r1[d] = g.L[d] - 1;
  }
}

int check_tile(const rk_world* w, int tile)
{
  if(!w) return fail(RK_EINVAL, "null world");
  if(tile < -1 || tile >= w->g.ntiles()) return fail(RK_EINVAL, "tile index out of range");
  if(!w->E[0]) return fail(RK_EINVAL, "world fields not set");
  return RK_OK;
}

int check_particles(const rk_world* w)
{
  for(int s = 0; s < w->d.n_species; ++s)
    if(!w->sp[s].bin_start) return fail(RK_EINVAL, "particle pool of a species is not set");
  return RK_OK;
}

// a consumer of J closes the current deposit generation
void close_j(rk_world* w) { w->j_accumulating = false; }

int read_counter(rk_world* w, int which, uint32_t& v, hipStream_t s)
{
  RK_HIP(hipMemcpyAsync(&v, &w->counters[which], sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  RK_HIP(hipStreamSynchronize(s));
  return RK_OK;
}

// Record buffers and the overflow list may only be replaced while no
// overflow record is pending: the list indexes the current buffers.
int check_no_pending(rk_world* w)
{
  if(!w->counters) return RK_OK;
  uint32_t v = 0;
  RK_HIP(hipDeviceSynchronize());
  RK_HIP(hipMemcpy(&v, &w->counters[C_OVF], sizeof(uint32_t), hipMemcpyDeviceToHost));
  if(v) return fail(RK_ESTATE, "overflow records pending: repack and retry the append before replacing record buffers");
  return RK_OK;
}

int check_lost(rk_world* w, hipStream_t s)
{
  uint32_t lost = 0;
  int rc = read_counter(w, C_MIGLOST, lost, s);
  if(rc) return rc;
  if(lost) return fail(RK_EOVERFLOW, "migration record buffer too small (" + std::to_string(lost) +
                                         " particles could not migrate); grow it and re-pack");
  return RK_OK;
}

}  // namespace

namespace {

// Partial sums of rk_energies_device: one file-scope device array, of which
// every GPU holds its own copy until the process ends.  Layout: 2 doubles per
// x plane of the region (sum E^2, sum B^2), then ENERGY_KIN doubles per
// species (kinetic partials, one per block).  Host code never names the array
// in a launch or copy: its address on each device comes from
// hipGetSymbolAddress, resolved once per device (with that device current)
// when the first world is created there.
constexpr int ENERGY_PLANES = 8192;
constexpr int ENERGY_KIN = 1024;  // >= kinetic_partial_count() for any tile or world
constexpr int ENERGY_DOUBLES = 2 * ENERGY_PLANES + RK_MAX_SPECIES * ENERGY_KIN;  // 160 KiB
static __device__ __align__(256) double scratch[ENERGY_DOUBLES];

int resolve_energy_scratch(double** out)
{
  static std::mutex mu;
  static std::map<int, double*> per_device;
  int dev = 0;
  RK_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(mu);
  auto it = per_device.find(dev);
  if(it == per_device.end()) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if(e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(energy scratch)");
    if(!p) return fail(RK_EHIP, "hipGetSymbolAddress(energy scratch) returned null");
    it = per_device.emplace(dev, static_cast<double*>(p)).first;
  }
  *out = it->second;
  return RK_OK;
}

// The host's sums of rk_energies in the same order (planes in order from 0,
// then / 2; kinetic partials in order from 0), so the values are bit-identical.
__global__ void k_energy_finish(const double* __restrict__ part, int planes, int kin_blocks, int nsp, double* out)
{
  if(blockIdx.x != 0 || threadIdx.x != 0) return;
  double e = 0.0, b = 0.0;
  for(int p = 0; p < planes; ++p) {
    e += part[2 * p];
    b += part[2 * p + 1];
  }
  out[0] = e / 2.0;
  out[1] = b / 2.0;
  for(int s = 0; s < nsp; ++s) {
    double acc = 0.0;
    const double* k = part + 2 * ENERGY_PLANES + s * ENERGY_KIN;
    for(int q = 0; q < kin_blocks; ++q) acc += k[q];
    out[2 + s] = acc;
  }
}

}  // namespace

extern "C" {

const char* rk_last_error(void) { return rk::g_err.c_str(); }

int rk_version(void) { return 1; }

// x/cfl through a reciprocal and an FMA correction is used only where it was
// checked bit-identical to the division for every float input (cached per cfl)
static int fast_division_ok(float cfl, float rcp, int& ok_div, int& ok_sqrt)
{
  static std::mutex mu;
  static std::map<uint32_t, std::pair<int, int>> cache;
  uint32_t key;
  std::memcpy(&key, &cfl, sizeof key);
  std::lock_guard<std::mutex> lk(mu);
  auto it = cache.find(key);
  if(it == cache.end()) {
    uint64_t bad[2] = {0, 0};
    if(int rc = verify_fast_division(cfl, rcp, bad)) return rc;
    it = cache.emplace(key, std::make_pair(bad[0] == 0 ? 1 : 0, bad[1] == 0 ? 1 : 0)).first;
  }
  ok_div = it->second.first;
  ok_sqrt = it->second.second;
  return RK_OK;
}

int rk_world_create(rk_world** out, const rk_world_desc* d)
{
  if(!out || !d) return fail(RK_EINVAL, "null argument");
  *out = nullptr;
  if(d->n_species < 0 || d->n_species > RK_MAX_SPECIES) return fail(RK_EINVAL, "at most 4 particle species");
  for(int a = 0; a < 3; ++a) {
    if(d->n_cells[a] < HALO)
      return fail(RK_EINVAL, "Yee Lattice extents (" + std::to_string(d->n_cells[0]) + ", " +
                               std::to_string(d->n_cells[1]) + ", " + std::to_string(d->n_cells[2]) +
                               ") are assumed to be at least halo size: 3");
    if(d->n_tiles[a] < 1 || d->grid_tiles[a] < 1) return fail(RK_EINVAL, "tile counts must be positive");
    if(d->brick[a] < 1 || d->n_cells[a] % d->brick[a]) return fail(RK_EINVAL, "brick must divide n_cells");
  }
  if(!(d->cfl > 0)) return fail(RK_EINVAL, "cfl must be positive");
  auto* w = new(std::nothrow) rk_world();
  if(!w) return fail(RK_EINVAL, "out of host memory");
  w->d = *d;
  Geo& g = w->g;
  int nb = 1;
  for(int a = 0; a < 3; ++a) {
    g.n[a] = d->n_cells[a];
    g.nt[a] = d->n_tiles[a];
    g.N[a] = d->n_cells[a] * d->n_tiles[a];
    g.L[a] = g.N[a] + 2 * HALO;
    g.br[a] = d->brick[a];
    g.nb[a] = d->n_cells[a] / d->brick[a];
    g.toff[a] = d->tile_offset[a];
    g.wmin[a] = 0.0f;
    g.wmax[a] = float(int64_t(d->grid_tiles[a]) * d->n_cells[a]);
    nb *= g.nb[a];
  }
  g.bpt = nb;
  g.nbricks = nb * g.ntiles();
  g.periodic = d->periodic;
  g.cfl = float(d->cfl);
  g.rcp_cfl = 1.0f / g.cfl;
  g.fast_div = 0;
  g.fast_sqrt = 0;
  for(int a = 0; a < 3; ++a) {
    if(g.n[a] >= (1 << 16)) {
      delete w;
      return fail(RK_EINVAL, "at most 65535 cells per tile and dimension");
    }
    g.br_magic[a] = g.br[a] > 1 ? uint32_t(((uint64_t(1) << 32) + uint64_t(g.br[a]) - 1) / uint64_t(g.br[a])) : 0u;
  }
  if(int64_t(g.L[0]) * g.L[1] * g.L[2] >= (int64_t(1) << 32)) {
    delete w;
    return fail(RK_EINVAL, "world lattice too large for 32-bit node indices");
  }
  if(g.nbricks >= (1 << 24)) {
    delete w;
    return fail(RK_EINVAL, "too many bricks (>= 2^24)");
  }
  double maxq = 0.0;
  for(int s = 0; s < d->n_species; ++s) {
    w->sp[s].qm = static_cast<float>(double(sign_of(d->charge[s])) / d->mass[s]);
    w->sp[s].charge = static_cast<float>(d->charge[s]);
    maxq = std::max(maxq, std::fabs(double(w->sp[s].charge)));
  }
  // every contribution is |q| * (a move < 1 cell) * weights <= 1: scale so
  // that max|q| * jscale <= 2^30 (a contribution rounds to an int32); the
  // int64 window sums are exact up to 2^63, and converting a sum back to FP32
  // rounds once, like any FP32 result
  {
    int e = 0;
    if(maxq > 0.0 && std::isfinite(maxq)) (void)std::frexp(maxq, &e);  // maxq < 2^e
    g.jscale = std::ldexp(1.0, 30 - e);
    g.jinv = std::ldexp(1.0, e - 30);
  }
  hipError_t e = hipMalloc(reinterpret_cast<void**>(&w->counters), sizeof(uint32_t) * C_NUM);
  if(e == hipSuccess) e = hipMemset(w->counters, 0, sizeof(uint32_t) * C_NUM);
  w->red_cap = std::max<int64_t>(2 * int64_t(g.L[0]) + 16, 2048);
  if(e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&w->red), sizeof(double) * size_t(w->red_cap));
  if(e != hipSuccess) {
    (void)hipFree(w->counters);
    (void)hipFree(w->red);
    delete w;
    return hip_fail(e, "rk_world_create");
  }
  w->filter_done.assign(size_t(g.ntiles()), 0);
  if(int rc = resolve_energy_scratch(&w->energy_scratch)) {
    rk_world_destroy(w);
    return rc;
  }
  if(int rc = fast_division_ok(g.cfl, g.rcp_cfl, g.fast_div, g.fast_sqrt)) {
    rk_world_destroy(w);
    return rc;
  }
  *out = w;
  return RK_OK;
}

int rk_world_destroy(rk_world* w)
{
  if(!w) return RK_OK;
  for(auto& r : w->prof) {
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  (void)hipFree(w->counters);
  (void)hipFree(w->red);
  (void)hipFree(w->ovf_list);
  (void)hipFree(w->mbox_count);
  (void)hipFree(w->mig_table);
  for(int c = 0; c < 3; ++c) (void)hipFree(w->corrJ[c]);
  (void)hipFree(w->walls_dev);
  delete w;
  return RK_OK;
}

int64_t rk_world_field_nodes(const rk_world* w) { return w ? w->g.nodes() : 0; }

int32_t rk_world_num_bricks(const rk_world* w) { return w ? w->g.nbricks : 0; }

int rk_world_set_fields(rk_world* w, float* const E[3], float* const B[3], float* const J[3], float* const Jf[3])
{
  if(!w) return fail(RK_EINVAL, "null world");
  for(int d = 0; d < 3; ++d) {
    if(!E[d] || !B[d] || !J[d] || !Jf[d]) return fail(RK_EINVAL, "null field pointer");
    w->E[d] = E[d];
    w->B[d] = B[d];
    w->J[d] = J[d];
    w->Jf[d] = Jf[d];
  }
  return RK_OK;
}

int rk_world_set_species(rk_world* w, int32_t s, int64_t capacity, float* const pos[3], float* const vel[3],
                         uint64_t* id, int64_t* bin_start, uint32_t* bin_cap, uint32_t* bin_count)
{
  if(!w || s < 0 || s >= w->d.n_species) return fail(RK_EINVAL, "invalid species");
  auto& p = w->sp[s];
  p.capacity = capacity;
  for(int d = 0; d < 3; ++d) {
    p.pos[d] = pos[d];
    p.vel[d] = vel[d];
  }
  p.id = id;
  p.bin_start = bin_start;
  p.bin_cap = bin_cap;
  p.bin_count = bin_count;
  return RK_OK;
}

int rk_world_set_migration(rk_world* w, int32_t set, int64_t capacity, float* const pos[3], float* const vel[3],
                           uint64_t* id, uint32_t* meta)
{
  if(!w || set < 0 || set > 1) return fail(RK_EINVAL, "invalid record set");
  if(capacity >= (int64_t(1) << 30)) return fail(RK_EINVAL, "record set capacity must be < 2^30");
  if(int rc = check_no_pending(w)) return rc;
  auto& m = w->mig[set];
  m.capacity = capacity;
  for(int d = 0; d < 3; ++d) {
    m.pos[d] = pos[d];
    m.vel[d] = vel[d];
  }
  m.id = id;
  m.meta = meta;
  const int64_t need = w->mig[0].capacity + w->mig[1].capacity + w->mbox.capacity;
  if(need > w->ovf_cap) {
    (void)hipFree(w->ovf_list);
    w->ovf_list = nullptr;
    RK_HIP(hipMalloc(reinterpret_cast<void**>(&w->ovf_list), sizeof(uint32_t) * size_t(std::max<int64_t>(need, 1))));
    w->ovf_cap = need;
  }
  return upload_mig_table(w);
}

int rk_world_set_current_staging(rk_world* w, float* buf, int64_t stride)
{
  if(!w) return fail(RK_EINVAL, "null world");
  const int64_t need = 3 * int64_t(w->g.br[0] + 3) * (w->g.br[1] + 3) * (w->g.br[2] + 3);
  if(!buf || stride < need) return fail(RK_EINVAL, "current staging: null buffer or stride below 3*(br+3)^3");
  w->jstage = buf;
  w->jstage_stride = stride;
  return RK_OK;
}

int rk_world_set_mailbox(rk_world* w, int64_t per_brick, float* const pos[3], float* const vel[3], uint64_t* id,
                         uint32_t* meta)
{
  if(!w) return fail(RK_EINVAL, "null world");
  if(int rc = check_no_pending(w)) return rc;
  const int64_t cap = per_brick * int64_t(w->g.nbricks);
  if(per_brick <= 0 || cap >= (int64_t(1) << 30)) return fail(RK_EINVAL, "mailbox capacity must be in (0, 2^30)");
  auto& m = w->mbox;
  m.capacity = cap;
  for(int d = 0; d < 3; ++d) {
    m.pos[d] = pos[d];
    m.vel[d] = vel[d];
  }
  m.id = id;
  m.meta = meta;
  w->mbox_per_brick = per_brick;
  if(!w->mbox_count) {
    RK_HIP(hipMalloc(reinterpret_cast<void**>(&w->mbox_count), sizeof(uint32_t) * size_t(w->g.nbricks)));
    RK_HIP(hipMemset(w->mbox_count, 0, sizeof(uint32_t) * size_t(w->g.nbricks)));
  }
  const int64_t need = w->mig[0].capacity + w->mig[1].capacity + w->mbox.capacity;
  if(need > w->ovf_cap) {
    (void)hipFree(w->ovf_list);
    w->ovf_list = nullptr;
    RK_HIP(hipMalloc(reinterpret_cast<void**>(&w->ovf_list), sizeof(uint32_t) * size_t(need)));
    w->ovf_cap = need;
  }
  return upload_mig_table(w);
}

int rk_push_half_b(rk_world* w, int32_t tile, void* stream)
{
  if(int rc = check_tile(w, tile)) return rc;
  int r0[3], r1[3];
  interior(w, tile, r0, r1);
  ProfScope ps(w, 5, S(stream));
  const float dt = static_cast<float>(w->d.cfl / 2);
  if(w->stencil_on) return launch_curl_b_stencil(w, r0, r1, dt, S(stream));
  return launch_curl_b(w, r0, r1, dt, S(stream));
}

int rk_world_set_schemes(rk_world* w, int32_t particle_pusher, const float* stencil)
{
  if(!w) return fail(RK_EINVAL, "null world");
  if(particle_pusher < 0 || particle_pusher > 2) return fail(RK_EINVAL, "unknown particle pusher");
  w->g.pusher = particle_pusher;
  w->stencil_on = stencil != nullptr;
  if(stencil) std::memcpy(w->stencil, stencil, sizeof(w->stencil));
  return RK_OK;
}

int rk_push_e(rk_world* w, int32_t tile, void* stream)
{
  if(int rc = check_tile(w, tile)) return rc;
  int r0[3], r1[3];
  interior(w, tile, r0, r1);
  ProfScope ps(w, 5, S(stream));
  return launch_curl_e(w, r0, r1, static_cast<float>(w->d.cfl), false, S(stream));
}

int rk_push_e_add_current(rk_world* w, int32_t tile, void* stream)
{
  if(int rc = check_tile(w, tile)) return rc;
  close_j(w);
  int r0[3], r1[3];
  interior(w, tile, r0, r1);
  ProfScope ps(w, 5, S(stream));
  return launch_curl_e(w, r0, r1, static_cast<float>(w->d.cfl), true, S(stream));
}

int rk_add_current(rk_world* w, int32_t tile, void* stream)
{
  if(int rc = check_tile(w, tile)) return rc;
  close_j(w);
  int r0[3], r1[3];
  interior(w, tile, r0, r1);
  ProfScope ps(w, 5, S(stream));
  return launch_add_current(w, r0, r1, S(stream));
}

int rk_filter_current(rk_world* w, int32_t tile, void* stream)
{
  if(int rc = check_tile(w, tile)) return rc;
  close_j(w);
  int r0[3], r1[3];
  filter_region(w, tile, r0, r1);
  auto swap = [&] {
    for(int d = 0; d < 3; ++d) std::swap(w->J[d], w->Jf[d]);
  };
  if(tile < 0) {
    if(w->filter_open)
      return fail(RK_ESTATE, "filter_current: a per-tile filter pass is open; finish it on every tile first");
    {
      ProfScope ps(w, 6, S(stream));
      if(int rc = launch_filter(w, r0, r1, S(stream))) return rc;
    }
    swap();
    return RK_OK;
  }
  if(!w->filter_open) {
    std::fill(w->filter_done.begin(), w->filter_done.end(), 0);
    w->filter_done_count = 0;
    w->filter_open = true;
  }
  if(w->filter_done[size_t(tile)])
    return fail(RK_ESTATE,
                "filter_current: tile filtered twice in one pass (every tile of the grid must be filtered once "
                "per pass)");
  {
    ProfScope ps(w, 6, S(stream));
    if(int rc = launch_filter(w, r0, r1, S(stream))) return rc;
  }
  w->filter_done[size_t(tile)] = 1;
  if(++w->filter_done_count == w->g.ntiles()) {
    swap();
    w->filter_open = false;
  }
  return RK_OK;
}

int rk_filter_current_passes(rk_world* w, int32_t passes, void* stream)
{
  if(int rc = check_tile(w, -1)) return rc;
  if(passes < 0) return fail(RK_EINVAL, "filter_current_passes: negative pass count");
  if(w->filter_open)
    return fail(RK_ESTATE, "filter_current: a per-tile filter pass is open; finish it on every tile first");
  if(passes == 3 && w->g.periodic && w->d.current_filter == 0) {
    close_j(w);
    {
      ProfScope ps(w, 6, S(stream));
      if(int rc = launch_filter_x3(w, S(stream))) return rc;
    }
    for(int d = 0; d < 3; ++d) std::swap(w->J[d], w->Jf[d]);
    // the new J's ghost layer: periodic images of the filtered interior
    ProfScope ps(w, 7, S(stream));
    return launch_ghost_copy(w, w->J, S(stream));
  }
  for(int p = 0; p < passes; ++p) {
    if(p > 0 && w->g.periodic) {
      ProfScope ps(w, 7, S(stream));
      if(int rc = launch_ghost_copy(w, w->J, S(stream))) return rc;
    }
    if(int rc = rk_filter_current(w, -1, stream)) return rc;
  }
  return RK_OK;
}

int rk_push_particles(rk_world* w, int32_t tile, void* stream)
{
  if(int rc = check_tile(w, tile)) return rc;
  if(int rc = check_particles(w)) return rc;
  ProfScope ps(w, 1, S(stream));
  return launch_push(w, tile, S(stream));
}

int rk_deposit_current(rk_world* w, int32_t tile, void* stream)
{
  if(int rc = check_tile(w, tile)) return rc;
  if(int rc = check_particles(w)) return rc;
  if(w->filter_open) return fail(RK_ESTATE, "deposit_current during an open filter pass");
  if(!w->j_accumulating) {
    // clear_current + zeroed generated J: J = 0 on the whole lattice (tile.c++:369-398)
    ProfScope ps(w, 8, S(stream));
    if(int rc = launch_fill(w->J, w->g.nodes(), 0.0f, S(stream))) return rc;
    w->j_accumulating = true;
  }
  ProfScope ps(w, 2, S(stream));
  if(int rc = launch_deposit(w, tile, S(stream))) return rc;
  // pending reflector-wall correction currents, pic/tile.c++:411-414
  return apply_corrections(w, tile, S(stream));
}

int rk_pack_outgoing_particles(rk_world* w, int32_t tile, void* stream)
{
  if(int rc = check_tile(w, tile)) return rc;
  if(int rc = check_particles(w)) return rc;
  if(!w->mig[0].meta || !w->mig[1].meta) return fail(RK_EINVAL, "migration buffers not set");
  hipStream_t s = S(stream);
  {
    ProfScope ps(w, 3, s);
    if(int rc = launch_pack(w, tile, s)) return rc;
  }
  // brick changes inside the tile are settled at once
  {
    ProfScope ps(w, 4, s);
    if(int rc = launch_append(w, 1, s)) return rc;
    if(int rc = launch_reset_counter(w, C_MIG1, s)) return rc;
  }
  if(int rc = check_lost(w, s)) return rc;
  uint32_t ovf = 0;
  if(int rc = read_counter(w, C_OVF, ovf, s)) return rc;
  if(ovf) return fail(RK_EOVERFLOW, "particle bins full");
  return RK_OK;
}

int rk_sort_particles(rk_world* w, int32_t tile, void* stream)
{
  if(int rc = check_tile(w, tile)) return rc;
  if(int rc = check_particles(w)) return rc;
  ProfScope ps(w, 8, S(stream));
  return launch_sort(w, tile, S(stream));
}

int rk_world_set_fused_launch(rk_world* w, int32_t mode)
{
  if(!w) return fail(RK_EINVAL, "null world");
  if(mode != 0 && mode != 1) return fail(RK_EINVAL, "fused launch: 0 (one workgroup per brick) or 1 (persistent queue)");
  w->fused_queue = mode;
  return RK_OK;
}

int rk_world_set_fused_workgroup(rk_world* w, int32_t threads)
{
  if(!w) return fail(RK_EINVAL, "null world");
  if(threads != 256 && threads != 512) return fail(RK_EINVAL, "fused workgroup: 256 or 512 threads");
  w->fused_wg = threads;
  return RK_OK;
}

int rk_particle_step_fused(rk_world* w, void* stream) { return rk_particle_step_fused_walls(w, nullptr, 0, stream); }

namespace {
int fused_step(rk_world* w, const rk_reflector_wall* walls, int32_t n_walls, bool device_walls, void* stream);
}

int rk_particle_step_fused_walls(rk_world* w, const rk_reflector_wall* walls, int32_t n_walls, void* stream)
{
  if(!w) return fail(RK_EINVAL, "null world");
  if(n_walls < 0 || n_walls > MAX_WALLS) return fail(RK_EINVAL, "at most 8 reflector walls");
  if(n_walls > 0 && !walls) return fail(RK_EINVAL, "null wall array");
  return fused_step(w, walls, n_walls, false, stream);
}

int rk_particle_step_fused_device_walls(rk_world* w, void* stream)
{
  if(!w) return fail(RK_EINVAL, "null world");
  return fused_step(w, nullptr, w->walls_dev_n, true, stream);
}

}  // extern "C"

namespace {
int fused_step(rk_world* w, const rk_reflector_wall* walls, int32_t n_walls, bool device_walls, void* stream)
{
  if(int rc = check_tile(w, -1)) return rc;
  if(int rc = check_particles(w)) return rc;
  if(!w->g.periodic) return fail(RK_EINVAL, "the fused particle step needs a periodic tile grid");
  if(!w->mbox.meta || !w->mig[0].meta || !w->mig[1].meta || !w->jstage)
    return fail(RK_EINVAL, "fused particle step: mailbox, migration or current staging buffers not set");
  if(w->filter_open) return fail(RK_ESTATE, "particle step during an open filter pass");
  WallSet ws {};
  if(device_walls) ws.n = n_walls;  // the kernel reads the walls from device memory
  else
    for(int i = 0; i < n_walls; ++i) ws.w[ws.n++] = walls[i];
  if(ws.n > 0 && !device_walls)
    if(int rc = mark_wall_tiles(w, ws)) return rc;
  hipStream_t s = S(stream);
  {
    ProfScope ps(w, 8, s);
    if(int rc = launch_fill(w->J, w->g.nodes(), 0.0f, s)) return rc;
  }
  w->j_accumulating = true;
  {
    ProfScope ps(w, 0, s);
    if(int rc = launch_fused(w, s, &ws, device_walls ? w->walls_dev : nullptr)) return rc;
  }
  {
    ProfScope ps(w, 2, s);
    if(int rc = launch_current_gather(w, s)) return rc;
    if(int rc = apply_corrections(w, -1, s)) return rc;
    if(device_walls && ws.n > 0)
      if(int rc = add_corrections_world(w, s)) return rc;
  }
  ProfScope ps(w, 4, s);
  if(int rc = launch_append_mailbox(w, s)) return rc;
  return launch_append_records(w, s);
}
}  // namespace

extern "C" {

int rk_comm_local(rk_world* w, int32_t mode, void* stream)
{
  if(int rc = check_tile(w, -1)) return rc;
  if(!w->g.periodic) return fail(RK_EINVAL, "comm_local needs a tile grid (stand-alone tile)");
  hipStream_t s = S(stream);
  ProfScope ps(w, mode == RK_COMM_PIC_PARTICLE ? 4 : 7, s);
  switch(mode) {
    case RK_COMM_EMF_E: return launch_ghost_copy(w, w->E, s);
    case RK_COMM_EMF_B: return launch_ghost_copy(w, w->B, s);
    case RK_COMM_EMF_J: close_j(w); return launch_ghost_copy(w, w->J, s);
    case RK_COMM_EMF_J_EXCHANGE: close_j(w); return launch_ghost_fold(w, w->J, s);
    case RK_COMM_PIC_PARTICLE: {
      if(int rc = check_particles(w)) return rc;
      if(int rc = launch_append(w, 0, s)) return rc;
      if(int rc = launch_reset_counter(w, C_MIG0, s)) return rc;
      if(int rc = check_lost(w, s)) return rc;
      uint32_t ovf = 0;
      if(int rc = read_counter(w, C_OVF, ovf, s)) return rc;
      if(ovf) return fail(RK_EOVERFLOW, "particle bins full");
      return RK_OK;
    }
    default:
      return fail(RK_EINVAL, "pic::Tile::local_communication does not support given communication mode: " +
                               std::to_string(mode));
  }
}

int rk_inject(rk_world* w, int32_t tile, int32_t s, int64_t n, const float* const pos[3], const float* const vel[3],
              const uint64_t* id, void* stream)
{
  if(int rc = check_tile(w, tile)) return rc;
  if(tile < 0) return fail(RK_EINVAL, "inject needs one tile");
  if(s < 0 || s >= w->d.n_species) return fail(RK_EINVAL, "invalid species");
  if(int rc = check_particles(w)) return rc;
  if(int rc = launch_inject(w, tile, s, n, pos, vel, id, S(stream))) return rc;
  uint32_t ovf = 0;
  if(int rc = read_counter(w, C_OVF, ovf, S(stream))) return rc;
  if(ovf) return fail(RK_EOVERFLOW, "particle bins full");
  return RK_OK;
}

int rk_pending_overflow(rk_world* w, int64_t* n_pending, void* stream)
{
  if(!w || !n_pending) return fail(RK_EINVAL, "null argument");
  if(int rc = check_lost(w, S(stream))) return rc;
  uint32_t v = 0;
  if(int rc = read_counter(w, C_OVF, v, S(stream))) return rc;
  *n_pending = v;
  return RK_OK;
}

int rk_bin_demand(rk_world* w, int32_t s, uint32_t* demand_dev, void* stream)
{
  if(!w || s < 0 || s >= w->d.n_species) return fail(RK_EINVAL, "invalid species");
  RK_HIP(hipMemcpyAsync(demand_dev, w->sp[s].bin_count, sizeof(uint32_t) * size_t(w->g.nbricks),
                        hipMemcpyDeviceToDevice, S(stream)));
  return RK_OK;
}

int rk_repack(rk_world* w, int32_t s, int64_t new_capacity, float* const pos[3], float* const vel[3], uint64_t* id,
              int64_t* bin_start, uint32_t* bin_cap, void* stream)
{
  if(!w || s < 0 || s >= w->d.n_species) return fail(RK_EINVAL, "invalid species");
  Species n = w->sp[s];
  n.capacity = new_capacity;
  for(int d = 0; d < 3; ++d) {
    n.pos[d] = pos[d];
    n.vel[d] = vel[d];
  }
  n.id = id;
  n.bin_start = bin_start;
  n.bin_cap = bin_cap;
  if(int rc = launch_repack(w, s, n, S(stream))) return rc;
  w->sp[s] = n;
  return RK_OK;
}

int rk_move_bins(rk_world* w, int32_t s, int32_t n, const int32_t* bins, const int64_t* new_start,
                 const uint32_t* new_cap, void* stream)
{
  if(!w || s < 0 || s >= w->d.n_species) return fail(RK_EINVAL, "invalid species");
  if(n < 0 || (n > 0 && (!bins || !new_start || !new_cap))) return fail(RK_EINVAL, "invalid bin list");
  return launch_move_bins(w, s, n, bins, new_start, new_cap, S(stream));
}

int rk_relayout(rk_world* w, int32_t s, const uint32_t* new_cap, void* tmp, int64_t tmp_bytes, int64_t* new_start,
                void* stream)
{
  if(!w || s < 0 || s >= w->d.n_species) return fail(RK_EINVAL, "invalid species");
  if(!new_cap || !new_start) return fail(RK_EINVAL, "null argument");
  return relayout_species(w, s, new_cap, tmp, tmp_bytes, new_start, S(stream));
}

int rk_retry_append(rk_world* w, void* stream)
{
  if(!w) return fail(RK_EINVAL, "null world");
  if(int rc = launch_retry(w, S(stream))) return rc;
  uint32_t ovf = 0;
  if(int rc = read_counter(w, C_OVF, ovf, S(stream))) return rc;
  if(ovf) return fail(RK_EOVERFLOW, "particle bins still full after repack");
  return RK_OK;
}

int rk_tile_particle_count(rk_world* w, int32_t tile, int32_t s, int64_t* count_out, void* stream)
{
  if(int rc = check_tile(w, tile)) return rc;
  if(s < 0 || s >= w->d.n_species) return fail(RK_EINVAL, "invalid species");
  const int b0 = tile < 0 ? 0 : tile * w->g.bpt;
  const int nb = tile < 0 ? w->g.nbricks : w->g.bpt;
  std::vector<uint32_t> c(static_cast<size_t>(nb));
  RK_HIP(hipMemcpyAsync(c.data(), w->sp[s].bin_count + b0, sizeof(uint32_t) * size_t(nb), hipMemcpyDeviceToHost,
                        S(stream)));
  RK_HIP(hipStreamSynchronize(S(stream)));
  int64_t t = 0;
  for(auto v : c) t += v;
  *count_out = t;
  return RK_OK;
}

int rk_tile_particles(rk_world* w, int32_t tile, int32_t s, float* const pos[3], float* const vel[3], uint64_t* id,
                      int64_t capacity, int64_t* count_out, void* stream)
{
  if(int rc = check_tile(w, tile)) return rc;
  if(tile < 0) return fail(RK_EINVAL, "readback needs one tile");
  if(s < 0 || s >= w->d.n_species) return fail(RK_EINVAL, "invalid species");
  hipStream_t st = S(stream);
  const int nb = w->g.bpt;
  std::vector<uint32_t> c(static_cast<size_t>(nb));
  RK_HIP(hipMemcpyAsync(c.data(), w->sp[s].bin_count + size_t(tile) * nb, sizeof(uint32_t) * size_t(nb),
                        hipMemcpyDeviceToHost, st));
  RK_HIP(hipStreamSynchronize(st));
  std::vector<int64_t> off(size_t(nb) + 1, 0);
  for(int b = 0; b < nb; ++b) off[size_t(b) + 1] = off[size_t(b)] + c[size_t(b)];
  *count_out = off[size_t(nb)];
  if(off[size_t(nb)] > capacity) return fail(RK_EINVAL, "readback buffer too small");
  if(off[size_t(nb)] == 0) return RK_OK;
  int64_t* doff = nullptr;
  RK_HIP(hipMallocAsync(reinterpret_cast<void**>(&doff), sizeof(int64_t) * off.size(), st));
  RK_HIP(hipMemcpyAsync(doff, off.data(), sizeof(int64_t) * off.size(), hipMemcpyHostToDevice, st));
  int rc = launch_gather_tile(w, tile, s, pos, vel, id, doff, st);
  RK_HIP(hipFreeAsync(doff, st));
  RK_HIP(hipStreamSynchronize(st));
  return rc;
}

int rk_outgoing(rk_world* w, float* const pos[3], float* const vel[3], uint64_t* id, int32_t* src_tile, int32_t* dir,
                int32_t* species, int64_t capacity, int64_t* count_out, void* stream)
{
  if(!w) return fail(RK_EINVAL, "null world");
  hipStream_t st = S(stream);
  uint32_t n = 0;
  if(int rc = read_counter(w, C_MIG0, n, st)) return rc;
  n = uint32_t(std::min<int64_t>(n, w->mig[0].capacity));
  *count_out = n;
  if(int64_t(n) > capacity) return fail(RK_EINVAL, "outgoing buffer too small");
  if(n == 0) return RK_OK;
  const auto& m = w->mig[0];
  for(int d = 0; d < 3; ++d) {
    RK_HIP(hipMemcpyAsync(pos[d], m.pos[d], sizeof(float) * n, hipMemcpyDeviceToDevice, st));
    RK_HIP(hipMemcpyAsync(vel[d], m.vel[d], sizeof(float) * n, hipMemcpyDeviceToDevice, st));
  }
  RK_HIP(hipMemcpyAsync(id, m.id, sizeof(uint64_t) * n, hipMemcpyDeviceToDevice, st));
  std::vector<uint32_t> meta(n);
  RK_HIP(hipMemcpyAsync(meta.data(), m.meta, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, st));
  RK_HIP(hipStreamSynchronize(st));
  for(uint32_t r = 0; r < n; ++r) {
    src_tile[r] = int32_t(meta_bin(meta[r]) / uint32_t(w->g.bpt));
    dir[r] = meta_dir(meta[r]);
    species[r] = meta_species(meta[r]);
  }
  return RK_OK;
}

int rk_energies(rk_world* w, int32_t tile, double* out, void* stream)
{
  if(int rc = check_tile(w, tile)) return rc;
  int r0[3], r1[3];
  interior(w, tile, r0, r1);
  double fe[2];
  if(int rc = field_energies(w, r0, r1, fe, S(stream))) return rc;
  out[0] = fe[0];
  out[1] = fe[1];
  if(w->d.n_species > 0 && w->sp[0].bin_start) return kinetic_energies(w, tile, out + 2, S(stream));
  for(int s = 0; s < w->d.n_species; ++s) out[2 + s] = 0.0;
  return RK_OK;
}

int rk_energies_device(rk_world* w, int32_t tile, double* out_dev, void* stream)
{
  if(int rc = check_tile(w, tile)) return rc;
  if(!out_dev) return fail(RK_EINVAL, "rk_energies_device: null output");
  if(!w->energy_scratch) return fail(RK_ESTATE, "rk_energies_device: energy scratch not resolved");
  int r0[3], r1[3];
  interior(w, tile, r0, r1);
  const bool empty = r1[0] <= r0[0] || r1[1] <= r0[1] || r1[2] <= r0[2];
  const int planes = empty ? 0 : r1[0] - r0[0];  // no partials are written for an empty region
  if(planes > ENERGY_PLANES)
    return fail(RK_EINVAL, "rk_energies_device: the region has " + std::to_string(planes) +
                             " x planes; the static partial-sum array holds at most " +
                             std::to_string(ENERGY_PLANES));
  double* part = w->energy_scratch;
  const hipStream_t st = S(stream);
  if(int rc = launch_field_energy_partials(w, r0, r1, part, st)) return rc;
  const bool kin = w->d.n_species > 0 && w->sp[0].bin_start;
  const int kin_blocks = kin ? kinetic_partial_count(w, tile) : 0;
  if(kin_blocks > ENERGY_KIN) return fail(RK_EINVAL, "rk_energies_device: too many kinetic partials");
  for(int s = 0; kin && s < w->d.n_species; ++s)
    if(int rc = launch_kinetic_partials(w, tile, s, part + 2 * ENERGY_PLANES + s * ENERGY_KIN, st)) return rc;
  hipLaunchKernelGGL(k_energy_finish, dim3(1), dim3(64), 0, st, part, planes, kin_blocks, w->d.n_species, out_dev);
  RK_HIP(hipGetLastError());
  return RK_OK;
}

int rk_gauss_residual(rk_world* w, int32_t filter_passes, double* out, void* stream)
{
  if(int rc = check_tile(w, -1)) return rc;
  if(int rc = check_particles(w)) return rc;
  return gauss_residual(w, filter_passes, out, S(stream));
}

int rk_error_counters(rk_world* w, int64_t* out2, void* stream)
{
  if(!w) return fail(RK_EINVAL, "null world");
  uint32_t a = 0, b = 0;
  if(int rc = read_counter(w, C_STENCIL, a, S(stream))) return rc;
  if(int rc = read_counter(w, C_MIGLOST, b, S(stream))) return rc;
  out2[0] = a;
  out2[1] = b;
  return RK_OK;
}

int rk_profile_enable(rk_world* w, int32_t enable)
{
  if(!w) return fail(RK_EINVAL, "null world");
  w->prof_on = enable != 0;
  return RK_OK;
}

int rk_profile_read(rk_world* w, double* ms, int64_t* launches)
{
  if(!w) return fail(RK_EINVAL, "null world");
  for(auto& r : w->prof) {
    RK_HIP(hipEventSynchronize(r.b));
    float t = 0.0f;
    RK_HIP(hipEventElapsedTime(&t, r.a, r.b));
    w->prof_ms[r.cat] += double(t);
    w->prof_n[r.cat] += 1;
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  w->prof.clear();
  for(int c = 0; c < RK_PROF_NCAT; ++c) {
    if(ms) ms[c] = w->prof_ms[c];
    if(launches) launches[c] = w->prof_n[c];
    w->prof_ms[c] = 0;
    w->prof_n[c] = 0;
  }
  return RK_OK;
}

const void* rk_world_j_pointer(const rk_world* w, int32_t comp) { return w ? w->J[comp] : nullptr; }

}  // extern "C"
