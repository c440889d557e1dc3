// Internal definitions shared by the runko_amd HIP translation units.
//
// Data layout in HBM (one world = the whole tile grid, or one stand-alone tile):
//  * Fields: one haloed lattice of L = N+6 nodes per dimension (N = n_tiles *
//    n_cells), one planar float array per component, index (i*Ly + j)*Lz + k.
//    Haloed index i <-> global coordinate gmin + i - 3.  A tile's own haloed
//    lattice (runko emf/yee_lattice.h:67-108) is the window starting at
//    tile*n_cells, so its interior is the world interior of that tile and its
//    halo is the neighbouring tiles' interior (or the world ghost shell).
//  * Particles: per species one SoA pool (x,y,z,ux,uy,uz float, id uint64)
//    split into bins; a bin is one brick of the tile (brick = sub-block of
//    cells that fits a workgroup's LDS).  Bins are numbered tile-major, then
//    layout_right inside the tile; bin b occupies [start[b], start[b]+count[b])
//    with capacity cap[b].  Bins are always compact (no dead slots).
//  * Migration records: SoA (pos, vel, id, meta), set 0 = tile leavers waiting
//    for comm_local(pic_particle), set 1 = brick changes inside a tile.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/runko_amd.h"

namespace rk {

constexpr int HALO = 3;
constexpr uint64_t DEAD_ID = UINT64_MAX;
constexpr int MAX_HOLES = 2048;  // swap-remove hole list per workgroup (LDS)

// meta word of a migration record
//   bits 0-1 species, bit 2 kind (0 tile leaver, 1 brick change), bits 3-7 dir, bits 8-31 source bin
__host__ __device__ inline uint32_t meta_pack(int s, int kind, int dir, uint32_t bin)
{
  return uint32_t(s) | (uint32_t(kind) << 2) | (uint32_t(dir) << 3) | (bin << 8);
}
__host__ __device__ inline int meta_species(uint32_t m) { return int(m & 3u); }
__host__ __device__ inline int meta_kind(uint32_t m) { return int((m >> 2) & 1u); }
__host__ __device__ inline int meta_dir(uint32_t m) { return int((m >> 3) & 31u); }
__host__ __device__ inline uint32_t meta_bin(uint32_t m) { return m >> 8; }

// Geometry passed by value to kernels.
struct Geo {
  int n[3];       // cells per tile
  int nt[3];      // tiles in the world
  int N[3];       // world interior cells
  int L[3];       // haloed extents N+6
  int br[3];      // brick shape (cells)
  int nb[3];      // bricks per tile per dimension
  int bpt;        // bricks per tile
  int nbricks;    // bricks in the world
  int toff[3];    // grid index of world tile (0,0,0)
  int periodic;
  float wmin[3];  // periodic wrap box of the full grid (global coordinates)
  float wmax[3];
  float cfl;
  float rcp_cfl;      // RN(1/cfl)
  int fast_div;       // 1: x/cfl as q = x*r; q += fma(-q, cfl, x)*r for 2^-100 <= |x| <= 2^100
  int fast_sqrt;      // 1: cfl/sqrt(X), 2/t, 1/sqrt(t) from hardware sqrt/rcp estimates with exact
                      //    corrections on their ranges (both flags set only when the forms were
                      //    verified bit-identical to the IEEE operations for every float of the
                      //    range at world creation)
  double jscale;      // fixed-point scale of the fused current windows: 2^30 / (max|q| rounded up to 2^k)
  double jinv;        // 1 / jscale (exact)
  uint32_t br_magic[3];  // ceil(2^32 / br) for br > 1: c / br == umulhi(c, magic) for 0 <= c < 2^16
  int pusher;            // 0 boris, 1 higuera_cary, 2 faraday (pic/tile.c++:74-86)

  __host__ __device__ int div_br(int c, int d) const
  {
#ifdef __HIP_DEVICE_COMPILE__
    return br[d] == 1 ? c : int(__umulhi(uint32_t(c), br_magic[d]));
#else
    return c / br[d];
#endif
  }

  __host__ __device__ int64_t nodes() const { return int64_t(L[0]) * L[1] * L[2]; }
  __host__ __device__ uint32_t at(int i, int j, int k) const
  {
    return (uint32_t(i) * uint32_t(L[1]) + uint32_t(j)) * uint32_t(L[2]) + uint32_t(k);
  }
  __host__ __device__ int ntiles() const { return nt[0] * nt[1] * nt[2]; }
  __host__ __device__ void tile_coords(int t, int c[3]) const
  {
    c[2] = t % nt[2];
    c[1] = (t / nt[2]) % nt[1];
    c[0] = t / (nt[1] * nt[2]);
  }
  __host__ __device__ int tile_lin(int a, int b, int c) const { return (a * nt[1] + b) * nt[2] + c; }
  // float(tile mins) of world tile t, dimension d (exact: integers < 2^24)
  __host__ __device__ float tile_min(int tc, int d) const { return float((toff[d] + tc) * n[d]); }
  // brick b -> tile, brick coords inside the tile
  __host__ __device__ void brick_coords(int b, int& tile, int lb[3]) const
  {
    tile = b / bpt;
    int r = b % bpt;
    lb[2] = r % nb[2];
    lb[1] = (r / nb[2]) % nb[1];
    lb[0] = r / (nb[1] * nb[2]);
  }
};

struct Species {
  int64_t capacity = 0;
  float* pos[3] = {nullptr, nullptr, nullptr};
  float* vel[3] = {nullptr, nullptr, nullptr};
  uint64_t* id = nullptr;
  int64_t* bin_start = nullptr;
  uint32_t* bin_cap = nullptr;
  uint32_t* bin_count = nullptr;
  float qm = 0;      // (float)(sign(q)/m), particle_boris.h:26-27
  float charge = 0;  // (float)q, particle_current_zigzag_1st.c++:238
};

struct MigSet {
  int64_t capacity = 0;
  float* pos[3] = {nullptr, nullptr, nullptr};
  float* vel[3] = {nullptr, nullptr, nullptr};
  uint64_t* id = nullptr;
  uint32_t* meta = nullptr;
};

// device counters (uint32 unless noted), in one scratch allocation
enum Counter {
  C_MIG0 = 0,        // records in set 0
  C_MIG1 = 1,        // records in set 1
  C_OVF = 2,         // overflowed records (indices in ovf_list)
  C_MIGLOST = 3,     // records that did not fit the record buffer (particle kept in place)
  C_STENCIL = 4,     // particles whose stencil left the haloed lattice
  C_SLOWCOMPACT = 5, // bins compacted on the slow path
  C_APPDONE = 6,     // workgroups done in k_append_records (the last one clears C_MIG0/1 and itself)
  C_QUEUE = 8,       // 8 per-XCD brick queues of the persistent fused step
  C_NUM = 16
};

}  // namespace rk

struct rk_world {
  rk_world_desc d {};
  rk::Geo g {};
  float* E[3] {};
  float* B[3] {};
  float* J[3] {};
  float* Jf[3] {};
  rk::Species sp[RK_MAX_SPECIES];
  rk::MigSet mig[2];
  // per-brick current windows of the fused step: brick b writes its
  // 3 x (br+3)^3 partial current at jstage + b*jstage_stride, a gather kernel
  // sums the overlapping windows into J
  float* jstage = nullptr;
  int64_t jstage_stride = 0;
  // per-brick mailboxes of the fused particle step: brick b owns records
  // [b*mbox_per_brick, (b+1)*mbox_per_brick) of mbox, mbox_count[b] used
  rk::MigSet mbox;
  int64_t mbox_per_brick = 0;
  uint32_t* mbox_count = nullptr;
  void* mig_table = nullptr;  // device copy of the record-set pointers (upload_mig_table)
  // reflector-wall correction current (world lattice, 3 components), added to
  // J and cleared by the next deposit of each pending tile
  float* corrJ[3] = {nullptr, nullptr, nullptr};
  // B propagator: extended stencil (emf/yee_lattice_stencil.c++) when set,
  // [axis][row][column] coefficient matrices with M[0][0] normalised
  bool stencil_on = false;
  int fused_queue = 0;  // fused step launch: 0 one workgroup per brick, 1 persistent queue
  int fused_wg = 512;   // fused step threads per workgroup (256 for Boris 8^3 bricks, else 512)
  float stencil[3][3][5] = {};
  std::vector<char> corr_pending;
  rk_reflector_wall* walls_dev = nullptr;  // walls held on the device (rk_world_set_walls)
  int walls_dev_n = 0;
  int corr_any = 0;
  // library scratch
  uint32_t* counters = nullptr;   // rk::C_NUM
  uint32_t* ovf_list = nullptr;   // overflowed record refs: (set << 30) | index, set 2 = mailbox
  int64_t ovf_cap = 0;
  double* red = nullptr;          // reduction partials
  int64_t red_cap = 0;
  double* energy_scratch = nullptr;  // this device's static partial-sum array (rk_energies_device)
  uint32_t* host_counters = nullptr;  // pinned mirror
  // kernel timing (rk_profile_*)
  bool prof_on = false;
  struct ProfRec {
    int cat;
    hipEvent_t a, b;
  };
  std::vector<ProfRec> prof;
  double prof_ms[RK_PROF_NCAT] = {};
  int64_t prof_n[RK_PROF_NCAT] = {};
  // host-side call-sequence state
  bool j_accumulating = false;  // a deposit generation is open (J holds partial sums)
  bool filter_open = false;
  int filter_done_count = 0;
  std::vector<uint8_t> filter_done;
};

namespace rk {

void set_error(const std::string& msg);

// Brackets the launches issued in its lifetime with events when timing is on.
struct ProfScope {
  rk_world* w;
  int cat;
  hipStream_t s;
  hipEvent_t a = nullptr;
  ProfScope(rk_world* w_, int cat_, hipStream_t s_) : w(w_), cat(cat_), s(s_)
  {
    if(w->prof_on && hipEventCreate(&a) == hipSuccess) (void)hipEventRecord(a, s);
    else a = nullptr;
  }
  ~ProfScope()
  {
    if(!a) return;
    hipEvent_t b = nullptr;
    if(hipEventCreate(&b) == hipSuccess) {
      (void)hipEventRecord(b, s);
      w->prof.push_back({cat, a, b});
    } else {
      (void)hipEventDestroy(a);
    }
  }
};
int hip_fail(hipError_t e, const char* where);

#define RK_HIP(call)                                       \
  do {                                                     \
    hipError_t e_ = (call);                                \
    if(e_ != hipSuccess) return rk::hip_fail(e_, #call);   \
  } while(0)

// launchers implemented in rk_fields.hip / rk_particles.hip
int launch_curl_b(rk_world* w, const int r0[3], const int r1[3], float dt, hipStream_t s);
int launch_curl_b_stencil(rk_world* w, const int r0[3], const int r1[3], float dt, hipStream_t s);
int launch_curl_e(rk_world* w, const int r0[3], const int r1[3], float dt, bool add_j, hipStream_t s);
int launch_add_current(rk_world* w, const int r0[3], const int r1[3], hipStream_t s);
int launch_filter(rk_world* w, const int r0[3], const int r1[3], hipStream_t s);
int launch_ghost_copy(rk_world* w, float* const F[3], hipStream_t s);
int launch_ghost_fold(rk_world* w, float* const F[3], hipStream_t s);
int launch_fill(float* const F[3], int64_t n, float v, hipStream_t s);
int field_energies(rk_world* w, const int r0[3], const int r1[3], double out[2], hipStream_t s);

int launch_push(rk_world* w, int tile, hipStream_t s);
int launch_deposit(rk_world* w, int tile, hipStream_t s);
int launch_pack(rk_world* w, int tile, hipStream_t s);
int launch_append(rk_world* w, int set, hipStream_t s);
int launch_append_mailbox(rk_world* w, hipStream_t s);
int launch_append_records(rk_world* w, hipStream_t s);  // sets 0 and 1, then clears their counters
int launch_current_gather(rk_world* w, hipStream_t s);
int verify_fast_division(float cfl, float rcp, uint64_t mismatches[2]);
int upload_mig_table(rk_world* w);
int launch_reset_counter(rk_world* w, int which, hipStream_t s);
int apply_corrections(rk_world* w, int tile, hipStream_t s);  // rk_wall.hip
struct WallSet;                                               // rk_reflect.h
int mark_wall_tiles(rk_world* w, const WallSet& ws);          // rk_wall.hip
int launch_sort(rk_world* w, int tile, hipStream_t s);
int launch_fused(rk_world* w, hipStream_t s, const WallSet* walls = nullptr,
                 const rk_reflector_wall* wdev = nullptr);
int add_corrections_world(rk_world* w, hipStream_t s);      // rk_wall.hip
int kinetic_energies(rk_world* w, int tile, double* out, hipStream_t s);
int launch_inject(rk_world* w, int tile, int s, int64_t n, const float* const pos[3],
                  const float* const vel[3], const uint64_t* id, hipStream_t st);
int launch_retry(rk_world* w, hipStream_t s);
int launch_repack(rk_world* w, int s, const rk::Species& nsp, hipStream_t st);
int launch_move_bins(rk_world* w, int s, int n, const int32_t* bins, const int64_t* new_start,
                     const uint32_t* new_cap, hipStream_t st);
// three binomial2 passes J -> Jf over the interior of a periodic world (J ghosts current)
int launch_filter_x3(rk_world* w, hipStream_t s);
int relayout_species(rk_world* w, int s, const uint32_t* new_cap, void* tmp, int64_t tmp_bytes,
                     int64_t* new_start, hipStream_t st);
int launch_gather_tile(rk_world* w, int tile, int s, float* const pos[3], float* const vel[3], uint64_t* id,
                       const int64_t* offsets_dev, hipStream_t st);
int gauss_residual(rk_world* w, int passes, double* out, hipStream_t s);
// energy partial sums into a caller buffer (rk_energies and its allocation-free form)
int launch_field_energy_partials(rk_world* w, const int r0[3], const int r1[3], double* partial, hipStream_t s);
int kinetic_partial_count(rk_world* w, int tile);
int launch_kinetic_partials(rk_world* w, int tile, int s, double* partial, hipStream_t st);

}  // namespace rk
