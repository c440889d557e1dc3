/* runko_amd -- MI355X-native (gfx950) PIC time-step hot path behind runko's
 * pic/emf tile-method API.
 *
 * C ABI only: plain pointers, sizes and status codes; no C++ or torch types.
 * Device buffers (fields, particle pools, migration records) are allocated by
 * the caller (the Python layer allocates them as torch tensors) and borrowed by
 * the library through the rk_world_set_* calls; the library owns only small
 * scratch (counters, reduction partials, sort staging).
 *
 * Every entry point returns 0 on success or a nonzero RK_E* code; the message
 * of the last failure on the calling thread is returned by rk_last_error().
 * Calls enqueue work on the given HIP stream (hipStream_t passed as void*) and
 * return without waiting, except where a comment says "synchronous".
 *
 * Tile arguments: tile >= 0 names one tile (layout_right index over n_tiles,
 * as the reference's tile tag); tile == -1 runs one launch over every tile
 * (the batched form).
 *
 * Reference interfaces replaced (hel-astro-lab/runko v5.0.0):
 *   pic::Tile<3> methods   src/runko/bindings/pypic.c++:105-136
 *   emf::Tile<3> methods   src/runko/bindings/pyemf.c++:225-260
 *   comm_mode enum         src/runko/communication_common.h:30-38
 *   corgi local_communication (pycorgi.Grid.local_communication), called from
 *                          runko/simulation.py:292-293
 */
#ifndef RUNKO_AMD_H
#define RUNKO_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RK_MAX_SPECIES 4

enum rk_status {
  RK_OK = 0,
  RK_EINVAL = 1,     /* invalid argument / precondition (Python: RuntimeError) */
  RK_EHIP = 2,       /* HIP runtime error */
  RK_EOVERFLOW = 3,  /* a particle bin or the migration buffer is full: repack */
  RK_ESTATE = 4      /* call sequence not allowed (e.g. filter twice in a pass) */
};

/* comm_mode values, src/runko/communication_common.h:30-38 */
enum rk_comm_mode {
  RK_COMM_EMF_J = 0,
  RK_COMM_EMF_E = 1,
  RK_COMM_EMF_B = 2,
  RK_COMM_PIC_PARTICLE = 3,
  RK_COMM_PIC_PARTICLE_EXTRA = 4,
  RK_COMM_NUMBER_OF_PARTICLES = 5,
  RK_COMM_EMF_J_EXCHANGE = 6
};

typedef struct rk_world rk_world;

/* Geometry and physics constants of one world.  A world is either the whole
 * tile grid (periodic = 1) or one stand-alone tile (periodic = 0, n_tiles =
 * {1,1,1}, tile_offset = its grid index).  Fields live on one haloed lattice of
 * (N+6)^3 nodes per component, N = n_tiles * n_cells (the reference's per-tile
 * YeeLattice, emf/yee_lattice.h:67-108, scaled to the whole domain). */
typedef struct rk_world_desc {
  int32_t n_cells[3];      /* cells per tile                      (config n_cells_per_tile) */
  int32_t n_tiles[3];      /* tiles in this world                  (config n_tiles or 1,1,1) */
  int32_t tile_offset[3];  /* grid index of this world's tile (0,0,0) */
  int32_t grid_tiles[3];   /* tiles of the full grid (tile tags and periodic wrap box) */
  int32_t brick[3];        /* particle bin shape in cells; divides n_cells */
  int32_t periodic;        /* 1: comm_local allowed (periodic ghost sync) */
  int32_t n_species;
  int32_t current_filter;  /* 0 binomial2 (27-tap), 1 binomial2_unrolled */
  double cfl;
  double charge[RK_MAX_SPECIES];
  double mass[RK_MAX_SPECIES];
} rk_world_desc;

const char* rk_last_error(void);
int rk_version(void);

/* Creates a world (synchronous).  Allocates library scratch only. */
int rk_world_create(rk_world** out, const rk_world_desc* desc);
int rk_world_destroy(rk_world* w);

/* Number of haloed lattice nodes per component, and particle bins (species x bricks). */
int64_t rk_world_field_nodes(const rk_world* w);
int32_t rk_world_num_bricks(const rk_world* w);

/* Borrow the field arrays: E, B, J, Jf (filter target), each 3 device
 * pointers of rk_world_field_nodes() floats, layout [x][y][z], z fastest. */
int rk_world_set_fields(rk_world* w, float* const E[3], float* const B[3], float* const J[3],
                        float* const Jf[3]);

/* Device pointer of component comp of the buffer that currently holds J
 * (filter passes write the other buffer and swap the two). */
const void* rk_world_j_pointer(const rk_world* w, int32_t comp);

/* Borrow the particle pool of one species.  pos/vel: 3 device pointers each of
 * capacity floats; id: capacity uint64.  bin_start (int64), bin_cap and
 * bin_count (uint32) have rk_world_num_bricks() entries (device memory). */
int rk_world_set_species(rk_world* w, int32_t s, int64_t capacity, float* const pos[3],
                         float* const vel[3], uint64_t* id, int64_t* bin_start, uint32_t* bin_cap,
                         uint32_t* bin_count);

/* Borrow the migration record buffer (capacity records): pos[3], vel[3] float,
 * id uint64, meta uint32 (species | kind | direction | source bin).  Two
 * record sets are used: set 0 holds tile leavers (consumed by comm_local
 * pic_particle), set 1 holds in-tile brick changes (consumed at once). */
int rk_world_set_migration(rk_world* w, int32_t set, int64_t capacity, float* const pos[3],
                           float* const vel[3], uint64_t* id, uint32_t* meta);

/* Borrow the per-brick mailboxes of the fused particle step: per_brick records
 * for every brick (same SoA record layout as above).  A brick writes its
 * leavers at consecutive offsets of its own mailbox, so the fused step needs
 * no global atomics to place them; leavers past per_brick spill into the
 * record sets. */
int rk_world_set_mailbox(rk_world* w, int64_t per_brick, float* const pos[3], float* const vel[3],
                         uint64_t* id, uint32_t* meta);

/* Borrow the current staging buffer of the fused particle step: one slot of
 * stride floats per brick (stride >= 3 * (br+3)^3 for brick shape br), where
 * every brick leaves its current window; a gather pass then sums the windows
 * into J.  Required by rk_particle_step_fused. */
int rk_world_set_current_staging(rk_world* w, float* buf, int64_t stride);

/* Numerical schemes of every tile (the configuration keys particle_pusher and
 * field_propagator, src/runko/pic/tile.c++:74-86,337-350 and
 * src/runko/emf/tile.c++:29-141,363-372): particle_pusher 0 boris,
 * 1 higuera_cary (particle_higuera_cary.h:15-77), 2 faraday
 * (particle_faraday.h:37-109); stencil = null for the fdtd2 B push, else 45
 * floats [axis x,y,z][row 0..2][column 0..4], the StencilAxisCoeffs matrices
 * with M[0][0] normalised (stencil_coefficients.h:42-60), for push_b_stencil
 * (emf/yee_lattice_stencil.c++:17-297).  The fused particle step runs every
 * pusher, with or without reflecting walls. */
int rk_world_set_schemes(rk_world* w, int32_t particle_pusher, const float* stencil);

/* Reflecting wall (conducting piston) normal to x, pic::reflector_wall
 * (src/runko/pic/reflector_wall.h:13-25): position, velocity beta, Lorentz factor. */
typedef struct rk_reflector_wall {
  float walloc;
  float betawall;
  float gammawall;
} rk_reflector_wall;

/* pic::Tile::reflect_particles for the given walls (the tile's registered
 * walls, in registration order; src/runko/pic/reflector_wall.c++:125-298;
 * tile -1: every tile).  At most 8 walls.
 * Particles behind a wall are reflected or, when they did not cross it this
 * step, removed; the +q / -q correction currents are kept per world and added
 * to J by the next deposit_current of each affected tile (tile.c++:411-414). */
int rk_reflect_particles(rk_world* w, int32_t tile, const rk_reflector_wall* walls, int32_t n_walls,
                         void* stream);

/* emf::edge_bc (src/runko/emf/edge_bc.h:22-38): set field components in the
 * region left (side 0) or right (side 1) of a global position along one axis. */
typedef struct rk_edge_bc {
  int32_t direction;  /* 0 x, 1 y, 2 z */
  int32_t side;       /* 0 left of position, 1 right */
  float position;
  float E[3], B[3], J[3];
  int32_t E_components, B_components, J_components;  /* bit c: set component c */
} rk_edge_bc;

/* emf::Tile::apply_edge_bc for comm_mode emf_E, emf_B or emf_J
 * (src/runko/emf/tile.c++:808-847, yee_lattice.c++:263-306; tile -1: every
 * tile).  Sets the region on the tile's interior cells and on the part of its
 * haloed region that lies in the world's outer halo; halo cells inside the
 * world are neighbour cells that the same width rule puts in the region. */
int rk_apply_edge_bc(rk_world* w, int32_t tile, const rk_edge_bc* bc, int32_t mode, void* stream);

/* ---- emf::Tile methods (src/runko/emf/tile.c++:358-426) ---- */
int rk_push_half_b(rk_world* w, int32_t tile, void* stream);   /* B += cfl/2 curl E, fdtd2 */
int rk_push_e(rk_world* w, int32_t tile, void* stream);        /* E += cfl curl B, fdtd2 */
int rk_add_current(rk_world* w, int32_t tile, void* stream);   /* E -= J */
int rk_filter_current(rk_world* w, int32_t tile, void* stream);/* binomial (1,2,1)^3 pass */

/* Batched form of the driver's filter sequence (projects/pic-turbulence/pic.py:208-212:
 * filter_current, comm_local(emf_J), filter_current, filter_current on every
 * tile): `passes` filter passes over every tile, starting from a current J
 * ghost layer (comm_local(emf_J) after the deposit), with the J ghost layer
 * synced between passes.  On a periodic world with three binomial2 passes
 * the three run as one launch that reads J and writes the result once; the
 * interior equals the passes run one by one bit for bit.  On return the J
 * ghost layer of a periodic world holds the periodic images of the filtered
 * interior. */
int rk_filter_current_passes(rk_world* w, int32_t passes, void* stream);

/* ---- pic::Tile methods (src/runko/pic/tile.c++:324-438, tile_communication.c++:66-96) ---- */
int rk_push_particles(rk_world* w, int32_t tile, void* stream);  /* interpolation + Boris */
int rk_deposit_current(rk_world* w, int32_t tile, void* stream); /* zigzag, first order */
int rk_pack_outgoing_particles(rk_world* w, int32_t tile, void* stream);
int rk_sort_particles(rk_world* w, int32_t tile, void* stream);

/* Fused batched particle phase: interpolation + push (the world's pusher:
 * Boris, Higuera-Cary or Faraday) + tile/brick classification + zigzag
 * deposit of every particle of the world in one launch, followed by the
 * migration append.  Equivalent to push_particles,
 * pack_outgoing_particles, comm_local(pic_particle), deposit_current on every
 * tile.  No counterpart in the reference (its driver issues those four). */
int rk_particle_step_fused(rk_world* w, void* stream);

/* How the fused particle step is launched: mode 0, one workgroup per brick
 * (bricks dealt to XCDs by whole tile-column groups); mode 1, persistent
 * workgroups (as many as stay resident) that take bricks from per-XCD
 * queues in the same order, for Boris with 8^3 bricks (other shapes and
 * pushers launch as mode 0).  Results do not depend on the mode.  No
 * counterpart in the reference. */
int rk_world_set_fused_launch(rk_world* w, int32_t mode);

/* Threads per workgroup of the fused particle step: 512 (default) or 256; 256
 * applies to Boris with 8^3 bricks launched one workgroup per brick (other
 * cases use 512).  Results do not depend on it.  No counterpart in the
 * reference. */
int rk_world_set_fused_workgroup(rk_world* w, int32_t threads);

/* The fused particle phase with reflecting walls: the tile methods
 * push_particles, reflect_particles (every wall in registration order on the
 * tiles it lies in, src/runko/pic/reflector_wall.c++:253-289),
 * pack_outgoing_particles, comm_local(pic_particle), deposit_current of the
 * reference's shock lap (projects/pic-shock/pic.py:259-270) in one launch;
 * the correction currents of the reflections are added to J before it
 * returns.  n_walls = 0 is rk_particle_step_fused.  At most 8 walls.  No
 * counterpart in the reference (its driver issues those five). */
int rk_particle_step_fused_walls(rk_world* w, const rk_reflector_wall* walls, int32_t n_walls, void* stream);

/* Walls held in device memory, so that a moving wall can advance inside a
 * captured HIP graph: rk_world_set_walls copies the walls to the device
 * (synchronous; not under capture), rk_particle_step_fused_device_walls is
 * rk_particle_step_fused_walls with those walls (its correction currents are
 * added over the whole lattice, as the host does not track which tiles the
 * walls are in), rk_advance_walls is Tile::advance_reflector_walls
 * (src/runko/pic/reflector_wall.c++:293-300) for them as a kernel, and
 * rk_read_walls copies them back (synchronous).  At most 8 walls. */
int rk_world_set_walls(rk_world* w, const rk_reflector_wall* walls, int32_t n_walls);
int rk_particle_step_fused_device_walls(rk_world* w, void* stream);
int rk_advance_walls(rk_world* w, void* stream);
int rk_read_walls(rk_world* w, rk_reflector_wall* out, int32_t n_walls, void* stream);

/* E += cfl curl B followed by E -= J in one pass (same two roundings per
 * node as rk_push_e then rk_add_current).  Batched form of that pair. */
int rk_push_e_add_current(rk_world* w, int32_t tile, void* stream);

/* Kernel timing: when enabled, every launch is bracketed by HIP events on its
 * stream, accumulated per category; rk_profile_read (synchronous) returns and
 * clears the totals.  Categories: 0 fused particle step kernel, 1 push,
 * 2 deposit, 3 pack, 4 migration append, 5 fdtd curl (+E -= J), 6 filter,
 * 7 ghost copy/fold, 8 other (fills, counters, sort).  Not for use while a
 * stream is being captured into a graph. */
#define RK_PROF_NCAT 9
int rk_profile_enable(rk_world* w, int32_t enable);
int rk_profile_read(rk_world* w, double* ms, int64_t* launches);

/* corgi local communication for one comm_mode (runko/simulation.py:292-293).
 * Fields: periodic ghost copy (E, B, J) or ghost fold-add (J_exchange).
 * pic_particle: append every pending tile-leaver record to its destination
 * bin with the global periodic wrap (src/runko/pic/particle.h:449-572). */
int rk_comm_local(rk_world* w, int32_t mode, void* stream);

/* Particle injection: n particles of species s (device pointers, float32 pos/vel,
 * uint64 ids) appended to the bins of tile by position.  May return
 * RK_EOVERFLOW: then the caller repacks (rk_repack) and calls rk_retry_append. */
int rk_inject(rk_world* w, int32_t tile, int32_t s, int64_t n, const float* const pos[3],
              const float* const vel[3], const uint64_t* id, void* stream);

/* Overflow handling (synchronous).  rk_pending_overflow reports how many
 * records are waiting because their bin was full, and fills per-bin demand
 * (device uint32[n_bricks] per species, counts clamped to capacity). */
int rk_pending_overflow(rk_world* w, int64_t* n_pending, void* stream);
int rk_bin_demand(rk_world* w, int32_t s, uint32_t* demand_dev, void* stream);
/* Copies every bin of species s from the borrowed pool into a new pool with
 * new bin starts/caps (device), then borrows the new pool. */
int rk_repack(rk_world* w, int32_t s, int64_t new_capacity, float* const pos[3], float* const vel[3],
              uint64_t* id, int64_t* bin_start, uint32_t* bin_cap, void* stream);
int rk_retry_append(rk_world* w, void* stream);
/* Moves n listed bins of species s (device int32 bin indices) to new starts
 * and capacities (device int64 / uint32, one per listed bin) inside the
 * borrowed pool; the targets must be disjoint from every bin (the pool's
 * spare tail).  Buffer addresses stay, so captured graphs stay valid.
 * Replaces the reference's container growth on append
 * (src/runko/pic/particle.h:471-509) for bricks that outgrow their bin. */
int rk_move_bins(rk_world* w, int32_t s, int32_t n, const int32_t* bins, const int64_t* new_start,
                 const uint32_t* new_cap, void* stream);
/* In-place relayout of species s (synchronous): every bin b gets capacity
 * new_cap[b] (host uint32[n_bricks], >= its stored count), bins packed in the
 * order of their current starts; *new_start (host int64[n_bricks]) receives
 * the new starts; every bin is copied once, through tmp.  tmp is a caller
 * device buffer of tmp_bytes (at least the tables, 28 B per brick, plus 32 B
 * per particle of the largest bin; more means fewer batches).  RK_EOVERFLOW
 * if the capacities exceed the pool.
 * Also the in-place counterpart of particle.h:471-509: no second pool. */
int rk_relayout(rk_world* w, int32_t s, const uint32_t* new_cap, void* tmp, int64_t tmp_bytes,
                int64_t* new_start, void* stream);

/* Copies of one tile's alive particles of species s in bin order into caller
 * device buffers (capacity >= count); *count_out set synchronously. */
int rk_tile_particle_count(rk_world* w, int32_t tile, int32_t s, int64_t* count_out, void* stream);
int rk_tile_particles(rk_world* w, int32_t tile, int32_t s, float* const pos[3], float* const vel[3],
                      uint64_t* id, int64_t capacity, int64_t* count_out, void* stream);

/* Pending tile-leaver records (after pack, before comm): count and copies
 * (pos, vel, id, source tile, direction index, species).  Synchronous. */
int rk_outgoing(rk_world* w, float* const pos[3], float* const vel[3], uint64_t* id, int32_t* src_tile,
                int32_t* dir, int32_t* species, int64_t capacity, int64_t* count_out, void* stream);

/* FP64 reductions over the tile (or world interior for tile = -1):
 * out[0] = sum(E^2)/2, out[1] = sum(B^2)/2, out[2+s] = sum_alive(gamma-1).
 * Synchronous.  src/runko/emf/yee_lattice.c++:383-428, src/runko/pic/particle.c++:351-377 */
int rk_energies(rk_world* w, int32_t tile, double* out, void* stream);
/* rk_energies without allocation, host copy or synchronisation: the same
 * values, bit for bit, written to out_dev (device double[2 + n_species]) on
 * `stream`.  No counterpart in the reference: it is for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc).  Its partial sums live in one 160 KiB
 * static device array per GPU (a region of at most 8192 x planes; RK_EINVAL
 * beyond), held until the process ends.  Calls that overlap on one GPU race
 * on that array. */
int rk_energies_device(rk_world* w, int32_t tile, double* out_dev, void* stream);

/* Gauss-law residual (FP64): rho = first-order (CIC) charge on the nodes where
 * div E is centred, optionally passed through the same number of binomial
 * passes as the current; out[0] = max|div E - rho|, out[1] = rms(rho),
 * out[2] = rms(div E).  Synchronous. */
int rk_gauss_residual(rk_world* w, int32_t filter_passes, double* out, void* stream);

/* Device error counters: particles whose stencil left the haloed lattice
 * (reference behaviour undefined; clamped here).  Synchronous. */
int rk_error_counters(rk_world* w, int64_t* out2, void* stream);

/* ---- Snapshot packing (the "RNKO" v3 output files) ----------------------
 * Device side of the reference's MPI-IO writers (src/runko/io/snapshots/,
 * bound at src/runko/bindings/pyemf.c++:271-296); the host writes the
 * 512-byte header (mpiio_header.h:55-103) and the arrays. */

/* mpiio::FieldsWriter::pack_tile for every tile (mpiio_fields.c++:218-321):
 * out = float [9 + min(nspecies, 5)][nz][ny][nx] in file order (z outermost),
 * n_axis = tiles_axis * max(1, cells_axis / stride): E and B at the stride
 * points, J summed over stride^3 cells, then per-species particle counts at
 * floor((x - tile mins) / stride).  1 <= stride <= cells per tile. */
int rk_snapshot_fields(rk_world* w, int32_t stride, int32_t nspecies, float* out, void* stream);

/* mpiio::SpectraWriter::histogram_tile for every tile (mpiio_spectra.c++:130-209):
 * out = float [4 * min(nspecies, 3)][tiles_z][tiles_y][tiles_x * nxt][nbins],
 * nxt = max(1, cells_x / stride): per species log10|u| bins on [umin, umax]
 * and beta_x, beta_y, beta_z bins on [-1, 1], values outside in the edge bins. */
int rk_snapshot_spectra(rk_world* w, int32_t stride, int32_t nbins, float umin, float umax, int32_t nspecies,
                        float* out, void* stream);

/* pic::Tile::EB_linear_1st_interpolator at n global positions of one tile
 * (emf/yee_lattice_interpolate_linear_1st.h:57-138, used by
 * mpiio_particles.c++:226-251): E and B written to the six output arrays. */
int rk_interpolate_at(rk_world* w, int32_t tile, int64_t n, const float* const pos[3], float* const E[3],
                      float* const B[3], void* stream);

/* ---- Differentiable periodic box --------------------------------------
 * No counterpart in the reference (runko has no differentiation).  One
 * periodic box of n[0] x n[1] x n[2] cells with the same discrete step as one
 * lap of projects/pic-turbulence/pic.py:187-221 (half B, push, deposit,
 * filter passes, half B, E, E -= J), without tiles: fields are planar
 * [3][n0][n1][n2] arrays (z fastest, Yee-staggered as the tile lattice),
 * particles one flat SoA array per coordinate holding species 0, 1, ... in
 * consecutive ranges, in a fixed order (no sort, no bins), so that a step's
 * adjoint can address every particle by its index.  Values are float
 * (precision 4) or double (precision 8). */
typedef struct rk_diff_desc {
  int32_t n[3];
  int32_t n_species;
  int64_t count[RK_MAX_SPECIES];  /* particles per species */
  double charge[RK_MAX_SPECIES];
  double mass[RK_MAX_SPECIES];
  double cfl;
  int32_t filter_passes;          /* binomial2 passes per step */
  int32_t precision;              /* 4 or 8 bytes per value */
} rk_diff_desc;

/* A state (or the adjoint of a state): E, B, positions, momenta. */
typedef struct rk_diff_state {
  void* E[3];
  void* B[3];
  void* x[3];
  void* u[3];
} rk_diff_state;

typedef struct rk_diff rk_diff;

int rk_diff_create(rk_diff** out, const rk_diff_desc* desc);

/* An ensemble of independent periodic boxes stepped together: every kernel of
 * a step covers all boxes in one launch.  No counterpart in the reference.
 * Boxes may differ in sides, species, counts, charges and masses; precision,
 * cfl and filter passes are common.  A state holds the boxes one after
 * another: the fields of box i start at cell sum_{j<i} n0 n1 n2 of box j (each
 * box [n0][n1][n2] inside), its particles at sum_{j<i} particles of box j.
 * rk_diff_step / rk_diff_step_adjoint / rk_diff_error_count / rk_diff_destroy
 * take the handle as for one box (one box = an ensemble of one). */
int rk_diff_create_ensemble(rk_diff** out, const rk_diff_desc* descs, int32_t n_boxes);
int rk_diff_destroy(rk_diff* d);

/* One step: out = step(in).  out may alias in. */
int rk_diff_step(rk_diff* d, const rk_diff_state* in, rk_diff_state* out, void* stream);

/* Adjoint of one step at state `in`: given the adjoint of the step's output
 * (adj_out), writes the adjoint of its input (adj_in), i.e. the
 * vector-Jacobian product adj_in = J(in)^T adj_out.  adj_in may alias adj_out. */
int rk_diff_step_adjoint(rk_diff* d, const rk_diff_state* in, const rk_diff_state* adj_out,
                         rk_diff_state* adj_in, void* stream);

/* Particles whose stencil found a position outside [-1, n+1] (skipped). */
int rk_diff_error_count(rk_diff* d, int64_t* out, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* RUNKO_AMD_H */
