// Reflecting wall, per particle: the reflection of one pushed particle
// through an ordered set of walls and the correction current of every
// reflection, shared by the per-tile reflect kernel and the fused particle
// step.
//
// Reference semantics followed (hel-astro-lab/runko v5.0.0):
//   reflect_at_wall     src/runko/pic/reflector_wall.c++:125-237
//   correction current  src/runko/pic/reflector_wall.c++:36-118
//
// Every expression is evaluated in the order the reference writes it (the
// library builds with -ffp-contract=off), so reflected positions and
// velocities are bit-identical to the CPU restatement wherever this is
// inlined.  The correction currents go to a world-lattice buffer with float
// atomics (the reference also adds them with atomics).
#pragma once

#include "rk_internal.h"

namespace rk {

constexpr int MAX_WALLS = 8;

struct WallSet {
  rk_reflector_wall w[MAX_WALLS];
  int n;
};

__device__ inline float dot3w(const float a[3], const float b[3])
{
  float r = a[0] * b[0];
  r = r + a[1] * b[1];
  r = r + a[2] * b[2];
  return r;
}

// zigzag deposit of one sub-trajectory x1 -> x2 (tile-lattice coordinates)
// into the correction buffer, reflector_wall.c++:36-118
__device__ inline void zigzag_single(const Geo& g, float* const Jc[3], const int toffh[3], const float x1[3],
                                     const float x2[3], float charge, uint32_t* err)
{
  const float one = 1.0f;
  float fi1[3], fi2[3], xr[3], F1[3], F2[3], W1[3], W2[3];
#pragma unroll
  for(int d = 0; d < 3; ++d) {
    fi1[d] = floorf(x1[d]);
    fi2[d] = floorf(x2[d]);
  }
#pragma unroll
  for(int d = 0; d < 3; ++d) {
    const float a = fminf(fi1[d], fi2[d]) + one;
    const float b = fmaxf(fmaxf(fi1[d], fi2[d]), 0.5f * (x1[d] + x2[d]));
    xr[d] = fminf(a, b);
  }
#pragma unroll
  for(int d = 0; d < 3; ++d) {
    F1[d] = charge * (xr[d] - x1[d]);
    F2[d] = charge * (x2[d] - xr[d]);
    W1[d] = 0.5f * (x1[d] + xr[d]) - fi1[d];
    W2[d] = 0.5f * (x2[d] + xr[d]) - fi2[d];
  }
  const int a1 = int(uint32_t(fi1[0])), b1 = int(uint32_t(fi1[1])), c1 = int(uint32_t(fi1[2]));
  const int a2 = int(uint32_t(fi2[0])), b2 = int(uint32_t(fi2[1])), c2 = int(uint32_t(fi2[2]));
  auto st = [&](int i, int j, int k, float jx, float jy, float jz) {
    const int I = toffh[0] + i, J = toffh[1] + j, K = toffh[2] + k;
    if(I < 0 || J < 0 || K < 0 || I >= g.L[0] || J >= g.L[1] || K >= g.L[2]) {
      atomicAdd(err, 1u);
      return;
    }
    const uint32_t q = g.at(I, J, K);
    unsafeAtomicAdd(&Jc[0][q], jx);
    unsafeAtomicAdd(&Jc[1][q], jy);
    unsafeAtomicAdd(&Jc[2][q], jz);
  };
  const float Fx1 = F1[0], Fy1 = F1[1], Fz1 = F1[2], Fx2 = F2[0], Fy2 = F2[1], Fz2 = F2[2];
  const float Wx1 = W1[0], Wy1 = W1[1], Wz1 = W1[2], Wx2 = W2[0], Wy2 = W2[1], Wz2 = W2[2];
  st(a1, b1, c1, Fx1 * (one - Wy1) * (one - Wz1), Fy1 * (one - Wx1) * (one - Wz1), Fz1 * (one - Wx1) * (one - Wy1));
  st(a2, b2, c2, Fx2 * (one - Wy2) * (one - Wz2), Fy2 * (one - Wx2) * (one - Wz2), Fz2 * (one - Wx2) * (one - Wy2));
  st(a1 + 1, b1, c1, 0.0f, Fy1 * Wx1 * (one - Wz1), Fz1 * Wx1 * (one - Wy1));
  st(a2 + 1, b2, c2, 0.0f, Fy2 * Wx2 * (one - Wz2), Fz2 * Wx2 * (one - Wy2));
  st(a1, b1 + 1, c1, Fx1 * Wy1 * (one - Wz1), 0.0f, Fz1 * (one - Wx1) * Wy1);
  st(a2, b2 + 1, c2, Fx2 * Wy2 * (one - Wz2), 0.0f, Fz2 * (one - Wx2) * Wy2);
  st(a1, b1, c1 + 1, Fx1 * (one - Wy1) * Wz1, Fy1 * (one - Wx1) * Wz1, 0.0f);
  st(a2, b2, c2 + 1, Fx2 * (one - Wy2) * Wz2, Fy2 * (one - Wx2) * Wz2, 0.0f);
  st(a1, b1 + 1, c1 + 1, Fx1 * Wy1 * Wz1, 0.0f, 0.0f);
  st(a2, b2 + 1, c2 + 1, Fx2 * Wy2 * Wz2, 0.0f, 0.0f);
  st(a1 + 1, b1, c1 + 1, 0.0f, Fy1 * Wx1 * Wz1, 0.0f);
  st(a2 + 1, b2, c2 + 1, 0.0f, Fy2 * Wx2 * Wz2, 0.0f);
  st(a1 + 1, b1 + 1, c1, 0.0f, 0.0f, Fz1 * Wx1 * Wy1);
  st(a2 + 1, b2 + 1, c2, 0.0f, 0.0f, Fz2 * Wx2 * Wy2);
}

// wall_is_in_tile, reflector_wall.c++:261-264: bit i set when wall i lies in
// [tmin - cfl, tmax] along x, tmin and tmax bounding the tile
__host__ __device__ inline unsigned walls_in_tile(const WallSet& ws, float tmin, float tmax, float cfl)
{
  unsigned m = 0;
  for(int i = 0; i < ws.n; ++i)
    if(ws.w[i].walloc >= tmin - cfl && ws.w[i].walloc <= tmax) m |= 1u << i;
  return m;
}

// the largest wall position: particles at or beyond it meet no wall
__device__ inline float walls_max_x(const WallSet& ws)
{
  float m = -INFINITY;
  for(int i = 0; i < ws.n; ++i) m = fmaxf(m, ws.w[i].walloc);
  return m;
}

// One pushed particle (global position pos, 4-velocity u) through the walls
// of `mask` in registration order.  origo: the tile lattice origo (tile mins
// - 3); toffh: the world haloed index of the tile's haloed index 0.  Returns
// false when a wall parks the particle (behind the wall without crossing it
// this step): the caller removes it.
__device__ inline bool reflect_particle(const Geo& g, const WallSet& walls, unsigned mask, float cfl,
                                        const float origo[3], const int toffh[3], float* const Jc[3], float charge,
                                        float pos[3], float u[3], uint32_t* err)
{
  const float EPS = 1e-10f;
  const float one = 1.0f;
  const float c = cfl;
  for(int wi = 0; wi < walls.n; ++wi) {
    if(!((mask >> wi) & 1u)) continue;
    const float walloc = walls.w[wi].walloc, betawall = walls.w[wi].betawall, gammawall = walls.w[wi].gammawall;
    if(pos[0] >= walloc) continue;  // in front of the wall: untouched (all masks zero)
    const float walloc0 = walloc - betawall * c;
    const float pos1[3] = {pos[0], pos[1], pos[2]};
    const float gam = sqrtf(one + dot3w(u, u));
    const float invgam = one / gam;
    float pos0[3];
#pragma unroll
    for(int d = 0; d < 3; ++d) pos0[d] = pos1[d] - c * invgam * u[d];
    const float mask_close = (walloc0 - pos0[0] <= c) ? one : 0.0f;
    const float denom = betawall * c - c * u[0] * invgam;
    const float dt = fabsf((pos0[0] - walloc0) / (denom + EPS));
    const float mask_crossed = (dt <= one) ? one : 0.0f;
    const float mask_refl = mask_close * mask_crossed;
    const float mask_park = one - mask_refl;
    if(mask_refl > 0.0f) {
      float pcol[3];
#pragma unroll
      for(int d = 0; d < 3; ++d) pcol[d] = pos0[d] + c * dt * invgam * u[d];
      const float ux_new =
        gammawall * gammawall * gam * (2.0f * betawall - u[0] * invgam * (one + betawall * betawall));
      const float un[3] = {ux_new, u[1], u[2]};
      const float gam_new = sqrtf(one + dot3w(un, un));
      const float invgam_new = one / gam_new;
      const float dt_refl = fminf(one, fabsf((pos1[0] - pcol[0]) / (pos1[0] - pos0[0] + EPS)));
      float prefl[3], p1l[3], f0[3], f1[3], r0[3], r1[3];
#pragma unroll
      for(int d = 0; d < 3; ++d) {
        prefl[d] = pcol[d] + c * dt_refl * invgam_new * un[d];
        p1l[d] = pos1[d] - origo[d];
      }
#pragma unroll
      for(int d = 0; d < 3; ++d) {
        f0[d] = p1l[d] + mask_refl * (pos0[d] - pos1[d]);
        f1[d] = p1l[d] + mask_refl * (pcol[d] - pos1[d]);
      }
      zigzag_single(g, Jc, toffh, f0, f1, mask_refl * charge, err);
#pragma unroll
      for(int d = 0; d < 3; ++d) {
        const float x1d = prefl[d] - c * invgam_new * un[d];
        r0[d] = p1l[d] + mask_refl * (x1d - pos1[d]);
        r1[d] = p1l[d] + mask_refl * (pcol[d] - pos1[d]);
      }
      zigzag_single(g, Jc, toffh, r0, r1, mask_refl * (-charge), err);
#pragma unroll
      for(int d = 0; d < 3; ++d) pos[d] = (one - mask_refl) * pos1[d] + mask_refl * prefl[d];
      u[0] = (one - mask_refl) * u[0] + mask_refl * ux_new;
    }
    if(mask_park > 0.5f) return false;
  }
  return true;
}

}  // namespace rk
