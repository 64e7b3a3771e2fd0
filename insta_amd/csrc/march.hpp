// march.hpp -- ray/box, occupancy-grid stepping, Sobol jitter and
// wave-aggregated list helpers shared by the render and training marches
// (instant-ngp/src/testbed_nerf.cu:186-208, 313-337, 634-686;
// random_val.cuh:159-288).
#pragma once

#include "context.hpp"

namespace insta {

struct RenderArgs {
    float M[12];
    float cx, cy, fx, fy;
    float off_x, off_y;
    uint32_t width, height, spp, G, max_steps;
    float dt, min_transmittance, density_bias;
    float lo[3], hi[3];
    const uint8_t* grid;
};

// ---- Sobol / Owen scrambling (random_val.cuh:159-288) ---------------------
INSTA_DEV uint32_t lk_permutation(uint32_t x, uint32_t seed) {
    x += seed;
    x ^= x * 0x6c50b47cu;
    x ^= x * 0xb82f1e52u;
    x ^= x * 0xc7afe638u;
    x ^= x * 0x8d22f6e6u;
    return x;
}
INSTA_DEV uint32_t owen_scramble(uint32_t x, uint32_t seed) {
    return __builtin_bitreverse32(lk_permutation(__builtin_bitreverse32(x), seed));
}
INSTA_DEV uint32_t hash_combine(uint32_t seed, uint32_t v) { return seed ^ (v + (seed << 6) + (seed >> 2)); }
// dimension 0 of the Sobol sequence is the bit-reversed index
INSTA_DEV float ld_random_val0(uint32_t index, uint32_t seed) {
    const float S = (float)(1.0 / 4294967296.0);
    index = owen_scramble(index, seed);
    return (float)owen_scramble(__builtin_bitreverse32(index), hash_combine(seed, 0u)) * S;
}

// ---- march helpers (testbed_nerf.cu:186-208, 313-337) --------------------
INSTA_DEV bool box_contains(const RenderArgs& a, V3 p) {
    return p.x >= a.lo[0] && p.x <= a.hi[0] && p.y >= a.lo[1] && p.y <= a.hi[1] && p.z >= a.lo[2] && p.z <= a.hi[2];
}

// density_grid_occupied_at for one cascade (testbed_nerf.cu:313-337), read
// from the linear-order copy of the Morton bitfield (same bits).
INSTA_DEV bool occupied(const uint8_t* __restrict__ grid, uint32_t G, V3 pos) {
    const float g = (float)G;
    const float px = ((pos.x - 0.5f) * 1.0f + 0.5f) * g;
    const float py = ((pos.y - 0.5f) * 1.0f + 0.5f) * g;
    const float pz = ((pos.z - 0.5f) * 1.0f + 0.5f) * g;
    int ix = (int)px, iy = (int)py, iz = (int)pz;
    const int hi = (int)G - 1;
    ix = ix < 0 ? 0 : (ix > hi ? hi : ix);
    iy = iy < 0 ? 0 : (iy > hi ? hi : iy);
    iz = iz < 0 ? 0 : (iz > hi ? hi : iz);
    const uint32_t idx = (uint32_t)ix + G * ((uint32_t)iy + G * (uint32_t)iz);
    return (grid[idx >> 3] >> (idx & 7u)) & 1u;
}

INSTA_DEV float advance_to_next_voxel(float t, float dt, V3 pos, V3 dir, V3 idir, uint32_t res) {
    const float r = (float)res;
    const V3 p = v3(r * pos.x, r * pos.y, r * pos.z);
    const float tx = (floorf(p.x + 0.5f + 0.5f * signf(dir.x)) - p.x) * idir.x;
    const float ty = (floorf(p.y + 0.5f + 0.5f * signf(dir.y)) - p.y) * idir.y;
    const float tz = (floorf(p.z + 0.5f + 0.5f * signf(dir.z)) - p.z) * idir.z;
    const float tn = fminf(fminf(tx, ty), tz);
    const float t_target = t + fmaxf(tn / r, 0.0f);
    do {
        t += dt;
    } while (t < t_target);
    return t;
}

// Walk from t to the next occupied position inside the box.  Returns false
// if the ray leaves the box first.
INSTA_DEV bool skip_empty(const RenderArgs& a, V3 o, V3 d, V3 idir, float& t, V3& pos) {
    while (true) {
        pos = o + d * t;
        if (!box_contains(a, pos)) return false;
        if (occupied(a.grid, a.G, pos)) return true;
        t = advance_to_next_voxel(t, a.dt, pos, d, idir, a.G);
    }
}

INSTA_DEV float box_ray_tmin(const RenderArgs& a, V3 o, V3 d) {
    float tmin = (a.lo[0] - o.x) / d.x, tmax = (a.hi[0] - o.x) / d.x;
    if (tmin > tmax) { float s = tmin; tmin = tmax; tmax = s; }
    float tymin = (a.lo[1] - o.y) / d.y, tymax = (a.hi[1] - o.y) / d.y;
    if (tymin > tymax) { float s = tymin; tymin = tymax; tymax = s; }
    if (tmin > tymax || tymin > tmax) return 3.40282347e38f;
    if (tymin > tmin) tmin = tymin;
    if (tymax < tmax) tmax = tymax;
    float tzmin = (a.lo[2] - o.z) / d.z, tzmax = (a.hi[2] - o.z) / d.z;
    if (tzmin > tzmax) { float s = tzmin; tzmin = tzmax; tzmax = s; }
    if (tmin > tzmax || tzmin > tmax) return 3.40282347e38f;
    if (tzmin > tmin) tmin = tzmin;
    return tmin;
}

// wave-aggregated append: returns this lane's slot (only meaningful if `want`)
INSTA_DEV uint32_t wave_append(uint32_t* counter, bool want) {
    const uint64_t mask = __ballot(want);
    if (mask == 0) return 0;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t leader = __ffsll((unsigned long long)mask) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(mask));
    base = __shfl(base, leader);
    const uint64_t below = mask & ((1ull << lane) - 1ull);
    return base + (uint32_t)__popcll(below);
}

// wave-wide exclusive prefix sum of `v` plus one atomic reservation
INSTA_DEV uint32_t wave_reserve(uint32_t* counter, uint32_t v) {
    const uint32_t lane = threadIdx.x & 63u;
    uint32_t incl = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t up = __shfl_up(incl, o);
        if (lane >= (uint32_t)o) incl += up;
    }
    const uint32_t total = __shfl(incl, 63);
    uint32_t base = 0;
    if (lane == 63 && total > 0) base = atomicAdd(counter, total);
    base = __shfl(base, 63);
    return base + incl - v;
}

INSTA_DEV uint32_t wave_max(uint32_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, (uint32_t)__shfl_xor(v, o));
    return v;
}

// Step-major order within a wave's sample range: sample j of lane l sits at
// wave_base + (samples of steps < j over the wave's lanes) + (lanes below l
// with more than j samples).  Lanes are adjacent rays, so neighbouring
// samples are the same step of neighbouring pixels.  Every lane of the wave
// calls this with the same j sequence (j < the wave's largest count).
struct StepMajor {
    uint32_t running = 0;
    INSTA_DEV uint32_t next(uint32_t j, uint32_t count, bool& mine) {
        const uint64_t act = __ballot(j < count);
        const uint32_t lane = threadIdx.x & 63u;
        const uint64_t below = act & ((1ull << lane) - 1ull);
        const uint32_t slot = running + (uint32_t)__popcll(below);
        running += (uint32_t)__popcll(act);
        mine = j < count;
        return slot;
    }
};

}  // namespace insta
