// encoding.hpp -- hash-grid encoding, spherical-harmonics direction
// encoding and the bias-free dense layer, shared by the inference and
// training kernels.
//
// Hash grid convention (the reference delegates it to tiny-cuda-nn, absent
// from its tree; Instant-NGP, Mueller et al. 2022): level l uses scale s_l =
// b^l * N_min - 1 and resolution ceil(s_l) + 1; position x maps to
// x * s_l + 0.5, whose floor/fraction give the cell and trilinear weights;
// dense levels index x + y r + z r^2, hashed levels
// (x * 1 ^ y * 2654435761 ^ z * 805459861) mod T.
#pragma once

#include "context.hpp"

namespace insta {

// table entry of corner cg on a level: dense levels index x + y r + z r^2
// (wrapped into the level's table), hashed levels use the spatial hash
INSTA_DEV uint32_t grid_index(const uint32_t cg[3], bool dense, uint32_t res, uint32_t size) {
    if (dense) {
        uint32_t idx = cg[0] + cg[1] * res + cg[2] * (res * res);
        if (idx >= size) idx %= size;
        return idx;
    }
    return (cg[0] ^ (cg[1] * 2654435761u) ^ (cg[2] * 805459861u)) & (size - 1u);
}

template <int F>
INSTA_DEV void encode_level(const float* __restrict__ table, const GridParams& gp, uint32_t l, V3 x, float acc[F]) {
    const float sc = gp.scale[l];
    const float px = x.x * sc + 0.5f, py = x.y * sc + 0.5f, pz = x.z * sc + 0.5f;
    const float fx = floorf(px), fy = floorf(py), fz = floorf(pz);
    const float frac[3] = {px - fx, py - fy, pz - fz};
    const uint32_t g[3] = {(uint32_t)(int32_t)fx, (uint32_t)(int32_t)fy, (uint32_t)(int32_t)fz};
    const bool dense = (gp.dense_mask >> l) & 1u;
    const uint32_t res = gp.res[l], size = gp.size[l];
    const float* base = table + (size_t)gp.offset[l] * F;
#pragma unroll
    for (int f = 0; f < F; ++f) acc[f] = 0.0f;
#pragma unroll
    for (uint32_t c = 0; c < 8; ++c) {
        float w = 1.0f;
        uint32_t cg[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const bool up = (c >> d) & 1u;
            w = w * (up ? frac[d] : 1.0f - frac[d]);
            cg[d] = g[d] + (up ? 1u : 0u);
        }
        const uint32_t idx = grid_index(cg, dense, res, size);
        const float* e = base + (size_t)idx * F;
        if constexpr (F == 2) {
            const float2 v = *reinterpret_cast<const float2*>(e);
            acc[0] = acc[0] + w * v.x;
            acc[1] = acc[1] + w * v.y;
        } else if constexpr (F == 4) {
            const float4 v = *reinterpret_cast<const float4*>(e);
            acc[0] = acc[0] + w * v.x;
            acc[1] = acc[1] + w * v.y;
            acc[2] = acc[2] + w * v.z;
            acc[3] = acc[3] + w * v.w;
        } else {
#pragma unroll
            for (int f = 0; f < F; ++f) acc[f] = acc[f] + w * e[f];
        }
    }
}

// real SH bands 0..3 of a direction in [-1,1]^3 (see the oracle for the
// normalisation constants' closed forms)
INSTA_DEV void sh16(float x, float y, float z, float* o) {
    const float xy = x * y, xz = x * z, yz = y * z, x2 = x * x, y2 = y * y, z2 = z * z;
    o[0] = 0.28209479177387814f;
    o[1] = -0.48860251190291987f * y;
    o[2] = 0.48860251190291987f * z;
    o[3] = -0.48860251190291987f * x;
    o[4] = 1.0925484305920792f * xy;
    o[5] = -1.0925484305920792f * yz;
    o[6] = 0.94617469575755997f * z2 - 0.31539156525251999f;
    o[7] = -1.0925484305920792f * xz;
    o[8] = 0.54627421529603959f * x2 - 0.54627421529603959f * y2;
    o[9] = 0.59004358992664352f * y * (-3.0f * x2 + y2);
    o[10] = 2.8906114426405538f * xy * z;
    o[11] = 0.45704579946446572f * y * (1.0f - 5.0f * z2);
    o[12] = 0.3731763325901154f * z * (5.0f * z2 - 3.0f);
    o[13] = 0.45704579946446572f * x * (1.0f - 5.0f * z2);
    o[14] = 1.4453057213202769f * z * (x2 - y2);
    o[15] = 0.59004358992664352f * x * (-x2 + 3.0f * y2);
}

// bias-free dense layer, row-major W[out][in]; weights are wave-uniform and
// reach the FMAs as scalar operands
template <int NIN, int NOUT, bool RELU>
INSTA_DEV void dense(const float* __restrict__ W, const float* x, float* y) {
#pragma unroll
    for (int o = 0; o < NOUT; ++o) {
        float acc = 0.0f;
#pragma unroll
        for (int i = 0; i < NIN; ++i) acc = __builtin_fmaf(W[o * NIN + i], x[i], acc);
        y[o] = RELU ? (acc > 0.0f ? acc : 0.0f) : acc;
    }
}

}  // namespace insta
