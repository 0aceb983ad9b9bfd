// Device building blocks of the pose search: quaternion/angle helpers,
// Philox-4x32-10, scoring terms, grid sampling, energy+gradient of one pose
// per lane group, BFGS with Armijo line search.
//
// Execution model (gfx950, wave64): a workgroup is ONE wavefront of 64 lanes
// split into G = 64 / L groups of L lanes; each group owns one pose / chain,
// all groups of a workgroup work on the same ligand (its atoms and torsion
// tree are staged once in LDS).  Control flow is uniform across the
// wavefront: loops run while any group still needs them and per-group
// predicates decide which results are committed, so every __syncthreads()
// is reached by all lanes.
#pragma once
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include "ud_internal.h"

namespace ud {

static constexpr int kWG = 64;
static constexpr int kNMax = 6 + UD_MAX_TORSIONS;   // gradient length capacity
static constexpr int kPoseStride = 40;              // >= UD_POSE_FLOATS, LDS pose buffers
static constexpr float kPi = 3.14159265358979f;     // common.h:204 (fl pi)
static constexpr float kEpsFl = 1.1920928955078125e-07f;  // epsilon_fl (common.h:227-230)
static constexpr float kMaxFl = 3.402823466e+38f;

// ------------------------------------------------------------ lane groups
template <int L> __device__ __forceinline__ float gsum(float v) {
#pragma unroll
    for (int o = L / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWG);
    return v;
}
template <int L> __device__ __forceinline__ double gsum_d(double v) {
#pragma unroll
    for (int o = L / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWG);
    return v;
}
__device__ __forceinline__ bool wave_any(bool p) { return __ballot(p) != 0; }

// ------------------------------------------------------------ angles / quaternions
// normalize_angle (common.h:260-281)
__device__ __forceinline__ float normalize_angle(float x) {
    for (int it = 0; it < 8; ++it) {
        if (x > 3.f * kPi) {
            float n = (x - kPi) / (2.f * kPi);
            x -= 2.f * kPi * ceilf(n);
        } else if (x < -3.f * kPi) {
            float n = (-x - kPi) / (2.f * kPi);
            x += 2.f * kPi * ceilf(n);
        } else {
            if (x > kPi) x -= 2.f * kPi;
            else if (x < -kPi) x += 2.f * kPi;
            break;
        }
    }
    return x;
}
// Hamilton product p * q, (w, x, y, z)
__device__ __forceinline__ float4 qmul(float4 p, float4 q) {
    return make_float4(p.x * q.x - p.y * q.y - p.z * q.z - p.w * q.w,
                       p.x * q.y + p.y * q.x + p.z * q.w - p.w * q.z,
                       p.x * q.z - p.y * q.w + p.z * q.x + p.w * q.y,
                       p.x * q.w + p.y * q.z - p.z * q.y + p.w * q.x);
}
// components stored as float4 (x=w, y=x, z=y, w=z)
// quaternion_to_r3 (quaternion.cpp:77-114), row-major m[9]
__device__ __forceinline__ void q_to_mat(float4 q, float* m) {
    const float a = q.x, b = q.y, c = q.z, d = q.w;
    const float aa = a * a, ab = a * b, ac = a * c, ad = a * d, bb = b * b, bc = b * c, bd = b * d,
                cc = c * c, cd = c * d, dd = d * d;
    m[0] = aa + bb - cc - dd; m[1] = 2.f * (-ad + bc); m[2] = 2.f * (ac + bd);
    m[3] = 2.f * (ad + bc); m[4] = aa - bb + cc - dd; m[5] = 2.f * (-ab + cd);
    m[6] = 2.f * (-ac + bd); m[7] = 2.f * (ab + cd); m[8] = aa - bb - cc + dd;
}
// quaternion_normalize_approx (quaternion.h:105-117)
__device__ __forceinline__ float4 q_normalize_approx(float4 q) {
    const float s = q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w;
    if (fabsf(s - 1.f) < 1e-6f) return q;
    const float inv = 1.f / sqrtf(s);
    return make_float4(q.x * inv, q.y * inv, q.z * inv, q.w * inv);
}
// angle_to_quaternion(axis, angle) (quaternion.cpp:37-47)
__device__ __forceinline__ float4 axis_angle_q(float ax, float ay, float az, float angle) {
    angle = normalize_angle(angle);
    float s, c;
    sincosf(angle * 0.5f, &s, &c);
    return make_float4(c, s * ax, s * ay, s * az);
}
// quaternion_increment (quaternion.cpp:128-133) with angle_to_quaternion(rotation)
__device__ __forceinline__ float4 q_increment(float4 q, float rx, float ry, float rz) {
    const float angle = sqrtf(rx * rx + ry * ry + rz * rz);
    float4 r = make_float4(1.f, 0.f, 0.f, 0.f);
    if (angle > kEpsFl) r = axis_angle_q(rx / angle, ry / angle, rz / angle, angle);
    return q_normalize_approx(qmul(r, q));
}

// ------------------------------------------------------------ Philox-4x32-10
// Salmon et al., SC'11 ("Parallel random numbers: as easy as 1, 2, 3").
__host__ __device__ __forceinline__ uint4 philox4x32_10(uint4 c, uint2 k) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = uint64_t(0xD2511F53u) * c.x;
        const uint64_t p1 = uint64_t(0xCD9E8D57u) * c.z;
        const uint32_t hi0 = uint32_t(p0 >> 32), lo0 = uint32_t(p0);
        const uint32_t hi1 = uint32_t(p1 >> 32), lo1 = uint32_t(p1);
        c = make_uint4(hi1 ^ c.y ^ k.x, lo1, hi0 ^ c.w ^ k.y, lo0);
        k.x += 0x9E3779B9u;
        k.y += 0xBB67AE85u;
    }
    return c;
}
// 24-bit uniform in [0, 1)
__host__ __device__ __forceinline__ float u01(uint32_t x) { return float(x >> 8) * (1.0f / 16777216.0f); }

// ------------------------------------------------------------ scoring terms
// slope_step (potentials.h:30-39)
__device__ __forceinline__ float slope_step(float bad, float good, float x) {
    if (bad < good) {
        if (x <= bad) return 0.f;
        if (x >= good) return 1.f;
    } else {
        if (x >= bad) return 0.f;
        if (x <= good) return 1.f;
    }
    return (x - bad) / (good - bad);
}
__device__ __forceinline__ bool xs_hydrophobic(int t) { return t == 0 || t == 12 || t == 13 || t == 14 || t == 15; }
__device__ __forceinline__ bool xs_acceptor(int t) { return t == 4 || t == 5 || t == 8 || t == 9 || t == 32 || t == 33; }
__device__ __forceinline__ bool xs_donor(int t) { return t == 3 || t == 5 || t == 7 || t == 9 || t == 18 || t == 35 || t == 34; }
__device__ __forceinline__ bool xs_hbond(int a, int b) { return (xs_donor(a) && xs_acceptor(b)) || (xs_donor(b) && xs_acceptor(a)); }
__device__ __forceinline__ bool xs_glue(int t) { return t == 21 || t == 24 || t == 27 || t == 30; }
__device__ __forceinline__ bool xs_glued(int a, int b) {
    // potentials.h:106-126: G_k with C_H_CGk / C_P_CGk
    const int ga[4] = {21, 24, 27, 30}, ch[4] = {19, 22, 25, 28}, cp[4] = {20, 23, 26, 29};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (a == ga[k] && (b == ch[k] || b == cp[k])) return true;
        if (b == ga[k] && (a == ch[k] || a == cp[k])) return true;
    }
    return false;
}

}  // namespace ud
