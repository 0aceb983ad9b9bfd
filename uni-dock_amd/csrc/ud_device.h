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

#include <type_traits>

#include "ud_internal.h"

namespace ud {

static constexpr int kWG = 64;
static constexpr int kNMax = 6 + UD_MAX_TORSIONS;   // gradient length capacity
static constexpr float kPi = 3.14159265358979f;     // common.h:204 (fl pi)
static constexpr float kEpsFl = 1.1920928955078125e-07f;  // epsilon_fl (common.h:227-230)
static constexpr float kMaxFl = 3.402823466e+38f;

// ------------------------------------------------------------ lane groups
// Sums over aligned groups of lanes with DPP moves, which run in the vector
// ALU (a lane shuffle would go through the LDS pipe, the busiest unit of the
// search loop).  Within a 16-lane row: quad_perm [1,0,3,2] and [2,3,0,1],
// then half-row and row mirrors (after the quad step every lane of a quad
// holds the quad's sum, so mirrored partners always sit in the other half).
// Across rows: row_bcast:15 / row_bcast:31 accumulate row totals into the
// last lane of the group, which readlane broadcasts.  Every lane of the
// wavefront must be active.
enum : int {
    kDppQuadXor1 = 0xB1,     // quad_perm [1,0,3,2]
    kDppQuadXor2 = 0x4E,     // quad_perm [2,3,0,1]
    kDppRowHalfMirror = 0x141,
    kDppRowMirror = 0x140,
    kDppRowBcast15 = 0x142,
    kDppRowBcast31 = 0x143,
};
template <int CTRL, int ROWS = 0xF> __device__ __forceinline__ float dpp_f(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROWS, 0xF, false));
}
template <int CTRL, int ROWS = 0xF> __device__ __forceinline__ double dpp_d(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, int(b), CTRL, ROWS, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, int(b >> 32), CTRL, ROWS, 0xF, false);
    return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
}
__device__ __forceinline__ float readlane_f(float v, int lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}
__device__ __forceinline__ double readlane_d(double v, int lane) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane(int(b), lane);
    const int hi = __builtin_amdgcn_readlane(int(b >> 32), lane);
    return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
}

// sum over the S-lane subgroup (S in 1, 2, 4, ..., 64; runtime, wave-uniform)
template <class T> __device__ __forceinline__ T sub_sum(T v, int S) {
    auto mv = [](T x, auto ctrl_tag, auto rows_tag) -> T {
        constexpr int C = decltype(ctrl_tag)::value, R = decltype(rows_tag)::value;
        if constexpr (sizeof(T) == 8) return dpp_d<C, R>(x);
        else return dpp_f<C, R>(x);
    };
    using I1 = std::integral_constant<int, kDppQuadXor1>;
    using I2 = std::integral_constant<int, kDppQuadXor2>;
    using I4 = std::integral_constant<int, kDppRowHalfMirror>;
    using I8 = std::integral_constant<int, kDppRowMirror>;
    using B15 = std::integral_constant<int, kDppRowBcast15>;
    using B31 = std::integral_constant<int, kDppRowBcast31>;
    using RA = std::integral_constant<int, 0xF>;
    using R13 = std::integral_constant<int, 0xA>;
    using R23 = std::integral_constant<int, 0xC>;
    if (S >= 2) v += mv(v, I1{}, RA{});
    if (S >= 4) v += mv(v, I2{}, RA{});
    if (S >= 8) v += mv(v, I4{}, RA{});
    if (S >= 16) v += mv(v, I8{}, RA{});
    if (S >= 32) {
        v += mv(v, B15{}, R13{});  // rows 1 and 3 add the totals of rows 0 and 2
        if (S == 64) {
            v += mv(v, B31{}, R23{});  // rows 2 and 3 add lane 31
            if constexpr (sizeof(T) == 8) v = readlane_d(v, 63);
            else v = readlane_f(v, 63);
        } else {
            T a, b;
            if constexpr (sizeof(T) == 8) { a = readlane_d(v, 31); b = readlane_d(v, 63); }
            else { a = readlane_f(v, 31); b = readlane_f(v, 63); }
            v = (threadIdx.x < 32) ? a : b;
        }
    }
    return v;
}
// Inclusive prefix sum over the L-lane group (lane gl gets the sum of group
// lanes 0..gl): row_shr steps within 16-lane rows (lanes whose source lies
// before the row start receive 0; for groups narrower than a row the lane's
// group index masks the steps that would cross into the previous group), then
// row_bcast:15 / row_bcast:31 carry row totals into the rows above.
template <int K, int L> __device__ __forceinline__ float scan_step(float v, int gl) {
    const float t = dpp_f<0x110 + K>(v);  // row_shr:K
    return v + ((L >= 16 || gl >= K) ? t : 0.f);
}
template <int L> __device__ __forceinline__ float gscan(float v, int gl) {
    if constexpr (L >= 2) v = scan_step<1, L>(v, gl);
    if constexpr (L >= 4) v = scan_step<2, L>(v, gl);
    if constexpr (L >= 8) v = scan_step<4, L>(v, gl);
    if constexpr (L >= 16) v = scan_step<8, L>(v, gl);
    if constexpr (L >= 32) v += dpp_f<kDppRowBcast15, 0xA>(v);  // rows 1 and 3 add rows 0 and 2
    if constexpr (L == 64) v += dpp_f<kDppRowBcast31, 0xC>(v);  // rows 2 and 3 add lane 31
    return v;
}

template <int L> __device__ __forceinline__ float gsum(float v) { return sub_sum<float>(v, L); }
template <int L> __device__ __forceinline__ double gsum_d(double v) { return sub_sum<double>(v, L); }
__device__ __forceinline__ bool wave_any(bool p) { return __ballot(p) != 0; }

// ------------------------------------------------------------ angles / quaternions
// normalize_angle (common.h:260-281).  Angles beyond +-3 pi take the
// reference's "how many 2 pi's" steps (with its IEEE division) in a function
// of their own, so that the common path inlined everywhere is two compares.
__device__ __attribute__((noinline)) float normalize_angle_far(float x) {
    for (int it = 0; it < 8; ++it) {
        if (x > 3.f * kPi) {
            float n = (x - kPi) / (2.f * kPi);
            x -= 2.f * kPi * ceilf(n);
        } else if (x < -3.f * kPi) {
            float n = (-x - kPi) / (2.f * kPi);
            x += 2.f * kPi * ceilf(n);
        } else {
            break;
        }
    }
    return x;
}
__device__ __forceinline__ float normalize_angle(float x) {
    if (!(fabsf(x) <= 3.f * kPi)) x = normalize_angle_far(x);
    if (x > kPi) x -= 2.f * kPi;
    else if (x < -kPi) x += 2.f * kPi;
    return x;
}
// sin and cos of x in [-pi/2, pi/2] (half of a normalised angle) from their
// Taylor series through x^13 / x^14: the first omitted terms are below
// 7e-10 there, under half an ulp of the results, so the error is the FP32
// evaluation's (a few ulp).  No range reduction is needed for these
// arguments.
__device__ __forceinline__ void sincos_half(float x, float* s, float* c) {
    const float z = x * x;
    float ps = 1.f / 6227020800.f;                     // 1/13!
    ps = fmaf(ps, z, -1.f / 39916800.f);               // 1/11!
    ps = fmaf(ps, z, 1.f / 362880.f);                  // 1/9!
    ps = fmaf(ps, z, -1.f / 5040.f);                   // 1/7!
    ps = fmaf(ps, z, 1.f / 120.f);
    ps = fmaf(ps, z, -1.f / 6.f);
    *s = fmaf(ps * z, x, x);
    float pc = -1.f / 87178291200.f;                   // 1/14!
    pc = fmaf(pc, z, 1.f / 479001600.f);               // 1/12!
    pc = fmaf(pc, z, -1.f / 3628800.f);                // 1/10!
    pc = fmaf(pc, z, 1.f / 40320.f);                   // 1/8!
    pc = fmaf(pc, z, -1.f / 720.f);
    pc = fmaf(pc, z, 1.f / 24.f);
    pc = fmaf(pc, z, -0.5f);
    *c = fmaf(pc, z, 1.f);
}
// Hamilton product p * q, components (w, x, y, z) stored as float4 (x, y, z, w)
__device__ __forceinline__ float4 qmul(float4 p, float4 q) {
    return make_float4(p.x * q.x - p.y * q.y - p.z * q.z - p.w * q.w,
                       p.x * q.y + p.y * q.x + p.z * q.w - p.w * q.z,
                       p.x * q.z - p.y * q.w + p.z * q.x + p.w * q.y,
                       p.x * q.w + p.y * q.z - p.z * q.y + p.w * q.x);
}
// Quaternions split over the four lanes of a quad: lane c holds component c
// (w, x, y, z).  quad_perm<S0, S1, S2, S3>(v): lane i of each quad receives
// v of quad lane Si (all four lanes of the quad must be active).
template <int S0, int S1, int S2, int S3> __device__ __forceinline__ float quad_perm(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), S0 | (S1 << 2) | (S2 << 4) | (S3 << 6), 0xF, 0xF, true));
}
// x with its sign bit flipped where the mask's is set
__device__ __forceinline__ float sign_xor(float x, uint32_t m) { return __int_as_float(__float_as_int(x) ^ m); }
// Per-lane sign masks of the quad forms below (lane c = gl % 4): the Hamilton
// product's terms 1..3 of component c, and the b^2, c^2, d^2 terms of the
// diagonal of matrix row min(c, 2).  Computed arithmetically from c.
struct QuadSigns {
    uint32_t e1, e2, e3, db, dc, dd;
    __device__ __forceinline__ explicit QuadSigns(uint32_t c) {
        e1 = ((c & 1u) ^ 1u) << 31;                  // + for c = 1, 3
        e2 = (((c ^ (c >> 1)) & 1u) ^ 1u) << 31;     // + for c = 1, 2
        e3 = (((c >> 1) & 1u) ^ 1u) << 31;           // + for c = 2, 3
        db = ((c | (c >> 1)) & 1u) << 31;            // + for c = 0
        dc = (((c ^ 1u) | ((c ^ 1u) >> 1)) & 1u) << 31;  // + for c = 1
        dd = e3;                                     // + for c = 2, 3
    }
};
// Row r = min(c, 2) of the rotation matrix of the quad's quaternion
// (quaternion_to_r3, quaternion.cpp:77-114): with (a, b, c, d) the components,
// v = (b, c, d) and (r, s, t) = (r, r + 1, r + 2) mod 3,
//   m[r][r] = ((a^2 + -+b^2) + -+c^2) + -+d^2   (+ for the row's own axis),
//   m[r][s] = 2 (v_r v_s - a v_t),   m[r][t] = 2 (a v_s + v_r v_t),
// the reference's terms, every square and product rounded on its own.
// Lane 3 computes a copy of row 2.
__device__ __forceinline__ void quad_mat_row(float q, const QuadSigns& sg, float& mrr, float& mrs, float& mrt) {
#pragma clang fp contract(off)
    const float sq = q * q;
    const float aa = quad_perm<0, 0, 0, 0>(sq), bb = quad_perm<1, 1, 1, 1>(sq);
    const float cc = quad_perm<2, 2, 2, 2>(sq), dd = quad_perm<3, 3, 3, 3>(sq);
    const float a = quad_perm<0, 0, 0, 0>(q), vr = quad_perm<1, 2, 3, 3>(q);
    const float vs = quad_perm<2, 3, 1, 1>(q), vt = quad_perm<3, 1, 2, 2>(q);
    mrr = ((aa + sign_xor(bb, sg.db)) + sign_xor(cc, sg.dc)) + sign_xor(dd, sg.dd);
    const float p1 = vr * vs, p2 = vr * vt;
    mrs = 2.f * fmaf(-a, vt, p1);
    mrt = 2.f * fmaf(a, vs, p2);
}
// quaternion_normalize_approx (quaternion.h:105-117)
__device__ __forceinline__ float4 q_normalize_approx(float4 q) {
    const float s = q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w;
    if (fabsf(s - 1.f) < 1e-6f) return q;
    const float inv = __builtin_amdgcn_rsqf(s);
    return make_float4(q.x * inv, q.y * inv, q.z * inv, q.w * inv);
}
// angle_to_quaternion(axis, angle) (quaternion.cpp:37-47)
__device__ __forceinline__ float4 axis_angle_q(float ax, float ay, float az, float angle) {
    angle = normalize_angle(angle);
    float s, c;
    sincos_half(angle * 0.5f, &s, &c);
    return make_float4(c, s * ax, s * ay, s * az);
}
// quaternion_increment (quaternion.cpp:128-133) with angle_to_quaternion(rotation);
// the axis is the rotation times the angle's reciprocal.  Square roots and
// reciprocals on the per-iteration path are the hardware's (about 1 ulp), as
// the reference's GPU build computes them (--use_fast_math).
__device__ __forceinline__ float4 q_increment(float4 q, float rx, float ry, float rz) {
    const float angle = __builtin_amdgcn_sqrtf(rx * rx + ry * ry + rz * rz);
    float4 r = make_float4(1.f, 0.f, 0.f, 0.f);
    if (angle > kEpsFl) {
        const float inv = __builtin_amdgcn_rcpf(angle);
        r = axis_angle_q(rx * inv, ry * inv, rz * inv, angle);
    }
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
