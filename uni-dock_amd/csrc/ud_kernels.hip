// HIP kernels of the batched pose search for gfx950 (wave64).
//
//   table_kernel        per-type-pair interpolation tables (precalculate)
//   grid_build_kernel   receptor affinity maps (cache::populate_no_bias)
//   grid_absmax_kernel / grid_to_half_kernel   FP16 map storage
//   score_kernel        energy + gradient of poses (model::eval_deriv)
//   optimize_kernel     BFGS from given starts (quasi_newton / bfgs)
//   dock_kernel         Monte Carlo chains (kernel<32, Config>)
//   terms_kernel        final-score components (Vina::score_gpu)
//   coords_kernel       atom coordinates of poses (model::set)
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include "ud_device.h"
#include "ud_launch.h"

namespace ud {

// ====================================================================== tables
struct SfWeights {
    float w[7];
};

// ScoringFunction::eval(t1, t2, r) (scoring_function.h:114-118) with the
// Vina / Vinardo term lists of scoring_function.h:47-70 and the closed forms
// of potentials.h:137-343, 487-507.
__device__ float sf_eval(int sf, int t1, int t2, float r, const SfWeights& W) {
    // xs_vdw_radii / xs_vinardo_vdw_radii looked up by index (atom_constants.h:313-396)
    const float rv[36] = {1.9f, 1.9f, 1.8f, 1.8f, 1.8f, 1.8f, 1.7f, 1.7f, 1.7f, 1.7f, 2.0f, 2.1f,
                          1.5f, 1.8f, 2.0f, 2.2f, 2.2f, 2.3f, 1.2f, 1.9f, 1.9f, 1.9f, 1.9f, 1.9f,
                          1.9f, 1.9f, 1.9f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 1.7f, 1.8f, 1.7f, 1.8f};
    const float rd[36] = {2.0f, 2.0f, 1.7f, 1.7f, 1.7f, 1.7f, 1.6f, 1.6f, 1.6f, 1.6f, 2.0f, 2.1f,
                          1.5f, 1.8f, 2.0f, 2.2f, 2.2f, 2.3f, 1.2f, 2.0f, 2.0f, 2.0f, 2.0f, 2.0f,
                          2.0f, 2.0f, 2.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 1.6f, 1.7f, 1.6f, 1.7f};
    const bool glue = xs_glue(t1) || xs_glue(t2);
    float acc = 0.f;
    if (sf == UD_SF_VINA) {
        const float od = glue ? 0.f : rv[t1] + rv[t2];
        float term = 0.f;
        if (r < 8.f) { float x = (r - (od + 0.f)) / 0.5f; term = expf(-(x * x)); }
        acc += W.w[0] * term;
        term = 0.f;
        if (r < 8.f) { float x = (r - (od + 3.f)) / 2.0f; term = expf(-(x * x)); }
        acc += W.w[1] * term;
        term = 0.f;
        if (r < 8.f) { float d = r - (od + 0.f); term = (d > 0.f) ? 0.f : d * d; }
        acc += W.w[2] * term;
        term = 0.f;
        if (r < 8.f && xs_hydrophobic(t1) && xs_hydrophobic(t2)) term = slope_step(1.5f, 0.5f, r - od);
        acc += W.w[3] * term;
        term = 0.f;
        if (r < 8.f && xs_hbond(t1, t2)) {
            term = slope_step(0.f, -0.7f, r - od);
            if ((t1 >= 32 && t1 <= 35) || (t2 >= 32 && t2 <= 35)) term = 10.0f * term;
        }
        acc += W.w[4] * term;
        term = 0.f;
        if (r < 20.f && xs_glued(t1, t2)) term = r;
        acc += W.w[5] * term;
    } else {
        const float od = glue ? 0.f : rd[t1] + rd[t2];
        float term = 0.f;
        if (r < 8.f) { float x = (r - (od + 0.f)) / 0.8f; term = expf(-(x * x)); }
        acc += W.w[0] * term;
        term = 0.f;
        if (r < 8.f) { float d = r - (od + 0.f); term = (d > 0.f) ? 0.f : d * d; }
        acc += W.w[1] * term;
        term = 0.f;
        if (r < 8.f && xs_hydrophobic(t1) && xs_hydrophobic(t2)) term = slope_step(2.5f, 0.f, r - od);
        acc += W.w[2] * term;
        term = 0.f;
        if (r < 8.f && xs_hbond(t1, t2)) term = slope_step(0.f, -0.6f, r - od);
        acc += W.w[3] * term;
        term = 0.f;
        if (r < 20.f && xs_glued(t1, t2)) term = r;
        acc += W.w[4] * term;
    }
    return acc;
}

// One workgroup per type pair: smooth e at rs[k] = sqrt(k / 32), dor by
// central difference, fast = midpoint (precalculate.h:66-85, 160-171).
__global__ __launch_bounds__(256) void table_kernel(int sf, SfWeights W, float4* deriv_out, float* fast_out) {
    __shared__ float e[kTableN];
    __shared__ float rs[kTableN];
    const int slot = blockIdx.x;
    int t2 = 0;
    while ((t2 + 1) * (t2 + 2) / 2 <= slot) ++t2;
    const int t1 = slot - t2 * (t2 + 1) / 2;
    for (int k = threadIdx.x; k < kTableN; k += blockDim.x) {
        const float r = sqrtf(float(k) / 32.f);
        rs[k] = r;
        e[k] = fminf(kMaxFl, sf_eval(sf, t1, t2, r, W));
    }
    __syncthreads();
    auto dor = [&](int k) -> float {
        if (k == 0 || k == kTableN - 1) return 0.f;
        const float delta = rs[k + 1] - rs[k - 1];
        return (e[k + 1] - e[k - 1]) / (delta * rs[k]);
    };
    for (int k = threadIdx.x; k < kTableStride; k += blockDim.x) {
        const float d0 = dor(k), d1 = dor(k + 1);
        deriv_out[size_t(slot) * kTableStride + k] = make_float4(e[k], d0, e[k + 1] - e[k], d1 - d0);
        fast_out[size_t(slot) * kTableStride + k] = (e[k + 1] + e[k]) / 2.f;
    }
}

// ====================================================================== grid maps
__device__ __forceinline__ int tri_index(int i, int j) { return (i <= j) ? i + j * (j + 1) / 2 : j + i * (i + 1) / 2; }

// Each thread owns one sample point and up to 8 map types (blockIdx.y picks
// the chunk of types).  Receptor atoms stream through LDS in index order, so
// every map value is summed in the same order as the reference's loop over
// szv_grid possibilities (cache.cpp:518-548).
__global__ __launch_bounds__(256) void grid_build_kernel(const float4* rec, int n_rec, GridBuildArgs a,
                                                         const float* table_fast, float* maps_out) {
    __shared__ float4 tile[256];
    const int pt = blockIdx.x * blockDim.x + threadIdx.x;
    const int chunk = blockIdx.y;
    const int m0 = chunk * 8;
    const int nm = min(8, a.n_maps - m0);
    int tm[8];
#pragma unroll
    for (int m = 0; m < 8; ++m) tm[m] = (m < nm) ? a.types[m0 + m] : 0;
    const int x = pt % a.n[0], y = (pt / a.n[0]) % a.n[1], z = pt / (a.n[0] * a.n[1]);
    const float px = a.init[0] + a.factor_inv[0] * float(x);
    const float py = a.init[1] + a.factor_inv[1] * float(y);
    const float pz = a.init[2] + a.factor_inv[2] * float(z);
    float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    for (int base = 0; base < n_rec; base += 256) {
        const int j = base + threadIdx.x;
        if (j < n_rec) tile[threadIdx.x] = rec[j];
        __syncthreads();
        const int cnt = min(256, n_rec - base);
        if (pt < a.n_points) {
            for (int q = 0; q < cnt; ++q) {
                const float4 at = tile[q];
                const float dx = at.x - px, dy = at.y - py, dz = at.z - pz;
                const float r2 = dx * dx + dy * dy + dz * dz;
                if (r2 <= 64.f) {
                    const int k = int(32.f * r2);
                    const int t1 = __float_as_int(at.w);
#pragma unroll
                    for (int m = 0; m < 8; ++m)
                        if (m < nm) acc[m] += table_fast[size_t(tri_index(t1, tm[m])) * kTableStride + k];
                }
            }
        }
        __syncthreads();
    }
    if (pt < a.n_points) {
#pragma unroll
        for (int m = 0; m < 8; ++m)
            if (m < nm) maps_out[size_t(m0 + m) * a.n_points + pt] = acc[m];
    }
}

// max |value| per map (as ordered uint bits of a non-negative float)
__global__ __launch_bounds__(256) void grid_absmax_kernel(const float* maps, int n_points, unsigned int* out) {
    const int m = blockIdx.y;
    float mx = 0.f;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n_points; i += gridDim.x * blockDim.x)
        mx = fmaxf(mx, fabsf(maps[size_t(m) * n_points + i]));
    for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(out + m, __float_as_uint(mx));
}
__global__ __launch_bounds__(256) void grid_to_half_kernel(const float* maps, int n_points, const float* inv_scale,
                                                           __half* out) {
    const int m = blockIdx.y;
    const float s = inv_scale[m];
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n_points; i += gridDim.x * blockDim.x)
        out[size_t(m) * n_points + i] = __float2half_rn(maps[size_t(m) * n_points + i] * s);
}

// ====================================================================== per-group machinery
// LDS layout of one workgroup: the ligand's atoms and nodes, then one block
// per lane group (coordinates, minus-forces, node frames, node force/torque
// accumulators, six pose buffers, three broadcast vectors, inverse Hessian).
struct Layout {
    int na, nn, nc, nh;
    int o_coords, o_mf, o_frames, o_nacc, o_poses, o_bc, o_H, group_floats, shared_floats;
    __host__ __device__ void init(int na_, int nn_) {
        na = na_;
        nn = nn_;
        nc = 6 + nn - 1;
        nh = (nc * (nc + 1) / 2 + 3) & ~3;
        o_coords = 0;
        o_mf = 4 * na;
        o_frames = 8 * na;
        o_nacc = o_frames + 20 * nn;
        o_poses = o_nacc + 8 * nn;
        o_bc = o_poses + 6 * kPoseStride;
        o_H = o_bc + 3 * kPoseStride;
        group_floats = (o_H + nh + 3) & ~3;
        shared_floats = 4 * na + 8 * nn;
    }
    __host__ __device__ size_t bytes(int groups) const { return size_t(shared_floats + groups * group_floats) * 4; }
};
enum { PX = 0, PXN = 1, PXO = 2, PTMP = 3, PBEST = 4, PCAND = 5 };

struct KArgs {
    LigandSetDev ls;
    GridDev grid;
    int na_cap, nn_cap;
};

template <int L> struct Grp {
    static constexpr int G = kWG / L;
    static constexpr int KV = (kNMax + L - 1) / L;
    const AtomRec* atoms;  // LDS (shared)
    const NodeRec* nodes;  // LDS (shared)
    float* gb;             // LDS group block
    Layout lay;
    LigHdr h;
    int gl, g;
    __device__ float4* coords() const { return reinterpret_cast<float4*>(gb + lay.o_coords); }
    __device__ float4* mf() const { return reinterpret_cast<float4*>(gb + lay.o_mf); }
    __device__ float* frames() const { return gb + lay.o_frames; }
    __device__ float* nacc() const { return gb + lay.o_nacc; }
    __device__ float* pose(int k) const { return gb + lay.o_poses + k * kPoseStride; }
    __device__ float* bc(int k) const { return gb + lay.o_bc + k * kPoseStride; }
    __device__ float* H() const { return gb + lay.o_H; }
    __device__ int ntors() const { return h.n_nodes - 1; }
    __device__ int ndof() const { return 6 + h.n_nodes - 1; }
};

// Stage the ligand of `lig` into LDS and set up the lane's group.
template <int L>
__device__ Grp<L> setup_group(const KArgs& ka, int lig, float* lds) {
    Grp<L> s;
    s.lay.init(ka.na_cap, ka.nn_cap);
    s.h = ka.ls.hdr[lig];
    AtomRec* atoms = reinterpret_cast<AtomRec*>(lds);
    NodeRec* nodes = reinterpret_cast<NodeRec*>(lds + 4 * s.lay.na);
    for (int i = threadIdx.x; i < s.h.n_atoms; i += kWG) atoms[i] = ka.ls.atoms[s.h.atom_begin + i];
    for (int i = threadIdx.x; i < s.h.n_nodes; i += kWG) nodes[i] = ka.ls.nodes[s.h.node_begin + i];
    s.atoms = atoms;
    s.nodes = nodes;
    s.g = threadIdx.x / L;
    s.gl = threadIdx.x % L;
    s.gb = lds + s.lay.shared_floats + s.g * s.lay.group_floats;
    __syncthreads();
    return s;
}

// Forward kinematics (model::set, tree.h:121-144, 176-214): root frame from
// the pose, then tree levels parent-first, then atom coordinates.  Coordinates
// land in LDS; node frames (origin, axis, quaternion, rotation matrix) too.
template <int L>
__device__ void fk_frames(const Grp<L>& s, const float* pose) {
    float* fr = s.frames();
    if (s.gl == 0) {
        const float4 q = make_float4(pose[3], pose[4], pose[5], pose[6]);
        fr[0] = pose[0]; fr[1] = pose[1]; fr[2] = pose[2];
        fr[3] = 0.f; fr[4] = 0.f; fr[5] = 0.f;
        fr[6] = q.x; fr[7] = q.y; fr[8] = q.z; fr[9] = q.w;
        q_to_mat(q, fr + 10);
    }
    __syncthreads();
    for (int d = 1; d <= s.h.max_depth; ++d) {
        for (int k = 1 + s.gl; k < s.h.n_nodes; k += L) {
            const NodeRec nd = s.nodes[k];
            if (int(nd.packed >> 24) != d) continue;
            const float* pf = fr + 20 * nd.parent;
            const float* R = pf + 10;
            float* f = fr + 20 * k;
            const float ox = pf[0] + (R[0] * nd.ro[0] + R[1] * nd.ro[1] + R[2] * nd.ro[2]);
            const float oy = pf[1] + (R[3] * nd.ro[0] + R[4] * nd.ro[1] + R[5] * nd.ro[2]);
            const float oz = pf[2] + (R[6] * nd.ro[0] + R[7] * nd.ro[1] + R[8] * nd.ro[2]);
            const float axx = R[0] * nd.ax[0] + R[1] * nd.ax[1] + R[2] * nd.ax[2];
            const float axy = R[3] * nd.ax[0] + R[4] * nd.ax[1] + R[5] * nd.ax[2];
            const float axz = R[6] * nd.ax[0] + R[7] * nd.ax[1] + R[8] * nd.ax[2];
            const float4 qp = make_float4(pf[6], pf[7], pf[8], pf[9]);
            const float4 q = q_normalize_approx(qmul(axis_angle_q(axx, axy, axz, pose[7 + k - 1]), qp));
            f[0] = ox; f[1] = oy; f[2] = oz;
            f[3] = axx; f[4] = axy; f[5] = axz;
            f[6] = q.x; f[7] = q.y; f[8] = q.z; f[9] = q.w;
            q_to_mat(q, f + 10);
        }
        __syncthreads();
    }
}

__device__ __forceinline__ float3 frame_apply(const float* f, const AtomRec& a) {
    const float* R = f + 10;
    return make_float3(f[0] + (R[0] * a.x + R[1] * a.y + R[2] * a.z),
                       f[1] + (R[3] * a.x + R[4] * a.y + R[5] * a.z),
                       f[2] + (R[6] * a.x + R[7] * a.y + R[8] * a.z));
}

// grid::evaluate_aux (grid.cpp:39-129) + curl (curl.h:28-36); energy-only
// when d == nullptr.  FP16 maps interpolate the stored halves and apply the
// power-of-two map scale to value and gradient before the curl.
template <int GF>
__device__ __forceinline__ float grid_sample(const GridDev& g, int slot, float3 p, float v, float3* d) {
    float s[3] = {(p.x - g.init[0]) * g.factor[0], (p.y - g.init[1]) * g.factor[1], (p.z - g.init[2]) * g.factor[2]};
    float miss[3];
    int reg[3], a[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        if (s[i] < 0.f) {
            miss[i] = -s[i]; reg[i] = -1; a[i] = 0; s[i] = 0.f;
        } else if (s[i] >= g.dimm1[i]) {
            miss[i] = s[i] - g.dimm1[i]; reg[i] = 1; a[i] = g.n[i] - 2; s[i] = 1.f;
        } else {
            miss[i] = 0.f; reg[i] = 0; a[i] = int(s[i]); s[i] -= float(a[i]);
        }
        a[i] = min(max(a[i], 0), g.n[i] - 2);  // a NaN coordinate must not index outside the map
    }
    const float penalty = g.slope * (miss[0] * g.factor_inv[0] + miss[1] * g.factor_inv[1] + miss[2] * g.factor_inv[2]);
    const size_t sy = size_t(g.n[0]), sz = size_t(g.n[0]) * g.n[1];
    const size_t base = size_t(slot) * g.n_points + a[0] + sy * a[1] + sz * a[2];
    float f000, f100, f010, f110, f001, f101, f011, f111;
    if (GF == UD_GRID_F32) {
        const float* M = reinterpret_cast<const float*>(g.maps);
        f000 = M[base]; f100 = M[base + 1]; f010 = M[base + sy]; f110 = M[base + sy + 1];
        f001 = M[base + sz]; f101 = M[base + sz + 1]; f011 = M[base + sy + sz]; f111 = M[base + sy + sz + 1];
    } else {
        const __half* M = reinterpret_cast<const __half*>(g.maps);
        f000 = __half2float(M[base]); f100 = __half2float(M[base + 1]);
        f010 = __half2float(M[base + sy]); f110 = __half2float(M[base + sy + 1]);
        f001 = __half2float(M[base + sz]); f101 = __half2float(M[base + sz + 1]);
        f011 = __half2float(M[base + sy + sz]); f111 = __half2float(M[base + sy + sz + 1]);
    }
    const float x = s[0], y = s[1], z = s[2];
    const float mx = 1.f - x, my = 1.f - y, mz = 1.f - z;
    float f = f000 * mx * my * mz + f100 * x * my * mz + f010 * mx * y * mz + f110 * x * y * mz +
              f001 * mx * my * z + f101 * x * my * z + f011 * mx * y * z + f111 * x * y * z;
    if (GF == UD_GRID_F16) f *= g.scale[slot];
    if (d) {
        float gx = -f000 * my * mz + f100 * my * mz - f010 * y * mz + f110 * y * mz - f001 * my * z + f101 * my * z -
                   f011 * y * z + f111 * y * z;
        float gy = -f000 * mx * mz - f100 * x * mz + f010 * mx * mz + f110 * x * mz - f001 * mx * z - f101 * x * z +
                   f011 * mx * z + f111 * x * z;
        float gz = -f000 * mx * my - f100 * x * my - f010 * mx * y - f110 * x * y + f001 * mx * my + f101 * x * my +
                   f011 * mx * y + f111 * x * y;
        if (GF == UD_GRID_F16) {
            const float sc = g.scale[slot];
            gx *= sc; gy *= sc; gz *= sc;
        }
        if (f > 0.f && v < 0.1f * kMaxFl) {
            const float tmp = (v < kEpsFl) ? 0.f : (v / (v + f));
            f *= tmp;
            const float t2 = tmp * tmp;
            gx *= t2; gy *= t2; gz *= t2;
        }
        d->x = g.factor[0] * (reg[0] == 0 ? gx : 0.f) + g.slope * float(reg[0]);
        d->y = g.factor[1] * (reg[1] == 0 ? gy : 0.f) + g.slope * float(reg[1]);
        d->z = g.factor[2] * (reg[2] == 0 ? gz : 0.f) + g.slope * float(reg[2]);
    } else {
        if (f > 0.f && v < 0.1f * kMaxFl) {
            const float tmp = (v < kEpsFl) ? 0.f : (v / (v + f));
            f *= tmp;
        }
    }
    return f + penalty;
}

// Energy (+ gradient) of the group's pose: model::eval_deriv (model.cpp:954-981)
// = grid term per atom, intramolecular pairs (model.cpp:907-934, pair forces
// accumulated with LDS atomics), then the torsion-tree derivative
// (tree.h:75-84, 152-193): per-node force/torque over its own atoms, then
// children pushed into parents level by level, deepest first.
// gout[j] receives gradient element gl + j * L.
template <int L, int GF>
__device__ float eval_group(const KArgs& ka, const Grp<L>& s, const float* pose, float v_intra, float v_inter,
                            float (&gout)[Grp<L>::KV], bool want_grad) {
    fk_frames(s, pose);
    float4* coords = s.coords();
    float4* mf = s.mf();
    float* fr = s.frames();
    const int N = s.h.n_atoms;
    const int c = (N + L - 1) / L;
    const int i0 = s.gl * c, i1 = min(N, i0 + c);
    float e = 0.f;
    for (int i = i0; i < i1; ++i) {
        const AtomRec a = s.atoms[i];
        const float3 p = frame_apply(fr + 20 * (a.meta & 0xff), a);
        coords[i] = make_float4(p.x, p.y, p.z, 0.f);
        const int slot = (a.meta >> 8) & 0xff;
        float3 d = make_float3(0.f, 0.f, 0.f);
        if (slot != 0xff) e += grid_sample<GF>(ka.grid, slot, p, v_inter, &d);
        mf[i] = make_float4(d.x, d.y, d.z, 0.f);
    }
    __syncthreads();
    const PairRec* prs = ka.ls.pairs + s.h.pair_begin;
    for (int q = s.gl; q < s.h.n_pairs; q += L) {
        const PairRec rec = prs[q];
        const int a = rec & 0xff, b = (rec >> 8) & 0xff;
        const float4 ca = coords[a], cb = coords[b];
        const float rx = cb.x - ca.x, ry = cb.y - ca.y, rz = cb.z - ca.z;
        const float r2 = rx * rx + ry * ry + rz * rz;
        if (r2 < 64.f) {
            const float rf = 32.f * r2;
            const int i1t = int(rf);
            const float rem = rf - float(i1t);
            const float4 t = ka.ls.table_deriv[size_t(rec >> 16) * kTableStride + i1t];
            float pe = t.x + rem * t.z;
            const float dor = t.y + rem * t.w;
            float fx = dor * rx, fy = dor * ry, fz = dor * rz;
            if (pe > 0.f && v_intra < 0.1f * kMaxFl) {
                const float tmp = (v_intra < kEpsFl) ? 0.f : (v_intra / (v_intra + pe));
                pe *= tmp;
                const float t2 = tmp * tmp;
                fx *= t2; fy *= t2; fz *= t2;
            }
            e += pe;
            atomicAdd(&mf[a].x, -fx); atomicAdd(&mf[a].y, -fy); atomicAdd(&mf[a].z, -fz);
            atomicAdd(&mf[b].x, fx); atomicAdd(&mf[b].y, fy); atomicAdd(&mf[b].z, fz);
        }
    }
    e = gsum<L>(e);
    __syncthreads();
    if (!want_grad) return e;
    float* nacc = s.nacc();
    float* gbuf = s.bc(2);
    for (int k = s.gl; k < s.h.n_nodes; k += L) {
        const NodeRec nd = s.nodes[k];
        const int ab = nd.packed & 0xff, ae = (nd.packed >> 8) & 0xff;
        const float* f = fr + 20 * k;
        float F0 = 0.f, F1 = 0.f, F2 = 0.f, T0 = 0.f, T1 = 0.f, T2 = 0.f;
        for (int i = ab; i < ae; ++i) {
            const float4 fi = mf[i], ci = coords[i];
            const float dx = ci.x - f[0], dy = ci.y - f[1], dz = ci.z - f[2];
            F0 += fi.x; F1 += fi.y; F2 += fi.z;
            T0 += dy * fi.z - dz * fi.y;
            T1 += dz * fi.x - dx * fi.z;
            T2 += dx * fi.y - dy * fi.x;
        }
        float* na = nacc + 8 * k;
        na[0] = F0; na[1] = F1; na[2] = F2; na[3] = T0; na[4] = T1; na[5] = T2;
    }
    __syncthreads();
    for (int d = s.h.max_depth; d >= 1; --d) {
        for (int k = 1 + s.gl; k < s.h.n_nodes; k += L) {
            const NodeRec nd = s.nodes[k];
            if (int(nd.packed >> 24) != d) continue;
            const float* na = nacc + 8 * k;
            const float* f = fr + 20 * k;
            const float* pf = fr + 20 * nd.parent;
            gbuf[5 + k] = na[3] * f[3] + na[4] * f[4] + na[5] * f[5];
            const float rx = f[0] - pf[0], ry = f[1] - pf[1], rz = f[2] - pf[2];
            float* pa = nacc + 8 * nd.parent;
            atomicAdd(pa + 0, na[0]); atomicAdd(pa + 1, na[1]); atomicAdd(pa + 2, na[2]);
            atomicAdd(pa + 3, (ry * na[2] - rz * na[1]) + na[3]);
            atomicAdd(pa + 4, (rz * na[0] - rx * na[2]) + na[4]);
            atomicAdd(pa + 5, (rx * na[1] - ry * na[0]) + na[5]);
        }
        __syncthreads();
    }
    if (s.gl < 6) gbuf[s.gl] = nacc[s.gl];
    __syncthreads();
    const int n = s.ndof();
#pragma unroll
    for (int j = 0; j < Grp<L>::KV; ++j) {
        const int idx = s.gl + j * L;
        gout[j] = (idx < n) ? gbuf[idx] : 0.f;
    }
    __syncthreads();
    return e;
}

template <int L>
__device__ __forceinline__ float gdot(const float (&a)[Grp<L>::KV], const float (&b)[Grp<L>::KV]) {
    float s = 0.f;
#pragma unroll
    for (int j = 0; j < Grp<L>::KV; ++j) s += a[j] * b[j];  // padded entries are zero
    return gsum<L>(s);
}
template <int L>
__device__ __forceinline__ double gdot_d(const float (&a)[Grp<L>::KV], const float (&b)[Grp<L>::KV]) {
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < Grp<L>::KV; ++j) s += double(a[j]) * double(b[j]);
    return gsum_d<L>(s);
}

// out = -H * in over the packed upper triangle (minus_mat_vec_product, bfgs.h:30-38)
template <int L>
__device__ void minus_hmatvec(const Grp<L>& s, const float (&in)[Grp<L>::KV], float (&out)[Grp<L>::KV], int n) {
    float* bc = s.bc(0);
    const float* H = s.H();
#pragma unroll
    for (int j = 0; j < Grp<L>::KV; ++j) {
        const int idx = s.gl + j * L;
        if (idx < n) bc[idx] = in[j];
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < Grp<L>::KV; ++j) {
        const int i = s.gl + j * L;
        float sum = 0.f;
        if (i < n) {
            for (int c2 = 0; c2 < n; ++c2) sum += H[tri_index(i, c2)] * bc[c2];
        }
        out[j] = (i < n) ? -sum : 0.f;
    }
    __syncthreads();
}

// conf::increment (conf.h:98-104, 50-56, 169-172): position += a p,
// orientation <- exp(a rot) * q (normalised), torsions wrapped.
template <int L>
__device__ void pose_increment(const Grp<L>& s, float* dst, const float* src, const float* p, float alpha, bool pred) {
    if (!pred) return;
    const int T = s.ntors();
    for (int idx = s.gl; idx < 4 + T; idx += L) {
        if (idx < 3) {
            dst[idx] = src[idx] + alpha * p[idx];
        } else if (idx == 3) {
            const float4 q = q_increment(make_float4(src[3], src[4], src[5], src[6]), alpha * p[3], alpha * p[4], alpha * p[5]);
            dst[3] = q.x; dst[4] = q.y; dst[5] = q.z; dst[6] = q.w;
        } else {
            const int k = idx - 4;
            dst[7 + k] = normalize_angle(src[7 + k] + normalize_angle(alpha * p[6 + k]));
        }
    }
}

template <int L>
__device__ __forceinline__ void copy_pose(const Grp<L>& s, float* dst, const float* src, bool pred) {
    if (!pred) return;
    for (int idx = s.gl; idx < 7 + s.ntors(); idx += L) dst[idx] = src[idx];
}

// bfgs() (bfgs.h:97-147) with line_search (bfgs.h:65-85) and bfgs_update
// (bfgs.h:46-63).  x is the group's pose buffer (in/out).  The curvature
// scalar y.p is accumulated in FP64.  Groups with active == false run along
// (uniform control flow) without committing anything.
template <int L, int GF>
__device__ float bfgs_group(const KArgs& ka, const Grp<L>& s, float* x, float v_intra, float v_inter, int max_steps,
                            bool active, int& evals) {
    constexpr int KV = Grp<L>::KV;
    float* xn = s.pose(PXN);
    float* xo = s.pose(PXO);
    float* H = s.H();
    const int n = s.ndof();
    const int nh = n * (n + 1) / 2;
    for (int idx = s.gl; idx < nh; idx += L) H[idx] = 0.f;
    for (int i = s.gl; i < n; i += L) H[i + i * (i + 1) / 2] = 1.f;
    copy_pose(s, xn, x, true);
    copy_pose(s, xo, x, true);
    __syncthreads();
    float g[KV], gn[KV], p[KV], y[KV], go[KV], t[KV];
    float f0 = eval_group<L, GF>(ka, s, x, v_intra, v_inter, g, true);
    if (active) ++evals;
    const float f_orig = f0;
#pragma unroll
    for (int j = 0; j < KV; ++j) { go[j] = g[j]; gn[j] = g[j]; y[j] = 0.f; }
    bool done = !active;
    for (int step = 0; step < max_steps; ++step) {
        if (!wave_any(!done)) break;
        minus_hmatvec<L>(s, g, p, n);
        const float pg = gdot<L>(p, g);
        float* pb = s.bc(1);
#pragma unroll
        for (int j = 0; j < KV; ++j) {
            const int idx = s.gl + j * L;
            if (idx < n) pb[idx] = p[j];
        }
        __syncthreads();
        float alpha = 1.f, f1 = 0.f;
        bool srch = !done;
        for (int trial = 0; trial < 10; ++trial) {
            if (!wave_any(srch)) break;
            pose_increment(s, xn, x, pb, alpha, srch);
            __syncthreads();
            const float ft = eval_group<L, GF>(ka, s, xn, v_intra, v_inter, t, true);
            if (srch) {
                ++evals;
                f1 = ft;
#pragma unroll
                for (int j = 0; j < KV; ++j) gn[j] = t[j];
                if (f1 - f0 < 0.0001f * alpha * pg) srch = false;
                else alpha *= 0.5f;
            }
        }
        if (!done) {
#pragma unroll
            for (int j = 0; j < KV; ++j) y[j] = gn[j] - g[j];
            f0 = f1;
        }
        copy_pose(s, x, xn, !done);
        __syncthreads();
        const float gg = gdot<L>(g, g);
        if (!done && !(sqrtf(gg) >= 1e-5f)) done = true;  // breaks before g = g_new (bfgs.h:131-132)
        const bool upd = !done;
        if (upd) {
#pragma unroll
            for (int j = 0; j < KV; ++j) g[j] = gn[j];
        }
        if (step == 0) {
            const float yy = gdot<L>(y, y);
            const float yp0 = gdot<L>(y, p);
            if (upd && fabsf(yy) > kEpsFl) {
                const float dv = alpha * yp0 / yy;
                for (int i = s.gl; i < n; i += L) H[i + i * (i + 1) / 2] = dv;
            }
            __syncthreads();
        }
        const float yp = float(gdot_d<L>(y, p));
        const bool do_upd = upd && !(alpha * yp < kEpsFl);
        minus_hmatvec<L>(s, y, t, n);
        const float yhy = -gdot<L>(y, t);
        const float r = 1.f / (alpha * yp);
        float* ma = s.bc(0);
        float* pp = s.bc(1);
#pragma unroll
        for (int j = 0; j < KV; ++j) {
            const int idx = s.gl + j * L;
            if (idx < n) { ma[idx] = t[j]; pp[idx] = p[j]; }
        }
        __syncthreads();
        if (do_upd) {
            const float c1 = alpha * r, c2 = alpha * alpha * (r * r * yhy + r);
            for (int idx = s.gl; idx < nh; idx += L) {
                int jj = int((sqrtf(8.f * float(idx) + 1.f) - 1.f) * 0.5f);
                while (jj * (jj + 1) / 2 > idx) --jj;
                while ((jj + 1) * (jj + 2) / 2 <= idx) ++jj;
                const int ii = idx - jj * (jj + 1) / 2;
                H[idx] += c1 * (ma[ii] * pp[jj] + ma[jj] * pp[ii]) + c2 * pp[ii] * pp[jj];
            }
        }
        __syncthreads();
    }
    const bool revert = active && !(f0 <= f_orig);
    if (revert) f0 = f_orig;
    copy_pose(s, x, xo, revert);
    __syncthreads();
    return f0;
}

// ====================================================================== kernels
template <int L, int GF>
__global__ __launch_bounds__(64) void score_kernel(KArgs ka, const WorkItem* work, const int32_t* perm,
                                                   const float* poses, float v_intra, float v_inter, float* e_out,
                                                   float* g_out) {
    extern __shared__ float4 lds_raw[];
    float* lds = reinterpret_cast<float*>(lds_raw);
    const WorkItem w = work[blockIdx.x];
    Grp<L> s = setup_group<L>(ka, w.lig, lds);
    const bool active = s.g < w.count;
    const int pi = active ? perm[w.first + s.g] : 0;
    float* x = s.pose(PX);
    const int np = 7 + s.ntors();
    for (int i = s.gl; i < np; i += L) x[i] = active ? poses[size_t(pi) * UD_POSE_FLOATS + i] : (i == 3 ? 1.f : 0.f);
    __syncthreads();
    float gout[Grp<L>::KV];
    const float e = eval_group<L, GF>(ka, s, x, v_intra, v_inter, gout, true);
    if (active) {
        if (s.gl == 0) e_out[pi] = e;
#pragma unroll
        for (int j = 0; j < Grp<L>::KV; ++j) {
            const int idx = s.gl + j * L;
            if (idx < UD_GRAD_FLOATS) g_out[size_t(pi) * UD_GRAD_FLOATS + idx] = gout[j];
        }
    }
}

template <int L, int GF>
__global__ __launch_bounds__(64) void optimize_kernel(KArgs ka, const WorkItem* work, const int32_t* perm, float* poses,
                                                      float v_intra, float v_inter, int max_steps, float* e_out,
                                                      int32_t* evals_out) {
    extern __shared__ float4 lds_raw[];
    float* lds = reinterpret_cast<float*>(lds_raw);
    const WorkItem w = work[blockIdx.x];
    Grp<L> s = setup_group<L>(ka, w.lig, lds);
    const bool active = s.g < w.count;
    const int pi = active ? perm[w.first + s.g] : 0;
    float* x = s.pose(PX);
    const int np = 7 + s.ntors();
    for (int i = s.gl; i < np; i += L) x[i] = active ? poses[size_t(pi) * UD_POSE_FLOATS + i] : (i == 3 ? 1.f : 0.f);
    __syncthreads();
    const int steps = max_steps > 0 ? max_steps : s.h.local_steps;
    int evals = 0;
    const float e = bfgs_group<L, GF>(ka, s, x, v_intra, v_inter, steps, active, evals);
    if (active) {
        for (int i = s.gl; i < np; i += L) poses[size_t(pi) * UD_POSE_FLOATS + i] = x[i];
        if (s.gl == 0) {
            if (e_out) e_out[pi] = e;
            if (evals_out) evals_out[pi] = evals;
        }
    }
}

// conf::randomize (conf.h:109-112, 152-155): position uniform in the box,
// orientation a normalised 4-D Gaussian, torsions U(-pi, pi); drawn from
// Philox with counter (0, chain, purpose, 0).
__device__ void random_start(float* pose, int T, const SearchDev& sp, uint32_t chain) {
    const uint2 key = make_uint2(sp.seed_lo, sp.seed_hi);
    uint4 o = philox4x32_10(make_uint4(0u, chain, 1u, 0u), key);
    pose[0] = sp.box_lo[0] + u01(o.x) * (sp.box_hi[0] - sp.box_lo[0]);
    pose[1] = sp.box_lo[1] + u01(o.y) * (sp.box_hi[1] - sp.box_lo[1]);
    pose[2] = sp.box_lo[2] + u01(o.z) * (sp.box_hi[2] - sp.box_lo[2]);
    o = philox4x32_10(make_uint4(0u, chain, 2u, 0u), key);
    float nq[4];
    const uint32_t ov[4] = {o.x, o.y, o.z, o.w};
    for (int j = 0; j < 2; ++j) {
        const float u1 = 1.f - u01(ov[2 * j]);
        const float u2 = u01(ov[2 * j + 1]);
        const float rad = sqrtf(-2.f * logf(u1));
        const float ang = 2.f * kPi * u2;
        nq[2 * j] = rad * cosf(ang);
        nq[2 * j + 1] = rad * sinf(ang);
    }
    float nrm = sqrtf(nq[0] * nq[0] + nq[1] * nq[1] + nq[2] * nq[2] + nq[3] * nq[3]);
    if (!(nrm > kEpsFl)) { nq[0] = 1.f; nq[1] = nq[2] = nq[3] = 0.f; nrm = 1.f; }
    for (int j = 0; j < 4; ++j) pose[3 + j] = nq[j] / nrm;
    uint32_t tv[4] = {0, 0, 0, 0};
    for (int k = 0; k < T; ++k) {
        if (k % 4 == 0) {
            const uint4 t = philox4x32_10(make_uint4(0u, chain, 3u + uint32_t(k / 4), 0u), key);
            tv[0] = t.x; tv[1] = t.y; tv[2] = t.z; tv[3] = t.w;
        }
        pose[7 + k] = -kPi + 2.f * kPi * u01(tv[k % 4]);
    }
}

// mutate_conf (mutate.cpp:35-81; device twin mutate_conf_cuda, common.cuh:207-258)
__device__ void mutate(float* pose, int T, float rgyr, const SearchDev& sp, uint32_t chain, uint32_t step) {
    const uint4 o = philox4x32_10(make_uint4(step + 1u, chain, 1u, 0u), make_uint2(sp.seed_lo, sp.seed_hi));
    const uint32_t count = 2u + uint32_t(T);
    const uint32_t which = uint32_t((uint64_t(o.x) * count) >> 32);
    const float z = 2.f * u01(o.y) - 1.f;
    const float phi = 2.f * kPi * u01(o.z);
    const float rr = cbrtf(u01(o.w));
    const float sxy = sqrtf(fmaxf(0.f, 1.f - z * z));
    const float ux = rr * sxy * cosf(phi), uy = rr * sxy * sinf(phi), uz = rr * z;
    const float amp = sp.amplitude;
    if (which == 0) {
        pose[0] += amp * ux; pose[1] += amp * uy; pose[2] += amp * uz;
    } else if (which == 1) {
        if (rgyr > kEpsFl) {
            const float f = amp / rgyr;
            const float4 q = q_increment(make_float4(pose[3], pose[4], pose[5], pose[6]), f * ux, f * uy, f * uz);
            pose[3] = q.x; pose[4] = q.y; pose[5] = q.z; pose[6] = q.w;
        }
    } else {
        pose[7 + (which - 2)] = -kPi + 2.f * kPi * u01(o.y);
    }
}

// gyration radius of the group's pose (model::gyration_radius, model.cpp:871-883)
template <int L>
__device__ float gyration_group(const Grp<L>& s, const float* pose) {
    fk_frames(s, pose);
    const float* fr = s.frames();
    float acc = 0.f;
    for (int i = s.gl; i < s.h.n_atoms; i += L) {
        const AtomRec a = s.atoms[i];
        if (!((a.meta >> 16) & 1)) continue;
        const float3 p = frame_apply(fr + 20 * (a.meta & 0xff), a);
        const float dx = p.x - pose[0], dy = p.y - pose[1], dz = p.z - pose[2];
        acc += dx * dx + dy * dy + dz * dz;
    }
    acc = gsum<L>(acc);
    __syncthreads();
    return s.h.n_heavy > 0 ? sqrtf(acc / float(s.h.n_heavy)) : 0.f;
}

// Monte Carlo chains (kernel<32, Config>, monte_carlo.cu:236-356): one chain
// per lane group, all groups of the workgroup on one ligand.
template <int L, int GF>
__global__ __launch_bounds__(64) void dock_kernel(KArgs ka, SearchDev sp, const WorkItem* work, float* out_pose,
                                                  float* out_e, long long* out_evals) {
    extern __shared__ float4 lds_raw[];
    float* lds = reinterpret_cast<float*>(lds_raw);
    const WorkItem w = work[blockIdx.x];
    Grp<L> s = setup_group<L>(ka, w.lig, lds);
    const bool active = s.g < w.count;
    const uint32_t chain = uint32_t(w.first + (active ? s.g : 0));
    const int T = s.ntors();
    const int np = 7 + T;
    float* tmp = s.pose(PTMP);
    float* best = s.pose(PBEST);
    float* cand = s.pose(PCAND);
    if (s.gl == 0) random_start(tmp, T, sp, chain);
    __syncthreads();
    float rgyr = gyration_group<L>(s, tmp);
    copy_pose(s, best, tmp, true);
    float tmp_e = 0.f, best_e = __builtin_huge_valf();
    int evals = 0;
    const int lsteps = sp.local_steps_override > 0 ? sp.local_steps_override : s.h.local_steps;
    const uint2 key = make_uint2(sp.seed_lo, sp.seed_hi);
    for (int step = 0; step < sp.global_steps; ++step) {
        copy_pose(s, cand, tmp, true);
        __syncthreads();
        if (s.gl == 0) mutate(cand, T, rgyr, sp, chain, uint32_t(step));
        __syncthreads();
        const float cand_e = bfgs_group<L, GF>(ka, s, cand, sp.hunt[0], sp.hunt[1], lsteps, active, evals);
        const uint4 o = philox4x32_10(make_uint4(uint32_t(step) + 1u, chain, 2u, 0u), key);
        bool acc = cand_e < tmp_e;
        if (!acc) acc = u01(o.x) < expf((tmp_e - cand_e) / sp.temperature);  // metropolis_accept (monte_carlo.cu:92-97)
        const bool take = active && (step == 0 || acc);
        copy_pose(s, tmp, cand, take);
        if (take) tmp_e = cand_e;
        __syncthreads();
        const bool need = take && tmp_e < best_e;
        if (wave_any(need)) {
            const float e2 = bfgs_group<L, GF>(ka, s, tmp, sp.authentic[0], sp.authentic[1], lsteps, need, evals);
            if (need) tmp_e = e2;
        }
        const bool better = need && tmp_e < best_e;
        copy_pose(s, best, tmp, better);
        if (better) best_e = tmp_e;
        __syncthreads();
        if (wave_any(take)) {
            const float rg = gyration_group<L>(s, tmp);
            if (take) rgyr = rg;
        }
    }
    if (active) {
        for (int i = s.gl; i < UD_POSE_FLOATS; i += L) out_pose[size_t(chain) * UD_POSE_FLOATS + i] = (i < np) ? best[i] : 0.f;
        if (s.gl == 0) {
            out_e[chain] = best_e;
            out_evals[chain] = evals;
        }
    }
}

// Final-score components with authentic caps (no gradient): inter = grid
// (cache::eval), intra = pairs through the piecewise-constant table with
// curl (eval_interacting_pairs, model.cpp:885-905).
template <int L, int GF>
__global__ __launch_bounds__(64) void terms_kernel(KArgs ka, const WorkItem* work, const int32_t* perm,
                                                   const float* poses, float v, float* inter_out, float* intra_out) {
    extern __shared__ float4 lds_raw[];
    float* lds = reinterpret_cast<float*>(lds_raw);
    const WorkItem w = work[blockIdx.x];
    Grp<L> s = setup_group<L>(ka, w.lig, lds);
    const bool active = s.g < w.count;
    const int pi = active ? perm[w.first + s.g] : 0;
    float* x = s.pose(PX);
    const int np = 7 + s.ntors();
    for (int i = s.gl; i < np; i += L) x[i] = active ? poses[size_t(pi) * UD_POSE_FLOATS + i] : (i == 3 ? 1.f : 0.f);
    __syncthreads();
    fk_frames(s, x);
    float4* coords = s.coords();
    const float* fr = s.frames();
    float ei = 0.f, ea = 0.f;
    for (int i = s.gl; i < s.h.n_atoms; i += L) {
        const AtomRec a = s.atoms[i];
        const float3 p = frame_apply(fr + 20 * (a.meta & 0xff), a);
        coords[i] = make_float4(p.x, p.y, p.z, 0.f);
        const int slot = (a.meta >> 8) & 0xff;
        if (slot != 0xff) ei += grid_sample<GF>(ka.grid, slot, p, v, nullptr);
    }
    __syncthreads();
    const PairRec* prs = ka.ls.pairs + s.h.pair_begin;
    for (int q = s.gl; q < s.h.n_pairs; q += L) {
        const PairRec rec = prs[q];
        const float4 ca = coords[rec & 0xff], cb = coords[(rec >> 8) & 0xff];
        const float rx = cb.x - ca.x, ry = cb.y - ca.y, rz = cb.z - ca.z;
        const float r2 = rx * rx + ry * ry + rz * rz;
        if (r2 < 64.f) {
            float t = ka.ls.table_fast[size_t(rec >> 16) * kTableStride + int(32.f * r2)];
            if (t > 0.f && v < 0.1f * kMaxFl) t *= (v < kEpsFl) ? 0.f : v / (v + t);
            ea += t;
        }
    }
    ei = gsum<L>(ei);
    ea = gsum<L>(ea);
    if (active && s.gl == 0) {
        inter_out[pi] = ei;
        intra_out[pi] = ea;
    }
}

template <int L>
__global__ __launch_bounds__(64) void coords_kernel(KArgs ka, const WorkItem* work, const int32_t* perm,
                                                    const float* poses, float* coords_out) {
    extern __shared__ float4 lds_raw[];
    float* lds = reinterpret_cast<float*>(lds_raw);
    const WorkItem w = work[blockIdx.x];
    Grp<L> s = setup_group<L>(ka, w.lig, lds);
    const bool active = s.g < w.count;
    const int pi = active ? perm[w.first + s.g] : 0;
    float* x = s.pose(PX);
    const int np = 7 + s.ntors();
    for (int i = s.gl; i < np; i += L) x[i] = active ? poses[size_t(pi) * UD_POSE_FLOATS + i] : (i == 3 ? 1.f : 0.f);
    __syncthreads();
    fk_frames(s, x);
    const float* fr = s.frames();
    if (active) {
        for (int i = s.gl; i < UD_MAX_ATOMS; i += L) {
            float3 p = make_float3(0.f, 0.f, 0.f);
            if (i < s.h.n_atoms) {
                const AtomRec a = s.atoms[i];
                p = frame_apply(fr + 20 * (a.meta & 0xff), a);
            }
            float* o = coords_out + (size_t(pi) * UD_MAX_ATOMS + i) * 3;
            o[0] = p.x; o[1] = p.y; o[2] = p.z;
        }
    }
}

// ====================================================================== launchers
size_t lds_bytes(int lanes, int na_cap, int nn_cap) {
    Layout l;
    l.init(na_cap, nn_cap);
    return l.bytes(kWG / lanes);
}

static KArgs make_kargs(const LigandSetDev& ls, const GridDev& g, int na_cap, int nn_cap) {
    KArgs k;
    k.ls = ls;
    k.grid = g;
    k.na_cap = na_cap;
    k.nn_cap = nn_cap;
    return k;
}

#define UD_DISPATCH_L(lanes, BODY)                 \
    switch (lanes) {                               \
        case 8: { constexpr int L = 8; BODY; } break;   \
        case 16: { constexpr int L = 16; BODY; } break; \
        case 32: { constexpr int L = 32; BODY; } break; \
        default: { constexpr int L = 64; BODY; } break; \
    }

hipError_t launch_tables(int sf, const float* w, float4* deriv, float* fast, hipStream_t st) {
    SfWeights W;
    for (int i = 0; i < 7; ++i) W.w[i] = w[i];
    hipLaunchKernelGGL(table_kernel, dim3(kTypePairs), dim3(256), 0, st, sf, W, deriv, fast);
    return hipGetLastError();
}

hipError_t launch_grid_build(const float4* rec, int n_rec, const GridBuildArgs& a, const float* table_fast,
                             float* maps_out, hipStream_t st) {
    dim3 grid((a.n_points + 255) / 256, (a.n_maps + 7) / 8);
    hipLaunchKernelGGL(grid_build_kernel, grid, dim3(256), 0, st, rec, n_rec, a, table_fast, maps_out);
    return hipGetLastError();
}

hipError_t launch_grid_absmax(const float* maps, int n_points, int n_maps, unsigned int* out, hipStream_t st) {
    hipLaunchKernelGGL(grid_absmax_kernel, dim3(256, n_maps), dim3(256), 0, st, maps, n_points, out);
    return hipGetLastError();
}

hipError_t launch_grid_to_half(const float* maps, int n_points, int n_maps, const float* inv_scale, void* out,
                               hipStream_t st) {
    hipLaunchKernelGGL(grid_to_half_kernel, dim3(256, n_maps), dim3(256), 0, st, maps, n_points, inv_scale,
                       reinterpret_cast<__half*>(out));
    return hipGetLastError();
}

hipError_t launch_score(int lanes, int gf, const LigandSetDev& ls, const GridDev& g, int na_cap, int nn_cap,
                        const WorkItem* work, int n_work, const int32_t* perm, const float* poses, float v_intra,
                        float v_inter, float* e_out, float* g_out, hipStream_t st) {
    const KArgs ka = make_kargs(ls, g, na_cap, nn_cap);
    const size_t shm = lds_bytes(lanes, na_cap, nn_cap);
    UD_DISPATCH_L(lanes, {
        if (gf == UD_GRID_F16)
            hipLaunchKernelGGL((score_kernel<L, UD_GRID_F16>), dim3(n_work), dim3(kWG), shm, st, ka, work, perm, poses,
                               v_intra, v_inter, e_out, g_out);
        else
            hipLaunchKernelGGL((score_kernel<L, UD_GRID_F32>), dim3(n_work), dim3(kWG), shm, st, ka, work, perm, poses,
                               v_intra, v_inter, e_out, g_out);
    });
    return hipGetLastError();
}

hipError_t launch_optimize(int lanes, int gf, const LigandSetDev& ls, const GridDev& g, int na_cap, int nn_cap,
                           const WorkItem* work, int n_work, const int32_t* perm, float* poses, float v_intra,
                           float v_inter, int max_steps, float* e_out, int32_t* evals_out, hipStream_t st) {
    const KArgs ka = make_kargs(ls, g, na_cap, nn_cap);
    const size_t shm = lds_bytes(lanes, na_cap, nn_cap);
    UD_DISPATCH_L(lanes, {
        if (gf == UD_GRID_F16)
            hipLaunchKernelGGL((optimize_kernel<L, UD_GRID_F16>), dim3(n_work), dim3(kWG), shm, st, ka, work, perm,
                               poses, v_intra, v_inter, max_steps, e_out, evals_out);
        else
            hipLaunchKernelGGL((optimize_kernel<L, UD_GRID_F32>), dim3(n_work), dim3(kWG), shm, st, ka, work, perm,
                               poses, v_intra, v_inter, max_steps, e_out, evals_out);
    });
    return hipGetLastError();
}

hipError_t launch_dock(int lanes, int gf, const LigandSetDev& ls, const GridDev& g, int na_cap, int nn_cap,
                       const SearchDev& sp, const WorkItem* work, int n_work, float* out_pose, float* out_e,
                       long long* out_evals, hipStream_t st) {
    const KArgs ka = make_kargs(ls, g, na_cap, nn_cap);
    const size_t shm = lds_bytes(lanes, na_cap, nn_cap);
    UD_DISPATCH_L(lanes, {
        if (gf == UD_GRID_F16)
            hipLaunchKernelGGL((dock_kernel<L, UD_GRID_F16>), dim3(n_work), dim3(kWG), shm, st, ka, sp, work, out_pose,
                               out_e, out_evals);
        else
            hipLaunchKernelGGL((dock_kernel<L, UD_GRID_F32>), dim3(n_work), dim3(kWG), shm, st, ka, sp, work, out_pose,
                               out_e, out_evals);
    });
    return hipGetLastError();
}

hipError_t launch_terms(int lanes, int gf, const LigandSetDev& ls, const GridDev& g, int na_cap, int nn_cap,
                        const WorkItem* work, int n_work, const int32_t* perm, const float* poses, float v,
                        float* inter_out, float* intra_out, hipStream_t st) {
    const KArgs ka = make_kargs(ls, g, na_cap, nn_cap);
    const size_t shm = lds_bytes(lanes, na_cap, nn_cap);
    UD_DISPATCH_L(lanes, {
        if (gf == UD_GRID_F16)
            hipLaunchKernelGGL((terms_kernel<L, UD_GRID_F16>), dim3(n_work), dim3(kWG), shm, st, ka, work, perm, poses,
                               v, inter_out, intra_out);
        else
            hipLaunchKernelGGL((terms_kernel<L, UD_GRID_F32>), dim3(n_work), dim3(kWG), shm, st, ka, work, perm, poses,
                               v, inter_out, intra_out);
    });
    return hipGetLastError();
}

hipError_t launch_coords(int lanes, const LigandSetDev& ls, const GridDev& g, int na_cap, int nn_cap,
                         const WorkItem* work, int n_work, const int32_t* perm, const float* poses, float* coords_out,
                         hipStream_t st) {
    const KArgs ka = make_kargs(ls, g, na_cap, nn_cap);
    const size_t shm = lds_bytes(lanes, na_cap, nn_cap);
    UD_DISPATCH_L(lanes, {
        hipLaunchKernelGGL((coords_kernel<L>), dim3(n_work), dim3(kWG), shm, st, ka, work, perm, poses, coords_out);
    });
    return hipGetLastError();
}

}  // namespace ud
