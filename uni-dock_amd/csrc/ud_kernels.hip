// HIP kernels of the batched pose search for gfx950 (wave64).
//
//   table_kernel        per-type-pair interpolation tables (precalculate)
//   grid_build_kernel   receptor affinity maps (cache::populate_no_bias)
//   grid_absmax_kernel / grid_to_half_kernel   FP16 map storage
//   score_kernel        energy + gradient of poses (model::eval_deriv)
//   search_kernel       BFGS from given starts (quasi_newton / bfgs) and,
//                       with MC, Monte Carlo chains (kernel<32, Config>)
//   terms_kernel        final-score components (Vina::score_gpu)
//   energy_grad_kernel  energy with pose-coordinate and term-weight gradients
//   coords_kernel       atom coordinates of poses (model::set)
#include <hip/hip_fp16.h>
#include <cstring>
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

// AutoDock 4.2 pair terms that depend on the two AD types only:
// w_vdw ad4_vdw + w_hb ad4_hb (potentials.h:67-79, 421-480), both with
// smoothing 0.5, cap 100000 and cutoff 8 A; a pair is an H-bond pair (12-10)
// when the product of the H-bond depths is negative, else a vdW pair (12-6).
// The charge-dependent electrostatic and desolvation terms are sampled
// separately (shared radial tables times per-pair coefficients).
__device__ float ad4_type_pair(int t1, int t2, float r, const Ad4Params& P, float w_vdw, float w_hb) {
    if (r >= 8.f) return 0.f;
    const float hbd = P.hb_depth[t1] * P.hb_depth[t2];
    const bool hb = hbd < 0.f;
    const float rij = hb ? P.hb_r[t1] + P.hb_r[t2] : P.radius[t1] + P.radius[t2];
    const float sm = 0.25f;  // smoothing / 2
    const float rs = (r > rij + sm) ? r - sm : (r < rij - sm) ? r + sm : rij;
    const float r2 = rs * rs, r4 = r2 * r2, r6 = r4 * r2, r10 = r6 * r4, r12 = r6 * r6;
    const float q2 = rij * rij, q4 = q2 * q2, q6 = q4 * q2, q10 = q6 * q4, q12 = q6 * q6;
    float val;
    if (hb) {
        const float c12 = q12 * -hbd * 10.f / 2.f, c10 = q10 * -hbd * 12.f / 2.f;
        val = (r12 > kEpsFl && r10 > kEpsFl) ? fminf(100000.f, c12 / r12 - c10 / r10) : 100000.f;
        return w_hb * val;
    }
    const float depth = sqrtf(P.depth[t1] * P.depth[t2]);
    const float c12 = q12 * depth, c6 = q6 * depth * 2.f;
    val = (r12 > kEpsFl && r6 > kEpsFl) ? fminf(100000.f, c12 / r12 - c6 / r6) : 100000.f;
    return w_vdw * val;
}

// One workgroup per type pair: smooth e at rs[k] = sqrt(k / 32), dor by
// central difference, fast = midpoint (precalculate.h:66-85, 160-171).
// The derivative table holds (e, dor) per sample (8 B); the interpolation
// forms the differences to the next sample itself, as eval_deriv does.
// AD4.2 type pairs hold the vdW / H-bond part (zero beyond 8 A).
__global__ __launch_bounds__(256) void table_kernel(int sf, SfWeights W, Ad4Params P, float2* deriv_out,
                                                    float* fast_out) {
    __shared__ float e[kTableN];
    __shared__ float rs[kTableN];
    const int slot = blockIdx.x;
    int t2 = 0;
    while ((t2 + 1) * (t2 + 2) / 2 <= slot) ++t2;
    const int t1 = slot - t2 * (t2 + 1) / 2;
    for (int k = threadIdx.x; k < kTableN; k += blockDim.x) {
        const float r = sqrtf(float(k) / 32.f);
        rs[k] = r;
        float v;
        if (sf == UD_SF_AD4) v = (t1 < kAd4Types && t2 < kAd4Types) ? ad4_type_pair(t1, t2, r, P, W.w[0], W.w[1]) : 0.f;
        else v = sf_eval(sf, t1, t2, r, W);
        e[k] = fminf(kMaxFl, v);
    }
    __syncthreads();
    auto dor = [&](int k) -> float {
        if (k == 0 || k == kTableN - 1) return 0.f;
        const float delta = rs[k + 1] - rs[k - 1];
        return (e[k + 1] - e[k - 1]) / (delta * rs[k]);
    };
    for (int k = threadIdx.x; k < kTableStride; k += blockDim.x) {
        deriv_out[size_t(slot) * kTableStride + k] = make_float2(e[k], dor(k));
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

// AD4.2 maps (the AutoGrid decomposition of the pair potentials, see
// launch_grid_build_ad4): each thread owns one sample point and up to 8
// affinity maps (blockIdx.y picks the chunk); chunk 0 also sums the
// electrostatic and desolvation maps.  Receptor atoms stream through LDS.
//   affinity[t] = sum_j w_vdw vdw + w_hb hb + w_dsolv (S_t V_j + (S_j + Q|q_j|) V_t) G(r)
//   elec        = sum_j w_elec q_j 332 min(100, 1 / (r eps(r)))
//   dsolv       = sum_j w_dsolv Q V_j G(r)
// with G(r) = exp(-r^2 / (2 3.6^2)), Q = 0.01097 and the sigmoidal
// dielectric eps(r) of ad4_electrostatic (potentials.h:350-400).
__global__ __launch_bounds__(256) void grid_build_ad4_kernel(const float4* rec, const float* rec_q, int n_rec,
                                                             GridBuildArgs a, Ad4Params P, SfWeights W,
                                                             float* maps_out) {
    __shared__ float4 tile[256];
    __shared__ float qtile[256];
    const int pt = blockIdx.x * blockDim.x + threadIdx.x;
    const int chunk = blockIdx.y;
    const int m0 = chunk * 8;
    const int nm = max(0, min(8, a.n_maps - m0));
    int tm[8];
#pragma unroll
    for (int m = 0; m < 8; ++m) tm[m] = (m < nm) ? a.types[m0 + m] : 0;
    const int x = pt % a.n[0], y = (pt / a.n[0]) % a.n[1], z = pt / (a.n[0] * a.n[1]);
    const float px = a.init[0] + a.factor_inv[0] * float(x);
    const float py = a.init[1] + a.factor_inv[1] * float(y);
    const float pz = a.init[2] + a.factor_inv[2] * float(z);
    const float cut2 = kAd4Cutoff * kAd4Cutoff;
    const float B = 78.4f + 8.5525f, lB = -B * 0.003627f;
    float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    float el = 0.f, ds = 0.f;
    for (int base = 0; base < n_rec; base += 256) {
        const int j = base + threadIdx.x;
        if (j < n_rec) {
            tile[threadIdx.x] = rec[j];
            qtile[threadIdx.x] = rec_q[j];
        }
        __syncthreads();
        const int cnt = min(256, n_rec - base);
        if (pt < a.n_points) {
            for (int k = 0; k < cnt; ++k) {
                const float4 at = tile[k];
                const float dx = at.x - px, dy = at.y - py, dz = at.z - pz;
                const float r2 = dx * dx + dy * dy + dz * dz;
                if (!(r2 < cut2)) continue;
                const float r = sqrtf(r2);
                const int tj = __float_as_int(at.w);
                const float qj = qtile[k];
                const float G = expf(-0.5f * (r / 3.6f) * (r / 3.6f));
                const float Sj = P.solv[tj] + 0.01097f * fabsf(qj);
#pragma unroll
                for (int m = 0; m < 8; ++m)
                    if (m < nm) {
                        const int t = tm[m];
                        acc[m] += ad4_type_pair(t, tj, r, P, W.w[0], W.w[1]) +
                                  W.w[3] * (P.solv[t] * P.vol[tj] + Sj * P.vol[t]) * G;
                    }
                if (chunk == 0) {
                    const float diel = -8.5525f + B / (1.f + 7.7839f * expf(lB * r));
                    const float ef = (r < kEpsFl) ? 100.f / diel : fminf(100.f, 1.f / (r * diel));
                    el += W.w[2] * qj * 332.f * ef;
                    ds += W.w[3] * 0.01097f * P.vol[tj] * G;
                }
            }
        }
        __syncthreads();
    }
    if (pt < a.n_points) {
#pragma unroll
        for (int m = 0; m < 8; ++m)
            if (m < nm) maps_out[size_t(m0 + m) * a.n_points + pt] = acc[m];
        if (chunk == 0) {
            maps_out[size_t(a.n_maps) * a.n_points + pt] = el;
            maps_out[size_t(a.n_maps + 1) * a.n_points + pt] = ds;
        }
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
// Launches whose ligands have at most 10 torsions size the search vectors 16
// (kNSmall); others use the full capacity.  With kNSmall vectors and groups
// of at most 32 lanes the inverse Hessian lives in registers, otherwise in LDS
// (full-wavefront groups are register-bound already).
// Compile-time switches of the search kernel (defaults measured fastest on
// MI355X): the occupancy targets of 32- and 16-lane groups (waves per SIMD), the
// register form of the small inverse Hessian, subtree sums from prefix
// sums, whether pass B's first record batch is loaded before pass A (off:
// the 8 registers it holds through pass A cost more than the latency saved),
// and each lane quad's torsion-tree node record kept in registers across the
// forward-kinematics levels.
#ifndef UD_WAVES_L32
#define UD_WAVES_L32 4
#endif
#ifndef UD_WAVES_L16
#define UD_WAVES_L16 4
#endif
#ifndef UD_WAVES_L64
#define UD_WAVES_L64 4
#endif
#ifndef UD_HESS_IN_REGS
#define UD_HESS_IN_REGS 1
#endif
#ifndef UD_TORQUE_SCAN
#define UD_TORQUE_SCAN 1
#endif
#ifndef UD_PREFETCH_RECORDS
#define UD_PREFETCH_RECORDS 0
#endif
#ifndef UD_FK_NODE_REGS
#define UD_FK_NODE_REGS 1
#endif

// Line-search trials per BFGS step (bfgs.h:70-83): trial k takes step
// 2^-k along the search direction.
static constexpr int kTrials = 10;

// LDS layout of one workgroup: the ligand's atoms and nodes, then one block
// per lane group (coordinates, minus-forces, node frames, five pose buffers,
// three broadcast vectors, the orientations of the line search's trials,
// inverse Hessian).  Pose and vector strides follow
// the launch's node capacity, so small ligands take little LDS per chain and
// many chains fit a compute unit.  Pair records stay in global memory.
struct Layout {
    int na, nn, nc, nh, ps;
    int o_coords, o_mf, o_frames, o_poses, o_bc, o_qt, o_H, o_pf, group_floats, shared_floats, o_apl;
    // pf: per-pair buffer (one float per pair, then a zero) for pair-parallel forces
    __host__ __device__ void init(int na_, int nn_, int np_, bool pf, int lanes) {
        na = na_;
        nn = nn_;
        nc = 6 + nn - 1;
        const bool hreg = UD_HESS_IN_REGS && nc <= kNSmall && lanes <= 32;  // inverse Hessian in registers
        ps = (7 + nn - 1 + 3) & ~3;  // pose floats (7 + torsions), >= nc
        if (hreg && ps < kNSmall) ps = kNSmall;  // broadcast vectors hold all kNSmall entries
        nh = hreg ? 0 : (nc * (nc | 1) + 3) & ~3;  // full symmetric inverse Hessian, odd row stride
        o_coords = 0;
        o_mf = 4 * na;
        o_frames = 8 * na;
        o_poses = o_frames + 20 * nn;
        o_bc = o_poses + 5 * ps;
        o_qt = o_bc + 3 * ps;
        o_H = o_qt + 4 * kTrials;
        o_pf = pf ? o_H + nh : -1;
        const int end = pf ? o_pf + ((np_ + 1 + 3) & ~3) : o_H + nh;  // + a zero entry after the pairs
        // group blocks start 4 floats apart modulo 32 banks, so that the
        // same offset in up to 8 groups falls into different LDS banks
        group_floats = ((end + 31) & ~31) + 4;
        // shared: atoms, nodes, and with the pair buffer each atom's partner
        // list descriptor (read by the second pair pass)
        o_apl = 4 * na + 8 * nn;
        shared_floats = o_apl + (pf ? (na + 3) & ~3 : 0);
    }
    __host__ __device__ size_t bytes(int groups) const { return size_t(shared_floats + groups * group_floats) * 4; }
};
enum { PX = 0, PXN = 1, PXO = 2, PTMP = 3, PBEST = 4 };

struct KArgs {
    LigandSetDev ls;
    GridDev grid;
    RecDev rec;  // explicit receptor (refinement launches only)
    int na_cap, nn_cap, np_cap;
    int pair_forces;
};

// NM: compile-time capacity of the gradient / search vectors (6 + torsions);
// each lane keeps KV = ceil(NM / L) of their entries in registers.
template <int L, int NM = kNMax> struct Grp {
    static constexpr int G = kWG / L;
    static constexpr int KV = (NM + L - 1) / L;
    const AtomRec* atoms;  // LDS (shared)
    const NodeRec* nodes;  // LDS (shared)
    const AtomPartners* apl;  // LDS (shared; with the pair buffer only)
    const PairRec* pairs;  // global memory
    float* gb;             // LDS group block
    Layout lay;
    LigHdr h;
    int gl, g;
    __device__ float4* coords() const { return reinterpret_cast<float4*>(gb + lay.o_coords); }
    __device__ float4* mf() const { return reinterpret_cast<float4*>(gb + lay.o_mf); }
    __device__ float* frames() const { return gb + lay.o_frames; }
    __device__ float* pose(int k) const { return gb + lay.o_poses + k * lay.ps; }
    __device__ float* bc(int k) const { return gb + lay.o_bc + k * lay.ps; }
    __device__ float* H() const { return gb + lay.o_H; }
    __device__ float* qt() const { return gb + lay.o_qt; }
    __device__ float* pf() const { return gb + lay.o_pf; }
    __device__ int ntors() const { return h.n_nodes - 1; }
    __device__ int ndof() const { return 6 + h.n_nodes - 1; }
};

// Stage the ligand of `lig` into LDS and set up the lane's group.
template <int L, int NM>
__device__ Grp<L, NM> setup_group(const KArgs& ka, int lig, float* lds) {
    Grp<L, NM> s;
    s.lay.init(ka.na_cap, ka.nn_cap, ka.np_cap, ka.pair_forces != 0, L);
    s.h = ka.ls.hdr[lig];
    AtomRec* atoms = reinterpret_cast<AtomRec*>(lds);
    NodeRec* nodes = reinterpret_cast<NodeRec*>(lds + 4 * s.lay.na);
    for (int i = threadIdx.x; i < s.h.n_atoms; i += kWG) atoms[i] = ka.ls.atoms[s.h.atom_begin + i];
    for (int i = threadIdx.x; i < s.h.n_nodes; i += kWG) nodes[i] = ka.ls.nodes[s.h.node_begin + i];
    AtomPartners* apl = reinterpret_cast<AtomPartners*>(lds + s.lay.o_apl);
    if (ka.pair_forces)
        for (int i = threadIdx.x; i < s.h.n_atoms; i += kWG) apl[i] = ka.ls.atom_partners[s.h.atom_begin + i];
    s.atoms = atoms;
    s.nodes = nodes;
    s.apl = apl;
    s.pairs = ka.ls.pairs + s.h.pair_begin;
    s.g = threadIdx.x / L;
    s.gl = threadIdx.x % L;
    s.gb = lds + s.lay.shared_floats + s.g * s.lay.group_floats;
    if (s.lay.o_pf >= 0 && s.gl == 0) s.pf()[s.h.n_pairs] = 0.f;  // the padding records' pair
    __syncthreads();
    return s;
}

// Forward kinematics (model::set, tree.h:121-144, 176-214): root frame from
// the pose, then tree levels parent-first.  The node frames (origin, axis,
// quaternion, rotation matrix) land in LDS, from which atoms_coords places
// the atoms.
template <int L, int NM>
__device__ void fk_frames(const Grp<L, NM>& s, const float* pose) {
    // Each node's frame is computed by the four lanes of a quad: lane c
    // (component c) does coordinate c of the origin and of the axis (c < 3),
    // quaternion component c and matrix row c (c < 3), and the lanes
    // exchange components through DPP.  Every product and sum is rounded on
    // its own in a fixed order (explicit fused multiply-adds), so that all
    // kernels compute the same frames.
#pragma clang fp contract(off)
    float* fr = s.frames();
    // the lane's quad position and what derives from it are recomputed per
    // call (the empty asm keeps them from being hoisted out of the search
    // loop, where they would hold registers for the whole kernel)
    int c = s.gl & 3;
    asm volatile("" : "+v"(c));
    const int r = c < 3 ? c : 2;
    const QuadSigns sg{uint32_t(c)};
    if (s.gl < 4) {
        // root: origin and orientation of the pose, rotation matrix
        const float q = pose[3 + c];
        float mrr, mrs, mrt;
        quad_mat_row(q, sg, mrr, mrs, mrt);
        fr[6 + c] = q;
        if (c < 3) {
            fr[c] = pose[c];
            fr[3 + c] = 0.f;
            float* m = fr + 10 + 3 * r;
            m[r] = mrr;
            m[r == 2 ? 0 : r + 1] = mrs;
            m[r == 0 ? 2 : r - 1] = mrt;
        }
    }
    // half-angle sine and cosine of each node's torsion (one node per lane),
    // which do not depend on the parent frames; they wait in the last two
    // entries of the node's frame (matrix entry 8, not yet written) until
    // its level
    constexpr int KT = (NM - 6 + L - 1) / L;
#pragma unroll
    for (int m = 0; m < KT; ++m) {
        const int k = 1 + s.gl + m * L;
        if (k < s.h.n_nodes) {
            float sn, cs;
            sincos_half(normalize_angle(pose[7 + k - 1]) * 0.5f, &sn, &cs);
            fr[20 * k + 18] = cs;
            fr[20 * k + 19] = sn;
        }
    }
    // the quad of lanes 4j..4j+3 handles nodes 1 + j + m * Q
    constexpr int Q = L / 4;
    constexpr int KN = (NM - 6 + Q - 1) / Q;
    NodeRec nd1;
    if (UD_FK_NODE_REGS) nd1 = s.nodes[min(1 + (s.gl >> 2), s.h.n_nodes - 1)];
    __syncthreads();
    for (int d = 1; d <= s.h.max_depth; ++d) {
#pragma unroll
        for (int m = 0; m < KN; ++m) {
            if (1 + m * Q >= s.h.n_nodes) break;
            const int k = 1 + (s.gl >> 2) + m * Q;
            if (k >= s.h.n_nodes) continue;
            const NodeRec nd = (UD_FK_NODE_REGS && m == 0) ? nd1 : s.nodes[k];
            if (node_depth(nd) != d) continue;
            // the quad's four lanes share k and so take this branch together
            const float* pf = fr + 20 * node_parent(nd);
            const float* R = pf + 10 + 3 * r;
            float* f = fr + 20 * k;
            const float hc = f[18], hs = f[19];
            // origin and axis coordinate r: parent origin + row r of R times
            // the relative origin, row r of R times the relative axis
            const float o = pf[r] + fmaf(R[2], nd.ro[2], fmaf(R[1], nd.ro[1], R[0] * nd.ro[0]));
            const float ac = fmaf(R[2], nd.ax[2], fmaf(R[1], nd.ax[1], R[0] * nd.ax[0]));
            const float ax = quad_perm<0, 0, 0, 0>(ac), ay = quad_perm<1, 1, 1, 1>(ac), az = quad_perm<2, 2, 2, 2>(ac);
            // component c of angle_to_quaternion(axis, angle) (quaternion.cpp:37-47)
            // times the parent's: sum over j of +-qa_j qp_(c xor j)
            const float qp = pf[6 + c];
            const float q1 = quad_perm<1, 0, 3, 2>(qp), q2 = quad_perm<2, 3, 0, 1>(qp), q3 = quad_perm<3, 2, 1, 0>(qp);
            const float p1 = sign_xor(hs, sg.e1) * ax, p2 = sign_xor(hs, sg.e2) * ay, p3 = sign_xor(hs, sg.e3) * az;
            float q = fmaf(p3, q3, fmaf(p2, q2, fmaf(p1, q1, hc * qp)));
            // quaternion_normalize_approx (quaternion.h:105-117)
            const float sq = q * q;
            float n2 = sq + quad_perm<1, 0, 3, 2>(sq);
            n2 = n2 + quad_perm<2, 3, 0, 1>(n2);
            if (!(fabsf(n2 - 1.f) < 1e-6f)) q = q * __builtin_amdgcn_rsqf(n2);
            float mrr, mrs, mrt;
            quad_mat_row(q, sg, mrr, mrs, mrt);
            f[6 + c] = q;
            if (c < 3) {
                f[r] = o;
                f[3 + r] = ac;
                float* mm = f + 10 + 3 * r;
                mm[r] = mrr;
                mm[r == 2 ? 0 : r + 1] = mrs;
                mm[r == 0 ? 2 : r - 1] = mrt;
            }
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

// grid::evaluate_aux (grid.cpp:39-129) + curl (curl.h:28-36), split in three
// steps so that the corner gathers of several atoms are in flight together:
// locate (cell, fractions, out-of-box penalty), gather (8 loads), finish.
// curl's factor v / (v + e) (curl.h:28-36) with the hardware reciprocal
// (1 ulp), as the reference's GPU build divides (--use_fast_math)
__device__ __forceinline__ float curl_factor(float v, float e) {
    return (v < kEpsFl) ? 0.f : v * __builtin_amdgcn_rcpf(v + e);
}

struct GridPt {
    uint32_t base;  // map element of corner 000 (maps hold fewer than 2^32 values)
    float x, y, z, penalty;
    int reg[3];
};

// grid coordinates of p (in units of the spacing, from the box's first sample)
__device__ __forceinline__ void grid_scale(const GridDev& g, float3 p, float (&s)[3]) {
    s[0] = (p.x - g.init[0]) * g.factor[0];
    s[1] = (p.y - g.init[1]) * g.factor[1];
    s[2] = (p.z - g.init[2]) * g.factor[2];
}
// inside the box on every axis (0 <= s < n - 1), or NaN
__device__ __forceinline__ bool grid_inside(const GridDev& g, const float (&s)[3]) {
    return !(s[0] < 0.f || s[0] >= g.dimm1[0] || s[1] < 0.f || s[1] >= g.dimm1[1] || s[2] < 0.f || s[2] >= g.dimm1[2]);
}
// grid_locate for a point inside the box: the same cell and fractions, no
// penalty (a NaN coordinate converts to cell 0 and keeps a NaN fraction)
__device__ __forceinline__ GridPt grid_locate_in(const GridDev& g, int slot, const float (&s)[3]) {
    GridPt q;
    int a[3];
    float f[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        a[i] = min(max(int(s[i]), 0), g.n[i] - 2);
        f[i] = s[i] - float(a[i]);
        q.reg[i] = 0;
    }
    q.penalty = 0.f;
    q.base = uint32_t(slot) * uint32_t(g.n_points) + uint32_t(a[0] + g.n[0] * (a[1] + g.n[1] * a[2]));
    q.x = f[0]; q.y = f[1]; q.z = f[2];
    return q;
}

__device__ __forceinline__ GridPt grid_locate(const GridDev& g, int slot, float3 p) {
    float s[3];
    grid_scale(g, p, s);
    float miss[3];
    int a[3];
    GridPt q;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        // below the box: corner 0, fraction 0; at or beyond its last sample:
        // cell n - 2, fraction 1.  Branch-free: a clamp (v_med3) and selects.
        // The out-of-box distance is selected, not formed as |s - clamp(s)|,
        // which FMA contraction with s's product would make a rounding
        // residue (times the 1e6 penalty slope) inside the box.
        const float si = s[i];
        const float sc = __builtin_amdgcn_fmed3f(si, 0.f, g.dimm1[i]);
        const bool hi = si >= g.dimm1[i];
        miss[i] = hi ? si - g.dimm1[i] : (si < 0.f ? -si : 0.f);
        q.reg[i] = hi ? 1 : (si < 0.f ? -1 : 0);
        a[i] = min(max(int(sc), 0), g.n[i] - 2);  // a NaN coordinate must not index outside the map
        // the fraction is formed from s itself, so that a NaN coordinate
        // still gives a NaN energy (which the line search rejects)
        s[i] = hi ? 1.f : (si < 0.f ? 0.f : si - float(a[i]));
    }
    q.penalty = g.slope * (miss[0] * g.factor_inv[0] + miss[1] * g.factor_inv[1] + miss[2] * g.factor_inv[2]);
    q.base = uint32_t(slot) * uint32_t(g.n_points) + uint32_t(a[0] + g.n[0] * (a[1] + g.n[1] * a[2]));
    q.x = s[0]; q.y = s[1]; q.z = s[2];
    return q;
}

// Storage format (UD_GRID_F32 / F16) and scoring family of a grid-format
// template value: 0 / 1 maps, 2 the explicit receptor, 3 / 4 AD4.2 maps.
template <int GF> struct Fmt {
    static constexpr int store = (GF >= kGridAd4F32) ? GF - kGridAd4F32 : GF;
    static constexpr bool ad4 = GF >= kGridAd4F32;
};

// Load of a T at a 32-bit byte offset from a wave-uniform base: the address
// is the base (scalar registers) plus a zero-extended 32-bit offset, which
// the global load instructions take directly (constant parts become their
// immediate offset), instead of 64-bit address arithmetic per load.  Every
// buffer read this way is below 4 GiB (maps at most 2^24 points x 32 maps x 4
// B, tables and ligand records far less).
template <class T> __device__ __forceinline__ T ld_at(const void* base, uint32_t byte_off) {
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + byte_off);
}
// two consecutive (e, dor) samples of a pair table in one 16-byte load (the
// table's entries are 8-byte aligned; global memory takes unaligned vectors)
typedef float f4a8 __attribute__((ext_vector_type(4), aligned(8)));
__device__ __forceinline__ void ld_samples(const float2* table, uint32_t idx, float2& t0, float2& t1) {
    const f4a8 v = ld_at<f4a8>(table, idx * 8u);
    t0 = make_float2(v.x, v.y);
    t1 = make_float2(v.z, v.w);
}

template <int GF>
__device__ __forceinline__ void grid_gather(const GridDev& g, uint32_t base, float (&f)[8]) {
    constexpr uint32_t B = (Fmt<GF>::store == UD_GRID_F32) ? 4u : 2u;  // bytes per value
    const uint32_t sy = uint32_t(g.n[0]) * B, sz = uint32_t(g.n[0] * g.n[1]) * B;
    const uint32_t o0 = base * B, o1 = o0 + sy, o2 = o0 + sz, o3 = o2 + sy;
    // the x-neighbours of each corner pair are adjacent values: one load per
    // pair (at the value's alignment, which global loads accept)
    if (Fmt<GF>::store == UD_GRID_F32) {
        typedef float f2a4 __attribute__((ext_vector_type(2), aligned(4)));
        const uint32_t o[4] = {o0, o1, o2, o3};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const f2a4 v = ld_at<f2a4>(g.maps, o[k]);
            f[2 * k] = v.x;
            f[2 * k + 1] = v.y;
        }
    } else {
        typedef uint32_t u32a2 __attribute__((aligned(2)));
        const uint32_t w0 = ld_at<u32a2>(g.maps, o0), w1 = ld_at<u32a2>(g.maps, o1);
        const uint32_t w2 = ld_at<u32a2>(g.maps, o2), w3 = ld_at<u32a2>(g.maps, o3);
        const uint32_t w[4] = {w0, w1, w2, w3};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            f[2 * k] = __half2float(__ushort_as_half(uint16_t(w[k] & 0xFFFFu)));
            f[2 * k + 1] = __half2float(__ushort_as_half(uint16_t(w[k] >> 16)));
        }
    }
}

// corners f = (000, 100, 010, 110, 001, 101, 011, 111): the trilinear value
// of grid::evaluate_aux (grid.cpp:39-129) and its gradient, formed as nested
// linear interpolations (x, then y, then z) so that value and gradient share
// their partial sums; FP16 maps apply the power-of-two map scale to value and
// gradient before the curl
template <int GF, bool INBOX = false>
__device__ __forceinline__ float grid_finish(const GridDev& g, int slot, const GridPt& q, const float (&f)[8], float v,
                                             float3* d) {
    const float x = q.x, y = q.y, z = q.z;
    const float dx0 = f[1] - f[0], dx1 = f[3] - f[2], dx2 = f[5] - f[4], dx3 = f[7] - f[6];
    const float a0 = f[0] + x * dx0, a1 = f[2] + x * dx1, a2 = f[4] + x * dx2, a3 = f[6] + x * dx3;
    const float dy0 = a1 - a0, dy1 = a3 - a2;
    const float b0 = a0 + y * dy0, b1 = a2 + y * dy1;
    const float dz = b1 - b0;
    float e = b0 + z * dz;
    if (Fmt<GF>::store == UD_GRID_F16) e *= g.scale[slot];
    if (d) {
        const float ex0 = dx0 + y * (dx1 - dx0), ex1 = dx2 + y * (dx3 - dx2);
        float gx = ex0 + z * (ex1 - ex0);
        float gy = dy0 + z * (dy1 - dy0);
        float gz = dz;
        if (Fmt<GF>::store == UD_GRID_F16) {
            const float sc = g.scale[slot];
            gx *= sc; gy *= sc; gz *= sc;
        }
        if constexpr (INBOX) {
            // curl as a factor of 1 where it does not apply (exact), without
            // a branch
            const float tmp = (e > 0.f && v < 0.1f * kMaxFl) ? curl_factor(v, e) : 1.f;
            e = __fmul_rn(e, tmp);
            const float t2 = tmp * tmp;
            gx *= t2; gy *= t2; gz *= t2;
        } else if (e > 0.f && v < 0.1f * kMaxFl) {
            // (conditional here: the out-of-box penalty is added to this
            // value, and AD4.2 adds charge-scaled copies of it whose sums
            // cancel, so the rounding of that sum is kept as it is)
            const float tmp = curl_factor(v, e);
            e *= tmp;
            const float t2 = tmp * tmp;
            gx *= t2; gy *= t2; gz *= t2;
        }
        if (INBOX) {
            d->x = g.factor[0] * gx;
            d->y = g.factor[1] * gy;
            d->z = g.factor[2] * gz;
        } else {
            d->x = g.factor[0] * (q.reg[0] == 0 ? gx : 0.f) + g.slope * float(q.reg[0]);
            d->y = g.factor[1] * (q.reg[1] == 0 ? gy : 0.f) + g.slope * float(q.reg[1]);
            d->z = g.factor[2] * (q.reg[2] == 0 ? gz : 0.f) + g.slope * float(q.reg[2]);
        }
    } else if (e > 0.f && v < 0.1f * kMaxFl) {
        e *= curl_factor(v, e);
    }
    return INBOX ? e : e + q.penalty;
}

// ad4cache::eval_deriv (ad4cache.cpp:223-262) for an atom of charge q whose
// affinity map is `slot`: that map, plus q times the electrostatic map and
// |q| times the desolvation map, each through grid::evaluate with its own
// curl and out-of-box penalty; the three share the cell and fractions.
template <int GF>
__device__ __forceinline__ float ad4_grid_term(const GridDev& g, int slot, float q, float3 p, float v, float3* d) {
    const GridPt gq = grid_locate(g, slot, p);
    const uint32_t cell = gq.base - uint32_t(slot) * uint32_t(g.n_points);
    float f[8];
    grid_gather<GF>(g, gq.base, f);
    float3 dt = make_float3(0.f, 0.f, 0.f);
    float e = grid_finish<GF>(g, slot, gq, f, v, d ? &dt : nullptr);
    const float fac[2] = {q, fabsf(q)};
    const int sl[2] = {g.elec_slot, g.dsolv_slot};
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        grid_gather<GF>(g, cell + uint32_t(sl[k]) * uint32_t(g.n_points), f);
        float3 dk = make_float3(0.f, 0.f, 0.f);
        e += fac[k] * grid_finish<GF>(g, sl[k], gq, f, v, d ? &dk : nullptr);
        dt.x += fac[k] * dk.x; dt.y += fac[k] * dk.y; dt.z += fac[k] * dk.z;
    }
    if (d) *d = dt;
    return e;
}

// Grid-format value of the refinement launches: the inter term is scored
// against the explicit receptor atoms instead of the maps.
static constexpr int kGridExplicit = 2;

// Intramolecular pair cutoff^2 (sqr(sf.get_cutoff()), scoring_function.h:56-83)
template <int GF> __device__ __forceinline__ constexpr float pair_cut2() {
    return Fmt<GF>::ad4 ? kAd4Cutoff * kAd4Cutoff : 64.f;
}

// non_cache::eval_deriv / non_cache::eval (non_cache.cpp:420-860, 34-255)
// for one ligand atom of XS type t at p: the coordinate is clamped into the
// box (penalty slope * distance outside, gradient slope * sign there), every
// receptor atom of the clamped point's cell within the cutoff contributes its
// type-pair table value (interpolated e and dor with a gradient, the fast
// table without), then curl; the pair gradient is kept in clamped directions
// as in the reference.  The type pair uses t itself (get_type_pair_index).
// The atom's clamped position, out-of-box penalty and receptor cell list.
struct ExplicitAtom {
    float a[3], oob[3], pen;
    int k0, k1;
};
__device__ __forceinline__ ExplicitAtom explicit_locate(const KArgs& ka, float3 p) {
    const RecDev& r = ka.rec;
    ExplicitAtom x;
    x.a[0] = p.x; x.a[1] = p.y; x.a[2] = p.z;
    x.pen = 0.f;
    int c[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        x.oob[i] = 0.f;
        if (x.a[i] < r.init[i]) {
            x.pen += r.init[i] - x.a[i];
            x.a[i] = r.init[i];
            x.oob[i] = -1.f;
        } else if (x.a[i] > r.end[i]) {
            x.pen += x.a[i] - r.end[i];
            x.a[i] = r.end[i];
            x.oob[i] = 1.f;
        }
        // szv_grid::possibilities (szv_grid.cpp:65-74) with fl_to_sz (common.h:205-211)
        const float xs = (x.a[i] - r.init[i]) * float(r.dim[i]) / r.range[i];
        c[i] = (xs <= 0.f) ? 0 : (xs >= float(r.dim[i] - 1)) ? r.dim[i] - 1 : max(0, min(int(xs), r.dim[i] - 1));
    }
    const int cell = c[0] + r.dim[0] * (c[1] + r.dim[1] * c[2]);
    x.k0 = r.cell_start[cell];
    x.k1 = r.cell_start[cell + 1];
    return x;
}
// Sums over the receptor atoms of batches part, part + parts, ... of the
// atom's cell list (kRecBatch atoms each): energy and the gradient's sum of
// dor * (a - b) (with d == false the fast table's energy only).  Each batch
// is loaded while the previous one is summed and its table samples loaded
// together, all unconditionally (atoms past the cell's list are the next
// cell's or the array's far padding and are masked; sample 0 outside the
// cutoff), so that each batch waits only for its own loads.
static constexpr int kRecBatch = 4;
static_assert(kCellPad >= 9 * kRecBatch, "look-ahead of up to 8 interleaved parts");
__device__ __forceinline__ void explicit_sum(const KArgs& ka, int t, const ExplicitAtom& x, int part, int parts, bool d,
                                             float& e, float& gx, float& gy, float& gz) {
    const RecDev& r = ka.rec;
    const int step = parts * kRecBatch;
    float4 b[kRecBatch];
#pragma unroll
    for (int u = 0; u < kRecBatch; ++u) b[u] = r.atoms[x.k0 + part * kRecBatch + u];
    for (int k = x.k0 + part * kRecBatch; k < x.k1; k += step) {
        float4 nb[kRecBatch];
#pragma unroll
        for (int u = 0; u < kRecBatch; ++u) nb[u] = r.atoms[k + step + u];
        bool in[kRecBatch];
        float rx[kRecBatch], ry[kRecBatch], rz[kRecBatch], rf[kRecBatch];
        uint32_t idx[kRecBatch];
#pragma unroll
        for (int u = 0; u < kRecBatch; ++u) {
            rx[u] = x.a[0] - b[u].x; ry[u] = x.a[1] - b[u].y; rz[u] = x.a[2] - b[u].z;
            const float r2 = rx[u] * rx[u] + ry[u] * ry[u] + rz[u] * rz[u];
            in[u] = (k + u < x.k1) && r2 < 64.f;
            const int t2 = __float_as_int(b[u].w);
            rf[u] = 32.f * r2;
            idx[u] = in[u] ? uint32_t(t <= t2 ? t + t2 * (t2 + 1) / 2 : t2 + t * (t + 1) / 2) * uint32_t(kTableStride) +
                                 uint32_t(int(rf[u]))
                           : 0u;
        }
        if (d) {
            float2 t0[kRecBatch], t1[kRecBatch];
#pragma unroll
            for (int u = 0; u < kRecBatch; ++u) { t0[u] = ka.ls.table_deriv[idx[u]]; t1[u] = ka.ls.table_deriv[idx[u] + 1]; }
#pragma unroll
            for (int u = 0; u < kRecBatch; ++u)
                if (in[u]) {
                    const float rem = rf[u] - float(int(rf[u]));
                    e += t0[u].x + rem * (t1[u].x - t0[u].x);
                    const float dor = t0[u].y + rem * (t1[u].y - t0[u].y);
                    gx += dor * rx[u]; gy += dor * ry[u]; gz += dor * rz[u];
                }
        } else {
            float tf[kRecBatch];
#pragma unroll
            for (int u = 0; u < kRecBatch; ++u) tf[u] = ka.ls.table_fast[idx[u]];
#pragma unroll
            for (int u = 0; u < kRecBatch; ++u)
                if (in[u]) e += tf[u];
        }
#pragma unroll
        for (int u = 0; u < kRecBatch; ++u) b[u] = nb[u];
    }
}
// Curl of the summed energy, then the out-of-box penalty and its gradient.
// slope < 0: the launch's (ka.grid.slope).
__device__ __forceinline__ float explicit_finish(const KArgs& ka, const ExplicitAtom& x, float v, float e, float gx,
                                                 float gy, float gz, float3* d, float slope_in = -1.f) {
    const float slope = slope_in < 0.f ? ka.grid.slope : slope_in;
    if (e > 0.f && v < 0.1f * kMaxFl) {
        const float tmp = curl_factor(v, e);
        e *= tmp;
        const float t2 = tmp * tmp;
        gx *= t2; gy *= t2; gz *= t2;
    }
    if (d) *d = make_float3(gx + slope * x.oob[0], gy + slope * x.oob[1], gz + slope * x.oob[2]);
    return e + slope * x.pen;
}
__device__ __forceinline__ float explicit_term(const KArgs& ka, int t, float3 p, float v, float3* d,
                                               float slope = -1.f) {
    const ExplicitAtom x = explicit_locate(ka, p);
    float e = 0.f, gx = 0.f, gy = 0.f, gz = 0.f;
    explicit_sum(ka, t, x, 0, 1, d != nullptr, e, gx, gy, gz);
    return explicit_finish(ka, x, v, e, gx, gy, gz, d, slope);
}

// Compile-time tuning of the search kernel: load batching and the occupancy
// target handed to the register allocator (waves per SIMD; 1 = no target).
// Full-wavefront, 32- and 16-lane groups target 4 waves per SIMD (UD_WAVES_L32,
// UD_WAVES_L16), 8-lane groups UD_SEARCH_WAVES_PER_EU; one pair per lane per
// table-load batch: measured fastest so.
#ifndef UD_ATOM_BATCH
#define UD_ATOM_BATCH 1
#endif
#ifndef UD_PAIR_PIPELINE
#define UD_PAIR_PIPELINE 1  // pass A's table loads one batch ahead of its math
#endif
#ifndef UD_PAIR_BATCH
#define UD_PAIR_BATCH 1
#endif
#ifndef UD_SEARCH_WAVES_PER_EU
#define UD_SEARCH_WAVES_PER_EU 3
#endif
static constexpr int kAtomBatch = UD_ATOM_BATCH;  // atoms per lane whose corner gathers are issued together
static constexpr int kPairBatch = UD_PAIR_BATCH;  // pairs per lane whose table loads are issued together

// Atom coordinates from the node frames (model::set, tree.h:176-214), lane
// gl handling atoms gl, gl + L, ...; also the heavy-atom sum of squared
// distances to the root origin (model::gyration_radius, model.cpp:871-883).
template <int L, int NM>
__device__ __forceinline__ void atoms_coords(const Grp<L, NM>& s, float& rg2, bool want_rg = true) {
    float4* coords = s.coords();
    const float* fr = s.frames();
    for (int i = s.gl; i < s.h.n_atoms; i += L) {
        const AtomRec a = s.atoms[i];
        const float3 p = frame_apply(fr + 20 * (a.meta & 0xff), a);
        coords[i] = make_float4(p.x, p.y, p.z, 0.f);
        if (want_rg && ((a.meta >> 16) & 1)) {
            const float dx = p.x - fr[0], dy = p.y - fr[1], dz = p.z - fr[2];
            rg2 += dx * dx + dy * dy + dz * dz;
        }
    }
}

// AD4.2 charge terms of one pair at fraction `rem` between samples i1 and
// i1 + 1 of the shared radial tables: (q_a q_b) times the electrostatic
// sample and the desolvation coefficient times the Gaussian sample, added to
// the type-pair table's (e, dor).  The reference interpolates one table of
// the pair's total (precalculate_element::eval_deriv, precalculate.h:48-65);
// the total, its samples and its central differences are linear in these
// parts, so the sum is the same interpolation.
__device__ __forceinline__ void ad4_charge_terms(float2 cf, float4 u0, float4 u1, float rem, float& pe, float& dor) {
    pe += cf.x * (u0.x + rem * (u1.x - u0.x)) + cf.y * (u0.z + rem * (u1.z - u0.z));
    dor += cf.x * (u0.y + rem * (u1.y - u0.y)) + cf.y * (u0.w + rem * (u1.w - u0.w));
}

// Energy and gradient of each of the lane's atoms (atoms gl, gl + L, ...):
// the grid term (cache::eval_deriv, cache.cpp:160-199) plus every
// intramolecular pair the atom belongs to (eval_interacting_pairs_deriv,
// model.cpp:907-934, with precalculate_element::eval_deriv,
// precalculate.h:48-65).  Each pair is evaluated by the owners of both of
// its atoms; each keeps the force on its own atom (the two results are exact
// negatives of each other), so every gradient is written once by one lane
// and no atomics are needed.  A pair's energy is counted on the side of its
// lower atom index.
template <int L, int NM, int GF>
__device__ __forceinline__ float atoms_forces_2s(const KArgs& ka, const Grp<L, NM>& s, float v_intra, float v_inter,
                                                 float slope = -1.f) {
    const float4* coords = s.coords();
    float4* mf = s.mf();
    const PartnerRec* part = ka.ls.partners + s.h.partner_begin;
    const AtomPartners* apl = ka.ls.atom_partners + s.h.atom_begin;
    float e = 0.f;
    for (int i = s.gl; i < s.h.n_atoms; i += L) {
        const AtomRec a = s.atoms[i];
        const float4 ci = coords[i];
        const int slot = (a.meta >> 8) & 0xff;
        GridPt gq;
        float gf[8];
        if constexpr (GF != kGridExplicit && !Fmt<GF>::ad4) {
            if (slot != 0xff) {
                gq = grid_locate(ka.grid, slot, make_float3(ci.x, ci.y, ci.z));
                grid_gather<GF>(ka.grid, gq.base, gf);
            }
        }
        const uint32_t ap = apl[i];
        const int k0 = int(ap & 0xFFFFFu), kn = int(ap >> 20);
        float gx = 0.f, gy = 0.f, gz = 0.f;
        for (int k = 0; k < kn; k += kPairBatch) {
            int other[kPairBatch], idx[kPairBatch], it[kPairBatch], pq[kPairBatch];
            float rx[kPairBatch], ry[kPairBatch], rz[kPairBatch], rem[kPairBatch];
            bool in[kPairBatch];
            float2 t0[kPairBatch], t1[kPairBatch], cf[kPairBatch];
            float4 u0[kPairBatch], u1[kPairBatch];
#pragma unroll
            for (int u = 0; u < kPairBatch; ++u) {
                in[u] = false;
                other[u] = 0; idx[u] = 0; it[u] = 0; pq[u] = 0; rem[u] = 0.f; rx[u] = ry[u] = rz[u] = 0.f;
                if (k + u < kn) {
                    const PartnerRec rec = ld_at<uint32_t>(part, uint32_t(k0 + k + u) * 4u);
                    other[u] = int(partner_other(rec));
                    const float4 cb = coords[other[u]];
                    rx[u] = cb.x - ci.x; ry[u] = cb.y - ci.y; rz[u] = cb.z - ci.z;
                    const float r2 = rx[u] * rx[u] + ry[u] * ry[u] + rz[u] * rz[u];
                    if (r2 < pair_cut2<GF>()) {
                        const float rf = 32.f * r2;
                        const int i1t = int(rf);
                        rem[u] = rf - float(i1t);
                        idx[u] = int(partner_slot(rec)) * kTableStride + i1t;
                        it[u] = i1t;
                        pq[u] = int(partner_pair(rec));
                        in[u] = true;
                    }
                }
            }
#pragma unroll
            for (int u = 0; u < kPairBatch; ++u)
                if (in[u]) {
                    if (!Fmt<GF>::ad4 || it[u] < kTableStride - 1) {
                        ld_samples(ka.ls.table_deriv, uint32_t(idx[u]), t0[u], t1[u]);
                    } else {
                        t0[u] = t1[u] = make_float2(0.f, 0.f);
                    }
                    if constexpr (Fmt<GF>::ad4) {
                        u0[u] = ld_at<float4>(ka.ls.ad4_deriv, uint32_t(it[u]) * 16u);
                        u1[u] = ld_at<float4>(ka.ls.ad4_deriv, uint32_t(it[u]) * 16u + 16u);
                        cf[u] = ka.ls.pair_coef[s.h.pair_begin + pq[u]];
                    }
                }
#pragma unroll
            for (int u = 0; u < kPairBatch; ++u) {
                if (!in[u]) continue;
                float pe = t0[u].x + rem[u] * (t1[u].x - t0[u].x);
                float dor = t0[u].y + rem[u] * (t1[u].y - t0[u].y);
                if constexpr (Fmt<GF>::ad4) ad4_charge_terms(cf[u], u0[u], u1[u], rem[u], pe, dor);
                float fx = dor * rx[u], fy = dor * ry[u], fz = dor * rz[u];
                if (pe > 0.f && v_intra < 0.1f * kMaxFl) {
                    const float tmp = curl_factor(v_intra, pe);
                    pe *= tmp;
                    const float t2 = tmp * tmp;
                    fx *= t2; fy *= t2; fz *= t2;
                }
                if (other[u] > i) e += pe;
                gx -= fx; gy -= fy; gz -= fz;
            }
        }
        float3 d = make_float3(0.f, 0.f, 0.f);
        if constexpr (GF == kGridExplicit) {
            if (slot != 0xff) e += explicit_term(ka, int(a.meta >> 24), make_float3(ci.x, ci.y, ci.z), v_inter, &d, slope);
        } else if constexpr (Fmt<GF>::ad4) {
            if (slot != 0xff)
                e += ad4_grid_term<GF>(ka.grid, slot, ka.ls.charge[s.h.atom_begin + i], make_float3(ci.x, ci.y, ci.z),
                                       v_inter, &d);
        } else {
            if (slot != 0xff) e += grid_finish<GF>(ka.grid, slot, gq, gf, v_inter, &d);
        }
        mf[i] = make_float4(d.x + gx, d.y + gy, d.z + gz, 0.f);
    }
    return e;
}

// Same result as atoms_forces_2s with every pair evaluated once, in two
// passes.  Pass A (atoms_forces_pp_a) evaluates the grid term of the lane's
// atoms and the pairs gl, gl + L, ..., storing each pair's force scale d (dor
// with the curl factor; the force on the pair's second atom is d (c_b - c_a))
// in the group's per-pair buffer; it gives the energy.  Pass B
// (atoms_forces_pp_b), needed only for the gradient, adds to each atom's grid
// gradient d (c_i - c_other) for its pairs in pair order (forces[a] -= f,
// forces[b] += f, model.cpp:923-931; the difference is formed from the same
// coordinates, so both ends see exact negatives).  Lanes stay balanced
// whatever the atoms' pair counts, and the per-atom accumulation order is the
// reference's.
template <int L, int NM, int GF>
__device__ __forceinline__ float atoms_forces_pp_a(const KArgs& ka, const Grp<L, NM>& s, float v_intra, float v_inter,
                                                   const uint64_t (&rec0)[kPairBatch], float slope = -1.f) {
    const float4* coords = s.coords();
    float4* mf = s.mf();
    float* pf = s.pf();
    const int N = s.h.n_atoms;
    // pass A's pair batches: r, table index and fraction, then the sample
    // loads (issue), then energy and force scale into pf (finish); the first
    // batch is issued before the grid term so that its table loads and the
    // map gathers are in flight together
    const int P = s.h.n_pairs;
    struct Batch {
        float rem[kPairBatch];
        int idx[kPairBatch], it[kPairBatch];
        bool in[kPairBatch];
        float2 t0[kPairBatch], t1[kPairBatch], cf[kPairBatch];
        float4 u0[kPairBatch], u1[kPairBatch];
    } b, b2;
    // pre-decoded pair records (coordinate byte offsets, table base), each
    // batch's loaded one batch ahead
    // The lane's next record is read through a per-lane address in vector
    // registers, advanced by one batch per fetch: the uniform base would
    // otherwise compete for scalar registers with the search state.
    typedef const uint64_t __attribute__((address_space(1)))* GlobalRecPtr;  // global memory (not flat) loads
    GlobalRecPtr rp = (GlobalRecPtr)(ka.ls.pair_ofs + s.h.pair_begin + s.gl);
    asm volatile("" : "+v"(rp));
    const char* cbase = reinterpret_cast<const char*>(coords);
    uint2 nrec[kPairBatch];
    auto fetch = [&](int q0) {
#pragma unroll
        for (int u = 0; u < kPairBatch; ++u) {
            // (past the ligand's last pair this reads the next ligand's
            // records or the array's zero padding; they are not used)
            const uint64_t r = rp[u * L];
            nrec[u] = make_uint2(uint32_t(r), uint32_t(r >> 32));
        }
        rp += kPairBatch * L;
    };
    auto issue = [&](Batch& b, int q0) {
        uint2 rec[kPairBatch];
#pragma unroll
        for (int u = 0; u < kPairBatch; ++u) rec[u] = nrec[u];
        fetch(q0 + kPairBatch * L);
#pragma unroll
        for (int u = 0; u < kPairBatch; ++u) {
            const int q = q0 + u * L;
            b.in[u] = false;  // rem and it are read only for pairs within the cutoff
            b.idx[u] = 0;
            if (q < P) {
                const float4 ca = *reinterpret_cast<const float4*>(cbase + (rec[u].x & 0xFFFFu));
                const float4 cb = *reinterpret_cast<const float4*>(cbase + (rec[u].x >> 16));
                const float rx = cb.x - ca.x, ry = cb.y - ca.y, rz = cb.z - ca.z;
                const float r2 = rx * rx + ry * ry + rz * rz;
                if (r2 < pair_cut2<GF>()) {
                    const float rf = 32.f * r2;
                    const int i1t = int(rf);
                    b.rem[u] = rf - float(i1t);
                    b.idx[u] = int(rec[u].y) + i1t;
                    b.it[u] = i1t;
                    b.in[u] = true;
                }
            }
        }
#pragma unroll
        for (int u = 0; u < kPairBatch; ++u)
            if constexpr (!Fmt<GF>::ad4) {
                // every lane loads (sample 0 of table 0 outside the cutoff),
                // so that no load sits in a branch and each batch waits
                // only for its own loads
                ld_samples(ka.ls.table_deriv, uint32_t(b.idx[u]), b.t0[u], b.t1[u]);
            } else if (b.in[u]) {
                if (b.it[u] < kTableStride - 1) {
                    ld_samples(ka.ls.table_deriv, uint32_t(b.idx[u]), b.t0[u], b.t1[u]);
                } else {
                    b.t0[u] = b.t1[u] = make_float2(0.f, 0.f);
                }
                if constexpr (Fmt<GF>::ad4) {
                    b.u0[u] = ld_at<float4>(ka.ls.ad4_deriv, uint32_t(b.it[u]) * 16u);
                    b.u1[u] = ld_at<float4>(ka.ls.ad4_deriv, uint32_t(b.it[u]) * 16u + 16u);
                    b.cf[u] = ka.ls.pair_coef[s.h.pair_begin + q0 + u * L];
                }
            }
    };
    auto finish = [&](const Batch& b, int q0) {
        float ep = -0.f;  // -0 + pe is pe: no addition for the first pair
#pragma unroll
        for (int u = 0; u < kPairBatch; ++u) {
            const int q = q0 + u * L;
            if (q >= P) continue;
            float d = 0.f;
            if (b.in[u]) {
                float pe = b.t0[u].x + b.rem[u] * (b.t1[u].x - b.t0[u].x);
                d = b.t0[u].y + b.rem[u] * (b.t1[u].y - b.t0[u].y);
                if constexpr (Fmt<GF>::ad4) ad4_charge_terms(b.cf[u], b.u0[u], b.u1[u], b.rem[u], pe, d);
                // curl, as a factor of 1 where it does not apply (exact); the
                // product is rounded before it is summed, as a conditional
                // multiply would be (no fused multiply-add into the sum)
                const float tmp = (pe > 0.f && v_intra < 0.1f * kMaxFl) ? curl_factor(v_intra, pe) : 1.f;
                pe = __fmul_rn(pe, tmp);
                d *= tmp * tmp;
                ep += pe;
            }
            pf[q] = d;
        }
        return ep;
    };
    // the first batch's records, loaded before the forward kinematics
#pragma unroll
    for (int u = 0; u < kPairBatch; ++u) nrec[u] = make_uint2(uint32_t(rec0[u]), uint32_t(rec0[u] >> 32));
    rp += kPairBatch * L;
    issue(b, s.gl);
    float e = 0.f;
    if constexpr (GF == kGridExplicit) {
        // the explicit receptor: S lanes per atom (the largest power of two
        // up to 8 with S * N <= L) share its receptor list, batch by batch in
        // turn, and add their sums before the curl
        int S = 1;
        while (S < 8 && 2 * S * N <= L) S *= 2;
        const int per = L / S, ia = s.gl / S, jj = s.gl % S;
        for (int i0 = 0; i0 < N; i0 += per) {
            const int i = i0 + ia;
            const bool have = i < N;
            const AtomRec a = s.atoms[have ? i : 0];
            const bool typed = have && ((a.meta >> 8) & 0xff) != 0xff;
            const float4 ci = coords[have ? i : 0];
            const ExplicitAtom x = explicit_locate(ka, make_float3(ci.x, ci.y, ci.z));
            float ep = 0.f, gx = 0.f, gy = 0.f, gz = 0.f;
            if (typed) explicit_sum(ka, int(a.meta >> 24), x, jj, S, true, ep, gx, gy, gz);
            ep = sub_sum(ep, S); gx = sub_sum(gx, S); gy = sub_sum(gy, S); gz = sub_sum(gz, S);
            if (have && jj == 0) {
                float3 d = make_float3(0.f, 0.f, 0.f);
                if (typed) e += explicit_finish(ka, x, v_inter, ep, gx, gy, gz, &d, slope);
                mf[i] = make_float4(d.x, d.y, d.z, 0.f);
            }
        }
    }
    for (int i = s.gl; i < N; i += L) {
        if constexpr (GF == kGridExplicit) break;  // done above
        const AtomRec a = s.atoms[i];
        const int slot = (a.meta >> 8) & 0xff;
        float3 d = make_float3(0.f, 0.f, 0.f);
        if (slot != 0xff) {
            const float4 ci = coords[i];
            if constexpr (GF == kGridExplicit) {
            } else if constexpr (Fmt<GF>::ad4) {
                e += ad4_grid_term<GF>(ka.grid, slot, ka.ls.charge[s.h.atom_begin + i], make_float3(ci.x, ci.y, ci.z),
                                       v_inter, &d);
            } else {
                // with every atom of the wavefront's pass inside the box, the
                // clamping and out-of-box terms (zero there) are left out
                float sg[3];
                grid_scale(ka.grid, make_float3(ci.x, ci.y, ci.z), sg);
                float gf[8];
                if (!wave_any(!grid_inside(ka.grid, sg))) {
                    const GridPt gq = grid_locate_in(ka.grid, slot, sg);
                    grid_gather<GF>(ka.grid, gq.base, gf);
                    e += grid_finish<GF, true>(ka.grid, slot, gq, gf, v_inter, &d);
                } else {
                    const GridPt gq = grid_locate(ka.grid, slot, make_float3(ci.x, ci.y, ci.z));
                    grid_gather<GF>(ka.grid, gq.base, gf);
                    e += grid_finish<GF>(ka.grid, slot, gq, gf, v_inter, &d);
                }
            }
        }
        mf[i] = make_float4(d.x, d.y, d.z, 0.f);
    }
    if constexpr (UD_PAIR_PIPELINE && !Fmt<GF>::ad4) {
        // software-pipelined: each batch's table loads are issued before the
        // previous batch is finished, so they are in flight during its math
        // (not with AD4.2's larger batches, whose second copy would spill)
        constexpr int QB = kPairBatch * L;
        for (int q0 = s.gl; q0 < P; q0 += 2 * QB) {
            issue(b2, q0 + QB);
            e += finish(b, q0);
            if (q0 + QB >= P) break;
            issue(b, q0 + 2 * QB);
            e += finish(b2, q0 + QB);
        }
    } else {
        for (int q0 = s.gl; q0 < P; q0 += kPairBatch * L) {
            if (q0 != s.gl) issue(b, q0);
            e += finish(b, q0);
        }
    }
    return e;
}

// Pass B of the pair-parallel forces (after a barrier that completes pass A):
// each owned atom's grid gradient plus its pairs' forces.
template <int L, int NM, int GF>
__device__ __forceinline__ void atoms_forces_pp_b(const KArgs& ka, const Grp<L, NM>& s) {
    const float4* coords = s.coords();
    float4* mf = s.mf();
    const float* pf = s.pf();
    const int N = s.h.n_atoms;
    // pre-decoded partner records: byte offsets of the other atom's
    // coordinates and of the pair's force scale
    const uint32_t* part = ka.ls.partners + s.h.partner_ofs_begin;
    const char* cbase = reinterpret_cast<const char*>(coords);
    const char* pbase = reinterpret_cast<const char*>(pf);
    // partner records fetched together: each atom's list is padded to a
    // multiple of kPartnerPad (8) records and starts at one, so a batch is
    // one or two aligned 16-byte loads; the search's 32- and 64-lane groups,
    // with few registers to spare, take half batches
    constexpr int kRec = ((L == 32 || L == 64) && !Fmt<GF>::ad4) ? 4 : kPartnerPad;
    static_assert(kPartnerPad == 8 && kPartnerPad % kRec == 0, "record batches are one or two uint4");
    auto load_batch = [&](uint32_t first_rec, uint32_t (&r)[kRec]) {
        const uint4 lo = ld_at<uint4>(part, first_rec * 4u);
        r[0] = lo.x; r[1] = lo.y; r[2] = lo.z; r[3] = lo.w;
        if constexpr (kRec == 8) {
            const uint4 hi = ld_at<uint4>(part, first_rec * 4u + 16u);
            r[4] = hi.x; r[5] = hi.y; r[6] = hi.z; r[7] = hi.w;
        }
    };
    for (int i = s.gl; i < N; i += L) {
        const uint32_t ap = s.apl[i];  // staged in LDS with the ligand
        const int k0 = int(ap & 0xFFFFFu), kn = int(ap >> 20);
        // each next record batch is loaded (unconditionally) while the
        // current one is summed; the load after the last batch reads the
        // next atom's list or the array's zero padding and is not used.
        // AD4.2 kernels, with fewer registers to spare, load each batch when
        // it is needed (at 32 lanes a look-ahead of whole 8-record batches
        // cost 14 spilled VGPRs and 6 % of the launch's time, at 64 lanes 7
        // spilled VGPRs and 1 %, hence the half batches there)
        constexpr bool kAhead = !Fmt<GF>::ad4;
        uint32_t rec[kRec], nrec[kAhead ? kRec : 1];
        if constexpr (kAhead) load_batch(uint32_t(k0), rec);
        float4 g = mf[i];
        const float4 ci = coords[i];
        // sum d (c_i - c_o) = c_i sum d - sum d c_o over the atom's records,
        // in whole batches (padding records carry d = 0)
        float sd = 0.f, sx = 0.f, sy = 0.f, sz = 0.f;
        for (int k = 0; k < kn; k += kRec) {
            if constexpr (kAhead) load_batch(uint32_t(k0 + k + kRec), nrec);
            else load_batch(uint32_t(k0 + k), rec);
#pragma unroll
            for (int u = 0; u < kRec; ++u) {
                const float d = *reinterpret_cast<const float*>(pbase + (rec[u] >> 16));
                const float4 co = *reinterpret_cast<const float4*>(cbase + (rec[u] & 0xFFFFu));
                sd += d;
                sx += d * co.x;
                sy += d * co.y;
                sz += d * co.z;
            }
            if constexpr (kAhead) {
#pragma unroll
                for (int u = 0; u < kRec; ++u) rec[u] = nrec[u];
            }
        }
        g.x += ci.x * sd - sx;
        g.y += ci.y * sd - sy;
        g.z += ci.z * sd - sz;
        mf[i] = g;
    }
}

// Energy of the group's pose (model::eval_deriv, model.cpp:954-981, first
// half): forward kinematics, the grid term and the intramolecular pairs.
// Leaves in LDS what eval_grad needs: coordinates, each atom's grid gradient
// (or, without the pair buffer, its whole force) and the per-pair force
// scales.  rgyr (if not null) receives the gyration radius.
template <int L, int NM, int GF>
__device__ float eval_energy(const KArgs& ka, const Grp<L, NM>& s, const float* pose, float v_intra, float v_inter,
                             float* rgyr = nullptr, float slope = -1.f) {
    // the pair pass's first record batch does not depend on the pose: its
    // loads are issued here and complete during the forward kinematics (the
    // record array is padded, so lanes past the last pair read zeros)
    uint64_t rec0[kPairBatch];
    if (ka.pair_forces) {
#pragma unroll
        for (int u = 0; u < kPairBatch; ++u)
            rec0[u] = ld_at<uint64_t>(ka.ls.pair_ofs + s.h.pair_begin, uint32_t(s.gl + u * L) * 8u);
    }
    fk_frames(s, pose);
    float rg2 = 0.f;
    atoms_coords(s, rg2, rgyr != nullptr);
    __syncthreads();
    float e = ka.pair_forces ? atoms_forces_pp_a<L, NM, GF>(ka, s, v_intra, v_inter, rec0, slope)
                             : atoms_forces_2s<L, NM, GF>(ka, s, v_intra, v_inter, slope);
    __syncthreads();
    e = gsum<L>(e);
    if (rgyr) {
        // model::gyration_radius (model.cpp:871-883): heavy atoms about the root origin
        rg2 = gsum<L>(rg2);
        *rgyr = s.h.n_heavy > 0 ? __builtin_amdgcn_sqrtf(rg2 * __builtin_amdgcn_rcpf(float(s.h.n_heavy))) : 0.f;
    }
    return e;
}

// Gyration radius of the pose evaluated by the last eval_energy, from its
// coordinates and root frame still in LDS: the same sum, in the same order,
// as eval_energy's (model::gyration_radius, model.cpp:871-883).
template <int L, int NM>
__device__ __forceinline__ float gyration(const Grp<L, NM>& s) {
    const float4* coords = s.coords();
    const float* fr = s.frames();
    float rg2 = 0.f;
    for (int i = s.gl; i < s.h.n_atoms; i += L) {
        if ((s.atoms[i].meta >> 16) & 1) {
            const float4 p = coords[i];
            const float dx = p.x - fr[0], dy = p.y - fr[1], dz = p.z - fr[2];
            rg2 += dx * dx + dy * dy + dz * dz;
        }
    }
    rg2 = gsum<L>(rg2);
    return s.h.n_heavy > 0 ? __builtin_amdgcn_sqrtf(rg2 * __builtin_amdgcn_rcpf(float(s.h.n_heavy))) : 0.f;
}

// Gradient of the pose evaluated by the last eval_energy (model::eval_deriv,
// second half): the pair forces of pass B, then the torsion-tree derivative
// (tree.h:75-84, 152-193).  Each node's force and torque are sums over the
// atoms of its subtree (a contiguous range in pre-order), which is the
// reference's bottom-up accumulation written out:
//   F_k = sum f_i,  T_k = sum (c_i - o_k) x f_i  over atoms i below node k,
//   dE/dtorsion_k = T_k . axis_k,  root: (F_0, T_0).
// With at most one atom per lane they come from group prefix sums; larger
// ligands sum each subtree directly.
// gout[j] receives gradient element gl + j * L.
template <int L, int NM, int GF>
__device__ void eval_grad(const KArgs& ka, const Grp<L, NM>& s, float (&gout)[Grp<L, NM>::KV]) {
    if (ka.pair_forces) {
        atoms_forces_pp_b<L, NM, GF>(ka, s);
        __syncthreads();
    }
    float4* coords = s.coords();
    float4* mf = s.mf();
    const float* fr = s.frames();
    float* gbuf = s.bc(2);
    const int nn = s.h.n_nodes;
#if UD_TORQUE_SCAN
    if (s.h.n_atoms <= L) {
        // One atom per lane: prefix sums over the pre-order atom list of the
        // force and of its moment about the root origin o0, m_i = (c_i - o0) x f_i.
        // Node k's subtree is atoms [ab, se), so F_k = PF[se-1] - PF[ab-1] and
        // T_k = (PM[se-1] - PM[ab-1]) + (o0 - o_k) x F_k; each node's lane
        // reads the two prefix lanes it needs.
        float v[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        if (s.gl < s.h.n_atoms) {
            const float4 fi = mf[s.gl], ci = coords[s.gl];
            const float dx = ci.x - fr[0], dy = ci.y - fr[1], dz = ci.z - fr[2];
            v[0] = fi.x; v[1] = fi.y; v[2] = fi.z;
            v[3] = dy * fi.z - dz * fi.y;
            v[4] = dz * fi.x - dx * fi.z;
            v[5] = dx * fi.y - dy * fi.x;
        }
#pragma unroll
        for (int c = 0; c < 6; ++c) v[c] = gscan<L>(v[c], s.gl);
        for (int k0 = 0; k0 < nn; k0 += L) {
            const int k = k0 + s.gl;
            const NodeRec nd = s.nodes[min(k, nn - 1)];
            const int ab = node_abegin(nd), se = node_subtree_end(nd);
            float d[6];
#pragma unroll
            for (int c = 0; c < 6; ++c) {
                const float hi = __shfl(v[c], se - 1, L), lo = __shfl(v[c], max(ab - 1, 0), L);
                d[c] = hi - (ab > 0 ? lo : 0.f);
            }
            if (k < nn) {
                if (k == 0) {
#pragma unroll
                    for (int c = 0; c < 6; ++c) gbuf[c] = d[c];
                } else {
                    const float* f = fr + 20 * k;
                    const float ex = fr[0] - f[0], ey = fr[1] - f[1], ez = fr[2] - f[2];
                    const float T0 = d[3] + (ey * d[2] - ez * d[1]);
                    const float T1 = d[4] + (ez * d[0] - ex * d[2]);
                    const float T2 = d[5] + (ex * d[1] - ey * d[0]);
                    gbuf[5 + k] = T0 * f[3] + T1 * f[4] + T2 * f[5];
                }
            }
        }
    } else
#endif
    {
    // S lanes share one node's subtree sum (strided over its atoms, then a
    // shuffle reduction); L / S nodes are handled per pass
    int S = L;
    while (S > 1 && S * nn > L) S >>= 1;
    const int per = L / S;
    const int kk = s.gl / S, jj = s.gl % S;
    for (int k0 = 0; k0 < nn; k0 += per) {
        const int k = k0 + kk;
        const float* f = fr + 20 * min(k, nn - 1);
        float F0 = 0.f, F1 = 0.f, F2 = 0.f, T0 = 0.f, T1 = 0.f, T2 = 0.f;
        if (k < nn) {
            const NodeRec nd = s.nodes[k];
            const int ab = node_abegin(nd), se = node_subtree_end(nd);
            for (int i = ab + jj; i < se; i += S) {
                const float4 fi = mf[i], ci = coords[i];
                const float dx = ci.x - f[0], dy = ci.y - f[1], dz = ci.z - f[2];
                F0 += fi.x; F1 += fi.y; F2 += fi.z;
                T0 += dy * fi.z - dz * fi.y;
                T1 += dz * fi.x - dx * fi.z;
                T2 += dx * fi.y - dy * fi.x;
            }
        }
        F0 = sub_sum(F0, S); F1 = sub_sum(F1, S); F2 = sub_sum(F2, S);
        T0 = sub_sum(T0, S); T1 = sub_sum(T1, S); T2 = sub_sum(T2, S);
        if (k < nn && jj == 0) {
            if (k == 0) {
                gbuf[0] = F0; gbuf[1] = F1; gbuf[2] = F2;
                gbuf[3] = T0; gbuf[4] = T1; gbuf[5] = T2;
            } else {
                gbuf[5 + k] = T0 * f[3] + T1 * f[4] + T2 * f[5];
            }
        }
    }
    }
    __syncthreads();
    const int n = s.ndof();
#pragma unroll
    for (int j = 0; j < Grp<L, NM>::KV; ++j) {
        const int idx = s.gl + j * L;
        gout[j] = (idx < n) ? gbuf[idx] : 0.f;
    }
    __syncthreads();
}

// Energy and gradient of the group's pose (model::eval_deriv).  gout[j]
// receives gradient element gl + j * L.
template <int L, int NM, int GF>
__device__ float eval_group(const KArgs& ka, const Grp<L, NM>& s, const float* pose, float v_intra, float v_inter,
                            float (&gout)[Grp<L, NM>::KV], float* rgyr = nullptr) {
    const float e = eval_energy<L, NM, GF>(ka, s, pose, v_intra, v_inter, rgyr);
    eval_grad<L, NM, GF>(ka, s, gout);
    return e;
}

template <int L, int KV>
__device__ __forceinline__ float gdot(const float (&a)[KV], const float (&b)[KV]) {
    float s = 0.f;
#pragma unroll
    for (int j = 0; j < KV; ++j) s += a[j] * b[j];  // padded entries are zero
    return gsum<L>(s);
}
template <int L, int KV>
__device__ __forceinline__ double gdot_d(const float (&a)[KV], const float (&b)[KV]) {
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < KV; ++j) s += double(a[j]) * double(b[j]);
    return gsum_d<L>(s);
}

// out = -H * in over the packed upper triangle (minus_mat_vec_product, bfgs.h:30-38)
template <int L, int NM>
__device__ __forceinline__ void minus_hmatvec(const Grp<L, NM>& s, const float (&in)[Grp<L, NM>::KV],
                                              float (&out)[Grp<L, NM>::KV], int n) {
    float* bc = s.bc(0);
    const float* H = s.H();
#pragma unroll
    for (int j = 0; j < Grp<L, NM>::KV; ++j) {
        const int idx = s.gl + j * L;
        if (idx < n) bc[idx] = in[j];
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < Grp<L, NM>::KV; ++j) {
        const int i = s.gl + j * L;
        float sum = 0.f;
        if (i < n) {
            const float* row = H + i * (n | 1);
            for (int c2 = 0; c2 < n; ++c2) sum += row[c2] * bc[c2];
        }
        out[j] = (i < n) ? -sum : 0.f;
    }
    __syncthreads();
}

// Register form of the inverse Hessian (NM <= kNSmall): lane gl keeps rows
// gl + j L, all NM columns, so every index is a compile-time constant.
// Vectors are broadcast through the group's bc(0) with entries n..NM-1
// zero, which keeps the columns beyond n at zero.
// Register form of the inverse Hessian.  HS = 1: lane gl + j L holds row
// gl + j L.  HS = 2 (groups of 2 NM lanes): lanes gl and gl + NM share row gl,
// columns [0, NM/2) and [NM/2, NM), and their partial products are added
// across the pair; half the registers per lane.
template <int L, int NM, int HS>
__device__ __forceinline__ int hess_row(int gl, int j) { return HS > 1 ? gl % NM : gl + j * L; }
template <int L, int NM, int HS>
__device__ __forceinline__ int hess_col0(int gl) { return HS > 1 ? (gl / NM) * (NM / HS) : 0; }

template <int L, int NM, int KV, int HS>
__device__ __forceinline__ void minus_hmatvec_reg(const Grp<L, NM>& s, const float (&Hr)[KV][NM / HS],
                                                  const float (&in)[KV], float (&out)[KV], int n) {
    static_assert(HS == 1 || (HS == 2 && KV == 1 && L == 2 * NM), "row split of two lanes per row");
    constexpr int HC = NM / HS;
    float* bc = s.bc(0);
#pragma unroll
    for (int j = 0; j < KV; ++j) {
        const int idx = s.gl + j * L;
        if (idx < NM) bc[idx] = (idx < n) ? in[j] : 0.f;
    }
    __syncthreads();
    const int c0 = hess_col0<L, NM, HS>(s.gl);
    float v[HC];
#pragma unroll
    for (int c = 0; c < HC; ++c) v[c] = bc[c0 + c];
#pragma unroll
    for (int j = 0; j < KV; ++j) {
        const int i = s.gl + j * L;
        float sum = 0.f;
#pragma unroll
        for (int c = 0; c < HC; ++c) sum += Hr[j][c] * v[c];
        if constexpr (HS > 1) sum += __shfl_xor(sum, NM);  // the row's other half (lane gl ^ NM)
        out[j] = (i < n) ? -sum : 0.f;
    }
    __syncthreads();
}

// rank-2 update of the register rows (bfgs_update, bfgs.h:46-63).  The mixed
// term ma_i p_c + ma_c p_i is formed without contraction, so entries (i, c)
// and (c, i) get identical values and H stays exactly symmetric.
template <int L, int NM, int KV, int HS>
__device__ __forceinline__ void hess_update_reg(const Grp<L, NM>& s, float (&Hr)[KV][NM / HS], const float (&mhy)[KV],
                                                const float (&p)[KV], bool pred, float c1, float c2,
                                                const float* ma, const float* pb) {
    constexpr int HC = NM / HS;
    const int c0 = hess_col0<L, NM, HS>(s.gl);
    float mav[HC], pbv[HC];
#pragma unroll
    for (int c = 0; c < HC; ++c) { mav[c] = ma[c0 + c]; pbv[c] = pb[c0 + c]; }
#pragma unroll
    for (int j = 0; j < KV; ++j) {
        // the row's entries of mhy and p (the broadcast buffers hold them for lanes of the second half)
        const float mi = HS > 1 ? ma[s.gl % NM] : mhy[j], pi = HS > 1 ? pb[s.gl % NM] : p[j];
#pragma unroll
        for (int c = 0; c < HC; ++c) {
            const float t = __fadd_rn(__fmul_rn(mi, pbv[c]), __fmul_rn(mav[c], pi));
            const float nv = Hr[j][c] + (c1 * t + c2 * (pi * pbv[c]));
            Hr[j][c] = pred ? nv : Hr[j][c];
        }
    }
}

// Orientations of the line search's trials: entry k = the increment of
// src's orientation by 2^-k times the direction's rotation part
// (quaternion_increment, conf.h:50-56), the value trial k's pose_increment
// would form, computed once per line search with one trial per lane.
template <int L, int NM>
__device__ __forceinline__ void trial_orientations(const Grp<L, NM>& s, const float* src, const float* p, bool pred) {
    if (!pred) return;
    float* qt = s.qt();
    for (int k = s.gl; k < kTrials; k += L) {
        const float a = __builtin_ldexpf(1.f, -k);  // the trial's alpha, exactly
        const float4 q = q_increment(make_float4(src[3], src[4], src[5], src[6]), a * p[3], a * p[4], a * p[5]);
        qt[4 * k] = q.x; qt[4 * k + 1] = q.y; qt[4 * k + 2] = q.z; qt[4 * k + 3] = q.w;
    }
}

// conf::increment (conf.h:98-104, 50-56, 169-172) for line-search trial
// `trial` (alpha = 2^-trial): position += alpha p, orientation from the
// trial table, torsions wrapped.
template <int L, int NM>
__device__ __forceinline__ void pose_increment(const Grp<L, NM>& s, float* dst, const float* src, const float* p, float alpha,
                                               int trial, bool pred) {
    if (!pred) return;
    const int T = s.ntors();
    const float* qt = s.qt() + 4 * trial;
    for (int idx = s.gl; idx < 7 + T; idx += L) {
        float v;
        if (idx < 3) v = src[idx] + alpha * p[idx];
        else if (idx < 7) v = qt[idx - 3];
        else v = normalize_angle(src[idx] + normalize_angle(alpha * p[idx - 1]));
        dst[idx] = v;
    }
}

template <int L, int NM>
__device__ __forceinline__ void copy_pose(const Grp<L, NM>& s, float* dst, const float* src, bool pred) {
    if (!pred) return;
    for (int idx = s.gl; idx < 7 + s.ntors(); idx += L) dst[idx] = src[idx];
}

template <int L, int GF>
__global__ __launch_bounds__(64) void score_kernel(KArgs ka, const WorkItem* work, const int32_t* perm,
                                                   const float* poses, float v_intra, float v_inter, float* e_out,
                                                   float* g_out) {
    extern __shared__ float4 lds_raw[];
    float* lds = reinterpret_cast<float*>(lds_raw);
    const WorkItem w = work[blockIdx.x];
    Grp<L> s = setup_group<L, kNMax>(ka, w.lig, lds);
    const bool active = s.g < w.count;
    const int pi = active ? perm[w.first + s.g] : 0;
    float* x = s.pose(PX);
    const int np = 7 + s.ntors();
    for (int i = s.gl; i < np; i += L) x[i] = active ? poses[size_t(pi) * UD_POSE_FLOATS + i] : (i == 3 ? 1.f : 0.f);
    __syncthreads();
    float gout[Grp<L>::KV];
    const float e = eval_group<L, kNMax, GF>(ka, s, x, v_intra, v_inter, gout);
    if (active) {
        if (s.gl == 0) e_out[pi] = e;
#pragma unroll
        for (int j = 0; j < Grp<L>::KV; ++j) {
            const int idx = s.gl + j * L;
            if (idx < UD_GRAD_FLOATS) g_out[size_t(pi) * UD_GRAD_FLOATS + idx] = gout[j];
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
        // cos / sin of 2 pi u2 through h = pi (u2 - 1/2) in [-pi/2, pi/2):
        // cos(2h + pi) = 2 sin^2 h - 1, sin(2h + pi) = -2 sin h cos h
        float sh, ch;
        sincos_half(kPi * (u2 - 0.5f), &sh, &ch);
        nq[2 * j] = rad * (2.f * sh * sh - 1.f);
        nq[2 * j + 1] = rad * (-2.f * sh * ch);
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
    // phi = 2 pi u: cos phi = 2 sin^2 h - 1, sin phi = -2 sin h cos h, h = pi (u - 1/2)
    float sh, ch;
    sincos_half(kPi * (u01(o.z) - 0.5f), &sh, &ch);
    const float rr = cbrtf(u01(o.w));
    const float sxy = sqrtf(fmaxf(0.f, 1.f - z * z));
    const float ux = rr * sxy * (2.f * sh * sh - 1.f), uy = rr * sxy * (-2.f * sh * ch), uz = rr * z;
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

// States of a lane group's search.  Every loop iteration evaluates exactly
// one pose per group (X in ST_INIT / ST_RGYR, the line-search trial XN in
// ST_TRIAL), then advances each group's own state, so groups never wait for
// one another's line searches and the energy+gradient code exists once.
enum { ST_RGYR = 0, ST_INIT = 1, ST_TRIAL = 2, ST_IDLE = 3 };

// BFGS local search (bfgs() with line_search and bfgs_update, bfgs.h:30-147;
// the reference's GPU BFGS is bfgs_warp, warp_ops.cuh:1051-1119, whose first
// line search compares with f0 = 0 where bfgs.h:107 uses f(x): SearchDev's
// ls_f0_zero selects that rule, off by default) and, when MC is set, the
// Monte Carlo chain around it (kernel<32, Config>, monte_carlo.cu:236-356):
// mutate -> BFGS with hunt_cap -> Metropolis at T -> if the accepted pose
// beats the chain's best, BFGS with authentic_v and keep it.
// MC == false runs one BFGS per group on the given poses (quasi_newton).
// The curvature scalar y.p is accumulated in FP64.
template <int L, int GF, bool MC, int NM>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(L == 64 ? UD_WAVES_L64 : L == 32 ? UD_WAVES_L32 : L == 16 ? UD_WAVES_L16 : UD_SEARCH_WAVES_PER_EU)))
void search_kernel(KArgs ka, SearchDev sp, const WorkItem* work, const int32_t* perm,
                                                    float* poses_io, float v_intra_opt, float v_inter_opt,
                                                    int max_steps_opt, float* out_pose, float* out_e,
                                                    long long* out_evals, int32_t* out_evals32) {
    constexpr int KV = Grp<L, NM>::KV;
    extern __shared__ float4 lds_raw[];
    float* lds = reinterpret_cast<float*>(lds_raw);
    const WorkItem w = work[blockIdx.x];
    Grp<L, NM> s = setup_group<L, NM>(ka, w.lig, lds);
    const bool active = s.g < w.count;
    const int T = s.ntors();
    const int np = 7 + T;
    const int n = s.ndof();
    const int hs = n | 1;  // row stride of the full inverse Hessian
    const int nh = n * hs;
    float* X = s.pose(PX);
    float* XN = s.pose(PXN);
    float* XO = s.pose(PXO);
    float* TMP = s.pose(PTMP);
    float* BEST = s.pose(PBEST);
    float* H = s.H();
    float* pb = s.bc(1);  // search direction, read by the pose increment
    constexpr bool HREG = UD_HESS_IN_REGS && NM <= kNSmall && L <= 32;
    constexpr int HS = (HREG && L == 2 * NM) ? 2 : 1;  // lanes per Hessian row
    float Hr[HREG ? KV : 1][HREG ? NM / HS : 1];
    // MC: the work item's `count` chains are taken by the groups in turn; a
    // group whose chain ends writes its result and starts the next one
    uint32_t chain = uint32_t(w.first + (active ? s.g : 0));
    int next_chain = Grp<L, NM>::G;  // wave-uniform
    const int pi = (!MC && active) ? perm[w.first + s.g] : 0;
    int lsteps;
    if (MC) lsteps = sp.local_steps_override > 0 ? sp.local_steps_override : s.h.local_steps;
    else lsteps = max_steps_opt > 0 ? max_steps_opt : s.h.local_steps;
    const uint2 key = make_uint2(sp.seed_lo, sp.seed_hi);

    // ---- per-group state (group-uniform registers)
    int state = ST_INIT, phase = 0, mc_step = 0, step = 0, trial = 0, evals = 0;
    float vI = v_intra_opt, vE = v_inter_opt;
    float f0 = 0.f, f_orig = 0.f, alpha = 1.f, pg = 0.f;
    float tmp_e = 0.f, best_e = __builtin_huge_valf(), rg_tmp = 0.f, rg_x = 0.f, rg_o = 0.f;
    // refinement pass and its penalty slope (explicit receptor only)
    int pass = 0;
    float slope = (!MC && GF == kGridExplicit && sp.refine_passes > 1) ? sp.refine_slope[0] : -1.f;
    float g[KV], p[KV], t[KV], y[KV], mhy[KV];
#pragma unroll
    for (int j = 0; j < KV; ++j) { g[j] = 0.f; p[j] = 0.f; y[j] = 0.f; }

    if (MC && sp.global_steps == 0) {
        // no Monte Carlo steps: each chain reports its random start
        for (int c = s.g; c < w.count; c += Grp<L, NM>::G) {
            const uint32_t ch = uint32_t(w.first + c);
            if (s.gl == 0) random_start(TMP, T, sp, ch);
            __syncthreads();
            for (int i = s.gl; i < UD_POSE_FLOATS; i += L) out_pose[size_t(ch) * UD_POSE_FLOATS + i] = (i < np) ? TMP[i] : 0.f;
            if (s.gl == 0) {
                out_e[ch] = __builtin_huge_valf();
                out_evals[ch] = 0;
            }
            __syncthreads();
        }
        return;
    }
    if (MC) {
        if (s.gl == 0 && active) random_start(TMP, T, sp, chain);
        __syncthreads();
        copy_pose(s, X, TMP, true);
        copy_pose(s, BEST, TMP, true);
        state = ST_RGYR;  // the first evaluation only measures the start's gyration radius
    } else {
        for (int i = s.gl; i < np; i += L) X[i] = active ? poses_io[size_t(pi) * UD_POSE_FLOATS + i] : (i == 3 ? 1.f : 0.f);
    }
    if (!active) state = ST_IDLE;
    __syncthreads();

    for (;;) {
        if (!wave_any(state != ST_IDLE)) break;
        // ---------------------------------------------------------------- evaluate
        const bool trialing = state == ST_TRIAL;
        const float f = eval_energy<L, NM, GF>(ka, s, trialing ? XN : X, vI, vE, nullptr, slope);
        if (state != ST_IDLE) ++evals;
        const bool is_init = state == ST_INIT;
        const bool ok = trialing && (f - f0 < 0.0001f * alpha * pg);
        const bool ls_end = trialing && (ok || trial == kTrials - 1);  // 10 trials (bfgs.h:70-83)
        // Most evaluations are line-search trials that fail the Armijo test
        // before the last one: their gradient is never used and their group's
        // state only halves alpha.  When that holds for every group of the
        // wavefront, take the next trial at once.
        if (!wave_any(is_init || ls_end || state == ST_RGYR)) {
            if (trialing) {
                alpha *= 0.5f;
                ++trial;
            }
            pose_increment(s, XN, X, pb, alpha, trial, trialing);  // pb still holds each trialing group's direction
            __syncthreads();
            continue;
        }
        // the gyration radius is used only at the ends of line searches and
        // of the start's measurement, all of which take this path
        const float rg = gyration<L, NM>(s);
        if (wave_any(is_init || ls_end)) eval_grad<L, NM, GF>(ka, s, t);
        // ---------------------------------------------------------------- line search / BFGS bookkeeping
        // (each product below runs only when some group of the wavefront uses it)
        const float gg = wave_any(ls_end) ? gdot<L>(g, g) : 0.f;
        if (trialing && !ok) alpha *= 0.5f;
        if (trialing && !ls_end) ++trial;
        const int step_now = step;
        bool finished = false;
        if (ls_end) {
#pragma unroll
            for (int j = 0; j < KV; ++j) y[j] = t[j] - g[j];
            f0 = f;
            rg_x = rg;
        }
        const bool conv = ls_end && !(__builtin_amdgcn_sqrtf(gg) >= 1e-5f);  // uses g before g = g_new (bfgs.h:131-132)
        const bool cont = ls_end && !conv;
        if (cont) {
#pragma unroll
            for (int j = 0; j < KV; ++j) g[j] = t[j];
            ++step;
        }
        if (conv || (cont && step >= lsteps)) finished = true;
        if (is_init) {
            // the first line search compares with f(x) (bfgs.h:107) or, as
            // the reference's GPU BFGS, with 0 (warp_ops.cuh:1062-1076)
            f0 = sp.ls_f0_zero ? 0.f : f;
            f_orig = f;
            rg_x = rg;
            rg_o = rg;
            step = 0;
#pragma unroll
            for (int j = 0; j < KV; ++j) g[j] = t[j];
            if (lsteps <= 0) finished = true;
        }
        // X <- XN after a finished line search; XO <- X and H <- I at BFGS start
        copy_pose(s, X, XN, ls_end);
        copy_pose(s, XO, X, is_init);
        if constexpr (HREG) {
            if (wave_any(is_init)) {
#pragma unroll
                for (int j = 0; j < KV; ++j)
#pragma unroll
                    for (int c = 0; c < NM / HS; ++c)
                        Hr[j][c] = is_init ? ((hess_col0<L, NM, HS>(s.gl) + c == hess_row<L, NM, HS>(s.gl, j)) ? 1.f : 0.f)
                                           : Hr[j][c];
            }
        } else if (is_init) {
            for (int idx = s.gl; idx < nh; idx += L) H[idx] = 0.f;
            for (int i = s.gl; i < n; i += L) H[i * hs + i] = 1.f;
        }
        __syncthreads();
        // initial scaling H = (y.p / y.y) I after the first step (bfgs.h:135-138)
        // y.y and y.p only matter after a chain's first BFGS step
        float yy = 0.f, yp0 = 0.f;
        if (wave_any(cont && step_now == 0)) {
            yy = gdot<L>(y, y);
            yp0 = gdot<L>(y, p);
        }
        const bool scale0 = cont && step_now == 0 && fabsf(yy) > kEpsFl;
        if constexpr (HREG) {
            if (wave_any(scale0)) {
                const float dv = alpha * yp0 * __builtin_amdgcn_rcpf(yy);
#pragma unroll
                for (int j = 0; j < KV; ++j)
#pragma unroll
                    for (int c = 0; c < NM / HS; ++c)
                        Hr[j][c] = (scale0 && hess_col0<L, NM, HS>(s.gl) + c == hess_row<L, NM, HS>(s.gl, j)) ? dv : Hr[j][c];
            }
        } else if (scale0) {
            const float dv = alpha * yp0 * __builtin_amdgcn_rcpf(yy);
            for (int i = s.gl; i < n; i += L) H[i * hs + i] = dv;
        }
        __syncthreads();
        // rank-2 update (bfgs_update, bfgs.h:46-63); skipped on the last step,
        // where the reference's update is never used
        const bool upd_pre = cont && !finished;
        const float yp = wave_any(upd_pre) ? float(gdot_d<L>(y, p)) : 0.f;
        const bool do_upd = upd_pre && !(alpha * yp < kEpsFl);
        if (wave_any(do_upd)) {
            if constexpr (HREG) minus_hmatvec_reg<L, NM, KV, HS>(s, Hr, y, mhy, n);
            else minus_hmatvec<L, NM>(s, y, mhy, n);
            const float yhy = -gdot<L>(y, mhy);
            float* ma = s.bc(0);
#pragma unroll
            for (int j = 0; j < KV; ++j) {
                const int idx = s.gl + j * L;
                // entries n..NM-1 of mhy and p are zero
                if (idx < (HREG ? NM : n)) { ma[idx] = mhy[j]; pb[idx] = p[j]; }
            }
            __syncthreads();
            if constexpr (HREG) {
                const float r = __builtin_amdgcn_rcpf(alpha * yp);
                const float c1 = alpha * r, c2 = alpha * alpha * (r * r * yhy + r);
                hess_update_reg<L, NM, KV, HS>(s, Hr, mhy, p, do_upd, c1, c2, ma, pb);
            } else if (do_upd) {
                const float r = __builtin_amdgcn_rcpf(alpha * yp);
                const float c1 = alpha * r, c2 = alpha * alpha * (r * r * yhy + r);
                // one row per lane; each entry uses the (min, max) index order of
                // the reference's upper-triangle update, so H stays exactly symmetric
                for (int i = s.gl; i < n; i += L) {
                    float* row = H + i * hs;
                    for (int j = 0; j < n; ++j) {
                        const int a = min(i, j), b = max(i, j);
                        row[j] += c1 * (ma[a] * pb[b] + ma[b] * pb[a]) + c2 * pb[a] * pb[b];
                    }
                }
            }
            __syncthreads();
        }
        // new search direction p = -H g (minus_mat_vec_product, bfgs.h:30-38)
        const bool new_dir = !finished && (is_init || cont);
        if (wave_any(new_dir)) {
            if constexpr (HREG) minus_hmatvec_reg<L, NM, KV, HS>(s, Hr, g, t, n);
            else minus_hmatvec<L, NM>(s, g, t, n);
            if (new_dir) {
#pragma unroll
                for (int j = 0; j < KV; ++j) p[j] = t[j];
            }
            const float pg_new = gdot<L>(p, g);
            if (new_dir) {
                pg = pg_new;
                alpha = 1.f;
                trial = 0;
                state = ST_TRIAL;
            }
        }
        // ---------------------------------------------------------------- BFGS end
        bool revert = false;
        float e_res = 0.f;
        if (finished) {
            revert = !(f0 <= f_orig);  // bfgs.h:140-144
            if (revert) { f0 = f_orig; rg_x = rg_o; }
            e_res = f0;
        }
        copy_pose(s, X, XO, revert);
        __syncthreads();
        // ---------------------------------------------------------------- Monte Carlo transitions
        bool start_bfgs = false, next_step = false;
        if (!MC) {
            if (finished) state = ST_IDLE;
            if constexpr (GF == kGridExplicit) {
                // refinement passes (vina.h:249-272): a pose whose heavy
                // atoms are not all inside the box (non_cache::within,
                // non_cache.cpp:408-418, with its margin) after a pass runs
                // the next pass's BFGS from its refined pose with the next,
                // steeper penalty slope
                const bool check = finished && pass + 1 < sp.refine_passes;
                if (wave_any(check)) {
                    fk_frames(s, X);  // every group's (the others' frames are not used again)
                    const float* fr = s.frames();
                    float outside = 0.f;
                    for (int i = s.gl; i < s.h.n_atoms; i += L) {
                        const AtomRec a = s.atoms[i];
                        if (!((a.meta >> 16) & 1)) continue;
                        const float3 q = frame_apply(fr + 20 * (a.meta & 0xff), a);
                        const float c[3] = {q.x, q.y, q.z};
#pragma unroll
                        for (int d = 0; d < 3; ++d)
                            if (c[d] < sp.box_lo[d] || c[d] > sp.box_hi[d]) outside = 1.f;
                    }
                    outside = gsum<L>(outside);
                    if (check && outside != 0.f) {
                        ++pass;
                        slope = sp.refine_slope[pass];
                        state = ST_INIT;
                    }
                    __syncthreads();
                }
            }
        } else {
            if (state == ST_RGYR) {
                rg_tmp = rg;
                next_step = true;
                mc_step = -1;
            }
            if (finished && phase == 0) {
                const uint4 o = philox4x32_10(make_uint4(uint32_t(mc_step) + 1u, chain, 2u, 0u), key);
                bool acc = e_res < tmp_e;
                if (!acc) acc = u01(o.x) < expf((tmp_e - e_res) / sp.temperature);  // metropolis_accept (monte_carlo.cu:92-97)
                const bool take = mc_step == 0 || acc;
                if (take) { tmp_e = e_res; rg_tmp = rg_x; }
                if (take && tmp_e < best_e) {
                    phase = 1;
                    start_bfgs = true;
                } else {
                    next_step = true;
                }
                copy_pose(s, TMP, X, take);
            } else if (finished && phase == 1) {
                tmp_e = e_res;
                rg_tmp = rg_x;
                copy_pose(s, TMP, X, true);
                const bool better = tmp_e < best_e;
                if (better) best_e = tmp_e;
                copy_pose(s, BEST, X, better);
                phase = 0;
                next_step = true;
            }
        }
        __syncthreads();
        bool chain_done = false;
        if (MC && next_step) {
            ++mc_step;
            if (mc_step >= sp.global_steps) {
                state = ST_IDLE;
                next_step = false;
                chain_done = true;
            } else {
                copy_pose(s, X, TMP, true);
            }
        }
        if (MC && wave_any(chain_done)) {
            // report finished chains, then hand out the item's next chains in
            // group order (one claim per finished group)
            if (chain_done) {
                for (int i = s.gl; i < UD_POSE_FLOATS; i += L)
                    out_pose[size_t(chain) * UD_POSE_FLOATS + i] = (i < np) ? BEST[i] : 0.f;
                if (s.gl == 0) {
                    out_e[chain] = best_e;
                    out_evals[chain] = evals;
                }
            }
            const uint64_t fin = __ballot(chain_done && s.gl == 0);
            const int rank = __popcll(fin & ((uint64_t(1) << (s.g * L)) - 1));
            const int c = next_chain + rank;
            next_chain += __popcll(fin);
            const bool refill = chain_done && c < w.count;
            if (refill) {
                chain = uint32_t(w.first + c);
                if (s.gl == 0) random_start(TMP, T, sp, chain);
                state = ST_RGYR;
                phase = 0;
                evals = 0;
                best_e = __builtin_huge_valf();
            }
            __syncthreads();
            copy_pose(s, X, TMP, refill);
            copy_pose(s, BEST, TMP, refill);
        }
        __syncthreads();
        if (MC && next_step && s.gl == 0) mutate(X, T, rg_tmp, sp, chain, uint32_t(mc_step));
        if (MC && (next_step || start_bfgs)) {
            state = ST_INIT;
            vI = phase == 0 ? sp.hunt[0] : sp.authentic[0];
            vE = phase == 0 ? sp.hunt[1] : sp.authentic[1];
        }
        // ---------------------------------------------------------------- next trial pose
        if (state == ST_TRIAL) {
#pragma unroll
            for (int j = 0; j < KV; ++j) {
                const int idx = s.gl + j * L;
                if (idx < n) pb[idx] = p[j];
            }
        }
        __syncthreads();
        // a new line search: its trials' orientations
        const bool ls_start = state == ST_TRIAL && trial == 0;
        if (wave_any(ls_start)) {
            trial_orientations(s, X, pb, ls_start);
            __syncthreads();
        }
        pose_increment(s, XN, X, pb, alpha, trial, state == ST_TRIAL);
        __syncthreads();
    }

    if (!active) return;
    if (!MC) {
        for (int i = s.gl; i < np; i += L) poses_io[size_t(pi) * UD_POSE_FLOATS + i] = X[i];
        if (s.gl == 0) {
            if (out_e) out_e[pi] = f0;
            if (out_evals32) out_evals32[pi] = evals;
        }
    }
}

// curl(e, v) (curl.h:28-43) in FP64
__device__ __forceinline__ double curl_d(double e, double v) {
    if (e > 0.0 && v < 0.1 * double(kMaxFl)) e *= (v < double(kEpsFl)) ? 0.0 : v / (v + e);
    return e;
}

// Final-score components with authentic caps (no gradient), the rescoring of
// Vina::score_gpu (vina.cpp:1245-1327): inter = the receptor term of each
// atom (cache::eval through the maps, or non_cache::eval against the
// explicit atoms), intra = the ligand pairs through the piecewise-constant
// table (eval_interacting_pairs, model.cpp:885-905).  Coordinates, map and
// table values are the FP32 / FP16 inputs; interpolation, curl and every sum
// are FP64.
template <int L, int GF>
__global__ __launch_bounds__(64) void terms_kernel(KArgs ka, const WorkItem* work, const int32_t* perm,
                                                   const float* poses, float v, float* inter_out, float* intra_out) {
    extern __shared__ float4 lds_raw[];
    float* lds = reinterpret_cast<float*>(lds_raw);
    const WorkItem w = work[blockIdx.x];
    Grp<L> s = setup_group<L, kNMax>(ka, w.lig, lds);
    const bool active = s.g < w.count;
    const int pi = active ? perm[w.first + s.g] : 0;
    float* x = s.pose(PX);
    const int np = 7 + s.ntors();
    for (int i = s.gl; i < np; i += L) x[i] = active ? poses[size_t(pi) * UD_POSE_FLOATS + i] : (i == 3 ? 1.f : 0.f);
    __syncthreads();
    fk_frames(s, x);
    float4* coords = s.coords();
    float rg2 = 0.f;
    atoms_coords(s, rg2);
    __syncthreads();
    const double vd = double(v);
    double ei = 0.0, ea = 0.0;
    for (int i = s.gl; i < s.h.n_atoms; i += L) {
        const AtomRec a = s.atoms[i];
        const int slot = (a.meta >> 8) & 0xff;
        if (slot == 0xff) continue;
        const float4 c = coords[i];
        if constexpr (GF == kGridExplicit) {
            // non_cache::eval (non_cache.cpp:34-255): clamp into the box, the
            // fast table of every receptor atom within the cutoff, curl, penalty
            const RecDev& r = ka.rec;
            const int t = int(a.meta >> 24);
            float p[3] = {c.x, c.y, c.z};
            double pen = 0.0;
            int cc[3];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                if (p[d] < r.init[d]) { pen += double(r.init[d]) - double(p[d]); p[d] = r.init[d]; }
                else if (p[d] > r.end[d]) { pen += double(p[d]) - double(r.end[d]); p[d] = r.end[d]; }
                const float xx = (p[d] - r.init[d]) * float(r.dim[d]) / r.range[d];
                cc[d] = (xx <= 0.f) ? 0 : (xx >= float(r.dim[d] - 1)) ? r.dim[d] - 1 : max(0, min(int(xx), r.dim[d] - 1));
            }
            const int cell = cc[0] + r.dim[0] * (cc[1] + r.dim[1] * cc[2]);
            double e = 0.0;
            for (int k = r.cell_start[cell]; k < r.cell_start[cell + 1]; ++k) {
                const float4 b = r.atoms[k];
                const float rx = p[0] - b.x, ry = p[1] - b.y, rz = p[2] - b.z;
                const float r2 = rx * rx + ry * ry + rz * rz;
                if (r2 < 64.f) {
                    const int t2 = __float_as_int(b.w);
                    const int tp = t <= t2 ? t + t2 * (t2 + 1) / 2 : t2 + t * (t + 1) / 2;
                    e += double(ka.ls.table_fast[size_t(tp) * kTableStride + int(32.f * r2)]);
                }
            }
            ei += curl_d(e, vd) + double(ka.grid.slope) * pen;
        } else {
            // grid::evaluate_aux (grid.cpp:39-129) energy path; AD4.2 adds the
            // electrostatic map times q and the desolvation map times |q|
            // (ad4cache::eval, ad4cache.cpp:151-185)
            const GridPt q = grid_locate(ka.grid, slot, make_float3(c.x, c.y, c.z));
            const int nm = Fmt<GF>::ad4 ? 3 : 1;
            const float qa = Fmt<GF>::ad4 ? ka.ls.charge[s.h.atom_begin + i] : 0.f;
            const uint32_t cell = q.base - uint32_t(slot) * uint32_t(ka.grid.n_points);
            for (int m = 0; m < nm; ++m) {
                const int sl = (m == 0) ? slot : (m == 1) ? ka.grid.elec_slot : ka.grid.dsolv_slot;
                const double fac = (m == 0) ? 1.0 : (m == 1) ? double(qa) : fabs(double(qa));
                float f[8];
                grid_gather<GF>(ka.grid, cell + uint32_t(sl) * uint32_t(ka.grid.n_points), f);
                const double fx = q.x, fy = q.y, fz = q.z, mx = 1.0 - fx, my = 1.0 - fy, mz = 1.0 - fz;
                double e = f[0] * mx * my * mz + f[1] * fx * my * mz + f[2] * mx * fy * mz + f[3] * fx * fy * mz +
                           f[4] * mx * my * fz + f[5] * fx * my * fz + f[6] * mx * fy * fz + f[7] * fx * fy * fz;
                if (Fmt<GF>::store == UD_GRID_F16) e *= double(ka.grid.scale[sl]);
                ei += fac * (curl_d(e, vd) + double(q.penalty));
            }
        }
    }
    const PairRec* prs = s.pairs;
    for (int q = s.gl; q < s.h.n_pairs; q += L) {
        const PairRec rec = prs[q];
        const float4 ca = coords[pair_a(rec)], cb = coords[pair_b(rec)];
        const float rx = cb.x - ca.x, ry = cb.y - ca.y, rz = cb.z - ca.z;
        const float r2 = rx * rx + ry * ry + rz * rz;
        if (r2 < pair_cut2<GF>()) {
            // precalculate_element::eval_fast (precalculate.h:40-47); AD4.2:
            // type-pair part plus the charge terms of the pair
            const int k = int(32.f * r2);
            double t = (k < kTableStride) ? double(ka.ls.table_fast[size_t(pair_slot(rec)) * kTableStride + k]) : 0.0;
            if constexpr (Fmt<GF>::ad4) {
                const float2 cf = ka.ls.pair_coef[s.h.pair_begin + q];
                const float2 uf = ka.ls.ad4_fast[k];
                t += double(cf.x) * double(uf.x) + double(cf.y) * double(uf.y);
            }
            ea += curl_d(t, vd);
        }
    }
    ei = gsum_d<L>(ei);
    ea = gsum_d<L>(ea);
    if (active && s.gl == 0) {
        inter_out[pi] = float(ei);
        intra_out[pi] = float(ea);
    }
}

// d curl(e) / de (curl.h:28-36): (v / (v + e))^2 for e > 0, else 1
__device__ __forceinline__ float curl_slope(float e, float v) {
    if (!(e > 0.f) || !(v < 0.1f * kMaxFl)) return 1.f;
    if (v < kEpsFl) return 0.f;
    const float t = v / (v + e);
    return t * t;
}

__device__ __forceinline__ float trilinear(const float (&f)[8], const GridPt& q) {
    const float x = q.x, y = q.y, z = q.z;
    const float mx = 1.f - x, my = 1.f - y, mz = 1.f - z;
    return f[0] * mx * my * mz + f[1] * x * my * mz + f[2] * mx * y * mz + f[3] * x * y * mz + f[4] * mx * my * z +
           f[5] * x * my * z + f[6] * mx * y * z + f[7] * x * y * z;
}

// Energy of poses with its gradient in the pose's own coordinates and in the
// term weights (the differentiable-energy binding).
//   * The quaternion is normalised before the evaluation, so the energy is
//     E(q / |q|); its gradient is 2 (0, tau) q^ / |q| with tau the torque
//     about the root origin (orientation part of model::eval_deriv's change,
//     model.cpp:954-981), since a rotation w moves q^ to exp(w / 2) q^
//     (quaternion_increment, quaternion.cpp:77-84) and changes E by tau . w.
//   * Position and torsion entries are the change's force and torsion parts.
//   * Map and pair energies are sums of w_k term_k before curl
//     (precalculate.h:203-241, cache.cpp:518-548), so dE/dw_k is the sum over
//     atoms of curl'(e) m_k(x) plus the sum over pairs of curl'(e) s_k(r^2),
//     m_k / s_k being term k's own maps and tables interpolated exactly as the
//     energy's.  Out-of-box penalties do not depend on the weights.
// gw_out holds UD_WEIGHT_FLOATS entries per pose in the order of the weights
// array; the conf-independent (rot) entry is 0: the energy does not contain it.
// dE/dw of the AD4.2 terms (vdW, H-bond, electrostatic, desolvation) for the
// lane's atoms and pairs, at the coordinates in LDS.  Each atom's energy is
// curl(A) + q curl(El) + |q| curl(D) over its affinity map A and the
// electrostatic / desolvation maps El, D (ad4cache::eval_deriv,
// ad4cache.cpp:223-262); each pair's is curl(type-pair table + q_a q_b E(r)
// + c_ab G(r)) (precalculate_element::eval_deriv).  Maps and tables are
// linear in the weights, so dE/dw_k is the curl slope of each weighted
// value times its unit-weight counterpart.
template <int L, int NM, int GF>
__device__ void ad4_weight_grad(const KArgs& ka, const TermsDev& td, const Grp<L, NM>& s, float v_intra,
                                float v_inter, float (&dw)[UD_WEIGHT_FLOATS]) {
    const float4* coords = s.coords();
    const GridDev& g = ka.grid;
    const size_t map_stride = size_t(td.n_maps) * g.n_points;
    const size_t sy = size_t(g.n[0]), sz = sy * g.n[1];
    for (int i = s.gl; i < s.h.n_atoms; i += L) {
        const AtomRec a = s.atoms[i];
        const int slot = (a.meta >> 8) & 0xff;
        if (slot == 0xff) continue;
        const float4 ci = coords[i];
        const float q = ka.ls.charge[s.h.atom_begin + i];
        const GridPt gq = grid_locate(g, slot, make_float3(ci.x, ci.y, ci.z));
        const uint32_t cell = gq.base - uint32_t(slot) * uint32_t(g.n_points);
        const int sl[3] = {slot, g.elec_slot, g.dsolv_slot};
        const float fac[3] = {1.f, q, fabsf(q)};
#pragma unroll
        for (int m = 0; m < 3; ++m) {
            float f[8];
            grid_gather<GF>(g, cell + uint32_t(sl[m]) * uint32_t(g.n_points), f);
            float er = trilinear(f, gq);
            if (Fmt<GF>::store == UD_GRID_F16) er *= g.scale[sl[m]];
            const float c = fac[m] * curl_slope(er, v_inter);
            for (int k = 0; k < td.n_terms; ++k) {
                const float* M = td.maps + size_t(k) * map_stride + size_t(sl[m]) * g.n_points + cell;
                const float t[8] = {M[0], M[1], M[sy], M[sy + 1], M[sz], M[sz + 1], M[sy + sz], M[sy + sz + 1]};
                dw[k] += c * trilinear(t, gq);
            }
        }
    }
    const size_t tab_stride = size_t(kTypePairs) * kTableStride;
    for (int q = s.gl; q < s.h.n_pairs; q += L) {
        const PairRec rec = s.pairs[q];
        const float4 ca = coords[pair_a(rec)], cb = coords[pair_b(rec)];
        const float rx = cb.x - ca.x, ry = cb.y - ca.y, rz = cb.z - ca.z;
        const float r2 = rx * rx + ry * ry + rz * rz;
        if (!(r2 < pair_cut2<GF>())) continue;
        const float rf = 32.f * r2;
        const int it = int(rf);
        const float rem = rf - float(it);
        const size_t idx = size_t(pair_slot(rec)) * kTableStride + it;
        // the weighted pair energy, as the search evaluates it
        float pe = 0.f, dor = 0.f;
        float2 t0 = make_float2(0.f, 0.f), t1 = t0;
        if (it < kTableStride - 1) {
            t0 = ka.ls.table_deriv[idx];
            t1 = ka.ls.table_deriv[idx + 1];
        }
        pe = t0.x + rem * (t1.x - t0.x);
        const float4 u0 = ka.ls.ad4_deriv[it], u1 = ka.ls.ad4_deriv[it + 1];
        const float2 cf = ka.ls.pair_coef[s.h.pair_begin + q];
        ad4_charge_terms(cf, u0, u1, rem, pe, dor);
        const float c = curl_slope(pe, v_intra);
        // unit-weight parts: type-pair tables of vdW and H-bond, and the
        // electrostatic / desolvation radial samples times the pair's coefficients
        const float4 w0 = td.ad4_unit[it], w1 = td.ad4_unit[it + 1];
        float part[4];
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            float e0 = 0.f, e1 = 0.f;
            if (it < kTableStride - 1) {
                const float2* T = td.tables + size_t(k) * tab_stride + idx;
                e0 = T[0].x;
                e1 = T[1].x;
            }
            part[k] = e0 + rem * (e1 - e0);
        }
        part[2] = cf.x * (w0.x + rem * (w1.x - w0.x));
        part[3] = cf.y * (w0.z + rem * (w1.z - w0.z));
#pragma unroll
        for (int k = 0; k < 4; ++k) dw[k] += c * part[k];
    }
}

template <int L, int GF>
__global__ __launch_bounds__(64) void energy_grad_kernel(KArgs ka, TermsDev td, const WorkItem* work,
                                                         const int32_t* perm, const float* poses, float v_intra,
                                                         float v_inter, float* e_out, float* gpose_out, float* gw_out) {
    extern __shared__ float4 lds_raw[];
    float* lds = reinterpret_cast<float*>(lds_raw);
    const WorkItem w = work[blockIdx.x];
    Grp<L> s = setup_group<L, kNMax>(ka, w.lig, lds);
    const bool active = s.g < w.count;
    const int pi = active ? perm[w.first + s.g] : 0;
    float* x = s.pose(PX);
    const int np = 7 + s.ntors();
    for (int i = s.gl; i < np; i += L) x[i] = active ? poses[size_t(pi) * UD_POSE_FLOATS + i] : (i == 3 ? 1.f : 0.f);
    __syncthreads();
    const float qn = sqrtf(x[3] * x[3] + x[4] * x[4] + x[5] * x[5] + x[6] * x[6]);
    const bool qok = qn > kEpsFl;  // a zero quaternion is taken as the identity, with zero gradient
    float qh[4] = {1.f, 0.f, 0.f, 0.f};
    if (qok)
        for (int j = 0; j < 4; ++j) qh[j] = x[3 + j] / qn;
    __syncthreads();
    if (s.gl == 0)
        for (int j = 0; j < 4; ++j) x[3 + j] = qh[j];
    __syncthreads();
    float gout[Grp<L>::KV];
    const float e = eval_group<L, kNMax, GF>(ka, s, x, v_intra, v_inter, gout);
    const float* gb = s.bc(2);  // change-space gradient left by eval_group
    if (active) {
        const float tx = gb[3], ty = gb[4], tz = gb[5];
        const float sc = qok ? 2.f / qn : 0.f;
        for (int i = s.gl; i < UD_POSE_FLOATS; i += L) {
            float v = 0.f;
            if (i < 3) {
                v = gb[i];
            } else if (i == 3) {
                v = -sc * (tx * qh[1] + ty * qh[2] + tz * qh[3]);
            } else if (i < 7) {
                // vector part of (0, tau) q^ = w tau + tau x v
                const float vx = qh[1], vy = qh[2], vz = qh[3];
                v = (i == 4) ? qh[0] * tx + (ty * vz - tz * vy)
                  : (i == 5) ? qh[0] * ty + (tz * vx - tx * vz)
                             : qh[0] * tz + (tx * vy - ty * vx);
                v *= sc;
            } else if (i < np) {
                v = gb[6 + (i - 7)];
            }
            gpose_out[size_t(pi) * UD_POSE_FLOATS + i] = v;
        }
        if (s.gl == 0) e_out[pi] = e;
    }
    // term weights
    float dw[UD_WEIGHT_FLOATS];
#pragma unroll
    for (int k = 0; k < UD_WEIGHT_FLOATS; ++k) dw[k] = 0.f;
    const float4* coords = s.coords();
    const size_t map_stride = size_t(td.n_maps) * ka.grid.n_points;
    if constexpr (Fmt<GF>::ad4) {
        ad4_weight_grad<L, kNMax, GF>(ka, td, s, v_intra, v_inter, dw);
    } else {
    for (int i = s.gl; i < s.h.n_atoms; i += L) {
        const AtomRec a = s.atoms[i];
        const int slot = (a.meta >> 8) & 0xff;
        if (slot == 0xff) continue;
        const float4 ci = coords[i];
        const GridPt q = grid_locate(ka.grid, slot, make_float3(ci.x, ci.y, ci.z));
        float f[8];
        grid_gather<GF>(ka.grid, q.base, f);
        float er = trilinear(f, q);
        if (GF == UD_GRID_F16) er *= ka.grid.scale[slot];
        const float c = curl_slope(er, v_inter);
        const size_t sy = size_t(ka.grid.n[0]), sz = sy * ka.grid.n[1];
#pragma unroll
        for (int k = 0; k < UD_WEIGHT_FLOATS; ++k) {
            if (k >= td.n_terms) break;
            const float* M = td.maps + size_t(k) * map_stride + q.base;
            const float t[8] = {M[0], M[1], M[sy], M[sy + 1], M[sz], M[sz + 1], M[sy + sz], M[sy + sz + 1]};
            dw[k] += c * trilinear(t, q);
        }
    }
    const size_t tab_stride = size_t(kTypePairs) * kTableStride;
    for (int q = s.gl; q < s.h.n_pairs; q += L) {
        const PairRec rec = s.pairs[q];
        const float4 ca = coords[pair_a(rec)], cb = coords[pair_b(rec)];
        const float rx = cb.x - ca.x, ry = cb.y - ca.y, rz = cb.z - ca.z;
        const float r2 = rx * rx + ry * ry + rz * rz;
        if (!(r2 < 64.f)) continue;
        const float rf = 32.f * r2;
        const int it = int(rf);
        const float rem = rf - float(it);
        const size_t idx = size_t(pair_slot(rec)) * kTableStride + it;
        const float2 t0 = ka.ls.table_deriv[idx], t1 = ka.ls.table_deriv[idx + 1];
        const float c = curl_slope(t0.x + rem * (t1.x - t0.x), v_intra);
#pragma unroll
        for (int k = 0; k < UD_WEIGHT_FLOATS; ++k) {
            if (k >= td.n_terms) break;
            const float2* T = td.tables + size_t(k) * tab_stride + idx;
            const float e0 = T[0].x, e1 = T[1].x;
            dw[k] += c * (e0 + rem * (e1 - e0));
        }
    }
    }
#pragma unroll
    for (int k = 0; k < UD_WEIGHT_FLOATS; ++k) dw[k] = gsum<L>(dw[k]);
    if (active && s.gl == 0)
        for (int k = 0; k < UD_WEIGHT_FLOATS; ++k) gw_out[size_t(pi) * UD_WEIGHT_FLOATS + k] = dw[k];
}

// Atom coordinates of poses.  With compact_exh == 0: pose perm[w.first + g]
// of the work item, UD_MAX_ATOMS atoms per pose.  With compact_exh > 0 the
// poses are docking chains (chain w.first + g of ligand w.lig, perm unused)
// and chain k of ligand l writes its n_atoms atoms at atom
// atom_begin(l) * compact_exh + k * n_atoms, so that all chains of all
// ligands take sum(n_atoms) * exh entries.
template <int L>
__global__ __launch_bounds__(64) void coords_kernel(KArgs ka, const WorkItem* work, const int32_t* perm,
                                                    const float* poses, float* coords_out, int compact_exh,
                                                    BoxLimits lim, int32_t* within_out) {
    extern __shared__ float4 lds_raw[];
    float* lds = reinterpret_cast<float*>(lds_raw);
    const WorkItem w = work[blockIdx.x];
    Grp<L> s = setup_group<L, kNMax>(ka, w.lig, lds);
    const bool active = s.g < w.count;
    const int pi = active ? (compact_exh > 0 ? w.first + s.g : perm[w.first + s.g]) : 0;
    float* x = s.pose(PX);
    const int np = 7 + s.ntors();
    for (int i = s.gl; i < np; i += L) x[i] = active ? poses[size_t(pi) * UD_POSE_FLOATS + i] : (i == 3 ? 1.f : 0.f);
    __syncthreads();
    fk_frames(s, x);
    const float* fr = s.frames();
    if (within_out) {
        // non_cache::within (non_cache.cpp:408-418): heavy atoms in the box
        // with its margin (a NaN coordinate does not count as outside)
        float outside = 0.f;
        for (int i = s.gl; i < s.h.n_atoms; i += L) {
            const AtomRec a = s.atoms[i];
            if (!((a.meta >> 16) & 1)) continue;
            const float3 p = frame_apply(fr + 20 * (a.meta & 0xff), a);
            const float c[3] = {p.x, p.y, p.z};
#pragma unroll
            for (int d = 0; d < 3; ++d)
                if (c[d] < lim.lo[d] || c[d] > lim.hi[d]) outside = 1.f;
        }
        outside = gsum<L>(outside);
        if (active && s.gl == 0) within_out[pi] = outside == 0.f ? 1 : 0;
        return;
    }
    if (active && compact_exh > 0) {
        const size_t first = size_t(s.h.atom_begin) * compact_exh + size_t(pi - w.lig * compact_exh) * s.h.n_atoms;
        for (int i = s.gl; i < s.h.n_atoms; i += L) {
            const AtomRec a = s.atoms[i];
            const float3 p = frame_apply(fr + 20 * (a.meta & 0xff), a);
            float* o = coords_out + (first + i) * 3;
            o[0] = p.x; o[1] = p.y; o[2] = p.z;
        }
    } else if (active) {
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

// ====================================================================== clustering and RMSD bounds
static constexpr int kMaxModes = UD_MAX_MODES;  // output modes per ligand of ud_postprocess
// rmsd_upper_bound (coords.cpp:21-30) over the heavy atoms `heavy[0..nh)` of
// two coordinate sets, in atom order with every product and sum rounded to
// FP32 (no contraction), as the reference's fl accumulation.
__device__ float rmsd_ub_ordered(const float* a, const float* b, const int* heavy, int nh) {
    float acc = 0.f;
    for (int k = 0; k < nh; ++k) {
        const int i = 3 * heavy[k];
        const float dx = __fsub_rn(a[i], b[i]), dy = __fsub_rn(a[i + 1], b[i + 1]), dz = __fsub_rn(a[i + 2], b[i + 2]);
        acc = __fadd_rn(acc, __fadd_rn(__fadd_rn(__fmul_rn(dx, dx), __fmul_rn(dy, dy)), __fmul_rn(dz, dz)));
    }
    return nh > 0 ? __fsqrt_rn(__fdiv_rn(acc, float(nh))) : 0.f;
}

// One add_to_output_container step (coords.cpp:42-68) with candidate t of
// energy e on the container kept[0..nk) (sorted by energy, stably): the
// closest kept pose (the first of equal RMSDs) absorbs the candidate if it
// is within min_rmsd, else a full container replaces its worst entry if the
// candidate is better.  Lanes compute the RMSDs; lane 0 updates.
__device__ void container_add(int t, float e, const float* base, int na, const int* heavy, int nh, const float* en,
                              float min_rmsd, int max_size, int* kept, int& nk, float* rr) {
    if (e < kEpsFl && e > -kEpsFl) return;  // an energy of exactly 0 marks a failed chain
    for (int k = threadIdx.x; k < nk; k += kWG)
        rr[k] = rmsd_ub_ordered(base + size_t(t) * na * 3, base + size_t(kept[k]) * na * 3, heavy, nh);
    __syncthreads();
    if (threadIdx.x == 0) {
        int closest = -1;
        float cr = 0.f;
        for (int k = 0; k < nk; ++k)
            if (k == 0 || rr[k] < cr) { closest = k; cr = rr[k]; }
        int pos = -1;
        if (closest >= 0 && cr < min_rmsd) {
            if (e < en[kept[closest]]) { kept[closest] = t; pos = closest; }
        } else if (nk < max_size) {
            kept[nk] = t;
            pos = nk++;
        } else if (nk > 0 && e < en[kept[nk - 1]]) {
            kept[nk - 1] = t;
            pos = nk - 1;
        }
        // stable sort by energy: only the changed entry is out of place
        if (pos >= 0) {
            while (pos > 0 && en[kept[pos]] < en[kept[pos - 1]]) { const int x = kept[pos]; kept[pos] = kept[pos - 1]; kept[--pos] = x; }
            while (pos + 1 < nk && en[kept[pos + 1]] < en[kept[pos]]) { const int x = kept[pos]; kept[pos] = kept[pos + 1]; kept[++pos] = x; }
        }
    }
    __syncthreads();
}

// Per ligand (one workgroup): the chains in chain order through
// add_to_output_container (as monte_carlo.cu:2940-2943 merges the chains'
// results), then remove_redundant (vina.cpp:1465-1470): the kept poses, in
// their order, through the same step into an empty container.  coords: the
// compact chain coordinates of coords_kernel; energy: the chains' best
// energies.  kept_out[lig * max_size + k] = chain index within the ligand,
// sorted by energy; n_out[lig] their number.
__global__ __launch_bounds__(64) void cluster_kernel(const LigHdr* hdr, const AtomRec* atoms, const float* coords,
                                                     const float* energy, int exh, int max_size, float min_rmsd,
                                                     int32_t* kept_out, int32_t* n_out) {
    __shared__ int heavy[UD_MAX_ATOMS];
    __shared__ int kept[kMaxModes], kept2[kMaxModes];
    __shared__ float rr[kMaxModes];
    __shared__ int nh_s;
    const int lig = blockIdx.x;
    const LigHdr h = hdr[lig];
    if (threadIdx.x == 0) {
        int nh = 0;
        for (int i = 0; i < h.n_atoms; ++i)
            if ((atoms[h.atom_begin + i].meta >> 16) & 1) heavy[nh++] = i;
        nh_s = nh;
    }
    __syncthreads();
    const int nh = nh_s, na = h.n_atoms;
    const float* base = coords + size_t(h.atom_begin) * exh * 3;
    const float* en = energy + size_t(lig) * exh;
    int nk = 0;
    for (int t = 0; t < exh; ++t) {
        container_add(t, en[t], base, na, heavy, nh, en, min_rmsd, max_size, kept, nk, rr);
        nk = __shfl(nk, 0);
    }
    int nk2 = 0;
    for (int j = 0; j < nk; ++j) {
        container_add(kept[j], en[kept[j]], base, na, heavy, nh, en, min_rmsd, nk, kept2, nk2, rr);
        nk2 = __shfl(nk2, 0);
    }
    if (threadIdx.x == 0) n_out[lig] = nk2;
    for (int k = threadIdx.x; k < nk2; k += kWG) kept_out[size_t(lig) * max_size + k] = kept2[k];
}

// RMSD bounds of mode m against mode ref[m] (model::rmsd_upper_bound /
// rmsd_lower_bound, model.cpp:1021-1064) over the heavy atoms of ligand
// lig[m]: the upper bound pairs atoms by index; the lower bound takes for
// every atom the nearest atom of the same element in the other pose, both
// ways, the larger of the two.  coords: UD_MAX_ATOMS atoms per mode.
__global__ __launch_bounds__(64) void bounds_kernel(const LigHdr* hdr, const AtomRec* atoms, const int32_t* elem,
                                                    const float* coords, const int32_t* mode_lig,
                                                    const int32_t* mode_ref, float* ub_out, float* lb_out) {
    __shared__ int heavy[UD_MAX_ATOMS];
    __shared__ int nh_s;
    const int m = blockIdx.x;
    const LigHdr h = hdr[mode_lig[m]];
    if (threadIdx.x == 0) {
        int nh = 0;
        for (int i = 0; i < h.n_atoms; ++i)
            if ((atoms[h.atom_begin + i].meta >> 16) & 1) heavy[nh++] = i;
        nh_s = nh;
    }
    __syncthreads();
    const int nh = nh_s;
    const float* a = coords + size_t(m) * UD_MAX_ATOMS * 3;
    const float* b = coords + size_t(mode_ref[m]) * UD_MAX_ATOMS * 3;
    const int32_t* el = elem + h.atom_begin;
    auto one_way = [&](const float* x, const float* y) {
        float acc = 0.f;
        for (int k = threadIdx.x; k < nh; k += kWG) {
            const int i = heavy[k];
            float best = kMaxFl;
            for (int q = 0; q < nh; ++q) {
                const int j = heavy[q];
                if (el[j] != el[i]) continue;
                const float dx = x[3 * i] - y[3 * j], dy = x[3 * i + 1] - y[3 * j + 1], dz = x[3 * i + 2] - y[3 * j + 2];
                best = fminf(best, dx * dx + dy * dy + dz * dz);
            }
            acc += best;
        }
        acc = gsum<kWG>(acc);
        return nh > 0 ? sqrtf(acc / float(nh)) : 0.f;
    };
    const float lb = fmaxf(one_way(a, b), one_way(b, a));
    if (threadIdx.x == 0) {
        ub_out[m] = rmsd_ub_ordered(a, b, heavy, nh);
        lb_out[m] = lb;
    }
}

// The kept chains' best poses and energies, one slot per (ligand, mode):
// slot l * max_size + k holds chain kept[slot] of ligand l for k < n_kept[l].
__global__ __launch_bounds__(256) void gather_kept_kernel(const float* poses, const float* energy,
                                                          const int32_t* kept, const int32_t* n_kept, int n_lig,
                                                          int exh, int max_size, float* poses_out, float* e_out) {
    const size_t t = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const size_t slot = t / UD_POSE_FLOATS;
    const int f = int(t % UD_POSE_FLOATS);
    if (slot >= size_t(n_lig) * max_size) return;
    const int l = int(slot / max_size), k = int(slot % max_size);
    if (k >= n_kept[l]) return;
    const int ch = kept[slot];
    if (ch < 0 || ch >= exh) return;
    const size_t src = size_t(l) * exh + ch;
    poses_out[slot * UD_POSE_FLOATS + f] = poses[src * UD_POSE_FLOATS + f];
    if (f == 0) e_out[slot] = energy[src];
}

// ====================================================================== launchers
size_t lds_bytes(int lanes, const LaunchCap& cap) {
    Layout l;
    l.init(cap.n_atoms, cap.n_nodes, cap.n_pairs, cap.pair_forces != 0, lanes);
    return l.bytes(kWG / lanes);
}

static KArgs make_kargs(const LigandSetDev& ls, const GridDev& g, const LaunchCap& cap, const RecDev* rec = nullptr) {
    KArgs k;
    std::memset(&k.rec, 0, sizeof(k.rec));
    if (rec) k.rec = *rec;
    k.ls = ls;
    k.grid = g;
    k.na_cap = cap.n_atoms;
    k.nn_cap = cap.n_nodes;
    k.np_cap = cap.n_pairs;
    k.pair_forces = cap.pair_forces;
    return k;
}

#define UD_DISPATCH_L(lanes, ...)                          \
    switch (lanes) {                                       \
        case 8: { constexpr int L = 8; __VA_ARGS__; } break;   \
        case 16: { constexpr int L = 16; __VA_ARGS__; } break; \
        case 32: { constexpr int L = 32; __VA_ARGS__; } break; \
        default: { constexpr int L = 64; __VA_ARGS__; } break; \
    }

// grid-format template values dispatched at run time: FP32 / FP16 maps and
// AD4.2 over FP32 / FP16 maps
#define UD_DISPATCH_GF(gf, ...)                                                      \
    switch (gf) {                                                                    \
        case UD_GRID_F16: { constexpr int GFV = UD_GRID_F16; __VA_ARGS__; } break;   \
        case kGridAd4F32: { constexpr int GFV = kGridAd4F32; __VA_ARGS__; } break;   \
        case kGridAd4F16: { constexpr int GFV = kGridAd4F16; __VA_ARGS__; } break;   \
        default: { constexpr int GFV = UD_GRID_F32; __VA_ARGS__; } break;            \
    }

hipError_t launch_tables(int sf, const float* w, float2* deriv, float* fast, hipStream_t st) {
    SfWeights W;
    for (int i = 0; i < 7; ++i) W.w[i] = w[i];
    hipLaunchKernelGGL(table_kernel, dim3(kTypePairs), dim3(256), 0, st, sf, W, ad4_params(), deriv, fast);
    return hipGetLastError();
}

hipError_t launch_grid_build_ad4(const float4* rec, const float* rec_q, int n_rec, const GridBuildArgs& a,
                                 const float* w, float* maps_out, hipStream_t st) {
    SfWeights W;
    for (int i = 0; i < 7; ++i) W.w[i] = (i < 4) ? w[i] : 0.f;
    dim3 grid((a.n_points + 255) / 256, std::max(1, (a.n_maps + 7) / 8));
    hipLaunchKernelGGL(grid_build_ad4_kernel, grid, dim3(256), 0, st, rec, rec_q, n_rec, a, ad4_params(), W, maps_out);
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

hipError_t launch_score(int lanes, int gf, const LigandSetDev& ls, const GridDev& g, const LaunchCap& cap,
                        const WorkItem* work, int n_work, const int32_t* perm, const float* poses, float v_intra,
                        float v_inter, float* e_out, float* g_out, hipStream_t st) {
    const KArgs ka = make_kargs(ls, g, cap);
    const size_t shm = lds_bytes(lanes, cap);
    UD_DISPATCH_L(lanes, {
        UD_DISPATCH_GF(gf, hipLaunchKernelGGL((score_kernel<L, GFV>), dim3(n_work), dim3(kWG), shm, st, ka, work, perm,
                                              poses, v_intra, v_inter, e_out, g_out));
    });
    return hipGetLastError();
}


// NM must be kNSmall exactly when the launch's ligands have at most
// kNSmall - 6 torsions (5 + nodes <= kNSmall): Layout::init allocates no LDS
// inverse Hessian for such launches at <= 32 lanes, which keep it in registers.
template <int L, int GF, bool MC, int NM>
static void launch_search(const KArgs& ka, const SearchDev& sp, const WorkItem* work, int n_work, size_t shm,
                          hipStream_t st, const int32_t* perm, float* poses, float v_intra, float v_inter,
                          int max_steps, float* out_pose, float* out_e, long long* out_evals, int32_t* evals32) {
    hipLaunchKernelGGL((search_kernel<L, GF, MC, NM>), dim3(n_work), dim3(kWG), shm, st, ka, sp, work, perm, poses,
                       v_intra, v_inter, max_steps, out_pose, out_e, out_evals, evals32);
}

template <int L, bool MC>
static void launch_search_nm(int gf, const LaunchCap& cap, const KArgs& ka, const SearchDev& sp,
                             const WorkItem* work, int n_work, size_t shm, hipStream_t st, const int32_t* perm,
                             float* poses, float v_intra, float v_inter, int max_steps, float* out_pose, float* out_e,
                             long long* out_evals, int32_t* evals32) {
    const bool small = 6 + cap.n_nodes - 1 <= kNSmall;
    UD_DISPATCH_GF(gf, {
        if (small)
            launch_search<L, GFV, MC, kNSmall>(ka, sp, work, n_work, shm, st, perm, poses, v_intra, v_inter, max_steps,
                                               out_pose, out_e, out_evals, evals32);
        else
            launch_search<L, GFV, MC, kNMax>(ka, sp, work, n_work, shm, st, perm, poses, v_intra, v_inter, max_steps,
                                             out_pose, out_e, out_evals, evals32);
    });
}

hipError_t launch_optimize(int lanes, int gf, const LigandSetDev& ls, const GridDev& g, const LaunchCap& cap,
                           const WorkItem* work, int n_work, const int32_t* perm, float* poses, float v_intra,
                           float v_inter, int max_steps, float* e_out, int32_t* evals_out, hipStream_t st) {
    const KArgs ka = make_kargs(ls, g, cap);
    const size_t shm = lds_bytes(lanes, cap);
    SearchDev sp = {};
    UD_DISPATCH_L(lanes, {
        launch_search_nm<L, false>(gf, cap, ka, sp, work, n_work, shm, st, perm, poses, v_intra, v_inter, max_steps,
                                   nullptr, e_out, nullptr, evals_out);
    });
    return hipGetLastError();
}

// BFGS of given poses against the explicit receptor (the refinement of
// Vina::global_search_gpu, vina.h:249-272), all passes in one launch; with
// one pass ka.grid.slope is its slope
hipError_t launch_refine(int lanes, const LigandSetDev& ls, const GridDev& g, const RecDev& rec, const LaunchCap& cap,
                         const WorkItem* work, int n_work, const int32_t* perm, float* poses, float v_intra,
                         float v_inter, int max_steps, float* e_out, int32_t* evals_out, hipStream_t st,
                         int passes, const float* slopes, const BoxLimits& lim) {
    if (passes < 1 || passes > kMaxRefinePasses) return hipErrorInvalidValue;
    const KArgs ka = make_kargs(ls, g, cap, &rec);
    const size_t shm = lds_bytes(lanes, cap);
    SearchDev sp = {};
    sp.refine_passes = passes;
    for (int k = 0; k < passes; ++k) sp.refine_slope[k] = slopes[k];
    for (int d = 0; d < 3; ++d) {
        sp.box_lo[d] = lim.lo[d];
        sp.box_hi[d] = lim.hi[d];
    }
    // the vector capacity must follow the LDS layout's choice (Layout::init):
    // small launches keep the inverse Hessian in registers at <= 32 lanes
    const bool small = 6 + cap.n_nodes - 1 <= kNSmall;
    UD_DISPATCH_L(lanes, {
        if (small)
            launch_search<L, kGridExplicit, false, kNSmall>(ka, sp, work, n_work, shm, st, perm, poses, v_intra,
                                                            v_inter, max_steps, nullptr, e_out, nullptr, evals_out);
        else
            launch_search<L, kGridExplicit, false, kNMax>(ka, sp, work, n_work, shm, st, perm, poses, v_intra, v_inter,
                                                          max_steps, nullptr, e_out, nullptr, evals_out);
    });
    return hipGetLastError();
}

hipError_t launch_score_explicit(int lanes, const LigandSetDev& ls, const GridDev& g, const RecDev& rec,
                                 const LaunchCap& cap, const WorkItem* work, int n_work, const int32_t* perm,
                                 const float* poses, float v_intra, float v_inter, float* e_out, float* g_out,
                                 hipStream_t st) {
    const KArgs ka = make_kargs(ls, g, cap, &rec);
    const size_t shm = lds_bytes(lanes, cap);
    UD_DISPATCH_L(lanes, {
        hipLaunchKernelGGL((score_kernel<L, kGridExplicit>), dim3(n_work), dim3(kWG), shm, st, ka, work, perm, poses,
                           v_intra, v_inter, e_out, g_out);
    });
    return hipGetLastError();
}

hipError_t launch_terms_explicit(int lanes, const LigandSetDev& ls, const GridDev& g, const RecDev& rec,
                                 const LaunchCap& cap, const WorkItem* work, int n_work, const int32_t* perm,
                                 const float* poses, float v, float* inter_out, float* intra_out, hipStream_t st) {
    const KArgs ka = make_kargs(ls, g, cap, &rec);
    const size_t shm = lds_bytes(lanes, cap);
    UD_DISPATCH_L(lanes, {
        hipLaunchKernelGGL((terms_kernel<L, kGridExplicit>), dim3(n_work), dim3(kWG), shm, st, ka, work, perm, poses,
                           v, inter_out, intra_out);
    });
    return hipGetLastError();
}

hipError_t launch_dock(int lanes, int gf, const LigandSetDev& ls, const GridDev& g, const LaunchCap& cap,
                       const SearchDev& sp, const WorkItem* work, int n_work, float* out_pose, float* out_e,
                       long long* out_evals, hipStream_t st) {
    const KArgs ka = make_kargs(ls, g, cap);
    const size_t shm = lds_bytes(lanes, cap);
    UD_DISPATCH_L(lanes, {
        launch_search_nm<L, true>(gf, cap, ka, sp, work, n_work, shm, st, nullptr, nullptr, sp.hunt[0], sp.hunt[1], 0,
                                  out_pose, out_e, out_evals, nullptr);
    });
    return hipGetLastError();
}

hipError_t launch_terms(int lanes, int gf, const LigandSetDev& ls, const GridDev& g, const LaunchCap& cap,
                        const WorkItem* work, int n_work, const int32_t* perm, const float* poses, float v,
                        float* inter_out, float* intra_out, hipStream_t st) {
    const KArgs ka = make_kargs(ls, g, cap);
    const size_t shm = lds_bytes(lanes, cap);
    UD_DISPATCH_L(lanes, {
        UD_DISPATCH_GF(gf, hipLaunchKernelGGL((terms_kernel<L, GFV>), dim3(n_work), dim3(kWG), shm, st, ka, work, perm,
                                              poses, v, inter_out, intra_out));
    });
    return hipGetLastError();
}

hipError_t launch_energy_grad(int lanes, int gf, const LigandSetDev& ls, const GridDev& g, const TermsDev& td,
                              const LaunchCap& cap, const WorkItem* work, int n_work, const int32_t* perm,
                              const float* poses, float v_intra, float v_inter, float* e_out, float* gpose_out,
                              float* gw_out, hipStream_t st) {
    const KArgs ka = make_kargs(ls, g, cap);
    const size_t shm = lds_bytes(lanes, cap);
    UD_DISPATCH_L(lanes, {
        if (gf == UD_GRID_F16)
            hipLaunchKernelGGL((energy_grad_kernel<L, UD_GRID_F16>), dim3(n_work), dim3(kWG), shm, st, ka, td, work,
                               perm, poses, v_intra, v_inter, e_out, gpose_out, gw_out);
        else if (gf == kGridAd4F32)
            hipLaunchKernelGGL((energy_grad_kernel<L, kGridAd4F32>), dim3(n_work), dim3(kWG), shm, st, ka, td, work,
                               perm, poses, v_intra, v_inter, e_out, gpose_out, gw_out);
        else if (gf == kGridAd4F16)
            hipLaunchKernelGGL((energy_grad_kernel<L, kGridAd4F16>), dim3(n_work), dim3(kWG), shm, st, ka, td, work,
                               perm, poses, v_intra, v_inter, e_out, gpose_out, gw_out);
        else
            hipLaunchKernelGGL((energy_grad_kernel<L, UD_GRID_F32>), dim3(n_work), dim3(kWG), shm, st, ka, td, work,
                               perm, poses, v_intra, v_inter, e_out, gpose_out, gw_out);
    });
    return hipGetLastError();
}

hipError_t launch_coords(int lanes, const LigandSetDev& ls, const GridDev& g, const LaunchCap& cap,
                         const WorkItem* work, int n_work, const int32_t* perm, const float* poses, float* coords_out,
                         hipStream_t st, int compact_exh, const BoxLimits* lim, int32_t* within_out) {
    const KArgs ka = make_kargs(ls, g, cap);
    const size_t shm = lds_bytes(lanes, cap);
    const BoxLimits bl = lim ? *lim : BoxLimits{};
    UD_DISPATCH_L(lanes, {
        hipLaunchKernelGGL((coords_kernel<L>), dim3(n_work), dim3(kWG), shm, st, ka, work, perm, poses, coords_out,
                           compact_exh, bl, within_out);
    });
    return hipGetLastError();
}

hipError_t launch_cluster(const LigandSetDev& ls, int n_lig, const float* coords, const float* energy, int exh,
                          int max_size, float min_rmsd, int32_t* kept_out, int32_t* n_out, hipStream_t st) {
    hipLaunchKernelGGL(cluster_kernel, dim3(n_lig), dim3(kWG), 0, st, ls.hdr, ls.atoms, coords, energy, exh, max_size,
                       min_rmsd, kept_out, n_out);
    return hipGetLastError();
}

hipError_t launch_gather_kept(const float* poses, const float* energy, const int32_t* kept, const int32_t* n_kept,
                              int n_lig, int exh, int max_size, float* poses_out, float* e_out, hipStream_t st) {
    const size_t n = size_t(n_lig) * max_size * UD_POSE_FLOATS;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(gather_kept_kernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0, st, poses, energy, kept,
                       n_kept, n_lig, exh, max_size, poses_out, e_out);
    return hipGetLastError();
}

hipError_t launch_bounds(const LigandSetDev& ls, const int32_t* elem, int n_modes, const float* coords,
                         const int32_t* mode_lig, const int32_t* mode_ref, float* ub_out, float* lb_out,
                         hipStream_t st) {
    hipLaunchKernelGGL(bounds_kernel, dim3(n_modes), dim3(kWG), 0, st, ls.hdr, ls.atoms, elem, coords, mode_lig,
                       mode_ref, ub_out, lb_out);
    return hipGetLastError();
}

}  // namespace ud
