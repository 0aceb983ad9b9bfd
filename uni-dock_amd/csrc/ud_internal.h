// Internal device-side data layout shared by the kernels and the host side
// of the library (not part of the C ABI).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/unidock_amd.h"

namespace ud {

// Per-ligand header in HBM.
struct LigHdr {
    int32_t atom_begin;   // into the atom array
    int32_t n_atoms;
    int32_t node_begin;   // into the node array
    int32_t n_nodes;      // torsions = n_nodes - 1
    int32_t pair_begin;   // into the pair array
    int32_t n_pairs;
    int32_t max_depth;    // deepest node level (root = 0)
    int32_t local_steps;  // BFGS iterations, (25 + n_atoms) / 3 (vina.h:197)
    float num_tors;       // conf_independent torsion count
    int32_t n_heavy;
    int32_t partner_begin;  // into the partner array (2 * n_pairs records)
    int32_t partner_ofs_begin;  // the same list pre-decoded (partner_ofs), in the same array
};
static_assert(sizeof(LigHdr) == 48, "LigHdr layout");

// Atom record (16 B): local coordinates + packed metadata
//   bits 0..7   owning node
//   bits 8..15  grid map slot (0xFF = no grid term)
//   bit  16     heavy atom
//   bits 24..31 XS type (0xFF = untyped)
// Atom indices in the node, pair and partner records take 9 or 10 bits
// (UD_MAX_ATOMS = 300); nodes 8 bits (UD_MAX_TORSIONS + 1 = 49).
struct AtomRec {
    float x, y, z;
    uint32_t meta;
};

// Node record (32 B): relative origin / axis in the parent frame
// (segment::relative_origin / relative_axis, tree.h:121-131)
//   link   bits 0..15 parent node, 16..31 depth (root: parent 0, depth 0)
//   packed bits 0..9 atom begin, 10..19 atom end, 20..29 subtree atom end
struct NodeRec {
    float ro[3];
    uint32_t link;
    float ax[3];
    uint32_t packed;
};
__host__ __device__ inline uint32_t node_link(uint32_t parent, uint32_t depth) { return parent | (depth << 16); }
__host__ __device__ inline int node_parent(const NodeRec& n) { return int(n.link & 0xFFFFu); }
__host__ __device__ inline int node_depth(const NodeRec& n) { return int(n.link >> 16); }
__host__ __device__ inline uint32_t node_packed(uint32_t ab, uint32_t ae, uint32_t se) { return ab | (ae << 10) | (se << 20); }
__host__ __device__ inline int node_abegin(const NodeRec& n) { return int(n.packed & 0x3FFu); }
__host__ __device__ inline int node_aend(const NodeRec& n) { return int((n.packed >> 10) & 0x3FFu); }
__host__ __device__ inline int node_subtree_end(const NodeRec& n) { return int((n.packed >> 20) & 0x3FFu); }

// Pair record: a (bits 0..8) | b (9..17) | table slot (18..27)
typedef uint32_t PairRec;
__host__ __device__ inline uint32_t pair_rec(uint32_t a, uint32_t b, uint32_t slot) { return a | (b << 9) | (slot << 18); }
__host__ __device__ inline uint32_t pair_a(uint32_t r) { return r & 0x1FFu; }
__host__ __device__ inline uint32_t pair_b(uint32_t r) { return (r >> 9) & 0x1FFu; }
__host__ __device__ inline uint32_t pair_slot(uint32_t r) { return r >> 18; }
// The same pair pre-decoded for the search's pair pass: x = byte offsets of
// both atoms' coordinates in the group's float4 coordinate array (a * 16 |
// b * 16 << 16), y = first sample of the pair's table (slot * kTableStride).
__host__ __device__ inline uint2 pair_ofs(uint32_t a, uint32_t b, uint32_t slot, uint32_t stride) {
    return make_uint2((a * 16u) | ((b * 16u) << 16), slot * stride);
}
// Zero records after the last ligand's pre-decoded pairs: the pair pass
// prefetches records up to two batches of 64 lanes past a ligand's last
// pair without a bound check (those records are never used).
static constexpr int kPairOfsPad = 256;
// Partner record (each pair listed once under each of its atoms, in pair
// order): bits 0..8 other atom, 9..21 pair index, 22..31 table slot.  Per
// atom, AtomPartners = first partner (relative to the ligand's
// partner_begin) | count << 20.  Each atom's list is padded to a multiple of
// kPartnerPad records with records naming the atom itself and pair n_pairs,
// whose force-scale slot holds 0 (count excludes them), so that record
// batches of that size need no per-record bound.
typedef uint32_t PartnerRec;
__host__ __device__ inline uint32_t partner_rec(uint32_t other, uint32_t pair, uint32_t slot) {
    return other | (pair << 9) | (slot << 22);
}
__host__ __device__ inline uint32_t partner_other(uint32_t r) { return r & 0x1FFu; }
__host__ __device__ inline uint32_t partner_pair(uint32_t r) { return (r >> 9) & 0x1FFFu; }
__host__ __device__ inline uint32_t partner_slot(uint32_t r) { return r >> 22; }
typedef uint32_t AtomPartners;
static constexpr int kPartnerPad = 8;
// Gradient-vector capacity of the small search variant (6 + at most 10
// torsions), whose inverse Hessian fits in registers.
static constexpr int kNSmall = 16;
// The same partner record pre-decoded for the search's second pair pass:
// byte offset of the other atom's coordinates in the group's float4 array
// (other * 16, bits 0..15) | byte offset of the pair's force scale in the
// group's float array (pair * 4, bits 16..31).  Both fit 16 bits for 300
// atoms and 8,128 + 1 pairs.
__host__ __device__ inline uint32_t partner_ofs(uint32_t r) {
    return (partner_other(r) * 16u) | ((partner_pair(r) * 4u) << 16);
}
static_assert(UD_MAX_ATOMS * 16 <= 0x10000 && (UD_MAX_PAIRS + 1) * 4 <= 0x10000, "partner offsets fit 16 bits");

// Pair-table entry for the derivative path: e_k, dor_k, e_{k+1} - e_k,
// dor_{k+1} - dor_k (precalculate_element::eval_deriv, precalculate.h:48-65)
static constexpr int kTableN = 2051;     // sz(32 * 64) + 3 (precalculate.h:142-145)
// entries 0..2047 are reachable with r^2 < 64 (2048 is kept for r^2 == 64 in map builds)

// AutoDock 4.2 (SF_AD42): samples of its 20.48 A cutoff, sz(32 * 20.48^2) + 3
static constexpr int kAd4TableN = 13424;
static constexpr float kAd4Cutoff = 20.48f;  // scoring_function.h:82-83
static constexpr int kAd4Types = 35;         // AD_TYPE_SIZE (atom_constants.h:81)
static constexpr int kAd4MapElec = UD_AD4_MAP_ELEC, kAd4MapDsolv = UD_AD4_MAP_DSOLV;
// Per AutoDock type (atom_constants.h:155-194): vdW radius and depth,
// H-bond depth and radius, solvation parameter, atomic volume.
struct Ad4Params {
    float radius[kAd4Types], depth[kAd4Types], hb_depth[kAd4Types], hb_r[kAd4Types], solv[kAd4Types],
        vol[kAd4Types];
};
const Ad4Params& ad4_params();  // host copy
// AD type of an atom's grid term (ad4cache.cpp:223-240): G types none, CG
// types the C map
inline int ad4_grid_type(int t) {
    if (t < 0 || t >= kAd4Types) return -1;
    if (t >= 22 && t <= 25) return -1;
    if (t >= 26 && t <= 29) return 0;
    return t;
}

struct GridDev {
    const void* maps;      // n_maps * n_points values (float or __half bits)
    float init[3];
    float factor[3];
    float factor_inv[3];
    float dimm1[3];
    int32_t n[3];
    int32_t n_points;
    float slope;
    float scale[UD_MAX_MAPS];  // FP16 maps: value = half * scale
    int32_t elec_slot, dsolv_slot;  // AD4.2: slots of the electrostatic / desolvation maps
};

struct LigandSetDev {
    const LigHdr* hdr;
    const AtomRec* atoms;
    const NodeRec* nodes;
    const PairRec* pairs;
    const uint2* pair_ofs;  // pre-decoded pairs, indexed like pairs
    // partner records of every ligand, then the same records pre-decoded
    // (partner_ofs); LigHdr::partner_begin and partner_ofs_begin index both
    const PartnerRec* partners;
    const AtomPartners* atom_partners;  // per atom, indexed like atoms
    const float2* table_deriv;  // (e, dor) per sample, slot * kTableStride
    const float* table_fast;    // slot * kTableUsed
    // AD4.2: per-atom partial charges (indexed like atoms), per-pair charge
    // coefficients (q_a q_b, desolvation coefficient; indexed like pairs), and
    // the shared electrostatic / desolvation radial tables
    const float* charge;
    const float2* pair_coef;
    const float4* ad4_deriv;  // (E e, E dor, G e, G dor) per sample, kAd4TableN
    const float2* ad4_fast;   // (E fast, G fast) per sample
};

// Receptor atoms for scoring against the explicit receptor (non_cache,
// non_cache.cpp) in the refinement: cells of about 3 A over the box
// (szv_grid_dims, szv_grid.cpp:86-96), each listing in receptor order the
// typed atoms within the cutoff of the cell (szv_grid.cpp:24-55).
static constexpr int kCellPad = 64;  // far-away atoms after the cell lists (see build_cells)
struct RecDev {
    const float4* atoms;        // x, y, z, XS type bits; the cell lists concatenated, then kCellPad far atoms
    const int32_t* cell_start;  // n_cells + 1 offsets into atoms
    int32_t dim[3];             // cells per axis
    float init[3], range[3], end[3];  // box begin, end - begin, end
};

// Maps and pair tables of each scoring term alone (weight 1, the others 0),
// for term-weight gradients: the energy is linear in the weights before curl.
struct TermsDev {
    const float* maps;     // [term][slot][n_points], FP32, same slots as GridDev
    const float2* tables;  // [term][type pair][kTableStride] (e, dor); e is read
    const float4* ad4_unit;  // AD4.2: electrostatic / desolvation radial tables at unit weights
    int32_t n_terms;       // 6 (Vina) or 5 (Vinardo), glue included; 4 (AD4.2: vdW, H-bond, elec, desolvation)
    int32_t n_maps;
};

// Work item: one ligand and a run of consecutive poses/chains of it.
struct WorkItem {
    int32_t lig;
    int32_t first;   // first pose / chain index (global)
    int32_t count;   // <= groups per workgroup
    int32_t pad;
};

static constexpr int kMaxRefinePasses = 8;
struct SearchDev {
    int32_t exh;
    int32_t global_steps;
    int32_t local_steps_override;  // > 0 overrides the per-ligand value
    uint32_t seed_lo, seed_hi;
    float hunt[3];
    float authentic[3];
    float temperature;
    float amplitude;
    float box_lo[3], box_hi[3];
    int32_t ls_f0_zero;  // first line search of each BFGS against f0 = 0 (bfgs_warp) instead of f(x)
    // refinement against the explicit receptor: passes (> 1: a pose whose
    // heavy atoms leave [box_lo, box_hi] after a pass runs the next one) and
    // each pass's out-of-box penalty slope
    int32_t refine_passes;
    float refine_slope[kMaxRefinePasses];
};

// Capacity of one launch (dynamic LDS is sized from it)
struct LaunchCap {
    int32_t n_atoms;  // max atoms
    int32_t n_nodes;  // max nodes
    int32_t n_pairs;  // max intramolecular pairs
    int32_t pair_forces;  // 1: per-pair force buffer in LDS (pair-parallel forces)
};

}  // namespace ud
