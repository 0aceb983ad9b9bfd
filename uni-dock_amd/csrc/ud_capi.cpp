// Host side of the C ABI (include/unidock_amd.h): context, uploads, size
// buckets, work lists and launches.  Every entry point returns a status code;
// nothing throws across the ABI.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <limits>
#include <cmath>
#include <cstring>
#include <string>
#include <thread>
#include <mutex>
#include <vector>

#include "../../include/unidock_amd.h"
#include "ud_internal.h"
#include "ud_launch.h"
#include "ud_post.h"

using namespace ud;

namespace ud {
// AutoDock 4.2 atom kinds by AD type index (atom_constants.h:155-194)
const Ad4Params& ad4_params() {
    static const Ad4Params P = [] {
        const float t[kAd4Types][6] = {
            {2.000f, 0.150f, 0.f, 0.f, -0.00143f, 33.5103f},  {2.000f, 0.150f, 0.f, 0.f, -0.00052f, 33.5103f},   // C A
            {1.750f, 0.160f, 0.f, 0.f, -0.00162f, 22.4493f},  {1.600f, 0.200f, 0.f, 0.f, -0.00251f, 17.1573f},   // N O
            {2.100f, 0.200f, 0.f, 0.f, -0.00110f, 38.7924f},  {2.000f, 0.200f, 0.f, 0.f, -0.00214f, 33.5103f},   // P S
            {1.000f, 0.020f, 0.f, 0.f, 0.00051f, 0.f},        {1.545f, 0.080f, 0.f, 0.f, -0.00110f, 15.4480f},   // H F
            {2.360f, 0.550f, 0.f, 0.f, -0.00110f, 55.0585f},                                                     // I
            {1.750f, 0.160f, -5.f, 1.9f, -0.00162f, 22.4493f}, {1.600f, 0.200f, -5.f, 1.9f, -0.00251f, 17.1573f},  // NA OA
            {2.000f, 0.200f, -1.f, 2.5f, -0.00214f, 33.5103f}, {1.000f, 0.020f, 1.f, 0.f, 0.00051f, 0.f},          // SA HD
            {0.650f, 0.875f, 0.f, 0.f, -0.00110f, 1.560f},    {0.650f, 0.875f, 0.f, 0.f, -0.00110f, 2.140f},     // Mg Mn
            {0.740f, 0.550f, 0.f, 0.f, -0.00110f, 1.700f},    {0.990f, 0.550f, 0.f, 0.f, -0.00110f, 2.770f},     // Zn Ca
            {0.650f, 0.010f, 0.f, 0.f, -0.00110f, 1.840f},    {2.045f, 0.276f, 0.f, 0.f, -0.00110f, 35.8235f},   // Fe Cl
            {2.165f, 0.389f, 0.f, 0.f, -0.00110f, 42.5661f},  {2.300f, 0.200f, 0.f, 0.f, -0.00143f, 50.9650f},   // Br Si
            {2.400f, 0.550f, 0.f, 0.f, -0.00110f, 57.9058f},                                                     // At
            {0, 0, 0, 0, 0, 0}, {0, 0, 0, 0, 0, 0}, {0, 0, 0, 0, 0, 0}, {0, 0, 0, 0, 0, 0},                      // G0-G3
            {2.000f, 0.150f, 0.f, 0.f, -0.00143f, 33.5103f},  {2.000f, 0.150f, 0.f, 0.f, -0.00143f, 33.5103f},   // CG0 CG1
            {2.000f, 0.150f, 0.f, 0.f, -0.00143f, 33.5103f},  {2.000f, 0.150f, 0.f, 0.f, -0.00143f, 33.5103f},   // CG2 CG3
            {0, 0, 0, 0, 0, 0},                                                                                  // W
            {1.600f, 0.200f, 0.f, 0.f, -0.00251f, 17.1573f},  {1.750f, 0.160f, 0.f, 0.f, -0.00162f, 22.4493f},   // OXA NXA
            {1.600f, 0.200f, 0.f, 0.f, -0.00251f, 17.1573f},  {1.750f, 0.160f, 0.f, 0.f, -0.00162f, 22.4493f},   // OXD NXD
        };
        Ad4Params p;
        for (int i = 0; i < kAd4Types; ++i) {
            p.radius[i] = t[i][0]; p.depth[i] = t[i][1]; p.hb_depth[i] = t[i][2];
            p.hb_r[i] = t[i][3]; p.solv[i] = t[i][4]; p.vol[i] = t[i][5];
        }
        return p;
    }();
    return P;
}
}  // namespace ud

namespace {
// The charge-dependent AD4.2 pair terms as shared radial tables, sampled like
// precalculate_byatom (precalculate.h:206-241, 66-85) at rs[k] = sqrt(k / 32),
// k < kAd4TableN: E(r) = w_elec 332 min(100, 1 / (r eps(r))) (times q_a q_b per
// pair) and G(r) = w_dsolv exp(-r^2 / (2 3.6^2)) (times the pair's
// desolvation coefficient), both 0 from 20.48 A (ad4_electrostatic,
// ad4_solvation, potentials.h:346-416).  deriv = (E, dE/r, G, dG/r) with
// the central-difference dor and 0 at both ends; fast = midpoints.
void ad4_radial_tables(float w_elec, float w_dsolv, std::vector<float4>& deriv, std::vector<float2>& fast) {
    const int n = kAd4TableN;
    std::vector<double> rs(n), E(n), G(n);
    const double B = 78.4 + 8.5525, lB = -B * 0.003627;
    for (int k = 0; k < n; ++k) {
        const double r = std::sqrt(double(k) / 32.0);
        rs[k] = r;
        if (r >= double(kAd4Cutoff)) { E[k] = G[k] = 0.0; continue; }
        const double diel = -8.5525 + B / (1.0 + 7.7839 * std::exp(lB * r));
        E[k] = double(w_elec) * 332.0 * ((r < 1.1920928955078125e-07) ? 100.0 / diel : std::min(100.0, 1.0 / (r * diel)));
        G[k] = double(w_dsolv) * std::exp(-0.5 * (r / 3.6) * (r / 3.6));
    }
    auto dor = [&](const std::vector<double>& e, int k) {
        if (k == 0 || k == n - 1) return 0.0;
        return (e[k + 1] - e[k - 1]) / ((rs[k + 1] - rs[k - 1]) * rs[k]);
    };
    deriv.resize(n);
    fast.resize(n);
    for (int k = 0; k < n; ++k) {
        deriv[k] = make_float4(float(E[k]), float(dor(E, k)), float(G[k]), float(dor(G, k)));
        const double e1 = (k + 1 < n) ? E[k + 1] : 0.0, g1 = (k + 1 < n) ? G[k + 1] : 0.0;
        fast[k] = make_float2(float((E[k] + e1) / 2.0), float((G[k] + g1) / 2.0));
    }
}
}  // namespace

struct ud_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    // scoring
    int sf = -1;
    float w[7] = {0, 0, 0, 0, 0, 0, 0};  // term weights (glue at [5] for Vina, [4] for Vinardo) + internal rot
    float rot_internal = 0.f;
    float2* d_table_deriv = nullptr;
    float* d_table_fast = nullptr;
    // AD4.2 shared radial tables (electrostatic, desolvation Gaussian)
    float4* d_ad4_deriv = nullptr;
    float2* d_ad4_fast = nullptr;
    // grid
    bool box_set = false;
    float init[3], factor[3], factor_inv[3], dimm1[3];
    float gend[3];  // box end (grid_dims end, vina.cpp:536-538)
    int n[3] = {0, 0, 0};
    int n_points = 0;
    int precision = UD_GRID_F32;
    int n_maps = 0;
    int map_types[UD_MAX_MAPS];
    int slot_of_type[40];  // XS types 0..35, or AD types 0..34 and the AD4.2 elec / dsolv maps 35, 36
    bool maps_ad4 = false;  // the maps are AD4.2 maps
    float scale[UD_MAX_MAPS];
    void* d_maps = nullptr;
    // receptor atoms of the last ud_compute_maps (x, y, z, XS type bits)
    float4* d_rec = nullptr;
    size_t rec_cap = 0;
    int n_rec = 0;
    bool have_rec = false;
    // receptor cell lists of the explicit-receptor scoring (refinement)
    float4* d_cells = nullptr;
    int32_t* d_cell_start = nullptr;
    size_t cells_cap = 0, cstart_cap = 0;
    int cell_dim[3] = {1, 1, 1};
    // receptor atoms (x, y, z, AD type bits) and charges of the last
    // ud_compute_maps_ad4: the input of the AD4.2 per-term maps
    float4* d_rec_ad4 = nullptr;
    float* d_rec_q = nullptr;
    size_t rec_ad4_cap = 0, rec_q_cap = 0;
    int n_rec_ad4 = 0;
    bool have_rec_ad4 = false;
    // per-term maps and tables for term-weight gradients (built on demand);
    // AD4.2 adds the electrostatic / desolvation radial tables at unit weights
    float2* d_term_tab = nullptr;
    float* d_term_maps = nullptr;
    float4* d_term_ad4 = nullptr;
    size_t term_tab_cap = 0, term_maps_cap = 0, term_ad4_cap = 0;
    bool terms_ok = false;
    // ligands
    int n_lig = 0;
    std::vector<LigHdr> hdr;
    std::vector<int> lig_nat, lig_nn;  // per ligand atoms / nodes
    std::vector<std::vector<int>> lig_xs;  // per ligand the atom types (XS, or AD for AD4.2)
    bool lig_ad4 = false;                  // ligands were set with AD4.2 typing
    std::vector<int> h_elem;               // element class per atom (RMSD lower bound)
    float* d_charge = nullptr;             // AD4.2 charges, per atom
    float2* d_pair_coef = nullptr;         // AD4.2 charge coefficients, per pair
    LigHdr* d_hdr = nullptr;
    AtomRec* d_atoms = nullptr;
    NodeRec* d_nodes = nullptr;
    PairRec* d_pairs = nullptr;
    uint2* d_pair_ofs = nullptr;
    PartnerRec* d_partners = nullptr;  // records, then the same pre-decoded
    AtomPartners* d_apar = nullptr;
    // scratch / results
    WorkItem* d_work = nullptr;
    size_t work_cap = 0;
    int32_t* d_perm = nullptr;
    size_t perm_cap = 0;
    float* d_poses = nullptr;
    size_t poses_cap = 0;
    float* d_f1 = nullptr;
    float* d_f2 = nullptr;
    float* d_f3 = nullptr;
    size_t f_cap = 0, f2_cap = 0, f3_cap = 0;
    std::vector<AtomRec> h_atoms;  // host copy of the atom records
    bool slots_dirty = true;
    int32_t* d_i1 = nullptr;
    size_t i_cap = 0;
    // post-processing on the device: compact chain coordinates, cluster
    // lists, per-atom element classes, mode coordinates, ligands, reference
    // modes and RMSD bounds
    float* d_cc = nullptr;
    size_t cc_cap = 0;
    int32_t* d_kept = nullptr;
    size_t kept_cap = 0;
    int32_t* d_nkept = nullptr;
    size_t nkept_cap = 0;
    float* d_kp = nullptr;  // kept chains' poses and energies, one slot per (ligand, mode)
    float* d_ke = nullptr;
    size_t kp_cap = 0, ke_cap = 0;
    int32_t* d_elem = nullptr;
    size_t elem_cap = 0;
    float* d_mc = nullptr;
    size_t mc_cap = 0;
    int32_t* d_mlig = nullptr;
    int32_t* d_mref = nullptr;
    size_t mlig_cap = 0, mref_cap = 0;
    float* d_ub = nullptr;
    float* d_lb = nullptr;
    size_t ub_cap = 0, lb_cap = 0;
    float* d_out_pose = nullptr;
    float* d_out_e = nullptr;
    long long* d_out_evals = nullptr;
    size_t out_cap = 0;
    size_t out_n = 0;
    int last_exh = 0;  // exhaustiveness of the last dock
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_ms = 0;
    int last_launches = 0;
    // ud_dock runs its size-bucket launches concurrently, launch k on
    // bstream[k] (forked from and joined back to `stream` with events);
    // lev0[k] / lev[k] bracket launch k
    hipStream_t bstream[UD_MAX_LAUNCHES] = {};
    hipEvent_t lev0[UD_MAX_LAUNCHES] = {};
    hipEvent_t lev[UD_MAX_LAUNCHES] = {};
    ud_launch_info linfo[UD_MAX_LAUNCHES] = {};
    // lanes per chain per size bucket (ud_autotune / ud_set_bucket_lanes; 0 = default)
    int tuned_lanes[UD_N_BUCKETS] = {};
    int tuned_pb[UD_N_BUCKETS] = {-1, -1, -1, -1, -1, -1, -1, -1};  // per-pair LDS buffer: -1 default, 0 off, 1 on
    int forced_lanes = 0;  // ud_set_lanes
};

namespace {

int fail(ud_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    return code;
}
#define UD_HIP(c, call)                                                                         \
    do {                                                                                        \
        hipError_t e_ = (call);                                                                 \
        if (e_ != hipSuccess) return fail(c, UD_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <class T> int grow(ud_ctx* c, T*& p, size_t& cap, size_t n) {
    if (n <= cap && p) return UD_OK;
    if (p) (void)hipFree(p);
    p = nullptr;
    size_t nn = std::max<size_t>(n, 1);
    if (hipMalloc(reinterpret_cast<void**>(&p), nn * sizeof(T)) != hipSuccess) {
        cap = 0;
        return fail(c, UD_ERR_NOMEM, "device allocation failed");
    }
    cap = nn;
    return UD_OK;
}

int tri_index(int i, int j) { return (i <= j) ? i + j * (j + 1) / 2 : j + i * (i + 1) / 2; }

// cache.cpp:160-199: G types carry no grid term, CG carbons use C_H / C_P
int grid_type_of(int t) {
    if (t < 0 || t >= 36) return -1;
    if (t == 21 || t == 24 || t == 27 || t == 30) return -1;
    if (t == 19 || t == 22 || t == 25 || t == 28) return 0;
    if (t == 20 || t == 23 || t == 26 || t == 29) return 1;
    return t;
}

// the scoring family of the maps / ligands must match the scoring function
bool sf_ad4(const ud_ctx* c) { return c->sf == UD_SF_AD4; }
int grid_format(const ud_ctx* c) { return c->precision + (sf_ad4(c) ? kGridAd4F32 : 0); }
int map_type_of(const ud_ctx* c, int t) { return sf_ad4(c) ? ad4_grid_type(t) : grid_type_of(t); }

GridDev grid_dev(const ud_ctx* c) {
    GridDev g;
    std::memset(&g, 0, sizeof(g));
    g.maps = c->d_maps;
    for (int i = 0; i < 3; ++i) {
        g.init[i] = c->init[i];
        g.factor[i] = c->factor[i];
        g.factor_inv[i] = c->factor_inv[i];
        g.dimm1[i] = c->dimm1[i];
        g.n[i] = c->n[i];
    }
    g.n_points = c->n_points;
    g.slope = 1e6f;  // vina.cpp:515
    for (int i = 0; i < UD_MAX_MAPS; ++i) g.scale[i] = c->scale[i];
    g.elec_slot = c->maps_ad4 ? c->slot_of_type[kAd4MapElec] : 0;
    g.dsolv_slot = c->maps_ad4 ? c->slot_of_type[kAd4MapDsolv] : 0;
    return g;
}

LigandSetDev lig_dev(const ud_ctx* c) {
    LigandSetDev l;
    l.hdr = c->d_hdr;
    l.atoms = c->d_atoms;
    l.nodes = c->d_nodes;
    l.pairs = c->d_pairs;
    l.pair_ofs = c->d_pair_ofs;
    l.partners = c->d_partners;
    l.atom_partners = c->d_apar;
    l.table_deriv = c->d_table_deriv;
    l.table_fast = c->d_table_fast;
    l.charge = c->d_charge;
    l.pair_coef = c->d_pair_coef;
    l.ad4_deriv = c->d_ad4_deriv;
    l.ad4_fast = c->d_ad4_fast;
    return l;
}

// largest lanes-per-group whose LDS footprint fits one workgroup (64 KiB);
// the per-pair force buffer is used when the launch still fits with it,
// unless pair_buffer is 0
int fit_lanes(int lanes, LaunchCap& cap, int pair_buffer = -1) {
    cap.pair_forces = 0;
    while (lanes < 64 && lds_bytes(lanes, cap) > 65536) lanes *= 2;
    LaunchCap with = cap;
    with.pair_forces = 1;
    if (pair_buffer != 0 && lds_bytes(lanes, with) <= 65536) cap.pair_forces = 1;
    return lanes;
}

int ready_for_eval(ud_ctx* c) {
    if (c->sf < 0) return fail(c, UD_ERR_STATE, "scoring function not set");
    if (!c->d_maps) return fail(c, UD_ERR_STATE, "grid maps not computed");
    if (c->n_lig == 0) return fail(c, UD_ERR_STATE, "no ligands set");
    if (c->maps_ad4 != sf_ad4(c)) return fail(c, UD_ERR_STATE, "the maps were made for another scoring function");
    if (c->lig_ad4 != sf_ad4(c)) return fail(c, UD_ERR_STATE, "the ligands were set for another scoring function");
    // ad4cache::are_atom_types_grid_initialized (ad4cache.cpp:290-320) and its Vina counterpart
    for (int l = 0; l < c->n_lig; ++l)
        for (int t : c->lig_xs[l]) {
            int gt = map_type_of(c, t);
            if (gt >= 0 && c->slot_of_type[gt] < 0)
                return fail(c, UD_ERR_STATE, "affinity map for atom type " + std::to_string(gt) + " is not present");
        }
    return UD_OK;
}

// Upload atoms with their grid slots (depends on the current maps).
int upload_atoms(ud_ctx* c, const std::vector<AtomRec>& atoms) {
    size_t cap = 0;
    AtomRec* p = nullptr;
    if (c->d_atoms) (void)hipFree(c->d_atoms);
    c->d_atoms = nullptr;
    int rc = grow(c, p, cap, atoms.size());
    if (rc) return rc;
    c->d_atoms = p;
    UD_HIP(c, hipMemcpy(c->d_atoms, atoms.data(), atoms.size() * sizeof(AtomRec), hipMemcpyHostToDevice));
    return UD_OK;
}

// Build poses grouped by ligand: perm (sorted pose indices) and work items
// of up to `groups` poses of one ligand each.
void make_pose_work(int n_poses, const int32_t* pose_ligand, int groups, std::vector<int32_t>& perm,
                    std::vector<WorkItem>& work) {
    perm.resize(n_poses);
    for (int i = 0; i < n_poses; ++i) perm[i] = i;
    std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) { return pose_ligand[a] < pose_ligand[b]; });
    work.clear();
    int i = 0;
    while (i < n_poses) {
        int lig = pose_ligand[perm[i]];
        int j = i;
        while (j < n_poses && pose_ligand[perm[j]] == lig && j - i < groups) ++j;
        work.push_back(WorkItem{lig, i, j - i, 0});
        i = j;
    }
}

LaunchCap caps_of(const ud_ctx* c, const std::vector<int>* ligs) {
    LaunchCap cap{1, 1, 0, 0};
    const int n = ligs ? int(ligs->size()) : c->n_lig;
    for (int k = 0; k < n; ++k) {
        const int l = ligs ? (*ligs)[k] : k;
        cap.n_atoms = std::max(cap.n_atoms, c->lig_nat[l]);
        cap.n_nodes = std::max(cap.n_nodes, c->lig_nn[l]);
        cap.n_pairs = std::max(cap.n_pairs, c->hdr[l].n_pairs);
    }
    return cap;
}
LaunchCap caps_all(const ud_ctx* c) { return caps_of(c, nullptr); }

int check_pose_args(ud_ctx* c, int n_poses, const int32_t* pose_ligand) {
    if (n_poses < 0 || (n_poses > 0 && !pose_ligand)) return fail(c, UD_ERR_INVALID, "bad pose arguments");
    for (int i = 0; i < n_poses; ++i)
        if (pose_ligand[i] < 0 || pose_ligand[i] >= c->n_lig) return fail(c, UD_ERR_INVALID, "pose ligand index out of range");
    return UD_OK;
}

}  // namespace

extern "C" {

void ud_default_search_params(ud_search_params* p) {
    if (!p) return;
    p->exhaustiveness = 8;
    p->global_steps = 20;
    p->local_steps = 0;
    p->lanes_per_chain = 0;
    p->seed = 181129;
    for (int i = 0; i < 3; ++i) {
        p->hunt_cap[i] = 10.f;
        p->authentic_v[i] = 1000.f;
    }
    p->temperature = 1.2f;
    p->mutation_amplitude = 2.f;
    p->chains_per_group = 0;
    p->first_ls_f0_zero = 0;
    p->single_launch = 0;
}

}  // extern "C"

// ---- scratch of the allocation-free scoring entry point
// The pose permutation (4 B per pose) then the work list (16 B per item, at
// most one item per pose) of ud_score_grad_device_noalloc.  Host code must
// not pass `scratch` itself to a kernel or a copy: its device address comes
// from hipGetSymbolAddress, once per device, kept in g_scratch.
static_assert(sizeof(WorkItem) == 16, "work items are 16 B");
static_assert(UD_NOALLOC_SCRATCH_BYTES == size_t(UD_NOALLOC_MAX_POSES) * (4 + sizeof(WorkItem)), "scratch size");
static __device__ __align__(256) unsigned char scratch[UD_NOALLOC_SCRATCH_BYTES];
namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch[kMaxDevices] = {};

// Resolve this device's copy of `scratch` (the device must be current).
int scratch_init(int device) {
    if (device < 0 || device >= kMaxDevices) return UD_ERR_INVALID;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch[device]) return UD_OK;
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return UD_ERR_HIP;
    g_scratch[device] = p;
    return UD_OK;
}
void* scratch_of(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return (device >= 0 && device < kMaxDevices) ? g_scratch[device] : nullptr;
}
}  // namespace

extern "C" {

int ud_create(int device_id, void* stream, ud_ctx** out) {
    if (!out) return UD_ERR_INVALID;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return UD_ERR_HIP;
    if (device_id < 0 || device_id >= ndev) return UD_ERR_INVALID;
    if (hipSetDevice(device_id) != hipSuccess) return UD_ERR_HIP;
    if (device_id >= kMaxDevices) return UD_ERR_INVALID;
    if (scratch_init(device_id) != UD_OK) return UD_ERR_HIP;
    ud_ctx* c = new ud_ctx();
    c->device = device_id;
    c->stream = reinterpret_cast<hipStream_t>(stream);
    for (int i = 0; i < 36; ++i) c->slot_of_type[i] = -1;
    for (int i = 0; i < UD_MAX_MAPS; ++i) c->scale[i] = 1.f;
    bool ev_ok = hipEventCreate(&c->ev0) == hipSuccess && hipEventCreate(&c->ev1) == hipSuccess;
    for (int k = 0; k < UD_MAX_LAUNCHES && ev_ok; ++k)
        ev_ok = hipEventCreate(&c->lev[k]) == hipSuccess && hipEventCreate(&c->lev0[k]) == hipSuccess &&
                hipStreamCreateWithFlags(&c->bstream[k], hipStreamNonBlocking) == hipSuccess;
    if (!ev_ok) {
        ud_destroy(c);
        return UD_ERR_HIP;
    }
    *out = c;
    return UD_OK;
}

int ud_destroy(ud_ctx* c) {
    if (!c) return UD_ERR_INVALID;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    void* ptrs[] = {c->d_table_deriv, c->d_table_fast, c->d_maps, c->d_rec, c->d_term_tab, c->d_term_maps, c->d_cells,
                    c->d_rec_ad4, c->d_rec_q, c->d_term_ad4,
                    c->d_cell_start,
                    c->d_hdr, c->d_atoms, c->d_nodes, c->d_pairs, c->d_pair_ofs, c->d_partners, c->d_apar,
                    c->d_work, c->d_perm, c->d_poses, c->d_f1, c->d_f2, c->d_f3, c->d_i1, c->d_out_pose, c->d_out_e,
                    c->d_out_evals, c->d_ad4_deriv, c->d_ad4_fast, c->d_charge, c->d_pair_coef, c->d_cc, c->d_kept,
                    c->d_nkept, c->d_kp, c->d_ke, c->d_elem, c->d_mc, c->d_mlig, c->d_mref, c->d_ub, c->d_lb};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    for (int k = 0; k < UD_MAX_LAUNCHES; ++k) {
        if (c->lev[k]) (void)hipEventDestroy(c->lev[k]);
        if (c->lev0[k]) (void)hipEventDestroy(c->lev0[k]);
        if (c->bstream[k]) (void)hipStreamDestroy(c->bstream[k]);
    }
    delete c;
    return UD_OK;
}

const char* ud_last_error(const ud_ctx* c) { return c ? c->err.c_str() : "null context"; }

int ud_set_scoring(ud_ctx* c, int sf, const double* weights, int n_weights) {
    if (!c) return UD_ERR_INVALID;
    if (sf != UD_SF_VINA && sf != UD_SF_VINARDO && sf != UD_SF_AD4)
        return fail(c, UD_ERR_INVALID, "unknown scoring function");
    const int need = (sf == UD_SF_VINA) ? 7 : 6;
    if (!weights || n_weights != need) return fail(c, UD_ERR_INVALID, "wrong number of weights");
    UD_HIP(c, hipSetDevice(c->device));
    c->sf = sf;
    c->terms_ok = false;
    for (int i = 0; i < 7; ++i) c->w[i] = 0.f;
    for (int i = 0; i < need - 1; ++i) c->w[i] = float(weights[i]);
    // internal rotational weight 5 * w_rot / 0.1 - 1 (vina.cpp:404, 425); the
    // AD4.2 torsional weight is used as given (vina.cpp:435-450)
    c->rot_internal = (sf == UD_SF_AD4) ? float(weights[need - 1]) : float(5.0 * weights[need - 1] / 0.1 - 1.0);
    size_t cap1 = 0, cap2 = 0;
    if (!c->d_table_deriv) {
        int rc = grow(c, c->d_table_deriv, cap1, size_t(kTypePairs) * kTableStride);
        if (rc) return rc;
        rc = grow(c, c->d_table_fast, cap2, size_t(kTypePairs) * kTableStride);
        if (rc) return rc;
    }
    UD_HIP(c, launch_tables(sf, c->w, c->d_table_deriv, c->d_table_fast, c->stream));
    if (sf == UD_SF_AD4) {
        std::vector<float4> deriv;
        std::vector<float2> fast;
        ad4_radial_tables(c->w[2], c->w[3], deriv, fast);
        size_t cap3 = 0, cap4 = 0;
        if (!c->d_ad4_deriv) {
            int rc = grow(c, c->d_ad4_deriv, cap3, size_t(kAd4TableN));
            if (rc) return rc;
            rc = grow(c, c->d_ad4_fast, cap4, size_t(kAd4TableN));
            if (rc) return rc;
        }
        UD_HIP(c, hipMemcpyAsync(c->d_ad4_deriv, deriv.data(), sizeof(float4) * kAd4TableN, hipMemcpyHostToDevice,
                                 c->stream));
        UD_HIP(c, hipMemcpyAsync(c->d_ad4_fast, fast.data(), sizeof(float2) * kAd4TableN, hipMemcpyHostToDevice,
                                 c->stream));
    }
    UD_HIP(c, hipStreamSynchronize(c->stream));
    return UD_OK;
}

int ud_set_box(ud_ctx* c, const double* center, const double* size, double granularity, int32_t* n_out) {
    if (!c || !center || !size) return UD_ERR_INVALID;
    if (!(granularity > 0)) return fail(c, UD_ERR_INVALID, "granularity must be positive");
    long long pts = 1;
    for (int i = 0; i < 3; ++i) {
        if (!(size[i] > 0)) return fail(c, UD_ERR_INVALID, "grid box dimensions must be greater than 0");
        const long nv = long(std::ceil(size[i] / granularity));
        // compute_vina_maps geometry (vina.cpp:527-539) and grid::init (grid.cpp:24-37), in float
        const float real_span = float(granularity * double(nv));
        const float begin = float(center[i]) - real_span / 2.f;
        const float end = begin + real_span;
        c->n[i] = int(nv) + 1;
        c->init[i] = begin;
        c->gend[i] = end;
        const float range = end - begin;
        c->dimm1[i] = float(c->n[i] - 1);
        c->factor[i] = c->dimm1[i] / range;
        c->factor_inv[i] = 1.f / c->factor[i];
        pts *= c->n[i];
        if (n_out) n_out[i] = c->n[i];
    }
    if (pts > UD_MAX_GRID_POINTS) return fail(c, UD_ERR_CAPACITY, "grid has more points than UD_MAX_GRID_POINTS");
    c->n_points = int(pts);
    c->box_set = true;
    c->have_rec = false;
    c->have_rec_ad4 = false;
    c->terms_ok = false;
    if (c->d_maps) {
        (void)hipFree(c->d_maps);
        c->d_maps = nullptr;
    }
    c->n_maps = 0;
    for (int i = 0; i < 40; ++i) c->slot_of_type[i] = -1;
    return UD_OK;
}

static int store_maps(ud_ctx* c, float* d_f32, int n_types, const int32_t* types, int precision) {
    // d_f32: n_types maps of n_points floats on the device (consumed)
    if (c->d_maps) {
        (void)hipFree(c->d_maps);
        c->d_maps = nullptr;
    }
    for (int i = 0; i < 40; ++i) c->slot_of_type[i] = -1;
    c->maps_ad4 = sf_ad4(c);
    c->n_maps = n_types;
    for (int m = 0; m < n_types; ++m) {
        c->map_types[m] = types[m];
        c->slot_of_type[types[m]] = m;
        c->scale[m] = 1.f;
    }
    c->precision = precision;
    c->slots_dirty = true;
    c->terms_ok = false;
    if (precision == UD_GRID_F32) {
        c->d_maps = d_f32;
        return UD_OK;
    }
    // FP16 with a power-of-two scale per map: max |v| / scale in [16384, 32768)
    unsigned int* d_mx = nullptr;
    UD_HIP(c, hipMalloc(&d_mx, sizeof(unsigned int) * n_types));
    UD_HIP(c, hipMemsetAsync(d_mx, 0, sizeof(unsigned int) * n_types, c->stream));
    UD_HIP(c, launch_grid_absmax(d_f32, c->n_points, n_types, d_mx, c->stream));
    std::vector<unsigned int> mx(n_types);
    UD_HIP(c, hipMemcpyAsync(mx.data(), d_mx, sizeof(unsigned int) * n_types, hipMemcpyDeviceToHost, c->stream));
    UD_HIP(c, hipStreamSynchronize(c->stream));
    std::vector<float> inv(n_types);
    for (int m = 0; m < n_types; ++m) {
        float v;
        std::memcpy(&v, &mx[m], sizeof(float));
        int e = 0;
        if (v > 0.f && std::isfinite(v)) {
            std::frexp(v, &e);  // v = f * 2^e, f in [0.5, 1)
            e -= 15;            // v / 2^e in [16384, 32768)
        }
        c->scale[m] = std::ldexp(1.f, e);
        inv[m] = std::ldexp(1.f, -e);
    }
    float* d_inv = nullptr;
    UD_HIP(c, hipMalloc(&d_inv, sizeof(float) * n_types));
    UD_HIP(c, hipMemcpyAsync(d_inv, inv.data(), sizeof(float) * n_types, hipMemcpyHostToDevice, c->stream));
    void* d_h = nullptr;
    if (hipMalloc(&d_h, size_t(2) * n_types * c->n_points) != hipSuccess) return fail(c, UD_ERR_NOMEM, "map allocation failed");
    UD_HIP(c, launch_grid_to_half(d_f32, c->n_points, n_types, d_inv, d_h, c->stream));
    UD_HIP(c, hipStreamSynchronize(c->stream));
    (void)hipFree(d_inv);
    (void)hipFree(d_mx);
    (void)hipFree(d_f32);
    c->d_maps = d_h;
    return UD_OK;
}

// squared distance from v to the box [lo, hi] (brick_distance_sqr, brick.h:27-48)
static float brick_distance_sqr(const float* lo, const float* hi, const float4& v) {
    const float p[3] = {v.x, v.y, v.z};
    float d2 = 0.f;
    for (int i = 0; i < 3; ++i) {
        const float cl = (p[i] <= lo[i]) ? lo[i] : (p[i] >= hi[i]) ? hi[i] : p[i];
        d2 += (cl - p[i]) * (cl - p[i]);
    }
    return d2;
}

// Cell lists of the explicit-receptor scoring, szv_grid (szv_grid.cpp:24-55,
// 76-96): about 3 A cells over the box; each lists, in receptor order, the
// typed atoms closer than the cutoff to the cell (of those closer than the
// cutoff to the box).
static int build_cells(ud_ctx* c, const std::vector<float4>& rec) {
    float range[3];
    for (int i = 0; i < 3; ++i) {
        range[i] = c->gend[i] - c->init[i];
        const int n_int = int(range[i] / 3.f);
        c->cell_dim[i] = n_int < 1 ? 1 : n_int;
    }
    const float cut2 = 64.f;
    std::vector<int> relevant;
    for (size_t i = 0; i < rec.size(); ++i)
        if (brick_distance_sqr(c->init, c->gend, rec[i]) < cut2) relevant.push_back(int(i));
    const int nx = c->cell_dim[0], ny = c->cell_dim[1], nz = c->cell_dim[2];
    auto coord = [&](int d, int k) { return c->init[d] + range[d] * float(k) / float(c->cell_dim[d]); };
    std::vector<int32_t> start(size_t(nx) * ny * nz + 1, 0);
    std::vector<float4> cells;
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                const float lo[3] = {coord(0, x), coord(1, y), coord(2, z)};
                const float hi[3] = {coord(0, x + 1), coord(1, y + 1), coord(2, z + 1)};
                for (int i : relevant)
                    if (brick_distance_sqr(lo, hi, rec[i]) < cut2) cells.push_back(rec[i]);
                start[size_t(x) + size_t(nx) * (y + size_t(ny) * z) + 1] = int32_t(cells.size());
            }
    // far-away atoms after the last cell: the refinement loads one batch of
    // atoms past a cell's list without a bound check (beyond the cutoff of
    // any point, and masked by the list bound anyway)
    cells.insert(cells.end(), kCellPad, make_float4(1e30f, 1e30f, 1e30f, 0.f));
    int rc = grow(c, c->d_cells, c->cells_cap, cells.size());
    if (rc) return rc;
    rc = grow(c, c->d_cell_start, c->cstart_cap, start.size());
    if (rc) return rc;
    UD_HIP(c, hipMemcpyAsync(c->d_cells, cells.data(), sizeof(float4) * cells.size(), hipMemcpyHostToDevice, c->stream));
    UD_HIP(c, hipMemcpyAsync(c->d_cell_start, start.data(), sizeof(int32_t) * start.size(), hipMemcpyHostToDevice, c->stream));
    UD_HIP(c, hipStreamSynchronize(c->stream));
    return UD_OK;
}

static RecDev rec_dev(const ud_ctx* c) {
    RecDev r;
    r.atoms = c->d_cells;
    r.cell_start = c->d_cell_start;
    for (int i = 0; i < 3; ++i) {
        r.dim[i] = c->cell_dim[i];
        r.init[i] = c->init[i];
        r.end[i] = c->gend[i];
        r.range[i] = c->gend[i] - c->init[i];
    }
    return r;
}

static GridBuildArgs build_args(const ud_ctx* c, int n_types, const int32_t* types) {
    GridBuildArgs a;
    std::memset(&a, 0, sizeof(a));
    for (int i = 0; i < 3; ++i) {
        a.init[i] = c->init[i];
        a.factor_inv[i] = c->factor_inv[i];
        a.n[i] = c->n[i];
    }
    a.n_points = c->n_points;
    a.n_maps = n_types;
    for (int m = 0; m < n_types; ++m) a.types[m] = types[m];
    return a;
}

// AD4.2 terms alone (vdW, H-bond, electrostatic, desolvation; ad4cache.cpp
// and potentials.h:350-480): the type-pair tables, the affinity /
// electrostatic / desolvation maps from the receptor of the last
// ud_compute_maps_ad4, and the electrostatic / desolvation radial tables,
// each at unit weight.  The builders are linear in the weights, so these are
// the derivatives of the weighted maps and tables.
static int ensure_terms_ad4(ud_ctx* c) {
    if (!c->have_rec_ad4 || c->n_maps < 2)
        return fail(c, UD_ERR_STATE, "term-weight gradients need maps built by ud_compute_maps_ad4");
    const int nt = 4;
    const size_t tab = size_t(kTypePairs) * kTableStride;
    const size_t mp = size_t(c->n_maps) * c->n_points;
    int rc = grow(c, c->d_term_tab, c->term_tab_cap, nt * tab);
    if (rc) return rc;
    rc = grow(c, c->d_term_maps, c->term_maps_cap, nt * mp);
    if (rc) return rc;
    rc = grow(c, c->d_term_ad4, c->term_ad4_cap, size_t(kAd4TableN));
    if (rc) return rc;
    float* d_fast = nullptr;
    if (hipMalloc(&d_fast, sizeof(float) * tab) != hipSuccess) return fail(c, UD_ERR_NOMEM, "table allocation failed");
    // the affinity maps in slot order, then the electrostatic and desolvation
    // maps, as ud_compute_maps_ad4 stored them
    const GridBuildArgs a = build_args(c, c->n_maps - 2, c->map_types);
    hipError_t e = hipSuccess;
    for (int k = 0; k < nt && e == hipSuccess; ++k) {
        float w1[7] = {0, 0, 0, 0, 0, 0, 0};
        w1[k] = 1.f;
        e = launch_tables(UD_SF_AD4, w1, c->d_term_tab + k * tab, d_fast, c->stream);
        if (e == hipSuccess)
            e = launch_grid_build_ad4(c->d_rec_ad4, c->d_rec_q, c->n_rec_ad4, a, w1, c->d_term_maps + k * mp, c->stream);
    }
    std::vector<float4> deriv;
    std::vector<float2> fast;
    ad4_radial_tables(1.f, 1.f, deriv, fast);
    if (e == hipSuccess)
        e = hipMemcpyAsync(c->d_term_ad4, deriv.data(), sizeof(float4) * kAd4TableN, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(d_fast);
    if (e != hipSuccess) return fail(c, UD_ERR_HIP, hipGetErrorString(e));
    c->terms_ok = true;
    return UD_OK;
}

// Maps and tables of each scoring term alone (its weight 1, the others 0),
// from the receptor of the last ud_compute_maps: the inputs of the
// term-weight gradients.  Built once per scoring / box / map set.
static int ensure_terms(ud_ctx* c) {
    if (c->terms_ok) return UD_OK;
    if (sf_ad4(c)) return ensure_terms_ad4(c);
    if (!c->have_rec || c->n_maps == 0)
        return fail(c, UD_ERR_STATE, "term-weight gradients need maps built by ud_compute_maps");
    const int nt = (c->sf == UD_SF_VINA) ? 6 : 5;
    const size_t tab = size_t(kTypePairs) * kTableStride;
    const size_t mp = size_t(c->n_maps) * c->n_points;
    int rc = grow(c, c->d_term_tab, c->term_tab_cap, nt * tab);
    if (rc) return rc;
    rc = grow(c, c->d_term_maps, c->term_maps_cap, nt * mp);
    if (rc) return rc;
    float* d_fast = nullptr;
    if (hipMalloc(&d_fast, sizeof(float) * tab) != hipSuccess) return fail(c, UD_ERR_NOMEM, "table allocation failed");
    const GridBuildArgs a = build_args(c, c->n_maps, c->map_types);
    hipError_t e = hipSuccess;
    for (int k = 0; k < nt && e == hipSuccess; ++k) {
        float w1[7] = {0, 0, 0, 0, 0, 0, 0};
        w1[k] = 1.f;
        e = launch_tables(c->sf, w1, c->d_term_tab + k * tab, d_fast, c->stream);
        if (e == hipSuccess) e = launch_grid_build(c->d_rec, c->n_rec, a, d_fast, c->d_term_maps + k * mp, c->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(d_fast);
    if (e != hipSuccess) return fail(c, UD_ERR_HIP, hipGetErrorString(e));
    c->terms_ok = true;
    return UD_OK;
}

// extra: the AD4.2 electrostatic and desolvation maps may (and, for a full
// AD4.2 map set, must) appear in the list
static int check_types(ud_ctx* c, int n_types, const int32_t* types, bool ad4_extra) {
    if (n_types <= 0 || n_types > UD_MAX_MAPS || !types) return fail(c, UD_ERR_INVALID, "bad map type list");
    bool have_e = false, have_d = false;
    for (int m = 0; m < n_types; ++m) {
        const int t = types[m];
        if (sf_ad4(c)) {
            if (ad4_extra && t == kAd4MapElec) have_e = true;
            else if (ad4_extra && t == kAd4MapDsolv) have_d = true;
            else if (t < 0 || t >= kAd4Types || ad4_grid_type(t) != t)
                return fail(c, UD_ERR_INVALID, "map type must be a plain AutoDock type");
        } else if (t < 0 || t >= 36 || grid_type_of(t) != t) {
            return fail(c, UD_ERR_INVALID, "map type must be a plain XS type");
        }
        for (int k = 0; k < m; ++k)
            if (types[k] == t) return fail(c, UD_ERR_INVALID, "duplicate map type");
    }
    if (ad4_extra && !(have_e && have_d))
        return fail(c, UD_ERR_INVALID, "AD4.2 maps need the electrostatic and desolvation maps");
    return UD_OK;
}

int ud_compute_maps(ud_ctx* c, int n_rec, const float* rec_xyz, const int32_t* rec_xs, int n_types,
                    const int32_t* types, int precision) {
    if (!c) return UD_ERR_INVALID;
    if (!c->box_set) return fail(c, UD_ERR_STATE, "box not set");
    if (c->sf < 0) return fail(c, UD_ERR_STATE, "scoring function not set");
    if (n_rec < 0 || (n_rec > 0 && (!rec_xyz || !rec_xs))) return fail(c, UD_ERR_INVALID, "bad receptor arrays");
    if (precision != UD_GRID_F32 && precision != UD_GRID_F16) return fail(c, UD_ERR_INVALID, "bad precision");
    if (sf_ad4(c)) return fail(c, UD_ERR_STATE, "AD4.2 maps are built by ud_compute_maps_ad4");
    int rc = check_types(c, n_types, types, false);
    if (rc) return rc;
    UD_HIP(c, hipSetDevice(c->device));
    std::vector<float4> rec;
    rec.reserve(n_rec);
    for (int i = 0; i < n_rec; ++i) {
        if (rec_xs[i] < 0 || rec_xs[i] >= 36) continue;  // untyped atoms do not contribute (cache.cpp:527)
        int32_t t = rec_xs[i];
        float tf;
        std::memcpy(&tf, &t, 4);
        rec.push_back(make_float4(rec_xyz[3 * i], rec_xyz[3 * i + 1], rec_xyz[3 * i + 2], tf));
    }
    c->have_rec = false;
    rc = grow(c, c->d_rec, c->rec_cap, std::max<size_t>(rec.size(), 1));
    if (rc) return rc;
    if (!rec.empty())
        UD_HIP(c, hipMemcpyAsync(c->d_rec, rec.data(), sizeof(float4) * rec.size(), hipMemcpyHostToDevice, c->stream));
    c->n_rec = int(rec.size());
    rc = build_cells(c, rec);
    if (rc) return rc;
    float* d_f32 = nullptr;
    if (hipMalloc(&d_f32, sizeof(float) * size_t(n_types) * c->n_points) != hipSuccess)
        return fail(c, UD_ERR_NOMEM, "map allocation failed");
    const GridBuildArgs a = build_args(c, n_types, types);
    UD_HIP(c, hipEventRecord(c->ev0, c->stream));
    UD_HIP(c, launch_grid_build(c->d_rec, c->n_rec, a, c->d_table_fast, d_f32, c->stream));
    UD_HIP(c, hipEventRecord(c->ev1, c->stream));
    UD_HIP(c, hipStreamSynchronize(c->stream));
    float ms = 0;
    (void)hipEventElapsedTime(&ms, c->ev0, c->ev1);
    c->last_ms = ms;
    c->last_launches = 1;
    c->have_rec_ad4 = false;
    rc = store_maps(c, d_f32, n_types, types, precision);
    if (rc == UD_OK) c->have_rec = true;
    return rc;
}

int ud_set_maps(ud_ctx* c, int n_types, const int32_t* types, const float* values, int precision) {
    if (!c) return UD_ERR_INVALID;
    if (!c->box_set) return fail(c, UD_ERR_STATE, "box not set");
    if (!values) return fail(c, UD_ERR_INVALID, "null map values");
    if (precision != UD_GRID_F32 && precision != UD_GRID_F16) return fail(c, UD_ERR_INVALID, "bad precision");
    if (c->sf < 0) return fail(c, UD_ERR_STATE, "scoring function not set");
    int rc = check_types(c, n_types, types, sf_ad4(c));
    if (rc) return rc;
    UD_HIP(c, hipSetDevice(c->device));
    float* d_f32 = nullptr;
    if (hipMalloc(&d_f32, sizeof(float) * size_t(n_types) * c->n_points) != hipSuccess)
        return fail(c, UD_ERR_NOMEM, "map allocation failed");
    UD_HIP(c, hipMemcpy(d_f32, values, sizeof(float) * size_t(n_types) * c->n_points, hipMemcpyHostToDevice));
    c->have_rec = false;
    c->have_rec_ad4 = false;
    return store_maps(c, d_f32, n_types, types, precision);
}

int ud_compute_maps_ad4(ud_ctx* c, int n_rec, const float* rec_xyz, const int32_t* rec_ad, const float* rec_charge,
                        int n_types, const int32_t* types, int precision) {
    if (!c) return UD_ERR_INVALID;
    if (!c->box_set) return fail(c, UD_ERR_STATE, "box not set");
    if (!sf_ad4(c)) return fail(c, UD_ERR_STATE, "ud_compute_maps_ad4 needs the AD4.2 scoring function");
    if (n_rec < 0 || (n_rec > 0 && (!rec_xyz || !rec_ad || !rec_charge)))
        return fail(c, UD_ERR_INVALID, "bad receptor arrays");
    if (precision != UD_GRID_F32 && precision != UD_GRID_F16) return fail(c, UD_ERR_INVALID, "bad precision");
    int rc = check_types(c, n_types, types, false);
    if (rc) return rc;
    if (n_types + 2 > UD_MAX_MAPS) return fail(c, UD_ERR_CAPACITY, "more than UD_MAX_MAPS maps");
    UD_HIP(c, hipSetDevice(c->device));
    std::vector<float4> rec;
    std::vector<float> q;
    for (int i = 0; i < n_rec; ++i) {
        const int32_t t = rec_ad[i];
        if (t < 0 || t >= kAd4Types) return fail(c, UD_ERR_INVALID, "receptor atom type out of range");
        float tf;
        std::memcpy(&tf, &t, 4);
        rec.push_back(make_float4(rec_xyz[3 * i], rec_xyz[3 * i + 1], rec_xyz[3 * i + 2], tf));
        q.push_back(rec_charge[i]);
    }
    c->have_rec = false;
    c->have_rec_ad4 = false;
    const size_t nr = std::max<size_t>(rec.size(), 1);
    rc = grow(c, c->d_rec_ad4, c->rec_ad4_cap, nr);
    if (rc) return rc;
    rc = grow(c, c->d_rec_q, c->rec_q_cap, nr);
    if (rc) return rc;
    float* d_f32 = nullptr;
    if (hipMalloc(&d_f32, sizeof(float) * size_t(n_types + 2) * c->n_points) != hipSuccess)
        return fail(c, UD_ERR_NOMEM, "map allocation failed");
    if (!rec.empty()) {
        UD_HIP(c, hipMemcpyAsync(c->d_rec_ad4, rec.data(), sizeof(float4) * rec.size(), hipMemcpyHostToDevice,
                                 c->stream));
        UD_HIP(c, hipMemcpyAsync(c->d_rec_q, q.data(), sizeof(float) * q.size(), hipMemcpyHostToDevice, c->stream));
    }
    c->n_rec_ad4 = int(rec.size());
    float4* d_r = c->d_rec_ad4;
    float* d_q = c->d_rec_q;
    const GridBuildArgs a = build_args(c, n_types, types);
    UD_HIP(c, hipEventRecord(c->ev0, c->stream));
    UD_HIP(c, launch_grid_build_ad4(d_r, d_q, int(rec.size()), a, c->w, d_f32, c->stream));
    UD_HIP(c, hipEventRecord(c->ev1, c->stream));
    UD_HIP(c, hipStreamSynchronize(c->stream));
    float ms = 0;
    (void)hipEventElapsedTime(&ms, c->ev0, c->ev1);
    c->last_ms = ms;
    c->last_launches = 1;
    std::vector<int32_t> all(types, types + n_types);
    all.push_back(kAd4MapElec);
    all.push_back(kAd4MapDsolv);
    rc = store_maps(c, d_f32, n_types + 2, all.data(), precision);
    if (rc == UD_OK) c->have_rec_ad4 = true;
    return rc;
}

int ud_get_map(ud_ctx* c, int xs_type, float* out) {
    if (!c || !out) return UD_ERR_INVALID;
    if (xs_type < 0 || xs_type >= 40 || c->slot_of_type[xs_type] < 0) return fail(c, UD_ERR_INVALID, "no map for type");
    const int m = c->slot_of_type[xs_type];
    UD_HIP(c, hipSetDevice(c->device));
    if (c->precision == UD_GRID_F32) {
        UD_HIP(c, hipMemcpy(out, static_cast<float*>(c->d_maps) + size_t(m) * c->n_points, sizeof(float) * c->n_points,
                            hipMemcpyDeviceToHost));
    } else {
        std::vector<uint16_t> h(c->n_points);
        UD_HIP(c, hipMemcpy(h.data(), static_cast<uint16_t*>(c->d_maps) + size_t(m) * c->n_points, 2 * size_t(c->n_points),
                            hipMemcpyDeviceToHost));
        for (int i = 0; i < c->n_points; ++i) {
            // IEEE binary16 -> float
            const uint16_t b = h[i];
            const uint32_t sign = (b >> 15) & 1, ex = (b >> 10) & 31, man = b & 1023;
            float v;
            if (ex == 0) v = std::ldexp(float(man), -24);
            else if (ex == 31) v = man ? NAN : INFINITY;
            else v = std::ldexp(float(man | 1024), int(ex) - 25);
            out[i] = (sign ? -v : v) * c->scale[m];
        }
    }
    return UD_OK;
}

int ud_set_ligands(ud_ctx* c, const ud_ligand_batch* b) {
    if (!c || !b) return UD_ERR_INVALID;
    if (b->n_ligands <= 0) return fail(c, UD_ERR_INVALID, "empty ligand batch");
    if (!b->atom_offset || !b->node_offset || !b->pair_offset || !b->atom_local || !b->atom_xs || !b->atom_heavy ||
        !b->node_parent || !b->node_abegin || !b->node_aend || !b->node_origin || !b->node_axis || !b->num_tors)
        return fail(c, UD_ERR_INVALID, "null array in ligand batch");
    // AD4.2 scoring types atoms by their AutoDock type and needs charges and TORSDOF
    const bool ad = sf_ad4(c);
    if (ad && (!b->atom_ad || !b->atom_charge || !b->torsdof))
        return fail(c, UD_ERR_INVALID, "AD4.2 scoring needs atom_ad, atom_charge and torsdof");
    UD_HIP(c, hipSetDevice(c->device));
    const int L = b->n_ligands;
    std::vector<float> charge;
    std::vector<float2> coef;
    std::vector<int> elem;
    const Ad4Params& A4 = ad4_params();
    std::vector<LigHdr> hdr(L);
    std::vector<AtomRec> atoms;
    std::vector<NodeRec> nodes;
    std::vector<PairRec> pairs;
    std::vector<uint2> pofs;
    std::vector<PartnerRec> partners;
    std::vector<AtomPartners> apar;
    std::vector<int> nat(L), nnd(L);
    std::vector<std::vector<int>> xs(L);
    for (int l = 0; l < L; ++l) {
        const int a0 = b->atom_offset[l], a1 = b->atom_offset[l + 1];
        const int k0 = b->node_offset[l], k1 = b->node_offset[l + 1];
        const int p0 = b->pair_offset[l], p1 = b->pair_offset[l + 1];
        const int na = a1 - a0, nn = k1 - k0, np = p1 - p0;
        const std::string tag = "ligand " + std::to_string(l) + ": ";
        if (na <= 0 || nn <= 0 || np < 0) return fail(c, UD_ERR_INVALID, tag + "empty or negative ranges");
        if (a0 != int(atoms.size()) || k0 != int(nodes.size()) || p0 != int(pairs.size()))
            return fail(c, UD_ERR_INVALID, tag + "offsets must start at 0 and be contiguous");
        if (na > UD_MAX_ATOMS) return fail(c, UD_ERR_CAPACITY, tag + "more than UD_MAX_ATOMS atoms");
        if (nn - 1 > UD_MAX_TORSIONS) return fail(c, UD_ERR_CAPACITY, tag + "more than UD_MAX_TORSIONS torsions");
        if (np > UD_MAX_PAIRS) return fail(c, UD_ERR_CAPACITY, tag + "more than UD_MAX_PAIRS pairs");
        if (np > 0 && !b->pairs) return fail(c, UD_ERR_INVALID, tag + "null pair array");
        // nodes: pre-order, contiguous atom ranges covering [0, na)
        std::vector<int> depth(nn, 0), sub_end(nn, 0), node_of(na, -1);
        int expect = 0;
        for (int k = 0; k < nn; ++k) {
            const int par = b->node_parent[k0 + k];
            const int ab = b->node_abegin[k0 + k], ae = b->node_aend[k0 + k];
            if (k == 0 ? par != -1 : (par < 0 || par >= k)) return fail(c, UD_ERR_INVALID, tag + "nodes must be in pre-order with the root first");
            if (ab != expect || ae <= ab || ae > na) return fail(c, UD_ERR_INVALID, tag + "node atom ranges must be contiguous and non-empty");
            expect = ae;
            depth[k] = (k == 0) ? 0 : depth[par] + 1;
            for (int i = ab; i < ae; ++i) node_of[i] = k;
        }
        if (expect != na) return fail(c, UD_ERR_INVALID, tag + "node atom ranges do not cover all atoms");
        // subtree atom end: pre-order subtrees are contiguous
        for (int k = nn - 1; k >= 0; --k) {
            int e = b->node_aend[k0 + k];
            for (int j = k + 1; j < nn; ++j) {
                // descendants of k are the following nodes until depth returns to <= depth[k]
                if (depth[j] <= depth[k]) break;
                e = std::max(e, b->node_aend[k0 + j]);
            }
            sub_end[k] = e;
        }
        int maxd = 0;
        for (int k = 0; k < nn; ++k) maxd = std::max(maxd, depth[k]);
        // check pre-order consistency: parent of j must lie on the path
        for (int k = 1; k < nn; ++k) {
            const int par = b->node_parent[k0 + k];
            if (depth[k] != depth[par] + 1) return fail(c, UD_ERR_INVALID, tag + "inconsistent tree depth");
        }
        for (int k = 0; k < nn; ++k) {
            NodeRec r;
            const int par = b->node_parent[k0 + k];
            for (int d = 0; d < 3; ++d) {
                // segment c'tor (tree.h:123-131): relative origin/axis at the identity parent orientation
                r.ro[d] = (k == 0) ? 0.f : b->node_origin[3 * (k0 + k) + d] - b->node_origin[3 * (k0 + par) + d];
                r.ax[d] = (k == 0) ? 0.f : b->node_axis[3 * (k0 + k) + d];
            }
            if (k > 0) {
                const float nrm = std::sqrt(r.ax[0] * r.ax[0] + r.ax[1] * r.ax[1] + r.ax[2] * r.ax[2]);
                if (!(std::fabs(nrm - 1.f) < 0.005f)) return fail(c, UD_ERR_INVALID, tag + "node axis is not a unit vector");
            }
            r.link = node_link(uint32_t(k == 0 ? 0 : par), uint32_t(depth[k]));
            r.packed = node_packed(uint32_t(b->node_abegin[k0 + k]), uint32_t(b->node_aend[k0 + k]), uint32_t(sub_end[k]));
            nodes.push_back(r);
        }
        int n_heavy = 0;
        for (int i = 0; i < na; ++i) {
            const int t = ad ? b->atom_ad[a0 + i] : b->atom_xs[a0 + i];
            if (ad ? (t < 0 || t >= kAd4Types) : (t < -1 || t >= 36))
                return fail(c, UD_ERR_INVALID, tag + "atom type out of range");
            xs[l].push_back(t);
            elem.push_back(ad ? ad4_element(t) : xs_element(t));
            charge.push_back(ad ? b->atom_charge[a0 + i] : 0.f);
            AtomRec a;
            a.x = b->atom_local[3 * (a0 + i)];
            a.y = b->atom_local[3 * (a0 + i) + 1];
            a.z = b->atom_local[3 * (a0 + i) + 2];
            const int heavy = b->atom_heavy[a0 + i] ? 1 : 0;
            n_heavy += heavy;
            a.meta = uint32_t(node_of[i]) | (0xFFu << 8) | (uint32_t(heavy) << 16) | (uint32_t(t < 0 ? 0xFF : t) << 24);
            atoms.push_back(a);
        }
        for (int q = 0; q < np; ++q) {
            const int pa = b->pairs[2 * (p0 + q)], pb = b->pairs[2 * (p0 + q) + 1];
            if (pa < 0 || pb < 0 || pa >= na || pb >= na || pa == pb) return fail(c, UD_ERR_INVALID, tag + "pair index out of range");
            const int ta = xs[l][pa], tb = xs[l][pb];
            if (ta < 0 || tb < 0) return fail(c, UD_ERR_INVALID, tag + "pair with an untyped atom");
            pairs.push_back(pair_rec(uint32_t(pa), uint32_t(pb), uint32_t(tri_index(ta, tb))));
            pofs.push_back(pair_ofs(uint32_t(pa), uint32_t(pb), uint32_t(tri_index(ta, tb)), uint32_t(kTableStride)));
            if (ad) {
                // ad4_electrostatic's charge product and ad4_solvation's
                // coefficient of the Gaussian (potentials.h:350-400)
                const float qa = b->atom_charge[a0 + pa], qb = b->atom_charge[a0 + pb];
                const float cs = (A4.solv[ta] + 0.01097f * std::fabs(qa)) * A4.vol[tb] +
                                 (A4.solv[tb] + 0.01097f * std::fabs(qb)) * A4.vol[ta];
                coef.push_back(make_float2(qa * qb, cs));
            }
        }
        // every pair under both of its atoms, in pair-list order
        std::vector<std::vector<PartnerRec>> part(na);
        for (int q = 0; q < np; ++q) {
            const int pa = b->pairs[2 * (p0 + q)], pb = b->pairs[2 * (p0 + q) + 1];
            const uint32_t slot = uint32_t(tri_index(xs[l][pa], xs[l][pb]));
            part[pa].push_back(partner_rec(uint32_t(pb), uint32_t(q), slot));
            part[pb].push_back(partner_rec(uint32_t(pa), uint32_t(q), slot));
        }
        const size_t pbeg = partners.size();
        // the search kernel loads record batches as aligned 16-byte pairs
        if (pbeg % kPartnerPad) return fail(c, UD_ERR_INVALID, "internal: partner list misaligned");
        for (int i = 0; i < na; ++i) {
            apar.push_back(uint32_t(partners.size() - pbeg) | (uint32_t(part[i].size()) << 20));
            partners.insert(partners.end(), part[i].begin(), part[i].end());
            while ((partners.size() - pbeg) % kPartnerPad) partners.push_back(partner_rec(uint32_t(i), uint32_t(np), 0));
        }
        LigHdr& h = hdr[l];
        std::memset(&h, 0, sizeof(h));
        h.atom_begin = a0;
        h.n_atoms = na;
        h.node_begin = k0;
        h.n_nodes = nn;
        h.pair_begin = p0;
        h.n_pairs = np;
        h.max_depth = maxd;
        h.local_steps = (25 + na) / 3;  // vina.h:197
        h.num_tors = ad ? b->torsdof[l] : b->num_tors[l];  // AD4.2: TORSDOF (conf_independent.cpp:183-187)
        h.n_heavy = n_heavy;
        h.partner_begin = int32_t(pbeg);
        nat[l] = na;
        nnd[l] = nn;
    }
    // commit
    for (void* p : {static_cast<void*>(c->d_hdr), static_cast<void*>(c->d_nodes), static_cast<void*>(c->d_pairs),
                    static_cast<void*>(c->d_pair_ofs), static_cast<void*>(c->d_partners), static_cast<void*>(c->d_apar)})
        if (p) (void)hipFree(p);
    c->d_hdr = nullptr;
    c->d_nodes = nullptr;
    c->d_pairs = nullptr;
    c->d_pair_ofs = nullptr;
    c->d_partners = nullptr;
    c->d_apar = nullptr;
    size_t cap = 0;
    int rc = grow(c, c->d_hdr, cap, hdr.size());
    if (rc) return rc;
    cap = 0;
    rc = grow(c, c->d_nodes, cap, nodes.size());
    if (rc) return rc;
    cap = 0;
    rc = grow(c, c->d_pairs, cap, std::max<size_t>(pairs.size(), 1));
    if (rc) return rc;
    cap = 0;
    pofs.insert(pofs.end(), kPairOfsPad, make_uint2(0u, 0u));
    rc = grow(c, c->d_pair_ofs, cap, pofs.size());
    if (rc) return rc;
    cap = 0;
    // the records, then the same records pre-decoded
    const size_t n_part = partners.size();
    if (2 * n_part > size_t(INT32_MAX)) return fail(c, UD_ERR_CAPACITY, "too many intramolecular pairs in one ligand set");
    for (size_t k = 0; k < n_part; ++k) partners.push_back(partner_ofs(partners[k]));
    // zero records after the pre-decoded lists: the second pair pass loads
    // one record batch past an atom's list without a bound check
    partners.insert(partners.end(), 2 * kPartnerPad, 0u);
    for (LigHdr& h : hdr) h.partner_ofs_begin = int32_t(n_part) + h.partner_begin;
    rc = grow(c, c->d_partners, cap, std::max<size_t>(partners.size(), 1));
    if (rc) return rc;
    cap = 0;
    rc = grow(c, c->d_apar, cap, apar.size());
    if (rc) return rc;
    if (!partners.empty())
        UD_HIP(c, hipMemcpy(c->d_partners, partners.data(), partners.size() * sizeof(PartnerRec), hipMemcpyHostToDevice));
    UD_HIP(c, hipMemcpy(c->d_apar, apar.data(), apar.size() * sizeof(AtomPartners), hipMemcpyHostToDevice));
    UD_HIP(c, hipMemcpy(c->d_hdr, hdr.data(), hdr.size() * sizeof(LigHdr), hipMemcpyHostToDevice));
    UD_HIP(c, hipMemcpy(c->d_nodes, nodes.data(), nodes.size() * sizeof(NodeRec), hipMemcpyHostToDevice));
    if (!pairs.empty())
        UD_HIP(c, hipMemcpy(c->d_pairs, pairs.data(), pairs.size() * sizeof(PairRec), hipMemcpyHostToDevice));
    UD_HIP(c, hipMemcpy(c->d_pair_ofs, pofs.data(), pofs.size() * sizeof(uint2), hipMemcpyHostToDevice));
    for (void* p : {static_cast<void*>(c->d_charge), static_cast<void*>(c->d_pair_coef)})
        if (p) (void)hipFree(p);
    c->d_charge = nullptr;
    c->d_pair_coef = nullptr;
    if (ad) {
        cap = 0;
        rc = grow(c, c->d_charge, cap, charge.size());
        if (rc) return rc;
        cap = 0;
        rc = grow(c, c->d_pair_coef, cap, std::max<size_t>(coef.size(), 1));
        if (rc) return rc;
        UD_HIP(c, hipMemcpy(c->d_charge, charge.data(), charge.size() * sizeof(float), hipMemcpyHostToDevice));
        if (!coef.empty())
            UD_HIP(c, hipMemcpy(c->d_pair_coef, coef.data(), coef.size() * sizeof(float2), hipMemcpyHostToDevice));
    }
    c->lig_ad4 = ad;
    c->h_elem = elem;
    rc = grow(c, c->d_elem, c->elem_cap, std::max<size_t>(elem.size(), 1));
    if (rc) return rc;
    UD_HIP(c, hipMemcpy(c->d_elem, elem.data(), elem.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    c->hdr = hdr;
    c->n_lig = L;
    c->lig_nat = nat;
    c->lig_nn = nnd;
    c->lig_xs = xs;
    // atoms carry grid slots; they are (re)written before each evaluation
    rc = upload_atoms(c, atoms);
    if (rc) return rc;
    c->h_atoms = atoms;
    c->slots_dirty = true;
    return UD_OK;
}

}  // extern "C"

namespace {
// refresh the grid slots of the uploaded atoms for the current maps
int refresh_slots(ud_ctx* c) {
    int rc = ready_for_eval(c);
    if (rc) return rc;
    if (!c->slots_dirty) return UD_OK;
    std::vector<AtomRec>& atoms = c->h_atoms;
    const size_t total = atoms.size();
    bool changed = false;
    for (auto& a : atoms) {
        const uint32_t t = a.meta >> 24;
        const int gt = (t == 0xFF) ? -1 : map_type_of(c, int(t));
        const uint32_t slot = (gt < 0) ? 0xFFu : uint32_t(c->slot_of_type[gt]);
        const uint32_t meta = (a.meta & ~(0xFFu << 8)) | (slot << 8);
        if (meta != a.meta) { a.meta = meta; changed = true; }
    }
    if (changed) UD_HIP(c, hipMemcpy(c->d_atoms, atoms.data(), total * sizeof(AtomRec), hipMemcpyHostToDevice));
    c->slots_dirty = false;
    return UD_OK;
}

int pose_upload(ud_ctx* c, int n_poses, const float* poses) {
    int rc = grow(c, c->d_poses, c->poses_cap, size_t(n_poses) * UD_POSE_FLOATS);
    if (rc) return rc;
    UD_HIP(c, hipMemcpyAsync(c->d_poses, poses, sizeof(float) * n_poses * UD_POSE_FLOATS, hipMemcpyHostToDevice, c->stream));
    return UD_OK;
}

int work_upload(ud_ctx* c, const std::vector<int32_t>& perm, const std::vector<WorkItem>& work) {
    int rc = grow(c, c->d_perm, c->perm_cap, std::max<size_t>(perm.size(), 1));
    if (rc) return rc;
    rc = grow(c, c->d_work, c->work_cap, std::max<size_t>(work.size(), 1));
    if (rc) return rc;
    if (!perm.empty()) UD_HIP(c, hipMemcpyAsync(c->d_perm, perm.data(), perm.size() * 4, hipMemcpyHostToDevice, c->stream));
    if (!work.empty()) UD_HIP(c, hipMemcpyAsync(c->d_work, work.data(), work.size() * sizeof(WorkItem), hipMemcpyHostToDevice, c->stream));
    return UD_OK;
}

// default lanes per chain by atom capacity (replaced by autotuned values)
// lanes per chain without tuning, from MI355X timings of the search kernel
// over ligand size buckets (32 lanes up to 32 atoms, a full wavefront above)
int default_lanes(int na) { return na <= 32 ? 32 : 64; }

constexpr float kMaxFloat = std::numeric_limits<float>::max();  // max_fl
}  // namespace

extern "C" {

int ud_score_grad_device(ud_ctx* c, int n_poses, const int32_t* pose_ligand, const float* d_poses, const float* v,
                         float* d_energy, float* d_grad) {
    if (!c || !v) return UD_ERR_INVALID;
    int rc = check_pose_args(c, n_poses, pose_ligand);
    if (rc) return rc;
    if (n_poses == 0) return UD_OK;
    UD_HIP(c, hipSetDevice(c->device));
    rc = refresh_slots(c);
    if (rc) return rc;
    LaunchCap cap = caps_all(c);
    const int na = cap.n_atoms;
    const int lanes = fit_lanes(c->forced_lanes ? c->forced_lanes : 16, cap);
    std::vector<int32_t> perm;
    std::vector<WorkItem> work;
    make_pose_work(n_poses, pose_ligand, 64 / lanes, perm, work);
    rc = work_upload(c, perm, work);
    if (rc) return rc;
    UD_HIP(c, hipEventRecord(c->ev0, c->stream));
    UD_HIP(c, launch_score(lanes, grid_format(c), lig_dev(c), grid_dev(c), cap, c->d_work, int(work.size()), c->d_perm,
                           d_poses, v[0], v[1], d_energy, d_grad, c->stream));
    UD_HIP(c, hipEventRecord(c->ev1, c->stream));
    c->last_launches = 1;
    return UD_OK;
}

int ud_score_grad_device_noalloc(ud_ctx* c, int n_poses, const int32_t* pose_ligand, const float* d_poses,
                                 const float* v, float* d_energy, float* d_grad) {
    if (!c || !v) return UD_ERR_INVALID;
    if (n_poses > UD_NOALLOC_MAX_POSES)
        return fail(c, UD_ERR_CAPACITY, "ud_score_grad_device_noalloc takes at most " +
                                            std::to_string(UD_NOALLOC_MAX_POSES) + " poses");
    int rc = check_pose_args(c, n_poses, pose_ligand);
    if (rc) return rc;
    if (n_poses == 0) return UD_OK;
    if (!d_poses || !d_energy || !d_grad) return fail(c, UD_ERR_INVALID, "null device pointer");
    UD_HIP(c, hipSetDevice(c->device));
    unsigned char* scr = static_cast<unsigned char*>(scratch_of(c->device));
    if (!scr) return fail(c, UD_ERR_STATE, "scratch of this device is not resolved");
    rc = refresh_slots(c);
    if (rc) return rc;
    LaunchCap cap = caps_all(c);
    const int lanes = fit_lanes(c->forced_lanes ? c->forced_lanes : 16, cap);
    std::vector<int32_t> perm;
    std::vector<WorkItem> work;
    make_pose_work(n_poses, pose_ligand, 64 / lanes, perm, work);
    // the permutation and the work list overwrite the array's earlier contents
    int32_t* d_perm = reinterpret_cast<int32_t*>(scr);
    WorkItem* d_work = reinterpret_cast<WorkItem*>(scr + size_t(UD_NOALLOC_MAX_POSES) * 4);
    UD_HIP(c, hipMemcpyAsync(d_perm, perm.data(), perm.size() * 4, hipMemcpyHostToDevice, c->stream));
    UD_HIP(c, hipMemcpyAsync(d_work, work.data(), work.size() * sizeof(WorkItem), hipMemcpyHostToDevice, c->stream));
    UD_HIP(c, hipEventRecord(c->ev0, c->stream));
    UD_HIP(c, launch_score(lanes, grid_format(c), lig_dev(c), grid_dev(c), cap, d_work, int(work.size()), d_perm,
                           d_poses, v[0], v[1], d_energy, d_grad, c->stream));
    UD_HIP(c, hipEventRecord(c->ev1, c->stream));
    c->last_launches = 1;
    return UD_OK;
}

int ud_score_grad(ud_ctx* c, int n_poses, const int32_t* pose_ligand, const float* poses, const float* v,
                  float* energy_out, float* grad_out) {
    if (!c || !v || (n_poses > 0 && (!poses || !energy_out))) return UD_ERR_INVALID;
    int rc = check_pose_args(c, n_poses, pose_ligand);
    if (rc || n_poses == 0) return rc;
    UD_HIP(c, hipSetDevice(c->device));
    rc = pose_upload(c, n_poses, poses);
    if (rc) return rc;
    rc = grow(c, c->d_f1, c->f_cap, size_t(n_poses) * UD_GRAD_FLOATS);
    if (rc) return rc;
    rc = grow(c, c->d_f2, c->f2_cap, size_t(n_poses));
    if (rc) return rc;
    float* d_e = c->d_f2;
    float* d_g = c->d_f1;
    rc = ud_score_grad_device(c, n_poses, pose_ligand, c->d_poses, v, d_e, d_g);
    if (rc) return rc;
    UD_HIP(c, hipMemcpyAsync(energy_out, d_e, sizeof(float) * n_poses, hipMemcpyDeviceToHost, c->stream));
    if (grad_out)
        UD_HIP(c, hipMemcpyAsync(grad_out, d_g, sizeof(float) * size_t(n_poses) * UD_GRAD_FLOATS, hipMemcpyDeviceToHost, c->stream));
    UD_HIP(c, hipStreamSynchronize(c->stream));
    float ms = 0;
    (void)hipEventElapsedTime(&ms, c->ev0, c->ev1);
    c->last_ms = ms;
    return UD_OK;
}

int ud_energy_grad_device(ud_ctx* c, int n_poses, const int32_t* pose_ligand, const float* d_poses, const float* v,
                          float* d_energy, float* d_pose_grad, float* d_weight_grad) {
    if (!c || !v) return UD_ERR_INVALID;
    int rc = check_pose_args(c, n_poses, pose_ligand);
    if (rc) return rc;
    if (n_poses == 0) return UD_OK;
    if (!d_poses || !d_energy || !d_pose_grad || !d_weight_grad) return fail(c, UD_ERR_INVALID, "null device pointer");
    UD_HIP(c, hipSetDevice(c->device));
    rc = refresh_slots(c);
    if (rc) return rc;
    rc = ensure_terms(c);
    if (rc) return rc;
    LaunchCap cap = caps_all(c);
    const int lanes = fit_lanes(c->forced_lanes ? c->forced_lanes : 16, cap);
    std::vector<int32_t> perm;
    std::vector<WorkItem> work;
    make_pose_work(n_poses, pose_ligand, 64 / lanes, perm, work);
    rc = work_upload(c, perm, work);
    if (rc) return rc;
    TermsDev td;
    td.maps = c->d_term_maps;
    td.tables = c->d_term_tab;
    td.ad4_unit = c->d_term_ad4;
    td.n_terms = sf_ad4(c) ? 4 : (c->sf == UD_SF_VINA) ? 6 : 5;
    td.n_maps = c->n_maps;
    UD_HIP(c, hipEventRecord(c->ev0, c->stream));
    UD_HIP(c, launch_energy_grad(lanes, grid_format(c), lig_dev(c), grid_dev(c), td, cap, c->d_work, int(work.size()),
                                 c->d_perm, d_poses, v[0], v[1], d_energy, d_pose_grad, d_weight_grad, c->stream));
    UD_HIP(c, hipEventRecord(c->ev1, c->stream));
    c->last_launches = 1;
    return UD_OK;
}

int ud_energy_grad(ud_ctx* c, int n_poses, const int32_t* pose_ligand, const float* poses, const float* v,
                   float* energy_out, float* pose_grad_out, float* weight_grad_out) {
    if (!c || !v || (n_poses > 0 && (!poses || !energy_out || !pose_grad_out || !weight_grad_out)))
        return UD_ERR_INVALID;
    int rc = check_pose_args(c, n_poses, pose_ligand);
    if (rc || n_poses == 0) return rc;
    UD_HIP(c, hipSetDevice(c->device));
    rc = pose_upload(c, n_poses, poses);
    if (rc) return rc;
    rc = grow(c, c->d_f1, c->f_cap, size_t(n_poses) * UD_POSE_FLOATS);
    if (rc) return rc;
    rc = grow(c, c->d_f2, c->f2_cap, size_t(n_poses));
    if (rc) return rc;
    rc = grow(c, c->d_f3, c->f3_cap, size_t(n_poses) * UD_WEIGHT_FLOATS);
    if (rc) return rc;
    rc = ud_energy_grad_device(c, n_poses, pose_ligand, c->d_poses, v, c->d_f2, c->d_f1, c->d_f3);
    if (rc) return rc;
    UD_HIP(c, hipMemcpyAsync(energy_out, c->d_f2, sizeof(float) * n_poses, hipMemcpyDeviceToHost, c->stream));
    UD_HIP(c, hipMemcpyAsync(pose_grad_out, c->d_f1, sizeof(float) * size_t(n_poses) * UD_POSE_FLOATS,
                             hipMemcpyDeviceToHost, c->stream));
    UD_HIP(c, hipMemcpyAsync(weight_grad_out, c->d_f3, sizeof(float) * size_t(n_poses) * UD_WEIGHT_FLOATS,
                             hipMemcpyDeviceToHost, c->stream));
    UD_HIP(c, hipStreamSynchronize(c->stream));
    float ms = 0;
    (void)hipEventElapsedTime(&ms, c->ev0, c->ev1);
    c->last_ms = ms;
    return UD_OK;
}

int ud_optimize(ud_ctx* c, int n_poses, const int32_t* pose_ligand, float* poses, const float* v, int max_steps,
                float* energy_out, int32_t* evals_out) {
    if (!c || !v || (n_poses > 0 && !poses)) return UD_ERR_INVALID;
    int rc = check_pose_args(c, n_poses, pose_ligand);
    if (rc || n_poses == 0) return rc;
    UD_HIP(c, hipSetDevice(c->device));
    rc = refresh_slots(c);
    if (rc) return rc;
    rc = pose_upload(c, n_poses, poses);
    if (rc) return rc;
    rc = grow(c, c->d_f1, c->f_cap, size_t(n_poses) * UD_GRAD_FLOATS);
    if (rc) return rc;
    rc = grow(c, c->d_i1, c->i_cap, size_t(n_poses));
    if (rc) return rc;
    LaunchCap cap = caps_all(c);
    const int na = cap.n_atoms;
    const int lanes = fit_lanes(c->forced_lanes ? c->forced_lanes : default_lanes(na), cap);
    std::vector<int32_t> perm;
    std::vector<WorkItem> work;
    make_pose_work(n_poses, pose_ligand, 64 / lanes, perm, work);
    rc = work_upload(c, perm, work);
    if (rc) return rc;
    UD_HIP(c, hipEventRecord(c->ev0, c->stream));
    UD_HIP(c, launch_optimize(lanes, grid_format(c), lig_dev(c), grid_dev(c), cap, c->d_work, int(work.size()), c->d_perm,
                              c->d_poses, v[0], v[1], max_steps, c->d_f1, c->d_i1, c->stream));
    UD_HIP(c, hipEventRecord(c->ev1, c->stream));
    UD_HIP(c, hipMemcpyAsync(poses, c->d_poses, sizeof(float) * size_t(n_poses) * UD_POSE_FLOATS, hipMemcpyDeviceToHost, c->stream));
    if (energy_out) UD_HIP(c, hipMemcpyAsync(energy_out, c->d_f1, sizeof(float) * n_poses, hipMemcpyDeviceToHost, c->stream));
    if (evals_out) UD_HIP(c, hipMemcpyAsync(evals_out, c->d_i1, sizeof(int32_t) * n_poses, hipMemcpyDeviceToHost, c->stream));
    UD_HIP(c, hipStreamSynchronize(c->stream));
    float ms = 0;
    (void)hipEventElapsedTime(&ms, c->ev0, c->ev1);
    c->last_ms = ms;
    c->last_launches = 1;
    return UD_OK;
}

}  // extern "C"

namespace {

const int kBucketCaps[UD_N_BUCKETS] = {16, 32, 48, 64, 96, 128, 192, 300};
// one chain per lane group unless the caller asks for more: many small
// workgroups balance the end of a launch (and of concurrent bucket launches)
// better than fewer workgroups whose groups refill
const int kChainsPerGroup = 1;

int bucket_of(int n_atoms) {
    int bi = 0;
    while (bi < UD_N_BUCKETS - 1 && kBucketCaps[bi] < n_atoms) ++bi;
    return bi;
}

int check_search_params(ud_ctx* c, const ud_search_params* p) {
    if (p->exhaustiveness < 1) return fail(c, UD_ERR_INVALID, "exhaustiveness must be 1 or greater");
    if (p->global_steps < 0) return fail(c, UD_ERR_INVALID, "global_steps must be >= 0");
    if (p->lanes_per_chain != 0 && p->lanes_per_chain != 8 && p->lanes_per_chain != 16 && p->lanes_per_chain != 32 &&
        p->lanes_per_chain != 64)
        return fail(c, UD_ERR_INVALID, "lanes_per_chain must be 0, 8, 16, 32 or 64");
    if (p->chains_per_group < 0 || p->chains_per_group > 64)
        return fail(c, UD_ERR_INVALID, "chains_per_group must be in 0..64");
    return UD_OK;
}

// Monte Carlo search of the ligands in `subset` (all ligands when null): one
// launch per size bucket, largest bucket first.  Chain results land at
// ligand * exhaustiveness + chain of the context's output buffers, so a
// subset run leaves the other ligands' slots untouched.  lanes_all != 0
// forces the lanes per chain of every bucket.
// lanes_all / pb_all: one lane count and LDS layout for every bucket (the
// autotuner's trials), else each bucket's setting (lanes_all 0, pb_all -2)
int dock_run(ud_ctx* c, const ud_search_params* p, const std::vector<int>* subset, int lanes_all, int pb_all = -2) {
    UD_HIP(c, hipSetDevice(c->device));
    int rc = refresh_slots(c);
    if (rc) return rc;
    const int exh = p->exhaustiveness;
    const size_t nch = size_t(c->n_lig) * exh;
    if (nch > size_t(1) << 31) return fail(c, UD_ERR_CAPACITY, "too many chains");
    if (nch > c->out_cap) {
        for (void* q : {static_cast<void*>(c->d_out_pose), static_cast<void*>(c->d_out_e), static_cast<void*>(c->d_out_evals)})
            if (q) (void)hipFree(q);
        c->d_out_pose = nullptr;
        c->d_out_e = nullptr;
        c->d_out_evals = nullptr;
        if (hipMalloc(&c->d_out_pose, nch * UD_POSE_FLOATS * sizeof(float)) != hipSuccess ||
            hipMalloc(&c->d_out_e, nch * sizeof(float)) != hipSuccess ||
            hipMalloc(&c->d_out_evals, nch * sizeof(long long)) != hipSuccess) {
            c->out_cap = 0;
            return fail(c, UD_ERR_NOMEM, "result allocation failed");
        }
        c->out_cap = nch;
    }
    c->out_n = nch;
    c->last_exh = exh;
    SearchDev sp;
    std::memset(&sp, 0, sizeof(sp));
    sp.exh = exh;
    sp.global_steps = p->global_steps;
    sp.local_steps_override = p->local_steps;
    sp.seed_lo = uint32_t(p->seed);
    sp.seed_hi = uint32_t(p->seed >> 32);
    for (int i = 0; i < 3; ++i) {
        sp.hunt[i] = p->hunt_cap[i];
        sp.authentic[i] = p->authentic_v[i];
        // random_in_box corners: the grid's first and last sample points
        sp.box_lo[i] = c->init[i];
        sp.box_hi[i] = c->init[i] + c->factor_inv[i] * float(c->n[i] - 1);
    }
    sp.temperature = p->temperature;
    sp.amplitude = p->mutation_amplitude;
    sp.ls_f0_zero = p->first_ls_f0_zero ? 1 : 0;
    // size buckets by atom capacity; each bucket is one launch, largest first
    std::vector<std::vector<int>> bl(UD_N_BUCKETS);
    if (subset) {
        for (int l : *subset) bl[p->single_launch ? UD_N_BUCKETS - 1 : bucket_of(c->lig_nat[l])].push_back(l);
    } else {
        for (int l = 0; l < c->n_lig; ++l) bl[p->single_launch ? UD_N_BUCKETS - 1 : bucket_of(c->lig_nat[l])].push_back(l);
    }
    std::vector<std::vector<WorkItem>> works;
    std::vector<int> lanes_v;
    std::vector<LaunchCap> cap_v;
    int nl = 0;
    size_t total_work = 0;
    // each bucket's ligands with at most kNSmall - 6 torsions (the search
    // keeps their inverse Hessian in registers) and those with more are
    // launched separately, so that a few flexible ligands do not move the
    // whole bucket to the larger-vector kernel
    std::vector<std::pair<int, std::vector<int>>> parts;
    for (int bi = UD_N_BUCKETS - 1; bi >= 0; --bi) {
        if (bl[bi].empty()) continue;
        std::vector<int> few, many;
        for (int l : bl[bi]) (c->lig_nn[l] - 1 <= kNSmall - 6 ? few : many).push_back(l);
        if (p->single_launch) {
            parts.emplace_back(bi, bl[bi]);
            continue;
        }
        if (!many.empty()) parts.emplace_back(bi, std::move(many));
        if (!few.empty()) parts.emplace_back(bi, std::move(few));
    }
    for (auto& part : parts) {
        const int bi = part.first;
        const std::vector<int>& members = part.second;
        LaunchCap cap = caps_of(c, &members);
        int lanes = lanes_all ? lanes_all
                              : (p->lanes_per_chain ? p->lanes_per_chain
                                                    : (c->tuned_lanes[bi] ? c->tuned_lanes[bi] : default_lanes(cap.n_atoms)));
        lanes = fit_lanes(lanes, cap, pb_all != -2 ? pb_all : (p->single_launch ? -1 : c->tuned_pb[bi]));
        if (lds_bytes(lanes, cap) > 65536) return fail(c, UD_ERR_CAPACITY, "ligand too large for one workgroup's LDS");
        const int G = 64 / lanes;
        // larger ligands first inside the bucket (tail balance)
        std::vector<int> order = members;
        std::stable_sort(order.begin(), order.end(), [&](int a, int b2) {
            return c->lig_nat[a] * c->lig_nn[a] > c->lig_nat[b2] * c->lig_nn[b2];
        });
        // a work item holds up to chains_per_group chains per lane group; a
        // group that finishes a chain takes the item's next one
        const int per_item = G * (p->chains_per_group > 0 ? p->chains_per_group : kChainsPerGroup);
        std::vector<WorkItem> w;
        for (int l : order)
            for (int c0 = 0; c0 < exh; c0 += per_item)
                w.push_back(WorkItem{l, l * exh + c0, std::min(per_item, exh - c0), 0});
        total_work += w.size();
        ud_launch_info& li = c->linfo[nl++];
        li.atom_cap = kBucketCaps[bi];
        li.lanes = lanes;
        li.n_ligands = int(order.size());
        li.n_workgroups = int(w.size());
        li.pair_buffer = cap.pair_forces;
        li.lds_bytes = int(lds_bytes(lanes, cap));
        li.max_torsions = cap.n_nodes - 1;
        li.ms = 0.0;
        works.push_back(std::move(w));
        lanes_v.push_back(lanes);
        cap_v.push_back(cap);
    }
    rc = grow(c, c->d_work, c->work_cap, total_work);
    if (rc) return rc;
    std::vector<WorkItem> all;
    for (auto& w : works) all.insert(all.end(), w.begin(), w.end());
    UD_HIP(c, hipMemcpyAsync(c->d_work, all.data(), all.size() * sizeof(WorkItem), hipMemcpyHostToDevice, c->stream));
    UD_HIP(c, hipEventRecord(c->ev0, c->stream));
    size_t off = 0;
    for (size_t k = 0; k < works.size(); ++k) {
        hipStream_t bs = c->bstream[k];
        UD_HIP(c, hipStreamWaitEvent(bs, c->ev0, 0));
        UD_HIP(c, hipEventRecord(c->lev0[k], bs));
        UD_HIP(c, launch_dock(lanes_v[k], grid_format(c), lig_dev(c), grid_dev(c), cap_v[k], sp, c->d_work + off,
                              int(works[k].size()), c->d_out_pose, c->d_out_e, c->d_out_evals, bs));
        UD_HIP(c, hipEventRecord(c->lev[k], bs));
        off += works[k].size();
    }
    for (size_t k = 0; k < works.size(); ++k) UD_HIP(c, hipStreamWaitEvent(c->stream, c->lev[k], 0));
    UD_HIP(c, hipEventRecord(c->ev1, c->stream));
    c->last_launches = nl;
    return UD_OK;
}

// Resolve the timing events of the last dock (after a stream synchronize).
void dock_times(ud_ctx* c) {
    float ms = 0;
    (void)hipEventElapsedTime(&ms, c->ev0, c->ev1);
    c->last_ms = ms;
    for (int k = 0; k < c->last_launches; ++k) {
        float m = 0;
        (void)hipEventElapsedTime(&m, c->lev0[k], c->lev[k]);
        c->linfo[k].ms = m;
    }
}

}  // namespace

extern "C" {

int ud_dock(ud_ctx* c, const ud_search_params* p, float* best_poses, float* best_e, int64_t* evals) {
    if (!c || !p) return UD_ERR_INVALID;
    int rc = check_search_params(c, p);
    if (rc) return rc;
    rc = dock_run(c, p, nullptr, 0);
    if (rc) return rc;
    if (best_poses || best_e || evals) return ud_fetch_dock_results(c, best_poses, best_e, evals);
    return UD_OK;
}

int ud_fetch_dock_results(ud_ctx* c, float* best_poses, float* best_e, int64_t* evals) {
    if (!c) return UD_ERR_INVALID;
    if (c->out_n == 0) return fail(c, UD_ERR_STATE, "no docking results");
    UD_HIP(c, hipSetDevice(c->device));
    const size_t n = c->out_n;
    if (best_poses) UD_HIP(c, hipMemcpyAsync(best_poses, c->d_out_pose, n * UD_POSE_FLOATS * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    if (best_e) UD_HIP(c, hipMemcpyAsync(best_e, c->d_out_e, n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    if (evals) UD_HIP(c, hipMemcpyAsync(evals, c->d_out_evals, n * sizeof(long long), hipMemcpyDeviceToHost, c->stream));
    UD_HIP(c, hipStreamSynchronize(c->stream));
    dock_times(c);
    return UD_OK;
}

int ud_last_launches(const ud_ctx* c, int max_launches, ud_launch_info* out, int32_t* n_out) {
    if (!c || (max_launches > 0 && !out)) return UD_ERR_INVALID;
    const int n = std::min(max_launches, c->last_launches);
    for (int k = 0; k < n; ++k) out[k] = c->linfo[k];
    if (n_out) *n_out = c->last_launches;
    return UD_OK;
}

int ud_set_bucket_lanes(ud_ctx* c, const int32_t* lanes) {
    if (!c || !lanes) return UD_ERR_INVALID;
    for (int b = 0; b < UD_N_BUCKETS; ++b)
        if (lanes[b] != 0 && lanes[b] != 8 && lanes[b] != 16 && lanes[b] != 32 && lanes[b] != 64)
            return fail(c, UD_ERR_INVALID, "bucket lanes must be 0, 8, 16, 32 or 64");
    for (int b = 0; b < UD_N_BUCKETS; ++b) c->tuned_lanes[b] = lanes[b];
    return UD_OK;
}

int ud_set_bucket_pair_buffer(ud_ctx* c, const int32_t* pair_buffer) {
    if (!c || !pair_buffer) return UD_ERR_INVALID;
    for (int b = 0; b < UD_N_BUCKETS; ++b)
        if (pair_buffer[b] < -1 || pair_buffer[b] > 1) return fail(c, UD_ERR_INVALID, "pair buffer must be -1, 0 or 1");
    for (int b = 0; b < UD_N_BUCKETS; ++b) c->tuned_pb[b] = pair_buffer[b];
    return UD_OK;
}

int ud_get_bucket_layout(const ud_ctx* c, int32_t* lanes_out, int32_t* pair_buffer_out) {
    if (!c) return UD_ERR_INVALID;
    for (int b = 0; b < UD_N_BUCKETS; ++b) {
        if (lanes_out) lanes_out[b] = c->tuned_lanes[b];
        if (pair_buffer_out) pair_buffer_out[b] = c->tuned_pb[b];
    }
    return UD_OK;
}

int ud_autotune(ud_ctx* c, const ud_search_params* p, int sample_ligands, int32_t* lanes_out) {
    if (!c || !p || sample_ligands < 1) return UD_ERR_INVALID;
    int rc = check_search_params(c, p);
    if (rc) return rc;
    rc = ready_for_eval(c);
    if (rc) return rc;
    const int exh = p->exhaustiveness;
    std::vector<std::vector<int>> bl(UD_N_BUCKETS);
    for (int l = 0; l < c->n_lig; ++l) bl[bucket_of(c->lig_nat[l])].push_back(l);
    std::vector<long long> ev;
    for (int bi = 0; bi < UD_N_BUCKETS; ++bi) {
        if (bl[bi].empty()) continue;
        // evenly spaced sample of the bucket's ligands
        std::vector<int> sub;
        const int m = std::min<int>(sample_ligands, int(bl[bi].size()));
        for (int k = 0; k < m; ++k) sub.push_back(bl[bi][size_t(k) * bl[bi].size() / m]);
        LaunchCap cap = caps_of(c, &sub);
        int best = 0, best_pb = -1;
        double best_rate = -1.0;
        for (int lanes : {8, 16, 32, 64}) {
            if (fit_lanes(lanes, cap) != lanes) continue;
            for (int pb : {1, 0}) {
                LaunchCap trial = cap;
                (void)fit_lanes(lanes, trial, pb);
                if (trial.pair_forces != pb) continue;  // the buffer does not fit at this lane count
                double ms_min = 1e300;
                long long tot = 0;
                for (int rep = 0; rep < 2; ++rep) {  // first run warms caches and code
                    rc = dock_run(c, p, &sub, lanes, pb);
                    if (rc) return rc;
                    UD_HIP(c, hipStreamSynchronize(c->stream));
                    dock_times(c);
                    ms_min = std::min(ms_min, c->last_ms);
                }
                ev.resize(c->out_n);
                UD_HIP(c, hipMemcpy(ev.data(), c->d_out_evals, ev.size() * sizeof(long long), hipMemcpyDeviceToHost));
                for (int l : sub)
                    for (int k = 0; k < exh; ++k) tot += ev[size_t(l) * exh + k];
                const double rate = double(tot) / std::max(ms_min, 1e-6);  // evaluations per ms
                if (rate > best_rate) {
                    best_rate = rate;
                    best = lanes;
                    best_pb = pb;
                }
            }
        }
        c->tuned_lanes[bi] = best;
        c->tuned_pb[bi] = best_pb;
    }
    if (lanes_out)
        for (int b = 0; b < UD_N_BUCKETS; ++b) lanes_out[b] = c->tuned_lanes[b];
    c->out_n = 0;  // the tuning runs' results are not docking results
    return UD_OK;
}

int ud_score_terms(ud_ctx* c, int n_poses, const int32_t* pose_ligand, const float* poses, float* inter_out,
                   float* intra_out) {
    if (!c || (n_poses > 0 && (!poses || !inter_out || !intra_out))) return UD_ERR_INVALID;
    int rc = check_pose_args(c, n_poses, pose_ligand);
    if (rc || n_poses == 0) return rc;
    UD_HIP(c, hipSetDevice(c->device));
    rc = refresh_slots(c);
    if (rc) return rc;
    rc = pose_upload(c, n_poses, poses);
    if (rc) return rc;
    rc = grow(c, c->d_f1, c->f_cap, size_t(n_poses) * 2);
    if (rc) return rc;
    LaunchCap cap = caps_all(c);
    const int na = cap.n_atoms;
    const int lanes = fit_lanes(c->forced_lanes ? c->forced_lanes : 16, cap);
    if (lds_bytes(lanes, cap) > 65536) return fail(c, UD_ERR_CAPACITY, "ligand too large for one workgroup's LDS");
    std::vector<int32_t> perm;
    std::vector<WorkItem> work;
    make_pose_work(n_poses, pose_ligand, 64 / lanes, perm, work);
    rc = work_upload(c, perm, work);
    if (rc) return rc;
    UD_HIP(c, launch_terms(lanes, grid_format(c), lig_dev(c), grid_dev(c), cap, c->d_work, int(work.size()), c->d_perm,
                           c->d_poses, 1000.f, c->d_f1, c->d_f1 + n_poses, c->stream));
    UD_HIP(c, hipMemcpyAsync(inter_out, c->d_f1, sizeof(float) * n_poses, hipMemcpyDeviceToHost, c->stream));
    UD_HIP(c, hipMemcpyAsync(intra_out, c->d_f1 + n_poses, sizeof(float) * n_poses, hipMemcpyDeviceToHost, c->stream));
    UD_HIP(c, hipStreamSynchronize(c->stream));
    return UD_OK;
}

int ud_pose_coords(ud_ctx* c, int n_poses, const int32_t* pose_ligand, const float* poses, float* coords_out) {
    if (!c || (n_poses > 0 && (!poses || !coords_out))) return UD_ERR_INVALID;
    int rc = check_pose_args(c, n_poses, pose_ligand);
    if (rc || n_poses == 0) return rc;
    if (c->n_lig == 0) return fail(c, UD_ERR_STATE, "no ligands set");
    UD_HIP(c, hipSetDevice(c->device));
    rc = pose_upload(c, n_poses, poses);
    if (rc) return rc;
    rc = grow(c, c->d_f1, c->f_cap, size_t(n_poses) * UD_MAX_ATOMS * 3);
    if (rc) return rc;
    LaunchCap cap = caps_all(c);
    const int na = cap.n_atoms;
    const int lanes = fit_lanes(c->forced_lanes ? c->forced_lanes : 16, cap);
    if (lds_bytes(lanes, cap) > 65536) return fail(c, UD_ERR_CAPACITY, "ligand too large for one workgroup's LDS");
    std::vector<int32_t> perm;
    std::vector<WorkItem> work;
    make_pose_work(n_poses, pose_ligand, 64 / lanes, perm, work);
    rc = work_upload(c, perm, work);
    if (rc) return rc;
    UD_HIP(c, launch_coords(lanes, lig_dev(c), grid_dev(c), cap, c->d_work, int(work.size()), c->d_perm, c->d_poses, c->d_f1,
                            c->stream));
    UD_HIP(c, hipMemcpyAsync(coords_out, c->d_f1, sizeof(float) * size_t(n_poses) * UD_MAX_ATOMS * 3, hipMemcpyDeviceToHost, c->stream));
    UD_HIP(c, hipStreamSynchronize(c->stream));
    return UD_OK;
}

int ud_score_terms_receptor(ud_ctx* c, int n_poses, const int32_t* pose_ligand, const float* poses, float* inter_out,
                            float* intra_out) {
    if (!c || (n_poses > 0 && (!poses || !inter_out || !intra_out))) return UD_ERR_INVALID;
    int rc = check_pose_args(c, n_poses, pose_ligand);
    if (rc || n_poses == 0) return rc;
    if (!c->have_rec) return fail(c, UD_ERR_STATE, "explicit-receptor scoring needs the receptor of ud_compute_maps");
    UD_HIP(c, hipSetDevice(c->device));
    rc = refresh_slots(c);
    if (rc) return rc;
    rc = pose_upload(c, n_poses, poses);
    if (rc) return rc;
    rc = grow(c, c->d_f1, c->f_cap, size_t(n_poses) * 2);
    if (rc) return rc;
    LaunchCap cap = caps_all(c);
    const int lanes = fit_lanes(c->forced_lanes ? c->forced_lanes : 16, cap);
    if (lds_bytes(lanes, cap) > 65536) return fail(c, UD_ERR_CAPACITY, "ligand too large for one workgroup's LDS");
    std::vector<int32_t> perm;
    std::vector<WorkItem> work;
    make_pose_work(n_poses, pose_ligand, 64 / lanes, perm, work);
    rc = work_upload(c, perm, work);
    if (rc) return rc;
    UD_HIP(c, launch_terms_explicit(lanes, lig_dev(c), grid_dev(c), rec_dev(c), cap, c->d_work, int(work.size()),
                                    c->d_perm, c->d_poses, 1000.f, c->d_f1, c->d_f1 + n_poses, c->stream));
    UD_HIP(c, hipMemcpyAsync(inter_out, c->d_f1, sizeof(float) * n_poses, hipMemcpyDeviceToHost, c->stream));
    UD_HIP(c, hipMemcpyAsync(intra_out, c->d_f1 + n_poses, sizeof(float) * n_poses, hipMemcpyDeviceToHost, c->stream));
    UD_HIP(c, hipStreamSynchronize(c->stream));
    return UD_OK;
}

int ud_score_grad_receptor(ud_ctx* c, int n_poses, const int32_t* pose_ligand, const float* poses, const float* v,
                           double slope, float* energy_out, float* grad_out) {
    if (!c || !v || (n_poses > 0 && (!poses || !energy_out))) return UD_ERR_INVALID;
    if (!(slope >= 0.0)) return fail(c, UD_ERR_INVALID, "slope must be >= 0");
    int rc = check_pose_args(c, n_poses, pose_ligand);
    if (rc || n_poses == 0) return rc;
    if (!c->have_rec) return fail(c, UD_ERR_STATE, "explicit-receptor scoring needs the receptor of ud_compute_maps");
    UD_HIP(c, hipSetDevice(c->device));
    rc = refresh_slots(c);
    if (rc) return rc;
    rc = pose_upload(c, n_poses, poses);
    if (rc) return rc;
    rc = grow(c, c->d_f1, c->f_cap, size_t(n_poses) * UD_GRAD_FLOATS);
    if (rc) return rc;
    rc = grow(c, c->d_f2, c->f2_cap, size_t(n_poses));
    if (rc) return rc;
    LaunchCap cap = caps_all(c);
    const int lanes = fit_lanes(c->forced_lanes ? c->forced_lanes : 16, cap);
    std::vector<int32_t> perm;
    std::vector<WorkItem> work;
    make_pose_work(n_poses, pose_ligand, 64 / lanes, perm, work);
    rc = work_upload(c, perm, work);
    if (rc) return rc;
    GridDev g = grid_dev(c);
    g.slope = float(slope);
    UD_HIP(c, launch_score_explicit(lanes, lig_dev(c), g, rec_dev(c), cap, c->d_work, int(work.size()), c->d_perm,
                                    c->d_poses, v[0], v[1], c->d_f2, c->d_f1, c->stream));
    UD_HIP(c, hipMemcpyAsync(energy_out, c->d_f2, sizeof(float) * n_poses, hipMemcpyDeviceToHost, c->stream));
    if (grad_out)
        UD_HIP(c, hipMemcpyAsync(grad_out, c->d_f1, sizeof(float) * size_t(n_poses) * UD_GRAD_FLOATS,
                                 hipMemcpyDeviceToHost, c->stream));
    UD_HIP(c, hipStreamSynchronize(c->stream));
    return UD_OK;
}

int ud_refine(ud_ctx* c, int n_poses, const int32_t* pose_ligand, float* poses, int refine_step, float* energy_out,
              int32_t* within_out) {
    if (!c || (n_poses > 0 && (!poses || !energy_out))) return UD_ERR_INVALID;
    if (refine_step < 1 || refine_step > 8) return fail(c, UD_ERR_INVALID, "refine_step must be in 1..8");
    int rc = check_pose_args(c, n_poses, pose_ligand);
    if (rc || n_poses == 0) return rc;
    if (!c->have_rec) return fail(c, UD_ERR_STATE, "refinement needs the receptor of ud_compute_maps");
    UD_HIP(c, hipSetDevice(c->device));
    rc = refresh_slots(c);
    if (rc) return rc;
    LaunchCap cap = caps_all(c);
    const int lanes = fit_lanes(c->forced_lanes ? c->forced_lanes : default_lanes(cap.n_atoms), cap);
    // all passes in one launch: a pose runs pass k + 1 (slope 100 * 10^2k,
    // vina.h:261) only while its heavy atoms are outside the box
    // (non_cache::within, decided on the device after each pass), as the
    // reference stops a pose's passes once it is within the box
    std::vector<float> slopes(refine_step);
    for (int k = 0; k < refine_step; ++k) slopes[k] = float(100.0 * std::pow(10.0, 2.0 * k));
    const float margin = 1e-4f;  // non_cache::within default (non_cache.h:41)
    BoxLimits lim;
    for (int d = 0; d < 3; ++d) {
        lim.lo[d] = c->init[d] - margin;
        lim.hi[d] = c->gend[d] + margin;
    }
    rc = pose_upload(c, n_poses, poses);
    if (rc) return rc;
    rc = grow(c, c->d_f1, c->f_cap, size_t(n_poses));
    if (rc) return rc;
    rc = grow(c, c->d_i1, c->i_cap, size_t(n_poses));
    if (rc) return rc;
    std::vector<int32_t> perm;
    std::vector<WorkItem> work;
    make_pose_work(n_poses, pose_ligand, 64 / lanes, perm, work);
    rc = work_upload(c, perm, work);
    if (rc) return rc;
    GridDev g = grid_dev(c);
    g.slope = slopes[0];
    UD_HIP(c, hipEventRecord(c->ev0, c->stream));
    UD_HIP(c, launch_refine(lanes, lig_dev(c), g, rec_dev(c), cap, c->d_work, int(work.size()), c->d_perm, c->d_poses,
                            1000.f, 1000.f, 0, c->d_f1, c->d_i1, c->stream, refine_step, slopes.data(), lim));
    UD_HIP(c, hipEventRecord(c->ev1, c->stream));
    // the within flags of the refined poses (the refinement's evaluation
    // counts in d_i1 are not needed)
    UD_HIP(c, launch_coords(lanes, lig_dev(c), g, cap, c->d_work, int(work.size()), c->d_perm, c->d_poses, nullptr,
                            c->stream, 0, &lim, c->d_i1));
    std::vector<float> e(n_poses);
    std::vector<int32_t> inside(n_poses);
    UD_HIP(c, hipMemcpyAsync(poses, c->d_poses, sizeof(float) * size_t(n_poses) * UD_POSE_FLOATS,
                             hipMemcpyDeviceToHost, c->stream));
    UD_HIP(c, hipMemcpyAsync(e.data(), c->d_f1, sizeof(float) * n_poses, hipMemcpyDeviceToHost, c->stream));
    UD_HIP(c, hipMemcpyAsync(inside.data(), c->d_i1, sizeof(int32_t) * n_poses, hipMemcpyDeviceToHost, c->stream));
    UD_HIP(c, hipStreamSynchronize(c->stream));
    float ms_total = 0;
    (void)hipEventElapsedTime(&ms_total, c->ev0, c->ev1);
    for (int i = 0; i < n_poses; ++i) {
        energy_out[i] = inside[i] ? e[i] : kMaxFloat;  // vina.h:268
        if (within_out) within_out[i] = inside[i];
    }
    c->last_ms = ms_total;
    return UD_OK;
}

int ud_set_lanes(ud_ctx* c, int lanes) {
    if (!c) return UD_ERR_INVALID;
    if (lanes != 0 && lanes != 8 && lanes != 16 && lanes != 32 && lanes != 64)
        return fail(c, UD_ERR_INVALID, "lanes must be 0, 8, 16, 32 or 64");
    c->forced_lanes = lanes;
    return UD_OK;
}

int ud_cluster_modes(int n_cand, int n_heavy, const float* heavy_xyz, const float* energy, double min_rmsd,
                     int max_size, int32_t* kept_out, int32_t* n_kept_out) {
    if (n_cand < 0 || n_heavy < 0 || max_size < 1 || !(min_rmsd >= 0.0) || !kept_out || !n_kept_out ||
        (n_cand > 0 && (!energy || (n_heavy > 0 && !heavy_xyz))))
        return UD_ERR_INVALID;
    std::vector<int> kept;
    cluster_modes(n_cand, n_heavy, heavy_xyz, energy, float(min_rmsd), max_size, kept);
    for (size_t k = 0; k < kept.size(); ++k) kept_out[k] = kept[k];
    *n_kept_out = int32_t(kept.size());
    return UD_OK;
}

int ud_postprocess(ud_ctx* c, int max_modes, double min_rmsd, int refine_step, ud_mode* modes_out,
                   int32_t* n_modes_out) {
    if (!c || max_modes < 1 || !(min_rmsd >= 0.0) || !modes_out || !n_modes_out) return UD_ERR_INVALID;
    if (max_modes > UD_MAX_MODES) return fail(c, UD_ERR_INVALID, "max_modes must be at most UD_MAX_MODES");
    if (refine_step < 0 || refine_step > 8) return fail(c, UD_ERR_INVALID, "refine_step must be in 0..8");
    if (c->out_n == 0 || c->last_exh <= 0) return fail(c, UD_ERR_STATE, "no docking results");
    const int exh = c->last_exh, L = c->n_lig;
    UD_HIP(c, hipSetDevice(c->device));
    // every chain's best-pose coordinates, packed per ligand, then the
    // clustering of each ligand's chains, on the device
    size_t n_atoms_all = 0;
    for (int l = 0; l < L; ++l) n_atoms_all += size_t(c->lig_nat[l]);
    int rc = grow(c, c->d_cc, c->cc_cap, n_atoms_all * exh * 3);
    if (rc) return rc;
    rc = grow(c, c->d_kept, c->kept_cap, size_t(L) * max_modes);
    if (rc) return rc;
    rc = grow(c, c->d_nkept, c->nkept_cap, size_t(L));
    if (rc) return rc;
    {
        LaunchCap cap = caps_all(c);
        const int lanes = fit_lanes(c->forced_lanes ? c->forced_lanes : default_lanes(cap.n_atoms), cap);
        const int G = 64 / lanes;
        std::vector<WorkItem> work;
        for (int l = 0; l < L; ++l)
            for (int c0 = 0; c0 < exh; c0 += G) work.push_back(WorkItem{l, l * exh + c0, std::min(G, exh - c0), 0});
        rc = work_upload(c, std::vector<int32_t>(), work);
        if (rc) return rc;
        UD_HIP(c, launch_coords(lanes, lig_dev(c), grid_dev(c), cap, c->d_work, int(work.size()), nullptr,
                                c->d_out_pose, c->d_cc, c->stream, exh));
        UD_HIP(c, launch_cluster(lig_dev(c), L, c->d_cc, c->d_out_e, exh, max_modes, float(min_rmsd), c->d_kept,
                                 c->d_nkept, c->stream));
    }
    // only the kept chains' poses and energies come back
    const size_t slots = size_t(L) * max_modes;
    rc = grow(c, c->d_kp, c->kp_cap, slots * UD_POSE_FLOATS);
    if (rc) return rc;
    rc = grow(c, c->d_ke, c->ke_cap, slots);
    if (rc) return rc;
    UD_HIP(c, launch_gather_kept(c->d_out_pose, c->d_out_e, c->d_kept, c->d_nkept, L, exh, max_modes, c->d_kp, c->d_ke,
                                 c->stream));
    std::vector<int32_t> kept_all(slots), nkept(L);
    std::vector<float> KP(slots * UD_POSE_FLOATS), KE(slots);
    UD_HIP(c, hipMemcpyAsync(kept_all.data(), c->d_kept, kept_all.size() * 4, hipMemcpyDeviceToHost, c->stream));
    UD_HIP(c, hipMemcpyAsync(nkept.data(), c->d_nkept, nkept.size() * 4, hipMemcpyDeviceToHost, c->stream));
    UD_HIP(c, hipMemcpyAsync(KP.data(), c->d_kp, KP.size() * 4, hipMemcpyDeviceToHost, c->stream));
    UD_HIP(c, hipMemcpyAsync(KE.data(), c->d_ke, KE.size() * 4, hipMemcpyDeviceToHost, c->stream));
    UD_HIP(c, hipStreamSynchronize(c->stream));
    dock_times(c);
    // kept modes of every ligand in cluster order
    std::vector<int32_t> kl, kchain;
    std::vector<float> kp, ke;
    for (int l = 0; l < L; ++l)
        for (int k = 0; k < nkept[l]; ++k) {
            const size_t slot = size_t(l) * max_modes + k;
            kl.push_back(l);
            kchain.push_back(kept_all[slot]);
            kp.insert(kp.end(), KP.begin() + slot * UD_POSE_FLOATS, KP.begin() + (slot + 1) * UD_POSE_FLOATS);
            ke.push_back(KE[slot]);
        }
    const int nk_all = int(kl.size());
    // refinement against the explicit receptor when it is known (vina.h:248-272),
    // then rescoring with authentic caps against the same receptor model
    // (AD4.2 modes are not refined, vina.h:247-249)
    const bool refine = refine_step > 0 && c->have_rec && !sf_ad4(c);
    std::vector<float> rkey(nk_all, 0.f), inter(nk_all), intra(nk_all);
    if (nk_all > 0) {
        if (refine) {
            rc = ud_refine(c, nk_all, kl.data(), kp.data(), refine_step, rkey.data(), nullptr);
            if (rc) return rc;
            rc = ud_score_terms_receptor(c, nk_all, kl.data(), kp.data(), inter.data(), intra.data());
        } else {
            rc = ud_score_terms(c, nk_all, kl.data(), kp.data(), inter.data(), intra.data());
        }
        if (rc) return rc;
    }
    // final scores and order per ligand
    std::vector<float> mpose(size_t(nk_all) * UD_POSE_FLOATS);
    std::vector<int32_t> mlig(nk_all), mref(nk_all);
    size_t q = 0, m0 = 0;
    for (int l = 0; l < L; ++l) {
        const int nk = nkept[l];
        ud_mode* out = modes_out + size_t(l) * max_modes;
        n_modes_out[l] = nk;
        if (nk == 0) continue;
        // refined modes are ordered by their refined energy before the best one
        // gives the intramolecular reference (vina.h:274-283)
        std::vector<int> ord(nk);
        for (int k = 0; k < nk; ++k) ord[k] = k;
        if (refine)
            std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return rkey[q + a] < rkey[q + b]; });
        const float intra_best = sf_ad4(c) ? 0.f : intra[q + ord[0]];
        for (int k = 0; k < nk; ++k) {
            ud_mode& m = out[k];
            const size_t i = q + ord[k];
            std::copy(kp.begin() + i * UD_POSE_FLOATS, kp.begin() + (i + 1) * UD_POSE_FLOATS, m.pose);
            m.inter = inter[i];
            m.intra = intra[i];
            if (sf_ad4(c)) {
                // AD4.2 branch of score_gpu (vina.cpp:1289-1305, 1316-1320):
                // total = inter + w_tors TORSDOF, unbound = -intra
                m.conf_independent = c->rot_internal * c->hdr[l].num_tors;
                m.energy = m.inter + m.conf_independent;
                m.unbound = -m.intra;
            } else {
                const float x = m.inter + m.intra - intra_best;
                m.energy = conf_independent(x, c->hdr[l].num_tors, c->rot_internal);
                m.conf_independent = m.energy - x;
                m.unbound = intra_best;
            }
            m.search_energy = ke[i];
            m.chain = int32_t(size_t(l) * exh + kchain[i]);
        }
        std::stable_sort(out, out + nk, [](const ud_mode& a, const ud_mode& b) { return a.energy < b.energy; });
        for (int k = 0; k < nk; ++k) {
            std::copy(out[k].pose, out[k].pose + UD_POSE_FLOATS, mpose.begin() + (m0 + k) * UD_POSE_FLOATS);
            mlig[m0 + k] = l;
            mref[m0 + k] = int32_t(m0);  // RMSD bounds from the ligand's best mode
        }
        q += nk;
        m0 += nk;
    }
    if (nk_all == 0) return UD_OK;
    // RMSD bounds of every mode against its ligand's best mode, on the device
    LaunchCap cap = caps_all(c);
    const int lanes = fit_lanes(c->forced_lanes ? c->forced_lanes : default_lanes(cap.n_atoms), cap);
    std::vector<int32_t> perm;
    std::vector<WorkItem> work;
    make_pose_work(nk_all, mlig.data(), 64 / lanes, perm, work);
    rc = pose_upload(c, nk_all, mpose.data());
    if (rc) return rc;
    rc = work_upload(c, perm, work);
    if (rc) return rc;
    rc = grow(c, c->d_mc, c->mc_cap, size_t(nk_all) * UD_MAX_ATOMS * 3);
    if (rc) return rc;
    rc = grow(c, c->d_mlig, c->mlig_cap, size_t(nk_all));
    if (rc) return rc;
    rc = grow(c, c->d_mref, c->mref_cap, size_t(nk_all));
    if (rc) return rc;
    rc = grow(c, c->d_ub, c->ub_cap, size_t(nk_all));
    if (rc) return rc;
    rc = grow(c, c->d_lb, c->lb_cap, size_t(nk_all));
    if (rc) return rc;
    UD_HIP(c, hipMemcpyAsync(c->d_mlig, mlig.data(), mlig.size() * 4, hipMemcpyHostToDevice, c->stream));
    UD_HIP(c, hipMemcpyAsync(c->d_mref, mref.data(), mref.size() * 4, hipMemcpyHostToDevice, c->stream));
    UD_HIP(c, launch_coords(lanes, lig_dev(c), grid_dev(c), cap, c->d_work, int(work.size()), c->d_perm, c->d_poses,
                            c->d_mc, c->stream));
    UD_HIP(c, launch_bounds(lig_dev(c), c->d_elem, nk_all, c->d_mc, c->d_mlig, c->d_mref, c->d_ub, c->d_lb, c->stream));
    std::vector<float> ub(nk_all), lb(nk_all);
    UD_HIP(c, hipMemcpyAsync(ub.data(), c->d_ub, ub.size() * 4, hipMemcpyDeviceToHost, c->stream));
    UD_HIP(c, hipMemcpyAsync(lb.data(), c->d_lb, lb.size() * 4, hipMemcpyDeviceToHost, c->stream));
    UD_HIP(c, hipStreamSynchronize(c->stream));
    m0 = 0;
    for (int l = 0; l < L; ++l) {
        ud_mode* out = modes_out + size_t(l) * max_modes;
        for (int k = 0; k < nkept[l]; ++k) {
            out[k].rmsd_ub = ub[m0 + k];
            out[k].rmsd_lb = lb[m0 + k];
        }
        m0 += nkept[l];
    }
    return UD_OK;
}

int ud_last_kernel_time(const ud_ctx* c, double* ms, int32_t* launches) {
    if (!c) return UD_ERR_INVALID;
    if (ms) *ms = c->last_ms;
    if (launches) *launches = c->last_launches;
    return UD_OK;
}

}  // extern "C"
