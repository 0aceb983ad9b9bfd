// Host-side launch wrappers of the kernels in ud_kernels.hip.
#pragma once
#include <hip/hip_runtime.h>

#include "ud_internal.h"

namespace ud {

static constexpr int kTableStride = 2050;  // table entries stored per type pair
static constexpr int kTypePairs = 36 * 37 / 2;  // XS types; the 35 AD types need 630
// grid-format template values of the kernels: FP32 / FP16 maps (UD_GRID_*),
// the explicit receptor, and AD4.2 over FP32 / FP16 maps
static constexpr int kGridAd4F32 = 3, kGridAd4F16 = 4;

struct GridBuildArgs {
    float init[3];
    float factor_inv[3];
    int32_t n[3];
    int32_t n_points;
    int32_t n_maps;
    int32_t types[UD_MAX_MAPS];
};

size_t lds_bytes(int lanes, const LaunchCap& cap);

hipError_t launch_tables(int sf, const float* w, float2* deriv, float* fast, hipStream_t st);
hipError_t launch_grid_build(const float4* rec, int n_rec, const GridBuildArgs& a, const float* table_fast,
                             float* maps_out, hipStream_t st);
// AD4.2 maps: a.n_maps affinity maps, then the electrostatic and desolvation
// maps (rec.w holds the AD type bits); w = (vdw, hb, elec, dsolv)
hipError_t launch_grid_build_ad4(const float4* rec, const float* rec_q, int n_rec, const GridBuildArgs& a,
                                 const float* w, float* maps_out, hipStream_t st);
hipError_t launch_grid_absmax(const float* maps, int n_points, int n_maps, unsigned int* out, hipStream_t st);
hipError_t launch_grid_to_half(const float* maps, int n_points, int n_maps, const float* inv_scale, void* out,
                               hipStream_t st);
hipError_t launch_score(int lanes, int gf, const LigandSetDev& ls, const GridDev& g, const LaunchCap& cap,
                        const WorkItem* work, int n_work, const int32_t* perm, const float* poses, float v_intra,
                        float v_inter, float* e_out, float* g_out, hipStream_t st);
hipError_t launch_optimize(int lanes, int gf, const LigandSetDev& ls, const GridDev& g, const LaunchCap& cap,
                           const WorkItem* work, int n_work, const int32_t* perm, float* poses, float v_intra,
                           float v_inter, int max_steps, float* e_out, int32_t* evals_out, hipStream_t st);
hipError_t launch_dock(int lanes, int gf, const LigandSetDev& ls, const GridDev& g, const LaunchCap& cap,
                       const SearchDev& sp, const WorkItem* work, int n_work, float* out_pose, float* out_e,
                       long long* out_evals, hipStream_t st);
hipError_t launch_terms(int lanes, int gf, const LigandSetDev& ls, const GridDev& g, const LaunchCap& cap, const WorkItem* work,
                        int n_work, const int32_t* perm, const float* poses, float v, float* inter_out,
                        float* intra_out, hipStream_t st);
hipError_t launch_energy_grad(int lanes, int gf, const LigandSetDev& ls, const GridDev& g, const TermsDev& td,
                              const LaunchCap& cap, const WorkItem* work, int n_work, const int32_t* perm,
                              const float* poses, float v_intra, float v_inter, float* e_out, float* gpose_out,
                              float* gw_out, hipStream_t st);
// Box limits of a within test: [lo, hi] per axis (the box with its margin).
struct BoxLimits {
    float lo[3], hi[3];
};
// Refinement BFGS against the explicit receptor: `passes` passes, pass k
// with penalty slope slopes[k]; a pose whose heavy atoms stay inside lim
// after a pass stops there.
hipError_t launch_refine(int lanes, const LigandSetDev& ls, const GridDev& g, const RecDev& rec, const LaunchCap& cap,
                         const WorkItem* work, int n_work, const int32_t* perm, float* poses, float v_intra,
                         float v_inter, int max_steps, float* e_out, int32_t* evals_out, hipStream_t st,
                         int passes, const float* slopes, const BoxLimits& lim);
hipError_t launch_score_explicit(int lanes, const LigandSetDev& ls, const GridDev& g, const RecDev& rec,
                                 const LaunchCap& cap, const WorkItem* work, int n_work, const int32_t* perm,
                                 const float* poses, float v_intra, float v_inter, float* e_out, float* g_out,
                                 hipStream_t st);
hipError_t launch_terms_explicit(int lanes, const LigandSetDev& ls, const GridDev& g, const RecDev& rec,
                                 const LaunchCap& cap, const WorkItem* work, int n_work, const int32_t* perm,
                                 const float* poses, float v, float* inter_out, float* intra_out, hipStream_t st);
// coords_out: UD_MAX_ATOMS atoms per pose, or with compact_exh > 0 the
// chains' atoms packed per ligand (coords_kernel).  With within_out set, no
// coordinates are written: within_out[pose] = 1 if every heavy atom lies in
// [lo - margin, hi + margin] on every axis (non_cache::within), else 0.
hipError_t launch_coords(int lanes, const LigandSetDev& ls, const GridDev& g, const LaunchCap& cap, const WorkItem* work,
                         int n_work, const int32_t* perm, const float* poses, float* coords_out, hipStream_t st,
                         int compact_exh = 0, const BoxLimits* lim = nullptr, int32_t* within_out = nullptr);
// add_to_output_container + remove_redundant per ligand (cluster_kernel)
hipError_t launch_cluster(const LigandSetDev& ls, int n_lig, const float* coords, const float* energy, int exh,
                          int max_size, float min_rmsd, int32_t* kept_out, int32_t* n_out, hipStream_t st);
// RMSD bounds of modes against reference modes (bounds_kernel)
hipError_t launch_gather_kept(const float* poses, const float* energy, const int32_t* kept, const int32_t* n_kept,
                              int n_lig, int exh, int max_size, float* poses_out, float* e_out, hipStream_t st);
hipError_t launch_bounds(const LigandSetDev& ls, const int32_t* elem, int n_modes, const float* coords,
                         const int32_t* mode_lig, const int32_t* mode_ref, float* ub_out, float* lb_out,
                         hipStream_t st);

}  // namespace ud
