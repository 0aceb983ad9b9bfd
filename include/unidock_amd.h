/* unidock_amd.h -- C ABI of the MI355X (gfx950) batched pose-search library.
 *
 * The reference has no formal plugin API; its GPU seam is C++:
 *   monte_carlo_template::do_docking<Config>(models, outputs, precalculate_byatom,
 *       m_data_list_gpu, igrid, corner1, corner2, rng, verbosity, seed, bias)
 *       -- unidock/src/lib/monte_carlo.h:121-126, body unidock/src/cuda/monte_carlo.cu:2397-2974
 *   precalculate_parallel(m_data_list, precalculate_byatom, sf, models, rs, thread)
 *       -- unidock/src/lib/precalculate.h:315-318, body unidock/src/cuda/precalculate.cu:169-326
 * called from Vina::global_search_gpu<Config> (unidock/src/lib/vina.h:154-358) and
 * Vina::set_ligand_from_object_gpu (unidock/src/lib/vina.cpp:246-303).
 * Each entry point below cites the reference interface it replaces.
 *
 * Conventions: plain pointers and sizes, no exceptions, no exit().  Every
 * call returns UD_OK (0) or an error code; ud_last_error() describes the last
 * failure on a context.  Host pointers are caller-owned and only read/written
 * during the call; device buffers are owned by the context and reused across
 * calls.  A context is bound to one device and one HIP stream and must not be
 * used from two threads at once; separate contexts are independent (the
 * reference's paired mode runs one Vina object per std::thread,
 * unidock/src/main/simulation_container.h:419-449).
 */
#ifndef UNIDOCK_AMD_H
#define UNIDOCK_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
enum {
    UD_OK = 0,
    UD_ERR_INVALID = 1,   /* bad argument or descriptor */
    UD_ERR_CAPACITY = 2,  /* a ligand/grid exceeds a documented capacity */
    UD_ERR_HIP = 3,       /* HIP runtime failure */
    UD_ERR_STATE = 4,     /* call order violated (e.g. dock before maps) */
    UD_ERR_NOMEM = 5      /* device allocation failed */
};

/* scoring functions (scoring_function.h:34) */
enum { UD_SF_VINA = 0, UD_SF_VINARDO = 1, UD_SF_AD4 = 2 };
/* AutoDock 4.2 map indices beside the AutoDock atom types 0..34
 * (atom_constants.h:46-81): ad4cache's m_grids[AD_TYPE_SIZE] and
 * m_grids[AD_TYPE_SIZE + 1] (ad4cache.cpp:309-318) */
#define UD_AD4_MAP_ELEC 35
#define UD_AD4_MAP_DSOLV 36
/* grid-map storage: FP32, or FP16 with a power-of-two scale per map; all
 * interpolation and accumulation is FP32 either way */
enum { UD_GRID_F32 = 0, UD_GRID_F16 = 1 };

/* capacities: the atoms and torsions of the reference's largest size class
 * (MaxConfig: 300 atoms, 48 torsions, 4,096 pairs, unidock/src/cuda/kernel.h:
 * 242-277), and more pairs; exceeding them returns UD_ERR_CAPACITY, never
 * undefined behaviour */
#define UD_MAX_ATOMS 300
#define UD_MAX_TORSIONS 48
#define UD_MAX_PAIRS 8128
#define UD_MAX_MAPS 32
#define UD_MAX_MODES 64  /* output modes per ligand of ud_postprocess */
#define UD_MAX_GRID_POINTS (1 << 24)

/* A pose is UD_POSE_FLOATS floats: position[3], orientation quaternion
 * (w, x, y, z)[4], torsions[UD_MAX_TORSIONS] (pre-order of the torsion tree);
 * the layout of output_type_cuda_t minus its coordinates (kernel.h:296-306). */
#define UD_POSE_FLOATS (7 + UD_MAX_TORSIONS)
/* A gradient is UD_GRAD_FLOATS floats: force[3], torque[3], d/dtorsion[T]
 * (the change vector of conf.h:182-227). */
#define UD_GRAD_FLOATS (6 + UD_MAX_TORSIONS)
/* Term-weight gradients: UD_WEIGHT_FLOATS floats per pose, in the order of
 * the weights array of ud_set_scoring (Vina: gauss1, gauss2, repulsion,
 * hydrophobic, hydrogen, glue, rot; Vinardo: gauss1, repulsion, hydrophobic,
 * hydrogen, glue, rot, unused). */
#define UD_WEIGHT_FLOATS 7

/* ud_dock groups ligands into size buckets by heavy+hydrogen atom count
 * (<= 16, 32, 48, 64, 96, 128, 192, 300 atoms) and launches each bucket once; the
 * reference groups ligands by size class too (SmallConfig / MediumConfig /
 * LargeConfig ..., kernel.h:242-277, chosen in main.cpp:267-275). */
#define UD_N_BUCKETS 8
/* Launches of one ud_dock: each size bucket's ligands with at most 10
 * torsions (whose search keeps the inverse Hessian in registers) and those
 * with more are launched separately. */
#define UD_MAX_LAUNCHES (2 * UD_N_BUCKETS)

/* Timing of one launch of the last ud_dock (ud_last_launches). */
typedef struct ud_launch_info {
    int32_t atom_cap;      /* atom capacity of the bucket */
    int32_t lanes;         /* lanes per chain */
    int32_t n_ligands;
    int32_t n_workgroups;  /* one wavefront each */
    int32_t pair_buffer;   /* 1: pair-parallel forces through a per-pair LDS buffer; 0: per-atom pair loops */
    int32_t lds_bytes;     /* LDS per workgroup */
    int32_t max_torsions;  /* most torsions of a ligand in the launch */
    double ms;             /* kernel time, HIP events around the launch on its stream */
} ud_launch_info;

typedef struct ud_ctx ud_ctx;

/* A batch of ligands in flat arrays (offsets have n_ligands + 1 entries).
 * Per ligand, atoms are grouped by rigid node and nodes are in depth-first
 * pre-order with the root first; node k > 0 carries torsion k - 1.
 *   atom_local[3a]   coordinates relative to the owning node's origin in the
 *                    input frame (parse_pdbqt.cpp:206-211)
 *   atom_xs[a]       X-Score type (0..35) or -1 for untyped atoms (hydrogens)
 *   atom_heavy[a]    1 for heavy atoms (gyration radius)
 *   node_parent[k]   parent node (local index), -1 for the root
 *   node_abegin/aend atom range of the node (local indices)
 *   node_origin[3k]  node origin in the input frame; node_axis[3k] unit
 *                    rotation axis (ignored for the root) (tree.h:101-131)
 *   pairs[2p]        intramolecular pairs (local atom indices)
 *                    (model.cpp:592-630, ligand.pairs)
 *   num_tors[l]      torsion count of conf_independent (conf_independent.cpp:79-103)
 * For UD_SF_AD4 (read at ud_set_ligands; may be NULL otherwise):
 *   atom_ad[a]       AutoDock type 0..34 (atom_constants.h:46-81); every atom
 *                    is typed, hydrogens included
 *   atom_charge[a]   partial charge (parse_pdbqt.cpp:112-114, 678-680)
 *   torsdof[l]       TORSDOF of ad4_tors_add (conf_independent.cpp:183-187)
 *   and `pairs` lists the pairs of AutoDock-typed atoms (model.cpp:608-612;
 *   ud_ligset_batch_sf builds either list).
 * Replaces the per-ligand model / m_cuda_t marshalling of do_docking_base
 * (monte_carlo.cu:2487-2631). */
typedef struct {
    int32_t n_ligands;
    const int32_t* atom_offset;
    const int32_t* node_offset;
    const int32_t* pair_offset;
    const float* atom_local;
    const int32_t* atom_xs;
    const int32_t* atom_heavy;
    const int32_t* node_parent;
    const int32_t* node_abegin;
    const int32_t* node_aend;
    const float* node_origin;
    const float* node_axis;
    const int32_t* pairs;
    const float* num_tors;
    const int32_t* atom_ad;
    const float* atom_charge;
    const float* torsdof;
} ud_ligand_batch;

/* Monte Carlo search parameters: the public fields of monte_carlo_template
 * (monte_carlo.h:75-98) as set by Vina::global_search_gpu (vina.h:189-212). */
typedef struct {
    int32_t exhaustiveness;      /* chains per ligand */
    int32_t global_steps;        /* MC steps per chain (min(heuristic, max_step)) */
    int32_t local_steps;         /* BFGS iterations; <= 0: (25 + n_atoms) / 3 per ligand */
    int32_t lanes_per_chain;     /* 0: autotuned per size bucket, else 8/16/32/64 */
    uint64_t seed;
    float hunt_cap[3];           /* curl caps during the search (10, 10, 10) */
    float authentic_v[3];        /* caps of the refinement BFGS (1000, 1000, 1000) */
    float temperature;           /* Metropolis temperature (1.2) */
    float mutation_amplitude;    /* 2 */
    int32_t chains_per_group;    /* chains each lane group of a workgroup runs in turn, taking
                                    the workgroup's next chain when one ends; 0: 1 (one chain
                                    per group, the finest workgroups).  No counterpart in the
                                    reference, whose blocks each run one chain. */
    int32_t first_ls_f0_zero;    /* 1: the first line search of each BFGS compares with
                                    f0 = 0, as the reference's GPU BFGS does (bfgs_warp,
                                    warp_ops.cuh:1062-1076); 0: with f0 = f(x), the CPU rule
                                    (bfgs.h:107) */
    int32_t single_launch;       /* 1: one launch over the whole batch with capacities of its
                                    largest ligand, as the reference launches one kernel per
                                    batch (monte_carlo.cu:2877-2882); 0: one launch per size
                                    bucket */
} ud_search_params;

/* fill *p with the reference defaults (exhaustiveness 8, global_steps 20) */
void ud_default_search_params(ud_search_params* p);

/* ---- context ---------------------------------------------------------- */
/* Create a context on device_id using HIP stream `stream` (NULL = the null
 * stream).  Replaces the implicit CUDA device/stream use of do_docking_base
 * (monte_carlo.cu:2414-2480) and mc_stream's own stream (monte_carlo.cu:367-368). */
int ud_create(int device_id, void* stream, ud_ctx** out);
int ud_destroy(ud_ctx* ctx);
const char* ud_last_error(const ud_ctx* ctx);

/* Scoring function and weights.  Vina: 7 weights (gauss1, gauss2, repulsion,
 * hydrophobic, hydrogen, glue, rot); Vinardo: 6 (gauss1, repulsion,
 * hydrophobic, hydrogen, glue, rot); AD4.2: 6 (vdw, hb, elec, dsolv, glue,
 * rot).  Replaces Vina::set_vina_weights / set_vinardo_weights /
 * set_ad4_weights + set_forcefield (vina.cpp:393-450) and builds the
 * per-type-pair tables that precalculate_parallel builds per atom pair
 * (precalculate.cu:14-167, precalculate.h:203-241); AD4.2's charge-dependent
 * terms become two shared radial tables times per-pair coefficients.
 * Ligands and maps are typed for the scoring function in force when they are
 * set: after switching between AD4.2 and Vina / Vinardo set both again. */
int ud_set_scoring(ud_ctx* ctx, int sf, const double* weights, int n_weights);

/* Search box: n_voxels = ceil(size / granularity), begin = center - span / 2
 * (Vina::compute_vina_maps geometry, vina.cpp:527-539).  n_points_out (3) may
 * be NULL. */
int ud_set_box(ud_ctx* ctx, const double* center, const double* size, double granularity, int32_t* n_points_out);

/* Build grid maps on the GPU for the listed XS types from the receptor atoms
 * (x, y, z per atom; types -1 are skipped).  Replaces cache::populate_no_bias
 * (cache.cpp:471-549) called from Vina::compute_vina_maps (vina.cpp:491-567).
 * precision: UD_GRID_F32 or UD_GRID_F16. */
int ud_compute_maps(ud_ctx* ctx, int n_rec, const float* rec_xyz, const int32_t* rec_xs, int n_types,
                    const int32_t* types, int precision);

/* Load precomputed FP32 map values (x fastest, one map per listed type) and
 * store them at `precision` (Vina::load_maps, vina.cpp:569-607). */
int ud_set_maps(ud_ctx* ctx, int n_types, const int32_t* types, const float* values, int precision);

/* AD4.2 maps on the GPU from receptor atoms with AutoDock types and partial
 * charges: an affinity map per listed AutoDock type plus the electrostatic
 * (UD_AD4_MAP_ELEC) and desolvation (UD_AD4_MAP_DSOLV) maps, the AutoGrid
 * decomposition of the reference's AD4.2 pair potentials (potentials.h:
 * 346-484), so that an atom of type t and charge q at p scores
 * map_t(p) + q elec(p) + |q| dsolv(p).  The reference reads such maps from
 * AutoGrid files (ad4cache::read, ad4cache.cpp:391-440; ud_set_maps takes
 * them with both extra maps in `types`).  Needs UD_SF_AD4. */
int ud_compute_maps_ad4(ud_ctx* ctx, int n_rec, const float* rec_xyz, const int32_t* rec_ad, const float* rec_charge,
                        int n_types, const int32_t* types, int precision);

/* Read back one map as FP32 (dequantised for FP16 storage). */
int ud_get_map(ud_ctx* ctx, int xs_type, float* out);

/* Set the ligand batch (copied to the device; replaces
 * Vina::set_ligand_from_object_gpu, vina.cpp:246-303). */
int ud_set_ligands(ud_ctx* ctx, const ud_ligand_batch* batch);

/* Energy and gradient of n_poses poses (pose_ligand[i] = ligand index).
 * v = curl caps (intra, inter, other) as in model::eval_deriv (model.cpp:954-981).
 * energy_out[n_poses], grad_out[n_poses * UD_GRAD_FLOATS]; host pointers.
 * The device path of the reference is m_eval_deriv_warp (warp_ops.cuh:836-860). */
int ud_score_grad(ud_ctx* ctx, int n_poses, const int32_t* pose_ligand, const float* poses, const float* v,
                  float* energy_out, float* grad_out);
/* Same with device pointers, asynchronous on the context's stream. */
int ud_score_grad_device(ud_ctx* ctx, int n_poses, const int32_t* d_pose_ligand, const float* d_poses,
                         const float* v, float* d_energy, float* d_grad);

/* ud_score_grad_device for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  The same kernel and results
 * as ud_score_grad_device; its pose permutation and work list live in a
 * static device array of the library (UD_NOALLOC_SCRATCH_BYTES per device,
 * held until the process ends) instead of context buffers, and the call
 * makes no hipMalloc / hipFree.  At most UD_NOALLOC_MAX_POSES poses
 * (UD_ERR_CAPACITY beyond).  Calls that overlap on one device (other
 * contexts or streams of that device) share the array and race: serialise
 * them.  Maps and ligands must be set as for ud_score_grad_device. */
#define UD_NOALLOC_MAX_POSES 65536
#define UD_NOALLOC_SCRATCH_BYTES (UD_NOALLOC_MAX_POSES * 20)
int ud_score_grad_device_noalloc(ud_ctx* ctx, int n_poses, const int32_t* pose_ligand, const float* d_poses,
                                 const float* v, float* d_energy, float* d_grad);

/* Energy of poses with its gradient in the pose's own coordinates and in
 * the scoring-term weights, for differentiable use (torch.autograd).  The
 * energy is model::eval_deriv's (model.cpp:954-981) at the pose with its
 * quaternion normalised, i.e. E(position, q / |q|, torsions).
 * pose_grad_out[n_poses * UD_POSE_FLOATS]: dE/dposition[3], dE/dq[4]
 * (w, x, y, z; the gradient of E(q / |q|)), dE/dtorsion[T], zeros after.
 * weight_grad_out[n_poses * UD_WEIGHT_FLOATS]: dE/dw per term weight; the rot
 * entry is 0 (the energy has no conf-independent part).  Needs maps built by
 * ud_compute_maps (the per-term maps are built from its receptor on first
 * use).  No counterpart in the reference, whose eval_deriv returns the
 * change-space gradient only (ud_score_grad); its `ud_score_grad(...,
 * dE_dw_out)` seam is this call.  Vina and Vinardo (UD_ERR_STATE for
 * AD4.2).  Host pointers; pose_ligand is on the host. */
int ud_energy_grad(ud_ctx* ctx, int n_poses, const int32_t* pose_ligand, const float* poses, const float* v,
                   float* energy_out, float* pose_grad_out, float* weight_grad_out);
/* Same with device pointers for poses and outputs (pose_ligand stays on the
 * host), asynchronous on the context's stream. */
int ud_energy_grad_device(ud_ctx* ctx, int n_poses, const int32_t* pose_ligand, const float* d_poses, const float* v,
                          float* d_energy, float* d_pose_grad, float* d_weight_grad);

/* BFGS local optimisation from given starts (poses in/out), quasi_newton +
 * bfgs (quasi_newton.cpp:43-53, bfgs.h:97-147; device: bfgs_warp,
 * warp_ops.cuh:1051-1119).  max_steps <= 0: (25 + n_atoms) / 3 per ligand.
 * energy_out and evals_out may be NULL. */
int ud_optimize(ud_ctx* ctx, int n_poses, const int32_t* pose_ligand, float* poses, const float* v, int max_steps,
                float* energy_out, int32_t* evals_out);

/* Monte Carlo docking of every ligand of the batch, exhaustiveness chains
 * each, in one launch per size bucket (kernel<32, Config>,
 * monte_carlo.cu:236-356, launched at monte_carlo.cu:2877-2882).  The
 * bucket launches run concurrently on streams of the context, ordered after
 * earlier work on the context's stream and before later work on it.
 * Outputs per chain (ligand-major, chain-minor): best_poses
 * [n_ligands * exh * UD_POSE_FLOATS], best_e [n_ligands * exh], evals
 * [n_ligands * exh] (energy+gradient evaluations).  Any output may be NULL:
 * results then stay on the device for ud_fetch_dock_results. */
int ud_dock(ud_ctx* ctx, const ud_search_params* params, float* best_poses, float* best_e, int64_t* evals);
int ud_fetch_dock_results(ud_ctx* ctx, float* best_poses, float* best_e, int64_t* evals);

/* Final-score components of poses with authentic caps (v = 1000): inter =
 * grid energy (cache::eval, cache.cpp:84-114), intra = intramolecular pairs
 * with the piecewise-constant table (eval_interacting_pairs,
 * model.cpp:885-905).  AD4.2: inter also holds the electrostatic and
 * desolvation maps times q and |q| (ad4cache::eval, ad4cache.cpp:151-185).
 * Used by the host post-processing (Vina::score_gpu, vina.cpp:1245-1327). */
int ud_score_terms(ud_ctx* ctx, int n_poses, const int32_t* pose_ligand, const float* poses, float* inter_out,
                   float* intra_out);

/* Heavy-atom coordinates of poses (model::set + get_heavy_atom_movable_coords,
 * model.cpp:699-706), coords_out[n_poses * UD_MAX_ATOMS * 3], all atoms in
 * ligand order. */
int ud_pose_coords(ud_ctx* ctx, int n_poses, const int32_t* pose_ligand, const float* poses, float* coords_out);

/* Lanes per pose for ud_score_grad / ud_optimize / ud_score_terms /
 * ud_pose_coords on this context: 8, 16, 32 or 64; 0 restores the automatic
 * choice.  (The reference fixes 32 threads per pose, monte_carlo.cu:117-118;
 * here a wavefront of 64 lanes holds 64 / lanes poses.) */
int ud_set_lanes(ud_ctx* ctx, int lanes);

/* Kernel timing of the last ud_dock / ud_score_grad call on the context's
 * stream (HIP events around the launches): total milliseconds and number of
 * launches. */
int ud_last_kernel_time(const ud_ctx* ctx, double* ms, int32_t* launches);

/* Per-launch timing of the last ud_dock (resolved by ud_fetch_dock_results,
 * or by ud_dock when it fetches): up to max_launches entries (at most
 * UD_MAX_LAUNCHES launches), launch order (largest bucket first); n_out
 * receives the number of launches.  The
 * reference times its docking loop on the host only (main.cpp:272-276). */
int ud_last_launches(const ud_ctx* ctx, int max_launches, ud_launch_info* out, int32_t* n_out);

/* Lanes per chain of each size bucket for ud_dock (0 = built-in default).
 * The reference fixes 32 threads per chain for every size class
 * (monte_carlo.cu:117-118, 2877-2882). */
int ud_set_bucket_lanes(ud_ctx* ctx, const int32_t* lanes /* [UD_N_BUCKETS] */);

/* LDS layout of each size bucket for ud_dock: 1 = intramolecular pair
 * forces evaluated once per pair through a per-pair buffer in LDS (when it
 * fits), 0 = per-atom pair loops without that buffer (each pair evaluated by
 * both of its atoms; less LDS per workgroup, more resident workgroups for
 * large ligands), -1 = built-in default (the buffer when it fits).  The
 * reference keeps its per-thread pair work in registers and global memory
 * (kernel.h:98-277 sizes its shared arrays per size class). */
int ud_set_bucket_pair_buffer(ud_ctx* ctx, const int32_t* pair_buffer /* [UD_N_BUCKETS] */);

/* The current per-bucket settings: lanes per chain (0 = default) and LDS
 * layout (-1, 0, 1 as ud_set_bucket_pair_buffer).  Either pointer may be NULL. */
int ud_get_bucket_layout(const ud_ctx* ctx, int32_t* lanes_out, int32_t* pair_buffer_out);

/* Autotuner: for every size bucket of the current ligand set, docks an
 * evenly spaced sample of up to sample_ligands ligands with params at 8, 16,
 * 32 and 64 lanes per chain, each with and without the per-pair LDS buffer
 * (the layouts that fit the LDS), keeps the combination with the most
 * energy+gradient evaluations per second and stores it as the bucket's
 * setting (lanes also written to lanes_out[UD_N_BUCKETS] if non-NULL;
 * ud_get_bucket_layout reads both).  Discards earlier docking results.  No
 * counterpart in the reference (fixed 32 threads per chain). */
int ud_autotune(ud_ctx* ctx, const ud_search_params* params, int sample_ligands, int32_t* lanes_out);

/* One output mode of a ligand after ud_postprocess. */
typedef struct ud_mode {
    float pose[UD_POSE_FLOATS];
    float energy;            /* final affinity: conf-independent form of inter + intra - intra_best */
    float inter;             /* ligand-receptor energy (caps 1000; maps, or explicit atoms if refined) */
    float intra;             /* ligand intramolecular energy (caps 1000) */
    float conf_independent;  /* energy - (inter + intra - intra_best), the torsion term */
    float search_energy;     /* the chain's best energy from the search (before refinement) */
    float rmsd_lb, rmsd_ub;  /* from the best mode (model::rmsd_lower/upper_bound) */
    float unbound;           /* intramolecular energy of the best pose (subtracted above) */
    int32_t chain;           /* chain the mode comes from (ligand * exhaustiveness + k) */
} ud_mode;

/* Post-processing of the last ud_dock: the part of Vina::global_search_gpu
 * after do_docking (vina.h:232-345).  Per ligand: the chains' best poses in
 * chain order go through add_to_output_container (coords.cpp:42-68;
 * heavy-atom RMSD upper bound < min_rmsd keeps the better of two, at most
 * max_modes kept) as monte_carlo.cu:2940-2943 does, then remove_redundant
 * (vina.cpp:1465-1470).  With refine_step > 0 and a receptor from
 * ud_compute_maps, every kept mode is refined against the explicit receptor
 * (ud_refine, vina.h:249-272) and the modes are ordered by refined energy;
 * refine_step 0 (or maps from ud_set_maps) is the reference's no_refine path.
 * The intramolecular energy of the first mode is subtracted and every mode is
 * rescored (score_gpu, vina.cpp:1245-1327: maps, or the explicit receptor
 * after refinement), sorted by the final energy and given RMSD bounds from
 * the best mode.  modes_out[l * max_modes + k], n_modes_out[l].  The
 * reference CLI's default refine_step is 3 (main.cpp:557-558).  AD4.2 modes
 * are not refined and score inter + w_rot TORSDOF with unbound = -intra
 * (score_gpu's AD4.2 branch, vina.cpp:1289-1320). */
int ud_postprocess(ud_ctx* ctx, int max_modes, double min_rmsd, int refine_step, ud_mode* modes_out,
                   int32_t* n_modes_out);

/* Refinement of poses against the explicit receptor atoms (non_cache,
 * non_cache.cpp:420-860) as Vina::global_search_gpu does for the kept modes
 * (vina.h:249-272): BFGS with caps 1000 and (25 + atoms) / 3 steps, for pass
 * p = 0 .. refine_step - 1 with box slope 100 * 10^(2p), until every heavy
 * atom is within the box (non_cache::within, margin 1e-4).  poses are updated
 * in place; energy_out gets the last BFGS energy, or FLT_MAX for a pose that
 * never got inside (vina.h:268); within_out (may be NULL) 1 / 0.  Needs the
 * receptor of ud_compute_maps. */
int ud_refine(ud_ctx* ctx, int n_poses, const int32_t* pose_ligand, float* poses, int refine_step, float* energy_out,
              int32_t* within_out);

/* Energy and gradient (as ud_score_grad) against the explicit receptor
 * atoms with box slope `slope`: model::eval_deriv with non_cache as the
 * receptor term (non_cache.cpp:420-860), what the refinement minimises. */
int ud_score_grad_receptor(ud_ctx* ctx, int n_poses, const int32_t* pose_ligand, const float* poses, const float* v,
                           double slope, float* energy_out, float* grad_out);

/* Final-score components against the explicit receptor (score_gpu's
 * refined branch, vina.cpp:1262-1283): inter = non_cache::eval with caps
 * 1000 and box slope 1e6 (non_cache.cpp:34-255, fast table), intra = the
 * ligand pairs as in ud_score_terms. */
int ud_score_terms_receptor(ud_ctx* ctx, int n_poses, const int32_t* pose_ligand, const float* poses,
                            float* inter_out, float* intra_out);

/* add_to_output_container over n_cand candidates in order (coords.cpp:42-68):
 * heavy_xyz[n_cand][n_heavy][3], energy[n_cand]; writes the kept candidate
 * indices sorted by energy to kept_out[max_size] and their number to
 * n_kept_out.  Host only. */
int ud_cluster_modes(int n_cand, int n_heavy, const float* heavy_xyz, const float* energy, double min_rmsd,
                     int max_size, int32_t* kept_out, int32_t* n_kept_out);

/* ---- ingestion (host only) ----------------------------------------------
 * A ligand set built from PDBQT text: the torsion tree of the ROOT / BRANCH
 * blocks in the reference's pre-order (parse_pdbqt.cpp:440-540), bonds,
 * X-Score types and the intramolecular pair list (model::assign_bonds /
 * assign_types / initialize_pairs, model.cpp:343-640) and the
 * conf-independent torsion count (conf_independent.cpp:36-103). */
typedef struct ud_ligset ud_ligset;
int ud_ligset_create(ud_ligset** out);
int ud_ligset_destroy(ud_ligset* set);
const char* ud_ligset_last_error(const ud_ligset* set);
/* Parse one ligand (a single MODEL); index_out receives its index. */
int ud_ligset_add_pdbqt(ud_ligset* set, const char* text, int32_t* index_out);
/* Add every record of SDF text (records end at "$$$$" lines) as a ligand:
 * parse_sdf_aux / parse_sdf_branch (parse_pdbqt.cpp:376-420, 611-864,
 * 962-1000).  Atoms come from the V2000 atom block; the torsion tree from the
 * fragInfo / torsionInfo data fields (fragment 0 is the root; a torsion line
 * "a b fa fb" joins atom a of fragment fa to atom b of fragment fb), AutoDock
 * types from atomInfo where present, else from the element; a record without
 * fragInfo is rigid.  Bonds, X-Score types, pairs and the torsion count are
 * built as for PDBQT; hydrogens are kept.  Either every record is added or
 * none: first_index_out / count_out (may be NULL) receive the index of the
 * first added ligand and the number of records. */
int ud_ligset_add_sdf(ud_ligset* set, const char* text, int32_t* first_index_out, int32_t* count_out);
int ud_ligset_count(const ud_ligset* set);
/* Fill *batch with arrays owned by the set (valid until the next add or
 * destroy), ready for ud_set_ligands. */
int ud_ligset_batch(ud_ligset* set, ud_ligand_batch* batch);
/* Same, with the pair list of scoring function sf (UD_SF_AD4: the pairs of
 * AutoDock-typed atoms, hydrogens included); ud_ligset_batch is sf Vina. */
int ud_ligset_batch_sf(ud_ligset* set, int sf, ud_ligand_batch* batch);
/* Input coordinates and atom numbers of ligand `lig` in the batch's atom
 * order (for writing poses back out); n_atoms_out first, arrays may be NULL. */
int ud_ligset_atoms(const ud_ligset* set, int lig, int32_t* n_atoms_out, float* xyz_out, int32_t* serial_out);

/* PDBQT text of ligand `lig` with its modes (Vina::get_poses and
 * vina_remarks, vina.cpp:725-830; model::write_model, model.cpp:771-847):
 * per mode MODEL k, the REMARK VINA RESULT / INTER + INTRA / INTER / INTRA /
 * UNBOUND lines, the ligand's input lines with atom coordinates from
 * coords[mode][atom][3] (batch atom order) written into columns 31-54, and
 * ENDMDL.  Writing stops after how_many modes or at the first mode whose
 * energy exceeds the first one's by more than energy_range.  out may be NULL:
 * *len_out receives the text length (without the terminating 0) either way;
 * UD_ERR_CAPACITY if cap is too small. */
int ud_ligset_write_pdbqt(const ud_ligset* set, int lig, int n_modes, const ud_mode* modes, const float* coords,
                          int how_many, double energy_range, char* out, size_t cap, size_t* len_out);

/* SDF text of ligand `lig` (read from SDF) with its modes
 * (Vina::get_sdf_poses_gpu and sdf_remarks, vina.cpp:751-770, 925-970;
 * model::write_sdf_model, model.cpp:758-862): per mode the record's input
 * lines with atom coordinates from coords[mode][atom][3] (batch atom order)
 * written into columns 1-30, the "> <Uni-Dock RESULT>" field (ENERGY /
 * LOWER_BOUND / UPPER_BOUND, INTER + INTRA, INTER, INTRA, UNBOUND) and
 * "$$$$".  Stops like ud_ligset_write_pdbqt.  UD_ERR_STATE for a ligand read
 * from PDBQT (and ud_ligset_write_pdbqt returns it for one read from SDF). */
int ud_ligset_write_sdf(const ud_ligset* set, int lig, int n_modes, const ud_mode* modes, const float* coords,
                        int how_many, double energy_range, char* out, size_t cap, size_t* len_out);

/* Receptor heavy atoms with X-Score types from PDBQT text (the rigid
 * receptor of Vina::set_receptor, typed like the ligand: bonds decide C_P /
 * C_H and donors).  n_out receives the atom count; with capacity too small
 * (or NULL arrays) nothing is written beyond it and UD_ERR_CAPACITY is
 * returned when arrays were given. */
int ud_receptor_from_pdbqt(const char* text, int32_t capacity, int32_t* n_out, float* xyz_out, int32_t* xs_out);
/* Every receptor atom of PDBQT text with its AutoDock type and partial charge
 * (hydrogens included), the input of ud_compute_maps_ad4; capacity rules as
 * above. */
int ud_receptor_ad4_from_pdbqt(const char* text, int32_t capacity, int32_t* n_out, float* xyz_out, int32_t* ad_out,
                               float* charge_out);

/* Seeded random sequential addition of n_atoms points in [0, edge)^3 with a
 * minimum spacing (synthetic receptors); host-only. */
int ud_synth_receptor(int n_atoms, double edge, double min_dist, uint64_t seed, long long max_attempts,
                      double* xyz_out);

#ifdef __cplusplus
}
#endif
#endif /* UNIDOCK_AMD_H */
