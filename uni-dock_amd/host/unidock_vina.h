// unidock_vina.h -- Vina-style C++ host over the uni-dock_amd C ABI.
//
// The GPU docking surface of the reference's `Vina` class
// (unidock/src/lib/vina.h:117-381): scoring function and weights, receptor
// and batched ligands from PDBQT / SDF files or strings, affinity maps,
// global_search_gpu and the pose writers, with the same method names,
// argument meanings and defaults.  Everything runs through
// include/unidock_amd.h; this file holds no GPU code.
//
// Errors: where the reference prints "ERROR: ..." and exits
// (vina.h:156-166, vina.cpp:491-510), these methods throw std::runtime_error
// with the same message, so that a host can report it and exit as the
// reference's main does; failures inside the library carry its
// ud_last_error text.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "unidock_amd.h"

namespace unidock_amd {

// --search_mode presets (main.cpp:644-656): "fast" sets exhaustiveness 128
// and max_step 20, "balance" / "balanced" 384 and 40, "detail" / "detailed"
// 512 and 40.  Any other name leaves both as given, as the reference does;
// the return value says whether a preset applied.
bool search_mode_preset(const std::string& search_mode, int& exhaustiveness, int& max_step);

class Vina {
public:
    // sf_name: "vina", "vinardo" or "ad4" (vina.h:75-110).  seed 0 draws a
    // random seed (generate_seed).  no_refine: post-processing without the
    // explicit-receptor refinement (--no_refine).  device: HIP device;
    // map_precision: UD_GRID_F32 (the reference's) or UD_GRID_F16.
    explicit Vina(const std::string& sf_name = "vina", int seed = 0, int verbosity = 1, bool no_refine = false,
                  int device = 0, int map_precision = UD_GRID_F32);
    ~Vina();
    Vina(const Vina&) = delete;
    Vina& operator=(const Vina&) = delete;

    int seed() const { return m_seed; }

    // Rigid receptor from a PDBQT file (Vina::set_receptor); flexible
    // residues are not built (a non-empty flex_name throws).
    void set_receptor(const std::string& rigid_name = std::string(), const std::string& flex_name = std::string());

    // Batched ligands (Vina::set_ligand_from_file_gpu /
    // set_ligand_from_string_gpu): PDBQT, or SDF for files ending in .sdf and
    // strings holding "$$$$" records (every record becomes a ligand).
    // Replaces the ligands set before.
    void set_ligand_from_file_gpu(const std::vector<std::string>& ligand_name);
    void set_ligand_from_string_gpu(const std::vector<std::string>& ligand_string);

    // Weights (vina.h:129-138); they apply to maps computed afterwards.
    void set_vina_weights(double weight_gauss1 = -0.035579, double weight_gauss2 = -0.005156,
                          double weight_repulsion = 0.840245, double weight_hydrophobic = -0.035069,
                          double weight_hydrogen = -0.587439, double weight_glue = 50, double weight_rot = 0.05846);
    void set_vinardo_weights(double weight_gauss1 = -0.045, double weight_repulsion = 0.8,
                             double weight_hydrophobic = -0.035, double weight_hydrogen = -0.600,
                             double weight_glue = 50, double weight_rot = 0.05846);
    void set_ad4_weights(double weight_ad4_vdw = 0.1662, double weight_ad4_hb = 0.1209,
                         double weight_ad4_elec = 0.1406, double weight_ad4_dsolv = 0.1322,
                         double weight_glue = 50, double weight_ad4_rot = 0.2983);

    // Vina / Vinardo maps on the GPU (Vina::compute_vina_maps, vina.cpp:491-567):
    // n_voxels = ceil(size / granularity) per axis (made even on request),
    // centred on the centre; the ligands' map types when ligands are set,
    // else every type of the scoring function.
    void compute_vina_maps(double center_x, double center_y, double center_z, double size_x, double size_y,
                           double size_z, double granularity = 0.5, bool force_even_voxels = false);
    // AD4.2 maps built on the GPU from the receptor's AutoDock types and
    // charges, the same geometry (no counterpart: the reference reads AutoGrid
    // files for AD4.2).
    void compute_ad4_maps(double center_x, double center_y, double center_z, double size_x, double size_y,
                          double size_z, double granularity = 0.375, bool force_even_voxels = false);
    // AutoGrid maps <maps>.<type>.map, <maps>.e.map, <maps>.d.map for the ad4
    // scoring function (Vina::load_maps, vina.cpp:570-590); needs the ligands
    // set first (their AutoDock types name the maps).
    void load_maps(const std::string& maps);

    // Monte Carlo search of every ligand set, then clustering, refinement
    // (refine_step > 0 unless no_refine; never under AD4.2) and the final
    // score (Vina::global_search_gpu, vina.h:154-358).  max_step > 0 caps the
    // MC steps; n_poses is capped at UD_MAX_MODES.  local_only optimises the
    // input poses instead (one mode each).  max_evals is accepted and unused,
    // as on the reference's GPU path.
    void global_search_gpu(int exhaustiveness = 8, int n_poses = 20, double min_rmsd = 1.0, int max_evals = 0,
                           int max_step = 0, int num_of_ligands = 1, unsigned long long seed = 181129,
                           int refine_step = 5, bool local_only = false);

    // Modes of ligand ligand_id as PDBQT / SDF text (Vina::get_poses_gpu,
    // get_sdf_poses_gpu): at most how_many, within energy_range of the best.
    std::string get_poses_gpu(int ligand_id, int how_many = 9, double energy_range = 3.0);
    std::string get_sdf_poses_gpu(int ligand_id, int how_many = 9, double energy_range = 3.0);
    // {energy, inter, intra, conf_independent, unbound} per written mode
    // (Vina::get_poses_energies, vina.cpp:683-720).
    std::vector<std::vector<double>> get_poses_energies_gpu(int ligand_id, int how_many = 9,
                                                            double energy_range = 3.0);
    // One output file per ligand (Vina::write_poses_gpu); SDF for names
    // ending in .sdf.
    void write_poses_gpu(const std::vector<std::string>& gpu_output_name, int how_many = 9,
                         double energy_range = 3.0);

    // Tune lanes per chain and LDS layout per size bucket (ud_autotune) at
    // the start of each global_search_gpu; off by default, as the reference
    // launches one fixed configuration.
    void set_autotune(bool on) { m_autotune = on; }

    int num_ligands() const { return m_n_lig; }
    ud_ctx* context() { return m_ctx; }

private:
    void check(int rc, const char* what) const;
    void apply_weights();
    void finish_ligands();
    void box(double cx, double cy, double cz, double sx, double sy, double sz, double granularity, bool even);
    std::vector<float> mode_coords(int ligand_id, const std::vector<ud_mode>& modes) const;
    const std::vector<ud_mode>& modes_of(int ligand_id) const;

    ud_ctx* m_ctx = nullptr;
    ud_ligset* m_set = nullptr;
    int m_sf = UD_SF_VINA;
    std::vector<double> m_weights;
    int m_seed = 0;
    int m_verbosity = 1;
    bool m_no_refine = false;
    bool m_autotune = false;
    int m_precision = UD_GRID_F32;
    // receptor atoms: X-Score types (Vina / Vinardo) or AutoDock types and charges
    std::vector<float> m_rec_xyz, m_rec_q;
    std::vector<int32_t> m_rec_type;
    bool m_receptor_initialized = false;
    bool m_ligand_initialized = false;
    bool m_map_initialized = false;
    int m_n_lig = 0;
    std::vector<int32_t> m_lig_atoms, m_lig_tors;
    std::vector<float> m_lig_origin;
    std::vector<int32_t> m_types;  // map types the ligands need
    std::vector<std::vector<ud_mode>> m_modes;
};

}  // namespace unidock_amd
