// unidock_vina.cpp -- Vina-style C++ host over the uni-dock_amd C ABI
// (see unidock_vina.h).  Host code only.
#include "unidock_vina.h"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <random>
#include <set>
#include <sstream>
#include <stdexcept>

namespace unidock_amd {
namespace {

// AutoDock type names (atom_constants.h:46-81), the suffixes of AutoGrid map files
const char* const kAdNames[35] = {"C",  "A",   "N",   "O",   "P",   "S",   "H",   "F",   "I",   "NA",  "OA",  "SA",
                                  "HD", "Mg",  "Mn",  "Zn",  "Ca",  "Fe",  "Cl",  "Br",  "Si",  "At",  "G0",  "G1",
                                  "G2", "G3",  "CG0", "CG1", "CG2", "CG3", "W",   "OXA", "NXA", "OXD", "NXD"};

// X-Score type whose map an atom of type t reads: CG carbons the plain
// carbon maps, G atoms and untyped atoms none (cache.cpp:172-191)
int xs_map_type(int t) {
    if (t < 0 || t >= 36 || t == 21 || t == 24 || t == 27 || t == 30) return -1;
    if (t == 19 || t == 22 || t == 25 || t == 28) return 0;
    if (t == 20 || t == 23 || t == 26 || t == 29) return 1;
    return t;
}
// AutoDock type whose affinity map an atom of type t reads: CG carbons the C
// map, G atoms none (ad4cache.cpp:223-240)
int ad_map_type(int t) {
    if (t < 0 || t >= 35 || (t >= 22 && t <= 25)) return -1;
    if (t >= 26 && t <= 29) return 0;
    return t;
}

std::string read_file(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("ERROR: Could not open " + path + " for reading.");
    std::stringstream ss;
    ss << f.rdbuf();
    return ss.str();
}

bool ends_with(const std::string& s, const std::string& suffix) {
    return s.size() >= suffix.size() && std::equal(suffix.rbegin(), suffix.rend(), s.rbegin(),
                                                   [](char a, char b) {
                                                       return std::tolower(static_cast<unsigned char>(a)) ==
                                                              std::tolower(static_cast<unsigned char>(b));
                                                   });
}

// One AutoGrid map file (ad4cache.cpp:332-389): header lines 4-6 hold
// SPACING, NELEMENTS (even counts) and CENTER; the values follow, x fastest.
struct AutoGridMap {
    double spacing = 0, center[3] = {0, 0, 0};
    int nv[3] = {0, 0, 0};
    std::vector<float> values;
};
AutoGridMap read_autogrid_map(const std::string& path) {
    std::istringstream in(read_file(path));
    AutoGridMap m;
    std::string line, key;
    for (int k = 0; k < 6; ++k) {
        if (!std::getline(in, line)) throw std::runtime_error("ERROR: " + path + ": truncated map header");
        std::istringstream ls(line);
        ls >> key;
        if (k == 3) ls >> m.spacing;
        if (k == 4) ls >> m.nv[0] >> m.nv[1] >> m.nv[2];
        if (k == 5) ls >> m.center[0] >> m.center[1] >> m.center[2];
    }
    for (int d = 0; d < 3; ++d)
        if (m.nv[d] <= 0 || m.nv[d] % 2) throw std::runtime_error("ERROR: " + path + ": number of voxels must be even");
    const size_t n = size_t(m.nv[0] + 1) * (m.nv[1] + 1) * (m.nv[2] + 1);
    m.values.reserve(n);
    double v;
    while (in >> v) m.values.push_back(float(v));
    if (m.values.size() != n) throw std::runtime_error("ERROR: " + path + ": wrong number of map values");
    return m;
}

}  // namespace

bool search_mode_preset(const std::string& search_mode, int& exhaustiveness, int& max_step) {
    struct Preset {
        const char* name;
        int exhaustiveness, max_step;
    };
    static const Preset presets[] = {{"fast", 128, 20},   {"balance", 384, 40},  {"balanced", 384, 40},
                                     {"detail", 512, 40}, {"detailed", 512, 40}};
    for (const Preset& p : presets)
        if (search_mode == p.name) {
            exhaustiveness = p.exhaustiveness;
            max_step = p.max_step;
            return true;
        }
    return false;
}

Vina::Vina(const std::string& sf_name, int seed, int verbosity, bool no_refine, int device, int map_precision)
    : m_verbosity(verbosity), m_no_refine(no_refine), m_precision(map_precision) {
    // generate_seed (vina.h:431): 0 asks for a random seed
    m_seed = seed != 0 ? seed : int(std::random_device{}() & 0x7fffffff);
    if (sf_name == "vina") m_sf = UD_SF_VINA;
    else if (sf_name == "vinardo") m_sf = UD_SF_VINARDO;
    else if (sf_name == "ad4") m_sf = UD_SF_AD4;
    else throw std::runtime_error("ERROR: Scoring function " + sf_name +
                                  " not implemented (choices: vina, vinardo or ad4)");
    if (map_precision != UD_GRID_F32 && map_precision != UD_GRID_F16)
        throw std::runtime_error("ERROR: map precision must be UD_GRID_F32 or UD_GRID_F16");
    int rc = ud_create(device, nullptr, &m_ctx);
    if (rc != UD_OK) throw std::runtime_error("ERROR: could not create a uni-dock_amd context on device " +
                                              std::to_string(device) + " (code " + std::to_string(rc) + ")");
    rc = ud_ligset_create(&m_set);
    if (rc != UD_OK) {
        ud_destroy(m_ctx);
        throw std::runtime_error("ERROR: could not create a ligand set");
    }
    try {
        if (m_sf == UD_SF_VINA) set_vina_weights();
        else if (m_sf == UD_SF_VINARDO) set_vinardo_weights();
        else set_ad4_weights();
    } catch (...) {
        ud_ligset_destroy(m_set);
        ud_destroy(m_ctx);
        throw;
    }
}

Vina::~Vina() {
    if (m_set) ud_ligset_destroy(m_set);
    if (m_ctx) ud_destroy(m_ctx);
}

void Vina::check(int rc, const char* what) const {
    if (rc != UD_OK) throw std::runtime_error(std::string("ERROR: ") + what + ": " + ud_last_error(m_ctx));
}

void Vina::apply_weights() {
    check(ud_set_scoring(m_ctx, m_sf, m_weights.data(), int(m_weights.size())), "set scoring");
    // maps depend on the weights; ligands are typed for the scoring function
    m_map_initialized = false;
    if (m_ligand_initialized) finish_ligands();
}

void Vina::set_vina_weights(double g1, double g2, double rep, double hyd, double hb, double glue, double rot) {
    if (m_sf != UD_SF_VINA) return;  // the weights of another scoring function are stored only by it
    m_weights = {g1, g2, rep, hyd, hb, glue, rot};
    apply_weights();
}

void Vina::set_vinardo_weights(double g1, double rep, double hyd, double hb, double glue, double rot) {
    if (m_sf != UD_SF_VINARDO) return;
    m_weights = {g1, rep, hyd, hb, glue, rot};
    apply_weights();
}

void Vina::set_ad4_weights(double vdw, double hb, double elec, double dsolv, double glue, double rot) {
    if (m_sf != UD_SF_AD4) return;
    m_weights = {vdw, hb, elec, dsolv, glue, rot};
    apply_weights();
}

void Vina::set_receptor(const std::string& rigid_name, const std::string& flex_name) {
    if (!flex_name.empty()) throw std::runtime_error("ERROR: flexible residues are not supported");
    if (rigid_name.empty()) throw std::runtime_error("ERROR: No (rigid) receptor was specified");
    const std::string text = read_file(rigid_name);
    int32_t n = 0;
    const int rc = (m_sf == UD_SF_AD4) ? ud_receptor_ad4_from_pdbqt(text.c_str(), 0, &n, nullptr, nullptr, nullptr)
                                       : ud_receptor_from_pdbqt(text.c_str(), 0, &n, nullptr, nullptr);
    if (rc != UD_OK) throw std::runtime_error("ERROR: could not read receptor " + rigid_name);
    m_rec_xyz.assign(size_t(n) * 3, 0.f);
    m_rec_type.assign(size_t(n), 0);
    m_rec_q.assign(size_t(n), 0.f);
    const int rc2 = (m_sf == UD_SF_AD4)
                        ? ud_receptor_ad4_from_pdbqt(text.c_str(), n, &n, m_rec_xyz.data(), m_rec_type.data(),
                                                     m_rec_q.data())
                        : ud_receptor_from_pdbqt(text.c_str(), n, &n, m_rec_xyz.data(), m_rec_type.data());
    if (rc2 != UD_OK) throw std::runtime_error("ERROR: could not read receptor " + rigid_name);
    m_receptor_initialized = true;
}

// Parse every ligand into a new set; the current ligands stay in place if any fails.
static ud_ligset* parse_ligands(const std::vector<std::string>& texts, const std::vector<bool>& sdf) {
    ud_ligset* set = nullptr;
    if (ud_ligset_create(&set) != UD_OK) throw std::runtime_error("ERROR: could not create a ligand set");
    for (size_t k = 0; k < texts.size(); ++k) {
        const int rc = sdf[k] ? ud_ligset_add_sdf(set, texts[k].c_str(), nullptr, nullptr)
                              : ud_ligset_add_pdbqt(set, texts[k].c_str(), nullptr);
        if (rc != UD_OK) {
            const std::string msg = std::string("ERROR: ") + ud_ligset_last_error(set);
            ud_ligset_destroy(set);
            throw std::runtime_error(msg);
        }
    }
    return set;
}

void Vina::set_ligand_from_file_gpu(const std::vector<std::string>& ligand_name) {
    std::vector<std::string> texts;
    std::vector<bool> sdf;
    for (const std::string& f : ligand_name) {
        texts.push_back(read_file(f));
        sdf.push_back(ends_with(f, ".sdf"));
    }
    ud_ligset* set = parse_ligands(texts, sdf);
    ud_ligset_destroy(m_set);
    m_set = set;
    finish_ligands();
}

void Vina::set_ligand_from_string_gpu(const std::vector<std::string>& ligand_string) {
    std::vector<bool> sdf;
    for (const std::string& t : ligand_string) sdf.push_back(t.find("$$$$") != std::string::npos);
    ud_ligset* set = parse_ligands(ligand_string, sdf);
    ud_ligset_destroy(m_set);
    m_set = set;
    finish_ligands();
}

void Vina::finish_ligands() {
    ud_ligand_batch b;
    if (ud_ligset_batch_sf(m_set, m_sf, &b) != UD_OK) throw std::runtime_error(std::string("ERROR: ") + ud_ligset_last_error(m_set));
    if (b.n_ligands < 1) throw std::runtime_error("ERROR: no ligands were read");
    check(ud_set_ligands(m_ctx, &b), "set ligands");
    m_n_lig = b.n_ligands;
    m_lig_atoms.resize(m_n_lig);
    m_lig_tors.resize(m_n_lig);
    m_lig_origin.resize(size_t(m_n_lig) * 3);
    std::set<int> types;
    for (int l = 0; l < m_n_lig; ++l) {
        m_lig_atoms[l] = b.atom_offset[l + 1] - b.atom_offset[l];
        m_lig_tors[l] = b.node_offset[l + 1] - b.node_offset[l] - 1;
        for (int d = 0; d < 3; ++d) m_lig_origin[3 * l + d] = b.node_origin[3 * b.node_offset[l] + d];
        for (int a = b.atom_offset[l]; a < b.atom_offset[l + 1]; ++a) {
            const int t = (m_sf == UD_SF_AD4) ? ad_map_type(b.atom_ad[a]) : xs_map_type(b.atom_xs[a]);
            if (t >= 0) types.insert(t);
        }
    }
    m_types.assign(types.begin(), types.end());
    m_modes.clear();
    m_ligand_initialized = true;
}

void Vina::box(double cx, double cy, double cz, double sx, double sy, double sz, double granularity, bool even) {
    if (sx <= 0 || sy <= 0 || sz <= 0)
        throw std::runtime_error("ERROR: Grid box dimensions must be greater than 0 Angstrom.");
    if (sx * sy * sz > 27e3 && m_verbosity > 0)
        fprintf(stderr, "WARNING: Search space volume is greater than 27000 Angstrom^3 (See FAQ)\n");
    double size[3] = {sx, sy, sz};
    if (even)
        for (double& s : size) {
            // n_voxels odd -> + 1 (vina.cpp:531-533): the span that gives the even count
            const long nv = long(std::ceil(s / granularity));
            if (nv % 2) s = double(nv + 1) * granularity;
        }
    const double center[3] = {cx, cy, cz};
    check(ud_set_box(m_ctx, center, size, granularity, nullptr), "set box");
}

void Vina::compute_vina_maps(double cx, double cy, double cz, double sx, double sy, double sz, double granularity,
                             bool force_even_voxels) {
    if (m_sf == UD_SF_AD4)
        throw std::runtime_error("ERROR: Cannot compute Vina affinity maps using the AD4 scoring function.");
    if (!m_receptor_initialized)
        throw std::runtime_error(
            "ERROR: Cannot compute Vina or Vinardo affinity maps. The (rigid) receptor was not initialized.");
    box(cx, cy, cz, sx, sy, sz, granularity, force_even_voxels);
    std::vector<int32_t> types = m_types;
    if (!m_ligand_initialized) {
        types.clear();
        for (int t = 0; t < 36; ++t)
            if (xs_map_type(t) == t) types.push_back(t);
    }
    check(ud_compute_maps(m_ctx, int(m_rec_type.size()), m_rec_xyz.data(), m_rec_type.data(), int(types.size()),
                          types.data(), m_precision),
          "compute maps");
    m_map_initialized = true;
}

void Vina::compute_ad4_maps(double cx, double cy, double cz, double sx, double sy, double sz, double granularity,
                            bool force_even_voxels) {
    if (m_sf != UD_SF_AD4) throw std::runtime_error("ERROR: AD4.2 maps need the ad4 scoring function.");
    if (!m_receptor_initialized)
        throw std::runtime_error("ERROR: Cannot compute AD4 maps. The (rigid) receptor was not initialized.");
    box(cx, cy, cz, sx, sy, sz, granularity, force_even_voxels);
    std::vector<int32_t> types = m_types;
    if (!m_ligand_initialized) {
        types.clear();
        for (int t = 0; t < 35; ++t)
            if (ad_map_type(t) == t) types.push_back(t);
    }
    check(ud_compute_maps_ad4(m_ctx, int(m_rec_type.size()), m_rec_xyz.data(), m_rec_type.data(), m_rec_q.data(),
                              int(types.size()), types.data(), m_precision),
          "compute AD4 maps");
    m_map_initialized = true;
}

void Vina::load_maps(const std::string& maps) {
    if (m_sf != UD_SF_AD4) throw std::runtime_error("ERROR: load_maps reads AutoGrid maps of the ad4 scoring function");
    if (!m_ligand_initialized) throw std::runtime_error("ERROR: set the ligands before load_maps (their types name the maps)");
    std::vector<int32_t> types = m_types;
    std::vector<std::string> names;
    for (int t : types) names.push_back(maps + "." + kAdNames[t] + ".map");
    names.push_back(maps + ".e.map");
    names.push_back(maps + ".d.map");
    types.push_back(UD_AD4_MAP_ELEC);
    types.push_back(UD_AD4_MAP_DSOLV);
    std::vector<float> values;
    AutoGridMap first;
    for (size_t k = 0; k < names.size(); ++k) {
        AutoGridMap m = read_autogrid_map(names[k]);
        if (k == 0) {
            first = m;
        } else {
            for (int d = 0; d < 3; ++d)
                if (m.nv[d] != first.nv[d] || std::fabs(m.center[d] - first.center[d]) > 1e-6 ||
                    m.spacing != first.spacing)
                    throw std::runtime_error("ERROR: " + names[k] + ": grid differs from " + names[0]);
        }
        values.insert(values.end(), m.values.begin(), m.values.end());
    }
    const double size[3] = {first.nv[0] * first.spacing, first.nv[1] * first.spacing, first.nv[2] * first.spacing};
    check(ud_set_box(m_ctx, first.center, size, first.spacing, nullptr), "set box");
    check(ud_set_maps(m_ctx, int(types.size()), types.data(), values.data(), m_precision), "set maps");
    m_map_initialized = true;
}

void Vina::global_search_gpu(int exhaustiveness, int n_poses, double min_rmsd, int /*max_evals*/, int max_step,
                             int num_of_ligands, unsigned long long seed, int refine_step, bool local_only) {
    if (!m_ligand_initialized)
        throw std::runtime_error("ERROR: Cannot do the global search. Ligand(s) was(ere) not initialized.");
    if (!m_map_initialized)
        throw std::runtime_error("ERROR: Cannot do the global search. Affinity maps were not initialized.");
    if (exhaustiveness < 1) throw std::runtime_error("ERROR: Exhaustiveness must be 1 or greater");
    if (num_of_ligands > m_n_lig)
        throw std::runtime_error("ERROR: num_of_ligands exceeds the ligands set (" + std::to_string(m_n_lig) + ")");
    if (n_poses < 1) throw std::runtime_error("ERROR: number of poses asked must be greater than zero.");
    const int max_modes = std::min(n_poses, UD_MAX_MODES);
    m_modes.assign(m_n_lig, {});
    if (local_only) {
        // the input pose of each ligand optimised (Vina::optimize,
        // vina.cpp:1379-1430), scored with authentic caps, as one mode
        std::vector<int32_t> pl(m_n_lig);
        std::vector<float> P(size_t(m_n_lig) * UD_POSE_FLOATS, 0.f), e(m_n_lig), inter(m_n_lig), intra(m_n_lig);
        for (int l = 0; l < m_n_lig; ++l) {
            pl[l] = l;
            for (int d = 0; d < 3; ++d) P[size_t(l) * UD_POSE_FLOATS + d] = m_lig_origin[3 * l + d];
            P[size_t(l) * UD_POSE_FLOATS + 3] = 1.f;
        }
        const float v[3] = {1000.f, 1000.f, 1000.f};
        check(ud_optimize(m_ctx, m_n_lig, pl.data(), P.data(), v, 0, e.data(), nullptr), "optimize");
        check(ud_score_terms(m_ctx, m_n_lig, pl.data(), P.data(), inter.data(), intra.data()), "score terms");
        ud_ligand_batch b;
        if (ud_ligset_batch_sf(m_set, m_sf, &b) != UD_OK) throw std::runtime_error("ERROR: ligand batch");
        for (int l = 0; l < m_n_lig; ++l) {
            ud_mode m{};
            std::copy(P.begin() + size_t(l) * UD_POSE_FLOATS, P.begin() + size_t(l + 1) * UD_POSE_FLOATS, m.pose);
            m.inter = inter[l];
            m.intra = intra[l];
            if (m_sf == UD_SF_AD4) {
                // AD4.2: inter + w_rot TORSDOF (conf_independent.cpp:183-187)
                m.conf_independent = float(m_weights.back() * b.torsdof[l]);
                m.energy = m.inter + m.conf_independent;
                m.unbound = -m.intra;
            } else {
                // num_tors_div (conf_independent.cpp:153-156): inter / (1 + w_rot N)
                m.energy = float(m.inter / (1.0 + m_weights.back() * b.num_tors[l]));
                m.conf_independent = m.energy - m.inter;
                m.unbound = m.intra;
            }
            m.search_energy = e[l];
            m.chain = l;
            m_modes[l].push_back(m);
        }
        return;
    }
    // Vina::global_search_gpu step counts (vina.h:192-204)
    long heur = 0;
    for (int l = 0; l < m_n_lig; ++l) heur = std::max<long>(heur, m_lig_atoms[l] + 10L * (6 + m_lig_tors[l]));
    long gs = 70L * 3 * (50 + heur) / 2;
    if (max_step > 0) gs = std::min<long>(gs, max_step);
    ud_search_params p;
    ud_default_search_params(&p);
    p.exhaustiveness = exhaustiveness;
    p.seed = seed;
    if (m_autotune) {
        // lanes per chain and LDS layout per size bucket from 2-step trial docks
        p.global_steps = 2;
        check(ud_autotune(m_ctx, &p, 1 << 20, nullptr), "autotune");
    }
    p.global_steps = int32_t(gs);
    check(ud_dock(m_ctx, &p, nullptr, nullptr, nullptr), "dock");
    std::vector<ud_mode> modes(size_t(m_n_lig) * max_modes);
    std::vector<int32_t> count(m_n_lig);
    const int rs = (m_no_refine || m_sf == UD_SF_AD4) ? 0 : refine_step;
    check(ud_postprocess(m_ctx, max_modes, min_rmsd, rs, modes.data(), count.data()), "post-process");
    for (int l = 0; l < m_n_lig; ++l)
        m_modes[l].assign(modes.begin() + size_t(l) * max_modes, modes.begin() + size_t(l) * max_modes + count[l]);
}

const std::vector<ud_mode>& Vina::modes_of(int ligand_id) const {
    if (ligand_id < 0 || ligand_id >= int(m_modes.size()))
        throw std::runtime_error("ERROR: no poses for ligand " + std::to_string(ligand_id));
    return m_modes[ligand_id];
}

std::vector<float> Vina::mode_coords(int ligand_id, const std::vector<ud_mode>& modes) const {
    const int n = int(modes.size());
    std::vector<int32_t> pl(n, ligand_id);
    std::vector<float> P(size_t(n) * UD_POSE_FLOATS);
    for (int k = 0; k < n; ++k) std::copy(modes[k].pose, modes[k].pose + UD_POSE_FLOATS, P.begin() + size_t(k) * UD_POSE_FLOATS);
    std::vector<float> all(size_t(n) * UD_MAX_ATOMS * 3);
    if (n > 0) check(ud_pose_coords(m_ctx, n, pl.data(), P.data(), all.data()), "pose coordinates");
    // the writers take coords[mode][atom][3] over the ligand's atoms
    const int na = m_lig_atoms[ligand_id];
    std::vector<float> c(size_t(n) * na * 3);
    for (int k = 0; k < n; ++k)
        std::copy(all.begin() + size_t(k) * UD_MAX_ATOMS * 3, all.begin() + size_t(k) * UD_MAX_ATOMS * 3 + size_t(na) * 3,
                  c.begin() + size_t(k) * na * 3);
    return c;
}

std::string Vina::get_poses_gpu(int ligand_id, int how_many, double energy_range) {
    const std::vector<ud_mode>& m = modes_of(ligand_id);
    const std::vector<float> c = mode_coords(ligand_id, m);
    size_t len = 0;
    int rc = ud_ligset_write_pdbqt(m_set, ligand_id, int(m.size()), m.data(), c.data(), how_many, energy_range, nullptr,
                                   0, &len);
    if (rc != UD_OK && rc != UD_ERR_CAPACITY) throw std::runtime_error(std::string("ERROR: ") + ud_ligset_last_error(m_set));
    std::string out(len + 1, '\0');
    rc = ud_ligset_write_pdbqt(m_set, ligand_id, int(m.size()), m.data(), c.data(), how_many, energy_range, &out[0],
                               out.size(), &len);
    if (rc != UD_OK) throw std::runtime_error(std::string("ERROR: ") + ud_ligset_last_error(m_set));
    out.resize(len);
    return out;
}

std::string Vina::get_sdf_poses_gpu(int ligand_id, int how_many, double energy_range) {
    const std::vector<ud_mode>& m = modes_of(ligand_id);
    const std::vector<float> c = mode_coords(ligand_id, m);
    size_t len = 0;
    int rc = ud_ligset_write_sdf(m_set, ligand_id, int(m.size()), m.data(), c.data(), how_many, energy_range, nullptr,
                                 0, &len);
    if (rc != UD_OK && rc != UD_ERR_CAPACITY) throw std::runtime_error(std::string("ERROR: ") + ud_ligset_last_error(m_set));
    std::string out(len + 1, '\0');
    rc = ud_ligset_write_sdf(m_set, ligand_id, int(m.size()), m.data(), c.data(), how_many, energy_range, &out[0],
                             out.size(), &len);
    if (rc != UD_OK) throw std::runtime_error(std::string("ERROR: ") + ud_ligset_last_error(m_set));
    out.resize(len);
    return out;
}

std::vector<std::vector<double>> Vina::get_poses_energies_gpu(int ligand_id, int how_many, double energy_range) {
    if (how_many < 0) throw std::runtime_error("Error: number of poses asked must be greater than zero.");
    if (energy_range < 0) throw std::runtime_error("Error: energy range must be greater than zero.");
    const std::vector<ud_mode>& m = modes_of(ligand_id);
    std::vector<std::vector<double>> out;
    for (size_t k = 0; k < m.size(); ++k) {
        // the writers' stop rule (vina.cpp:700-710)
        if (int(k) >= how_many || !(std::fabs(m[k].energy) < 3.4e38f) || m[k].energy > m[0].energy + energy_range) break;
        out.push_back({m[k].energy, m[k].inter, m[k].intra, m[k].conf_independent, m[k].unbound});
    }
    return out;
}

void Vina::write_poses_gpu(const std::vector<std::string>& gpu_output_name, int how_many, double energy_range) {
    if (int(gpu_output_name.size()) < m_n_lig)
        throw std::runtime_error("ERROR: one output name per ligand is needed");
    for (int l = 0; l < m_n_lig; ++l) {
        if (modes_of(l).empty()) continue;
        const std::string& name = gpu_output_name[l];
        const std::string text = ends_with(name, ".sdf") ? get_sdf_poses_gpu(l, how_many, energy_range)
                                                         : get_poses_gpu(l, how_many, energy_range);
        std::ofstream f(name, std::ios::binary);
        if (!f) throw std::runtime_error("ERROR: Could not open " + name + " for writing.");
        f << text;
    }
}

}  // namespace unidock_amd
