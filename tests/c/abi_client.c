/* A C client of include/unidock_amd.h: what a maintainer's C or C++ host of
 * the library compiles against (the reference's host is C++, vina.h:117-381).
 *
 *   abi_client layout                      struct sizes and field offsets as JSON
 *   abi_client dock RECEPTOR LIGAND CX CY CZ SIZE
 *                                          dock a PDBQT ligand into a PDBQT receptor
 *                                          (Vina, exhaustiveness 64, 20 MC steps) and print
 *                                          the best mode as JSON
 */
#include <stddef.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "unidock_amd.h"

#define FIELD(T, f) printf("\"%s.%s\": [%zu, %zu], ", #T, #f, offsetof(T, f), sizeof(((T*)0)->f))

static void layout(void) {
    printf("{");
    printf("\"sizeof.ud_search_params\": %zu, \"sizeof.ud_mode\": %zu, \"sizeof.ud_ligand_batch\": %zu, "
           "\"sizeof.ud_launch_info\": %zu, ",
           sizeof(ud_search_params), sizeof(ud_mode), sizeof(ud_ligand_batch), sizeof(ud_launch_info));
    FIELD(ud_search_params, exhaustiveness); FIELD(ud_search_params, global_steps);
    FIELD(ud_search_params, local_steps); FIELD(ud_search_params, lanes_per_chain);
    FIELD(ud_search_params, seed); FIELD(ud_search_params, hunt_cap); FIELD(ud_search_params, authentic_v);
    FIELD(ud_search_params, temperature); FIELD(ud_search_params, mutation_amplitude);
    FIELD(ud_search_params, chains_per_group); FIELD(ud_search_params, first_ls_f0_zero);
    FIELD(ud_search_params, single_launch);
    FIELD(ud_mode, pose); FIELD(ud_mode, energy); FIELD(ud_mode, inter); FIELD(ud_mode, intra);
    FIELD(ud_mode, conf_independent); FIELD(ud_mode, search_energy); FIELD(ud_mode, rmsd_lb);
    FIELD(ud_mode, rmsd_ub); FIELD(ud_mode, unbound); FIELD(ud_mode, chain);
    FIELD(ud_ligand_batch, n_ligands); FIELD(ud_ligand_batch, atom_offset); FIELD(ud_ligand_batch, node_offset);
    FIELD(ud_ligand_batch, pair_offset); FIELD(ud_ligand_batch, atom_local); FIELD(ud_ligand_batch, atom_xs);
    FIELD(ud_ligand_batch, atom_heavy); FIELD(ud_ligand_batch, node_parent); FIELD(ud_ligand_batch, node_abegin);
    FIELD(ud_ligand_batch, node_aend); FIELD(ud_ligand_batch, node_origin); FIELD(ud_ligand_batch, node_axis);
    FIELD(ud_ligand_batch, pairs); FIELD(ud_ligand_batch, num_tors); FIELD(ud_ligand_batch, atom_ad);
    FIELD(ud_ligand_batch, atom_charge); FIELD(ud_ligand_batch, torsdof);
    FIELD(ud_launch_info, atom_cap); FIELD(ud_launch_info, lanes); FIELD(ud_launch_info, n_ligands);
    FIELD(ud_launch_info, n_workgroups); FIELD(ud_launch_info, pair_buffer); FIELD(ud_launch_info, lds_bytes);
    FIELD(ud_launch_info, max_torsions); FIELD(ud_launch_info, ms);
    printf("\"UD_POSE_FLOATS\": %d, \"UD_GRAD_FLOATS\": %d, \"UD_MAX_ATOMS\": %d, \"UD_MAX_TORSIONS\": %d, "
           "\"UD_MAX_PAIRS\": %d, \"UD_N_BUCKETS\": %d, \"UD_MAX_LAUNCHES\": %d, \"UD_MAX_MODES\": %d, "
           "\"UD_NOALLOC_MAX_POSES\": %d}\n",
           UD_POSE_FLOATS, UD_GRAD_FLOATS, UD_MAX_ATOMS, UD_MAX_TORSIONS, UD_MAX_PAIRS, UD_N_BUCKETS, UD_MAX_LAUNCHES,
           UD_MAX_MODES,
           UD_NOALLOC_MAX_POSES);
}

static char* slurp(const char* path) {
    FILE* f = fopen(path, "rb");
    if (!f) return NULL;
    fseek(f, 0, SEEK_END);
    long n = ftell(f);
    fseek(f, 0, SEEK_SET);
    char* s = (char*)malloc((size_t)n + 1);
    if (s && fread(s, 1, (size_t)n, f) != (size_t)n) { free(s); s = NULL; }
    if (s) s[n] = 0;
    fclose(f);
    return s;
}

#define CHECK(call)                                                                      \
    do {                                                                                 \
        int rc_ = (call);                                                                \
        if (rc_ != UD_OK) {                                                              \
            fprintf(stderr, "%s failed: %d %s\n", #call, rc_, ctx ? ud_last_error(ctx) : ""); \
            return 1;                                                                    \
        }                                                                                \
    } while (0)

/* X-Score type of an atom's grid term (cache.cpp:160-199): G types none,
 * CG carbons the C_H / C_P maps */
static int grid_type(int t) {
    if (t < 0 || t > 35 || t == 21 || t == 24 || t == 27 || t == 30) return -1;
    if (t == 19 || t == 22 || t == 25 || t == 28) return 0;
    if (t == 20 || t == 23 || t == 26 || t == 29) return 1;
    return t;
}

static int dock(const char* rec_path, const char* lig_path, const double center[3], double size) {
    ud_ctx* ctx = NULL;
    char* rec_text = slurp(rec_path);
    char* lig_text = slurp(lig_path);
    if (!rec_text || !lig_text) { fprintf(stderr, "cannot read inputs\n"); return 1; }
    int32_t n_rec = 0;
    if (ud_receptor_from_pdbqt(rec_text, 0, &n_rec, NULL, NULL) != UD_OK || n_rec <= 0) return 1;
    float* rxyz = (float*)malloc(sizeof(float) * 3 * (size_t)n_rec);
    int32_t* rxs = (int32_t*)malloc(sizeof(int32_t) * (size_t)n_rec);
    CHECK(ud_receptor_from_pdbqt(rec_text, n_rec, &n_rec, rxyz, rxs));
    ud_ligset* set = NULL;
    int32_t lig_index = -1;
    CHECK(ud_ligset_create(&set));
    CHECK(ud_ligset_add_pdbqt(set, lig_text, &lig_index));
    ud_ligand_batch batch;
    CHECK(ud_ligset_batch(set, &batch));
    CHECK(ud_create(0, NULL, &ctx));
    /* Vina weights (vina.h:129-132): gauss1, gauss2, repulsion, hydrophobic, hydrogen, glue, rot */
    const double w[7] = {-0.035579, -0.005156, 0.840245, -0.035069, -0.587439, 50.0, 0.05846};
    CHECK(ud_set_scoring(ctx, UD_SF_VINA, w, 7));
    const double sz[3] = {size, size, size};
    int32_t npts[3];
    CHECK(ud_set_box(ctx, center, sz, 0.375, npts));
    int32_t types[36], n_types = 0;
    for (int a = 0; a < batch.atom_offset[batch.n_ligands]; ++a) {
        int t = grid_type(batch.atom_xs[a]), seen = 0;
        for (int k = 0; k < n_types; ++k) seen |= types[k] == t;
        if (t >= 0 && !seen) types[n_types++] = t;
    }
    CHECK(ud_compute_maps(ctx, n_rec, rxyz, rxs, n_types, types, UD_GRID_F16));
    CHECK(ud_set_ligands(ctx, &batch));
    ud_search_params p;
    ud_default_search_params(&p);
    p.exhaustiveness = 64;
    p.global_steps = 20;
    p.seed = 181129;
    CHECK(ud_dock(ctx, &p, NULL, NULL, NULL));
    ud_mode modes[9];
    int32_t n_modes = 0;
    CHECK(ud_postprocess(ctx, 9, 1.0, 0, modes, &n_modes));
    /* the best mode re-scored through the ABI: its components add up */
    float inter = 0.f, intra = 0.f;
    int32_t zero = 0;
    CHECK(ud_score_terms(ctx, 1, &zero, modes[0].pose, &inter, &intra));
    printf("{\"n_modes\": %d, \"energy\": %.6f, \"inter\": %.6f, \"intra\": %.6f, \"rescored_inter\": %.6f, "
           "\"rescored_intra\": %.6f, \"grid\": [%d, %d, %d]}\n",
           n_modes, modes[0].energy, modes[0].inter, modes[0].intra, inter, intra, npts[0], npts[1], npts[2]);
    ud_destroy(ctx);
    ud_ligset_destroy(set);
    free(rxyz); free(rxs); free(rec_text); free(lig_text);
    return 0;
}

int main(int argc, char** argv) {
    if (argc >= 2 && strcmp(argv[1], "layout") == 0) { layout(); return 0; }
    if (argc >= 8 && strcmp(argv[1], "dock") == 0) {
        const double c[3] = {atof(argv[4]), atof(argv[5]), atof(argv[6])};
        return dock(argv[2], argv[3], c, atof(argv[7]));
    }
    fprintf(stderr, "usage: abi_client layout | dock RECEPTOR LIGAND CX CY CZ SIZE\n");
    return 2;
}
