// A C++ host of the Vina-style class over the C ABI (uni-dock_amd/host):
//   dock RECEPTOR LIGAND CX CY CZ SIZE EXH MAX_STEP SEED GRANULARITY OUT SF N_POSES REFINE
//   ad4 LIGAND MAP_PREFIX EXH MAX_STEP SEED N_POSES OUT   (AutoGrid maps)
//   errors RECEPTOR LIGAND
//   modes MODE...   (--search_mode presets: applied?, exhaustiveness, max_step)
// prints one JSON line (energies of every mode of ligand 0, or the error
// messages caught).
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <vector>

#include "unidock_vina.h"

using unidock_amd::Vina;

static void print_modes(Vina& v, int n_poses) {
    const std::vector<std::vector<double>> e = v.get_poses_energies_gpu(0, n_poses, 1e30);
    std::printf("{\"n_modes\": %zu, \"energies\": [", e.size());
    for (size_t k = 0; k < e.size(); ++k)
        std::printf("%s[%.9g, %.9g, %.9g, %.9g, %.9g]", k ? ", " : "", e[k][0], e[k][1], e[k][2], e[k][3], e[k][4]);
    std::printf("]}\n");
}

int main(int argc, char** argv) {
    if (argc < 2) return 2;
    const std::string cmd = argv[1];
    try {
        if (cmd == "dock" && argc == 16) {
            // REC LIG CX CY CZ SIZE EXH MAX_STEP SEED GRANULARITY OUT SF N_POSES REFINE
            const unsigned long long seed = std::strtoull(argv[10], nullptr, 10);
            const int n_poses = std::atoi(argv[14]);
            Vina v(argv[13], int(seed));
            v.set_receptor(argv[2]);
            v.set_ligand_from_file_gpu({argv[3]});
            const double size = std::atof(argv[7]);
            v.compute_vina_maps(std::atof(argv[4]), std::atof(argv[5]), std::atof(argv[6]), size, size, size,
                                std::atof(argv[11]));
            v.global_search_gpu(std::atoi(argv[8]), n_poses, 1.0, 0, std::atoi(argv[9]), 1, seed, std::atoi(argv[15]));
            v.write_poses_gpu({argv[12]}, n_poses, 1e30);
            print_modes(v, n_poses);
            return 0;
        }
        if (cmd == "ad4" && argc == 9) {
            Vina v("ad4", std::atoi(argv[6]));
            v.set_ligand_from_file_gpu({argv[2]});
            v.load_maps(argv[3]);
            v.global_search_gpu(std::atoi(argv[4]), std::atoi(argv[7]), 1.0, 0, std::atoi(argv[5]), 1,
                                std::strtoull(argv[6], nullptr, 10), 0);
            v.write_poses_gpu({argv[8]}, std::atoi(argv[7]), 1e30);
            print_modes(v, std::atoi(argv[7]));
            return 0;
        }
        if (cmd == "modes") {
            std::printf("{\"modes\": [");
            for (int k = 2; k < argc; ++k) {
                int exh = 8, max_step = 0;  // the values before --search_mode is read
                const bool applied = unidock_amd::search_mode_preset(argv[k], exh, max_step);
                std::printf("%s[%d, %d, %d]", k > 2 ? ", " : "", applied ? 1 : 0, exh, max_step);
            }
            std::printf("]}\n");
            return 0;
        }
        if (cmd == "errors" && argc == 4) {
            std::vector<std::string> msgs;
            auto expect_throw = [&](auto&& f) {
                try {
                    f();
                    msgs.push_back("no error");
                } catch (const std::exception& e) {
                    msgs.push_back(e.what());
                }
            };
            expect_throw([] { Vina v("ad5"); });
            Vina v("vina", 1);
            expect_throw([&] { v.global_search_gpu(8); });
            expect_throw([&] { v.compute_vina_maps(0, 0, 0, 20, 20, 20); });
            v.set_receptor(argv[2]);
            expect_throw([&] { v.compute_vina_maps(0, 0, 0, 20, -1, 20); });
            v.set_ligand_from_file_gpu({argv[3]});
            expect_throw([&] { v.global_search_gpu(8); });
            v.compute_vina_maps(15.19, 53.903, 16.917, 20, 20, 20, 0.375);
            expect_throw([&] { v.global_search_gpu(0); });
            expect_throw([&] { v.get_poses_gpu(0); });
            expect_throw([&] { v.load_maps("nowhere"); });
            expect_throw([&] { v.set_ligand_from_file_gpu({"/nonexistent.pdbqt"}); });
            std::printf("{\"errors\": [");
            for (size_t k = 0; k < msgs.size(); ++k) {
                std::string m = msgs[k];
                for (char& c : m)
                    if (c == '"' || c == '\\' || c == '\n') c = ' ';
                std::printf("%s\"%s\"", k ? ", " : "", m.c_str());
            }
            std::printf("]}\n");
            return 0;
        }
    } catch (const std::exception& e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    std::fprintf(stderr, "usage: vina_host_client dock ... | ad4 ... | errors RECEPTOR LIGAND\n");
    return 2;
}
