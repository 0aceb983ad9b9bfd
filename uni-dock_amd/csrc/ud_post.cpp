// Host post-processing of docking results: pose clustering by RMSD and the
// final (conf-independent) score, as Vina::global_search_gpu does after the
// search (vina.h:232-345).  Pure host code; the device supplies coordinates
// and energy terms.
#include "ud_post.h"

#include <algorithm>
#include <cmath>

namespace ud {

// rmsd_upper_bound over equally ordered heavy-atom coordinates
// (coords.cpp:21-30; model::rmsd_upper_bound, model.cpp:1049-1064)
float rmsd_upper_bound(const float* a, const float* b, int n) {
    float acc = 0.f;  // float accumulation, as the reference's fl
    for (int i = 0; i < 3 * n; i += 3) {
        const float dx = a[i] - b[i], dy = a[i + 1] - b[i + 1], dz = a[i + 2] - b[i + 2];
        acc += dx * dx + dy * dy + dz * dz;
    }
    return n > 0 ? std::sqrt(acc / float(n)) : 0.f;
}

// model::rmsd_lower_bound (model.cpp:1021-1047): for every heavy atom the
// nearest heavy atom of the same element in the other pose, both directions,
// the larger of the two.
float rmsd_lower_bound(const float* a, const float* b, const int* element, int n) {
    auto one_way = [&](const float* x, const float* y) {
        float acc = 0.f;
        for (int i = 0; i < n; ++i) {
            float best = 3.4028234e38f;
            for (int j = 0; j < n; ++j) {
                if (element[i] != element[j]) continue;
                const float dx = x[3 * i] - y[3 * j], dy = x[3 * i + 1] - y[3 * j + 1], dz = x[3 * i + 2] - y[3 * j + 2];
                best = std::min(best, dx * dx + dy * dy + dz * dz);
            }
            acc += best;
        }
        return n > 0 ? std::sqrt(acc / float(n)) : 0.f;
    };
    return std::max(one_way(a, b), one_way(b, a));
}

// add_to_output_container (coords.cpp:42-68) applied to the candidates in
// order; `kept` holds candidate indices sorted by energy (output_container
// sort, stable here so that equal energies keep their insertion order).
void cluster_modes(int n_cand, int n_heavy, const float* xyz, const float* energy, float min_rmsd, int max_size,
                   std::vector<int>& kept) {
    kept.clear();
    const size_t stride = size_t(3) * n_heavy;
    for (int t = 0; t < n_cand; ++t) {
        const float e = energy[t];
        if (e < kEpsFloat && e > -kEpsFloat) continue;  // an energy of exactly 0 marks an error
        int closest = -1;
        float closest_rmsd = 3.4028234e38f;
        for (size_t k = 0; k < kept.size(); ++k) {
            const float r = rmsd_upper_bound(xyz + stride * t, xyz + stride * kept[k], n_heavy);
            if (k == 0 || r < closest_rmsd) {
                closest = int(k);
                closest_rmsd = r;
            }
        }
        if (closest >= 0 && closest_rmsd < min_rmsd) {
            if (e < energy[kept[closest]]) kept[closest] = t;
        } else if (int(kept.size()) < max_size) {
            kept.push_back(t);
        } else if (!kept.empty() && e < energy[kept.back()]) {
            kept.back() = t;
        }
        std::stable_sort(kept.begin(), kept.end(), [&](int a, int b) { return energy[a] < energy[b]; });
    }
}

// num_tors_div (conf_independent.cpp:153-156) with the internal rotational
// weight w (5 w_rot / 0.1 - 1, vina.cpp:404): x / (1 + 0.1 (w + 1) N / 5)
float conf_independent(float x, float num_tors, float rot_internal) {
    const float weight = 0.1f * (rot_internal + 1.f);
    const float y = 1.f + weight * num_tors / 5.f;
    if (std::fabs(x) < kEpsFloat) return 0.f;
    if (std::fabs(y) < kEpsFloat) return (x * y > 0) ? 3.4028234e38f : -3.4028234e38f;
    return x / y;
}

// element class of an X-Score type for same_element (atom_constants.h:84-120):
// the CG carbons are carbons, N_* / O_* (incl. X variants) share their element
int xs_element(int t) {
    switch (t) {
        case 0: case 1: case 19: case 20: case 22: case 23: case 25: case 26: case 28: case 29: return 6;
        case 2: case 3: case 4: case 5: case 33: case 35: return 7;
        case 6: case 7: case 8: case 9: case 32: case 34: return 8;
        default: return 100 + t;
    }
}

// element class of an AutoDock type (ad_type_to_el_type, atom_constants.h
// and atom_type.h): the same classes as xs_element for the same atoms
int ad4_element(int t) {
    switch (t) {
        case 0: case 1: case 26: case 27: case 28: case 29: return 6;   // C A CG0-3
        case 2: case 9: case 32: case 34: return 7;                    // N NA NXA NXD
        case 3: case 10: case 31: case 33: return 8;                   // O OA OXA OXD
        case 5: case 11: return 110;                                   // S SA
        case 13: case 14: case 15: case 16: case 17: return 118;       // metals
        default: return 200 + t;
    }
}

}  // namespace ud
