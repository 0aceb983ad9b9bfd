// Seeded random sequential addition of receptor atoms in a cube with a
// minimum centre-to-centre spacing (the synthetic receptors of the docking
// workloads).  Host-only helper exported through the C ABI so Python callers
// and the CLI share one generator.
//
// Blind rejection sampling cannot reach 2,000 atoms at 3.0 A spacing in a
// 40 A cube in reasonable time (that is ~93% of the random-sequential-
// addition jamming density).  This version draws each new atom uniformly
// among the still-admissible sites of a 0.2 A lattice (each site jittered
// inside its cell and re-checked), which is the same process without the
// wasted draws.
#include <cmath>
#include <cstdint>
#include <vector>

#include "../../include/unidock_amd.h"

namespace {
// splitmix64 (Steele, Lea, Flood, "Fast splittable pseudorandom number
// generators", OOPSLA 2014)
struct SplitMix {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return double(next() >> 11) * (1.0 / 9007199254740992.0); }
    uint64_t below(uint64_t n) { return uint64_t((unsigned __int128)next() * n >> 64); }
};
}  // namespace

extern "C" int ud_synth_receptor(int n_atoms, double edge, double min_dist, uint64_t seed,
                                 long long max_attempts, double* xyz_out) {
    if (n_atoms <= 0 || edge <= 0 || min_dist <= 0 || !xyz_out) return UD_ERR_INVALID;
    const double h = 0.2;                       // site lattice spacing
    const int ns = int(std::ceil(edge / h));    // sites per axis
    if (double(ns) * ns * ns > 2.0e8) return UD_ERR_CAPACITY;
    const size_t nsite = size_t(ns) * ns * ns;
    // available-site list with O(1) removal
    std::vector<uint32_t> avail(nsite), where(nsite);
    for (size_t i = 0; i < nsite; ++i) { avail[i] = uint32_t(i); where[i] = uint32_t(i); }
    size_t n_avail = nsite;
    auto remove = [&](size_t s) {
        uint32_t w = where[s];
        if (w >= n_avail) return;  // already removed
        uint32_t last = avail[n_avail - 1];
        avail[w] = last; where[last] = w;
        avail[n_avail - 1] = uint32_t(s); where[s] = uint32_t(n_avail - 1);
        --n_avail;
    };
    // cell list of placed atoms for the jitter re-check
    const double cell = min_dist;
    const int nc = int(std::ceil(edge / cell));
    std::vector<std::vector<int>> grid(size_t(nc) * nc * nc);
    SplitMix rng{seed};
    const double md2 = min_dist * min_dist;
    int k = 0;
    long long attempts = 0;
    while (k < n_atoms) {
        if (n_avail == 0) return UD_ERR_CAPACITY;
        if (++attempts > max_attempts) return UD_ERR_CAPACITY;
        size_t s = avail[rng.below(n_avail)];
        int sx = int(s % ns), sy = int((s / ns) % ns), sz = int(s / (size_t(ns) * ns));
        double p[3] = {(sx + rng.uniform()) * h, (sy + rng.uniform()) * h, (sz + rng.uniform()) * h};
        bool inside = p[0] < edge && p[1] < edge && p[2] < edge;
        int c[3];
        for (int i = 0; i < 3; ++i) {
            c[i] = int(p[i] / cell);
            if (c[i] >= nc) c[i] = nc - 1;
        }
        bool ok = inside;
        for (int dx = -1; dx <= 1 && ok; ++dx)
            for (int dy = -1; dy <= 1 && ok; ++dy)
                for (int dz = -1; dz <= 1 && ok; ++dz) {
                    int x = c[0] + dx, y = c[1] + dy, z = c[2] + dz;
                    if (x < 0 || y < 0 || z < 0 || x >= nc || y >= nc || z >= nc) continue;
                    for (int j : grid[(size_t(z) * nc + y) * nc + x]) {
                        double d0 = xyz_out[3 * j] - p[0], d1 = xyz_out[3 * j + 1] - p[1],
                               d2 = xyz_out[3 * j + 2] - p[2];
                        if (d0 * d0 + d1 * d1 + d2 * d2 < md2) { ok = false; break; }
                    }
                }
        if (!ok) {
            // the jittered point failed; the site stays available only if
            // its centre region is still admissible -- drop it when its
            // whole cell is within min_dist of an atom is costly to test, so
            // retry the site a few times by leaving it in the list
            continue;
        }
        xyz_out[3 * k] = p[0]; xyz_out[3 * k + 1] = p[1]; xyz_out[3 * k + 2] = p[2];
        grid[(size_t(c[2]) * nc + c[1]) * nc + c[0]].push_back(k);
        ++k;
        // remove sites whose whole cell lies within min_dist of the new atom
        // (cell corners all closer than min_dist), and sites whose centre is
        // within min_dist - half diagonal
        const double rr = min_dist + h;  // candidate radius
        int lo[3], hi[3];
        for (int i = 0; i < 3; ++i) {
            lo[i] = std::max(0, int((p[i] - rr) / h));
            hi[i] = std::min(ns - 1, int((p[i] + rr) / h));
        }
        for (int z = lo[2]; z <= hi[2]; ++z)
            for (int y = lo[1]; y <= hi[1]; ++y)
                for (int x = lo[0]; x <= hi[0]; ++x) {
                    // farthest point of the site cell from p
                    double fx = std::max(std::fabs(x * h - p[0]), std::fabs((x + 1) * h - p[0]));
                    double fy = std::max(std::fabs(y * h - p[1]), std::fabs((y + 1) * h - p[1]));
                    double fz = std::max(std::fabs(z * h - p[2]), std::fabs((z + 1) * h - p[2]));
                    if (fx * fx + fy * fy + fz * fz < md2) remove((size_t(z) * ns + y) * ns + x);
                }
    }
    return UD_OK;
}
