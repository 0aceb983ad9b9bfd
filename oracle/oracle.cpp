// oracle.cpp -- CPU restatement of Uni-Dock's batched pose-search path.
//
// TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py may load this library, and only to check the
// GPU's results (or as the timed CPU baseline).  The product (uni-dock_amd/) never links,
// loads or calls it.
//
// Every function is templated on the real type R and instantiated twice:
//   R = double : the FP64 oracle (parity target for energies/gradients/minima)
//   R = float  : a restatement of the reference's CPU path, which is FP32
//                throughout (`typedef float fl`, unidock/src/lib/common.h:45);
//                this is the "reference analytic derivative" and the CPU
//                baseline that bench.py times.
// Citations are to files of the reference tree (unidock/src/...).
//
// Pinning (what checks each function, see tests/test_oracle.py):
//   potentials        closed-form invariants (gauss(0)=1 at optimum, repulsion 0
//                     for d>0, slope_step endpoints, cutoff), hand values
//   pair tables       recomputation from the closed form at the sample radii,
//                     dor = central difference, fast = midpoint
//   grid build        brute force definition; numpy recomputation on tiny receptors
//   grid eval         affine fields reproduced exactly, central FD inside cells,
//                     out-of-box penalty slope*miss/factor
//   FK                bond lengths / rigid distances preserved, zero pose
//                     reproduces the input, rotation equivariance
//   eval_deriv        central finite differences (h=1e-3 .. 1e-6) in FP64
//   BFGS/line search  quadratic-free checks: monotone decrease, stationarity,
//                     FP32 vs FP64 agreement from identical starts
//   Philox-4x32-10    Random123 published known-answer vectors
//   MC chain          seeded determinism; FP32/FP64 agreement of early steps
//   final score       closed form of num_tors_div
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <atomic>
#include <thread>
#include <vector>

namespace orc {

// ---------------------------------------------------------------- constants
// epsilon_fl / max_fl are the float limits in the reference (common.h:227-230);
// the FP64 oracle keeps the same thresholds because they are part of the
// algorithm (BFGS curvature test, gradient cut-off, curl guard).
static const double kEpsFl = 1.1920928955078125e-07;
static const double kMaxFl = 3.4028234663852886e+38;
static const double kPi = 3.1415926535897931;  // common.h:204 (stored as float there)
// pose stride of the flat pose arrays exchanged with tests and the product's
// C ABI: position 3, quaternion 4, torsions 48 (UD_POSE_FLOATS)
static const int kPose = 7 + 48;

enum { XS_C_H = 0, XS_C_P, XS_N_P, XS_N_D, XS_N_A, XS_N_DA, XS_O_P, XS_O_D, XS_O_A, XS_O_DA,
       XS_S_P, XS_P_P, XS_F_H, XS_Cl_H, XS_Br_H, XS_I_H, XS_Si, XS_At, XS_Met_D,
       XS_C_H_CG0, XS_C_P_CG0, XS_G0, XS_C_H_CG1, XS_C_P_CG1, XS_G1, XS_C_H_CG2, XS_C_P_CG2,
       XS_G2, XS_C_H_CG3, XS_C_P_CG3, XS_G3, XS_W, XS_O_XA, XS_N_XA, XS_O_XD, XS_N_XD,
       XS_SIZE };  // atom_constants.h:84-120

// xs_vdw_radii / xs_vinardo_vdw_radii, looked up BY INDEX exactly like
// xs_radius() (atom_constants.h:313-396).  The table's order after Met_D is
// C_H_CG0, C_P_CG0, C_H_CG1, ... G0..G3, W, which differs from the enum; the
// index lookup is what the reference computes, so it is kept.
static const double kXsRadius[XS_SIZE] = {
    1.9, 1.9, 1.8, 1.8, 1.8, 1.8, 1.7, 1.7, 1.7, 1.7, 2.0, 2.1, 1.5, 1.8, 2.0, 2.2, 2.2, 2.3,
    1.2, 1.9, 1.9, 1.9, 1.9, 1.9, 1.9, 1.9, 1.9, 0.0, 0.0, 0.0, 0.0, 0.0, 1.7, 1.8, 1.7, 1.8};
static const double kXsVinardoRadius[XS_SIZE] = {
    2.0, 2.0, 1.7, 1.7, 1.7, 1.7, 1.6, 1.6, 1.6, 1.6, 2.0, 2.1, 1.5, 1.8, 2.0, 2.2, 2.2, 2.3,
    1.2, 2.0, 2.0, 2.0, 2.0, 2.0, 2.0, 2.0, 2.0, 0.0, 0.0, 0.0, 0.0, 0.0, 1.6, 1.7, 1.6, 1.7};

// atom_constants.h:420-444
static bool xs_is_hydrophobic(int t) {
    return t == XS_C_H || t == XS_F_H || t == XS_Cl_H || t == XS_Br_H || t == XS_I_H;
}
static bool xs_is_acceptor(int t) {
    return t == XS_N_A || t == XS_N_DA || t == XS_O_A || t == XS_O_DA || t == XS_O_XA ||
           t == XS_N_XA;
}
static bool xs_is_donor(int t) {
    return t == XS_N_D || t == XS_N_DA || t == XS_O_D || t == XS_O_DA || t == XS_Met_D ||
           t == XS_N_XD || t == XS_O_XD;
}
static bool xs_h_bond_possible(int a, int b) {
    return (xs_is_donor(a) && xs_is_acceptor(b)) || (xs_is_donor(b) && xs_is_acceptor(a));
}
// potentials.h:41-46, 106-126
static bool is_glue_type(int t) { return t == XS_G0 || t == XS_G1 || t == XS_G2 || t == XS_G3; }
static bool is_glued(int a, int b) {
    const int G[4] = {XS_G0, XS_G1, XS_G2, XS_G3};
    const int CH[4] = {XS_C_H_CG0, XS_C_H_CG1, XS_C_H_CG2, XS_C_H_CG3};
    const int CP[4] = {XS_C_P_CG0, XS_C_P_CG1, XS_C_P_CG2, XS_C_P_CG3};
    for (int k = 0; k < 4; ++k) {
        if (a == G[k] && (b == CH[k] || b == CP[k])) return true;
        if (b == G[k] && (a == CH[k] || a == CP[k])) return true;
    }
    return false;
}

enum { SF_VINA = 0, SF_VINARDO = 1, SF_AD4 = 2 };

// ---------------------------------------------------------------- potentials
// slope_step, potentials.h:30-39
template <class R> R slope_step(R x_bad, R x_good, R x) {
    if (x_bad < x_good) {
        if (x <= x_bad) return R(0);
        if (x >= x_good) return R(1);
    } else {
        if (x >= x_bad) return R(0);
        if (x <= x_good) return R(1);
    }
    return (x - x_bad) / (x_good - x_bad);
}

template <class R> R optimal_distance(int sf, int t1, int t2) {
    // potentials.h:48-51 (vina), 61-64 (vinardo)
    if (is_glue_type(t1) || is_glue_type(t2)) return R(0);
    const double* rad = (sf == SF_VINA) ? kXsRadius : kXsVinardoRadius;
    return R(rad[t1]) + R(rad[t2]);
}

template <class R> R sqr(R x) { return x * x; }

// ScoringFunction::eval(t1, t2, r) = sum_i w_i * term_i(t1, t2, r)
// (scoring_function.h:114-118) with the term list of scoring_function.h:47-70.
// The atom-based overload used for intramolecular tables (precalculate_byatom,
// precalculate.h:211-228) reduces to the same function of the two XS types for
// Vina and Vinardo (potentials.h:137-343), so one routine serves both tables.
template <class R> R sf_eval(int sf, int t1, int t2, R r, const R* w) {
    R acc = 0;
    if (sf == SF_VINA) {
        const R od = optimal_distance<R>(SF_VINA, t1, t2);
        // vina_gaussian(0, 0.5, 8) potentials.h:137-158
        R term = 0;
        if (r < R(8)) term = std::exp(-sqr((r - (od + R(0))) / R(0.5)));
        acc += w[0] * term;
        // vina_gaussian(3, 2.0, 8)
        term = 0;
        if (r < R(8)) term = std::exp(-sqr((r - (od + R(3))) / R(2.0)));
        acc += w[1] * term;
        // vina_repulsion(0, 8) potentials.h:160-181
        term = 0;
        if (r < R(8)) {
            R d = r - (od + R(0));
            term = (d > R(0)) ? R(0) : d * d;
        }
        acc += w[2] * term;
        // vina_hydrophobic(0.5, 1.5, 8) potentials.h:183-207
        term = 0;
        if (r < R(8) && xs_is_hydrophobic(t1) && xs_is_hydrophobic(t2))
            term = slope_step<R>(R(1.5), R(0.5), r - od);
        acc += w[3] * term;
        // vina_non_dir_h_bond(-0.7, 0, 8) potentials.h:209-243
        term = 0;
        if (r < R(8) && xs_h_bond_possible(t1, t2)) {
            bool x_type = (t1 >= XS_O_XA && t1 <= XS_N_XD) || (t2 >= XS_O_XA && t2 <= XS_N_XD);
            term = slope_step<R>(R(0), R(-0.7), r - od);
            if (x_type) term = R(10.0) * term;
        }
        acc += w[4] * term;
        // linearattraction(20) potentials.h:487-507
        term = 0;
        if (r < R(20) && is_glued(t1, t2)) term = r;
        acc += w[5] * term;
    } else {
        const R od = optimal_distance<R>(SF_VINARDO, t1, t2);
        // vinardo_gaussian(0, 0.8, 8) potentials.h:246-269
        R term = 0;
        if (r < R(8)) term = std::exp(-sqr((r - (od + R(0))) / R(0.8)));
        acc += w[0] * term;
        // vinardo_repulsion(0, 8) potentials.h:271-292
        term = 0;
        if (r < R(8)) {
            R d = r - (od + R(0));
            term = (d > R(0)) ? R(0) : d * d;
        }
        acc += w[1] * term;
        // vinardo_hydrophobic(0, 2.5, 8) potentials.h:294-318
        term = 0;
        if (r < R(8) && xs_is_hydrophobic(t1) && xs_is_hydrophobic(t2))
            term = slope_step<R>(R(2.5), R(0), r - od);
        acc += w[2] * term;
        // vinardo_non_dir_h_bond(-0.6, 0, 8) potentials.h:320-343
        term = 0;
        if (r < R(8) && xs_h_bond_possible(t1, t2)) term = slope_step<R>(R(0), R(-0.6), r - od);
        acc += w[3] * term;
        // linearattraction(20)
        term = 0;
        if (r < R(20) && is_glued(t1, t2)) term = r;
        acc += w[4] * term;
    }
    return acc;
}

// ---------------------------------------------------------------- AD4.2
// AutoDock 4.2 atom kinds, by AD type index (atom_constants.h:46-81, 155-194):
// vdW radius, vdW depth, H-bond depth, H-bond radius, solvation parameter,
// atomic volume.
enum { AD_C = 0, AD_A, AD_N, AD_O, AD_P, AD_S, AD_H, AD_F, AD_I, AD_NA, AD_OA, AD_SA, AD_HD, AD_Mg, AD_Mn,
       AD_Zn, AD_Ca, AD_Fe, AD_Cl, AD_Br, AD_Si, AD_At, AD_G0, AD_G1, AD_G2, AD_G3, AD_CG0, AD_CG1, AD_CG2,
       AD_CG3, AD_W, AD_OXA, AD_NXA, AD_OXD, AD_NXD, AD_SIZE };
struct AdKind { double radius, depth, hb_depth, hb_r, solv, vol; };
static const AdKind kAd[AD_SIZE] = {
    {2.000, 0.150, 0, 0, -0.00143, 33.5103}, {2.000, 0.150, 0, 0, -0.00052, 33.5103},  // C, A
    {1.750, 0.160, 0, 0, -0.00162, 22.4493}, {1.600, 0.200, 0, 0, -0.00251, 17.1573},  // N, O
    {2.100, 0.200, 0, 0, -0.00110, 38.7924}, {2.000, 0.200, 0, 0, -0.00214, 33.5103},  // P, S
    {1.000, 0.020, 0, 0, 0.00051, 0.0},      {1.545, 0.080, 0, 0, -0.00110, 15.4480},  // H, F
    {2.360, 0.550, 0, 0, -0.00110, 55.0585},                                            // I
    {1.750, 0.160, -5, 1.9, -0.00162, 22.4493}, {1.600, 0.200, -5, 1.9, -0.00251, 17.1573},  // NA, OA
    {2.000, 0.200, -1, 2.5, -0.00214, 33.5103}, {1.000, 0.020, 1, 0, 0.00051, 0.0},          // SA, HD
    {0.650, 0.875, 0, 0, -0.00110, 1.560},  {0.650, 0.875, 0, 0, -0.00110, 2.140},   // Mg, Mn
    {0.740, 0.550, 0, 0, -0.00110, 1.700},  {0.990, 0.550, 0, 0, -0.00110, 2.770},   // Zn, Ca
    {0.650, 0.010, 0, 0, -0.00110, 1.840},  {2.045, 0.276, 0, 0, -0.00110, 35.8235},  // Fe, Cl
    {2.165, 0.389, 0, 0, -0.00110, 42.5661}, {2.300, 0.200, 0, 0, -0.00143, 50.9650},  // Br, Si
    {2.400, 0.550, 0, 0, -0.00110, 57.9058},                                            // At
    {0, 0, 0, 0, 0, 0}, {0, 0, 0, 0, 0, 0}, {0, 0, 0, 0, 0, 0}, {0, 0, 0, 0, 0, 0},     // G0..G3
    {2.000, 0.150, 0, 0, -0.00143, 33.5103}, {2.000, 0.150, 0, 0, -0.00143, 33.5103},  // CG0, CG1
    {2.000, 0.150, 0, 0, -0.00143, 33.5103}, {2.000, 0.150, 0, 0, -0.00143, 33.5103},  // CG2, CG3
    {0, 0, 0, 0, 0, 0},                                                                 // W
    {1.600, 0.200, 0, 0, -0.00251, 17.1573}, {1.750, 0.160, 0, 0, -0.00162, 22.4493},  // OXA, NXA
    {1.600, 0.200, 0, 0, -0.00251, 17.1573}, {1.750, 0.160, 0, 0, -0.00162, 22.4493},  // OXD, NXD
};
static const double kAd4Cutoff = 20.48;   // scoring_function.h:75-83
static const double kAd4VdwCutoff = 8.0, kAd4Cap = 100000.0, kAd4Smoothing = 0.5;
static const double kAd4ElecCap = 100.0, kAd4Sigma = 3.6, kAd4SolvQ = 0.01097;

template <class R> R ipow(R x, int n) { R y = 1; for (int i = 0; i < n; ++i) y *= x; return y; }
// smoothen (potentials.h:67-79)
template <class R> R ad4_smoothen(R r, R rij, R smoothing) {
    smoothing *= R(0.5);
    if (r > rij + smoothing) return r - smoothing;
    if (r < rij - smoothing) return r + smoothing;
    return rij;
}
// ad4_vdw::eval (potentials.h:421-447): 12-6 with smoothing, not for H-bond pairs
template <class R> R ad4_vdw(int t1, int t2, R r) {
    if (r >= R(kAd4VdwCutoff)) return 0;
    const R hb_depth = R(kAd[t1].hb_depth) * R(kAd[t2].hb_depth);
    const R rij = R(kAd[t1].radius) + R(kAd[t2].radius);
    const R depth = std::sqrt(R(kAd[t1].depth) * R(kAd[t2].depth));
    if (hb_depth < 0) return 0;
    r = ad4_smoothen(r, rij, R(kAd4Smoothing));
    const R c12 = ipow(rij, 12) * depth, c6 = ipow(rij, 6) * depth * R(2);
    const R r6 = ipow(r, 6), r12 = ipow(r, 12);
    if (r12 > R(kEpsFl) && r6 > R(kEpsFl)) return std::min(R(kAd4Cap), c12 / r12 - c6 / r6);
    return R(kAd4Cap);
}
// ad4_hb::eval (potentials.h:454-480): 12-10 with smoothing, H-bond pairs only
template <class R> R ad4_hb(int t1, int t2, R r) {
    if (r >= R(kAd4VdwCutoff)) return 0;
    const R rij = R(kAd[t1].hb_r) + R(kAd[t2].hb_r);
    const R hb_depth = R(kAd[t1].hb_depth) * R(kAd[t2].hb_depth);
    if (hb_depth >= 0) return 0;
    r = ad4_smoothen(r, rij, R(kAd4Smoothing));
    const R c12 = ipow(rij, 12) * -hb_depth * R(10) / R(2.0), c10 = ipow(rij, 10) * -hb_depth * R(12) / R(2.0);
    const R r10 = ipow(r, 10), r12 = ipow(r, 12);
    if (r12 > R(kEpsFl) && r10 > R(kEpsFl)) return std::min(R(kAd4Cap), c12 / r12 - c10 / r10);
    return R(kAd4Cap);
}
// ad4_electrostatic::eval (potentials.h:350-366) without the charges:
// 332 min(cap, 1 / (r eps(r))) with the sigmoidal distance-dependent dielectric
template <class R> R ad4_elec_r(R r) {
    if (r >= R(kAd4Cutoff)) return 0;
    const R B = R(78.4) + R(8.5525);
    const R lB = -B * R(0.003627);
    const R diel = R(-8.5525) + (B / (R(1) + R(7.7839) * std::exp(lB * r)));
    if (r < R(kEpsFl)) return R(332.0) * R(kAd4ElecCap) / diel;
    return R(332.0) * std::min(R(kAd4ElecCap), R(1.0) / (r * diel));
}
// the Gaussian of ad4_solvation::eval (potentials.h:380-400)
template <class R> R ad4_solv_r(R r) {
    if (r >= R(kAd4Cutoff)) return 0;
    return std::exp(R(-0.5) * sqr(r / R(kAd4Sigma)));
}
// ad4_solvation's coefficient of the Gaussian for atoms of types t1, t2 and
// charges q1, q2 (charge dependent, potentials.h:384-396)
template <class R> R ad4_solv_coef(int t1, R q1, int t2, R q2) {
    return (R(kAd[t1].solv) + R(kAd4SolvQ) * std::abs(q1)) * R(kAd[t2].vol) +
           (R(kAd[t2].solv) + R(kAd4SolvQ) * std::abs(q2)) * R(kAd[t1].vol);
}
// ScoringFunction::eval(atom a, atom b, r) for AD4.2 (scoring_function.h:74-85):
// w = (vdw, hb, elec, dsolv); the glue term only joins G / CG closure atoms,
// whose pairs are never ligand pairs.
template <class R> R sf_eval_ad4(int t1, R q1, int t2, R q2, R r, const R* w) {
    R acc = w[0] * ad4_vdw(t1, t2, r);
    acc += w[1] * ad4_hb(t1, t2, r);
    acc += w[2] * q1 * q2 * ad4_elec_r(r);
    acc += w[3] * ad4_solv_coef(t1, q1, t2, q2) * ad4_solv_r(r);
    return acc;
}
// AD type remapping of ad4cache::eval_deriv (ad4cache.cpp:223-240): G types
// carry no grid term, CG types use the C map.
static int ad4_grid_type(int t) {
    if (t < 0 || t >= AD_SIZE) return -1;
    if (t >= AD_G0 && t <= AD_G3) return -1;
    if (t >= AD_CG0 && t <= AD_CG3) return AD_C;
    return t;
}
static const int kMapElec = AD_SIZE, kMapDsolv = AD_SIZE + 1;  // m_grids[AD_TYPE_SIZE (+1)]

// Number of distance samples: sz(factor * max_cutoff_sqr) + 3 (precalculate.h:142-145)
static const int kFactor = 32;
static const int kTableN = 32 * 64 + 3;  // 2051
static int table_n(int sf) {  // 13424 for the 20.48 A cutoff of AD4.2
    if (sf != SF_AD4) return kTableN;
    const float c = float(kAd4Cutoff);
    return int(float(kFactor) * (c * c)) + 3;
}

// precalculate_element::init_from_smooth_fst (precalculate.h:66-85) applied to
// smooth[].first = min(v, sf.eval(t1, t2, rs[i])) with v = max_fl
// (precalculate.h:160-171).  rs[i] = sqrt(i / factor) (precalculate.h:191-196).
template <class R> struct PairTable {
    std::vector<R> e, dor, fast;
    void build(int sf, int t1, int t2, const R* w) {
        build_from(kTableN, [&](R r) { return sf_eval<R>(sf, t1, t2, r, w); });
    }
    // precalculate_byatom (precalculate.h:206-241): the table of one atom
    // pair, sampled from the atom-based ScoringFunction::eval
    void build_ad4(int t1, R q1, int t2, R q2, const R* w) {
        build_from(table_n(SF_AD4), [&](R r) { return sf_eval_ad4<R>(t1, q1, t2, q2, r, w); });
    }
    template <class F> void build_from(int n, F f) {
        e.assign(n, 0); dor.assign(n, 0); fast.assign(n, 0);
        std::vector<R> rs(n);
        for (int i = 0; i < n; ++i) rs[i] = std::sqrt(R(i) / R(kFactor));
        for (int i = 0; i < n; ++i) {
            R v = f(rs[i]);
            e[i] = std::min<R>(R(kMaxFl), v);
        }
        for (int i = 0; i < n; ++i) {
            if (i == 0 || i == n - 1)
                dor[i] = 0;
            else {
                R delta = rs[i + 1] - rs[i - 1];
                R r = rs[i];
                dor[i] = (e[i + 1] - e[i - 1]) / (delta * r);
            }
            R f1 = e[i];
            R f2 = (i + 1 >= n) ? R(0) : e[i + 1];
            fast[i] = (f2 + f1) / R(2);
        }
    }
    // precalculate_element::eval_fast, precalculate.h:40-47
    R eval_fast(R r2) const { return fast[int(R(kFactor) * r2)]; }
    // precalculate_element::eval_deriv, precalculate.h:48-65
    void eval_deriv(R r2, R& oe, R& odor) const {
        R rf = R(kFactor) * r2;
        int i1 = int(rf);
        int i2 = i1 + 1;
        R rem = rf - R(i1);
        oe = e[i1] + rem * (e[i2] - e[i1]);
        odor = dor[i1] + rem * (dor[i2] - dor[i1]);
    }
};

// triangular_matrix_index_permissive (triangular_matrix_index.h:28-37)
static int tri_index(int i, int j) { return (i <= j) ? i + j * (j + 1) / 2 : j + i * (i + 1) / 2; }

// ---------------------------------------------------------------- vectors
template <class R> struct V3 {
    R x[3];
    V3() { x[0] = x[1] = x[2] = 0; }
    V3(R a, R b, R c) { x[0] = a; x[1] = b; x[2] = c; }
    R& operator[](int i) { return x[i]; }
    const R& operator[](int i) const { return x[i]; }
};
template <class R> V3<R> operator+(const V3<R>& a, const V3<R>& b) { return V3<R>(a[0] + b[0], a[1] + b[1], a[2] + b[2]); }
template <class R> V3<R> operator-(const V3<R>& a, const V3<R>& b) { return V3<R>(a[0] - b[0], a[1] - b[1], a[2] - b[2]); }
template <class R> V3<R> operator*(R s, const V3<R>& a) { return V3<R>(s * a[0], s * a[1], s * a[2]); }
template <class R> R dot(const V3<R>& a, const V3<R>& b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
template <class R> V3<R> cross(const V3<R>& a, const V3<R>& b) {
    return V3<R>(a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]);
}
template <class R> R norm_sqr(const V3<R>& a) { return sqr(a[0]) + sqr(a[1]) + sqr(a[2]); }

// quaternion (a, b, c, d) = a + bi + cj + dk, Hamilton product.
template <class R> struct Q {
    R a, b, c, d;
};
template <class R> Q<R> qmul(const Q<R>& p, const Q<R>& q) {
    return Q<R>{p.a * q.a - p.b * q.b - p.c * q.c - p.d * q.d,
                p.a * q.b + p.b * q.a + p.c * q.d - p.d * q.c,
                p.a * q.c - p.b * q.d + p.c * q.a + p.d * q.b,
                p.a * q.d + p.b * q.c - p.c * q.b + p.d * q.a};
}
// Matrix stored row-major m[3*i + j].
// quaternion_to_r3, quaternion.cpp:77-114
template <class R> void q_to_mat(const Q<R>& q, R* m) {
    const R aa = q.a * q.a, ab = q.a * q.b, ac = q.a * q.c, ad = q.a * q.d;
    const R bb = q.b * q.b, bc = q.b * q.c, bd = q.b * q.d;
    const R cc = q.c * q.c, cd = q.c * q.d, dd = q.d * q.d;
    m[0] = (aa + bb - cc - dd); m[1] = R(2) * (-ad + bc); m[2] = R(2) * (ac + bd);
    m[3] = R(2) * (ad + bc); m[4] = (aa - bb + cc - dd); m[5] = R(2) * (-ab + cd);
    m[6] = R(2) * (-ac + bd); m[7] = R(2) * (ab + cd); m[8] = (aa - bb - cc + dd);
}
template <class R> V3<R> matvec(const R* m, const V3<R>& v) {
    return V3<R>(m[0] * v[0] + m[1] * v[1] + m[2] * v[2], m[3] * v[0] + m[4] * v[1] + m[5] * v[2],
                 m[6] * v[0] + m[7] * v[1] + m[8] * v[2]);
}
// normalize_angle, common.h:260-281
template <class R> void normalize_angle(R& x) {
    const R pi = R(kPi);
    if (x > R(3) * pi) {
        R n = (x - pi) / (R(2) * pi);
        x -= R(2) * pi * std::ceil(n);
        normalize_angle(x);
    } else if (x < -R(3) * pi) {
        R n = (-x - pi) / (R(2) * pi);
        x += R(2) * pi * std::ceil(n);
        normalize_angle(x);
    } else if (x > pi) {
        x -= R(2) * pi;
    } else if (x < -pi) {
        x += R(2) * pi;
    }
}
template <class R> R normalized_angle(R x) { normalize_angle(x); return x; }
// angle_to_quaternion(axis, angle), quaternion.cpp:37-47
template <class R> Q<R> axis_angle_q(const V3<R>& axis, R angle) {
    normalize_angle(angle);
    R c = std::cos(angle / R(2));
    R s = std::sin(angle / R(2));
    return Q<R>{c, s * axis[0], s * axis[1], s * axis[2]};
}
// angle_to_quaternion(rotation), quaternion.cpp:49-59
template <class R> Q<R> rotation_q(const V3<R>& rot) {
    R angle = std::sqrt(norm_sqr(rot));
    if (angle > R(kEpsFl)) {
        V3<R> axis(rot[0] / angle, rot[1] / angle, rot[2] / angle);
        return axis_angle_q(axis, angle);
    }
    return Q<R>{1, 0, 0, 0};
}
// quaternion_normalize_approx, quaternion.h:105-117
template <class R> void q_normalize_approx(Q<R>& q) {
    const R s = q.a * q.a + q.b * q.b + q.c * q.c + q.d * q.d;
    if (std::abs(s - R(1)) < R(1e-6)) return;
    const R a = std::sqrt(s);
    const R inv = R(1) / a;
    q.a *= inv; q.b *= inv; q.c *= inv; q.d *= inv;
}
// quaternion_increment, quaternion.cpp:128-133
template <class R> void q_increment(Q<R>& q, const V3<R>& rot) {
    q = qmul(rotation_q(rot), q);
    q_normalize_approx(q);
}

// ---------------------------------------------------------------- grid
// grid::init (grid.cpp:24-37), compute_vina_maps geometry (vina.cpp:527-539)
template <class R> struct Grid {
    int n[3] = {0, 0, 0};   // sample points per axis (= n_voxels + 1)
    R init[3], factor[3], factor_inv[3], dim_minus_1[3];
    R hi[3];                // box end (grid_dims end, vina.cpp:536-538)
    std::vector<std::vector<R>> maps;  // one per XS type (index = type), empty if absent
    void set_geometry(const double center[3], const double size[3], double granularity) {
        for (int i = 0; i < 3; ++i) {
            long nv = long(std::ceil(size[i] / granularity));
            R real_span = R(granularity * double(nv));
            R begin = R(center[i]) - real_span / R(2);
            R end = begin + real_span;
            n[i] = int(nv) + 1;
            init[i] = begin;
            hi[i] = end;
            R range = end - begin;
            dim_minus_1[i] = R(n[i] - 1);
            factor[i] = dim_minus_1[i] / range;
            factor_inv[i] = R(1) / factor[i];
        }
        maps.assign(std::max<int>(XS_SIZE, AD_SIZE + 2), std::vector<R>());
    }
    R& at(int t, int x, int y, int z) { return maps[t][x + n[0] * (y + n[1] * size_t(z))]; }
    R at(int t, int x, int y, int z) const { return maps[t][x + n[0] * (y + n[1] * size_t(z))]; }
    // grid::index_to_argument (grid.h:45-48)
    V3<R> point(int x, int y, int z) const {
        return V3<R>(init[0] + factor_inv[0] * R(x), init[1] + factor_inv[1] * R(y),
                     init[2] + factor_inv[2] * R(z));
    }
};

// curl(e, deriv, v) and curl(e, v), curl.h:28-43
template <class R> void curl(R& e, V3<R>& deriv, R v) {
    if (e > R(0) && v < R(0.1) * R(kMaxFl)) {
        R tmp = (v < R(kEpsFl)) ? R(0) : (v / (v + e));
        e *= tmp;
        R t2 = tmp * tmp;
        deriv = t2 * deriv;
    }
}
template <class R> void curl(R& e, R v) {
    if (e > R(0) && v < R(0.1) * R(kMaxFl)) {
        R tmp = (v < R(kEpsFl)) ? R(0) : (v / (v + e));
        e *= tmp;
    }
}

// grid::evaluate_aux (grid.cpp:39-129).  deriv == nullptr -> energy only.
template <class R> R grid_eval(const Grid<R>& g, int t, const V3<R>& loc, R slope, R v, V3<R>* deriv) {
    R s[3], miss[3];
    int region[3], a[3];
    for (int i = 0; i < 3; ++i) {
        s[i] = (loc[i] - g.init[i]) * g.factor[i];
        miss[i] = 0;
        // `!(s >= 0)` also routes a NaN coordinate (an overflowed FP32 pose)
        // to the clamped branch instead of an undefined int(NaN) index
        if (!(s[i] >= R(0))) {
            miss[i] = -s[i]; region[i] = -1; a[i] = 0; s[i] = 0;
        } else if (s[i] >= g.dim_minus_1[i]) {
            miss[i] = s[i] - g.dim_minus_1[i]; region[i] = 1; a[i] = g.n[i] - 2; s[i] = 1;
        } else {
            region[i] = 0; a[i] = int(s[i]); s[i] -= R(a[i]);
        }
    }
    const R penalty = slope * (miss[0] * g.factor_inv[0] + miss[1] * g.factor_inv[1] + miss[2] * g.factor_inv[2]);
    const int x0 = a[0], y0 = a[1], z0 = a[2], x1 = x0 + 1, y1 = y0 + 1, z1 = z0 + 1;
    const R f000 = g.at(t, x0, y0, z0), f100 = g.at(t, x1, y0, z0), f010 = g.at(t, x0, y1, z0),
            f110 = g.at(t, x1, y1, z0), f001 = g.at(t, x0, y0, z1), f101 = g.at(t, x1, y0, z1),
            f011 = g.at(t, x0, y1, z1), f111 = g.at(t, x1, y1, z1);
    const R x = s[0], y = s[1], z = s[2];
    const R mx = R(1) - x, my = R(1) - y, mz = R(1) - z;
    R f = f000 * mx * my * mz + f100 * x * my * mz + f010 * mx * y * mz + f110 * x * y * mz +
          f001 * mx * my * z + f101 * x * my * z + f011 * mx * y * z + f111 * x * y * z;
    if (deriv) {
        const R x_g = f000 * (-1) * my * mz + f100 * 1 * my * mz + f010 * (-1) * y * mz + f110 * 1 * y * mz +
                      f001 * (-1) * my * z + f101 * 1 * my * z + f011 * (-1) * y * z + f111 * 1 * y * z;
        const R y_g = f000 * mx * (-1) * mz + f100 * x * (-1) * mz + f010 * mx * 1 * mz + f110 * x * 1 * mz +
                      f001 * mx * (-1) * z + f101 * x * (-1) * z + f011 * mx * 1 * z + f111 * x * 1 * z;
        const R z_g = f000 * mx * my * (-1) + f100 * x * my * (-1) + f010 * mx * y * (-1) + f110 * x * y * (-1) +
                      f001 * mx * my * 1 + f101 * x * my * 1 + f011 * mx * y * 1 + f111 * x * y * 1;
        V3<R> grad(x_g, y_g, z_g);
        curl(f, grad, v);
        for (int i = 0; i < 3; ++i) {
            R ge = (region[i] == 0) ? grad[i] : R(0);
            (*deriv)[i] = g.factor[i] * ge + slope * R(region[i]);
        }
        return f + penalty;
    }
    curl(f, v);
    return f + penalty;
}

// ---------------------------------------------------------------- ligand
// A ligand in the flat layout shared with the C-ABI: atoms grouped by rigid
// node (contiguous ranges), nodes in depth-first pre-order (parents first),
// local atom coordinates relative to the owning node's origin in the input
// frame (parse_pdbqt.cpp:206-211), relative origin/axis per node
// (tree.h:121-131), intramolecular pairs (model.cpp:592-630).
template <class R> struct Ligand {
    int n_atoms = 0, n_nodes = 0, n_pairs = 0;
    std::vector<V3<R>> local;       // per atom
    std::vector<int> xs;            // per atom, -1 = untyped (hydrogen etc.)
    std::vector<int> heavy;         // per atom
    std::vector<int> node_of;       // per atom
    std::vector<int> parent, abegin, aend;  // per node
    std::vector<V3<R>> rel_origin, rel_axis, origin0;  // per node (origin0: input-frame origin)
    std::vector<int> pa, pb, ptp;   // per pair
    R num_tors = 0;
    // AD4.2: AutoDock types, partial charges, TORSDOF and one table per pair
    std::vector<int> ad;
    std::vector<R> q;
    R torsdof = 0;
    std::vector<PairTable<R>> ptab;
    int n_tors() const { return n_nodes - 1; }
};

// Pose: position[3], orientation q[4] (a,b,c,d), torsions[T] -> flat 7+T.
template <class R> struct Pose {
    V3<R> pos;
    Q<R> q{1, 0, 0, 0};
    std::vector<R> tors;
};

template <class R> struct Ctx {
    int sf = SF_VINA;
    R w[7];
    std::vector<PairTable<R>> tables;  // indexed by tri_index(t1, t2), built lazily
    std::vector<char> have;
    Grid<R> grid;
    R slope = R(1e6);  // vina.cpp:515
    std::vector<Ligand<R>> ligs;
    // typed receptor atoms of build_maps, for the explicit-receptor scoring
    std::vector<V3<R>> rec;
    std::vector<int> rec_t;
    bool use_nc = false;   // eval_deriv scores against rec (non_cache) instead of the maps
    R nc_slope = R(1e6);
    // first line search of every BFGS against f0 = 0 (the reference GPU's
    // bfgs_warp, warp_ops.cuh:1062-1076) instead of f(x) (bfgs.h:107)
    bool ls_f0_zero = false;
    const PairTable<R>& table(int t1, int t2) {
        int k = tri_index(t1, t2);
        if (!have[k]) { tables[k].build(sf, t1, t2, w); have[k] = 1; }
        return tables[k];
    }
    void init(int sf_, const double* weights) {
        sf = sf_;
        int nw = (sf == SF_VINA) ? 7 : 6;
        for (int i = 0; i < 7; ++i) w[i] = 0;
        if (sf == SF_AD4) {  // vdw, hb, elec, dsolv, glue, rot used as given (vina.cpp:435-450)
            for (int i = 0; i < 6; ++i) w[i] = R(weights[i]);
            int ntp = XS_SIZE * (XS_SIZE + 1) / 2;
            tables.assign(ntp, PairTable<R>());
            have.assign(ntp, 0);
            return;
        }
        // weights as given to set_vina_weights/set_vinardo_weights; the last
        // is weight_rot, turned into the internal 5*w_rot/0.1 - 1 (vina.cpp:404, 425)
        for (int i = 0; i < nw - 1; ++i) w[i] = R(weights[i]);
        w[nw - 1] = R(5) * R(weights[nw - 1]) / R(0.1) - R(1);
        if (sf == SF_VINARDO) { w[6] = w[5]; w[5] = 0; }
        int ntp = XS_SIZE * (XS_SIZE + 1) / 2;
        tables.assign(ntp, PairTable<R>());
        have.assign(ntp, 0);
    }
    R rot_weight_internal() const { return (sf == SF_VINA) ? w[6] : w[6]; }
};

// Grid type remapping of cache::eval_deriv (cache.cpp:160-199): G types carry
// no grid term, C_*_CG* map to C_H / C_P.
static int grid_type(int t) {
    if (t < 0 || t >= XS_SIZE) return -1;
    switch (t) {
        case XS_G0: case XS_G1: case XS_G2: case XS_G3: return -1;
        case XS_C_H_CG0: case XS_C_H_CG1: case XS_C_H_CG2: case XS_C_H_CG3: return XS_C_H;
        case XS_C_P_CG0: case XS_C_P_CG1: case XS_C_P_CG2: case XS_C_P_CG3: return XS_C_P;
        default: return t;
    }
}

// model::set -> heterotree/tree set_conf (tree.h:121-144, 176-214): root
// frame from (position, orientation), each child in pre-order from its parent.
template <class R> void forward_kinematics(const Ligand<R>& L, const Pose<R>& P, std::vector<V3<R>>& coords,
                                           std::vector<V3<R>>& origins, std::vector<V3<R>>& axes,
                                           std::vector<std::array<R, 9>>* mats_out = nullptr) {
    const int nn = L.n_nodes;
    origins.assign(nn, V3<R>());
    axes.assign(nn, V3<R>());
    std::vector<Q<R>> qs(nn);
    std::vector<std::array<R, 9>> ms(nn);
    coords.assign(L.n_atoms, V3<R>());
    for (int k = 0; k < nn; ++k) {
        if (k == 0) {
            origins[0] = P.pos;
            qs[0] = P.q;  // rigid_body::set_conf does not normalise (tree.h:95-99)
        } else {
            int p = L.parent[k];
            origins[k] = origins[p] + matvec(ms[p].data(), L.rel_origin[k]);
            axes[k] = matvec(ms[p].data(), L.rel_axis[k]);
            Q<R> t = qmul(axis_angle_q(axes[k], P.tors[k - 1]), qs[p]);
            q_normalize_approx(t);
            qs[k] = t;
        }
        q_to_mat(qs[k], ms[k].data());
        for (int i = L.abegin[k]; i < L.aend[k]; ++i) coords[i] = origins[k] + matvec(ms[k].data(), L.local[i]);
    }
    if (mats_out) *mats_out = ms;
}

// non_cache::eval_deriv / non_cache::eval (non_cache.cpp:420-860, 34-255) for
// one ligand atom of XS type t at p: the coordinate clamped into the box
// (penalty slope * |miss|, gradient slope * sign), the type-pair table value
// of every typed receptor atom within the cutoff of the clamped point
// (eval_deriv with a gradient, eval_fast without), curl, and the pair
// gradient kept in clamped directions.  Brute force over the receptor atoms:
// the reference's szv_grid (szv_grid.cpp:24-84) only lists, per cell, the
// atoms that can pass the same r^2 test, in receptor order.
template <class R> R nc_atom(Ctx<R>& C, int t, const V3<R>& p, R slope, R v, V3<R>* deriv) {
    const Grid<R>& g = C.grid;
    V3<R> a = p, oob;
    R pen = 0;
    for (int i = 0; i < 3; ++i) {
        if (p[i] < g.init[i]) { a[i] = g.init[i]; oob[i] = R(-1); pen += std::abs(p[i] - g.init[i]); }
        else if (p[i] > g.hi[i]) { a[i] = g.hi[i]; oob[i] = R(1); pen += std::abs(p[i] - g.hi[i]); }
    }
    pen *= slope;
    R e = 0;
    V3<R> d;
    for (size_t j = 0; j < C.rec.size(); ++j) {
        const V3<R> r = a - C.rec[j];
        const R r2 = norm_sqr(r);
        if (r2 < R(64)) {
            const PairTable<R>& tab = C.table(t, C.rec_t[j]);
            if (deriv) {
                R pe, dor;
                tab.eval_deriv(r2, pe, dor);
                e += pe;
                d = d + dor * r;
            } else {
                e += tab.eval_fast(r2);
            }
        }
    }
    if (deriv) {
        curl(e, d, v);
        *deriv = d + slope * oob;
    } else {
        curl(e, v);
    }
    return e + pen;
}

// ad4cache::eval_deriv / ad4cache::eval (ad4cache.cpp:151-270) for one atom
// of AD type t and charge q at p: the affinity map of its type, the
// electrostatic map times q and the desolvation map times |q|, each
// evaluated (with its own curl and out-of-box penalty) by grid::evaluate.
template <class R> R ad4_grid_atom(const Ctx<R>& C, int t, R q, const V3<R>& p, R v, V3<R>* deriv) {
    t = ad4_grid_type(t);
    if (t < 0 || C.grid.maps[t].empty()) return 0;
    V3<R> d;
    R e = grid_eval(C.grid, t, p, C.slope, v, deriv ? &d : nullptr);
    if (deriv) *deriv = d;
    const R f[2] = {q, std::abs(q)};
    const int m[2] = {kMapElec, kMapDsolv};
    for (int k = 0; k < 2; ++k) {
        e += grid_eval(C.grid, m[k], p, C.slope, v, deriv ? &d : nullptr) * f[k];
        if (deriv) *deriv = *deriv + f[k] * d;
    }
    return e;
}

// model::eval_deriv (model.cpp:954-981) for a single ligand without flex
// residues: grid term (cache::eval_deriv), intramolecular pairs
// (eval_interacting_pairs_deriv, model.cpp:907-934), then the torsion-tree
// derivative (tree.h:75-84, 164-193, 216-222).  g has 6 + T entries:
// [force(3), torque(3), torsion derivatives].
template <class R> R eval_deriv(Ctx<R>& C, const Ligand<R>& L, const Pose<R>& P, const R v[3], R* g,
                                std::vector<V3<R>>* coords_out = nullptr, R* e_inter = nullptr, R* e_intra = nullptr) {
    std::vector<V3<R>> coords, origins, axes;
    forward_kinematics(L, P, coords, origins, axes);
    std::vector<V3<R>> mf(L.n_atoms);
    R e = 0;
    for (int i = 0; i < L.n_atoms && C.sf == SF_AD4; ++i) {
        mf[i] = V3<R>();
        e += ad4_grid_atom(C, L.ad[i], L.q[i], coords[i], v[1], &mf[i]);
    }
    for (int i = 0; i < L.n_atoms && C.sf != SF_AD4; ++i) {
        int t = grid_type(L.xs[i]);
        if (C.use_nc) {  // the explicit receptor: G and untyped atoms skipped, table of the atom's own type
            V3<R> d;
            if (t >= 0) e += nc_atom(C, L.xs[i], coords[i], C.nc_slope, v[1], &d);
            mf[i] = d;
            continue;
        }
        if (t < 0 || C.grid.maps[t].empty()) { mf[i] = V3<R>(); continue; }
        V3<R> d;
        e += grid_eval(C.grid, t, coords[i], C.slope, v[1], &d);
        mf[i] = d;
    }
    R einter = e;
    R eintra = 0;
    // sqr(sf.get_cutoff()), scoring_function.h:56,66,82
    const R cutoff_sqr = (C.sf == SF_AD4) ? sqr(R(kAd4Cutoff)) : R(64);
    for (int k = 0; k < L.n_pairs; ++k) {
        const int a = L.pa[k], b = L.pb[k];
        V3<R> r = coords[b] - coords[a];
        R r2 = norm_sqr(r);
        if (r2 < cutoff_sqr) {
            R pe, dor;
            if (C.sf == SF_AD4) L.ptab[k].eval_deriv(r2, pe, dor);
            else C.table(L.xs[a], L.xs[b]).eval_deriv(r2, pe, dor);
            V3<R> force = dor * r;
            curl(pe, force, v[0]);
            eintra += pe;
            mf[a] = mf[a] - force;
            mf[b] = mf[b] + force;
        }
    }
    e += eintra;
    // tree derivative: post-order accumulation of force and torque
    const int nn = L.n_nodes;
    std::vector<V3<R>> F(nn), T(nn);
    for (int k = 0; k < nn; ++k) {
        for (int i = L.abegin[k]; i < L.aend[k]; ++i) {
            F[k] = F[k] + mf[i];
            T[k] = T[k] + cross(coords[i] - origins[k], mf[i]);
        }
    }
    // children appear after their parent in pre-order, so a reverse sweep
    // visits every child before its parent; within one parent the children
    // are added in reverse order, which only reorders a floating-point sum
    // relative to branches_derivative (tree.h:152-162).
    for (int k = nn - 1; k >= 1; --k) {
        int p = L.parent[k];
        F[p] = F[p] + F[k];
        T[p] = T[p] + (cross(origins[k] - origins[p], F[k]) + T[k]);
    }
    g[0] = F[0][0]; g[1] = F[0][1]; g[2] = F[0][2];
    g[3] = T[0][0]; g[4] = T[0][1]; g[5] = T[0][2];
    for (int k = 1; k < nn; ++k) g[5 + k] = dot(T[k], axes[k]);
    if (coords_out) *coords_out = coords;
    if (e_inter) *e_inter = einter;
    if (e_intra) *e_intra = eintra;
    return e;
}

// conf::increment (conf.h:98-104, 50-56, 169-172)
template <class R> void pose_increment(Pose<R>& x, const R* p, R alpha) {
    for (int i = 0; i < 3; ++i) x.pos[i] += alpha * p[i];
    V3<R> rot(alpha * p[3], alpha * p[4], alpha * p[5]);
    q_increment(x.q, rot);
    for (size_t k = 0; k < x.tors.size(); ++k) {
        x.tors[k] += normalized_angle(alpha * p[6 + k]);
        normalize_angle(x.tors[k]);
    }
}

// Per-step record of a BFGS run (test hook): the accepted step length, the
// energy after the line search, the number of line-search evaluations, and
// after the step's update the inverse Hessian (full n x n) with the step's
// s = alpha p and y = g_new - g.
struct BfgsTrace {
    int steps = 0;
    std::vector<double> alpha, f, H, s, y;
    std::vector<int> trials;
};

// bfgs() with line_search() and bfgs_update() (bfgs.h:30-147) over any
// objective: eval(x, g) returns f(x) and fills g (n entries); incr(x, p, a)
// moves x by a p (conf::increment for poses, conf.h:98-104).  The line search
// compares with f0 = f(x) (bfgs.h:107); with f0_zero_first, the first line
// search compares with f0 = 0 instead, as the reference's GPU BFGS does
// (bfgs_warp, warp_ops.cuh:1062-1076: f0 starts at 0 and f_orig is kept
// separately).  Packed upper-triangular inverse Hessian (triangular_matrix,
// matrix.h), index tri_index.
template <class R, class X, class Eval, class Incr>
R bfgs_core(Eval& eval, Incr& incr, X& x, int n, int max_steps, int* evalcount, bool f0_zero_first,
            BfgsTrace* tr = nullptr) {
    std::vector<R> h(size_t(n) * (n + 1) / 2, R(0));
    auto H = [&](int i, int j) -> R& { return h[tri_index(i, j)]; };
    for (int i = 0; i < n; ++i) H(i, i) = 1;
    std::vector<R> g(n), g_new(n), p(n), y(n), minus_hy(n);
    X x_new = x;
    const R f_orig = eval(x, g.data());
    ++*evalcount;
    R f0 = f0_zero_first ? R(0) : f_orig;
    const X x_orig = x;
    auto dotn = [&](const std::vector<R>& a, const std::vector<R>& b) {
        R s = 0;
        for (int i = 0; i < n; ++i) s += a[i] * b[i];
        return s;
    };
    for (int step = 0; step < max_steps; ++step) {
        for (int i = 0; i < n; ++i) {  // minus_mat_vec_product (bfgs.h:30-38)
            R sum = 0;
            for (int j = 0; j < n; ++j) sum += H(i, j) * g[j];
            p[i] = -sum;
        }
        // line_search (bfgs.h:65-85): c0 = 1e-4, multiplier 0.5, 10 trials
        R f1 = 0;
        R alpha = 1;
        const R pg = dotn(p, g);
        int trials = 0;
        for (int trial = 0; trial < 10; ++trial) {
            x_new = x;
            incr(x_new, p.data(), alpha);
            f1 = eval(x_new, g_new.data());
            ++*evalcount;
            ++trials;
            if (f1 - f0 < R(0.0001) * alpha * pg) break;
            alpha *= R(0.5);
        }
        for (int i = 0; i < n; ++i) y[i] = g_new[i] - g[i];
        f0 = f1;
        x = x_new;
        if (tr) {
            tr->steps = step + 1;
            tr->alpha.push_back(double(alpha));
            tr->f.push_back(double(f1));
            tr->trials.push_back(trials);
            for (int i = 0; i < n; ++i) tr->s.push_back(double(alpha * p[i]));
            for (int i = 0; i < n; ++i) tr->y.push_back(double(y[i]));
        }
        // the convergence test reads g before g = g_new (bfgs.h:131-132)
        const bool stop = !(std::sqrt(dotn(g, g)) >= R(1e-5));
        if (!stop) {
            g = g_new;
            if (step == 0) {  // initial scaling (bfgs.h:135-138)
                const R yy = dotn(y, y);
                if (std::abs(yy) > R(kEpsFl)) {
                    const R d = alpha * dotn(y, p) / yy;
                    for (int i = 0; i < n; ++i) H(i, i) = d;
                }
            }
            // bfgs_update (bfgs.h:46-63)
            const R yp = dotn(y, p);
            if (!(alpha * yp < R(kEpsFl))) {
                for (int i = 0; i < n; ++i) {
                    R sum = 0;
                    for (int j = 0; j < n; ++j) sum += H(i, j) * y[j];
                    minus_hy[i] = -sum;
                }
                const R yhy = -dotn(y, minus_hy);
                const R r = R(1) / (alpha * yp);
                for (int i = 0; i < n; ++i)
                    for (int j = i; j < n; ++j)
                        H(i, j) += alpha * r * (minus_hy[i] * p[j] + minus_hy[j] * p[i]) +
                                   alpha * alpha * (r * r * yhy + r) * p[i] * p[j];
            }
        }
        if (tr)
            for (int i = 0; i < n; ++i)
                for (int j = 0; j < n; ++j) tr->H.push_back(double(H(i, j)));
        if (stop) break;
    }
    // revert when nothing improved (bfgs.h:140-144)
    if (!(f0 <= f_orig)) {
        f0 = f_orig;
        x = x_orig;
    }
    return f0;
}

// quasi_newton (quasi_newton.cpp:43-53) on a pose: eval_deriv with caps v,
// conf::increment.  C.ls_f0_zero selects the GPU's first-line-search rule.
template <class R> R bfgs(Ctx<R>& C, const Ligand<R>& L, Pose<R>& x, const R v[3], int max_steps, int* evalcount) {
    auto ev = [&](const Pose<R>& P, R* g) { return eval_deriv(C, L, P, v, g); };
    auto inc = [&](Pose<R>& P, const R* p, R a) { pose_increment(P, p, a); };
    return bfgs_core<R>(ev, inc, x, 6 + L.n_tors(), max_steps, evalcount, C.ls_f0_zero);
}

// ---------------------------------------------------------------- Philox
// Philox-4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
// easy as 1, 2, 3", SC'11): multipliers 0xD2511F53 / 0xCD9E8D57, Weyl key
// increments 0x9E3779B9 / 0xBB67AE85, ten rounds.
static void philox(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = uint64_t(0xD2511F53u) * c0;
        uint64_t p1 = uint64_t(0xCD9E8D57u) * c2;
        uint32_t hi0 = uint32_t(p0 >> 32), lo0 = uint32_t(p0);
        uint32_t hi1 = uint32_t(p1 >> 32), lo1 = uint32_t(p1);
        uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}
// Stream layout used by the search (identical in the HIP kernels):
//   counter = (step, chain, purpose, 0), key = (seed_lo, seed_hi)
//   step 0 is the random start; MC step s uses counter step s + 1.
//   purpose 1: start position / mutation, 2: start orientation / Metropolis,
//   3 + j: start torsions 4j .. 4j+3.
static inline double u01(uint32_t x) { return double(x >> 8) * (1.0 / 16777216.0); }

template <class R> void random_start(const Ligand<R>& L, const R lo[3], const R hi[3], uint64_t seed, uint32_t chain, Pose<R>& P) {
    uint32_t key[2] = {uint32_t(seed), uint32_t(seed >> 32)};
    uint32_t o[4];
    uint32_t c[4] = {0, chain, 1, 0};
    philox(c, key, o);
    // random_in_box (random.cpp:78-84): uniform per axis between the corners
    for (int i = 0; i < 3; ++i) P.pos[i] = lo[i] + R(u01(o[i])) * (hi[i] - lo[i]);
    // random_orientation (quaternion.cpp:116-126): normalised 4-D Gaussian
    c[2] = 2;
    philox(c, key, o);
    R nrm4[4];
    for (int j = 0; j < 2; ++j) {
        R u1 = R(1.0 - u01(o[2 * j]));  // (0, 1]
        R u2 = R(u01(o[2 * j + 1]));
        R rad = std::sqrt(R(-2) * std::log(u1));
        R ang = R(2 * kPi) * u2;
        nrm4[2 * j] = rad * std::cos(ang);
        nrm4[2 * j + 1] = rad * std::sin(ang);
    }
    R nq = std::sqrt(nrm4[0] * nrm4[0] + nrm4[1] * nrm4[1] + nrm4[2] * nrm4[2] + nrm4[3] * nrm4[3]);
    if (!(nq > R(kEpsFl))) { nrm4[0] = 1; nrm4[1] = nrm4[2] = nrm4[3] = 0; nq = 1; }
    P.q = Q<R>{nrm4[0] / nq, nrm4[1] / nq, nrm4[2] / nq, nrm4[3] / nq};
    // torsions_randomize (conf.h:58-61): U(-pi, pi)
    P.tors.assign(L.n_tors(), R(0));
    for (int k = 0; k < L.n_tors(); ++k) {
        if (k % 4 == 0) { c[2] = 3 + k / 4; philox(c, key, o); }
        P.tors[k] = R(-kPi) + R(2 * kPi) * R(u01(o[k % 4]));
    }
}

// model::gyration_radius (model.cpp:871-883) of the pose being mutated:
// RMS distance of heavy atoms from the root origin.
template <class R> R gyration_radius(const Ligand<R>& L, const std::vector<V3<R>>& coords, const V3<R>& root) {
    R acc = 0;
    int cnt = 0;
    for (int i = 0; i < L.n_atoms; ++i)
        if (L.heavy[i]) { acc += norm_sqr(coords[i] - root); ++cnt; }
    return cnt > 0 ? std::sqrt(acc / R(cnt)) : R(0);
}

// mutate_conf (mutate.cpp:35-81): one of {position, orientation, one torsion}
// chosen uniformly; position += amplitude * u, orientation incremented by
// (amplitude / r_gyr) * u, u uniform in the unit ball; torsion := U(-pi, pi).
template <class R> void mutate(const Ligand<R>& L, Pose<R>& P, R rgyr, R amplitude, uint64_t seed, uint32_t chain, uint32_t step) {
    uint32_t key[2] = {uint32_t(seed), uint32_t(seed >> 32)};
    uint32_t c[4] = {step + 1, chain, 1, 0};
    uint32_t o[4];
    philox(c, key, o);
    const uint32_t count = 2 + L.n_tors();
    uint32_t which = uint32_t((uint64_t(o[0]) * count) >> 32);
    // unit ball: direction from (z, phi), radius cbrt(u)
    R z = R(2) * R(u01(o[1])) - R(1);
    R phi = R(2 * kPi) * R(u01(o[2]));
    R rr = std::cbrt(R(u01(o[3])));
    R sxy = std::sqrt(std::max(R(0), R(1) - z * z));
    V3<R> u(rr * sxy * std::cos(phi), rr * sxy * std::sin(phi), rr * z);
    if (which == 0) {
        P.pos = P.pos + amplitude * u;
        return;
    }
    if (which == 1) {
        if (rgyr > R(kEpsFl)) {
            V3<R> rot = (amplitude / rgyr) * u;
            q_increment(P.q, rot);
        }
        return;
    }
    P.tors[which - 2] = R(-kPi) + R(2 * kPi) * R(u01(o[1]));
}

// metropolis_accept (monte_carlo.cu:92-97)
template <class R> bool metropolis(R old_f, R new_f, R temperature, R u) {
    if (new_f < old_f) return true;
    const R acc = std::exp((old_f - new_f) / temperature);
    return u < acc;
}

// One Monte Carlo chain as run per chain by kernel<32, Config>
// (monte_carlo.cu:236-356): random start, then for each step: mutate a copy,
// BFGS with hunt_cap, Metropolis at T = 1.2, and when the accepted pose beats
// the chain's best, BFGS again with authentic_v = 1000 and keep it.  The BFGS
// is the CPU semantics of bfgs.h (see bfgs above).
template <class R> R mc_chain(Ctx<R>& C, const Ligand<R>& L, uint64_t seed, uint32_t chain, int global_steps,
                              int local_steps, R amplitude, R temperature, Pose<R>& best, int* evalcount,
                              R* trace = nullptr) {
    const R hunt[3] = {10, 10, 10};     // vina.h:207
    const R authentic[3] = {1000, 1000, 1000};
    R lo[3], hi[3];
    for (int i = 0; i < 3; ++i) {
        lo[i] = C.grid.init[i];
        hi[i] = C.grid.init[i] + C.grid.factor_inv[i] * R(C.grid.n[i] - 1);
    }
    Pose<R> tmp;
    random_start(L, lo, hi, seed, chain, tmp);
    R tmp_e = 0;
    R best_e = std::numeric_limits<R>::infinity();
    best = tmp;
    std::vector<V3<R>> coords, origins, axes;
    forward_kinematics(L, tmp, coords, origins, axes);
    R rgyr = gyration_radius(L, coords, tmp.pos);
    uint32_t key[2] = {uint32_t(seed), uint32_t(seed >> 32)};
    for (int step = 0; step < global_steps; ++step) {
        Pose<R> cand = tmp;
        mutate(L, cand, rgyr, amplitude, seed, chain, uint32_t(step));
        R cand_e = bfgs(C, L, cand, hunt, local_steps, evalcount);
        uint32_t c[4] = {uint32_t(step + 1), chain, 2, 0}, o[4];
        philox(c, key, o);
        bool acc = metropolis(tmp_e, cand_e, temperature, R(u01(o[0])));
        if (step == 0 || acc) {
            tmp = cand;
            tmp_e = cand_e;
            forward_kinematics(L, tmp, coords, origins, axes);
            rgyr = gyration_radius(L, coords, tmp.pos);
            if (tmp_e < best_e) {
                tmp_e = bfgs(C, L, tmp, authentic, local_steps, evalcount);
                forward_kinematics(L, tmp, coords, origins, axes);
                rgyr = gyration_radius(L, coords, tmp.pos);
                if (tmp_e < best_e) {
                    best = tmp;
                    best_e = tmp_e;
                }
            }
        }
        if (trace) trace[step] = best_e;
    }
    return best_e;
}

// Final affinity of a pose (Vina::score_gpu, vina.cpp:1245-1327, no flex, no
// refinement): inter = grid energy with authentic_v (cache::eval, energy-only
// path, cache.cpp:84-114), intra = ligand pairs with eval_fast and curl at
// authentic_v (eval_interacting_pairs, model.cpp:885-905), then
// num_tors_div (conf_independent.cpp:153-156):
//   total = (inter + intra - intra_best) / (1 + 0.1 (w_rot_internal + 1) num_tors / 5)
template <class R> void score_terms(Ctx<R>& C, const Ligand<R>& L, const Pose<R>& P, R* inter, R* intra) {
    std::vector<V3<R>> coords, origins, axes;
    forward_kinematics(L, P, coords, origins, axes);
    R e = 0;
    for (int i = 0; i < L.n_atoms; ++i) {
        if (C.sf == SF_AD4) {
            e += ad4_grid_atom(C, L.ad[i], L.q[i], coords[i], R(1000), (V3<R>*)nullptr);
            continue;
        }
        int t = grid_type(L.xs[i]);
        if (t < 0 || C.grid.maps[t].empty()) continue;
        e += grid_eval(C.grid, t, coords[i], C.slope, R(1000), (V3<R>*)nullptr);
    }
    R ei = 0;
    const R cutoff_sqr = (C.sf == SF_AD4) ? sqr(R(kAd4Cutoff)) : R(64);
    for (int k = 0; k < L.n_pairs; ++k) {
        const int a = L.pa[k], b = L.pb[k];
        R r2 = norm_sqr(coords[b] - coords[a]);
        if (r2 < cutoff_sqr) {
            R t = (C.sf == SF_AD4) ? L.ptab[k].eval_fast(r2) : C.table(L.xs[a], L.xs[b]).eval_fast(r2);
            curl(t, R(1000));
            ei += t;
        }
    }
    *inter = e;
    *intra = ei;
}
template <class R> R conf_independent(const Ctx<R>& C, R num_tors, R x) {
    // AD4.2: ad4_tors_add (conf_independent.cpp:183-187), num_tors = TORSDOF
    if (C.sf == SF_AD4) return x + C.w[5] * num_tors;
    const R weight = R(0.1) * (C.w[6] + R(1));
    const R y = R(1) + weight * num_tors / R(5.0);
    if (std::abs(x) < R(kEpsFl)) return 0;
    if (std::abs(y) < R(kEpsFl)) return (x * y > 0) ? R(kMaxFl) : R(-kMaxFl);
    return x / y;
}

// non_cache::within (non_cache.cpp:408-418, margin 1e-4): every heavy atom
// inside the box.
template <class R> bool nc_within(Ctx<R>& C, const Ligand<R>& L, const Pose<R>& P) {
    std::vector<V3<R>> coords, origins, axes;
    forward_kinematics(L, P, coords, origins, axes);
    const R margin = R(0.0001);
    for (int i = 0; i < L.n_atoms; ++i) {
        if (!L.heavy[i]) continue;
        for (int d = 0; d < 3; ++d)
            if (coords[i][d] < C.grid.init[d] - margin || coords[i][d] > C.grid.hi[d] + margin) return false;
    }
    return true;
}

// Refinement of one kept pose (Vina::global_search_gpu, vina.h:249-272):
// quasi_newton with authentic_v = 1000 and (25 + atoms) / 3 steps against the
// explicit receptor, pass p with slope 100 * 10^(2p), stopping once the pose
// is within the box; a pose never within gets max_fl.
template <class R> R refine(Ctx<R>& C, const Ligand<R>& L, Pose<R>& P, int refine_step, bool* within) {
    const R v[3] = {R(1000), R(1000), R(1000)};
    const int max_steps = (25 + L.n_atoms) / 3;
    C.use_nc = true;
    R e = 0;
    bool in = false;
    for (int p = 0; p < refine_step; ++p) {
        C.nc_slope = R(100.0 * std::pow(10.0, 2.0 * p));
        int ec = 0;
        e = bfgs(C, L, P, v, max_steps, &ec);
        in = nc_within(C, L, P);
        if (in) break;
    }
    C.use_nc = false;
    *within = in;
    return in ? e : R(kMaxFl);
}

// score_gpu's refined branch (vina.cpp:1262-1283): inter = non_cache::eval
// with authentic_v and slope 1e6, intra as in score_terms.
template <class R> void score_terms_nc(Ctx<R>& C, const Ligand<R>& L, const Pose<R>& P, R* inter, R* intra) {
    std::vector<V3<R>> coords, origins, axes;
    forward_kinematics(L, P, coords, origins, axes);
    R e = 0;
    for (int i = 0; i < L.n_atoms; ++i)
        if (grid_type(L.xs[i]) >= 0) e += nc_atom(C, L.xs[i], coords[i], R(1e6), R(1000), (V3<R>*)nullptr);
    R dummy;
    score_terms(C, L, P, &dummy, intra);
    *inter = e;
}

// Grid population, cache::populate_no_bias (cache.cpp:471-549): for each
// sample point and needed type t2, the sum over typed receptor atoms with
// r^2 <= cutoff^2 of eval_fast(t1, t2, r^2) of the TYPE-based table
// (precalculate, precalculate.h:130-201).  Brute force over all atoms; the
// reference's szv_grid only prunes atoms that fail the same r^2 test.
template <class R> void build_maps(Ctx<R>& C, int n_rec, const double* rec_xyz, const int* rec_xs, int n_types, const int* types) {
    Grid<R>& g = C.grid;
    std::vector<V3<R>> ra;
    std::vector<int> rt;
    for (int i = 0; i < n_rec; ++i) {
        if (rec_xs[i] < 0 || rec_xs[i] >= XS_SIZE) continue;
        ra.push_back(V3<R>(R(rec_xyz[3 * i]), R(rec_xyz[3 * i + 1]), R(rec_xyz[3 * i + 2])));
        rt.push_back(rec_xs[i]);
    }
    C.rec = ra;
    C.rec_t = rt;
    for (int j = 0; j < n_types; ++j) {
        g.maps[types[j]].assign(size_t(g.n[0]) * g.n[1] * g.n[2], R(0));
    }
    const R cutoff_sqr = R(64);
    // tables are built lazily: build every (receptor type, map type) pair
    // before the z-slices are shared out among threads
    for (int t : rt)
        for (int j = 0; j < n_types; ++j) C.table(t, types[j]);
    auto slice = [&](int z) {
        std::vector<R> acc(n_types);
        for (int y = 0; y < g.n[1]; ++y)
            for (int x = 0; x < g.n[0]; ++x) {
                V3<R> p = g.point(x, y, z);
                std::fill(acc.begin(), acc.end(), R(0));
                for (size_t i = 0; i < ra.size(); ++i) {
                    R r2 = norm_sqr(ra[i] - p);
                    if (r2 <= cutoff_sqr)
                        for (int j = 0; j < n_types; ++j) acc[j] += C.table(rt[i], types[j]).eval_fast(r2);
                }
                for (int j = 0; j < n_types; ++j) g.at(types[j], x, y, z) = acc[j];
            }
    };
    const int nt = std::max(1, std::min<int>(16, int(std::thread::hardware_concurrency())));
    std::atomic<int> next(0);
    std::vector<std::thread> th;
    for (int i = 0; i < nt; ++i)
        th.emplace_back([&] {
            for (int z = next.fetch_add(1); z < g.n[2]; z = next.fetch_add(1)) slice(z);
        });
    for (auto& t : th) t.join();
}

// AutoGrid-style AD4.2 maps from the pair potentials above, so that a
// ligand atom of type t and charge q at p scores
//   map_t(p) + q elec(p) + |q| dsolv(p) = sum over receptor atoms j of
//   sf_eval_ad4(t, q, t_j, q_j, r):
//   map_t  = sum_j w_vdw vdw + w_hb hb + w_dsolv (solv_t V_j + (solv_j + Q |q_j|) V_t) G(r)
//   elec   = sum_j w_elec q_j 332 / (r eps(r))   (capped)
//   dsolv  = sum_j w_dsolv Q V_j G(r)
// (the reference reads such maps from AutoGrid files, ad4cache::read,
// ad4cache.cpp:391-440).  Brute force over every receptor atom.
template <class R> void build_maps_ad4(Ctx<R>& C, int n_rec, const double* rec_xyz, const int* rec_ad,
                                       const double* rec_q, int n_types, const int* types) {
    Grid<R>& g = C.grid;
    std::vector<V3<R>> ra;
    std::vector<int> rt;
    std::vector<R> rq;
    for (int i = 0; i < n_rec; ++i) {
        if (rec_ad[i] < 0 || rec_ad[i] >= AD_SIZE) continue;
        ra.push_back(V3<R>(R(rec_xyz[3 * i]), R(rec_xyz[3 * i + 1]), R(rec_xyz[3 * i + 2])));
        rt.push_back(rec_ad[i]);
        rq.push_back(R(rec_q[i]));
    }
    const size_t np = size_t(g.n[0]) * g.n[1] * g.n[2];
    for (int j = 0; j < n_types; ++j) g.maps[types[j]].assign(np, R(0));
    g.maps[kMapElec].assign(np, R(0));
    g.maps[kMapDsolv].assign(np, R(0));
    const R* w = C.w;
    const R cut2 = sqr(R(kAd4Cutoff));
    auto slice = [&](int z) {
        std::vector<R> acc(n_types + 2);
        for (int y = 0; y < g.n[1]; ++y)
            for (int x = 0; x < g.n[0]; ++x) {
                V3<R> p = g.point(x, y, z);
                std::fill(acc.begin(), acc.end(), R(0));
                for (size_t i = 0; i < ra.size(); ++i) {
                    const R r2 = norm_sqr(ra[i] - p);
                    if (!(r2 < cut2)) continue;
                    const R r = std::sqrt(r2);
                    const R G = ad4_solv_r(r);
                    const int tj = rt[i];
                    const R Sj = R(kAd[tj].solv) + R(kAd4SolvQ) * std::abs(rq[i]);
                    for (int j = 0; j < n_types; ++j) {
                        const int t = types[j];
                        acc[j] += w[0] * ad4_vdw(t, tj, r) + w[1] * ad4_hb(t, tj, r) +
                                  w[3] * (R(kAd[t].solv) * R(kAd[tj].vol) + Sj * R(kAd[t].vol)) * G;
                    }
                    acc[n_types] += w[2] * rq[i] * ad4_elec_r(r);
                    acc[n_types + 1] += w[3] * R(kAd4SolvQ) * R(kAd[tj].vol) * G;
                }
                for (int j = 0; j < n_types; ++j) g.at(types[j], x, y, z) = acc[j];
                g.at(kMapElec, x, y, z) = acc[n_types];
                g.at(kMapDsolv, x, y, z) = acc[n_types + 1];
            }
    };
    const int nt = std::max(1, std::min<int>(16, int(std::thread::hardware_concurrency())));
    std::atomic<int> next(0);
    std::vector<std::thread> th;
    for (int i = 0; i < nt; ++i)
        th.emplace_back([&] {
            for (int z = next.fetch_add(1); z < g.n[2]; z = next.fetch_add(1)) slice(z);
        });
    for (auto& t : th) t.join();
}

}  // namespace orc

// ====================================================================== C ABI
// Opaque handle; precision selected at creation (32 or 64).
using namespace orc;

struct OrcHandle {
    int prec;
    Ctx<double> d;
    Ctx<float> f;
};

template <class R> static Ctx<R>& ctx_of(OrcHandle* h);
template <> Ctx<double>& ctx_of<double>(OrcHandle* h) { return h->d; }
template <> Ctx<float>& ctx_of<float>(OrcHandle* h) { return h->f; }

template <class R> static void load_pose(const Ligand<R>& L, const double* p, Pose<R>& P) {
    P.pos = V3<R>(R(p[0]), R(p[1]), R(p[2]));
    P.q = Q<R>{R(p[3]), R(p[4]), R(p[5]), R(p[6])};
    P.tors.resize(L.n_tors());
    for (int k = 0; k < L.n_tors(); ++k) P.tors[k] = R(p[7 + k]);
}
template <class R> static void store_pose(const Pose<R>& P, double* p) {
    p[0] = P.pos[0]; p[1] = P.pos[1]; p[2] = P.pos[2];
    p[3] = P.q.a; p[4] = P.q.b; p[5] = P.q.c; p[6] = P.q.d;
    for (size_t k = 0; k < P.tors.size(); ++k) p[7 + k] = P.tors[k];
}


extern "C" void orc_philox(const uint32_t* ctr, const uint32_t* key, uint32_t* out) { philox(ctr, key, out); }

extern "C" OrcHandle* orc_new(int prec, int sf, const double* weights) {
    OrcHandle* h = new OrcHandle();
    h->prec = prec;
    h->d.init(sf, weights);
    h->f.init(sf, weights);
    return h;
}
extern "C" void orc_free(OrcHandle* h) { delete h; }

// sample tables of one type pair: e, dor, fast (each 2051 doubles)
extern "C" void orc_pair_table(OrcHandle* h, int t1, int t2, double* e, double* dor, double* fast) {
    if (h->prec == 64) {
        const auto& t = h->d.table(t1, t2);
        for (int i = 0; i < kTableN; ++i) { e[i] = t.e[i]; dor[i] = t.dor[i]; fast[i] = t.fast[i]; }
    } else {
        const auto& t = h->f.table(t1, t2);
        for (int i = 0; i < kTableN; ++i) { e[i] = t.e[i]; dor[i] = t.dor[i]; fast[i] = t.fast[i]; }
    }
}
extern "C" double orc_sf_eval(OrcHandle* h, int t1, int t2, double r) {
    if (h->prec == 64) return sf_eval<double>(h->d.sf, t1, t2, r, h->d.w);
    return sf_eval<float>(h->f.sf, t1, t2, float(r), h->f.w);
}

// grid geometry -> n[3], init[3], factor[3], factor_inv[3]
extern "C" void orc_set_grid(OrcHandle* h, const double* center, const double* size, double granularity, int* n_out, double* geo_out) {
    h->d.grid.set_geometry(center, size, granularity);
    h->f.grid.set_geometry(center, size, granularity);
    for (int i = 0; i < 3; ++i) {
        n_out[i] = h->d.grid.n[i];
        if (h->prec == 64) {
            geo_out[i] = h->d.grid.init[i]; geo_out[3 + i] = h->d.grid.factor[i]; geo_out[6 + i] = h->d.grid.factor_inv[i];
        } else {
            geo_out[i] = h->f.grid.init[i]; geo_out[3 + i] = h->f.grid.factor[i]; geo_out[6 + i] = h->f.grid.factor_inv[i];
        }
    }
}
extern "C" void orc_build_maps(OrcHandle* h, int n_rec, const double* xyz, const int* xs, int n_types, const int* types) {
    if (h->prec == 64) build_maps(h->d, n_rec, xyz, xs, n_types, types);
    else build_maps(h->f, n_rec, xyz, xs, n_types, types);
}
// copy maps in/out (x fastest); values as double
extern "C" void orc_get_map(OrcHandle* h, int t, double* out) {
    if (h->prec == 64) { auto& m = h->d.grid.maps[t]; for (size_t i = 0; i < m.size(); ++i) out[i] = m[i]; }
    else { auto& m = h->f.grid.maps[t]; for (size_t i = 0; i < m.size(); ++i) out[i] = m[i]; }
}
extern "C" void orc_set_map(OrcHandle* h, int t, const double* in) {
    size_t n = size_t(h->d.grid.n[0]) * h->d.grid.n[1] * h->d.grid.n[2];
    h->d.grid.maps[t].assign(in, in + n);
    h->f.grid.maps[t].resize(n);
    for (size_t i = 0; i < n; ++i) h->f.grid.maps[t][i] = float(in[i]);
}
extern "C" double orc_grid_eval(OrcHandle* h, int t, const double* loc, double v, double* deriv) {
    if (h->prec == 64) {
        V3<double> d;
        double e = grid_eval(h->d.grid, t, V3<double>(loc[0], loc[1], loc[2]), h->d.slope, v, deriv ? &d : nullptr);
        if (deriv) for (int i = 0; i < 3; ++i) deriv[i] = d[i];
        return e;
    }
    V3<float> d;
    float e = grid_eval(h->f.grid, t, V3<float>(float(loc[0]), float(loc[1]), float(loc[2])), h->f.slope, float(v), deriv ? &d : nullptr);
    if (deriv) for (int i = 0; i < 3; ++i) deriv[i] = d[i];
    return e;
}

// Add one ligand in the flat layout.  Arrays:
//   local[3*n_atoms], xs[n_atoms], heavy[n_atoms]
//   parent[n_nodes], abegin[n_nodes], aend[n_nodes], origin0[3*n_nodes], axis0[3*n_nodes]
//   pairs[2*n_pairs]
// Returns the ligand index.
template <class R> static int add_lig(Ctx<R>& C, int na, const double* local, const int* xs, const int* heavy, int nn,
                                      const int* parent, const int* ab, const int* ae, const double* origin0,
                                      const double* axis0, int np, const int* pairs, double num_tors) {
    Ligand<R> L;
    L.n_atoms = na; L.n_nodes = nn; L.n_pairs = np;
    L.local.resize(na); L.xs.assign(xs, xs + na); L.heavy.assign(heavy, heavy + na);
    L.node_of.assign(na, 0);
    for (int i = 0; i < na; ++i) L.local[i] = V3<R>(R(local[3 * i]), R(local[3 * i + 1]), R(local[3 * i + 2]));
    L.parent.assign(parent, parent + nn); L.abegin.assign(ab, ab + nn); L.aend.assign(ae, ae + nn);
    L.rel_origin.resize(nn); L.rel_axis.resize(nn); L.origin0.resize(nn);
    for (int k = 0; k < nn; ++k) {
        L.origin0[k] = V3<R>(R(origin0[3 * k]), R(origin0[3 * k + 1]), R(origin0[3 * k + 2]));
        for (int i = ab[k]; i < ae[k]; ++i) L.node_of[i] = k;
    }
    for (int k = 0; k < nn; ++k) {
        if (k == 0) continue;
        // segment c'tor (tree.h:123-131): relative_origin = origin - parent origin,
        // relative_axis = axis (parent orientation is the identity at set-up)
        L.rel_origin[k] = L.origin0[k] - L.origin0[parent[k]];
        L.rel_axis[k] = V3<R>(R(axis0[3 * k]), R(axis0[3 * k + 1]), R(axis0[3 * k + 2]));
    }
    L.pa.resize(np); L.pb.resize(np); L.ptp.resize(np);
    for (int k = 0; k < np; ++k) { L.pa[k] = pairs[2 * k]; L.pb[k] = pairs[2 * k + 1]; L.ptp[k] = tri_index(xs[pairs[2 * k]], xs[pairs[2 * k + 1]]); }
    L.num_tors = R(num_tors);
    C.ligs.push_back(L);
    return int(C.ligs.size()) - 1;
}
extern "C" int orc_add_ligand(OrcHandle* h, int na, const double* local, const int* xs, const int* heavy, int nn, const int* parent,
                   const int* ab, const int* ae, const double* origin0, const double* axis0, int np, const int* pairs, double num_tors) {
    add_lig(h->f, na, local, xs, heavy, nn, parent, ab, ae, origin0, axis0, np, pairs, num_tors);
    return add_lig(h->d, na, local, xs, heavy, nn, parent, ab, ae, origin0, axis0, np, pairs, num_tors);
}

template <class R> static double eval_impl(Ctx<R>& C, int lig, const double* pose, const double* v, double* grad, double* coords_out, double* terms) {
    const Ligand<R>& L = C.ligs[lig];
    Pose<R> P;
    load_pose(L, pose, P);
    R vv[3] = {R(v[0]), R(v[1]), R(v[2])};
    std::vector<R> g(6 + L.n_tors());
    std::vector<V3<R>> coords;
    R ei, ea;
    R e = eval_deriv(C, L, P, vv, g.data(), &coords, &ei, &ea);
    if (grad) for (size_t i = 0; i < g.size(); ++i) grad[i] = g[i];
    if (coords_out) for (int i = 0; i < L.n_atoms; ++i) for (int k = 0; k < 3; ++k) coords_out[3 * i + k] = coords[i][k];
    if (terms) { terms[0] = ei; terms[1] = ea; }
    return e;
}
// energy + gradient (6 + T) of one pose; terms[2] = (grid part, intramolecular part)
extern "C" double orc_eval_deriv(OrcHandle* h, int lig, const double* pose, const double* v, double* grad, double* coords, double* terms) {
    if (h->prec == 64) return eval_impl(h->d, lig, pose, v, grad, coords, terms);
    return eval_impl(h->f, lig, pose, v, grad, coords, terms);
}

template <class R> static double bfgs_impl(Ctx<R>& C, int lig, double* pose, const double* v, int max_steps, int* evals) {
    const Ligand<R>& L = C.ligs[lig];
    Pose<R> P;
    load_pose(L, pose, P);
    R vv[3] = {R(v[0]), R(v[1]), R(v[2])};
    int ec = 0;
    R e = bfgs(C, L, P, vv, max_steps, &ec);
    store_pose(P, pose);
    if (evals) *evals = ec;
    return e;
}
extern "C" double orc_bfgs(OrcHandle* h, int lig, double* pose, const double* v, int max_steps, int* evals) {
    if (h->prec == 64) return bfgs_impl(h->d, lig, pose, v, max_steps, evals);
    return bfgs_impl(h->f, lig, pose, v, max_steps, evals);
}

// BFGS on the quadratic f(x) = 1/2 x^T A x - b^T x + c (A row-major n x n),
// x moved by x + alpha p: the search itself, free of poses, for tests with
// hand-computable iterates.  prec 32 or 64.  Per step s (up to max_steps):
// alpha[s], f[s], trials[s], H[s][n][n] after the step's update, sv[s][n] =
// alpha p, yv[s][n] = g_new - g.  Returns the final f; x in place.
template <class R> static double bfgs_quad_impl(int n, const double* A, const double* b, double c, double* x,
                                                int max_steps, int f0_zero, int* evals, int* steps, double* alpha,
                                                double* f, int* trials, double* H, double* sv, double* yv) {
    std::vector<R> xv(x, x + n);
    auto ev = [&](const std::vector<R>& z, R* g) {
        R fz = R(c);
        for (int i = 0; i < n; ++i) {
            R ai = 0;
            for (int j = 0; j < n; ++j) ai += R(A[i * n + j]) * z[j];
            g[i] = ai - R(b[i]);
            fz += R(0.5) * z[i] * ai - R(b[i]) * z[i];
        }
        return fz;
    };
    auto inc = [&](std::vector<R>& z, const R* p, R a) {
        for (int i = 0; i < n; ++i) z[i] += a * p[i];
    };
    BfgsTrace tr;
    int ec = 0;
    R e = bfgs_core<R>(ev, inc, xv, n, max_steps, &ec, f0_zero != 0, &tr);
    for (int i = 0; i < n; ++i) x[i] = double(xv[i]);
    if (evals) *evals = ec;
    if (steps) *steps = tr.steps;
    for (int k = 0; k < tr.steps; ++k) {
        if (alpha) alpha[k] = tr.alpha[k];
        if (f) f[k] = tr.f[k];
        if (trials) trials[k] = tr.trials[k];
    }
    if (H) std::copy(tr.H.begin(), tr.H.end(), H);
    if (sv) std::copy(tr.s.begin(), tr.s.end(), sv);
    if (yv) std::copy(tr.y.begin(), tr.y.end(), yv);
    return double(e);
}
extern "C" double orc_bfgs_quadratic(int prec, int n, const double* A, const double* b, double c, double* x,
                                     int max_steps, int f0_zero, int* evals, int* steps, double* alpha, double* f,
                                     int* trials, double* H, double* sv, double* yv) {
    if (prec == 64) return bfgs_quad_impl<double>(n, A, b, c, x, max_steps, f0_zero, evals, steps, alpha, f, trials, H, sv, yv);
    return bfgs_quad_impl<float>(n, A, b, c, x, max_steps, f0_zero, evals, steps, alpha, f, trials, H, sv, yv);
}
// first-line-search rule of every BFGS on this handle: 0 = f0 = f(x)
// (bfgs.h:107), 1 = f0 = 0 (bfgs_warp, warp_ops.cuh:1062-1076)
extern "C" void orc_set_ls_f0_zero(OrcHandle* h, int on) {
    h->d.ls_f0_zero = on != 0;
    h->f.ls_f0_zero = on != 0;
}
// metropolis_accept (monte_carlo.cu:92-97) with uniform u
extern "C" int orc_metropolis(double old_f, double new_f, double temperature, double u) {
    return metropolis<double>(old_f, new_f, temperature, u) ? 1 : 0;
}

template <class R> static double mc_impl(Ctx<R>& C, int lig, uint64_t seed, uint32_t chain, int gsteps, int lsteps, double* best_pose, int* evals, double* trace) {
    const Ligand<R>& L = C.ligs[lig];
    Pose<R> best;
    int ec = 0;
    std::vector<R> tr(gsteps);
    R e = mc_chain(C, L, seed, chain, gsteps, lsteps, R(2), R(1.2), best, &ec, tr.data());
    store_pose(best, best_pose);
    if (evals) *evals = ec;
    if (trace) for (int i = 0; i < gsteps; ++i) trace[i] = tr[i];
    return e;
}
// one MC chain (global chain id `chain`); returns best energy, writes best pose
extern "C" double orc_mc_chain(OrcHandle* h, int lig, uint64_t seed, uint32_t chain, int gsteps, int lsteps, double* best_pose, int* evals, double* trace) {
    if (h->prec == 64) return mc_impl(h->d, lig, seed, chain, gsteps, lsteps, best_pose, evals, trace);
    return mc_impl(h->f, lig, seed, chain, gsteps, lsteps, best_pose, evals, trace);
}
// random start pose of a chain (exposed to check the device generator)
extern "C" void orc_random_start(OrcHandle* h, int lig, uint64_t seed, uint32_t chain, double* pose) {
    const Ligand<double>& L = h->d.ligs[lig];
    double lo[3], hi[3];
    for (int i = 0; i < 3; ++i) { lo[i] = h->d.grid.init[i]; hi[i] = h->d.grid.init[i] + h->d.grid.factor_inv[i] * (h->d.grid.n[i] - 1); }
    Pose<double> P;
    random_start(L, lo, hi, seed, chain, P);
    store_pose(P, pose);
}
// mutation of a given pose (exposed to check the device mutation)
extern "C" void orc_mutate(OrcHandle* h, int lig, double* pose, double rgyr, uint64_t seed, uint32_t chain, uint32_t step) {
    const Ligand<double>& L = h->d.ligs[lig];
    Pose<double> P;
    load_pose(L, pose, P);
    mutate(L, P, rgyr, 2.0, seed, chain, step);
    store_pose(P, pose);
}
// inter / intra terms with authentic caps (for the final affinity)
extern "C" void orc_score_terms(OrcHandle* h, int lig, const double* pose, double* inter, double* intra) {
    if (h->prec == 64) {
        Pose<double> P; load_pose(h->d.ligs[lig], pose, P);
        score_terms(h->d, h->d.ligs[lig], P, inter, intra);
    } else {
        Pose<float> P; load_pose(h->f.ligs[lig], pose, P);
        float a, b;
        score_terms(h->f, h->f.ligs[lig], P, &a, &b);
        *inter = a; *intra = b;
    }
}
template <class R> static double refine_impl(Ctx<R>& C, int lig, double* pose, int refine_step, int* within) {
    const Ligand<R>& L = C.ligs[lig];
    Pose<R> P;
    load_pose(L, pose, P);
    bool in = false;
    R e = refine(C, L, P, refine_step, &in);
    store_pose(P, pose);
    *within = in ? 1 : 0;
    return e;
}
// refinement of one pose against the explicit receptor (pose in/out)
extern "C" double orc_refine(OrcHandle* h, int lig, double* pose, int refine_step, int* within) {
    if (h->prec == 64) return refine_impl(h->d, lig, pose, refine_step, within);
    return refine_impl(h->f, lig, pose, refine_step, within);
}
template <class R> static void terms_nc_impl(Ctx<R>& C, int lig, const double* pose, double* inter, double* intra) {
    Pose<R> P;
    load_pose(C.ligs[lig], pose, P);
    R a, b;
    score_terms_nc(C, C.ligs[lig], P, &a, &b);
    *inter = a;
    *intra = b;
}
extern "C" void orc_score_terms_nc(OrcHandle* h, int lig, const double* pose, double* inter, double* intra) {
    if (h->prec == 64) terms_nc_impl(h->d, lig, pose, inter, intra);
    else terms_nc_impl(h->f, lig, pose, inter, intra);
}
// energy + gradient against the explicit receptor with box slope `slope`
extern "C" double orc_eval_deriv_nc(OrcHandle* h, int lig, const double* pose, const double* v, double slope,
                                    double* grad) {
    double r;
    if (h->prec == 64) {
        h->d.use_nc = true; h->d.nc_slope = slope;
        r = eval_impl(h->d, lig, pose, v, grad, nullptr, nullptr);
        h->d.use_nc = false;
    } else {
        h->f.use_nc = true; h->f.nc_slope = float(slope);
        r = eval_impl(h->f, lig, pose, v, grad, nullptr, nullptr);
        h->f.use_nc = false;
    }
    return r;
}
// one atom of XS type t at loc against the explicit receptor (deriv NULL: the
// energy-only form with the fast table)
extern "C" double orc_nc_atom(OrcHandle* h, int t, const double* loc, double slope, double v, double* deriv) {
    if (h->prec == 64) {
        V3<double> d;
        double e = nc_atom(h->d, t, V3<double>(loc[0], loc[1], loc[2]), slope, v, deriv ? &d : nullptr);
        if (deriv) for (int k = 0; k < 3; ++k) deriv[k] = d[k];
        return e;
    }
    V3<float> d;
    float e = nc_atom(h->f, t, V3<float>(float(loc[0]), float(loc[1]), float(loc[2])), float(slope), float(v),
                      deriv ? &d : nullptr);
    if (deriv) for (int k = 0; k < 3; ++k) deriv[k] = d[k];
    return e;
}
// AD4.2 data of ligand `lig` (AD types, charges, TORSDOF); builds the
// per-pair tables.  The context must have been created with SF_AD4.
template <class R> static void set_ad4(Ctx<R>& C, int lig, const int* ad, const double* q, double torsdof) {
    Ligand<R>& L = C.ligs[lig];
    L.ad.assign(ad, ad + L.n_atoms);
    L.q.resize(L.n_atoms);
    for (int i = 0; i < L.n_atoms; ++i) L.q[i] = R(q[i]);
    L.torsdof = R(torsdof);
    L.ptab.assign(L.n_pairs, PairTable<R>());
    std::atomic<int> next(0);
    auto work = [&] {
        for (int k = next.fetch_add(1); k < L.n_pairs; k = next.fetch_add(1))
            L.ptab[k].build_ad4(L.ad[L.pa[k]], L.q[L.pa[k]], L.ad[L.pb[k]], L.q[L.pb[k]], C.w);
    };
    std::vector<std::thread> th;
    const int nt = std::max(1, std::min<int>(16, int(std::thread::hardware_concurrency())));
    for (int i = 0; i < nt; ++i) th.emplace_back(work);
    for (auto& t : th) t.join();
}
extern "C" int orc_set_ad4(OrcHandle* h, int lig, const int* ad, const double* q, double torsdof) {
    if (h->d.sf != SF_AD4 || lig < 0 || lig >= int(h->d.ligs.size())) return 1;
    set_ad4(h->f, lig, ad, q, torsdof);
    set_ad4(h->d, lig, ad, q, torsdof);
    return 0;
}
extern "C" void orc_build_maps_ad4(OrcHandle* h, int n_rec, const double* xyz, const int* ad, const double* q,
                                   int n_types, const int* types) {
    if (h->prec == 64) build_maps_ad4(h->d, n_rec, xyz, ad, q, n_types, types);
    else build_maps_ad4(h->f, n_rec, xyz, ad, q, n_types, types);
}
// AD4.2 potentials at r for atoms (t1, q1), (t2, q2): out = vdw, hb, elec
// (with the charges), solvation, weighted sum
extern "C" void orc_ad4_terms(OrcHandle* h, int t1, double q1, int t2, double q2, double r, double* out) {
    out[0] = ad4_vdw<double>(t1, t2, r);
    out[1] = ad4_hb<double>(t1, t2, r);
    out[2] = q1 * q2 * ad4_elec_r<double>(r);
    out[3] = ad4_solv_coef<double>(t1, q1, t2, q2) * ad4_solv_r<double>(r);
    out[4] = sf_eval_ad4<double>(t1, q1, t2, q2, r, h->d.w);
}
// the AD4.2 kind of AD type t: radius, depth, hb_depth, hb_r, solv, vol
extern "C" void orc_ad4_kind(int t, double* out) {
    const AdKind& k = kAd[t];
    out[0] = k.radius; out[1] = k.depth; out[2] = k.hb_depth; out[3] = k.hb_r; out[4] = k.solv; out[5] = k.vol;
}
// sample tables of one AD4.2 atom pair (table_n(SF_AD4) doubles each)
extern "C" int orc_pair_table_ad4(OrcHandle* h, int t1, double q1, int t2, double q2, double* e, double* dor, double* fast) {
    PairTable<double> T;
    T.build_ad4(t1, q1, t2, q2, h->d.w);
    for (size_t i = 0; i < T.e.size(); ++i) { e[i] = T.e[i]; dor[i] = T.dor[i]; fast[i] = T.fast[i]; }
    return int(T.e.size());
}
extern "C" double orc_conf_independent(OrcHandle* h, double num_tors, double x) {
    if (h->prec == 64) return conf_independent(h->d, num_tors, x);
    return conf_independent(h->f, float(num_tors), float(x));
}

// CPU docking baseline: exhaustiveness chains of every listed ligand spread
// over n_threads std::threads (parallel_mc, parallel_mc.cpp:318-329).

#include <atomic>
#include <thread>
// energies of many poses (FP64 or FP32 per the handle), n_threads threads;
// poses[i * kPose] in the pose layout, ligs[i] the ligand of pose i
extern "C" void orc_eval_batch(OrcHandle* h, int n, const int* ligs, const double* poses, const double* v,
                               int n_threads, double* e_out) {
    for (int i = 0; i < n && h->d.sf != SF_AD4; ++i) {  // build the lazily made type-pair tables first
        if (h->prec == 64) { auto& L = h->d.ligs[ligs[i]]; for (int k = 0; k < L.n_pairs; ++k) h->d.table(L.xs[L.pa[k]], L.xs[L.pb[k]]); }
        else { auto& L = h->f.ligs[ligs[i]]; for (int k = 0; k < L.n_pairs; ++k) h->f.table(L.xs[L.pa[k]], L.xs[L.pb[k]]); }
    }
    std::atomic<int> next(0);
    auto work = [&] {
        for (int i = next.fetch_add(1); i < n; i = next.fetch_add(1)) {
            std::vector<double> g(6 + 64);
            e_out[i] = (h->prec == 64) ? eval_impl(h->d, ligs[i], poses + size_t(i) * kPose, v, g.data(), nullptr, nullptr)
                                       : eval_impl(h->f, ligs[i], poses + size_t(i) * kPose, v, g.data(), nullptr, nullptr);
        }
    };
    std::vector<std::thread> th;
    for (int i = 0; i < std::max(1, n_threads); ++i) th.emplace_back(work);
    for (auto& t : th) t.join();
}

// Every chain of n_lig ligands (chain ids ligand_position * exh + k), on
// n_threads threads: best energies, evaluation count and (when best_pose_out
// is not NULL) the best pose of each chain, kPose values per chain.
extern "C" double orc_dock_cpu(OrcHandle* h, int n_lig, const int* ligs, int exh, uint64_t seed, int gsteps,
                               const int* lsteps, int n_threads, double* best_e_out, long long* evals_out,
                               double* best_pose_out) {
    std::atomic<int> next(0);
    const int total = n_lig * exh;
    std::vector<long long> evals(n_threads, 0);
    std::vector<double> be(total, 0);
    auto work = [&](int tid) {
        for (;;) {
            int t = next.fetch_add(1);
            if (t >= total) break;
            int li = t / exh, ci = t % exh;
            int lig = ligs[li];
            int ec = 0;
            std::vector<double> pose(7 + 64);
            double e;
            if (h->prec == 64) e = mc_impl(h->d, lig, seed, uint32_t(li * exh + ci), gsteps, lsteps[li], pose.data(), &ec, nullptr);
            else e = mc_impl(h->f, lig, seed, uint32_t(li * exh + ci), gsteps, lsteps[li], pose.data(), &ec, nullptr);
            be[t] = e;
            evals[tid] += ec;
            if (best_pose_out) std::copy(pose.begin(), pose.begin() + kPose, best_pose_out + size_t(t) * kPose);
        }
    };
    // tables are built lazily; build them for all pairs used before threading
    for (int li = 0; li < n_lig; ++li) {
        if (h->prec == 64) { auto& L = h->d.ligs[ligs[li]]; for (int k = 0; k < L.n_pairs; ++k) h->d.table(L.xs[L.pa[k]], L.xs[L.pb[k]]); }
        else { auto& L = h->f.ligs[ligs[li]]; for (int k = 0; k < L.n_pairs; ++k) h->f.table(L.xs[L.pa[k]], L.xs[L.pb[k]]); }
    }
    std::vector<std::thread> th;
    for (int i = 0; i < n_threads; ++i) th.emplace_back(work, i);
    for (auto& t : th) t.join();
    long long ev = 0;
    for (auto e : evals) ev += e;
    if (evals_out) *evals_out = ev;
    if (best_e_out) for (int i = 0; i < total; ++i) best_e_out[i] = be[i];
    return 0;
}
