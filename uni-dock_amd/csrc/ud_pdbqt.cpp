// Ligand and receptor ingestion from PDBQT text: the torsion tree of
// ROOT / BRANCH blocks, bonds, X-Score atom types, the intramolecular pair
// list and the conf-independent torsion count, as the reference builds them
// when it parses a ligand (parse_pdbqt.cpp:107-126, 440-540) and initialises
// its model (model::assign_bonds / assign_types / initialize_pairs,
// model.cpp:343-640; conf_independent_inputs, conf_independent.cpp:36-103).
// Host code only; the result feeds ud_set_ligands / ud_compute_maps.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/unidock_amd.h"

namespace {

// AutoDock atom kinds (atom_constants.h:156-192): name, element class,
// covalent radius.  Element classes: 0 H, 1 C, 2 N, 3 O, 4 S, 5 P, 6 F, 7 Cl,
// 8 Br, 9 I, 10 Si, 11 At, 12 metal, 13 dummy.
struct AdKind {
    const char* name;
    int el;
    float cov;
};
enum { EL_H = 0, EL_C, EL_N, EL_O, EL_S, EL_P, EL_F, EL_Cl, EL_Br, EL_I, EL_Si, EL_At, EL_Met, EL_Dummy };
enum {
    AD_C = 0, AD_A, AD_N, AD_O, AD_P, AD_S, AD_H, AD_F, AD_I, AD_NA, AD_OA, AD_SA, AD_HD, AD_Mg, AD_Mn, AD_Zn,
    AD_Ca, AD_Fe, AD_Cl, AD_Br, AD_Si, AD_At, AD_G0, AD_G1, AD_G2, AD_G3, AD_CG0, AD_CG1, AD_CG2, AD_CG3, AD_W,
    AD_OXA, AD_NXA, AD_OXD, AD_NXD, AD_SIZE
};
const AdKind kAd[AD_SIZE] = {
    {"C", EL_C, 0.77f},   {"A", EL_C, 0.77f},   {"N", EL_N, 0.75f},    {"O", EL_O, 0.73f},   {"P", EL_P, 1.06f},
    {"S", EL_S, 1.02f},   {"H", EL_H, 0.37f},   {"F", EL_F, 0.71f},    {"I", EL_I, 1.33f},   {"NA", EL_N, 0.75f},
    {"OA", EL_O, 0.73f},  {"SA", EL_S, 1.02f},  {"HD", EL_H, 0.37f},   {"Mg", EL_Met, 1.30f}, {"Mn", EL_Met, 1.39f},
    {"Zn", EL_Met, 1.31f}, {"Ca", EL_Met, 1.74f}, {"Fe", EL_Met, 1.25f}, {"Cl", EL_Cl, 0.99f}, {"Br", EL_Br, 1.14f},
    {"Si", EL_Si, 1.11f}, {"At", EL_At, 1.44f}, {"G0", EL_Dummy, 0.77f}, {"G1", EL_Dummy, 0.77f},
    {"G2", EL_Dummy, 0.77f}, {"G3", EL_Dummy, 0.77f}, {"CG0", EL_C, 0.77f}, {"CG1", EL_C, 0.77f},
    {"CG2", EL_C, 0.77f}, {"CG3", EL_C, 0.77f}, {"W", EL_Dummy, 0.0f}, {"OXA", EL_O, 0.73f},
    {"NXA", EL_N, 0.75f}, {"OXD", EL_O, 0.73f}, {"NXD", EL_N, 0.75f}};
const float kMaxCov = 1.74f;  // largest covalent radius of the table
const float kBondAllowance = 1.1f;

// X-Score types (atom_constants.h:84-120)
enum {
    XS_C_H = 0, XS_C_P, XS_N_P, XS_N_D, XS_N_A, XS_N_DA, XS_O_P, XS_O_D, XS_O_A, XS_O_DA, XS_S_P, XS_P_P, XS_F_H,
    XS_Cl_H, XS_Br_H, XS_I_H, XS_Si, XS_At, XS_Met_D, XS_C_H_CG0, XS_C_P_CG0, XS_G0, XS_C_H_CG1, XS_C_P_CG1, XS_G1,
    XS_C_H_CG2, XS_C_P_CG2, XS_G2, XS_C_H_CG3, XS_C_P_CG3, XS_G3, XS_W, XS_O_XA, XS_N_XA, XS_O_XD, XS_N_XD
};

int ad_of(const std::string& name) {
    std::string n = name;
    if (n == "Se") n = "S";  // atom_equivalence_data (atom_constants.h:208)
    if (n == "CL") n = "Cl";
    for (int i = 0; i < AD_SIZE; ++i)
        if (n == kAd[i].name) return i;
    return -1;
}
// ad_is_heteroatom (atom_constants.h:227-230): everything but C, A, H, HD
bool ad_is_heteroatom(int ad) { return ad != AD_A && ad != AD_C && ad != AD_H && ad != AD_HD; }

struct Atom {
    int serial = 0;
    int ad = 0;
    int el = 0;
    float x = 0, y = 0, z = 0;
    int node = 0;
    int xs = -1;
    float charge = 0.f;  // partial charge (AD4.2 scoring)
};

std::string trim(const std::string& s) {
    size_t a = s.find_first_not_of(" \t\r"), b = s.find_last_not_of(" \t\r");
    return a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
}
// 1-based inclusive columns, as checked_convert_substring (parse_pdbqt.cpp:107-126)
std::string cols(const std::string& s, size_t i, size_t j) {
    if (s.size() < i) return std::string();
    return s.substr(i - 1, std::min(j, s.size()) - (i - 1));
}

bool parse_atom_line(const std::string& ln, Atom& a, std::string& err) {
    char* end = nullptr;
    const std::string sn = trim(cols(ln, 7, 11));
    a.serial = int(std::strtol(sn.c_str(), &end, 10));
    if (sn.empty() || *end) { err = "bad atom number: " + ln; return false; }
    float v[3];
    for (int d = 0; d < 3; ++d) {
        const std::string c = trim(cols(ln, 31 + 8 * d, 38 + 8 * d));
        v[d] = std::strtof(c.c_str(), &end);
        if (c.empty() || *end) { err = "bad coordinate: " + ln; return false; }
    }
    a.x = v[0]; a.y = v[1]; a.z = v[2];
    // partial charge, columns 69-76 when not blank (parse_pdbqt_atom_string, parse_pdbqt.cpp:112-114)
    const std::string q = trim(cols(ln, 69, 76));
    a.charge = 0.f;
    if (!q.empty()) {
        a.charge = std::strtof(q.c_str(), &end);
        if (*end) { err = "Charge \"" + q + "\" is not valid: " + ln; return false; }
    }
    const std::string t = trim(cols(ln, 78, 79));
    a.ad = ad_of(t);
    if (a.ad < 0) { err = "atom type " + t + " is not a valid AutoDock type: " + ln; return false; }
    a.el = kAd[a.ad].el;
    return true;
}

float dist2(const Atom& a, const Atom& b) {
    const float dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
    return dx * dx + dy * dy + dz * dz;
}

// model::assign_bonds (model.cpp:343-409): i-j is a bond when their mobility
// is not variable, r < 1.1 (R_i + R_j), and no atom immobile relative to both
// lies closer to both of them than they are to each other.  `fixed(i, j)`
// tells whether two atoms cannot move relative to each other.
template <class Fixed>
void assign_bonds(const std::vector<Atom>& at, Fixed fixed, std::vector<std::vector<int>>& bonds) {
    const int n = int(at.size());
    bonds.assign(n, {});
    // cell list of edge 2 * 1.1 * kMaxCov, enough for every bond candidate
    const float cell = 2.f * kBondAllowance * kMaxCov;
    std::map<long long, std::vector<int>> grid;
    auto key = [&](int kx, int ky, int kz) { return (static_cast<long long>(kx) * 1000003LL + ky) * 1000003LL + kz; };
    std::vector<int> cx(n), cy(n), cz(n);
    for (int i = 0; i < n; ++i) {
        cx[i] = int(std::floor(at[i].x / cell));
        cy[i] = int(std::floor(at[i].y / cell));
        cz[i] = int(std::floor(at[i].z / cell));
        grid[key(cx[i], cy[i], cz[i])].push_back(i);
    }
    for (int i = 0; i < n; ++i) {
        const float ri = kAd[at[i].ad].cov;
        std::vector<int> rel;  // relevant atoms: immobile relative to i and near it
        const float cut2 = (kBondAllowance * (ri + kMaxCov)) * (kBondAllowance * (ri + kMaxCov));
        for (int dx = -1; dx <= 1; ++dx)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dz = -1; dz <= 1; ++dz) {
                    auto it = grid.find(key(cx[i] + dx, cy[i] + dy, cz[i] + dz));
                    if (it == grid.end()) continue;
                    for (int j : it->second)
                        if (j != i && fixed(i, j) && dist2(at[i], at[j]) < cut2) rel.push_back(j);
                }
        std::sort(rel.begin(), rel.end());
        for (int j : rel) {
            if (j <= i) continue;
            const float bl = ri + kAd[at[j].ad].cov;
            const float r2 = dist2(at[i], at[j]);
            if (!(r2 < (kBondAllowance * bl) * (kBondAllowance * bl))) continue;
            bool between = false;  // atom_exists_between (model.cpp:307-321)
            for (int c : rel) {
                if (c == j) continue;
                if (fixed(j, c) && dist2(at[i], at[c]) < r2 && dist2(at[j], at[c]) < r2) {
                    between = true;
                    break;
                }
            }
            if (between) continue;
            bonds[i].push_back(j);
            bonds[j].push_back(i);
        }
    }
}

// model::assign_types (model.cpp:427-530)
void assign_types(std::vector<Atom>& at, const std::vector<std::vector<int>>& bonds) {
    for (size_t i = 0; i < at.size(); ++i) {
        Atom& a = at[i];
        bool hd = false, hetero = false;
        for (int j : bonds[i]) {
            if (at[j].ad == AD_HD) hd = true;
            if (ad_is_heteroatom(at[j].ad)) hetero = true;
        }
        const bool acceptor = a.ad == AD_OA || a.ad == AD_NA;
        const bool donor = a.el == EL_Met || hd;
        switch (a.el) {
            case EL_H: a.xs = -1; break;
            case EL_C:
                if (a.ad == AD_CG0) a.xs = hetero ? XS_C_P_CG0 : XS_C_H_CG0;
                else if (a.ad == AD_CG1) a.xs = hetero ? XS_C_P_CG1 : XS_C_H_CG1;
                else if (a.ad == AD_CG2) a.xs = hetero ? XS_C_P_CG2 : XS_C_H_CG2;
                else if (a.ad == AD_CG3) a.xs = hetero ? XS_C_P_CG3 : XS_C_H_CG3;
                else a.xs = hetero ? XS_C_P : XS_C_H;
                break;
            case EL_N:
                if (a.ad == AD_NXA) a.xs = XS_N_XA;
                else if (a.ad == AD_NXD) a.xs = XS_N_XD;
                else a.xs = (acceptor && donor) ? XS_N_DA : (acceptor ? XS_N_A : (donor ? XS_N_D : XS_N_P));
                break;
            case EL_O:
                // the reference maps OXA / OXD to the N_X types (model.cpp:465-470)
                if (a.ad == AD_OXA) a.xs = XS_N_XA;
                else if (a.ad == AD_OXD) a.xs = XS_N_XD;
                else a.xs = (acceptor && donor) ? XS_O_DA : (acceptor ? XS_O_A : (donor ? XS_O_D : XS_O_P));
                break;
            case EL_S: a.xs = XS_S_P; break;
            case EL_P: a.xs = XS_P_P; break;
            case EL_F: a.xs = XS_F_H; break;
            case EL_Cl: a.xs = XS_Cl_H; break;
            case EL_Br: a.xs = XS_Br_H; break;
            case EL_I: a.xs = XS_I_H; break;
            case EL_Si: a.xs = XS_Si; break;
            case EL_At: a.xs = XS_At; break;
            case EL_Met: a.xs = XS_Met_D; break;
            default:  // dummies: closure atoms G0..G3 keep their types, W has none
                if (a.ad == AD_G0) a.xs = XS_G0;
                else if (a.ad == AD_G1) a.xs = XS_G1;
                else if (a.ad == AD_G2) a.xs = XS_G2;
                else if (a.ad == AD_G3) a.xs = XS_G3;
                else a.xs = -1;
        }
    }
}

struct Node {
    int parent = -1;
    int axis_begin = -1;  // atom (in the parent) the rotatable bond starts from
    int axis_end = -1;    // the branch's atom on the other end of that bond: the node's origin
    std::vector<int> atoms;  // input order
};

// A parsed torsion tree: atoms in input order, nodes in creation order (root
// first), each branch created after its parent.
struct Tree {
    std::vector<Atom> fa;
    std::vector<Node> nodes;
    int torsdof = -1;
};

struct Ligand {
    int format = 0;                  // 0 PDBQT, 1 SDF
    std::vector<std::string> lines;  // input lines kept for output (trailing blank lines dropped)
    std::vector<int> line_atom;      // batch atom index of an atom line, else -1
    std::vector<Atom> atoms;  // reference order (nodes in pre-order, input order inside)
    std::vector<int> node_parent, node_abegin, node_aend;
    std::vector<float> node_origin, node_axis;
    std::vector<int> pairs;      // X-Score typed atoms (Vina, Vinardo)
    std::vector<int> pairs_ad4;  // AutoDock typed atoms, hydrogens included (AD4.2)
    float num_tors = 0;
    int torsdof = 0;
};

// ROOT / BRANCH blocks of a PDBQT ligand (parse_pdbqt_aux / parse_pdbqt_branch,
// parse_pdbqt.cpp:440-540).
bool parse_pdbqt_tree(const std::string& text, Tree& T, std::string& err) {
    std::vector<Atom>& fa = T.fa;
    std::vector<Node>& nodes = T.nodes;
    std::vector<int> stack;
    std::map<int, int> by_serial;
    bool root_seen = false;
    std::istringstream in(text);
    std::string ln;
    while (std::getline(in, ln)) {
        if (!ln.empty() && ln.back() == '\r') ln.pop_back();
        if (ln.rfind("ATOM", 0) == 0 || ln.rfind("HETATM", 0) == 0) {
            if (stack.empty()) { err = "atom outside ROOT / BRANCH: " + ln; return false; }
            Atom a;
            if (!parse_atom_line(ln, a, err)) return false;
            a.node = stack.back();
            if (by_serial.count(a.serial)) { err = "duplicate atom number: " + ln; return false; }
            by_serial[a.serial] = int(fa.size());
            nodes[a.node].atoms.push_back(int(fa.size()));
            fa.push_back(a);
        } else if (ln.rfind("ROOT", 0) == 0) {
            if (root_seen) { err = "second ROOT"; return false; }
            root_seen = true;
            nodes.push_back(Node());
            stack.push_back(0);
        } else if (ln.rfind("ENDROOT", 0) == 0) {
            if (stack.size() != 1) { err = "misplaced ENDROOT"; return false; }
            stack.pop_back();
        } else if (ln.rfind("BRANCH", 0) == 0) {
            std::istringstream ss(ln.substr(6));
            int a = 0, b = 0;
            if (!(ss >> a >> b)) { err = "bad BRANCH line: " + ln; return false; }
            auto it = by_serial.find(a);
            if (it == by_serial.end()) { err = "BRANCH from an unknown atom: " + ln; return false; }
            Node nd;
            nd.parent = fa[it->second].node;
            nd.axis_begin = it->second;
            nd.axis_end = b;  // serial for now, resolved after parsing
            nodes.push_back(nd);
            stack.push_back(int(nodes.size()) - 1);
        } else if (ln.rfind("ENDBRANCH", 0) == 0) {
            if (stack.empty() || stack.back() == 0) { err = "misplaced ENDBRANCH"; return false; }
            stack.pop_back();
        } else if (ln.rfind("TORSDOF", 0) == 0) {
            T.torsdof = std::atoi(ln.c_str() + 7);
        } else if (ln.rfind("MODEL", 0) == 0) {
            err = "multi-MODEL ligand text: split it first";
            return false;
        }
    }
    if (!root_seen || nodes.empty() || nodes[0].atoms.empty()) { err = "no ROOT atoms"; return false; }
    if (!stack.empty()) { err = "unterminated ROOT or BRANCH"; return false; }
    for (size_t k = 1; k < nodes.size(); ++k) {
        auto it = by_serial.find(nodes[k].axis_end);
        if (it == by_serial.end() || fa[it->second].node != int(k)) {
            err = "atom " + std::to_string(nodes[k].axis_end) + " has not been found in this branch";
            return false;
        }
        nodes[k].axis_end = it->second;
    }
    return true;
}

// The model the reference builds from a parsed tree (postprocess_ligand /
// postprocess_branch, parse_pdbqt.cpp:498-566, then model::assign_bonds,
// assign_types, initialize_pairs and the conf-independent torsion count).
bool finish_ligand(const Tree& T, Ligand& L, std::string& err) {
    const std::vector<Atom>& fa = T.fa;
    const std::vector<Node>& nodes = T.nodes;
    // pre-order: a node's children ordered by the position of their first axis
    // atom in the node, then creation order (postprocess_branch, parse_pdbqt.cpp:498-533)
    std::vector<std::vector<int>> kids(nodes.size());
    for (size_t k = 1; k < nodes.size(); ++k) kids[nodes[k].parent].push_back(int(k));
    for (size_t k = 0; k < nodes.size(); ++k) {
        const std::vector<int>& na = nodes[k].atoms;
        auto pos = [&](int child) { return int(std::find(na.begin(), na.end(), nodes[child].axis_begin) - na.begin()); };
        std::stable_sort(kids[k].begin(), kids[k].end(), [&](int a, int b) { return pos(a) < pos(b); });
    }
    std::vector<int> order;
    std::vector<int> st{0};
    while (!st.empty()) {
        const int k = st.back();
        st.pop_back();
        order.push_back(k);
        for (auto it = kids[k].rbegin(); it != kids[k].rend(); ++it) st.push_back(*it);
    }
    std::vector<int> new_node(nodes.size()), new_atom(fa.size(), -1);
    for (size_t k = 0; k < order.size(); ++k) new_node[order[k]] = int(k);
    L.atoms.clear();
    for (int k : order) {
        const int nk = new_node[k];
        L.node_abegin.push_back(int(L.atoms.size()));
        for (int i : nodes[k].atoms) {
            new_atom[i] = int(L.atoms.size());
            Atom a = fa[i];
            a.node = nk;
            L.atoms.push_back(a);
        }
        L.node_aend.push_back(int(L.atoms.size()));
        L.node_parent.push_back(k == 0 ? -1 : new_node[nodes[k].parent]);
    }
    const int n = int(L.atoms.size());
    // node origins and axes: the root at its first atom, a branch at its
    // second axis atom with the axis from the parent's bond atom (segment,
    // tree.h:101-131; postprocess_ligand / postprocess_branch)
    for (int k : order) {
        const Atom& o = (k == 0) ? fa[nodes[0].atoms.front()] : fa[nodes[k].axis_end];
        L.node_origin.insert(L.node_origin.end(), {o.x, o.y, o.z});
        if (k == 0) {
            L.node_axis.insert(L.node_axis.end(), {1.f, 0.f, 0.f});
        } else {
            const Atom& b0 = fa[nodes[k].axis_begin];
            float ax = o.x - b0.x, ay = o.y - b0.y, az = o.z - b0.z;
            const float nr = std::sqrt(ax * ax + ay * ay + az * az);
            if (!(nr > 1e-6f)) { err = "zero-length rotatable bond"; return false; }
            L.node_axis.insert(L.node_axis.end(), {ax / nr, ay / nr, az / nr});
        }
    }
    // mobility (parse_pdbqt.cpp:498-533): the reference stores a branch's
    // second axis atom with its parent's atoms (it lies on the rotation axis,
    // node::insert_immobiles), marks both axis atoms fixed relative to every
    // atom stored in the branch (add_bonds) and the axis pair as the rotor.
    const int nn = int(order.size());
    std::vector<int> axis_b(nn, -1), axis_e(nn, -1);
    for (int k = 1; k < int(nodes.size()); ++k) {
        axis_b[new_node[k]] = new_atom[nodes[k].axis_begin];
        axis_e[new_node[k]] = new_atom[nodes[k].axis_end];
    }
    std::vector<int> body(n);
    for (int i = 0; i < n; ++i) body[i] = L.atoms[i].node;
    for (int k = 1; k < nn; ++k) body[axis_e[k]] = L.node_parent[k];
    auto axis_of = [&](int i, int k) { return k > 0 && (i == axis_b[k] || i == axis_e[k]); };
    auto fixed = [&](int i, int j) { return body[i] == body[j] || axis_of(i, body[j]) || axis_of(j, body[i]); };
    std::vector<std::vector<int>> bonds;
    assign_bonds(L.atoms, fixed, bonds);
    assign_types(L.atoms, bonds);
    auto rotatable = [&](int i, int j) {
        const int ni = L.atoms[i].node, nj = L.atoms[j].node;
        return (axis_b[nj] == i && axis_e[nj] == j) || (axis_b[ni] == j && axis_e[ni] == i);
    };
    // pairs (model::initialize_pairs, model.cpp:592-630): variable mobility,
    // more than 3 bonds apart, both atoms X-Score typed, not a closure clash
    // or a glue pair
    auto within3 = [&](int i) {
        std::vector<int> seen{i}, frontier{i};
        for (int d = 0; d < 3; ++d) {
            std::vector<int> next;
            for (int u : frontier)
                for (int v : bonds[u])
                    if (std::find(seen.begin(), seen.end(), v) == seen.end()) {
                        seen.push_back(v);
                        next.push_back(v);
                    }
            frontier.swap(next);
        }
        return seen;
    };
    auto is_cg = [](int ad, int k) { return ad == AD_CG0 + k; };
    auto is_g = [](int ad, int k) { return ad == AD_G0 + k; };
    auto glue = [&](int i, int j) {
        for (int k = 0; k < 4; ++k)
            if ((is_cg(L.atoms[i].ad, k) && is_g(L.atoms[j].ad, k)) || (is_cg(L.atoms[j].ad, k) && is_g(L.atoms[i].ad, k)))
                return true;
        return false;
    };
    auto closure_clash = [&](int i, int j) {  // model::is_closure_clash (model.cpp:552-582)
        if (glue(i, j)) return false;
        bool has[4] = {false, false, false, false};
        std::vector<int> ni{i};
        ni.insert(ni.end(), bonds[i].begin(), bonds[i].end());
        for (int u : ni)
            for (int k = 0; k < 4; ++k)
                if (is_cg(L.atoms[u].ad, k)) has[k] = true;
        std::vector<int> nj{j};
        nj.insert(nj.end(), bonds[j].begin(), bonds[j].end());
        for (int u : nj)
            for (int k = 0; k < 4; ++k)
                if (is_cg(L.atoms[u].ad, k) && has[k]) return true;
        return false;
    };
    // the typing in use decides which atoms are typed (model.cpp:608-612):
    // every atom has an AutoDock type; hydrogens have no X-Score type
    for (int i = 0; i < n; ++i) {
        const std::vector<int> near = within3(i);
        for (int j = i + 1; j < n; ++j) {
            if (fixed(i, j)) continue;
            if (std::find(near.begin(), near.end(), j) != near.end()) continue;
            if (closure_clash(i, j)) continue;
            if (glue(i, j)) continue;  // glue pairs are not ligand pairs
            L.pairs_ad4.push_back(i);
            L.pairs_ad4.push_back(j);
            if (L.atoms[i].xs < 0 || L.atoms[j].xs < 0) continue;
            L.pairs.push_back(i);
            L.pairs.push_back(j);
        }
    }
    // conf_independent_inputs::num_tors (conf_independent.cpp:36-103): half the
    // rotatable bonds of heavy atoms to heavy atoms that have more than one
    // heavy neighbour
    auto heavy_deg = [&](int i) {
        int d = 0;
        for (int j : bonds[i]) d += L.atoms[j].el != EL_H;
        return d;
    };
    float nt = 0.f;
    for (int i = 0; i < n; ++i) {
        if (L.atoms[i].el == EL_H) continue;
        int ar = 0;
        for (int j : bonds[i])
            if (rotatable(i, j) && L.atoms[j].el != EL_H && heavy_deg(j) > 1) ++ar;
        nt += 0.5f * float(ar);
    }
    L.num_tors = nt;
    L.torsdof = T.torsdof < 0 ? int(nodes.size()) - 1 : T.torsdof;
    return true;
}

bool build_ligand(const std::string& text, Ligand& L, std::string& err) {
    Tree T;
    if (!parse_pdbqt_tree(text, T, err) || !finish_ligand(T, L, err)) return false;
    L.format = 0;
    // the input lines, for writing poses back out
    std::istringstream in2(text);
    std::string l2;
    std::map<int, int> batch_of_serial;
    for (size_t i = 0; i < L.atoms.size(); ++i) batch_of_serial[L.atoms[i].serial] = int(i);
    while (std::getline(in2, l2)) {
        if (!l2.empty() && l2.back() == '\r') l2.pop_back();
        int idx = -1;
        if (l2.rfind("ATOM", 0) == 0 || l2.rfind("HETATM", 0) == 0) {
            const std::string sn = trim(cols(l2, 7, 11));
            idx = batch_of_serial[std::atoi(sn.c_str())];
        }
        L.lines.push_back(l2);
        L.line_atom.push_back(idx);
    }
    while (!L.lines.empty() && trim(L.lines.back()).empty()) {
        L.lines.pop_back();
        L.line_atom.pop_back();
    }
    return true;
}

// whitespace-separated integers of one line of an SDF data field
bool sdf_ints(const std::string& ln, std::vector<int>& out) {
    out.clear();
    std::istringstream ss(ln);
    std::string tok;
    while (ss >> tok) {
        char* end = nullptr;
        const long v = std::strtol(tok.c_str(), &end, 10);
        if (*end) return false;
        out.push_back(int(v));
    }
    return true;
}

// One SDF record (lines without the "$$$$" terminator) with the torsion tree
// of its fragInfo / torsionInfo / atomInfo data fields (parse_sdf_aux,
// parse_sdf_branch_aux / parse_sdf_branch, parse_pdbqt.cpp:129-145, 376-420,
// 611-864, 962-1000): atoms from the V2000 atom block (coordinates in columns
// 1-10, 11-20, 21-30, element at 32-33), AutoDock types from atomInfo where
// given, the root = fragment 0 in its listed order, and each torsion
// (a, b, fragment of a, fragment of b) that starts in an expanded fragment
// opening the other fragment as a branch, depth first, fragments in their
// listed order.  Without fragInfo the ligand is rigid (every atom in the
// root, torsdof 0).  Bond lines are not read: bonds come from the geometry as
// for PDBQT.  Hydrogens are kept (keep_nonpolar_H, main.cpp:412, 531).
bool build_ligand_sdf(const std::vector<std::string>& lines, Ligand& L, std::string& err) {
    if (lines.size() < 4) { err = "SDF record shorter than its header"; return false; }
    const std::string& counts = lines[3];
    char* end = nullptr;
    const std::string sa = trim(cols(counts, 1, 3)), sb = trim(cols(counts, 4, 6));
    const long n_atoms = std::strtol(sa.c_str(), &end, 10);
    if (sa.empty() || *end || n_atoms <= 0) { err = "bad SDF counts line: " + counts; return false; }
    const long n_bonds = std::strtol(sb.c_str(), &end, 10);
    if (sb.empty() || *end || n_bonds < 0) { err = "bad SDF counts line: " + counts; return false; }
    if (lines.size() < size_t(4 + n_atoms + n_bonds)) { err = "SDF atom / bond block truncated"; return false; }
    Tree T;
    std::vector<Atom> all(n_atoms);
    for (long i = 0; i < n_atoms; ++i) {
        const std::string& ln = lines[4 + i];
        Atom& a = all[i];
        a.serial = int(i + 1);
        float v[3];
        for (int d = 0; d < 3; ++d) {
            const std::string c = trim(cols(ln, 1 + 10 * d, 10 + 10 * d));
            v[d] = std::strtof(c.c_str(), &end);
            if (c.empty() || *end) { err = "bad SDF coordinate: " + ln; return false; }
        }
        a.x = v[0]; a.y = v[1]; a.z = v[2];
        std::string el = ln.size() > 32 ? ln.substr(31, 2) : (ln.size() > 31 ? ln.substr(31, 1) : "");
        if (el.size() == 2 && el[1] == ' ') el = el.substr(0, 1);
        a.ad = ad_of(el);
        if (a.ad < 0) { err = "element " + el + " is not a valid AutoDock type: " + ln; return false; }
    }
    // data fields after "M  END"
    size_t li = size_t(4 + n_atoms + n_bonds);
    while (li < lines.size() && lines[li].find("M  END") == std::string::npos) ++li;
    std::vector<std::vector<int>> frags, tors;
    std::vector<int> ints;
    for (++li; li < lines.size(); ++li) {
        const std::string& hd = lines[li];
        if (hd.empty() || hd[0] != '>') continue;
        const bool info = hd.find("atomInfo") != std::string::npos || hd.find("atom_info") != std::string::npos;
        const bool tor = !info && hd.find("torsion") != std::string::npos;
        const bool frag = !info && !tor && hd.find("frag") != std::string::npos;
        for (++li; li < lines.size() && !lines[li].empty(); ++li) {
            const std::string& ln = lines[li];
            if (info) {
                const size_t sp = ln.find(' ');
                const std::string sid = trim(cols(ln, 1, std::min<size_t>(sp == std::string::npos ? ln.size() : sp, 3)));
                const long id = std::strtol(sid.c_str(), &end, 10);
                if (sid.empty() || *end || id < 1 || id > n_atoms) { err = "bad atomInfo line: " + ln; return false; }
                const std::string name = trim(ln.size() >= 14 ? ln.substr(13) : std::string());
                const int ad = ad_of(name);
                if (ad < 0) { err = "atom type " + name + " is not a valid AutoDock type: " + ln; return false; }
                all[id - 1].ad = ad;
                // partial charge in columns 4-13 (parse_sdf_aux, parse_pdbqt.cpp:678-680)
                const std::string qs = trim(cols(ln, 4, 13));
                const float qv = std::strtof(qs.c_str(), &end);
                if (qs.empty() || *end) { err = "bad partial charge in atomInfo line: " + ln; return false; }
                all[id - 1].charge = qv;
            } else if (tor || frag) {
                if (!sdf_ints(ln, ints) || ints.empty()) { err = "bad SDF data line: " + ln; return false; }
                if (tor && ints.size() != 4) { err = "torsionInfo lines hold 4 numbers: " + ln; return false; }
                (tor ? tors : frags).push_back(ints);
            }
        }
    }
    for (Atom& a : all) a.el = kAd[a.ad].el;
    if (frags.empty()) {
        T.nodes.push_back(Node());
        for (long i = 0; i < n_atoms; ++i) {
            T.nodes[0].atoms.push_back(int(i));
            T.fa.push_back(all[i]);
        }
        T.torsdof = 0;
    } else {
        const int nf = int(frags.size());
        for (const std::vector<int>& f : frags)
            for (int id : f)
                if (id < 1 || id > n_atoms) { err = "fragInfo names atom " + std::to_string(id) + " out of range"; return false; }
        for (const std::vector<int>& t : tors)
            if (t[0] < 1 || t[0] > n_atoms || t[1] < 1 || t[1] > n_atoms || t[2] < 0 || t[2] >= nf || t[3] < 0 ||
                t[3] >= nf) {
                err = "torsionInfo entry out of range";
                return false;
            }
        std::vector<int> tree_of(n_atoms, -1);  // SDF atom -> tree atom
        std::vector<char> been(nf, 0);
        auto add_frag = [&](int f, int node, int to_id, int& to_idx) -> bool {
            for (int id : frags[f]) {
                if (tree_of[id - 1] >= 0) { err = "atom " + std::to_string(id) + " is in two fragments"; return false; }
                tree_of[id - 1] = int(T.fa.size());
                if (id == to_id) to_idx = int(T.fa.size());
                Atom a = all[id - 1];
                a.node = node;
                T.nodes[node].atoms.push_back(int(T.fa.size()));
                T.fa.push_back(a);
            }
            return true;
        };
        T.nodes.push_back(Node());
        been[0] = 1;
        int dummy = -1;
        if (!add_frag(0, 0, -1, dummy)) return false;
        // depth-first expansion, as parse_sdf_branch recurses inside its loops
        std::function<bool(int, int)> expand = [&](int node, int f) -> bool {
            for (int id : frags[f])
                for (const std::vector<int>& t : tors) {
                    int from, to, next;
                    if (t[0] == id) { from = t[0]; to = t[1]; next = t[3]; }
                    else if (t[1] == id) { from = t[1]; to = t[0]; next = t[2]; }
                    else continue;
                    if (been[next]) continue;
                    been[next] = 1;
                    Node nd;
                    nd.parent = node;
                    nd.axis_begin = tree_of[from - 1];
                    T.nodes.push_back(nd);
                    const int k = int(T.nodes.size()) - 1;
                    int to_idx = -1;
                    if (!add_frag(next, k, to, to_idx)) return false;
                    if (to_idx < 0) { err = "atom " + std::to_string(to) + " has not been found in this branch"; return false; }
                    T.nodes[k].axis_end = to_idx;
                    if (!expand(k, next)) return false;
                }
            return true;
        };
        if (!expand(0, 0)) return false;
        T.torsdof = int(tors.size());
    }
    if (!finish_ligand(T, L, err)) return false;
    L.format = 1;
    // SDF atom index -> batch atom index (the tree's atom order was changed by finish_ligand)
    std::map<int, int> batch_of_serial;
    for (size_t i = 0; i < L.atoms.size(); ++i) batch_of_serial[L.atoms[i].serial] = int(i);
    for (size_t i = 0; i < lines.size(); ++i) {
        int idx = -1;
        if (i >= 4 && long(i) < 4 + n_atoms) {
            auto it = batch_of_serial.find(int(i - 4 + 1));
            if (it != batch_of_serial.end()) idx = it->second;
        }
        L.lines.push_back(lines[i]);
        L.line_atom.push_back(idx);
    }
    return true;
}

}  // namespace

struct ud_ligset {
    std::vector<Ligand> ligs;
    std::string err;
    // flattened batch arrays (rebuilt by ud_ligset_batch)
    std::vector<int32_t> atom_offset, node_offset, pair_offset, atom_xs, atom_heavy, node_parent, node_abegin,
        node_aend, pairs, atom_ad;
    std::vector<float> atom_local, node_origin, node_axis, num_tors, atom_charge, torsdof;
};

extern "C" {

int ud_ligset_create(ud_ligset** out) {
    if (!out) return UD_ERR_INVALID;
    *out = new ud_ligset();
    return UD_OK;
}

int ud_ligset_destroy(ud_ligset* s) {
    delete s;
    return UD_OK;
}

const char* ud_ligset_last_error(const ud_ligset* s) { return s ? s->err.c_str() : "null ligand set"; }

int ud_ligset_add_pdbqt(ud_ligset* s, const char* text, int32_t* index_out) {
    if (!s || !text) return UD_ERR_INVALID;
    Ligand L;
    std::string err;
    if (!build_ligand(text, L, err)) {
        s->err = err;
        return UD_ERR_INVALID;
    }
    if (int(L.atoms.size()) > UD_MAX_ATOMS || int(L.node_parent.size()) - 1 > UD_MAX_TORSIONS ||
        int(L.pairs_ad4.size() / 2) > UD_MAX_PAIRS) {
        s->err = "ligand exceeds the library's capacities";
        return UD_ERR_CAPACITY;
    }
    s->ligs.push_back(std::move(L));
    if (index_out) *index_out = int32_t(s->ligs.size() - 1);
    return UD_OK;
}

int ud_ligset_add_sdf(ud_ligset* s, const char* text, int32_t* first_index_out, int32_t* count_out) {
    if (!s || !text) return UD_ERR_INVALID;
    // records end at "$$$$" lines, which the reference does not keep
    std::vector<std::vector<std::string>> recs(1);
    std::istringstream in(text);
    std::string ln;
    while (std::getline(in, ln)) {
        if (!ln.empty() && ln.back() == '\r') ln.pop_back();
        if (ln.find("$$$$") != std::string::npos) {
            recs.emplace_back();
            continue;
        }
        recs.back().push_back(ln);
    }
    while (!recs.empty()) {  // drop a trailing empty record
        bool blank = true;
        for (const std::string& l : recs.back()) blank = blank && trim(l).empty();
        if (!blank) break;
        recs.pop_back();
    }
    if (recs.empty()) {
        s->err = "no SDF record";
        return UD_ERR_INVALID;
    }
    std::vector<Ligand> parsed(recs.size());
    for (size_t r = 0; r < recs.size(); ++r) {
        std::string err;
        if (!build_ligand_sdf(recs[r], parsed[r], err)) {
            s->err = "SDF record " + std::to_string(r + 1) + ": " + err;
            return UD_ERR_INVALID;
        }
        const Ligand& L = parsed[r];
        if (int(L.atoms.size()) > UD_MAX_ATOMS || int(L.node_parent.size()) - 1 > UD_MAX_TORSIONS ||
            int(L.pairs_ad4.size() / 2) > UD_MAX_PAIRS) {
            s->err = "SDF record " + std::to_string(r + 1) + ": ligand exceeds the library's capacities";
            return UD_ERR_CAPACITY;
        }
    }
    if (first_index_out) *first_index_out = int32_t(s->ligs.size());
    if (count_out) *count_out = int32_t(parsed.size());
    for (Ligand& L : parsed) s->ligs.push_back(std::move(L));
    return UD_OK;
}

int ud_ligset_count(const ud_ligset* s) { return s ? int(s->ligs.size()) : 0; }

int ud_ligset_batch(ud_ligset* s, ud_ligand_batch* b) { return ud_ligset_batch_sf(s, UD_SF_VINA, b); }

int ud_ligset_batch_sf(ud_ligset* s, int sf, ud_ligand_batch* b) {
    if (!s || !b || (sf != UD_SF_VINA && sf != UD_SF_VINARDO && sf != UD_SF_AD4)) return UD_ERR_INVALID;
    s->atom_offset.assign(1, 0);
    s->node_offset.assign(1, 0);
    s->pair_offset.assign(1, 0);
    s->atom_xs.clear(); s->atom_heavy.clear(); s->node_parent.clear(); s->node_abegin.clear();
    s->node_aend.clear(); s->pairs.clear(); s->atom_local.clear(); s->node_origin.clear();
    s->node_axis.clear(); s->num_tors.clear(); s->atom_ad.clear(); s->atom_charge.clear(); s->torsdof.clear();
    for (const Ligand& L : s->ligs) {
        for (const Atom& a : L.atoms) {
            const float* o = &L.node_origin[3 * a.node];
            s->atom_local.insert(s->atom_local.end(), {a.x - o[0], a.y - o[1], a.z - o[2]});
            s->atom_xs.push_back(a.xs);
            s->atom_heavy.push_back(a.el != EL_H ? 1 : 0);
            s->atom_ad.push_back(a.ad);
            s->atom_charge.push_back(a.charge);
        }
        for (size_t k = 0; k < L.node_parent.size(); ++k) {
            s->node_parent.push_back(L.node_parent[k]);
            s->node_abegin.push_back(L.node_abegin[k]);
            s->node_aend.push_back(L.node_aend[k]);
        }
        s->node_origin.insert(s->node_origin.end(), L.node_origin.begin(), L.node_origin.end());
        s->node_axis.insert(s->node_axis.end(), L.node_axis.begin(), L.node_axis.end());
        const std::vector<int>& pr = (sf == UD_SF_AD4) ? L.pairs_ad4 : L.pairs;
        s->pairs.insert(s->pairs.end(), pr.begin(), pr.end());
        s->num_tors.push_back(L.num_tors);
        s->torsdof.push_back(float(L.torsdof));
        s->atom_offset.push_back(s->atom_offset.back() + int32_t(L.atoms.size()));
        s->node_offset.push_back(s->node_offset.back() + int32_t(L.node_parent.size()));
        s->pair_offset.push_back(s->pair_offset.back() + int32_t(pr.size() / 2));
    }
    b->n_ligands = int32_t(s->ligs.size());
    b->atom_offset = s->atom_offset.data();
    b->node_offset = s->node_offset.data();
    b->pair_offset = s->pair_offset.data();
    b->atom_local = s->atom_local.data();
    b->atom_xs = s->atom_xs.data();
    b->atom_heavy = s->atom_heavy.data();
    b->node_parent = s->node_parent.data();
    b->node_abegin = s->node_abegin.data();
    b->node_aend = s->node_aend.data();
    b->node_origin = s->node_origin.data();
    b->node_axis = s->node_axis.data();
    b->pairs = s->pairs.data();
    b->num_tors = s->num_tors.data();
    b->atom_ad = s->atom_ad.data();
    b->atom_charge = s->atom_charge.data();
    b->torsdof = s->torsdof.data();
    return UD_OK;
}

int ud_ligset_atoms(const ud_ligset* s, int lig, int32_t* n_atoms_out, float* xyz_out, int32_t* serial_out) {
    if (!s || lig < 0 || lig >= int(s->ligs.size()) || !n_atoms_out) return UD_ERR_INVALID;
    const Ligand& L = s->ligs[lig];
    *n_atoms_out = int32_t(L.atoms.size());
    for (size_t i = 0; i < L.atoms.size(); ++i) {
        if (xyz_out) {
            xyz_out[3 * i] = L.atoms[i].x;
            xyz_out[3 * i + 1] = L.atoms[i].y;
            xyz_out[3 * i + 2] = L.atoms[i].z;
        }
        if (serial_out) serial_out[i] = L.atoms[i].serial;
    }
    return UD_OK;
}

int ud_ligset_write_pdbqt(const ud_ligset* s, int lig, int n_modes, const ud_mode* modes, const float* coords,
                          int how_many, double energy_range, char* out, size_t cap, size_t* len_out) {
    if (!s || lig < 0 || lig >= int(s->ligs.size()) || n_modes < 0 || (n_modes > 0 && (!modes || !coords)) ||
        how_many < 0 || !(energy_range >= 0.0) || !len_out)
        return UD_ERR_INVALID;
    const Ligand& L = s->ligs[lig];
    if (L.format != 0) return UD_ERR_STATE;  // the text written back is the PDBQT input's
    const size_t na = L.atoms.size();
    std::string txt;
    char buf[160];
    int written = 0;
    for (int k = 0; k < n_modes && written < how_many; ++k) {
        const ud_mode& m = modes[k];
        if (!(m.energy < 3.0e38f) || m.energy > modes[0].energy + energy_range) break;
        ++written;
        std::snprintf(buf, sizeof(buf), "MODEL %d\n", written);
        txt += buf;
        std::snprintf(buf, sizeof(buf), "REMARK VINA RESULT: %9.3f  %9.3f  %9.3f\n", m.energy, m.rmsd_lb, m.rmsd_ub);
        txt += buf;
        std::snprintf(buf, sizeof(buf), "REMARK INTER + INTRA:    %12.3f\n", m.inter + m.intra);
        txt += buf;
        std::snprintf(buf, sizeof(buf), "REMARK INTER:            %12.3f\n", m.inter);
        txt += buf;
        std::snprintf(buf, sizeof(buf), "REMARK INTRA:            %12.3f\n", m.intra);
        txt += buf;
        std::snprintf(buf, sizeof(buf), "REMARK UNBOUND:          %12.3f\n", m.unbound);
        txt += buf;
        const float* c = coords + size_t(k) * na * 3;
        for (size_t i = 0; i < L.lines.size(); ++i) {
            std::string ln = L.lines[i];
            const int a = L.line_atom[i];
            if (a >= 0) {  // string_write_coord at columns 31, 39, 47 (model.cpp:771-777)
                if (ln.size() < 54) ln.resize(54, ' ');
                for (int d = 0; d < 3; ++d) {
                    std::snprintf(buf, sizeof(buf), "%8.3f", c[3 * a + d]);
                    ln.replace(30 + 8 * d, 8, std::string(buf).substr(0, 8));
                }
            }
            txt += ln;
            txt += '\n';
        }
        txt += "ENDMDL\n";
    }
    *len_out = txt.size();
    if (!out) return UD_OK;
    if (cap < txt.size() + 1) return UD_ERR_CAPACITY;
    std::memcpy(out, txt.c_str(), txt.size() + 1);
    return UD_OK;
}

int ud_ligset_write_sdf(const ud_ligset* s, int lig, int n_modes, const ud_mode* modes, const float* coords,
                        int how_many, double energy_range, char* out, size_t cap, size_t* len_out) {
    if (!s || lig < 0 || lig >= int(s->ligs.size()) || n_modes < 0 || (n_modes > 0 && (!modes || !coords)) ||
        how_many < 0 || !(energy_range >= 0.0) || !len_out)
        return UD_ERR_INVALID;
    const Ligand& L = s->ligs[lig];
    if (L.format != 1) return UD_ERR_STATE;  // the text written back is the SDF input's
    const size_t na = L.atoms.size();
    std::string txt;
    char buf[200];
    int written = 0;
    for (int k = 0; k < n_modes && written < how_many; ++k) {
        const ud_mode& m = modes[k];
        if (!(m.energy < 3.0e38f) || m.energy > modes[0].energy + energy_range) break;
        ++written;
        const float* c = coords + size_t(k) * na * 3;
        for (size_t i = 0; i < L.lines.size(); ++i) {
            std::string ln = L.lines[i];
            const int a = L.line_atom[i];
            if (a >= 0) {  // string_write_sdf_coord at columns 1, 11, 21 (model.cpp:758-785)
                if (ln.size() < 30) ln.resize(30, ' ');
                for (int d = 0; d < 3; ++d) {
                    std::snprintf(buf, sizeof(buf), "%10.4f", c[3 * a + d]);
                    ln.replace(10 * d, 10, std::string(buf).substr(0, 10));
                }
            }
            txt += ln;
            txt += '\n';
        }
        // sdf_remarks (vina.cpp:751-770)
        std::snprintf(buf, sizeof(buf), "> <Uni-Dock RESULT> \nENERGY=%9.3f  LOWER_BOUND=%9.3f  UPPER_BOUND=%9.3f\n",
                      m.energy, m.rmsd_lb, m.rmsd_ub);
        txt += buf;
        std::snprintf(buf, sizeof(buf), "INTER + INTRA:    %12.3f\n", m.inter + m.intra);
        txt += buf;
        std::snprintf(buf, sizeof(buf), "INTER:            %12.3f\n", m.inter);
        txt += buf;
        std::snprintf(buf, sizeof(buf), "INTRA:            %12.3f\n", m.intra);
        txt += buf;
        std::snprintf(buf, sizeof(buf), "UNBOUND:          %12.3f\n\n", m.unbound);
        txt += buf;
        txt += "$$$$\n";
    }
    *len_out = txt.size();
    if (!out) return UD_OK;
    if (cap < txt.size() + 1) return UD_ERR_CAPACITY;
    std::memcpy(out, txt.c_str(), txt.size() + 1);
    return UD_OK;
}

int ud_receptor_ad4_from_pdbqt(const char* text, int32_t capacity, int32_t* n_out, float* xyz_out, int32_t* ad_out,
                               float* charge_out) {
    if (!text || !n_out) return UD_ERR_INVALID;
    std::istringstream in(text);
    std::string ln, err;
    int n = 0;
    const bool write = xyz_out && ad_out && charge_out;
    while (std::getline(in, ln)) {
        if (!ln.empty() && ln.back() == '\r') ln.pop_back();
        if (ln.rfind("ATOM", 0) != 0 && ln.rfind("HETATM", 0) != 0) continue;
        Atom a;
        if (!parse_atom_line(ln, a, err)) return UD_ERR_INVALID;
        if (write && n < capacity) {
            xyz_out[3 * n] = a.x;
            xyz_out[3 * n + 1] = a.y;
            xyz_out[3 * n + 2] = a.z;
            ad_out[n] = a.ad;
            charge_out[n] = a.charge;
        }
        ++n;
    }
    *n_out = n;
    if ((xyz_out || ad_out || charge_out) && n > capacity) return UD_ERR_CAPACITY;
    return UD_OK;
}

int ud_receptor_from_pdbqt(const char* text, int32_t capacity, int32_t* n_out, float* xyz_out, int32_t* xs_out) {
    if (!text || !n_out) return UD_ERR_INVALID;
    std::vector<Atom> at;
    std::istringstream in(text);
    std::string ln, err;
    while (std::getline(in, ln)) {
        if (!ln.empty() && ln.back() == '\r') ln.pop_back();
        if (ln.rfind("ATOM", 0) == 0 || ln.rfind("HETATM", 0) == 0) {
            Atom a;
            if (!parse_atom_line(ln, a, err)) return UD_ERR_INVALID;
            at.push_back(a);
        }
    }
    std::vector<std::vector<int>> bonds;
    assign_bonds(at, [](int, int) { return true; }, bonds);  // a rigid receptor: all atoms fixed
    assign_types(at, bonds);
    int n = 0;
    for (const Atom& a : at) {
        if (a.xs < 0) continue;  // hydrogens and untyped atoms have no grid term
        if (xyz_out && xs_out && n < capacity) {
            xyz_out[3 * n] = a.x;
            xyz_out[3 * n + 1] = a.y;
            xyz_out[3 * n + 2] = a.z;
            xs_out[n] = a.xs;
        }
        ++n;
    }
    *n_out = n;
    if ((xyz_out || xs_out) && n > capacity) return UD_ERR_CAPACITY;
    return UD_OK;
}

}  // extern "C"
