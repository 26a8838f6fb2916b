// AM1-BCC charges on the host (am1-rs src/bcc/mod.rs:99-150): the AM1 Mulliken charges from the
// GPU SCF, then molecular-graph perception from the geometry (src/topology.rs), antechamber BCC
// atom types (src/bcc/atomtype.rs, which interprets ATOMTYPE_BCC.DEF) and the bond charge
// corrections of BCCPARM.DAT. O(N^2) in the atoms, microseconds per molecule next to the SCF.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <set>
#include <sstream>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/am1g.h"

namespace {

struct BccParam {
    int ti, tj, bt;
    double delta;
};

#include "../data/bcc_tables.inc"

constexpr double kBohrToAngstrom = 0.529167;  // constants.rs:10

// Single-bond covalent radii in Angstrom (Cordero et al., Dalton Trans. 2008, 2832), the table of
// constants.rs:37-47; 1.5 beyond it.
double covalent_radius(int z) {
    static const double r[87] = {
        0.0,  0.31, 0.28, 1.28, 0.96, 0.84, 0.76, 0.71, 0.66, 0.57, 0.58, 1.66, 1.41, 1.21, 1.11, 1.07, 1.05, 1.02,
        1.06, 2.03, 1.76, 1.70, 1.60, 1.53, 1.39, 1.39, 1.32, 1.26, 1.24, 1.32, 1.22, 1.22, 1.20, 1.19, 1.20, 1.20,
        1.16, 2.20, 1.95, 1.90, 1.75, 1.64, 1.54, 1.47, 1.46, 1.42, 1.39, 1.45, 1.44, 1.42, 1.39, 1.39, 1.38, 1.39,
        1.40, 2.44, 2.15, 2.07, 2.04, 2.03, 2.01, 1.99, 1.98, 1.98, 1.96, 1.94, 1.92, 1.92, 1.89, 1.90, 1.87, 1.87,
        1.75, 1.70, 1.62, 1.51, 1.44, 1.41, 1.36, 1.36, 1.32, 1.45, 1.46, 1.48, 1.40, 1.50, 1.50};
    return (z >= 0 && z < 87) ? r[z] : 1.5;
}

const char* symbol(int z) {
    static const char* s[] = {"X",  "H",  "He", "Li", "Be", "B",  "C",  "N",  "O",  "F",  "Ne", "Na", "Mg", "Al", "Si",
                              "P",  "S",  "Cl", "Ar", "K",  "Ca", "Sc", "Ti", "V",  "Cr", "Mn", "Fe", "Co", "Ni", "Cu",
                              "Zn", "Ga", "Ge", "As", "Se", "Br", "Kr", "Rb", "Sr", "Y",  "Zr", "Nb", "Mo", "Tc", "Ru",
                              "Rh", "Pd", "Ag", "Cd", "In", "Sn", "Sb", "Te", "I",  "Xe", "Cs", "Ba", "La", "Ce", "Pr",
                              "Nd", "Pm", "Sm", "Eu", "Gd", "Tb", "Dy", "Ho", "Er", "Tm", "Yb", "Lu", "Hf", "Ta", "W",
                              "Re", "Os", "Ir", "Pt", "Au", "Hg", "Tl", "Pb", "Bi", "Po", "At", "Rn"};
    return (z > 0 && z < 87) ? s[z] : "?";
}

// ------------------------------------------------------------------------------------ topology
enum class Order { kSingle, kDouble, kTriple, kAromatic, kDelocalized };
enum class Hyb { kSp3, kSp2, kSp, kNone };

struct Vec3 {
    double x, y, z;
};
Vec3 sub(Vec3 a, Vec3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
double dot(Vec3 a, Vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
Vec3 cross(Vec3 a, Vec3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
double norm(Vec3 a) { return std::sqrt(dot(a, a)); }

struct Bond {
    int i, j;
    Order order;
    double length;  // Angstrom
};

struct Topo {
    int nat = 0;
    std::vector<int> z;
    std::vector<Vec3> pos;  // Angstrom
    std::vector<Bond> bonds;
    std::vector<std::vector<int>> nbr;
    std::vector<Hyb> hyb;
    std::vector<char> aromatic, in_ring;
    std::vector<std::vector<int>> rings;  // sorted by size
    std::vector<char> ring_aromatic;
    std::vector<char> kekule_double;
    bool kekule_unique = true;
    std::vector<std::string> warnings;

    int bond_between(int a, int b) const {
        for (size_t k = 0; k < bonds.size(); ++k)
            if ((bonds[k].i == a && bonds[k].j == b) || (bonds[k].i == b && bonds[k].j == a)) return (int)k;
        return -1;
    }
    // (single-like, double-like, triple) of bond k: aromatic bonds by their Kekule assignment,
    // delocalized bonds count as single (topology.rs bond_kinds)
    void kinds(int k, bool& sb, bool& db, bool& tb) const {
        sb = db = tb = false;
        switch (bonds[k].order) {
            case Order::kSingle:
            case Order::kDelocalized: sb = true; break;
            case Order::kDouble: db = true; break;
            case Order::kTriple: tb = true; break;
            case Order::kAromatic: (kekule_double[k] ? db : sb) = true; break;
        }
    }
    bool in_ring_of_size(int a, int n) const {
        for (const auto& r : rings)
            if ((int)r.size() == n && std::find(r.begin(), r.end(), a) != r.end()) return true;
        return false;
    }
};

// Breadth-first shortest path from `from` to `to` that does not use the edge (from, to) itself,
// as the list [to, ..., from]; empty when there is none.
std::vector<int> path_avoiding_edge(const std::vector<std::vector<int>>& adj, int from, int to) {
    const int n = (int)adj.size();
    std::vector<int> prev(n, -1);
    std::vector<char> seen(n, 0);
    std::deque<int> q;
    seen[from] = 1;
    q.push_back(from);
    while (!q.empty()) {
        const int x = q.front();
        q.pop_front();
        for (int y : adj[x]) {
            if (x == from && y == to) continue;
            if (y == to) {
                std::vector<int> path{to, x};
                int c = x;
                while (c != from) {
                    c = prev[c];
                    path.push_back(c);
                }
                return path;
            }
            if (!seen[y]) {
                seen[y] = 1;
                prev[y] = x;
                q.push_back(y);
            }
        }
    }
    return {};
}

// Typical single / double / triple bond lengths (Angstrom) of the pairs whose order is guessed
// from the distance (topology.rs reference_lengths); other pairs are single.
bool reference_lengths(int zi, int zj, double& s, double& d, double& t) {
    const int a = std::min(zi, zj), b = std::max(zi, zj);
    struct L {
        int a, b;
        double s, d, t;
    };
    static const L tab[] = {{6, 6, 1.54, 1.34, 1.20},   {6, 7, 1.47, 1.28, 1.16},   {6, 8, 1.43, 1.21, 1.13},
                            {7, 7, 1.45, 1.25, 1.10},   {7, 8, 1.40, 1.21, 1.06},   {8, 8, 1.48, 1.21, 1.21},
                            {6, 15, 1.84, 1.66, 1.54},  {6, 16, 1.82, 1.60, 1.53},  {7, 15, 1.77, 1.58, 1.48},
                            {7, 16, 1.71, 1.54, 1.44},  {8, 15, 1.63, 1.48, 1.48},  {8, 16, 1.57, 1.44, 1.44},
                            {15, 15, 2.21, 2.02, 1.89}, {15, 16, 2.10, 1.92, 1.92}, {16, 16, 2.05, 1.88, 1.88}};
    for (const L& l : tab)
        if (l.a == a && l.b == b) {
            s = l.s;
            d = l.d;
            t = l.t;
            return true;
        }
    return false;
}

Order guess_order(int zi, int zj, double dist) {
    double s, d, t;
    if (!reference_lengths(zi, zj, s, d, t)) return Order::kSingle;
    const double ds = std::fabs(dist - s), dd = std::fabs(dist - d), dt = std::fabs(dist - t);
    if (t < d && dt < dd && dt < ds) return Order::kTriple;
    if (dd < ds) return Order::kDouble;
    return Order::kSingle;
}

// Three or more neighbours whose three widest angles sum to more than 350 degrees.
bool planar_centre(const Topo& T, int i) {
    const auto& nb = T.nbr[i];
    if (nb.size() < 3) return false;
    std::vector<double> ang;
    for (size_t a = 0; a < nb.size(); ++a)
        for (size_t b = a + 1; b < nb.size(); ++b) {
            Vec3 va = sub(T.pos[nb[a]], T.pos[i]), vb = sub(T.pos[nb[b]], T.pos[i]);
            const double c = dot(va, vb) / (norm(va) * norm(vb));
            ang.push_back(std::acos(std::max(-1.0, std::min(1.0, c))));
        }
    std::sort(ang.begin(), ang.end(), std::greater<double>());
    double sum = 0.0;
    for (size_t k = 0; k < ang.size() && k < 3; ++k) sum += ang[k];
    return sum > 350.0 * M_PI / 180.0;
}

Hyb hybridization(const Topo& T, int i) {
    const int z = T.z[i], cn = (int)T.nbr[i].size();
    switch (z) {
        case 1: return Hyb::kNone;
        case 6: return cn == 3 ? Hyb::kSp2 : cn == 2 ? Hyb::kSp : Hyb::kSp3;
        case 7:
            if (cn == 3 || cn == 4) return planar_centre(T, i) ? Hyb::kSp2 : Hyb::kSp3;
            return cn == 2 ? Hyb::kSp2 : cn == 1 ? Hyb::kSp : Hyb::kSp3;
        case 8: return cn == 1 ? Hyb::kSp2 : Hyb::kSp3;
        case 16:
            if (cn == 1) return Hyb::kSp2;
            if (cn == 2) return planar_centre(T, i) ? Hyb::kSp2 : Hyb::kSp3;
            return Hyb::kSp3;
        default: return Hyb::kSp3;
    }
}

bool contains(const std::vector<int>& v, int x) { return std::find(v.begin(), v.end(), x) != v.end(); }

// pi electrons atom `a` gives ring `ring` (Hueckel count), or -1 when it cannot take part.
int pi_electrons(const Topo& T, int a, const std::vector<int>& ring) {
    int in_ring = 0;
    for (int n : T.nbr[a]) in_ring += contains(ring, n);
    if (in_ring != 2) return -1;
    const int cn = (int)T.nbr[a].size();
    switch (T.z[a]) {
        case 6: {
            if (T.hyb[a] != Hyb::kSp2) return -1;
            for (int n : T.nbr[a])  // an exocyclic carbonyl (or thione) takes the carbon's pi electron
                if (!contains(ring, n) && (T.z[n] == 8 || T.z[n] == 16) && T.nbr[n].size() == 1) return 0;
            return 1;
        }
        case 7: return cn == 2 ? 1 : cn == 3 ? 2 : -1;
        case 8:
        case 16: return (cn == 2 && ring.size() == 5) ? 2 : -1;
        default: return -1;
    }
}

bool ring_planar(const Topo& T, const std::vector<int>& ring) {
    Vec3 c{0, 0, 0};
    for (int a : ring) c = {c.x + T.pos[a].x, c.y + T.pos[a].y, c.z + T.pos[a].z};
    const double inv = 1.0 / ring.size();
    c = {c.x * inv, c.y * inv, c.z * inv};
    Vec3 nrm{0, 0, 0};
    for (size_t k = 0; k < ring.size(); ++k) {
        const Vec3 x = cross(sub(T.pos[ring[k]], c), sub(T.pos[ring[(k + 1) % ring.size()]], c));
        nrm = {nrm.x + x.x, nrm.y + x.y, nrm.z + x.z};
    }
    const double nn = norm(nrm);
    if (nn < 1.0e-8) return false;
    nrm = {nrm.x / nn, nrm.y / nn, nrm.z / nn};
    for (int a : ring)
        if (std::fabs(dot(sub(T.pos[a], c), nrm)) >= 0.15) return false;
    return true;
}

bool ring_is_aromatic(const Topo& T, const std::vector<int>& ring) {
    if (ring.size() != 5 && ring.size() != 6) return false;
    int total = 0;
    for (int a : ring) {
        const int p = pi_electrons(T, a, ring);
        if (p < 0) return false;
        total += p;
    }
    if (total < 2 || (total - 2) % 4 != 0) return false;  // 4n + 2
    return ring_planar(T, ring);
}

// Backtracking perfect matching of the atoms in `pending` over aromatic bonds (no bond in
// `forbidden`); chosen[k] marks the double bonds.
bool assign_kekule(const Topo& T, const std::vector<int>& pending, size_t at, const std::vector<std::vector<int>>& inc,
                   std::vector<char>& matched, std::vector<char>& chosen, int forbidden) {
    if (at == pending.size()) return true;
    const int atom = pending[at];
    if (matched[atom]) return assign_kekule(T, pending, at + 1, inc, matched, chosen, forbidden);
    for (int k : inc[atom]) {
        if (k == forbidden) continue;
        const int other = T.bonds[k].i == atom ? T.bonds[k].j : T.bonds[k].i;
        if (matched[other] || !contains(pending, other)) continue;
        matched[atom] = matched[other] = 1;
        chosen[k] = 1;
        if (assign_kekule(T, pending, at + 1, inc, matched, chosen, forbidden)) return true;
        matched[atom] = matched[other] = 0;
        chosen[k] = 0;
    }
    return false;
}

// Topology::perceive (topology.rs): bonds by covalent radii, smallest rings, hybridization, bond
// orders from lengths, aromatic 5/6 rings (5-rings fused to an aromatic 6-ring demoted, as
// antechamber does for AM1-BCC), delocalized X-O/S bonds, Kekule double bonds.
Topo perceive(int nat, const uint8_t* z, const double* xyz_bohr) {
    Topo T;
    T.nat = nat;
    T.z.assign(z, z + nat);
    T.pos.resize(nat);
    for (int i = 0; i < nat; ++i)
        T.pos[i] = {xyz_bohr[3 * i] * kBohrToAngstrom, xyz_bohr[3 * i + 1] * kBohrToAngstrom,
                    xyz_bohr[3 * i + 2] * kBohrToAngstrom};
    T.nbr.assign(nat, {});
    for (int i = 0; i < nat; ++i)
        for (int j = i + 1; j < nat; ++j) {
            const double d = norm(sub(T.pos[j], T.pos[i]));
            if (d < 1.3 * (covalent_radius(T.z[i]) + covalent_radius(T.z[j])) && d > 0.4) {
                T.nbr[i].push_back(j);
                T.nbr[j].push_back(i);
                T.bonds.push_back({i, j, Order::kSingle, d});
            }
        }
    // smallest ring through each bond (3 to 12 atoms), each ring once
    {
        std::set<std::vector<int>> seen;
        std::vector<std::vector<int>> found;
        for (const Bond& b : T.bonds) {
            std::vector<int> path = path_avoiding_edge(T.nbr, b.i, b.j);
            if (path.size() < 3 || path.size() > 12) continue;
            std::vector<int> key = path;
            std::sort(key.begin(), key.end());
            if (seen.insert(key).second) found.push_back(path);
        }
        std::stable_sort(found.begin(), found.end(),
                         [](const std::vector<int>& a, const std::vector<int>& b) { return a.size() < b.size(); });
        T.rings = found;
    }
    T.in_ring.assign(nat, 0);
    for (const auto& r : T.rings)
        for (int a : r) T.in_ring[a] = 1;
    T.hyb.resize(nat);
    for (int i = 0; i < nat; ++i) T.hyb[i] = hybridization(T, i);
    for (Bond& b : T.bonds) b.order = guess_order(T.z[b.i], T.z[b.j], b.length);
    T.ring_aromatic.assign(T.rings.size(), 0);
    for (size_t k = 0; k < T.rings.size(); ++k) T.ring_aromatic[k] = ring_is_aromatic(T, T.rings[k]);
    {
        std::vector<size_t> demote;
        for (size_t k = 0; k < T.rings.size(); ++k) {
            if (!T.ring_aromatic[k] || T.rings[k].size() != 5) continue;
            for (size_t m = 0; m < T.rings.size(); ++m) {
                if (m == k || !T.ring_aromatic[m] || T.rings[m].size() != 6) continue;
                int shared = 0;
                for (int a : T.rings[k]) shared += contains(T.rings[m], a);
                if (shared >= 2) {
                    demote.push_back(k);
                    break;
                }
            }
        }
        for (size_t k : demote) T.ring_aromatic[k] = 0;
    }
    T.aromatic.assign(nat, 0);
    for (size_t k = 0; k < T.rings.size(); ++k)
        if (T.ring_aromatic[k])
            for (int a : T.rings[k]) T.aromatic[a] = 1;
    for (Bond& b : T.bonds) {
        if (!T.aromatic[b.i] || !T.aromatic[b.j]) continue;
        for (size_t k = 0; k < T.rings.size(); ++k)
            if (T.ring_aromatic[k] && contains(T.rings[k], b.i) && contains(T.rings[k], b.j)) {
                b.order = Order::kAromatic;
                break;
            }
    }
    // delocalized centres: C, N, P or S with two or more terminal O/S, or an N-oxide
    {
        std::vector<char> centre(nat, 0);
        for (int c = 0; c < nat; ++c) {
            const int zc = T.z[c];
            if (zc != 6 && zc != 7 && zc != 15 && zc != 16) continue;
            int term = 0;
            for (int n : T.nbr[c]) term += (T.z[n] == 8 || T.z[n] == 16) && T.nbr[n].size() == 1;
            const bool n_oxide = zc == 7 && T.nbr[c].size() == 4 && term == 1;
            if (term >= 2 || n_oxide) centre[c] = 1;
        }
        for (Bond& b : T.bonds) {
            int c = b.i, t = b.j;
            if (!centre[b.i]) std::swap(c, t);
            if (!centre[c]) continue;
            if ((T.z[t] == 8 || T.z[t] == 16) && T.nbr[t].size() == 1) b.order = Order::kDelocalized;
        }
    }
    // Kekule structure of the aromatic systems: a perfect matching of the atoms that give one pi
    // electron, over the aromatic bonds; unique when no chosen bond can be avoided
    T.kekule_double.assign(T.bonds.size(), 0);
    for (size_t k = 0; k < T.bonds.size(); ++k) T.kekule_double[k] = T.bonds[k].order == Order::kDouble;
    {
        std::vector<char> needs(nat, 0);
        bool any = false;
        for (size_t r = 0; r < T.rings.size(); ++r) {
            if (!T.ring_aromatic[r]) continue;
            for (int a : T.rings[r]) {
                any = true;
                if (pi_electrons(T, a, T.rings[r]) == 1) needs[a] = 1;
            }
        }
        if (any) {
            std::vector<std::vector<int>> inc(nat);
            for (size_t k = 0; k < T.bonds.size(); ++k)
                if (T.bonds[k].order == Order::kAromatic) {
                    inc[T.bonds[k].i].push_back((int)k);
                    inc[T.bonds[k].j].push_back((int)k);
                }
            std::vector<int> pending;
            for (int a = 0; a < nat; ++a)
                if (needs[a]) pending.push_back(a);
            std::vector<char> matched(nat, 0), chosen(T.bonds.size(), 0);
            if (!assign_kekule(T, pending, 0, inc, matched, chosen, -1)) {
                char buf[256];
                std::snprintf(buf, sizeof(buf),
                              "no Kekule structure was found for the aromatic system spanning %zu atoms; its bonds are "
                              "treated as formally single, which can change the atom type of a two-coordinate ring "
                              "nitrogen",
                              pending.size());
                T.warnings.push_back(buf);
                T.kekule_unique = false;
            } else {
                for (size_t k = 0; k < chosen.size(); ++k)
                    if (chosen[k]) T.kekule_double[k] = 1;
                for (size_t k = 0; k < chosen.size() && T.kekule_unique; ++k) {
                    if (!chosen[k]) continue;
                    std::vector<char> m2(nat, 0), c2(T.bonds.size(), 0);
                    if (assign_kekule(T, pending, 0, inc, m2, c2, (int)k)) T.kekule_unique = false;
                }
            }
        }
    }
    for (int i = 0; i < nat; ++i) {
        const int zi = T.z[i];
        if (!(zi == 1 || zi == 6 || zi == 7 || zi == 8 || zi == 9 || zi == 14 || zi == 15 || zi == 16 || zi == 17 ||
              zi == 35 || zi == 53)) {
            char buf[256];
            std::snprintf(buf, sizeof(buf),
                          "atom %d (%s) has no AM1-BCC atom type; its charge will be the raw AM1 Mulliken value and "
                          "every bond to it is left uncorrected",
                          i, symbol(zi));
            T.warnings.push_back(buf);
        }
    }
    return T;
}

// ------------------------------------------------------------------------------------ atom types
// The rule language of ATOMTYPE_BCC.DEF as atomtype.rs reads it: element specs (symbol plus
// optional connection count, or a wild-atom set), atomic properties in [...] (',' = and, '.' =
// or; AR1..AR5 aromatic, RG / RGn ring, NR non-ring, sb/db/tb bond kinds with counts, a trailing
// ' for a bond to the preceding atom in the chain and '' for its absence, SB/DB/TB/AB/DL strict
// orders) and a chemical environment of nested (...) lists matched to distinct neighbours.
enum class BK { kSb, kDb, kTb, kSS, kSD, kST, kSA, kSDl };
struct Prop {
    int kind;  // 0 aromatic, 1 in ring (size or 0), 2 non-ring, 3 bond
    int ring;
    BK bk;
    int count;
    int pred;  // -1 any, 1 bond to the predecessor, 0 no such bond
};
using PropExpr = std::vector<std::vector<Prop>>;  // and of ors
struct Alt {
    int z;     // 0: any
    int conn;  // -1: any
};
struct Spec {
    std::vector<Alt> alts;  // empty: any atom
    PropExpr props;
};
struct EnvItem {
    Spec spec;
    std::vector<EnvItem> kids;
};
struct Rule {
    int type;
    int z, conn;  // 0 / -1 when not constrained
    PropExpr props;
    std::vector<EnvItem> env;
};

int element_z(const std::string& sym) {
    for (int z = 1; z < 87; ++z)
        if (sym == symbol(z)) return z;
    return -1;
}

bool parse_prop(const std::string& text, Prop& p) {
    size_t i = 0;
    while (i < text.size() && std::isdigit((unsigned char)text[i])) ++i;
    const int count = i ? std::stoi(text.substr(0, i)) : 1;
    std::string rest = text.substr(i);
    int primes = 0;
    while (!rest.empty() && rest.back() == '\'') {
        rest.pop_back();
        ++primes;
    }
    p = Prop{3, 0, BK::kSb, count, primes == 0 ? -1 : primes == 1 ? 1 : 0};
    if (rest.size() == 3 && rest.compare(0, 2, "AR") == 0) {
        p.kind = 0;
        return true;
    }
    if (rest == "NR") {
        p.kind = 2;
        return true;
    }
    if (rest.compare(0, 2, "RG") == 0) {
        p.kind = 1;
        p.ring = rest.size() > 2 ? std::atoi(rest.c_str() + 2) : 0;
        return true;
    }
    static const std::map<std::string, BK> bonds = {{"sb", BK::kSb}, {"db", BK::kDb}, {"tb", BK::kTb},
                                                    {"SB", BK::kSS}, {"DB", BK::kSD}, {"TB", BK::kST},
                                                    {"AB", BK::kSA}, {"DL", BK::kSDl}};
    auto it = bonds.find(rest);
    if (it == bonds.end()) return false;
    p.bk = it->second;
    return true;
}

PropExpr parse_props(std::string text) {
    PropExpr out;
    if (!text.empty() && text.front() == '[') text.erase(0, 1);
    if (!text.empty() && text.back() == ']') text.pop_back();
    std::stringstream groups(text);
    std::string g;
    while (std::getline(groups, g, ',')) {
        std::vector<Prop> ors;
        std::stringstream alts(g);
        std::string a;
        while (std::getline(alts, a, '.')) {
            Prop p;
            if (!a.empty() && parse_prop(a, p)) ors.push_back(p);
        }
        if (!ors.empty()) out.push_back(ors);
    }
    return out;
}

using WildMap = std::map<std::string, std::vector<Alt>>;

bool parse_alt(const std::string& text, const WildMap& wild, std::vector<Alt>& out) {
    out.clear();
    if (text.empty() || text == "*") return true;
    size_t i = 0;
    while (i < text.size() && std::isalpha((unsigned char)text[i])) ++i;
    const std::string name = text.substr(0, i);
    const int conn = i < text.size() ? std::atoi(text.c_str() + i) : -1;
    auto w = wild.find(name);
    if (w != wild.end()) {
        for (Alt a : w->second) out.push_back({a.z, a.conn >= 0 ? a.conn : conn});
        return true;
    }
    const int z = element_z(name);
    if (z < 0) return false;
    out.push_back({z, conn});
    return true;
}

std::vector<std::string> split_top(const std::string& text, char sep) {
    std::vector<std::string> out;
    int depth = 0;
    size_t start = 0;
    for (size_t i = 0; i < text.size(); ++i) {
        if (text[i] == '(') ++depth;
        else if (text[i] == ')') depth = std::max(0, depth - 1);
        else if (text[i] == sep && depth == 0) {
            out.push_back(text.substr(start, i - start));
            start = i + 1;
        }
    }
    if (start < text.size()) out.push_back(text.substr(start));
    return out;
}

std::string strip_parens(const std::string& t) {
    if (t.size() >= 2 && t.front() == '(' && t.back() == ')') {
        int depth = 0;
        for (size_t i = 0; i < t.size(); ++i) {
            if (t[i] == '(') ++depth;
            else if (t[i] == ')' && --depth == 0) return i == t.size() - 1 ? t.substr(1, t.size() - 2) : t;
        }
    }
    return t;
}

std::vector<EnvItem> parse_env_list(const std::string& text, const WildMap& wild);

bool parse_env_item(const std::string& text, const WildMap& wild, EnvItem& item) {
    int depth = 0;
    size_t split = text.size();
    for (size_t i = 0; i < text.size(); ++i) {
        if (text[i] == '(') {
            if (depth == 0) split = i;
            ++depth;
        } else if (text[i] == ')') {
            depth = std::max(0, depth - 1);
        }
    }
    std::string head = text.substr(0, split), tail = text.substr(split);
    const size_t tag = head.find('<');  // tags (<a1>) are not used by the matching
    if (tag != std::string::npos) head = head.substr(0, tag);
    const size_t br = head.find('[');
    std::string el = br == std::string::npos ? head : head.substr(0, br);
    item.spec.props = br == std::string::npos ? PropExpr{} : parse_props(head.substr(br));
    if (!parse_alt(el, wild, item.spec.alts)) return false;
    if (!tail.empty()) item.kids = parse_env_list(strip_parens(tail), wild);
    return true;
}

std::vector<EnvItem> parse_env_list(const std::string& text, const WildMap& wild) {
    std::vector<EnvItem> out;
    for (const std::string& piece : split_top(text, ',')) {
        EnvItem it;
        if (parse_env_item(piece, wild, it)) out.push_back(it);
    }
    return out;
}

const std::vector<Rule>& rules() {
    static std::vector<Rule> R;
    static std::once_flag once;
    std::call_once(once, [] {
        WildMap wild;
        for (const char* line : kBccTypeRules) {
            std::stringstream ss(line);
            std::vector<std::string> tok;
            std::string t;
            while (ss >> t) tok.push_back(t);
            if (tok.empty()) continue;
            if (tok[0] == "WILDATOM" && tok.size() >= 3) {
                std::vector<Alt> all, one;
                for (size_t k = 2; k < tok.size(); ++k)
                    if (parse_alt(tok[k], wild, one)) all.insert(all.end(), one.begin(), one.end());
                wild[tok[1]] = all;
            } else if (tok[0] == "ATD") {
                if (tok.back() == "&") tok.pop_back();
                auto field = [&](size_t i) -> std::string {
                    return (i < tok.size() && tok[i] != "*" && tok[i] != "&") ? tok[i] : std::string();
                };
                Rule r;
                r.type = tok[1] == "DU" ? 0 : std::atoi(tok[1].c_str());
                r.z = field(3).empty() ? 0 : std::atoi(field(3).c_str());
                r.conn = field(4).empty() ? -1 : std::atoi(field(4).c_str());
                if (!field(7).empty()) r.props = parse_props(field(7));
                if (!field(8).empty()) r.env = parse_env_list(strip_parens(field(8)), wild);
                R.push_back(r);
            }
        }
    });
    return R;
}

bool bond_is(const Topo& T, int k, BK kind) {
    bool sb, db, tb;
    T.kinds(k, sb, db, tb);
    switch (kind) {
        case BK::kSb: return sb;
        case BK::kDb: return db;
        case BK::kTb: return tb;
        case BK::kSS: return T.bonds[k].order == Order::kSingle;
        case BK::kSD: return T.bonds[k].order == Order::kDouble;
        case BK::kST: return T.bonds[k].order == Order::kTriple;
        case BK::kSA: return T.bonds[k].order == Order::kAromatic;
        case BK::kSDl: return T.bonds[k].order == Order::kDelocalized;
    }
    return false;
}

bool prop_holds(const Prop& p, const Topo& T, int atom, int pred) {
    switch (p.kind) {
        case 0: return T.aromatic[atom];
        case 1: return p.ring ? T.in_ring_of_size(atom, p.ring) : T.in_ring[atom];
        case 2: return !T.in_ring[atom];
        default: break;
    }
    if (p.pred < 0) {
        int have = 0;
        for (size_t k = 0; k < T.bonds.size(); ++k)
            if ((T.bonds[k].i == atom || T.bonds[k].j == atom) && bond_is(T, (int)k, p.bk)) ++have;
        return have >= p.count;
    }
    if (pred < 0) return p.pred == 0;
    const int k = T.bond_between(atom, pred);
    const bool held = k >= 0 && bond_is(T, k, p.bk);
    return held == (p.pred == 1);
}

bool props_hold(const PropExpr& e, const Topo& T, int atom, int pred) {
    for (const auto& group : e) {
        bool any = false;
        for (const Prop& p : group) any = any || prop_holds(p, T, atom, pred);
        if (!any) return false;
    }
    return true;
}

bool spec_matches(const Spec& s, const Topo& T, int atom, int pred) {
    const int z = T.z[atom], cn = (int)T.nbr[atom].size();
    bool ok = s.alts.empty();
    for (const Alt& a : s.alts) ok = ok || ((a.z == 0 || a.z == z) && (a.conn < 0 || a.conn == cn));
    return ok && props_hold(s.props, T, atom, pred);
}

bool match_env(const std::vector<EnvItem>& items, const Topo& T, int atom, int pred);

// each item to a distinct neighbour of `parent` (not its predecessor), with backtracking
bool assign_items(const std::vector<EnvItem>& items, size_t at, const std::vector<int>& cand, std::vector<char>& used,
                  const Topo& T, int parent) {
    if (at == items.size()) return true;
    for (size_t s = 0; s < cand.size(); ++s) {
        if (used[s]) continue;
        if (!spec_matches(items[at].spec, T, cand[s], parent)) continue;
        if (!match_env(items[at].kids, T, cand[s], parent)) continue;
        used[s] = 1;
        if (assign_items(items, at + 1, cand, used, T, parent)) return true;
        used[s] = 0;
    }
    return false;
}

bool match_env(const std::vector<EnvItem>& items, const Topo& T, int atom, int pred) {
    if (items.empty()) return true;
    std::vector<int> cand;
    for (int n : T.nbr[atom])
        if (n != pred) cand.push_back(n);
    std::vector<char> used(cand.size(), 0);
    return assign_items(items, 0, cand, used, T, atom);
}

int bcc_type(const Topo& T, int atom) {
    for (const Rule& r : rules()) {
        if (r.z && T.z[atom] != r.z) continue;
        if (r.conn >= 0 && (int)T.nbr[atom].size() != r.conn) continue;
        if (!props_hold(r.props, T, atom, -1)) continue;
        if (match_env(r.env, T, atom, -1)) return r.type;
    }
    return 0;
}

// ------------------------------------------------------------------------------------ corrections
const std::map<std::tuple<int, int, int>, double>& bcc_table() {
    static std::map<std::tuple<int, int, int>, double> m;
    static std::once_flag once;
    std::call_once(once, [] {
        for (const BccParam& p : kBccParams) m[std::make_tuple(p.ti, p.tj, p.bt)] = p.delta;
    });
    return m;
}

// Antechamber bond-type codes of bond k, most specific first (bcc/mod.rs antechamber_bond_codes):
// 1/2/3 single/double/triple, aromatic 7 (single) / 8 (double) by the Kekule structure or 10
// when it is not unique (falling back to 7), delocalized 6 (nitrogen centre) or 9, and 11 last
// when both ends have the same type.
std::vector<int> bond_codes(const Topo& T, int k, int ta, int tb) {
    const Bond& b = T.bonds[k];
    int primary = 1;
    switch (b.order) {
        case Order::kSingle: primary = 1; break;
        case Order::kDouble: primary = 2; break;
        case Order::kTriple: primary = 3; break;
        case Order::kAromatic: primary = !T.kekule_unique ? 10 : T.kekule_double[k] ? 8 : 7; break;
        case Order::kDelocalized: {
            const bool terminal_j = (T.z[b.j] == 8 || T.z[b.j] == 16) && T.nbr[b.j].size() == 1;
            primary = T.z[terminal_j ? b.i : b.j] == 7 ? 6 : 9;
            break;
        }
    }
    std::vector<int> chain{primary};
    if (primary == 8 || primary == 10) chain.push_back(7);
    if (ta == tb && ta != 0) chain.push_back(11);
    return chain;
}

}  // namespace

extern "C" int am1g_bcc_charges(int n_atoms, const uint8_t* z, const double* xyz_bohr, const double* mulliken,
                                double* charges, int* atom_types, int* n_warnings, char* warnings, int warnings_len) {
    if (n_atoms < 1 || !z || !xyz_bohr || !mulliken || !charges) return AM1G_INVALID_INPUT;
    Topo T = perceive(n_atoms, z, xyz_bohr);
    std::vector<int> types(n_atoms);
    for (int i = 0; i < n_atoms; ++i) types[i] = bcc_type(T, i);
    std::vector<std::string> warn = T.warnings;
    for (int i = 0; i < n_atoms; ++i) charges[i] = mulliken[i];
    const auto& tab = bcc_table();
    for (size_t k = 0; k < T.bonds.size(); ++k) {
        const Bond& b = T.bonds[k];
        const int ta = types[b.i], tb = types[b.j];
        const std::vector<int> codes = bond_codes(T, (int)k, ta, tb);
        bool found = false;
        double delta = 0.0;
        for (int code : codes) {  // directional: (ta, tb) shifts +delta onto i, (tb, ta) the opposite
            auto f = tab.find(std::make_tuple(ta, tb, code));
            if (f != tab.end()) {
                delta = f->second;
                found = true;
                break;
            }
            f = tab.find(std::make_tuple(tb, ta, code));
            if (f != tab.end()) {
                delta = -f->second;
                found = true;
                break;
            }
        }
        if (found) {
            charges[b.i] += delta;
            charges[b.j] -= delta;
        } else if (ta != 0 && tb != 0) {
            char buf[256];
            std::snprintf(buf, sizeof(buf),
                          "no bond charge correction for atoms %d-%d (types %d-%d, bond type %d); that bond is left at "
                          "its raw AM1 Mulliken charges",
                          b.i, b.j, ta, tb, codes[0]);
            warn.push_back(buf);
        }
    }
    if (atom_types)
        for (int i = 0; i < n_atoms; ++i) atom_types[i] = types[i];
    if (warnings && warnings_len > 0) {
        std::string all;
        for (const std::string& w : warn) all += w + "\n";
        std::snprintf(warnings, (size_t)warnings_len, "%s", all.c_str());
    }
    if (n_warnings) *n_warnings = (int)warn.size();
    return AM1G_OK;
}
