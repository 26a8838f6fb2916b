// Host-side element tables: the published AM1/RM1 parameter rows turned into the DevElem records
// the kernels read. Derivations follow am1-rs params.rs:139-330 (dd, qq, the rho0/rho1/rho2
// additive terms by 30-step secant solves) and data_tables.rs:53-224 (valence configuration,
// isolated-atom energy coefficients, atomic heats of formation).
#include <cmath>
#include <cstring>

#include "../data/nddo_params.inc"
#include "am1g_common.hpp"

namespace am1g {
namespace {

struct Valence {
    int z;
    double ns, np, eheat_kcal;
};

// Neutral valence occupations (data_tables.rs:53-87) and gas-phase atomic heats of formation
// in kcal/mol (data_tables.rs:89-100) for every element either parameter set covers.
constexpr Valence kValence[] = {
    {1, 1, 0, 52.102},  {4, 2, 0, 76.960},  {5, 2, 1, 135.700}, {6, 2, 2, 170.890}, {7, 2, 3, 113.000},
    {8, 2, 4, 59.559},  {9, 2, 5, 18.890},  {13, 2, 1, 79.490}, {14, 2, 2, 108.390}, {15, 2, 3, 75.570},
    {16, 2, 4, 66.400}, {17, 2, 5, 28.990}, {30, 2, 0, 31.170}, {32, 2, 2, 89.500}, {33, 2, 3, 72.300},
    {34, 2, 4, 54.300}, {35, 2, 5, 26.740}, {51, 2, 3, 63.200}, {52, 2, 4, 47.000}, {53, 2, 5, 25.517},
    {80, 2, 0, 14.690},
};

int principal_n(int z) {
    return z <= 2 ? 1 : z <= 10 ? 2 : z <= 18 ? 3 : z <= 36 ? 4 : z <= 54 ? 5 : 6;
}

// Secant iteration shared by the two additive-term solves: find x with g(x) = target.
template <typename G>
double secant(G g, double target, double start) {
    double a = start, b = start + 0.04;
    for (int it = 0; it < 30; ++it) {
        double ga = g(a), gb = g(b);
        double c = std::fabs(gb - ga) > 1.0e-16 ? a + (b - a) * (target - ga) / (gb - ga) : b;
        a = b;
        b = c;
    }
    return b;
}

DevElem make(const NddoParamRow& row) {
    DevElem e;
    std::memset(&e, 0, sizeof(e));
    const double* v = row.v;
    e.uss = v[0]; e.upp = v[1]; e.zs = v[2]; e.zp = v[3]; e.bs = v[4]; e.bp = v[5];
    e.gss = v[6]; e.gsp = v[7]; e.gpp = v[8]; e.gp2 = v[9]; e.hsp = v[10]; e.alpha = v[11];
    e.n = principal_n(row.z);
    e.norb = e.zp > 0.0 ? 4 : 1;
    for (int k = 0; k < 4; ++k) {
        double K = v[12 + 3 * k], L = v[13 + 3 * k], M = v[14 + 3 * k];
        if (K != 0.0 || L != 0.0) {
            e.gk[e.ng] = K;
            e.gl[e.ng] = L;
            e.gm[e.ng] = M;
            ++e.ng;
        }
    }
    const double n = e.n;
    if (e.norb == 4) {
        // charge separations of the dipole and quadrupole (MOPAC ddpo.f form)
        e.dd = (2.0 * n + 1.0) * std::pow(4.0 * e.zs * e.zp, n + 0.5) / std::pow(e.zs + e.zp, 2.0 * n + 2.0) /
               std::sqrt(3.0);
        e.qq = std::sqrt((4.0 * n * n + 6.0 * n + 2.0) / 20.0) / e.zp;
        const double d1 = e.dd, d2 = e.qq;
        const double hsp = e.hsp / kEv;
        double a1 = std::cbrt(std::fabs(hsp) / (d1 * d1));
        if (hsp < 0.0) a1 = -a1;
        double x1 = secant([&](double d) { return 0.5 * d - 0.5 / std::sqrt(4.0 * d1 * d1 + 1.0 / (d * d)); }, hsp, a1);
        e.rho1 = 0.5 / x1;
        const double hpp = std::fmax(0.5 * (e.gpp - e.gp2), 0.1) / kEv;
        double a2 = std::pow(std::fabs(hpp) / (3.0 * d2 * d2 * d2 * d2), 0.2);
        if (hpp < 0.0) a2 = -a2;
        double x2 = secant(
            [&](double q) {
                return 0.25 * q - 0.5 / std::sqrt(4.0 * d2 * d2 + 1.0 / (q * q)) +
                       0.25 / std::sqrt(8.0 * d2 * d2 + 1.0 / (q * q));
            },
            hpp, a2);
        e.rho2 = 0.5 / x2;
    }
    e.rho0 = 0.5 * kEv / e.gss;
    double ns = 0.0, np = 0.0, eheat = 0.0;
    for (const Valence& vl : kValence)
        if (vl.z == row.z) {
            ns = vl.ns;
            np = vl.np;
            eheat = vl.eheat_kcal;
        }
    // average-of-configuration coefficients of the isolated atom (MOPAC calpar.f closed forms)
    const int npi = (int)np;
    const double c_gss = ns * (ns - 1.0) / 2.0, c_gsp = ns * np, c_hsp = npi == 6 ? 0.0 : -np;
    const double c_gp2 = npi == 2 ? 1.5 : npi == 3 ? 4.5 : npi == 4 ? 6.5 : npi == 5 ? 10.0 : 0.0;
    const double c_gpp = npi == 2 ? -0.5 : npi == 3 ? -1.5 : npi == 4 ? -0.5 : 0.0;
    e.e_isol = e.uss * ns + e.upp * np + e.gss * c_gss + e.gpp * c_gpp + e.gsp * c_gsp + e.gp2 * c_gp2 + e.hsp * c_hsp;
    e.core = ns + np;
    e.eheat_ev = eheat / kEvToKcal;
    e.valid = 1;
    return e;
}

}  // namespace

// tab[method * kMaxZ + z]; invalid entries have valid == 0.
void build_element_tables(DevElem* tab) {
    std::memset(tab, 0, sizeof(DevElem) * kMethods * kMaxZ);
    for (const NddoParamRow& r : kAM1Rows) tab[0 * kMaxZ + r.z] = make(r);
    for (const NddoParamRow& r : kRM1Rows) tab[1 * kMaxZ + r.z] = make(r);
}

}  // namespace am1g
