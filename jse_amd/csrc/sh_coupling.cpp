// Coupling terms of the L3 / L4 invariants (see sh_coupling.h).  The
// invariants' order is the reference's column order (the order of its
// L3_INDEX / L4_INDEX tables, basis_SphericalUtil0.hpp:10-500).
#include "sh_coupling.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <complex>
#include <cstdlib>
#include <map>

#include "common.h"

namespace jnn {
namespace {

using cplx = std::complex<double>;

const int kL3List[23][3] = {
    {2, 2, 2}, {1, 1, 2},                                                    // l3max 2
    {2, 3, 3}, {1, 2, 3},                                                    // 3
    {4, 4, 4}, {2, 2, 4}, {3, 3, 4}, {2, 4, 4}, {1, 3, 4},                   // 4
    {2, 5, 5}, {4, 5, 5}, {2, 3, 5}, {1, 4, 5}, {3, 4, 5},                   // 5
    {6, 6, 6}, {3, 3, 6}, {4, 4, 6}, {5, 5, 6}, {2, 6, 6}, {4, 6, 6}, {2, 4, 6}, {1, 5, 6}, {3, 5, 6}};  // 6
const int kL4List[9][4] = {
    {1, 1, 1, 1},                                                            // l4max 1
    {2, 2, 2, 2}, {1, 1, 2, 2},                                              // 2
    {3, 3, 3, 3}, {1, 1, 3, 3}, {2, 2, 3, 3}, {1, 1, 1, 3}, {1, 3, 3, 3}, {1, 2, 2, 3}};  // 3

double factorial(int n) {
    double f = 1.0;
    for (int k = 2; k <= n; ++k) f *= k;
    return f;
}

// Wigner 3j symbol (j1 j2 j3; m1 m2 m3) by the Racah formula
double wigner3j(int j1, int j2, int j3, int m1, int m2, int m3) {
    if (m1 + m2 + m3 != 0 || j3 < std::abs(j1 - j2) || j3 > j1 + j2 || std::abs(m1) > j1 || std::abs(m2) > j2 ||
        std::abs(m3) > j3)
        return 0.0;
    const double tri = factorial(j1 + j2 - j3) * factorial(j1 - j2 + j3) * factorial(-j1 + j2 + j3) /
                       factorial(j1 + j2 + j3 + 1);
    const double pre = std::sqrt(tri * factorial(j1 + m1) * factorial(j1 - m1) * factorial(j2 + m2) *
                                 factorial(j2 - m2) * factorial(j3 + m3) * factorial(j3 - m3));
    const int kmin = std::max({0, j2 - j3 - m1, j1 - j3 + m2});
    const int kmax = std::min({j1 + j2 - j3, j1 - m1, j2 + m2});
    double s = 0.0;
    for (int k = kmin; k <= kmax; ++k)
        s += ((k & 1) ? -1.0 : 1.0) / (factorial(k) * factorial(j1 + j2 - j3 - k) * factorial(j1 - m1 - k) *
                                       factorial(j2 + m2 - k) * factorial(j3 - j2 + m1 + k) *
                                       factorial(j3 - j1 - m2 + k));
    return (((j1 - j2 - m3) % 2) ? -1.0 : 1.0) * pre * s;
}

// real harmonic m as a combination of complex ones: Y_m = sum_k u[k] Y^c_{mk[k]}
int real_to_complex(int m, int *mk, cplx *u) {
    if (m == 0) {
        mk[0] = 0;
        u[0] = 1.0;
        return 1;
    }
    const int a = std::abs(m);
    const double h = std::sqrt(0.5), sg = (a & 1) ? -1.0 : 1.0;
    if (m > 0) {
        mk[0] = a;   u[0] = h;
        mk[1] = -a;  u[1] = sg * h;
    } else {
        mk[0] = -a;  u[0] = cplx(0.0, sg * h);
        mk[1] = a;   u[1] = cplx(0.0, -h);
    }
    return 2;
}

// R(l1 l2 l3) as [(2l1+1)][(2l2+1)][(2l3+1)] (entry a <-> m = a - l)
std::vector<cplx> real_3j(int l1, int l2, int l3) {
    const int n1 = 2 * l1 + 1, n2 = 2 * l2 + 1, n3 = 2 * l3 + 1;
    std::vector<cplx> R((size_t)n1 * n2 * n3);
    for (int a = 0; a < n1; ++a)
        for (int b = 0; b < n2; ++b)
            for (int c = 0; c < n3; ++c) {
                int k1[2], k2[2], k3[2];
                cplx u1[2], u2[2], u3[2], s = 0.0;
                const int r1 = real_to_complex(a - l1, k1, u1), r2 = real_to_complex(b - l2, k2, u2),
                          r3 = real_to_complex(c - l3, k3, u3);
                for (int x = 0; x < r1; ++x)
                    for (int y = 0; y < r2; ++y)
                        for (int z = 0; z < r3; ++z)
                            if (k1[x] + k2[y] + k3[z] == 0)
                                s += wigner3j(l1, l2, l3, k1[x], k2[y], k3[z]) *
                                     std::conj(u1[x] * u2[y] * u3[z]);
                R[((size_t)a * n2 + b) * n3 + c] = s;
            }
    return R;
}

using Key = std::array<int, 4>;  // sorted A indices, -1 (unused) first
constexpr double kDrop = 1e-14;  // merged coefficients below this are zeros of the expansion

void add_term(std::map<Key, double> &poly, std::array<int, 4> ix, int nf, double c) {
    Key k{-1, -1, -1, -1};
    for (int f = 0; f < nf; ++f) k[4 - nf + f] = ix[f];
    std::sort(k.begin(), k.end());
    poly[k] += c;
}

}  // namespace

CouplingTerms coupling_terms(int l3max, int l4max, int M) {
    JNN_CHECK(l3max >= 0 && l3max <= 6 && l4max >= 0 && l4max <= 3, ERR_ARG, "l3max / l4max out of range");
    const int lm = std::max(l3max, l4max);
    JNN_CHECK(M >= (lm + 1) * (lm + 1), ERR_ARG, "coupling terms: A row too short for l3max / l4max");
    CouplingTerms t;
    t.ninv = kL3Cols[l3max] + kL4Cols[l4max];
    std::vector<std::map<Key, double>> polys(t.ninv);
    auto base = [](int l) { return l * l; };
    for (int k = 0; k < kL3Cols[l3max]; ++k) {
        const int *l = kL3List[k];
        const std::vector<cplx> R = real_3j(l[0], l[1], l[2]);
        const int n2 = 2 * l[1] + 1, n3 = 2 * l[2] + 1;
        for (int a = 0; a < 2 * l[0] + 1; ++a)
            for (int b = 0; b < n2; ++b)
                for (int c = 0; c < n3; ++c) {
                    const double v = R[((size_t)a * n2 + b) * n3 + c].real();
                    if (v != 0.0) add_term(polys[k], {base(l[0]) + a, base(l[1]) + b, base(l[2]) + c, 0}, 3, v);
                }
    }
    for (int k = 0; k < kL4Cols[l4max]; ++k) {
        const int *l = kL4List[k];
        const int n1 = 2 * l[0] + 1, n2 = 2 * l[1] + 1, n3 = 2 * l[2] + 1, n4 = 2 * l[3] + 1;
        std::vector<double> T((size_t)n1 * n2 * n3 * n4, 0.0);
        const int L0 = std::max(std::abs(l[0] - l[1]), std::abs(l[2] - l[3]));
        const int L1 = std::min(l[0] + l[1], l[2] + l[3]);
        for (int L = L0; L <= L1; ++L) {
            const std::vector<cplx> Ra = real_3j(l[0], l[1], L), Rb = real_3j(l[2], l[3], L);
            const int nL = 2 * L + 1;
            for (int a = 0; a < n1; ++a)
                for (int b = 0; b < n2; ++b)
                    for (int c = 0; c < n3; ++c)
                        for (int d = 0; d < n4; ++d) {
                            cplx s = 0.0;
                            for (int q = 0; q < nL; ++q)
                                s += Ra[((size_t)a * n2 + b) * nL + q] * Rb[((size_t)c * n4 + d) * nL + q];
                            T[(((size_t)a * n2 + b) * n3 + c) * n4 + d] += s.real();
                        }
        }
        for (int a = 0; a < n1; ++a)
            for (int b = 0; b < n2; ++b)
                for (int c = 0; c < n3; ++c)
                    for (int d = 0; d < n4; ++d) {
                        const double v = T[(((size_t)a * n2 + b) * n3 + c) * n4 + d];
                        if (v != 0.0)
                            add_term(polys[kL3Cols[l3max] + k],
                                     {base(l[0]) + a, base(l[1]) + b, base(l[2]) + c, base(l[3]) + d}, 4, v);
                    }
    }
    // forward terms, and the derivative terms grouped by A index
    std::vector<std::map<std::array<int, 4>, double>> grad(M);  // key: other factors (sorted, -1 padded), k
    t.off.push_back(0);
    for (int k = 0; k < t.ninv; ++k) {
        for (const auto &kv : polys[k]) {
            if (std::fabs(kv.second) < kDrop) continue;
            const Key &ix = kv.first;  // a cubic term's key is (-1, i0, i1, i2)
            const int nf = ix[0] < 0 ? 3 : 4;
            const int *f = ix.data() + (4 - nf);
            t.idx.insert(t.idx.end(), {f[0], f[1], f[2], nf == 4 ? f[3] : -1});
            t.coef.push_back(kv.second);
            for (int q = 0; q < nf; ++q) {
                if (q > 0 && f[q] == f[q - 1]) continue;  // each distinct index once
                int mult = 0, rest[3], nr = 0;
                bool skipped = false;
                for (int r = 0; r < nf; ++r) {
                    if (f[r] == f[q]) ++mult;
                    if (f[r] == f[q] && !skipped) skipped = true;
                    else rest[nr++] = f[r];
                }
                std::array<int, 4> g{rest[0], rest[1], nf == 4 ? rest[2] : -1, k};
                grad[f[q]][g] += kv.second * mult;
            }
        }
        t.off.push_back((int)t.coef.size());
    }
    t.goff.push_back(0);
    for (int m = 0; m < M; ++m) {
        for (const auto &kv : grad[m]) {
            t.gidx.insert(t.gidx.end(), kv.first.begin(), kv.first.end());
            t.gcoef.push_back(kv.second);
        }
        t.goff.push_back((int)t.gcoef.size());
    }
    return t;
}

}  // namespace jnn
