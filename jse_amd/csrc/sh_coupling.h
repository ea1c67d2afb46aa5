// Coupling terms of the L3 / L4 invariants of a spherical-Chebyshev basis
// (basis_SphericalUtil.hpp:395-467).  Per radial channel, after its lmax + 1
// L2 values, the descriptor carries kL3Cols[l3max] cubic and kL4Cols[l4max]
// quartic invariants of the channel's A row (index l^2 + l + m):
//   L3(l1 l2 l3)    = sum_{m} R(l1 l2 l3)_{m1 m2 m3} A_{l1 m1} A_{l2 m2} A_{l3 m3}
//   L4(l1 l2 l3 l4) = sum_L sum_M u_{LM} v_{LM},  u_{LM} = sum R(l1 l2 L)_{m1 m2 M} A_{l1 m1} A_{l2 m2},
//                     v_{LM} likewise with (l3 l4 L), L over every value both couplings allow,
// where R is the Wigner 3j symbol carried to the real harmonics (real Y_{l,m>0}
// = sqrt2 Re Y^c_{l,m}, Y_{l,-m} = sqrt2 Im Y^c_{l,m}, Condon-Shortley Y^c).
// Expanded, each invariant is a sparse polynomial in A; these are its terms
// (indices sorted, equal index tuples merged), and the terms of its
// derivative by each A entry, for the kernels.
#pragma once
#include <vector>

#include "common.h"

namespace jnn {

struct CouplingTerms {
    int ninv = 0;               // kL3Cols[l3max] + kL4Cols[l4max]
    std::vector<int> off;       // [ninv + 1]: terms of invariant k are off[k] .. off[k+1]-1
    std::vector<int> idx;       // 4 per term: A indices; the fourth is -1 for a cubic term
    std::vector<double> coef;   // per term
    std::vector<int> goff;      // [M + 1]: terms of d/dA[m] are goff[m] .. goff[m+1]-1
    std::vector<int> gidx;      // 4 per term: the other factors' A indices (the third -1 for a
                                // cubic invariant), then the invariant k
    std::vector<double> gcoef;  // per term (the term's coefficient times the multiplicity of m)
};

// The terms for one basis; M = (lmax' + 1)^2 with lmax' >= max(l3max, l4max).
CouplingTerms coupling_terms(int l3max, int l4max, int M);

}  // namespace jnn
