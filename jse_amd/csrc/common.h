// Shared host-side definitions of libjsenn_hip.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace jnn {

// Error raised inside the library; the C-ABI turns it into a status code and
// a message readable with jse_nnap_last_error().
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

enum Status : int {
    OK = 0,
    ERR_ARG = 1,        // invalid argument / model
    ERR_UNSUPPORTED = 2,// model shape without a compiled kernel
    ERR_HIP = 3,        // HIP runtime error
    ERR_CAPACITY = 4,   // neighbour capacity exceeded (internal retry failed)
    ERR_TYPE = 5,       // atom type outside [1, ntypes]
    ERR_MEMORY = 6,     // device memory budget exhausted
};

#define JNN_HIP(call)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess)                                                                  \
            throw ::jnn::Error(::jnn::ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define JNN_CHECK(cond, code, msg)                  \
    do {                                            \
        if (!(cond)) throw ::jnn::Error((code), (msg)); \
    } while (0)

// radial mixing specialisations of the kernels
enum MixKind : int { MIX_NONE = 0, MIX_DEFAULT = 1, MIX_DENSE = 2 };

constexpr int kMaxLayers = 8;
// columns the L3 / L4 invariants add per radial channel, by l3max (0..6) and
// l4max (0..3) (SphericalChebyshev.java:15-16)
constexpr int kL3Cols[7] = {0, 0, 2, 4, 9, 14, 23};
constexpr int kL4Cols[4] = {0, 1, 3, 9};
constexpr int kMaxSpecies = 15;  // a neighbour's species travels in 4 bits of its pair tag

// One species of a model, FP64 blocks in the reference's packing plus derived shapes.
struct SpeciesHost {
    std::string symbol;
    int nmax = 5, lmax = 6, P = 6, wtype = 0, typed_wtype = 0;
    double rcut = 6.0;
    std::vector<int> hidden;          // hidden layer widths
    std::vector<double> W;            // [ntypes][P][K]
    std::vector<double> nn;           // [hidden W | out W | hidden b | out b]
    std::vector<double> norm;         // [muE+refE, sigmaE, mu[S], sigma[S]]
    int l3max = 0, l4max = 0;         // L3 / L4 invariants (SphericalChebyshev.java:33-42)
    int K() const { return nmax + 1; }
    // the harmonics' and A's lmax: max(lmax, l3max, l4max)
    int lmm() const { return lmax > l3max ? (lmax > l4max ? lmax : l4max) : (l3max > l4max ? l3max : l4max); }
    int M() const { return (lmm() + 1) * (lmm() + 1); }
    bool l34() const { return l3max > 0 || l4max > 0; }
    // W is a parameter block (fuse, rfuse, exfuse; WTypeBasis.parameterSize)
    bool mix_trainable() const { return typed_wtype == 4 || typed_wtype == 5 || typed_wtype == 6; }
    // descriptor columns per radial channel: lmax + 1 L2 values, then the L3 and L4 invariants
    int SL() const { return lmax + 1 + kL3Cols[l3max] + kL4Cols[l4max]; }
    int mix = MIX_DENSE;              // kernel specialisation for W
    int mirror = 0;                   // mirror basis: for this species' centres, neighbour species
                                      // `mirror` and this species swap roles (0: none)
    bool radial = false;              // radial-only chebyshev basis: fp[p] = sum_j fc Rnp[p] (lmax 0 sizes)
    // merge basis (MergedBasis.java): the parts in fp order, each with the basis
    // fields above; the species' own basis fields then hold part 0 and rcut the
    // largest part cutoff.  Empty for a single basis.
    std::vector<SpeciesHost> parts;
    // descriptor size (all parts of a merge basis)
    int S() const {
        if (parts.empty()) return P * SL();
        int s = 0;
        for (const SpeciesHost &q : parts) s += q.P * q.SL();
        return s;
    }
    // the bases evaluated for this species, in fp order
    int nbases() const { return parts.empty() ? 1 : (int)parts.size(); }
    const SpeciesHost &basis(int k) const { return parts.empty() ? *this : parts[k]; }
    // A floats per row, all bases
    int pm() const {
        int s = 0;
        for (int k = 0; k < nbases(); ++k) s += basis(k).P * basis(k).M();
        return s;
    }
};

struct ModelHost {
    std::vector<SpeciesHost> species;
    int ntypes() const { return (int)species.size(); }
    double rcut_max() const {
        double r = 0;
        for (auto &s : species) r = s.rcut > r ? s.rcut : r;
        return r;
    }
};

// model_host.cpp
ModelHost parse_model_json(const std::string &text);
// The JSON text of a YAML model file's data (the YAML subset a model file
// needs; status 1 naming the line for anything outside it).
std::string yaml_to_json(const std::string &yaml);
void finalize_species(ModelHost &m);   // derive mix kind, validate sizes

}  // namespace jnn
