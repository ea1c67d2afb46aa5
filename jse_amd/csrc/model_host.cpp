// Model-file loading for libjsenn_hip: a small JSON reader and the
// construction of the packed per-species blocks, following the reference's
// NNAP constructor (src/nnap/java/jsex/nnap/NNAP.java:129-307), the basis
// loader (basis/SphericalChebyshev.java:60-103, basis/WTypeBasis.java:202-354)
// and FeedForward.load (nn/FeedForward.java:131-221).
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <map>
#include <memory>

#include "common.h"

namespace jnn {
namespace {

struct JVal {
    enum Kind { NUL, BOOL, NUM, STR, ARR, OBJ } kind = NUL;
    bool b = false;
    double num = 0;
    std::string str;
    std::vector<JVal> arr;
    std::vector<std::pair<std::string, JVal>> obj;
    const JVal *get(const char *k) const {
        if (kind != OBJ) return nullptr;
        for (auto &kv : obj)
            if (kv.first == k) return kv.second.kind == NUL ? nullptr : &kv.second;
        return nullptr;
    }
};

struct JParser {
    const std::string &s;
    size_t p = 0;
    int depth = 0;  // open values: a model file nests a few levels; past 64 the text is refused
    explicit JParser(const std::string &t) : s(t) {}
    [[noreturn]] void fail(const char *what) {
        throw Error(ERR_ARG, std::string("model JSON: ") + what + " at offset " + std::to_string(p));
    }
    void ws() {
        while (p < s.size() && (s[p] == ' ' || s[p] == '\n' || s[p] == '\r' || s[p] == '\t')) ++p;
    }
    struct Nest {
        int &d;
        explicit Nest(JParser &j) : d(j.depth) {
            if (d >= 64) j.fail("nesting deeper than 64 levels");
            ++d;
        }
        ~Nest() { --d; }
    };
    JVal value() {
        const Nest nest(*this);
        ws();
        if (p >= s.size()) fail("unexpected end");
        char c = s[p];
        JVal v;
        if (c == '{') {
            v.kind = JVal::OBJ;
            ++p;
            ws();
            if (p < s.size() && s[p] == '}') { ++p; return v; }
            for (;;) {
                ws();
                if (p >= s.size() || s[p] != '"') fail("expected key");
                std::string k = str();
                ws();
                if (p >= s.size() || s[p] != ':') fail("expected ':'");
                ++p;
                v.obj.emplace_back(k, value());
                ws();
                if (p < s.size() && s[p] == ',') { ++p; continue; }
                if (p < s.size() && s[p] == '}') { ++p; break; }
                fail("expected ',' or '}'");
            }
        } else if (c == '[') {
            v.kind = JVal::ARR;
            ++p;
            ws();
            if (p < s.size() && s[p] == ']') { ++p; return v; }
            for (;;) {
                v.arr.push_back(value());
                ws();
                if (p < s.size() && s[p] == ',') { ++p; continue; }
                if (p < s.size() && s[p] == ']') { ++p; break; }
                fail("expected ',' or ']'");
            }
        } else if (c == '"') {
            v.kind = JVal::STR;
            v.str = str();
        } else if (s.compare(p, 4, "true") == 0) {
            v.kind = JVal::BOOL; v.b = true; p += 4;
        } else if (s.compare(p, 5, "false") == 0) {
            v.kind = JVal::BOOL; v.b = false; p += 5;
        } else if (s.compare(p, 4, "null") == 0) {
            p += 4;
        } else if (s.compare(p, 3, "NaN") == 0) {
            v.kind = JVal::NUM; v.num = NAN; p += 3;
        } else {
            const char *b = s.c_str() + p;
            char *e = nullptr;
            v.num = std::strtod(b, &e);
            if (e == b) fail("bad value");
            v.kind = JVal::NUM;
            p += (size_t)(e - b);
        }
        return v;
    }
    std::string str() {
        ++p;  // opening quote
        std::string out;
        while (p < s.size() && s[p] != '"') {
            if (s[p] == '\\' && p + 1 < s.size()) {
                ++p;
                char e = s[p];
                if (e == 'n') out += '\n';
                else if (e == 't') out += '\t';
                else if (e == 'u') { out += '?'; p += 4; }
                else out += e;
                ++p;
            } else {
                out += s[p++];
            }
        }
        if (p >= s.size()) fail("unterminated string");
        ++p;
        return out;
    }
};

double num(const JVal *v, double dflt) {
    if (!v) return dflt;
    if (v->kind == JVal::NUM) return v->num;
    if (v->kind == JVal::BOOL) return v->b ? 1.0 : 0.0;
    throw Error(ERR_ARG, "model JSON: expected a number");
}
std::vector<double> vec(const JVal *v) {
    std::vector<double> out;
    if (!v) return out;
    if (v->kind != JVal::ARR) throw Error(ERR_ARG, "model JSON: expected an array");
    out.reserve(v->arr.size());
    for (auto &e : v->arr) out.push_back(num(&e, 0));
    return out;
}
const JVal *first(const JVal &m, std::initializer_list<const char *> ks) {
    for (auto k : ks)
        if (auto v = m.get(k)) return v;
    return nullptr;
}

// weight standardisation of the rfuse weight: each row's norm starts at this
// epsilon (WTypeBasis.java:20, used at :216)
constexpr double kWsEps = 1e-5;

constexpr int WT_DEFAULT = 0, WT_NONE = -1, WT_FULL = 2, WT_EXFULL = 3, WT_FUSE = 4, WT_RFUSE = 5, WT_EXFUSE = 6;

int size_n(int wtype, int ntypes, int nmax, int fuse) {  // WTypeBasis.java:331-354
    switch (wtype) {
        case WT_EXFULL: return ntypes > 1 ? (ntypes + 1) * (nmax + 1) : nmax + 1;
        case WT_FULL: case WT_RFUSE: return ntypes * (nmax + 1);
        case WT_NONE: return nmax + 1;
        case WT_DEFAULT: return ntypes > 1 ? 2 * nmax + 2 : nmax + 1;
        case WT_FUSE: return fuse * (nmax + 1);
        case WT_EXFUSE: return (fuse + 1) * (nmax + 1);
    }
    throw Error(ERR_ARG, "Input wtype MUST be in {-1, 0, 2, 3, 4, 5, 6}, input: " + std::to_string(wtype));
}

}  // namespace

// the single-basis fields of `dst` from `src` (a mirror / share species, or a
// merge species' part 0)
static void set_basis_fields(SpeciesHost &dst, const SpeciesHost &src) {
    dst.nmax = src.nmax;
    dst.lmax = src.lmax;
    dst.P = src.P;
    dst.wtype = src.wtype;
    dst.typed_wtype = src.typed_wtype;
    dst.rcut = src.rcut;
    dst.W = src.W;
    dst.radial = src.radial;
    dst.l3max = src.l3max;
    dst.l4max = src.l4max;
}

// SphericalChebyshev.load / Chebyshev.load (SphericalChebyshev.java:60-103,
// Chebyshev.java) with the WTypeBasis mixing (WTypeBasis.java:202-354): one
// basis entry `b` of type `type` into the basis fields of `sp`.
static void parse_basis(const JVal *b, const std::string &type, int ntypes, SpeciesHost &sp) {
    // Chebyshev.java: the radial-only basis (no angular part)
    sp.radial = type == "chebyshev" || type == "cheby";
    JNN_CHECK(sp.radial || type == "spherical_chebyshev" || type == "sph_cheby", ERR_UNSUPPORTED,
              "Unsupported basis type: " + type);
    sp.nmax = (int)num(b->get("nmax"), 5);
    sp.lmax = sp.radial ? 0 : (int)num(b->get("lmax"), 6);
    JNN_CHECK(sp.lmax >= 0 && sp.lmax <= 12, ERR_ARG, "Input lmax MUST be in [0, 12], input: " + std::to_string(sp.lmax));
    // SphericalChebyshev.java:33-34 (the radial basis has no angular part)
    sp.l3max = sp.radial ? 0 : (int)num(b->get("l3max"), 0);
    sp.l4max = sp.radial ? 0 : (int)num(b->get("l4max"), 0);
    JNN_CHECK(sp.l3max >= 0 && sp.l3max <= 6, ERR_ARG, "Input l3max MUST be in [0, 6], input: " + std::to_string(sp.l3max));
    JNN_CHECK(sp.l4max >= 0 && sp.l4max <= 3, ERR_ARG, "Input l4max MUST be in [0, 3], input: " + std::to_string(sp.l4max));
    JNN_CHECK(sp.nmax >= 0 && sp.nmax <= 20, ERR_ARG,  // WTypeBasis.java:56
              "Input nmax MUST be in [0, 20], input: " + std::to_string(sp.nmax));
    sp.rcut = num(b->get("rcut"), 6.0);
    JNN_CHECK(sp.rcut > 0, ERR_ARG, "rcut must be positive");
    int wtype = WT_DEFAULT;
    if (auto w = b->get("wtype")) {
        if (w->kind == JVal::NUM) wtype = (int)w->num;
        else {
            static const std::map<std::string, int> names = {{"default", WT_DEFAULT}, {"none", WT_NONE},
                                                             {"full", WT_FULL},       {"exfull", WT_EXFULL},
                                                             {"fuse", WT_FUSE},       {"rfuse", WT_RFUSE},
                                                             {"exfuse", WT_EXFUSE}};
            auto it = names.find(w->str);
            JNN_CHECK(it != names.end(), ERR_ARG, "Input wtype MUST be in {-1, 0, 2, 3, 4, 5, 6}, input: " + w->str);
            wtype = it->second;
        }
    }
    sp.wtype = wtype;
    JNN_CHECK(wtype == WT_DEFAULT || wtype == WT_NONE || (wtype >= WT_FULL && wtype <= WT_EXFUSE), ERR_ARG,
              "Input wtype MUST be in {-1, 0, 2, 3, 4, 5, 6}, input: " + std::to_string(wtype));
    // the fuse weight is read for the fuse wtypes only and the rfuse weight for
    // rfuse only; other wtypes ignore those keys (WTypeBasis.java:365-410)
    std::vector<double> fuse_w, rfuse_w;
    if (wtype == WT_FUSE || wtype == WT_EXFUSE) {
        const JVal *fw = b->get("fuse_weight");
        const JVal *fs = b->get("fuse_size");
        if (fw) {
            JNN_CHECK(fw->kind == JVal::ARR && !fw->arr.empty(), ERR_ARG, "model JSON: fuse_weight must be a non-empty array");
            const bool nested = fw->arr[0].kind == JVal::ARR;
            if (!fs && nested) {
                // an older file: the inner lists are the columns of the [ntypes][fuse] matrix
                const size_t ncol = fw->arr.size(), nrow = fw->arr[0].arr.size();
                fuse_w.assign(nrow * ncol, 0.0);
                for (size_t c = 0; c < ncol; ++c) {
                    std::vector<double> col = vec(&fw->arr[c]);
                    JNN_CHECK(col.size() == nrow, ERR_ARG, "Size of fuse weight mismatch");
                    for (size_t r = 0; r < nrow; ++r) fuse_w[r * ncol + c] = col[r];
                }
            } else if (nested) {
                for (auto &row : fw->arr) {
                    std::vector<double> rr = vec(&row);
                    JNN_CHECK((int)rr.size() == (int)num(fs, 0), ERR_ARG, "Column number of fuse weight mismatch");
                    fuse_w.insert(fuse_w.end(), rr.begin(), rr.end());
                }
            } else {
                fuse_w = vec(fw);
                if (fs) JNN_CHECK((long)fuse_w.size() == (long)num(fs, 0) * ntypes, ERR_ARG, "Size of fuse weight mismatch");
            }
        } else {
            JNN_CHECK(fs, ERR_ARG, "Key `fuse_weight` or `fuse_size` required for fuse wtype");
            fuse_w.assign((size_t)num(fs, 0) * ntypes, 0.0);
        }
        JNN_CHECK(!fuse_w.empty() && fuse_w.size() % ntypes == 0, ERR_ARG, "Size of fuse weight mismatch");
    }
    const int fuse0 = (wtype == WT_FUSE || wtype == WT_EXFUSE) ? (int)fuse_w.size() / ntypes : 0;
    if (wtype == WT_RFUSE) {
        JNN_CHECK(!b->get("post_fuse_size") && !b->get("post_fuse_weight"), ERR_ARG,
                  "Key `post_fuse_size` or `post_fuse_weight` is invalid for wtype=='rfuse'");
        const int sn0 = size_n(wtype, ntypes, sp.nmax, 0);
        const JVal *rs = b->get("rfuse_size");
        if (const JVal *rw = b->get("rfuse_weight")) {
            rfuse_w = vec(rw);
            if (rs) JNN_CHECK((long)rfuse_w.size() == (long)num(rs, 0) * sn0, ERR_ARG, "Size of rfuse weight mismatch");
        } else {
            JNN_CHECK(rs, ERR_ARG, "Key `rfuse_weight` or `rfuse_size` required for wtype=='rfuse'");
            rfuse_w.assign((size_t)num(rs, 0) * sn0, 0.0);
        }
        JNN_CHECK(!rfuse_w.empty() && rfuse_w.size() % sn0 == 0, ERR_ARG, "Size of rfuse weight mismatch");
    }
    double rfuse_scale = num(b->get("rfuse_scale"), 1.0);
    bool ws = num(first(*b, {"weight_standardization", "ws"}), 0.0) != 0.0;
    const int K = sp.K();
    // base mixing W[t][q][n] of a non-rfuse wtype (sizeN rows), WTypeBasis.java:202-304
    auto base_mixing = [&](int typed_, int sn_, int fuse_) {
        std::vector<double> Wb((size_t)ntypes * sn_ * K, 0.0);
        auto at = [&](int t, int q, int n) -> double & { return Wb[((size_t)(t - 1) * sn_ + q) * K + n]; };
        for (int t = 1; t <= ntypes; ++t)
            for (int n = 0; n < K; ++n) {
                switch (typed_) {
                    case WT_NONE: at(t, n, n) = 1.0; break;
                    case WT_DEFAULT: at(t, n, n) = 1.0; at(t, K + n, n) = (t & 1) ? (double)t : -(double)t; break;
                    case WT_FULL: at(t, (t - 1) * K + n, n) = 1.0; break;
                    case WT_EXFULL: at(t, n, n) = 1.0; at(t, t * K + n, n) = 1.0; break;
                    case WT_FUSE:
                        for (int k = 0; k < fuse_; ++k) at(t, k * K + n, n) = fuse_w[(size_t)(t - 1) * fuse_ + k];
                        break;
                    case WT_EXFUSE:
                        at(t, n, n) = 1.0;
                        for (int k = 0; k < fuse_; ++k) at(t, (k + 1) * K + n, n) = fuse_w[(size_t)(t - 1) * fuse_ + k];
                        break;
                    default: throw Error(ERR_ARG, "unsupported wtype");
                }
            }
        return Wb;
    };
    auto typed_of = [&](int wt_) {
        return (ntypes == 1 && (wt_ == WT_EXFULL || wt_ == WT_FULL || wt_ == WT_NONE || wt_ == WT_DEFAULT)) ? WT_NONE
                                                                                                        : wt_;
    };
    // an older file's post-fuse weight becomes the rfuse weight it equals, rfuse[t] = post . W_base[t]
    // (SphericalChebyshev.java:81-94; WTypeBasis.getPostFuseWeight_ / postFuse2RFuse_, :391-540)
    const bool post_keys = b->get("post_fuse_size") || b->get("post_fuse_weight");
    if (wtype == WT_RFUSE) {
        JNN_CHECK(!post_keys, ERR_ARG, "Key `post_fuse_size` or `post_fuse_weight` is invalid for wtype=='rfuse'");
    } else {
        const JVal *pf = b->get("post_fuse");
        const bool off = pf && pf->kind == JVal::BOOL && !pf->b;
        if (!off && (pf || post_keys)) {
            const std::vector<double> post = vec(b->get("post_fuse_weight"));
            JNN_CHECK(!post.empty(), ERR_ARG,
                      "Key `post_fuse_weight` always required for post_fuse, training with post_fuse is invalid now");
            const int fz = fuse0;
            const int sn0 = size_n(wtype, ntypes, sp.nmax, fz);
            if (auto ps = b->get("post_fuse_size"))
                JNN_CHECK(post.size() == (size_t)num(ps, 0) * sn0, ERR_ARG, "Size of post fuse weight mismatch");
            const int pp = (int)post.size() / sn0;
            const std::vector<double> Wb = base_mixing(typed_of(wtype), sn0, fz);
            rfuse_w.assign((size_t)ntypes * pp * K, 0.0);
            for (int t = 0; t < ntypes; ++t)
                for (int p = 0; p < pp; ++p)
                    for (int q = 0; q < sn0; ++q) {
                        const double w = post[(size_t)p * sn0 + q];
                        if (w == 0.0) continue;
                        for (int n = 0; n < K; ++n)
                            rfuse_w[((size_t)t * pp + p) * K + n] += w * Wb[((size_t)t * sn0 + q) * K + n];
                    }
            rfuse_scale = num(b->get("post_fuse_scale"), 1.0);
            wtype = WT_RFUSE;
            sp.wtype = wtype;
            fuse_w.clear();
        }
    }
    int typed = typed_of(wtype);
    sp.typed_wtype = typed;
    const int fuse = (wtype == WT_FUSE || wtype == WT_EXFUSE) ? (int)fuse_w.size() / ntypes : 0;
    const int sn = size_n(wtype, ntypes, sp.nmax, fuse);
    sp.P = rfuse_w.empty() ? sn : (int)rfuse_w.size() / sn;
    JNN_CHECK(sp.P >= 1 && sp.P <= 128, ERR_ARG, "basis size_np out of range");
    sp.W.assign((size_t)ntypes * sp.P * K, 0.0);
    if (!rfuse_w.empty()) {
        JNN_CHECK(rfuse_w.size() == (size_t)sp.P * sn, ERR_ARG, "Size of rfuse weight mismatch");
        std::vector<double> rw = rfuse_w;
        if (ws) {
            for (int p = 0; p < sp.P; ++p) {
                double norm = kWsEps;
                for (int t = 1; t <= ntypes; ++t)
                    for (int n = 0; n < K; ++n) norm += std::fabs(rfuse_w[((size_t)(t - 1) * sp.P + p) * K + n]);
                norm /= sn;
                for (int t = 1; t <= ntypes; ++t)
                    for (int n = 0; n < K; ++n) rw[((size_t)(t - 1) * sp.P + p) * K + n] /= norm;
            }
        }
        for (size_t q = 0; q < rw.size(); ++q) sp.W[q] = rfuse_scale * rw[q];
    } else {
        sp.W = base_mixing(typed, sn, fuse);
    }
}

ModelHost parse_model_json(const std::string &text) {
    JParser jp(text);
    JVal root = jp.value();
    if (root.kind != JVal::OBJ) throw Error(ERR_ARG, "model JSON: top level must be an object");
    if (auto v = root.get("version"))
        if ((int)num(v, 0) > 6) throw Error(ERR_ARG, "Unsupported version: " + std::to_string((int)num(v, 0)));
    const JVal *models = root.get("models");
    if (!models || models->kind != JVal::ARR) throw Error(ERR_ARG, "No models in ModelInfo");
    const int ntypes = (int)models->arr.size();
    JNN_CHECK(ntypes >= 1, ERR_ARG, "No models in ModelInfo");
    double sig_e = 1.0, mu_e = 0.0;
    bool have_s = false, have_m = false;
    for (auto &m : models->arr) {  // NNAP.java:272-278
        if (!have_s && m.get("norm_sigma_eng")) { sig_e = num(m.get("norm_sigma_eng"), 1.0); have_s = true; }
        if (!have_m && m.get("norm_mu_eng")) { mu_e = num(m.get("norm_mu_eng"), 0.0); have_m = true; }
    }
    ModelHost out;
    out.species.resize(ntypes);
    // pass 1 (Basis.java:22-50): symbols and the species' own bases; a mirror or
    // share entry only records the species whose basis it takes
    std::vector<int> ref(ntypes, 0);
    std::vector<char> is_mirror(ntypes, 0);
    for (int i = 0; i < ntypes; ++i) {
        const JVal &m = models->arr[i];
        SpeciesHost &sp = out.species[i];
        const JVal *sym = m.get("symbol");
        JNN_CHECK(sym && sym->kind == JVal::STR, ERR_ARG, "No symbol in model");
        sp.symbol = sym->str;
        JNN_CHECK(!m.get("torch"), ERR_ARG, "torch model is invalid now.");
        JVal dflt_basis;
        dflt_basis.kind = JVal::OBJ;
        const JVal *b = m.get("basis");
        if (!b) b = &dflt_basis;
        std::string btype = "spherical_chebyshev";
        if (auto t = b->get("type")) {
            JNN_CHECK(t->kind == JVal::STR, ERR_UNSUPPORTED, "Unsupported basis type: ?");
            const bool mir = t->str == "mirror" || t->str == "mirror_basis";
            if (mir || t->str == "share" || t->str == "shared_basis") {
                // MirrorBasis.java:61-66, SharedBasis.java:61-66: key `mirror` / `share`
                const JVal *k = b->get(mir ? "mirror" : "share");
                JNN_CHECK(k, ERR_ARG, std::string("Key `") + (mir ? "mirror" : "share") + "` required for " +
                                          (mir ? "mirror_basis" : "shared_basis"));
                ref[i] = (int)num(k, 0);
                JNN_CHECK(ref[i] >= 1 && ref[i] <= ntypes, ERR_ARG, "mirror/share type out of range");
                is_mirror[i] = mir;
                continue;
            }
            btype = t->str;
        }
        if (btype == "merge") {  // MergedBasis.load (MergedBasis.java:204-229)
            const JVal *lst = b->get("basis");
            JNN_CHECK(lst, ERR_ARG, "Key `basis` required for merge load");
            JNN_CHECK(lst->kind == JVal::ARR && !lst->arr.empty(), ERR_ARG, "Merge basis can not be null or empty");
            for (const JVal &pb : lst->arr) {
                JNN_CHECK(pb.kind == JVal::OBJ, ERR_ARG, "model JSON: a merge part must be an object");
                std::string pt = "spherical_chebyshev";
                if (auto t2 = pb.get("type")) {
                    JNN_CHECK(t2->kind == JVal::STR, ERR_UNSUPPORTED, "Unsupported basis type: ?");
                    pt = t2->str;
                }
                SpeciesHost part;
                parse_basis(&pb, pt, ntypes, part);
                sp.parts.push_back(std::move(part));
            }
            // the species' single-basis fields hold part 0, rcut the largest (MergedBasis.rcutMax)
            set_basis_fields(sp, sp.parts[0]);
            for (const SpeciesHost &q : sp.parts) sp.rcut = std::max(sp.rcut, q.rcut);
            continue;
        }
        parse_basis(b, btype, ntypes, sp);
    }
    // pass 2 (Basis.java:51-66): mirror / share take the named species' basis
    for (int i = 0; i < ntypes; ++i) {
        if (!ref[i]) continue;
        JNN_CHECK(!ref[ref[i] - 1], ERR_UNSUPPORTED, "a mirror/share basis of a mirror/share basis is not supported");
        SpeciesHost &sp = out.species[i];
        const SpeciesHost &src = out.species[ref[i] - 1];
        set_basis_fields(sp, src);
        sp.parts = src.parts;
        sp.mirror = is_mirror[i] ? ref[i] : 0;
    }
    // pass 3 (NeuralNetwork.java:24-65): the networks.  feed_forward entries
    // first; then a mirror species takes the full network of the species it
    // mirrors (SharedFeedForward.full, its `nn` must be absent) and a
    // shared_feed_forward entry takes another species' network with its own
    // leading rows per layer (SharedFeedForward.java:163-198, 235-353)
    std::vector<char> shared_nn(ntypes, 0);
    for (int i = 0; i < ntypes; ++i) {
        const JVal &m = models->arr[i];
        SpeciesHost &sp = out.species[i];
        const int S = sp.S();
        if (is_mirror[i]) {
            JNN_CHECK(!m.get("nn"), ERR_ARG, "nn data in mirror ModelInfo MUST be empty");
            continue;
        }
        JVal dflt_nn;
        dflt_nn.kind = JVal::OBJ;
        const JVal *nn = m.get("nn");
        if (!nn) nn = &dflt_nn;
        if (auto t = nn->get("type")) {
            JNN_CHECK(t->kind == JVal::STR, ERR_UNSUPPORTED, "Unsupported nn type: ?");
            if (t->str == "shared_feed_forward") {
                shared_nn[i] = 1;
                continue;
            }
            JNN_CHECK(t->str == "feed_forward", ERR_UNSUPPORTED, "Unsupported nn type: " + t->str);
        }
        if (auto d = nn->get("input_dim"))
            JNN_CHECK((int)num(d, 0) == S, ERR_ARG,
                      "Input dimension mismatch: " + std::to_string((int)num(d, 0)) + " vs " + std::to_string(S));
        if (auto h = first(*nn, {"hidden_dims", "nnarch"})) {
            for (double v : vec(h)) sp.hidden.push_back((int)v);
        } else {
            sp.hidden = {32, 32};
        }
        JNN_CHECK(!sp.hidden.empty(), ERR_ARG, "At least one hidden layer is required");
        JNN_CHECK((int)sp.hidden.size() <= kMaxLayers, ERR_UNSUPPORTED, "too many hidden layers");
        std::vector<double> hw, hb;
        const JVal *hws = nn->get("hidden_weights");
        const JVal *hbs = nn->get("hidden_biases");
        if (hws) JNN_CHECK(hws->kind == JVal::ARR && hws->arr.size() == sp.hidden.size(), ERR_ARG, "The number of hidden weights mismatch");
        if (hbs) JNN_CHECK(hbs->kind == JVal::ARR && hbs->arr.size() == sp.hidden.size(), ERR_ARG, "The number of hidden biases mismatch");
        int cols = S;
        for (size_t li = 0; li < sp.hidden.size(); ++li) {
            int h = sp.hidden[li];
            if (hws) {
                const JVal &rows = hws->arr[li];
                JNN_CHECK(rows.kind == JVal::ARR && (int)rows.arr.size() == h, ERR_ARG,
                          "nrows of hidden weight '" + std::to_string(li) + "' mismatch");
                for (auto &r : rows.arr) {
                    std::vector<double> rr = vec(&r);
                    JNN_CHECK((int)rr.size() == cols, ERR_ARG, "ncols of hidden weight '" + std::to_string(li) + "' mismatch");
                    hw.insert(hw.end(), rr.begin(), rr.end());
                }
            } else {
                hw.insert(hw.end(), (size_t)h * cols, 0.0);
            }
            if (hbs) {
                std::vector<double> bb = vec(&hbs->arr[li]);
                JNN_CHECK((int)bb.size() == h, ERR_ARG, "Size of hidden bias '" + std::to_string(li) + "' mismatch");
                hb.insert(hb.end(), bb.begin(), bb.end());
            } else {
                hb.insert(hb.end(), (size_t)h, 0.0);
            }
            cols = h;
        }
        std::vector<double> ow = vec(nn->get("output_weight"));
        if (ow.empty()) ow.assign((size_t)cols, 0.0);
        JNN_CHECK((int)ow.size() == cols, ERR_ARG, "Size of output weight mismatch");
        double ob = num(nn->get("output_bias"), 0.0);
        sp.nn = hw;
        sp.nn.insert(sp.nn.end(), ow.begin(), ow.end());
        sp.nn.insert(sp.nn.end(), hb.begin(), hb.end());
        sp.nn.push_back(ob);
    }
    // the species whose own feed_forward network another one takes
    auto base_of = [&](int i, int b) -> const SpeciesHost & {
        JNN_CHECK(b >= 1 && b <= ntypes && !is_mirror[b - 1] && !shared_nn[b - 1], ERR_ARG,
                  "species " + std::to_string(i + 1) + " takes the network of species " + std::to_string(b) +
                      ", which has no feed_forward network of its own");
        return out.species[b - 1];
    };
    for (int i = 0; i < ntypes; ++i) {
        SpeciesHost &sp = out.species[i];
        if (is_mirror[i]) {
            const SpeciesHost &bs = base_of(i, ref[i]);
            sp.hidden = bs.hidden;
            sp.nn = bs.nn;
            continue;
        }
        if (!shared_nn[i]) continue;
        const JVal &nn = *models->arr[i].get("nn");
        const JVal *sh = nn.get("share");
        JNN_CHECK(sh, ERR_ARG, "Key `share` required for shared_feed_forward");
        const SpeciesHost &bs = base_of(i, (int)num(sh, 0));
        JNN_CHECK(!nn.get("shared_flags"), ERR_ARG, "shared in layer nn is invalid now");
        const int S = sp.S();
        if (auto d = nn.get("input_dim"))
            JNN_CHECK((int)num(d, 0) == S, ERR_ARG,
                      "Input dimension mismatch: " + std::to_string((int)num(d, 0)) + " vs " + std::to_string(S));
        JNN_CHECK(bs.S() == S, ERR_ARG, "Input dimensions mismatch for shared nn");
        const JVal *shd = first(nn, {"shared_hidden_dims", "shared_nnarch"});
        JNN_CHECK(shd, ERR_ARG, "Key `shared_hidden_dims` required for shared_feed_forward");
        std::vector<double> shv = vec(shd);
        const int L = (int)bs.hidden.size();
        if ((int)shv.size() == L) shv.push_back(0.0);  // the output layer defaults to not shared
        JNN_CHECK((int)shv.size() == L + 1, ERR_ARG,
                  "The number of hidden layers mismatch: " + std::to_string((int)shv.size() - 1) + " vs " + std::to_string(L));
        // start from the base network; each layer's first hidden - shared rows are this species' own
        sp.hidden = bs.hidden;
        sp.nn = bs.nn;
        size_t hw_size = 0, hb_size = 0;
        {
            int cols = S;
            for (int h : bs.hidden) { hw_size += (size_t)h * cols; hb_size += h; cols = h; }
        }
        const size_t ow_off = hw_size, hb_off = hw_size + bs.hidden.back(), ob_off = hb_off + hb_size;
        const JVal *hws = nn.get("hidden_weights");
        const JVal *hbs = nn.get("hidden_biases");
        if (hws) JNN_CHECK(hws->kind == JVal::ARR && (int)hws->arr.size() == L, ERR_ARG, "The number of hidden weights mismatch");
        if (hbs) JNN_CHECK(hbs->kind == JVal::ARR && (int)hbs->arr.size() == L, ERR_ARG, "The number of hidden biases mismatch");
        size_t woff = 0, boff = 0;
        int cols = S;
        for (int li = 0; li < L; ++li) {
            const int h = bs.hidden[li], shared = (int)shv[li];
            JNN_CHECK(shared >= 0 && shared <= h, ERR_ARG,
                      "invalid shared dimension: " + std::to_string(shared) + " vs " + std::to_string(h));
            const int own = h - shared;
            if (hws) {
                const JVal &rows = hws->arr[li];
                JNN_CHECK(rows.kind == JVal::ARR && (int)rows.arr.size() == own, ERR_ARG,
                          "nrows of hidden weight '" + std::to_string(li) + "' mismatch");
                for (int r = 0; r < own; ++r) {
                    std::vector<double> rr = vec(&rows.arr[r]);
                    JNN_CHECK((int)rr.size() == cols, ERR_ARG, "ncols of hidden weight '" + std::to_string(li) + "' mismatch");
                    std::copy(rr.begin(), rr.end(), sp.nn.begin() + woff + (size_t)r * cols);
                }
            } else {
                std::fill(sp.nn.begin() + woff, sp.nn.begin() + woff + (size_t)own * cols, 0.0);
            }
            if (hbs) {
                std::vector<double> bb = vec(&hbs->arr[li]);
                JNN_CHECK((int)bb.size() == own, ERR_ARG, "Size of hidden bias '" + std::to_string(li) + "' mismatch");
                std::copy(bb.begin(), bb.end(), sp.nn.begin() + hb_off + boff);
            } else {
                std::fill(sp.nn.begin() + hb_off + boff, sp.nn.begin() + hb_off + boff + own, 0.0);
            }
            woff += (size_t)h * cols;
            boff += h;
            cols = h;
        }
        const int oshared = (int)shv[L];
        JNN_CHECK(oshared == 0 || oshared == 1, ERR_ARG, "invalid shared dimension: " + std::to_string(oshared) + " vs 1");
        if (!oshared) {
            std::vector<double> ow = vec(nn.get("output_weight"));
            if (ow.empty()) ow.assign((size_t)cols, 0.0);
            JNN_CHECK((int)ow.size() == cols, ERR_ARG, "Size of output weight mismatch");
            std::copy(ow.begin(), ow.end(), sp.nn.begin() + ow_off);
            sp.nn[ob_off] = num(nn.get("output_bias"), 0.0);
        }
    }
    // normalisation (NNAP.java:272-306): a mirror species must not carry its
    // own and uses that of the species it mirrors
    for (int i = 0; i < ntypes; ++i) {
        SpeciesHost &sp = out.species[i];
        const int S = sp.S();
        const JVal *mm = &models->arr[i];
        if (is_mirror[i]) {
            JNN_CHECK(!mm->get("ref_eng"), ERR_ARG, "ref_eng in mirror_basis MUST be empty");
            JNN_CHECK(!first(*mm, {"norm_vec", "norm_sigma", "norm_mu"}), ERR_ARG,
                      "norm_vec/norm_sigma/norm_mu in mirror_basis MUST be empty");
            mm = &models->arr[ref[i] - 1];
        }
        const JVal &m = *mm;
        double refe = num(m.get("ref_eng"), 0.0);
        std::vector<double> mu = vec(m.get("norm_mu"));
        std::vector<double> sg = vec(first(m, {"norm_sigma", "norm_vec"}));
        if (mu.empty()) mu.assign((size_t)S, 0.0);
        if (sg.empty()) sg.assign((size_t)S, 1.0);
        JNN_CHECK((int)mu.size() == S && (int)sg.size() == S, ERR_ARG, "norm_mu/norm_sigma size mismatch");
        sp.norm = {mu_e + refe, sig_e};
        sp.norm.insert(sp.norm.end(), mu.begin(), mu.end());
        sp.norm.insert(sp.norm.end(), sg.begin(), sg.end());
    }
    finalize_species(out);
    return out;
}

void finalize_species(ModelHost &m) {
    const int nt = m.ntypes();
    // a neighbour's species travels in 4 bits of its pair tag
    JNN_CHECK(nt >= 1 && nt <= kMaxSpecies, ERR_UNSUPPORTED,
              "at most " + std::to_string(kMaxSpecies) + " species are supported, the model has " + std::to_string(nt));
    // recognise the structured mixings so the kernels can skip the dense product
    auto mix_of = [&](SpeciesHost &b) {
        const int K = b.K(), P = b.P;
        JNN_CHECK((int)b.W.size() == nt * P * K, ERR_ARG, "fp_param size mismatch");
        bool none = (P == K), def = (P == 2 * K);
        for (int t = 1; t <= nt; ++t)
            for (int p = 0; p < P; ++p)
                for (int n = 0; n < K; ++n) {
                    double w = b.W[((size_t)(t - 1) * P + p) * K + n];
                    if (none && w != (p == n ? 1.0 : 0.0)) none = false;
                    double wt = (t & 1) ? (double)t : -(double)t;
                    double want = (p == n) ? 1.0 : (p == K + n ? wt : 0.0);
                    if (def && w != want) def = false;
                }
        b.mix = none ? MIX_NONE : (def ? MIX_DEFAULT : MIX_DENSE);
    };
    for (auto &sp : m.species) {
        size_t need = 0;
        int cols = sp.S();
        for (int h : sp.hidden) { need += (size_t)h * cols + h; cols = h; }
        need += (size_t)cols + 1;
        JNN_CHECK(sp.nn.size() == need, ERR_ARG, "nn_param size mismatch");
        JNN_CHECK((int)sp.norm.size() == 2 + 2 * sp.S(), ERR_ARG, "norm_param size mismatch");
        mix_of(sp);
        for (auto &q : sp.parts) mix_of(q);
    }
}

}  // namespace jnn
