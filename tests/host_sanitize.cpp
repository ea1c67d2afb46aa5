// Host-side checks of the model parsers under AddressSanitizer and
// UndefinedBehaviorSanitizer (built with g++ by tests/test_host_sanitizers.py):
//   host_sanitize NMUT SEED FILE...
// parses every file (YAML when its name ends in .yaml / .yml, else JSON),
// which must succeed, then NMUT seeded byte-level mutations of each, which
// must either parse or be refused with jnn::Error; and builds the L3 / L4
// coupling terms of every (l3max, l4max).  A crash or a sanitizer report
// fails the run.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <string>

#include "../jse_amd/csrc/common.h"
#include "../jse_amd/csrc/sh_coupling.h"

namespace {

uint64_t g_state = 1;
uint64_t next_u64() {  // xorshift64*
    g_state ^= g_state >> 12;
    g_state ^= g_state << 25;
    g_state ^= g_state >> 27;
    return g_state * 0x2545F4914F6CDD1DULL;
}
size_t pick(size_t n) { return n ? (size_t)(next_u64() % n) : 0; }

bool ends_with(const std::string &s, const char *suf) {
    const std::string t(suf);
    return s.size() >= t.size() && s.compare(s.size() - t.size(), t.size(), t) == 0;
}

// 0 parsed, 1 refused with a library error
int parse(const std::string &text, bool yaml) {
    try {
        const jnn::ModelHost h = jnn::parse_model_json(yaml ? jnn::yaml_to_json(text) : text);
        return h.ntypes() > 0 ? 0 : 1;
    } catch (const jnn::Error &) {
        return 1;
    }
}

std::string mutate(const std::string &src) {
    static const char kChars[] = ":-[]{},'\"#\n \t0123456789.eE+-~&*!|>abcxyz";
    std::string s = src;
    const int edits = 1 + (int)pick(4);
    for (int e = 0; e < edits && !s.empty(); ++e) {
        const size_t at = pick(s.size());
        switch (pick(5)) {
            case 0: s.erase(at, 1 + pick(8)); break;                                      // delete a run
            case 1: s.insert(at, 1, kChars[pick(sizeof kChars - 1)]); break;              // insert a character
            case 2: s[at] = kChars[pick(sizeof kChars - 1)]; break;                       // replace a character
            case 3: {                                                                     // duplicate a line
                const size_t b = s.rfind('\n', at), f = s.find('\n', at);
                const size_t lb = b == std::string::npos ? 0 : b + 1;
                const size_t lf = f == std::string::npos ? s.size() : f + 1;
                s.insert(lf, s.substr(lb, lf - lb));
                break;
            }
            default: s.resize(at); break;                                                 // truncate
        }
    }
    return s;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc < 4) {
        std::fprintf(stderr, "usage: host_sanitize NMUT SEED FILE...\n");
        return 2;
    }
    const int nmut = std::atoi(argv[1]);
    g_state = (uint64_t)std::strtoull(argv[2], nullptr, 10) * 2654435761ULL + 1;
    long parsed = 0, refused = 0;
    for (int a = 3; a < argc; ++a) {
        const std::string path = argv[a];
        std::ifstream f(path);
        std::stringstream ss;
        ss << f.rdbuf();
        const std::string text = ss.str();
        const bool yaml = ends_with(path, ".yaml") || ends_with(path, ".yml");
        if (parse(text, yaml) != 0) {
            std::fprintf(stderr, "%s: the unmutated file does not parse\n", path.c_str());
            return 3;
        }
        for (int k = 0; k < nmut; ++k) (parse(mutate(text), yaml) == 0 ? parsed : refused)++;
    }
    long terms = 0;
    for (int l3 = 0; l3 <= 6; ++l3)
        for (int l4 = 0; l4 <= 3; ++l4) {
            const int lm = l3 > l4 ? l3 : l4;
            terms += (long)jnn::coupling_terms(l3, l4, (lm + 1) * (lm + 1)).coef.size();
        }
    std::printf("mutations parsed %ld refused %ld; coupling terms %ld\n", parsed, refused, terms);
    return 0;
}
