// YAML model files (NNAP(String) reads a path ending in .yaml / .yml with
// IO.yaml2map, NNAP.java:407-409; IO.java:853-855, 2473-2477): the YAML text
// is turned into the JSON text of the same data, which the model parser then
// reads.  The subset covers what a model file needs and what YAML writers
// emit for one: block mappings and sequences (a sequence may sit at its key's
// indentation, items may open with an inline mapping or another item), flow
// collections ([...], {...}, also over several lines, so JSON itself is read),
// plain, single- and double-quoted scalars, comments, and the document markers
// "---" / "..." and directives.  Plain scalars follow the YAML 1.2 core schema:
// null / ~ / empty, true / false, integers (decimal, 0x, 0o), floats; anything
// else is a string.  Anchors, aliases, tags, block scalars (| >), multi-line
// plain or quoted scalars, complex keys and non-finite numbers are refused
// with the line number.
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <set>
#include <string>
#include <vector>

#include "common.h"

namespace jnn {
namespace {

[[noreturn]] void fail(int no, const std::string &msg) {
    throw Error(ERR_ARG, "YAML model line " + std::to_string(no) + ": " + msg);
}

struct Line {
    int indent;
    std::string text;
    int no;
};

// A quote opens a quoted scalar only where a scalar starts.
bool scalar_start(const std::string &s, size_t i) {
    size_t k = i;
    while (k > 0 && s[k - 1] == ' ') --k;
    if (k == 0) return true;
    const char c = s[k - 1];
    return c == '-' || c == ':' || c == ',' || c == '[' || c == '{' || c == '?';
}

// Index just past the quoted scalar opening at s[i] (npos: unterminated).
size_t quote_end(const std::string &s, size_t i) {
    const char q = s[i];
    for (size_t k = i + 1; k < s.size(); ++k) {
        if (q == '"' && s[k] == '\\') {
            ++k;
        } else if (s[k] == q) {
            if (q == '\'' && k + 1 < s.size() && s[k + 1] == '\'') {
                ++k;
                continue;
            }
            return k + 1;
        }
    }
    return std::string::npos;
}

std::string strip_comment(const std::string &s, int no) {
    for (size_t i = 0; i < s.size(); ++i) {
        if ((s[i] == '"' || s[i] == '\'') && scalar_start(s, i)) {
            const size_t e = quote_end(s, i);
            if (e == std::string::npos) fail(no, "unterminated quoted scalar (multi-line quoted scalars are not supported)");
            i = e - 1;
        } else if (s[i] == '#' && (i == 0 || s[i - 1] == ' ')) {
            return s.substr(0, i);
        }
    }
    return s;
}

std::string rtrim(const std::string &s) {
    size_t e = s.find_last_not_of(" \r");
    return e == std::string::npos ? std::string() : s.substr(0, e + 1);
}

void json_escape_char(std::string &out, unsigned char c) {
    switch (c) {
        case '"': out += "\\\""; break;
        case '\\': out += "\\\\"; break;
        case '\n': out += "\\n"; break;
        case '\t': out += "\\t"; break;
        case '\r': out += "\\r"; break;
        case '\b': out += "\\b"; break;
        case '\f': out += "\\f"; break;
        default:
            if (c < 0x20) {
                char b[8];
                std::snprintf(b, sizeof b, "\\u%04x", c);
                out += b;
            } else {
                out += (char)c;
            }
    }
}

std::string json_string(const std::string &s) {
    std::string out = "\"";
    for (unsigned char c : s) json_escape_char(out, c);
    return out + "\"";
}

// JSON text of the quoted scalar s[i, e).
std::string quoted_json(const std::string &s, size_t i, size_t e, int no) {
    std::string out = "\"";
    if (s[i] == '\'') {
        for (size_t k = i + 1; k + 1 < e; ++k) {
            if (s[k] == '\'') ++k;  // '' is one quote
            json_escape_char(out, (unsigned char)s[k]);
        }
        return out + "\"";
    }
    for (size_t k = i + 1; k + 1 < e; ++k) {
        const unsigned char c = (unsigned char)s[k];
        if (c != '\\') {
            json_escape_char(out, c);
            continue;
        }
        const char x = s[++k];
        switch (x) {
            case '"': out += "\\\""; break;
            case '\\': out += "\\\\"; break;
            case '/': out += "/"; break;
            case 'n': out += "\\n"; break;
            case 't': out += "\\t"; break;
            case 'r': out += "\\r"; break;
            case 'b': out += "\\b"; break;
            case 'f': out += "\\f"; break;
            case '0': out += "\\u0000"; break;
            case ' ': out += " "; break;
            case 'x': {
                if (k + 2 >= e) fail(no, "bad \\x escape");
                out += "\\u00" + s.substr(k + 1, 2);
                k += 2;
                break;
            }
            case 'u': {
                if (k + 4 >= e) fail(no, "bad \\u escape");
                out += "\\u" + s.substr(k + 1, 4);
                k += 4;
                break;
            }
            default: fail(no, std::string("unsupported escape \\") + x);
        }
    }
    return out + "\"";
}

bool is_digits(const std::string &s, size_t a, size_t b, bool hex = false) {
    if (a >= b) return false;
    for (size_t k = a; k < b; ++k) {
        const char c = s[k];
        const bool ok = (c >= '0' && c <= '9') || (hex && ((c >= 'a' && c <= 'f') || (c >= 'A' && c <= 'F')));
        if (!ok) return false;
    }
    return true;
}

// JSON text of a plain scalar (YAML 1.2 core schema).
std::string plain_json(const std::string &t, int no) {
    if (t.empty() || t == "~" || t == "null" || t == "Null" || t == "NULL") return "null";
    if (t == "true" || t == "True" || t == "TRUE") return "true";
    if (t == "false" || t == "False" || t == "FALSE") return "false";
    const char c0 = t[0];
    if (c0 == '&' || c0 == '*' || c0 == '!') fail(no, "anchors, aliases and tags are not supported");
    if (c0 == '|' || c0 == '>') fail(no, "block scalars are not supported");
    if (c0 == '@' || c0 == '`') fail(no, "reserved indicator");
    size_t a = (c0 == '+' || c0 == '-') ? 1 : 0;
    const std::string low = t.substr(a);
    if (low == ".inf" || low == ".Inf" || low == ".INF" || t == ".nan" || t == ".NaN" || t == ".NAN")
        fail(no, "non-finite numbers are not supported in a model file");
    // integers: decimal, 0x hex, 0o octal
    if (a == 0 && t.size() > 2 && t[0] == '0' && (t[1] == 'x' || t[1] == 'o')) {
        const bool hex = t[1] == 'x';
        if (is_digits(t, 2, t.size(), hex)) {
            errno = 0;
            const long long v = std::strtoll(t.c_str() + 2, nullptr, hex ? 16 : 8);
            if (errno == ERANGE) fail(no, "integer out of range");
            return std::to_string(v);
        }
    }
    if (is_digits(t, a, t.size())) {
        errno = 0;
        const long long v = std::strtoll(t.c_str(), nullptr, 10);
        if (errno == ERANGE) fail(no, "integer out of range");
        return std::to_string(v);
    }
    // floats: [-+]? ( . digits | digits [ . digits* ] ) ( [eE] [-+]? digits )?
    size_t k = a;
    const size_t ip = k;
    while (k < t.size() && t[k] >= '0' && t[k] <= '9') ++k;
    bool mant = k > ip;
    if (k < t.size() && t[k] == '.') {
        ++k;
        const size_t fp = k;
        while (k < t.size() && t[k] >= '0' && t[k] <= '9') ++k;
        mant = mant || k > fp;
    }
    bool num = mant;
    if (num && k < t.size() && (t[k] == 'e' || t[k] == 'E')) {
        ++k;
        if (k < t.size() && (t[k] == '+' || t[k] == '-')) ++k;
        const size_t ep = k;
        while (k < t.size() && t[k] >= '0' && t[k] <= '9') ++k;
        num = k > ep;
    }
    if (num && k == t.size()) {
        errno = 0;
        const double v = std::strtod(t.c_str(), nullptr);
        if (errno == ERANGE && (v > 1.0 || v < -1.0)) fail(no, "number out of range");
        char b[40];
        std::snprintf(b, sizeof b, "%.17g", v);
        std::string r(b);
        if (r.find_first_of(".eEn") == std::string::npos) r += ".0";  // keep it a float
        return r;
    }
    return json_string(t);
}

class Converter {
  public:
    explicit Converter(const std::string &text) {
        int no = 0;
        size_t pos = 0;
        bool body = false;
        while (pos <= text.size()) {
            size_t e = text.find('\n', pos);
            if (e == std::string::npos) e = text.size();
            std::string raw = text.substr(pos, e - pos);
            pos = e + 1;
            ++no;
            const std::string s = rtrim(strip_comment(raw, no));
            const size_t ind = s.find_first_not_of(' ');
            if (ind == std::string::npos) continue;
            if (s[ind] == '\t') fail(no, "tabs in indentation");
            if (ind == 0 && s[0] == '%' && !body) continue;  // a directive
            if (ind == 0 && (s == "---" || s == "...")) {
                if (s == "---" && body) fail(no, "more than one document");
                continue;
            }
            if (ind == 0 && s.compare(0, 4, "--- ") == 0) fail(no, "content after the document marker");
            body = true;
            L_.push_back({(int)ind, s.substr(ind), no});
        }
    }
    std::string run() {
        if (L_.empty()) fail(0, "empty document");
        node(-1);
        if (p_ < L_.size()) fail(L_[p_].no, "unexpected content (check the indentation)");
        return out_;
    }

  private:
    std::vector<Line> L_;
    size_t p_ = 0;
    std::string out_;
    int depth_ = 0;  // open collections: past 64 the text is refused
    struct Nest {
        int &d;
        Nest(int &depth, int no) : d(depth) {
            if (d >= 64) fail(no, "nesting deeper than 64 levels");
            ++d;
        }
        ~Nest() { --d; }
    };

    static bool seq_item(const std::string &t) { return t == "-" || (t.size() > 1 && t[0] == '-' && t[1] == ' '); }
    // The colon of a block mapping entry "key: value" / "key:", npos if none.
    static size_t map_colon(const std::string &t, int no) {
        if (t.empty() || t[0] == '[' || t[0] == '{') return std::string::npos;
        size_t i = 0;
        if (t[0] == '"' || t[0] == '\'') {
            i = quote_end(t, 0);
            if (i == std::string::npos) fail(no, "unterminated quoted scalar");
            while (i < t.size() && t[i] == ' ') ++i;
            return (i < t.size() && t[i] == ':' && (i + 1 == t.size() || t[i + 1] == ' ')) ? i : std::string::npos;
        }
        if (t[0] == '?') fail(no, "complex keys are not supported");
        for (; i < t.size(); ++i)
            if (t[i] == ':' && (i + 1 == t.size() || t[i + 1] == ' ')) return i;
        return std::string::npos;
    }

    void node(int parent) {
        const Nest nest(depth_, p_ < L_.size() ? L_[p_].no : 0);
        if (p_ >= L_.size() || L_[p_].indent <= parent) {
            out_ += "null";
            return;
        }
        const Line &ln = L_[p_];
        if (seq_item(ln.text)) sequence(ln.indent);
        else if (map_colon(ln.text, ln.no) != std::string::npos) mapping(ln.indent);
        else inline_value(parent);
    }

    void sequence(int ind) {
        out_ += '[';
        bool first = true;
        while (p_ < L_.size() && L_[p_].indent == ind && seq_item(L_[p_].text)) {
            if (!first) out_ += ',';
            first = false;
            Line &ln = L_[p_];
            const std::string rest = ln.text.substr(1);
            const size_t k = rest.find_first_not_of(' ');
            if (k == std::string::npos) {
                ++p_;
                node(ind);
            } else {  // the item starts on this line: read it as a line of its own at its column
                ln.indent = ind + 1 + (int)k;
                ln.text = rest.substr(k);
                node(ind);
            }
        }
        if (p_ < L_.size() && L_[p_].indent > ind) fail(L_[p_].no, "bad indentation in a sequence");
        out_ += ']';
    }

    void mapping(int ind) {
        out_ += '{';
        bool first = true;
        std::set<std::string> keys;
        while (p_ < L_.size() && L_[p_].indent == ind && !seq_item(L_[p_].text)) {
            Line &ln = L_[p_];
            const size_t c = map_colon(ln.text, ln.no);
            if (c == std::string::npos) fail(ln.no, "expected 'key: value'");
            std::string kt = rtrim(ln.text.substr(0, c));
            std::string key = (kt[0] == '"' || kt[0] == '\'') ? quoted_json(kt, 0, kt.size(), ln.no) : json_string(kt);
            if (!kt.empty() && (kt[0] == '&' || kt[0] == '*' || kt[0] == '!'))
                fail(ln.no, "anchors, aliases and tags are not supported");
            if (!keys.insert(key).second) fail(ln.no, "duplicate key " + key);
            if (!first) out_ += ',';
            first = false;
            out_ += key;
            out_ += ':';
            const std::string rest = ln.text.substr(c + 1);
            const size_t k = rest.find_first_not_of(' ');
            if (k == std::string::npos) {
                ++p_;
                // a block sequence may sit at its key's own indentation
                if (p_ < L_.size() && L_[p_].indent == ind && seq_item(L_[p_].text)) sequence(ind);
                else node(ind);
            } else {
                ln.indent = ind + (int)(c + 1 + k);
                ln.text = rest.substr(k);
                if (seq_item(ln.text) || map_colon(ln.text, ln.no) != std::string::npos)
                    fail(ln.no, "a block collection cannot start on its key's line");
                inline_value(ind);
            }
        }
        if (p_ < L_.size() && L_[p_].indent > ind) fail(L_[p_].no, "bad indentation in a mapping");
        out_ += '}';
    }

    // A scalar or a flow collection starting at line p_ (a flow collection may
    // continue over the following lines until its brackets close); `parent` is
    // the indentation of the collection that holds it.
    void inline_value(int parent) {
        const Line &ln = L_[p_];
        std::string s = ln.text;
        const int no = ln.no;
        ++p_;
        if (s[0] == '[' || s[0] == '{') {
            while (!balanced(s, no)) {
                if (p_ >= L_.size()) fail(no, "unterminated flow collection");
                s += ' ';
                s += L_[p_].text;
                ++p_;
            }
            size_t i = 0;
            flow(s, i, no);
            skip_ws(s, i);
            if (i != s.size()) fail(no, "content after a flow collection");
            return;
        }
        if (s[0] == '"' || s[0] == '\'') {
            const size_t e = quote_end(s, 0);
            if (e != s.size()) fail(no, "content after a quoted scalar");
            out_ += quoted_json(s, 0, e, no);
            return;
        }
        if (s[0] == '|' || s[0] == '>') fail(no, "block scalars are not supported");
        if (p_ < L_.size() && L_[p_].indent > parent && !seq_item(L_[p_].text) &&
            map_colon(L_[p_].text, L_[p_].no) == std::string::npos)
            fail(L_[p_].no, "multi-line plain scalars are not supported");
        out_ += plain_json(s, no);
    }

    static bool balanced(const std::string &s, int no) {
        int depth = 0;
        for (size_t i = 0; i < s.size(); ++i) {
            const char c = s[i];
            if ((c == '"' || c == '\'') && scalar_start(s, i)) {
                const size_t e = quote_end(s, i);
                if (e == std::string::npos) fail(no, "unterminated quoted scalar");
                i = e - 1;
            } else if (c == '[' || c == '{') {
                ++depth;
            } else if (c == ']' || c == '}') {
                --depth;
            }
        }
        return depth <= 0;
    }
    static void skip_ws(const std::string &s, size_t &i) {
        while (i < s.size() && s[i] == ' ') ++i;
    }
    // A flow scalar: quoted, or plain up to , ] } (or ": " / a final ':' for a key).
    std::string flow_scalar(const std::string &s, size_t &i, int no, bool key) {
        skip_ws(s, i);
        if (i < s.size() && (s[i] == '"' || s[i] == '\'')) {
            const size_t e = quote_end(s, i);
            if (e == std::string::npos) fail(no, "unterminated quoted scalar");
            std::string r = quoted_json(s, i, e, no);
            i = e;
            return r;
        }
        const size_t a = i;
        while (i < s.size() && s[i] != ',' && s[i] != ']' && s[i] != '}' &&
               !(s[i] == ':' && (i + 1 == s.size() || s[i + 1] == ' ' || s[i + 1] == ',' || s[i + 1] == '}' ||
                                 s[i + 1] == ']'))) {
            if (s[i] == '[' || s[i] == '{') fail(no, "unexpected bracket in a flow scalar");
            ++i;
        }
        const std::string t = rtrim(s.substr(a, i - a));
        return key ? json_string(t) : plain_json(t, no);
    }
    void flow(const std::string &s, size_t &i, int no) {
        const Nest nest(depth_, no);
        skip_ws(s, i);
        if (i >= s.size()) fail(no, "missing value in a flow collection");
        if (s[i] == '[') {
            ++i;
            out_ += '[';
            bool first = true;
            for (;;) {
                skip_ws(s, i);
                if (i >= s.size()) fail(no, "unterminated flow sequence");
                if (s[i] == ']') {
                    ++i;
                    break;
                }
                if (!first) out_ += ',';
                first = false;
                flow(s, i, no);
                skip_ws(s, i);
                if (i < s.size() && s[i] == ',') ++i;
                else if (i < s.size() && s[i] == ']') continue;
                else fail(no, "expected ',' or ']' in a flow sequence");
            }
            out_ += ']';
        } else if (s[i] == '{') {
            ++i;
            out_ += '{';
            bool first = true;
            std::set<std::string> keys;
            for (;;) {
                skip_ws(s, i);
                if (i >= s.size()) fail(no, "unterminated flow mapping");
                if (s[i] == '}') {
                    ++i;
                    break;
                }
                std::string key = flow_scalar(s, i, no, true);
                if (!keys.insert(key).second) fail(no, "duplicate key " + key);
                skip_ws(s, i);
                if (i >= s.size() || s[i] != ':') fail(no, "expected ':' in a flow mapping");
                ++i;
                if (!first) out_ += ',';
                first = false;
                out_ += key;
                out_ += ':';
                skip_ws(s, i);
                if (i < s.size() && (s[i] == ',' || s[i] == '}')) out_ += "null";
                else flow(s, i, no);
                skip_ws(s, i);
                if (i < s.size() && s[i] == ',') ++i;
                else if (i < s.size() && s[i] == '}') continue;
                else fail(no, "expected ',' or '}' in a flow mapping");
            }
            out_ += '}';
        } else {
            out_ += flow_scalar(s, i, no, false);
        }
    }
};

}  // namespace

std::string yaml_to_json(const std::string &yaml) { return Converter(yaml).run(); }

}  // namespace jnn
