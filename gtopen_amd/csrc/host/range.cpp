// Range text -> 1326 combo weights (crates/solver/src/range.rs:28-303).
// Tokens are separated by ',' or ' ' (newline, tab and ';' act as ','),
// each optionally suffixed ':w'; w > 1 is a percentage; later tokens
// overwrite earlier ones for the combos they cover.
#include <cctype>
#include <cmath>
#include <cstdlib>

#include "host.h"

namespace gto {

namespace {

enum class Suited { Pair, Suited, Offsuit };
typedef std::vector<std::pair<Card, Card>> Combos;

Combos class_combos(int hi, int lo, Suited s) {  // range.rs:144-179
    Combos out;
    if (s == Suited::Pair) {
        if (hi != lo) return out;
        for (int s1 = 0; s1 < 4; ++s1)
            for (int s2 = 0; s2 < s1; ++s2) out.push_back({make_card(hi, s1), make_card(lo, s2)});
    } else if (s == Suited::Suited) {
        if (hi == lo) return out;
        for (int su = 0; su < 4; ++su) out.push_back({make_card(hi, su), make_card(lo, su)});
    } else {
        if (hi == lo) return out;
        for (int s1 = 0; s1 < 4; ++s1)
            for (int s2 = 0; s2 < 4; ++s2)
                if (s1 != s2) out.push_back({make_card(hi, s1), make_card(lo, s2)});
    }
    return out;
}

void append(Combos& dst, const Combos& src) { dst.insert(dst.end(), src.begin(), src.end()); }

std::string rank_pair(int a, int b) {
    static const char r[] = "23456789TJQKA";
    return std::string(1, r[a]) + r[b];
}

Combos expand_class(int hi, int lo, char marker) {  // range.rs:181-202; marker 0 = none
    Combos out;
    if (hi == lo) {
        if (marker) throw Error("pairs cannot have a suitedness marker: " + rank_pair(hi, lo));
        append(out, class_combos(hi, lo, Suited::Pair));
    } else if (marker == 's') {
        append(out, class_combos(hi, lo, Suited::Suited));
    } else if (marker == 'o') {
        append(out, class_combos(hi, lo, Suited::Offsuit));
    } else if (marker == 0) {
        append(out, class_combos(hi, lo, Suited::Suited));
        append(out, class_combos(hi, lo, Suited::Offsuit));
    } else {
        throw Error(std::string("invalid suitedness marker '") + marker + "'");
    }
    return out;
}

// "AKs" / "TT" -> (hi, lo, marker); false if malformed (range.rs:205-222)
bool parse_class(const std::string& s, int* hi, int* lo, char* marker) {
    if (s.size() < 2 || s.size() > 3) return false;
    int r1 = rank_from_char(s[0]), r2 = rank_from_char(s[1]);
    if (r1 < 0 || r2 < 0) return false;
    *marker = 0;
    if (s.size() == 3) {
        char m = (char)std::tolower((unsigned char)s[2]);
        if (m != 's' && m != 'o') return false;
        *marker = m;
    }
    *hi = r1 >= r2 ? r1 : r2;
    *lo = r1 >= r2 ? r2 : r1;
    return true;
}

std::string trim(const std::string& s) {
    size_t a = 0, b = s.size();
    while (a < b && std::isspace((unsigned char)s[a])) ++a;
    while (b > a && std::isspace((unsigned char)s[b - 1])) --b;
    return s.substr(a, b - a);
}

Combos expand_item(const std::string& raw) {  // range.rs:224-303
    std::string item = trim(raw);
    if (item.size() == 4 && rank_from_char(item[0]) >= 0 && suit_from_char(item[1]) >= 0 &&
        rank_from_char(item[2]) >= 0 && suit_from_char(item[3]) >= 0) {
        Card c1 = card_from_str(item.substr(0, 2)), c2 = card_from_str(item.substr(2, 2));
        if (c1 == c2) throw Error("combo with duplicate card: \"" + item + "\"");
        return Combos{{c1, c2}};
    }
    size_t dash = item.find('-');
    if (dash != std::string::npos) {
        int h1, l1, h2, l2;
        char m1, m2;
        if (!parse_class(trim(item.substr(0, dash)), &h1, &l1, &m1))
            throw Error("invalid range start: \"" + item + "\"");
        if (!parse_class(trim(item.substr(dash + 1)), &h2, &l2, &m2))
            throw Error("invalid range end: \"" + item + "\"");
        if (m1 != m2) throw Error("mismatched suitedness in range: \"" + item + "\"");
        Combos out;
        if (h1 == l1 && h2 == l2) {
            int top = h1 >= h2 ? h1 : h2, bot = h1 >= h2 ? h2 : h1;
            for (int r = bot; r <= top; ++r) append(out, expand_class(r, r, m1));
        } else if (h1 == h2) {
            int top = l1 >= l2 ? l1 : l2, bot = l1 >= l2 ? l2 : l1;
            for (int r = bot; r <= top; ++r) {
                if (r == h1) continue;
                append(out, expand_class(h1, r, m1));
            }
        } else if (h1 - l1 == h2 - l2) {
            int top = h1 >= h2 ? h1 : h2, bot = h1 >= h2 ? h2 : h1;
            int gap = h1 - l1;
            for (int h = bot; h <= top; ++h) append(out, expand_class(h, h - gap, m1));
        } else {
            throw Error("unsupported range shape: \"" + item + "\"");
        }
        return out;
    }
    if (!item.empty() && item.back() == '+') {
        int hi, lo;
        char marker;
        if (!parse_class(trim(item.substr(0, item.size() - 1)), &hi, &lo, &marker))
            throw Error("invalid item: \"" + item + "\"");
        Combos out;
        if (hi == lo) {
            for (int r = hi; r < 13; ++r) append(out, expand_class(r, r, marker));
        } else {
            for (int l = lo; l < hi; ++l) append(out, expand_class(hi, l, marker));
        }
        return out;
    }
    int hi, lo;
    char marker;
    if (!parse_class(item, &hi, &lo, &marker)) throw Error("invalid range token: \"" + item + "\"");
    return expand_class(hi, lo, marker);
}

// Strict decimal float parse of the whole token (Rust `str::parse::<f32>`
// accepts an optional sign, digits, '.', exponent, "inf" and "nan").
bool parse_f32(const std::string& s, float* out) {
    if (s.empty()) return false;
    char* end = nullptr;
    float v = std::strtof(s.c_str(), &end);
    if (end != s.c_str() + s.size()) return false;
    // strtof also accepts hex floats and leading spaces, which Rust rejects
    for (char ch : s)
        if (ch == 'x' || ch == 'X' || std::isspace((unsigned char)ch)) return false;
    *out = v;
    return true;
}

}  // namespace

std::vector<float> parse_range(const std::string& text) {
    std::vector<float> w(kNumCombos, 0.0f);
    std::string norm = text;
    for (char& ch : norm)
        if (ch == '\n' || ch == '\t' || ch == ';') ch = ',';
    size_t i = 0;
    while (i <= norm.size()) {
        size_t j = i;
        while (j < norm.size() && norm[j] != ',' && norm[j] != ' ') ++j;
        std::string token = trim(norm.substr(i, j - i));
        i = j + 1;
        if (token.empty()) continue;
        std::string item = token;
        float weight = 1.0f;
        size_t colon = token.rfind(':');
        if (colon != std::string::npos) {
            item = trim(token.substr(0, colon));
            if (!parse_f32(trim(token.substr(colon + 1)), &weight))
                throw Error("invalid weight in token \"" + token + "\"");
            if (weight > 1.0f) weight /= 100.0f;
            if (!(weight >= 0.0f && weight <= 1.0f))
                throw Error("weight out of range in token \"" + token + "\"");
        }
        for (auto& c : expand_item(item)) w[combo_index(c.first, c.second)] = weight;
    }
    return w;
}

}  // namespace gto
