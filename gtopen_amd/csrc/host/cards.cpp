// Cards, combos, canonical flops and the 7-card evaluator.
// Card = rank*4 + suit, ranks 0=deuce..12=ace, suits c,d,h,s
// (crates/solver/src/cards.rs:1-6). Evaluator value = cat<<20 | five
// 4-bit tiebreak ranks, higher is stronger (crates/solver/src/evaluator.rs:1-7).
#include <algorithm>
#include <cctype>

#include "host.h"

namespace gto {

static const char kRanks[] = "23456789TJQKA";
static const char kSuits[] = "cdhs";

void combo_from_index(uint32_t idx, Card* hi, Card* lo) {
    // hi is the largest h with h*(h-1)/2 <= idx (cards.rs:106-114)
    uint32_t h = 1;
    while ((h + 1) * h / 2 <= idx) ++h;
    *hi = (Card)h;
    *lo = (Card)(idx - h * (h - 1) / 2);
}

int rank_from_char(char ch) {
    ch = (char)std::toupper((unsigned char)ch);
    for (int i = 0; i < 13; ++i)
        if (kRanks[i] == ch) return i;
    return -1;
}

int suit_from_char(char ch) {
    ch = (char)std::tolower((unsigned char)ch);
    for (int i = 0; i < 4; ++i)
        if (kSuits[i] == ch) return i;
    return -1;
}

static std::string trim(const std::string& s) {
    size_t a = 0, b = s.size();
    while (a < b && std::isspace((unsigned char)s[a])) ++a;
    while (b > a && std::isspace((unsigned char)s[b - 1])) --b;
    return s.substr(a, b - a);
}

Card card_from_str(const std::string& raw) {
    std::string s = trim(raw);
    if (s.size() != 2) throw Error("invalid card: \"" + raw + "\"");
    int r = rank_from_char(s[0]);
    if (r < 0) throw Error("invalid rank in card: \"" + raw + "\"");
    int su = suit_from_char(s[1]);
    if (su < 0) throw Error("invalid suit in card: \"" + raw + "\"");
    return make_card((uint8_t)r, (uint8_t)su);
}

std::string card_to_string(Card c) {
    return std::string(1, kRanks[rank_of(c)]) + kSuits[suit_of(c)];
}

std::vector<Card> parse_cards(const std::string& s) {
    std::string cleaned;
    for (char ch : s)
        if (!std::isspace((unsigned char)ch) && ch != ',') cleaned.push_back(ch);
    if (cleaned.size() % 2 != 0) throw Error("invalid card list: \"" + s + "\"");
    std::vector<Card> out;
    uint64_t seen = 0;
    for (size_t i = 0; i < cleaned.size(); i += 2) {
        Card c = card_from_str(cleaned.substr(i, 2));
        if (seen & card_mask(c)) throw Error("duplicate card in list: \"" + s + "\"");
        seen |= card_mask(c);
        out.push_back(c);
    }
    return out;
}

// Suit-isomorphism classes of the 22,100 raw flops: the canonical member is
// the lexicographically smallest descending-sorted triple over the 24 suit
// permutations; classes are returned high-to-low (cards.rs:261-311).
std::vector<FlopClass> canonical_flops() {
    std::vector<std::array<uint8_t, 4>> perms;
    for (uint8_t a = 0; a < 4; ++a)
        for (uint8_t b = 0; b < 4; ++b) {
            if (b == a) continue;
            for (uint8_t c = 0; c < 4; ++c) {
                if (c == a || c == b) continue;
                perms.push_back({a, b, c, (uint8_t)(6 - a - b - c)});
            }
        }
    std::map<std::array<uint8_t, 3>, uint32_t> classes;
    for (int x = 0; x < 52; ++x)
        for (int y = 0; y < x; ++y)
            for (int z = 0; z < y; ++z) {
                std::array<uint8_t, 3> best = {255, 255, 255};
                for (auto& p : perms) {
                    std::array<uint8_t, 3> t = {make_card(rank_of(x), p[suit_of(x)]),
                                                make_card(rank_of(y), p[suit_of(y)]),
                                                make_card(rank_of(z), p[suit_of(z)])};
                    std::sort(t.begin(), t.end(), [](uint8_t u, uint8_t v) { return u > v; });
                    if (t < best) best = t;
                }
                classes[best] += 1;
            }
    std::vector<FlopClass> out;
    for (auto it = classes.rbegin(); it != classes.rend(); ++it) {
        FlopClass f;
        for (int i = 0; i < 3; ++i) f.cards[i] = it->first[i];
        f.weight = it->second;
        out.push_back(f);
    }
    return out;
}

// Weighted systematic subset over n equal-weight strata; a class emitted
// once with its stratum multiplicity (cards.rs:236-259).
std::vector<FlopClass> canonical_flops_subset(size_t n) {
    std::vector<FlopClass> all = canonical_flops();
    if (n == 0 || n >= all.size()) return all;
    double total = 0.0;
    for (auto& f : all) total += (double)f.weight;
    std::vector<FlopClass> out;
    double cum = 0.0;
    size_t ti = 0;
    for (auto& f : all) {
        cum += (double)f.weight;
        uint32_t mult = 0;
        while (ti < n && (((double)ti + 0.5) / (double)n) * total <= cum) {
            ++mult;
            ++ti;
        }
        if (mult > 0) {
            FlopClass g = f;
            g.weight = mult;
            out.push_back(g);
        }
    }
    return out;
}

// ---- evaluator -----------------------------------------------------------

enum { CAT_HIGH = 0, CAT_PAIR, CAT_TWO_PAIR, CAT_TRIPS, CAT_STRAIGHT, CAT_FLUSH, CAT_FULL_HOUSE,
       CAT_QUADS, CAT_STRAIGHT_FLUSH };

static inline uint32_t pack(uint32_t cat, const uint8_t* r, int n) {
    uint32_t v = cat;
    for (int slot = 0; slot < 5; ++slot) v = (v << 4) | (slot < n ? r[slot] : 0u);
    return v;
}

// Top rank of the best 5-long run in a 13-bit rank mask; the ace also
// counts below the deuce (wheel -> 3). Returns -1 if none.
static inline int best_straight(uint32_t mask13) {
    uint32_t m = (mask13 << 1) | ((mask13 >> 12) & 1u);  // bit i+1 = rank i, bit 0 = ace-low
    uint32_t run = m & (m >> 1) & (m >> 2) & (m >> 3) & (m >> 4);  // bit j: ranks j-1..j+3 (shifted)
    if (!run) return -1;
    int top_bit = 31 - __builtin_clz(run);  // lowest card of the run sits at bit top_bit
    return top_bit + 3;                      // top rank = (top_bit - 1) + 4
}

static inline int take_top(uint32_t mask13, int n, uint8_t* out) {
    int k = 0;
    for (int r = 12; r >= 0 && k < n; --r)
        if (mask13 & (1u << r)) out[k++] = (uint8_t)r;
    return k;
}

uint32_t evaluate7(const Card* cards) {
    uint32_t suit_mask[4] = {0, 0, 0, 0};
    uint8_t count[13] = {0};
    uint32_t ranks = 0;
    for (int k = 0; k < 7; ++k) {
        int r = cards[k] >> 2;
        suit_mask[cards[k] & 3] |= 1u << r;
        count[r]++;
        ranks |= 1u << r;
    }
    int flush = -1;
    for (int s = 0; s < 4; ++s)
        if (__builtin_popcount(suit_mask[s]) >= 5) {
            flush = s;
            break;
        }
    if (flush >= 0) {
        int top = best_straight(suit_mask[flush]);
        if (top >= 0) {
            uint8_t r[1] = {(uint8_t)top};
            return pack(CAT_STRAIGHT_FLUSH, r, 1);
        }
    }
    int quad = -1, ntrip = 0, npair = 0;
    uint8_t trips[2], pairs[3];
    for (int r = 12; r >= 0; --r) {
        if (count[r] == 4) quad = r;
        else if (count[r] == 3) {
            if (ntrip < 2) trips[ntrip++] = (uint8_t)r;
        } else if (count[r] == 2) {
            if (npair < 3) pairs[npair++] = (uint8_t)r;
        }
    }
    if (quad >= 0) {
        uint8_t r[2] = {(uint8_t)quad, 0};
        take_top(ranks & ~(1u << quad), 1, r + 1);
        return pack(CAT_QUADS, r, 2);
    }
    if (ntrip >= 1 && (ntrip >= 2 || npair >= 1)) {
        uint8_t r[2] = {trips[0], ntrip >= 2 ? trips[1] : pairs[0]};
        return pack(CAT_FULL_HOUSE, r, 2);
    }
    if (flush >= 0) {
        uint8_t r[5];
        take_top(suit_mask[flush], 5, r);
        return pack(CAT_FLUSH, r, 5);
    }
    {
        int top = best_straight(ranks);
        if (top >= 0) {
            uint8_t r[1] = {(uint8_t)top};
            return pack(CAT_STRAIGHT, r, 1);
        }
    }
    if (ntrip >= 1) {
        uint8_t r[3] = {trips[0], 0, 0};
        take_top(ranks & ~(1u << trips[0]), 2, r + 1);
        return pack(CAT_TRIPS, r, 3);
    }
    if (npair >= 2) {
        uint8_t r[3] = {pairs[0], pairs[1], 0};
        take_top(ranks & ~(1u << pairs[0]) & ~(1u << pairs[1]), 1, r + 2);
        return pack(CAT_TWO_PAIR, r, 3);
    }
    if (npair == 1) {
        uint8_t r[4] = {pairs[0], 0, 0, 0};
        take_top(ranks & ~(1u << pairs[0]), 3, r + 1);
        return pack(CAT_PAIR, r, 4);
    }
    uint8_t r[5];
    take_top(ranks, 5, r);
    return pack(CAT_HIGH, r, 5);
}

}  // namespace gto
