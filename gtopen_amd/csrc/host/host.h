// Host-side spot model: cards, 7-card evaluator, range parser, betting-tree
// builder and the Spot (hands, same-combo cross index, river strength
// tables). Restates the reference's Rust host code in C++ so the built tree
// is node-for-node identical (same node order, data offsets and payoffs):
// crates/solver/src/{cards,evaluator,range,tree,game}.rs.
#pragma once

#include <array>
#include <cstdint>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace gto {

using Card = uint8_t;
constexpr int kNumCards = 52;
constexpr int kNumCombos = 1326;
constexpr uint32_t kSentinel = 0xFFFFFFFFu;

struct Error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

// ---- cards.rs ------------------------------------------------------------
inline uint8_t rank_of(Card c) { return c >> 2; }
inline uint8_t suit_of(Card c) { return c & 3; }
inline Card make_card(uint8_t r, uint8_t s) { return (Card)((r << 2) | s); }
inline uint64_t card_mask(Card c) { return 1ull << c; }
// cards.rs:100-103
inline uint32_t combo_index(Card a, Card b) {
    uint32_t hi = a > b ? a : b, lo = a > b ? b : a;
    return hi * (hi - 1) / 2 + lo;
}
void combo_from_index(uint32_t idx, Card* hi, Card* lo);  // cards.rs:106-114
int rank_from_char(char ch);
int suit_from_char(char ch);
Card card_from_str(const std::string& s);                 // cards.rs:46-55
std::vector<Card> parse_cards(const std::string& s);      // cards.rs:66-85
std::string card_to_string(Card c);
struct FlopClass {
    Card cards[3];
    uint32_t weight;
};
std::vector<FlopClass> canonical_flops();                  // cards.rs:261-311
std::vector<FlopClass> canonical_flops_subset(size_t n);   // cards.rs:236-259

// ---- evaluator.rs --------------------------------------------------------
uint32_t evaluate7(const Card* cards7);                    // evaluator.rs:68-175

// ---- range.rs ------------------------------------------------------------
std::vector<float> parse_range(const std::string& text);   // range.rs:28-58

// ---- tree.rs -------------------------------------------------------------
enum class SizeKind { PotPct, PrevMult, AllIn };
struct BetSize {
    SizeKind kind;
    double value;
};
struct StreetSizing {
    std::vector<BetSize> bet, raise, donk;
};
struct TreeConfig {  // tree.rs:70-102
    double starting_pot = 60.0;
    double effective_stack = 970.0;
    double rake_pct = 0.0;
    double rake_cap = 0.0;
    StreetSizing oop[3], ip[3];
    double allin_threshold = 0.85;
    bool add_allin = false;
    uint8_t max_raises = 10;
};
enum NodeKind : uint8_t { KIND_ACTION = 0, KIND_CHANCE = 1, KIND_FOLD = 2, KIND_SHOWDOWN = 3 };
enum class ActionKind : uint8_t { Fold, Check, Call, Bet, Raise };
struct Action {
    ActionKind kind;
    double amount;
};
struct Tree {  // tree.rs:201-211, nodes as parallel arrays (tree.rs:178-198)
    uint8_t root_street = 0;
    std::vector<uint8_t> kind, player, street, num_children;
    std::vector<uint32_t> children_start, actions_start;
    std::vector<uint64_t> data_offset;
    std::vector<double> put;  // 2 per node
    std::vector<double> t_win, t_lose, t_tie;
    std::vector<uint32_t> children;
    std::vector<Action> actions;
    uint64_t data_size[2] = {0, 0};
    size_t num_nodes() const { return kind.size(); }
};
enum class Strictness { Strict, LenientLoad };
Tree build_tree(const TreeConfig& cfg, const std::vector<Card>& board, const size_t num_hands[2],
                uint64_t max_nodes, Strictness strictness);  // tree.rs:287-402

// ---- game.rs -------------------------------------------------------------
struct SpotConfig {
    std::string board, range_oop, range_ip;
    TreeConfig tree;
};
SpotConfig parse_spot_config(const std::string& json);  // serde layout of game.rs:21-27

struct Spot {
    SpotConfig config;
    std::vector<Card> board;
    uint64_t board_mask = 0;
    Tree tree;
    std::vector<Card> hand_c1[2], hand_c2[2];  // c1 = higher card
    std::vector<uint64_t> hand_mask[2];
    std::vector<float> weights[2];
    std::vector<uint32_t> same_combo[2];
    uint32_t num_river_keys = 0;
    std::vector<int64_t> river_off[2];
    std::vector<uint32_t> river_cnt[2];
    std::vector<uint32_t> river_str[2];
    std::vector<uint16_t> river_idx[2];
    std::vector<std::array<uint8_t, 4>> suit_perms;
    std::vector<uint8_t> suit_perms_flat;  // 4 per symmetry
    // [k * nh + i]: index of hand i's image under suit_perms[k] (game.rs:294-306)
    std::vector<uint16_t> hand_perm[2];
    size_t num_hands(int p) const { return hand_c1[p].size(); }
};
std::unique_ptr<Spot> build_spot(const SpotConfig& cfg, uint64_t max_nodes, Strictness strictness);

// ---- save.rs -------------------------------------------------------------
struct SaveHeader {  // save.rs:17-28
    SpotConfig config;
    uint32_t iteration = 0;
    std::vector<std::pair<uint32_t, std::vector<float>>> locks;
    std::vector<std::pair<uint32_t, std::string>> lock_labels;
};
std::string spot_config_json(const SpotConfig& c);  // serde layout, every field
std::string save_header_json(const SaveHeader& h);
SaveHeader parse_save_header(const std::string& line);
// save.rs:42-70; throws Error with the reference's messages
void validate_locks(const Spot& spot, const std::vector<std::pair<uint32_t, std::vector<float>>>& locks);

}  // namespace gto
