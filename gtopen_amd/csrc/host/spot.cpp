// Spot construction (crates/solver/src/game.rs:135-331, 384-462) and the
// gto_spot_* C ABI. Hands are listed per player in combo_index order with
// zero-weight and board-conflicting combos dropped; river strength tables
// hold (evaluate7, hand index) sorted ascending for every reachable river
// board.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <unordered_map>

#include "abi.h"

namespace gto {

static void build_river_tables(Spot& sp) {  // game.rs:384-462
    const int root = sp.tree.root_street;
    std::vector<std::vector<Card>> extras;  // per key (empty vector when no entry)
    std::vector<uint32_t> keys;
    if (root == 2) {
        sp.num_river_keys = 1;
        keys.push_back(0);
        extras.push_back({});
    } else if (root == 1) {
        sp.num_river_keys = 52;
        for (int c = 0; c < 52; ++c)
            if (!(sp.board_mask & card_mask(c))) {
                keys.push_back(c);
                extras.push_back({(Card)c});
            }
    } else {
        sp.num_river_keys = 52 * 52;
        for (int a = 0; a < 52; ++a) {
            if (sp.board_mask & card_mask(a)) continue;
            for (int b = a + 1; b < 52; ++b) {
                if (sp.board_mask & card_mask(b)) continue;
                keys.push_back(a * 52 + b);
                extras.push_back({(Card)a, (Card)b});
            }
        }
    }
    const size_t nk = keys.size();
    // evaluate in parallel over keys, then lay out in key order
    std::vector<std::vector<std::pair<uint32_t, uint16_t>>> sorted[2];
    sorted[0].resize(nk);
    sorted[1].resize(nk);
    auto work = [&](size_t k0, size_t k1) {
        for (size_t k = k0; k < k1; ++k) {
            Card c7[7];
            for (size_t i = 0; i < sp.board.size(); ++i) c7[i] = sp.board[i];
            uint64_t xm = 0;
            for (size_t i = 0; i < extras[k].size(); ++i) {
                c7[sp.board.size() + i] = extras[k][i];
                xm |= card_mask(extras[k][i]);
            }
            for (int p = 0; p < 2; ++p) {
                auto& v = sorted[p][k];
                v.reserve(sp.num_hands(p));
                for (size_t i = 0; i < sp.num_hands(p); ++i) {
                    if (sp.hand_mask[p][i] & xm) continue;
                    c7[5] = sp.hand_c1[p][i];
                    c7[6] = sp.hand_c2[p][i];
                    v.push_back({evaluate7(c7), (uint16_t)i});
                }
                std::sort(v.begin(), v.end());
            }
        }
    };
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (nk < 64) nt = 1;
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; ++t) th.emplace_back(work, nk * t / nt, nk * (t + 1) / nt);
    for (auto& t : th) t.join();
    for (int p = 0; p < 2; ++p) {
        sp.river_off[p].assign(sp.num_river_keys, -1);
        sp.river_cnt[p].assign(sp.num_river_keys, 0);
        for (size_t k = 0; k < nk; ++k) {
            sp.river_off[p][keys[k]] = (int64_t)sp.river_str[p].size();
            sp.river_cnt[p][keys[k]] = (uint32_t)sorted[p][k].size();
            for (auto& e : sorted[p][k]) {
                sp.river_str[p].push_back(e.first);
                sp.river_idx[p].push_back(e.second);
            }
        }
    }
}

static Card permute(Card c, const std::array<uint8_t, 4>& pm) { return make_card(rank_of(c), pm[suit_of(c)]); }

std::unique_ptr<Spot> build_spot(const SpotConfig& cfg, uint64_t max_nodes, Strictness strictness) {
    const TreeConfig& tc = cfg.tree;
    // rake validation (game.rs:141-160)
    if (tc.rake_pct >= 1.0)
        throw Error("rake_pct is a FRACTION of the pot (0.05 = 5%), got " + std::to_string(tc.rake_pct) +
                    " — did you pass a percent?");
    if (tc.rake_pct < 0.0) throw Error("rake_pct must be >= 0, got " + std::to_string(tc.rake_pct));
    if (tc.rake_cap < 0.0) throw Error("rake_cap must be >= 0, got " + std::to_string(tc.rake_cap));
    auto sp = std::make_unique<Spot>();
    sp->config = cfg;
    sp->board = parse_cards(cfg.board);
    if (sp->board.size() < 3 || sp->board.size() > 5) throw Error("board must have 3 to 5 cards");
    for (Card c : sp->board) sp->board_mask |= card_mask(c);

    std::vector<float> ranges[2] = {parse_range(cfg.range_oop), parse_range(cfg.range_ip)};
    for (int p = 0; p < 2; ++p) {
        for (uint32_t idx = 0; idx < (uint32_t)kNumCombos; ++idx) {
            float w = ranges[p][idx];
            if (w <= 0.0f) continue;
            Card hi, lo;
            combo_from_index(idx, &hi, &lo);
            uint64_t m = card_mask(hi) | card_mask(lo);
            if (m & sp->board_mask) continue;
            sp->hand_c1[p].push_back(hi);
            sp->hand_c2[p].push_back(lo);
            sp->hand_mask[p].push_back(m);
            sp->weights[p].push_back(w);
        }
        if (sp->hand_c1[p].empty())
            throw Error(std::string(p == 0 ? "OOP" : "IP") + " range has no combos left after board card removal");
        if (sp->hand_c1[p].size() > 65535) throw Error("range too large");
    }
    // mutual impossibility (game.rs:211-226)
    double pair_mass = 0.0;
    for (size_t i = 0; i < sp->num_hands(0); ++i)
        for (size_t j = 0; j < sp->num_hands(1); ++j)
            if (!(sp->hand_mask[0][i] & sp->hand_mask[1][j]))
                pair_mass += (double)sp->weights[0][i] * (double)sp->weights[1][j];
    if (pair_mass <= 0.0)
        throw Error(
            "OOP and IP ranges are mutually impossible: every OOP combo shares a card with every IP combo "
            "after board removal, so no deal exists (zero compatible pair mass)");
    // same_combo cross index (game.rs:229-244)
    for (int p = 0; p < 2; ++p) {
        std::unordered_map<uint32_t, uint32_t> opp;
        for (size_t j = 0; j < sp->num_hands(1 - p); ++j)
            opp[combo_index(sp->hand_c1[1 - p][j], sp->hand_c2[1 - p][j])] = (uint32_t)j;
        for (size_t i = 0; i < sp->num_hands(p); ++i) {
            auto it = opp.find(combo_index(sp->hand_c1[p][i], sp->hand_c2[p][i]));
            sp->same_combo[p].push_back(it == opp.end() ? kSentinel : it->second);
        }
    }
    size_t nh[2] = {sp->num_hands(0), sp->num_hands(1)};
    sp->tree = build_tree(tc, sp->board, nh, max_nodes, strictness);
    build_river_tables(*sp);

    // suit symmetry group: permutations fixing the board pointwise with both
    // ranges invariant (game.rs:262-317); identity first
    for (uint8_t a = 0; a < 4; ++a)
        for (uint8_t b = 0; b < 4; ++b)
            for (uint8_t c = 0; c < 4; ++c)
                for (uint8_t d = 0; d < 4; ++d) {
                    std::array<uint8_t, 4> pm = {a, b, c, d};
                    bool seen[4] = {false, false, false, false}, ok = true;
                    for (uint8_t s : pm) {
                        if (seen[s]) ok = false;
                        seen[s] = true;
                    }
                    if (!ok) continue;
                    for (Card bc : sp->board)
                        if (permute(bc, pm) != bc) ok = false;
                    for (int p = 0; p < 2 && ok; ++p) {
                        for (size_t i = 0; i < nh[p] && ok; ++i) {
                            uint32_t pc = combo_index(permute(sp->hand_c1[p][i], pm), permute(sp->hand_c2[p][i], pm));
                            Card hi, lo;
                            combo_from_index(pc, &hi, &lo);
                            float w = ranges[p][pc];
                            uint64_t m = card_mask(hi) | card_mask(lo);
                            if (w <= 0.0f || (m & sp->board_mask) || std::fabs(w - sp->weights[p][i]) >= 1e-6f)
                                ok = false;
                        }
                    }
                    if (ok) sp->suit_perms.push_back(pm);
                }
    for (size_t k = 0; k < sp->suit_perms.size(); ++k)
        if (sp->suit_perms[k] == std::array<uint8_t, 4>{0, 1, 2, 3}) std::swap(sp->suit_perms[0], sp->suit_perms[k]);
    // hand tables of each symmetry (every image is in the range: checked above)
    for (int p = 0; p < 2; ++p) {
        std::vector<int32_t> idx_of(1326, -1);
        for (size_t i = 0; i < nh[p]; ++i) idx_of[combo_index(sp->hand_c1[p][i], sp->hand_c2[p][i])] = (int32_t)i;
        for (const auto& pm : sp->suit_perms)
            for (size_t i = 0; i < nh[p]; ++i)
                sp->hand_perm[p].push_back(
                    (uint16_t)idx_of[combo_index(permute(sp->hand_c1[p][i], pm), permute(sp->hand_c2[p][i], pm))]);
    }
    for (const auto& pm : sp->suit_perms) sp->suit_perms_flat.insert(sp->suit_perms_flat.end(), pm.begin(), pm.end());
    return sp;
}

}  // namespace gto

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

namespace gto {
thread_local std::string g_last_error;
void set_last_error(const std::string& msg) { g_last_error = msg; }
}  // namespace gto


extern "C" {

const char* gto_last_error(void) { return gto::g_last_error.c_str(); }

int gto_spot_build(const char* config_json, int strictness, uint64_t max_nodes, gto_spot** out) {
    if (!config_json || !out) {
        gto::set_last_error("gto_spot_build: null argument");
        return GTO_ERR_INVALID;
    }
    *out = nullptr;
    try {
        gto::SpotConfig cfg = gto::parse_spot_config(config_json);
        auto h = new gto_spot();
        h->spot = gto::build_spot(cfg, max_nodes,
                                  strictness == GTO_LENIENT_LOAD ? gto::Strictness::LenientLoad : gto::Strictness::Strict);
        const gto::Spot& s = *h->spot;
        gto_spot_view& v = h->view;
        std::memset(&v, 0, sizeof(v));
        v.num_nodes = (uint32_t)s.tree.num_nodes();
        v.node_kind = s.tree.kind.data();
        v.node_player = s.tree.player.data();
        v.node_street = s.tree.street.data();
        v.node_num_children = s.tree.num_children.data();
        v.node_children_start = s.tree.children_start.data();
        v.node_data_offset = s.tree.data_offset.data();
        v.node_t_win = s.tree.t_win.data();
        v.node_t_lose = s.tree.t_lose.data();
        v.node_t_tie = s.tree.t_tie.data();
        v.node_put = s.tree.put.data();
        v.num_children_total = (uint32_t)s.tree.children.size();
        v.children = s.tree.children.data();
        v.data_size[0] = s.tree.data_size[0];
        v.data_size[1] = s.tree.data_size[1];
        v.root_street = s.tree.root_street;
        v.starting_pot = s.config.tree.starting_pot;
        v.effective_stack = s.config.tree.effective_stack;
        v.rake_pct = s.config.tree.rake_pct;
        v.rake_cap = s.config.tree.rake_cap;
        v.board_len = (uint32_t)s.board.size();
        for (size_t i = 0; i < s.board.size(); ++i) v.board[i] = s.board[i];
        v.board_mask = s.board_mask;
        for (int p = 0; p < 2; ++p) {
            v.num_hands[p] = (uint32_t)s.num_hands(p);
            v.hand_c1[p] = s.hand_c1[p].data();
            v.hand_c2[p] = s.hand_c2[p].data();
            v.weights[p] = s.weights[p].data();
            v.same_combo[p] = s.same_combo[p].data();
            v.river_off[p] = s.river_off[p].data();
            v.river_cnt[p] = s.river_cnt[p].data();
            v.river_str[p] = s.river_str[p].data();
            v.river_idx[p] = s.river_idx[p].data();
        }
        v.num_river_keys = s.num_river_keys;
        v.num_suit_perms = (uint32_t)s.suit_perms.size();
        v.suit_perms = s.suit_perms_flat.data();
        for (int p = 0; p < 2; ++p) v.hand_perm[p] = s.hand_perm[p].data();
        *out = h;
        return GTO_OK;
    } catch (const std::bad_alloc&) {
        gto::set_last_error("out of host memory while building the spot");
        return GTO_ERR_MEMORY;
    } catch (const std::exception& e) {
        gto::set_last_error(e.what());
        return GTO_ERR_INVALID;
    }
}

void gto_spot_free(gto_spot* spot) { delete spot; }

const gto_spot_view* gto_spot_view_of(const gto_spot* spot) { return spot ? &spot->view : nullptr; }

uint64_t gto_spot_num_action_nodes(const gto_spot* spot) {
    uint64_t n = 0;
    for (uint8_t k : spot->spot->tree.kind) n += k == gto::KIND_ACTION;
    return n;
}

uint64_t gto_spot_arena_bytes(const gto_spot* spot, int storage) {
    const gto::Tree& t = spot->spot->tree;
    uint64_t entries = t.data_size[0] + t.data_size[1];
    if (storage == GTO_STORAGE_COMPRESSED) return entries * 2 * 2 + (uint64_t)t.num_nodes() * 16;
    // f16 arenas: 2 bytes per entry, per-node scales and maxima
    if (storage == GTO_STORAGE_HALF || storage == GTO_STORAGE_HALF_ROWS)
        return entries * 2 * 2 + (uint64_t)t.num_nodes() * 32;
    return entries * 2 * 4;
}

uint32_t gto_spot_num_suit_perms(const gto_spot* spot) { return (uint32_t)spot->spot->suit_perms.size(); }

uint32_t gto_combo_index(uint8_t c1, uint8_t c2) { return gto::combo_index(c1, c2); }

uint32_t gto_evaluate7(const uint8_t* cards7) { return gto::evaluate7(cards7); }

int gto_range_parse(const char* text, float* weights1326) {
    try {
        std::vector<float> w = gto::parse_range(text ? text : "");
        std::memcpy(weights1326, w.data(), sizeof(float) * gto::kNumCombos);
        return GTO_OK;
    } catch (const std::exception& e) {
        gto::set_last_error(e.what());
        return GTO_ERR_INVALID;
    }
}

int gto_canonical_flops(uint32_t n_subset, uint8_t* cards3, uint32_t* weights, uint32_t cap, uint32_t* n) {
    std::vector<gto::FlopClass> v = gto::canonical_flops_subset(n_subset);
    uint32_t k = 0;
    for (; k < v.size() && k < cap; ++k) {
        for (int i = 0; i < 3; ++i) cards3[3 * k + i] = v[k].cards[i];
        weights[k] = v[k].weight;
    }
    *n = (uint32_t)v.size();
    return GTO_OK;
}

}  // extern "C"
