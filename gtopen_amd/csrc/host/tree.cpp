// Betting-tree builder. Produces the same node sequence as the reference
// builder (crates/solver/src/tree.rs:287-787): depth-first, a node is pushed
// before its children, chance nodes own 52 child slots (board cards hold
// SENTINEL; already-dealt turn cards still get a subtree, tree.rs:692-697),
// action nodes reserve na*nh_actor arena elements in push order
// (tree.rs:546-548). All amounts are f64, as in the reference.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "host.h"

namespace gto {

namespace {

struct State {  // tree.rs:213-225
    uint8_t street;
    uint8_t to_act;
    double put[2];
    double street_bet[2];
    double last_increment;
    uint8_t num_raises;
    int last_aggressor;  // -1 = none
    bool checked;
};

void dedupe_amounts(std::vector<double>& v) {  // tree.rs:781-787
    std::sort(v.begin(), v.end());
    std::vector<double> out;
    for (double x : v)
        if (out.empty() || !(std::fabs(x - out.back()) < 1e-6)) out.push_back(x);
    v.swap(out);
}

class Builder {
   public:
    Builder(const TreeConfig& cfg, uint8_t root_street, uint64_t board_mask, uint64_t max_nodes,
            const size_t num_hands[2])
        : cfg_(cfg), root_street_(root_street), board_mask_(board_mask), max_nodes_(max_nodes) {
        nh_[0] = num_hands[0];
        nh_[1] = num_hands[1];
        t_.root_street = root_street;
    }

    Tree finish() { return std::move(t_); }

    uint32_t action_node(const State& st) {  // tree.rs:535-574
        check_budget();
        std::vector<Action> acts = legal_actions(st);
        size_t n = acts.size();
        if (n == 0 || n > 250) throw Error("invalid action count " + std::to_string(n) + " during tree build");
        uint32_t actions_start = (uint32_t)t_.actions.size();
        t_.actions.insert(t_.actions.end(), acts.begin(), acts.end());
        int me = st.to_act;
        uint64_t data_offset = t_.data_size[me];
        t_.data_size[me] += (uint64_t)n * nh_[me];
        uint32_t idx = push(KIND_ACTION, st.to_act, st.street, (uint8_t)n, actions_start, data_offset,
                            st.put, 0.0, 0.0, 0.0);
        std::vector<uint32_t> kids;
        kids.reserve(n);
        for (const Action& a : acts) kids.push_back(apply_action(st, a));
        t_.children_start[idx] = (uint32_t)t_.children.size();
        t_.children.insert(t_.children.end(), kids.begin(), kids.end());
        return idx;
    }

   private:
    const TreeConfig& cfg_;
    uint8_t root_street_;
    uint64_t board_mask_;
    uint64_t max_nodes_;
    uint64_t nh_[2];
    Tree t_;

    uint32_t push(uint8_t kind, uint8_t player, uint8_t street, uint8_t nc, uint32_t actions_start,
                  uint64_t data_offset, const double put[2], double tw, double tl, double tt) {
        uint32_t idx = (uint32_t)t_.kind.size();
        t_.kind.push_back(kind);
        t_.player.push_back(player);
        t_.street.push_back(street);
        t_.num_children.push_back(nc);
        t_.children_start.push_back(0);
        t_.actions_start.push_back(actions_start);
        t_.data_offset.push_back(data_offset);
        t_.put.push_back(put[0]);
        t_.put.push_back(put[1]);
        t_.t_win.push_back(tw);
        t_.t_lose.push_back(tl);
        t_.t_tie.push_back(tt);
        return idx;
    }

    void check_budget() {  // tree.rs:412-420
        if (max_nodes_ && t_.kind.size() >= max_nodes_)
            throw Error("tree too large: over the max_nodes budget of " + std::to_string(max_nodes_) +
                        "; reduce bet/raise sizes, max_raises, or effective stack");
    }

    const StreetSizing& sizing(uint8_t player, uint8_t street) const {
        return player == 0 ? cfg_.oop[street] : cfg_.ip[street];
    }

    double rake(double pot) const {  // tree.rs:526-533
        double r = cfg_.rake_pct * pot;
        if (cfg_.rake_cap > 0.0) r = std::min(r, cfg_.rake_cap);
        return r;
    }

    std::vector<Action> legal_actions(const State& st) const {  // tree.rs:432-524
        int me = st.to_act, opp = 1 - me;
        double stack_me = cfg_.effective_stack - (st.put[me] - cfg_.starting_pot / 2.0);
        std::vector<Action> out;
        double facing = st.street_bet[opp] - st.street_bet[me];
        if (facing > 1e-9) {
            out.push_back({ActionKind::Fold, 0.0});
            out.push_back({ActionKind::Call, st.street_bet[opp]});
            bool can_raise = stack_me > facing + 1e-9 && st.num_raises < cfg_.max_raises;
            if (can_raise) {
                double pot_after_call = st.put[0] + st.put[1] + facing;
                double max_to = st.street_bet[me] + stack_me;
                std::vector<double> cand;
                for (const BetSize& s : sizing(st.to_act, st.street).raise) {
                    if (s.kind == SizeKind::PotPct)
                        cand.push_back(st.street_bet[opp] + s.value / 100.0 * pot_after_call);
                    else if (s.kind == SizeKind::PrevMult)
                        cand.push_back(st.street_bet[opp] * s.value);
                    else
                        cand.push_back(max_to);
                }
                if (cfg_.add_allin) cand.push_back(max_to);
                double min_to = st.street_bet[opp] + std::max(st.last_increment, 1e-9);
                std::vector<double> tos;
                for (double to : cand) {
                    if (to < min_to) to = min_to;
                    if (to >= max_to - 1e-9 || to >= cfg_.allin_threshold * max_to - 1e-9) to = max_to;
                    if (to > st.street_bet[opp] + 1e-9) tos.push_back(to);
                }
                dedupe_amounts(tos);
                for (double to : tos) out.push_back({ActionKind::Raise, to});
            }
        } else {
            out.push_back({ActionKind::Check, 0.0});
            if (stack_me > 1e-9) {
                const StreetSizing& sz = sizing(st.to_act, st.street);
                bool donking = st.to_act == 0 && st.street > root_street_ && st.last_aggressor == 1;
                const std::vector<BetSize>& list = donking ? sz.donk : sz.bet;
                double pot = st.put[0] + st.put[1];
                double max_to = stack_me;
                std::vector<double> cand;
                for (const BetSize& s : list) {
                    if (s.kind == SizeKind::PotPct) cand.push_back(s.value / 100.0 * pot);
                    else if (s.kind == SizeKind::PrevMult) continue;  // raise-only size: dropped
                    else cand.push_back(max_to);
                }
                if (cfg_.add_allin && (!list.empty() || !donking)) cand.push_back(max_to);
                std::vector<double> tos;
                for (double to : cand) {
                    if (to <= 1e-9) continue;
                    if (to >= max_to - 1e-9 || to >= cfg_.allin_threshold * max_to - 1e-9) to = max_to;
                    tos.push_back(to);
                }
                dedupe_amounts(tos);
                for (double to : tos) out.push_back({ActionKind::Bet, to});
            }
        }
        return out;
    }

    uint32_t apply_action(const State& st, const Action& a) {  // tree.rs:576-659
        int me = st.to_act, opp = 1 - me;
        switch (a.kind) {
            case ActionKind::Fold: {
                double f = st.put[me];
                double rk = rake(2.0 * f);
                return push(KIND_FOLD, st.to_act, st.street, 0, 0, 0, st.put, f - rk, -f, 0.0);
            }
            case ActionKind::Check: {
                if (st.checked) return street_end(st.street, st.put, st.last_aggressor, false);
                State ns = st;
                ns.to_act = st.to_act ^ 1;
                ns.checked = true;
                return action_node(ns);
            }
            case ActionKind::Call: {
                double put[2] = {st.put[0], st.put[1]};
                put[me] += a.amount - st.street_bet[me];
                double stack_after = cfg_.effective_stack - (put[me] - cfg_.starting_pot / 2.0);
                double opp_after = cfg_.effective_stack - (put[opp] - cfg_.starting_pot / 2.0);
                bool allin = stack_after <= 1e-9 || opp_after <= 1e-9;
                return street_end(st.street, put, st.to_act ^ 1, allin);
            }
            default: {  // Bet / Raise
                double to = a.amount;
                State ns = st;
                ns.put[me] += to - st.street_bet[me];
                ns.street_bet[me] = to;
                double increment = to - std::max(st.street_bet[opp], st.street_bet[me]);
                ns.to_act = st.to_act ^ 1;
                ns.last_increment = std::max(increment, st.last_increment);
                ns.num_raises = st.num_raises + (a.kind == ActionKind::Raise ? 1 : 0);
                return action_node(ns);
            }
        }
    }

    uint32_t street_end(uint8_t street, const double put[2], int aggressor, bool allin) {  // tree.rs:662-723
        if (street == 2) return showdown_node(put);
        if (allin) return runout_chance(street, put);
        check_budget();
        uint8_t next = street + 1;
        uint32_t idx = push(KIND_CHANCE, 0, next, 0, 0, 0, put, 0.0, 0.0, 0.0);
        std::vector<uint32_t> kids(52, kSentinel);
        for (int card = 0; card < 52; ++card) {
            if (board_mask_ & (1ull << card)) continue;
            State cs;
            cs.street = next;
            cs.to_act = 0;
            cs.put[0] = put[0];
            cs.put[1] = put[1];
            cs.street_bet[0] = cs.street_bet[1] = 0.0;
            cs.last_increment = 0.0;
            cs.num_raises = 0;
            cs.last_aggressor = aggressor;
            cs.checked = false;
            kids[card] = action_node(cs);
        }
        t_.children_start[idx] = (uint32_t)t_.children.size();
        t_.children.insert(t_.children.end(), kids.begin(), kids.end());
        t_.num_children[idx] = 52;
        return idx;
    }

    uint32_t runout_chance(uint8_t street, const double put[2]) {  // tree.rs:725-759
        check_budget();
        uint8_t next = street + 1;
        uint32_t idx = push(KIND_CHANCE, 0, next, 0, 0, 0, put, 0.0, 0.0, 0.0);
        std::vector<uint32_t> kids(52, kSentinel);
        for (int card = 0; card < 52; ++card) {
            if (board_mask_ & (1ull << card)) continue;
            kids[card] = next == 2 ? showdown_node(put) : runout_chance(next, put);
        }
        t_.children_start[idx] = (uint32_t)t_.children.size();
        t_.children.insert(t_.children.end(), kids.begin(), kids.end());
        t_.num_children[idx] = 52;
        return idx;
    }

    uint32_t showdown_node(const double put[2]) {  // tree.rs:761-778
        double pot = put[0] + put[1];
        double rk = rake(pot);
        return push(KIND_SHOWDOWN, 0, 2, 0, 0, 0, put, put[0] - rk, -put[0], -rk / 2.0);
    }
};

const char* street_name(int s) { return s == 0 ? "flop" : s == 1 ? "turn" : "river"; }

}  // namespace

Tree build_tree(const TreeConfig& cfg, const std::vector<Card>& board, const size_t num_hands[2],
                uint64_t max_nodes, Strictness strictness) {
    if (board.size() < 3 || board.size() > 5) throw Error("board must have 3 to 5 cards");
    if (!(cfg.starting_pot > 0.0 && std::isfinite(cfg.starting_pot)) ||
        !(cfg.effective_stack > 0.0 && std::isfinite(cfg.effective_stack)))
        throw Error("pot and stacks must be positive finite numbers");
    uint8_t root_street = (uint8_t)(board.size() - 3);
    // non-finite sizes are rejected in every list (tree.rs:316-347)
    for (int pl = 0; pl < 2; ++pl)
        for (int s = 0; s < 3; ++s) {
            const StreetSizing& sz = pl == 0 ? cfg.oop[s] : cfg.ip[s];
            const std::vector<BetSize>* lists[3] = {&sz.bet, &sz.raise, &sz.donk};
            const char* names[3] = {"bet", "raise", "donk"};
            for (int l = 0; l < 3; ++l)
                for (const BetSize& b : *lists[l])
                    if (b.kind != SizeKind::AllIn && !std::isfinite(b.value))
                        throw Error(std::string(street_name(s)) + (pl == 0 ? " OOP " : " IP ") + names[l] +
                                    ": size is not a finite number");
        }
    // raise-only sizes in consumed bet/donk lists (tree.rs:348-380)
    if (strictness == Strictness::Strict) {
        for (int pl = 0; pl < 2; ++pl)
            for (int s = root_street; s < 3; ++s) {
                const StreetSizing& sz = pl == 0 ? cfg.oop[s] : cfg.ip[s];
                bool donk_used = pl == 0 && s > root_street;
                for (int l = 0; l < 2; ++l) {
                    if (l == 1 && !donk_used) continue;
                    const std::vector<BetSize>& list = l == 0 ? sz.bet : sz.donk;
                    for (const BetSize& b : list)
                        if (b.kind == SizeKind::PrevMult) {
                            char m[32];
                            for (int p = 1; p <= 17; ++p) {  // shortest form that reads back
                                std::snprintf(m, sizeof m, "%.*g", p, b.value);
                                if (std::strtod(m, nullptr) == b.value) break;
                            }
                            throw Error(std::string(street_name(s)) + (pl == 0 ? " OOP " : " IP ") +
                                        (l == 0 ? "bet" : "donk") + ": '" + m +
                                        "x' is a raise-only size — use % of pot or 'a'");
                        }
                }
            }
    }
    uint64_t board_mask = 0;
    for (Card c : board) board_mask |= card_mask(c);
    Builder b(cfg, root_street, board_mask, max_nodes, num_hands);
    State st;
    st.street = root_street;
    st.to_act = 0;
    st.put[0] = st.put[1] = cfg.starting_pot / 2.0;
    st.street_bet[0] = st.street_bet[1] = 0.0;
    st.last_increment = 0.0;
    st.num_raises = 0;
    st.last_aggressor = -1;
    st.checked = false;
    b.action_node(st);
    return b.finish();
}

}  // namespace gto
