// Level plan construction (the role of crates/solver/src/gpu/plan.rs:100-387,
// redesigned: compact reach-owner slots, per-board showdown rank tables and
// per-card prefix indexes instead of per-hand binary searches and blocker
// loops).
#include "plan.h"

#include <algorithm>
#include <cmath>
#include <map>
#include <string>
#include <stdexcept>

namespace gto {

namespace {
constexpr uint32_t kSent = 0xFFFFFFFFu;

struct Item {
    uint32_t node;
    uint32_t level;
    uint8_t dealt[2];
    uint8_t ndealt;
    uint32_t src[2];
    double cw;  // product of the chance probabilities on the path from the root
};

uint32_t river_key_of(const gto_spot_view& v, const Item& it) {
    if (v.root_street == 2) return 0;
    if (v.root_street == 1) return it.dealt[0];
    uint32_t a = std::min(it.dealt[0], it.dealt[1]), b = std::max(it.dealt[0], it.dealt[1]);
    return a * 52 + b;
}
// permute_card, cards.rs:94-96 (card = 4 * rank + suit)
inline uint8_t permute_card(uint8_t c, const uint8_t* pm) { return (uint8_t)((c & ~3u) | pm[c & 3u]); }

// runs of at most `cap` board-sorted records sharing a board: (board, first, count, 0)
void runs(const std::vector<ShowRec>& recs, uint32_t cap, std::vector<uint32_t>& out) {
    for (uint32_t k = 0; k < recs.size();) {
        uint32_t e = k;
        while (e < recs.size() && e - k < cap && recs[e].board == recs[k].board) ++e;
        out.insert(out.end(), {recs[k].board, k, e - k, 0u});
        k = e;
    }
}

// The program of one river subtree for traverser p: its depth-first walk as
// RiverOps with every LDS row assigned on a stack (a row is live from its
// op until the walk leaves the node that needed it).
struct RiverProgram {
    const gto_spot_view& v;
    int p;
    uint32_t root;
    uint64_t base[2];
    std::vector<RiverOp> ops;
    uint32_t top = 0, rows = 0;
    bool ok = true;

    RiverProgram(const gto_spot_view& v_, int p_, uint32_t root_, const uint64_t b[2]) : v(v_), p(p_), root(root_) {
        base[0] = b[0];
        base[1] = b[1];
    }
    uint32_t alloc(uint32_t k) {
        const uint32_t r = top;
        top += k;
        rows = std::max(rows, top);
        return r;
    }
    RiverOp make(uint8_t kind, uint32_t n) {
        RiverOp op{};
        op.op = kind;
        op.rel = n - root;
        return op;
    }
    // node n with reach rows rowP (traverser) and rowO (opponent); its value
    // goes into row dst with the given mode
    void emit(uint32_t n, uint32_t rowP, uint32_t rowO, uint32_t dst, uint8_t mode) {
        const uint8_t kind = v.node_kind[n];
        if (kind == GTO_KIND_FOLD || kind == GTO_KIND_SHOWDOWN) {
            RiverOp op = make(kind == GTO_KIND_FOLD ? kOpFold : kOpShow, n);
            op.q = v.node_player[n];
            op.src = (uint8_t)rowO;
            op.dst = (uint8_t)dst;
            op.mode = mode;
            op.w[0] = (float)v.node_t_win[n];
            op.w[1] = (float)v.node_t_lose[n];
            op.w[2] = (float)v.node_t_tie[n];
            ops.push_back(op);
            return;
        }
        if (kind != GTO_KIND_ACTION) {  // a chance node below the river: not a river subtree
            ok = false;
            return;
        }
        const int q = v.node_player[n];
        const uint32_t na = v.node_num_children[n];
        const uint64_t doff = v.node_data_offset[n] - base[q];
        if (doff > 0xFFFFFFFFull || na > 250) {
            ok = false;
            return;
        }
        const uint32_t cs = v.node_children_start[n];
        if (q == p) {
            const uint32_t vbase = alloc(na);
            for (uint32_t a = 0; a < na; ++a) {
                const uint32_t c = v.children[cs + a];
                if (v.node_kind[c] == GTO_KIND_ACTION) {
                    const uint32_t save = top, nr = alloc(1);
                    RiverOp op = make(kOpDescend, n);
                    op.na = (uint8_t)na;
                    op.a = (uint8_t)a;
                    op.q = (uint8_t)q;
                    op.src = (uint8_t)rowP;
                    op.dst = (uint8_t)nr;
                    op.doff = (uint32_t)doff;
                    ops.push_back(op);
                    emit(c, nr, rowO, vbase + a, 0);
                    top = save;
                } else {
                    emit(c, rowP, rowO, vbase + a, 0);  // a terminal reads only the opponent's reach
                }
            }
            RiverOp op = make(kOpPostP, n);
            op.na = (uint8_t)na;
            op.src = (uint8_t)rowP;
            op.vrow = (uint8_t)vbase;
            op.dst = (uint8_t)dst;
            op.mode = mode;
            op.doff = (uint32_t)doff;
            ops.push_back(op);
            top = vbase;
        } else {
            // the children's values are summed in action order, 0 + v0 + v1 ..
            // (directly into dst unless dst itself accumulates)
            const bool sep = mode == 2;
            const uint32_t arow = sep ? alloc(1) : dst;
            for (uint32_t a = 0; a < na; ++a) {
                const uint32_t c = v.children[cs + a];
                const uint32_t save = top, nr = alloc(1);
                const size_t at = ops.size();
                RiverOp op = make(kOpDescend, n);
                op.na = (uint8_t)na;
                op.a = (uint8_t)a;
                op.q = (uint8_t)q;
                op.src = (uint8_t)rowO;
                op.dst = (uint8_t)nr;
                op.doff = (uint32_t)doff;
                op.cdst = (uint8_t)arow;
                op.cmode = a == 0 ? 1 : 2;
                ops.push_back(op);
                emit(c, rowP, nr, arow, a == 0 ? 1 : 2);
                const size_t skip = ops.size() - at - 1;
                if (skip > 0xFFFF) ok = false;
                ops[at].skip = (uint16_t)skip;
                top = save;
            }
            if (sep) {
                RiverOp op = make(kOpPostO, n);
                op.vrow = (uint8_t)arow;
                op.dst = (uint8_t)dst;
                op.mode = 2;
                ops.push_back(op);
                top = arow;
            }
        }
    }
};
}  // namespace

// River subtrees for the fused river kernel: the traversed children of the
// chance nodes that deal the river (or the root of a river spot), each with
// its per-traverser program (shared by subtrees with identical programs),
// and the level lists of the rest of the tree.
static void build_river(const gto_spot_view& v, Plan& pl, const std::vector<uint64_t>& dmask) {
    const uint32_t N = v.num_nodes;
    std::vector<uint8_t> in_river(N, 0);
    struct Root {
        uint32_t node, src[2], card;
    };
    std::vector<Root> roots;
    if (v.root_street == 2) {
        roots.push_back({0u, {0u, 0u}, 52u});
    } else {
        for (const LevelLists& ll : pl.levels)
            for (size_t e = 0; e < ll.edge_parent.size(); ++e) {
                const uint32_t c = ll.edge_parent[e];
                if (v.node_street[c] != 2) continue;  // chance nodes dealing the turn
                roots.push_back({ll.edge_child[e], {pl.owner[0][c], pl.owner[1][c]}, ll.edge_card[e]});
            }
    }
    pl.river_ok = true;
    std::map<std::string, uint32_t> prog_of;
    std::vector<uint32_t> stack;
    for (const Root& r : roots) {
        // the subtree: pre-order, contiguous node ids
        uint64_t base[2] = {~0ull, ~0ull}, end[2] = {0, 0};
        uint32_t board = kSent;
        stack.assign(1, r.node);
        while (!stack.empty()) {
            const uint32_t n = stack.back();
            stack.pop_back();
            in_river[n] = 1;
            const uint8_t kind = v.node_kind[n];
            if (kind == GTO_KIND_ACTION) {
                const int q = v.node_player[n];
                base[q] = std::min<uint64_t>(base[q], v.node_data_offset[n]);
                end[q] = std::max<uint64_t>(end[q], v.node_data_offset[n] + (uint64_t)v.node_num_children[n] * v.num_hands[q]);
                pl.river_entries += (uint64_t)v.node_num_children[n] * v.num_hands[q];
                for (uint32_t a = 0; a < v.node_num_children[n]; ++a)
                    stack.push_back(v.children[v.node_children_start[n] + a]);
            } else if (kind == GTO_KIND_SHOWDOWN) {
                if (board != kSent && board != pl.show_slot[n]) pl.river_ok = false;
                board = pl.show_slot[n];
            } else if (kind == GTO_KIND_CHANCE) {
                pl.river_ok = false;
            }
        }
        for (int q = 0; q < 2; ++q)
            if (base[q] == ~0ull) base[q] = 0;
        RiverSub sub{};
        sub.base[0] = base[0];
        sub.base[1] = base[1];
        sub.dealt = dmask[r.node];
        sub.node = r.node;
        sub.src[0] = r.src[0];
        sub.src[1] = r.src[1];
        sub.card = r.card;
        sub.cfv = pl.cfv_slot[r.node];
        sub.board = board;
        for (int q = 0; q < 2; ++q) {
            if (end[q] - base[q] > 0xFFFFFFFFull) pl.river_ok = false;
            sub.size[q] = end[q] > base[q] ? (uint32_t)(end[q] - base[q]) : 0u;
        }
        for (int p = 0; p < 2; ++p) {
            RiverProgram prog(v, p, r.node, base);
            const uint32_t out = prog.alloc(1);  // row 0: the subtree's value
            const uint32_t rp = prog.alloc(1), ro = prog.alloc(1);  // rows 1, 2: the root's reach
            prog.emit(r.node, rp, ro, out, 0);
            if (!prog.ok || prog.rows > 255) {
                pl.river_ok = false;
                continue;
            }
            pl.river_rows = std::max(pl.river_rows, prog.rows);
            const std::string key((const char*)prog.ops.data(), prog.ops.size() * sizeof(RiverOp));
            auto it = prog_of.find(key);
            if (it == prog_of.end()) {
                it = prog_of.emplace(key, (uint32_t)pl.river_ops.size()).first;
                pl.river_ops.insert(pl.river_ops.end(), prog.ops.begin(), prog.ops.end());
            }
            sub.prog[p] = it->second;
            sub.nops[p] = (uint32_t)prog.ops.size();
        }
        pl.river_subs.push_back(sub);
    }
    if (!pl.river_ok) {
        pl.river_subs.clear();
        pl.river_ops.clear();
        return;
    }
    // the rest of the tree: level records of nodes outside river subtrees
    for (LevelLists& ll : pl.levels) {
        for (int q = 0; q < 2; ++q)
            for (const ActRec& r : ll.act_rec[q])
                if (!in_river[r.node]) ll.act_up[q].push_back(r);
        for (size_t e = 0; e < ll.edge_rec.size(); ++e)
            if (!in_river[ll.edge_child[e]]) ll.edge_up.push_back(ll.edge_rec[e]);
        for (size_t k = 0; k < ll.fold.size(); ++k)
            if (!in_river[ll.fold[k]]) pl.fold_up.push_back(ll.fold_rec[k]);
        for (size_t k = 0; k < ll.show.size(); ++k)
            if (!in_river[ll.show[k]]) pl.show_up.push_back(ll.show_rec[k]);
    }
    std::stable_sort(pl.show_up.begin(), pl.show_up.end(),
                     [](const ShowRec& a, const ShowRec& b) { return a.board < b.board; });
    for (int c = 0; c < 3; ++c) runs(pl.show_up, kShowCaps[c], pl.show_up_chunks[c]);
}

Plan build_plan(const gto_spot_view& v, bool iso) {
    Plan pl;
    iso = iso && v.num_suit_perms > 1;
    pl.iso = iso;
    const uint32_t N = v.num_nodes;
    pl.num_nodes = N;
    pl.nh[0] = v.num_hands[0];
    pl.nh[1] = v.num_hands[1];
    pl.nh_max = std::max(pl.nh[0], pl.nh[1]);
    pl.owner[0].assign(N, 0);
    pl.owner[1].assign(N, 0);
    pl.cfv_slot.assign(N, kSent);
    pl.ch_off.assign(N, 0);
    pl.ch_cnt.assign(N, 0);
    pl.show_slot.assign(N, kSent);
    std::vector<uint32_t> level(N, kSent);
    std::vector<uint32_t> key(N, kSent);
    std::vector<float> cw(N, 0.0f);
    std::vector<uint64_t> dmask(N, 0);  // cards dealt by chance nodes above the node

    // depth-first walk over the reachable tree (chance nodes skip cards
    // already dealt, as cfr.rs:556-565 does)
    std::vector<Item> stack;
    pl.num_slots[0] = pl.num_slots[1] = 1;  // slot 0 = root reach
    Item root{0, 0, {0, 0}, 0, {0, 0}, 1.0};
    stack.push_back(root);
    uint32_t max_level = 0;
    std::vector<std::vector<uint32_t>> edges_p, edges_c, edges_card;
    while (!stack.empty()) {
        Item it = stack.back();
        stack.pop_back();
        const uint32_t n = it.node;
        level[n] = it.level;
        pl.owner[0][n] = it.src[0];
        pl.owner[1][n] = it.src[1];
        cw[n] = (float)it.cw;
        for (int d = 0; d < it.ndealt; ++d) dmask[n] |= 1ull << it.dealt[d];
        max_level = std::max(max_level, it.level);
        const uint8_t kind = v.node_kind[n];
        const uint32_t cs = v.node_children_start[n];
        if (kind == GTO_KIND_ACTION) {
            const int actor = v.node_player[n];
            const uint32_t na = v.node_num_children[n];
            pl.reached_entries += (uint64_t)na * v.num_hands[actor];
            pl.reached_entries_by_player[actor] += (uint64_t)na * v.num_hands[actor];
            for (uint32_t a = na; a-- > 0;) {
                Item c = it;
                c.node = v.children[cs + a];
                c.level = it.level + 1;
                c.src[actor] = pl.num_slots[actor]++;
                stack.push_back(c);
            }
        } else if (kind == GTO_KIND_CHANCE) {
            pl.ch_off[n] = (uint32_t)pl.ch_child.size();
            // suit isomorphism (plan.rs:154-191, cfr.rs:568-603): symmetries
            // fixing the dealt cards; each card's orbit representative is its
            // smallest image, and only representatives are traversed
            std::vector<uint32_t> valid;
            if (iso)
                for (uint32_t k = 0; k < v.num_suit_perms; ++k) {
                    bool fix = true;
                    for (int d = 0; d < it.ndealt; ++d)
                        if (permute_card(it.dealt[d], v.suit_perms + 4 * k) != it.dealt[d]) fix = false;
                    if (fix) valid.push_back(k);
                }
            std::vector<Item> kids;
            for (uint32_t c = 0; c < 52; ++c) {
                uint32_t child = v.children[cs + c];
                if (child == kSent) continue;
                bool dealt = false;
                for (int d = 0; d < it.ndealt; ++d) dealt |= it.dealt[d] == c;
                if (dealt) continue;
                uint8_t rep = (uint8_t)c;
                uint32_t k_rep = 0;
                for (uint32_t k : valid) {
                    const uint8_t pc = permute_card((uint8_t)c, v.suit_perms + 4 * k);
                    if (pc < rep) {
                        rep = pc;
                        k_rep = k;
                    }
                }
                // the value row of the representative, read through the hand
                // table of the symmetry that maps this card onto it
                pl.ch_child.push_back(v.children[cs + rep]);
                pl.ch_card.push_back((uint8_t)c);
                pl.ch_perm.push_back((uint8_t)k_rep);
                if (rep != c) continue;
                Item k = it;
                k.node = child;
                k.level = it.level + 1;
                k.dealt[k.ndealt++] = (uint8_t)c;
                k.src[0] = pl.num_slots[0]++;
                k.src[1] = pl.num_slots[1]++;
                k.cw = it.cw / (double)(46 - (int)v.node_street[n]);
                kids.push_back(k);
                if (edges_p.size() <= it.level) {
                    edges_p.resize(it.level + 1);
                    edges_c.resize(it.level + 1);
                    edges_card.resize(it.level + 1);
                }
                edges_p[it.level].push_back(n);
                edges_c[it.level].push_back(child);
                edges_card[it.level].push_back(c);
            }
            pl.ch_cnt[n] = (uint32_t)(pl.ch_child.size() - pl.ch_off[n]);
            for (size_t k = kids.size(); k-- > 0;) stack.push_back(kids[k]);
        } else if (kind == GTO_KIND_SHOWDOWN) {
            key[n] = river_key_of(v, it);
        }
    }

    // level lists (node-index order within a level)
    const uint32_t L = max_level + 1;
    pl.levels.resize(L);
    uint32_t next_cfv = 0;
    for (uint32_t l = 0; l < L; ++l) {
        LevelLists& ll = pl.levels[l];
        if (l < edges_p.size()) {
            ll.edge_parent = edges_p[l];
            ll.edge_child = edges_c[l];
            ll.edge_card = edges_card[l];
        }
    }
    for (uint32_t n = 0; n < N; ++n) {
        if (level[n] == kSent) continue;
        LevelLists& ll = pl.levels[level[n]];
        switch (v.node_kind[n]) {
            case GTO_KIND_ACTION: ll.act[v.node_player[n]].push_back(n); break;
            case GTO_KIND_CHANCE: ll.chance.push_back(n); break;
            case GTO_KIND_FOLD: ll.fold.push_back(n); break;
            default: ll.show.push_back(n); break;
        }
    }
    // cfv slots in level order so one launch touches a contiguous range
    for (uint32_t l = 0; l < L; ++l) {
        LevelLists& ll = pl.levels[l];
        for (auto* lst : {&ll.act[0], &ll.act[1], &ll.chance, &ll.fold, &ll.show})
            for (uint32_t n : *lst) pl.cfv_slot[n] = next_cfv++;
    }
    pl.num_reached = next_cfv;

    // river boards: dense slots in first-seen order of showdown nodes; the
    // showdown lists are then sorted by board so neighbouring workgroups
    // share a board's tables in cache
    std::map<uint32_t, uint32_t> slot_of;
    for (uint32_t n = 0; n < N; ++n) {
        if (key[n] == kSent) continue;
        auto it = slot_of.find(key[n]);
        if (it == slot_of.end()) {
            uint32_t s = (uint32_t)pl.board_key.size();
            slot_of[key[n]] = s;
            pl.board_key.push_back(key[n]);
            pl.show_slot[n] = s;
        } else {
            pl.show_slot[n] = it->second;
        }
    }
    for (auto& ll : pl.levels)
        std::stable_sort(ll.show.begin(), ll.show.end(),
                         [&](uint32_t a, uint32_t b) { return pl.show_slot[a] < pl.show_slot[b]; });
    pl.num_boards = (uint32_t)pl.board_key.size();

    for (int q = 0; q < 2; ++q) {
        pl.sd_off[q].resize(pl.num_boards);
        pl.sd_cnt[q].resize(pl.num_boards);
        pl.sd_coff[q].resize((size_t)pl.num_boards * 53);
        pl.sd_cpos_base[q].resize(pl.num_boards);
        for (uint32_t b = 0; b < pl.num_boards; ++b) {
            uint32_t k = pl.board_key[b];
            int64_t off = v.river_off[q][k];
            if (off < 0) throw std::runtime_error("river table missing for a reachable board");
            uint32_t cnt = v.river_cnt[q][k];
            pl.sd_off[q][b] = (uint32_t)pl.sd_idx[q].size();
            pl.sd_cnt[q][b] = cnt;
            pl.max_board_cnt[q] = std::max(pl.max_board_cnt[q], cnt);
            for (uint32_t t = 0; t < cnt; ++t) pl.sd_idx[q].push_back(v.river_idx[q][off + t]);
            // per-card position lists of q's sorted hands on this board
            std::vector<std::vector<uint16_t>> by_card(52);
            for (uint32_t t = 0; t < cnt; ++t) {
                uint16_t h = v.river_idx[q][off + t];
                by_card[v.hand_c1[q][h]].push_back((uint16_t)t);
                by_card[v.hand_c2[q][h]].push_back((uint16_t)t);
            }
            pl.sd_cpos_base[q][b] = (uint32_t)pl.sd_cpos[q].size();
            // the device finds a board's lists at 2 * sd_off (two entries per hand)
            if (pl.sd_cpos_base[q][b] != 2 * pl.sd_off[q][b]) throw std::runtime_error("per-card list offsets out of step");
            uint32_t acc = 0;
            for (int c = 0; c < 52; ++c) {
                pl.sd_coff[q][(size_t)b * 53 + c] = acc;
                for (uint16_t pos : by_card[c]) pl.sd_cpos[q].push_back(pos);
                acc += (uint32_t)by_card[c].size();
            }
            pl.sd_coff[q][(size_t)b * 53 + 52] = acc;
        }
    }
    // rank indexes of each sorted hand of p against the opponent's sorted list
    for (int p = 0; p < 2; ++p) {
        const int o = 1 - p;
        pl.sd_lbub[p].resize(pl.sd_idx[p].size());
        pl.sd_cnt4[p].resize(pl.sd_idx[p].size());
        for (uint32_t b = 0; b < pl.num_boards; ++b) {
            uint32_t k = pl.board_key[b];
            const uint32_t* pstr = v.river_str[p] + v.river_off[p][k];
            const uint32_t* ostr = v.river_str[o] + v.river_off[o][k];
            uint32_t mp = pl.sd_cnt[p][b], mo = pl.sd_cnt[o][b];
            const uint32_t* coff = &pl.sd_coff[o][(size_t)b * 53];
            const uint16_t* cpos = pl.sd_cpos[o].data() + pl.sd_cpos_base[o][b];
            for (uint32_t t = 0; t < mp; ++t) {
                uint32_t m = pstr[t];
                uint32_t lb = (uint32_t)(std::lower_bound(ostr, ostr + mo, m) - ostr);
                uint32_t ub = (uint32_t)(std::upper_bound(ostr, ostr + mo, m) - ostr);
                uint16_t h = pl.sd_idx[p][pl.sd_off[p][b] + t];
                uint8_t cc[2] = {v.hand_c1[p][h], v.hand_c2[p][h]};
                uint32_t counts[4];
                for (int x = 0; x < 2; ++x) {
                    const uint16_t* s0 = cpos + coff[cc[x]];
                    const uint16_t* s1 = cpos + coff[cc[x] + 1];
                    counts[x] = (uint32_t)(std::lower_bound(s0, s1, (uint16_t)lb) - s0);
                    counts[2 + x] = (uint32_t)(std::lower_bound(s0, s1, (uint16_t)ub) - s0);
                }
                for (int x = 0; x < 4; ++x)
                    if (counts[x] > 255) throw std::runtime_error("per-card list longer than 255");
                pl.sd_lbub[p][pl.sd_off[p][b] + t] = lb | (ub << 16);
                pl.sd_cnt4[p][pl.sd_off[p][b] + t] = counts[0] | (counts[1] << 8) | (counts[2] << 16) | (counts[3] << 24);
            }
        }
    }
    // showdown tables: 16 bytes per (board, sorted hand of p), at sd_off[p][b]
    //   x = lb | ub << 16
    //   y = seg(c1) | seg(c2) << 16   start of each card's list in the
    //       opponent's per-card arrays (the card offset plus the card index)
    //   z = entries before lb (c1, c2), before ub (c1, c2), one byte each
    //   w = list length (c1) | (c2) << 8 | opponent sorted position of the
    //       identical combo << 16 (0xFFFF: none)
    for (int p = 0; p < 2; ++p) {
        const int o = 1 - p;
        pl.sd_tab[p].assign(pl.sd_idx[p].size() * 4, 0u);
        for (uint32_t b = 0; b < pl.num_boards; ++b) {
            const uint32_t* coff = &pl.sd_coff[o][(size_t)b * 53];
            std::vector<int> pos_of(v.num_hands[o], -1);
            for (uint32_t t = 0; t < pl.sd_cnt[o][b]; ++t) pos_of[pl.sd_idx[o][pl.sd_off[o][b] + t]] = (int)t;
            for (uint32_t t = 0; t < pl.sd_cnt[p][b]; ++t) {
                uint32_t k = pl.sd_off[p][b] + t;
                uint16_t h = pl.sd_idx[p][k];
                uint32_t a1 = v.hand_c1[p][h], a2 = v.hand_c2[p][h];
                uint32_t l1 = coff[a1 + 1] - coff[a1], l2 = coff[a2 + 1] - coff[a2];
                uint32_t sp = 0xFFFFu;
                uint32_t sm = v.same_combo[p][h];
                if (sm != kSent) {
                    if (pos_of[sm] < 0) throw std::runtime_error("identical combo missing from a board list");
                    sp = (uint32_t)pos_of[sm];
                }
                if (coff[52] + 52 > 0xFFFFu || l1 > 255 || l2 > 255) throw std::runtime_error("showdown table overflow");
                uint32_t* e = &pl.sd_tab[p][4 * (size_t)k];
                e[0] = pl.sd_lbub[p][k];
                e[1] = (coff[a1] + a1) | ((coff[a2] + a2) << 16);
                e[2] = pl.sd_cnt4[p][k];
                e[3] = l1 | (l2 << 8) | (sp << 16);
            }
        }
    }
    // absorbed opponent nodes (records.h): an action node of one player with
    // at most kMaxAbsorb actions whose parent is an action node of the other
    // player with at most kMaxAbsorbParent actions
    std::vector<uint8_t> absorbed(v.num_nodes, 0);
    for (uint32_t l = 0; l < L; ++l)
        for (int q = 0; q < 2; ++q)
            for (uint32_t n : pl.levels[l].act[q]) {
                const uint32_t cs = v.node_children_start[n];
                if (v.node_num_children[n] > (uint32_t)kMaxAbsorbParent) continue;
                for (uint32_t a = 0; a < v.node_num_children[n]; ++a) {
                    const uint32_t m = v.children[cs + a];
                    if (v.node_kind[m] == GTO_KIND_ACTION && v.node_player[m] != q &&
                        v.node_num_children[m] <= (uint32_t)kMaxAbsorb) {
                        absorbed[m] = 1;
                        ++pl.absorbed[1 - q];
                    }
                }
            }
    // work records
    for (uint32_t l = 0; l < L; ++l) {
        LevelLists& ll = pl.levels[l];
        for (int q = 0; q < 2; ++q)
            for (uint32_t n : ll.act[q]) {
                ActRec r{};
                r.doff = v.node_data_offset[n];
                r.lock_off = -1;
                r.node = n;
                r.own[0] = pl.owner[0][n];
                r.own[1] = pl.owner[1][n];
                r.cfv = pl.cfv_slot[n];
                r.na = v.node_num_children[n];
                r.kid = (uint32_t)pl.kids.size();
                r.cw = cw[n];
                pl.max_actions = std::max(pl.max_actions, r.na);
                const uint32_t cs = v.node_children_start[n];
                for (uint32_t a = 0; a < r.na; ++a) pl.kids.push_back(pl.cfv_slot[v.children[cs + a]]);
                for (uint32_t a = 0; a < r.na; ++a) pl.kids.push_back(pl.owner[q][v.children[cs + a]]);
                r.xk = kNoAbsorb;
                for (uint32_t a = 0; a < r.na; ++a)
                    if (absorbed[v.children[cs + a]]) r.xk = 0;
                if (r.xk == 0) {
                    r.xk = (uint32_t)pl.xkids.size();
                    for (uint32_t a = 0; a < r.na; ++a) {
                        const uint32_t m = v.children[cs + a];
                        if (!absorbed[m]) {
                            pl.xkids.push_back(0u);
                            continue;
                        }
                        const uint32_t ms = v.node_children_start[m], mn = v.node_num_children[m];
                        pl.xkids.push_back(mn);
                        pl.xkids.push_back(pl.owner[1 - q][m]);
                        for (uint32_t b = 0; b < mn; ++b) pl.xkids.push_back(pl.cfv_slot[v.children[ms + b]]);
                    }
                }
                ll.act_rec[q].push_back(r);
                if (!absorbed[n]) ll.act_keep[q].push_back(r);
            }

        for (uint32_t n : ll.fold) {
            FoldRec r{};
            r.cfv = pl.cfv_slot[n];
            r.own[0] = pl.owner[0][n];
            r.own[1] = pl.owner[1][n];
            r.folder = v.node_player[n];
            r.t_win = (float)v.node_t_win[n];
            r.t_lose = (float)v.node_t_lose[n];
            ll.fold_rec.push_back(r);
        }
        for (uint32_t n : ll.show) {
            ShowRec r{};
            r.cfv = pl.cfv_slot[n];
            r.own[0] = pl.owner[0][n];
            r.own[1] = pl.owner[1][n];
            r.board = pl.show_slot[n];
            r.win = (float)v.node_t_win[n];
            r.lose = (float)v.node_t_lose[n];
            r.tie = (float)v.node_t_tie[n];
            ll.show_rec.push_back(r);
        }
        for (uint32_t n : ll.chance) {
            ChanceRec r{};
            r.cfv = pl.cfv_slot[n];
            r.own[0] = pl.owner[0][n];
            r.own[1] = pl.owner[1][n];
            r.off = (uint32_t)pl.chr_cfv.size();
            r.cnt = pl.ch_cnt[n];
            r.inv = 1.0f / (float)(46 - (int)v.node_street[n]);
            for (uint32_t t = 0; t < pl.ch_cnt[n]; ++t) {
                pl.chr_card.push_back(pl.ch_card[pl.ch_off[n] + t]);
                pl.chr_cfv.push_back(pl.cfv_slot[pl.ch_child[pl.ch_off[n] + t]]);
                pl.chr_perm.push_back(pl.ch_perm[pl.ch_off[n] + t]);
            }
            ll.chance_rec.push_back(r);
        }
        for (size_t e = 0; e < ll.edge_parent.size(); ++e) {
            EdgeRec r{};
            for (int q = 0; q < 2; ++q) {
                r.src[q] = pl.owner[q][ll.edge_parent[e]];
                r.dst[q] = pl.owner[q][ll.edge_child[e]];
            }
            r.card = ll.edge_card[e];
            ll.edge_rec.push_back(r);
        }
    }

    // the absorbed-children reader may form (not follow) addresses up to
    // kMaxAbsorb entries past a node's list
    if (!pl.xkids.empty()) pl.xkids.insert(pl.xkids.end(), (size_t)kMaxAbsorb, 0u);

    // all-level terminal lists; showdown runs per board
    for (const LevelLists& ll : pl.levels) {
        pl.fold_all.insert(pl.fold_all.end(), ll.fold_rec.begin(), ll.fold_rec.end());
        pl.show_all.insert(pl.show_all.end(), ll.show_rec.begin(), ll.show_rec.end());
    }
    std::stable_sort(pl.show_all.begin(), pl.show_all.end(),
                     [](const ShowRec& a, const ShowRec& b) { return a.board < b.board; });
    for (int c = 0; c < 3; ++c) {
        runs(pl.show_all, kShowCaps[c], pl.show_chunks[c]);
        for (LevelLists& ll : pl.levels) runs(ll.show_rec, kShowCaps[c], ll.show_chunks[c]);
    }
    runs(pl.show_all, kMmaNodes, pl.show_chunks_mma);

    build_river(v, pl, dmask);
    return pl;
}

}  // namespace gto
