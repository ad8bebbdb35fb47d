// Level plan: the spot's reachable tree flattened into per-depth work lists
// for level-synchronous sweeps, plus the static tables the showdown and
// fold kernels index. Built once per engine on the host.
#pragma once

#include <cstdint>
#include <vector>

#include "../../../include/gtopen_amd.h"
#include "records.h"

namespace gto {


struct LevelLists {
    // action nodes split by actor, fold, showdown and chance nodes at this depth
    std::vector<uint32_t> act[2], fold, show, chance;
    // chance edges whose parent sits at this depth: (parent, child, card)
    std::vector<uint32_t> edge_parent, edge_child, edge_card;
    // work records, in the same order as the lists above
    std::vector<ActRec> act_rec[2];
    // q's action nodes that no action node of the other player absorbs
    // (their parent is a chance node or the root): the opponent's list of a
    // DCFR sweep with absorbed opponent nodes
    std::vector<ActRec> act_keep[2];
    // the same without the nodes inside river subtrees (fused river sweeps)
    std::vector<ActRec> act_up[2];
    std::vector<EdgeRec> edge_up;
    std::vector<FoldRec> fold_rec;
    std::vector<ShowRec> show_rec;  // sorted by board slot
    std::vector<ChanceRec> chance_rec;
    std::vector<EdgeRec> edge_rec;
    // runs of show_rec sharing a board, one list per entry of kShowCaps:
    // (board, first record, count, 0)
    std::vector<uint32_t> show_chunks[3];
};

constexpr uint32_t kShowChunk = 16;    // most showdown nodes of one board per workgroup run
constexpr uint32_t kShowCaps[3] = {16, 8, 4};  // run lengths a schedule may choose
constexpr uint32_t kMmaNodes = 32;     // showdown nodes per MFMA workgroup (one 32-wide N tile)

struct Plan {
    uint32_t num_nodes = 0;
    uint32_t nh[2] = {0, 0};
    uint32_t nh_max = 0;
    uint32_t num_reached = 0;
    std::vector<LevelLists> levels;

    // Reach vectors are stored once per "owner": the root, every child of an
    // action node for its actor, every chance child for both players. A
    // node reads player q's reach (and its alive flag) from owner[q][node].
    uint32_t num_slots[2] = {0, 0};
    std::vector<uint32_t> owner[2];
    // compact per-node index into the counterfactual-value buffer
    std::vector<uint32_t> cfv_slot;

    // chance nodes: valid (card, child) lists in card order
    std::vector<uint32_t> ch_off, ch_cnt;  // per node
    // per valid card of a chance node: the value row's node (the card's
    // orbit representative under suit isomorphism, else its own child), the
    // card, and the symmetry mapping the card onto its representative
    // (0 = identity)
    std::vector<uint32_t> ch_child;
    std::vector<uint8_t> ch_card;
    std::vector<uint8_t> ch_perm;
    bool iso = false;  // suit isomorphism active (symmetries beyond the identity)

    // showdown nodes -> dense river-board slot
    std::vector<uint32_t> show_slot;  // per node (kSentinel elsewhere)
    uint32_t num_boards = 0;
    std::vector<uint32_t> board_key;   // per slot
    // per player q and board slot: q's strength-sorted hands on that board
    std::vector<uint32_t> sd_off[2], sd_cnt[2];
    std::vector<uint16_t> sd_idx[2];   // hand index, sorted ascending by strength
    // per sorted hand of q (as the traverser): lb | ub << 16, where lb/ub =
    // number of opponent hands strictly weaker / weaker-or-equal
    std::vector<uint32_t> sd_lbub[2];
    // per sorted hand of q: counts of entries before lb (c1, c2) and before
    // ub (c1, c2) in the opponent's per-card lists, one byte each
    std::vector<uint32_t> sd_cnt4[2];
    // per player q and board slot: 53 offsets (relative to sd_cpos_base) of
    // the per-card lists of q's sorted positions holding card c
    std::vector<uint32_t> sd_coff[2];
    std::vector<uint32_t> sd_cpos_base[2];
    std::vector<uint16_t> sd_cpos[2];
    uint32_t max_board_cnt[2] = {0, 0};


    // children of action nodes referenced by ActRec::kid
    std::vector<uint32_t> kids;
    // absorbed opponent children referenced by ActRec::xk (records.h)
    std::vector<uint32_t> xkids;
    uint64_t absorbed[2] = {0, 0};  // q's action nodes absorbed by the other player's nodes
    // chance children referenced by ChanceRec: dealt card and value row
    std::vector<uint8_t> chr_card;
    std::vector<uint32_t> chr_cfv;
    std::vector<uint8_t> chr_perm;  // symmetry index per entry (0 = identity)
    // per board slot and sorted hand of q (at sd_off[q][b]), four words:
    // rank bounds, per-card list segments, per-card counts, list lengths +
    // identical combo position (layout documented in build_plan)
    std::vector<uint32_t> sd_tab[2];
    uint32_t max_actions = 0;

    // terminals of every level in one list each: their values depend only on
    // the down pass, so one launch per sweep evaluates all of them
    std::vector<FoldRec> fold_all;
    std::vector<ShowRec> show_all;  // sorted by board slot
    // runs of show_all records sharing a board, one list per entry of
    // kShowCaps: (board, first record, count, 0)
    std::vector<uint32_t> show_chunks[3];
    // the same runs capped at kMmaNodes nodes (the MFMA showdown's N tile)
    std::vector<uint32_t> show_chunks_mma;

    // fused river sweeps: the subtrees, their programs (deduplicated), the
    // LDS rows the programs need, and the terminals outside river subtrees
    // (all levels, as fold_all / show_all)
    std::vector<RiverSub> river_subs;
    std::vector<RiverOp> river_ops;
    uint32_t river_rows = 0;
    bool river_ok = false;                 // every reached river subtree has a program
    uint64_t river_entries = 0;            // reached arena entries inside river subtrees
    std::vector<FoldRec> fold_up;
    std::vector<ShowRec> show_up;
    std::vector<uint32_t> show_up_chunks[3];

    // algorithmic arena entries over reached action nodes (sum na*nh_actor)
    uint64_t reached_entries = 0;
    uint64_t reached_entries_by_player[2] = {0, 0};
};

// iso: traverse one chance branch per suit orbit (cfr.rs:84-87)
Plan build_plan(const gto_spot_view& v, bool iso = true);

}  // namespace gto
