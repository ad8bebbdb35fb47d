// Per-node work records of the level plan, shared by the host plan builder
// and the device kernels. One record holds everything a kernel needs to
// locate a node's vectors, so a wavefront reaches its data after a single
// dependent load of the record instead of walking node -> children -> slots.
#pragma once

#include <stdint.h>

namespace gto {

// Action node. Child c's value row is cfv slot kids[kid + c]; its reach
// vector for the acting player is reach slot kids[kid + na + c].
struct ActRec {
    uint64_t doff;      // arena element offset (tree data_offset)
    int64_t lock_off;   // offset into the lock sigma array, -1 = unlocked
    uint32_t node;      // tree node index (per-node scales of compressed arenas)
    uint32_t own[2];    // reach slots of the node for OOP and IP
    uint32_t cfv;       // value row of the node
    uint32_t na;        // number of actions
    uint32_t kid;       // index into the kids array
    float cw;           // product of the chance probabilities above the node
    uint32_t xk;        // absorbed children (kNoAbsorb: none): index into DevTree::xkids
};
static_assert(sizeof(ActRec) == 48, "ActRec layout");

// Absorbed opponent nodes. In the DCFR sweep whose traverser acts at node n,
// a child of n that is the opponent's action node m (with at most
// kMaxAbsorb actions) needs no value row of its own: m's value is the sum of
// its children's rows (the opponent's reach is already in them), so n adds
// them up itself. xkids[n.xk ..] holds, per action a of n: 0 (the child's own
// row, kids[kid + a]), or the count c of m's actions, then m's reach slot for
// the opponent (whose alive flag zeroes the value) and the c value rows of
// m's children in action order.
// Only nodes n with at most kMaxAbsorbParent actions absorb (the update's
// per-action-count specialisations, 1-4, read xkids; wider nodes read their
// children's own rows).
constexpr uint32_t kNoAbsorb = 0xFFFFFFFFu;
constexpr int kMaxAbsorb = 3;
constexpr int kMaxAbsorbParent = 4;

struct FoldRec {
    uint32_t cfv;
    uint32_t own[2];
    uint32_t folder;  // player who folded
    float t_win, t_lose;
    uint32_t pad[2];
};
static_assert(sizeof(FoldRec) == 32, "FoldRec layout");

struct ShowRec {
    uint32_t cfv;
    uint32_t own[2];
    uint32_t board;  // river-board slot of the showdown tables
    float win, lose, tie;
    uint32_t pad;
};
static_assert(sizeof(ShowRec) == 32, "ShowRec layout");

// Chance node: children (card, value row) at ch_card/ch_cfv[off .. off+cnt).
struct ChanceRec {
    uint32_t cfv;
    uint32_t own[2];
    uint32_t off, cnt;
    float inv;  // 1 / (46 - street)
    uint32_t pad[2];
};
static_assert(sizeof(ChanceRec) == 32, "ChanceRec layout");

// Chance edge: reach slots of the parent (src) and of the child (dst).
struct EdgeRec {
    uint32_t src[2], dst[2];
    uint32_t card;
    uint32_t pad[3];
};
static_assert(sizeof(EdgeRec) == 32, "EdgeRec layout");

// River subtree, the unit of the fused river kernel: a child of a chance
// node dealing the river (or the root of a spot that starts on the river).
// Every node below it shares one board, and the tree builder lays its nodes
// and arena blocks out contiguously in pre-order, so a subtree is its root
// plus per-player arena bases and a program of ops (RiverOp) shared by all
// subtrees of the same shape and payoffs.
struct RiverSub {
    uint64_t base[2];   // arena offset of each player's first action node in the subtree
    uint64_t dealt;     // cards dealt on the turn and river (the traverser's hands holding one are off the board)
    uint32_t node;      // tree node of the subtree root (op node ids are relative to it)
    uint32_t prog[2];   // first op of the program, per traverser
    uint32_t nops[2];   // op count, per traverser
    uint32_t src[2];    // reach slots of the parent (the river chance node), per player
    uint32_t card;      // river card dealt into the subtree (52: none)
    uint32_t cfv;       // value row of the subtree root
    uint32_t board;     // river-board slot of its showdowns (0xFFFFFFFF: none)
    uint32_t size[2];   // arena entries of each player's action nodes in the subtree (from base)
};
static_assert(sizeof(RiverSub) == 72, "RiverSub layout");

// One step of a river subtree program for one traverser p. Programs are
// the subtree's depth-first walk with every LDS row assigned statically
// (stack discipline): reach rows of both players and value rows of the
// traverser, nh_max floats each.
//   kOpDescend: reach row `dst` of the actor q's child a = row `src` times
//               q's current strategy for action a at the node. When q is the
//               opponent and the new row is all zero (zero-reach pruning,
//               cfr.rs:424-430) the child's value is 0, written into row
//               `cdst` with mode `cmode`, and the next `skip` ops (the
//               child's subtree) are skipped.
//   kOpFold / kOpShow: terminal values from the opponent's reach row `src`
//               into row `dst` (w: fold t_win, t_lose; showdown win, lose,
//               tie; `q`: the player who folded).
//   kOpPostP:   the traverser's node: out = sum sigma v over the child rows
//               vrow.., regret and strategy-sum update with reach row `src`,
//               out into row `dst`.
//   kOpPostO:   an opponent node whose children summed into row `vrow`: its
//               value added into row `dst`.
// Destination modes: 0 = store v, 1 = store 0 + v (the first term of a sum),
// 2 = add v.
enum RiverOpKind : uint8_t { kOpDescend = 0, kOpFold = 1, kOpShow = 2, kOpPostP = 3, kOpPostO = 4 };
struct RiverOp {
    uint8_t op, na, a, q;
    uint8_t src, dst, mode, vrow;
    uint8_t cdst, cmode;
    uint16_t skip;
    uint32_t rel;   // node id minus the subtree root's
    uint32_t doff;  // arena offset minus the subtree's base of the actor
    float w[3];
};
static_assert(sizeof(RiverOp) == 32, "RiverOp layout");

}  // namespace gto
