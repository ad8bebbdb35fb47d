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
    uint32_t pad;       // keeps the record at 48 bytes
};
static_assert(sizeof(ActRec) == 48, "ActRec layout");

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

}  // namespace gto
