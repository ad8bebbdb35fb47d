// Device-side views shared by the level-sweep kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "records.h"

namespace gto {

constexpr int kBlock = 256;       // threads per workgroup (4 wavefronts of 64)
constexpr int kWave = 64;         // CDNA wavefront
constexpr int kWaves = kBlock / kWave;
constexpr int kMaxUnroll = 8;     // actions cached in registers per hand
constexpr int kMaxActions = 64;   // engine limit on actions per node (one ballot bit each)

// Arena storage inside the kernels: f32; the reference's i16/u16 quants
// against a per-node block max; f16 values times a per-node power of two.
// kStoreF32P: f32 arenas under PCFR+ (the update also writes Q, below)
enum Store { kStoreF32 = 0, kStoreQ = 1, kStoreH = 2, kStoreF32P = 3 };

// Power-of-two scale for f16 storage of a node whose values are bounded by
// b: b / scale <= 2^15, well inside f16's range (65504), so the bound may
// overshoot by 2x without overflow; relative precision is f16's 2^-11 at
// every magnitude down to ~2^-24 of the scale.
__host__ __device__ inline float half_scale(float b) {
    if (!(b > 0.0f) || !(b < 3.0e38f)) return 1.0f;
    int e;
    (void)frexpf(b, &e);  // b = m * 2^e, m in [0.5, 1)
    e -= 15;
    if (e < -126) e = -126;
    return ldexpf(1.0f, e);
}

// Static hand and table data (read-only for the whole solve).
struct DevTree {
    const float* lock_sigma;
    float lock_bound;          // max(1, largest lock sigma): a locked child's reach / value bound factor
    const uint32_t* kids;      // ActRec children: value rows, then reach slots
    const uint32_t* xkids;     // ActRec absorbed opponent children (records.h)
    const float* zero_row;     // nh_max zeros: the rows an absorbed sum pads with
    const uint8_t* chr_card;   // ChanceRec children: dealt card
    const uint32_t* chr_cfv;   // ChanceRec children: value row (the orbit representative's)
    const uint8_t* chr_perm;   // ChanceRec children: suit symmetry onto the representative (0 = identity)
    const uint16_t* hand_perm[2];  // [k * nh_q + i]: hand i's image under symmetry k
    uint32_t nh[2];
    uint32_t nh_max;
    const uint8_t* c1[2];
    const uint8_t* c2[2];
    const uint64_t* hmask[2];
    const uint32_t* same[2];
    const uint32_t* hpack[2];    // c1 | c2 << 8 | identical opponent combo << 16 (0xFFFF: none)
    const float* weights[2];
    // showdown tables, per river-board slot
    const uint32_t* sd_off[2];        // start of q's sorted list
    const uint32_t* sd_cnt[2];        // its length
    const uint16_t* sd_idx[2];        // hand index, ascending strength
    const uint32_t* sd_coff[2];       // 53 per board: per-card list offsets
    const uint4* sd_tab[2];           // [sorted hand] (at sd_off): packed rank/segment table
    const uint16_t* sd_cpos[2];       // per board at 2 * sd_off: sorted positions holding card 0, 1, ...
    // MFMA showdown: per board, for each sorted hand of q, a bitset over the
    // opponent's sorted hands (1 = no shared card), padded to 32-bit words
    // and stored pair-split (bit 2i of a word at bit i, bit 2i+1 at 16+i)
    const uint32_t* sd_djbits[2];
    const uint32_t* sd_djoff[2];      // per board: first word of the board's bitsets
    const uint32_t* sd_lbub[2];       // [sorted hand]: lb | ub << 16 (opponent hands weaker / not stronger)
};

// LDS layout of the showdown kernel for one board (byte offsets): the
// per-card offsets stay for a run of nodes (the other static tables are
// held in registers); the per-node rows (the hand-order reach row, the
// strength-order reach and its prefix P, per-card prefixes, the value row)
// are rebuilt per node.
struct ShowLds {
    uint32_t P, cpre, creg, wsum, exr, rs, coff, nodes, O, rhs, bytes;
};
// per node of a run, staged once: value row, opponent reach slot, live
// flag, win, lose, tie
struct ShowNode {
    uint32_t cfv, oslot, live;
    float win, lose, tie;
};
// mo: the opponent's hands on the board; nho / nhp: all hands of each player
__host__ __device__ inline ShowLds show_lds_layout(uint32_t mo, uint32_t nhp, uint32_t nho) {
    ShowLds L;
    const uint32_t creg = 2 * mo + 52;
    // the last slot of P, cpre and rs is a spare: padded lanes store there
    // (P, cpre) or read a zero there (rs) instead of branching
    L.P = 0;                                    // (mo + 2) doubles
    L.cpre = L.P + 8 * (mo + 2);                // creg + 1 doubles
    L.creg = creg;
    L.wsum = L.cpre + 8 * (creg + 1);           // per wave: scan total (double), row max (float)
    L.exr = L.wsum + 48;                        // per wave: a float (sequential sums' exactness test)
    L.rs = L.exr + 16;                          // mo + 1 floats: reach in strength order, then 0
    L.coff = L.rs + 4 * (mo + 1);               // 53 u32
    L.nodes = L.coff + 53 * 4;                  // up to 16 ShowNode
    L.O = L.nodes + 16 * sizeof(ShowNode);      // nhp floats: the value row in hand order
    L.rhs = L.O + 4 * nhp;                      // nho floats: the opponent's reach row in hand order
    L.bytes = (L.rhs + 4 * nho + 15) & ~15u;
    return L;
}

// LDS layout of the fused river kernel: `nrows` rows of nh_max floats
// (reach and value rows of the subtree program), the board's 53 per-card
// offsets, reduction slots, and the terminal scratch: the showdown's f64
// prefix arrays for up to mo_max opponent hands on a board, or the fold's
// 52 x 52 card matrix and card sums.
struct RiverLds {
    uint32_t rstride, coff, red, scr, bytes;
};
__host__ __device__ inline RiverLds river_lds_layout(uint32_t nrows, uint32_t nh_max, uint32_t mo_max) {
    RiverLds L;
    L.rstride = (nh_max + 3) & ~3u;
    L.coff = nrows * L.rstride * 4;
    L.red = L.coff + 224;                 // 16 floats of reductions, then 4 doubles (scan totals)
    L.scr = L.red + 96;
    const uint32_t show = 8 * (mo_max + 2) + 8 * (2 * mo_max + 53) + 4 * (mo_max + 1);
    const uint32_t fold = 52 * 52 * 4 + 53 * 8;
    L.bytes = (L.scr + (show > fold ? show : fold) + 15) & ~15u;
    return L;
}

// LDS layout of the MFMA showdown for one board: opponent hand masks in
// strength order, then the reach of up to 32 nodes in strength order as
// bf16 ([node][k], k padded to 16), the nodes' table, per-node max |value|.
constexpr size_t kLdsBudget = 160 * 1024;  // LDS per workgroup on gfx950

struct MmaLds {
    uint32_t B, bsc, nodes, vmx, onb, oidx, pidx, lbub, O, os, bytes, kpad, rs, words;
};
// fp8 == false: B in bf16, [node][k], k padded to 16, row stride rs
// elements. fp8 == true: B in e4m3 bytes, k padded to 32 and permuted within
// each 32-k word (the order the A operand's byte extraction yields), row
// stride rs bytes, with a per-node power-of-two scale. nout != 0: the node
// rows (nout values each) are assembled in LDS and written out whole;
// nout == 0: values are stored to the rows directly.
__host__ __device__ inline MmaLds mma_lds_layout(uint32_t mo, uint32_t mp, uint32_t nout, bool fp8 = false) {
    MmaLds L;
    L.kpad = fp8 ? (mo + 31) & ~31u : (mo + 15) & ~15u;
    L.words = (L.kpad + 31) / 32;         // bitset words per hand
    // 16-byte (bf16) or 8-byte (fp8) fragment reads of the 32 node rows
    // then start 4 banks apart
    L.rs = fp8 ? ((L.kpad + 255) & ~255u) + 16 : ((L.kpad + 127) & ~127u) + 8;
    L.B = 0;                              // 32 x rs elements
    L.bsc = L.B + (fp8 ? 1 : 2) * 32 * L.rs;  // 32 f32: per-node inverse scale of B
    L.nodes = L.bsc + 32 * 4;             // 32 ShowNode
    L.vmx = L.nodes + 32 * sizeof(ShowNode);  // 32 u32
    L.onb = L.vmx + 32 * 4;               // 42 u32: bitmap of p's hands on the board (nout == 0)
    L.lbub = L.onb + 42 * 4;              // mp u32: rank bounds of p's sorted position i
    L.oidx = L.lbub + 4 * mp;             // kpad u16: opponent hand of sorted position k
    L.pidx = L.oidx + 2 * L.kpad;         // mp u16: p's hand of sorted position i
    // 32 x os f32 output rows; os odd so the 32 node rows of one scattered
    // write start in different banks
    L.os = nout ? (((nout + 63) & ~63u) + 1) : 0;
    L.O = (L.pidx + 2 * mp + 15) & ~15u;
    L.bytes = L.O + ((32 * L.os * 4 + 15) & ~15u);
    return L;
}
// Mutable solver state.
struct DevState {
    float* R[2];          // f32 regrets (storage F32)
    // f32 arena that current-strategy regret matching reads: R, or under
    // PCFR+ Q = max(R,0) + last instantaneous regret (the value cfr.rs:
    // 259-266 matches on, stored by the update that produced both)
    float* cur[2];
    float* Q[2];          // PCFR+ only (null otherwise)
    float* S[2];          // f32 strategy sums
    int16_t* Rq[2];       // compressed regrets
    uint16_t* Sq[2];      // compressed strategy sums
    float* Rs[2];         // per-node regret scale
    float* Ss[2];         // per-node strategy scale
    float* Rm[2];         // half storage: per-node max |regret| after the last update
    float* Sm[2];         // half storage: per-node max strategy sum
    float* reach[2];      // [slot][nh_q]
    uint8_t* alive[2];    // [slot]: any reach > 0
    float* rmax[2];       // [slot]: max of the reach row
    float* vmax;          // [value row]: max |value| of the row (f16 engines; null otherwise)
    float* cfv;           // [cfv_slot][nh_max], traverser's values
    // f16 rows (DCFR sweeps of GTO_STORAGE_HALF_ROWS): reach and value rows
    // hold f16 entries at the same slot positions (half the bytes), each
    // row with a power-of-two decode scale: value = f16 entry x scale
    float* rsc[2];        // [slot]: reach row decode scale
    float* vsc;           // [cfv_slot]: value row decode scale
    const float* disc;    // [iteration][3]: pos, neg, strat
    uint32_t* iter;       // device iteration counter (1-based after the first begin)
};

}  // namespace gto
