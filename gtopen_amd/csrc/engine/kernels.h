// Device-side views shared by the level-sweep kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gto {

constexpr int kBlock = 256;       // threads per workgroup (4 wavefronts of 64)
constexpr int kMaxUnroll = 8;     // actions cached in registers per hand
constexpr int kMaxActions = 256;  // tree builder caps action counts at 250

// Static tree, hand and table data (read-only for the whole solve).
struct DevTree {
    const uint8_t* player;
    const uint8_t* num_children;
    const uint8_t* street;
    const uint32_t* children_start;
    const uint64_t* data_offset;
    const float* t_win;
    const float* t_lose;
    const float* t_tie;
    const uint32_t* children;
    const uint32_t* owner[2];
    const uint32_t* cfv_slot;
    const uint32_t* ch_off;
    const uint32_t* ch_cnt;
    const uint32_t* ch_child;
    const uint8_t* ch_card;
    const uint32_t* show_slot;
    const int64_t* lock_off;
    const float* lock_sigma;
    // hands
    uint32_t nh[2];
    uint32_t nh_max;
    const uint8_t* c1[2];
    const uint8_t* c2[2];
    const uint64_t* hmask[2];
    const uint32_t* same[2];
    const float* weights[2];
    const uint32_t* hc_off[2];
    const uint16_t* hc_idx[2];
    // showdown tables
    const uint32_t* sd_off[2];
    const uint32_t* sd_cnt[2];
    const uint16_t* sd_idx[2];
    const uint32_t* sd_lbub[2];
    const uint32_t* sd_cnt4[2];
    const uint32_t* sd_coff[2];
    const uint32_t* sd_cpos_base[2];
    const uint16_t* sd_cpos[2];
};

// Mutable solver state.
struct DevState {
    float* R[2];          // f32 regrets (storage F32)
    float* S[2];          // f32 strategy sums
    int16_t* Rq[2];       // compressed regrets
    uint16_t* Sq[2];      // compressed strategy sums
    float* Rs[2];         // per-node regret scale
    float* Ss[2];         // per-node strategy scale
    float* reach[2];      // [slot][nh_q]
    uint8_t* alive[2];    // [slot]: any reach > 0
    float* cfv;           // [cfv_slot][nh_max], traverser's values
    const float* disc;    // [iteration][3]: pos, neg, strat
    uint32_t* iter;       // device iteration counter (1-based after the first begin)
};

}  // namespace gto
