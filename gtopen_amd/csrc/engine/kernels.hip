// Level-synchronous DCFR sweep kernels for CDNA4 (gfx950).
//
// One workgroup of 256 threads per tree node of a depth level; threads
// stride over the hands of the player whose vector the node produces, so
// every arena, reach and value access is a contiguous 256-wide run.
// Arithmetic mirrors the CPU engine (crates/solver/src/cfr.rs,
// best_response.rs, game.rs) operation for operation in f32, with f64
// accumulation where the CPU accumulates in f64 (fold and showdown sums,
// average-strategy normalisation). The file is compiled with
// -ffp-contract=off so no multiply-add is fused.
#include <type_traits>

#include "kernels.h"

namespace gto {

// Visit actions 0..na-1; the first kMaxUnroll visits get a compile-time
// action index (cached values live in registers), the rest re-read memory.
template <class F>
__device__ __forceinline__ void for_actions(int na, F&& f) {
#pragma unroll
    for (int a = 0; a < kMaxUnroll; ++a)
        if (a < na) f(a, std::true_type{});
    for (int a = kMaxUnroll; a < na; ++a) f(a, std::false_type{});
}

// Round half to even, then saturate as Rust's `as i16` / `as u16` do
// (NaN -> 0), store.rs:171-195.
__device__ __forceinline__ int16_t quant_i16(float x) {
    float r = rintf(x);
    if (!(r == r)) return 0;
    r = fminf(fmaxf(r, -32768.0f), 32767.0f);
    return (int16_t)r;
}
__device__ __forceinline__ uint16_t quant_u16(float x) {
    float r = rintf(x);
    if (!(r == r)) return 0;
    r = fminf(fmaxf(r, 0.0f), 65535.0f);
    return (uint16_t)r;
}

// ---------------------------------------------------------------------------
// Strategy at one hand of an action node.
//   CURRENT: regret matching, sigma = max(R,0) * (1/sum), uniform if
//            sum <= 1e-12 (cfr.rs:875-895); compressed: integer sums of the
//            i16 quants, uniform if the sum is 0 (cfr.rs:899-919).
//   AVERAGE: max(S,0)/sum with an f64 sum, uniform if sum <= 1e-12
//            (cfr.rs:679-702); compressed: u32 sums of the u16 quants
//            (cfr.rs:704-722).
// Values of the first kMaxUnroll actions are kept in registers.
// ---------------------------------------------------------------------------
enum SigmaKind { kCurrentF32 = 0, kAverageF32 = 1, kCurrentQ = 2, kAverageQ = 3 };

template <int KIND>
struct SigmaHand;

template <>
struct SigmaHand<kCurrentF32> {
    const float* base;
    uint32_t nh, i;
    int na;
    float inv, u;
    bool uniform;
    float c[kMaxUnroll];
    template <bool C>
    __device__ float raw(int a) const { if constexpr (C) return c[a]; else return base[(size_t)a * nh + i]; }
    __device__ void init(const float* b, uint32_t nh_, uint32_t i_, int na_) {
        base = b;
        nh = nh_;
        i = i_;
        na = na_;
        float sum = 0.0f;
#pragma unroll
        for (int a = 0; a < kMaxUnroll; ++a)
            if (a < na) {
                c[a] = base[(size_t)a * nh + i];
                sum += fmaxf(c[a], 0.0f);
            }
        for (int a = kMaxUnroll; a < na; ++a) sum += fmaxf(base[(size_t)a * nh + i], 0.0f);
        uniform = !(sum > 1e-12f);
        inv = uniform ? 0.0f : 1.0f / sum;
        u = 1.0f / (float)na;
    }
    template <bool C>
    __device__ float sigma(int a) const { return uniform ? u : fmaxf(raw<C>(a), 0.0f) * inv; }
};

template <>
struct SigmaHand<kAverageF32> {
    const float* base;
    uint32_t nh, i;
    int na;
    double sum;
    float u;
    bool uniform;
    float c[kMaxUnroll];
    template <bool C>
    __device__ float raw(int a) const { if constexpr (C) return c[a]; else return base[(size_t)a * nh + i]; }
    __device__ void init(const float* b, uint32_t nh_, uint32_t i_, int na_) {
        base = b;
        nh = nh_;
        i = i_;
        na = na_;
        double s = 0.0;
#pragma unroll
        for (int a = 0; a < kMaxUnroll; ++a)
            if (a < na) {
                c[a] = base[(size_t)a * nh + i];
                s += (double)fmaxf(c[a], 0.0f);
            }
        for (int a = kMaxUnroll; a < na; ++a) s += (double)fmaxf(base[(size_t)a * nh + i], 0.0f);
        sum = s;
        uniform = !(s > 1e-12);
        u = 1.0f / (float)na;
    }
    template <bool C>
    __device__ float sigma(int a) const { return uniform ? u : (float)((double)fmaxf(raw<C>(a), 0.0f) / sum); }
};

template <>
struct SigmaHand<kCurrentQ> {
    const int16_t* base;
    uint32_t nh, i;
    int na;
    float inv, u;
    bool uniform;
    int c[kMaxUnroll];
    template <bool C>
    __device__ int raw(int a) const { if constexpr (C) return c[a]; else return (int)base[(size_t)a * nh + i]; }
    __device__ void init(const int16_t* b, uint32_t nh_, uint32_t i_, int na_) {
        base = b;
        nh = nh_;
        i = i_;
        na = na_;
        int sum = 0;
#pragma unroll
        for (int a = 0; a < kMaxUnroll; ++a)
            if (a < na) {
                c[a] = base[(size_t)a * nh + i];
                sum += max(c[a], 0);
            }
        for (int a = kMaxUnroll; a < na; ++a) sum += max((int)base[(size_t)a * nh + i], 0);
        uniform = !(sum > 0);
        inv = uniform ? 0.0f : 1.0f / (float)sum;
        u = 1.0f / (float)na;
    }
    template <bool C>
    __device__ float sigma(int a) const { return uniform ? u : (float)max(raw<C>(a), 0) * inv; }
};

template <>
struct SigmaHand<kAverageQ> {
    const uint16_t* base;
    uint32_t nh, i;
    int na;
    float inv, u;
    bool uniform;
    uint32_t c[kMaxUnroll];
    template <bool C>
    __device__ uint32_t raw(int a) const { if constexpr (C) return c[a]; else return (uint32_t)base[(size_t)a * nh + i]; }
    __device__ void init(const uint16_t* b, uint32_t nh_, uint32_t i_, int na_) {
        base = b;
        nh = nh_;
        i = i_;
        na = na_;
        uint32_t sum = 0;
#pragma unroll
        for (int a = 0; a < kMaxUnroll; ++a)
            if (a < na) {
                c[a] = base[(size_t)a * nh + i];
                sum += c[a];
            }
        for (int a = kMaxUnroll; a < na; ++a) sum += base[(size_t)a * nh + i];
        uniform = sum == 0;
        inv = uniform ? 0.0f : 1.0f / (float)sum;
        u = 1.0f / (float)na;
    }
    template <bool C>
    __device__ float sigma(int a) const { return uniform ? u : (float)raw<C>(a) * inv; }
};

// ---------------------------------------------------------------------------
// Root: reach = range weights (gpu/mod.rs:640 in the reference's terms).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_root(DevTree T, DevState S, int mask) {
    for (int q = 0; q < 2; ++q) {
        if (!((mask >> q) & 1)) continue;
        bool any = false;
        for (uint32_t i = threadIdx.x; i < T.nh[q]; i += kBlock) {
            float w = T.weights[q][i];
            S.reach[q][i] = w;
            any |= w > 0.0f;
        }
        any = __syncthreads_or(any);
        if (threadIdx.x == 0) S.alive[q][0] = any ? 1 : 0;
    }
}

// Advance the device iteration counter (discount table index).
__global__ void k_begin_iteration(DevState S) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *S.iter += 1;
}

// ---------------------------------------------------------------------------
// Down sweep, action nodes of actor q: child reach = parent reach x sigma
// (cfr.rs:478-492, 509-515) and the child's "any reach > 0" flag, which is
// the zero-reach pruning test of cfr.rs:424-430 for the subtree below.
// ---------------------------------------------------------------------------
template <int KIND>
__global__ __launch_bounds__(kBlock) void k_down_action(DevTree T, DevState S, const uint32_t* nodes, int q) {
    const uint32_t n = nodes[blockIdx.x];
    const uint32_t nh = T.nh[q];
    const int na = T.num_children[n];
    const uint32_t cs = T.children_start[n];
    const uint32_t src = T.owner[q][n];
    __shared__ int s_any[kMaxActions];
    for (int a = threadIdx.x; a < na; a += kBlock) s_any[a] = 0;
    __syncthreads();
    float* reach = S.reach[q];
    if (!S.alive[q][src]) {
        for (int a = 0; a < na; ++a) {
            float* dst = reach + (size_t)T.owner[q][T.children[cs + a]] * nh;
            for (uint32_t i = threadIdx.x; i < nh; i += kBlock) dst[i] = 0.0f;
        }
        for (int a = threadIdx.x; a < na; a += kBlock) S.alive[q][T.owner[q][T.children[cs + a]]] = 0;
        return;
    }
    const float* pr = reach + (size_t)src * nh;
    const int64_t loff = T.lock_off[n];
    const uint64_t off = T.data_offset[n];
    float* dst_c[kMaxUnroll];
    for_actions(na, [&](int a, auto cached) {
        if constexpr (decltype(cached)::value) dst_c[a] = reach + (size_t)T.owner[q][T.children[cs + a]] * nh;
    });
    auto dst = [&](int a, auto cached) -> float* {
        if constexpr (decltype(cached)::value) return dst_c[a];
        else return reach + (size_t)T.owner[q][T.children[cs + a]] * nh;
    };
    for (uint32_t i = threadIdx.x; i < nh; i += kBlock) {
        const float p = pr[i];
        if (loff >= 0) {
            const float* sig = T.lock_sigma + loff;
            for_actions(na, [&](int a, auto cached) {
                float v = p * sig[(size_t)a * nh + i];
                dst(a, cached)[i] = v;
                if (v > 0.0f) s_any[a] = 1;
            });
            continue;
        }
        SigmaHand<KIND> sh;
        if constexpr (KIND == kCurrentF32) sh.init(S.R[q] + off, nh, i, na);
        else if constexpr (KIND == kAverageF32) sh.init(S.S[q] + off, nh, i, na);
        else if constexpr (KIND == kCurrentQ) sh.init(S.Rq[q] + off, nh, i, na);
        else sh.init(S.Sq[q] + off, nh, i, na);
        for_actions(na, [&](int a, auto cached) {
            float v = p * sh.template sigma<decltype(cached)::value>(a);
            dst(a, cached)[i] = v;
            if (v > 0.0f) s_any[a] = 1;
        });
    }
    __syncthreads();
    for (int a = threadIdx.x; a < na; a += kBlock) S.alive[q][T.owner[q][T.children[cs + a]]] = s_any[a] ? 1 : 0;
}

// ---------------------------------------------------------------------------
// Down sweep, chance edges: child reach = parent reach with the hands that
// hold the dealt card zeroed (cfr.rs:615-629), for the players in `mask`.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_down_chance(DevTree T, DevState S, const uint32_t* parents,
                                                        const uint32_t* childs, const uint32_t* cards, int mask) {
    const uint32_t e = blockIdx.x;
    const uint32_t pn = parents[e], cn = childs[e];
    const uint64_t cm = 1ull << cards[e];
    for (int q = 0; q < 2; ++q) {
        if (!((mask >> q) & 1)) continue;
        const uint32_t nh = T.nh[q];
        const uint32_t src = T.owner[q][pn], dst = T.owner[q][cn];
        float* out = S.reach[q] + (size_t)dst * nh;
        const bool alive = S.alive[q][src];
        const float* in = S.reach[q] + (size_t)src * nh;
        bool any = false;
        for (uint32_t i = threadIdx.x; i < nh; i += kBlock) {
            float v = (!alive || (T.hmask[q][i] & cm)) ? 0.0f : in[i];
            out[i] = v;
            any |= v > 0.0f;
        }
        any = __syncthreads_or(any);
        if (threadIdx.x == 0) S.alive[q][dst] = any ? 1 : 0;
    }
}

// ---------------------------------------------------------------------------
// Block-wide f64 reduction helper (deterministic tree order).
// ---------------------------------------------------------------------------
__device__ inline double block_sum_f64(double v, double* sh) {
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    double r = sh[0];
    __syncthreads();
    return r;
}

// ---------------------------------------------------------------------------
// Up sweep, fold terminals (game.rs:470-496): value = amount * (opponent
// reach compatible with the hand), with f64 totals. Per-card sums run
// serially in hand order, as the CPU accumulates them.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_up_fold(DevTree T, DevState S, const uint32_t* nodes, int p) {
    const int o = 1 - p;
    const uint32_t n = nodes[blockIdx.x];
    const uint32_t nhp = T.nh[p], nho = T.nh[o];
    float* out = S.cfv + (size_t)T.cfv_slot[n] * T.nh_max;
    const uint32_t osrc = T.owner[o][n];
    if (!S.alive[o][osrc]) {
        for (uint32_t i = threadIdx.x; i < nhp; i += kBlock) out[i] = 0.0f;
        return;
    }
    const float* ro = S.reach[o] + (size_t)osrc * nho;
    __shared__ double s_card[52];
    __shared__ double s_red[kBlock];
    if (threadIdx.x < 52) {
        const uint32_t c = threadIdx.x;
        double acc = 0.0;
        for (uint32_t k = T.hc_off[o][c]; k < T.hc_off[o][c + 1]; ++k) acc += (double)ro[T.hc_idx[o][k]];
        s_card[c] = acc;
    }
    double part = 0.0;
    for (uint32_t j = threadIdx.x; j < nho; j += kBlock) part += (double)ro[j];
    const double tot = block_sum_f64(part, s_red);  // includes the barrier that publishes s_card
    const float amount = T.player[n] == p ? T.t_lose[n] : T.t_win[n];
    for (uint32_t i = threadIdx.x; i < nhp; i += kBlock) {
        uint32_t sm = T.same[p][i];
        float same_r = sm != 0xFFFFFFFFu ? ro[sm] : 0.0f;
        float valid = (float)(tot - s_card[T.c1[p][i]] - s_card[T.c2[p][i]]) + same_r;
        out[i] = amount * valid;
    }
}

// ---------------------------------------------------------------------------
// Up sweep, showdown terminals (game.rs:501-577). The opponent's reach is
// staged in LDS in strength order; f64 exclusive prefix and suffix sums
// give the strictly-weaker / strictly-stronger totals at each hand's static
// rank (lb, ub); per-card prefix and suffix sums (serial, in the CPU's
// accumulation order) remove the hands that share a card with ours.
// LDS: m_o floats + 2(m_o+1) doubles + 2(E+52) doubles, E = 2 m_o.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_up_show(DevTree T, DevState S, const uint32_t* nodes, int p) {
    extern __shared__ double lds[];
    __shared__ double s_part[kBlock];
    __shared__ double s_scan[kBlock];
    const int o = 1 - p;
    const uint32_t n = nodes[blockIdx.x];
    const uint32_t nhp = T.nh[p], nho = T.nh[o];
    float* out = S.cfv + (size_t)T.cfv_slot[n] * T.nh_max;
    const uint32_t osrc = T.owner[o][n];
    for (uint32_t i = threadIdx.x; i < nhp; i += kBlock) out[i] = 0.0f;
    if (!S.alive[o][osrc]) return;
    const float* ro = S.reach[o] + (size_t)osrc * nho;
    const uint32_t b = T.show_slot[n];
    const uint32_t mo = T.sd_cnt[o][b], mp = T.sd_cnt[p][b];
    const uint16_t* oidx = T.sd_idx[o] + T.sd_off[o][b];
    const uint32_t* coff = T.sd_coff[o] + (size_t)b * 53;
    const uint16_t* cpos = T.sd_cpos[o] + T.sd_cpos_base[o][b];
    const uint32_t E = coff[52];
    double* P = lds;                  // mo + 1
    double* Q = P + (mo + 1);         // mo + 1
    double* cpre = Q + (mo + 1);      // E + 52
    double* csuf = cpre + (E + 52);   // E + 52
    float* rs = (float*)(csuf + (E + 52));  // mo
    for (uint32_t k = threadIdx.x; k < mo; k += kBlock) rs[k] = ro[oidx[k]];
    __syncthreads();
    // chunked scan: thread t owns [t*ch, min((t+1)*ch, mo))
    const uint32_t ch = (mo + kBlock - 1) / kBlock;
    const uint32_t lo = min(threadIdx.x * ch, mo), hi = min(lo + ch, mo);
    double acc = 0.0;
    for (uint32_t k = lo; k < hi; ++k) acc += (double)rs[k];
    s_part[threadIdx.x] = acc;
    s_scan[threadIdx.x] = acc;
    __syncthreads();
    // inclusive Hillis-Steele scan of the chunk totals
    for (int s = 1; s < kBlock; s <<= 1) {
        double add = (int)threadIdx.x >= s ? s_scan[threadIdx.x - s] : 0.0;
        __syncthreads();
        s_scan[threadIdx.x] += add;
        __syncthreads();
    }
    const double total = s_scan[kBlock - 1];
    {
        double run = s_scan[threadIdx.x] - s_part[threadIdx.x];  // exclusive base
        for (uint32_t k = lo; k < hi; ++k) {
            P[k] = run;
            run += (double)rs[k];
        }
        if (threadIdx.x == kBlock - 1) P[mo] = total;
        double suf = total - s_scan[threadIdx.x];  // sum of later chunks
        // suffix within the chunk, accumulated from its end
        double runs = 0.0;
        for (uint32_t k = hi; k-- > lo;) {
            runs += (double)rs[k];
            Q[k] = runs + suf;
        }
        if (threadIdx.x == 0) Q[mo] = 0.0;
    }
    if (threadIdx.x < 52) {
        const uint32_t c = threadIdx.x;
        const uint32_t s0 = coff[c], s1 = coff[c + 1];
        const uint32_t base = s0 + c;
        double a = 0.0;
        for (uint32_t t = s0; t < s1; ++t) {
            cpre[base + (t - s0)] = a;
            a += (double)rs[cpos[t]];
        }
        cpre[base + (s1 - s0)] = a;
        double d = 0.0;
        csuf[base + (s1 - s0)] = 0.0;
        for (uint32_t t = s1; t-- > s0;) {
            d += (double)rs[cpos[t]];
            csuf[base + (t - s0)] = d;
        }
    }
    __syncthreads();
    const float win = T.t_win[n], lose = T.t_lose[n], tie = T.t_tie[n];
    const uint16_t* pidx = T.sd_idx[p] + T.sd_off[p][b];
    const uint32_t* lbub = T.sd_lbub[p] + T.sd_off[p][b];
    const uint32_t* cnt4 = T.sd_cnt4[p] + T.sd_off[p][b];
    for (uint32_t k = threadIdx.x; k < mp; k += kBlock) {
        const uint32_t i = pidx[k];
        const uint32_t lb = lbub[k] & 0xFFFFu, ub = lbub[k] >> 16;
        const uint32_t c4 = cnt4[k];
        const uint32_t a1 = T.c1[p][i], a2 = T.c2[p][i];
        const uint32_t b1 = coff[a1] + a1, b2 = coff[a2] + a2;
        const uint32_t l1 = coff[a1 + 1] - coff[a1], l2 = coff[a2 + 1] - coff[a2];
        const float lower = (float)(P[lb] - cpre[b1 + (c4 & 0xFF)] - cpre[b2 + ((c4 >> 8) & 0xFF)]);
        const float higher = (float)(Q[ub] - csuf[b1 + ((c4 >> 16) & 0xFF)] - csuf[b2 + (c4 >> 24)]);
        const uint32_t sm = T.same[p][i];
        const float same_r = sm != 0xFFFFFFFFu ? ro[sm] : 0.0f;
        const float valid = (float)(total - cpre[b1 + l1] - cpre[b2 + l2]) + same_r;
        const float tied = valid - lower - higher;
        out[i] = win * lower + lose * higher + tie * tied;
    }
}

// ---------------------------------------------------------------------------
// Up sweep, chance nodes (cfr.rs:637-663): sum the children's values over
// the dealt cards in card order, skipping hands that hold the card, then
// scale by 1/(46 - street).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_up_chance(DevTree T, DevState S, const uint32_t* nodes, int p) {
    const int o = 1 - p;
    const uint32_t n = nodes[blockIdx.x];
    const uint32_t nhp = T.nh[p];
    float* out = S.cfv + (size_t)T.cfv_slot[n] * T.nh_max;
    if (!S.alive[o][T.owner[o][n]]) {
        for (uint32_t i = threadIdx.x; i < nhp; i += kBlock) out[i] = 0.0f;
        return;
    }
    const uint32_t off = T.ch_off[n], cnt = T.ch_cnt[n];
    const float inv = 1.0f / (float)(46 - (int)T.street[n]);
    for (uint32_t i = threadIdx.x; i < nhp; i += kBlock) {
        const uint64_t hm = T.hmask[p][i];
        float acc = 0.0f;
        for (uint32_t t = 0; t < cnt; ++t) {
            if ((hm >> T.ch_card[off + t]) & 1ull) continue;
            acc += S.cfv[(size_t)T.cfv_slot[T.ch_child[off + t]] * T.nh_max + i];
        }
        out[i] = acc * inv;
    }
}

// ---------------------------------------------------------------------------
// Up sweep, action nodes.
//   MODE 0 (DCFR/CFR+ iteration, f32 arenas): at the traverser's nodes
//     out = sum_a sigma_a v_a, then R <- R*(R>0 ? pos : neg) + (v_a - out),
//     S <- S*strat + reach*sigma_a (cfr.rs:297-336, 493-502); locked nodes
//     play the lock without updating. Opponent nodes sum their children.
//   MODE 1 (best response): max over children at the traverser's nodes,
//     lock-weighted where locked (best_response.rs:67-95).
//   MODE 2 (average strategy values): average-strategy weighting
//     (best_response.rs:180-195).
//   MODE 3 (current strategy values): regret-matched weighting, no update.
// Subtrees whose opponent reach is all zero produce zeros and no update.
// ---------------------------------------------------------------------------
template <int MODE, bool COMPRESSED>
__global__ __launch_bounds__(kBlock) void k_up_action(DevTree T, DevState S, const uint32_t* nodes, int p, int q) {
    const int o = 1 - p;
    const uint32_t n = nodes[blockIdx.x];
    const uint32_t nhp = T.nh[p];
    const uint32_t nh_max = T.nh_max;
    const int na = T.num_children[n];
    const uint32_t cs = T.children_start[n];
    float* out = S.cfv + (size_t)T.cfv_slot[n] * nh_max;
    if (!S.alive[o][T.owner[o][n]]) {
        for (uint32_t i = threadIdx.x; i < nhp; i += kBlock) out[i] = 0.0f;
        return;
    }
    const float* vb[kMaxUnroll];
    for_actions(na, [&](int a, auto cached) {
        if constexpr (decltype(cached)::value) vb[a] = S.cfv + (size_t)T.cfv_slot[T.children[cs + a]] * nh_max;
    });
    auto val = [&](int a, auto cached, uint32_t i) -> float {
        if constexpr (decltype(cached)::value) return vb[a][i];
        else return S.cfv[(size_t)T.cfv_slot[T.children[cs + a]] * nh_max + i];
    };
    if (q != p) {
        for (uint32_t i = threadIdx.x; i < nhp; i += kBlock) {
            float acc = 0.0f;
            for_actions(na, [&](int a, auto c) { acc += val(a, c, i); });
            out[i] = acc;
        }
        return;
    }
    const int64_t loff = T.lock_off[n];
    if (loff >= 0) {
        const float* sig = T.lock_sigma + loff;
        for (uint32_t i = threadIdx.x; i < nhp; i += kBlock) {
            float acc = 0.0f;
            for_actions(na, [&](int a, auto c) { acc += sig[(size_t)a * nhp + i] * val(a, c, i); });
            out[i] = acc;
        }
        return;
    }
    const uint64_t off = T.data_offset[n];
    if constexpr (MODE == 1) {
        for (uint32_t i = threadIdx.x; i < nhp; i += kBlock) {
            float best = 0.0f;
            for_actions(na, [&](int a, auto c) {
                float v = val(a, c, i);
                if (a == 0 || v > best) best = v;
            });
            out[i] = best;
        }
        return;
    } else if constexpr (MODE == 2) {
        for (uint32_t i = threadIdx.x; i < nhp; i += kBlock) {
            float acc = 0.0f;
            if constexpr (COMPRESSED) {
                SigmaHand<kAverageQ> sh;
                sh.init(S.Sq[p] + off, nhp, i, na);
                for_actions(na, [&](int a, auto c) {
                    acc += sh.template sigma<decltype(c)::value>(a) * val(a, c, i);
                });
            } else {
                SigmaHand<kAverageF32> sh;
                sh.init(S.S[p] + off, nhp, i, na);
                for_actions(na, [&](int a, auto c) {
                    acc += sh.template sigma<decltype(c)::value>(a) * val(a, c, i);
                });
            }
            out[i] = acc;
        }
        return;
    } else if constexpr (MODE == 3) {
        for (uint32_t i = threadIdx.x; i < nhp; i += kBlock) {
            float acc = 0.0f;
            if constexpr (COMPRESSED) {
                SigmaHand<kCurrentQ> sh;
                sh.init(S.Rq[p] + off, nhp, i, na);
                for_actions(na, [&](int a, auto c) {
                    acc += sh.template sigma<decltype(c)::value>(a) * val(a, c, i);
                });
            } else {
                SigmaHand<kCurrentF32> sh;
                sh.init(S.R[p] + off, nhp, i, na);
                for_actions(na, [&](int a, auto c) {
                    acc += sh.template sigma<decltype(c)::value>(a) * val(a, c, i);
                });
            }
            out[i] = acc;
        }
        return;
    } else {
        const uint32_t t = *S.iter;
        const float pos = S.disc[3 * t + 0], neg = S.disc[3 * t + 1], ds = S.disc[3 * t + 2];
        const float* rp = S.reach[p] + (size_t)T.owner[p][n] * nhp;
        if constexpr (!COMPRESSED) {
            float* R = S.R[p] + off;
            float* Sm = S.S[p] + off;
            for (uint32_t i = threadIdx.x; i < nhp; i += kBlock) {
                SigmaHand<kCurrentF32> sh;
                sh.init(R, nhp, i, na);
                float v[kMaxUnroll];
                float acc = 0.0f;
                for_actions(na, [&](int a, auto c) {
                    constexpr bool C = decltype(c)::value;
                    float va = val(a, c, i);
                    if constexpr (C) v[a] = va;
                    acc += sh.template sigma<C>(a) * va;
                });
                const float reach = rp[i];
                for_actions(na, [&](int a, auto c) {
                    constexpr bool C = decltype(c)::value;
                    const size_t idx = (size_t)a * nhp + i;
                    const float r = sh.template raw<C>(a);
                    float va;
                    if constexpr (C) va = v[a];
                    else va = val(a, c, i);
                    const float d = r > 0.0f ? pos : neg;
                    R[idx] = r * d + (va - acc);
                    Sm[idx] = Sm[idx] * ds + reach * sh.template sigma<C>(a);
                });
                out[i] = acc;
            }
        } else {
            // Compressed arenas (cfr.rs:350-385): decode + discount + update
            // with a block-wide max, then quantize against the new maxima.
            // Pass 2 recomputes the same values (deterministically) instead
            // of staging na*nh floats.
            __shared__ float s_rmax[kBlock], s_smax[kBlock];
            int16_t* Rq = S.Rq[p] + off;
            uint16_t* Sq = S.Sq[p] + off;
            const float kr = S.Rs[p][n] / 32767.0f;
            const float kss = S.Ss[p][n] / 65535.0f * ds;
            float rmax = 0.0f, smax = 0.0f;
            auto update_vals = [&](uint32_t i, auto&& sink) {
                SigmaHand<kCurrentQ> sh;
                sh.init(Rq, nhp, i, na);
                float v[kMaxUnroll];
                float acc = 0.0f;
                for_actions(na, [&](int a, auto c) {
                    constexpr bool C = decltype(c)::value;
                    float va = val(a, c, i);
                    if constexpr (C) v[a] = va;
                    acc += sh.template sigma<C>(a) * va;
                });
                const float reach = rp[i];
                for_actions(na, [&](int a, auto c) {
                    constexpr bool C = decltype(c)::value;
                    const size_t idx = (size_t)a * nhp + i;
                    const int qv = sh.template raw<C>(a);
                    float va;
                    if constexpr (C) va = v[a];
                    else va = val(a, c, i);
                    const float d = qv > 0 ? pos : neg;
                    const float nr = (float)qv * kr * d + (va - acc);
                    const float ns = (float)Sq[idx] * kss + reach * sh.template sigma<C>(a);
                    sink(idx, nr, ns);
                });
                return acc;
            };
            for (uint32_t i = threadIdx.x; i < nhp; i += kBlock) {
                out[i] = update_vals(i, [&](size_t, float nr, float ns) {
                    rmax = fmaxf(rmax, fabsf(nr));
                    smax = fmaxf(smax, ns);
                });
            }
            s_rmax[threadIdx.x] = rmax;
            s_smax[threadIdx.x] = smax;
            __syncthreads();
            for (int st = kBlock / 2; st > 0; st >>= 1) {
                if ((int)threadIdx.x < st) {
                    s_rmax[threadIdx.x] = fmaxf(s_rmax[threadIdx.x], s_rmax[threadIdx.x + st]);
                    s_smax[threadIdx.x] = fmaxf(s_smax[threadIdx.x], s_smax[threadIdx.x + st]);
                }
                __syncthreads();
            }
            const float RM = s_rmax[0], SM = s_smax[0];
            const bool rok = RM > 0.0f && isfinite(RM), sok = SM > 0.0f && isfinite(SM);
            const float rk = rok ? 32767.0f / RM : 0.0f, sk = sok ? 65535.0f / SM : 0.0f;
            for (uint32_t i = threadIdx.x; i < nhp; i += kBlock) {
                update_vals(i, [&](size_t idx, float nr, float ns) {
                    Rq[idx] = rok ? quant_i16(nr * rk) : (int16_t)0;
                    Sq[idx] = sok ? quant_u16(fmaxf(ns, 0.0f) * sk) : (uint16_t)0;
                });
            }
            if (threadIdx.x == 0) {
                S.Rs[p][n] = rok ? RM : 0.0f;
                S.Ss[p][n] = sok ? SM : 0.0f;
            }
        }
    }
}

}  // namespace gto
