// Level-synchronous DCFR sweep kernels for CDNA4 (gfx950).
//
// One wavefront (64 lanes) per tree node of a depth level, four nodes per
// workgroup; lanes stride over the hands of the player whose vector the
// node produces, so arena, reach and value accesses are contiguous 64-wide
// runs, and reductions are wavefront-level (ballots, shuffles) with no
// workgroup barriers. Each wavefront finds its node's rows through one
// work record (records.h). Workgroups are mapped to XCD-contiguous slices
// of the work list.
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
template <int A, class F>
__device__ __forceinline__ void for_cached(int na, F& f) {
    if constexpr (A < kMaxUnroll) {
        if (A < na) {
            f(A, std::true_type{});
            for_cached<A + 1>(na, f);
        }
    }
}
template <class F>
__device__ __forceinline__ void for_actions(int na, F&& f) {
    for_cached<0>(na, f);
    for (int a = kMaxUnroll; a < na; ++a) f(a, std::false_type{});
}

// Element q (0 or 1) of a two-entry record field without dynamic indexing
// (which would place the record in scratch memory).
__device__ __forceinline__ uint32_t pick(const uint32_t (&v)[2], int q) { return q ? v[1] : v[0]; }

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
//   RAW:     the strategy arena read as a behavioural strategy, sigma = S
//            as stored (the game-value pass, whose strategy is an input).
//   *H:      f16 storage: the same arithmetic on the stored f16 values as
//            f32 (the per-node power-of-two scale cancels exactly), uniform
//            where the sum is 0.
enum SigmaKind { kCurrentF32 = 0, kAverageF32 = 1, kCurrentQ = 2, kAverageQ = 3, kRawF32 = 4, kCurrentH = 5,
                 kAverageH = 6 };

__device__ __forceinline__ float h2f(uint16_t u) { return (float)__builtin_bit_cast(_Float16, u); }
// round to nearest even (IEEE conversion)
__device__ __forceinline__ uint16_t f2h(float x) { return __builtin_bit_cast(uint16_t, (_Float16)x); }

// Reach and value rows of the DCFR sweeps. H16 = false: f32 entries. H16 =
// true (GTO_STORAGE_HALF_ROWS): f16 entries at the same slot positions, the
// row's value = entry x its power-of-two decode scale; a writer picks the
// scale from a bound of the row's magnitude (half_scale: bound / scale <
// 2^15), so encoding is one rounding to f16 (2^-11 relative) and decoding
// is exact.
// raw(i) / dec(raw, scale): an entry as stored and its value, for loads
// issued ahead of their use (a prefetched row stays raw in registers until
// it is used, so converting it does not wait for the load early).
// f16-row selections of the kernels that read or write both kinds of row
constexpr int kRows16Reach = 1, kRows16Value = 2;

template <bool H16>
struct RowIn {
    using Raw = float;
    const float* __restrict__ p;
    float sc;
    __device__ __forceinline__ RowIn() : p(nullptr), sc(1.0f) {}
    __device__ __forceinline__ RowIn(const float* base, size_t off, float) : p(base + off), sc(1.0f) {}
    __device__ __forceinline__ float operator[](uint32_t i) const { return p[i]; }
    __device__ __forceinline__ Raw raw(uint32_t i) const { return p[i]; }
    __device__ static __forceinline__ float dec(Raw x, float) { return x; }
};
template <>
struct RowIn<true> {
    // a full register per entry: entries packed two to a register would
    // make every conditional load wait for the previous one
    using Raw = uint32_t;
    const uint16_t* __restrict__ p;
    float sc;
    __device__ __forceinline__ RowIn() : p(nullptr), sc(1.0f) {}
    __device__ __forceinline__ RowIn(const float* base, size_t off, float s)
        : p((const uint16_t*)base + off), sc(s) {}
    __device__ __forceinline__ float operator[](uint32_t i) const { return h2f(p[i]) * sc; }
    __device__ __forceinline__ Raw raw(uint32_t i) const { return p[i]; }
    __device__ static __forceinline__ float dec(Raw x, float s) { return h2f((uint16_t)x) * s; }
};
// decode scale and encode factor of rows whose |entries| are <= bound
struct RowEnc {
    float sc, inv;
    __device__ __forceinline__ explicit RowEnc(float bound) : sc(half_scale(bound)), inv(1.0f / sc) {}
};
template <bool H16>
struct RowOut {
    float* __restrict__ p;
    __device__ __forceinline__ RowOut() : p(nullptr) {}
    __device__ __forceinline__ RowOut(float* base, size_t off, const RowEnc&) : p(base + off) {}
    __device__ __forceinline__ void set(uint32_t i, float x) const { p[i] = x; }
};
template <>
struct RowOut<true> {
    uint16_t* __restrict__ p;
    float inv;
    __device__ __forceinline__ RowOut() : p(nullptr), inv(1.0f) {}
    __device__ __forceinline__ RowOut(float* base, size_t off, const RowEnc& e)
        : p((uint16_t*)base + off), inv(e.inv) {}
    __device__ __forceinline__ void set(uint32_t i, float x) const { p[i] = f2h(x * inv); }
};
// row of reach slot `slot` of player q / of value row `slot`
template <bool H16>
__device__ __forceinline__ RowIn<H16> reach_in(const DevTree& T, const DevState& S, int q, uint32_t slot) {
    return RowIn<H16>(S.reach[q], (size_t)slot * T.nh[q], H16 ? S.rsc[q][slot] : 1.0f);
}
template <bool H16>
__device__ __forceinline__ RowIn<H16> cfv_in(const DevTree& T, const DevState& S, uint32_t slot) {
    return RowIn<H16>(S.cfv, (size_t)slot * T.nh_max, H16 ? S.vsc[slot] : 1.0f);
}

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

template <>
struct SigmaHand<kCurrentH> {
    const uint16_t* base;
    uint32_t nh, i;
    float inv, u;
    bool uniform;
    __device__ void init(const uint16_t* b, uint32_t nh_, uint32_t i_, int na) {
        base = b;
        nh = nh_;
        i = i_;
        float sum = 0.0f;
        for (int a = 0; a < na; ++a) sum += fmaxf(h2f(base[(size_t)a * nh + i]), 0.0f);
        uniform = !(sum > 0.0f);
        inv = uniform ? 0.0f : 1.0f / sum;
        u = 1.0f / (float)na;
    }
    template <bool C>
    __device__ float sigma(int a) const { return uniform ? u : fmaxf(h2f(base[(size_t)a * nh + i]), 0.0f) * inv; }
};

template <>
struct SigmaHand<kAverageH> {
    const uint16_t* base;
    uint32_t nh, i;
    double sum;
    float u;
    bool uniform;
    __device__ void init(const uint16_t* b, uint32_t nh_, uint32_t i_, int na) {
        base = b;
        nh = nh_;
        i = i_;
        double s2 = 0.0;
        for (int a = 0; a < na; ++a) s2 += (double)fmaxf(h2f(base[(size_t)a * nh + i]), 0.0f);
        sum = s2;
        uniform = !(s2 > 0.0);
        u = 1.0f / (float)na;
    }
    template <bool C>
    __device__ float sigma(int a) const {
        return uniform ? u : (float)((double)fmaxf(h2f(base[(size_t)a * nh + i]), 0.0f) / sum);
    }
};

template <>
struct SigmaHand<kRawF32> {
    const float* base;
    uint32_t nh, i;
    __device__ void init(const float* b, uint32_t nh_, uint32_t i_, int) {
        base = b;
        nh = nh_;
        i = i_;
    }
    template <bool C>
    __device__ float sigma(int a) const { return base[(size_t)a * nh + i]; }
};

// ---------------------------------------------------------------------------
// Strategy of one hand at a node with a compile-time action count NA: all
// NA arena values are loaded before any arithmetic (no per-action
// branches), then the same regret-matching / averaging arithmetic as
// SigmaHand, with selects instead of branches.
// ---------------------------------------------------------------------------
template <int NA, int KIND>
struct HandSigma;

// fetch(): issue the NA arena loads; finish(): strategy from them; load():
// both. Splitting lets a caller issue loads for several hands first.
template <int NA>
struct HandSigma<NA, kCurrentF32> {
    float raw[NA], sig[NA];
    __device__ __forceinline__ void fetch(const float* __restrict__ b, uint32_t nh, uint32_t i) {
#pragma unroll
        for (int a = 0; a < NA; ++a) raw[a] = b[(size_t)a * nh + i];
    }
    __device__ __forceinline__ void finish() {
        float sum = 0.0f;
#pragma unroll
        for (int a = 0; a < NA; ++a) sum += fmaxf(raw[a], 0.0f);
        const bool uni = !(sum > 1e-12f);
        const float inv = 1.0f / sum;
        const float u = 1.0f / (float)NA;
#pragma unroll
        for (int a = 0; a < NA; ++a) sig[a] = uni ? u : fmaxf(raw[a], 0.0f) * inv;
    }
    __device__ __forceinline__ void load(const float* __restrict__ b, uint32_t nh, uint32_t i) {
        fetch(b, nh, i);
        finish();
    }
};

template <int NA>
struct HandSigma<NA, kAverageF32> {
    float raw[NA], sig[NA];
    __device__ __forceinline__ void fetch(const float* __restrict__ b, uint32_t nh, uint32_t i) {
#pragma unroll
        for (int a = 0; a < NA; ++a) raw[a] = b[(size_t)a * nh + i];
    }
    __device__ __forceinline__ void finish() {
        double sum = 0.0;
#pragma unroll
        for (int a = 0; a < NA; ++a) sum += (double)fmaxf(raw[a], 0.0f);
        const bool uni = !(sum > 1e-12);
        const float u = 1.0f / (float)NA;
#pragma unroll
        for (int a = 0; a < NA; ++a) sig[a] = uni ? u : (float)((double)fmaxf(raw[a], 0.0f) / sum);
    }
    __device__ __forceinline__ void load(const float* __restrict__ b, uint32_t nh, uint32_t i) {
        fetch(b, nh, i);
        finish();
    }
};

template <int NA>
struct HandSigma<NA, kCurrentQ> {
    int raw[NA];
    float sig[NA];
    __device__ __forceinline__ void fetch(const int16_t* __restrict__ b, uint32_t nh, uint32_t i) {
#pragma unroll
        for (int a = 0; a < NA; ++a) raw[a] = b[(size_t)a * nh + i];
    }
    __device__ __forceinline__ void finish() {
        int sum = 0;
#pragma unroll
        for (int a = 0; a < NA; ++a) sum += max(raw[a], 0);
        const bool uni = !(sum > 0);
        const float inv = 1.0f / (float)sum;
        const float u = 1.0f / (float)NA;
#pragma unroll
        for (int a = 0; a < NA; ++a) sig[a] = uni ? u : (float)max(raw[a], 0) * inv;
    }
    __device__ __forceinline__ void load(const int16_t* __restrict__ b, uint32_t nh, uint32_t i) {
        fetch(b, nh, i);
        finish();
    }
};

template <int NA>
struct HandSigma<NA, kAverageQ> {
    uint32_t raw[NA];
    float sig[NA];
    __device__ __forceinline__ void fetch(const uint16_t* __restrict__ b, uint32_t nh, uint32_t i) {
#pragma unroll
        for (int a = 0; a < NA; ++a) raw[a] = b[(size_t)a * nh + i];
    }
    __device__ __forceinline__ void finish() {
        uint32_t sum = 0;
#pragma unroll
        for (int a = 0; a < NA; ++a) sum += raw[a];
        const bool uni = sum == 0;
        const float inv = 1.0f / (float)sum;
        const float u = 1.0f / (float)NA;
#pragma unroll
        for (int a = 0; a < NA; ++a) sig[a] = uni ? u : (float)raw[a] * inv;
    }
    __device__ __forceinline__ void load(const uint16_t* __restrict__ b, uint32_t nh, uint32_t i) {
        fetch(b, nh, i);
        finish();
    }
};

template <int NA>
struct HandSigma<NA, kCurrentH> {
    float raw[NA], sig[NA];
    __device__ __forceinline__ void fetch(const uint16_t* __restrict__ b, uint32_t nh, uint32_t i) {
#pragma unroll
        for (int a = 0; a < NA; ++a) raw[a] = h2f(b[(size_t)a * nh + i]);
    }
    __device__ __forceinline__ void finish() {
        float sum = 0.0f;
#pragma unroll
        for (int a = 0; a < NA; ++a) sum += fmaxf(raw[a], 0.0f);
        const bool uni = !(sum > 0.0f);
        const float inv = 1.0f / sum;
        const float u = 1.0f / (float)NA;
#pragma unroll
        for (int a = 0; a < NA; ++a) sig[a] = uni ? u : fmaxf(raw[a], 0.0f) * inv;
    }
    __device__ __forceinline__ void load(const uint16_t* __restrict__ b, uint32_t nh, uint32_t i) {
        fetch(b, nh, i);
        finish();
    }
};

template <int NA>
struct HandSigma<NA, kAverageH> {
    float raw[NA], sig[NA];
    __device__ __forceinline__ void fetch(const uint16_t* __restrict__ b, uint32_t nh, uint32_t i) {
#pragma unroll
        for (int a = 0; a < NA; ++a) raw[a] = h2f(b[(size_t)a * nh + i]);
    }
    __device__ __forceinline__ void finish() {
        double sum = 0.0;
#pragma unroll
        for (int a = 0; a < NA; ++a) sum += (double)fmaxf(raw[a], 0.0f);
        const bool uni = !(sum > 0.0);
        const float u = 1.0f / (float)NA;
#pragma unroll
        for (int a = 0; a < NA; ++a) sig[a] = uni ? u : (float)((double)fmaxf(raw[a], 0.0f) / sum);
    }
    __device__ __forceinline__ void load(const uint16_t* __restrict__ b, uint32_t nh, uint32_t i) {
        fetch(b, nh, i);
        finish();
    }
};

template <int NA>
struct HandSigma<NA, kRawF32> {
    float sig[NA];
    __device__ __forceinline__ void fetch(const float* __restrict__ b, uint32_t nh, uint32_t i) {
#pragma unroll
        for (int a = 0; a < NA; ++a) sig[a] = b[(size_t)a * nh + i];
    }
    __device__ __forceinline__ void finish() {}
    __device__ __forceinline__ void load(const float* __restrict__ b, uint32_t nh, uint32_t i) { fetch(b, nh, i); }
};

// Calls f(std::integral_constant<int, NA>) for na in 1..8, f(0) otherwise
// (runtime action count).
template <class F>
__device__ __forceinline__ void dispatch_na(int na, F&& f) {
    switch (na) {
        case 1: f(std::integral_constant<int, 1>{}); break;
        case 2: f(std::integral_constant<int, 2>{}); break;
        case 3: f(std::integral_constant<int, 3>{}); break;
        case 4: f(std::integral_constant<int, 4>{}); break;
        case 5: f(std::integral_constant<int, 5>{}); break;
        case 6: f(std::integral_constant<int, 6>{}); break;
        case 7: f(std::integral_constant<int, 7>{}); break;
        case 8: f(std::integral_constant<int, 8>{}); break;
        default: f(std::integral_constant<int, 0>{}); break;
    }
}

// The same for the up sweep's kernels with at most 4 actions specialised:
// those kernels hold four hands' values per lane, and the register demand of
// the wider cases would lower every launch's occupancy; nodes with more
// actions take the generic loop.
template <class F>
__device__ __forceinline__ void dispatch_na4(int na, F&& f) {
    switch (na) {
        case 1: f(std::integral_constant<int, 1>{}); break;
        case 2: f(std::integral_constant<int, 2>{}); break;
        case 3: f(std::integral_constant<int, 3>{}); break;
        case 4: f(std::integral_constant<int, 4>{}); break;
        default: f(std::integral_constant<int, 0>{}); break;
    }
}

template <int KIND>
struct ArenaOf;
template <>
struct ArenaOf<kCurrentF32> {
    static __device__ const float* get(const DevState& S, int q) { return S.cur[q]; }
};
template <>
struct ArenaOf<kAverageF32> {
    static __device__ const float* get(const DevState& S, int q) { return S.S[q]; }
};
template <>
struct ArenaOf<kCurrentQ> {
    static __device__ const int16_t* get(const DevState& S, int q) { return S.Rq[q]; }
};
template <>
struct ArenaOf<kAverageQ> {
    static __device__ const uint16_t* get(const DevState& S, int q) { return S.Sq[q]; }
};
template <>
struct ArenaOf<kCurrentH> {
    static __device__ const uint16_t* get(const DevState& S, int q) { return (const uint16_t*)S.Rq[q]; }
};
template <>
struct ArenaOf<kAverageH> {
    static __device__ const uint16_t* get(const DevState& S, int q) { return S.Sq[q]; }
};
template <>
struct ArenaOf<kRawF32> {
    static __device__ const float* get(const DevState& S, int q) { return S.S[q]; }
};

// ---------------------------------------------------------------------------
// Launch-shape helpers.
// ---------------------------------------------------------------------------
// The dispatcher hands workgroup b to XCD b % 8; the host pads the grid to a
// multiple of 8 and this gives each XCD a contiguous slice of the list, so
// nodes sharing a river board's tables share that XCD's L2.
__device__ __forceinline__ uint32_t xcd_slice(uint32_t b, uint32_t grid) {
    const uint32_t per = grid >> 3;
    return (b & 7u) * per + (b >> 3);
}

__device__ __forceinline__ uint32_t wave_item() {
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    return xcd_slice(blockIdx.x, gridDim.x) * (blockDim.x >> 6) + wave;
}

// Hands per lane whose loads are issued together in the hand loops (a
// load's first use otherwise waits for it before the next one is issued).
constexpr int kBatch = 4;
// Hands per lane in one step of the up-sweep action kernels: every load
// of all of them is issued before the first store. The f16-row update
// takes three (93 VGPRs, five waves per SIMD: 3.40 ms per flop SRP
// iteration against 3.59 with four); the f32-row updates are bound by
// HBM bandwidth, where three hands (7 waves) change nothing.
constexpr int kStepHands = 4;
constexpr int kStepHands16 = 3;
// With f16 rows and an even hand count the update reads and writes hands in
// pairs (one 32-bit access for two f16 entries; 3.40 -> 3.26 ms per flop
// SRP iteration, bit-identical): pairs per lane per step. The down sweep
// measured the same either way and keeps single hands.
constexpr int kStepPairs = 2;
// Opponent-node sums (no arena traffic, fewer registers) take six.
constexpr int kSumHands = 6;
// The same in the down-sweep action kernel (two: 2.09 ms per f32 iteration,
// three: 1.93, four: 1.96).
constexpr int kDownHands = 3;

// Orders this wavefront's LDS writes before its later LDS reads.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Cross-lane moves on the VALU (DPP) rather than through the LDS crossbar
// (ds_bpermute), which the showdown and fold kernels keep busy. Controls:
// row_shr:n = 0x110 + n (within 16-lane rows), row_bcast:15 = 0x142 and
// row_bcast:31 = 0x143 (the last lane of a row, or of lanes 0-31, into the
// rows that ROWS enables), wave_shr:1 = 0x138, quad_perm = 2 bits per lane.
// dpp_keep: lanes without a source (or in rows not enabled) keep x.
// dpp_zero: they get 0.
template <int CTRL, int ROWS = 0xF>
__device__ __forceinline__ float dpp_keep(float x) {
    const int u = __builtin_bit_cast(int, x);
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(u, u, CTRL, ROWS, 0xF, false));
}
template <int CTRL, int ROWS = 0xF>
__device__ __forceinline__ double dpp_zero(double x) {
    const uint64_t u = __builtin_bit_cast(uint64_t, x);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)u, CTRL, ROWS, 0xF, false);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(u >> 32), CTRL, ROWS, 0xF, false);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}
template <int CTRL, int ROWS = 0xF>
__device__ __forceinline__ float dpp_zero(float x) {
    const int u = __builtin_bit_cast(int, x);
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, u, CTRL, ROWS, 0xF, false));
}
__device__ __forceinline__ double readlane_f64(double x, int l) {
    const uint64_t u = __builtin_bit_cast(uint64_t, x);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, l);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), l);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// Inclusive prefix sum over the wavefront's 64 lanes (all lanes active),
// in f64 or f32.
template <class X>
__device__ __forceinline__ X wave_incl_sum(X x) {
    x += dpp_zero<0x111>(x);
    x += dpp_zero<0x112>(x);
    x += dpp_zero<0x114>(x);
    x += dpp_zero<0x118>(x);
    x += dpp_zero<0x142, 0xA>(x);
    x += dpp_zero<0x143, 0xC>(x);
    return x;
}
__device__ __forceinline__ double wave_incl_sum_f64(double x) {
    x += dpp_zero<0x111>(x);
    x += dpp_zero<0x112>(x);
    x += dpp_zero<0x114>(x);
    x += dpp_zero<0x118>(x);
    x += dpp_zero<0x142, 0xA>(x);
    x += dpp_zero<0x143, 0xC>(x);
    return x;
}

// Sum over the wavefront (all lanes active), the same in every lane.
__device__ __forceinline__ double wave_sum_f64(double v) { return readlane_f64(wave_incl_sum_f64(v), kWave - 1); }
__device__ __forceinline__ double wave_sum(double v) { return wave_sum_f64(v); }
__device__ __forceinline__ float wave_sum(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, wave_incl_sum(v)), kWave - 1));
}

// Max over the wavefront (all lanes active), the same in every lane.
__device__ __forceinline__ float wave_max_f32(float x) {
    x = fmaxf(x, dpp_keep<0x111>(x));
    x = fmaxf(x, dpp_keep<0x112>(x));
    x = fmaxf(x, dpp_keep<0x114>(x));
    x = fmaxf(x, dpp_keep<0x118>(x));
    x = fmaxf(x, dpp_keep<0x142, 0xA>(x));
    x = fmaxf(x, dpp_keep<0x143, 0xC>(x));
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, x), kWave - 1));
}

// ---------------------------------------------------------------------------
// Root: reach = range weights.
// ---------------------------------------------------------------------------
template <bool H16>
__global__ __launch_bounds__(kBlock) void k_root(DevTree T, DevState S, int mask) {
    __shared__ float pmax[kWaves];
    for (int q = 0; q < 2; ++q) {
        if (!((mask >> q) & 1)) continue;
        bool any = false;
        float mx = 0.0f;
        for (uint32_t i = threadIdx.x; i < T.nh[q]; i += kBlock) {
            float w = T.weights[q][i];
            if constexpr (!H16) S.reach[q][i] = w;
            any |= w > 0.0f;
            mx = fmaxf(mx, w);
        }
        mx = wave_max_f32(mx);
        if ((threadIdx.x & 63) == 0) pmax[threadIdx.x >> 6] = mx;
        any = __syncthreads_or(any);
        mx = fmaxf(fmaxf(pmax[0], pmax[1]), fmaxf(pmax[2], pmax[3]));
        if constexpr (H16) {  // the row, encoded with the scale of its maximum
            const RowEnc enc(mx);
            const RowOut<true> out(S.reach[q], 0, enc);
            for (uint32_t i = threadIdx.x; i < T.nh[q]; i += kBlock) out.set(i, T.weights[q][i]);
            if (threadIdx.x == 0) S.rsc[q][0] = enc.sc;
        }
        if (threadIdx.x == 0) {
            S.alive[q][0] = any ? 1 : 0;
            S.rmax[q][0] = mx;
        }
        __syncthreads();
    }
}

// Fill n floats with v (clears device buffers where a memset node must not
// be used: code a caller may capture into a graph).
__global__ __launch_bounds__(kBlock) void k_fill(float* __restrict__ p, uint64_t n, float v) {
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kBlock) p[i] = v;
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
// PART (k_down_node): the wavefront takes hands [hbeg, hend) of the row and,
// instead of writing each child's alive flag and reach maximum, leaves its
// share in part (bit a of part->any: some reach of child a > 0 among its
// hands; part->mx[a]: their maximum) for the workgroup to combine.
struct DownPart {
    uint64_t any;
    float mx[kMaxUnroll];
};
template <int KIND, bool H16 = false, bool PART = false>
__device__ __forceinline__ void down_action_node(const DevTree& T, const DevState& S, const ActRec& r, int q,
                                                 uint32_t lane, uint32_t hbeg = 0, uint32_t hend = 0xFFFFFFFFu,
                                                 DownPart* part = nullptr) {
    static_assert(!(PART && H16), "split rows: f32 reach rows only");
    const uint32_t nh = T.nh[q];
    const uint32_t h1 = min(hend, nh);  // end of this wavefront's hands
    const int na = r.na;
    const uint32_t* kid_own = T.kids + r.kid + na;
    const uint32_t src = pick(r.own, q);
    // f16 rows: the children share one scale, bounded by the parent's
    // maximum (sigma <= 1; a lock's entries by the lock bound)
    const RowEnc enc(H16 ? S.rmax[q][src] * (r.lock_off >= 0 ? T.lock_bound : 1.0f) : 0.0f);
    auto child = [&](int a) { return RowOut<H16>(S.reach[q], (size_t)kid_own[a] * nh, enc); };
    if (!S.alive[q][src]) {
        for (int a = 0; a < na; ++a) {
            const RowOut<H16> dst = child(a);
            for (uint32_t i = hbeg + lane; i < h1; i += kWave) dst.set(i, 0.0f);
        }
        if (PART) {  // every wavefront of the node sees the same flag
            if (lane == 0) part->any = 0;
            if (lane < kMaxUnroll) part->mx[lane] = 0.0f;
            return;
        }
        for (int a = lane; a < na; a += kWave) {
            S.alive[q][kid_own[a]] = 0;
            S.rmax[q][kid_own[a]] = 0.0f;
            if constexpr (H16) S.rsc[q][kid_own[a]] = enc.sc;
        }
        return;
    }
    const RowIn<H16> pr = reach_in<H16>(T, S, q, src);
    uint64_t anym = 0;
    // per-child reach maxima (the f16 scale bounds); the first kMaxUnroll
    // children in registers, the rest from memory below
    float mmax[kMaxUnroll];
#pragma unroll
    for (int a = 0; a < kMaxUnroll; ++a) mmax[a] = 0.0f;
    if (r.lock_off >= 0) {
        const float* sig = T.lock_sigma + r.lock_off;
        for (uint32_t i = hbeg + lane; i < h1; i += kWave) {
            const float pv = pr[i];
            for (int a = 0; a < na; ++a) {
                float v = pv * sig[(size_t)a * nh + i];
                child(a).set(i, v);
                anym |= (uint64_t)(v > 0.0f) << a;
            }
        }
    } else {
        const auto* arena = ArenaOf<KIND>::get(S, q) + r.doff;
        dispatch_na(na, [&](auto NAc) {
            constexpr int NA = decltype(NAc)::value;
            if constexpr (NA > 0) {
                RowOut<H16> dst[NA];
#pragma unroll
                for (int a = 0; a < NA; ++a) dst[a] = child(a);
                for (uint32_t i0 = hbeg + lane; i0 < h1; i0 += kDownHands * kWave) {
                    HandSigma<NA, KIND> hs[kDownHands];
                    float pv[kDownHands];
#pragma unroll
                    for (int h = 0; h < kDownHands; ++h) {
                        const uint32_t ii = min(i0 + h * kWave, nh - 1);
                        hs[h].fetch(arena, nh, ii);
                        pv[h] = pr[ii];
                    }
#pragma unroll
                    for (int h = 0; h < kDownHands; ++h) {
                        const uint32_t i = i0 + h * kWave;
                        if (i >= h1) break;
                        hs[h].finish();
#pragma unroll
                        for (int a = 0; a < NA; ++a) {
                            float v = pv[h] * hs[h].sig[a];
                            dst[a].set(i, v);
                            anym |= (uint64_t)(v > 0.0f) << a;
                            mmax[a] = fmaxf(mmax[a], v);
                        }
                    }
                }
            } else {
                for (uint32_t i = hbeg + lane; i < h1; i += kWave) {
                    SigmaHand<KIND> sh;
                    sh.init(arena, nh, i, na);
                    const float pv = pr[i];
                    for (int a = 0; a < na; ++a) {
                        float v = pv * sh.template sigma<false>(a);
                        child(a).set(i, v);
                        anym |= (uint64_t)(v > 0.0f) << a;
                    }
                }
            }
        });
    }
    const bool fast = r.lock_off < 0 && na <= kMaxUnroll;  // mmax filled in the loops above
    if constexpr (PART) {
        // na <= kMaxUnroll (the launch checks it); locked nodes read back
        // their own hands' writes
        uint64_t anyw = 0;
        for (int a = 0; a < na; ++a) {
            if (__ballot((anym >> a) & 1ull) != 0ull) anyw |= 1ull << a;
            float mx = 0.0f;
            if (fast) {
#pragma unroll
                for (int a2 = 0; a2 < kMaxUnroll; ++a2) mx = a2 == a ? mmax[a2] : mx;
            } else {
                __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
                const RowIn<H16> row(S.reach[q], (size_t)kid_own[a] * nh, enc.sc);
                for (uint32_t i = hbeg + lane; i < h1; i += kWave) mx = fmaxf(mx, row[i]);
            }
            mx = wave_max_f32(mx);
            if (lane == 0) part->mx[a] = mx;
        }
        if (lane == 0) part->any = anyw;
        return;
    }
    for (int a = 0; a < na; ++a) {
        const bool any = __ballot((anym >> a) & 1ull) != 0ull;
        float mx = 0.0f;
        if (fast) {
#pragma unroll
            for (int a2 = 0; a2 < kMaxUnroll; ++a2) mx = a2 == a ? mmax[a2] : mx;
        } else {
            // this wavefront's own writes above: read back in order
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
            const RowIn<H16> row(S.reach[q], (size_t)kid_own[a] * nh, enc.sc);
            for (uint32_t i = hbeg + lane; i < h1; i += kWave) mx = fmaxf(mx, row[i]);
        }
        mx = wave_max_f32(mx);
        if (lane == 0) {
            S.alive[q][kid_own[a]] = any ? 1 : 0;
            S.rmax[q][kid_own[a]] = mx;
            if constexpr (H16) S.rsc[q][kid_own[a]] = enc.sc;
        }
    }
}

// ---------------------------------------------------------------------------
// Down sweep, chance edges: child reach = parent reach with the hands that
// hold the dealt card zeroed (cfr.rs:615-629), for the players in `mask`.
// ---------------------------------------------------------------------------
template <bool H16 = false>
__device__ __forceinline__ void down_edge(const DevTree& T, const DevState& S, const EdgeRec& r, int mask,
                                          uint32_t lane) {
    const uint64_t cm = 1ull << r.card;
    for (int q = 0; q < 2; ++q) {
        if (!((mask >> q) & 1)) continue;
        const uint32_t nh = T.nh[q];
        const bool alive = S.alive[q][pick(r.src, q)];
        // f16 rows: the child keeps the parent's scale (its entries are a
        // subset), so entries are copied as stored
        const RowIn<H16> in = reach_in<H16>(T, S, q, pick(r.src, q));
        RowEnc enc(0.0f);
        if constexpr (H16) enc.sc = in.sc, enc.inv = 1.0f / in.sc;
        const RowOut<H16> out(S.reach[q], (size_t)pick(r.dst, q) * nh, enc);
        bool any = false;
        float mmx = 0.0f;
        for (uint32_t i0 = lane; i0 < nh; i0 += kBatch * kWave) {
            float v[kBatch];
            uint64_t m[kBatch];
#pragma unroll
            for (int u = 0; u < kBatch; ++u) {
                const uint32_t ii = min(i0 + u * kWave, nh - 1);
                v[u] = in[ii];
                m[u] = T.hmask[q][ii];
            }
#pragma unroll
            for (int u = 0; u < kBatch; ++u) {
                const uint32_t i = i0 + u * kWave;
                if (i >= nh) break;
                const float x = (!alive || (m[u] & cm)) ? 0.0f : v[u];
                out.set(i, x);  // f16: x * inv is the stored entry, exactly
                any |= x > 0.0f;
                mmx = fmaxf(mmx, x);
            }
        }
        any = __ballot(any) != 0ull;
        mmx = wave_max_f32(mmx);
        if (lane == 0) {
            S.alive[q][pick(r.dst, q)] = any ? 1 : 0;
            S.rmax[q][pick(r.dst, q)] = mmx;
            if constexpr (H16) S.rsc[q][pick(r.dst, q)] = enc.sc;
        }
    }
}

template <bool H16 = false>
__global__ __launch_bounds__(kBlock) void k_down_chance(DevTree T, DevState S, const EdgeRec* __restrict__ recs,
                                                        uint32_t count, int mask) {
    const uint32_t w = wave_item();
    if (w >= count) return;
    down_edge<H16>(T, S, recs[w], mask, threadIdx.x & 63);
}

// A level of the down sweep in one launch: actor q's action nodes and the
// chance edges below the level's chance nodes (independent: each reads the
// level's own rows and writes its children's), the two lists interleaved
// in proportion over the grid as in k_up_action2.
template <int KIND, bool H16 = false>
__global__ __launch_bounds__(kBlock) void k_down_level(DevTree T, DevState S, const ActRec* __restrict__ acts,
                                                       uint32_t na, int q, const EdgeRec* __restrict__ edges,
                                                       uint32_t ne, int mask) {
    const uint32_t m = wave_item();
    const uint32_t tot = na + ne;
    if (m >= tot) return;
    const uint32_t k0 = (uint32_t)(((uint64_t)m * na) / tot);
    const uint32_t k1 = (uint32_t)(((uint64_t)(m + 1) * na) / tot);
    const uint32_t lane = threadIdx.x & 63;
    if (__builtin_amdgcn_readfirstlane(k1 > k0 ? 1u : 0u)) down_action_node<KIND, H16>(T, S, acts[k0], q, lane);
    else down_edge<H16>(T, S, edges[m - k0], mask, lane);
}

// A level of few action nodes, f32 reach rows, at most kMaxUnroll actions:
// one workgroup per node, each wavefront a quarter of the hands
// (64-aligned), so the level fills more of the GPU; the children's alive
// flags and reach maxima are combined in LDS. Every hand's arithmetic is
// down_action_node's. The level's chance edges follow, a wavefront each
// (workgroups na.., four edges each).
template <int KIND>
__global__ __launch_bounds__(kBlock) void k_down_node(DevTree T, DevState S, const ActRec* __restrict__ acts,
                                                      uint32_t na, int q, const EdgeRec* __restrict__ edges,
                                                      uint32_t ne, int mask) {
    __shared__ DownPart part[kWaves];
    const uint32_t n = xcd_slice(blockIdx.x, gridDim.x);
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (n >= na) {  // the whole workgroup: edges
        const uint32_t e = (n - na) * kWaves + wave;
        if (e < ne) down_edge<false>(T, S, edges[e], mask, lane);
        return;
    }
    const uint32_t ch = ((T.nh[q] + kWaves - 1) / kWaves + kWave - 1) & ~(uint32_t)(kWave - 1);
    const ActRec r = acts[n];
    down_action_node<KIND, false, true>(T, S, r, q, lane, wave * ch, (wave + 1) * ch, &part[wave]);
    __syncthreads();
    if (wave == 0 && lane < (uint32_t)r.na) {
        const uint32_t a = lane;
        const uint32_t kid = T.kids[r.kid + r.na + a];
        bool any = false;
        float mx = 0.0f;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) {
            any = any || ((part[w].any >> a) & 1ull);
            mx = fmaxf(mx, part[w].mx[a]);
        }
        S.alive[q][kid] = any ? 1 : 0;
        S.rmax[q][kid] = mx;
    }
}

// Exactness test for the f64 sums of the fold and showdown terminals. The
// nonzero entries of an opponent reach row are all multiples of u, the
// smallest of their lowest set bits; so is every sum the terminals form
// from them (totals, prefixes, per-card sums, and the differences of
// those), whose magnitude is at most S, the row's total (entries are >= 0).
// When S < 2^53 u each of these is exact in f64 in any order of addition,
// and the GPU's chunked sums equal the CPU engine's sequential ones bit for
// bit. S is bounded by the f64 total as computed (its relative error is far
// below 2^-40). lsb_exp: the exponent of x's lowest set bit (x > 0,
// finite).
__device__ __forceinline__ int lsb_exp(float x) {
    const uint32_t b = __builtin_bit_cast(uint32_t, x);
    const uint32_t e = (b >> 23) & 0xFFu, m = e ? (b & 0x7FFFFFu) | 0x800000u : b & 0x7FFFFFu;
    return (int)(e ? e : 1u) - 150 + __builtin_ctz(m);
}
__device__ __forceinline__ bool exact_in_any_order(double total, int lsb_min) {
    return total * (1.0 + 0x1p-40) < ldexp(1.0, 53 + lsb_min);
}

// t + x(lane 0) + x(lane 1) + ... + x(lane m - 1), one addition at a time
// (the same in every lane)
__device__ __forceinline__ double add_lanes_in_order(double t, float x, uint32_t m) {
    const int xi = __builtin_bit_cast(int, x);
    for (uint32_t l = 0; l < m; ++l) t += (double)__builtin_bit_cast(float, __builtin_amdgcn_readlane(xi, (int)l));
    return t;
}

// ---------------------------------------------------------------------------
// Up sweep, fold terminals (game.rs:470-496), all levels in one launch:
// value = amount * (opponent reach compatible with the hand), with f64
// totals (a wavefront sum; under SEQ the CPU's sum in hand order for a row
// that fails the exactness test above). One wavefront takes `per` (at most
// kFoldNodes; fewer on small trees, so they still fill the GPU)
// consecutive nodes. The opponent's row is scattered into a symmetric
// 52 x 52 card matrix in LDS, M[a][b] = M[b][a] = reach of the hand {a, b}
// (0 where the opponent has no such hand, and on the diagonal; those zeros
// are written once per wavefront, since every node writes the same hand
// set). Card c's sum is row c read
// in column order, i.e. its hands in ascending order of the other card,
// which is hand (combo) order: the order the CPU accumulates them in
// (game.rs:478-485). Rows are 52 floats, so lane c's 16-byte reads of row
// c fall in distinct banks. Each hand then reads its two card sums and
// its own combo's entry. LDS per wavefront: 52 x 52 floats + 52 doubles.
// ---------------------------------------------------------------------------
constexpr int kMaxH = 21;         // hand rows per lane: 21 * 64 >= 1326
constexpr int kRowPer = (1326 + kBlock - 1) / kBlock;  // hand rows per thread of a workgroup
constexpr uint32_t kFoldNodes = 16;                     // fold nodes per wavefront, at most
constexpr size_t kFoldWaveBytes = 52 * 52 * 4 + 52 * 8;  // card matrix + per-card sums

// combo_index (cards.rs:100-103) of two distinct cards, hi > lo
// Reach sums of the fold and showdown terminals: f64 over f32 reach rows
// (the CPU engine's sums, game.rs:477-484, 517-574); f32 over f16 reach rows, whose own rounding
// (2^-11 of each entry) is far above f32's accumulation error, at half the
// LDS bytes of the showdown's random prefix reads.
template <bool R16>
struct ShowAcc {
    using T = double;
};
template <>
struct ShowAcc<true> {
    using T = float;
};

__device__ __forceinline__ uint32_t combo_of(uint32_t hi, uint32_t lo) { return hi * (hi - 1) / 2 + lo; }

template <int ROWS = 0, bool SEQ = false>
__global__ __launch_bounds__(kBlock) void k_up_fold(DevTree T, DevState S, const FoldRec* __restrict__ recs,
                                                    uint32_t count, int p, int swap, uint32_t per) {
    constexpr bool R16 = (ROWS & kRows16Reach) != 0, V16 = (ROWS & kRows16Value) != 0;
    static_assert(!(SEQ && R16), "sequential totals are f64 sums of f32 reach rows");
    // totals and card sums in f64 (the CPU's, game.rs:477-484), f32 over
    // f16 reach rows (as the showdown's prefix sums)
    using Acc = typename ShowAcc<R16>::T;
    extern __shared__ double lds[];
    const int o = 1 - p;
    const uint32_t nho = T.nh[o], nhp = T.nh[p];
    const uint32_t first = wave_item() * per;
    if (first >= count) return;
    const uint32_t last = min(first + per, count);
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t wave = threadIdx.x >> 6;
    char* base = (char*)lds + wave * kFoldWaveBytes;
    float* M = (float*)base;
    Acc* s_card = (Acc*)(base + 52 * 52 * 4);
    // packed hand words (c1 | c2 << 8 | ...) of both players: static for the
    // wavefront's nodes; the opponent's as the two matrix positions of the
    // hand (a * 52 + b | (b * 52 + a) << 16)
    const uint32_t* __restrict__ hpp = T.hpack[p];
    const uint32_t* __restrict__ hpo = T.hpack[o];
    uint32_t mo_pos[kMaxH], hw[kMaxH];
#pragma unroll
    for (int u = 0; u < kMaxH; ++u)
        if (u * kWave < (int)nho) {
            const uint32_t h = hpo[min(lane + u * kWave, nho - 1)];
            const uint32_t a = h & 0xFFu, b = (h >> 8) & 0xFFu;
            mo_pos[u] = (a * 52 + b) | ((b * 52 + a) << 16);
        }
#pragma unroll
    for (int u = 0; u < kMaxH; ++u)
        if (u * kWave < (int)nhp) hw[u] = hpp[min(lane + u * kWave, nhp - 1)];
    {
        float4* m4 = (float4*)M;
        for (uint32_t k = lane; k < 52 * 52 / 4; k += kWave) m4[k] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    typename RowIn<R16>::Raw rh[kMaxH];
    float rsc;  // rh's decode scale
    {
        const RowIn<R16> ro = reach_in<R16>(T, S, o, pick(recs[first].own, o));
        rsc = ro.sc;
#pragma unroll
        for (int u = 0; u < kMaxH; ++u)
            if (u * kWave < (int)nho) rh[u] = ro.raw(min(lane + u * kWave, nho - 1));
    }
#pragma unroll 1
    for (uint32_t n = first; n < last; ++n) {
        const FoldRec r = recs[n];
        const bool live = S.alive[o][pick(r.own, o)] != 0;
        Acc part = 0;
        uint32_t emin = 255;  // SEQ: smallest exponent field of the row's nonzero entries (subnormals: 1)
        if (live) {
            wave_sync();  // the previous node's reads of M and s_card are done
#pragma unroll
            for (int u = 0; u < kMaxH; ++u)
                if (lane + u * kWave < nho) {
                    const float x = RowIn<R16>::dec(rh[u], rsc);
                    M[mo_pos[u] & 0xFFFFu] = x;
                    M[mo_pos[u] >> 16] = x;
                    part += (Acc)x;
                    if constexpr (SEQ) {
                        const uint32_t b = __builtin_bit_cast(uint32_t, x);
                        emin = min(emin, b ? max((b >> 23) & 0xFFu, 1u) : 255u);
                    }
                }
        }
        // the next node's row, in flight during the rest of this node
        if (n + 1 < last) {
            const RowIn<R16> ro = reach_in<R16>(T, S, o, pick(recs[n + 1].own, o));
            rsc = ro.sc;
#pragma unroll
            for (int u = 0; u < kMaxH; ++u)
                if (u * kWave < (int)nho) rh[u] = ro.raw(min(lane + u * kWave, nho - 1));
        }
        if (!live) {
            const RowOut<V16> out(S.cfv, (size_t)r.cfv * T.nh_max, RowEnc(0.0f));
            for (uint32_t i = lane; i < nhp; i += kWave) out.set(i, 0.0f);
            if (S.vmax && lane == 0) S.vmax[r.cfv] = 0.0f;
            if (V16 && lane == 0) S.vsc[r.cfv] = 1.0f;
            continue;
        }
        Acc tot = wave_sum(part);
        wave_sync();
        if constexpr (SEQ) {
            // the CPU engine's total (hand order, one addition at a time)
            // where the wavefront's sum could round differently; every
            // entry's lowest set bit is at least 2^(exponent field - 150)
            const int lmin = (int)-wave_max_f32(-(float)emin) - 150;
            if (!exact_in_any_order(tot, lmin)) {
                double t = 0.0;
#pragma unroll
                for (int u = 0; u < kMaxH; ++u) {
                    if (u * kWave >= (int)nho) break;
                    const float x = lane + u * kWave < nho ? M[mo_pos[u] & 0xFFFFu] : 0.0f;
                    t = add_lanes_in_order(t, x, min(nho - u * kWave, (uint32_t)kWave));
                }
                tot = t;
            }
        }
        if (lane < 52) {
            const float4* row = (const float4*)(M + lane * 52);
            Acc acc = 0;
#pragma unroll
            for (int t0 = 0; t0 < 13; t0 += 4) {
                float4 v[4];
#pragma unroll
                for (int t = 0; t < 4; ++t)
                    if (t0 + t < 13) v[t] = row[t0 + t];
#pragma unroll
                for (int t = 0; t < 4; ++t)
                    if (t0 + t < 13) {
                        acc += (Acc)v[t].x;
                        acc += (Acc)v[t].y;
                        acc += (Acc)v[t].z;
                        acc += (Acc)v[t].w;
                    }
            }
            s_card[lane] = acc;
        }
        wave_sync();
        // swap: the opponent's payoff (win and lose exchanged), for game-value
        // gradients with respect to the opponent's range and strategy
        const float amount = (r.folder == (uint32_t)p) != (swap != 0) ? r.t_lose : r.t_win;
        // f16 rows: |value| <= |amount| x the opponent's total reach
        const RowEnc enc(V16 ? fabsf(amount) * (float)tot : 0.0f);
        const RowOut<V16> out(S.cfv, (size_t)r.cfv * T.nh_max, enc);
        if (V16 && lane == 0) S.vsc[r.cfv] = enc.sc;
        float vm = 0.0f;
        // a rolled loop (the unrolled form, with its early exit removed,
        // holds every hand's loads at once and runs slower)
        for (int u = 0; u < kMaxH; ++u) {
            const uint32_t i = lane + u * kWave;
            if (u * kWave >= (int)nhp) break;
            if (i < nhp) {
                const uint32_t h = hw[u];
                const uint32_t a1 = h & 0xFFu, a2 = (h >> 8) & 0xFFu;
                const float same_r = M[a1 * 52 + a2];
                const float valid = (float)(tot - s_card[a1] - s_card[a2]) + same_r;
                out.set(i, amount * valid);
                vm = fmaxf(vm, fabsf(amount * valid));
            }
        }
        if (S.vmax) {
            vm = wave_max_f32(vm);
            if (lane == 0) S.vmax[r.cfv] = vm;
        }
    }
}

// One quad lane's turn in a card list's sequential sum (below): lane R of
// each quad adds its kSeg entries to the running sum y and stores the
// prefixes, then y is passed to the whole quad.
template <int R, bool DESC, int kSeg>
__device__ __forceinline__ double card_list_turn(double y, uint32_t seg, const float (&x)[kSeg],
                                                 const uint32_t (&cw)[kSeg], double* cpre) {
    if (seg == (uint32_t)R) {
        if constexpr (!DESC) {
#pragma unroll
            for (int u = 0; u < kSeg; ++u) {
                cpre[cw[u] >> 16] = y;
                y += (double)x[u];
            }
        } else {
#pragma unroll
            for (int u = kSeg - 1; u >= 0; --u) {
                y += (double)x[u];
                cpre[cw[u] >> 16] = y;
            }
        }
    }
    return dpp_zero<R * 0x55>(y);  // quad_perm: lane R into all four
}

// The showdown sums in the CPU engine's own order (game.rs:513-574), for a
// node whose reach row fails the exactness test above. Ascending (DESC
// false): P[k] = rs[0] + rs[1] + ... + rs[k-1] added from the bottom, P[mo]
// the total, and per card c of the board the same over the reach of the
// opponent hands holding it (cpre[coff[c] + c + k], the list's total at k =
// its length); descending (DESC true): P[k] = rs[mo-1] + ... + rs[k] added
// from the top, per card likewise (P[mo] and the list totals are left as
// they are: the positions past each end read as 0). Each card list's four
// quad lanes take turns over their quarters (segments as in the showdown's
// card prefixes: cw, ctot = the list total's slot); thread kChainThread,
// past the card lanes, adds up P. The caller's barrier publishes the
// sums.
constexpr uint32_t kChainThread = 4 * 52;  // the first thread without a card list
template <bool DESC, int kSeg>
__device__ __forceinline__ void cpu_order_sums(uint32_t wave, uint32_t lane, uint32_t card, uint32_t seg,
                                               uint32_t mo, const float* rs, double* P, double* cpre,
                                               const uint32_t (&cw)[kSeg], uint32_t ctot) {
    if (card < 52) {
        float x[kSeg];
#pragma unroll
        for (int u = 0; u < kSeg; ++u) x[u] = rs[cw[u] & 0xFFFFu];  // past the list: the 0 in rs[mo]
        double y = 0.0;
        if constexpr (!DESC) {
            y = card_list_turn<0, false>(y, seg, x, cw, cpre);
            y = card_list_turn<1, false>(y, seg, x, cw, cpre);
            y = card_list_turn<2, false>(y, seg, x, cw, cpre);
            y = card_list_turn<3, false>(y, seg, x, cw, cpre);
            if (seg == 3) cpre[ctot] = y;
        } else {
            y = card_list_turn<3, true>(y, seg, x, cw, cpre);
            y = card_list_turn<2, true>(y, seg, x, cw, cpre);
            y = card_list_turn<1, true>(y, seg, x, cw, cpre);
            (void)card_list_turn<0, true>(y, seg, x, cw, cpre);
        }
    }
    if (wave * kWave + lane == kChainThread) {
        // one lane (of the wavefront with the fewest card lists) runs P's
        // chain through LDS, reading kB entries ahead of the additions
        constexpr uint32_t kB = 16;
        double t = 0.0;
        if constexpr (!DESC) {
            uint32_t k = 0;
            for (; k + kB <= mo; k += kB) {
                float x[kB];
#pragma unroll
                for (uint32_t i = 0; i < kB; ++i) x[i] = rs[k + i];
#pragma unroll
                for (uint32_t i = 0; i < kB; ++i) {
                    P[k + i] = t;
                    t += (double)x[i];
                }
            }
            for (; k < mo; ++k) {
                P[k] = t;
                t += (double)rs[k];
            }
            P[mo] = t;
        } else {
            uint32_t k = mo;
            for (; k >= kB; k -= kB) {
                float x[kB];
#pragma unroll
                for (uint32_t i = 0; i < kB; ++i) x[i] = rs[k - 1 - i];
#pragma unroll
                for (uint32_t i = 0; i < kB; ++i) {
                    t += (double)x[i];
                    P[k - 1 - i] = t;
                }
            }
            for (; k > 0; --k) {
                t += (double)rs[k - 1];
                P[k - 1] = t;
            }
        }
    }
}

// Phases 2-4 of the showdown (k_up_show) in the CPU engine's order
// (game.rs:513-574) for a reach row that fails the exactness test: the
// ascending sums give each hand's L (staged in O), then the descending sums
// its H, and V from the ascending totals, which the descending pass leaves
// in place (P[mo], each card list's total). Returns the thread's max
// |value|. All threads of the workgroup call it.
template <int kR, int kSeg>
__device__ __forceinline__ float show_in_cpu_order(uint32_t tid, uint32_t wave, uint32_t lane, uint32_t card,
                                                   uint32_t seg, uint32_t mo, uint32_t mp, const float* rs, double* P,
                                                   double* cpre, float* O, const uint32_t (&cw)[kSeg], uint32_t ctot,
                                                   const uint4 (&ent)[kR], const uint32_t (&pidx)[kR], float win,
                                                   float lose, float tie) {
    cpu_order_sums<false>(wave, lane, card, seg, mo, rs, P, cpre, cw, ctot);
    __syncthreads();
    const double total = P[mo];
#pragma unroll
    for (int u = 0; u < kR; ++u) {
        const uint32_t t = tid + u * kBlock;
        if (t >= mp) break;
        const uint4 e = ent[u];
        const double* cp1 = cpre + (e.y & 0xFFFFu);
        const double* cp2 = cpre + (e.y >> 16);
        O[pidx[u]] = (float)(P[e.x & 0xFFFFu] - cp1[e.z & 0xFFu] - cp2[(e.z >> 8) & 0xFFu]);
    }
    __syncthreads();
    cpu_order_sums<true>(wave, lane, card, seg, mo, rs, P, cpre, cw, ctot);
    __syncthreads();
    float vm = 0.0f;
#pragma unroll
    for (int u = 0; u < kR; ++u) {
        const uint32_t t = tid + u * kBlock;
        if (t >= mp) break;
        const uint4 e = ent[u];
        const uint32_t ub = e.x >> 16, n1 = e.w & 0xFFu, n2 = (e.w >> 8) & 0xFFu;
        const uint32_t h1 = (e.z >> 16) & 0xFFu, h2 = e.z >> 24;
        const double* cp1 = cpre + (e.y & 0xFFFFu);
        const double* cp2 = cpre + (e.y >> 16);
        const float lower = O[pidx[u]];
        const float higher = (float)((ub < mo ? P[ub] : 0.0) - (h1 < n1 ? cp1[h1] : 0.0) - (h2 < n2 ? cp2[h2] : 0.0));
        const uint32_t sp = e.w >> 16;
        const float same_r = sp != 0xFFFFu ? rs[sp] : 0.0f;
        const float valid = (float)(total - cp1[n1] - cp2[n2]) + same_r;
        const float tied = valid - lower - higher;
        const float val = win * lower + lose * higher + tie * tied;
        O[pidx[u]] = val;
        vm = fmaxf(vm, fabsf(val));
    }
    return vm;
}

// ---------------------------------------------------------------------------
// Up sweep, showdown terminals (game.rs:501-577), all levels in one launch:
// value = win * L + lose * H + tie * (V - L - H) with L/H/V the opponent
// reach strictly weaker / strictly stronger / compatible, blockers removed.
// One workgroup per run of nodes sharing a river board: the board's static
// tables are read once per run (into registers; the per-card offsets into
// LDS), then per node (all 256 lanes on every phase):
//   1. the opponent's reach row (hand order) into LDS; the next node's row
//      is loaded meanwhile,
//   2. reach in strength order + its exclusive prefix P (f64, f32 over f16
//      reach rows; per-thread chunks and a workgroup scan),
//   3. per card, the exclusive prefix of the reach of the opponent hands
//      holding it, in strength order (as P; four lanes per card, each a
//      quarter of the card's list, read from the strength-order row at
//      positions held in registers for the run),
//   4. per hand (the traverser's hands in strength order), O(1) lookups:
//      L = P[lb] - (card prefixes at lb),
//      H = (total - P[ub]) - (card totals - card prefixes at ub),
//      V = total - card totals + identical combo; the row is assembled in
//      LDS in hand order and written out after the next node's first
//      barrier.
// A reach row that fails the exactness test (f64 sums only) takes the CPU's
// order for phases 2-4 instead: its ascending sums give L and V, then its
// descending ones H (cpu_order_sums).
// ---------------------------------------------------------------------------
constexpr int kMaxCardList = 48;  // opponent hands holding one card on a river board: <= 46
constexpr int kShortSeg = 9;      // card-list entries per lane when no list exceeds 36 (partial ranges)
// kR: hand rows per thread, kR * kBlock >= either player's hand count (also
// the strength-order chunk per thread in phase 2); kSeg: entries per lane
// of a card's list (four lanes per card), 4 kSeg >= the longest list

template <int kR, int kSeg, int ROWS = 0, bool SEQ = false>
__global__ __launch_bounds__(kBlock) void k_up_show(DevTree T, DevState S, const ShowRec* __restrict__ recs,
                                                    const uint4* __restrict__ chunks, uint32_t n_chunks, int p,
                                                    int swap) {
    constexpr bool R16 = (ROWS & kRows16Reach) != 0, V16 = (ROWS & kRows16Value) != 0;
    static_assert(!(SEQ && R16), "sequential sums are f64 sums of f32 reach rows");
    using Acc = typename ShowAcc<R16>::T;
    extern __shared__ double lds[];
    const uint32_t c = xcd_slice(blockIdx.x, gridDim.x);
    if (c >= n_chunks) return;
    const uint4 ck = chunks[c];
    const uint32_t b = ck.x, first = ck.y, cnt = ck.z;
    const int o = 1 - p;
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t nhp = T.nh[p], nho = T.nh[o];
    const uint32_t mo = T.sd_cnt[o][b], mp = T.sd_cnt[p][b];
    const ShowLds L = show_lds_layout(mo, nhp, nho);
    char* base = (char*)lds;
    Acc* P = (Acc*)(base + L.P);
    Acc* cpre = (Acc*)(base + L.cpre);
    Acc* wsum = (Acc*)(base + L.wsum);         // per wave: prefix scan total
    float* wmax = (float*)(wsum + kWaves);      // per wave: max |value| of the row
    float* rs = (float*)(base + L.rs);
    float* rhs = (float*)(base + L.rhs);
    uint32_t* coff = (uint32_t*)(base + L.coff);
    ShowNode* nodes = (ShowNode*)(base + L.nodes);
    float* O = (float*)(base + L.O);
    float* exr = (float*)(base + L.exr);        // per wave: smallest lowest set bit of the reach (SEQ)

    // the run's records and alive flags, then the board statics
    if (tid < cnt) {
        const ShowRec rr = recs[first + tid];
        const uint32_t os = pick(rr.own, o);
        ShowNode nd;
        nd.cfv = rr.cfv;
        nd.oslot = os;
        nd.live = (S.alive[o][os] && mo != 0) ? 1u : 0u;
        nd.win = swap ? rr.lose : rr.win;
        nd.lose = swap ? rr.win : rr.lose;
        nd.tie = rr.tie;
        nodes[tid] = nd;
    }
    // static for the run, in registers: this thread's sorted traverser hands
    // (table entry, hand index) and its strength-order chunk's opponent hands
    const uint32_t cs = (mo + kBlock - 1) / kBlock;
    const uint32_t lo = min(tid * cs, mo), hi = min(lo + cs, mo);
    uint4 ent[kR];
    uint32_t pidx[kR], oidx[kR];
    {
        const uint32_t sp = T.sd_off[p][b];
        const uint4* __restrict__ tab = T.sd_tab[p] + sp;
        const uint16_t* __restrict__ gp = T.sd_idx[p] + sp;
        const uint32_t mlast = mp > 0 ? mp - 1 : 0;
#pragma unroll
        for (int u = 0; u < kR; ++u) {
            const uint32_t t = min(tid + u * kBlock, mlast);
            ent[u] = mp > 0 ? tab[t] : make_uint4(0u, 0u, 0u, 0u);
            pidx[u] = mp > 0 ? gp[t] : 0u;
        }
        const uint32_t so = T.sd_off[o][b];
        const uint16_t* __restrict__ gi = T.sd_idx[o] + so;
        const uint32_t olast = mo > 0 ? mo - 1 : 0;
#pragma unroll
        for (int u = 0; u < kR; ++u) oidx[u] = mo > 0 ? gi[min(lo + u, olast)] : 0u;
        // hands not on the board keep 0 in the staged value row
        for (uint32_t i = tid; i < nhp; i += kBlock) O[i] = 0.0f;
        if (tid == 0) rs[mo] = 0.0f;
        if (tid < 53) coff[tid] = T.sd_coff[o][(size_t)b * 53 + tid];
    }
    __syncthreads();
    // reach row of the first node (as stored, and its decode scale)
    typename RowIn<R16>::Raw rh[kR];
    float rsc;
    {
        const RowIn<R16> ro = reach_in<R16>(T, S, o, nodes[0].oslot);
        rsc = ro.sc;
#pragma unroll
        for (int u = 0; u < kR; ++u) rh[u] = ro.raw(min(tid + u * kBlock, nho - 1));
    }
    // per-card segment of this thread (cards 0..51, four lanes each)
    const uint32_t card = tid >> 2, seg = tid & 3;
    uint32_t c0 = 0, len = 0, s0 = 0, sn = 0;
    if (card < 52) {
        c0 = coff[card];
        len = coff[card + 1] - c0;
        s0 = min(seg * (uint32_t)kSeg, len);
        sn = min((uint32_t)kSeg, len - s0);
    }
    // this thread's segment of its card's list, static for the run: the
    // strength-order position of each entry and the cpre slot of its prefix
    // (position | slot << 16); slots past the segment read the zero in rs's
    // spare slot and store to cpre's
    uint32_t cw[kSeg];
    {
        const uint16_t* __restrict__ gpos = T.sd_cpos[o] + 2 * (size_t)T.sd_off[o][b];
        const uint32_t last = mo > 0 ? 2 * mo - 1 : 0;
#pragma unroll
        for (int u = 0; u < (int)kSeg; ++u)
            cw[u] = (uint32_t)u < sn ? gpos[min(c0 + s0 + u, last)] | ((c0 + card + s0 + u) << 16)
                                     : mo | (L.creg << 16);
    }

    // the last evaluated node's row, staged in O, waiting for a barrier
    bool pending = false;
    uint32_t pend_cfv = 0;
    auto flush = [&]() {
        const float m = S.vmax ? fmaxf(fmaxf(wmax[0], wmax[1]), fmaxf(wmax[2], wmax[3])) : 0.0f;
        const RowEnc enc(V16 ? m : 0.0f);  // f16 value rows: the scale of the row's maximum
        const RowOut<V16> po(S.cfv, (size_t)pend_cfv * T.nh_max, enc);
        for (uint32_t i = tid; i < nhp; i += kBlock) po.set(i, O[i]);
        if (S.vmax && tid == 0) S.vmax[pend_cfv] = m;
        if (V16 && tid == 0) S.vsc[pend_cfv] = enc.sc;
        pending = false;
    };
#pragma unroll 1
    for (uint32_t n = 0; n < cnt; ++n) {
        const ShowNode nd = nodes[n];
        const bool live = nd.live != 0;
        if (live) {
#pragma unroll
            for (int u = 0; u < kR; ++u)
                if (tid + u * kBlock < nho) rhs[tid + u * kBlock] = RowIn<R16>::dec(rh[u], rsc);
        }
        // 1. the next node's reach row, in flight during this node
        if (n + 1 < cnt) {
            const RowIn<R16> ro = reach_in<R16>(T, S, o, nodes[n + 1].oslot);
            rsc = ro.sc;
#pragma unroll
            for (int u = 0; u < kR; ++u) rh[u] = ro.raw(min(tid + u * kBlock, nho - 1));
        }
        if (!live) {  // no reach, or no opponent hand on this board: all 0
            if (pending) {
                __syncthreads();
                flush();
            }
            const RowOut<V16> out(S.cfv, (size_t)nd.cfv * T.nh_max, RowEnc(0.0f));
            for (uint32_t i = tid; i < nhp; i += kBlock) out.set(i, 0.0f);
            if (S.vmax && tid == 0) S.vmax[nd.cfv] = 0.0f;
            if (V16 && tid == 0) S.vsc[nd.cfv] = 1.0f;
            continue;
        }
        __syncthreads();
        // the previous node's row, complete since every thread passed the barrier
        if (pending) flush();
        // 2. strength order and the chunk sums of its exclusive prefix
        float v[kR];
        Acc part = 0;
#pragma unroll
        for (int u = 0; u < kR; ++u) v[u] = rhs[oidx[u]];
#pragma unroll
        for (int u = 0; u < kR; ++u) {
            // past the chunk: 0, stored to the spare slot (which stays 0)
            const uint32_t k = lo + u;
            v[u] = k < hi ? v[u] : 0.0f;
            rs[k < hi ? k : mo] = v[u];
            part += (Acc)v[u];
        }
        const Acc incl = wave_incl_sum(part);
        const Acc excl = dpp_zero<0x138>(incl);  // the previous lane's; 0 in lane 0
        if (lane == kWave - 1) wsum[wave] = incl;
        if constexpr (SEQ) {
            // the wavefront's smallest lowest set bit (negated: a max)
            int nl = -1024;
#pragma unroll
            for (int u = 0; u < kR; ++u)
                if (v[u] != 0.0f) nl = max(nl, -lsb_exp(v[u]));
            const float wl = wave_max_f32((float)nl);
            if (lane == 0) exr[wave] = wl;
        }
        __syncthreads();
        // SEQ: a row whose chunked f64 sums could round differently from the
        // CPU engine's takes the CPU's order
        bool seq = false;
        if constexpr (SEQ) {
            float l = -1024.0f;
            Acc tot = 0;
#pragma unroll
            for (int w2 = 0; w2 < kWaves; ++w2) {
                l = fmaxf(l, exr[w2]);
                tot += wsum[w2];
            }
            seq = !exact_in_any_order(tot, -(int)l);
        }
        const float win = nd.win, lose = nd.lose, tie = nd.tie;
        float vm = 0.0f;
        if constexpr (SEQ) {
            if (seq)
                vm = show_in_cpu_order<kR, kSeg>(tid, wave, lane, card, seg, mo, mp, rs, (double*)P, (double*)cpre, O, cw,
                                                 c0 + card + len, ent, pidx, win, lose, tie);
        }
        if (!seq) {
            Acc before = 0, total = 0;
#pragma unroll
            for (int w2 = 0; w2 < kWaves; ++w2) {
                const Acc t = wsum[w2];
                before += (uint32_t)w2 < wave ? t : 0.0;
                total += t;
            }
            Acc run = before + excl;
#pragma unroll
            for (int u = 0; u < kR; ++u) {
                const uint32_t k = lo + u;
                P[k < hi ? k : mo + 1] = run;
                run += (Acc)v[u];
            }
            if (tid == 0) P[mo] = total;
            // 3. per-card exclusive prefixes in strength order
            if (card < 52) {
                float x[kSeg];
#pragma unroll
                for (int u = 0; u < (int)kSeg; ++u) x[u] = rs[cw[u] & 0xFFFFu];
                Acc a = 0;
#pragma unroll
                for (int u = 0; u < (int)kSeg; ++u) a += (Acc)x[u];
                // exclusive scan over the card's four segments (lanes 4*card..+3,
                // one quad: quad_perm broadcasts each lane's sum to the quad)
                const Acc a0 = dpp_zero<0x00>(a), a1 = dpp_zero<0x55>(a), a2 = dpp_zero<0xAA>(a), a3 = dpp_zero<0xFF>(a);
                Acc y = 0;
                if (seg > 0) y += a0;
                if (seg > 1) y += a1;
                if (seg > 2) y += a2;
#pragma unroll
                for (int u = 0; u < (int)kSeg; ++u) {
                    cpre[cw[u] >> 16] = y;
                    y += (Acc)x[u];
                }
                if (seg == 3) cpre[c0 + card + len] = ((a0 + a1) + a2) + a3;
            }
            __syncthreads();
            // 4. per-hand values, the traverser's hands in strength order (the
            //    rank bounds, hence the prefix reads, rise with the thread index),
            //    staged in hand order in LDS
#pragma unroll
            for (int u = 0; u < kR; ++u) {
                const uint32_t t = tid + u * kBlock;
                if (t >= mp) break;
                const uint4 e = ent[u];
                const uint32_t lb = e.x & 0xFFFFu, ub = e.x >> 16;
                const Acc* cp1 = cpre + (e.y & 0xFFFFu);
                const Acc* cp2 = cpre + (e.y >> 16);
                const Acc S1 = cp1[e.w & 0xFFu], S2 = cp2[(e.w >> 8) & 0xFFu];
                const float lower = (float)(P[lb] - cp1[e.z & 0xFFu] - cp2[(e.z >> 8) & 0xFFu]);
                const float higher =
                    (float)((total - P[ub]) - (S1 - cp1[(e.z >> 16) & 0xFFu]) - (S2 - cp2[e.z >> 24]));
                const uint32_t sp = e.w >> 16;
                const float same_r = sp != 0xFFFFu ? rs[sp] : 0.0f;
                const float valid = (float)(total - S1 - S2) + same_r;
                const float tied = valid - lower - higher;
                const float val = win * lower + lose * higher + tie * tied;
                O[pidx[u]] = val;
                vm = fmaxf(vm, fabsf(val));
            }
        }
        if (S.vmax) {
            vm = wave_max_f32(vm);
            if (lane == 0) wmax[wave] = vm;
        }
        // the row goes out after the next barrier (the next node's first,
        // or the one below): the next node's writes before that barrier
        // touch only its hand-order row
        pending = true;
        pend_cfv = nd.cfv;
    }
    if (pending) {
        __syncthreads();
        flush();
    }
}

// ---------------------------------------------------------------------------
// Fused river sweep (mode 0, DCFR iteration): one workgroup per river
// subtree runs its program (RiverOp, records.h) — the depth-first walk of
// the subtree with both players' reach rows and the traverser's value rows
// in LDS — so river-level reach and value rows never go to HBM: only the
// arenas stream, the parent's reach rows are read once and the subtree
// root's value row is written once (for the river chance node's merge).
// Each thread owns hands tid + 256 u of both players; reach, strategy and
// update arithmetic per hand is thread-local and needs no barrier. Only the
// terminals (and the opponent's all-zero test after each of its actions)
// read across hands, behind workgroup barriers. Every per-hand expression,
// the fold and showdown association included, is the level-sweep kernels'
// (down_action_node, k_up_fold, k_up_show<kR, kSeg>, up_action_node<0, STORE>),
// so the fused and level-synchronous sweeps give bit-identical results.
// ---------------------------------------------------------------------------
template <int KIND, int NA>
__device__ __forceinline__ float pick_sigma(const HandSigma<NA, KIND>& hs, int a) {
    float s = 0.0f;
#pragma unroll
    for (int x = 0; x < NA; ++x) s = x == a ? hs.sig[x] : s;
    return s;
}

// destination modes of a value row entry: store, 0 + v, add
__device__ __forceinline__ void put_value(float* __restrict__ row, uint32_t i, float v, int mode) {
    if (mode == 0) row[i] = v;
    else if (mode == 1) row[i] = 0.0f + v;
    else row[i] = row[i] + v;
}

// max over the workgroup (all threads call it); slot: 4 floats of LDS
__device__ __forceinline__ float block_max_f32(float x, float* slot) {
    x = wave_max_f32(x);
    if ((threadIdx.x & 63) == 0) slot[threadIdx.x >> 6] = x;
    __syncthreads();
    return fmaxf(fmaxf(slot[0], slot[1]), fmaxf(slot[2], slot[3]));
}

// A river op's fields as wavefront-uniform values (scalar registers): the
// op index is uniform but depends on loaded data (skips), so the loads are
// made uniform explicitly rather than left to the compiler's analysis.
struct RiverOpU {
    uint32_t op, na, a, q, src, dst, mode, vrow, cdst, cmode, skip, rel, doff;
    float w[3];
};
__device__ __forceinline__ uint32_t uni(uint32_t x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ RiverOpU load_op(const RiverOp* __restrict__ p) {
    const uint32_t* w = (const uint32_t*)p;
    const uint32_t w0 = uni(w[0]), w1 = uni(w[1]), w2 = uni(w[2]);
    RiverOpU o;
    o.op = w0 & 0xFFu;
    o.na = (w0 >> 8) & 0xFFu;
    o.a = (w0 >> 16) & 0xFFu;
    o.q = w0 >> 24;
    o.src = w1 & 0xFFu;
    o.dst = (w1 >> 8) & 0xFFu;
    o.mode = (w1 >> 16) & 0xFFu;
    o.vrow = w1 >> 24;
    o.cdst = w2 & 0xFFu;
    o.cmode = (w2 >> 8) & 0xFFu;
    o.skip = w2 >> 16;
    o.rel = uni(w[3]);
    o.doff = uni(w[4]);
    o.w[0] = __builtin_bit_cast(float, uni(w[5]));
    o.w[1] = __builtin_bit_cast(float, uni(w[6]));
    o.w[2] = __builtin_bit_cast(float, uni(w[7]));
    return o;
}
// element q (uniform) of a two-entry array without dynamic indexing
template <class X>
__device__ __forceinline__ X sel(int q, X a0, X a1) { return q ? a1 : a0; }

// Three waves per SIMD (at most 168 VGPRs): three subtrees per CU hide each
// other's memory latency, which bounds this kernel (two per CU: 16.9 ms per
// flop SRP iteration, three: 13.2).
template <int kR, int kSeg, int STORE>
__global__ __launch_bounds__(kBlock, 3) void k_river(DevTree T, DevState S, const RiverSub* __restrict__ subs,
                                                  uint32_t nsub, const RiverOp* __restrict__ ops, int p,
                                                  uint32_t nrows, uint32_t mo_max) {
    static_assert(STORE == kStoreF32 || STORE == kStoreH, "fused river sweeps: f32 or f16 arenas");
    constexpr int KCUR = STORE == kStoreH ? kCurrentH : kCurrentF32;
    extern __shared__ double lds[];
    const uint32_t si = xcd_slice(blockIdx.x, gridDim.x);
    if (si >= nsub) return;
    // the subtree's record as uniform values
    struct {
        uint64_t base0, base1, dealt;
        uint32_t node, prog, nops, src0, src1, card, cfv, board, size0, size1;
    } sub;
    {
        const uint32_t* w = (const uint32_t*)(subs + si);
        sub.base0 = (uint64_t)uni(w[0]) | ((uint64_t)uni(w[1]) << 32);
        sub.base1 = (uint64_t)uni(w[2]) | ((uint64_t)uni(w[3]) << 32);
        sub.dealt = (uint64_t)uni(w[4]) | ((uint64_t)uni(w[5]) << 32);
        sub.node = uni(w[6]);
        sub.prog = uni(w[7 + p]);
        sub.nops = uni(w[9 + p]);
        sub.src0 = uni(w[11]);
        sub.src1 = uni(w[12]);
        sub.card = uni(w[13]);
        sub.cfv = uni(w[14]);
        sub.board = uni(w[15]);
        sub.size0 = uni(w[16]);
        sub.size1 = uni(w[17]);
    }
    // Warm the L2 with the subtree's arena ranges (both players' regrets, the
    // traverser's strategy sums): one byte per 128-byte line, all issued at
    // once, so the program's own loads later hit the cache instead of each
    // waiting on HBM in turn. The bytes are summed after the first barrier.
    constexpr uint32_t kEB = STORE == kStoreH ? 2u : 4u;  // bytes per arena entry
    constexpr int kWarm = 4;                               // lines per thread and range
    uint32_t wv[3][kWarm];
    {
        const uint8_t* rr[3] = {
            (const uint8_t*)(STORE == kStoreH ? (const void*)(S.Rq[0]) : (const void*)(S.R[0])) + sub.base0 * kEB,
            (const uint8_t*)(STORE == kStoreH ? (const void*)(S.Rq[1]) : (const void*)(S.R[1])) + sub.base1 * kEB,
            (const uint8_t*)(STORE == kStoreH ? (const void*)(S.Sq[p]) : (const void*)(S.S[p])) +
                (p ? sub.base1 : sub.base0) * kEB};
        const uint64_t nb[3] = {(uint64_t)sub.size0 * kEB, (uint64_t)sub.size1 * kEB,
                                (uint64_t)(p ? sub.size1 : sub.size0) * kEB};
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int u = 0; u < kWarm; ++u) {
                const uint64_t x = ((uint64_t)u * kBlock + threadIdx.x) * 128u;
                wv[r][u] = x < nb[r] ? rr[r][x] : 0u;
            }
    }
    const int o = 1 - p;
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t nhp = T.nh[p], nho = T.nh[o];
    const RiverLds L = river_lds_layout(nrows, T.nh_max, mo_max);
    char* base = (char*)lds;
    const uint32_t rs = L.rstride;
    float* const rows = (float*)base;
    auto row = [&](uint32_t r) -> float* { return rows + (size_t)r * rs; };
    uint32_t* coff = (uint32_t*)(base + L.coff);
    float* red = (float*)(base + L.red);            // 4 x 4 floats of reductions
    double* wsum = (double*)(base + L.red + 64);    // per wave: showdown scan total
    char* scr = base + L.scr;

    const uint32_t it = *S.iter;
    const float pos = S.disc[3 * it + 0], neg = S.disc[3 * it + 1], ds = S.disc[3 * it + 2];

    // root reach rows: 1 = traverser, 2 = opponent (the chance edge's
    // masking, k_down_chance)
    const uint64_t cm = sub.card < 52 ? 1ull << sub.card : 0ull;
    bool anyo = false;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const int q = k == 0 ? p : o;
        const uint32_t nh = sel(q, T.nh[0], T.nh[1]);
        const uint32_t src = sel(q, sub.src0, sub.src1);
        float* dst = row(1 + k);
        const bool alive = sel(q, S.alive[0], S.alive[1])[src] != 0;
        const float* __restrict__ in = sel(q, S.reach[0], S.reach[1]) + (size_t)src * nh;
        const uint64_t* __restrict__ hm = sel(q, T.hmask[0], T.hmask[1]);
#pragma unroll
        for (int u = 0; u < kR; ++u) {
            const uint32_t j = tid + u * kBlock;
            if (j < nh) {
                const float x = (!alive || (hm[j] & cm)) ? 0.0f : in[j];
                dst[j] = x;
                if (k == 1) anyo |= x > 0.0f;
            }
        }
    }
    // static for the subtree: the showdown board's tables (k_up_show) and the
    // fold kernel's hand words
    const uint32_t b = sub.board;
    const bool has_show = b != 0xFFFFFFFFu;
    const uint32_t mo = has_show ? T.sd_cnt[o][b] : 0u, mp = has_show ? T.sd_cnt[p][b] : 0u;
    if (has_show && tid < 53) coff[tid] = T.sd_coff[o][(size_t)b * 53 + tid];
    anyo = __syncthreads_or(anyo);
    uint32_t warm = 0;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int u = 0; u < kWarm; ++u) warm += wv[r][u];
    float* __restrict__ out_g = S.cfv + (size_t)sub.cfv * T.nh_max;
    if (!anyo) {  // no opponent reach: the subtree is pruned
#pragma unroll
        for (int u = 0; u < kR; ++u)
            if (tid + u * kBlock < nhp) out_g[tid + u * kBlock] = 0.0f;
        return;
    }
    const uint32_t cs = (mo + kBlock - 1) / kBlock;
    const uint32_t lo = min(tid * cs, mo), hi = min(lo + cs, mo);
    uint4 ent[kR];
    uint32_t pidx[kR], oidx[kR];
    uint32_t onb = 0;  // bit u: own traverser hand tid + 256 u is on the river board
    {
#pragma unroll
        for (int u = 0; u < kR; ++u) {
            const uint32_t i = tid + u * kBlock;
            if (i < nhp && !(T.hmask[p][i] & sub.dealt)) onb |= 1u << u;
        }
        const uint32_t sp = has_show ? T.sd_off[p][b] : 0u;
        const uint4* __restrict__ tab = T.sd_tab[p] + sp;
        const uint16_t* __restrict__ gp = T.sd_idx[p] + sp;
        const uint32_t mlast = mp > 0 ? mp - 1 : 0;
#pragma unroll
        for (int u = 0; u < kR; ++u) {
            const uint32_t t = min(tid + u * kBlock, mlast);
            ent[u] = mp > 0 ? tab[t] : make_uint4(0u, 0u, 0u, 0u);
            pidx[u] = mp > 0 ? gp[t] : 0u;
        }
        const uint32_t so = has_show ? T.sd_off[o][b] : 0u;
        const uint16_t* __restrict__ gi = T.sd_idx[o] + so;
        const uint32_t olast = mo > 0 ? mo - 1 : 0;
#pragma unroll
        for (int u = 0; u < kR; ++u) oidx[u] = mo > 0 ? gi[min(lo + u, olast)] : 0u;
    }
    const uint32_t card = tid >> 2, seg = tid & 3;
    uint32_t c0 = 0, len = 0, s0 = 0, sn = 0;
    if (has_show && card < 52) {
        c0 = coff[card];
        len = coff[card + 1] - c0;
        s0 = min(seg * (uint32_t)kSeg, len);
        sn = min((uint32_t)kSeg, len - s0);
    }
    const uint32_t creg = 2 * mo + 52;
    uint32_t cw[kSeg];
    {
        const uint16_t* __restrict__ gpos = T.sd_cpos[o] + 2 * (size_t)(has_show ? T.sd_off[o][b] : 0u);
        const uint32_t last = mo > 0 ? 2 * mo - 1 : 0;
#pragma unroll
        for (int u = 0; u < (int)kSeg; ++u)
            cw[u] = (uint32_t)u < sn ? gpos[min(c0 + s0 + u, last)] | ((c0 + card + s0 + u) << 16)
                                     : mo | (creg << 16);
    }

    const RiverOp* __restrict__ prog = ops + sub.prog;
    const uint32_t nops = sub.nops;
    for (uint32_t k = 0; k < nops;) {
        const RiverOpU op = load_op(prog + k);
        ++k;
        if (op.op == kOpDescend) {
            // child reach of the actor q: parent reach x current strategy
            const int q = (int)op.q;
            const uint32_t nh = sel(q, T.nh[0], T.nh[1]);
            const float* __restrict__ src = row(op.src);
            float* __restrict__ dst = row(op.dst);
            const auto* arena = sel(q, ArenaOf<KCUR>::get(S, 0), ArenaOf<KCUR>::get(S, 1)) + sel(q, sub.base0, sub.base1) +
                                op.doff;
            const int a = op.a;
            bool any = false;
            dispatch_na4(op.na, [&](auto NAc) {
                constexpr int NA = decltype(NAc)::value;
                if constexpr (NA > 0) {
                    HandSigma<NA, KCUR> hs[kR];
#pragma unroll
                    for (int u = 0; u < kR; ++u) hs[u].fetch(arena, nh, min(tid + u * kBlock, nh - 1));
#pragma unroll
                    for (int u = 0; u < kR; ++u) {
                        const uint32_t j = tid + u * kBlock;
                        if (j < nh) {
                            hs[u].finish();
                            const float v = src[j] * pick_sigma(hs[u], a);
                            dst[j] = v;
                            any |= v > 0.0f;
                        }
                    }
                } else {
#pragma unroll
                    for (int u = 0; u < kR; ++u) {
                        const uint32_t j = tid + u * kBlock;
                        if (j < nh) {
                            SigmaHand<KCUR> sh;
                            sh.init(arena, nh, j, op.na);
                            const float v = src[j] * sh.template sigma<false>(a);
                            dst[j] = v;
                            any |= v > 0.0f;
                        }
                    }
                }
            });
            if (q == o) {
                // the opponent's row is complete (and read across hands by the
                // terminals) after this barrier; an all-zero row prunes the child
                any = __syncthreads_or(any);
                if (!any) {
                    float* cd = row(op.cdst);
#pragma unroll
                    for (int u = 0; u < kR; ++u)
                        if (tid + u * kBlock < nhp) put_value(cd, tid + u * kBlock, 0.0f, op.cmode);
                    k += op.skip;
                }
            }
        } else if (op.op == kOpFold) {
            // k_up_fold on the opponent's row src
            const float* __restrict__ ro = row(op.src);
            float* __restrict__ dst = row(op.dst);
            float* M = (float*)scr;
            double* s_card = (double*)(scr + 52 * 52 * 4);
            double* s_tot = s_card + 52;
            // the opponent's hands as card-matrix positions, the traverser's
            // packed hand words
            uint32_t hw[kR], mpos[kR];
#pragma unroll
            for (int u = 0; u < kR; ++u) {
                mpos[u] = T.hpack[o][min(tid + u * kBlock, nho - 1)];
                hw[u] = T.hpack[p][min(tid + u * kBlock, nhp - 1)];
            }
            __syncthreads();  // earlier readers of the scratch are done
            {
                float4* m4 = (float4*)M;
                for (uint32_t x = tid; x < 52 * 52 / 4; x += kBlock) m4[x] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            }
            __syncthreads();
#pragma unroll
            for (int u = 0; u < kR; ++u)
                if (tid + u * kBlock < nho) {
                    const float r = ro[tid + u * kBlock];
                    const uint32_t a1 = mpos[u] & 0xFFu, a2 = (mpos[u] >> 8) & 0xFFu;
                    M[a1 * 52 + a2] = r;
                    M[a2 * 52 + a1] = r;
                }
            __syncthreads();
            if (wave == 0) {
                // the total as k_up_fold forms it: per lane over hands
                // lane + 64 u, then the wavefront sum
                double part = 0.0;
                for (uint32_t j = lane; j < nho; j += kWave) part += (double)ro[j];
                const double tot = wave_sum_f64(part);
                if (lane == 0) *s_tot = tot;
                if (lane < 52) {
                    const float4* mrow = (const float4*)(M + lane * 52);
                    double acc = 0.0;
#pragma unroll
                    for (int t0 = 0; t0 < 13; t0 += 4) {
                        float4 v[4];
#pragma unroll
                        for (int t = 0; t < 4; ++t)
                            if (t0 + t < 13) v[t] = mrow[t0 + t];
#pragma unroll
                        for (int t = 0; t < 4; ++t)
                            if (t0 + t < 13) {
                                acc += (double)v[t].x;
                                acc += (double)v[t].y;
                                acc += (double)v[t].z;
                                acc += (double)v[t].w;
                            }
                    }
                    s_card[lane] = acc;
                }
            }
            __syncthreads();
            const double tot = *s_tot;
            const float amount = op.q == (uint32_t)p ? op.w[1] : op.w[0];
#pragma unroll
            for (int u = 0; u < kR; ++u) {
                const uint32_t i = tid + u * kBlock;
                if (i < nhp) {
                    const uint32_t a1 = hw[u] & 0xFFu, a2 = (hw[u] >> 8) & 0xFFu;
                    const float same_r = M[a1 * 52 + a2];
                    const float valid = (float)(tot - s_card[a1] - s_card[a2]) + same_r;
                    put_value(dst, i, amount * valid, op.mode);
                }
            }
        } else if (op.op == kOpShow) {
            // k_up_show's phases 2-4 on the opponent's row src
            const float* __restrict__ rhs = row(op.src);
            float* __restrict__ dst = row(op.dst);
            const uint32_t sd_P = 0, sd_cpre = 8 * (mo + 2), sd_rs = sd_cpre + 8 * (creg + 1);
            double* P = (double*)(scr + sd_P);
            double* cpre = (double*)(scr + sd_cpre);
            float* rsw = (float*)(scr + sd_rs);
            const float win = op.w[0], lose = op.w[1], tie = op.w[2];
            // hands off the board take 0 (k_up_show leaves them at 0); with no
            // opponent hand on the board every hand does
#pragma unroll
            for (int u = 0; u < kR; ++u)
                if (tid + u * kBlock < nhp && (mo == 0 || !((onb >> u) & 1u)))
                    put_value(dst, tid + u * kBlock, 0.0f, op.mode);
            if (mo == 0) continue;
            __syncthreads();  // earlier readers of the scratch and writers of dst are done
            float v[kR];
            double part = 0.0;
#pragma unroll
            for (int u = 0; u < kR; ++u) v[u] = rhs[oidx[u]];
#pragma unroll
            for (int u = 0; u < kR; ++u) {
                const uint32_t kk = lo + u;
                v[u] = kk < hi ? v[u] : 0.0f;
                rsw[kk < hi ? kk : mo] = v[u];
                part += (double)v[u];
            }
            if (tid == 0) rsw[mo] = 0.0f;
            const double incl = wave_incl_sum_f64(part);
            const double excl = dpp_zero<0x138>(incl);
            if (lane == kWave - 1) wsum[wave] = incl;
            __syncthreads();
            double before = 0.0, total = 0.0;
#pragma unroll
            for (int w2 = 0; w2 < kWaves; ++w2) {
                const double t = wsum[w2];
                before += (uint32_t)w2 < wave ? t : 0.0;
                total += t;
            }
            double run = before + excl;
#pragma unroll
            for (int u = 0; u < kR; ++u) {
                const uint32_t kk = lo + u;
                P[kk < hi ? kk : mo + 1] = run;
                run += (double)v[u];
            }
            if (tid == 0) P[mo] = total;
            if (card < 52) {
                float x[kSeg];
#pragma unroll
                for (int u = 0; u < (int)kSeg; ++u) x[u] = rsw[cw[u] & 0xFFFFu];
                double a = 0.0;
#pragma unroll
                for (int u = 0; u < (int)kSeg; ++u) a += (double)x[u];
                const double a0 = dpp_zero<0x00>(a), a1 = dpp_zero<0x55>(a), a2 = dpp_zero<0xAA>(a),
                             a3 = dpp_zero<0xFF>(a);
                double y = 0.0;
                if (seg > 0) y += a0;
                if (seg > 1) y += a1;
                if (seg > 2) y += a2;
#pragma unroll
                for (int u = 0; u < (int)kSeg; ++u) {
                    cpre[cw[u] >> 16] = y;
                    y += (double)x[u];
                }
                if (seg == 3) cpre[c0 + card + len] = ((a0 + a1) + a2) + a3;
            }
            __syncthreads();
#pragma unroll
            for (int u = 0; u < kR; ++u) {
                const uint32_t t = tid + u * kBlock;
                if (t >= mp) continue;
                const uint4 e = ent[u];
                const uint32_t lb = e.x & 0xFFFFu, ub = e.x >> 16;
                const double* cp1 = cpre + (e.y & 0xFFFFu);
                const double* cp2 = cpre + (e.y >> 16);
                const double S1 = cp1[e.w & 0xFFu], S2 = cp2[(e.w >> 8) & 0xFFu];
                const float lower = (float)(P[lb] - cp1[e.z & 0xFFu] - cp2[(e.z >> 8) & 0xFFu]);
                const float higher =
                    (float)((total - P[ub]) - (S1 - cp1[(e.z >> 16) & 0xFFu]) - (S2 - cp2[e.z >> 24]));
                const uint32_t sp = e.w >> 16;
                const float same_r = sp != 0xFFFFu ? rsw[sp] : 0.0f;
                const float valid = (float)(total - S1 - S2) + same_r;
                const float tied = valid - lower - higher;
                put_value(dst, pidx[u], win * lower + lose * higher + tie * tied, op.mode);
            }
            __syncthreads();  // the scattered row is complete
        } else if (op.op == kOpPostP) {
            // the traverser's node: k_up_action<0, STORE> on LDS rows
            const int na = op.na;
            const float* __restrict__ rp = row(op.src);
            const float* __restrict__ V = row(op.vrow);
            float* __restrict__ dst = row(op.dst);
            if constexpr (STORE == kStoreF32) {
                const uint64_t off = sel(p, sub.base0, sub.base1) + op.doff;
                float* __restrict__ R = sel(p, S.R[0], S.R[1]) + off;
                float* __restrict__ Sm = sel(p, S.S[0], S.S[1]) + off;
                dispatch_na4(na, [&](auto NAc) {
                    constexpr int NA = decltype(NAc)::value;
                    if constexpr (NA > 0) {
                        HandSigma<NA, kCurrentF32> hs[kR];
                        float sm[kR][NA];
#pragma unroll
                        for (int u = 0; u < kR; ++u) {
                            const uint32_t ii = min(tid + u * kBlock, nhp - 1);
#pragma unroll
                            for (int x = 0; x < NA; ++x) {
                                hs[u].raw[x] = R[(size_t)x * nhp + ii];
                                sm[u][x] = Sm[(size_t)x * nhp + ii];
                            }
                        }
#pragma unroll
                        for (int u = 0; u < kR; ++u) {
                            const uint32_t i = tid + u * kBlock;
                            if (i >= nhp) continue;
                            hs[u].finish();
                            float v[NA];
#pragma unroll
                            for (int x = 0; x < NA; ++x) v[x] = V[(size_t)x * rs + i];
                            const float reach = rp[i];
                            float acc = 0.0f;
#pragma unroll
                            for (int x = 0; x < NA; ++x) acc += hs[u].sig[x] * v[x];
#pragma unroll
                            for (int x = 0; x < NA; ++x) {
                                const float rv = hs[u].raw[x];
                                const float d = rv > 0.0f ? pos : neg;
                                R[(size_t)x * nhp + i] = rv * d + (v[x] - acc);
                                Sm[(size_t)x * nhp + i] = sm[u][x] * ds + reach * hs[u].sig[x];
                            }
                            put_value(dst, i, acc, op.mode);
                        }
                    } else {
#pragma unroll
                        for (int u = 0; u < kR; ++u) {
                            const uint32_t i = tid + u * kBlock;
                            if (i >= nhp) continue;
                            SigmaHand<kCurrentF32> sh;
                            sh.init(R, nhp, i, na);
                            float acc = 0.0f;
                            for (int x = 0; x < na; ++x) acc += sh.template sigma<false>(x) * V[(size_t)x * rs + i];
                            const float reach = rp[i];
                            for (int x = 0; x < na; ++x) {
                                const size_t idx = (size_t)x * nhp + i;
                                const float rv = R[idx];
                                const float sg = sh.template sigma<false>(x);
                                const float d = rv > 0.0f ? pos : neg;
                                R[idx] = rv * d + (V[(size_t)x * rs + i] - acc);
                                Sm[idx] = Sm[idx] * ds + reach * sg;
                            }
                            put_value(dst, i, acc, op.mode);
                        }
                    }
                });
            } else {
                // f16 arenas: the scales come from the node's previous maxima
                // and the maxima of its children's value rows and of its
                // reach row (k_up_action<0, kStoreH>)
                const uint64_t off = sel(p, sub.base0, sub.base1) + op.doff;
                uint16_t* __restrict__ Rh = (uint16_t*)sel(p, S.Rq[0], S.Rq[1]) + off;
                uint16_t* __restrict__ Sh = sel(p, S.Sq[0], S.Sq[1]) + off;
                const uint32_t node = sub.node + op.rel;
                float* __restrict__ Rs_ = sel(p, S.Rs[0], S.Rs[1]);
                float* __restrict__ Ss_ = sel(p, S.Ss[0], S.Ss[1]);
                float* __restrict__ Rm_ = sel(p, S.Rm[0], S.Rm[1]);
                float* __restrict__ Sm_ = sel(p, S.Sm[0], S.Sm[1]);
                float mv = 0.0f, mr = 0.0f;
#pragma unroll
                for (int u = 0; u < kR; ++u) {
                    const uint32_t i = tid + u * kBlock;
                    if (i < nhp) {
                        for (int x = 0; x < na; ++x) mv = fmaxf(mv, fabsf(V[(size_t)x * rs + i]));
                        mr = fmaxf(mr, rp[i]);
                    }
                }
                mv = block_max_f32(mv, red);
                mr = block_max_f32(mr, red + 4);
                const float sR = Rs_[node], sS = Ss_[node];
                const float kRs = half_scale(Rm_[node] * fmaxf(pos, neg) + 2.0f * mv);
                const float kSs = half_scale(Sm_[node] * ds + mr);
                const float iR = 1.0f / kRs, iS = 1.0f / kSs;
                float xr = 0.0f, xs = 0.0f;
                dispatch_na4(na, [&](auto NAc) {
                    constexpr int NA = decltype(NAc)::value;
                    if constexpr (NA > 0) {
                        HandSigma<NA, kCurrentH> hs[kR];
                        float sm[kR][NA];
#pragma unroll
                        for (int u = 0; u < kR; ++u) {
                            const uint32_t ii = min(tid + u * kBlock, nhp - 1);
                            hs[u].fetch(Rh, nhp, ii);
#pragma unroll
                            for (int x = 0; x < NA; ++x) sm[u][x] = h2f(Sh[(size_t)x * nhp + ii]);
                        }
#pragma unroll
                        for (int u = 0; u < kR; ++u) {
                            const uint32_t i = tid + u * kBlock;
                            if (i >= nhp) continue;
                            hs[u].finish();
                            float v[NA];
#pragma unroll
                            for (int x = 0; x < NA; ++x) v[x] = V[(size_t)x * rs + i];
                            const float reach = rp[i];
                            float acc = 0.0f;
#pragma unroll
                            for (int x = 0; x < NA; ++x) acc += hs[u].sig[x] * v[x];
#pragma unroll
                            for (int x = 0; x < NA; ++x) {
                                const float rv = hs[u].raw[x] * sR;
                                const float d = rv > 0.0f ? pos : neg;
                                const float nr = rv * d + (v[x] - acc);
                                const float ns = (sm[u][x] * sS) * ds + reach * hs[u].sig[x];
                                Rh[(size_t)x * nhp + i] = f2h(nr * iR);
                                Sh[(size_t)x * nhp + i] = f2h(ns * iS);
                                xr = fmaxf(xr, fabsf(nr));
                                xs = fmaxf(xs, ns);
                            }
                            put_value(dst, i, acc, op.mode);
                        }
                    } else {
#pragma unroll
                        for (int u = 0; u < kR; ++u) {
                            const uint32_t i = tid + u * kBlock;
                            if (i >= nhp) continue;
                            SigmaHand<kCurrentH> sh;
                            sh.init(Rh, nhp, i, na);
                            float acc = 0.0f;
                            for (int x = 0; x < na; ++x) acc += sh.template sigma<false>(x) * V[(size_t)x * rs + i];
                            const float reach = rp[i];
                            for (int x = 0; x < na; ++x) {
                                const size_t idx = (size_t)x * nhp + i;
                                const float rv = h2f(Rh[idx]) * sR;
                                const float sg = sh.template sigma<false>(x);
                                const float d = rv > 0.0f ? pos : neg;
                                const float nr = rv * d + (V[(size_t)x * rs + i] - acc);
                                const float ns = (h2f(Sh[idx]) * sS) * ds + reach * sg;
                                Rh[idx] = f2h(nr * iR);
                                Sh[idx] = f2h(ns * iS);
                                xr = fmaxf(xr, fabsf(nr));
                                xs = fmaxf(xs, ns);
                            }
                            put_value(dst, i, acc, op.mode);
                        }
                    }
                });
                xr = wave_max_f32(xr);
                xs = wave_max_f32(xs);
                float* r2 = red + 8;  // a slot set of its own: red / red + 4 may still be read
                if (lane == 0) {
                    r2[wave] = xr;
                    r2[4 + wave] = xs;
                }
                __syncthreads();
                if (tid == 0) {
                    Rs_[node] = kRs;
                    Ss_[node] = kSs;
                    Rm_[node] = fmaxf(fmaxf(r2[0], r2[1]), fmaxf(r2[2], r2[3]));
                    Sm_[node] = fmaxf(fmaxf(r2[4], r2[5]), fmaxf(r2[6], r2[7]));
                }
            }
        } else {  // kOpPostO: an opponent node's sum added into its parent's row
            const float* __restrict__ A = row(op.vrow);
            float* __restrict__ dst = row(op.dst);
#pragma unroll
            for (int u = 0; u < kR; ++u)
                if (tid + u * kBlock < nhp) put_value(dst, tid + u * kBlock, A[tid + u * kBlock], 2);
        }
    }
    if (nops == 0u) red[15] = (float)warm;  // keeps the warming loads (a program is never empty)
    // the subtree root's row, for the river chance node's merge
    const float* __restrict__ res = row(0);
#pragma unroll
    for (int u = 0; u < kR; ++u)
        if (tid + u * kBlock < nhp) out_g[tid + u * kBlock] = res[tid + u * kBlock];
}

// ---------------------------------------------------------------------------
// Up sweep, showdown terminals as a matrix product on the matrix cores
// (the BF16 precision form; the exact form is k_up_show). For the nodes of
// one river board, sorted by strength on both sides:
//   L = W R, H = G R, Q = T R      (hands of p x nodes)
// with R[k][n] the opponent reach of sorted hand k at node n (bf16),
// D[i][k] = 1 if hands i and k share no card, W = D where k is strictly
// weaker than i (k < lb_i), G = D where k is strictly stronger (k >= ub_i),
// T = D - W - G (ties), and value = win L + lose H + tie Q: game.rs's
// showdown sums, stacked over the board's nodes. One workgroup per run of
// up to 32 nodes (one 32-wide N tile); each wavefront takes 32-hand M tiles
// and walks K in steps of 16 with v_mfma_f32_32x32x16_bf16. Rows of a tile
// are consecutive in strength, so most steps lie wholly below, above or
// inside every row's tie band and issue one product; only steps that
// straddle a bound issue three. Operand A comes from the card-disjointness
// bitsets (D); operand B is staged in LDS. Accumulation is f32.
// ---------------------------------------------------------------------------
constexpr int kMmaBlock = 512;                                     // 8 wavefronts share one B operand
constexpr int kMmaWaves = kMmaBlock / kWave;
constexpr int kMmaRowPer = (1326 + kMmaBlock - 1) / kMmaBlock;     // hand rows per thread
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x16_t __attribute__((ext_vector_type(16)));
typedef unsigned short u16x8_t __attribute__((ext_vector_type(8)));

// A-operand fragments of the 0/1 showdown matrices. Bitset words are stored
// pair-split (bit 2i of a 32-k word at bit i, bit 2i+1 at bit 16+i), so the
// two k of one bf16 pair sit 16 bits apart: a rotation puts them at bits 14
// and 30 and a mask leaves bf16 2.0 (0x4000) or 0 in each half. Products
// carry the factor 2, removed exactly in the epilogue. Lane l of a
// 32x32x16 step holds k = 8*(l/32) .. +7 of its row, i.e. pairs
// i0 .. i0+3 with i0 = 8*(step&1) + 4*(l/32) of the step's word.
__device__ __forceinline__ bf16x8_t pairs_to_bf16(uint32_t z, uint32_t i0) {
    uint32_t w[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) w[t] = __builtin_rotateright32(z, (i0 + t - 14) & 31u) & 0x40004000u;
    return __builtin_bit_cast(bf16x8_t, w);
}

// Pair-split mask of this lane's k (k = kb .. kb+7 at pairs i0 .. i0+3)
// with k < lim.
__device__ __forceinline__ uint32_t pairs_below(int lim, int kb, uint32_t i0) {
    const int d = lim - kb;
    const uint32_t ne = (uint32_t)min(max((d + 1) >> 1, 0), 4);  // even k: kb + 2t < lim
    const uint32_t no = (uint32_t)min(max(d >> 1, 0), 4);        // odd k: kb + 2t + 1 < lim
    return (((1u << ne) - 1u) << i0) | (((1u << no) - 1u) << (16 + i0));
}

// e4m3 A-operand fragment (OCP e4m3: 0x40 = 2.0) from a pair-split word.
// A rotation puts bits i, i+8, i+16, i+24 of the word at bits 6, 14, 22,
// 30, i.e. k = 2i, 2i+16, 2i+1, 2i+17 of the 32-k word, one per byte; the
// second dword does the same from bit i+1. Lane group g = 2*(step&1) +
// lane/32 takes i = 2g, so a step's two lane groups hold k = 4g..4g+3 and
// 4g+16..4g+19: the four groups partition the word's 32 k, and B is stored
// in this byte order (fp8_bpos).
__device__ __forceinline__ long pairs_to_fp8(uint32_t z, uint32_t i) {
    const uint32_t d0 = __builtin_rotateright32(z, (i - 6) & 31u) & 0x40404040u;
    const uint32_t d1 = __builtin_rotateright32(z, (i - 5) & 31u) & 0x40404040u;
    return (long)((uint64_t)d0 | ((uint64_t)d1 << 32));
}

// LDS byte position of k in an e4m3 B row (the order of pairs_to_fp8)
__device__ __forceinline__ uint32_t fp8_bpos(uint32_t k) {
    const uint32_t r = k & 31u, rr = r & 15u;
    return (k & ~31u) + 8 * (rr >> 2) + 2 * (rr & 3u) + (r >> 4);
}

// Pair-split mask of the lane's k (kw + 2i + {0..3} and kw + 2i + 16 +
// {0..3}, kw the word's first k) with k < lim.
__device__ __forceinline__ uint32_t fp8_below(int lim, int kw, uint32_t i) {
    const int d1 = lim - kw - 2 * (int)i, d2 = d1 - 16;
    const uint32_t c1 = (uint32_t)min(max(d1, 0), 4), c2 = (uint32_t)min(max(d2, 0), 4);
    return (((1u << ((c1 + 1) >> 1)) - 1u) << i) | (((1u << (c1 >> 1)) - 1u) << (16 + i)) |
           (((1u << ((c2 + 1) >> 1)) - 1u) << (i + 8)) | (((1u << (c2 >> 1)) - 1u) << (24 + i));
}

template <bool FP8>
__global__ __launch_bounds__(kMmaBlock) void k_up_show_mma(DevTree T, DevState S, const ShowRec* __restrict__ recs,
                                                           const uint4* __restrict__ chunks, uint32_t n_chunks,
                                                           int p, int swap, int staged) {
    extern __shared__ double lds[];
    const uint32_t c = xcd_slice(blockIdx.x, gridDim.x);
    if (c >= n_chunks) return;
    const uint4 ck = chunks[c];
    const uint32_t b = ck.x, first = ck.y, cnt = ck.z;
    const int o = 1 - p;
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t nhp = T.nh[p], nho = T.nh[o];
    const uint32_t mo = T.sd_cnt[o][b], mp = T.sd_cnt[p][b];
    const uint32_t so = T.sd_off[o][b], sp = T.sd_off[p][b];
    const MmaLds L = mma_lds_layout(mo, mp, staged ? nhp : 0, FP8);
    const uint32_t kp = L.kpad, rs = L.rs;
    char* base = (char*)lds;
    uint16_t* B = (uint16_t*)(base + L.B);
    uint8_t* B8 = (uint8_t*)(base + L.B);
    float* bsc = (float*)(base + L.bsc);
    ShowNode* nodes = (ShowNode*)(base + L.nodes);
    uint32_t* vmx = (uint32_t*)(base + L.vmx);
    uint32_t* lbub = (uint32_t*)(base + L.lbub);
    uint16_t* oidx = (uint16_t*)(base + L.oidx);
    uint16_t* pidx = (uint16_t*)(base + L.pidx);
    float* O = (float*)(base + L.O);
    const uint32_t col = lane & 31, half = lane >> 5;
    const uint32_t* __restrict__ bits = T.sd_djbits[p] + T.sd_djoff[p][b];
    const uint32_t nblk = (kp + 255) / 256;  // K blocks of 256 (8 bitset words)

    // this wavefront's first M tile: its first bitset block loads while the
    // operands are staged
    uint32_t cw[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) cw[j] = 0u;
    if (wave * 32 < mp) {
        const uint32_t* __restrict__ row = bits + (size_t)min(wave * 32 + col, mp - 1) * L.words;
#pragma unroll
        for (int j = 0; j < 8; ++j) cw[j] = row[j];
    }
    if (tid < 32) {
        ShowNode nd{0u, 0u, 0u, 0.0f, 0.0f, 0.0f};
        if (tid < cnt) {
            const ShowRec rr = recs[first + tid];
            const uint32_t os = pick(rr.own, o);
            nd.cfv = rr.cfv;
            nd.oslot = os;
            nd.live = (S.alive[o][os] && mo != 0) ? 1u : 0u;
            nd.win = swap ? rr.lose : rr.win;
            nd.lose = swap ? rr.win : rr.lose;
            nd.tie = rr.tie;
        }
        nodes[tid] = nd;
        vmx[tid] = 0u;
    }
    for (uint32_t k = tid; k < kp; k += kMmaBlock) oidx[k] = k < mo ? T.sd_idx[o][so + k] : (uint16_t)0;
    for (uint32_t i = tid; i < mp; i += kMmaBlock) {
        pidx[i] = T.sd_idx[p][sp + i];
        lbub[i] = T.sd_lbub[p][sp + i];
    }
    if (staged) {
        // output rows start at zero: hands that conflict with the board keep it
        float4* o4 = (float4*)O;
        for (uint32_t i = tid; i < (32 * L.os + 3) / 4; i += kMmaBlock) o4[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        __syncthreads();
    } else {
        // hands that conflict with the board (not in p's sorted list) get 0
        // in every node row; the products write the others
        uint32_t* onb = (uint32_t*)(base + L.onb);
        if (tid < 42) onb[tid] = 0u;
        __syncthreads();
        for (uint32_t i = tid; i < mp; i += kMmaBlock) {
            const uint32_t h = T.sd_idx[p][sp + i];
            atomicOr(&onb[h >> 5], 1u << (h & 31u));
        }
        __syncthreads();
        uint32_t bad = 0;
#pragma unroll
        for (int u = 0; u < kMmaRowPer; ++u) {
            const uint32_t h = tid + u * kMmaBlock;
            if (h < nhp && !((onb[h >> 5] >> (h & 31u)) & 1u)) bad |= 1u << u;
        }
        if (bad)
            for (uint32_t n = 0; n < cnt; ++n) {
                float* row = S.cfv + (size_t)nodes[n].cfv * T.nh_max;
#pragma unroll
                for (int u = 0; u < kMmaRowPer; ++u)
                    if ((bad >> u) & 1u) row[tid + u * kMmaBlock] = 0.0f;
            }
    }
    // R in strength order, [node][k] (row stride rs): wavefront per node,
    // every gather of the node in flight at once. bf16; or e4m3 after a
    // power-of-two scale that puts the row's maximum in [128, 256)
    for (uint32_t n = wave; n < 32; n += kMmaWaves) {
        const ShowNode nd = nodes[n];
        const bool use = n < cnt && nd.live;
        const float* __restrict__ ro = S.reach[o] + (size_t)nd.oslot * nho;
        float v[kMaxH];
#pragma unroll
        for (int u = 0; u < kMaxH; ++u) {
            const uint32_t k = lane + u * kWave;
            v[u] = 0.0f;
            if (use && k < mo) v[u] = ro[oidx[k]];
        }
        if constexpr (FP8) {
            float m = 0.0f;
#pragma unroll
            for (int u = 0; u < kMaxH; ++u) m = fmaxf(m, v[u]);
            m = wave_max_f32(m);
            int e = 0;
            if (m > 0.0f) {
                frexpf(m, &e);  // m < 2^e
                e = min(max(e, -100), 100);
            }
            const float sc = ldexpf(1.0f, 8 - e);
            if (lane == 0) bsc[n] = ldexpf(1.0f, e - 8);
#pragma unroll
            for (int u = 0; u < kMaxH; ++u) {
                const uint32_t k = lane + u * kWave;
                if (k < kp)
                    B8[n * rs + fp8_bpos(k)] = (uint8_t)(__builtin_amdgcn_cvt_pk_fp8_f32(v[u] * sc, 0.0f, 0, false) & 0xFF);
            }
        } else {
            if (lane == 0) bsc[n] = 1.0f;
#pragma unroll
            for (int u = 0; u < kMaxH; ++u) {
                const uint32_t k = lane + u * kWave;
                if (k < kp) B[n * rs + k] = __builtin_bit_cast(uint16_t, (__bf16)v[u]);
            }
        }
    }
    __syncthreads();

    const ShowNode nd = nodes[col];
    float* __restrict__ out = S.cfv + (size_t)nd.cfv * T.nh_max;
    // the operands' factor 2 (A = 2.0) and B's scale
    const float osc = 0.5f * bsc[col];
    float vm = 0.0f;
    for (uint32_t mt = wave; mt * 32 < mp; mt += kMmaWaves) {
        const uint32_t iv = min(mt * 32 + col, mp - 1);
        const uint32_t lb = lbub[iv] & 0xFFFFu, ub = lbub[iv] >> 16;
        const uint32_t* __restrict__ row = bits + (size_t)iv * L.words;
        // the next tile's rows: its first block loads during this tile's last
        const uint32_t* __restrict__ nrow = bits + (size_t)min((mt + kMmaWaves) * 32 + col, mp - 1) * L.words;
        uint32_t lbmin = lb, lbmax = lb, ubmin = ub, ubmax = ub;
        for (int off = 32; off > 0; off >>= 1) {
            lbmin = min(lbmin, (uint32_t)__shfl_xor((int)lbmin, off));
            lbmax = max(lbmax, (uint32_t)__shfl_xor((int)lbmax, off));
            ubmin = min(ubmin, (uint32_t)__shfl_xor((int)ubmin, off));
            ubmax = max(ubmax, (uint32_t)__shfl_xor((int)ubmax, off));
        }
        lbmin = __builtin_amdgcn_readfirstlane(lbmin);
        lbmax = __builtin_amdgcn_readfirstlane(lbmax);
        ubmin = __builtin_amdgcn_readfirstlane(ubmin);
        ubmax = __builtin_amdgcn_readfirstlane(ubmax);
        // win (k < lb), lose (k >= ub) and tie sums. A 16-k step wholly on
        // one side of every row's bounds feeds one product; steps that
        // straddle a bound split the row bits three ways.
        f32x16_t aW = {}, aG = {}, aT = {};
        uint32_t nw[8];
        for (uint32_t blk = 0; blk < nblk; ++blk) {
            // eight words from the rest of this row or the next tile's row
            // (the bitset array carries 8 words of padding, so a read past
            // a row's end stays in bounds; such words feed no product)
            const uint32_t* __restrict__ np = blk + 1 < nblk ? row + (blk + 1) * 8 : nrow;
#pragma unroll
            for (int j = 0; j < 8; ++j) nw[j] = np[j];
            // B fragments two steps ahead of their products (reads past kp
            // stay inside the workgroup's LDS and feed no product)
            if constexpr (FP8) {
                const uint8_t* __restrict__ bp = B8 + col * rs + blk * 256 + 8 * half;
                long bq[16];
#pragma unroll
                for (int st = 0; st < 2; ++st) bq[st] = *(const long*)(bp + st * 16);
#pragma unroll
                for (int st = 0; st < 16; ++st) {
                    // step st: word st/2, lane group g, k = kw + 2g*2 + {0..3, 16..19}
                    const uint32_t kw = blk * 256 + (st >> 1) * 32;
                    if (st + 2 < 16) bq[st + 2] = *(const long*)(bp + (st + 2) * 16);
                    if (kw < kp) {
                        const uint32_t z = cw[st >> 1], i = 2 * (2 * (st & 1) + half);
                        const long bv = bq[st];
                        const bool onlyw = kw + 32 <= lbmin, onlyg = kw >= ubmax;
                        const bool onlyt = kw >= lbmax && kw + 32 <= ubmin;
                        const bool split = !(onlyw || onlyg || onlyt);
                        uint32_t zw = z, zg = z, zt = z;
                        if (split) {
                            const uint32_t mw = fp8_below((int)lb, (int)kw, i);
                            const uint32_t mg = ~fp8_below((int)ub, (int)kw, i);
                            zw = z & mw;
                            zg = z & mg;
                            zt = z & ~(mw | mg);
                        }
                        if (onlyw || split)
                            aW = __builtin_amdgcn_mfma_f32_32x32x16_fp8_fp8(pairs_to_fp8(zw, i), bv, aW, 0, 0, 0);
                        if (onlyg || split)
                            aG = __builtin_amdgcn_mfma_f32_32x32x16_fp8_fp8(pairs_to_fp8(zg, i), bv, aG, 0, 0, 0);
                        if (onlyt || split)
                            aT = __builtin_amdgcn_mfma_f32_32x32x16_fp8_fp8(pairs_to_fp8(zt, i), bv, aT, 0, 0, 0);
                    }
                }
            } else {
                const uint16_t* __restrict__ bp = B + col * rs + blk * 256 + 8 * half;
                bf16x8_t bq[16];
#pragma unroll
                for (int st = 0; st < 2; ++st) bq[st] = *(const bf16x8_t*)(bp + st * 16);
#pragma unroll
                for (int st = 0; st < 16; ++st) {
                    const uint32_t k0 = blk * 256 + st * 16;
                    if (st + 2 < 16) bq[st + 2] = *(const bf16x8_t*)(bp + (st + 2) * 16);
                    if (k0 < kp) {
                        const uint32_t z = cw[st >> 1], i0 = 8 * (st & 1) + 4 * half;
                        const bf16x8_t bv = bq[st];
                        const bool onlyw = k0 + 16 <= lbmin, onlyg = k0 >= ubmax;
                        const bool onlyt = k0 >= lbmax && k0 + 16 <= ubmin;
                        const bool split = !(onlyw || onlyg || onlyt);
                        uint32_t zw = z, zg = z, zt = z;
                        if (split) {
                            const int kb = (int)(k0 + 8 * half);
                            const uint32_t mw = pairs_below((int)lb, kb, i0);
                            const uint32_t mg = ~pairs_below((int)ub, kb, i0);
                            zw = z & mw;
                            zg = z & mg;
                            zt = z & ~(mw | mg);
                        }
                        if (onlyw || split)
                            aW = __builtin_amdgcn_mfma_f32_32x32x16_bf16(pairs_to_bf16(zw, i0), bv, aW, 0, 0, 0);
                        if (onlyg || split)
                            aG = __builtin_amdgcn_mfma_f32_32x32x16_bf16(pairs_to_bf16(zg, i0), bv, aG, 0, 0, 0);
                        if (onlyt || split)
                            aT = __builtin_amdgcn_mfma_f32_32x32x16_bf16(pairs_to_bf16(zt, i0), bv, aT, 0, 0, 0);
                    }
                }
            }
            // the next block's words are waited for here, not at their first use
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                asm volatile("" : "+v"(nw[j]));
                cw[j] = nw[j];
            }
        }
        // D layout: element e of lane l is row 8*(e/4) + 4*(l/32) + e%4, column l%32
        if (col < cnt) {
#pragma unroll
            for (int e = 0; e < 16; ++e) {
                const uint32_t i = mt * 32 + 8 * (e / 4) + 4 * half + (e % 4);
                if (i < mp) {
                    const float val = osc * (nd.win * aW[e] + nd.lose * aG[e] + nd.tie * aT[e]);
                    if (staged) {
                        O[col * L.os + pidx[i]] = val;
                    } else {
                        out[pidx[i]] = val;
                        vm = fmaxf(vm, fabsf(val));
                    }
                }
            }
        }
    }
    if (staged) {
        // whole rows, in hand order, one wavefront per node
        __syncthreads();
        for (uint32_t n = wave; n < cnt; n += kMmaWaves) {
            const uint32_t slot = nodes[n].cfv;
            float* __restrict__ orow = S.cfv + (size_t)slot * T.nh_max;
            const float* src = O + n * L.os;
            float m = 0.0f;
            for (uint32_t h = lane; h < nhp; h += kWave) {
                const float v = src[h];
                orow[h] = v;
                m = fmaxf(m, fabsf(v));
            }
            if (S.vmax) {
                m = wave_max_f32(m);
                if (lane == 0) S.vmax[slot] = m;
            }
        }
    } else if (S.vmax) {
        if (col < cnt) atomicMax(&vmx[col], __float_as_uint(vm));
        __syncthreads();
        if (tid < cnt) S.vmax[nodes[tid].cfv] = __uint_as_float(vmx[tid]);
    }
}

// ---------------------------------------------------------------------------
// Up sweep, chance nodes (cfr.rs:637-663): sum the children's values over
// the dealt cards in card order, skipping hands that hold the card, then
// scale by 1/(46 - street). One workgroup of kChanceBlock threads per node
// (a node reads up to 49 child rows; there are few chance nodes per level,
// so a hand per thread keeps the most loads in flight); loads are issued in
// batches of 8 children ahead of the in-order adds.
// ---------------------------------------------------------------------------
constexpr int kChanceBlock = 1024;
static_assert(2 * kChanceBlock >= 1326, "k_up_chance holds at most two hands per thread");
template <bool H16 = false>
__global__ __launch_bounds__(kChanceBlock) void k_up_chance(DevTree T, DevState S, const ChanceRec* __restrict__ recs,
                                                            uint32_t count, int p) {
    const uint32_t w = xcd_slice(blockIdx.x, gridDim.x);
    if (w >= count) return;
    const int o = 1 - p;
    const ChanceRec r = recs[w];
    const uint32_t nhp = T.nh[p];
    if (!S.alive[o][pick(r.own, o)]) {
        const RowOut<H16> out(S.cfv, (size_t)r.cfv * T.nh_max, RowEnc(0.0f));
        for (uint32_t i = threadIdx.x; i < nhp; i += kChanceBlock) out.set(i, 0.0f);
        if (S.vmax && threadIdx.x == 0) S.vmax[r.cfv] = 0.0f;
        if (H16 && threadIdx.x == 0) S.vsc[r.cfv] = 1.0f;
        return;
    }
    float vm = 0.0f;
    const uint8_t* __restrict__ cards = T.chr_card + r.off;
    const uint32_t* __restrict__ rows = T.chr_cfv + r.off;
    const uint8_t* __restrict__ perms = T.chr_perm + r.off;
    const uint16_t* __restrict__ hperm = T.hand_perm[p];
    // entry ix of child t's row
    auto child = [&](uint32_t t, uint32_t ix) -> float {
        const size_t e = (size_t)rows[t] * T.nh_max + ix;
        if constexpr (H16) return h2f(((const uint16_t*)S.cfv)[e]) * S.vsc[rows[t]];
        else return S.cfv[e];
    };
    // under suit isomorphism entry t reads its representative's row at the
    // image of hand i under the entry's symmetry (identity: i itself;
    // cfr.rs:641-656); without it every entry reads its own child's row at i
    auto at = [&](uint32_t t, uint32_t i) -> uint32_t {
        if (!hperm) return i;
        return perms[t] ? hperm[(size_t)perms[t] * nhp + i] : i;
    };
    // f16 rows: the (at most two) values of this thread wait for the row's
    // maximum, which sets the row's scale; f32 rows are written at once
    float x0 = 0.0f, x1 = 0.0f;
    const RowOut<false> out32(S.cfv, (size_t)r.cfv * T.nh_max, RowEnc(0.0f));
    for (uint32_t i = threadIdx.x; i < nhp; i += kChanceBlock) {
        const uint64_t hm = T.hmask[p][i];
        float acc = 0.0f;
        uint32_t t = 0;
        if (!hperm) {
            for (; t + 8 <= r.cnt; t += 8) {
                float v[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) v[u] = child(t + u, i);
#pragma unroll
                for (int u = 0; u < 8; ++u)
                    if (!((hm >> cards[t + u]) & 1ull)) acc += v[u];
            }
        } else {
            for (; t + 8 <= r.cnt; t += 8) {
                uint32_t ix[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) ix[u] = at(t + u, i);
                float v[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) v[u] = child(t + u, ix[u]);
#pragma unroll
                for (int u = 0; u < 8; ++u)
                    if (!((hm >> cards[t + u]) & 1ull)) acc += v[u];
            }
        }
        for (; t < r.cnt; ++t)
            if (!((hm >> cards[t]) & 1ull)) acc += child(t, at(t, i));
        const float x = acc * r.inv;
        if constexpr (H16) {
            if (i < kChanceBlock) x0 = x;
            else x1 = x;
        } else {
            out32.set(i, x);
        }
        vm = fmaxf(vm, fabsf(x));
    }
    if (S.vmax) {
        constexpr int kW = kChanceBlock / kWave;
        __shared__ float wm[kW];
        vm = wave_max_f32(vm);
        if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = vm;
        __syncthreads();
        float m = wm[0];
        for (int k = 1; k < kW; ++k) m = fmaxf(m, wm[k]);
        if (threadIdx.x == 0) S.vmax[r.cfv] = m;
        if constexpr (H16) {
            const RowEnc enc(m);
            const RowOut<true> out(S.cfv, (size_t)r.cfv * T.nh_max, enc);
            if (threadIdx.x < nhp) out.set(threadIdx.x, x0);
            if (threadIdx.x + kChanceBlock < nhp) out.set(threadIdx.x + kChanceBlock, x1);
            if (threadIdx.x == 0) S.vsc[r.cfv] = enc.sc;
        }
    }
}

// ---------------------------------------------------------------------------
// Up sweep, action nodes.
//   MODE 0 (DCFR/CFR+ iteration): at the traverser's nodes
//     out = sum_a sigma_a v_a, then R <- R*(R>0 ? pos : neg) + (v_a - out),
//     S <- S*strat + reach*sigma_a (cfr.rs:297-336, 493-502); compressed
//     arenas decode, update, take the node's block max and re-quantize
//     (cfr.rs:350-385). Locked nodes play the lock without updating.
//     Opponent nodes sum their children.
//   MODE 1 (best response): max over children at the traverser's nodes,
//     lock-weighted where locked (best_response.rs:67-95).
//   MODE 2 (average strategy values): average-strategy weighting
//     (best_response.rs:180-195).
//   MODE 3 (current strategy values): regret-matched weighting, no update.
//   MODE 4 (game value): the strategy arena taken as the behavioural
//     strategy itself (kRawF32), no update.
// Subtrees whose opponent reach is all zero produce zeros and no update.
// ---------------------------------------------------------------------------
// Hands per lane in one step of the f32 update at a node with absorbed
// opponent children (up to kMaxAbsorb rows per action are read per hand;
// flop SRP update per iteration: two hands 4.98 ms, three 4.97, four 4.94).
constexpr int kStepAbsorb = 4;

// Value sources of node r's actions when its opponent children are absorbed
// (records.h): per action a, kMaxAbsorb f32 rows whose sum, taken from 0 in
// order, is the child's value (the child's own row and zero rows, or the
// absorbed child's children's rows and zero rows, or zero rows when the
// opponent's reach at the absorbed child is dead; adding 0 changes no sum).
// Returns max over actions of a bound of |value| (sums of the rows' recorded
// maxima) when S.vmax is kept, else 0.
template <int NA>
__device__ __forceinline__ float absorbed_sources(const DevTree& T, const DevState& S, const ActRec& r, int o,
                                                  const float* (&src)[NA][kMaxAbsorb]) {
    const uint32_t* __restrict__ xk = T.xkids + r.xk;
    const uint32_t* kid_cfv = T.kids + r.kid;
    const size_t nh_max = T.nh_max;
    uint32_t xp = 0;
    float mv = 0.0f;
#pragma unroll
    for (int a = 0; a < NA; ++a) {
        const uint32_t c = xk[xp++];
        float bound = 0.0f;
        if (c == 0) {
            src[a][0] = S.cfv + (size_t)kid_cfv[a] * nh_max;
#pragma unroll
            for (int b = 1; b < kMaxAbsorb; ++b) src[a][b] = T.zero_row;
            if (S.vmax) bound = S.vmax[kid_cfv[a]];
        } else {
            const bool live = S.alive[o][xk[xp++]] != 0;
#pragma unroll
            for (int b = 0; b < kMaxAbsorb; ++b) {
                const bool use = live && (uint32_t)b < c;
                src[a][b] = use ? S.cfv + (size_t)xk[xp + b] * nh_max : T.zero_row;
                if (S.vmax && use) bound += S.vmax[xk[xp + b]];
            }
            xp += c;
        }
        mv = fmaxf(mv, bound);
    }
    return mv;
}

template <int MODE, int STORE, int ROWS = 0>
// Returns this lane's max |value| over the row entries it wrote. [hbeg,
// hend): the traverser hands this wavefront takes (a slice of the row; the
// whole row by default; slices start at multiples of 64).
__device__ __forceinline__ float up_action_node(const DevTree& T, const DevState& S, const ActRec& r, int p, int q,
                                                uint32_t lane, int absorb = 0, uint32_t hbeg = 0,
                                                uint32_t hend = 0xFFFFFFFFu) {
    constexpr bool R16 = (ROWS & kRows16Reach) != 0, V16 = (ROWS & kRows16Value) != 0;
    static_assert(ROWS == 0 || (MODE == 0 && STORE == kStoreH), "f16 rows serve the DCFR sweeps of f16 arenas");
    float vm = 0.0f;
    const int o = 1 - p;
    const uint32_t nhp = T.nh[p];
    const uint32_t h1 = min(hend, nhp);  // end of this wavefront's hands
    const uint32_t nh_max = T.nh_max;
    const int na = r.na;
    const uint32_t* kid_cfv = T.kids + r.kid;
    if (!S.alive[o][pick(r.own, o)]) {
        const RowOut<V16> z(S.cfv, (size_t)r.cfv * nh_max, RowEnc(0.0f));
        for (uint32_t i = hbeg + lane; i < h1; i += kWave) z.set(i, 0.0f);
        if (V16 && lane == 0) S.vsc[r.cfv] = 1.0f;
        return 0.0f;
    }
    // f16 rows: |value| <= the sum of the children's row maxima (opponent
    // and locked nodes; a lock's weights are <= the lock bound), or their
    // max (the traverser's regret-matched sigma sums to 1)
    float vb = 0.0f;
    if constexpr (V16) {
        const bool sum = q != p || r.lock_off >= 0;
        for (int a = 0; a < na; ++a) {
            const float m = S.vmax[kid_cfv[a]];
            vb = sum ? vb + m : fmaxf(vb, m);
        }
        if (r.lock_off >= 0) vb *= T.lock_bound;
    }
    const RowEnc enc(vb);
    const RowOut<V16> out(S.cfv, (size_t)r.cfv * nh_max, enc);
    if (V16 && lane == 0) S.vsc[r.cfv] = enc.sc;
    auto row = [&](int a) { return cfv_in<V16>(T, S, kid_cfv[a]); };
    if (q != p || r.lock_off >= 0 || MODE == 1) {
        // opponent node: sum; locked node: lock-weighted; best response: max
        const float* sig = r.lock_off >= 0 ? T.lock_sigma + r.lock_off : nullptr;
        const int kind = q != p ? 0 : (sig ? 1 : 2);
        dispatch_na4(na, [&](auto NAc) {
            constexpr int NA = decltype(NAc)::value;
            if constexpr (NA > 0) {
                if constexpr (V16) {
                    if (kind == 0 && ((nhp | nh_max) & 1u) == 0u) {
                        // opponent node, f16 rows, hands in pairs: every row
                        // of the node starts at an even entry; the sums are
                        // the per-hand ones below
                        const uint32_t np = nhp >> 1;
                        const uint32_t* __restrict__ V2[NA];
                        float vs[NA];
#pragma unroll
                        for (int a = 0; a < NA; ++a) {
                            V2[a] = (const uint32_t*)((const uint16_t*)S.cfv + (size_t)kid_cfv[a] * nh_max);
                            vs[a] = S.vsc[kid_cfv[a]];
                        }
                        uint32_t* __restrict__ O2 = (uint32_t*)((uint16_t*)S.cfv + (size_t)r.cfv * nh_max);
                        for (uint32_t j0 = lane; j0 < np; j0 += kSumHands * kWave) {
                            uint32_t w[kSumHands][NA];
#pragma unroll
                            for (int h = 0; h < kSumHands; ++h) {
                                const uint32_t jj = min(j0 + h * kWave, np - 1);
#pragma unroll
                                for (int a = 0; a < NA; ++a) w[h][a] = V2[a][jj];
                            }
#pragma unroll
                            for (int h = 0; h < kSumHands; ++h) {
                                const uint32_t j = j0 + h * kWave;
                                if (j >= np) break;
                                uint32_t o2 = 0u;
#pragma unroll
                                for (int half = 0; half < 2; ++half) {
                                    float acc = 0.0f;
#pragma unroll
                                    for (int a = 0; a < NA; ++a) acc += h2f((uint16_t)(w[h][a] >> (16 * half))) * vs[a];
                                    o2 |= (uint32_t)f2h(acc * out.inv) << (16 * half);
                                    vm = fmaxf(vm, fabsf(acc));
                                }
                                O2[j] = o2;
                            }
                        }
                        return;
                    }
                }
                RowIn<V16> rw[NA];
#pragma unroll
                for (int a = 0; a < NA; ++a) rw[a] = row(a);
                for (uint32_t i0 = hbeg + lane; i0 < h1; i0 += kSumHands * kWave) {
                    float v[kSumHands][NA];
#pragma unroll
                    for (int h = 0; h < kSumHands; ++h) {
                        const uint32_t ii = min(i0 + h * kWave, nhp - 1);
#pragma unroll
                        for (int a = 0; a < NA; ++a) v[h][a] = rw[a][ii];
                    }
#pragma unroll
                    for (int h = 0; h < kSumHands; ++h) {
                        const uint32_t i = i0 + h * kWave;
                        if (i >= h1) break;
                        float acc = 0.0f;
                        if (kind == 0) {
#pragma unroll
                            for (int a = 0; a < NA; ++a) acc += v[h][a];
                        } else if (kind == 1) {
#pragma unroll
                            for (int a = 0; a < NA; ++a) acc += sig[(size_t)a * nhp + i] * v[h][a];
                        } else {
                            acc = v[h][0];
#pragma unroll
                            for (int a = 1; a < NA; ++a) acc = v[h][a] > acc ? v[h][a] : acc;
                        }
                        out.set(i, acc);
                    vm = fmaxf(vm, fabsf(acc));
                    }
                }
            } else {
                for (uint32_t i = hbeg + lane; i < h1; i += kWave) {
                    float acc = kind == 2 ? row(0)[i] : 0.0f;
                    for (int a = 0; a < na; ++a) {
                        const float v = row(a)[i];
                        if (kind == 0) acc += v;
                        else if (kind == 1) acc += sig[(size_t)a * nhp + i] * v;
                        else if (a > 0 && v > acc) acc = v;
                    }
                    out.set(i, acc);
                    vm = fmaxf(vm, fabsf(acc));
                }
            }
        });
        return vm;
    }
    const uint64_t off = r.doff;
    if constexpr (MODE >= 2) {
        // value of the average (2), current (3) or given (4) strategy, no update
        static_assert(MODE != 4 || STORE == kStoreF32, "game values read an f32 strategy arena");
        constexpr int KIND = MODE == 4 ? kRawF32
                             : MODE == 2
                                 ? (STORE == kStoreQ ? kAverageQ : STORE == kStoreH ? kAverageH : kAverageF32)
                                 : (STORE == kStoreQ ? kCurrentQ : STORE == kStoreH ? kCurrentH : kCurrentF32);
        const auto* arena = ArenaOf<KIND>::get(S, p) + off;
        dispatch_na4(na, [&](auto NAc) {
            constexpr int NA = decltype(NAc)::value;
            if constexpr (NA > 0) {
                for (uint32_t i = hbeg + lane; i < h1; i += kWave) {
                    HandSigma<NA, KIND> hs;
                    hs.load(arena, nhp, i);
                    float acc = 0.0f;
#pragma unroll
                    for (int a = 0; a < NA; ++a) acc += hs.sig[a] * row(a)[i];
                    out.set(i, acc);
                    vm = fmaxf(vm, fabsf(acc));
                }
            } else {
                for (uint32_t i = hbeg + lane; i < h1; i += kWave) {
                    SigmaHand<KIND> sh;
                    sh.init(arena, nhp, i, na);
                    float acc = 0.0f;
                    for (int a = 0; a < na; ++a) acc += sh.template sigma<false>(a) * row(a)[i];
                    out.set(i, acc);
                    vm = fmaxf(vm, fabsf(acc));
                }
            }
        });
        return vm;
    } else {
        const uint32_t t = *S.iter;
        const float pos = S.disc[3 * t + 0], neg = S.disc[3 * t + 1], ds = S.disc[3 * t + 2];
        const RowIn<R16> rp = reach_in<R16>(T, S, p, pick(r.own, p));
        if constexpr (STORE == kStoreF32 || STORE == kStoreF32P) {
            // PCFR+ (kStoreF32P): sigma from Q = max(R,0) + prediction, and
            // the update also stores Q' = max(R',0) + (v - out), the next
            // iteration's matching input (cfr.rs:259-266, 337-348)
            constexpr bool PRED = STORE == kStoreF32P;
            float* __restrict__ R = S.R[p] + off;
            float* __restrict__ Sm = S.S[p] + off;
            float* __restrict__ Qp = PRED ? S.Q[p] + off : nullptr;
            if constexpr (!PRED && ROWS == 0) {
                if (absorb && r.xk != kNoAbsorb) {
                    // absorbed opponent children: the child's value is the sum
                    // of its own children's rows, added in action order from 0
                    // as the opponent-node sum does (padding rows are zeros,
                    // which leave the sum unchanged); a dead opponent reach
                    // at the child gives 0
                    bool done = false;
                    dispatch_na4(na, [&](auto NAc) {
                        constexpr int NA = decltype(NAc)::value;
                        if constexpr (NA > 0) {
                            done = true;
                            const float* src[NA][kMaxAbsorb];
                            (void)absorbed_sources<NA>(T, S, r, o, src);
                            for (uint32_t i0 = hbeg + lane; i0 < h1; i0 += kStepAbsorb * kWave) {
                                HandSigma<NA, kCurrentF32> hs[kStepAbsorb];
                                float x[kStepAbsorb][NA][kMaxAbsorb], sm[kStepAbsorb][NA], reach[kStepAbsorb];
                                bool ok[kStepAbsorb];
#pragma unroll
                                for (int h = 0; h < kStepAbsorb; ++h) {
                                    const uint32_t i = i0 + h * kWave;
                                    ok[h] = i < h1;
                                    const uint32_t ii = ok[h] ? i : i0;
#pragma unroll
                                    for (int a = 0; a < NA; ++a) {
                                        hs[h].raw[a] = R[(size_t)a * nhp + ii];
                                        sm[h][a] = Sm[(size_t)a * nhp + ii];
#pragma unroll
                                        for (int b = 0; b < kMaxAbsorb; ++b) x[h][a][b] = src[a][b][ii];
                                    }
                                    reach[h] = rp[ii];
                                }
#pragma unroll
                                for (int h = 0; h < kStepAbsorb; ++h) {
                                    if (!ok[h]) continue;
                                    const uint32_t i = i0 + h * kWave;
                                    float v[NA];
#pragma unroll
                                    for (int a = 0; a < NA; ++a) {
                                        float t = 0.0f;
#pragma unroll
                                        for (int b = 0; b < kMaxAbsorb; ++b) t += x[h][a][b];
                                        v[a] = t;
                                    }
                                    hs[h].finish();
                                    float acc = 0.0f;
#pragma unroll
                                    for (int a = 0; a < NA; ++a) acc += hs[h].sig[a] * v[a];
#pragma unroll
                                    for (int a = 0; a < NA; ++a) {
                                        const float rv = hs[h].raw[a];
                                        const float d = rv > 0.0f ? pos : neg;
                                        R[(size_t)a * nhp + i] = rv * d + (v[a] - acc);
                                        Sm[(size_t)a * nhp + i] = sm[h][a] * ds + reach[h] * hs[h].sig[a];
                                    }
                                    out.set(i, acc);
                                    vm = fmaxf(vm, fabsf(acc));
                                }
                            }
                        }
                    });
                    if (done) return vm;
                }
            }
            dispatch_na4(na, [&](auto NAc) {
                constexpr int NA = decltype(NAc)::value;
                if constexpr (NA > 0) {
                    RowIn<V16> rw[NA];
#pragma unroll
                    for (int a = 0; a < NA; ++a) rw[a] = row(a);
                    for (uint32_t i0 = hbeg + lane; i0 < h1; i0 += kStepHands * kWave) {
                        HandSigma<NA, kCurrentF32> hs[kStepHands];
                        float v[kStepHands][NA], sm[kStepHands][NA], reach[kStepHands], rr[kStepHands][PRED ? NA : 1];
                        bool ok[kStepHands];
#pragma unroll
                        for (int h = 0; h < kStepHands; ++h) {
                            const uint32_t i = i0 + h * kWave;
                            ok[h] = i < h1;
                            const uint32_t ii = ok[h] ? i : i0;
#pragma unroll
                            for (int a = 0; a < NA; ++a) {
                                if constexpr (PRED) {
                                    hs[h].raw[a] = Qp[(size_t)a * nhp + ii];
                                    rr[h][a] = R[(size_t)a * nhp + ii];
                                } else {
                                    hs[h].raw[a] = R[(size_t)a * nhp + ii];
                                }
                                v[h][a] = rw[a][ii];
                                sm[h][a] = Sm[(size_t)a * nhp + ii];
                            }
                            reach[h] = rp[ii];
                        }
#pragma unroll
                        for (int h = 0; h < kStepHands; ++h) {
                            if (!ok[h]) continue;
                            const uint32_t i = i0 + h * kWave;
                            hs[h].finish();
                            float acc = 0.0f;
#pragma unroll
                            for (int a = 0; a < NA; ++a) acc += hs[h].sig[a] * v[h][a];
#pragma unroll
                            for (int a = 0; a < NA; ++a) {
                                const float rv = PRED ? rr[h][PRED ? a : 0] : hs[h].raw[a];
                                const float d = rv > 0.0f ? pos : neg;
                                const float nr = rv * d + (v[h][a] - acc);
                                R[(size_t)a * nhp + i] = nr;
                                Sm[(size_t)a * nhp + i] = sm[h][a] * ds + reach[h] * hs[h].sig[a];
                                if constexpr (PRED) Qp[(size_t)a * nhp + i] = fmaxf(nr, 0.0f) + (v[h][a] - acc);
                            }
                            out.set(i, acc);
                    vm = fmaxf(vm, fabsf(acc));
                        }
                    }
                } else {
                    for (uint32_t i = hbeg + lane; i < h1; i += kWave) {
                        SigmaHand<kCurrentF32> sh;
                        sh.init(PRED ? Qp : R, nhp, i, na);
                        float acc = 0.0f;
                        for (int a = 0; a < na; ++a) acc += sh.template sigma<false>(a) * row(a)[i];
                        const float reach = rp[i];
                        for (int a = 0; a < na; ++a) {
                            const size_t idx = (size_t)a * nhp + i;
                            const float rv = R[idx];
                            const float sg = sh.template sigma<false>(a);
                            const float d = rv > 0.0f ? pos : neg;
                            const float nr = rv * d + (row(a)[i] - acc);
                            R[idx] = nr;
                            Sm[idx] = Sm[idx] * ds + reach * sg;
                            if constexpr (PRED) Qp[idx] = fmaxf(nr, 0.0f) + (row(a)[i] - acc);
                        }
                        out.set(i, acc);
                    vm = fmaxf(vm, fabsf(acc));
                    }
                }
            });
        } else if constexpr (STORE == kStoreH) {
            constexpr int kStepH = ROWS ? kStepHands16 : kStepHands;
            // f16 arenas with a power-of-two scale per node (no counterpart in
            // the reference). The new scales come from bounds of the new
            // maxima, known before any value is read, so every value is
            // read and written once:
            //   |R'| <= max|R| max(pos, neg) + |v - out| <= mR max(pos,neg) + 2 max|v|
            //   S' = S ds + reach sigma <= mS ds + max reach
            // with max|v| the children's row maxima, recorded by the kernels
            // that wrote those rows.
            // Decoding multiplies by a power of two (exact); the f32
            // arithmetic is the f32 engine's on the decoded values.
            uint16_t* __restrict__ Rh = (uint16_t*)S.Rq[p] + off;
            uint16_t* __restrict__ Sh = S.Sq[p] + off;
            const float sR = S.Rs[p][r.node], sS = S.Ss[p][r.node];
            // max |v| over the children's rows (recorded by their writers);
            // reach * sigma <= max reach (recorded by the down sweep)
            float mv = 0.0f;
            for (int a = 0; a < na; ++a) mv = fmaxf(mv, S.vmax[kid_cfv[a]]);
            const float mr = S.rmax[p][pick(r.own, p)];
            const float kR = half_scale(S.Rm[p][r.node] * fmaxf(pos, neg) + 2.0f * mv);
            const float kS = half_scale(S.Sm[p][r.node] * ds + mr);
            const float iR = 1.0f / kR, iS = 1.0f / kS;
            float xr = 0.0f, xs = 0.0f;
            dispatch_na4(na, [&](auto NAc) {
                constexpr int NA = decltype(NAc)::value;
                if constexpr (NA > 0) {
                    if constexpr (R16 && V16) {
                        if (((nhp | nh_max) & 1u) == 0u) {
                            // f16 rows, hands in pairs (every arena segment and
                            // row of the node starts at an even entry); each
                            // hand's arithmetic is the per-hand loop's below
                            const uint32_t np = nhp >> 1;
                            uint32_t* __restrict__ R2 = (uint32_t*)Rh;
                            uint32_t* __restrict__ S2 = (uint32_t*)Sh;
                            const uint32_t* __restrict__ V2[NA];
                            float vs[NA];
#pragma unroll
                            for (int a = 0; a < NA; ++a) {
                                V2[a] = (const uint32_t*)((const uint16_t*)S.cfv + (size_t)kid_cfv[a] * nh_max);
                                vs[a] = S.vsc[kid_cfv[a]];
                            }
                            const uint32_t* __restrict__ P2 = (const uint32_t*)rp.p;
                            uint32_t* __restrict__ O2 = (uint32_t*)((uint16_t*)S.cfv + (size_t)r.cfv * nh_max);
                            for (uint32_t j0 = lane; j0 < np; j0 += kStepPairs * kWave) {
                                uint32_t rr[kStepPairs][NA], ss[kStepPairs][NA], ww[kStepPairs][NA], pp[kStepPairs];
                                bool ok[kStepPairs];
#pragma unroll
                                for (int h = 0; h < kStepPairs; ++h) {
                                    const uint32_t j = j0 + h * kWave;
                                    ok[h] = j < np;
                                    const uint32_t jj = ok[h] ? j : j0;
#pragma unroll
                                    for (int a = 0; a < NA; ++a) {
                                        rr[h][a] = R2[(size_t)a * np + jj];
                                        ss[h][a] = S2[(size_t)a * np + jj];
                                        ww[h][a] = V2[a][jj];
                                    }
                                    pp[h] = P2[jj];
                                }
#pragma unroll
                                for (int h = 0; h < kStepPairs; ++h) {
                                    if (!ok[h]) continue;
                                    const uint32_t j = j0 + h * kWave;
                                    uint32_t nr2[NA], ns2[NA], o2 = 0u;
#pragma unroll
                                    for (int a = 0; a < NA; ++a) nr2[a] = ns2[a] = 0u;
#pragma unroll
                                    for (int half = 0; half < 2; ++half) {
                                        const int sh = 16 * half;
                                        HandSigma<NA, kCurrentH> hs;
#pragma unroll
                                        for (int a = 0; a < NA; ++a) hs.raw[a] = h2f((uint16_t)(rr[h][a] >> sh));
                                        hs.finish();
                                        float v[NA];
#pragma unroll
                                        for (int a = 0; a < NA; ++a) v[a] = h2f((uint16_t)(ww[h][a] >> sh)) * vs[a];
                                        const float reach = h2f((uint16_t)(pp[h] >> sh)) * rp.sc;
                                        float acc = 0.0f;
#pragma unroll
                                        for (int a = 0; a < NA; ++a) acc += hs.sig[a] * v[a];
#pragma unroll
                                        for (int a = 0; a < NA; ++a) {
                                            const float rv = hs.raw[a] * sR;
                                            const float d = rv > 0.0f ? pos : neg;
                                            const float nr = rv * d + (v[a] - acc);
                                            const float ns = (h2f((uint16_t)(ss[h][a] >> sh)) * sS) * ds + reach * hs.sig[a];
                                            nr2[a] |= (uint32_t)f2h(nr * iR) << sh;
                                            ns2[a] |= (uint32_t)f2h(ns * iS) << sh;
                                            xr = fmaxf(xr, fabsf(nr));
                                            xs = fmaxf(xs, ns);
                                        }
                                        o2 |= (uint32_t)f2h(acc * out.inv) << sh;
                                        vm = fmaxf(vm, fabsf(acc));
                                    }
#pragma unroll
                                    for (int a = 0; a < NA; ++a) {
                                        R2[(size_t)a * np + j] = nr2[a];
                                        S2[(size_t)a * np + j] = ns2[a];
                                    }
                                    O2[j] = o2;
                                }
                            }
                            return;
                        }
                    }
                    RowIn<V16> rw[NA];
#pragma unroll
                    for (int a = 0; a < NA; ++a) rw[a] = row(a);
                    for (uint32_t i0 = hbeg + lane; i0 < h1; i0 += kStepH * kWave) {
                        HandSigma<NA, kCurrentH> hs[kStepH];
                        float v[kStepH][NA], sm[kStepH][NA], reach[kStepH];
                        bool ok[kStepH];
#pragma unroll
                        for (int h = 0; h < kStepH; ++h) {
                            const uint32_t i = i0 + h * kWave;
                            ok[h] = i < h1;
                            const uint32_t ii = ok[h] ? i : i0;
                            hs[h].fetch(Rh, nhp, ii);
#pragma unroll
                            for (int a = 0; a < NA; ++a) {
                                v[h][a] = rw[a][ii];
                                sm[h][a] = h2f(Sh[(size_t)a * nhp + ii]);
                            }
                            reach[h] = rp[ii];
                        }
#pragma unroll
                        for (int h = 0; h < kStepH; ++h) {
                            if (!ok[h]) continue;
                            const uint32_t i = i0 + h * kWave;
                            hs[h].finish();
                            float acc = 0.0f;
#pragma unroll
                            for (int a = 0; a < NA; ++a) acc += hs[h].sig[a] * v[h][a];
#pragma unroll
                            for (int a = 0; a < NA; ++a) {
                                const float rv = hs[h].raw[a] * sR;
                                const float d = rv > 0.0f ? pos : neg;
                                const float nr = rv * d + (v[h][a] - acc);
                                const float ns = (sm[h][a] * sS) * ds + reach[h] * hs[h].sig[a];
                                Rh[(size_t)a * nhp + i] = f2h(nr * iR);
                                Sh[(size_t)a * nhp + i] = f2h(ns * iS);
                                xr = fmaxf(xr, fabsf(nr));
                                xs = fmaxf(xs, ns);
                            }
                            out.set(i, acc);
                    vm = fmaxf(vm, fabsf(acc));
                        }
                    }
                } else {
                    for (uint32_t i = hbeg + lane; i < h1; i += kWave) {
                        SigmaHand<kCurrentH> sh;
                        sh.init(Rh, nhp, i, na);
                        float acc = 0.0f;
                        for (int a = 0; a < na; ++a) acc += sh.template sigma<false>(a) * row(a)[i];
                        const float reach = rp[i];
                        for (int a = 0; a < na; ++a) {
                            const size_t idx = (size_t)a * nhp + i;
                            const float rv = h2f(Rh[idx]) * sR;
                            const float sg = sh.template sigma<false>(a);
                            const float d = rv > 0.0f ? pos : neg;
                            const float nr = rv * d + (row(a)[i] - acc);
                            const float ns = (h2f(Sh[idx]) * sS) * ds + reach * sg;
                            Rh[idx] = f2h(nr * iR);
                            Sh[idx] = f2h(ns * iS);
                            xr = fmaxf(xr, fabsf(nr));
                            xs = fmaxf(xs, ns);
                        }
                        out.set(i, acc);
                    vm = fmaxf(vm, fabsf(acc));
                    }
                }
            });
            xr = wave_max_f32(xr);
            xs = wave_max_f32(xs);
            if (lane == 0) {
                S.Rs[p][r.node] = kR;
                S.Ss[p][r.node] = kS;
                S.Rm[p][r.node] = xr;
                S.Sm[p][r.node] = xs;
            }
        } else {
            // Compressed arenas (cfr.rs:350-385). Pass 1 finds the node's new
            // block maxima; pass 2 recomputes the same values
            // (deterministically) and quantizes them.
            int16_t* __restrict__ Rq = S.Rq[p] + off;
            uint16_t* __restrict__ Sq = S.Sq[p] + off;
            const float kr = S.Rs[p][r.node] / 32767.0f;
            const float kss = S.Ss[p][r.node] / 65535.0f * ds;
            float rmax = 0.0f, smax = 0.0f;
            float RM = 0.0f, SM = 0.0f, rk = 0.0f, sk = 0.0f;
            bool rok = false, sok = false;
            dispatch_na4(na, [&](auto NAc) {
                constexpr int NA = decltype(NAc)::value;
                auto hand = [&](uint32_t i, auto&& sink) -> float {
                    if constexpr (NA > 0) {
                        HandSigma<NA, kCurrentQ> hs;
                        hs.load(Rq, nhp, i);
                        float v[NA], sq[NA];
#pragma unroll
                        for (int a = 0; a < NA; ++a) {
                            v[a] = row(a)[i];
                            sq[a] = (float)Sq[(size_t)a * nhp + i];
                        }
                        const float reach = rp[i];
                        float acc = 0.0f;
#pragma unroll
                        for (int a = 0; a < NA; ++a) acc += hs.sig[a] * v[a];
#pragma unroll
                        for (int a = 0; a < NA; ++a) {
                            const int qv = hs.raw[a];
                            const float d = qv > 0 ? pos : neg;
                            const float nr = (float)qv * kr * d + (v[a] - acc);
                            const float ns = sq[a] * kss + reach * hs.sig[a];
                            sink((size_t)a * nhp + i, nr, ns);
                        }
                        return acc;
                    } else {
                        SigmaHand<kCurrentQ> sh;
                        sh.init(Rq, nhp, i, na);
                        float acc = 0.0f;
                        for (int a = 0; a < na; ++a) acc += sh.template sigma<false>(a) * row(a)[i];
                        const float reach = rp[i];
                        for (int a = 0; a < na; ++a) {
                            const size_t idx = (size_t)a * nhp + i;
                            const int qv = Rq[idx];
                            const float d = qv > 0 ? pos : neg;
                            const float nr = (float)qv * kr * d + (row(a)[i] - acc);
                            const float ns = (float)Sq[idx] * kss + reach * sh.template sigma<false>(a);
                            sink(idx, nr, ns);
                        }
                        return acc;
                    }
                };
                for (uint32_t i = hbeg + lane; i < h1; i += kWave) {
                    const float acc = hand(i, [&](size_t, float nr, float ns) {
                        rmax = fmaxf(rmax, fabsf(nr));
                        smax = fmaxf(smax, ns);
                    });
                    out.set(i, acc);
                    vm = fmaxf(vm, fabsf(acc));
                }
                RM = wave_max_f32(rmax);
                SM = wave_max_f32(smax);
                rok = RM > 0.0f && isfinite(RM);
                sok = SM > 0.0f && isfinite(SM);
                rk = rok ? 32767.0f / RM : 0.0f;
                sk = sok ? 65535.0f / SM : 0.0f;
                for (uint32_t i = hbeg + lane; i < h1; i += kWave) {
                    hand(i, [&](size_t idx, float nr, float ns) {
                        Rq[idx] = rok ? quant_i16(nr * rk) : (int16_t)0;
                        Sq[idx] = sok ? quant_u16(fmaxf(ns, 0.0f) * sk) : (uint16_t)0;
                    });
                }
            });
            if (lane == 0) {
                S.Rs[p][r.node] = rok ? RM : 0.0f;
                S.Ss[p][r.node] = sok ? SM : 0.0f;
            }
        }
    }
    return vm;
}

template <int MODE, int STORE, int ROWS = 0>
__global__ __launch_bounds__(kBlock) void k_up_action(DevTree T, DevState S, const ActRec* __restrict__ recs,
                                                      uint32_t count, int p, int q) {
    const uint32_t w = wave_item();
    if (w >= count) return;
    const uint32_t lane = threadIdx.x & 63;
    const ActRec r = recs[w];
    float vm = up_action_node<MODE, STORE, ROWS>(T, S, r, p, q, lane);
    if (S.vmax) {  // the row's max |value|, for the parent's f16 scale
        vm = wave_max_f32(vm);
        if (lane == 0) S.vmax[r.cfv] = vm;
    }
}

// Both players' action nodes of one level in one launch (they read only the
// level below and write disjoint rows and arenas): wavefront m takes item
// k0(m) = floor(m c0 / (c0 + c1)) of player 0's list when k0(m+1) > k0(m),
// else item m - k0(m) of player 1's, so every stretch of the grid (each
// XCD's slice among them) holds both lists in proportion and the lighter
// opponent-node sums fill in beside the traverser's updates.
// slices > 1 (levels with few nodes, f32 arenas and rows): each node's
// traverser hands split over `slices` wavefronts, 64-aligned, so a small
// level still fills the GPU; every hand's arithmetic is unchanged.
template <int MODE, int STORE, int ROWS = 0>
__global__ __launch_bounds__(kBlock) void k_up_action2(DevTree T, DevState S, const ActRec* __restrict__ r0,
                                                       uint32_t c0, const ActRec* __restrict__ r1, uint32_t c1, int p,
                                                       int absorb, uint32_t slices) {
    // f16 and i16/u16 arenas keep per-node maxima and f16 rows per-row
    // scales, reductions over the whole row: those take whole rows
    constexpr bool kSliceable = ROWS == 0 && (STORE == kStoreF32 || STORE == kStoreF32P);
    if constexpr (!kSliceable) slices = 1;
    const uint32_t w = wave_item();
    const uint32_t m = w / slices, sl = w - m * slices;
    const uint32_t tot = c0 + c1;
    if (m >= tot) return;
    uint32_t hb = 0, he = 0xFFFFFFFFu;
    if (slices > 1) {
        const uint32_t ch = ((T.nh[p] + slices - 1) / slices + kWave - 1) & ~(uint32_t)(kWave - 1);
        hb = sl * ch;
        he = hb + ch;
        if (hb >= T.nh[p]) return;
    }
    const uint32_t k0 = (uint32_t)(((uint64_t)m * c0) / tot);
    const uint32_t k1 = (uint32_t)(((uint64_t)(m + 1) * c0) / tot);
    const int q = __builtin_amdgcn_readfirstlane(k1 > k0 ? 0 : 1);
    const uint32_t lane = threadIdx.x & 63;
    const ActRec r = q == 0 ? r0[k0] : r1[m - k0];
    float vm = up_action_node<MODE, STORE, ROWS>(T, S, r, p, q, lane, absorb, hb, he);
    if (S.vmax) {  // engines with row maxima launch whole rows (slices == 1)
        vm = wave_max_f32(vm);
        if (lane == 0) S.vmax[r.cfv] = vm;
    }
}

// ---------------------------------------------------------------------------
// Game-value gradient with respect to actor q's behavioural strategy:
// dU/dsigma(node, a, i) = reach_q(node, i) * (chance weight of the path) *
// value of child a for hand i, read from the value rows of the sweep whose
// traverser is q (payoffs: the game value's player). Locked nodes play a
// fixed strategy: 0.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_grad_sigma(DevTree T, DevState S, const ActRec* __restrict__ recs,
                                                       uint32_t count, int q, float* __restrict__ grad) {
    const uint32_t w = wave_item();
    if (w >= count) return;
    const uint32_t lane = threadIdx.x & 63;
    const ActRec r = recs[w];
    const uint32_t nh = T.nh[q];
    const int na = r.na;
    float* __restrict__ g = grad + r.doff;
    const float* __restrict__ pr = S.reach[q] + (size_t)pick(r.own, q) * nh;
    const uint32_t* kid_cfv = T.kids + r.kid;
    const float k = r.lock_off >= 0 ? 0.0f : r.cw;
    for (int a = 0; a < na; ++a) {
        const float* __restrict__ v = S.cfv + (size_t)kid_cfv[a] * T.nh_max;
        for (uint32_t i = lane; i < nh; i += kWave) g[(size_t)a * nh + i] = pr[i] * k * v[i];
    }
}

}  // namespace gto
