// MI355X DCFR engine: device state, level-synchronous sweeps, hipGraph
// replay of steady-state iterations, best-response exploitability, and the
// gto_hip_* C ABI (replacing solver::gpu::GpuSolver,
// crates/solver/src/gpu/mod.rs:60-997).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <cstdlib>
#include <exception>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../../include/gtopen_amd.h"
#include "kernels.hip"
#include "plan.h"

namespace gto {
void set_last_error(const std::string& msg);
}

namespace {

using namespace gto;

// A device allocation failed for lack of memory: reported as GTO_ERR_MEMORY,
// so a caller can tell a VRAM shortfall from a device fault.
struct DeviceOutOfMemory : std::runtime_error {
    using std::runtime_error::runtime_error;
};

#define HIP_OK(expr)                                                                                       \
    do {                                                                                                   \
        hipError_t e_ = (expr);                                                                            \
        if (e_ == hipErrorOutOfMemory)                                                                     \
            throw DeviceOutOfMemory(std::string("hip: out of device memory in " #expr));                 \
        if (e_ != hipSuccess)                                                                              \
            throw std::runtime_error(std::string("hip: ") + hipGetErrorString(e_) + " in " #expr);     \
    } while (0)

// Static device scratch of the allocation-free game value
// (gto_hip_game_value_noalloc): one player's strategy gradient is staged
// here before the copy to the caller. Every GPU the library runs on holds
// its own copy (64 MiB) until the process ends; host code reaches a
// device's copy only through the address hipGetSymbolAddress returns with
// that device current (scratch_on), never through the array's name.
constexpr uint64_t kScratchFloats = uint64_t(16) << 20;
static __device__ __align__(256) float scratch[kScratchFloats];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
float* g_scratch_ptr[kMaxDevices] = {};

// A caller's buffer that must be device memory of `device` (the
// device-resident game value): anything else is an invalid argument. A
// failed attribute query leaves its error behind, so it is cleared here
// rather than reported by a later launch check.
void require_device_ptr(const void* ptr, int device, const char* what) {
    hipPointerAttribute_t a{};
    const hipError_t r = hipPointerGetAttributes(&a, ptr);
    if (r != hipSuccess) (void)hipGetLastError();
    if (r != hipSuccess || (a.type != hipMemoryTypeDevice && a.type != hipMemoryTypeManaged) || a.device != device)
        throw std::invalid_argument(std::string(what) + ": not device memory of the engine's GPU");
}

// This device's address of `scratch` (the device must be current); looked
// up once per device, at the first engine created on it.
float* scratch_on(int device) {
    if (device < 0 || device >= kMaxDevices) throw std::runtime_error("device ordinal out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        HIP_OK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) throw std::runtime_error("hipGetSymbolAddress returned no address for the scratch array");
        g_scratch_ptr[device] = (float*)p;
    }
    return g_scratch_ptr[device];
}

enum ProfKind { PK_ROOT = 0, PK_DOWN_ACTION, PK_DOWN_CHANCE, PK_UP_FOLD, PK_UP_SHOW, PK_UP_CHANCE, PK_UP_ACTION, PK_RIVER,
                PK_OTHER };
const char* kProfNames[GTO_PROF_KINDS] = {"root", "down_action", "down_chance", "up_fold", "up_show",
                                          "up_chance", "up_action", "river", "other"};

struct DevLevel {
    ActRec* act[2] = {nullptr, nullptr};
    uint32_t n_act[2] = {0, 0};
    uint32_t max_na[2] = {0, 0};  // most actions of a node in act[q] (and act_up[q])
    // q's nodes that no node of the other player absorbs (Plan act_keep)
    ActRec* act_keep[2] = {nullptr, nullptr};
    uint32_t n_act_keep[2] = {0, 0};
    FoldRec* fold = nullptr;
    uint32_t n_fold = 0;
    ShowRec* show = nullptr;
    uint32_t n_show = 0;
    uint4* show_chunks[3] = {nullptr, nullptr, nullptr};  // runs per entry of kShowCaps
    uint32_t n_show_chunks[3] = {0, 0, 0};
    ChanceRec* chance = nullptr;
    uint32_t n_chance = 0;
    EdgeRec* edge = nullptr;
    uint32_t n_edge = 0;
    // the same lists without river-subtree nodes (fused river sweeps)
    ActRec* act_up[2] = {nullptr, nullptr};
    uint32_t n_act_up[2] = {0, 0};
    EdgeRec* edge_up = nullptr;
    uint32_t n_edge_up = 0;
};

// IEEE half <-> float on the host (round to nearest even)
inline float half_to_float(uint16_t u) { return (float)__builtin_bit_cast(_Float16, u); }
inline uint16_t float_to_half(float x) { return __builtin_bit_cast(uint16_t, (_Float16)x); }

// Workgroups for a list of `count` one-wavefront items, padded to a
// multiple of the 8 XCDs (see xcd_slice).
uint32_t wave_grid(uint32_t count, uint32_t waves = kWaves) {
    uint32_t blocks = (count + waves - 1) / waves;
    return (blocks + 7u) & ~7u;
}

// Fold nodes per wavefront for a launch over n nodes that runs alone (a
// level's folds): kFoldNodes (the card matrix's zeroing and the hand
// tables amortised over 16 nodes) once that still gives about 8192
// wavefronts, fewer on small trees.
uint32_t fold_per(uint32_t n) { return std::max<uint32_t>(1, std::min<uint32_t>(kFoldNodes, (n + 8191) / 8192)); }
uint32_t fold_waves(uint32_t n) { return (n + fold_per(n) - 1) / fold_per(n); }

// Discounts::for_iteration (cfr.rs:44-63), computed in f64 and stored as f32.
void discounts(int algo, uint32_t it, float* out3) {
    double t = (double)it;
    double q = t / (t + 1.0);
    if (algo == GTO_ALGO_DCFR) {
        double ta = std::pow(t, 1.5);
        out3[0] = (float)(ta / (ta + 1.0));
        out3[1] = 0.5f;
    } else {
        out3[0] = 1.0f;
        out3[1] = 0.0f;
    }
    out3[2] = (float)(q * q);
}

class Engine {
   public:
    Engine(const gto_spot_view& v, const gto_hip_options& o) : opts_(o) {
        if (o.storage != GTO_STORAGE_F32 && o.storage != GTO_STORAGE_COMPRESSED && o.storage != GTO_STORAGE_HALF &&
            o.storage != GTO_STORAGE_HALF_ROWS)
            throw std::runtime_error("unknown storage mode");
        compressed_ = o.storage != GTO_STORAGE_F32;
        half_ = o.storage == GTO_STORAGE_HALF || o.storage == GTO_STORAGE_HALF_ROWS;
        rows_ = o.storage == GTO_STORAGE_HALF_ROWS ? (kRows16Reach | kRows16Value) : 0;
        pred_ = o.algo == GTO_ALGO_PCFRPLUS;
        if (pred_ && compressed_) throw std::runtime_error("PCFR+ runs with f32 arenas on the GPU");
        // fault injection: GTO_INJECT_OOM=1 reports a device allocation
        // failure before the device is touched, so a host-only test can
        // drive a caller's out-of-memory path
        if (const char* f = std::getenv("GTO_INJECT_OOM"); f && f[0] == '1')
            throw DeviceOutOfMemory("hip: out of device memory (injected by GTO_INJECT_OOM)");
        HIP_OK(hipSetDevice(o.device));
        HIP_OK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
        HIP_OK(hipStreamCreateWithFlags(&aux_, hipStreamNonBlocking));
        HIP_OK(hipEventCreateWithFlags(&fork_ev_, hipEventDisableTiming));
        HIP_OK(hipEventCreateWithFlags(&join_ev_, hipEventDisableTiming));
        lstream_ = stream_;
        plan_ = build_plan(v, o.isomorphism != 0);
        num_nodes_ = v.num_nodes;
        data_size_[0] = v.data_size[0];
        data_size_[1] = v.data_size[1];
        rake_ = v.rake_pct;
        starting_pot_ = v.starting_pot;
        node_kind_.assign(v.node_kind, v.node_kind + v.num_nodes);
        node_player_.assign(v.node_player, v.node_player + v.num_nodes);
        node_na_.assign(v.node_num_children, v.node_num_children + v.num_nodes);
        node_off_.assign(v.node_data_offset, v.node_data_offset + v.num_nodes);
        for (int q = 0; q < 2; ++q) {
            weights_[q].assign(v.weights[q], v.weights[q] + v.num_hands[q]);
            c1_[q].assign(v.hand_c1[q], v.hand_c1[q] + v.num_hands[q]);
            c2_[q].assign(v.hand_c2[q], v.hand_c2[q] + v.num_hands[q]);
            same_[q].assign(v.same_combo[q], v.same_combo[q] + v.num_hands[q]);
        }
        scratch_ = scratch_on(o.device);
        root_row_.assign(std::max<uint32_t>(plan_.nh_max, 1), 0.0f);
        upload_tree(v);
        alloc_state();
        reach_valid_[0] = reach_valid_[1] = false;
        if (plan_.iso) build_iso(v);
    }

    ~Engine() {
        drop_graph();
        for (void* p : allocs_) (void)hipFree(p);
        if (stream_) (void)hipStreamDestroy(stream_);
        if (aux_) (void)hipStreamDestroy(aux_);
        if (fork_ev_) (void)hipEventDestroy(fork_ev_);
        if (join_ev_) (void)hipEventDestroy(join_ev_);
        for (auto& e : evs_) (void)hipEventDestroy(e);
    }

    uint32_t iteration() const { return iteration_; }

    // --- batch support: several engines' iterations in one parent graph ---
    // Steady state: the regrets-changed reach of player 1 is the only stale
    // one, which is what one captured iteration assumes.
    bool steady() const { return reach_valid_[0] && !reach_valid_[1] && !prof_ && !(rows_stale_ && !fused()); }
    hipStream_t stream() const { return stream_; }
    int device() const { return opts_.device; }
    // bumped whenever a captured iteration's kernel arguments go stale
    uint64_t generation() const { return gen_; }
    void reserve_iterations(uint32_t n) { ensure_disc(iteration_ + n + 1); }
    // one steady-state iteration as an uninstantiated graph (caller owns it)
    hipGraph_t capture_iteration() {
        if (!steady()) throw std::runtime_error("engine is not in the steady state; iterate once first");
        HIP_OK(hipStreamSynchronize(stream_));
        HIP_OK(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
        try {
            launch_iteration();
        } catch (...) {
            hipGraph_t g;
            (void)hipStreamEndCapture(stream_, &g);
            if (g) (void)hipGraphDestroy(g);
            reach_valid_[0] = true;
            reach_valid_[1] = false;
            throw;
        }
        hipGraph_t g = nullptr;
        HIP_OK(hipStreamEndCapture(stream_, &g));
        reach_valid_[0] = true;
        reach_valid_[1] = false;
        return g;
    }
    // record n iterations run through a batch graph
    void advanced(uint32_t n) { iteration_ += n; }
    void set_iteration(uint32_t t) {
        iteration_ = t;
        ensure_disc(t + 1);
        HIP_OK(hipMemcpyAsync(S_.iter, &iteration_, sizeof(uint32_t), hipMemcpyHostToDevice, stream_));
        HIP_OK(hipStreamSynchronize(stream_));
    }

    void iterate(uint32_t n) {
        ensure_disc(iteration_ + n + 1);
        for (uint32_t k = 0; k < n; ++k) {
            if (opts_.use_graphs && steady()) {
                if (!graph_) capture_graph();
                HIP_OK(hipGraphLaunch(graph_, stream_));
                reach_valid_[0] = true;
                reach_valid_[1] = false;
            } else {
                launch_iteration();
            }
            iteration_ += 1;
        }
    }

    // Solver::run (cfr.rs:194-226): iterate until max_iterations, the stop
    // flag or the target; exploitability every check_every iterations and at
    // the last one. Iterations between checks are launched back to back and
    // synchronised once (gpu_secs: launch to completion).
    gto_progress run(const gto_run_options& o, const volatile int* stop, gto_progress_fn cb, void* user) {
        using clk = std::chrono::steady_clock;
        const auto t0 = clk::now();
        const uint32_t every = std::max(1u, o.check_every);
        double gpu_secs = 0.0;
        gto_progress last{iteration_, NAN, NAN, 0.0, 0.0};
        while (iteration_ < o.max_iterations && !(stop && *stop)) {
            const auto t = clk::now();
            do {
                iterate(1);
            } while (iteration_ % every != 0 && iteration_ < o.max_iterations && !(stop && *stop));
            synchronize();
            gpu_secs += std::chrono::duration<double>(clk::now() - t).count();
            if (iteration_ % every != 0 && iteration_ < o.max_iterations) break;  // stopped between checks
            const double e = exploitability();
            last.iteration = iteration_;
            last.exploit_chips = e;
            last.exploit_pct_pot = e / starting_pot_ * 100.0;
            last.elapsed_secs = std::chrono::duration<double>(clk::now() - t0).count();
            last.gpu_secs = gpu_secs;
            if (cb && cb(&last, user) == 0) break;
            if (last.exploit_pct_pot <= o.target_exploit_pct) break;
        }
        last.gpu_secs = gpu_secs;  // the record of the last check, as the reference returns it
        return last;
    }

    double exploitability() {
        double denom = pair_weight_sum();
        if (denom <= 0.0) return NAN;
        double br[2], val[2];
        std::vector<float> c;
        for (int p = 0; p < 2; ++p) {
            root_cfv(p, 0, c);
            br[p] = dot(p, c) / denom;
        }
        if (rake_ > 0.0) {
            for (int p = 0; p < 2; ++p) {
                root_cfv(p, 1, c);
                val[p] = dot(p, c) / denom;
            }
            return ((br[0] - val[0]) + (br[1] - val[1])) / 2.0;
        }
        return (br[0] + br[1]) / 2.0;
    }

    // mode 0 = best response, 1 = average strategy, 2 = current strategy
    void root_cfv(int p, int mode, std::vector<float>& out) {
        const int o = 1 - p;
        if (mode == 2) {
            // current strategy for both players: the opponent's reach under
            // regret matching, values weighted by the traverser's sigma
            launch_root(1 << o);
            launch_down(1 << o, kCurrentF32);
            reach_valid_[o] = !(rows_ & kRows16Reach);  // f32 rows: DCFR sweeps with f16 reach rows recompute them
            launch_up(p, 3);
        } else {
            launch_root(1 << o);
            launch_down(1 << o, kAverageF32);
            reach_valid_[o] = false;
            launch_up(p, mode == 0 ? 1 : 2);
        }
        out.resize(plan_.nh[p]);
        HIP_OK(hipMemcpyAsync(out.data(), S_.cfv + (size_t)plan_.cfv_slot[0] * plan_.nh_max,
                              sizeof(float) * plan_.nh[p], hipMemcpyDeviceToHost, stream_));
        HIP_OK(hipStreamSynchronize(stream_));
    }

    // Average strategy of one action node from the strategy arena, ignoring
    // any lock (solved_strategy_into, cfr.rs:679-726): max(S,0)/sum with an
    // f64 sum (f32 arenas) or q/sum over u16 quants (compressed), uniform
    // where the sum is <= 1e-12 / 0.
    void solved_strategy(uint32_t n, float* out) {
        if (n >= num_nodes_ || node_kind_[n] != GTO_KIND_ACTION)
            throw std::runtime_error("node " + std::to_string(n) + " is not an action node");
        const int p = node_player_[n];
        const size_t na = node_na_[n], nh = plan_.nh[p], cnt = na * nh;
        const uint64_t off = node_off_[n];
        if (iso_ && iso_src_[n] >= 0) {
            // a branch CFR does not traverse: its strategy sums are those of
            // the source it is copied from, suit-permuted (cfr.rs:806-840)
            const IsoCopy& c = iso_copies_[(size_t)iso_src_[n]];
            std::vector<float> base(cnt);
            if (const float* lk = lock_host(c.src)) {
                std::vector<float> sums(lk, lk + cnt);
                if (compressed_ && !half_) {  // the copy is stored through the u16 quantizer
                    std::vector<uint16_t> q(cnt);
                    encode_u16(sums.data(), cnt, q.data());
                    for (size_t e = 0; e < cnt; ++e) sums[e] = (float)q[e];
                }
                normalize_sums(sums.data(), na, nh, base.data(), compressed_ && !half_);
            } else {
                solved_strategy(c.src, base.data());
            }
            const uint16_t* tbl = &hand_perm_[p][(size_t)c.k * nh];
            for (size_t a = 0; a < na; ++a)
                for (size_t i = 0; i < nh; ++i) out[a * nh + i] = base[a * nh + tbl[i]];
            return;
        }
        if (!compressed_) {
            std::vector<float> b(cnt);
            HIP_OK(hipMemcpyAsync(b.data(), S_.S[p] + off, sizeof(float) * cnt, hipMemcpyDeviceToHost, stream_));
            HIP_OK(hipStreamSynchronize(stream_));
            for (size_t i = 0; i < nh; ++i) {
                double sum = 0.0;
                for (size_t a = 0; a < na; ++a) sum += (double)std::max(b[a * nh + i], 0.0f);
                for (size_t a = 0; a < na; ++a)
                    out[a * nh + i] =
                        sum > 1e-12 ? (float)((double)std::max(b[a * nh + i], 0.0f) / sum) : 1.0f / (float)na;
            }
        } else if (half_) {
            std::vector<uint16_t> q(cnt);
            HIP_OK(hipMemcpyAsync(q.data(), S_.Sq[p] + off, sizeof(uint16_t) * cnt, hipMemcpyDeviceToHost, stream_));
            HIP_OK(hipStreamSynchronize(stream_));
            for (size_t i = 0; i < nh; ++i) {
                double sum = 0.0;
                for (size_t a = 0; a < na; ++a) sum += (double)std::max(half_to_float(q[a * nh + i]), 0.0f);
                for (size_t a = 0; a < na; ++a)
                    out[a * nh + i] = sum > 0.0 ? (float)((double)std::max(half_to_float(q[a * nh + i]), 0.0f) / sum)
                                                : 1.0f / (float)na;
            }
        } else {
            std::vector<uint16_t> q(cnt);
            HIP_OK(hipMemcpyAsync(q.data(), S_.Sq[p] + off, sizeof(uint16_t) * cnt, hipMemcpyDeviceToHost, stream_));
            HIP_OK(hipStreamSynchronize(stream_));
            for (size_t i = 0; i < nh; ++i) {
                uint32_t sum = 0;
                for (size_t a = 0; a < na; ++a) sum += q[a * nh + i];
                const float inv = sum > 0 ? 1.0f / (float)sum : 0.0f;
                for (size_t a = 0; a < na; ++a) out[a * nh + i] = sum > 0 ? (float)q[a * nh + i] * inv : 1.0f / (float)na;
            }
        }
    }

    // Value rows of player p at the given nodes after one sweep: mode 0 =
    // best response (traverse_br, best_response.rs:19-124), 1 = average
    // strategy (traverse_avg, 128-223). The row of node n is what those
    // functions return for n given the opponent's reach at n.
    void node_values(int p, int mode, uint32_t count, const uint32_t* nodes, float* out) {
        for (uint32_t k = 0; k < count; ++k)
            if (nodes[k] >= num_nodes_ || plan_.cfv_slot[nodes[k]] == 0xFFFFFFFFu)
                throw std::runtime_error("node " + std::to_string(nodes[k]) + " is not in the reachable tree");
        const int o = 1 - p;
        launch_root(1 << o);
        if (mode == 2) {
            launch_down(1 << o, kCurrentF32);
            reach_valid_[o] = !(rows_ & kRows16Reach);  // f32 rows: DCFR sweeps with f16 reach rows recompute them
            launch_up(p, 3);
        } else {
            launch_down(1 << o, kAverageF32);
            reach_valid_[o] = false;
            launch_up(p, mode == 0 ? 1 : 2);
        }
        const size_t nh = plan_.nh[p];
        for (uint32_t k = 0; k < count; ++k)
            HIP_OK(hipMemcpyAsync(out + k * nh, S_.cfv + (size_t)plan_.cfv_slot[nodes[k]] * plan_.nh_max,
                                  sizeof(float) * nh, hipMemcpyDeviceToHost, stream_));
        HIP_OK(hipStreamSynchronize(stream_));
    }

    // Range weights (the root reach); the spot's hands stay as built, and a
    // zero weight leaves its hand with zero reach.
    void set_weights(int q, const float* w) {
        if (!w) return;
        const uint32_t nh = plan_.nh[q];
        for (uint32_t i = 0; i < nh; ++i)
            if (!(w[i] >= 0.0f) || !std::isfinite(w[i]))
                throw std::runtime_error("range weights must be finite and >= 0");
        HIP_OK(hipStreamSynchronize(stream_));
        HIP_OK(hipMemcpy((void*)T_.weights[q], w, sizeof(float) * nh, hipMemcpyHostToDevice));
        weights_[q].assign(w, w + nh);
        reach_valid_[0] = reach_valid_[1] = false;
    }

    // Game value of player p under the behavioural strategy held in the
    // strategy arenas (sigma as stored, no normalisation):
    //   U_p = sum_i w_p[i] * cfv_p(root)[i]   (best_response.rs:264-269)
    // and its gradient. U_p is linear in each range and in each node's
    // strategy, so dU/dw_p = cfv_p(root), dU/dw_o = the root values of o's
    // hands under p's payoffs, and dU/dsigma_q(n, a, i) = reach_q(n, i) *
    // chance weight * value of child a (k_grad_sigma).
    //
    // noalloc: the strategy gradient is staged in the library's static
    // device scratch (at most kScratchFloats entries per player, cleared by a
    // fill kernel) instead of a buffer the engine allocates on first use; no
    // device or host allocation happens in the call.
    //
    // dev: gw and gs are device memory of this GPU (checked); the root rows
    // are copied there and the strategy gradient is written in place, so
    // nothing but the value crosses to the host.
    void game_value(int p, double* U, float* gw[2], float* gs[2], bool noalloc = false, bool dev = false) {
        if (compressed_) throw std::runtime_error("game values need an engine with f32 arenas");
        if (dev)
            for (int q = 0; q < 2; ++q) {
                if (gw[q]) require_device_ptr(gw[q], opts_.device, "range gradient");
                if (gs[q] && data_size_[q]) require_device_ptr(gs[q], opts_.device, "strategy gradient");
            }
        // the strategy is an input per node; under suit isomorphism copied
        // branches would silently take their representatives' strategies
        if (iso_) throw std::runtime_error("game values need an engine created with isomorphism off");
        const int o = 1 - p;
        if (noalloc)
            for (int q = 0; q < 2; ++q)
                if (gs[q] && data_size_[q] > kScratchFloats)
                    throw std::runtime_error("gto_hip_game_value_noalloc: the strategy arena of player " +
                                             std::to_string(q) + " has " + std::to_string(data_size_[q]) +
                                             " entries, more than the " + std::to_string(kScratchFloats) +
                                             " the static device scratch holds");
        launch_root(3);
        launch_down(3, kRawF32);
        reach_valid_[0] = reach_valid_[1] = false;
        std::vector<float>& c = root_row_;  // sized at construction: no allocation here
        for (int pass = 0; pass < 2; ++pass) {
            const int t = pass == 0 ? p : o;  // traverser of this sweep
            if (pass == 1 && !gw[o] && !gs[o]) break;
            launch_up(t, 4, pass);
            const float* root = S_.cfv + (size_t)plan_.cfv_slot[0] * plan_.nh_max;
            if (dev && gw[t])
                HIP_OK(hipMemcpyAsync(gw[t], root, sizeof(float) * plan_.nh[t], hipMemcpyDeviceToDevice, stream_));
            if (!dev || (pass == 0 && U)) {
                HIP_OK(hipMemcpyAsync(c.data(), root, sizeof(float) * plan_.nh[t], hipMemcpyDeviceToHost, stream_));
                HIP_OK(hipStreamSynchronize(stream_));
            }
            if (pass == 0 && U) *U = dot(p, c.data(), plan_.nh[p]);
            if (!dev && gw[t]) std::copy(c.begin(), c.begin() + plan_.nh[t], gw[t]);
            if (gs[t] && data_size_[t]) {
                float* buf;
                if (dev) {
                    buf = gs[t];
                    const uint32_t blocks = (uint32_t)std::min<uint64_t>((data_size_[t] + kBlock - 1) / kBlock, 8192);
                    launch(PK_OTHER, k_fill, blocks, 0, buf, (uint64_t)data_size_[t], 0.0f);
                } else if (noalloc) {
                    buf = scratch_;
                    const uint32_t blocks = (uint32_t)std::min<uint64_t>((data_size_[t] + kBlock - 1) / kBlock, 8192);
                    launch(PK_OTHER, k_fill, blocks, 0, buf, (uint64_t)data_size_[t], 0.0f);
                } else {
                    if (!grad_buf_)
                        grad_buf_ = (float*)dalloc(sizeof(float) * std::max(data_size_[0], data_size_[1]));
                    HIP_OK(hipMemsetAsync(grad_buf_, 0, sizeof(float) * data_size_[t], stream_));
                    buf = grad_buf_;
                }
                for (size_t l = 0; l < lv_.size(); ++l) {
                    const DevLevel& D = lv_[l];
                    if (!D.n_act[t]) continue;
                    launch(PK_OTHER, k_grad_sigma, wave_grid(D.n_act[t]), 0, T_, S_, (const ActRec*)D.act[t],
                           D.n_act[t], t, buf);
                }
                if (!dev)
                    HIP_OK(hipMemcpyAsync(gs[t], buf, sizeof(float) * data_size_[t], hipMemcpyDeviceToHost, stream_));
            }
        }
        HIP_OK(hipStreamSynchronize(stream_));
    }

    // the strategy arenas (the game value's input) from device memory of
    // this GPU, f32 arenas only
    void upload_strategy_device(const float* s0, const float* s1) {
        if (compressed_) throw std::invalid_argument("device strategy uploads need an engine with f32 arenas");
        const float* src[2] = {s0, s1};
        for (int q = 0; q < 2; ++q)
            if (src[q] && data_size_[q]) require_device_ptr(src[q], opts_.device, "strategy");
        reach_valid_[0] = reach_valid_[1] = false;
        for (int q = 0; q < 2; ++q)
            if (src[q] && data_size_[q])
                HIP_OK(hipMemcpyAsync(S_.S[q], src[q], sizeof(float) * data_size_[q], hipMemcpyDeviceToDevice,
                                      stream_));
        HIP_OK(hipStreamSynchronize(stream_));
    }

    void download(int which, int p, float* dst) {
        if (!dst || data_size_[p] == 0) return;
        download_reps(which, p, dst);
        if (iso_) symmetrize(which, p, dst);
    }
    // the arenas as the device holds them: branches CFR does not traverse
    // (under suit isomorphism) are stale
    void download_reps(int which, int p, float* dst) {
        if (!compressed_) {
            HIP_OK(hipMemcpyAsync(dst, which == 0 ? S_.R[p] : S_.S[p], sizeof(float) * data_size_[p],
                                  hipMemcpyDeviceToHost, stream_));
            HIP_OK(hipStreamSynchronize(stream_));
            return;
        }
        // decode per action node: v = q * scale / 32767 (i16) or / 65535 (u16), store.rs:153-166
        std::vector<uint16_t> q(data_size_[p]);
        std::vector<float> sc(num_nodes_);
        HIP_OK(hipMemcpyAsync(q.data(), which == 0 ? (void*)S_.Rq[p] : (void*)S_.Sq[p],
                              sizeof(uint16_t) * data_size_[p], hipMemcpyDeviceToHost, stream_));
        HIP_OK(hipMemcpyAsync(sc.data(), which == 0 ? S_.Rs[p] : S_.Ss[p], sizeof(float) * num_nodes_,
                              hipMemcpyDeviceToHost, stream_));
        HIP_OK(hipStreamSynchronize(stream_));
        for (uint32_t n = 0; n < num_nodes_; ++n) {
            if (node_kind_[n] != GTO_KIND_ACTION || node_player_[n] != p) continue;
            size_t cnt = (size_t)node_na_[n] * plan_.nh[p];
            uint64_t off = node_off_[n];
            if (half_) {  // f16 value times the node's power-of-two scale (exact)
                for (size_t e = 0; e < cnt; ++e) dst[off + e] = half_to_float(q[off + e]) * sc[n];
                continue;
            }
            if (which == 0) {
                float k = sc[n] / 32767.0f;
                for (size_t e = 0; e < cnt; ++e) dst[off + e] = (float)(int16_t)q[off + e] * k;
            } else {
                float k = sc[n] / 65535.0f;
                for (size_t e = 0; e < cnt; ++e) dst[off + e] = (float)q[off + e] * k;
            }
        }
    }
    void download_raw(int p, int16_t* rq, uint16_t* sq, float* rs, float* ss) {
        if (!compressed_) throw std::runtime_error("engine was created with f32 arenas");
        if (rq) HIP_OK(hipMemcpyAsync(rq, S_.Rq[p], sizeof(int16_t) * data_size_[p], hipMemcpyDeviceToHost, stream_));
        if (sq) HIP_OK(hipMemcpyAsync(sq, S_.Sq[p], sizeof(uint16_t) * data_size_[p], hipMemcpyDeviceToHost, stream_));
        if (rs) HIP_OK(hipMemcpyAsync(rs, S_.Rs[p], sizeof(float) * num_nodes_, hipMemcpyDeviceToHost, stream_));
        if (ss) HIP_OK(hipMemcpyAsync(ss, S_.Ss[p], sizeof(float) * num_nodes_, hipMemcpyDeviceToHost, stream_));
        HIP_OK(hipStreamSynchronize(stream_));
        if (iso_ && !half_) symmetrize_raw(p, rq, sq, rs, ss);
    }
    void upload(int which, int p, const float* src) {
        if (!src || data_size_[p] == 0) return;
        reach_valid_[0] = reach_valid_[1] = false;
        if (!compressed_) {
            HIP_OK(hipMemcpyAsync(which == 0 ? S_.R[p] : S_.S[p], src, sizeof(float) * data_size_[p],
                                  hipMemcpyHostToDevice, stream_));
            if (pred_ && which == 0) {
                // regrets from outside carry no prediction (the save format
                // has none; the CPU engine starts its predictions at 0 too)
                std::vector<float> qv(data_size_[p]);
                for (size_t k = 0; k < qv.size(); ++k) qv[k] = std::max(src[k], 0.0f);
                HIP_OK(hipMemcpyAsync(S_.Q[p], qv.data(), sizeof(float) * qv.size(), hipMemcpyHostToDevice, stream_));
                HIP_OK(hipStreamSynchronize(stream_));  // qv is freed on return
            }
            HIP_OK(hipStreamSynchronize(stream_));
            return;
        }
        if (half_) {
            upload_half(which, p, src);
            return;
        }
        // encode per action node against its block max (store.rs:199-215)
        std::vector<uint16_t> q(data_size_[p], 0);
        std::vector<float> sc(num_nodes_, 0.0f);
        HIP_OK(hipMemcpyAsync(sc.data(), which == 0 ? S_.Rs[p] : S_.Ss[p], sizeof(float) * num_nodes_,
                              hipMemcpyDeviceToHost, stream_));
        HIP_OK(hipStreamSynchronize(stream_));
        for (uint32_t n = 0; n < num_nodes_; ++n) {
            if (node_kind_[n] != GTO_KIND_ACTION || node_player_[n] != p) continue;
            size_t cnt = (size_t)node_na_[n] * plan_.nh[p];
            uint64_t off = node_off_[n];
            float m = 0.0f;
            for (size_t e = 0; e < cnt; ++e) m = std::fmax(m, which == 0 ? std::fabs(src[off + e]) : src[off + e]);
            if (!(m > 0.0f && std::isfinite(m))) {
                sc[n] = 0.0f;
                continue;
            }
            sc[n] = m;
            if (which == 0) {
                float k = 32767.0f / m;
                for (size_t e = 0; e < cnt; ++e) {
                    float r = std::nearbyint(src[off + e] * k);
                    r = r != r ? 0.0f : std::fmin(std::fmax(r, -32768.0f), 32767.0f);
                    q[off + e] = (uint16_t)(int16_t)r;
                }
            } else {
                float k = 65535.0f / m;
                for (size_t e = 0; e < cnt; ++e) {
                    float r = std::nearbyint(std::fmax(src[off + e], 0.0f) * k);
                    r = r != r ? 0.0f : std::fmin(std::fmax(r, 0.0f), 65535.0f);
                    q[off + e] = (uint16_t)r;
                }
            }
        }
        HIP_OK(hipMemcpyAsync(which == 0 ? (void*)S_.Rq[p] : (void*)S_.Sq[p], q.data(),
                              sizeof(uint16_t) * data_size_[p], hipMemcpyHostToDevice, stream_));
        HIP_OK(hipMemcpyAsync(which == 0 ? S_.Rs[p] : S_.Ss[p], sc.data(), sizeof(float) * num_nodes_,
                              hipMemcpyHostToDevice, stream_));
        HIP_OK(hipStreamSynchronize(stream_));
    }

    // f16 arenas: per node, scale = half_scale(max |v|), values rounded to
    // nearest even; the node's max is kept for the next update's bound.
    void upload_half(int which, int p, const float* src) {
        std::vector<uint16_t> q(data_size_[p], 0);
        std::vector<float> sc(num_nodes_, 1.0f), mx(num_nodes_, 0.0f);
        HIP_OK(hipMemcpyAsync(sc.data(), which == 0 ? S_.Rs[p] : S_.Ss[p], sizeof(float) * num_nodes_,
                              hipMemcpyDeviceToHost, stream_));
        HIP_OK(hipMemcpyAsync(mx.data(), which == 0 ? S_.Rm[p] : S_.Sm[p], sizeof(float) * num_nodes_,
                              hipMemcpyDeviceToHost, stream_));
        HIP_OK(hipStreamSynchronize(stream_));
        for (uint32_t n = 0; n < num_nodes_; ++n) {
            if (node_kind_[n] != GTO_KIND_ACTION || node_player_[n] != p) continue;
            const size_t cnt = (size_t)node_na_[n] * plan_.nh[p];
            const uint64_t off = node_off_[n];
            float m = 0.0f;
            for (size_t e = 0; e < cnt; ++e) {
                const float v = which == 0 ? src[off + e] : std::fmax(src[off + e], 0.0f);
                if (!std::isfinite(v)) throw std::runtime_error("non-finite arena value for node " + std::to_string(n));
                m = std::fmax(m, std::fabs(v));
            }
            sc[n] = half_scale(m);
            mx[n] = m;
            const float inv = 1.0f / sc[n];
            for (size_t e = 0; e < cnt; ++e)
                q[off + e] = float_to_half((which == 0 ? src[off + e] : std::fmax(src[off + e], 0.0f)) * inv);
        }
        HIP_OK(hipMemcpyAsync(which == 0 ? (void*)S_.Rq[p] : (void*)S_.Sq[p], q.data(),
                              sizeof(uint16_t) * data_size_[p], hipMemcpyHostToDevice, stream_));
        HIP_OK(hipMemcpyAsync(which == 0 ? S_.Rs[p] : S_.Ss[p], sc.data(), sizeof(float) * num_nodes_,
                              hipMemcpyHostToDevice, stream_));
        HIP_OK(hipMemcpyAsync(which == 0 ? S_.Rm[p] : S_.Sm[p], mx.data(), sizeof(float) * num_nodes_,
                              hipMemcpyHostToDevice, stream_));
        HIP_OK(hipStreamSynchronize(stream_));
    }

    void update_locks(const int64_t* lock_off, const float* sigma, uint64_t n_sigma) {
        HIP_OK(hipStreamSynchronize(stream_));
        drop_graph();
        // locked nodes play their fixed strategy in the level sweeps; the
        // fused river sweep runs only while no node is locked
        locked_ = false;
        if (lock_off)
            for (uint32_t n = 0; n < num_nodes_; ++n) locked_ = locked_ || lock_off[n] >= 0;
        // host copies: symmetrize copies a locked source's strategy
        lock_off_h_.assign(num_nodes_, -1);
        if (lock_off) lock_off_h_.assign(lock_off, lock_off + num_nodes_);
        lock_sigma_h_.assign(sigma, sigma + n_sigma);
        if (lock_sigma_buf_) {
            HIP_OK(hipFree(lock_sigma_buf_));
            allocs_.erase(std::find(allocs_.begin(), allocs_.end(), (void*)lock_sigma_buf_));
            lock_sigma_buf_ = nullptr;
        }
        if (n_sigma) {
            lock_sigma_buf_ = (float*)dalloc(sizeof(float) * n_sigma);
            HIP_OK(hipMemcpy(lock_sigma_buf_, sigma, sizeof(float) * n_sigma, hipMemcpyHostToDevice));
        }
        T_.lock_sigma = lock_sigma_buf_;
        float lb = 1.0f;
        for (uint64_t k = 0; k < n_sigma; ++k)
            if (std::isfinite(sigma[k])) lb = std::max(lb, std::fabs(sigma[k]));
        T_.lock_bound = lb;
        // patch the action records' lock offsets
        for (size_t l = 0; l < plan_.levels.size(); ++l)
            for (int q = 0; q < 2; ++q) {
                std::vector<ActRec>& recs = plan_.levels[l].act_rec[q];
                if (recs.empty()) continue;
                for (ActRec& r : recs) {
                    int64_t lo = lock_off ? lock_off[r.node] : -1;
                    if (lo >= 0 && (uint64_t)lo + (uint64_t)r.na * plan_.nh[node_player_[r.node]] > n_sigma)
                        throw std::runtime_error("lock sigma block out of range for node " + std::to_string(r.node));
                    r.lock_off = lo;
                }
                HIP_OK(hipMemcpy(lv_[l].act[q], recs.data(), sizeof(ActRec) * recs.size(), hipMemcpyHostToDevice));
            }
        reach_valid_[0] = reach_valid_[1] = false;
    }

    void synchronize() { HIP_OK(hipStreamSynchronize(stream_)); }

    // Bitsets of card-compatible opponent hands for the MFMA showdown: per
    // board and sorted hand i of p, bit k set when i and the opponent's k-th
    // strongest-ordered hand share no card (padding bits 0).
    void build_mma_tables() {
        if (mma_tables_) return;
        for (int p = 0; p < 2; ++p) {
            const int o = 1 - p;
            std::vector<uint32_t> off(plan_.num_boards + 1, 0);
            for (uint32_t b = 0; b < plan_.num_boards; ++b)
                off[b + 1] = off[b] + plan_.sd_cnt[p][b] * mma_lds_layout(plan_.sd_cnt[o][b], 0, 0).words;
            // 8 words of padding: the kernel reads whole 8-word groups
            std::vector<uint32_t> bits(off[plan_.num_boards] + 8, 0u);
            auto work = [&](uint32_t b0, uint32_t b1) {
                std::vector<uint64_t> om;
                for (uint32_t b = b0; b < b1; ++b) {
                    const uint32_t mo = plan_.sd_cnt[o][b], mp = plan_.sd_cnt[p][b];
                    const uint32_t words = mma_lds_layout(mo, 0, 0).words;
                    om.resize(mo);
                    for (uint32_t k = 0; k < mo; ++k) {
                        const uint16_t j = plan_.sd_idx[o][plan_.sd_off[o][b] + k];
                        om[k] = (1ull << c1_[o][j]) | (1ull << c2_[o][j]);
                    }
                    for (uint32_t i = 0; i < mp; ++i) {
                        const uint16_t h = plan_.sd_idx[p][plan_.sd_off[p][b] + i];
                        const uint64_t mh = (1ull << c1_[p][h]) | (1ull << c2_[p][h]);
                        uint32_t* row = &bits[off[b] + (size_t)i * words];
                        for (uint32_t k = 0; k < mo; ++k)
                            if ((mh & om[k]) == 0) row[k >> 5] |= 1u << (k & 31);
                        // pair-split words: bit 2i -> bit i, bit 2i+1 -> bit 16+i
                        for (uint32_t w = 0; w < words; ++w) {
                            uint32_t z = 0;
                            for (uint32_t i = 0; i < 16; ++i)
                                z |= ((row[w] >> (2 * i)) & 1u) << i | ((row[w] >> (2 * i + 1)) & 1u) << (16 + i);
                            row[w] = z;
                        }
                    }
                }
            };
            const uint32_t nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
            std::vector<std::thread> pool;
            const uint32_t per = (plan_.num_boards + nt - 1) / nt;
            for (uint32_t t = 0; t < nt; ++t) {
                const uint32_t b0 = std::min(plan_.num_boards, t * per), b1 = std::min(plan_.num_boards, b0 + per);
                if (b0 < b1) pool.emplace_back(work, b0, b1);
            }
            for (auto& th : pool) th.join();
            T_.sd_djbits[p] = upload_vec(bits);
            T_.sd_lbub[p] = upload_vec(plan_.sd_lbub[p]);
            off.pop_back();
            T_.sd_djoff[p] = upload_vec(off);
        }
        mma_tables_ = true;
    }

    void set_showdown(int mode) {
        if (mode != GTO_SHOWDOWN_EXACT && mode != GTO_SHOWDOWN_BF16 && mode != GTO_SHOWDOWN_FP8 &&
            mode != GTO_SHOWDOWN_SEQUENTIAL)
            throw std::runtime_error("unknown showdown mode");
        if (mode == GTO_SHOWDOWN_SEQUENTIAL && rows_)
            throw std::invalid_argument("the sequential sums are f64 sums of f32 reach rows: not with f16 rows");
        if (mode != GTO_SHOWDOWN_EXACT && rows_)
            throw std::invalid_argument("the matrix-core showdown forms read f32 rows: not with f16 rows");
        if (mode == GTO_SHOWDOWN_BF16 || mode == GTO_SHOWDOWN_FP8) {
            const bool fp8 = mode == GTO_SHOWDOWN_FP8;
            build_mma_tables();
            // assemble output rows in LDS when they fit
            size_t direct = 0, staged = 0;
            for (uint32_t b = 0; b < plan_.num_boards; ++b)
                for (int q = 0; q < 2; ++q) {
                    const uint32_t mo = plan_.sd_cnt[1 - q][b], mp = plan_.sd_cnt[q][b];
                    direct = std::max<size_t>(direct, mma_lds_layout(mo, mp, 0, fp8).bytes);
                    staged = std::max<size_t>(staged, mma_lds_layout(mo, mp, plan_.nh[q], fp8).bytes);
                }
            const bool use_staged = staged <= kLdsBudget;
            const size_t need = use_staged ? staged : direct;
            if (need > kLdsBudget)
                throw std::runtime_error("showdown matrix operands exceed the LDS budget (too many hands per board)");
            const void* fn = fp8 ? (const void*)k_up_show_mma<true> : (const void*)k_up_show_mma<false>;
            HIP_OK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)std::max<size_t>(need, 1)));
            show_mma_staged_ = use_staged;
            show_mma_lds_ = need;
        }
        HIP_OK(hipStreamSynchronize(stream_));
        drop_graph();
        show_mode_ = mode;
    }

    // Launch schedule of the up sweep; every schedule computes the same
    // values (terminals depend only on the down pass, and a showdown run's
    // length does not change any node's arithmetic).
    void set_schedule(const gto_schedule& sc) {
        int c = -1;
        for (int k = 0; k < 3; ++k)
            if (kShowCaps[k] == sc.show_nodes) c = k;
        if (c < 0 || sc.per_level > 1 || sc.river > 1 || sc.absorb > 1) throw std::runtime_error("unknown schedule");
        if (c == cap_ && sc.per_level == sched_.per_level && sc.river == sched_.river && sc.absorb == sched_.absorb)
            return;
        HIP_OK(hipStreamSynchronize(stream_));
        drop_graph();
        cap_ = c;
        sched_ = sc;
    }
    gto_schedule schedule() const { return sched_; }
    // the fused river sweep runs (schedule, spot shape and engine options allow it)
    bool fused() const { return sched_.river && river_avail_ && show_mode_ == GTO_SHOWDOWN_EXACT && !locked_; }

    // Candidate i of the autotuner: i < 6 level-synchronous sweeps with
    // per-level terminals for i >= 3 and showdown runs of kShowCaps[i % 3]
    // nodes; i = 6 the fused river sweep.
    static gto_schedule candidate(int i) {
        if (i >= 6) return gto_schedule{kShowCaps[0], 0u, 1u, 1u};
        return gto_schedule{kShowCaps[i % 3], (uint32_t)(i / 3), 0u, 1u};
    }

    // GPU milliseconds per iteration of `iters` iterations after a warm-up
    // iteration (which captures the graph of a new schedule)
    float time_iterations(uint32_t iters) {
        iterate(1);
        hipEvent_t e[2];
        HIP_OK(hipEventCreate(&e[0]));
        if (hipEventCreate(&e[1]) != hipSuccess) {
            (void)hipEventDestroy(e[0]);
            throw std::runtime_error("hipEventCreate failed");
        }
        float ms = 0.0f;
        hipError_t rc = hipEventRecord(e[0], stream_);
        if (rc == hipSuccess) {
            iterate(iters);
            rc = hipEventRecord(e[1], stream_);
        }
        if (rc == hipSuccess) rc = hipEventSynchronize(e[1]);
        if (rc == hipSuccess) rc = hipEventElapsedTime(&ms, e[0], e[1]);
        (void)hipEventDestroy(e[0]);
        (void)hipEventDestroy(e[1]);
        HIP_OK(rc);
        return ms / (float)std::max(1u, iters);
    }

    void profile(uint32_t n, gto_hip_profile* out) {
        std::memset(out, 0, sizeof(*out));
        prof_ = out;
        ensure_disc(iteration_ + n + 1);
        for (uint32_t k = 0; k < n; ++k) {
            launch_iteration();
            iteration_ += 1;
        }
        HIP_OK(hipStreamSynchronize(stream_));
        flush_prof();
        prof_ = nullptr;
    }

    uint64_t reached_entries() const { return plan_.reached_entries; }
    uint64_t river_entries() const { return river_avail_ ? plan_.river_entries : 0; }

    // An upper bound of the device memory an engine allocates for `v` with
    // DCFR / CFR+ (PCFR+ adds 4 bytes per arena entry), the exact showdown,
    // counted buffer by buffer over every node (reached or not).
    static uint64_t estimate(const gto_spot_view& v, int storage) {
        const uint64_t n = v.num_nodes, nh[2] = {v.num_hands[0], v.num_hands[1]};
        const uint64_t nhm = std::max(nh[0], nh[1]);
        // reach slots: the root, every action child for its actor, every
        // chance child for both players
        uint64_t slots[2] = {1, 1};
        for (uint64_t i = 0; i < n; ++i) {
            const uint32_t cs = v.node_children_start[i];
            if (v.node_kind[i] == GTO_KIND_ACTION) {
                slots[v.node_player[i]] += v.node_num_children[i];
            } else if (v.node_kind[i] == GTO_KIND_CHANCE) {
                uint64_t c = 0;
                for (int x = 0; x < 52; ++x) c += v.children[cs + x] != 0xFFFFFFFFu;
                slots[0] += c;
                slots[1] += c;
            }
        }
        const uint64_t D = v.data_size[0] + v.data_size[1];
        uint64_t arenas = storage == GTO_STORAGE_F32 ? D * 8 : D * 4 + n * 16;
        if (storage == GTO_STORAGE_HALF || storage == GTO_STORAGE_HALF_ROWS) arenas += n * 16;  // per-node maxima
        uint64_t reach = 0;
        for (int q = 0; q < 2; ++q) reach += slots[q] * (nh[q] * 4 + 9);    // row, alive, max (and slack)
        const uint64_t cfv = n * (nhm * 4 + 4);                               // value rows and their maxima
        const uint64_t records = n * 96;                                      // work records, child lists
        const uint64_t keys = v.num_river_keys;
        const uint64_t boards = keys * ((nh[0] + nh[1]) * 30 + 2 * 52 * 2 * 48 + 2 * 53 * 4);  // showdown tables
        return arenas + reach + cfv + records + boards + (64ull << 20);
    }

   private:
    // ---- suit isomorphism (cfr.rs:568-660, 749-870) -------------------------
    // CFR traverses one chance branch per suit orbit. The other branches are
    // copies of their representative under a suit symmetry; iso_copies_ lists
    // those copies for action nodes in the order symmetrize performs them
    // (representatives' own subtrees first, so a source is complete before it
    // is copied), and downloads replay the list.
    struct IsoCopy {
        uint32_t dst, src, k;
    };
    struct IsoDealt {
        uint8_t c[2] = {0, 0};
        int n = 0;
    };
    bool iso_ = false;
    std::vector<uint32_t> children_start_, children_;
    std::vector<uint8_t> suit_perms_;
    std::vector<uint16_t> hand_perm_[2];
    std::vector<IsoCopy> iso_copies_;
    std::vector<int64_t> iso_src_;  // per node: its entry in iso_copies_, or -1
    std::vector<int64_t> lock_off_h_;
    std::vector<float> lock_sigma_h_;

    uint8_t permute_card(uint8_t c, uint32_t k) const { return (uint8_t)((c & ~3u) | suit_perms_[4 * k + (c & 3u)]); }
    const float* lock_host(uint32_t n) const {
        return (n < lock_off_h_.size() && lock_off_h_[n] >= 0) ? lock_sigma_h_.data() + lock_off_h_[n] : nullptr;
    }

    void build_iso(const gto_spot_view& v) {
        iso_ = true;
        children_start_.assign(v.node_children_start, v.node_children_start + v.num_nodes);
        children_.assign(v.children, v.children + v.num_children_total);
        suit_perms_.assign(v.suit_perms, v.suit_perms + 4 * (size_t)v.num_suit_perms);
        for (int q = 0; q < 2; ++q)
            hand_perm_[q].assign(v.hand_perm[q], v.hand_perm[q] + (size_t)v.num_suit_perms * v.num_hands[q]);
        iso_src_.assign(num_nodes_, -1);
        iso_walk(0, IsoDealt());
    }
    // symmetrize_node, cfr.rs:758-800
    void iso_walk(uint32_t n, IsoDealt d) {
        const uint32_t cs = children_start_[n];
        if (node_kind_[n] == GTO_KIND_ACTION) {
            for (uint32_t a = 0; a < node_na_[n]; ++a) iso_walk(children_[cs + a], d);
        } else if (node_kind_[n] == GTO_KIND_CHANCE) {
            std::vector<uint32_t> valid;
            for (uint32_t k = 0; k < suit_perms_.size() / 4; ++k) {
                bool fix = true;
                for (int i = 0; i < d.n; ++i)
                    if (permute_card(d.c[i], k) != d.c[i]) fix = false;
                if (fix) valid.push_back(k);
            }
            std::vector<std::array<uint32_t, 3>> orbit;  // card, rep, symmetry
            for (uint32_t c = 0; c < 52; ++c) {
                if (children_[cs + c] == 0xFFFFFFFFu) continue;
                bool dealt = false;
                for (int i = 0; i < d.n; ++i) dealt |= d.c[i] == c;
                if (dealt) continue;
                uint32_t rep = c, k_rep = 0;
                for (uint32_t k : valid) {
                    const uint32_t pc = permute_card((uint8_t)c, k);
                    if (pc < rep) {
                        rep = pc;
                        k_rep = k;
                    }
                }
                orbit.push_back({c, rep, k_rep});
            }
            for (const auto& o : orbit)
                if (o[0] == o[1]) {
                    IsoDealt nd = d;
                    nd.c[nd.n++] = (uint8_t)o[0];
                    iso_walk(children_[cs + o[0]], nd);
                }
            for (const auto& o : orbit)
                if (o[0] != o[1]) iso_copy(children_[cs + o[1]], children_[cs + o[0]], o[2]);
        }
    }
    // copy_branch, cfr.rs:806-870
    void iso_copy(uint32_t src, uint32_t dst, uint32_t k) {
        if (node_kind_[src] == GTO_KIND_ACTION) {
            iso_src_[dst] = (int64_t)iso_copies_.size();
            iso_copies_.push_back({dst, src, k});
            const uint32_t ss = children_start_[src], ds = children_start_[dst];
            for (uint32_t a = 0; a < node_na_[src]; ++a) iso_copy(children_[ss + a], children_[ds + a], k);
        } else if (node_kind_[src] == GTO_KIND_CHANCE) {
            const uint32_t ss = children_start_[src], ds = children_start_[dst];
            for (uint32_t x = 0; x < 52; ++x) {
                const uint32_t dc = children_[ds + x];
                if (dc == 0xFFFFFFFFu) continue;
                iso_copy(children_[ss + permute_card((uint8_t)x, k)], dc, k);
            }
        }
    }
    // u16 quantizer of the strategy arenas (store.rs:199-215); returns the scale
    static float encode_u16(const float* src, size_t cnt, uint16_t* q) {
        float m = 0.0f;
        for (size_t e = 0; e < cnt; ++e) m = std::fmax(m, src[e]);
        if (!(m > 0.0f && std::isfinite(m))) {
            std::fill(q, q + cnt, (uint16_t)0);
            return 0.0f;
        }
        const float k = 65535.0f / m;
        for (size_t e = 0; e < cnt; ++e) {
            float r = std::nearbyint(std::fmax(src[e], 0.0f) * k);
            r = r != r ? 0.0f : std::fmin(std::fmax(r, 0.0f), 65535.0f);
            q[e] = (uint16_t)r;
        }
        return m;
    }
    // solved_strategy's normalisation of one block of sums: f64 sums
    // (f32 store) or integer sums of u16 quants (cfr.rs:679-722)
    static void normalize_sums(const float* b, size_t na, size_t nh, float* out, bool quants) {
        for (size_t i = 0; i < nh; ++i) {
            if (quants) {
                uint32_t sum = 0;
                for (size_t a = 0; a < na; ++a) sum += (uint32_t)b[a * nh + i];
                const float inv = sum > 0 ? 1.0f / (float)sum : 0.0f;
                for (size_t a = 0; a < na; ++a) out[a * nh + i] = sum > 0 ? b[a * nh + i] * inv : 1.0f / (float)na;
            } else {
                double sum = 0.0;
                for (size_t a = 0; a < na; ++a) sum += (double)std::max(b[a * nh + i], 0.0f);
                for (size_t a = 0; a < na; ++a)
                    out[a * nh + i] =
                        sum > 1e-12 ? (float)((double)std::max(b[a * nh + i], 0.0f) / sum) : 1.0f / (float)na;
            }
        }
    }
    // ensure_symmetric on a downloaded f32 arena of player p (which 0 =
    // regrets, 1 = strategy sums; a locked source gives its strategy)
    void symmetrize(int which, int p, float* a) const {
        const size_t nh = plan_.nh[p];
        std::vector<float> tmp;
        for (const IsoCopy& c : iso_copies_) {
            if (node_player_[c.dst] != p) continue;
            const size_t na = node_na_[c.dst], cnt = na * nh;
            const uint16_t* tbl = &hand_perm_[p][(size_t)c.k * nh];
            const float* src = a + node_off_[c.src];
            const float* lk = which == 1 ? lock_host(c.src) : nullptr;
            if (lk) {
                tmp.assign(lk, lk + cnt);
                if (compressed_ && !half_) {  // stored through the u16 quantizer, then decoded
                    std::vector<uint16_t> q(cnt);
                    const float m = encode_u16(tmp.data(), cnt, q.data());
                    for (size_t e = 0; e < cnt; ++e) tmp[e] = (float)q[e] * (m / 65535.0f);
                }
                src = tmp.data();
            }
            float* d = a + node_off_[c.dst];
            for (size_t x = 0; x < na; ++x)
                for (size_t i = 0; i < nh; ++i) d[x * nh + i] = src[x * nh + tbl[i]];
        }
    }
    // the same on raw i16/u16 blocks: a permuted block keeps its max, so the
    // quants are permuted and the scale copied
    void symmetrize_raw(int p, int16_t* rq, uint16_t* sq, float* rs, float* ss) const {
        const size_t nh = plan_.nh[p];
        for (const IsoCopy& c : iso_copies_) {
            if (node_player_[c.dst] != p) continue;
            const size_t na = node_na_[c.dst], cnt = na * nh;
            const uint16_t* tbl = &hand_perm_[p][(size_t)c.k * nh];
            const uint64_t so = node_off_[c.src], dof = node_off_[c.dst];
            if (rq)
                for (size_t x = 0; x < na; ++x)
                    for (size_t i = 0; i < nh; ++i) rq[dof + x * nh + i] = rq[so + x * nh + tbl[i]];
            if (rs) rs[c.dst] = rs[c.src];
            if (const float* lk = lock_host(c.src)) {
                std::vector<float> perm(cnt);
                for (size_t x = 0; x < na; ++x)
                    for (size_t i = 0; i < nh; ++i) perm[x * nh + i] = lk[x * nh + tbl[i]];
                std::vector<uint16_t> q(cnt);
                const float m = encode_u16(perm.data(), cnt, q.data());
                if (sq) std::copy(q.begin(), q.end(), sq + dof);
                if (ss) ss[c.dst] = m;
            } else {
                if (sq)
                    for (size_t x = 0; x < na; ++x)
                        for (size_t i = 0; i < nh; ++i) sq[dof + x * nh + i] = sq[so + x * nh + tbl[i]];
                if (ss) ss[c.dst] = ss[c.src];
            }
        }
    }

    gto_hip_options opts_;
    bool compressed_ = false;  // 16-bit arenas (i16/u16 quants or f16)
    bool pred_ = false;        // PCFR+: current strategy matched on S_.Q
    bool half_ = false;        // f16 arenas with per-node power-of-two scales
    std::vector<uint8_t> node_kind_, node_player_, node_na_;
    std::vector<uint64_t> node_off_;
    hipStream_t stream_ = nullptr;
    uint32_t max_card_list_ = 0;  // longest per-card opponent list on any river board
    // a second stream for the fold terminals beside the showdown launch,
    // forked from and joined back into stream_ by events (captured graphs
    // get two parallel branches); lstream_ is the stream launches go to
    hipStream_t aux_ = nullptr, lstream_ = nullptr;
    hipEvent_t fork_ev_ = nullptr, join_ev_ = nullptr;
    Plan plan_;
    uint32_t num_nodes_ = 0;
    uint64_t data_size_[2] = {0, 0};
    double rake_ = 0.0;
    double starting_pot_ = 1.0;
    std::vector<float> weights_[2];
    std::vector<uint8_t> c1_[2], c2_[2];
    std::vector<uint32_t> same_[2];
    std::vector<void*> allocs_;
    DevTree T_{};
    DevState S_{};
    std::vector<DevLevel> lv_;
    uint32_t iteration_ = 0;
    bool reach_valid_[2] = {false, false};
    uint32_t disc_cap_ = 0;
    float* disc_buf_ = nullptr;
    float* lock_sigma_buf_ = nullptr;
    float* grad_buf_ = nullptr;
    float* scratch_ = nullptr;        // this device's copy of the static scratch array
    std::vector<float> root_row_;     // host copy of a root value row (nh_max floats)
    // all-level terminal lists
    FoldRec* fold_all_ = nullptr;
    uint32_t n_fold_all_ = 0;
    ShowRec* show_all_ = nullptr;
    uint4* show_chunks_[3] = {nullptr, nullptr, nullptr};  // all-level runs per entry of kShowCaps
    uint32_t n_show_chunks_[3] = {0, 0, 0};
    gto_schedule sched_{kShowCaps[0], 0, 0, 1};
    // fused river sweeps: subtree records and programs, the LDS they need,
    // the upper tree's terminals; river-level reach rows are left stale by
    // fused sweeps (level sweeps recompute both players' rows after them)
    bool river_avail_ = false, rows_stale_ = false, locked_ = false;
    // rows the DCFR sweeps keep in f16 (per-row power-of-two scales):
    // kRows16Reach | kRows16Value under GTO_STORAGE_HALF_ROWS; other sweeps
    // use f32 rows
    int rows_ = 0;
    RiverSub* river_subs_ = nullptr;
    RiverOp* river_ops_ = nullptr;
    uint32_t n_river_ = 0, river_mo_ = 0;
    size_t river_lds_ = 0;
    FoldRec* fold_up_ = nullptr;
    uint32_t n_fold_up_ = 0;
    ShowRec* show_up_ = nullptr;
    uint4* show_up_chunks_[3] = {nullptr, nullptr, nullptr};
    uint32_t n_show_up_chunks_[3] = {0, 0, 0};
    int cap_ = 0;  // index of sched_.show_nodes in kShowCaps
    uint4* show_chunks_mma_ = nullptr;
    uint32_t n_show_chunks_mma_ = 0;
    size_t show_mma_lds_ = 0;
    bool show_mma_staged_ = false;
    int show_mode_ = GTO_SHOWDOWN_EXACT;
    bool mma_tables_ = false;
    size_t show_lds_[2] = {0, 0};
    size_t fold_lds_[2] = {0, 0};
    hipGraphExec_t graph_ = nullptr;
    uint64_t gen_ = 0;
    gto_hip_profile* prof_ = nullptr;
    std::vector<hipEvent_t> evs_;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> pending_;

    void* dalloc(size_t bytes) {
        void* p = nullptr;
        if (bytes == 0) bytes = 16;
        HIP_OK(hipMalloc(&p, bytes));
        allocs_.push_back(p);
        return p;
    }
    template <class T>
    T* upload_vec(const std::vector<T>& v) {
        T* d = (T*)dalloc(sizeof(T) * std::max<size_t>(v.size(), 1));
        if (!v.empty()) HIP_OK(hipMemcpy(d, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
        return d;
    }
    template <class T>
    T* upload_arr(const T* src, size_t n) {
        T* d = (T*)dalloc(sizeof(T) * std::max<size_t>(n, 1));
        if (n) HIP_OK(hipMemcpy(d, src, sizeof(T) * n, hipMemcpyHostToDevice));
        return d;
    }

    void upload_tree(const gto_spot_view& v) {
        if (plan_.max_actions > (uint32_t)kMaxActions)
            throw std::runtime_error("a node has more than 64 actions; the GPU engine supports at most 64");
        T_.lock_sigma = nullptr;
        T_.lock_bound = 1.0f;
        T_.kids = upload_vec(plan_.kids);
        if (!plan_.xkids.empty()) T_.xkids = upload_vec(plan_.xkids);
        {
            float* z = (float*)dalloc(sizeof(float) * std::max<uint32_t>(plan_.nh_max, 1));
            HIP_OK(hipMemset(z, 0, sizeof(float) * std::max<uint32_t>(plan_.nh_max, 1)));
            T_.zero_row = z;
        }
        T_.chr_card = upload_vec(plan_.chr_card);
        T_.chr_cfv = upload_vec(plan_.chr_cfv);
        T_.chr_perm = upload_vec(plan_.chr_perm);
        if (plan_.iso)
            for (int q = 0; q < 2; ++q)
                T_.hand_perm[q] = upload_vec(std::vector<uint16_t>(
                    v.hand_perm[q], v.hand_perm[q] + (size_t)v.num_suit_perms * v.num_hands[q]));
        T_.nh_max = plan_.nh_max;
        for (int q = 0; q < 2; ++q) {
            const uint32_t nh = v.num_hands[q];
            T_.nh[q] = nh;
            T_.c1[q] = upload_arr(v.hand_c1[q], nh);
            T_.c2[q] = upload_arr(v.hand_c2[q], nh);
            std::vector<uint64_t> m(nh);
            for (uint32_t i = 0; i < nh; ++i) m[i] = (1ull << v.hand_c1[q][i]) | (1ull << v.hand_c2[q][i]);
            T_.hmask[q] = upload_vec(m);
            T_.same[q] = upload_arr(v.same_combo[q], nh);
            std::vector<uint32_t> hp(nh);
            for (uint32_t i = 0; i < nh; ++i) {
                uint32_t sm = v.same_combo[q][i];
                if (sm != 0xFFFFFFFFu && sm >= 0xFFFFu) throw std::runtime_error("too many hands for the packed hand table");
                hp[i] = v.hand_c1[q][i] | (v.hand_c2[q][i] << 8) | ((sm == 0xFFFFFFFFu ? 0xFFFFu : sm) << 16);
            }
            T_.hpack[q] = upload_vec(hp);
            T_.weights[q] = upload_arr(v.weights[q], nh);
            T_.sd_off[q] = upload_vec(plan_.sd_off[q]);
            T_.sd_cnt[q] = upload_vec(plan_.sd_cnt[q]);
            T_.sd_idx[q] = upload_vec(plan_.sd_idx[q]);
            T_.sd_coff[q] = upload_vec(plan_.sd_coff[q]);
            T_.sd_tab[q] = (const uint4*)upload_vec(plan_.sd_tab[q]);
            T_.sd_cpos[q] = upload_vec(plan_.sd_cpos[q]);
        }
        // LDS: showdown, per workgroup, for traverser p (largest board)
        for (int p = 0; p < 2; ++p) {
            const uint32_t o = 1 - p;
            size_t need = 0;
            for (uint32_t b = 0; b < plan_.num_boards; ++b) {
                const uint32_t* cf = &plan_.sd_coff[o][(size_t)b * 53];
                for (int c = 0; c < 52; ++c) {
                    if (cf[c + 1] - cf[c] > (uint32_t)kMaxCardList)
                        throw std::runtime_error("a card appears in more opponent hands than the showdown kernel holds");
                    max_card_list_ = std::max(max_card_list_, cf[c + 1] - cf[c]);
                }
                need = std::max<size_t>(need, show_lds_layout(plan_.sd_cnt[o][b], plan_.nh[p], plan_.nh[o]).bytes);
            }
            show_lds_[p] = need;
            fold_lds_[p] = kFoldWaveBytes * kWaves;
        }
        size_t lds_max = std::max(std::max(show_lds_[0], show_lds_[1]), std::max(fold_lds_[0], fold_lds_[1]));
        if (lds_max > kLdsBudget)
            throw std::runtime_error("showdown tables exceed the LDS budget (too many hands per board)");
        for (const int rows : {0, kRows16Reach, kRows16Reach | kRows16Value})
            for (const bool seq : {false, true}) {
                HIP_OK(hipFuncSetAttribute((const void*)show_kernel(rows, seq),
                                           hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)std::max(show_lds_[0], show_lds_[1])));
                HIP_OK(hipFuncSetAttribute((const void*)fold_kernel(rows, seq),
                                           hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)std::max(fold_lds_[0], fold_lds_[1])));
            }
        n_fold_all_ = (uint32_t)plan_.fold_all.size();
        if (n_fold_all_) fold_all_ = upload_vec(plan_.fold_all);
        n_show_chunks_mma_ = (uint32_t)(plan_.show_chunks_mma.size() / 4);
        if (!plan_.show_all.empty()) {
            show_all_ = upload_vec(plan_.show_all);
            for (int c = 0; c < 3; ++c) {
                n_show_chunks_[c] = (uint32_t)(plan_.show_chunks[c].size() / 4);
                show_chunks_[c] = (uint4*)upload_vec(plan_.show_chunks[c]);
            }
            show_chunks_mma_ = (uint4*)upload_vec(plan_.show_chunks_mma);
        }
        lv_.resize(plan_.levels.size());
        for (size_t l = 0; l < plan_.levels.size(); ++l) {
            const LevelLists& L = plan_.levels[l];
            DevLevel& D = lv_[l];
            for (int q = 0; q < 2; ++q) {
                D.n_act[q] = (uint32_t)L.act_rec[q].size();
                if (D.n_act[q]) D.act[q] = upload_vec(L.act_rec[q]);
                for (const ActRec& r : L.act_rec[q]) D.max_na[q] = std::max<uint32_t>(D.max_na[q], r.na);
                D.n_act_keep[q] = (uint32_t)L.act_keep[q].size();
                if (D.n_act_keep[q]) D.act_keep[q] = upload_vec(L.act_keep[q]);
            }
            D.n_fold = (uint32_t)L.fold_rec.size();
            if (D.n_fold) D.fold = upload_vec(L.fold_rec);
            D.n_show = (uint32_t)L.show_rec.size();
            if (D.n_show) {
                D.show = upload_vec(L.show_rec);
                for (int c = 0; c < 3; ++c) {
                    D.n_show_chunks[c] = (uint32_t)(L.show_chunks[c].size() / 4);
                    D.show_chunks[c] = (uint4*)upload_vec(L.show_chunks[c]);
                }
            }
            D.n_chance = (uint32_t)L.chance_rec.size();
            if (D.n_chance) D.chance = upload_vec(L.chance_rec);
            D.n_edge = (uint32_t)L.edge_rec.size();
            if (D.n_edge) D.edge = upload_vec(L.edge_rec);
            for (int q = 0; q < 2; ++q) {
                D.n_act_up[q] = (uint32_t)L.act_up[q].size();
                if (D.n_act_up[q]) D.act_up[q] = upload_vec(L.act_up[q]);
            }
            D.n_edge_up = (uint32_t)L.edge_up.size();
            if (D.n_edge_up) D.edge_up = upload_vec(L.edge_up);
        }
        upload_river();
    }

    // The fused river sweep needs the programs, f32 or f16 arenas without
    // PCFR+ (the other formats keep the level sweeps) and its LDS rows
    // within the workgroup's 160 KB.
    void upload_river() {
        river_avail_ = false;
        if (!plan_.river_ok || plan_.river_subs.empty() || (compressed_ && !half_) || pred_ || rows_) return;
        river_mo_ = std::max(plan_.max_board_cnt[0], plan_.max_board_cnt[1]);
        river_lds_ = river_lds_layout(plan_.river_rows, plan_.nh_max, river_mo_).bytes;
        if (river_lds_ > kLdsBudget) return;
        for (const void* f : {(const void*)river_kernel<kStoreF32>(), (const void*)river_kernel<kStoreH>()})
            HIP_OK(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)river_lds_));
        river_subs_ = upload_vec(plan_.river_subs);
        river_ops_ = upload_vec(plan_.river_ops);
        n_river_ = (uint32_t)plan_.river_subs.size();
        n_fold_up_ = (uint32_t)plan_.fold_up.size();
        if (n_fold_up_) fold_up_ = upload_vec(plan_.fold_up);
        if (!plan_.show_up.empty()) {
            show_up_ = upload_vec(plan_.show_up);
            for (int c = 0; c < 3; ++c) {
                n_show_up_chunks_[c] = (uint32_t)(plan_.show_up_chunks[c].size() / 4);
                show_up_chunks_[c] = (uint4*)upload_vec(plan_.show_up_chunks[c]);
            }
        }
        river_avail_ = true;
    }

    void alloc_state() {
        for (int q = 0; q < 2; ++q) {
            size_t ds = std::max<uint64_t>(data_size_[q], 1);
            if (compressed_) {
                S_.Rq[q] = (int16_t*)dalloc(sizeof(int16_t) * ds);
                S_.Sq[q] = (uint16_t*)dalloc(sizeof(uint16_t) * ds);
                S_.Rs[q] = (float*)dalloc(sizeof(float) * num_nodes_);
                S_.Ss[q] = (float*)dalloc(sizeof(float) * num_nodes_);
                HIP_OK(hipMemset(S_.Rq[q], 0, sizeof(int16_t) * ds));
                HIP_OK(hipMemset(S_.Sq[q], 0, sizeof(uint16_t) * ds));
                HIP_OK(hipMemset(S_.Rs[q], 0, sizeof(float) * num_nodes_));
                HIP_OK(hipMemset(S_.Ss[q], 0, sizeof(float) * num_nodes_));
                if (half_) {
                    S_.Rm[q] = (float*)dalloc(sizeof(float) * num_nodes_);
                    S_.Sm[q] = (float*)dalloc(sizeof(float) * num_nodes_);
                    HIP_OK(hipMemset(S_.Rm[q], 0, sizeof(float) * num_nodes_));
                    HIP_OK(hipMemset(S_.Sm[q], 0, sizeof(float) * num_nodes_));
                }
            } else {
                S_.R[q] = (float*)dalloc(sizeof(float) * ds);
                S_.S[q] = (float*)dalloc(sizeof(float) * ds);
                HIP_OK(hipMemset(S_.R[q], 0, sizeof(float) * ds));
                HIP_OK(hipMemset(S_.S[q], 0, sizeof(float) * ds));
                S_.cur[q] = S_.R[q];
                if (pred_) {
                    // Q = max(R,0) + prediction; R = 0 and no prediction yet
                    S_.Q[q] = (float*)dalloc(sizeof(float) * ds);
                    HIP_OK(hipMemset(S_.Q[q], 0, sizeof(float) * ds));
                    S_.cur[q] = S_.Q[q];
                }
            }
            S_.reach[q] = (float*)dalloc(sizeof(float) * (size_t)plan_.num_slots[q] * plan_.nh[q]);
            S_.alive[q] = (uint8_t*)dalloc(plan_.num_slots[q]);
            S_.rmax[q] = (float*)dalloc(sizeof(float) * plan_.num_slots[q]);
            HIP_OK(hipMemset(S_.rmax[q], 0, sizeof(float) * plan_.num_slots[q]));
            HIP_OK(hipMemset(S_.alive[q], 0, plan_.num_slots[q]));
        }
        S_.cfv = (float*)dalloc(sizeof(float) * (size_t)plan_.num_reached * plan_.nh_max);
        S_.vmax = nullptr;
        S_.rsc[0] = S_.rsc[1] = S_.vsc = nullptr;
        if (rows_) {
            for (int q = 0; q < 2; ++q) {
                S_.rsc[q] = (float*)dalloc(sizeof(float) * std::max<uint32_t>(plan_.num_slots[q], 1));
                HIP_OK(hipMemset(S_.rsc[q], 0, sizeof(float) * std::max<uint32_t>(plan_.num_slots[q], 1)));
            }
            S_.vsc = (float*)dalloc(sizeof(float) * std::max<uint32_t>(plan_.num_reached, 1));
            HIP_OK(hipMemset(S_.vsc, 0, sizeof(float) * std::max<uint32_t>(plan_.num_reached, 1)));
        }
        if (half_) {
            S_.vmax = (float*)dalloc(sizeof(float) * std::max<uint32_t>(plan_.num_reached, 1));
            HIP_OK(hipMemset(S_.vmax, 0, sizeof(float) * std::max<uint32_t>(plan_.num_reached, 1)));
        }
        S_.iter = (uint32_t*)dalloc(sizeof(uint32_t));
        HIP_OK(hipMemset(S_.iter, 0, sizeof(uint32_t)));
        ensure_disc(1024);
        HIP_OK(hipDeviceSynchronize());
    }

    void ensure_disc(uint32_t need) {
        if (need < disc_cap_) return;
        uint32_t cap = std::max<uint32_t>(disc_cap_ * 2, 1024);
        while (cap <= need) cap *= 2;
        std::vector<float> h((size_t)cap * 3);
        h[0] = h[1] = h[2] = 0.0f;
        for (uint32_t t = 1; t < cap; ++t) discounts(opts_.algo, t, &h[3 * (size_t)t]);
        HIP_OK(hipStreamSynchronize(stream_));
        float* d = (float*)dalloc(sizeof(float) * h.size());
        HIP_OK(hipMemcpy(d, h.data(), sizeof(float) * h.size(), hipMemcpyHostToDevice));
        if (disc_buf_) {
            drop_graph();
            HIP_OK(hipFree(disc_buf_));
            allocs_.erase(std::find(allocs_.begin(), allocs_.end(), (void*)disc_buf_));
        }
        disc_buf_ = d;
        S_.disc = d;
        disc_cap_ = cap;
    }

    void drop_graph() {
        ++gen_;
        if (graph_) {
            (void)hipGraphExecDestroy(graph_);
            graph_ = nullptr;
        }
    }

    // --- launch helpers with optional per-kind event timing ---------------
    template <class K, class... Args>
    void launch(int kind, K kernel, uint32_t grid, size_t lds, Args... args) {
        launch_b(kind, kernel, grid, kBlock, lds, args...);
    }
    template <class K, class... Args>
    void launch_b(int kind, K kernel, uint32_t grid, uint32_t block, size_t lds, Args... args) {
        if (grid == 0) return;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (prof_) {
            e0 = event();
            HIP_OK(hipEventRecord(e0, stream_));
        }
        hipLaunchKernelGGL(kernel, dim3(grid), dim3(block), lds, lstream_, args...);
        HIP_OK(hipGetLastError());
        if (prof_) {
            e1 = event();
            HIP_OK(hipEventRecord(e1, stream_));
            pending_.push_back({kind, {e0, e1}});
        }
    }
    hipEvent_t event() {
        hipEvent_t e;
        HIP_OK(hipEventCreate(&e));
        evs_.push_back(e);
        return e;
    }
    void flush_prof() {
        for (auto& pe : pending_) {
            float ms = 0.0f;
            HIP_OK(hipEventElapsedTime(&ms, pe.second.first, pe.second.second));
            prof_->ms[pe.first] += ms;
            prof_->launches[pe.first] += 1;
        }
        pending_.clear();
        for (auto& e : evs_) (void)hipEventDestroy(e);
        evs_.clear();
    }

    void launch_root(int mask, bool h16 = false) {
        if (h16) launch(PK_ROOT, k_root<true>, 1, 0, T_, S_, mask);
        else launch(PK_ROOT, k_root<false>, 1, 0, T_, S_, mask);
    }

    // upper: only the nodes outside river subtrees (fused river sweeps)
    // h16: f16 reach rows (the DCFR sweeps of GTO_STORAGE_HALF_ROWS)
    void launch_down(int mask, int kind, bool upper = false, bool h16 = false) {
        for (size_t l = 0; l < lv_.size(); ++l) {
            const DevLevel& D = lv_[l];
            const uint32_t ne = upper ? D.n_edge_up : D.n_edge;
            const EdgeRec* er = (const EdgeRec*)(upper ? D.edge_up : D.edge);
            // the level's chance edges ride along with its first action launch
            bool edges_done = ne == 0;
            for (int q = 0; q < 2; ++q) {
                const uint32_t n = upper ? D.n_act_up[q] : D.n_act[q];
                if (!((mask >> q) & 1) || !n) continue;
                const ActRec* recs = upper ? D.act_up[q] : D.act[q];
                const uint32_t e = edges_done ? 0u : ne;
                if (!h16 && n < kNodeWgMax && D.max_na[q] <= (uint32_t)kMaxUnroll) {
                    // few nodes: a workgroup per node, then the edges
                    launch(PK_DOWN_ACTION, node_kernel(kind), wave_grid(n + (e + kWaves - 1) / kWaves, 1), 0, T_, S_,
                           recs, n, q, er, e, mask);
                    edges_done = true;
                    continue;
                }
                launch(PK_DOWN_ACTION, down_kernel(kind, h16), wave_grid(n + e), 0, T_, S_, recs, n, q, er, e, mask);
                edges_done = true;
            }
            if (!edges_done) {
                if (h16) launch(PK_DOWN_CHANCE, k_down_chance<true>, wave_grid(ne), 0, T_, S_, er, ne, mask);
                else launch(PK_DOWN_CHANCE, k_down_chance<false>, wave_grid(ne), 0, T_, S_, er, ne, mask);
            }
        }
    }

    // levels with fewer action nodes than this (of one player) take a
    // workgroup per node in the down sweep (k_down_node)
    static constexpr uint32_t kNodeWgMax = 2048;
    using NodeFn = void (*)(DevTree, DevState, const ActRec*, uint32_t, int, const EdgeRec*, uint32_t, int);
    NodeFn node_kernel(int kind) const {
        if (kind == kRawF32) return k_down_node<kRawF32>;
        if (kind == kAverageF32)
            return half_ ? k_down_node<kAverageH> : compressed_ ? k_down_node<kAverageQ> : k_down_node<kAverageF32>;
        return half_ ? k_down_node<kCurrentH> : compressed_ ? k_down_node<kCurrentQ> : k_down_node<kCurrentF32>;
    }
    using DownFn = void (*)(DevTree, DevState, const ActRec*, uint32_t, int, const EdgeRec*, uint32_t, int);
    // the down-sweep level kernel for a strategy kind (current / average /
    // given) of this engine's arenas; h16: f16 reach rows (DCFR sweeps of
    // f16-row engines)
    DownFn down_kernel(int kind, bool h16) const {
        if (h16) {
            if (kind != kCurrentF32) throw std::runtime_error("f16 reach rows serve the DCFR sweeps only");
            return k_down_level<kCurrentH, true>;
        }
        if (kind == kRawF32) return k_down_level<kRawF32>;
        if (kind == kAverageF32)
            return half_ ? k_down_level<kAverageH> : compressed_ ? k_down_level<kAverageQ> : k_down_level<kAverageF32>;
        return half_ ? k_down_level<kCurrentH> : compressed_ ? k_down_level<kCurrentQ> : k_down_level<kCurrentF32>;
    }

    // mode 0 = DCFR update, 1 = best response, 2 = average values, 3 = current
    // values, 4 = values of the strategy arena as given (f32 only); swap:
    // terminal payoffs of the opponent instead of the traverser. A showdown
    // precision form serves the solve's own sweeps (0, 3); best response,
    // average-strategy values and the game value always take the exact
    // showdown, so exploitability judges every form with the same arithmetic.
    // the exact showdown kernel sized to the larger hand count (3 rows per
    // thread cover up to 768 hands) and to the longest per-card list
    // the fused river kernel instance: hands per thread and card-list
    // segments as the exact showdown kernel's
    using RiverFn = void (*)(DevTree, DevState, const RiverSub*, uint32_t, const RiverOp*, int, uint32_t, uint32_t);
    using ShowFn = void (*)(DevTree, DevState, const ShowRec*, const uint4*, uint32_t, int, int);
    template <int STORE>
    RiverFn river_kernel() const {
        if (std::max(plan_.nh[0], plan_.nh[1]) > 3u * kBlock) return k_river<kRowPer, kMaxCardList / 4, STORE>;
        return max_card_list_ <= 4u * kShortSeg ? k_river<3, kShortSeg, STORE> : k_river<3, kMaxCardList / 4, STORE>;
    }

    // Up sweep of a fused DCFR iteration for traverser p: the upper tree's
    // fold terminals on the second stream beside the river kernel (one
    // workgroup per river subtree: its reach, terminals and updates), then
    // the upper tree's levels bottom-up (the river chance nodes merge the
    // subtree values).
    void launch_up_fused(int p) {
        const bool fold = n_fold_up_ != 0, fork = fold && !prof_;
        if (fork) {
            HIP_OK(hipEventRecord(fork_ev_, stream_));
            HIP_OK(hipStreamWaitEvent(aux_, fork_ev_, 0));
            lstream_ = aux_;
        }
        std::exception_ptr err;
        try {
            if (fold)
                launch(PK_UP_FOLD, k_up_fold<0>, wave_grid(fold_waves(n_fold_up_)), fold_lds_[p], T_, S_,
                       (const FoldRec*)fold_up_, n_fold_up_, p, 0, fold_per(n_fold_up_));
        } catch (...) {
            err = std::current_exception();
        }
        lstream_ = stream_;
        if (!err) try {
                const uint32_t g = (n_river_ + 7u) & ~7u;
                if (half_)
                    launch(PK_RIVER, river_kernel<kStoreH>(), g, river_lds_, T_, S_, (const RiverSub*)river_subs_,
                           n_river_, (const RiverOp*)river_ops_, p, plan_.river_rows, river_mo_);
                else
                    launch(PK_RIVER, river_kernel<kStoreF32>(), g, river_lds_, T_, S_, (const RiverSub*)river_subs_,
                           n_river_, (const RiverOp*)river_ops_, p, plan_.river_rows, river_mo_);
                if (n_show_up_chunks_[0])
                    launch(PK_UP_SHOW, show_kernel(0), wave_grid(n_show_up_chunks_[0], 1), show_lds_[p], T_, S_,
                           (const ShowRec*)show_up_, (const uint4*)show_up_chunks_[0], n_show_up_chunks_[0], p, 0);
            } catch (...) {
                err = std::current_exception();
            }
        if (fork) {
            const hipError_t r1 = hipEventRecord(join_ev_, aux_);
            const hipError_t r2 = r1 == hipSuccess ? hipStreamWaitEvent(stream_, join_ev_, 0) : r1;
            if (!err) HIP_OK(r2);
        }
        if (err) std::rethrow_exception(err);
        for (size_t l = lv_.size(); l-- > 0;) {
            const DevLevel& D = lv_[l];
            if (D.n_chance)
                launch_b(PK_UP_CHANCE, k_up_chance<false>, wave_grid(D.n_chance, 1),
                         kChanceBlock, 0, T_, S_, (const ChanceRec*)D.chance, D.n_chance, p);
            for (int q = 0; q < 2; ++q) {
                if (!D.n_act_up[q]) continue;
                const ActRec* recs = D.act_up[q];
                const uint32_t n = D.n_act_up[q], g = wave_grid(n);
                if (half_) launch(PK_UP_ACTION, k_up_action<0, kStoreH>, g, 0, T_, S_, recs, n, p, q);
                else launch(PK_UP_ACTION, k_up_action<0, kStoreF32>, g, 0, T_, S_, recs, n, p, q);
            }
        }
    }

    // rows: the f16 rows of the sweep (0, kRows16Reach, or both kinds)
    // fold and showdown f64 sums in the CPU engine's order (f32 rows only)
    bool seq_sums() const { return show_mode_ == GTO_SHOWDOWN_SEQUENTIAL; }
    template <int ROWS, bool SEQ = false>
    ShowFn show_kernel_of() const {
        if (std::max(plan_.nh[0], plan_.nh[1]) > 3u * kBlock) return k_up_show<kRowPer, kMaxCardList / 4, ROWS, SEQ>;
        return max_card_list_ <= 4u * kShortSeg ? k_up_show<3, kShortSeg, ROWS, SEQ>
                                                 : k_up_show<3, kMaxCardList / 4, ROWS, SEQ>;
    }
    ShowFn show_kernel(int rows, bool seq = false) const {
        if (rows == kRows16Reach) return show_kernel_of<kRows16Reach>();
        if (rows) return show_kernel_of<kRows16Reach | kRows16Value>();
        return seq ? show_kernel_of<0, true>() : show_kernel_of<0>();
    }
    using FoldFn = void (*)(DevTree, DevState, const FoldRec*, uint32_t, int, int, uint32_t);
    static FoldFn fold_kernel(int rows, bool seq = false) {
        if (rows == kRows16Reach) return k_up_fold<kRows16Reach>;
        if (rows) return k_up_fold<kRows16Reach | kRows16Value>;
        return seq ? k_up_fold<0, true> : k_up_fold<0>;
    }

    void launch_up(int p, int mode, int swap = 0) {
        // the DCFR sweeps of f16-row engines read and write f16 rows
        const int rows = mode == 0 ? rows_ : 0;
        // DCFR sweeps of f32 engines without locks: the traverser's nodes sum
        // their absorbed opponent children's children themselves (f16 arenas
        // would bound those sums by their children's maxima, a different
        // scale than the child row's own maximum, so they keep the rows)
        const bool absorb = mode == 0 && sched_.absorb && !compressed_ && !pred_ && !locked_ && T_.xkids;
        const FoldFn fold_fn = fold_kernel(rows, seq_sums());
        const ShowFn show_fn = show_kernel(rows, seq_sums());
        const bool mma = n_show_chunks_mma_ && (show_mode_ == GTO_SHOWDOWN_BF16 || show_mode_ == GTO_SHOWDOWN_FP8) &&
                         (mode == 0 || mode == 3);
        if (mma)
            launch_b(PK_UP_SHOW, show_mode_ == GTO_SHOWDOWN_FP8 ? k_up_show_mma<true> : k_up_show_mma<false>,
                     wave_grid(n_show_chunks_mma_, 1), kMmaBlock, show_mma_lds_, T_, S_, (const ShowRec*)show_all_,
                     (const uint4*)show_chunks_mma_, n_show_chunks_mma_, p, swap, show_mma_staged_ ? 1 : 0);
        if (!sched_.per_level) {
            // terminals of every level first: they read only the down pass.
            // Fold and showdown are independent: the fold launch runs on the
            // second stream beside the showdown (not while kernels are timed
            // one by one)
            const bool fold = n_fold_all_ != 0, show = !mma && n_show_chunks_[cap_] != 0;
            const bool fork = fold && show && !prof_;
            if (fork) {
                HIP_OK(hipEventRecord(fork_ev_, stream_));
                HIP_OK(hipStreamWaitEvent(aux_, fork_ev_, 0));
                lstream_ = aux_;
            }
            // the branch is joined back whatever fails in between, so a
            // capture in progress can end and later work on stream_ still
            // orders after the fold
            std::exception_ptr err;
            try {
                if (fold)
                    // beside the showdown on the second stream: full batches
                    // (fewer waves measured faster here even on small trees:
                    // turn spot 0.853 against 0.891 ms per iteration)
                    launch(PK_UP_FOLD, fold_fn, wave_grid((n_fold_all_ + kFoldNodes - 1) / kFoldNodes), fold_lds_[p],
                           T_, S_, (const FoldRec*)fold_all_, n_fold_all_, p, swap, kFoldNodes);
            } catch (...) {
                err = std::current_exception();
            }
            lstream_ = stream_;
            if (!err) try {
                    if (show)
                        launch(PK_UP_SHOW, show_fn, wave_grid(n_show_chunks_[cap_], 1), show_lds_[p], T_, S_,
                               (const ShowRec*)show_all_, (const uint4*)show_chunks_[cap_], n_show_chunks_[cap_],
                               p, swap);
                } catch (...) {
                    err = std::current_exception();
                }
            if (fork) {
                const hipError_t r1 = hipEventRecord(join_ev_, aux_);
                const hipError_t r2 = r1 == hipSuccess ? hipStreamWaitEvent(stream_, join_ev_, 0) : r1;
                if (!err) HIP_OK(r2);
            }
            if (err) std::rethrow_exception(err);
        }
        for (size_t l = lv_.size(); l-- > 0;) {
            const DevLevel& D = lv_[l];
            if (sched_.per_level) {
                // this level's terminals, just before the kernels that read them
                if (D.n_fold)
                    launch(PK_UP_FOLD, fold_fn, wave_grid(fold_waves(D.n_fold)), fold_lds_[p], T_, S_, (const FoldRec*)D.fold,
                           D.n_fold, p, swap, fold_per(D.n_fold));
                if (!mma && D.n_show_chunks[cap_])
                    launch(PK_UP_SHOW, show_fn, wave_grid(D.n_show_chunks[cap_], 1), show_lds_[p], T_, S_,
                           (const ShowRec*)D.show, (const uint4*)D.show_chunks[cap_], D.n_show_chunks[cap_], p, swap);
            }
            if (D.n_chance)
                launch_b(PK_UP_CHANCE, (rows & kRows16Value) ? k_up_chance<true> : k_up_chance<false>, wave_grid(D.n_chance, 1),
                         kChanceBlock, 0, T_, S_, (const ChanceRec*)D.chance, D.n_chance, p);
            // both players' action nodes of the level in one launch; with
            // absorbed opponent nodes only the opponent's kept nodes
            const ActRec* l0 = D.act[0];
            const ActRec* l1 = D.act[1];
            uint32_t n0 = D.n_act[0], n1 = D.n_act[1];
            if (absorb && p == 0) l1 = D.act_keep[1], n1 = D.n_act_keep[1];
            if (absorb && p == 1) l0 = D.act_keep[0], n0 = D.n_act_keep[0];
            if (n0 + n1) {
                const uint32_t ns = up_slices(n0 + n1, rows);
                launch(PK_UP_ACTION, up_kernel(mode, rows), wave_grid((n0 + n1) * ns), 0, T_, S_, l0, n0, l1, n1, p,
                       absorb ? 1 : 0, ns);
            }
        }
    }

    using UpFn = void (*)(DevTree, DevState, const ActRec*, uint32_t, const ActRec*, uint32_t, int, int, uint32_t);
    // Wavefronts per node of an update launch over n nodes: a level with
    // few nodes (the upper levels; every level of a small tree) splits each
    // node's traverser hands over up to kMaxSlices wavefronts of at least
    // 128 hands, for about kFillWaves wavefronts in all (f32 arenas and
    // rows only: the other forms reduce over whole rows).
    static constexpr uint32_t kFillWaves = 8192, kMaxSlices = 8;
    uint32_t up_slices(uint32_t n, int rows) const {
        if (rows || compressed_ || half_ || n >= kFillWaves) return 1;
        const uint32_t by_hands = std::max<uint32_t>(1, std::max(plan_.nh[0], plan_.nh[1]) / 128);
        return std::max<uint32_t>(1, std::min({kFillWaves / n, by_hands, kMaxSlices}));
    }
    // the up-sweep action kernel for a sweep mode (0 DCFR update, 1 best
    // response, 2 average, 3 current, 4 given strategy) and the sweep's rows
    UpFn up_kernel(int mode, int rows) const {
        if (half_) {
            switch (mode) {
                case 0:
                    if (rows == kRows16Reach) return k_up_action2<0, kStoreH, kRows16Reach>;
                    if (rows) return k_up_action2<0, kStoreH, kRows16Reach | kRows16Value>;
                    return k_up_action2<0, kStoreH>;
                case 1: return k_up_action2<1, kStoreH>;
                case 2: return k_up_action2<2, kStoreH>;
                default: return k_up_action2<3, kStoreH>;
            }
        }
        if (compressed_) {
            switch (mode) {
                case 0: return k_up_action2<0, kStoreQ>;
                case 1: return k_up_action2<1, kStoreQ>;
                case 2: return k_up_action2<2, kStoreQ>;
                default: return k_up_action2<3, kStoreQ>;
            }
        }
        switch (mode) {
            case 0: return pred_ ? k_up_action2<0, kStoreF32P> : k_up_action2<0, kStoreF32>;
            case 1: return k_up_action2<1, kStoreF32>;
            case 2: return k_up_action2<2, kStoreF32>;
            case 4: return k_up_action2<4, kStoreF32>;
            default: return k_up_action2<3, kStoreF32>;
        }
    }

    // One DCFR iteration: traverser 0 then 1 (cfr.rs:170-190). Each sweep
    // recomputes only the reach vectors whose regrets changed since they
    // were last pushed down.
    void launch_iteration() {
        launch(PK_OTHER, k_begin_iteration, 1, 0, S_);
        const bool fu = fused();
        if (!fu && rows_stale_) {
            // fused sweeps leave the river-level reach rows behind: a level
            // sweep after them recomputes both players' rows
            reach_valid_[0] = reach_valid_[1] = false;
            rows_stale_ = false;
        }
        for (int p = 0; p < 2; ++p) {
            int mask = 0;
            for (int q = 0; q < 2; ++q)
                if (!reach_valid_[q]) mask |= 1 << q;
            if (mask) {
                launch_root(mask, (rows_ & kRows16Reach) != 0);
                launch_down(mask, kCurrentF32, fu, (rows_ & kRows16Reach) != 0);
            }
            reach_valid_[0] = reach_valid_[1] = true;
            if (fu) launch_up_fused(p);
            else launch_up(p, 0);
            reach_valid_[p] = false;
        }
        if (fu) rows_stale_ = true;
    }

    void capture_graph() {
        HIP_OK(hipStreamSynchronize(stream_));
        bool v0 = reach_valid_[0], v1 = reach_valid_[1];
        HIP_OK(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
        try {
            launch_iteration();
        } catch (...) {
            hipGraph_t g;
            (void)hipStreamEndCapture(stream_, &g);
            throw;
        }
        hipGraph_t g = nullptr;
        HIP_OK(hipStreamEndCapture(stream_, &g));
        HIP_OK(hipGraphInstantiate(&graph_, g, nullptr, nullptr, 0));
        HIP_OK(hipGraphDestroy(g));
        reach_valid_[0] = v0;
        reach_valid_[1] = v1;
    }

    double pair_weight_sum() const {  // best_response.rs:227-251
        double t = 0.0, s[52];
        std::fill(s, s + 52, 0.0);
        for (size_t j = 0; j < weights_[1].size(); ++j) {
            double r = (double)weights_[1][j];
            t += r;
            s[c1_[1][j]] += r;
            s[c2_[1][j]] += r;
        }
        double total = 0.0;
        for (size_t i = 0; i < weights_[0].size(); ++i) {
            uint32_t sm = same_[0][i];
            double same_r = sm != 0xFFFFFFFFu ? (double)weights_[1][sm] : 0.0;
            double valid = t - s[c1_[0][i]] - s[c2_[0][i]] + same_r;
            total += (double)weights_[0][i] * valid;
        }
        return total;
    }

    double dot(int p, const std::vector<float>& c) const { return dot(p, c.data(), c.size()); }
    double dot(int p, const float* c, size_t n) const {
        double acc = 0.0;
        for (size_t i = 0; i < n; ++i) acc += (double)weights_[p][i] * (double)c[i];
        return acc;
    }
};

template <class F>
int guarded(F&& f) {
    try {
        f();
        return GTO_OK;
    } catch (const std::bad_alloc&) {
        set_last_error("out of host memory");
        return GTO_ERR_MEMORY;
    } catch (const DeviceOutOfMemory& e) {
        set_last_error(e.what());
        return GTO_ERR_MEMORY;
    } catch (const std::invalid_argument& e) {
        set_last_error(e.what());
        return GTO_ERR_INVALID;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return GTO_ERR_DEVICE;
    }
}

}  // namespace

struct gto_hip {
    std::unique_ptr<Engine> e;
};

// A batch of engines (independent spots, e.g. the flops of a report) that
// iterate together: each batch iteration launches every engine's captured
// iteration graph on that engine's own stream, engines interleaved, and the
// batch synchronizes once at the end. Spots' kernels then share the device
// through its hardware queues. Measured on 16 report flops: 4.7 ms per
// flop-iteration, against 5.4 ms for the engines one after another and 5.6
// ms for child graphs of parent graphs (a parent graph's children ran one
// after another, 1.04 kernels in flight on average, even with one parent
// graph per stream).
struct gto_batch {
    std::vector<Engine*> eng;

    void iterate(uint32_t n) {
        for (Engine* e : eng) e->reserve_iterations(n);
        for (uint32_t k = 0; k < n; ++k)
            for (Engine* e : eng) e->iterate(1);
        for (Engine* e : eng) e->synchronize();
    }

    // wall milliseconds per batch iteration of `iters` iterations after a
    // warm-up iteration (which rebuilds the parent graph after a change)
    float time_iterations(uint32_t iters) {
        iterate(1);
        const auto t0 = std::chrono::steady_clock::now();
        iterate(iters);
        const std::chrono::duration<double, std::milli> dt = std::chrono::steady_clock::now() - t0;
        return (float)(dt.count() / std::max(1u, iters));
    }
};

// The autotuner: every candidate schedule for `iters` timed iterations,
// then the fastest is applied. Schedules do not change results, so the
// iterations spent tuning are ordinary iterations of the solve.
template <class Apply, class Time>
static gto_schedule autotune(uint32_t iters, Apply apply, Time time, float* ms) {
    int best = 0;
    float tbest = 0.0f;
    for (int i = 0; i < GTO_SCHEDULE_CANDIDATES; ++i) {
        apply(Engine::candidate(i));
        const float t = time(iters);
        if (ms) ms[i] = t;
        if (i == 0 || t < tbest) {
            tbest = t;
            best = i;
        }
    }
    apply(Engine::candidate(best));
    return Engine::candidate(best);
}

extern "C" {

int gto_hip_create(const gto_spot_view* spot, const gto_hip_options* opts, gto_hip** out) {
    if (!spot || !out) {
        set_last_error("gto_hip_create: null argument");
        return GTO_ERR_INVALID;
    }
    *out = nullptr;
    gto_hip_options o = {GTO_STORAGE_F32, GTO_ALGO_DCFR, 1, 0, 1};
    if (opts) o = *opts;
    if (o.algo != GTO_ALGO_DCFR && o.algo != GTO_ALGO_CFRPLUS && o.algo != GTO_ALGO_PCFRPLUS) {
        set_last_error("unknown algorithm");
        return GTO_ERR_INVALID;
    }
    return guarded([&] {
        auto h = std::make_unique<gto_hip>();
        h->e = std::make_unique<Engine>(*spot, o);
        *out = h.release();
    });
}

void gto_hip_destroy(gto_hip* h) { delete h; }

int gto_hip_iterate(gto_hip* h, uint32_t n) {
    return guarded([&] { h->e->iterate(n); });
}

int gto_hip_get_iteration(const gto_hip* h, uint32_t* it) {
    *it = h->e->iteration();
    return GTO_OK;
}

int gto_hip_set_iteration(gto_hip* h, uint32_t it) {
    return guarded([&] { h->e->set_iteration(it); });
}

int gto_hip_exploitability(gto_hip* h, double* chips) {
    return guarded([&] { *chips = h->e->exploitability(); });
}

int gto_hip_root_cfv(gto_hip* h, int p, int mode, float* out) {
    if (p < 0 || p > 1 || mode < 0 || mode > 2) {
        set_last_error("gto_hip_root_cfv: bad player or mode");
        return GTO_ERR_INVALID;
    }
    return guarded([&] {
        std::vector<float> c;
        h->e->root_cfv(p, mode, c);
        std::copy(c.begin(), c.end(), out);
    });
}

int gto_batch_create(gto_hip* const* engines, uint32_t n, gto_batch** out) {
    if (!engines || !out || n == 0) {
        set_last_error("gto_batch_create: need at least one engine");
        return GTO_ERR_INVALID;
    }
    *out = nullptr;
    return guarded([&] {
        auto b = std::make_unique<gto_batch>();
        for (uint32_t i = 0; i < n; ++i) {
            if (!engines[i]) throw std::runtime_error("gto_batch_create: null engine");
            for (Engine* e : b->eng)
                if (e == engines[i]->e.get()) throw std::runtime_error("gto_batch_create: an engine is listed twice");
            b->eng.push_back(engines[i]->e.get());
            if (b->eng.back()->device() != b->eng[0]->device())
                throw std::runtime_error("gto_batch_create: engines of one batch must share a device");
        }
        *out = b.release();
    });
}

int gto_batch_iterate(gto_batch* b, uint32_t n) {
    if (!b) {
        set_last_error("gto_batch_iterate: null batch");
        return GTO_ERR_INVALID;
    }
    return guarded([&] { b->iterate(n); });
}

void gto_batch_destroy(gto_batch* b) { delete b; }

int gto_hip_solved_strategy(gto_hip* h, uint32_t node, float* out) {
    if (!h || !out) {
        set_last_error("gto_hip_solved_strategy: null argument");
        return GTO_ERR_INVALID;
    }
    return guarded([&] { h->e->solved_strategy(node, out); });
}

int gto_hip_node_values(gto_hip* h, int p, int mode, uint32_t count, const uint32_t* nodes, float* out) {
    if (!h || p < 0 || p > 1 || mode < 0 || mode > 2 || (count && (!nodes || !out))) {
        set_last_error("gto_hip_node_values: bad argument");
        return GTO_ERR_INVALID;
    }
    return guarded([&] { h->e->node_values(p, mode, count, nodes, out); });
}

int gto_hip_set_showdown(gto_hip* h, int mode) {
    if (!h) {
        set_last_error("gto_hip_set_showdown: null handle");
        return GTO_ERR_INVALID;
    }
    return guarded([&] { h->e->set_showdown(mode); });
}

int gto_hip_set_schedule(gto_hip* h, const gto_schedule* sc) {
    if (!h || !sc) {
        set_last_error("gto_hip_set_schedule: null argument");
        return GTO_ERR_INVALID;
    }
    return guarded([&] { h->e->set_schedule(*sc); });
}

int gto_hip_get_schedule(const gto_hip* h, gto_schedule* sc) {
    if (!h || !sc) {
        set_last_error("gto_hip_get_schedule: null argument");
        return GTO_ERR_INVALID;
    }
    *sc = h->e->schedule();
    return GTO_OK;
}

int gto_hip_autotune(gto_hip* h, uint32_t iters, gto_schedule* best, float* ms) {
    if (!h || iters == 0) {
        set_last_error("gto_hip_autotune: null handle or zero iterations");
        return GTO_ERR_INVALID;
    }
    return guarded([&] {
        Engine* e = h->e.get();
        const gto_schedule b = autotune(
            iters, [&](const gto_schedule& sc) { e->set_schedule(sc); },
            [&](uint32_t n) { return e->time_iterations(n); }, ms);
        if (best) *best = b;
    });
}

int gto_batch_autotune(gto_batch* b, uint32_t iters, gto_schedule* best, float* ms) {
    if (!b || iters == 0) {
        set_last_error("gto_batch_autotune: null batch or zero iterations");
        return GTO_ERR_INVALID;
    }
    return guarded([&] {
        const gto_schedule r = autotune(
            iters,
            [&](const gto_schedule& sc) {
                for (Engine* e : b->eng) e->set_schedule(sc);
            },
            [&](uint32_t n) { return b->time_iterations(n); }, ms);
        if (best) *best = r;
    });
}

int gto_hip_set_weights(gto_hip* h, const float* w0, const float* w1) {
    if (!h) {
        set_last_error("gto_hip_set_weights: null handle");
        return GTO_ERR_INVALID;
    }
    return guarded([&] {
        h->e->set_weights(0, w0);
        h->e->set_weights(1, w1);
    });
}

int gto_hip_game_value(gto_hip* h, int player, double* value, float* grad_w0, float* grad_w1, float* grad_s0,
                       float* grad_s1) {
    if (!h || player < 0 || player > 1) {
        set_last_error("gto_hip_game_value: null handle or bad player");
        return GTO_ERR_INVALID;
    }
    return guarded([&] {
        float* gw[2] = {grad_w0, grad_w1};
        float* gs[2] = {grad_s0, grad_s1};
        h->e->game_value(player, value, gw, gs);
    });
}

int gto_hip_game_value_noalloc(gto_hip* h, int player, double* value, float* grad_w0, float* grad_w1,
                               float* grad_s0, float* grad_s1) {
    if (!h || player < 0 || player > 1) {
        set_last_error("gto_hip_game_value_noalloc: null handle or bad player");
        return GTO_ERR_INVALID;
    }
    return guarded([&] {
        float* gw[2] = {grad_w0, grad_w1};
        float* gs[2] = {grad_s0, grad_s1};
        h->e->game_value(player, value, gw, gs, true);
    });
}

int gto_hip_game_value_device(gto_hip* h, int player, double* value, float* grad_w0, float* grad_w1,
                              float* grad_s0, float* grad_s1) {
    if (!h || player < 0 || player > 1) {
        set_last_error("gto_hip_game_value_device: null handle or bad player");
        return GTO_ERR_INVALID;
    }
    return guarded([&] {
        float* gw[2] = {grad_w0, grad_w1};
        float* gs[2] = {grad_s0, grad_s1};
        h->e->game_value(player, value, gw, gs, false, true);
    });
}

int gto_hip_upload_strategy_device(gto_hip* h, const float* s0, const float* s1) {
    if (!h) {
        set_last_error("gto_hip_upload_strategy_device: null handle");
        return GTO_ERR_INVALID;
    }
    return guarded([&] { h->e->upload_strategy_device(s0, s1); });
}

int gto_hip_download_arenas(gto_hip* h, float* r0, float* r1, float* s0, float* s1) {
    return guarded([&] {
        h->e->download(0, 0, r0);
        h->e->download(0, 1, r1);
        h->e->download(1, 0, s0);
        h->e->download(1, 1, s1);
    });
}

int gto_hip_upload_arenas(gto_hip* h, const float* r0, const float* r1, const float* s0, const float* s1) {
    return guarded([&] {
        h->e->upload(0, 0, r0);
        h->e->upload(0, 1, r1);
        h->e->upload(1, 0, s0);
        h->e->upload(1, 1, s1);
    });
}

int gto_hip_download_compressed(gto_hip* h, int p, int16_t* rq, uint16_t* sq, float* rs, float* ss) {
    if (p < 0 || p > 1) {
        set_last_error("gto_hip_download_compressed: bad player");
        return GTO_ERR_INVALID;
    }
    return guarded([&] { h->e->download_raw(p, rq, sq, rs, ss); });
}

int gto_hip_update_locks(gto_hip* h, const int64_t* lock_off, const float* sigma, uint64_t n_sigma) {
    return guarded([&] { h->e->update_locks(lock_off, sigma, n_sigma); });
}

int gto_hip_synchronize(gto_hip* h) {
    return guarded([&] { h->e->synchronize(); });
}

const char* gto_hip_profile_kind_name(int k) { return (k >= 0 && k < GTO_PROF_KINDS) ? kProfNames[k] : "?"; }

int gto_hip_profile_iterations(gto_hip* h, uint32_t n, gto_hip_profile* out) {
    return guarded([&] { h->e->profile(n, out); });
}

int gto_hip_run(gto_hip* h, const gto_run_options* opts, const volatile int* stop, gto_progress_fn on_progress,
                void* user, gto_progress* out) {
    if (!h || !opts) {
        set_last_error("gto_hip_run: null handle or options");
        return GTO_ERR_INVALID;
    }
    return guarded([&] {
        const gto_progress p = h->e->run(*opts, stop, on_progress, user);
        if (out) *out = p;
    });
}

uint64_t gto_hip_reached_entries(const gto_hip* h) { return h->e->reached_entries(); }

uint64_t gto_hip_river_entries(const gto_hip* h) { return h->e->river_entries(); }

uint64_t gto_hip_estimate_vram(const gto_spot_view* spot, int storage) { return spot ? Engine::estimate(*spot, storage) : 0; }

int gto_hip_vram_info(int device, uint64_t* free_mb, uint64_t* total_mb) {
    return guarded([&] {
        HIP_OK(hipSetDevice(device));
        size_t fr = 0, tot = 0;
        HIP_OK(hipMemGetInfo(&fr, &tot));
        *free_mb = fr >> 20;
        *total_mb = tot >> 20;
    });
}

}  // extern "C"
