/*
 * gtopen_amd — C ABI of the MI355X DCFR engine for heads-up postflop spots.
 *
 * Two halves, one shared library (libgtopen_amd.so):
 *
 *   gto_spot_*  host-side spot construction (board, ranges, betting tree,
 *               river strength tables). Replaces `Spot::new*`
 *               (crates/solver/src/game.rs:103-133) and exposes the built
 *               spot as plain arrays through `gto_spot_view`, the same data a
 *               foreign caller (the reference's Rust `Spot`) would flatten.
 *
 *   gto_hip_*   the GPU engine. Replaces `solver::gpu::GpuSolver`
 *               (crates/solver/src/gpu/mod.rs:60-997): create from a spot,
 *               iterate, exploitability, arena download/upload, locks,
 *               synchronize, VRAM estimate/info, profile, game value.
 *
 * Conventions: every function returning `int` returns 0 on success and a
 * negative code on failure; `gto_last_error()` then holds a thread-local
 * message (the reference returns `Err(String)`, gpu/mod.rs:56-58). All
 * arrays are caller-owned host memory; nothing in this header uses C++ or
 * torch types. One engine handle must be used by one thread at a time
 * (the reference's `&mut self` rule).
 */
#ifndef GTOPEN_AMD_H
#define GTOPEN_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GTO_OK 0
#define GTO_ERR_INVALID (-1)
#define GTO_ERR_DEVICE (-2)
/* host or device memory exhausted (hipErrorOutOfMemory): a caller may fall
 * back to the CPU engine, as the reference's VRAM gate does (main.rs:654-662) */
#define GTO_ERR_MEMORY (-3)
#define GTO_ERR_IO (-4)

/* Node kinds (crates/solver/src/tree.rs:172-175). */
#define GTO_KIND_ACTION 0
#define GTO_KIND_CHANCE 1
#define GTO_KIND_FOLD 2
#define GTO_KIND_SHOWDOWN 3
#define GTO_SENTINEL 0xFFFFFFFFu

/* Arena storage (crates/solver/src/store.rs:69-76). */
#define GTO_STORAGE_F32 0
#define GTO_STORAGE_COMPRESSED 1 /* i16 regrets, u16 strategy sums, f32 scale per node */
/* f16 regrets and strategy sums times a per-node power-of-two scale: the same
 * 2 bytes per entry as COMPRESSED, 11-bit relative precision at every
 * magnitude. No counterpart in the reference (an engine-side format; saves
 * and downloads are f32 as always). */
#define GTO_STORAGE_HALF 2
/* GTO_STORAGE_HALF arenas, and the DCFR iterations' reach and value rows in
 * f16 too (each row with a power-of-two scale bounding its magnitude): half
 * the row traffic of the level sweeps. Best response, exploitability, node
 * values and the game value keep f32 rows; the matrix-core showdown forms
 * (gto_hip_set_showdown) are not available with it (GTO_ERR_INVALID). No
 * counterpart in the reference. */
#define GTO_STORAGE_HALF_ROWS 3

/* Algorithms (crates/solver/src/cfr.rs:13-35). PCFR+ (predictive CFR+:
 * CFR+ discounts, the current strategy matched on max(R,0) plus the last
 * instantaneous regret) runs on the GPU with f32 arenas; the reference runs
 * it on the CPU only (gpu/mod.rs:147-149). */
#define GTO_ALGO_DCFR 0
#define GTO_ALGO_CFRPLUS 1
#define GTO_ALGO_PCFRPLUS 2

/* Tree build strictness (crates/solver/src/tree.rs:213-243). */
#define GTO_STRICT 0
#define GTO_LENIENT_LOAD 1

/* Thread-local message of the last failed call on this thread ("" if none). */
const char* gto_last_error(void);

/* ------------------------------------------------------------------------ */
/* Spot: host-side construction                                              */
/* ------------------------------------------------------------------------ */

typedef struct gto_spot gto_spot;

/*
 * Flattened spot. Field meanings follow the reference structs:
 *   nodes   -> tree.rs:178-198 (`Node`), children -> tree.rs:201-211
 *   hands   -> game.rs:13-19 (`HandInfo`, combo_index order, game.rs:174-202)
 *   same_combo -> game.rs:93-95; river tables -> game.rs:29-61
 * River table keys follow `RiverTable::key` (game.rs:43-56): root street 2 has
 * one key (0); root street 1 is keyed by the river card (52 keys); root street
 * 0 by lo*52+hi of the two dealt cards (2704 keys). river_off[p][key] is -1
 * for keys without an entry; otherwise the entry occupies
 * [river_off, river_off + river_cnt) of river_str/river_idx, sorted ascending
 * by (strength, hand index).
 */
typedef struct gto_spot_view {
    uint32_t num_nodes;
    const uint8_t* node_kind;
    const uint8_t* node_player;   /* action: actor; fold: the player who folded */
    const uint8_t* node_street;   /* chance: street dealt into */
    const uint8_t* node_num_children;
    const uint32_t* node_children_start;
    const uint64_t* node_data_offset; /* action nodes: element offset into the actor's arenas */
    const double* node_t_win;
    const double* node_t_lose;
    const double* node_t_tie;
    const double* node_put;       /* 2 per node: chips put in by OOP, IP */
    uint32_t num_children_total;
    const uint32_t* children;     /* chance nodes own 52 slots (card -> child or SENTINEL) */
    uint64_t data_size[2];
    uint32_t root_street;
    double starting_pot;
    double effective_stack;
    double rake_pct;
    double rake_cap;
    uint32_t board_len;
    uint8_t board[5];
    uint64_t board_mask;
    uint32_t num_hands[2];
    const uint8_t* hand_c1[2];    /* higher card of the combo */
    const uint8_t* hand_c2[2];
    const float* weights[2];
    const uint32_t* same_combo[2];
    uint32_t num_river_keys;
    const int64_t* river_off[2];
    const uint32_t* river_cnt[2];
    const uint32_t* river_str[2];
    const uint16_t* river_idx[2];
    /* suit symmetries fixing the board pointwise with both ranges invariant
     * (game.rs:262-317), identity first: suit s -> suit_perms[4k + s] */
    uint32_t num_suit_perms;
    const uint8_t* suit_perms;
    const uint16_t* hand_perm[2]; /* [k * nh_q + i]: index of hand i's image under symmetry k */
} gto_spot_view;

/* Build a spot from the reference's SpotConfig JSON (game.rs:21-27, serde
 * field names and enum tags: {"PotPct":33}, {"PrevMult":2.5}, "AllIn").
 * `max_nodes` 0 = no node budget (game.rs:108-110). Replaces Spot::new,
 * Spot::new_with_limit, Spot::new_lenient[_with_limit] (game.rs:103-133). */
int gto_spot_build(const char* config_json, int strictness, uint64_t max_nodes, gto_spot** out);
void gto_spot_free(gto_spot* spot);
/* Stable view into the spot's arrays; valid until gto_spot_free. */
const gto_spot_view* gto_spot_view_of(const gto_spot* spot);
/* Number of action nodes (game.rs:375-381). */
uint64_t gto_spot_num_action_nodes(const gto_spot* spot);
/* Arena bytes under a storage mode (game.rs:341-358; the f16 modes, which
 * the reference lacks, count 2 bytes per entry plus per-node scales). */
uint64_t gto_spot_arena_bytes(const gto_spot* spot, int storage);
/* Suit-symmetry group size (game.rs:262-317); 1 = only the identity. */
uint32_t gto_spot_num_suit_perms(const gto_spot* spot);

/* Card helpers (crates/solver/src/cards.rs:100-114, evaluator.rs:68-175). */
uint32_t gto_combo_index(uint8_t c1, uint8_t c2);
uint32_t gto_evaluate7(const uint8_t* cards7);
/* Parse a range text into 1326 weights (range.rs:28-58). */
int gto_range_parse(const char* text, float* weights1326);
/* Canonical flop classes (cards.rs:236-311): writes up to `cap` rows of
 * 3 cards + weight; returns the row count through *n. n_subset 0 = all. */
int gto_canonical_flops(uint32_t n_subset, uint8_t* cards3, uint32_t* weights, uint32_t cap,
                        uint32_t* n);

/* ------------------------------------------------------------------------ */
/* GTOSOLVE2 save files (crates/solver/src/save.rs)                          */
/* ------------------------------------------------------------------------ */
/* "GTOSOLVE2\n", one JSON header line {"config": SpotConfig, "iteration",
 * "locks": [[node, [f32..]]..], "lock_labels": [[node, "label"]..]}, then
 * regrets OOP, IP, strategy sums OOP, IP as (u64 LE length, f32 LE data).
 * Always f32 on disk (compressed engines decode on download). Locks are
 * given as n_locks entries: node, offset and length into lock_sigma. */
typedef struct gto_save gto_save;

/* Solver::save / write_save (save.rs:94-143): staged to "<path>.tmp",
 * fsync'd, renamed over path. Arenas have data_size[0]/[1] entries. Locks
 * are validated first (save.rs:42-70). */
int gto_save_write(const char* path, const gto_spot* spot, uint32_t iteration, const float* regrets0,
                   const float* regrets1, const float* strat0, const float* strat1, uint32_t n_locks,
                   const uint32_t* lock_nodes, const uint64_t* lock_off, const uint64_t* lock_len,
                   const float* lock_sigma, uint32_t n_labels, const uint32_t* label_nodes,
                   const char* const* labels);
/* First half of Solver::load_with_storage (save.rs:149-168): checks the
 * magic ("not a solver save file") and parses the header ("bad header: ..").
 * The caller rebuilds the spot leniently from gto_save_config_json (every
 * SpotConfig field, serde layout), then validates locks and reads arenas. */
int gto_save_open(const char* path, gto_save** out);
void gto_save_close(gto_save* s);
const char* gto_save_config_json(const gto_save* s);
uint32_t gto_save_iteration(const gto_save* s);
uint32_t gto_save_num_locks(const gto_save* s);
int gto_save_lock(const gto_save* s, uint32_t k, uint32_t* node, const float** sigma, uint64_t* len);
uint32_t gto_save_num_labels(const gto_save* s);
int gto_save_label(const gto_save* s, uint32_t k, uint32_t* node, const char** label);
/* validate_locks on the header's locks against the rebuilt spot, errors
 * prefixed "bad lock section: " (save.rs:168-170). */
int gto_save_validate_locks(const gto_save* s, const gto_spot* spot);
/* The four arenas, sized against the rebuilt tree ("arena size mismatch:
 * file N, expected M (tree config changed?)", save.rs:177-192). NULL skips
 * an arena. */
int gto_save_read_arenas(const gto_save* s, const gto_spot* spot, float* regrets0, float* regrets1,
                         float* strat0, float* strat1);
/* validate_locks (save.rs:42-70) for a lock list. */
int gto_validate_locks(const gto_spot* spot, uint32_t n_locks, const uint32_t* lock_nodes,
                       const uint64_t* lock_off, const uint64_t* lock_len, const float* lock_sigma);

/* ------------------------------------------------------------------------ */
/* Engine: the GPU DCFR hot path                                             */
/* ------------------------------------------------------------------------ */

typedef struct gto_hip gto_hip;

typedef struct gto_hip_options {
    int storage;        /* GTO_STORAGE_F32, _COMPRESSED, _HALF or _HALF_ROWS (device arenas and rows) */
    int algo;           /* GTO_ALGO_DCFR, GTO_ALGO_CFRPLUS or GTO_ALGO_PCFRPLUS (f32 arenas) */
    int use_graphs;     /* capture each sweep into a hipGraph after the first iteration */
    int device;         /* HIP device ordinal */
    int isomorphism;    /* 1: one chance branch per suit orbit, as the reference's default
                         * use_isomorphism (cfr.rs:84-87, 143); 0: every branch */
} gto_hip_options;

/* Create an engine over `spot` (copied to the device; the view may be freed
 * afterwards). Arenas start at zero and iteration at 0, as a fresh
 * `Solver::with_storage` (cfr.rs:122-146). Replaces GpuSolver::new
 * (gpu/mod.rs:146-294). */
int gto_hip_create(const gto_spot_view* spot, const gto_hip_options* opts, gto_hip** out);
void gto_hip_destroy(gto_hip* h);

/* Run `n` DCFR iterations (alternating updates, traverser 0 then 1), as
 * GpuSolver::iterate (gpu/mod.rs:304-346) / Solver::iterate (cfr.rs:170-190). */
int gto_hip_iterate(gto_hip* h, uint32_t n);
int gto_hip_get_iteration(const gto_hip* h, uint32_t* iteration);
int gto_hip_set_iteration(gto_hip* h, uint32_t iteration);

/* Exploitability of the average strategy in chips (best-response sweeps on
 * the device, f64 root dot), as GpuSolver::exploitability (gpu/mod.rs:367-396)
 * and Solver::exploitability (best_response.rs:255-287). */
int gto_hip_exploitability(gto_hip* h, double* chips);

/* Root counterfactual values of player p (nh_p floats) under: mode 0 = p
 * best-responds, mode 1 = both play the average strategy, mode 2 = both play
 * the current (regret-matched) strategy. Mirrors traverse_br/traverse_avg
 * (best_response.rs:12-223). */
int gto_hip_root_cfv(gto_hip* h, int p, int mode, float* out);

/* Arena transfer as f32 (compressed device arenas are decoded/encoded with
 * the per-node scales). Layout = tree data_offset order (tree.rs:546-548);
 * lengths data_size[0]/[1]. Replaces sync_to_cpu / sync_strategy and the
 * upload in GpuSolver::new (gpu/mod.rs:188-199, 845-886, 976-997). Any
 * pointer may be NULL to skip that arena. */
int gto_hip_download_arenas(gto_hip* h, float* regrets0, float* regrets1, float* strat0,
                            float* strat1);
int gto_hip_upload_arenas(gto_hip* h, const float* regrets0, const float* regrets1,
                          const float* strat0, const float* strat1);
/* Raw compressed arenas (i16 regrets, u16 strategy sums, per-node f32
 * scales indexed by node) when storage is COMPRESSED; store.rs:79-92. */
int gto_hip_download_compressed(gto_hip* h, int p, int16_t* regrets_q, uint16_t* strat_q,
                                float* regret_scale, float* strat_scale);

/* Batches: several engines (independent spots on one device, e.g. the
 * flops of a report, main.rs:1583-1772) advanced together. A batch
 * iteration launches every engine's captured iteration graph on the
 * engine's own stream, engines interleaved, so the spots' kernels share the
 * device through its hardware queues. Engines stay owned by the caller and
 * must outlive the batch; between gto_batch_iterate calls each engine may
 * be used on its own (exploitability, downloads, locks: an engine
 * re-captures its graph when they change it). No reference counterpart:
 * the reference solves report flops one by one. */
typedef struct gto_batch gto_batch;
int gto_batch_create(gto_hip* const* engines, uint32_t n, gto_batch** out);
/* n iterations of every engine (an engine not in the steady state first
 * takes one iteration on its own stream); returns when all are done. */
int gto_batch_iterate(gto_batch* b, uint32_t n);
void gto_batch_destroy(gto_batch* b);

/* Average strategy of one action node from the device strategy arena,
 * ignoring locks: solved_strategy_into (cfr.rs:679-726), na*nh_actor floats
 * in arena layout. The query layer's walk_path / lock_node base
 * (query.rs:304-350, 645-734). */
int gto_hip_solved_strategy(gto_hip* h, uint32_t node, float* out);

/* Per-node value rows of player p (nh_p floats per node) after one sweep:
 * mode 0 = best response, mode 1 = average strategy; the row of node n is
 * traverse_br / traverse_avg (best_response.rs:19-223) at n given the
 * opponent's reach there. Feeds exploit_view (query.rs:800-961). Mode 2 =
 * both players' current (regret-matched) strategies, the values a DCFR
 * update sees, with the showdown form set by gto_hip_set_showdown (modes 0
 * and 1 always use the exact showdown). Nodes outside the reachable tree
 * are an error. */
int gto_hip_node_values(gto_hip* h, int p, int mode, uint32_t count, const uint32_t* nodes, float* out);

/* Showdown evaluation. EXACT (default): strength-order prefix sums in f64,
 * the CPU engine's arithmetic (game.rs:501-577) up to the association of
 * f64 sums. SEQUENTIAL: EXACT, with the f64 sums of the fold and showdown
 * terminals (game.rs:470-577) taken in the CPU engine's order wherever the
 * order could change a bit: the fold totals always, the showdown sums of
 * any reach row whose nonzero entries span more than 17 binary exponents
 * (below that every f64 sum of the row is exact in any order). Iterations,
 * values and exploitability then equal the CPU engine's bit for bit; f32
 * rows only (not with GTO_STORAGE_HALF_ROWS). BF16 / FP8: the showdown nodes of each river board stacked into
 * matrix products on the matrix cores, value = win*L + lose*H + tie*Q with
 * [L;H;Q] = [W;G;T] x R (0/1 matrices of strictly weaker, strictly stronger
 * and tied card-disjoint opponent hands), R = opponent reach in bf16, or in
 * e4m3 with a power-of-two scale per node, f32 accumulation (precision
 * forms; no counterpart in the reference). */
#define GTO_SHOWDOWN_EXACT 0
#define GTO_SHOWDOWN_BF16 1
#define GTO_SHOWDOWN_FP8 2
#define GTO_SHOWDOWN_SEQUENTIAL 3
int gto_hip_set_showdown(gto_hip* h, int mode);

/* Launch schedule of the up sweep. Every schedule computes the same values
 * bit for bit (terminal values depend only on the down pass; a showdown
 * run's length changes no node's arithmetic), so a schedule is a pure
 * performance choice. show_nodes: showdown nodes of one river board per
 * workgroup of the exact showdown kernel (16, 8 or 4). per_level: 0 = the
 * terminals of every level in one launch each per sweep, before the level
 * loop (the default); 1 = each level's terminals launched with that level,
 * just before the chance and action kernels that read their values. No
 * counterpart in the reference (its kernels have one fixed schedule). */
typedef struct gto_schedule {
    uint32_t show_nodes;
    uint32_t per_level;
    /* 1 = fused river sweeps: one workgroup per river subtree
     * runs the subtree's reach, terminals and updates with its rows in LDS,
     * so river-level reach and value rows stay off HBM; the level sweeps
     * cover the tree above. Applies to DCFR / CFR+ iterations with f32 or
     * f16 arenas, the exact showdown and no locked node; otherwise (and
     * with 0, the default) every level is swept level-synchronously.
     * Results are bit for bit those of the level sweeps. */
    uint32_t river;
    /* 1 (the default) = absorbed opponent nodes in the DCFR sweeps of f32
     * engines without locks: an opponent action node with at most 3
     * actions under a traverser node with at most 4 writes no value row;
     * its parent adds up the node's children's rows itself, in the order
     * the node would. 0 = every node writes its row. Bit-identical either
     * way; other engines and sweeps ignore it. */
    uint32_t absorb;
} gto_schedule;
int gto_hip_set_schedule(gto_hip* h, const gto_schedule* s);
int gto_hip_get_schedule(const gto_hip* h, gto_schedule* s);

/* Autotuner: for each of the GTO_SCHEDULE_CANDIDATES schedules (candidate
 * i < 6: level sweeps with show_nodes = {16, 8, 4}[i % 3], per_level =
 * i / 3; candidate 6: the fused river sweep), one warm-up
 * iteration (which captures the schedule's graph) and `iters` timed DCFR
 * iterations; the fastest schedule is then applied. The iterations are
 * ordinary iterations of the solve: it advances by
 * GTO_SCHEDULE_CANDIDATES * (iters + 1). ms (may be NULL): per-candidate
 * milliseconds per iteration (GPU time for an engine, wall time for a
 * batch); best (may be NULL): the chosen schedule. The batch form applies
 * each candidate to every engine of the batch and times the batch's
 * iterations. */
#define GTO_SCHEDULE_CANDIDATES 7
int gto_hip_autotune(gto_hip* h, uint32_t iters, gto_schedule* best, float* ms);
int gto_batch_autotune(gto_batch* b, uint32_t iters, gto_schedule* best, float* ms);

/* Range weights per hand (the root reach), replacing the spot's; NULL keeps
 * a player's weights. The hand lists stay as built: a weight of 0 leaves the
 * hand in place with zero reach (the reference's builder would drop it,
 * game.rs:177-180, which changes no value). Weights must be finite, >= 0. */
int gto_hip_set_weights(gto_hip* h, const float* w0, const float* w1);

/* Differentiable game value (the forward pass is traverse_avg,
 * best_response.rs:128-223, with the strategy given rather than averaged):
 * the strategy arenas (upload them with gto_hip_upload_arenas(h, NULL, NULL,
 * s0, s1)) are read as the behavioural strategy sigma itself, without
 * normalisation, and
 *   value = U_player = sum_i w_player[i] * cfv_player(root)[i]
 * (best_response.rs:264-269; chips, not divided by the pair weight sum).
 * Gradients (any pointer may be NULL to skip it):
 *   grad_w0/grad_w1 [num_hands[q]]: dU/dw_q; for q = player this is the
 *     root counterfactual value cfv_player(root), for the opponent it is the
 *     root value of the opponent's hands under the player's payoffs;
 *   grad_s0/grad_s1 [data_size[q]] (arena layout): dU/dsigma_q(node,a,i) =
 *     reach_q(node,i) * chance weight of the path * value of child a; 0 at
 *     locked nodes and nodes outside the reachable tree.
 * The reference has no backward; its CFR values are the reference gradient.
 * F32 storage only, engine created with isomorphism 0 (every node's strategy
 * is an input). */
int gto_hip_game_value(gto_hip* h, int player, double* value, float* grad_w0, float* grad_w1,
                       float* grad_s0, float* grad_s1);
/* The same value and gradients with no allocation in the call (no
 * counterpart in the reference): for callers that cannot allocate device
 * memory — code under graph capture, tight serving loops over range
 * perturbations, harnesses that forbid hipMalloc. gto_hip_game_value
 * allocates its strategy-gradient staging buffer on first use; this entry
 * stages it in a static device array of the library instead, 16777216
 * floats (64 MiB) per GPU, held until the process ends. A strategy gradient
 * requested for a player whose arena is larger is rejected with
 * GTO_ERR_DEVICE and a message. Every engine on one GPU shares that array:
 * calls that overlap on one GPU (engines used from several threads) race;
 * consecutive calls are independent (the array is cleared where needed). */
int gto_hip_game_value_noalloc(gto_hip* h, int player, double* value, float* grad_w0, float* grad_w1,
                               float* grad_s0, float* grad_s1);
/* The same value and gradients for callers that keep ranges, strategies
 * and gradients on the GPU, such as an autograd graph of device tensors (no
 * counterpart in the reference): grad_w0/grad_w1/grad_s0/grad_s1 are device
 * memory of the engine's GPU (any may be NULL); the root rows are copied
 * there and the strategy gradient is written in place, so only the value
 * crosses to the host. A pointer that is not device memory of that GPU is
 * GTO_ERR_INVALID. Returns when the gradients are written; the caller's
 * earlier writes to device memory must be complete. */
int gto_hip_game_value_device(gto_hip* h, int player, double* value, float* grad_w0, float* grad_w1,
                              float* grad_s0, float* grad_s1);
/* The strategy arenas (the game value's sigma) from device memory of the
 * engine's GPU, as gto_hip_upload_arenas(h, NULL, NULL, s0, s1) does from
 * host memory (f32 arenas; NULL keeps a player's; no counterpart in the
 * reference). Not device memory of that GPU: GTO_ERR_INVALID. */
int gto_hip_upload_strategy_device(gto_hip* h, const float* s0, const float* s1);

/* Node locks: lock_off[num_nodes] (-1 = unlocked, else offset into sigma),
 * sigma = concatenated na*nh_actor blocks. Replaces update_locks /
 * build_lock_table (gpu/mod.rs:894-905, 961-973). n_sigma may be 0. */
int gto_hip_update_locks(gto_hip* h, const int64_t* lock_off, const float* sigma,
                         uint64_t n_sigma);

int gto_hip_synchronize(gto_hip* h);

/* Per-kernel-kind device time over the last `gto_hip_profile_iterations`
 * call, in ms (HIP events around each launch group). Replaces
 * iterate_profiled / Profile (gpu/mod.rs:22-54, 351-362). */
#define GTO_PROF_KINDS 9
typedef struct gto_hip_profile {
    double ms[GTO_PROF_KINDS];
    uint32_t launches[GTO_PROF_KINDS];
} gto_hip_profile;
const char* gto_hip_profile_kind_name(int k);
int gto_hip_profile_iterations(gto_hip* h, uint32_t n, gto_hip_profile* out);

/* Run to a target: Solver::run (crates/solver/src/cfr.rs:194-226) with
 * RunOptions / Progress (cfr.rs:91-115). Iterates while iteration <
 * max_iterations and *stop (may be NULL) is 0; after every iteration that
 * is a multiple of check_every (or the last one) it measures exploitability,
 * fills a progress record and calls on_progress (may be NULL); a callback
 * returning 0 stops the run; the run also stops once exploit_pct_pot <=
 * target_exploit_pct. *out (may be NULL) receives the last check's record,
 * as Solver::run returns it (before any check: the starting iteration, NaN
 * exploitability, 0 s).
 * elapsed_secs is wall time from the call; gpu_secs the part spent in
 * iterations (launch to completion), as solve-cli's progress line reports
 * (solve_cli.rs:416-450). */
typedef struct gto_run_options {
    uint32_t max_iterations;
    double target_exploit_pct; /* percent of the starting pot, e.g. 0.3 */
    uint32_t check_every;      /* 0 is taken as 1 */
} gto_run_options;
typedef struct gto_progress {
    uint32_t iteration;
    double exploit_chips;
    double exploit_pct_pot;
    double elapsed_secs;
    double gpu_secs;
} gto_progress;
typedef int (*gto_progress_fn)(const gto_progress* progress, void* user);
int gto_hip_run(gto_hip* h, const gto_run_options* opts, const volatile int* stop, gto_progress_fn on_progress,
                void* user, gto_progress* out);

/* Algorithmic arena entries D = sum over reached action nodes of
 * na * nh_actor, the unit of the roofline bytes (20 B per entry per f32
 * iteration, 10 B compressed). */
uint64_t gto_hip_reached_entries(const gto_hip* h);
/* The part of D inside river subtrees, which the fused river kernel streams
 * (0 when the engine cannot run fused river sweeps). */
uint64_t gto_hip_river_entries(const gto_hip* h);

/* VRAM the engine would allocate for `spot` (gpu/mod.rs:946; game.rs:365-373). */
uint64_t gto_hip_estimate_vram(const gto_spot_view* spot, int storage);
/* Free/total device memory in MB (gpu/mod.rs:953-958). */
int gto_hip_vram_info(int device, uint64_t* free_mb, uint64_t* total_mb);

#ifdef __cplusplus
}
#endif
#endif /* GTOPEN_AMD_H */
