/*
 * CPU oracle for the DCFR hot path. TEST INFRASTRUCTURE ONLY: loaded by
 * tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg as
 * the checker and the timed CPU baseline; never by the product library.
 *
 * It restates the reference CPU engine (crates/solver/src/cfr.rs,
 * best_response.rs, the terminal math of game.rs:470-653 and the
 * quantizers of store.rs:153-215) over a spot given as a gto_spot_view.
 * Precision 32 follows the reference's f32 arithmetic operation for
 * operation (f64 running sums where the reference uses them); precision 64
 * runs every vector in f64 and is the FP64 oracle. Suit isomorphism
 * (cfr.rs:568-660, 749-870) is on by default, as in the reference
 * (cfr.rs:143): chance nodes traverse one card per orbit, and arenas are
 * symmetrized before they are read.
 */
#ifndef GTO_ORACLE_H
#define GTO_ORACLE_H

#include <stdint.h>

#include "../include/gtopen_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle oracle;

/* The view's arrays must stay alive for the oracle's lifetime. */
oracle* oracle_create(const gto_spot_view* spot, int precision, int storage, int algo);
void oracle_destroy(oracle* o);
void oracle_set_threads(int n);
int oracle_get_threads(void);
void oracle_iterate(oracle* o, uint32_t n);
uint32_t oracle_iteration(const oracle* o);
void oracle_set_iteration(oracle* o, uint32_t t);
double oracle_exploitability(oracle* o);
double oracle_pair_weight_sum(oracle* o);
/* mode 0: p best-responds vs the average strategy; 1: both average. */
void oracle_root_cfv(oracle* o, int p, int mode, double* out);
void oracle_get_arenas(oracle* o, float* r0, float* r1, float* s0, float* s1);
void oracle_set_arenas(oracle* o, const float* r0, const float* r1, const float* s0, const float* s1);
void oracle_get_arenas_f64(oracle* o, double* r0, double* r1, double* s0, double* s1);
/* -1 = unlocked per node; sigma blocks of na*nh_actor. */
void oracle_set_locks(oracle* o, const int64_t* lock_off, const float* sigma, uint64_t n_sigma);
/* Average strategy at a node (na*nh floats), lock-aware (cfr.rs:668-726). */
void oracle_average_strategy(oracle* o, uint32_t node, float* out);
/* use_isomorphism (cfr.rs:84-87); on by default. */
void oracle_set_isomorphism(oracle* o, int on);

/* Terminal math in isolation (game.rs:470-577), for the unit tests. */
void oracle_fold_cfv(const gto_spot_view* spot, int p, const float* reach_o, float amount, float* out);
void oracle_showdown_cfv(const gto_spot_view* spot, int p, uint32_t river_key, const float* reach_o,
                         float win, float lose, float tie, float* out);
/* Quantizers (store.rs:171-215); return the stored scale. */
float oracle_encode_i16(const float* src, uint64_t n, int16_t* dst);
float oracle_encode_u16(const float* src, uint64_t n, uint16_t* dst);

#ifdef __cplusplus
}
#endif
#endif
