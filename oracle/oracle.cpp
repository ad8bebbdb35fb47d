// CPU oracle: a restatement of the reference DCFR engine (TEST
// INFRASTRUCTURE: the checker and the timed CPU baseline, never part of
// the product library). Every function names the reference lines it
// follows. Built with -ffp-contract=off so that f32 expressions round
// exactly as the reference's (no fused multiply-add).
#include "oracle.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

namespace {

constexpr uint32_t kSent = 0xFFFFFFFFu;

// ---------------------------------------------------------------------------
// Fork-join pool with help-while-waiting (the role rayon's par_iter plays in
// cfr.rs:631-635). Results are merged by the caller in card order, so the
// outcome does not depend on the thread count (cfr.rs:637-663).
// ---------------------------------------------------------------------------
class Pool {
   public:
    explicit Pool(int n) { resize(n); }
    ~Pool() { shutdown(); }
    int threads() const { return nthreads_; }
    void resize(int n) {
        shutdown();
        nthreads_ = n < 1 ? 1 : n;
        stop_ = false;
        for (int i = 1; i < nthreads_; ++i) workers_.emplace_back([this] { worker(); });
    }
    void parallel_for(size_t n, const std::function<void(size_t)>& fn) {
        if (n == 0) return;
        if (nthreads_ == 1 || n == 1) {
            for (size_t i = 0; i < n; ++i) fn(i);
            return;
        }
        Batch b;
        b.fn = &fn;
        b.remaining.store(n);
        {
            std::lock_guard<std::mutex> g(m_);
            for (size_t i = 0; i < n; ++i) q_.push_back({&b, i});
        }
        cv_.notify_all();
        for (;;) {
            if (b.remaining.load() == 0) break;
            Task t;
            bool got = false;
            {
                std::unique_lock<std::mutex> g(m_);
                if (!q_.empty()) {
                    t = q_.back();
                    q_.pop_back();
                    got = true;
                } else {
                    done_cv_.wait(g, [&] { return b.remaining.load() == 0 || !q_.empty(); });
                }
            }
            if (got) run(t);
        }
    }

   private:
    struct Batch {
        const std::function<void(size_t)>* fn;
        std::atomic<size_t> remaining;
    };
    struct Task {
        Batch* b;
        size_t i;
    };
    int nthreads_ = 1;
    bool stop_ = false;
    std::vector<std::thread> workers_;
    std::mutex m_;
    std::condition_variable cv_, done_cv_;
    std::deque<Task> q_;

    void run(const Task& t) {
        (*t.b->fn)(t.i);
        if (t.b->remaining.fetch_sub(1) == 1) {
            std::lock_guard<std::mutex> g(m_);
            done_cv_.notify_all();
        }
    }
    void worker() {
        for (;;) {
            Task t;
            {
                std::unique_lock<std::mutex> g(m_);
                cv_.wait(g, [&] { return stop_ || !q_.empty(); });
                if (stop_ && q_.empty()) return;
                t = q_.front();
                q_.pop_front();
            }
            run(t);
            std::lock_guard<std::mutex> g(m_);
            done_cv_.notify_all();
        }
    }
    void shutdown() {
        {
            std::lock_guard<std::mutex> g(m_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto& w : workers_) w.join();
        workers_.clear();
    }
};

int default_threads() {
    const char* e = std::getenv("ORACLE_THREADS");
    if (e && std::atoi(e) > 0) return std::atoi(e);
    unsigned h = std::thread::hardware_concurrency();
    return h ? (int)h : 1;
}

Pool& pool() {
    static Pool p(default_threads());
    return p;
}

// Thread-local scratch vectors (the role of scratch.rs:19-84).
template <class T>
std::vector<std::vector<T>*>& free_list_of() {
    static thread_local std::vector<std::vector<T>*> fl;
    return fl;
}

template <class T>
struct Buf {
    std::vector<T>* v;
    size_t n;
    Buf(size_t len, bool zero) : n(len) {
        auto& free_list = free_list_of<T>();
        if (free_list.empty()) {
            v = new std::vector<T>();
        } else {
            v = free_list.back();
            free_list.pop_back();
        }
        if (v->size() < len) v->resize(len);
        if (zero) std::fill(v->begin(), v->begin() + len, T(0));
    }
    Buf(size_t len, T fillv) : Buf(len, false) { std::fill(v->begin(), v->begin() + len, fillv); }
    ~Buf() {
        auto& free_list = free_list_of<T>();
        if (free_list.size() < 64) free_list.push_back(v);
        else delete v;
    }
    T* data() { return v->data(); }
    T& operator[](size_t i) { return (*v)[i]; }
    Buf(const Buf&) = delete;
    Buf& operator=(const Buf&) = delete;
};

struct Dealt {  // game.rs:65-80
    uint8_t cards[2] = {0, 0};
    uint8_t len = 0;
    Dealt push(uint8_t c) const {
        Dealt d = *this;
        d.cards[d.len++] = c;
        return d;
    }
    bool contains(uint8_t c) const {
        for (int i = 0; i < len; ++i)
            if (cards[i] == c) return true;
        return false;
    }
};

// permute_card, cards.rs:94-96 (card = 4 * rank + suit)
inline uint8_t permute_card(uint8_t c, const uint8_t* pm) { return (uint8_t)((c & ~3u) | pm[c & 3u]); }

uint32_t river_key(const gto_spot_view* v, const Dealt& d) {  // game.rs:43-56
    if (v->root_street == 2) return 0;
    if (v->root_street == 1) return d.cards[0];
    uint8_t a = std::min(d.cards[0], d.cards[1]), b = std::max(d.cards[0], d.cards[1]);
    return (uint32_t)a * 52 + b;
}

// ---- quantizers (store.rs:153-215) -----------------------------------------
inline int16_t sat_i16(float x) {
    if (std::isnan(x)) return 0;
    if (x >= 32767.0f) return 32767;
    if (x <= -32768.0f) return -32768;
    return (int16_t)x;
}
inline uint16_t sat_u16(float x) {
    if (std::isnan(x)) return 0;
    if (x >= 65535.0f) return 65535;
    if (x <= 0.0f) return 0;
    return (uint16_t)x;
}
float quantize_i16(const float* src, size_t n, float mx, int16_t* dst) {
    if (!(mx > 0.0f && std::isfinite(mx))) {
        std::fill(dst, dst + n, (int16_t)0);
        return 0.0f;
    }
    float k = 32767.0f / mx;
    for (size_t i = 0; i < n; ++i) dst[i] = sat_i16(std::nearbyint(src[i] * k));
    return mx;
}
float quantize_u16(const float* src, size_t n, float mx, uint16_t* dst) {
    if (!(mx > 0.0f && std::isfinite(mx))) {
        std::fill(dst, dst + n, (uint16_t)0);
        return 0.0f;
    }
    float k = 65535.0f / mx;
    for (size_t i = 0; i < n; ++i) dst[i] = sat_u16(std::nearbyint(std::fmax(src[i], 0.0f) * k));
    return mx;
}

// ---- terminal math (game.rs:470-577) ----------------------------------------

// fold_cfv, game.rs:470-496: f64 totals, f32 (T) per-hand value.
template <class T>
void fold_cfv(const gto_spot_view* v, int p, const T* reach_o, T amount, T* out) {
    const int o = 1 - p;
    double t = 0.0, s[52];
    std::fill(s, s + 52, 0.0);
    const uint8_t *oc1 = v->hand_c1[o], *oc2 = v->hand_c2[o];
    for (uint32_t j = 0; j < v->num_hands[o]; ++j) {
        double r = (double)reach_o[j];
        t += r;
        s[oc1[j]] += r;
        s[oc2[j]] += r;
    }
    const uint8_t *pc1 = v->hand_c1[p], *pc2 = v->hand_c2[p];
    const uint32_t* same = v->same_combo[p];
    for (uint32_t i = 0; i < v->num_hands[p]; ++i) {
        T same_r = same[i] != kSent ? reach_o[same[i]] : T(0);
        T valid = (T)(t - s[pc1[i]] - s[pc2[i]]) + same_r;
        out[i] = amount * valid;
    }
}

// showdown_cfv, game.rs:501-577: ascending pass for strictly weaker
// opponent reach, descending pass for strictly stronger, card-removal
// corrections from per-card running sums. `out` must be zero on entry.
template <class T>
void showdown_cfv(const gto_spot_view* v, int p, uint32_t key, const T* reach_o, T win, T lose, T tie, T* out) {
    const int o = 1 - p;
    int64_t moff = v->river_off[p][key], ooff = v->river_off[o][key];
    uint32_t mcnt = v->river_cnt[p][key], ocnt = v->river_cnt[o][key];
    const uint32_t *mstr = v->river_str[p] + moff, *ostr = v->river_str[o] + ooff;
    const uint16_t *midx = v->river_idx[p] + moff, *oidx = v->river_idx[o] + ooff;
    const uint8_t *oc1 = v->hand_c1[o], *oc2 = v->hand_c2[o], *pc1 = v->hand_c1[p], *pc2 = v->hand_c2[p];
    const uint32_t* same = v->same_combo[p];
    double t_all = 0.0, s_all[52];
    std::fill(s_all, s_all + 52, 0.0);
    for (uint32_t k = 0; k < ocnt; ++k) {
        uint32_t j = oidx[k];
        double r = (double)reach_o[j];
        t_all += r;
        s_all[oc1[j]] += r;
        s_all[oc2[j]] += r;
    }
    Buf<T> lower(mcnt, false);
    {
        double t = 0.0, s[52];
        std::fill(s, s + 52, 0.0);
        uint32_t j = 0;
        for (uint32_t k = 0; k < mcnt; ++k) {
            uint32_t stren = mstr[k], i = midx[k];
            while (j < ocnt && ostr[j] < stren) {
                uint32_t jj = oidx[j];
                double r = (double)reach_o[jj];
                t += r;
                s[oc1[jj]] += r;
                s[oc2[jj]] += r;
                ++j;
            }
            lower[k] = (T)(t - s[pc1[i]] - s[pc2[i]]);
        }
    }
    {
        double t = 0.0, s[52];
        std::fill(s, s + 52, 0.0);
        uint32_t j = ocnt;
        for (uint32_t kk = mcnt; kk-- > 0;) {
            uint32_t stren = mstr[kk], i = midx[kk];
            while (j > 0 && ostr[j - 1] > stren) {
                --j;
                uint32_t jj = oidx[j];
                double r = (double)reach_o[jj];
                t += r;
                s[oc1[jj]] += r;
                s[oc2[jj]] += r;
            }
            T higher = (T)(t - s[pc1[i]] - s[pc2[i]]);
            T same_r = same[i] != kSent ? reach_o[same[i]] : T(0);
            T valid = (T)(t_all - s_all[pc1[i]] - s_all[pc2[i]]) + same_r;
            T lo = lower[kk];
            T ti = valid - lo - higher;
            out[i] = win * lo + lose * higher + tie * ti;
        }
    }
}

// ---------------------------------------------------------------------------
// Engine
// ---------------------------------------------------------------------------
template <class T>
struct Discounts {
    T pos, neg, strat;
};

// Discounts::for_iteration, cfr.rs:44-63 (computed in f64; the f32 engine
// stores them as f32).
template <class T>
Discounts<T> discounts(int algo, uint32_t it) {
    double t = (double)it;
    Discounts<T> d;
    if (algo == GTO_ALGO_DCFR) {
        double ta = std::pow(t, 1.5);
        d.pos = (T)(ta / (ta + 1.0));
        d.neg = (T)0.5;
        double q = t / (t + 1.0);
        d.strat = (T)(q * q);
    } else {
        double q = t / (t + 1.0);
        d.pos = (T)1.0;
        d.neg = (T)0.0;
        d.strat = (T)(q * q);
    }
    return d;
}

struct EngineBase {
    virtual ~EngineBase() {}
    virtual void iterate() = 0;
    virtual double exploitability() = 0;
    virtual double pair_weight_sum() = 0;
    virtual void root_cfv(int p, int mode, double* out) = 0;
    virtual void get_arenas(int which, int p, float* out) = 0;
    virtual void get_arenas_f64(int which, int p, double* out) = 0;
    virtual void set_arenas(int which, int p, const float* in) = 0;
    virtual void average_strategy(uint32_t node, float* out) = 0;
    virtual void set_isomorphism(bool on) = 0;
    virtual void mark_sym_dirty() = 0;
    uint32_t iteration = 0;
    std::vector<int64_t> lock_off;
    std::vector<float> lock_sigma;
};

template <class T>
struct Engine : EngineBase {
    const gto_spot_view* v;
    int storage, algo;
    size_t nh[2];
    std::vector<uint64_t> hmask[2];
    std::vector<T> R[2], S[2];
    std::vector<int16_t> Rq[2];
    std::vector<uint16_t> Sq[2];
    std::vector<float> Rs[2], Ss[2];
    // PCFR+ predictions, the last instantaneous regrets (cfr.rs:76-78,
    // 157-168), in the regrets' format
    std::vector<T> P[2];
    std::vector<int16_t> Pq[2];
    std::vector<float> Ps[2];
    bool compressed() const { return storage == GTO_STORAGE_COMPRESSED; }
    bool predictive() const { return algo == GTO_ALGO_PCFRPLUS; }
    // suit isomorphism (cfr.rs:84-87, 143): traverse one chance branch per
    // orbit; the other branches' arenas are stale until symmetrize()
    bool use_iso = true;
    bool sym_dirty = false;
    bool iso() const { return use_iso && v->num_suit_perms > 1; }
    const uint8_t* perm(size_t k) const { return v->suit_perms + 4 * k; }

    // Spot::perms_fixing, game.rs:335-344
    std::vector<size_t> perms_fixing(const Dealt& d) const {
        std::vector<size_t> ks;
        for (size_t k = 0; k < v->num_suit_perms; ++k) {
            bool ok = true;
            for (int i = 0; i < d.len; ++i)
                if (permute_card(d.cards[i], perm(k)) != d.cards[i]) ok = false;
            if (ok) ks.push_back(k);
        }
        return ks;
    }
    // orbit representative (smallest image) and the symmetry that maps the
    // card onto it, strictly-less scan over the valid symmetries
    // (cfr.rs:575-589; symmetrize_node 776-787)
    void orbit(uint8_t c, const std::vector<size_t>& valid, uint8_t* rep, size_t* k_rep) const {
        *rep = c;
        *k_rep = 0;
        for (size_t k : valid) {
            uint8_t pc = permute_card(c, perm(k));
            if (pc < *rep) {
                *rep = pc;
                *k_rep = k;
            }
        }
    }

    Engine(const gto_spot_view* view, int st, int al) : v(view), storage(st), algo(al) {
        for (int p = 0; p < 2; ++p) {
            nh[p] = v->num_hands[p];
            hmask[p].resize(nh[p]);
            for (size_t i = 0; i < nh[p]; ++i)
                hmask[p][i] = (1ull << v->hand_c1[p][i]) | (1ull << v->hand_c2[p][i]);
            if (compressed()) {
                Rq[p].assign(v->data_size[p], 0);
                Sq[p].assign(v->data_size[p], 0);
                Rs[p].assign(v->num_nodes, 0.0f);
                Ss[p].assign(v->num_nodes, 0.0f);
            } else {
                R[p].assign(v->data_size[p], T(0));
                S[p].assign(v->data_size[p], T(0));
            }
            if (predictive()) {
                if (compressed()) {
                    Pq[p].assign(v->data_size[p], 0);
                    Ps[p].assign(v->num_nodes, 0.0f);
                } else {
                    P[p].assign(v->data_size[p], T(0));
                }
            }
        }
        lock_off.assign(v->num_nodes, -1);
    }

    const float* lock_of(uint32_t n) const { return lock_off[n] >= 0 ? lock_sigma.data() + lock_off[n] : nullptr; }
    uint32_t child(uint32_t n, size_t a) const { return v->children[v->node_children_start[n] + a]; }

    // regret_match / regret_match_q / current_strategy (cfr.rs:240-292, 875-919)
    void current_strategy(uint32_t n, T* sigma) const {
        int pl = v->node_player[n];
        size_t na = v->node_num_children[n], h = nh[pl];
        if (const float* lk = lock_of(n)) {
            for (size_t k = 0; k < na * h; ++k) sigma[k] = (T)lk[k];
            return;
        }
        uint64_t off = v->node_data_offset[n];
        if (predictive()) {
            // PCFR+: sigma = max(R,0) + prediction, then regret_match_inplace
            // (cfr.rs:259-266, 276-288, 923-942)
            if (!compressed()) {
                const T* r = R[pl].data() + off;
                const T* pr = P[pl].data() + off;
                for (size_t k = 0; k < na * h; ++k) sigma[k] = std::max(r[k], T(0)) + pr[k];
            } else {
                const int16_t* qr = Rq[pl].data() + off;
                const int16_t* qp = Pq[pl].data() + off;
                const float kr = Rs[pl][n] / 32767.0f, kp = Ps[pl][n] / 32767.0f;
                for (size_t k = 0; k < na * h; ++k)
                    sigma[k] = (T)((float)std::max<int16_t>(qr[k], 0) * kr + (float)qp[k] * kp);
            }
            for (size_t i = 0; i < h; ++i) {
                T sum = 0;
                for (size_t a = 0; a < na; ++a) sum += std::max(sigma[a * h + i], T(0));
                if (sum > (T)1e-12f) {
                    T inv = T(1) / sum;
                    for (size_t a = 0; a < na; ++a) sigma[a * h + i] = std::max(sigma[a * h + i], T(0)) * inv;
                } else {
                    T u = T(1) / (T)na;
                    for (size_t a = 0; a < na; ++a) sigma[a * h + i] = u;
                }
            }
            return;
        }
        if (!compressed()) {
            const T* r = R[pl].data() + off;
            for (size_t i = 0; i < h; ++i) {
                T sum = 0;
                for (size_t a = 0; a < na; ++a) sum += std::max(r[a * h + i], T(0));
                if (sum > (T)1e-12f) {
                    T inv = T(1) / sum;
                    for (size_t a = 0; a < na; ++a) sigma[a * h + i] = std::max(r[a * h + i], T(0)) * inv;
                } else {
                    T u = T(1) / (T)na;
                    for (size_t a = 0; a < na; ++a) sigma[a * h + i] = u;
                }
            }
        } else {
            const int16_t* q = Rq[pl].data() + off;
            for (size_t i = 0; i < h; ++i) {
                int32_t sum = 0;
                for (size_t a = 0; a < na; ++a) sum += std::max<int32_t>(q[a * h + i], 0);
                if (sum > 0) {
                    T inv = T(1) / (T)sum;
                    for (size_t a = 0; a < na; ++a) sigma[a * h + i] = (T)std::max<int32_t>(q[a * h + i], 0) * inv;
                } else {
                    T u = T(1) / (T)na;
                    for (size_t a = 0; a < na; ++a) sigma[a * h + i] = u;
                }
            }
        }
    }

    // update_node, cfr.rs:297-409 (DCFR/CFR+ branches; PCFR+ also stores
    // the instantaneous regret as the next prediction, 337-348 and 386-406)
    void update_node(uint32_t n, int p, const T* sigma, const T* vals, const T* out, const T* reach_p,
                     const Discounts<T>& d) {
        size_t na = v->node_num_children[n], h = nh[p];
        uint64_t off = v->node_data_offset[n];
        if (!compressed()) {
            T* r = R[p].data() + off;
            T* s = S[p].data() + off;
            for (size_t a = 0; a < na; ++a) {
                size_t b = a * h;
                for (size_t i = 0; i < h; ++i) {
                    T rv = r[b + i];
                    T dd = rv > T(0) ? d.pos : d.neg;
                    r[b + i] = rv * dd + (vals[b + i] - out[i]);
                    s[b + i] = s[b + i] * d.strat + reach_p[i] * sigma[b + i];
                }
            }
            if (predictive()) {
                T* pr = P[p].data() + off;
                for (size_t a = 0; a < na; ++a)
                    for (size_t i = 0; i < h; ++i) pr[a * h + i] = vals[a * h + i] - out[i];
            }
            return;
        }
        // compressed: fused decode + update + block max, then quantize
        int16_t* rq = Rq[p].data() + off;
        uint16_t* sq = Sq[p].data() + off;
        float kr = Rs[p][n] / 32767.0f;
        float kss = Ss[p][n] / 65535.0f * (float)d.strat;
        Buf<float> rbuf(na * h, false), sbuf(na * h, false);
        float rmax = 0.0f, smax = 0.0f;
        for (size_t a = 0; a < na; ++a) {
            size_t b = a * h;
            for (size_t i = 0; i < h; ++i) {
                int16_t qv = rq[b + i];
                float dd = qv > 0 ? (float)d.pos : (float)d.neg;
                float nr = (float)qv * kr * dd + ((float)vals[b + i] - (float)out[i]);
                rbuf[b + i] = nr;
                rmax = std::fmax(rmax, std::fabs(nr));
            }
            for (size_t i = 0; i < h; ++i) {
                float sv = (float)sq[b + i] * kss + (float)reach_p[i] * (float)sigma[b + i];
                sbuf[b + i] = sv;
                smax = std::fmax(smax, sv);
            }
        }
        Rs[p][n] = quantize_i16(rbuf.data(), na * h, rmax, rq);
        Ss[p][n] = quantize_u16(sbuf.data(), na * h, smax, sq);
        if (predictive()) {
            float pmax = 0.0f;
            for (size_t a = 0; a < na; ++a)
                for (size_t i = 0; i < h; ++i) {
                    const float delta = (float)vals[a * h + i] - (float)out[i];
                    rbuf[a * h + i] = delta;
                    pmax = std::fmax(pmax, std::fabs(delta));
                }
            Ps[p][n] = quantize_i16(rbuf.data(), na * h, pmax, Pq[p].data() + off);
        }
    }

    static bool all_nonpositive(const T* r, size_t n) {
        for (size_t j = 0; j < n; ++j)
            if (r[j] > T(0)) return false;
        return true;
    }

    // chance_node, cfr.rs:539-664: one child per suit orbit is traversed
    // (every valid card is its own orbit when iso is off); the others take
    // the representative's values through the hand table of the symmetry
    // that maps them onto it. `rec(child, rp, ro, dealt, out)`.
    template <class F>
    void chance_node(uint32_t n, int p, const T* reach_p, const T* reach_o, Dealt dealt, T* out, const F& rec) {
        const int o = 1 - p;
        const size_t h = nh[p], ho = nh[o];
        T divisor = (T)(46 - (int)v->node_street[n]);
        uint8_t cards[52], rep_of[52], reps[52];
        size_t perm_of[52], slot_of[52];
        size_t nc = 0, nr = 0;
        uint32_t cs = v->node_children_start[n];
        for (int c = 0; c < 52; ++c)
            if (v->children[cs + c] != kSent && !dealt.contains((uint8_t)c)) cards[nc++] = (uint8_t)c;
        const std::vector<size_t> valid = iso() ? perms_fixing(dealt) : std::vector<size_t>{};
        for (size_t k = 0; k < nc; ++k) {
            uint8_t c = cards[k];
            rep_of[c] = c;
            perm_of[c] = 0;
            if (iso()) orbit(c, valid, &rep_of[c], &perm_of[c]);
            if (rep_of[c] == c) {
                slot_of[c] = nr;
                reps[nr++] = c;
            }
        }
        bool parallel = nc > 0 && v->node_kind[v->children[cs + cards[0]]] == GTO_KIND_ACTION;
        Buf<T> results(nr * h, true);
        auto run_child = [&](size_t k) {
            uint8_t c = reps[k];
            uint64_t cm = 1ull << c;
            Buf<T> rp(h, false), ro(ho, false);
            for (size_t i = 0; i < h; ++i) rp[i] = (hmask[p][i] & cm) ? T(0) : reach_p[i];
            for (size_t j = 0; j < ho; ++j) ro[j] = (hmask[o][j] & cm) ? T(0) : reach_o[j];
            rec(v->children[cs + c], rp.data(), ro.data(), dealt.push(c), results.data() + k * h);
        };
        if (parallel) pool().parallel_for(nr, run_child);
        else
            for (size_t k = 0; k < nr; ++k) run_child(k);
        for (size_t k = 0; k < nc; ++k) {
            uint8_t c = cards[k];
            uint64_t cm = 1ull << c;
            const T* co = results.data() + slot_of[rep_of[c]] * h;
            if (rep_of[c] == c) {
                for (size_t i = 0; i < h; ++i)
                    if (!(hmask[p][i] & cm)) out[i] += co[i];
            } else {
                const uint16_t* tbl = v->hand_perm[p] + perm_of[c] * h;
                for (size_t i = 0; i < h; ++i)
                    if (!(hmask[p][i] & cm)) out[i] += co[tbl[i]];
            }
        }
        T inv = T(1) / divisor;
        for (size_t i = 0; i < h; ++i) out[i] *= inv;
    }

    // ensure_symmetric / symmetrize_node / copy_branch, cfr.rs:749-870:
    // representatives' subtrees first, then each orbit member takes a
    // suit-permuted copy of its representative's subtree (regrets, strategy
    // sums, or a locked source's strategy; predictions are not copied)
    void ensure_symmetric() {
        if (!sym_dirty || !iso()) {
            sym_dirty = false;
            return;
        }
        symmetrize_node(0, Dealt());
        sym_dirty = false;
    }
    void symmetrize_node(uint32_t n, Dealt dealt) {
        uint32_t cs = v->node_children_start[n];
        if (v->node_kind[n] == GTO_KIND_ACTION) {
            for (size_t a = 0; a < v->node_num_children[n]; ++a) symmetrize_node(child(n, a), dealt);
        } else if (v->node_kind[n] == GTO_KIND_CHANCE) {
            const std::vector<size_t> valid = perms_fixing(dealt);
            std::vector<uint8_t> cards, reps;
            std::vector<size_t> ks;
            for (int c = 0; c < 52; ++c)
                if (v->children[cs + c] != kSent && !dealt.contains((uint8_t)c)) {
                    uint8_t rep;
                    size_t k;
                    orbit((uint8_t)c, valid, &rep, &k);
                    cards.push_back((uint8_t)c);
                    reps.push_back(rep);
                    ks.push_back(k);
                }
            for (size_t j = 0; j < cards.size(); ++j)
                if (cards[j] == reps[j]) symmetrize_node(v->children[cs + cards[j]], dealt.push(cards[j]));
            for (size_t j = 0; j < cards.size(); ++j)
                if (cards[j] != reps[j]) copy_branch(v->children[cs + reps[j]], v->children[cs + cards[j]], ks[j]);
        }
    }
    void copy_branch(uint32_t src, uint32_t dst, size_t k) {
        if (v->node_kind[src] == GTO_KIND_ACTION) {
            const int p = v->node_player[src];
            const size_t h = nh[p], na = v->node_num_children[src], cnt = na * h;
            const uint16_t* tbl = v->hand_perm[p] + k * h;
            const uint64_t so = v->node_data_offset[src], dof = v->node_data_offset[dst];
            const float* lk = lock_of(src);
            if (!compressed()) {
                for (size_t a = 0; a < na; ++a)
                    for (size_t i = 0; i < h; ++i) {
                        const size_t j = tbl[i];
                        R[p][dof + a * h + i] = R[p][so + a * h + j];
                        S[p][dof + a * h + i] = lk ? (T)lk[a * h + j] : S[p][so + a * h + j];
                    }
            } else {
                // permuting a block leaves its max, hence its scale and the
                // re-encoded quants, unchanged (store.rs:169-215)
                for (size_t a = 0; a < na; ++a)
                    for (size_t i = 0; i < h; ++i) Rq[p][dof + a * h + i] = Rq[p][so + a * h + tbl[i]];
                Rs[p][dst] = Rs[p][src];
                if (lk) {
                    std::vector<float> sb(cnt);
                    float smax = 0.0f;
                    for (size_t a = 0; a < na; ++a)
                        for (size_t i = 0; i < h; ++i) {
                            sb[a * h + i] = lk[a * h + tbl[i]];
                            smax = std::fmax(smax, sb[a * h + i]);
                        }
                    Ss[p][dst] = quantize_u16(sb.data(), cnt, smax, Sq[p].data() + dof);
                } else {
                    for (size_t a = 0; a < na; ++a)
                        for (size_t i = 0; i < h; ++i) Sq[p][dof + a * h + i] = Sq[p][so + a * h + tbl[i]];
                    Ss[p][dst] = Ss[p][src];
                }
            }
            for (size_t a = 0; a < na; ++a) copy_branch(child(src, a), child(dst, a), k);
        } else if (v->node_kind[src] == GTO_KIND_CHANCE) {
            const uint32_t ss = v->node_children_start[src], ds = v->node_children_start[dst];
            for (int x = 0; x < 52; ++x) {
                const uint32_t dc = v->children[ds + x];
                if (dc == kSent) continue;
                copy_branch(v->children[ss + permute_card((uint8_t)x, perm(k))], dc, k);
            }
        }
    }

    // cfr, cfr.rs:414-534. `out` is zero on entry.
    void cfr(uint32_t n, int p, const T* reach_p, const T* reach_o, Dealt dealt, const Discounts<T>& d, T* out) {
        const int o = 1 - p;
        if (all_nonpositive(reach_o, nh[o])) return;  // zero-reach pruning, cfr.rs:424-430
        switch (v->node_kind[n]) {
            case GTO_KIND_FOLD: {
                T amount = (T)(v->node_player[n] == p ? v->node_t_lose[n] : v->node_t_win[n]);
                fold_cfv<T>(v, p, reach_o, amount, out);
                return;
            }
            case GTO_KIND_SHOWDOWN:
                showdown_cfv<T>(v, p, river_key(v, dealt), reach_o, (T)v->node_t_win[n], (T)v->node_t_lose[n],
                                (T)v->node_t_tie[n], out);
                return;
            case GTO_KIND_CHANCE:
                chance_node(n, p, reach_p, reach_o, dealt, out,
                            [&](uint32_t c, const T* rp, const T* ro, Dealt dl, T* co) { cfr(c, p, rp, ro, dl, d, co); });
                return;
            default:
                break;
        }
        size_t na = v->node_num_children[n];
        if (v->node_player[n] == p) {
            size_t h = nh[p];
            bool locked = lock_off[n] >= 0;
            Buf<T> sigma(na * h, false), vals(na * h, true), reach(h, false);
            current_strategy(n, sigma.data());
            for (size_t a = 0; a < na; ++a) {
                const T* sg = sigma.data() + a * h;
                for (size_t i = 0; i < h; ++i) reach[i] = reach_p[i] * sg[i];
                cfr(child(n, a), p, reach.data(), reach_o, dealt, d, vals.data() + a * h);
            }
            for (size_t a = 0; a < na; ++a) {
                const T* sg = sigma.data() + a * h;
                const T* vl = vals.data() + a * h;
                for (size_t i = 0; i < h; ++i) out[i] += sg[i] * vl[i];
            }
            if (!locked) update_node(n, p, sigma.data(), vals.data(), out, reach_p, d);
        } else {
            size_t ho = nh[o], h = nh[p];
            Buf<T> sigma(na * ho, false), reach(ho, false), co(h, false);
            current_strategy(n, sigma.data());
            for (size_t a = 0; a < na; ++a) {
                const T* sg = sigma.data() + a * ho;
                for (size_t j = 0; j < ho; ++j) reach[j] = reach_o[j] * sg[j];
                std::fill(co.data(), co.data() + h, T(0));
                cfr(child(n, a), p, reach_p, reach.data(), dealt, d, co.data());
                for (size_t i = 0; i < h; ++i) out[i] += co[i];
            }
        }
    }

    void set_isomorphism(bool on) override {
        ensure_symmetric();
        use_iso = on;
    }
    void mark_sym_dirty() override { sym_dirty = true; }

    void iterate() override {  // cfr.rs:170-190
        sym_dirty = true;
        iteration += 1;
        Discounts<T> d = discounts<T>(algo, iteration);
        for (int p = 0; p < 2; ++p) {
            std::vector<T> w[2];
            for (int q = 0; q < 2; ++q) w[q].assign(v->weights[q], v->weights[q] + nh[q]);
            Buf<T> out(nh[p], true);
            cfr(0, p, w[p].data(), w[1 - p].data(), Dealt(), d, out.data());
        }
    }

    // solved_strategy_into / average_strategy_into, cfr.rs:668-726
    void avg_strategy(uint32_t n, T* sigma) const {
        int pl = v->node_player[n];
        size_t na = v->node_num_children[n], h = nh[pl];
        if (const float* lk = lock_of(n)) {
            for (size_t k = 0; k < na * h; ++k) sigma[k] = (T)lk[k];
            return;
        }
        uint64_t off = v->node_data_offset[n];
        if (!compressed()) {
            const T* s = S[pl].data() + off;
            for (size_t i = 0; i < h; ++i) {
                double sum = 0.0;
                for (size_t a = 0; a < na; ++a) sum += (double)std::max(s[a * h + i], T(0));
                if (sum > 1e-12) {
                    for (size_t a = 0; a < na; ++a) sigma[a * h + i] = (T)((double)std::max(s[a * h + i], T(0)) / sum);
                } else {
                    T u = T(1) / (T)na;
                    for (size_t a = 0; a < na; ++a) sigma[a * h + i] = u;
                }
            }
        } else {
            const uint16_t* q = Sq[pl].data() + off;
            for (size_t i = 0; i < h; ++i) {
                uint32_t sum = 0;
                for (size_t a = 0; a < na; ++a) sum += q[a * h + i];
                if (sum > 0) {
                    T inv = T(1) / (T)sum;
                    for (size_t a = 0; a < na; ++a) sigma[a * h + i] = (T)q[a * h + i] * inv;
                } else {
                    T u = T(1) / (T)na;
                    for (size_t a = 0; a < na; ++a) sigma[a * h + i] = u;
                }
            }
        }
    }

    // br_into (mode 0) and avg_into (mode 1), best_response.rs:19-223
    void eval(uint32_t n, int p, int mode, const T* reach_o, Dealt dealt, T* out) {
        const int o = 1 - p;
        if (all_nonpositive(reach_o, nh[o])) return;
        switch (v->node_kind[n]) {
            case GTO_KIND_FOLD: {
                T amount = (T)(v->node_player[n] == p ? v->node_t_lose[n] : v->node_t_win[n]);
                fold_cfv<T>(v, p, reach_o, amount, out);
                return;
            }
            case GTO_KIND_SHOWDOWN:
                showdown_cfv<T>(v, p, river_key(v, dealt), reach_o, (T)v->node_t_win[n], (T)v->node_t_lose[n],
                                (T)v->node_t_tie[n], out);
                return;
            case GTO_KIND_CHANCE: {
                Buf<T> dummy(nh[p], T(1));
                chance_node(n, p, dummy.data(), reach_o, dealt, out,
                            [&](uint32_t c, const T*, const T* ro, Dealt dl, T* co) { eval(c, p, mode, ro, dl, co); });
                return;
            }
            default:
                break;
        }
        size_t na = v->node_num_children[n];
        size_t h = nh[p];
        if (v->node_player[n] == p) {
            Buf<T> vals(h, false);
            const float* lk = lock_of(n);
            if (mode == 0 && !lk) {
                for (size_t a = 0; a < na; ++a) {
                    std::fill(vals.data(), vals.data() + h, T(0));
                    eval(child(n, a), p, mode, reach_o, dealt, vals.data());
                    if (a == 0) std::copy(vals.data(), vals.data() + h, out);
                    else
                        for (size_t i = 0; i < h; ++i)
                            if (vals[i] > out[i]) out[i] = vals[i];
                }
            } else {
                Buf<T> sigma(na * h, false);
                avg_strategy(n, sigma.data());  // lock-aware
                for (size_t a = 0; a < na; ++a) {
                    std::fill(vals.data(), vals.data() + h, T(0));
                    eval(child(n, a), p, mode, reach_o, dealt, vals.data());
                    const T* sg = sigma.data() + a * h;
                    for (size_t i = 0; i < h; ++i) out[i] += sg[i] * vals[i];
                }
            }
        } else {
            size_t ho = nh[o];
            Buf<T> sigma(na * ho, false), reach(ho, false), vals(h, false);
            avg_strategy(n, sigma.data());
            for (size_t a = 0; a < na; ++a) {
                const T* sg = sigma.data() + a * ho;
                for (size_t j = 0; j < ho; ++j) reach[j] = reach_o[j] * sg[j];
                std::fill(vals.data(), vals.data() + h, T(0));
                eval(child(n, a), p, mode, reach.data(), dealt, vals.data());
                for (size_t i = 0; i < h; ++i) out[i] += vals[i];
            }
        }
    }

    double pair_weight_sum() override {  // best_response.rs:227-251
        double t = 0.0, s[52];
        std::fill(s, s + 52, 0.0);
        for (size_t j = 0; j < nh[1]; ++j) {
            double r = (double)v->weights[1][j];
            t += r;
            s[v->hand_c1[1][j]] += r;
            s[v->hand_c2[1][j]] += r;
        }
        double total = 0.0;
        for (size_t i = 0; i < nh[0]; ++i) {
            uint32_t sm = v->same_combo[0][i];
            double same_r = sm != kSent ? (double)v->weights[1][sm] : 0.0;
            double valid = t - s[v->hand_c1[0][i]] - s[v->hand_c2[0][i]] + same_r;
            total += (double)v->weights[0][i] * valid;
        }
        return total;
    }

    void root_cfv(int p, int mode, double* out) override {
        std::vector<T> w(v->weights[1 - p], v->weights[1 - p] + nh[1 - p]);
        Buf<T> cfv(nh[p], true);
        eval(0, p, mode, w.data(), Dealt(), cfv.data());
        for (size_t i = 0; i < nh[p]; ++i) out[i] = (double)cfv[i];
    }

    double exploitability() override {  // best_response.rs:255-287
        double denom = pair_weight_sum();
        if (denom <= 0.0) return NAN;
        double br[2], val[2];
        for (int p = 0; p < 2; ++p) {
            std::vector<double> c(nh[p]);
            root_cfv(p, 0, c.data());
            double acc = 0.0;
            for (size_t i = 0; i < nh[p]; ++i) acc += (double)v->weights[p][i] * (double)(T)c[i];
            br[p] = acc / denom;
        }
        if (v->rake_pct > 0.0) {
            for (int p = 0; p < 2; ++p) {
                std::vector<double> c(nh[p]);
                root_cfv(p, 1, c.data());
                double acc = 0.0;
                for (size_t i = 0; i < nh[p]; ++i) acc += (double)v->weights[p][i] * (double)(T)c[i];
                val[p] = acc / denom;
            }
            return ((br[0] - val[0]) + (br[1] - val[1])) / 2.0;
        }
        return (br[0] + br[1]) / 2.0;
    }

    void get_arenas(int which, int p, float* out) override {
        ensure_symmetric();
        size_t len = v->data_size[p];
        if (!compressed()) {
            const std::vector<T>& a = which == 0 ? R[p] : S[p];
            for (size_t k = 0; k < len; ++k) out[k] = (float)a[k];
            return;
        }
        // decode per action node of player p (store.rs:153-166)
        for (uint32_t n = 0; n < v->num_nodes; ++n) {
            if (v->node_kind[n] != GTO_KIND_ACTION || v->node_player[n] != p) continue;
            size_t cnt = (size_t)v->node_num_children[n] * nh[p];
            uint64_t off = v->node_data_offset[n];
            if (which == 0) {
                float k = Rs[p][n] / 32767.0f;
                for (size_t e = 0; e < cnt; ++e) out[off + e] = (float)Rq[p][off + e] * k;
            } else {
                float k = Ss[p][n] / 65535.0f;
                for (size_t e = 0; e < cnt; ++e) out[off + e] = (float)Sq[p][off + e] * k;
            }
        }
    }
    void get_arenas_f64(int which, int p, double* out) override {
        ensure_symmetric();
        size_t len = v->data_size[p];
        if (!compressed()) {
            const std::vector<T>& a = which == 0 ? R[p] : S[p];
            for (size_t k = 0; k < len; ++k) out[k] = (double)a[k];
            return;
        }
        std::vector<float> tmp(len);
        get_arenas(which, p, tmp.data());
        for (size_t k = 0; k < len; ++k) out[k] = tmp[k];
    }
    void set_arenas(int which, int p, const float* in) override {
        size_t len = v->data_size[p];
        if (!compressed()) {
            std::vector<T>& a = which == 0 ? R[p] : S[p];
            for (size_t k = 0; k < len; ++k) a[k] = (T)in[k];
            return;
        }
        for (uint32_t n = 0; n < v->num_nodes; ++n) {
            if (v->node_kind[n] != GTO_KIND_ACTION || v->node_player[n] != p) continue;
            size_t cnt = (size_t)v->node_num_children[n] * nh[p];
            uint64_t off = v->node_data_offset[n];
            if (which == 0) Rs[p][n] = oracle_encode_i16(in + off, cnt, Rq[p].data() + off);
            else Ss[p][n] = oracle_encode_u16(in + off, cnt, Sq[p].data() + off);
        }
    }
    void average_strategy(uint32_t n, float* out) override {
        ensure_symmetric();
        int pl = v->node_player[n];
        size_t cnt = (size_t)v->node_num_children[n] * nh[pl];
        std::vector<T> s(cnt);
        avg_strategy(n, s.data());
        for (size_t k = 0; k < cnt; ++k) out[k] = (float)s[k];
    }
};

}  // namespace

struct oracle {
    EngineBase* e;
};

extern "C" {

oracle* oracle_create(const gto_spot_view* spot, int precision, int storage, int algo) {
    oracle* o = new oracle();
    if (precision == 64) o->e = new Engine<double>(spot, storage == GTO_STORAGE_COMPRESSED ? GTO_STORAGE_F32 : storage, algo);
    else o->e = new Engine<float>(spot, storage, algo);
    return o;
}
void oracle_destroy(oracle* o) {
    if (!o) return;
    delete o->e;
    delete o;
}
void oracle_set_threads(int n) { pool().resize(n); }
int oracle_get_threads(void) { return pool().threads(); }
void oracle_iterate(oracle* o, uint32_t n) {
    for (uint32_t k = 0; k < n; ++k) o->e->iterate();
}
uint32_t oracle_iteration(const oracle* o) { return o->e->iteration; }
void oracle_set_iteration(oracle* o, uint32_t t) { o->e->iteration = t; }
double oracle_exploitability(oracle* o) { return o->e->exploitability(); }
double oracle_pair_weight_sum(oracle* o) { return o->e->pair_weight_sum(); }
void oracle_root_cfv(oracle* o, int p, int mode, double* out) { o->e->root_cfv(p, mode, out); }
void oracle_get_arenas(oracle* o, float* r0, float* r1, float* s0, float* s1) {
    if (r0) o->e->get_arenas(0, 0, r0);
    if (r1) o->e->get_arenas(0, 1, r1);
    if (s0) o->e->get_arenas(1, 0, s0);
    if (s1) o->e->get_arenas(1, 1, s1);
}
void oracle_get_arenas_f64(oracle* o, double* r0, double* r1, double* s0, double* s1) {
    if (r0) o->e->get_arenas_f64(0, 0, r0);
    if (r1) o->e->get_arenas_f64(0, 1, r1);
    if (s0) o->e->get_arenas_f64(1, 0, s0);
    if (s1) o->e->get_arenas_f64(1, 1, s1);
}
void oracle_set_arenas(oracle* o, const float* r0, const float* r1, const float* s0, const float* s1) {
    if (r0) o->e->set_arenas(0, 0, r0);
    if (r1) o->e->set_arenas(0, 1, r1);
    if (s0) o->e->set_arenas(1, 0, s0);
    if (s1) o->e->set_arenas(1, 1, s1);
}
void oracle_set_locks(oracle* o, const int64_t* lock_off, const float* sigma, uint64_t n_sigma) {
    size_t nn = o->e->lock_off.size();
    if (lock_off) o->e->lock_off.assign(lock_off, lock_off + nn);
    else o->e->lock_off.assign(nn, -1);
    o->e->lock_sigma.assign(sigma, sigma + n_sigma);
    o->e->mark_sym_dirty();
}
void oracle_set_isomorphism(oracle* o, int on) { o->e->set_isomorphism(on != 0); }
void oracle_average_strategy(oracle* o, uint32_t node, float* out) { o->e->average_strategy(node, out); }

void oracle_fold_cfv(const gto_spot_view* spot, int p, const float* reach_o, float amount, float* out) {
    fold_cfv<float>(spot, p, reach_o, amount, out);
}
void oracle_showdown_cfv(const gto_spot_view* spot, int p, uint32_t river_key, const float* reach_o, float win,
                         float lose, float tie, float* out) {
    std::fill(out, out + spot->num_hands[p], 0.0f);
    showdown_cfv<float>(spot, p, river_key, reach_o, win, lose, tie, out);
}
float oracle_encode_i16(const float* src, uint64_t n, int16_t* dst) {
    float m = 0.0f;
    for (uint64_t k = 0; k < n; ++k) m = std::fmax(m, std::fabs(src[k]));
    return quantize_i16(src, n, m, dst);
}
float oracle_encode_u16(const float* src, uint64_t n, uint16_t* dst) {
    float m = 0.0f;
    for (uint64_t k = 0; k < n; ++k) m = std::fmax(m, src[k]);
    return quantize_u16(src, n, m, dst);
}

}  // extern "C"
