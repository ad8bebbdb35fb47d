// GTOSOLVE2 save files (crates/solver/src/save.rs): the magic line
// "GTOSOLVE2\n", one JSON header line {"config", "iteration", "locks",
// "lock_labels"}, then the four arenas -- regrets OOP, IP, strategy sums
// OOP, IP -- each as a little-endian u64 length and that many f32. The file
// is always f32 whatever the engine's storage (save.rs:1-5), so compressed
// and f32 solves write interchangeable files.
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fcntl.h>
#include <unistd.h>

#include "abi.h"

namespace gto {

void set_last_error(const std::string& msg);

namespace {

constexpr char kMagic[] = "GTOSOLVE2\n";
constexpr size_t kMagicLen = 10;

std::string sys_err(const std::string& what) { return what + ": " + std::strerror(errno); }

struct File {
    FILE* f = nullptr;
    ~File() {
        if (f) std::fclose(f);
    }
};

}  // namespace

void validate_locks(const Spot& spot, const std::vector<std::pair<uint32_t, std::vector<float>>>& locks) {
    const size_t n_nodes = spot.tree.num_nodes();
    for (const auto& [idx, sigma] : locks) {
        const std::string at = "lock at node " + std::to_string(idx) + ": ";
        if (idx >= n_nodes)
            throw Error(at + "index out of range (tree has " + std::to_string(n_nodes) + " nodes)");
        if (spot.tree.kind[idx] != KIND_ACTION) throw Error(at + "not an action node");
        const size_t na = spot.tree.num_children[idx];
        const size_t nh = spot.num_hands(spot.tree.player[idx]);
        if (sigma.size() != na * nh)
            throw Error(at + "sigma has " + std::to_string(sigma.size()) + " entries, expected " + std::to_string(na) +
                        " actions x " + std::to_string(nh) + " hands = " + std::to_string(na * nh));
        for (float v : sigma)
            if (!std::isfinite(v) || v < 0.0f) {
                char b[64];
                std::snprintf(b, sizeof b, "%g", (double)v);
                throw Error(at + "invalid frequency " + b + " (must be finite and >= 0)");
            }
    }
}

// Stage into "<path>.tmp", make it durable, then rename over the
// destination, so a failed write never destroys the previous save
// (save.rs:94-107, 141-143).
void write_save(const std::string& path, const SaveHeader& h, const float* const arena[4], const uint64_t len[4]) {
    const std::string tmp = path + ".tmp";
    const std::string hdr = save_header_json(h);
    bool ok = false;
    {
        File out;
        out.f = std::fopen(tmp.c_str(), "wb");
        if (!out.f) throw Error(sys_err("cannot create " + tmp));
        ok = std::fwrite(kMagic, 1, kMagicLen, out.f) == kMagicLen &&
             std::fwrite(hdr.data(), 1, hdr.size(), out.f) == hdr.size() && std::fputc('\n', out.f) != EOF;
        for (int k = 0; ok && k < 4; ++k) {
            uint8_t le[8];
            for (int b = 0; b < 8; ++b) le[b] = (uint8_t)(len[k] >> (8 * b));
            ok = std::fwrite(le, 1, 8, out.f) == 8 &&
                 (len[k] == 0 || std::fwrite(arena[k], sizeof(float), len[k], out.f) == len[k]);
        }
        ok = ok && std::fflush(out.f) == 0 && ::fsync(::fileno(out.f)) == 0;
        if (!ok) {
            std::string e = sys_err("write failed for " + tmp);
            std::fclose(out.f);
            out.f = nullptr;
            std::remove(tmp.c_str());
            throw Error(e);
        }
    }
    if (std::rename(tmp.c_str(), path.c_str()) != 0) {
        std::string e = sys_err("cannot rename " + tmp + " to " + path);
        std::remove(tmp.c_str());
        throw Error(e);
    }
}

struct SaveFile {
    std::string path;
    SaveHeader header;
    std::string config_json;
    long arena_pos = 0;
};

SaveFile open_save(const std::string& path) {
    File in;
    in.f = std::fopen(path.c_str(), "rb");
    if (!in.f) throw Error(sys_err("cannot open " + path));
    char magic[kMagicLen];
    if (std::fread(magic, 1, kMagicLen, in.f) != kMagicLen || std::memcmp(magic, kMagic, kMagicLen) != 0)
        throw Error("not a solver save file");
    std::string line;
    for (;;) {
        int c = std::fgetc(in.f);
        if (c == EOF) throw Error("not a solver save file: header line is not terminated");
        if (c == '\n') break;
        line.push_back((char)c);
    }
    SaveFile s;
    s.path = path;
    try {
        s.header = parse_save_header(line);
    } catch (const std::exception& e) {
        throw Error(std::string("bad header: ") + e.what());
    }
    s.config_json = spot_config_json(s.header.config);
    s.arena_pos = std::ftell(in.f);
    return s;
}

// Arenas against the rebuilt tree's sizes (save.rs:177-192); a NULL
// destination skips that arena.
void read_arenas(const SaveFile& s, const uint64_t data_size[2], float* const out[4]) {
    File in;
    in.f = std::fopen(s.path.c_str(), "rb");
    if (!in.f) throw Error(sys_err("cannot open " + s.path));
    if (std::fseek(in.f, s.arena_pos, SEEK_SET) != 0) throw Error(sys_err("seek failed in " + s.path));
    for (int k = 0; k < 4; ++k) {
        uint8_t le[8];
        if (std::fread(le, 1, 8, in.f) != 8) throw Error("truncated save file: missing arena " + std::to_string(k));
        uint64_t len = 0;
        for (int b = 0; b < 8; ++b) len |= (uint64_t)le[b] << (8 * b);
        const uint64_t expected = data_size[k % 2];
        if (len != expected)
            throw Error("arena size mismatch: file " + std::to_string(len) + ", expected " + std::to_string(expected) +
                        " (tree config changed?)");
        if (out[k]) {
            if (len && std::fread(out[k], sizeof(float), len, in.f) != len)
                throw Error("truncated save file: arena " + std::to_string(k));
        } else if (std::fseek(in.f, (long)(len * sizeof(float)), SEEK_CUR) != 0) {
            throw Error(sys_err("seek failed in " + s.path));
        }
    }
}

}  // namespace gto

struct gto_save {
    gto::SaveFile file;
};

namespace {
template <class F>
int guarded(F&& f) {
    try {
        f();
        return GTO_OK;
    } catch (const std::bad_alloc&) {
        gto::set_last_error("out of host memory");
        return GTO_ERR_MEMORY;
    } catch (const std::exception& e) {
        gto::set_last_error(e.what());
        return GTO_ERR_INVALID;
    }
}

std::vector<std::pair<uint32_t, std::vector<float>>> lock_list(uint32_t n, const uint32_t* nodes, const uint64_t* off,
                                                               const uint64_t* len, const float* sigma) {
    std::vector<std::pair<uint32_t, std::vector<float>>> v;
    for (uint32_t k = 0; k < n; ++k) v.emplace_back(nodes[k], std::vector<float>(sigma + off[k], sigma + off[k] + len[k]));
    return v;
}
}  // namespace

extern "C" {

int gto_save_write(const char* path, const gto_spot* spot, uint32_t iteration, const float* regrets0,
                   const float* regrets1, const float* strat0, const float* strat1, uint32_t n_locks,
                   const uint32_t* lock_nodes, const uint64_t* lock_off, const uint64_t* lock_len,
                   const float* lock_sigma, uint32_t n_labels, const uint32_t* label_nodes,
                   const char* const* labels) {
    if (!path || !spot || !regrets0 || !regrets1 || !strat0 || !strat1 ||
        (n_locks && (!lock_nodes || !lock_off || !lock_len || !lock_sigma)) ||
        (n_labels && (!label_nodes || !labels))) {
        gto::set_last_error("gto_save_write: null argument");
        return GTO_ERR_INVALID;
    }
    return guarded([&] {
        gto::SaveHeader h;
        h.config = spot->spot->config;
        h.iteration = iteration;
        h.locks = lock_list(n_locks, lock_nodes, lock_off, lock_len, lock_sigma);
        gto::validate_locks(*spot->spot, h.locks);
        for (uint32_t k = 0; k < n_labels; ++k) h.lock_labels.emplace_back(label_nodes[k], labels[k]);
        const float* arenas[4] = {regrets0, regrets1, strat0, strat1};
        const uint64_t* ds = spot->spot->tree.data_size;
        const uint64_t len[4] = {ds[0], ds[1], ds[0], ds[1]};
        gto::write_save(path, h, arenas, len);
    });
}

int gto_save_open(const char* path, gto_save** out) {
    if (!path || !out) {
        gto::set_last_error("gto_save_open: null argument");
        return GTO_ERR_INVALID;
    }
    *out = nullptr;
    return guarded([&] {
        auto s = std::make_unique<gto_save>();
        s->file = gto::open_save(path);
        *out = s.release();
    });
}

void gto_save_close(gto_save* s) { delete s; }

const char* gto_save_config_json(const gto_save* s) { return s ? s->file.config_json.c_str() : nullptr; }

uint32_t gto_save_iteration(const gto_save* s) { return s ? s->file.header.iteration : 0; }

uint32_t gto_save_num_locks(const gto_save* s) { return s ? (uint32_t)s->file.header.locks.size() : 0; }

int gto_save_lock(const gto_save* s, uint32_t k, uint32_t* node, const float** sigma, uint64_t* len) {
    if (!s || k >= s->file.header.locks.size()) {
        gto::set_last_error("gto_save_lock: bad handle or index");
        return GTO_ERR_INVALID;
    }
    const auto& e = s->file.header.locks[k];
    if (node) *node = e.first;
    if (sigma) *sigma = e.second.data();
    if (len) *len = e.second.size();
    return GTO_OK;
}

uint32_t gto_save_num_labels(const gto_save* s) { return s ? (uint32_t)s->file.header.lock_labels.size() : 0; }

int gto_save_label(const gto_save* s, uint32_t k, uint32_t* node, const char** label) {
    if (!s || k >= s->file.header.lock_labels.size()) {
        gto::set_last_error("gto_save_label: bad handle or index");
        return GTO_ERR_INVALID;
    }
    const auto& e = s->file.header.lock_labels[k];
    if (node) *node = e.first;
    if (label) *label = e.second.c_str();
    return GTO_OK;
}

int gto_save_validate_locks(const gto_save* s, const gto_spot* spot) {
    if (!s || !spot) {
        gto::set_last_error("gto_save_validate_locks: null argument");
        return GTO_ERR_INVALID;
    }
    return guarded([&] {
        try {
            gto::validate_locks(*spot->spot, s->file.header.locks);
        } catch (const std::exception& e) {
            throw gto::Error(std::string("bad lock section: ") + e.what());
        }
    });
}

int gto_save_read_arenas(const gto_save* s, const gto_spot* spot, float* regrets0, float* regrets1, float* strat0,
                         float* strat1) {
    if (!s || !spot) {
        gto::set_last_error("gto_save_read_arenas: null argument");
        return GTO_ERR_INVALID;
    }
    return guarded([&] {
        float* out[4] = {regrets0, regrets1, strat0, strat1};
        gto::read_arenas(s->file, spot->spot->tree.data_size, out);
    });
}

int gto_validate_locks(const gto_spot* spot, uint32_t n_locks, const uint32_t* lock_nodes, const uint64_t* lock_off,
                       const uint64_t* lock_len, const float* lock_sigma) {
    if (!spot || (n_locks && (!lock_nodes || !lock_off || !lock_len || !lock_sigma))) {
        gto::set_last_error("gto_validate_locks: null argument");
        return GTO_ERR_INVALID;
    }
    return guarded([&] { gto::validate_locks(*spot->spot, lock_list(n_locks, lock_nodes, lock_off, lock_len, lock_sigma)); });
}

}  // extern "C"
