// Minimal JSON reader for the SpotConfig wire format (the reference's serde
// layout, crates/solver/src/game.rs:21-27 and tree.rs:14-95): bet sizes are
// externally tagged enums ({"PotPct": 33}, {"PrevMult": 2.5}, "AllIn").
#include <cmath>
#include <cstdlib>
#include <map>

#include "host.h"

namespace gto {

namespace {

struct JVal {
    enum Type { Null, Bool, Num, Str, Arr, Obj } type = Null;
    bool b = false;
    double num = 0.0;
    std::string str;
    std::vector<JVal> arr;
    std::vector<std::pair<std::string, JVal>> obj;
    const JVal* get(const std::string& k) const {
        for (auto& kv : obj)
            if (kv.first == k) return &kv.second;
        return nullptr;
    }
};

class Parser {
   public:
    explicit Parser(const std::string& s) : s_(s) {}
    JVal parse() {
        JVal v = value();
        ws();
        if (i_ != s_.size()) fail("trailing characters");
        return v;
    }

   private:
    const std::string& s_;
    size_t i_ = 0;
    [[noreturn]] void fail(const char* what) {
        throw Error(std::string("invalid config JSON: ") + what + " at byte " + std::to_string(i_));
    }
    void ws() {
        while (i_ < s_.size() && (s_[i_] == ' ' || s_[i_] == '\n' || s_[i_] == '\t' || s_[i_] == '\r')) ++i_;
    }
    JVal value() {
        ws();
        if (i_ >= s_.size()) fail("unexpected end");
        char c = s_[i_];
        JVal v;
        if (c == '{') {
            v.type = JVal::Obj;
            ++i_;
            ws();
            if (i_ < s_.size() && s_[i_] == '}') {
                ++i_;
                return v;
            }
            for (;;) {
                ws();
                if (i_ >= s_.size() || s_[i_] != '"') fail("expected key");
                std::string k = string();
                ws();
                if (i_ >= s_.size() || s_[i_] != ':') fail("expected ':'");
                ++i_;
                v.obj.push_back({k, value()});
                ws();
                if (i_ < s_.size() && s_[i_] == ',') {
                    ++i_;
                    continue;
                }
                if (i_ < s_.size() && s_[i_] == '}') {
                    ++i_;
                    return v;
                }
                fail("expected ',' or '}'");
            }
        }
        if (c == '[') {
            v.type = JVal::Arr;
            ++i_;
            ws();
            if (i_ < s_.size() && s_[i_] == ']') {
                ++i_;
                return v;
            }
            for (;;) {
                v.arr.push_back(value());
                ws();
                if (i_ < s_.size() && s_[i_] == ',') {
                    ++i_;
                    continue;
                }
                if (i_ < s_.size() && s_[i_] == ']') {
                    ++i_;
                    return v;
                }
                fail("expected ',' or ']'");
            }
        }
        if (c == '"') {
            v.type = JVal::Str;
            v.str = string();
            return v;
        }
        if (s_.compare(i_, 4, "true") == 0) {
            v.type = JVal::Bool;
            v.b = true;
            i_ += 4;
            return v;
        }
        if (s_.compare(i_, 5, "false") == 0) {
            v.type = JVal::Bool;
            i_ += 5;
            return v;
        }
        if (s_.compare(i_, 4, "null") == 0) {
            i_ += 4;
            return v;
        }
        const char* begin = s_.c_str() + i_;
        char* end = nullptr;
        double d = std::strtod(begin, &end);
        if (end == begin) fail("unexpected token");
        i_ += (size_t)(end - begin);
        v.type = JVal::Num;
        v.num = d;
        return v;
    }
    std::string string() {
        ++i_;  // opening quote
        std::string out;
        while (i_ < s_.size() && s_[i_] != '"') {
            char c = s_[i_++];
            if (c == '\\') {
                if (i_ >= s_.size()) fail("bad escape");
                char e = s_[i_++];
                switch (e) {
                    case 'n': out.push_back('\n'); break;
                    case 't': out.push_back('\t'); break;
                    case 'r': out.push_back('\r'); break;
                    case 'b': out.push_back('\b'); break;
                    case 'f': out.push_back('\f'); break;
                    case 'u': {
                        if (i_ + 4 > s_.size()) fail("bad \\u escape");
                        unsigned cp = (unsigned)std::strtoul(s_.substr(i_, 4).c_str(), nullptr, 16);
                        i_ += 4;
                        if (cp >= 0xD800 && cp < 0xDC00 && i_ + 6 <= s_.size() && s_[i_] == '\\' &&
                            s_[i_ + 1] == 'u') {
                            unsigned lo = (unsigned)std::strtoul(s_.substr(i_ + 2, 4).c_str(), nullptr, 16);
                            if (lo >= 0xDC00 && lo < 0xE000) {
                                i_ += 6;
                                cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                                out.push_back((char)(0xF0 | (cp >> 18)));
                                out.push_back((char)(0x80 | ((cp >> 12) & 0x3F)));
                                out.push_back((char)(0x80 | ((cp >> 6) & 0x3F)));
                                out.push_back((char)(0x80 | (cp & 0x3F)));
                                break;
                            }
                        }
                        if (cp < 0x80) out.push_back((char)cp);
                        else if (cp < 0x800) {
                            out.push_back((char)(0xC0 | (cp >> 6)));
                            out.push_back((char)(0x80 | (cp & 0x3F)));
                        } else {
                            out.push_back((char)(0xE0 | (cp >> 12)));
                            out.push_back((char)(0x80 | ((cp >> 6) & 0x3F)));
                            out.push_back((char)(0x80 | (cp & 0x3F)));
                        }
                        break;
                    }
                    default: out.push_back(e);
                }
            } else {
                out.push_back(c);
            }
        }
        if (i_ >= s_.size()) fail("unterminated string");
        ++i_;
        return out;
    }
};

double num_or(const JVal* v, double dflt, const char* name) {
    if (!v || v->type == JVal::Null) return dflt;
    if (v->type != JVal::Num) throw Error(std::string("invalid config JSON: ") + name + " must be a number");
    return v->num;
}

std::string str_req(const JVal& o, const char* name) {
    const JVal* v = o.get(name);
    if (!v || v->type != JVal::Str) throw Error(std::string("invalid config JSON: missing string field ") + name);
    return v->str;
}

std::vector<BetSize> sizes(const JVal* v) {
    std::vector<BetSize> out;
    if (!v || v->type == JVal::Null) return out;
    if (v->type != JVal::Arr) throw Error("invalid config JSON: size list must be an array");
    for (const JVal& e : v->arr) {
        if (e.type == JVal::Str && e.str == "AllIn") {
            out.push_back({SizeKind::AllIn, 0.0});
        } else if (e.type == JVal::Obj && e.obj.size() == 1 && e.obj[0].second.type == JVal::Num) {
            const std::string& tag = e.obj[0].first;
            double val = e.obj[0].second.num;
            if (tag == "PotPct") out.push_back({SizeKind::PotPct, val});
            else if (tag == "PrevMult") out.push_back({SizeKind::PrevMult, val});
            else throw Error("invalid config JSON: unknown bet size variant " + tag);
        } else {
            throw Error("invalid config JSON: bad bet size entry");
        }
    }
    return out;
}

void street_sizings(const JVal* v, StreetSizing out[3], const char* name) {
    if (!v || v->type != JVal::Arr || v->arr.size() != 3)
        throw Error(std::string("invalid config JSON: ") + name + " must be an array of 3 street sizings");
    for (int s = 0; s < 3; ++s) {
        const JVal& o = v->arr[s];
        if (o.type != JVal::Obj) throw Error("invalid config JSON: street sizing must be an object");
        out[s].bet = sizes(o.get("bet"));
        out[s].raise = sizes(o.get("raise"));
        out[s].donk = sizes(o.get("donk"));
    }
}


SpotConfig spot_config_of(const JVal& root) {
    if (root.type != JVal::Obj) throw Error("invalid config JSON: expected an object");
    SpotConfig c;
    c.board = str_req(root, "board");
    c.range_oop = str_req(root, "range_oop");
    c.range_ip = str_req(root, "range_ip");
    const JVal* t = root.get("tree");
    if (!t || t->type != JVal::Obj) throw Error("invalid config JSON: missing tree");
    TreeConfig& tc = c.tree;
    if (!t->get("starting_pot") || !t->get("effective_stack"))
        throw Error("invalid config JSON: tree needs starting_pot and effective_stack");
    tc.starting_pot = num_or(t->get("starting_pot"), 0.0, "starting_pot");
    tc.effective_stack = num_or(t->get("effective_stack"), 0.0, "effective_stack");
    tc.rake_pct = num_or(t->get("rake_pct"), 0.0, "rake_pct");
    tc.rake_cap = num_or(t->get("rake_cap"), 0.0, "rake_cap");
    tc.allin_threshold = num_or(t->get("allin_threshold"), 0.85, "allin_threshold");
    const JVal* aa = t->get("add_allin");
    tc.add_allin = aa && aa->type == JVal::Bool && aa->b;
    double mr = num_or(t->get("max_raises"), 10.0, "max_raises");
    if (mr < 0 || mr > 255 || mr != std::floor(mr)) throw Error("invalid config JSON: max_raises must be a u8");
    tc.max_raises = (uint8_t)mr;
    street_sizings(t->get("oop"), tc.oop, "oop");
    street_sizings(t->get("ip"), tc.ip, "ip");
    return c;
}

// ---- writer ----------------------------------------------------------------
// Numbers are written in their shortest round-trip decimal form (what
// serde_json prints for f64/f32): integral f64 values keep a ".0", exponents
// carry no '+' or leading zeros.
template <class T>
std::string shortest(T x) {
    if (!std::isfinite((double)x)) throw Error("cannot write a non-finite number to JSON");
    const int maxp = sizeof(T) == 4 ? 9 : 17;
    char buf[40];
    for (int p = 1; p <= maxp; ++p) {
        std::snprintf(buf, sizeof buf, "%.*g", p, (double)x);
        T back = sizeof(T) == 4 ? (T)std::strtof(buf, nullptr) : (T)std::strtod(buf, nullptr);
        if (back == x) break;
    }
    std::string s(buf), mant = s, ex;
    size_t e = s.find('e');
    if (e != std::string::npos) {
        mant = s.substr(0, e);
        std::string d = s.substr(e + 1);
        bool neg = d[0] == '-';
        size_t k = (d[0] == '-' || d[0] == '+') ? 1 : 0;
        while (k + 1 < d.size() && d[k] == '0') ++k;
        ex = std::string("e") + (neg ? "-" : "") + d.substr(k);
    }
    if (ex.empty() && mant.find('.') == std::string::npos) mant += ".0";
    return mant + ex;
}

void put_str(std::string& o, const std::string& s) {
    o.push_back('"');
    for (unsigned char c : s) {
        switch (c) {
            case '"': o += "\\\""; break;
            case '\\': o += "\\\\"; break;
            case '\n': o += "\\n"; break;
            case '\r': o += "\\r"; break;
            case '\t': o += "\\t"; break;
            case '\b': o += "\\b"; break;
            case '\f': o += "\\f"; break;
            default:
                if (c < 0x20) {
                    char b[8];
                    std::snprintf(b, sizeof b, "\\u%04x", c);
                    o += b;
                } else {
                    o.push_back((char)c);
                }
        }
    }
    o.push_back('"');
}

void put_sizes(std::string& o, const std::vector<BetSize>& v) {
    o.push_back('[');
    for (size_t k = 0; k < v.size(); ++k) {
        if (k) o.push_back(',');
        switch (v[k].kind) {
            case SizeKind::AllIn: o += "\"AllIn\""; break;
            case SizeKind::PotPct: o += "{\"PotPct\":" + shortest(v[k].value) + "}"; break;
            default: o += "{\"PrevMult\":" + shortest(v[k].value) + "}"; break;
        }
    }
    o.push_back(']');
}

void put_streets(std::string& o, const StreetSizing (&s)[3]) {
    o.push_back('[');
    for (int k = 0; k < 3; ++k) {
        if (k) o.push_back(',');
        o += "{\"bet\":";
        put_sizes(o, s[k].bet);
        o += ",\"raise\":";
        put_sizes(o, s[k].raise);
        o += ",\"donk\":";
        put_sizes(o, s[k].donk);
        o.push_back('}');
    }
    o.push_back(']');
}

uint32_t u32_of(const JVal& v, const char* what) {
    if (v.type != JVal::Num || !(v.num >= 0.0) || v.num > 4294967295.0 || v.num != std::floor(v.num))
        throw Error(std::string(what) + " must be an unsigned 32-bit integer");
    return (uint32_t)v.num;
}

}  // namespace

SpotConfig parse_spot_config(const std::string& json) { return spot_config_of(Parser(json).parse()); }

std::string spot_config_json(const SpotConfig& c) {
    std::string o = "{\"board\":";
    put_str(o, c.board);
    o += ",\"range_oop\":";
    put_str(o, c.range_oop);
    o += ",\"range_ip\":";
    put_str(o, c.range_ip);
    const TreeConfig& t = c.tree;
    o += ",\"tree\":{\"starting_pot\":" + shortest(t.starting_pot) + ",\"effective_stack\":" +
         shortest(t.effective_stack) + ",\"rake_pct\":" + shortest(t.rake_pct) + ",\"rake_cap\":" +
         shortest(t.rake_cap) + ",\"oop\":";
    put_streets(o, t.oop);
    o += ",\"ip\":";
    put_streets(o, t.ip);
    o += ",\"allin_threshold\":" + shortest(t.allin_threshold) + ",\"add_allin\":" +
         (t.add_allin ? "true" : "false") + ",\"max_raises\":" + std::to_string((unsigned)t.max_raises) + "}}";
    return o;
}

std::string save_header_json(const SaveHeader& h) {
    std::string o = "{\"config\":" + spot_config_json(h.config) + ",\"iteration\":" + std::to_string(h.iteration) +
                    ",\"locks\":[";
    for (size_t k = 0; k < h.locks.size(); ++k) {
        if (k) o.push_back(',');
        o += "[" + std::to_string(h.locks[k].first) + ",[";
        const std::vector<float>& sg = h.locks[k].second;
        for (size_t j = 0; j < sg.size(); ++j) {
            if (j) o.push_back(',');
            o += shortest(sg[j]);
        }
        o += "]]";
    }
    o += "],\"lock_labels\":[";
    for (size_t k = 0; k < h.lock_labels.size(); ++k) {
        if (k) o.push_back(',');
        o += "[" + std::to_string(h.lock_labels[k].first) + ",";
        put_str(o, h.lock_labels[k].second);
        o.push_back(']');
    }
    o += "]}";
    return o;
}

SaveHeader parse_save_header(const std::string& line) {
    JVal root = Parser(line).parse();
    if (root.type != JVal::Obj) throw Error("expected a JSON object");
    SaveHeader h;
    const JVal* c = root.get("config");
    if (!c) throw Error("missing field `config`");
    h.config = spot_config_of(*c);
    const JVal* it = root.get("iteration");
    if (!it) throw Error("missing field `iteration`");
    h.iteration = u32_of(*it, "iteration");
    // serde defaults: saves written before locks existed carry neither field
    if (const JVal* l = root.get("locks"); l && l->type != JVal::Null) {
        if (l->type != JVal::Arr) throw Error("locks must be an array");
        for (const JVal& e : l->arr) {
            if (e.type != JVal::Arr || e.arr.size() != 2 || e.arr[1].type != JVal::Arr)
                throw Error("lock entries must be [node, [frequencies]]");
            std::vector<float> sg;
            sg.reserve(e.arr[1].arr.size());
            for (const JVal& x : e.arr[1].arr) {
                if (x.type != JVal::Num) throw Error("lock frequencies must be numbers");
                sg.push_back((float)x.num);
            }
            h.locks.emplace_back(u32_of(e.arr[0], "lock node"), std::move(sg));
        }
    }
    if (const JVal* l = root.get("lock_labels"); l && l->type != JVal::Null) {
        if (l->type != JVal::Arr) throw Error("lock_labels must be an array");
        for (const JVal& e : l->arr) {
            if (e.type != JVal::Arr || e.arr.size() != 2 || e.arr[1].type != JVal::Str)
                throw Error("lock label entries must be [node, \"label\"]");
            h.lock_labels.emplace_back(u32_of(e.arr[0], "lock label node"), e.arr[1].str);
        }
    }
    return h;
}

}  // namespace gto
