// model_io.cpp -- model files (model.cu:24-115) and weight snapshots
// (NeuralNetwork.cu:74-104, util.cuh:271-303).
#include "model_io.h"

#include <sys/stat.h>

#include <cerrno>
#include <cstdlib>
#include <fstream>
#include <iterator>
#include <sstream>

namespace pir {

namespace {

// ---------------------------------------------------------------------------
// A small JSON reader (RFC 8259 values; numbers through strtod) for the model
// files.

struct JValue {
    enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
    bool b = false;
    double num = 0.0;
    std::string str;
    std::vector<JValue> arr;
    std::vector<std::pair<std::string, JValue>> obj;

    const JValue *find(const std::string &key) const {
        for (const auto &kv : obj)
            if (kv.first == key) return &kv.second;
        return nullptr;
    }
};

class JsonReader {
  public:
    explicit JsonReader(const std::string &text) : s_(text) {}

    JValue document() {
        JValue v = value();
        space();
        if (i_ != s_.size()) fail("unexpected text after the document");
        return v;
    }

  private:
    [[noreturn]] void fail(const std::string &what) {
        set_error("model file: " + what + " at offset " + std::to_string(i_));
        throw PirError(ERR_ARG);
    }
    void space() {
        while (i_ < s_.size() && (s_[i_] == ' ' || s_[i_] == '\t' || s_[i_] == '\n' || s_[i_] == '\r')) i_++;
    }
    bool take(char c) {
        space();
        if (i_ < s_.size() && s_[i_] == c) {
            i_++;
            return true;
        }
        return false;
    }
    void expect(char c) {
        if (!take(c)) fail(std::string("expected '") + c + "'");
    }
    bool literal(const char *word) {
        size_t n = std::char_traits<char>::length(word);
        if (s_.compare(i_, n, word) == 0) {
            i_ += n;
            return true;
        }
        return false;
    }
    static void put_utf8(std::string &out, unsigned cp) {
        if (cp < 0x80) {
            out += (char)cp;
        } else if (cp < 0x800) {
            out += (char)(0xC0 | (cp >> 6));
            out += (char)(0x80 | (cp & 0x3F));
        } else if (cp < 0x10000) {
            out += (char)(0xE0 | (cp >> 12));
            out += (char)(0x80 | ((cp >> 6) & 0x3F));
            out += (char)(0x80 | (cp & 0x3F));
        } else {
            out += (char)(0xF0 | (cp >> 18));
            out += (char)(0x80 | ((cp >> 12) & 0x3F));
            out += (char)(0x80 | ((cp >> 6) & 0x3F));
            out += (char)(0x80 | (cp & 0x3F));
        }
    }
    unsigned hex4() {
        if (i_ + 4 > s_.size()) fail("truncated \\u escape");
        unsigned v = 0;
        for (int k = 0; k < 4; k++) {
            char c = s_[i_++];
            v <<= 4;
            if (c >= '0' && c <= '9') v |= (unsigned)(c - '0');
            else if (c >= 'a' && c <= 'f') v |= (unsigned)(c - 'a' + 10);
            else if (c >= 'A' && c <= 'F') v |= (unsigned)(c - 'A' + 10);
            else fail("bad \\u escape");
        }
        return v;
    }
    std::string string_body() {  // after the opening quote
        std::string out;
        while (true) {
            if (i_ >= s_.size()) fail("unterminated string");
            char c = s_[i_++];
            if (c == '"') return out;
            if ((unsigned char)c < 0x20) fail("control character in a string");
            if (c != '\\') {
                out += c;
                continue;
            }
            if (i_ >= s_.size()) fail("unterminated escape");
            char e = s_[i_++];
            switch (e) {
                case '"': out += '"'; break;
                case '\\': out += '\\'; break;
                case '/': out += '/'; break;
                case 'b': out += '\b'; break;
                case 'f': out += '\f'; break;
                case 'n': out += '\n'; break;
                case 'r': out += '\r'; break;
                case 't': out += '\t'; break;
                case 'u': {
                    unsigned cp = hex4();
                    if (cp >= 0xD800 && cp < 0xDC00 && s_.compare(i_, 2, "\\u") == 0) {
                        i_ += 2;
                        unsigned lo = hex4();
                        if (lo < 0xDC00 || lo > 0xDFFF) fail("bad surrogate pair");
                        cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                    }
                    put_utf8(out, cp);
                    break;
                }
                default: fail("bad escape");
            }
        }
    }
    JValue value() {
        space();
        if (i_ >= s_.size()) fail("unexpected end of the document");
        if (++depth_ > kMaxDepth) fail("nesting deeper than 64 levels");
        struct Leave {
            int &d;
            ~Leave() { --d; }
        } leave{depth_};
        JValue v;
        char c = s_[i_];
        if (c == '{') {
            i_++;
            v.kind = JValue::Object;
            if (take('}')) return v;
            do {
                space();
                if (i_ >= s_.size() || s_[i_] != '"') fail("expected a member name");
                i_++;
                std::string key = string_body();
                expect(':');
                v.obj.emplace_back(key, value());
            } while (take(','));
            expect('}');
        } else if (c == '[') {
            i_++;
            v.kind = JValue::Array;
            if (take(']')) return v;
            do v.arr.push_back(value());
            while (take(','));
            expect(']');
        } else if (c == '"') {
            i_++;
            v.kind = JValue::String;
            v.str = string_body();
        } else if (literal("true")) {
            v.kind = JValue::Bool;
            v.b = true;
        } else if (literal("false")) {
            v.kind = JValue::Bool;
        } else if (literal("null")) {
            v.kind = JValue::Null;
        } else if (c == '-' || (c >= '0' && c <= '9')) {
            const char *start = s_.c_str() + i_;
            char *end = nullptr;
            v.num = std::strtod(start, &end);
            if (end == start) fail("bad number");
            i_ += (size_t)(end - start);
            v.kind = JValue::Number;
        } else {
            fail("unexpected character");
        }
        return v;
    }

    static constexpr int kMaxDepth = 64;
    const std::string &s_;
    size_t i_ = 0;
    int depth_ = 0;
};

[[noreturn]] void model_error(const std::string &msg, int code = ERR_ARG) {
    set_error("model file: " + msg);
    throw PirError(code);
}

const JValue &member(const JValue &o, const char *key, const std::string &where) {
    const JValue *v = o.kind == JValue::Object ? o.find(key) : nullptr;
    if (!v) model_error(where + " has no \"" + key + "\"");
    return *v;
}

int int_of(const JValue &v, const std::string &what) {
    if (v.kind != JValue::Number || !(v.num >= -2147483648.0 && v.num <= 2147483647.0) ||
        v.num != (double)(long long)v.num)
        model_error(what + " must be an integer");
    return (int)v.num;
}

int int_member(const JValue &o, const char *key, const std::string &where) {
    return int_of(member(o, key, where), where + "." + key);
}

// first entry of a [h, w] pair (the reference reads only [0] of filter_hw and
// pool_hw, and both entries of input_hw)
int pair_entry(const JValue &o, const char *key, int idx, const std::string &where) {
    const JValue &v = member(o, key, where);
    if (v.kind != JValue::Array || (int)v.arr.size() <= idx) model_error(where + "." + key + " is not a pair");
    return int_of(v.arr[idx], where + "." + key);
}

}  // namespace

ModelFile load_model_file(const std::string &filename) {
    std::ifstream f(filename, std::ios::binary);
    if (!f) model_error("cannot open " + filename);
    std::stringstream ss;
    ss << f.rdbuf();
    const std::string text = ss.str();
    JValue doc = JsonReader(text).document();
    if (doc.kind != JValue::Object) model_error("the document is not an object");

    ModelFile m;
    m.input_size = int_member(doc, "input_size", "model");
    m.num_classes = int_member(doc, "num_classes", "model");
    m.batch_size = int_member(doc, "batch_size", "model");
    const JValue &ds = member(doc, "dataset", "model");
    if (ds.kind == JValue::String) m.dataset = ds.str;
    const JValue &layers = member(doc, "model", "model");
    if (layers.kind != JValue::Array) model_error("\"model\" is not an array");
    for (size_t i = 0; i < layers.arr.size(); i++) {
        const JValue &l = layers.arr[i];
        const std::string where = "layer " + std::to_string(i);
        const JValue &t = member(l, "layer", where);
        if (t.kind != JValue::String) model_error(where + ".layer is not a string");
        LayerSpec s{};
        if (t.str == "fc") {
            s.v[0] = 0;
            s.v[1] = int_member(l, "input_dim", where);
            s.v[2] = int_member(l, "output_dim", where);
        } else if (t.str == "relu" || t.str == "ln") {
            s.v[0] = t.str == "relu" ? 2 : 5;
            s.v[1] = int_member(l, "input_dim", where);
        } else if (t.str == "cnn") {
            s.v[0] = 1;
            s.v[1] = pair_entry(l, "input_hw", 0, where);
            s.v[2] = pair_entry(l, "input_hw", 1, where);
            s.v[3] = int_member(l, "in_channels", where);
            s.v[4] = int_member(l, "out_channels", where);
            s.v[5] = pair_entry(l, "filter_hw", 0, where);
            s.v[6] = int_member(l, "stride", where);
            s.v[7] = int_member(l, "padding", where);
        } else if (t.str == "maxpool" || t.str == "averagepool") {
            s.v[0] = t.str == "maxpool" ? 3 : 4;
            s.v[1] = pair_entry(l, "input_hw", 0, where);
            s.v[2] = pair_entry(l, "input_hw", 1, where);
            s.v[3] = int_member(l, "in_channels", where);
            s.v[4] = pair_entry(l, "pool_hw", 0, where);
            s.v[5] = int_member(l, "stride", where);
        } else if (t.str == "res") {
            s.v[0] = 6;
            s.v[1] = pair_entry(l, "input_hw", 0, where);
            s.v[2] = pair_entry(l, "input_hw", 1, where);
            s.v[3] = int_member(l, "in_planes", where);
            s.v[4] = int_member(l, "out_planes", where);
            s.v[5] = int_member(l, "num_blocks", where);
            s.v[6] = int_member(l, "stride", where);
            s.v[7] = int_member(l, "expansion", where);
        } else {
            model_error(where + ": unknown layer type \"" + t.str + "\"");
        }
        m.layers.push_back(s);
    }
    return m;
}

// ---------------------------------------------------------------------------
// Snapshots

namespace {

// copyToHost of a share (util.cuh:56-79): party A's reconstruction
// sA + sB + (C's sA), then fromFixed (double)(int64)v / 2^fp.
std::vector<double> reconstructed_doubles(const RSS64 &x, int fp) {
    const size_t n = x.size();
    std::vector<u64> h(6 * n);
    if (n) {
        hip_check(hipMemcpyAsync(h.data(), x.base(), 6 * n * sizeof(u64), hipMemcpyDeviceToHost, ctx().stream));
        hip_check(hipStreamSynchronize(ctx().stream));
    }
    std::vector<double> out(n);
    for (size_t k = 0; k < n; k++) {
        u64 v = h[0 * n + k] + h[1 * n + k] + h[4 * n + k];
        out[k] = (double)(s64)v / (double)(1 << fp);
    }
    return out;
}

// saveShareToFile, util.cuh:291-303
void save_share(const std::string &file, const RSS64 &x, int fp) {
    std::vector<double> vals = reconstructed_doubles(x, fp);
    std::ofstream f(file);
    if (!f) {
        set_error("snapshot: cannot write " + file);
        throw PirError(ERR_ARG);
    }
    std::copy(vals.begin(), vals.end(), std::ostream_iterator<double>(f, " "));
    if (!f) {
        set_error("snapshot: write failed for " + file);
        throw PirError(ERR_ARG);
    }
}

// loadShareFromFile, util.cuh:271-289
void load_share(const std::string &file, RSS64 &dest, int fp) {
    std::ifstream f(file);
    std::istream_iterator<double> it(f), eos;
    std::vector<double> vals(dest.size(), 0.0);
    size_t k = 0;
    while (it != eos && k < vals.size()) {
        vals[k++] = *it;
        ++it;
    }
    dest.setPublic(vals, fp);
}

}  // namespace

void save_snapshot(NeuralNetwork &net, const std::string &path) {
    if (mkdir(path.c_str(), S_IRWXU | S_IRWXG) != 0 && errno != EEXIST) {
        set_error("snapshot: cannot create directory " + path);
        throw PirError(ERR_ARG);
    }
    const int fp = ctx().frac_bits;
    for (auto &l : net.layers)
        for (auto &t : l->snapshotTensors()) save_share(path + "/" + t.first, *t.second, fp);
}

void load_snapshot(NeuralNetwork &net, const std::string &path) {
    const int fp = ctx().frac_bits;
    for (auto &l : net.layers) {
        auto tensors = l->snapshotTensors();
        for (auto &t : tensors) load_share(path + "/" + t.first, *t.second, fp);
        if (!tensors.empty()) l->weightsChanged();
    }
}

}  // namespace pir
