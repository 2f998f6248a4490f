// runtime.cpp -- context, caching allocator, error state, RSS host helpers.
#include "runtime.h"

#include "device_data.h"

namespace pir {

static thread_local std::string g_error;

void set_error(const std::string &msg) { g_error = msg; }
const char *last_error() { return g_error.c_str(); }

// The free-standing calls run in a per-thread default context, so threads
// calling the library at once share no stream setting or cached block.
static thread_local Context g_default_ctx;
static thread_local Context *g_ctx = nullptr;

Context &ctx() { return g_ctx ? *g_ctx : g_default_ctx; }
void set_ctx(Context *c) { g_ctx = c; }

PrgStream call_prg(const char *what) {
    Context &c = ctx();
    if (c.precompute != PRECOMPUTE_MASKED) return PrgStream{};
    if (c.mode != MODE_FUSED) {
        set_error(std::string(what) + ": the masked precompute mode runs in the fused form only");
        throw PirError(ERR_UNSUPPORTED);
    }
    return PrgStream{c.prg_seed, c.prg_call++, 1};
}

void require_placeholder(const char *what) {
    if (ctx().precompute == PRECOMPUTE_MASKED) {
        set_error(std::string(what) + ": no masked precompute form (placeholder precompute only)");
        throw PirError(ERR_UNSUPPORTED);
    }
}

Allocator::~Allocator() {
    for (auto &kv : free_) (void)hipFree(kv.second);
    for (auto &kv : used_) (void)hipFree(kv.first);
}

void *Allocator::alloc(size_t bytes, hipStream_t stream) {
    size_t sz = (bytes + 255) & ~(size_t)255;
    auto it = free_.find({stream, sz});
    if (it != free_.end()) {
        void *p = it->second;
        free_.erase(it);
        used_[p] = sz;
        return p;
    }
    if (capturing) {
        set_error("device allocation of " + std::to_string(sz) +
                  " bytes requested during hipGraph capture (run the forward once before capturing)");
        throw PirError(ERR_STATE);
    }
    void *p = nullptr;
    if (hipMalloc(&p, sz) != hipSuccess) {
        // release the cached blocks of every stream once their work is done, and retry once
        (void)hipDeviceSynchronize();
        trim_all();
        if (hipMalloc(&p, sz) != hipSuccess) {
            set_error("hipMalloc failed for " + std::to_string(sz) + " bytes");
            throw PirError(ERR_HIP);
        }
    }
    reserved_ += sz;
    used_[p] = sz;
    return p;
}

void Allocator::trim_stream(hipStream_t stream) {
    for (auto it = free_.begin(); it != free_.end();) {
        if (it->first.first == stream) {
            (void)hipFree(it->second);
            reserved_ -= it->first.second;
            it = free_.erase(it);
        } else {
            ++it;
        }
    }
}

void Allocator::trim_all() {
    for (auto &kv : free_) {
        (void)hipFree(kv.second);
        reserved_ -= kv.first.second;
    }
    free_.clear();
}

size_t Allocator::bytes_cached() const {
    size_t b = 0;
    for (auto &kv : free_) b += kv.first.second;
    return b;
}

void Allocator::release(void *p, hipStream_t stream) {
    auto it = used_.find(p);
    if (it == used_.end()) return;
    free_.emplace(std::make_pair(stream, it->second), p);
    used_.erase(it);
}

// ---------------------------------------------------------------------------
// RSS host-side members

template <typename T>
RSS<T>::RSS(std::initializer_list<double> il, bool convert) {
    resize(il.size());
    std::vector<T> h(6 * n_, 0);
    size_t k = 0;
    for (double f : il) {
        T v = convert ? (T)to_fixed(f, ctx().frac_bits) : (T)(u64)(s64)f;
        h[(2 * PARTY_A + 0) * n_ + k] = v;
        h[(2 * PARTY_C + 1) * n_ + k] = v;
        k++;
    }
    copy_from_host(h.data());
}

template <typename T>
void RSS<T>::setPublicFixed(const std::vector<T> &f) {
    std::vector<T> h(6 * n_, 0);
    for (size_t k = 0; k < n_ && k < f.size(); k++) {
        h[(2 * PARTY_A + 0) * n_ + k] = f[k];
        h[(2 * PARTY_B + 1) * n_ + k] = f[k];  // missing `break` after PARTY_B, RSS.inl:134-140
        h[(2 * PARTY_C + 1) * n_ + k] = f[k];
    }
    copy_from_host(h.data());
}

template <typename T>
void RSS<T>::setPublic(const std::vector<double> &v, int fp) {
    std::vector<T> f(v.size());
    for (size_t k = 0; k < v.size(); k++) f[k] = (T)to_fixed(v[k], fp);
    setPublicFixed(f);
}

template <typename T>
void RSS<T>::copy_to_host(T *h) const {
    for (int q = 0; q < 3; q++)
        for (int i = 0; i < 2; i++) s_[q][i].copy_to_host(h + (2 * q + i) * n_);
}

template <typename T>
void RSS<T>::copy_from_host(const T *h) {
    for (int q = 0; q < 3; q++)
        for (int i = 0; i < 2; i++) s_[q][i].copy_from_host(h + (2 * q + i) * n_);
}

#define PIR_RSS_HOST(T)                                                   \
    template RSS<T>::RSS(std::initializer_list<double>, bool);            \
    template void RSS<T>::setPublicFixed(const std::vector<T> &);         \
    template void RSS<T>::setPublic(const std::vector<double> &, int);   \
    template void RSS<T>::copy_to_host(T *) const;                        \
    template void RSS<T>::copy_from_host(const T *);
PIR_RSS_HOST(u64)
PIR_RSS_HOST(u8)
#undef PIR_RSS_HOST

}  // namespace pir
