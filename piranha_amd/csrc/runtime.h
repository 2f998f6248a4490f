// runtime.h -- execution context: stream, caching device allocator, mode.
//
// The reference allocates a fresh thrust::device_vector for every temporary
// and synchronises after every kernel (DeviceData.h:250-254, gemm.cuh:99).
// Here every temporary comes from a per-context caching allocator (exact-size
// free lists, never returned to HIP until the context dies), so a forward pass
// that ran once allocates nothing afterwards and can be captured in a hipGraph;
// work is only ordered on the context's stream, with no host synchronisation.
#pragma once

#include <map>
#include <unordered_map>
#include <vector>

#include "common.h"
#include "prg.h"

namespace pir {

enum ExecMode { MODE_FUSED = 0, MODE_PLAIN = 1 };
// Correlated randomness: the reference's placeholder constants
// (src/mpc/Precompute.h:73-134) or the masked mode's seeded Philox stream.
enum Precompute { PRECOMPUTE_PLACEHOLDER = 0, PRECOMPUTE_MASKED = 1 };
enum GemmAlgo { ALGO_PLAIN = 0, ALGO_LIMB = 1 };

// Freed blocks are cached per stream and handed out again only to requests on
// the same stream, whose order then keeps the block's earlier kernels ahead of
// its new ones; no block moves between streams while work may be pending.
class Allocator {
  public:
    ~Allocator();
    void *alloc(size_t bytes, hipStream_t stream);
    void release(void *p, hipStream_t stream);
    // Returns the cached free blocks of `stream` to HIP (the caller has
    // synchronised the stream, or is about to destroy it after doing so).
    void trim_stream(hipStream_t stream);
    // Returns every cached free block to HIP (after a device synchronisation).
    void trim_all();
    size_t bytes_reserved() const { return reserved_; }
    size_t bytes_cached() const;
    bool capturing = false;  // set while a hipGraph capture is open: no new hipMalloc allowed

  private:
    std::multimap<std::pair<hipStream_t, size_t>, void *> free_;
    std::unordered_map<void *, size_t> used_;
    size_t reserved_ = 0;
};

// Per-launch HIP events around the limb GEMM kernel, recorded only while `on`
// (an eager profiling forward; never during graph capture).
struct GemmTimer {
    struct Rec {
        hipEvent_t start, stop;
        double ops;   // algorithmic int8 ops of the launch: 2 * 36 * M * N * nseg * K
        int variant;  // A-operand form (0 dense row-major, 1 dense col-major, 2 conv, 3 conv Cin%32==0,
                      // 4 dgrad, 5 dgrad Cin%32==0, 6 wgrad), + 8 for the 32x128 workgroup tile
    };
    bool on = false;
    std::vector<Rec> recs;
};

struct Context {
    hipStream_t stream = nullptr;
    GemmTimer timer;
    bool tune = false;  // autotune untuned limb GEMM shapes on their first eager launch
    Allocator alloc;
    int mode = MODE_FUSED;
    int algo = ALGO_LIMB;
    int frac_bits = 26;  // FLOAT_PRECISION (globals.h:19-21), a runtime value here
    int log_lr = 5;      // log_learning_rate (main.cu:48), weight updates divide by 2^log_lr
    int error = 0;
    int precompute = PRECOMPUTE_PLACEHOLDER;
    u64 prg_seed = 0;
    u32 prg_call = 0;  // call number of the next masked protocol call
};

// The word stream of one protocol call that uses correlated randomness: in the
// masked precompute mode the next call number of the current context (the
// counter advances; the fused form only, anything else is refused with
// `what` in the message); a stream that is off in the placeholder mode.
PrgStream call_prg(const char *what);
// Refuses the masked precompute mode for a function that has no masked form.
void require_placeholder(const char *what);

// The context used by the C++ protocol API (the reference uses globals:
// partyNum, PrecomputeObject, piranha_config; main.cu:23-40).
Context &ctx();
void set_ctx(Context *c);

// RAII device buffer from the context allocator.  The buffer remembers the
// allocator and stream it came from and returns the block to exactly those,
// whichever context is current when it is destroyed.
class DevBuf {
  public:
    DevBuf() = default;
    explicit DevBuf(size_t bytes) : bytes_(bytes) {
        if (bytes) {
            alloc_ = &ctx().alloc;
            stream_ = ctx().stream;
            p_ = alloc_->alloc(bytes, stream_);
        }
    }
    ~DevBuf() { reset(); }
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : p_(o.p_), bytes_(o.bytes_), alloc_(o.alloc_), stream_(o.stream_) {
        o.p_ = nullptr;
        o.bytes_ = 0;
        o.alloc_ = nullptr;
    }
    DevBuf &operator=(DevBuf &&o) noexcept {
        if (this != &o) {
            reset();
            p_ = o.p_;
            bytes_ = o.bytes_;
            alloc_ = o.alloc_;
            stream_ = o.stream_;
            o.p_ = nullptr;
            o.bytes_ = 0;
            o.alloc_ = nullptr;
        }
        return *this;
    }
    void *get() const { return p_; }
    size_t bytes() const { return bytes_; }

  private:
    void reset() {
        if (p_ && alloc_) alloc_->release(p_, stream_);
        p_ = nullptr;
        bytes_ = 0;
        alloc_ = nullptr;
    }
    void *p_ = nullptr;
    size_t bytes_ = 0;
    Allocator *alloc_ = nullptr;
    hipStream_t stream_ = nullptr;
};

// Host-side failure inside the C++ protocol API (which mirrors the reference's
// void-returning functions): recorded in the context, reported by the C ABI.
class PirError : public std::exception {
  public:
    explicit PirError(int code) : code_(code) {}
    int code() const { return code_; }
    const char *what() const noexcept override { return last_error(); }

  private:
    int code_;
};

inline void check(int status) {
    if (status != 0) throw PirError(status);
}

inline void hip_check(hipError_t e) {
    if (e != hipSuccess) {
        set_error(std::string("hip: ") + hipGetErrorString(e));
        throw PirError(ERR_HIP);
    }
}

}  // namespace pir
