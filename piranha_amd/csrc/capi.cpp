// capi.cpp -- extern "C" entry points of libpir_hip.so (declared in include/pir.h).
#include "../../include/pir.h"

#include <atomic>
#include <memory>

#include "elementwise.h"
#include "fused_rss.h"
#include "layers.h"
#include "model_io.h"
#include "protocol.h"
#include "ring_gemm.h"
#include "tpc_fpc_compare.h"
#include "tpc_fpc.h"

using namespace pir;

static_assert(PIR_LIMB_VARIANTS == LIMB_VARIANTS && PIR_LIMB_WIDE == LIMB_VARIANT_WIDE,
              "kernel variant numbering of include/pir.h and ring_gemm.h");
static_assert(PIR_CONV_FPROP == kFprop && PIR_CONV_DGRAD == kDgrad && PIR_CONV_WGRAD == kWgrad,
              "convolution operator numbering");

namespace {

bool overlaps(const void *a, size_t abytes, const void *b, size_t bbytes) {
    const char *x = (const char *)a, *y = (const char *)b;
    return x < y + bbytes && y < x + abytes;
}

// Execution form of the free-standing pir_rss_* calls, process-wide (read by
// every call on any thread, written by pir_set_mode).
std::atomic<int> g_mode{MODE_FUSED};

struct Scope {
    explicit Scope(void *stream, int algo = ALGO_LIMB) {
        Context &c = ctx();
        c.stream = (hipStream_t)stream;
        c.mode = g_mode.load(std::memory_order_relaxed);
        c.algo = algo;
    }
};

// The TPC and FPC protocols keep the reference's placeholder precompute: they
// refuse the masked mode of the calling thread.
struct SchemeScope : Scope {
    explicit SchemeScope(void *stream, int algo = ALGO_LIMB) : Scope(stream, algo) {
        require_placeholder("TPC / FPC protocols");
    }
};

template <typename F>
int guarded(F &&f) {
    try {
        return f();
    } catch (const PirError &e) {
        return e.code();
    } catch (const std::exception &e) {
        set_error(e.what());
        return ERR_STATE;
    }
}

}  // namespace

struct pir_net {
    Context ctx;
    std::unique_ptr<NeuralNetwork> net;
    hipGraph_t graph = nullptr;  // captured forward
    hipGraphExec_t exec = nullptr;
    hipGraph_t tgraph = nullptr;  // captured training step (forward + backward)
    hipGraphExec_t texec = nullptr;
    RSS64 labels;  // label shares of the training-step API
    std::vector<LayerSpec> spec;
    bool own_stream = false;
};

namespace {
struct NetScope {
    explicit NetScope(pir_net *n) { set_ctx(&n->ctx); }
    ~NetScope() { set_ctx(nullptr); }
};

// A captured forward holds the weights packed at capture time: any weight
// change (set_weights, a backward pass) drops it, and replay then reports that
// the network must be captured again.
void drop_graph(pir_net *n) {
    for (hipGraphExec_t *e : {&n->exec, &n->texec})
        if (*e) {
            (void)hipGraphExecDestroy(*e);
            *e = nullptr;
        }
    for (hipGraph_t *g : {&n->graph, &n->tgraph})
        if (*g) {
            (void)hipGraphDestroy(*g);
            *g = nullptr;
        }
}

// Runs `body` with HIP events around every limb GEMM launch and sums the
// kernel times per variant (the layout of pir_net_profile_gemm).
template <typename F>
int timed_gemms(pir_net *n, double *ms, double *ops, int *launches, F &&body) {
    GemmTimer &t = n->ctx.timer;
    t.recs.clear();
    t.on = true;
    int rc = guarded(body);
    t.on = false;
    hipError_t e = hipStreamSynchronize(n->ctx.stream);
    for (int v = 0; v < PIR_LIMB_VARIANTS; v++) {
        ms[v] = 0;
        ops[v] = 0;
        launches[v] = 0;
    }
    for (auto &r : t.recs) {
        float dt = 0;
        if (rc == OK && e == hipSuccess && hipEventElapsedTime(&dt, r.start, r.stop) == hipSuccess) {
            ms[r.variant] += dt;
            ops[r.variant] += r.ops;
            launches[r.variant]++;
        }
        (void)hipEventDestroy(r.start);
        (void)hipEventDestroy(r.stop);
    }
    t.recs.clear();
    if (rc) return rc;
    PIR_HIP(e);
    return (int)OK;
}

void require_train_labels(pir_net *n) {
    if (n->net->layers.empty() || n->labels.size() != n->net->layers.back()->getActivation()->size()) {
        set_error("training step: set the labels first (pir_net_set_train_labels)");
        throw PirError(ERR_STATE);
    }
}

// one training step from the network's input buffer and label shares
void train_step(pir_net *n) {
    require_train_labels(n);
    n->net->run();
    n->net->backwardShares(n->labels);
}
}  // namespace

extern "C" {

int pir_version(void) { return 1; }
const char *pir_last_error(void) { return last_error(); }

int pir_malloc(void **ptr, size_t bytes) {
    PIR_HIP(hipMalloc(ptr, bytes ? bytes : 1));
    return OK;
}
int pir_free(void *ptr) {
    PIR_HIP(hipFree(ptr));
    return OK;
}
int pir_memcpy_h2d(void *dst, const void *src, size_t bytes, void *stream) {
    PIR_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, (hipStream_t)stream));
    PIR_HIP(hipStreamSynchronize((hipStream_t)stream));
    return OK;
}
int pir_memcpy_d2h(void *dst, const void *src, size_t bytes, void *stream) {
    PIR_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, (hipStream_t)stream));
    PIR_HIP(hipStreamSynchronize((hipStream_t)stream));
    return OK;
}
int pir_memcpy_d2d(void *dst, const void *src, size_t bytes, void *stream) {
    PIR_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream));
    return OK;
}
int pir_fill_u64(uint64_t *dst, uint64_t value, size_t n, void *stream) {
    return fill_u64(dst, value, n, (hipStream_t)stream);
}
int pir_fill_u8(uint8_t *dst, uint8_t value, size_t n, void *stream) { return fill_u8(dst, value, n, (hipStream_t)stream); }
int pir_stream_create(void **stream) {
    hipStream_t s;
    PIR_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = s;
    return OK;
}
int pir_stream_destroy(void *stream) {
    // the calling thread's cached temporaries of this stream can never be
    // handed out again: return them to HIP once the stream's work is done
    PIR_HIP(hipStreamSynchronize((hipStream_t)stream));
    ctx().alloc.trim_stream((hipStream_t)stream);
    PIR_HIP(hipStreamDestroy((hipStream_t)stream));
    return OK;
}
int pir_release_cached(void) {
    PIR_HIP(hipDeviceSynchronize());
    ctx().alloc.trim_all();
    return OK;
}
size_t pir_cached_bytes(void) { return ctx().alloc.bytes_cached(); }
int pir_stream_synchronize(void *stream) {
    PIR_HIP(hipStreamSynchronize((hipStream_t)stream));
    return OK;
}
int pir_device_synchronize(void) {
    PIR_HIP(hipDeviceSynchronize());
    return OK;
}
int pir_set_mode(int mode) {
    PIR_REQUIRE(mode == MODE_FUSED || mode == MODE_PLAIN, "pir_set_mode: unknown mode");
    g_mode.store(mode, std::memory_order_relaxed);
    return OK;
}

// ---- ring GEMM -------------------------------------------------------------

size_t pir_limb_packed_bytes(int K, int N, int nseg) { return limb_b_bytes(N, nseg, round_up(K, LIMB_BK)); }

int pir_limb_pack(const uint64_t *const *b0, const uint64_t *const *b1, int nseg, int K, int N, int row_major,
                  int8_t *packed, void *stream) {
    PIR_REQUIRE(nseg >= 1 && nseg <= LIMB_MAX_SEG && K > 0 && N > 0, "pir_limb_pack: invalid shape");
    LimbBSrc s{};
    for (int i = 0; i < nseg; i++) {
        s.b[i][0] = b0[i];
        s.b[i][1] = b1 ? b1[i] : nullptr;
        s.coef[i][0] = s.coef[i][1] = 1;
    }
    s.nseg = nseg;
    s.K = K;
    s.N = N;
    s.Kp = round_up(K, LIMB_BK);
    s.row_major = row_major;
    return limb_pack_b(s, packed, (hipStream_t)stream);
}

int pir_limb_gemm(const uint64_t *const *a, int nseg, int M, int K, int a_row_major, const int8_t *packed, int N,
                  uint64_t *C, int tc, int split, void *stream) {
    return guarded([&]() -> int {
        Scope sc(stream);
        PIR_REQUIRE(nseg >= 1 && nseg <= LIMB_MAX_SEG && M >= 0 && K > 0 && N >= 0, "pir_limb_gemm: invalid shape");
        LimbA la{};
        for (int i = 0; i < nseg; i++) la.seg[i] = a[i];
        la.nseg = nseg;
        la.M = M;
        la.K = K;
        la.Kp = round_up(K, LIMB_BK);
        la.lda = a_row_major ? K : M;
        la.row_major = a_row_major;
        Epilogue e{};
        e.mode = EPI_RAW;
        e.C = C;
        e.tc = tc;
        return limb_run(la, packed, N, e, split, (hipStream_t)stream);
    });
}

int pir_gemm(int M, int N, int K, const uint64_t *A, int ta, const uint64_t *B, int tb, uint64_t *C, int tc, int algo,
             void *stream) {
    if (algo == PIR_ALGO_PLAIN)
        return guarded([&]() -> int {
            Scope sc(stream, ALGO_PLAIN);
            return plain_gemm(M, N, K, A, ta, B, tb, C, tc, (hipStream_t)stream);
        });
    return guarded([&]() -> int {
        Scope sc(stream);
        return limb_ring_gemm(M, N, K, A, ta != 0, B, tb != 0, C, tc != 0, (hipStream_t)stream);
    });
}

int pir_conv_fprop(const uint64_t *x, const uint64_t *w, uint64_t *y, int b, int H, int W, int Din, int Dout, int fh,
                   int fw, int ph, int pw, int stride, int dilation, int algo, void *stream) {
    if (algo == PIR_ALGO_PLAIN)
        return guarded([&]() -> int {
            Scope sc(stream, ALGO_PLAIN);
            return plain_conv_fprop(x, w, y, b, H, W, Din, Dout, fh, fw, ph, pw, stride, dilation,
                                    (hipStream_t)stream);
        });
    return guarded([&]() -> int {
        Scope sc(stream);
        PIR_REQUIRE(dilation == 1, "pir_conv_fprop: the limb form supports dilation 1");
        return limb_ring_conv(x, w, y, b, H, W, Din, Dout, fh, fw, ph, pw, stride, (hipStream_t)stream);
    });
}

int pir_conv_dgrad(const uint64_t *dy, const uint64_t *w, uint64_t *dx, int b, int OH, int OW, int Dout, int fh,
                   int fw, int Din, int ph, int pw, int stride, int dilation, int H, int W, int algo, void *stream) {
    if (algo == PIR_ALGO_PLAIN)
        return guarded([&]() -> int {
            Scope sc(stream, ALGO_PLAIN);
            return plain_conv_dgrad(dy, w, dx, b, OH, OW, Dout, fh, fw, Din, ph, pw, stride, dilation, H, W,
                                    (hipStream_t)stream);
        });
    return guarded([&]() -> int {
        Scope sc(stream);
        PIR_REQUIRE(dilation == 1, "pir_conv_dgrad: the limb form supports dilation 1");
        return limb_ring_dgrad(dy, w, dx, b, OH, OW, Dout, fh, fw, Din, ph, pw, stride, H, W, (hipStream_t)stream);
    });
}

int pir_conv_wgrad(const uint64_t *dy, const uint64_t *x, uint64_t *dw, int b, int OH, int OW, int Dout, int H,
                   int W, int Din, int fh, int fw, int ph, int pw, int stride, int dilation, int algo, void *stream) {
    if (algo == PIR_ALGO_PLAIN)
        return guarded([&]() -> int {
            Scope sc(stream, ALGO_PLAIN);
            return plain_conv_wgrad(dy, x, dw, b, OH, OW, Dout, H, W, Din, fh, fw, ph, pw, stride, dilation,
                                    (hipStream_t)stream);
        });
    return guarded([&]() -> int {
        Scope sc(stream);
        PIR_REQUIRE(dilation == 1, "pir_conv_wgrad: the limb form supports dilation 1");
        return limb_ring_wgrad(dy, x, dw, b, OH, OW, Dout, H, W, Din, fh, fw, ph, pw, stride, (hipStream_t)stream);
    });
}

int pir_reduce_sum(const uint64_t *a, uint64_t *b, int reduce_rows, int rows, int cols, void *stream) {
    return reduce_sum(a, b, reduce_rows, rows, cols, (hipStream_t)stream);
}

int pir_vector_expand(const uint64_t *a, size_t n, uint64_t *b, int factor, void *stream) {
    return vector_expand(a, n, b, factor, (hipStream_t)stream);
}

int pir_maxpool_row2im(const uint64_t *rows, uint64_t *out, int image_height, int image_width, int f, int Din,
                       int batch, int stride, void *stream) {
    return maxpool_row2im(rows, out, image_height, image_width, f, Din, batch, stride, (hipStream_t)stream);
}

int pir_ring_gemm3(int algo, int M, int N, int K, const uint64_t *A0, const uint64_t *A1, const uint64_t *B0,
                   const uint64_t *B1, uint64_t *C, void *stream) {
    return guarded([&]() -> int {
        Scope sc(stream);
        PIR_REQUIRE(M > 0 && N > 0 && K > 0, "pir_ring_gemm3: invalid shape");
        hipStream_t s = (hipStream_t)stream;
        if (algo == PIR_ALGO_PLAIN) {
            // localMatMul's form (RSS.inl:1096-1110): three gemms and two accumulations
            DevBuf acc((size_t)M * N * 8);
            PIR_TRY(plain_gemm(M, N, K, A0, true, B0, true, C, true, s));
            PIR_TRY(plain_gemm(M, N, K, A0, true, B1, true, (u64 *)acc.get(), true, s));
            PIR_TRY(dd_op(OP_ADD, 8, C, 1, acc.get(), 1, 0, (size_t)M * N, s));
            PIR_TRY(plain_gemm(M, N, K, A1, true, B0, true, (u64 *)acc.get(), true, s));
            return dd_op(OP_ADD, 8, C, 1, acc.get(), 1, 0, (size_t)M * N, s);
        }
        PIR_REQUIRE(algo == PIR_ALGO_LIMB, "pir_ring_gemm3: unknown algorithm");
        DevBuf packed(pir_limb_packed_bytes(K, N, 2));
        const uint64_t *b0[2] = {B0, B0}, *b1[2] = {B1, nullptr};
        PIR_TRY(pir_limb_pack(b0, b1, 2, K, N, 1, (int8_t *)packed.get(), stream));
        const uint64_t *a[2] = {A0, A1};
        return pir_limb_gemm(a, 2, M, K, 1, (const int8_t *)packed.get(), N, C, 1, 0, stream);
    });
}

// ---- masked precompute mode and the seeded PRG --------------------------------

int pir_set_precompute(int mode, uint64_t seed, uint32_t call) {
    PIR_REQUIRE(mode == PRECOMPUTE_PLACEHOLDER || mode == PRECOMPUTE_MASKED, "pir_set_precompute: unknown mode");
    Context &c = ctx();
    c.precompute = mode;
    c.prg_seed = seed;
    c.prg_call = call;
    return OK;
}
int pir_get_prg_call(uint32_t *call) {
    PIR_REQUIRE(call, "pir_get_prg_call: null pointer");
    *call = ctx().prg_call;
    return OK;
}
int pir_prg_philox(const uint32_t *ctr, const uint32_t *key, size_t n, uint32_t *out, void *stream) {
    return prg_philox(ctr, key, n, out, (hipStream_t)stream);
}
int pir_prg_fill(uint64_t seed, uint32_t call, size_t n, uint64_t *out, void *stream) {
    return prg_fill(seed, call, n, out, (hipStream_t)stream);
}
int pir_rss_deal(const uint64_t *values, size_t n, uint64_t seed, uint32_t call, uint64_t *shares, void *stream) {
    PIR_REQUIRE(values, "pir_rss_deal: null values");
    return rss_deal(values, nullptr, n, seed, call, shares, (hipStream_t)stream);
}
int pir_rss_rerandomize(const uint64_t *in, size_t n, uint64_t seed, uint32_t call, uint64_t *out, void *stream) {
    PIR_REQUIRE(in, "pir_rss_rerandomize: null shares");
    return rss_deal(nullptr, in, n, seed, call, out, (hipStream_t)stream);
}

// ---- bit / layout kernels ----------------------------------------------------

int pir_bitexpand(const uint64_t *a, size_t n, void *bits, int eb, void *stream) {
    return bitexpand(a, n, bits, eb, (hipStream_t)stream);
}
int pir_set_carry_out_msb(void *rbits, void *abits, void *msb, size_t n, int bw, int xr, int eb, void *stream) {
    return set_carry_out_msb(rbits, abits, msb, n, bw, xr, eb, (hipStream_t)stream);
}
int pir_expand_compare(const void *b, const void *inv, void *out, size_t n, size_t out_n, int eb, void *stream) {
    return expand_compare(b, inv, out, n, out_n, eb, (hipStream_t)stream);
}
int pir_element_vector_add(uint64_t *a, const uint64_t *b, int rowwise, int rows, int cols, void *stream) {
    return element_vector_add(a, b, rowwise, rows, cols, (hipStream_t)stream);
}
int pir_truncate_cols(const void *a, size_t a_n, void *b, int sc, int ec, int eb, void *stream) {
    return truncate_cols(a, a_n, b, sc, ec, eb, (hipStream_t)stream);
}
int pir_stride_pad(const uint64_t *in, uint64_t *out, size_t n_in, int stride, int pad, uint64_t pv, void *stream) {
    return stride_pad(in, out, n_in, stride, pad, pv, (hipStream_t)stream);
}
size_t pir_maxpool_im2row_size(int H, int W, int f, int Din, int batch, int stride, int padding) {
    return maxpool_im2row_size(H, W, f, Din, batch, stride, padding);
}
int pir_maxpool_im2row(const uint64_t *im, uint64_t *out, int H, int W, int f, int Din, int batch, int stride,
                       int padding, uint64_t pv, void *stream) {
    return maxpool_im2row(im, out, H, W, f, Din, batch, stride, padding, pv, (hipStream_t)stream);
}
int pir_transpose(const void *a, void *b, int rows, int cols, int elem_bytes, void *stream) {
    return transpose(a, b, rows, cols, elem_bytes, (hipStream_t)stream);
}
int pir_get_msbs(const void *bits, void *msbs, size_t n_msbs, size_t bit_width, int elem_bytes, void *stream) {
    return get_msbs(bits, msbs, n_msbs, bit_width, elem_bytes, (hipStream_t)stream);
}
int pir_set_msbs(void *bits, size_t n_bits, uint64_t val, size_t bit_width, int elem_bytes, void *stream) {
    return set_msbs(bits, n_bits, val, bit_width, elem_bytes, (hipStream_t)stream);
}
size_t pir_averagepool_im2row_size(int image_width, int image_height, int f, int Din, int batch, int stride,
                                   int padding) {
    return averagepool_im2row_size(image_width, image_height, f, Din, batch, stride, padding);
}
int pir_averagepool_im2row(const uint64_t *im, uint64_t *out, int image_width, int image_height, int f, int Din,
                           int batch, int stride, int padding, void *stream) {
    return averagepool_im2row(im, out, image_width, image_height, f, Din, batch, stride, padding, (hipStream_t)stream);
}
int pir_averagepool_expand_delta(const uint64_t *in, size_t n_in, uint64_t *out, size_t out_capacity, int Din,
                                 int pool_size, void *stream) {
    return averagepool_expand_delta(in, n_in, out, out_capacity, Din, pool_size, (hipStream_t)stream);
}
int pir_dd_op(int op, int eb, void *x, long long xs, const void *y, long long ys, uint64_t scalar, size_t n,
              void *stream) {
    PIR_REQUIRE(op >= 0 && op <= PIR_OP_DIV, "pir_dd_op: unknown op");
    return dd_op(op, eb, x, xs, y, ys, scalar, n, (hipStream_t)stream);
}

// ---- RSS protocol ------------------------------------------------------------

int pir_rss_reconstruct(const void *x, size_t n, void *out, int eb, void *stream) {
    return guarded([&]() -> int {
        Scope sc(stream);
        if (eb == 8) {
            PartyData<u64> o;
            RSS<u64> in = RSS<u64>::wrap((u64 *)x, n);
            PartyData<u64> ov(DeviceData<u64>((u64 *)out, n), DeviceData<u64>((u64 *)out + n, n),
                              DeviceData<u64>((u64 *)out + 2 * n, n));
            if (g_mode == MODE_FUSED) return fused_reconstruct<u64>((const u64 *)x, n, (u64 *)out, (hipStream_t)stream);
            reconstruct(in, ov);
            return (int)OK;
        }
        PIR_REQUIRE(eb == 1, "pir_rss_reconstruct: element size must be 1 or 8");
        if (g_mode == MODE_FUSED) return fused_reconstruct<u8>((const u8 *)x, n, (u8 *)out, (hipStream_t)stream);
        RSS<u8> in = RSS<u8>::wrap((u8 *)x, n);
        PartyData<u8> ov(DeviceData<u8>((u8 *)out, n), DeviceData<u8>((u8 *)out + n, n),
                         DeviceData<u8>((u8 *)out + 2 * n, n));
        reconstruct(in, ov);
        return (int)OK;
    });
}

int pir_rss_reshare(const uint64_t *c, size_t n, uint64_t *out, void *stream) {
    return guarded([&]() -> int {
        Scope sc(stream);
        if (g_mode == MODE_FUSED) return fused_reshare(c, n, out, (hipStream_t)stream, call_prg("reshare"));
        PartyData<u64> cv(DeviceData<u64>((u64 *)c, n), DeviceData<u64>((u64 *)c + n, n),
                          DeviceData<u64>((u64 *)c + 2 * n, n));
        RSS<u64> o = RSS<u64>::wrap(out, n);
        reshare(cv, o);
        return (int)OK;
    });
}

int pir_rss_mult(uint64_t *x, const uint64_t *y, size_t n, void *stream) {
    return guarded([&]() -> int {
        Scope sc(stream);
        RSS<u64> a = RSS<u64>::wrap(x, n), b = RSS<u64>::wrap((u64 *)y, n);
        a *= b;
        return (int)OK;
    });
}

int pir_rss_and(void *x, const void *y, size_t n, int eb, void *stream) {
    return guarded([&]() -> int {
        Scope sc(stream);
        if (eb == 8) {
            RSS<u64> a = RSS<u64>::wrap((u64 *)x, n), b = RSS<u64>::wrap((u64 *)y, n);
            a &= b;
        } else {
            PIR_REQUIRE(eb == 1, "pir_rss_and: element size must be 1 or 8");
            RSS<u8> a = RSS<u8>::wrap((u8 *)x, n), b = RSS<u8>::wrap((u8 *)y, n);
            a &= b;
        }
        return (int)OK;
    });
}

int pir_rss_divide_public(uint64_t *a, size_t n, uint64_t d, const uint64_t *d_local, void *stream) {
    return guarded([&]() -> int {
        Scope sc(stream);
        RSS<u64> x = RSS<u64>::wrap(a, n);
        if (d_local) {
            PartyData<u64> dv(DeviceData<u64>((u64 *)d_local, n), DeviceData<u64>((u64 *)d_local + n, n),
                              DeviceData<u64>((u64 *)d_local + 2 * n, n));
            if (g_mode == MODE_FUSED) {
                // the fused kernel reads the contiguous [3][n] denominator block
                require_placeholder("dividePublic by per-party denominators");
                PIR_TRY(fused_divide_public(a, n, 0, d_local, (hipStream_t)stream));
                return (int)OK;
            }
            dividePublic(x, dv);
        } else {
            dividePublic(x, d);
        }
        return (int)OK;
    });
}

int pir_rss_divide_public_3out3(const uint64_t *a, size_t n, int k, uint64_t *out, void *stream) {
    return guarded([&]() -> int {
        Scope sc(stream);
        PIR_REQUIRE(k >= 0 && k <= 62, "dividePublic3out3: k must be in [0, 62]");
        if (g_mode == MODE_FUSED)
            return fused_divide_3out3(a, n, k, out, (hipStream_t)stream, call_prg("dividePublic3out3"));
        // the op-by-op form modifies its input (RSS.inl:1085): work on a copy
        PartyData<u64> av(n);
        PIR_HIP(hipMemcpyAsync(av.base(), a, 3 * n * 8, hipMemcpyDeviceToDevice, (hipStream_t)stream));
        RSS<u64> o = RSS<u64>::wrap(out, n);
        dividePublic3out3(av, (u64)1 << k, o);
        return (int)OK;
    });
}

int pir_rss_matmul(const uint64_t *a, const uint64_t *b, uint64_t *c, int M, int N, int K, int ta, int tb, int tc,
                   int trunc, int algo, void *stream) {
    return guarded([&]() -> int {
        Scope sc(stream, algo == PIR_ALGO_PLAIN ? ALGO_PLAIN : ALGO_LIMB);
        PIR_REQUIRE(M > 0 && N > 0 && K > 0 && trunc >= 0 && trunc <= 62, "pir_rss_matmul: invalid arguments");
        RSS<u64> A = RSS<u64>::wrap((u64 *)a, (size_t)M * K), B = RSS<u64>::wrap((u64 *)b, (size_t)K * N);
        RSS<u64> C = RSS<u64>::wrap(c, (size_t)M * N);
        matmul(A, B, C, M, N, K, ta != 0, tb != 0, tc != 0, (u64)trunc);
        return (int)OK;
    });
}

int pir_rss_convolution(const uint64_t *a, const uint64_t *w, uint64_t *c, int batch, int H, int W, int f, int Din,
                        int Dout, int stride, int pad, int trunc, int algo, void *stream) {
    return guarded([&]() -> int {
        Scope sc(stream, algo == PIR_ALGO_PLAIN ? ALGO_PLAIN : ALGO_LIMB);
        PIR_REQUIRE(batch > 0 && H > 0 && W > 0 && f > 0 && Din > 0 && Dout > 0 && stride > 0 && pad >= 0,
                    "pir_rss_convolution: invalid geometry");
        int OH = (H + 2 * pad - f) / stride + 1, OW = (W + 2 * pad - f) / stride + 1;
        PIR_REQUIRE(OH > 0 && OW > 0, "pir_rss_convolution: empty output");
        RSS<u64> A = RSS<u64>::wrap((u64 *)a, (size_t)batch * H * W * Din);
        RSS<u64> B = RSS<u64>::wrap((u64 *)w, (size_t)Dout * f * f * Din);
        RSS<u64> C = RSS<u64>::wrap(c, (size_t)batch * OH * OW * Dout);
        convolution(A, B, C, kFprop, batch, H, W, f, Din, Dout, stride, pad, trunc);
        return (int)OK;
    });
}

int pir_rss_convolution_op(int op, const uint64_t *a, const uint64_t *b, uint64_t *c, int batch, int H, int W, int f,
                           int Din, int Dout, int stride, int pad, int trunc, int algo, void *stream) {
    return guarded([&]() -> int {
        Scope sc(stream, algo == PIR_ALGO_PLAIN ? ALGO_PLAIN : ALGO_LIMB);
        PIR_REQUIRE(op == PIR_CONV_FPROP || op == PIR_CONV_DGRAD || op == PIR_CONV_WGRAD,
                    "pir_rss_convolution_op: unknown operator");
        PIR_REQUIRE(batch > 0 && H > 0 && W > 0 && f > 0 && Din > 0 && Dout > 0 && stride > 0 && pad >= 0,
                    "pir_rss_convolution_op: invalid geometry");
        PIR_REQUIRE(trunc >= 1 && trunc <= 62, "pir_rss_convolution_op: truncation must be in [1, 62]");
        int OH = (H + 2 * pad - f) / stride + 1, OW = (W + 2 * pad - f) / stride + 1;
        PIR_REQUIRE(H + 2 * pad >= f && W + 2 * pad >= f && OH > 0 && OW > 0, "pir_rss_convolution_op: empty output");
        size_t in_n = (size_t)batch * H * W * Din, w_n = (size_t)Dout * f * f * Din,
               out_n = (size_t)batch * OH * OW * Dout;
        size_t na = op == PIR_CONV_FPROP ? in_n : out_n;
        size_t nb = op == PIR_CONV_WGRAD ? in_n : w_n;
        size_t nc = op == PIR_CONV_FPROP ? out_n : (op == PIR_CONV_DGRAD ? in_n : w_n);
        RSS<u64> A = RSS<u64>::wrap((u64 *)a, na);
        RSS<u64> B = RSS<u64>::wrap((u64 *)b, nb);
        RSS<u64> C = RSS<u64>::wrap(c, nc);
        convolution(A, B, C, (ConvOperator)op, batch, H, W, f, Din, Dout, stride, pad, trunc);
        return (int)OK;
    });
}

// ---- TPC / FPC ----------------------------------------------------------------

int pir_tpc_reconstruct(const uint64_t *x, size_t n, uint64_t *out, void *stream) {
    return guarded([&]() -> int {
        SchemeScope sc(stream);
        tpc_reconstruct(x, n, out);
        return (int)OK;
    });
}

int pir_tpc_mult(uint64_t *x, const uint64_t *y, size_t n, void *stream) {
    return guarded([&]() -> int {
        SchemeScope sc(stream);
        return tpc_mult(x, y, n, (hipStream_t)stream);
    });
}

int pir_tpc_and(void *x, const void *y, size_t n, int eb, void *stream) {
    return guarded([&]() -> int {
        SchemeScope sc(stream);
        return tpc_and(x, y, n, eb, (hipStream_t)stream);
    });
}

int pir_tpc_drelu(const uint64_t *x, size_t n, void *result, int eb, void *stream) {
    return guarded([&]() -> int {
        SchemeScope sc(stream);
        return tpc_relu(x, n, result, nullptr, eb, (hipStream_t)stream);
    });
}

int pir_tpc_relu(const uint64_t *x, size_t n, uint64_t *y, void *dresult, int eb, void *stream) {
    return guarded([&]() -> int {
        SchemeScope sc(stream);
        PIR_REQUIRE(!overlaps(x, 16 * n, y, 16 * n), "pir_tpc_relu: y must not overlap x");
        return tpc_relu(x, n, dresult, y, eb, (hipStream_t)stream);
    });
}

int pir_tpc_select_share(const uint64_t *x, const uint64_t *y, const void *b, uint64_t *z, size_t n, int eb,
                         void *stream) {
    return guarded([&]() -> int {
        SchemeScope sc(stream);
        return tpc_select(x, y, b, z, n, eb, (hipStream_t)stream);
    });
}

int pir_tpc_maxpool(uint64_t *pools, size_t n, uint64_t *result, void *dresult, int k, int eb, void *stream) {
    return guarded([&]() -> int {
        SchemeScope sc(stream);
        return tpc_maxpool(pools, n, result, dresult, k, eb, (hipStream_t)stream);
    });
}

int pir_fpc_mult(uint64_t *x, const uint64_t *y, size_t n, void *stream) {
    return guarded([&]() -> int {
        SchemeScope sc(stream);
        return fpc_mult(x, y, n, (hipStream_t)stream);
    });
}

int pir_fpc_and(void *x, const void *y, size_t n, int eb, void *stream) {
    return guarded([&]() -> int {
        SchemeScope sc(stream);
        return fpc_and(x, y, n, eb, (hipStream_t)stream);
    });
}

int pir_fpc_drelu(const uint64_t *x, size_t n, void *result, int eb, void *stream) {
    return guarded([&]() -> int {
        SchemeScope sc(stream);
        return fpc_relu(x, n, result, nullptr, eb, (hipStream_t)stream);
    });
}

int pir_fpc_relu(const uint64_t *x, size_t n, uint64_t *y, void *dresult, int eb, void *stream) {
    return guarded([&]() -> int {
        SchemeScope sc(stream);
        PIR_REQUIRE(!overlaps(x, 96 * n, y, 96 * n), "pir_fpc_relu: y must not overlap x");
        return fpc_relu(x, n, dresult, y, eb, (hipStream_t)stream);
    });
}

int pir_fpc_select_share(const uint64_t *x, const uint64_t *y, const void *b, uint64_t *z, size_t n, int eb,
                         void *stream) {
    return guarded([&]() -> int {
        SchemeScope sc(stream);
        return fpc_select(x, y, b, z, n, eb, (hipStream_t)stream);
    });
}

int pir_fpc_maxpool(uint64_t *pools, size_t n, uint64_t *result, void *dresult, int k, int eb, void *stream) {
    return guarded([&]() -> int {
        SchemeScope sc(stream);
        return fpc_maxpool(pools, n, result, dresult, k, eb, (hipStream_t)stream);
    });
}

// getPowers / taylorSeries of TPC (P = 2, S = 1) and FPC (P = 4, S = 3) values [P][S][n]
namespace {
int scheme_get_powers_call(int P, int S, const uint64_t *x, size_t n, uint64_t *pow, void *stream, const char *what,
                           int (*f)(const u64 *, size_t, u64 *, hipStream_t)) {
    return guarded([&]() -> int {
        SchemeScope sc(stream);
        if (overlaps(x, (size_t)P * S * 8 * n, pow, (size_t)P * 8 * n)) {
            set_error(std::string(what) + ": pow must not overlap x");
            return (int)ERR_ARG;
        }
        return f(x, n, pow, (hipStream_t)stream);
    });
}
int scheme_taylor_call(int P, int S, const uint64_t *x, size_t n, uint64_t *out, double a0, double a1, double a2, int fn, int fp,
                       void *stream, const char *what,
                       int (*f)(const u64 *, size_t, u64 *, const TaylorTables &, hipStream_t)) {
    return guarded([&]() -> int {
        SchemeScope sc(stream);
        if (overlaps(x, (size_t)P * S * 8 * n, out, (size_t)P * S * 8 * n)) {
            set_error(std::string(what) + ": out must not overlap x");
            return (int)ERR_ARG;
        }
        const TaylorTables t = taylor_tables(a0, a1, a2, fn, fp);
        return f(x, n, out, t, (hipStream_t)stream);
    });
}
// RSS.inl:570-612, TPC.inl:615-642, FPC.inl:714-741
const double kTaylorCoeffs[3][3] = {{0.424, 0.584, 0.0}, {4.245, -5.857, 2.630}, {0.494286, 0.275589, -0.038751}};
}  // namespace

int pir_tpc_divide_public_local(uint64_t *a, size_t n, const uint64_t *d_local, void *stream) {
    return guarded([&]() -> int {
        SchemeScope sc(stream);
        return tpc_divide_public_vec(a, n, d_local, (hipStream_t)stream);
    });
}
int pir_fpc_divide_public_local(uint64_t *a, size_t n, const uint64_t *d_local, void *stream) {
    return guarded([&]() -> int {
        SchemeScope sc(stream);
        return fpc_divide_public_vec(a, n, d_local, (hipStream_t)stream);
    });
}

int pir_tpc_get_powers(const uint64_t *x, size_t n, uint64_t *pow, void *stream) {
    return scheme_get_powers_call(2, 1, x, n, pow, stream, "pir_tpc_get_powers", tpc_get_powers);
}
int pir_fpc_get_powers(const uint64_t *x, size_t n, uint64_t *pow, void *stream) {
    return scheme_get_powers_call(4, 3, x, n, pow, stream, "pir_fpc_get_powers", fpc_get_powers);
}
int pir_tpc_taylor_series(const uint64_t *x, size_t n, uint64_t *out, double a0, double a1, double a2, int fn, int fp,
                          void *stream) {
    return scheme_taylor_call(2, 1, x, n, out, a0, a1, a2, fn, fp, stream, "pir_tpc_taylor_series", tpc_taylor);
}
int pir_fpc_taylor_series(const uint64_t *x, size_t n, uint64_t *out, double a0, double a1, double a2, int fn, int fp,
                          void *stream) {
    return scheme_taylor_call(4, 3, x, n, out, a0, a1, a2, fn, fp, stream, "pir_fpc_taylor_series", fpc_taylor);
}
#define PIR_SCHEME_TAYLOR(scheme, P, S, name, FN)                                                             \
    int pir_##scheme##_##name(const uint64_t *x, size_t n, uint64_t *out, int fp, void *stream) {           \
        const double *c = kTaylorCoeffs[FN];                                                                 \
        return scheme_taylor_call(P, S, x, n, out, c[0], c[1], c[2], FN, fp, stream, "pir_" #scheme "_" #name, \
                                        scheme##_taylor);                                                    \
    }
PIR_SCHEME_TAYLOR(tpc, 2, 1, sqrt, TAYLOR_SQRT)
PIR_SCHEME_TAYLOR(tpc, 2, 1, inverse, TAYLOR_INVERSE)
PIR_SCHEME_TAYLOR(tpc, 2, 1, sigmoid, TAYLOR_SIGMOID)
PIR_SCHEME_TAYLOR(fpc, 4, 3, sqrt, TAYLOR_SQRT)
PIR_SCHEME_TAYLOR(fpc, 4, 3, inverse, TAYLOR_INVERSE)
PIR_SCHEME_TAYLOR(fpc, 4, 3, sigmoid, TAYLOR_SIGMOID)
#undef PIR_SCHEME_TAYLOR

int pir_tpc_divide_public(uint64_t *a, size_t n, uint64_t d, void *stream) {
    return guarded([&]() -> int {
        SchemeScope sc(stream);
        tpc_divide_public(a, n, d);
        return (int)OK;
    });
}

int pir_tpc_matmul(const uint64_t *a, const uint64_t *b, uint64_t *c, int M, int N, int K, int ta, int tb, int tc,
                   int trunc, int algo, void *stream) {
    return guarded([&]() -> int {
        SchemeScope sc(stream, algo == PIR_ALGO_PLAIN ? ALGO_PLAIN : ALGO_LIMB);
        PIR_REQUIRE(M > 0 && N > 0 && K > 0 && trunc >= 1 && trunc <= 62, "pir_tpc_matmul: invalid arguments");
        tpc_matmul(a, b, c, M, N, K, ta != 0, tb != 0, tc != 0, trunc);
        return (int)OK;
    });
}

int pir_tpc_convolution(const uint64_t *a, const uint64_t *w, uint64_t *c, int batch, int H, int W, int f, int Din,
                        int Dout, int stride, int pad, int trunc, int algo, void *stream) {
    return guarded([&]() -> int {
        SchemeScope sc(stream, algo == PIR_ALGO_PLAIN ? ALGO_PLAIN : ALGO_LIMB);
        PIR_REQUIRE(batch > 0 && H > 0 && W > 0 && f > 0 && Din > 0 && Dout > 0 && stride > 0 && pad >= 0 &&
                        trunc >= 1 && trunc <= 62 && (H + 2 * pad - f) >= 0 && (W + 2 * pad - f) >= 0,
                    "pir_tpc_convolution: invalid geometry");
        tpc_convolution(a, w, c, batch, H, W, f, Din, Dout, stride, pad, trunc);
        return (int)OK;
    });
}

int pir_fpc_reconstruct(const uint64_t *x, size_t n, uint64_t *out, void *stream) {
    return guarded([&]() -> int {
        SchemeScope sc(stream);
        fpc_reconstruct(x, n, out);
        return (int)OK;
    });
}

int pir_fpc_divide_public(uint64_t *a, size_t n, uint64_t d, void *stream) {
    return guarded([&]() -> int {
        SchemeScope sc(stream);
        fpc_divide_public(a, n, d);
        return (int)OK;
    });
}

int pir_fpc_reshare(const uint64_t *z, const uint64_t *zp, size_t n, int i, int j, uint64_t *out, void *stream) {
    return guarded([&]() -> int {
        SchemeScope sc(stream);
        fpc_reshare(z, zp, n, i, j, out);
        return (int)OK;
    });
}

int pir_fpc_matmul(const uint64_t *a, const uint64_t *b, uint64_t *c, int M, int N, int K, int ta, int tb, int tc,
                   int trunc, int algo, void *stream) {
    return guarded([&]() -> int {
        SchemeScope sc(stream, algo == PIR_ALGO_PLAIN ? ALGO_PLAIN : ALGO_LIMB);
        PIR_REQUIRE(M > 0 && N > 0 && K > 0 && trunc >= 1 && trunc <= 62, "pir_fpc_matmul: invalid arguments");
        fpc_matmul(a, b, c, M, N, K, ta != 0, tb != 0, tc != 0, trunc);
        return (int)OK;
    });
}

int pir_fpc_convolution(const uint64_t *a, const uint64_t *w, uint64_t *c, int batch, int H, int W, int f, int Din,
                        int Dout, int stride, int pad, int trunc, int algo, void *stream) {
    return guarded([&]() -> int {
        SchemeScope sc(stream, algo == PIR_ALGO_PLAIN ? ALGO_PLAIN : ALGO_LIMB);
        PIR_REQUIRE(batch > 0 && H > 0 && W > 0 && f > 0 && Din > 0 && Dout > 0 && stride > 0 && pad >= 0 &&
                        trunc >= 1 && trunc <= 62 && (H + 2 * pad - f) >= 0 && (W + 2 * pad - f) >= 0,
                    "pir_fpc_convolution: invalid geometry");
        fpc_convolution(a, w, c, batch, H, W, f, Din, Dout, stride, pad, trunc);
        return (int)OK;
    });
}

int pir_rss_drelu(const uint64_t *x, size_t n, void *result, int eb, void *stream) {
    return guarded([&]() -> int {
        Scope sc(stream);
        RSS<u64> in = RSS<u64>::wrap((u64 *)x, n);
        if (eb == 1) {
            RSS<u8> r = RSS<u8>::wrap((u8 *)result, n);
            dReLU(in, r);
        } else {
            PIR_REQUIRE(eb == 8, "pir_rss_drelu: element size must be 1 or 8");
            RSS<u64> r = RSS<u64>::wrap((u64 *)result, n);
            dReLU(in, r);
        }
        return (int)OK;
    });
}

int pir_rss_relu(const uint64_t *x, size_t n, uint64_t *y, void *dresult, int eb, void *stream) {
    return guarded([&]() -> int {
        Scope sc(stream);
        RSS<u64> in = RSS<u64>::wrap((u64 *)x, n), out = RSS<u64>::wrap(y, n);
        if (eb == 1) {
            RSS<u8> d = RSS<u8>::wrap((u8 *)dresult, n);
            ReLU(in, out, d);
        } else {
            PIR_REQUIRE(eb == 8, "pir_rss_relu: element size must be 1 or 8");
            RSS<u64> d = RSS<u64>::wrap((u64 *)dresult, n);
            ReLU(in, out, d);
        }
        return (int)OK;
    });
}

int pir_rss_select_share(const uint64_t *x, const uint64_t *y, const void *b, uint64_t *z, size_t n, int eb,
                         void *stream) {
    return guarded([&]() -> int {
        Scope sc(stream);
        RSS<u64> X = RSS<u64>::wrap((u64 *)x, n), Y = RSS<u64>::wrap((u64 *)y, n), Z = RSS<u64>::wrap(z, n);
        if (eb == 1) {
            RSS<u8> Bv = RSS<u8>::wrap((u8 *)b, n);
            selectShare(X, Y, Bv, Z);
        } else {
            PIR_REQUIRE(eb == 8, "pir_rss_select_share: element size must be 1 or 8");
            RSS<u64> Bv = RSS<u64>::wrap((u64 *)b, n);
            selectShare(X, Y, Bv, Z);
        }
        return (int)OK;
    });
}

int pir_rss_maxpool(uint64_t *pools, size_t n, uint64_t *result, void *dresult, int k, int eb, void *stream) {
    return guarded([&]() -> int {
        Scope sc(stream);
        PIR_REQUIRE(k >= 2 && (k & (k - 1)) == 0 && n % k == 0, "pir_rss_maxpool: k must be a power of two dividing n");
        RSS<u64> P = RSS<u64>::wrap(pools, n), R = RSS<u64>::wrap(result, n / k);
        if (eb == 1) {
            RSS<u8> D = RSS<u8>::wrap((u8 *)dresult, n);
            maxpool(P, R, D, k);
        } else {
            PIR_REQUIRE(eb == 8, "pir_rss_maxpool: element size must be 1 or 8");
            RSS<u64> D = RSS<u64>::wrap((u64 *)dresult, n);
            maxpool(P, R, D, k);
        }
        return (int)OK;
    });
}

int pir_rss_get_powers(const uint64_t *x, size_t n, uint64_t *pow, void *stream) {
    return guarded([&]() -> int {
        Scope sc(stream);
        PIR_REQUIRE(!overlaps(x, 6 * n * 8, pow, 3 * n * 8), "pir_rss_get_powers: pow must not overlap x");
        RSS<u64> X = RSS<u64>::wrap((u64 *)x, n);
        PartyData<u64> pw = PartyData<u64>::wrap(pow, n);
        getPowers(X, pw);
        return (int)OK;
    });
}

int pir_rss_taylor_series(const uint64_t *x, size_t n, uint64_t *out, double a0, double a1, double a2, int fn,
                          int fp, void *stream) {
    return guarded([&]() -> int {
        Scope sc(stream);
        PIR_REQUIRE(!overlaps(x, 6 * n * 8, out, 6 * n * 8), "pir_rss_taylor_series: out must not overlap x");
        RSS<u64> X = RSS<u64>::wrap((u64 *)x, n), Y = RSS<u64>::wrap(out, n);
        taylorSeries(X, Y, a0, a1, a2, fn, fp);
        return (int)OK;
    });
}

namespace {
int taylor_named(const uint64_t *x, size_t n, uint64_t *out, int fp, void *stream, const char *what,
                 void (*f)(const RSS<u64> &, RSS<u64> &, int)) {
    return guarded([&]() -> int {
        Scope sc(stream);
        if (overlaps(x, 6 * n * 8, out, 6 * n * 8)) {
            set_error(std::string(what) + ": out must not overlap x");
            return (int)ERR_ARG;
        }
        RSS<u64> X = RSS<u64>::wrap((u64 *)x, n), Y = RSS<u64>::wrap(out, n);
        f(X, Y, fp);
        return (int)OK;
    });
}
}  // namespace

int pir_rss_sqrt(const uint64_t *x, size_t n, uint64_t *out, int fp, void *stream) {
    return taylor_named(x, n, out, fp, stream, "pir_rss_sqrt", rss_sqrt);
}
int pir_rss_inverse(const uint64_t *x, size_t n, uint64_t *out, int fp, void *stream) {
    return taylor_named(x, n, out, fp, stream, "pir_rss_inverse", inverse);
}
int pir_rss_sigmoid(const uint64_t *x, size_t n, uint64_t *out, int fp, void *stream) {
    return taylor_named(x, n, out, fp, stream, "pir_rss_sigmoid", sigmoid);
}

// ---- network -------------------------------------------------------------------

namespace {
int create_net(pir_net **out, const std::vector<LayerSpec> &sp, int batch, int frac_bits, int seed, int mode,
               void *stream) {
    PIR_REQUIRE(frac_bits >= 1 && frac_bits <= 30, "pir_net_create: frac_bits must be in [1, 30]");
    PIR_REQUIRE(mode == MODE_FUSED || mode == MODE_PLAIN, "pir_net_create: unknown mode");
    pir_net *n = new pir_net();
    n->ctx.frac_bits = frac_bits;
    n->ctx.mode = mode;
    n->ctx.algo = mode == MODE_FUSED ? ALGO_LIMB : ALGO_PLAIN;
    n->spec = sp;
    if (stream) {
        n->ctx.stream = (hipStream_t)stream;
    } else {
        if (hipStreamCreateWithFlags(&n->ctx.stream, hipStreamNonBlocking) != hipSuccess) {
            delete n;
            set_error("pir_net_create: stream creation failed");
            return ERR_HIP;
        }
        n->own_stream = true;
    }
    int rc = guarded([&]() -> int {
        NetScope ns(n);
        n->net.reset(new NeuralNetwork(sp, batch, seed));
        return (int)OK;
    });
    if (rc) {
        pir_net_destroy(n);
        return rc;
    }
    *out = n;
    return OK;
}
}  // namespace

int pir_net_create(pir_net **out, const int *spec, int num_layers, int batch, int frac_bits, int seed, int mode,
                   void *stream) {
    PIR_REQUIRE(out && spec && num_layers > 0 && batch > 0, "pir_net_create: invalid arguments");
    std::vector<LayerSpec> sp(num_layers);
    for (int i = 0; i < num_layers; i++)
        for (int j = 0; j < 8; j++) sp[i].v[j] = spec[i * 8 + j];
    return create_net(out, sp, batch, frac_bits, seed, mode, stream);
}

int pir_model_read(const char *network_filename, int *input_size, int *num_classes, int *batch_size, int *num_layers,
                   int *spec, int max_layers) {
    PIR_REQUIRE(network_filename && num_layers && max_layers >= 0 && (spec || max_layers == 0),
                "pir_model_read: invalid arguments");
    return guarded([&]() -> int {
        ModelFile m = load_model_file(network_filename);
        if (input_size) *input_size = m.input_size;
        if (num_classes) *num_classes = m.num_classes;
        if (batch_size) *batch_size = m.batch_size;
        *num_layers = (int)m.layers.size();
        for (int i = 0; i < (int)m.layers.size() && i < max_layers; i++)
            for (int j = 0; j < 8; j++) spec[i * 8 + j] = m.layers[i].v[j];
        return (int)OK;
    });
}

int pir_net_load_model(pir_net **out, const char *network_filename, int batch, int frac_bits, int seed, int mode,
                       void *stream) {
    PIR_REQUIRE(out && network_filename, "pir_net_load_model: invalid arguments");
    ModelFile m;
    int rc = guarded([&]() -> int {
        m = load_model_file(network_filename);
        return (int)OK;
    });
    if (rc) return rc;
    PIR_REQUIRE(!m.layers.empty(), "pir_net_load_model: the model has no layers");
    if (batch <= 0) batch = m.batch_size;
    PIR_REQUIRE(batch > 0, "pir_net_load_model: the model's batch_size must be positive");
    return create_net(out, m.layers, batch, frac_bits, seed, mode, stream);
}

int pir_net_batch_size(const pir_net *n) { return n && n->net ? n->net->batch : 0; }

int pir_net_layer_spec(const pir_net *n, int layer, int *spec) {
    PIR_REQUIRE(n && spec && layer >= 0 && layer < (int)n->spec.size(), "pir_net_layer_spec: invalid arguments");
    for (int j = 0; j < 8; j++) spec[j] = n->spec[layer].v[j];
    return OK;
}

int pir_net_save_snapshot(pir_net *n, const char *path) {
    return guarded([&]() -> int {
        PIR_REQUIRE(n && path, "pir_net_save_snapshot: invalid arguments");
        NetScope ns(n);
        save_snapshot(*n->net, path);
        return (int)OK;
    });
}

int pir_net_load_snapshot(pir_net *n, const char *path) {
    return guarded([&]() -> int {
        PIR_REQUIRE(n && path, "pir_net_load_snapshot: invalid arguments");
        NetScope ns(n);
        load_snapshot(*n->net, path);
        drop_graph(n);
        return (int)OK;
    });
}

int pir_net_destroy(pir_net *n) {
    if (!n) return OK;
    {
        NetScope ns(n);
        (void)hipStreamSynchronize(n->ctx.stream);
        drop_graph(n);
        n->labels = RSS64();
        n->net.reset();
    }
    hipStream_t s = n->ctx.stream;
    bool own = n->own_stream;
    delete n;
    if (own) (void)hipStreamDestroy(s);
    return OK;
}

int pir_net_num_layers(const pir_net *n) { return n && n->net ? (int)n->net->layers.size() : 0; }

int pir_net_set_mode(pir_net *n, int mode) {
    PIR_REQUIRE(n && (mode == MODE_FUSED || mode == MODE_PLAIN), "pir_net_set_mode: invalid arguments");
    n->ctx.mode = mode;
    n->ctx.algo = mode == MODE_FUSED ? ALGO_LIMB : ALGO_PLAIN;
    return OK;
}

int pir_net_set_precompute(pir_net *n, int mode, uint64_t seed, uint32_t call) {
    PIR_REQUIRE(n && (mode == PRECOMPUTE_PLACEHOLDER || mode == PRECOMPUTE_MASKED),
                "pir_net_set_precompute: invalid arguments");
    n->ctx.precompute = mode;
    n->ctx.prg_seed = seed;
    n->ctx.prg_call = call;
    drop_graph(n);
    return OK;
}

int pir_net_set_weights(pir_net *n, int layer, const double *w, size_t nw, const double *b, size_t nb) {
    return guarded([&]() -> int {
        PIR_REQUIRE(n && layer >= 0 && layer < pir_net_num_layers(n), "pir_net_set_weights: bad layer");
        NetScope ns(n);
        Layer *L = n->net->layers[layer].get();
        PIR_REQUIRE(L->getWeights(), "pir_net_set_weights: layer has no weights");
        PIR_REQUIRE(nw == L->getWeights()->size(), "pir_net_set_weights: weight count mismatch");
        L->getWeights()->setPublic(std::vector<double>(w, w + nw), n->ctx.frac_bits);
        if (b) {
            PIR_REQUIRE(L->getBiases() && nb == L->getBiases()->size(), "pir_net_set_weights: bias count mismatch");
            L->getBiases()->setPublic(std::vector<double>(b, b + nb), n->ctx.frac_bits);
        }
        L->weightsChanged();
        drop_graph(n);
        return (int)OK;
    });
}

int pir_net_set_weight_shares(pir_net *n, int layer, const uint64_t *w, const uint64_t *b) {
    return guarded([&]() -> int {
        PIR_REQUIRE(n && w && layer >= 0 && layer < pir_net_num_layers(n), "pir_net_set_weight_shares: bad layer");
        NetScope ns(n);
        Layer *L = n->net->layers[layer].get();
        PIR_REQUIRE(L->getWeights(), "pir_net_set_weight_shares: layer has no weights");
        PIR_REQUIRE(!b || L->getBiases(), "pir_net_set_weight_shares: layer has no biases");
        L->getWeights()->copy_from_host(w);
        if (b) L->getBiases()->copy_from_host(b);
        L->weightsChanged();
        drop_graph(n);
        return (int)OK;
    });
}

int pir_net_set_log_learning_rate(pir_net *n, int log_lr) {
    PIR_REQUIRE(n && log_lr >= 1 && log_lr <= 30, "pir_net_set_log_learning_rate: log_lr must be in [1, 30]");
    n->ctx.log_lr = log_lr;
    return OK;
}

int pir_net_backward(pir_net *n, const double *labels, size_t count) {
    return guarded([&]() -> int {
        PIR_REQUIRE(n && labels, "pir_net_backward: invalid arguments");
        NetScope ns(n);
        drop_graph(n);
        n->net->backward(std::vector<double>(labels, labels + count));
        return (int)OK;
    });
}

int pir_net_backward_shares(pir_net *n, const uint64_t *label_shares) {
    return guarded([&]() -> int {
        PIR_REQUIRE(n && label_shares && !n->net->layers.empty(), "pir_net_backward_shares: invalid arguments");
        NetScope ns(n);
        drop_graph(n);
        size_t count = n->net->layers.back()->getActivation()->size();
        RSS64 labels = RSS64::wrap((u64 *)label_shares, count);
        n->net->backwardShares(labels);
        return (int)OK;
    });
}

int pir_net_layer_backward(pir_net *n, int layer, const uint64_t *delta, const uint64_t *forward_input) {
    return guarded([&]() -> int {
        PIR_REQUIRE(n && delta && forward_input && layer >= 0 && layer < pir_net_num_layers(n),
                    "pir_net_layer_backward: invalid arguments");
        NetScope ns(n);
        drop_graph(n);
        Layer *L = n->net->layers[layer].get();
        RSS64 d = RSS64::wrap((u64 *)delta, L->getActivation()->size());
        RSS64 x = RSS64::wrap((u64 *)forward_input, L->inputSize());
        L->backward(d, x);
        return (int)OK;
    });
}

int pir_net_delta(pir_net *n, int layer, const uint64_t **shares, size_t *count) {
    PIR_REQUIRE(n && shares && count && layer >= 0 && layer <= pir_net_num_layers(n), "pir_net_delta: bad layer");
    RSS64 *d = layer == pir_net_num_layers(n) ? &n->net->outputDelta : n->net->layers[layer]->getDelta();
    PIR_REQUIRE(d->size() > 0, "pir_net_delta: no backward pass has run");
    *shares = d->base();
    *count = d->size();
    return OK;
}

int pir_net_forward(pir_net *n, const double *input, size_t count) {
    return guarded([&]() -> int {
        PIR_REQUIRE(n && input, "pir_net_forward: invalid arguments");
        NetScope ns(n);
        PIR_REQUIRE(count == n->net->input.size(), "pir_net_forward: input size mismatch");
        n->net->forward(std::vector<double>(input, input + count));
        return (int)OK;
    });
}

int pir_net_forward_shares(pir_net *n, const uint64_t *shares) {
    return guarded([&]() -> int {
        PIR_REQUIRE(n, "pir_net_forward_shares: invalid arguments");
        NetScope ns(n);
        if (shares && shares != n->net->input.base())
            PIR_HIP(hipMemcpyAsync(n->net->input.base(), shares, 6 * n->net->input.size() * 8,
                                   hipMemcpyDeviceToDevice, n->ctx.stream));
        n->net->run();
        return (int)OK;
    });
}

int pir_net_input(pir_net *n, uint64_t **shares, size_t *count) {
    PIR_REQUIRE(n, "pir_net_input: invalid arguments");
    *shares = n->net->input.base();
    *count = n->net->input.size();
    return OK;
}

int pir_net_synchronize(pir_net *n) {
    PIR_REQUIRE(n, "pir_net_synchronize: invalid arguments");
    PIR_HIP(hipStreamSynchronize(n->ctx.stream));
    return OK;
}

int pir_net_activation(pir_net *n, int layer, const uint64_t **shares, size_t *count) {
    PIR_REQUIRE(n && layer >= 0 && layer < pir_net_num_layers(n), "pir_net_activation: bad layer");
    RSS64 *a = n->net->layers[layer]->getActivation();
    *shares = a->base();
    *count = a->size();
    return OK;
}

int pir_net_derivative(pir_net *n, int layer, const uint8_t **shares, size_t *count) {
    PIR_REQUIRE(n && layer >= 0 && layer < pir_net_num_layers(n), "pir_net_derivative: bad layer");
    RSS<u8> *d = n->net->layers[layer]->getDerivative();
    PIR_REQUIRE(d, "pir_net_derivative: layer has no derivative shares");
    *shares = d->base();
    *count = d->size();
    return OK;
}

int pir_net_weights(pir_net *n, int layer, const uint64_t **shares, size_t *count, const uint64_t **bias,
                    size_t *n_bias) {
    PIR_REQUIRE(n && layer >= 0 && layer < pir_net_num_layers(n), "pir_net_weights: bad layer");
    Layer *L = n->net->layers[layer].get();
    PIR_REQUIRE(L->getWeights(), "pir_net_weights: layer has no weights");
    *shares = L->getWeights()->base();
    *count = L->getWeights()->size();
    if (bias) *bias = L->getBiases() ? L->getBiases()->base() : nullptr;
    if (n_bias) *n_bias = L->getBiases() ? L->getBiases()->size() : 0;
    return OK;
}

namespace {
RSS64 *net_param(pir_net *n, int layer, int index, std::string *name) {
    if (!n || layer < 0 || layer >= pir_net_num_layers(n)) return nullptr;
    auto t = n->net->layers[layer]->snapshotTensors();
    if (index < 0 || index >= (int)t.size()) return nullptr;
    if (name) *name = t[index].first;
    return t[index].second;
}
}  // namespace

int pir_net_rerandomize(pir_net *n, uint64_t seed) {
    return guarded([&]() -> int {
        PIR_REQUIRE(n, "pir_net_rerandomize: invalid arguments");
        NetScope ns(n);
        u32 call = 0;
        for (auto &L : n->net->layers) {
            auto tensors = L->snapshotTensors();
            for (auto &t : tensors) {
                RSS64 *x = t.second;
                PIR_TRY(rss_deal(nullptr, x->base(), x->size(), seed, call++, x->base(), n->ctx.stream));
            }
            if (!tensors.empty()) L->weightsChanged();
        }
        RSS64 &in = n->net->input;
        PIR_TRY(rss_deal(nullptr, in.base(), in.size(), seed, call, in.base(), n->ctx.stream));
        drop_graph(n);
        return (int)OK;
    });
}

int pir_net_num_params(pir_net *n, int layer) {
    if (!n || layer < 0 || layer >= pir_net_num_layers(n)) return 0;
    return (int)n->net->layers[layer]->snapshotTensors().size();
}

int pir_net_param(pir_net *n, int layer, int index, const uint64_t **shares, size_t *count) {
    PIR_REQUIRE(shares && count, "pir_net_param: invalid arguments");
    RSS64 *t = net_param(n, layer, index, nullptr);
    PIR_REQUIRE(t, "pir_net_param: no such parameter");
    *shares = t->base();
    *count = t->size();
    return OK;
}

int pir_net_param_name(pir_net *n, int layer, int index, char *name, size_t capacity) {
    std::string s;
    PIR_REQUIRE(name && capacity > 0 && net_param(n, layer, index, &s), "pir_net_param_name: no such parameter");
    PIR_REQUIRE(s.size() < capacity, "pir_net_param_name: buffer too small");
    std::copy(s.begin(), s.end(), name);
    name[s.size()] = 0;
    return OK;
}

int pir_net_set_param_shares(pir_net *n, int layer, int index, const uint64_t *shares) {
    return guarded([&]() -> int {
        PIR_REQUIRE(shares, "pir_net_set_param_shares: invalid arguments");
        RSS64 *t = net_param(n, layer, index, nullptr);
        PIR_REQUIRE(t, "pir_net_set_param_shares: no such parameter");
        NetScope ns(n);
        t->copy_from_host(shares);
        n->net->layers[layer]->weightsChanged();
        drop_graph(n);
        return (int)OK;
    });
}

int pir_net_capture(pir_net *n, const uint64_t *shares) {
    return guarded([&]() -> int {
        PIR_REQUIRE(n, "pir_net_capture: invalid arguments");
        NetScope ns(n);
        hipStream_t s = n->ctx.stream;
        if (shares && shares != n->net->input.base())
            PIR_HIP(hipMemcpyAsync(n->net->input.base(), shares, 6 * n->net->input.size() * 8,
                                   hipMemcpyDeviceToDevice, s));
        // one eager pass packs the weights and warms the allocator cache; the
        // captured pass reuses its call numbers (masked precompute mode), so
        // every replay computes what that eager pass computed
        const u32 call_base = n->ctx.prg_call;
        n->net->run();
        PIR_HIP(hipStreamSynchronize(s));
        drop_graph(n);
        n->ctx.prg_call = call_base;
        PIR_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        n->ctx.alloc.capturing = true;
        int rc = guarded([&]() -> int {
            n->net->run();
            return (int)OK;
        });
        n->ctx.alloc.capturing = false;
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(s, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        PIR_HIP(e);
        n->graph = g;
        PIR_HIP(hipGraphInstantiate(&n->exec, g, nullptr, nullptr, 0));
        return (int)OK;
    });
}

int pir_net_profile_layers(pir_net *n, float *ms, int count) {
    return guarded([&]() -> int {
        PIR_REQUIRE(n && ms && count == pir_net_num_layers(n), "pir_net_profile_layers: invalid arguments");
        NetScope ns(n);
        hipStream_t s = n->ctx.stream;
        std::vector<hipEvent_t> ev(count + 1);
        for (auto &e : ev) PIR_HIP(hipEventCreate(&e));
        PIR_HIP(hipEventRecord(ev[0], s));
        // the same schedule as a forward (a ReLU fused into the layer before it
        // gets the event right after its producer's, i.e. ~0 ms)
        hipError_t rec = hipSuccess;
        n->net->run([&](size_t i) {
            hipError_t e = hipEventRecord(ev[i + 1], s);
            if (rec == hipSuccess) rec = e;
        });
        PIR_HIP(rec);
        PIR_HIP(hipEventSynchronize(ev[count]));
        for (int i = 0; i < count; i++) PIR_HIP(hipEventElapsedTime(&ms[i], ev[i], ev[i + 1]));
        for (auto &e : ev) (void)hipEventDestroy(e);
        return (int)OK;
    });
}

int pir_net_profile_gemm(pir_net *n, double *ms, double *ops, int *launches) {
    return guarded([&]() -> int {
        PIR_REQUIRE(n && ms && ops && launches, "pir_net_profile_gemm: invalid arguments");
        NetScope ns(n);
        return timed_gemms(n, ms, ops, launches, [&]() -> int {
            n->net->run();
            return (int)OK;
        });
    });
}

int pir_net_profile_train_gemm(pir_net *n, double *ms, double *ops, int *launches) {
    return guarded([&]() -> int {
        PIR_REQUIRE(n && ms && ops && launches, "pir_net_profile_train_gemm: invalid arguments");
        NetScope ns(n);
        drop_graph(n);
        return timed_gemms(n, ms, ops, launches, [&]() -> int {
            train_step(n);
            return (int)OK;
        });
    });
}

int pir_net_profile_train_layers(pir_net *n, float *ms, int count) {
    return guarded([&]() -> int {
        PIR_REQUIRE(n && ms && count == 2 * pir_net_num_layers(n) + 1,
                    "pir_net_profile_train_layers: invalid arguments");
        NetScope ns(n);
        drop_graph(n);
        require_train_labels(n);
        hipStream_t s = n->ctx.stream;
        const int L = pir_net_num_layers(n);
        std::vector<hipEvent_t> ev(count + 1);
        for (auto &e : ev) PIR_HIP(hipEventCreate(&e));
        hipError_t rec = hipSuccess;
        auto mark = [&](int k) {
            hipError_t e = hipEventRecord(ev[k], s);
            if (rec == hipSuccess) rec = e;
        };
        // events in issue order: forward layers 0..L-1, the softmax delta, then
        // the backward steps, which run from layer L-1 down to layer 0
        int next = 0;
        mark(next++);
        n->net->run([&](size_t) { mark(next++); });
        n->net->backwardShares(n->labels, [&](size_t) { mark(next++); });
        PIR_HIP(rec);
        PIR_REQUIRE(next == count + 1, "pir_net_profile_train_layers: unexpected step count");
        PIR_HIP(hipEventSynchronize(ev[count]));
        std::vector<float> dt(count);
        for (int j = 0; j < count; j++) PIR_HIP(hipEventElapsedTime(&dt[j], ev[j], ev[j + 1]));
        for (int j = 0; j <= L; j++) ms[j] = dt[j];                  // forward, softmax
        for (int t = 0; t < L; t++) ms[L + 1 + (L - 1 - t)] = dt[L + 1 + t];  // backward of layer L-1-t
        for (auto &e : ev) (void)hipEventDestroy(e);
        return (int)OK;
    });
}

int pir_net_set_train_labels(pir_net *n, const double *labels, size_t count) {
    return guarded([&]() -> int {
        PIR_REQUIRE(n && labels && !n->net->layers.empty(), "pir_net_set_train_labels: invalid arguments");
        PIR_REQUIRE(count == n->net->layers.back()->getActivation()->size(),
                    "pir_net_set_train_labels: labels must have batch x classes entries");
        NetScope ns(n);
        if (n->labels.size() != count) {
            drop_graph(n);  // a captured step reads the label buffer
            n->labels.resize(count);
        }
        n->labels.setPublic(std::vector<double>(labels, labels + count), n->ctx.frac_bits);
        n->net->prepareBackward();
        return (int)OK;
    });
}

int pir_net_train_labels(pir_net *n, uint64_t **shares, size_t *count) {
    PIR_REQUIRE(n && shares && count && n->labels.size() > 0, "pir_net_train_labels: no labels set");
    *shares = n->labels.base();
    *count = n->labels.size();
    return OK;
}

int pir_net_train_step(pir_net *n) {
    return guarded([&]() -> int {
        PIR_REQUIRE(n, "pir_net_train_step: invalid arguments");
        NetScope ns(n);
        drop_graph(n);
        train_step(n);
        return (int)OK;
    });
}

int pir_net_capture_train(pir_net *n) {
    return guarded([&]() -> int {
        PIR_REQUIRE(n, "pir_net_capture_train: invalid arguments");
        NetScope ns(n);
        hipStream_t s = n->ctx.stream;
        drop_graph(n);
        // persistent buffers first, then one eager step (a real training step)
        // fills the allocator cache with every temporary the capture needs
        n->net->prepareBackward();
        train_step(n);
        PIR_HIP(hipStreamSynchronize(s));
        PIR_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        n->ctx.alloc.capturing = true;
        int rc = guarded([&]() -> int {
            train_step(n);
            return (int)OK;
        });
        n->ctx.alloc.capturing = false;
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(s, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        PIR_HIP(e);
        n->tgraph = g;
        PIR_HIP(hipGraphInstantiate(&n->texec, g, nullptr, nullptr, 0));
        return (int)OK;
    });
}

int pir_net_replay_train(pir_net *n) {
    PIR_REQUIRE(n && n->texec, "pir_net_replay_train: no captured training step (capture again after a change)");
    PIR_HIP(hipGraphLaunch(n->texec, n->ctx.stream));
    return OK;
}

int pir_net_autotune_train(pir_net *n) {
    return guarded([&]() -> int {
        PIR_REQUIRE(n, "pir_net_autotune_train: invalid arguments");
        NetScope ns(n);
        drop_graph(n);
        n->ctx.tune = true;
        int rc = guarded([&]() -> int {
            train_step(n);
            return (int)OK;
        });
        n->ctx.tune = false;
        if (rc) return rc;
        PIR_HIP(hipStreamSynchronize(n->ctx.stream));
        return (int)OK;
    });
}

int pir_net_autotune(pir_net *n) {
    return guarded([&]() -> int {
        PIR_REQUIRE(n, "pir_net_autotune: invalid arguments");
        NetScope ns(n);
        n->ctx.tune = true;
        int rc = guarded([&]() -> int {
            n->net->run();
            return (int)OK;
        });
        n->ctx.tune = false;
        PIR_TRY(rc);
        PIR_HIP(hipStreamSynchronize(n->ctx.stream));
        return (int)OK;
    });
}

int pir_autotune_set(int on) {
    ctx().tune = on != 0;
    return OK;
}
int pir_autotune_entries(void) { return limb_tune_entries(); }
int pir_autotune_clear(void) {
    limb_tune_clear();
    return OK;
}
int pir_autotune_save(const char *path) { return limb_tune_save(path); }
int pir_autotune_load(const char *path) { return limb_tune_load(path); }

int pir_net_replay(pir_net *n) {
    PIR_REQUIRE(n && n->exec, "pir_net_replay: no captured graph (a weight change drops it; capture again)");
    PIR_HIP(hipGraphLaunch(n->exec, n->ctx.stream));
    return OK;
}

}  // extern "C"
