// peaks.hip -- measured peaks of the MI355X units that bound the path's
// kernels: HBM bandwidth (copy, read-only, write-only streams), the int8 MFMA
// rate of v_mfma_i32_32x32x32_i8 (the limb GEMM's instruction) and the rate of
// 64-bit integer multiply-adds on the vector ALUs (the plain GEMM's
// operation).  They are the measured denominators reported beside the
// datasheet peaks.  The read-only and write-only kernels, of known size, also
// calibrate the hardware counters of HBM read and write requests.
#include "../../include/pir.h"

#include "common.h"
#include "runtime.h"

namespace pir {
namespace {

typedef unsigned long long u64x2 __attribute__((ext_vector_type(2)));
typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

constexpr int PEAK_BLOCK = 256;

// One 16-byte item per thread, one launch over the whole buffer (measured
// faster on MI355X than grid-stride loops, which keep fewer bytes in flight per
// channel); the copy uses non-temporal accesses, the fastest copy form measured.
__global__ __launch_bounds__(PEAK_BLOCK) void peak_copy_kernel(const u64x2 *__restrict__ src,
                                                               u64x2 *__restrict__ dst, size_t n) {
    const size_t i = (size_t)blockIdx.x * PEAK_BLOCK + threadIdx.x;
    if (i < n) __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
}

// Reads n 16-byte items; the fold is stored only if it equals a value the
// filled buffer cannot produce, so the kernel writes nothing.
__global__ __launch_bounds__(PEAK_BLOCK) void peak_read_kernel(const u64x2 *__restrict__ src, size_t n,
                                                               u64 *__restrict__ sink) {
    const size_t i = (size_t)blockIdx.x * PEAK_BLOCK + threadIdx.x;
    if (i >= n) return;
    const u64x2 v = src[i];
    if ((v.x ^ v.y) == 0x5a5a5a5a5a5a5a5aull) sink[0] = v.x;
}

__global__ __launch_bounds__(PEAK_BLOCK) void peak_write_kernel(u64x2 *__restrict__ dst, size_t n, u64 value) {
    const size_t i = (size_t)blockIdx.x * PEAK_BLOCK + threadIdx.x;
    if (i < n) dst[i] = (u64x2){value, value ^ 0xffull};
}

// Each wave issues `iters` rounds of 8 independent v_mfma_i32_32x32x32_i8 (two
// accumulator uses of one register block are 8 instructions apart) and stores
// one lane's accumulator sum at the end.
constexpr int MFMA_CHAINS = 8;
// RND = 0: operands with few set bits per byte (lane-derived small values);
// RND = 1: full-entropy operand bytes, each of the 8 chains with its own pair.
// The difference is the MFMA array's data-dependent power draw: full-entropy
// operands, such as the digit planes of uniformly random shares, run at a
// lower sustained rate.
__device__ __forceinline__ unsigned peak_hash(unsigned x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

template <int RND>
__global__ __launch_bounds__(PEAK_BLOCK) void peak_mfma_kernel(int iters, int *__restrict__ out) {
    const int lane = threadIdx.x & 63;
    v4i a[MFMA_CHAINS], b[MFMA_CHAINS];
#pragma unroll
    for (int c = 0; c < MFMA_CHAINS; c++) {
        const unsigned h = ((unsigned)(blockIdx.x * PEAK_BLOCK + threadIdx.x) * MFMA_CHAINS + c) * 8u;
        if (RND) {
            a[c] = (v4i){(int)peak_hash(h), (int)peak_hash(h + 1), (int)peak_hash(h + 2), (int)peak_hash(h + 3)};
            b[c] = (v4i){(int)peak_hash(h + 4), (int)peak_hash(h + 5), (int)peak_hash(h + 6), (int)peak_hash(h + 7)};
        } else {
            a[c] = (v4i){lane, lane * 3 + 1, lane ^ 5, 7};
            b[c] = (v4i){lane + 2, lane * 5, 11, lane ^ 9};
        }
    }
    v16i acc[MFMA_CHAINS];
#pragma unroll
    for (int c = 0; c < MFMA_CHAINS; c++)
#pragma unroll
        for (int r = 0; r < 16; r++) acc[c][r] = 0;
    for (int it = 0; it < iters; it++) {
#pragma unroll
        for (int c = 0; c < MFMA_CHAINS; c++) acc[c] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a[c], b[c], acc[c], 0, 0, 0);
    }
    int s = 0;
#pragma unroll
    for (int c = 0; c < MFMA_CHAINS; c++)
#pragma unroll
        for (int r = 0; r < 16; r++) s += acc[c][r];
    out[(size_t)blockIdx.x * PEAK_BLOCK + threadIdx.x] = s;
}

// `iters` rounds of 8 independent 64-bit multiply-accumulates acc += a * b
// per lane, the plain ring GEMM's mod-2^64 inner-product step (a per lane and
// accumulator, b stepped by m each round so it is not a loop constant).
constexpr int VALU_CHAINS = 8;
__global__ __launch_bounds__(PEAK_BLOCK) void peak_valu_kernel(int iters, u64 m, u64 *__restrict__ out) {
    const u64 t = (size_t)blockIdx.x * PEAK_BLOCK + threadIdx.x;
    u64 a[VALU_CHAINS], acc[VALU_CHAINS];
#pragma unroll
    for (int c = 0; c < VALU_CHAINS; c++) {
        a[c] = t * 0x9e3779b97f4a7c15ull + (u64)c;
        acc[c] = 0;
    }
    u64 b = t ^ m;
    for (int it = 0; it < iters; it++) {
#pragma unroll
        for (int c = 0; c < VALU_CHAINS; c++) acc[c] += a[c] * b;
        b += m;
    }
    u64 s = 0;
#pragma unroll
    for (int c = 0; c < VALU_CHAINS; c++) s ^= acc[c];
    out[t] = s;
}

// Compute peaks: eight workgroups per CU, enough resident waves to fill the chip.
constexpr unsigned COMPUTE_GRID = 256 * 8;

int measure(int which, size_t bytes, int reps, double *rate, double *ms_out, hipStream_t s) {
    PIR_REQUIRE(reps >= 1 && rate && ms_out, "pir_measure_peak: reps >= 1 and result pointers needed");
    const bool hbm = which == PIR_PEAK_HBM_COPY || which == PIR_PEAK_HBM_READ || which == PIR_PEAK_HBM_WRITE;
    const bool mfma = which == PIR_PEAK_MFMA_I8 || which == PIR_PEAK_MFMA_I8_RANDOM;
    PIR_REQUIRE(hbm || mfma || which == PIR_PEAK_VALU_U64, "pir_measure_peak: unknown peak");
    PIR_REQUIRE(!hbm || (bytes >= 16 && bytes % 16 == 0 && bytes / 16 / PEAK_BLOCK < (1ull << 31)),
                "pir_measure_peak: bytes must be a positive multiple of 16 below 8 TiB");
    void *a = nullptr, *b = nullptr;
    const size_t n16 = bytes / 16;
    const unsigned hbm_grid = (unsigned)((n16 + PEAK_BLOCK - 1) / PEAK_BLOCK);
    const int mfma_iters = 4096, valu_iters = 4096;
    size_t sink_bytes = (size_t)COMPUTE_GRID * PEAK_BLOCK * 8;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = OK;
    auto fail = [&](hipError_t e, const char *what) {
        if (e != hipSuccess && rc == OK) {
            set_error(std::string("pir_measure_peak: ") + what + ": " + hipGetErrorString(e));
            rc = ERR_HIP;
        }
    };
    fail(hipMalloc(&b, hbm ? bytes : sink_bytes), "hipMalloc");
    if (rc == OK && which == PIR_PEAK_HBM_COPY) fail(hipMalloc(&a, bytes), "hipMalloc");
    if (rc == OK && (which == PIR_PEAK_HBM_COPY || which == PIR_PEAK_HBM_READ)) {
        // source contents: a fill of known words
        hipLaunchKernelGGL(peak_write_kernel, dim3(hbm_grid), dim3(PEAK_BLOCK), 0, s,
                           (u64x2 *)(which == PIR_PEAK_HBM_COPY ? a : b), n16, 0x0123456789abcdefull);
        fail(hipGetLastError(), "fill");
    }
    if (rc == OK) fail(hipEventCreate(&e0), "event");
    if (rc == OK) fail(hipEventCreate(&e1), "event");
    auto launch = [&]() {
        switch (which) {
            case PIR_PEAK_HBM_COPY:
                hipLaunchKernelGGL(peak_copy_kernel, dim3(hbm_grid), dim3(PEAK_BLOCK), 0, s, (const u64x2 *)a,
                                   (u64x2 *)b, n16);
                break;
            case PIR_PEAK_HBM_READ:
                hipLaunchKernelGGL(peak_read_kernel, dim3(hbm_grid), dim3(PEAK_BLOCK), 0, s, (const u64x2 *)b, n16,
                                   (u64 *)b);
                break;
            case PIR_PEAK_HBM_WRITE:
                hipLaunchKernelGGL(peak_write_kernel, dim3(hbm_grid), dim3(PEAK_BLOCK), 0, s, (u64x2 *)b, n16,
                                   0xfedcba9876543210ull);
                break;
            case PIR_PEAK_MFMA_I8:
                hipLaunchKernelGGL(peak_mfma_kernel<0>, dim3(COMPUTE_GRID), dim3(PEAK_BLOCK), 0, s, mfma_iters,
                                   (int *)b);
                break;
            case PIR_PEAK_MFMA_I8_RANDOM:
                hipLaunchKernelGGL(peak_mfma_kernel<1>, dim3(COMPUTE_GRID), dim3(PEAK_BLOCK), 0, s, mfma_iters,
                                   (int *)b);
                break;
            default:
                hipLaunchKernelGGL(peak_valu_kernel, dim3(COMPUTE_GRID), dim3(PEAK_BLOCK), 0, s, valu_iters,
                                   0x2545f4914f6cdd1dull, (u64 *)b);
                break;
        }
        fail(hipGetLastError(), "launch");
    };
    if (rc == OK) launch();  // warm-up
    if (rc == OK) fail(hipEventRecord(e0, s), "event");
    for (int r = 0; r < reps && rc == OK; r++) launch();
    if (rc == OK) fail(hipEventRecord(e1, s), "event");
    if (rc == OK) fail(hipEventSynchronize(e1), "event");
    float total = 0;
    if (rc == OK) fail(hipEventElapsedTime(&total, e0, e1), "event");
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    if (rc != OK) return rc;
    const double ms = (double)total / reps, sec = ms * 1e-3;
    double r = 0;
    switch (which) {
        case PIR_PEAK_HBM_COPY: r = 2.0 * bytes / sec / 1e9; break;  // GB/s read + written
        case PIR_PEAK_HBM_READ:
        case PIR_PEAK_HBM_WRITE: r = (double)bytes / sec / 1e9; break;
        case PIR_PEAK_MFMA_I8:  // TOP/s: 2 * 32 * 32 * 32 ops per instruction
        case PIR_PEAK_MFMA_I8_RANDOM:
            r = (double)COMPUTE_GRID * (PEAK_BLOCK / 64) * mfma_iters * MFMA_CHAINS * 65536.0 / sec / 1e12;
            break;
        default:  // G 64-bit multiply-adds per second
            r = (double)COMPUTE_GRID * PEAK_BLOCK * valu_iters * VALU_CHAINS / sec / 1e9;
            break;
    }
    *rate = r;
    *ms_out = ms;
    return OK;
}

}  // namespace
}  // namespace pir

extern "C" int pir_measure_peak(int which, size_t bytes, int reps, double *rate, double *ms, void *stream) {
    try {
        return pir::measure(which, bytes, reps, rate, ms, (hipStream_t)stream);
    } catch (const std::exception &e) {
        pir::set_error(e.what());
        return pir::ERR_STATE;
    }
}
