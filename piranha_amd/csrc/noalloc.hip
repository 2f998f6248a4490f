// noalloc.hip -- the allocation-free form of the RSS local product
// (pir_ring_gemm3's limb-MFMA form) for callers that cannot allocate device
// memory: code under hipGraph capture, tight serving loops, harnesses that
// forbid hipMalloc.
//
// The limb form needs device memory beyond its inputs and output: the packed
// B operand (the signed-digit limb planes of B0 + B1 and B0, 16 bytes per
// element of B) and, when K is split across workgroups, the 64-bit split-K
// partials.  Here both live in one file-scope device array of fixed size,
// `scratch`.  Each GPU holds its own copy of the array from the time the
// library's code is loaded on it until the process ends.
//
// In host code the array's name is only a host-side stand-in: its device
// address is obtained once per device with hipGetSymbolAddress (with that
// device current) and kept per device; kernels receive that address.
//
// Every call on one GPU shares that GPU's array and leaves the last call's
// contents in it.  Nothing in it has to start at zero: the pack kernel writes
// every byte of the packed operand it reads (padding included) and every
// split writes every partial it owns before the reduction reads it.  Calls
// that overlap on one GPU (two streams, or two threads, at once) race on the
// array; the caller orders them.
#include "../../include/pir.h"

#include <mutex>

#include "ring_gemm.h"
#include "runtime.h"

namespace pir {
namespace {

constexpr size_t SCRATCH_BYTES = (size_t)PIR_NOALLOC_SCRATCH_BYTES;
constexpr int MAX_DEVICES = 64;

static __device__ __align__(256) uint8_t scratch[SCRATCH_BYTES];

std::mutex g_scratch_mu;
void *g_scratch_dev[MAX_DEVICES];  // device address of `scratch` on each device, once resolved

// The device address of this device's `scratch`, resolved on the first call
// made with the device current.
int scratch_address(void **out) {
    int dev = -1;
    PIR_HIP(hipGetDevice(&dev));
    PIR_REQUIRE(dev >= 0 && dev < MAX_DEVICES, "pir_ring_gemm3_noalloc: device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void *p = nullptr;
        PIR_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        PIR_REQUIRE(p != nullptr, "pir_ring_gemm3_noalloc: no device address for the scratch array");
        g_scratch_dev[dev] = p;
    }
    *out = g_scratch_dev[dev];
    return OK;
}

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

int ring_gemm3_noalloc(int M, int N, int K, const u64 *A0, const u64 *A1, const u64 *B0, const u64 *B1, u64 *C,
                       hipStream_t s) {
    PIR_REQUIRE(M > 0 && N > 0 && K > 0, "pir_ring_gemm3_noalloc: invalid shape");
    PIR_REQUIRE(A0 && A1 && B0 && B1 && C, "pir_ring_gemm3_noalloc: null operand");
    LimbA la{};
    la.seg[0] = A0;
    la.seg[1] = A1;
    la.nseg = 2;
    la.M = M;
    la.K = K;
    la.Kp = round_up(K, LIMB_BK);
    la.lda = K;
    la.row_major = 1;
    la.conv = LIMB_DENSE;

    // the tuned or heuristic plan (a table lookup, no tuning run); if its
    // split-K partials do not fit beside the packed operand, the smallest split
    // the int32 accumulation bound allows (the result is the same for any split)
    const size_t packed_bytes = align256(limb_b_bytes(N, 2, la.Kp));
    LimbPlan plan = limb_plan(la, N);
    if (packed_bytes + limb_workspace_bytes(plan, M, N) > SCRATCH_BYTES)
        plan.split = limb_effective_split(2 * la.Kp / LIMB_BK, 1);
    const size_t need = packed_bytes + limb_workspace_bytes(plan, M, N);
    if (need > SCRATCH_BYTES) {
        set_error("pir_ring_gemm3_noalloc: M=" + std::to_string(M) + " N=" + std::to_string(N) +
                  " K=" + std::to_string(K) + " needs " + std::to_string(need) +
                  " bytes of scratch, more than the maximum of " + std::to_string(SCRATCH_BYTES) +
                  " (PIR_NOALLOC_SCRATCH_BYTES); use pir_ring_gemm3");
        return ERR_ARG;
    }

    void *base = nullptr;
    PIR_TRY(scratch_address(&base));
    int8_t *packed = (int8_t *)base;
    u64 *workspace = plan.split > 1 ? (u64 *)((uint8_t *)base + packed_bytes) : nullptr;

    LimbBSrc src{};
    src.b[0][0] = B0;  // segment 0 multiplies A0 by B0 + B1
    src.b[0][1] = B1;
    src.b[1][0] = B0;  // segment 1 multiplies A1 by B0
    src.coef[0][0] = src.coef[0][1] = src.coef[1][0] = 1;
    src.nseg = 2;
    src.K = K;
    src.N = N;
    src.Kp = la.Kp;
    src.row_major = 1;
    PIR_TRY(limb_pack_b(src, packed, s));

    Epilogue e{};
    e.mode = EPI_RAW;
    e.C = C;
    e.tc = 1;
    return limb_gemm(la, packed, N, e, plan, workspace, s);
}

}  // namespace
}  // namespace pir

extern "C" int pir_ring_gemm3_noalloc(int M, int N, int K, const uint64_t *A0, const uint64_t *A1,
                                      const uint64_t *B0, const uint64_t *B1, uint64_t *C, void *stream) {
    try {
        return pir::ring_gemm3_noalloc(M, N, K, A0, A1, B0, B1, C, (hipStream_t)stream);
    } catch (const pir::PirError &e) {
        return e.code();
    } catch (const std::exception &e) {
        pir::set_error(e.what());
        return pir::ERR_STATE;
    }
}
