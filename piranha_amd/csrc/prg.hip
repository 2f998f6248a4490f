// prg.hip -- device side of the seeded PRG (prg.h): raw Philox4x32-10 blocks,
// the word stream of one call, and the dealer of random RSS sharings (the
// masked precompute mode's input and weight sharing; the reference shares
// every public value as (v, 0, 0) through setPublic, RSS.inl:122-142).
#include "prg.h"

#include "rss_device.h"

namespace pir {

constexpr int PRG_BLOCK = 256;

__global__ __launch_bounds__(PRG_BLOCK) void philox_kernel(const u32 *__restrict__ ctr, const u32 *__restrict__ key,
                                                           size_t n, u32 *__restrict__ out) {
    size_t i = (size_t)blockIdx.x * PRG_BLOCK + threadIdx.x;
    if (i >= n) return;
    u32 c[4] = {ctr[4 * i], ctr[4 * i + 1], ctr[4 * i + 2], ctr[4 * i + 3]};
    u32 k[2] = {key[2 * i], key[2 * i + 1]};
    u32 o[4];
    philox4x32_10(c, k, o);
#pragma unroll
    for (int j = 0; j < 4; j++) out[4 * i + j] = o[j];
}

int prg_philox(const u32 *ctr, const u32 *key, size_t n, u32 *out, hipStream_t s) {
    if (n == 0) return OK;
    PIR_REQUIRE(ctr && key && out, "pir_prg_philox: null buffer");
    hipLaunchKernelGGL(philox_kernel, dim3(grid_for(n, PRG_BLOCK)), dim3(PRG_BLOCK), 0, s, ctr, key, n, out);
    PIR_LAUNCH_CHECK();
    return OK;
}

// out[2b + h] = word h of element b of the call (one Philox block per thread).
__global__ __launch_bounds__(PRG_BLOCK) void prg_fill_kernel(PrgStream g, size_t n, u64 *__restrict__ out) {
    size_t b = (size_t)blockIdx.x * PRG_BLOCK + threadIdx.x;
    if (2 * b >= n) return;
    u64 w0, w1;
    prg_pair(g, b, 0, w0, w1);
    out[2 * b] = w0;
    if (2 * b + 1 < n) out[2 * b + 1] = w1;
}

int prg_fill(u64 seed, u32 call, size_t n, u64 *out, hipStream_t s) {
    if (n == 0) return OK;
    PIR_REQUIRE(out, "pir_prg_fill: null buffer");
    const size_t blocks = (n + 1) / 2;
    hipLaunchKernelGGL(prg_fill_kernel, dim3(grid_for(blocks, PRG_BLOCK)), dim3(PRG_BLOCK), 0, s,
                       PrgStream{seed, call, 1}, n, out);
    PIR_LAUNCH_CHECK();
    return OK;
}

// Random consistent RSS sharing of v_i: x1 = word 0, x2 = word 1 of element i,
// x0 = v - x1 - x2, party p holding (x_p, x_{p+1}).  From shares (values ==
// nullptr), v_i is the value party A opens: its two shares plus party C's
// first share.
__global__ __launch_bounds__(PRG_BLOCK) void deal_kernel(const u64 *__restrict__ values,
                                                         const u64 *__restrict__ shares, size_t n, PrgStream g,
                                                         u64 *__restrict__ out) {
    size_t i = (size_t)blockIdx.x * PRG_BLOCK + threadIdx.x;
    if (i >= n) return;
    const u64 v = values ? values[i] : shares[i] + shares[n + i] + shares[4 * n + i];
    u64 x1, x2;
    prg_pair(g, i, 0, x1, x2);
    const u64 x[3] = {v - x1 - x2, x1, x2};
#pragma unroll
    for (int p = 0; p < 3; p++) {
        out[(size_t)(2 * p) * n + i] = x[p];
        out[(size_t)(2 * p + 1) * n + i] = x[next_party(p)];
    }
}

int rss_deal(const u64 *values, const u64 *shares, size_t n, u64 seed, u32 call, u64 *out, hipStream_t s) {
    if (n == 0) return OK;
    PIR_REQUIRE((values || shares) && out, "deal: null buffer");
    hipLaunchKernelGGL(deal_kernel, dim3(grid_for(n, PRG_BLOCK)), dim3(PRG_BLOCK), 0, s, values, shares, n,
                       PrgStream{seed, call, 1}, out);
    PIR_LAUNCH_CHECK();
    return OK;
}

}  // namespace pir
