// tpc_fpc_compare.h -- element-wise protocols of the reference's 2-party
// (TPC.inl) and 4-party (FPC.inl) schemes for all parties at once.  A TPC
// value is [2 parties][n] (one share each, party A first); an FPC value is
// [4 parties][3 shares][n], party q's share s holding component (q + s + 1) % 4.
// Bit values are uint8 (eb = 1) or uint64 (eb = 8).  Placeholder precompute as
// in the reference: TPC Beaver triples x = y = z = fill(1) (Precompute.h:93-105),
// FPC reshares with zero PRGs (FPC.inl:1183-1260).  Every launcher is
// asynchronous on `s`.
#pragma once

#include "common.h"

namespace pir {

// TPC *= TPC (TPC.inl:242-272): x [2][n] *= y [2][n]
int tpc_mult(u64 *x, const u64 *y, size_t n, hipStream_t s);
// TPC &= TPC (TPC.inl:281-312) on bit values
int tpc_and(void *x, const void *y, size_t n, int eb, hipStream_t s);
// dReLU (TPC.inl:828-868) into d [2][n]; with y set, ReLU (TPC.inl:870-884)
// also writes y = x * [x >= 0]
int tpc_relu(const u64 *x, size_t n, void *d, u64 *y, int eb, hipStream_t s);
// selectShare (TPC.inl:585-613): z = b ? y : x
int tpc_select(const u64 *x, const u64 *y, const void *b, u64 *z, size_t n, int eb, hipStream_t s);
// maxpool (TPC.inl:886-1030) over gathered pools [n / k windows][k], k in
// {2, 4, 8, 16}; result [2][n / k], d [2][n]; pools is modified like the
// reference's even views
int tpc_maxpool(u64 *pools, size_t n, u64 *result, void *d, int k, int eb, hipStream_t s);

// FPC *= FPC (FPC.inl:321-362): x [4][3][n] *= y
int fpc_mult(u64 *x, const u64 *y, size_t n, hipStream_t s);
// FPC &= FPC (FPC.inl:374-398) on bit values
int fpc_and(void *x, const void *y, size_t n, int eb, hipStream_t s);
// dReLU (FPC.inl:929-989) into d [4][3][n]; with y set, ReLU (:991-1002)
int fpc_relu(const u64 *x, size_t n, void *d, u64 *y, int eb, hipStream_t s);
// selectShare (FPC.inl:671-712): z = b ? y : x
int fpc_select(const u64 *x, const u64 *y, const void *b, u64 *z, size_t n, int eb, hipStream_t s);
// maxpool (FPC.inl:1004-1181) over gathered pools [n / k windows][k], k in
// {2, 4, 8, 16}; result [4][3][n / k], d [4][3][n]; pools modified in place
int fpc_maxpool(u64 *pools, size_t n, u64 *result, void *d, int k, int eb, hipStream_t s);

// dividePublic(Share, DeviceData) (TPC.inl:386-407, FPC.inl:566-580 with the
// vector getDividedShares, Precompute.h:121-134): d [2][n] / [4][n] per party
int tpc_divide_public_vec(u64 *a, size_t n, const u64 *d, hipStream_t s);
int fpc_divide_public_vec(u64 *a, size_t n, const u64 *d, hipStream_t s);
// getPowers (TPC.inl:1202-1226, FPC.inl:1418-1441): each party's opened
// power, pow [2][n] (TPC) or [4][n] (FPC)
int tpc_get_powers(const u64 *in, size_t n, u64 *pow, hipStream_t s);
int fpc_get_powers(const u64 *in, size_t n, u64 *pow, hipStream_t s);
// taylorSeries (TPC.inl:1228-1290, FPC.inl:1443-1526) with the tabulated
// coefficient functors of taylor_tables(); out must not overlap in
int tpc_taylor(const u64 *in, size_t n, u64 *out, const TaylorTables &t, hipStream_t s);
int fpc_taylor(const u64 *in, size_t n, u64 *out, const TaylorTables &t, hipStream_t s);

}  // namespace pir
