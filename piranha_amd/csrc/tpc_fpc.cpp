// tpc_fpc.cpp -- 2-party (TPC) and 4-party (FPC) matmul / convolution with the
// reference's placeholder precompute, for all parties at once (see tpc_fpc.h).
#include "tpc_fpc.h"

#include <vector>

namespace pir {

namespace {

bool fused_limb() { return ctx().mode == MODE_FUSED && ctx().algo == ALGO_LIMB; }
hipStream_t S() { return ctx().stream; }

// Array i of a contiguous block of n-element arrays.
DeviceData<u64> arr(const u64 *base, int i, size_t n) { return DeviceData<u64>((u64 *)base + (size_t)i * n, n); }

// `count` zeroed n-element scratch arrays in one allocation (DeviceData(n) of
// the reference zero-initialises, DeviceData.h:235-240).
class Block {
  public:
    Block(int count, size_t n) : buf_((size_t)count * n * sizeof(u64)), n_(n) {
        if (count && n) check(fill_u64(base(), 0, (size_t)count * n, S()));
    }
    u64 *base() const { return (u64 *)buf_.get(); }
    DeviceData<u64> operator[](int i) const { return arr(base(), i, n_); }

  private:
    DevBuf buf_;
    size_t n_;
};

// One multi-array launch of x_t op= y_t (y_t null: op= scalar_t).  The launch
// records data pointers when an array is added, so views may be temporaries.
class Ops {
  public:
    explicit Ops(int op) : l_(op) {}
    Ops &add(DeviceData<u64> x, const DeviceData<u64> *y, u64 scalar = 0) {
        l_.add(x, y, scalar);
        return *this;
    }
    Ops &add(DeviceData<u64> x, DeviceData<u64> y) {
        l_.add(x, &y);
        return *this;
    }
    void run() { l_.run(8); }

  private:
    MultiLaunch l_;
};

Ops ops(int op) { return Ops(op); }

// One ring product C = A.B mod 2^64 of the op-by-op form: gpu::gemm
// (gemm.cuh:64-112) with the plain VALU kernel, or the limb kernel.
void gemm_op(int M, int N, int K, const u64 *A, bool ta, const u64 *B, bool tb, u64 *C, bool tc) {
    if (ctx().algo == ALGO_PLAIN) check(plain_gemm(M, N, K, A, ta, B, tb, C, tc, S()));
    else check(limb_ring_gemm(M, N, K, A, ta, B, tb, C, tc, S()));
}

struct Geom {
    int batch, H, W, f, Din, Dout, stride, pad, OH, OW;
};

// gpu::conv_fprop (conv.cuh:380-401) of the op-by-op form.
void conv_op(const u64 *x, const u64 *w, u64 *y, const Geom &g) {
    if (ctx().algo == ALGO_PLAIN)
        check(plain_conv_fprop(x, w, y, g.batch, g.H, g.W, g.Din, g.Dout, g.f, g.f, g.pad, g.pad, g.stride, 1, S()));
    else
        check(limb_ring_conv(x, w, y, g.batch, g.H, g.W, g.Din, g.Dout, g.f, g.f, g.pad, g.pad, g.stride, S()));
}

// One term of a fused product: A segment times sum_t coef_t * B_t.
struct Seg {
    const u64 *a;
    const u64 *b[LIMB_MAX_TERMS];
    int coef[LIMB_MAX_TERMS];
};

// Fused form: raw = sum over segments of A_s . B_s (one K-concatenated limb
// GEMM), epilogue q = trunc((raw - 2^k) / 2^k) written as 1 + q into the
// arrays of `base_mask` among the `nout` output arrays at `out` (others 0).
void fused_product(const std::vector<Seg> &segs, bool ta, bool tb, const Geom *conv, int M, int N, int K, int trunc,
                   u64 *out, int nout, unsigned base_mask, bool tc) {
    if ((int)segs.size() > LIMB_MAX_SEG) {
        set_error("fused product: too many K-segments");
        throw PirError(ERR_STATE);
    }
    LimbBSrc s{};
    LimbA a{};
    int nseg = (int)segs.size();
    for (int i = 0; i < nseg; i++) {
        a.seg[i] = segs[i].a;
        for (int t = 0; t < LIMB_MAX_TERMS; t++) {
            s.b[i][t] = segs[i].b[t];
            s.coef[i][t] = segs[i].coef[t];
        }
    }
    s.nseg = a.nseg = nseg;
    s.K = a.K = K;
    s.N = N;
    s.Kp = a.Kp = round_up(K, LIMB_BK);
    s.row_major = conv ? 0 : (tb ? 1 : 0);
    a.M = M;
    if (conv) {
        a.conv = 1;
        a.row_major = 1;
        a.H = conv->H;
        a.W = conv->W;
        a.Cin = conv->Din;
        a.fh = a.fw = conv->f;
        a.ph = a.pw = conv->pad;
        a.stride = conv->stride;
        a.OH = conv->OH;
        a.OW = conv->OW;
    } else {
        a.lda = ta ? K : M;
        a.row_major = ta ? 1 : 0;
    }
    DevBuf packed(limb_b_bytes(N, nseg, s.Kp));
    check(limb_pack_b(s, (int8_t *)packed.get(), S()));
    Epilogue e{};
    e.mode = EPI_TRUNC;
    e.out = out;
    e.nout = nout;
    e.base_mask = base_mask;
    e.trunc_bits = trunc;
    e.tc = conv ? 1 : (tc ? 1 : 0);
    check(limb_run(a, (const int8_t *)packed.get(), N, e, 0, S()));
}

Geom make_geom(int batch, int H, int W, int f, int Din, int Dout, int stride, int pad) {
    Geom g{batch, H, W, f, Din, Dout, stride, pad, (H + 2 * pad - f) / stride + 1, (W + 2 * pad - f) / stride + 1};
    return g;
}

// ---------------------------------------------------------------------------
// TPC

constexpr int TA = 0, TB = 1;

// TPCBase::fill, TPC.inl:103-106: the value on party A, zero on party B.
void tpc_fill(const Block &x, u64 v) { ops(OP_SET).add(x[TA], nullptr, v).add(x[TB], nullptr, 0).run(); }

// The Beaver products of localMatMul / localFprop (TPC.inl:644-690, 1059-1090)
// in the reference's op-by-op form.  `product(x, y, out)` is one gemm/conv.
template <typename Product>
void tpc_local_product(const u64 *a, size_t an, const u64 *b, size_t bn, u64 *c, size_t cn, bool matrix_triple,
                       u64 K, Product product) {
    Block x(2, an), y(2, bn), z(2, cn), e(2, an), f(2, bn), temp(2, cn);
    if (matrix_triple) {  // getMatrixBeaverTriple: x = y = 1, z = shared dim (Precompute.h:73-91)
        tpc_fill(x, 1);
        tpc_fill(y, 1);
        tpc_fill(z, K);
    }  // getConvBeaverTriple_fprop: all zero (Precompute.h:16-33)
    // (one launch covers arrays of one length: x, e have a's size, y, f b's)
    ops(OP_ADD).add(x[TA], arr(a, TA, an)).add(x[TB], arr(a, TB, an)).run();
    ops(OP_ADD).add(y[TA], arr(b, TA, bn)).add(y[TB], arr(b, TB, bn)).run();
    // reconstruct(x, e), reconstruct(y, f): own + other's share
    ops(OP_ADD).add(e[TA], x[TA]).add(e[TB], x[TB]).run();
    ops(OP_ADD).add(e[TA], x[TB]).add(e[TB], x[TA]).run();
    ops(OP_ADD).add(f[TA], y[TA]).add(f[TB], y[TB]).run();
    ops(OP_ADD).add(f[TA], y[TB]).add(f[TB], y[TA]).run();
    ops(OP_SUB).add(x[TA], arr(a, TA, an)).add(x[TB], arr(a, TB, an)).run();
    ops(OP_SUB).add(y[TA], arr(b, TA, bn)).add(y[TB], arr(b, TB, bn)).run();
    ops(OP_SET).add(arr(c, TA, cn), nullptr, 0).add(arr(c, TB, cn), nullptr, 0).run();
    ops(OP_ADD).add(arr(c, TA, cn), z[TA]).add(arr(c, TB, cn), z[TB]).run();
    for (int q = 0; q < 2; q++) product(e[q].data(), f[q].data(), temp[q].data());
    ops(OP_ADD).add(arr(c, TA, cn), temp[TA]).run();  // TPC += DeviceData: party A only
    for (int q = 0; q < 2; q++) product(e[q].data(), y[q].data(), temp[q].data());
    if (matrix_triple) ops(OP_SUB).add(arr(c, TA, cn), temp[TA]).run();
    else ops(OP_SUB).add(arr(c, TA, cn), temp[TA]).add(arr(c, TB, cn), temp[TB]).run();  // *C.getShare(0) -=
    for (int q = 0; q < 2; q++) product(x[q].data(), f[q].data(), temp[q].data());
    if (matrix_triple) ops(OP_SUB).add(arr(c, TA, cn), temp[TA]).run();
    else ops(OP_SUB).add(arr(c, TA, cn), temp[TA]).add(arr(c, TB, cn), temp[TB]).run();
}

}  // namespace

void tpc_reconstruct(const u64 *x, size_t n, u64 *out) {
    ops(OP_SET).add(arr(out, TA, n), nullptr, 0).add(arr(out, TB, n), nullptr, 0).run();
    ops(OP_ADD).add(arr(out, TA, n), arr(x, TA, n)).add(arr(out, TB, n), arr(x, TB, n)).run();
    ops(OP_ADD).add(arr(out, TA, n), arr(x, TB, n)).add(arr(out, TB, n), arr(x, TA, n)).run();
}

void tpc_divide_public(u64 *a, size_t n, u64 d) {
    Block r(2, n), rPrime(2, n), rec(2, n);
    tpc_fill(rPrime, d);
    tpc_fill(r, 1);
    ops(OP_SUB).add(arr(a, TA, n), rPrime[TA]).add(arr(a, TB, n), rPrime[TB]).run();
    tpc_reconstruct(a, n, rec.base());
    ops(OP_DIV).add(rec[TA], nullptr, d).add(rec[TB], nullptr, d).run();
    ops(OP_SET).add(arr(a, TA, n), nullptr, 0).add(arr(a, TB, n), nullptr, 0).run();
    ops(OP_ADD).add(arr(a, TA, n), r[TA]).add(arr(a, TB, n), r[TB]).run();
    ops(OP_ADD).add(arr(a, TA, n), rec[TA]).run();  // TPC += DeviceData: party A only
}

void tpc_matmul(const u64 *a, const u64 *b, u64 *c, int M, int N, int K, bool ta, bool tb, bool tc, int trunc) {
    size_t an = (size_t)M * K, bn = (size_t)K * N, cn = (size_t)M * N;
    if (fused_limb()) {
        // party A ends with 1 + trunc((X.W - 2^k) / 2^k), X = a_A + a_B, W = b_A + b_B
        // (the Beaver identity K + e.f - e.1 - 1.f = X.W holds exactly mod 2^64)
        const u64 *bA = b, *bB = b + bn;
        std::vector<Seg> segs = {{a, {bA, bB}, {1, 1}}, {a + an, {bA, bB}, {1, 1}}};
        fused_product(segs, ta, tb, nullptr, M, N, K, trunc, c, 2, 1u << TA, tc);
        return;
    }
    tpc_local_product(a, an, b, bn, c, cn, true, (u64)K, [&](const u64 *x, const u64 *y, u64 *o) {
        gemm_op(M, N, K, x, ta, y, tb, o, tc);
    });
    tpc_divide_public(c, cn, (u64)1 << trunc);
}

void tpc_convolution(const u64 *A, const u64 *B, u64 *C, int batch, int H, int W, int f, int Din, int Dout,
                     int stride, int pad, int trunc) {
    Geom g = make_geom(batch, H, W, f, Din, Dout, stride, pad);
    size_t an = (size_t)batch * H * W * Din, bn = (size_t)Dout * f * f * Din, cn = (size_t)batch * g.OH * g.OW * Dout;
    if (fused_limb()) {
        const u64 *bA = B, *bB = B + bn;
        std::vector<Seg> segs = {{A, {bA, bB}, {1, 1}}, {A + an, {bA, bB}, {1, 1}}};
        fused_product(segs, true, false, &g, batch * g.OH * g.OW, Dout, f * f * Din, trunc, C, 2, 1u << TA, true);
        return;
    }
    ops(OP_SET).add(arr(C, TA, cn), nullptr, 0).add(arr(C, TB, cn), nullptr, 0).run();  // C.zero()
    tpc_local_product(A, an, B, bn, C, cn, false, 0, [&](const u64 *x, const u64 *w, u64 *o) { conv_op(x, w, o, g); });
    tpc_divide_public(C, cn, (u64)1 << trunc);
}

// ---------------------------------------------------------------------------
// FPC

namespace {

// the slot where party q keeps a public value / component x0 (FPC.inl:119-124,186-208)
int public_slot(int q) { return q == 1 ? 2 : (q == 2 ? 1 : (q == 3 ? 0 : -1)); }
int fpc_prev(int q) { return (q + 3) % 4; }
// FPC.inl:451-495
int partyG(int i, int j) {
    static const int g[4][4] = {{-1, 2, 1, 1}, {-1, -1, 3, 2}, {-1, -1, -1, 0}, {-1, -1, -1, -1}};
    return g[i][j];
}
int partyH(int i, int j) {
    static const int h[4][4] = {{-1, 3, 3, 2}, {-1, -1, 0, 0}, {-1, -1, -1, 1}, {-1, -1, -1, -1}};
    return h[i][j];
}

// FPCBase::fill (FPC.inl:119-124)
void fpc_fill(const Block &x, u64 v) {
    Ops o = ops(OP_SET);
    for (int q = 0; q < 4; q++)
        for (int s = 0; s < 3; s++) o.add(x[3 * q + s], nullptr, s == public_slot(q) ? v : 0);
    o.run();
}

// share-wise FPC op FPC over the 12 arrays of two [4][3][n] blocks
void fpc_share_op(int op, u64 *x, const u64 *y, size_t n) {
    Ops o = ops(op);
    for (int i = 0; i < 12; i++) o.add(arr(x, i, n), arr(y, i, n));
    o.run();
}

// Party q's raw 3-of-3 sum for dividePublic after matmul / convolution,
// expanded over the local products P_p(s, t) = a_{p,s} . b_{p,t} (FPC.inl
// :635-666 then reshareFPC :1183-1260 and reconstruct :582-604):
//   rec_B = P_B(0,0)+P_B(1,1)+P_B(2,2) + 2 z_B + z'_B + z_A + z'_A + P_A(0,0) + z_C
//   rec_C = P_C(0,0)+P_C(1,1)+P_C(2,2) + 2 z_C + z'_C + z'_B + P_B(0,0) + 2 z_A
//   rec_D = P_D(0,0)+P_D(1,1)+P_D(2,2) + 2 z_D + z'_D + z_B + z_A + P_C(0,0) + z'_C
// with z_p = P_p(0,1)+P_p(1,0), z'_p = P_p(0,2)+P_p(2,0); then minus 2^k.
std::vector<Seg> fpc_party_segments(int q, const u64 *a, size_t an, const u64 *b, size_t bn) {
    auto A = [&](int p, int s) { return a + (size_t)(3 * p + s) * an; };
    auto B = [&](int p, int s) { return b + (size_t)(3 * p + s) * bn; };
    // own products: P(0,0)+P(1,1)+P(2,2)+2z+z' grouped by A share
    std::vector<Seg> v = {
        {A(q, 0), {B(q, 0), B(q, 1), B(q, 2)}, {1, 2, 1}},
        {A(q, 1), {B(q, 1), B(q, 0)}, {1, 2}},
        {A(q, 2), {B(q, 2), B(q, 0)}, {1, 1}},
    };
    if (q == 1) {
        v.push_back({A(0, 0), {B(0, 0), B(0, 1), B(0, 2)}, {1, 1, 1}});  // P_A(0,0) + z_A + z'_A
        v.push_back({A(0, 1), {B(0, 0)}, {1}});
        v.push_back({A(0, 2), {B(0, 0)}, {1}});
        v.push_back({A(2, 0), {B(2, 1)}, {1}});  // z_C
        v.push_back({A(2, 1), {B(2, 0)}, {1}});
    } else if (q == 2) {
        v.push_back({A(1, 0), {B(1, 0), B(1, 2)}, {1, 1}});  // P_B(0,0) + z'_B
        v.push_back({A(1, 2), {B(1, 0)}, {1}});
        v.push_back({A(0, 0), {B(0, 1)}, {2}});  // 2 z_A
        v.push_back({A(0, 1), {B(0, 0)}, {2}});
    } else {
        v.push_back({A(1, 0), {B(1, 1)}, {1}});  // z_B
        v.push_back({A(1, 1), {B(1, 0)}, {1}});
        v.push_back({A(0, 0), {B(0, 1)}, {1}});  // z_A
        v.push_back({A(0, 1), {B(0, 0)}, {1}});
        v.push_back({A(2, 0), {B(2, 0), B(2, 2)}, {1, 1}});  // P_C(0,0) + z'_C
        v.push_back({A(2, 2), {B(2, 0)}, {1}});
    }
    return v;
}

void fpc_fused(const u64 *a, size_t an, const u64 *b, size_t bn, u64 *c, size_t cn, bool ta, bool tb,
               const Geom *conv, int M, int N, int K, int trunc, bool tc) {
    // party A: dividePublic leaves it all-zero (r and the opened value go to B, C, D)
    check(fill_u64(c, 0, 3 * cn, S()));
    for (int q = 1; q < 4; q++)
        fused_product(fpc_party_segments(q, a, an, b, bn), ta, tb, conv, M, N, K, trunc, c + (size_t)3 * q * cn, 3,
                      1u << public_slot(q), tc);
}

// the per-party local products of FPC matmul / convolution (FPC.inl:641-656,
// 744-782): c_{q,s} = a_{q,s} b_{q,s}; z_q, z'_q cross sums
template <typename Product>
void fpc_local_products(const u64 *a, size_t an, const u64 *b, size_t bn, u64 *c, size_t cn, Block &z, Block &zp,
                        Product product) {
    Block temp(1, cn);
    for (int q = 0; q < 4; q++) {
        for (int s = 0; s < 3; s++)
            product(a + (size_t)(3 * q + s) * an, b + (size_t)(3 * q + s) * bn, c + (size_t)(3 * q + s) * cn);
        const int zt[2][2] = {{0, 1}, {1, 0}}, zpt[2][2] = {{0, 2}, {2, 0}};
        for (auto &t : zt) {
            product(a + (size_t)(3 * q + t[0]) * an, b + (size_t)(3 * q + t[1]) * bn, temp.base());
            ops(OP_ADD).add(z[q], temp[0]).run();
        }
        for (auto &t : zpt) {
            product(a + (size_t)(3 * q + t[0]) * an, b + (size_t)(3 * q + t[1]) * bn, temp.base());
            ops(OP_ADD).add(zp[q], temp[0]).run();
        }
    }
}

void fpc_reshare_all(const Block &z, const Block &zp, size_t n, u64 *out) {
    const int pairs[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
    for (auto &pr : pairs) fpc_reshare(z.base(), zp.base(), n, pr[0], pr[1], out);
}

}  // namespace

void fpc_reconstruct(const u64 *x, size_t n, u64 *out) {
    Ops zero = ops(OP_SET);
    for (int q = 0; q < 4; q++) zero.add(arr(out, q, n), nullptr, 0);
    zero.run();
    for (int s = 0; s < 3; s++) {
        Ops o = ops(OP_ADD);
        for (int q = 0; q < 4; q++) o.add(arr(out, q, n), arr(x, 3 * q + s, n));
        o.run();
    }
    Ops rx = ops(OP_ADD);  // the previous party's share 0, received
    for (int q = 0; q < 4; q++) rx.add(arr(out, q, n), arr(x, 3 * fpc_prev(q), n));
    rx.run();
}

void fpc_divide_public(u64 *a, size_t n, u64 d) {
    Block r(12, n), rPrime(12, n), rec(4, n);
    fpc_fill(rPrime, d);
    fpc_fill(r, 1);
    fpc_share_op(OP_SUB, a, rPrime.base(), n);
    fpc_reconstruct(a, n, rec.base());
    Ops dv = ops(OP_DIV);
    for (int q = 0; q < 4; q++) dv.add(rec[q], nullptr, d);
    dv.run();
    Ops zero = ops(OP_SET);
    for (int i = 0; i < 12; i++) zero.add(arr(a, i, n), nullptr, 0);
    zero.run();
    fpc_share_op(OP_ADD, a, r.base(), n);
    Ops add = ops(OP_ADD);  // FPC += DeviceData: B -> shareC, C -> shareB, D -> shareA
    for (int q = 1; q < 4; q++) add.add(arr(a, 3 * q + public_slot(q), n), rec[q]);
    add.run();
}

void fpc_reshare(const u64 *z, const u64 *zp, size_t n, int i, int j, u64 *out) {
    if (!(i >= 0 && i < j && j < 4)) {
        set_error("reshareFPC: parties must satisfy 0 <= i < j < 4");
        throw PirError(ERR_ARG);
    }
    int G = partyG(i, j), H = partyH(i, j);
    bool opp = (j - i) == 2;
    const u64 *mask = opp ? zp : z;
    Ops o = ops(OP_ADD);
    for (int q = 0; q < 4; q++) {
        if (q == H) continue;
        int slot = (3 - q + H) % 4;  // shareH, FPC.inl:497-506
        int src = (q == G) ? i : q;  // G receives party i's mask
        o.add(arr(out, 3 * q + slot, n), arr(mask, src, n));
    }
    o.run();
}

void fpc_matmul(const u64 *a, const u64 *b, u64 *c, int M, int N, int K, bool ta, bool tb, bool tc, int trunc) {
    size_t an = (size_t)M * K, bn = (size_t)K * N, cn = (size_t)M * N;
    if (fused_limb()) {
        fpc_fused(a, an, b, bn, c, cn, ta, tb, nullptr, M, N, K, trunc, tc);
        return;
    }
    Block z(4, cn), zp(4, cn);
    fpc_local_products(a, an, b, bn, c, cn, z, zp,
                       [&](const u64 *x, const u64 *y, u64 *o) { gemm_op(M, N, K, x, ta, y, tb, o, tc); });
    fpc_reshare_all(z, zp, cn, c);
    fpc_divide_public(c, cn, (u64)1 << trunc);
}

void fpc_convolution(const u64 *A, const u64 *B, u64 *C, int batch, int H, int W, int f, int Din, int Dout,
                     int stride, int pad, int trunc) {
    Geom g = make_geom(batch, H, W, f, Din, Dout, stride, pad);
    size_t an = (size_t)batch * H * W * Din, bn = (size_t)Dout * f * f * Din, cn = (size_t)batch * g.OH * g.OW * Dout;
    if (fused_limb()) {
        fpc_fused(A, an, B, bn, C, cn, true, false, &g, batch * g.OH * g.OW, Dout, f * f * Din, trunc, true);
        return;
    }
    Block z(4, cn), zp(4, cn);
    fpc_local_products(A, an, B, bn, C, cn, z, zp, [&](const u64 *x, const u64 *w, u64 *o) { conv_op(x, w, o, g); });
    fpc_reshare_all(z, zp, cn, C);
    fpc_divide_public(C, cn, (u64)1 << trunc);
}

}  // namespace pir
