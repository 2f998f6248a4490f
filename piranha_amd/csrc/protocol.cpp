// protocol.cpp -- RSS protocol functions (src/mpc/RSS.inl) for all three
// parties, in the fused and the op-by-op execution modes (see protocol.h).
#include "protocol.h"

#include <cmath>

#include "fused_rss.h"

namespace pir {

static bool fused_mode() { return ctx().mode == MODE_FUSED; }
static hipStream_t S() { return ctx().stream; }

// A masked-mode call whose operands are not contiguous share tensors (strided
// views) would take the op-by-op form, which has no masked variant.
[[noreturn]] static void refuse_masked(const char *what) {
    set_error(std::string(what) + ": the masked precompute mode needs contiguous share tensors (fused form)");
    throw PirError(ERR_UNSUPPORTED);
}

static int log2_exact(u64 d) {
    if (d == 0 || (d & (d - 1)) != 0) {
        set_error("denominator must be a power of two");
        throw PirError(ERR_UNSUPPORTED);
    }
    int k = 0;
    while (((u64)1 << k) != d) k++;
    return k;
}

// ---------------------------------------------------------------------------
// reconstruct / reshare / reconstruct3out3

// RSS.inl:457-479: out_p = 0 + sA + sB + (prev party's sA, received)
template <typename T>
void reconstruct(const RSS<T> &in, PartyData<T> &out) {
    if (fused_mode() && in.base() && out.base()) {
        check(fused_reconstruct<T>(in.base(), in.size(), out.base(), S()));
        return;
    }
    out.zero();
    MultiLaunch l0(OP_ADD), l1(OP_ADD), l2(OP_ADD);
    for (int q = 0; q < 3; q++) {
        l0.add(out[q], in.getShare(q, 0));
        l1.add(out[q], in.getShare(q, 1));
        l2.add(out[q], in.getShare(prev_party(q), 0));
    }
    l0.run(sizeof(T));
    l1.run(sizeof(T));
    l2.run(sizeof(T));
}

// RSS.inl:1052-1074: out_p = (c_p, c_{next(p)}), no mask (placeholder precompute)
template <typename T>
void reshare(const PartyData<T> &c, RSS<T> &out) {
    const PrgStream g = call_prg("reshare");
    if (sizeof(T) == 8 && fused_mode() && c.base() && out.base()) {
        check(fused_reshare((const u64 *)c.base(), c.size(), (u64 *)out.base(), S(), g));
        return;
    }
    if (g.on) refuse_masked("reshare");
    PartyData<T> mask(c.size());
    mask += c;
    out.zero();
    MultiLaunch add(OP_ADD);
    for (int q = 0; q < 3; q++) add.add(*out.getShare(q, 0), &mask[q]);
    add.run(sizeof(T));
    MultiLaunch rx(OP_SET);  // party q receives shareB from next(q)'s transmitted shareA
    for (int q = 0; q < 3; q++) rx.add(*out.getShare(q, 1), out.getShare(next_party(q), 0));
    rx.run(sizeof(T));
}

// RSS.inl:1019-1050
void reconstruct3out3(const PartyData<u64> &in, PartyData<u64> &out) {
    out.zero();
    MultiLaunch l0(OP_ADD), l1(OP_ADD), l2(OP_ADD);
    for (int q = 0; q < 3; q++) {
        l0.add(out[q], &in[q]);
        l1.add(out[q], &in[prev_party(q)]);
        l2.add(out[q], &in[next_party(q)]);
    }
    l0.run(8);
    l1.run(8);
    l2.run(8);
}

// ---------------------------------------------------------------------------
// RSS *= and &= (RSS.inl:283-345)

template <typename T>
void rss_mult_impl(RSS<T> &x, const RSS<T> &y) {
    const PrgStream g = call_prg("RSS *=");
    if (sizeof(T) == 8 && fused_mode() && x.base() && y.base()) {
        check(fused_rss_mult((u64 *)x.base(), (const u64 *)y.base(), x.size(), S(), g));
        return;
    }
    if (g.on) refuse_masked("RSS *=");
    PartyData<T> summed(y.size());
    MultiLaunch s0(OP_ADD), s1(OP_ADD), m0(OP_MUL), m1(OP_MUL), a0(OP_ADD);
    for (int q = 0; q < 3; q++) {
        s0.add(summed[q], y.getShare(q, 0));
        s1.add(summed[q], y.getShare(q, 1));
        m0.add(*x.getShare(q, 0), &summed[q]);
        m1.add(*x.getShare(q, 1), y.getShare(q, 0));
        a0.add(*x.getShare(q, 0), x.getShare(q, 1));
    }
    s0.run(sizeof(T));
    s1.run(sizeof(T));
    m0.run(sizeof(T));
    m1.run(sizeof(T));
    a0.run(sizeof(T));
    PartyData<T> c(*x.getShare(0, 0), *x.getShare(1, 0), *x.getShare(2, 0));
    reshare(c, x);
}

template <typename T>
void rss_and_impl(RSS<T> &x, const RSS<T> &y) {
    const PrgStream g = call_prg("RSS &=");
    if (fused_mode() && x.base() && y.base()) {
        check(fused_rss_and<T>(x.base(), y.base(), x.size(), S(), g));
        return;
    }
    if (g.on) refuse_masked("RSS &=");
    PartyData<T> summed(y.size());
    MultiLaunch s0(OP_XOR), s1(OP_XOR), m0(OP_AND), m1(OP_AND), a0(OP_XOR);
    for (int q = 0; q < 3; q++) {
        s0.add(summed[q], y.getShare(q, 0));
        s1.add(summed[q], y.getShare(q, 1));
        m0.add(*x.getShare(q, 0), &summed[q]);
        m1.add(*x.getShare(q, 1), y.getShare(q, 0));
        a0.add(*x.getShare(q, 0), x.getShare(q, 1));
    }
    s0.run(sizeof(T));
    s1.run(sizeof(T));
    m0.run(sizeof(T));
    m1.run(sizeof(T));
    a0.run(sizeof(T));
    PartyData<T> c(*x.getShare(0, 0), *x.getShare(1, 0), *x.getShare(2, 0));
    reshare(c, x);
}

// ---------------------------------------------------------------------------
// Truncation

// RSS.inl:422-436
void dividePublic(RSS64 &a, u64 d) {
    const PrgStream g = call_prg("dividePublic");
    if (fused_mode() && a.base()) {
        check(fused_divide_public(a.base(), a.size(), d, nullptr, S(), g));
        return;
    }
    if (g.on) refuse_masked("dividePublic");
    size_t n = a.size();
    RSS64 r(n), rPrime(n);
    rPrime.fill(d);
    r.fill(1);
    a -= rPrime;
    PartyData<u64> rec(n);
    reconstruct(a, rec);
    rec /= d;
    a.zero();
    a += r;
    a += rec;
}

// RSS.inl:438-455
void dividePublic(RSS64 &a, PartyData<u64> &den) {
    require_placeholder("dividePublic by per-party denominators");
    size_t n = a.size();
    if (fused_mode() && a.base() && den.base()) {
        check(fused_divide_public(a.base(), n, 0, den.base(), S()));
        return;
    }
    RSS64 r(n), rPrime(n);
    rPrime.zero();
    rPrime += den;
    r.fill(1);
    a -= rPrime;
    PartyData<u64> rec(n);
    reconstruct(a, rec);
    MultiLaunch dv(OP_DIV);
    for (int q = 0; q < 3; q++) dv.add(rec[q], &den[q]);
    dv.run(8);
    a.zero();
    a += r;
    a += rec;
}

// RSS.inl:1076-1094 (`a` is left modified as in the reference in plain mode)
void dividePublic3out3(PartyData<u64> &a, u64 d, RSS64 &out) {
    int k = log2_exact(d);
    const PrgStream g = call_prg("dividePublic3out3");
    if (fused_mode() && a.base() && out.base()) {
        check(fused_divide_3out3(a.base(), a.size(), k, out.base(), S(), g));
        return;
    }
    if (g.on) refuse_masked("dividePublic3out3");
    size_t n = a.size();
    RSS64 r(n), rPrime(n);
    rPrime.fill(d);
    r.fill(1);
    MultiLaunch sub(OP_SUB);
    for (int q = 0; q < 3; q++) sub.add(a[q], rPrime.getShare(q, 0));
    sub.run(8);
    PartyData<u64> rec(n);
    reconstruct3out3(a, rec);
    rec /= d;
    out.zero();
    out += r;
    out += rec;
}

// ---------------------------------------------------------------------------
// Ring products

size_t packed_rss_weights_bytes(int K, int N) { return limb_b_bytes(N, 6, round_up(K, LIMB_BK)); }

int pack_rss_operand(const RSS64 &B, const LimbBSrc &geom, int8_t *packed) {
    LimbBSrc s = geom;
    for (int p = 0; p < 3; p++) {
        s.b[2 * p][0] = B.getShare(p, 0)->data();
        s.b[2 * p][1] = B.getShare(p, 1)->data();
        s.b[2 * p + 1][0] = B.getShare(p, 0)->data();
        s.coef[2 * p][0] = s.coef[2 * p][1] = s.coef[2 * p + 1][0] = 1;
    }
    s.nseg = 6;
    return limb_pack_b(s, packed, S());
}

int pack_rss_weights(const RSS64 &B, int K, int N, bool b_row_major, int8_t *packed) {
    LimbBSrc s{};
    s.K = K;
    s.N = N;
    s.Kp = round_up(K, LIMB_BK);
    s.row_major = b_row_major ? 1 : 0;
    return pack_rss_operand(B, s, packed);
}

Epilogue rss_trunc_epi(u64 *out, int trunc_bits, bool tc) {
    Epilogue e{};
    e.mode = EPI_TRUNC;
    e.out = out;
    e.nout = 6;
    e.base_mask = (1u << (2 * PARTY_A + 0)) | (1u << (2 * PARTY_C + 1));  // A(1+q, 0), C(0, 1+q)
    e.tc = tc ? 1 : 0;
    e.trunc_bits = trunc_bits;
    return e;
}

int rss_limb_run(const RSS64 &A, LimbA a, const int8_t *packed, int N, const Epilogue &e, int split) {
    for (int p = 0; p < 3; p++)
        for (int i = 0; i < 2; i++) a.seg[2 * p + i] = A.getShare(p, i)->data();
    a.nseg = 6;
    return limb_run(a, packed, N, e, split, S());
}

static LimbA conv_view(int conv, int batch, int H, int W, int f, int Cin, int stride, int pad) {
    LimbA a{};
    a.conv = conv;
    a.row_major = 1;
    a.H = H;
    a.W = W;
    a.Cin = Cin;
    a.fh = a.fw = f;
    a.ph = a.pw = pad;
    a.stride = stride;
    a.OH = (H + 2 * pad - f) / stride + 1;
    a.OW = (W + 2 * pad - f) / stride + 1;
    return a;
}

LimbA limb_dgrad_view(int batch, int H, int W, int f, int Din, int Dout, int stride, int pad) {
    LimbA a = conv_view(LIMB_DGRAD, batch, H, W, f, Dout, stride, pad);
    a.M = batch * H * W;
    a.K = f * f * Dout;
    a.Kp = round_up(a.K, LIMB_BK);
    return a;
}

LimbA limb_wgrad_view(int batch, int H, int W, int f, int Din, int Dout, int stride, int pad) {
    LimbA a = conv_view(LIMB_WGRAD, batch, H, W, f, Din, stride, pad);
    a.M = f * f * Din;
    a.K = batch * a.OH * a.OW;
    a.Kp = round_up(a.K, LIMB_BK);
    return a;
}

int rss_limb_product(const RSS64 &A, bool ta, const ConvShape *conv, int M, int N, int K, const int8_t *packed,
                     int trunc_bits, const RSS64 *bias, RSS64 &out, bool tc, int split, u64 *relu_out,
                     u8 *relu_d, const PrgStream &g, const PrgStream &gr) {
    LimbA a{};
    a.M = M;
    a.K = K;
    a.Kp = round_up(K, LIMB_BK);
    if (conv) {
        a.conv = LIMB_FPROP;
        a.row_major = 1;
        a.H = conv->H;
        a.W = conv->W;
        a.Cin = conv->Cin;
        a.fh = a.fw = conv->f;
        a.ph = a.pw = conv->pad;
        a.stride = conv->stride;
        a.OH = conv->OH;
        a.OW = conv->OW;
    } else {
        a.lda = ta ? K : M;
        a.row_major = ta ? 1 : 0;
    }
    Epilogue e = rss_trunc_epi(out.base(), trunc_bits, tc);
    e.bias = bias ? bias->base() : nullptr;
    e.relu_out = relu_out;
    e.relu_d = relu_d;
    e.prg = g;
    e.relu_prg = gr;
    return rss_limb_run(A, a, packed, N, e, split);
}

// RSS.inl:1096-1110: three gemms per party into a scratch accumulator
void localMatMul(const RSS64 &a, const RSS64 &b, PartyData<u64> &c, int M, int N, int K, bool ta, bool tb, bool tc) {
    PartyData<u64> acc(c.size());
    const int terms[3][2] = {{0, 0}, {0, 1}, {1, 0}};
    for (int t = 0; t < 3; t++) {
        for (int q = 0; q < 3; q++)
            check(plain_gemm(M, N, K, a.getShare(q, terms[t][0])->data(), ta, b.getShare(q, terms[t][1])->data(), tb,
                             acc[q].data(), tc, S()));
        c += acc;
        if (t < 2) acc.zero();
    }
}

// RSS.inl:481-540
void matmul(const RSS64 &a, const RSS64 &b, RSS64 &c, int M, int N, int K, bool ta, bool tb, bool tc, u64 truncation) {
    const PrgStream g = call_prg("matmul");
    if (fused_mode() && a.base() && b.base() && c.base()) {
        if (ctx().algo == ALGO_LIMB) {
            DevBuf packed(packed_rss_weights_bytes(K, N));
            check(pack_rss_weights(b, K, N, tb, (int8_t *)packed.get()));
            check(rss_limb_product(a, ta, nullptr, M, N, K, (const int8_t *)packed.get(), (int)truncation, nullptr, c,
                                   tc, 0, nullptr, nullptr, g));
            return;
        }
        PartyData<u64> raw((size_t)M * N);
        localMatMul(a, b, raw, M, N, K, ta, tb, tc);
        check(fused_divide_3out3(raw.base(), raw.size(), (int)truncation, c.base(), S(), g));
        return;
    }
    if (g.on) refuse_masked("matmul");
    PartyData<u64> raw((size_t)M * N);
    localMatMul(a, b, raw, M, N, K, ta, tb, tc);
    dividePublic3out3(raw, (u64)1 << truncation, c);
}

// RSS.inl:613-639
void localFprop(const RSS64 &A, const RSS64 &B, PartyData<u64> &C, int batch, int H, int W, int Din, int Dout,
                int fh, int fw, int ph, int pw, int stride, int dil) {
    PartyData<u64> acc(C.size());
    const int terms[3][2] = {{0, 0}, {0, 1}, {1, 0}};
    for (int t = 0; t < 3; t++) {
        for (int q = 0; q < 3; q++)
            check(plain_conv_fprop(A.getShare(q, terms[t][0])->data(), B.getShare(q, terms[t][1])->data(),
                                   acc[q].data(), batch, H, W, Din, Dout, fh, fw, ph, pw, stride, dil, S()));
        C += acc;
    }
}

// RSS.inl:641-672
void localDgrad(const RSS64 &A, const RSS64 &B, PartyData<u64> &C, int batch, int OH, int OW, int Dout, int fh,
                int fw, int Din, int ph, int pw, int stride, int dil, int H, int W) {
    PartyData<u64> acc(C.size());
    const int terms[3][2] = {{0, 0}, {0, 1}, {1, 0}};
    for (int t = 0; t < 3; t++) {
        for (int q = 0; q < 3; q++)
            check(plain_conv_dgrad(A.getShare(q, terms[t][0])->data(), B.getShare(q, terms[t][1])->data(),
                                   acc[q].data(), batch, OH, OW, Dout, fh, fw, Din, ph, pw, stride, dil, H, W, S()));
        C += acc;
    }
}

// RSS.inl:674-703
void localWgrad(const RSS64 &A, const RSS64 &B, PartyData<u64> &C, int batch, int OH, int OW, int Dout, int H, int W,
                int Din, int fh, int fw, int ph, int pw, int stride, int dil) {
    PartyData<u64> acc(C.size());
    const int terms[3][2] = {{0, 0}, {0, 1}, {1, 0}};
    for (int t = 0; t < 3; t++) {
        for (int q = 0; q < 3; q++)
            check(plain_conv_wgrad(A.getShare(q, terms[t][0])->data(), B.getShare(q, terms[t][1])->data(),
                                   acc[q].data(), batch, OH, OW, Dout, H, W, Din, fh, fw, ph, pw, stride, dil, S()));
        C += acc;
    }
}

static void local_conv(ConvOperator op, const RSS64 &A, const RSS64 &B, PartyData<u64> &local, int batch, int H,
                       int W, int f, int Din, int Dout, int stride, int pad, int OH, int OW) {
    switch (op) {
        case kFprop: localFprop(A, B, local, batch, H, W, Din, Dout, f, f, pad, pad, stride, 1); break;
        case kDgrad: localDgrad(A, B, local, batch, OH, OW, Dout, f, f, Din, pad, pad, stride, 1, H, W); break;
        default: localWgrad(A, B, local, batch, OH, OW, Dout, H, W, Din, f, f, pad, pad, stride, 1); break;
    }
}

// The dgrad and wgrad products in the fused limb form: one GEMM launch over the
// six share segments with the dividePublic3out3 epilogue.
void rss_conv_grad_limb(ConvOperator op, const RSS64 &A, const RSS64 &B, RSS64 &C, int batch, int H, int W, int f,
                        int Din, int Dout, int stride, int pad, int trunc, const u8 *bw_relu_d, u64 *bw_relu_out) {
    LimbBSrc g{};
    LimbA a;
    int N;
    if (op == kDgrad) {  // A = delta (gathered), B = filter in (r, s, dout)-major order
        a = limb_dgrad_view(batch, H, W, f, Din, Dout, stride, pad);
        N = Din;
        g.dgrad_filter = 1;
        g.kfh = g.kfw = f;
        g.kc = Dout;
    } else {  // A = forward input (transposed im2col), B = delta [b*OH*OW][Dout]
        a = limb_wgrad_view(batch, H, W, f, Din, Dout, stride, pad);
        N = Dout;
        g.row_major = 1;
    }
    g.K = a.K;
    g.N = N;
    g.Kp = a.Kp;
    DevBuf packed(limb_b_bytes(N, 6, g.Kp));
    const RSS64 &Aop = op == kDgrad ? A : B;
    const RSS64 &Bop = op == kDgrad ? B : A;
    check(pack_rss_operand(Bop, g, (int8_t *)packed.get()));
    Epilogue e = rss_trunc_epi(C.base(), trunc, op == kDgrad);
    e.bw_relu_d = bw_relu_d;
    e.bw_relu_out = bw_relu_out;
    check(rss_limb_run(Aop, a, (const int8_t *)packed.get(), N, e));
}

// RSS.inl:705-803
void convolution(const RSS64 &A, const RSS64 &B, RSS64 &C, ConvOperator op, int batch, int H, int W, int f, int Din,
                 int Dout, int stride, int pad, int trunc) {
    int OH = (H + 2 * pad - f) / stride + 1, OW = (W + 2 * pad - f) / stride + 1;
    if (op != kFprop) require_placeholder("convolution (dgrad / wgrad)");
    const PrgStream g = call_prg("convolution");
    if (fused_mode() && A.base() && B.base() && C.base()) {
        if (ctx().algo == ALGO_LIMB) {
            if (op != kFprop) {
                rss_conv_grad_limb(op, A, B, C, batch, H, W, f, Din, Dout, stride, pad, trunc);
                return;
            }
            int K = f * f * Din;
            DevBuf packed(packed_rss_weights_bytes(K, Dout));
            check(pack_rss_weights(B, K, Dout, false, (int8_t *)packed.get()));
            ConvShape cs{batch, H, W, Din, f, stride, pad, OH, OW};
            check(rss_limb_product(A, true, &cs, batch * OH * OW, Dout, K, (const int8_t *)packed.get(), trunc,
                                   nullptr, C, true, 0, nullptr, nullptr, g));
            return;
        }
        PartyData<u64> local(C.size());
        local_conv(op, A, B, local, batch, H, W, f, Din, Dout, stride, pad, OH, OW);
        check(fused_divide_3out3(local.base(), local.size(), trunc, C.base(), S(), g));
        return;
    }
    if (g.on) refuse_masked("convolution");
    PartyData<u64> local(C.size());
    local_conv(op, A, B, local, batch, H, W, f, Din, Dout, stride, pad, OH, OW);
    dividePublic3out3(local, (u64)1 << trunc, C);
}

// ---------------------------------------------------------------------------
// Comparison protocols

// RSS.inl:1112-1209 on strided views
template <typename T>
void carryOut(RSS<T> &p, RSS<T> &g, int k, RSS<T> &out) {
    int stride = 2, offset = 1;
    RSS<T> pEven = p.strided(0, stride), pOdd = p.strided(offset, stride);
    RSS<T> gEven = g.strided(0, stride), gOdd = g.strided(offset, stride);
    while (k > 1) {
        gEven &= pOdd;
        pEven &= pOdd;
        gOdd ^= gEven;
        gEven = g.strided(offset, stride * 2);
        offset += stride;
        gOdd = g.strided(offset, stride * 2);
        stride *= 2;
        pEven = p.strided(0, stride);
        pOdd = p.strided(stride / 2, stride);
        k /= 2;
    }
    RSS<T> oe = out.strided(0, 2), oo = out.strided(1, 2);
    MultiLaunch ce(OP_SET), co(OP_SET);
    for (int q = 0; q < 3; q++)
        for (int i = 0; i < 2; i++) {
            ce.add(*oe.getShare(q, i), gEven.getShare(q, i));
            co.add(*oo.getShare(q, i), gOdd.getShare(q, i));
        }
    ce.run(sizeof(T));
    co.run(sizeof(T));
}

// RSS.inl:806-852
template <typename U>
void dReLU(const RSS64 &input, RSS<U> &result) {
    const PrgStream prg = call_prg("dReLU");
    if (fused_mode() && input.base() && result.base()) {
        check(fused_drelu<U>(input.base(), input.size(), result.base(), S(), prg));
        return;
    }
    if (prg.on) refuse_masked("dReLU");
    const int bw = 64;
    size_t n = input.size();
    RSS64 r(n);
    RSS<U> rbits(n * bw);
    rbits.fill(1);
    PartyData<u64> a(n);
    r += input;
    reconstruct(r, a);
    a += 1;
    PartyData<U> abits(n * bw);
    for (int q = 0; q < 3; q++) check(bitexpand(a[q].data(), n, abits[q].data(), sizeof(U), S()));
    RSS<U> msb(n);
    for (int q = 0; q < 3; q++) {
        auto cm = [&](int share, bool xr) {
            check(set_carry_out_msb(rbits.getShare(q, share)->data(), abits[q].data(), msb.getShare(q, share)->data(),
                                    n, bw, xr, sizeof(U), S()));
        };
        if (q != PARTY_C) {
            cm(0, q == PARTY_A);
            cm(1, q == PARTY_C);
        } else {
            cm(1, q == PARTY_C);
            cm(0, q == PARTY_A);
        }
    }
    RSS<U> g(n * bw);
    g += rbits;
    g &= abits;
    RSS<U> p(n * bw);
    p += rbits;
    p ^= abits;
    RSS<U> pre(result.size());
    carryOut(p, g, bw, pre);
    pre ^= msb;
    result.fill(1);
    result -= pre;
}

// RSS.inl:542-569
template <typename U>
void selectShare(const RSS64 &x, const RSS64 &y, const RSS<U> &b, RSS64 &z) {
    const PrgStream g = call_prg("selectShare");
    if (fused_mode() && x.base() && y.base() && b.base() && z.base()) {
        check(fused_select<U>(x.base(), y.base(), b.base(), z.base(), x.size(), S(), g));
        return;
    }
    if (g.on) refuse_masked("selectShare");
    size_t n = x.size();
    RSS64 c(n);
    RSS<U> cbits(b.size());
    cbits ^= b;
    PartyData<U> e(cbits.size());
    reconstruct(cbits, e);
    RSS64 d(e.size());
    check(fused_convex_comb<U>(d.base(), c.base(), e.base(), n, S()));
    RSS64 result(n);
    result += y;
    result -= x;
    result *= d;
    result += x;
    z.zero();
    z += result;
}

// RSS.inl:854-865
template <typename U>
void ReLU(const RSS64 &input, RSS64 &result, RSS<U> &dresult) {
    const PrgStream g = call_prg("ReLU");
    if (fused_mode() && input.base() && result.base() && dresult.base()) {
        check(fused_relu<U>(input.base(), input.size(), result.base(), dresult.base(), S(), g));
        return;
    }
    if (g.on) refuse_masked("ReLU");
    dReLU(input, dresult);
    RSS64 zeros(input.size());
    selectShare(zeros, input, dresult, result);
}

// RSS.inl:867-1017 (selectShare writes into the even view of `input`)
template <typename U>
void maxpool(RSS64 &input, RSS64 &result, RSS<U> &dresult, int k) {
    const PrgStream g = call_prg("maxpool");
    if (fused_mode() && input.base() && result.base() && dresult.base()) {
        check(fused_maxpool_pools<U>(input.base(), input.size(), result.base(), dresult.base(), k, S(), g));
        return;
    }
    if (g.on) refuse_masked("maxpool");
    dresult.fill(1);
    int stride = 2, offset = 1;
    RSS64 even = input.strided(0, stride), odd = input.strided(offset, stride);
    auto round = [&](bool last) {
        RSS64 diff(even.size());
        diff += even;
        diff -= odd;
        RSS<U> b(even.size());
        dReLU(diff, b);
        selectShare(odd, even, b, even);
        if (!last) {
            stride *= 2;
            even = input.strided(0, stride);
            odd = input.strided(stride / 2, stride);
        } else {
            result.zero();
            result += even;
        }
        RSS<U> negated(b.size());
        negated.fill(1);
        negated -= b;
        RSS<U> expandedB(input.size());
        for (int q = 0; q < 3; q++)
            for (int i = 0; i < 2; i++)
                check(expand_compare(b.getShare(q, i)->data(), negated.getShare(q, i)->data(),
                                     expandedB.getShare(q, i)->data(), b.size(), expandedB.size(), sizeof(U), S()));
        dresult &= expandedB;
    };
    while (k > 2) {
        round(false);
        k /= 2;
    }
    round(true);
}

// ---------------------------------------------------------------------------
// getPowers and the Taylor-series functions

// RSS.inl:1211-1235
void getPowers(const RSS64 &in, PartyData<u64> &pow) {
    require_placeholder("getPowers / taylorSeries");
    if (fused_mode() && in.base() && pow.base()) {
        check(fused_get_powers(in.base(), in.size(), pow.base(), S()));
        return;
    }
    size_t n = in.size();
    RSS64 powers(pow.size()), cur(n), diff(n);
    RSS<u8> cmp(n);
    for (int bit = 0; bit < 64; bit++) {
        cur.fill((u64)bit);
        diff.zero();
        diff += in;
        diff -= (u64)1 << bit;
        cmp.zero();
        dReLU(diff, cmp);
        selectShare(powers, cur, cmp, powers);
    }
    reconstruct(powers, pow);
}

namespace {
// The reference's GPU: `1 << x` is a 32-bit int shift whose amount the
// hardware clamps (1 << 31 is INT_MIN, 0 from 32 on); floating-point ->
// integer conversions saturate to the destination range with NaN -> 0.
int int_shift_pow2(u64 x) { return x >= 32 ? 0 : (int)(1u << x); }
u64 saturate_s64(double v) {
    if (v != v) return 0;
    if (v >= 9223372036854775808.0) return (u64)INT64_MAX;
    if (v < -9223372036854775808.0) return (u64)INT64_MIN;
    return (u64)(s64)v;
}
u64 saturate_u64(float v) {
    if (!(v > 0.0f)) return 0;
    if (v >= 18446744073709551616.0f) return ~(u64)0;
    return (u64)v;
}
// tofixed_variable_precision_functor (functors.cuh:107-116): (T)(S)(a * (1 << x))
u64 tofixed_variable_precision(double a, u64 x) { return saturate_s64(a * (double)int_shift_pow2(x)); }
// calc_fn (functors.h:22-31) with sqrt_lambda / inv_lambda / sigmoid_lambda
// (functors.h:33-49): val = 2^(x - FP) in float, (T)(fn(val) * (1 << x)).
u64 calc_fn(int fn, u64 x, int fp) {
    // pow(2, float(x) - FP) is an exact power of two, +infinity past the float range
    float val = x > (u64)(fp + 127) ? INFINITY : std::ldexp(1.0f, (int)x - fp);
    float f = fn == TAYLOR_SQRT ? std::sqrt(val) : fn == TAYLOR_INVERSE ? 1.0f / val : 1.0f / (1.0f + std::exp(-val));
    return saturate_u64(f * (float)int_shift_pow2(x));
}
}  // namespace

TaylorTables taylor_tables(double a0, double a1, double a2, int fn, int fp) {
    auto fail = [](int code, const char *msg) {
        set_error(msg);
        throw PirError(code);
    };
    if (fn < TAYLOR_SQRT || fn > TAYLOR_SIGMOID) fail(ERR_ARG, "taylorSeries: unknown function");
    if (fp < 0 || fp > 62) fail(ERR_ARG, "taylorSeries: precision must be in [0, 62]");
    TaylorTables t{};
    auto fill = [&](TaylorTab &tab, auto f) {
        for (int x = 0; x < TAYLOR_TAB; x++) tab.v[x] = f((u64)x);
        // the last entry stands for every larger argument
        for (u64 x : {(u64)TAYLOR_TAB, (u64)63, (u64)64, (u64)fp + 128, ~(u64)0})
            if (f(x) != tab.v[TAYLOR_TAB - 1]) fail(ERR_STATE, "taylorSeries: coefficient table range");
    };
    fill(t.c2, [&](u64 x) { return tofixed_variable_precision(a2, x); });
    fill(t.c1, [&](u64 x) { return tofixed_variable_precision(a1, x); });
    fill(t.c0, [&](u64 x) { return tofixed_variable_precision(a0, x); });
    fill(t.f, [&](u64 x) { return calc_fn(fn, x, fp); });
    t.fp = fp;
    t.use_a2 = a2 != 0.0;
    t.use_a1 = a1 != 0.0;
    return t;
}

// RSS.inl:1237-1320.  The coefficient functors are applied by table lookup
// (their argument, the opened power plus one, takes at most 33 distinct values).
void taylorSeries(const RSS64 &in, RSS64 &out, double a0, double a1, double a2, int fn, int fp) {
    require_placeholder("getPowers / taylorSeries");
    const TaylorTables t = taylor_tables(a0, a1, a2, fn, fp);
    const size_t n = out.size();
    if (fused_mode() && in.base() && out.base()) {
        check(fused_taylor(in.base(), n, out.base(), t, S()));
        return;
    }
    out.zero();
    RSS64 scratch(n);
    PartyData<u64> pow(n), ones(n), coeff(n);
    getPowers(in, pow);
    pow += 1;
    ones.fill(1);
    ones <<= pow;
    auto lookup = [&](const TaylorTab &tab) { check(taylor_lookup(pow.base(), 3 * n, tab, coeff.base(), S())); };
    if (t.use_a2) {
        lookup(t.c2);
        scratch.zero();
        scratch += in;
        scratch *= in;
        dividePublic(scratch, ones);
        scratch *= coeff;
        dividePublic(scratch, ones);
        out += scratch;
    }
    if (t.use_a1) {
        lookup(t.c1);
        scratch.zero();
        scratch += in;
        scratch *= coeff;
        dividePublic(scratch, ones);
        out += scratch;
    }
    lookup(t.c0);
    out += coeff;
    lookup(t.f);
    out *= coeff;
    dividePublic(out, ones);
    // back to FP fractional bits
    pow += (u64)0 - (u64)fp;
    PartyData<u64> positivePow(n), negativePow(n);
    check(filter_powers(pow.base(), 3 * n, positivePow.base(), negativePow.base(), S()));
    ones.fill(1);
    ones <<= positivePow;
    dividePublic(out, ones);
    out <<= negativePow;  // both shares (lshift_functor)
}

// RSS.inl:570-612
void rss_sqrt(const RSS64 &in, RSS64 &out, int fp) { taylorSeries(in, out, 0.424, 0.584, 0.0, TAYLOR_SQRT, fp); }
void inverse(const RSS64 &in, RSS64 &out, int fp) { taylorSeries(in, out, 4.245, -5.857, 2.630, TAYLOR_INVERSE, fp); }
void sigmoid(const RSS64 &in, RSS64 &out, int fp) {
    taylorSeries(in, out, 0.494286, 0.275589, -0.038751, TAYLOR_SIGMOID, fp);
}

// Explicit instantiations
template void reconstruct<u64>(const RSS<u64> &, PartyData<u64> &);
template void reconstruct<u8>(const RSS<u8> &, PartyData<u8> &);
template void reshare<u64>(const PartyData<u64> &, RSS<u64> &);
template void reshare<u8>(const PartyData<u8> &, RSS<u8> &);
template void rss_mult_impl<u64>(RSS<u64> &, const RSS<u64> &);
template void rss_and_impl<u64>(RSS<u64> &, const RSS<u64> &);
template void rss_and_impl<u8>(RSS<u8> &, const RSS<u8> &);
template void carryOut<u8>(RSS<u8> &, RSS<u8> &, int, RSS<u8> &);
template void carryOut<u64>(RSS<u64> &, RSS<u64> &, int, RSS<u64> &);
template void dReLU<u8>(const RSS64 &, RSS<u8> &);
template void dReLU<u64>(const RSS64 &, RSS<u64> &);
template void selectShare<u8>(const RSS64 &, const RSS64 &, const RSS<u8> &, RSS64 &);
template void selectShare<u64>(const RSS64 &, const RSS64 &, const RSS<u64> &, RSS64 &);
template void ReLU<u8>(const RSS64 &, RSS64 &, RSS<u8> &);
template void ReLU<u64>(const RSS64 &, RSS64 &, RSS<u64> &);
template void maxpool<u8>(RSS64 &, RSS64 &, RSS<u8> &, int);
template void maxpool<u64>(RSS64 &, RSS64 &, RSS<u64> &, int);

}  // namespace pir
