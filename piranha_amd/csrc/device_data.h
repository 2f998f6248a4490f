// device_data.h -- share buffers mirroring the reference's DeviceData / RSS
// classes (src/gpu/DeviceData.h:52-277, src/mpc/RSS.h:15-101), with all three
// parties held by one object.
//
//   DeviceData<T>  one device array, owning or a strided view (the reference's
//                  thrust vector / StridedRange view, StridedRange.cuh:12-52);
//   PartyData<T>   one DeviceData per party: the party-local values each
//                  reference process keeps in a DeviceData (opened values,
//                  3-of-3 products, denominators);
//   RSS<T>         the six share arrays [party][shareA|shareB]; owning RSS
//                  objects keep them in one contiguous [3][2][n] allocation,
//                  which is the layout every fused kernel reads.
//
// Every operator is one kernel launch covering all parties, ordered on the
// context stream, with the reference's party semantics (RSS.inl:110-345).
#pragma once

#include <initializer_list>
#include <vector>

#include "elementwise.h"
#include "runtime.h"

namespace pir {

template <typename T>
struct ElemBytes {
    static constexpr int value = sizeof(T);
};

template <typename T>
class DeviceData {
  public:
    DeviceData() = default;
    explicit DeviceData(size_t n) : buf_(n * sizeof(T)), ptr_((T *)buf_.get()), n_(n), stride_(1) { zero(); }
    DeviceData(T *p, size_t n, long long stride = 1) : ptr_(p), n_(n), stride_(stride) {}
    DeviceData(DeviceData &&) = default;
    DeviceData &operator=(DeviceData &&) = default;
    DeviceData(const DeviceData &o) : ptr_(o.ptr_), n_(o.n_), stride_(o.stride_) {}  // copies are views
    DeviceData &operator=(const DeviceData &o) {
        buf_ = DevBuf();
        ptr_ = o.ptr_;
        n_ = o.n_;
        stride_ = o.stride_;
        return *this;
    }

    T *data() const { return ptr_; }
    size_t size() const { return n_; }
    long long stride() const { return stride_; }
    bool contiguous() const { return stride_ == 1; }

    // StridedRange(begin() + first, end(), stride) (StridedRange.cuh:36-38)
    DeviceData view(size_t first, size_t stride) const {
        size_t cnt = n_ > first ? ((n_ - first) + stride - 1) / stride : 0;
        return DeviceData(ptr_ + (long long)first * stride_, cnt, stride_ * (long long)stride);
    }

    void resize(size_t n) {
        buf_ = DevBuf(n * sizeof(T));
        ptr_ = (T *)buf_.get();
        n_ = n;
        stride_ = 1;
        zero();
    }

    void zero() { fill((T)0); }
    void fill(T v) {
        if (contiguous()) {
            if (sizeof(T) == 8) check(fill_u64((u64 *)ptr_, (u64)v, n_, ctx().stream));
            else check(fill_u8((u8 *)ptr_, (u8)v, n_, ctx().stream));
        } else {
            check(dd_op(OP_SET, sizeof(T), ptr_, stride_, nullptr, 0, (u64)v, n_, ctx().stream));
        }
    }

    void copy_from_host(const T *h) {
        check_contig();
        if (n_) check_hip(hipMemcpyAsync(ptr_, h, n_ * sizeof(T), hipMemcpyHostToDevice, ctx().stream), true);
    }
    void copy_to_host(T *h) const {
        check_contig();
        if (n_) check_hip(hipMemcpyAsync(h, ptr_, n_ * sizeof(T), hipMemcpyDeviceToHost, ctx().stream), true);
    }

    // DeviceData.h:131-209 -- scalar and element-wise overloads
    DeviceData &operator+=(T v) { return op_s(OP_ADD, v); }
    DeviceData &operator-=(T v) { return op_s(OP_SUB, v); }
    DeviceData &operator*=(T v) { return op_s(OP_MUL, v); }
    DeviceData &operator/=(T v) { return op_s(OP_DIV, v); }
    DeviceData &operator>>=(T v) { return op_s(OP_SHR, v); }
    DeviceData &operator<<=(T v) { return op_s(OP_SHL, v); }
    DeviceData &operator+=(const DeviceData &r) { return op_v(OP_ADD, r); }
    DeviceData &operator-=(const DeviceData &r) { return op_v(OP_SUB, r); }
    DeviceData &operator*=(const DeviceData &r) { return op_v(OP_MUL, r); }
    DeviceData &operator^=(const DeviceData &r) { return op_v(OP_XOR, r); }
    DeviceData &operator&=(const DeviceData &r) { return op_v(OP_AND, r); }
    DeviceData &operator>>=(const DeviceData &r) { return op_v(OP_SHR, r); }
    DeviceData &operator<<=(const DeviceData &r) { return op_v(OP_SHL, r); }

  private:
    DeviceData &op_s(int op, T v) {
        check(dd_op(op, sizeof(T), ptr_, stride_, nullptr, 0, (u64)v, n_, ctx().stream));
        return *this;
    }
    DeviceData &op_v(int op, const DeviceData &r) {
        check(dd_op(op, sizeof(T), ptr_, stride_, r.ptr_, r.stride_, 0, n_, ctx().stream));
        return *this;
    }
    void check_contig() const {
        if (!contiguous()) {
            set_error("host copy of a strided DeviceData view");
            throw PirError(ERR_ARG);
        }
    }
    static void check_hip(hipError_t e, bool sync) {
        if (e == hipSuccess && sync) e = hipStreamSynchronize(ctx().stream);
        if (e != hipSuccess) {
            set_error(std::string("hip: ") + hipGetErrorString(e));
            throw PirError(ERR_HIP);
        }
    }

    DevBuf buf_;
    T *ptr_ = nullptr;
    size_t n_ = 0;
    long long stride_ = 1;
};

template <typename T>
struct NoDeduce {
    typedef T type;
};

// Builds one multi-array launch.
class MultiLaunch {
  public:
    explicit MultiLaunch(int op) {
        m_ = MultiOp{};
        m_.op = op;
    }
    template <typename T>
    void add(DeviceData<T> &x, const typename NoDeduce<DeviceData<T>>::type *y, u64 scalar = 0) {
        if (m_.count == MULTI_OP_MAX) {
            set_error("MultiLaunch: more than 12 arrays in one launch");
            throw PirError(ERR_STATE);
        }
        if (m_.count > 0 && x.size() != m_.n) {  // one element count per launch
            set_error("MultiLaunch: arrays of different lengths in one launch");
            throw PirError(ERR_STATE);
        }
        int a = m_.count++;
        m_.x[a] = x.data();
        m_.xs[a] = x.stride();
        m_.y[a] = y ? y->data() : nullptr;
        m_.ys[a] = y ? y->stride() : 0;
        m_.scalar[a] = scalar;
        m_.n = x.size();
    }
    void run(int elem_bytes) {
        if (m_.count) check(multi_op(m_, elem_bytes, ctx().stream));
    }

  private:
    MultiOp m_;
};

template <typename T>
class PartyData {
  public:
    PartyData() = default;
    explicit PartyData(size_t n) { resize(n); }
    PartyData(const DeviceData<T> &a, const DeviceData<T> &b, const DeviceData<T> &c) {
        p_[0] = a;
        p_[1] = b;
        p_[2] = c;
        n_ = a.size();
    }
    // View over an external contiguous [3][n] array.
    static PartyData wrap(T *base, size_t n) {
        PartyData r;
        r.n_ = n;
        for (int q = 0; q < 3; q++) r.p_[q] = DeviceData<T>(base + q * n, n);
        r.ext_base_ = base;
        return r;
    }
    void resize(size_t n) {
        n_ = n;
        buf_ = DevBuf(3 * n * sizeof(T));
        ext_base_ = nullptr;
        for (int q = 0; q < 3; q++) p_[q] = DeviceData<T>((T *)buf_.get() + q * n, n);
        zero();
    }
    size_t size() const { return n_; }
    DeviceData<T> &operator[](int q) { return p_[q]; }
    const DeviceData<T> &operator[](int q) const { return p_[q]; }
    // [3][n] base pointer when the three arrays are contiguous, else null.
    T *base() const { return buf_.get() ? (T *)buf_.get() : ext_base_; }
    void zero() { fill((T)0); }
    void fill(T v) {
        MultiLaunch l(OP_SET);
        for (int q = 0; q < 3; q++) l.add(p_[q], nullptr, (u64)v);
        l.run(sizeof(T));
    }
    PartyData &operator+=(T v) { return ops(OP_ADD, v); }
    PartyData &operator/=(T v) { return ops(OP_DIV, v); }
    PartyData &operator<<=(const PartyData &r) { return opv(OP_SHL, r); }
    PartyData &operator+=(const PartyData &r) { return opv(OP_ADD, r); }
    PartyData &operator-=(const PartyData &r) { return opv(OP_SUB, r); }

  private:
    PartyData &ops(int op, T v) {
        MultiLaunch l(op);
        for (int q = 0; q < 3; q++) l.add(p_[q], nullptr, (u64)v);
        l.run(sizeof(T));
        return *this;
    }
    PartyData &opv(int op, const PartyData &r) {
        MultiLaunch l(op);
        for (int q = 0; q < 3; q++) l.add(p_[q], &r.p_[q]);
        l.run(sizeof(T));
        return *this;
    }
    DevBuf buf_;
    DeviceData<T> p_[3];
    size_t n_ = 0;
    T *ext_base_ = nullptr;
};

// Fixed-point encode (RSS.inl:126, util.cuh:28-33): (T)(v * (1 << FP)); on
// x86-64 a double below 2^63 converts through a signed 64-bit truncation.
inline u64 to_fixed(double v, int fp) { return (u64)(s64)(v * (double)(1 << fp)); }

template <typename T>
class RSS {
  public:
    enum Party { PARTY_A = 0, PARTY_B = 1, PARTY_C = 2 };
    static const int numParties = 3;
    static int numShares() { return 2; }
    static int nextParty(int p) { return next_party(p); }
    static int prevParty(int p) { return prev_party(p); }

    RSS() = default;
    explicit RSS(size_t n) { resize(n); }
    // Initializer-list constructor, RSS.inl:388-414: A.sA = C.sB = value.
    RSS(std::initializer_list<double> il, bool convertToFixedPoint = true);
    // View over six arrays (e.g. strided views for carryOut / maxpool).
    RSS(const DeviceData<T> (&s)[3][2]) {
        for (int q = 0; q < 3; q++)
            for (int i = 0; i < 2; i++) s_[q][i] = s[q][i];
        n_ = s[0][0].size();
    }
    // View over an external contiguous [3][2][n] array.
    static RSS wrap(T *base, size_t n) {
        RSS r;
        r.n_ = n;
        for (int q = 0; q < 3; q++)
            for (int i = 0; i < 2; i++) r.s_[q][i] = DeviceData<T>(base + (2 * q + i) * n, n);
        r.ext_base_ = base;
        return r;
    }
    RSS(RSS &&) = default;
    RSS &operator=(RSS &&) = default;

    void resize(size_t n) {
        n_ = n;
        buf_ = DevBuf(6 * n * sizeof(T));
        ext_base_ = nullptr;
        for (int q = 0; q < 3; q++)
            for (int i = 0; i < 2; i++) s_[q][i] = DeviceData<T>((T *)buf_.get() + (2 * q + i) * n, n);
        zero();
    }

    size_t size() const { return n_; }
    DeviceData<T> *getShare(int party, int i) { return &s_[party][i]; }
    const DeviceData<T> *getShare(int party, int i) const { return &s_[party][i]; }
    // [3][2][n] base pointer when the six arrays are contiguous, else null.
    T *base() const {
        if (buf_.get()) return (T *)buf_.get();
        return ext_base_;
    }
    RSS strided(size_t first, size_t stride) const {
        DeviceData<T> v[3][2];
        for (int q = 0; q < 3; q++)
            for (int i = 0; i < 2; i++) v[q][i] = s_[q][i].view(first, stride);
        return RSS(v);
    }

    void zero() {
        MultiLaunch l(OP_SET);
        for (int q = 0; q < 3; q++)
            for (int i = 0; i < 2; i++) l.add(s_[q][i], (const DeviceData<T> *)nullptr, 0);
        l.run(sizeof(T));
    }
    // RSS.inl:116-120
    void fill(T v) {
        MultiLaunch l(OP_SET);
        for (int q = 0; q < 3; q++)
            for (int i = 0; i < 2; i++) {
                bool on = (q == PARTY_A && i == 0) || (q == PARTY_C && i == 1);
                l.add(s_[q][i], (const DeviceData<T> *)nullptr, on ? (u64)v : 0);
            }
        l.run(sizeof(T));
    }
    // setPublic, RSS.inl:122-142 (party B keeps the missing-break result (0, v)).
    void setPublic(const std::vector<double> &v, int fp);
    void setPublicFixed(const std::vector<T> &fixed);

    // scalar ops: + - on A.sA and C.sB (RSS.inl:168-186); * >> on all shares
    RSS &operator+=(T v) { return sc(OP_ADD, v, false); }
    RSS &operator-=(T v) { return sc(OP_SUB, v, false); }
    RSS &operator*=(T v) { return sc(OP_MUL, v, true); }
    RSS &operator>>=(T v) { return sc(OP_SHR, v, true); }
    // party-local rhs: + - ^ on A.sA / C.sB, * & >> << on both (RSS.inl:202-265)
    RSS &operator+=(const PartyData<T> &r) { return lo(OP_ADD, r, false); }
    RSS &operator-=(const PartyData<T> &r) { return lo(OP_SUB, r, false); }
    RSS &operator^=(const PartyData<T> &r) { return lo(OP_XOR, r, false); }
    RSS &operator*=(const PartyData<T> &r) { return lo(OP_MUL, r, true); }
    RSS &operator&=(const PartyData<T> &r) { return lo(OP_AND, r, true); }
    RSS &operator>>=(const PartyData<T> &r) { return lo(OP_SHR, r, true); }
    RSS &operator<<=(const PartyData<T> &r) { return lo(OP_SHL, r, true); }
    // RSS rhs: + - ^ share-wise (RSS.inl:267-328); * and & are protocols
    RSS &operator+=(const RSS &r) { return rs(OP_ADD, r); }
    RSS &operator-=(const RSS &r) { return rs(OP_SUB, r); }
    RSS &operator^=(const RSS &r) { return rs(OP_XOR, r); }
    RSS &operator*=(const RSS &r);  // RSS.inl:283-320
    RSS &operator&=(const RSS &r);  // RSS.inl:330-345

    void copy_to_host(T *h) const;    // [3][2][n]
    void copy_from_host(const T *h);  // [3][2][n]

  private:
    RSS &sc(int op, T v, bool both) {
        MultiLaunch l(op);
        for (int q = 0; q < 3; q++)
            for (int i = 0; i < 2; i++)
                if (both || (q == PARTY_A && i == 0) || (q == PARTY_C && i == 1))
                    l.add(s_[q][i], (const DeviceData<T> *)nullptr, (u64)v);
        l.run(sizeof(T));
        return *this;
    }
    RSS &lo(int op, const PartyData<T> &r, bool both) {
        MultiLaunch l(op);
        for (int q = 0; q < 3; q++)
            for (int i = 0; i < 2; i++)
                if (both || (q == PARTY_A && i == 0) || (q == PARTY_C && i == 1)) l.add(s_[q][i], &r[q]);
        l.run(sizeof(T));
        return *this;
    }
    RSS &rs(int op, const RSS &r) {
        MultiLaunch l(op);
        for (int q = 0; q < 3; q++)
            for (int i = 0; i < 2; i++) l.add(s_[q][i], &r.s_[q][i]);
        l.run(sizeof(T));
        return *this;
    }

    DevBuf buf_;
    T *ext_base_ = nullptr;
    DeviceData<T> s_[3][2];
    size_t n_ = 0;
};

}  // namespace pir
