// Encode / Decode of every reference encoding.  Each function follows the
// reference encoding's host code (file:line cited) for output arity, sizes,
// empty-input behaviour and metadata layout; the per-row work runs in the
// kernels of k_*.hip.
#include "codecs.h"

#include <algorithm>
#include <cmath>
#include <limits>
#include <type_traits>

#include "dispatch.h"
#include "k_common.h"

namespace ddj {

Part own(DevBuf&& b) {
    auto sp = std::make_shared<DevBuf>(std::move(b));
    Part p;
    p.owner = sp;
    p.p = sp->bytes_ptr();
    p.n = sp->size();
    return p;
}

Part host_part(const void* h, size_t n) {
    Part p;
    auto v = std::make_shared<std::vector<uint8_t>>(n);
    if (n) memcpy(v->data(), h, n);
    p.host = std::move(v);
    p.n = n;
    return p;
}

Part borrow(DView v, std::shared_ptr<DevBuf> keep) {
    Part p;
    p.owner = std::move(keep);
    p.p = v.p;
    p.n = v.n;
    return p;
}

namespace {

Part empty() { return Part{}; }


template <typename T> T read_dev(Ctx* c, const uint8_t* p) {
    T v;
    read_small(c, &v, p, sizeof(T));
    return v;
}

// A copy at an aligned address when `v` is not aligned to `align`.
Part aligned_view(Ctx* c, DView v, size_t align) {
    if ((reinterpret_cast<uintptr_t>(v.p) % align) == 0) return borrow(v);
    DevBuf b(c, v.n);
    d2d(c, b.get(), v.p, v.n);
    return own(std::move(b));
}

// Integer view used by delta / scale for floating types
// (reference delta_encoding.hpp:57-64, scale_encoding.hpp:52-59).
int int_view(int dtype) {
    if (dtype == DDJ_D_FLOAT) return DDJ_D_INT;
    if (dtype == DDJ_D_DOUBLE) return DDJ_D_TIME;
    return dtype;
}

// n = floor(bytes / sizeof(T)): a CudaPtr<T> over a char buffer drops a
// trailing partial element (reference core/cuda_ptr.hpp:60).
size_t rows(int dtype, size_t bytes) { return bytes / dtype_size(dtype); }

template <typename T> uint64_t tbits(T v) {
    using U = std::make_unsigned_t<std::conditional_t<std::is_floating_point<T>::value,
                                                      std::conditional_t<sizeof(T) == 8, int64_t, int32_t>, T>>;
    U u;
    memcpy(&u, &v, sizeof(T));
    return (uint64_t)u;
}

struct MinMax {
    DevBuf buf;
    MinMaxResult host;
};
MinMax minmax(Ctx* c, int dtype, const void* x, size_t n) {
    if (dtype != DDJ_D_FLOAT && dtype != DDJ_D_DOUBLE) return MinMax{DevBuf(), minmax_to_host(c, dtype, x, n, false)};
    MinMax m{DevBuf(c, sizeof(MinMaxResult)), {}};
    launch_minmax(c, dtype, x, n, m.buf.get<MinMaxResult>());
    d2h(c, &m.host, m.buf.get(), sizeof(MinMaxResult));
    return m;
}

// ------------------------------------------------------------------- none
// NoneEncoding (reference compression/none/none_encoding.cu:13-49)
std::vector<Part> none_encode(Ctx* c, DView in) {
    size_t len = in.n;
    DevBuf d(c, in.n);
    d2d(c, d.get(), in.p, in.n);
    return {host_part(&len, sizeof len), own(std::move(d))};
}

// ------------------------------------------------------------------ delta
// DeltaEncoding (compression/delta/delta_encoding.cu:29-102)
std::vector<Part> delta_encode(Ctx* c, int dtype, DView in) {
    const size_t es = dtype_size(dtype), n = rows(dtype, in.n);
    if (n == 0) return {empty(), empty()};
    Part x = aligned_view(c, in, es);
    DevBuf data(c, (n - 1) * es);
    launch_delta_encode(c, (int)es, x.p, n, data.get());
    kernel_bytes(c, "k_delta_encode", (double)((2 * n - 1) * es));
    return {borrow(DView{x.p, es}, x.owner), own(std::move(data))};  // metadata: the first row, in place
}
Part delta_decode(Ctx* c, int dtype, const std::vector<DView>& in) {
    const size_t es = dtype_size(dtype);
    if (in[0].n == 0) return empty();
    if (in[0].n < es) fail(DDJ_ECORRUPT, "delta metadata shorter than one value");
    const size_t m = in[1].n / es;
    DevBuf out(c, (m + 1) * es);
    launch_delta_decode(c, (int)es, in[0].p, in[1].p, m, out.get());
    kernel_bytes(c, "k_delta_decode", (double)((2 * m + 1) * es));
    return own(std::move(out));
}

// ------------------------------------------------------------------ scale
// ScaleEncoding (compression/scale/scale_encoding.cu:30-108): m = min, but 0
// when the overflow guard (:49-50) fires; its std::abs<T> leaves T's minimum
// unchanged and T_MAX - max wraps for int/long.
template <typename T> T scale_min(T mn, T mx) {
    T m = mn;
    if (mn < 0) {
        long long tmax = (long long)std::numeric_limits<T>::max();
        long long diff;
        if (sizeof(T) >= 4)
            diff = (long long)(T)((uint64_t)(std::make_unsigned_t<T>)tmax - (uint64_t)(std::make_unsigned_t<T>)mx);
        else
            diff = tmax - (long long)mx;
        long long absmin = (mn == std::numeric_limits<T>::min()) ? (long long)mn : -(long long)mn;
        if (diff < absmin) m = 0;
    }
    return m;
}
std::vector<Part> scale_encode(Ctx* c, int dtype, DView in) {
    const int iv = int_view(dtype);
    const size_t es = dtype_size(dtype), n = rows(dtype, in.n);
    if (n == 0) return {empty(), empty()};
    Part x = aligned_view(c, in, es);
    MinMax mm = minmax(c, iv, x.p, n);
    return dispatch_int((int)es, [&](auto tag) -> std::vector<Part> {
        using T = decltype(tag);
        T mn, mx;
        decode_minmax(iv, mm.host, &mn, &mx);
        T m = scale_min(mn, mx);
        DevBuf data(c, n * es);
        launch_sub_scalar(c, (int)es, x.p, n, tbits(m), data.get());
        return {host_part(&m, sizeof m), own(std::move(data))};
    });
}
Part scale_decode(Ctx* c, int dtype, const std::vector<DView>& in) {
    const size_t es = dtype_size(dtype);
    if (in[1].n == 0) return empty();
    if (in[0].n < es) fail(DDJ_ECORRUPT, "scale metadata shorter than one value");
    const size_t n = in[1].n / es;
    uint64_t m = 0;
    read_small(c, &m, in[0].p, es);
    Part x = aligned_view(c, in[1], es);
    DevBuf out(c, n * es);
    launch_add_scalar(c, (int)es, x.p, n, m, out.get());
    return own(std::move(out));
}

// -------------------------------------------------------------------- afl
// AflEncoding::Encode<T> (compression/afl/afl_encoding.cu:15-56): metadata
// {minBit, rest, 0, 0}; data = ceil(n*b/W) words of T.
int bits_of(int dtype, const MinMaxResult& h) {
    return dispatch_int((int)dtype_size(dtype), [&](auto tag) {
        using T = decltype(tag);
        T mn, mx;
        decode_minmax(dtype, h, &mn, &mx);
        return bitlen_of<T>(mn, mx);
    });
}

// The FL data of n values at b bits: ceil(n*b/W) words of T.
uint64_t fl_bytes(size_t es, int b, size_t n) {
    const uint64_t W = 8 * es;
    return (b * (uint64_t)n + W - 1) / W * es;
}

std::vector<Part> afl_encode_int(Ctx* c, int dtype, DView in, const MinMaxResult* known = nullptr,
                                 bool defer = false, const MinMaxResult* known_host = nullptr) {
    const size_t es = dtype_size(dtype), n = rows(dtype, in.n);
    if (n == 0) return {empty(), empty()};
    Part x = aligned_view(c, in, es);
    if (known_host) {  // the width is known on the host: no min/max pass, no read-back
        const int b = bits_of(dtype, *known_host);
        Part data;
        if (defer) {
            auto job = std::make_shared<FlJob>();
            job->dtype = dtype;
            job->bits = b;
            job->n = n;
            job->x = x;
            data.n = (size_t)fl_bytes(es, b, n);
            data.fl = std::move(job);
        } else {
            DevBuf buf(c, ((size_t)fl_bytes(es, b, n) + 3) & ~size_t(3));
            launch_fl_pack(c, dtype, x.p, n, nullptr, b, buf.get(), buf.size());
            buf.set_size((size_t)fl_bytes(es, b, n));
            kernel_bytes(c, "k_fl_pack", (double)(n * es + buf.size()));
            data = own(std::move(buf));
        }
        const uint64_t bytes = data.n;
        uint8_t meta[4] = {(uint8_t)b, (uint8_t)((bytes * 8 - (uint64_t)n * b) & 0xff), 0, 0};
        return {host_part(meta, 4), data};
    }
    DevBuf mmb;
    const MinMaxResult* mm = known;
    MinMaxResult h;
    if (!mm && defer) {  // the width only, folded on the host
        h = minmax_to_host(c, dtype, x.p, n, false);
        kernel_bytes(c, "k_minmax", (double)(n * es));
    } else if (!mm) {
        mmb = DevBuf(c, sizeof(MinMaxResult));
        launch_minmax(c, dtype, x.p, n, mmb.get<MinMaxResult>());
        kernel_bytes(c, "k_minmax", (double)(n * es));
        mm = mmb.get<MinMaxResult>();
    }
    const uint64_t W = 8 * es;
    Part data;
    if (defer) {  // the width first; the pack runs when the blob is assembled
        if (mm) d2h(c, &h, mm, sizeof h);
        auto job = std::make_shared<FlJob>();
        job->dtype = dtype;
        job->bits = bits_of(dtype, h);
        job->n = n;
        job->x = x;
        data.n = (size_t)fl_bytes(es, job->bits, n);
        data.fl = std::move(job);
    } else {
        DevBuf buf(c, (size_t)((W * n + W - 1) / W) * es);  // worst case b = W
        launch_fl_pack(c, dtype, x.p, n, mm, 0, buf.get());
        d2h(c, &h, mm, sizeof h);
        buf.set_size((size_t)fl_bytes(es, bits_of(dtype, h), n));
        kernel_bytes(c, "k_fl_pack", (double)(n * es + buf.size()));
        data = own(std::move(buf));
    }
    const int b = bits_of(dtype, h);
    const uint64_t bytes = data.n;
    uint8_t meta[4] = {(uint8_t)b, (uint8_t)((bytes * 8 - (uint64_t)n * b) & 0xff), 0, 0};
    return {host_part(meta, 4), data};
}

Part fl_decode_raw(Ctx* c, int dtype, const uint8_t* data, size_t bytes, int b, int rest) {
    const int W = 8 * (int)dtype_size(dtype);
    if (b <= 0 || b > W) fail(DDJ_ECORRUPT, "AFL bit width out of range");
    long long bits = (long long)bytes * 8 - rest;
    if (bits < 0) fail(DDJ_ECORRUPT, "AFL rest larger than the data");
    const size_t n = (size_t)(bits / b);
    DevBuf out(c, n * dtype_size(dtype));
    launch_fl_unpack(c, dtype, data, bytes, n, b, out.get());
    kernel_bytes(c, "k_fl_unpack", (double)((n * b + 7) / 8 + n * dtype_size(dtype)));
    return own(std::move(out));
}

// DecodeAfl / AflEncoding::Decode<T> (afl_encoding.cu:74-113)
Part afl_decode_int(Ctx* c, int dtype, const std::vector<DView>& in) {
    if (in[1].n == 0) return empty();
    if (in[0].n < 2) fail(DDJ_ECORRUPT, "AFL metadata shorter than 2 bytes");
    int8_t m[2];
    read_small(c, m, in[0].p, 2);
    return fl_decode_raw(c, dtype, in[1].p, in[1].n, m[0], m[1]);
}

// AflEncoding::Encode<float> (afl_encoding.cu:115-167)
std::vector<Part> afl_encode_float(Ctx* c, DView in) {
    const size_t n = rows(DDJ_D_FLOAT, in.n);
    if (n == 0) return {empty(), empty()};
    Part x = aligned_view(c, in, 4);
    MinMax mm = minmax(c, DDJ_D_FLOAT, x.p, n);
    float mn, mx;
    decode_minmax(DDJ_D_FLOAT, mm.host, &mn, &mx);
    const char sign = (char)((mn >= 0 ? 1 : 0) + (mx < 0 ? 2 : 0));
    DevBuf mant(c, 4 * n), expo(c, 4 * n), sgn(c, 4 * n);
    launch_float_split(c, reinterpret_cast<const float*>(x.p), n, mant.get<int32_t>(), expo.get<int32_t>(),
                       sgn.get<int32_t>());
    auto a = afl_encode_int(c, DDJ_D_INT, DView{mant.bytes_ptr(), 4 * n});
    auto b = afl_encode_int(c, DDJ_D_INT, DView{expo.bytes_ptr(), 4 * n});
    size_t msize = a[1].n;
    const size_t stencil_bytes = sign == 0 ? 1 + (n + 7) / 8 : 0;
    DevBuf meta(c, sizeof(size_t) + 1 + stencil_bytes);
    uint8_t head[sizeof(size_t) + 1];
    memcpy(head, &msize, sizeof(size_t));
    head[sizeof(size_t)] = (uint8_t)sign;
    h2d(c, meta.get(), head, sizeof head);
    if (sign == 0) {
        PredArgs p{};
        p.kind = PRED_NOT_EQUAL;
        p.a_bits = 0;
        DevBuf cnt(c, 8);
        launch_split(c, DDJ_D_INT, sgn.get(), n, p, meta.bytes_ptr() + sizeof(size_t) + 1, nullptr, nullptr, nullptr,
                     cnt.get<uint64_t>());
    }
    DevBuf data(c, a[0].n + a[1].n + b[0].n + b[1].n);
    size_t off = 0;
    for (const Part* q : {&a[0], &a[1], &b[0], &b[1]}) {
        if (q->on_host())
            h2d(c, data.bytes_ptr() + off, q->host->data(), q->n);
        else
            d2d(c, data.bytes_ptr() + off, q->p, q->n);
        off += q->n;
    }
    return {own(std::move(meta)), own(std::move(data))};
}

// AflEncoding::Decode<float> (afl_encoding.cu:185-254)
Part afl_decode_float(Ctx* c, const std::vector<DView>& in) {
    if (in[1].n == 0) return empty();
    if (in[0].n < sizeof(size_t) + 1 || in[1].n < 8) fail(DDJ_ECORRUPT, "AFL float metadata too short");
    uint8_t head[sizeof(size_t) + 1];
    read_small(c, head, in[0].p, sizeof head);
    size_t msize;
    memcpy(&msize, head, sizeof(size_t));
    const int sign = (int8_t)head[sizeof(size_t)];
    if (msize + 8 > in[1].n) fail(DDJ_ECORRUPT, "AFL float mantissa size out of range");
    int8_t ma[2], ea[2];
    read_small(c, ma, in[1].p, 2);
    const uint8_t* eb = in[1].p + 4 + msize;
    read_small(c, ea, eb, 2);
    Part mant = fl_decode_raw(c, DDJ_D_INT, in[1].p + 4, msize, ma[0], ma[1]);
    Part expo = fl_decode_raw(c, DDJ_D_INT, eb + 4, in[1].n - msize - 8, ea[0], ea[1]);
    const size_t n = mant.n / 4;
    if (expo.n / 4 < n) fail(DDJ_ECORRUPT, "AFL float exponent shorter than mantissa");
    const uint8_t* bits = nullptr;
    if (sign == 0) {
        if (in[0].n < sizeof(size_t) + 2 + (n + 7) / 8) fail(DDJ_ECORRUPT, "AFL float sign stencil too short");
        bits = in[0].p + sizeof(size_t) + 2;
    }
    DevBuf out(c, 4 * n);
    launch_float_compose(c, reinterpret_cast<const int32_t*>(mant.p), reinterpret_cast<const int32_t*>(expo.p), bits,
                         sign - 1, n, out.get<float>());
    return own(std::move(out));
}

// -------------------------------------------------------------------- rle
// RleEncoding (compression/rle/rle_encode.cu:18-62): metadata = int32 run
// count (also for empty input), lengths int32[R], values T[R].
std::vector<Part> rle_encode(Ctx* c, int dtype, DView in) {
    const size_t es = dtype_size(dtype), n = rows(dtype, in.n);
    int32_t R = 0;
    if (n == 0) return {host_part(&R, 4), empty(), empty()};
    Part x = aligned_view(c, in, es);
    DevBuf lens(c, 4 * n), vals(c, es * n), runs(c, 8);
    // the run count comes back with both outputs' min/max (integer values):
    // FL encoders of the outputs then need no pass or read-back of their own
    const bool integral = dtype != DDJ_D_FLOAT && dtype != DDJ_D_DOUBLE;
    auto hl = std::make_shared<MinMaxResult>(), hv = std::make_shared<MinMaxResult>();
    const uint64_t r = launch_rle_encode(c, dtype, x.p, n, lens.get<int32_t>(), vals.get(), runs.get<uint64_t>(),
                                         integral ? hv.get() : nullptr, hl.get());
    R = (int32_t)r;
    kernel_bytes(c, "k_rle_encode", (double)(n * es + r * (4 + es)));
    lens.set_size(4 * r);
    vals.set_size(es * r);
    Part pl = own(std::move(lens)), pv = own(std::move(vals));
    if (r > 0) {
        pl.minmax_host = hl;
        if (integral) pv.minmax_host = hv;
    }
    return {host_part(&R, 4), pl, pv};
}
// rle_decode.cu:19-73
Part rle_decode(Ctx* c, int dtype, const std::vector<DView>& in) {
    const size_t es = dtype_size(dtype);
    if (in[1].n == 0) return empty();
    if (in[0].n < 4) fail(DDJ_ECORRUPT, "RLE metadata shorter than 4 bytes");
    const int32_t R = read_dev<int32_t>(c, in[0].p);
    if (R < 0 || (size_t)R * 4 > in[1].n || (size_t)R * es > in[2].n) fail(DDJ_ECORRUPT, "RLE run count out of range");
    if (R == 0) return empty();
    return own(rle_decode_dev(c, dtype, in[1].p, in[2].p, (size_t)R));
}

// ------------------------------------------------------- stencil helpers
size_t stencil_rows(Ctx* c, DView st) {
    // Stencil::unpack (util/stencil/stencil.cu:69-96)
    if (st.n < 1) fail(DDJ_ECORRUPT, "empty stencil");
    const int8_t rest = read_dev<int8_t>(c, st.p);
    long long n = (long long)(st.n - 1) * 8;
    if (rest) n -= 8 - rest;
    if (n < 0) fail(DDJ_ECORRUPT, "stencil length out of range");
    return (size_t)n;
}

// ------------------------------------------------------------------ patch
// PatchEncodingFactory::Get(data) + LowerPatchEncoding / OutsidePatchEncoding
// (compression/patch/patch_encoding_factory.hpp:40-57, lower_patch_encoding.cu:22-66,
// outside_patch_encoding.cu:22-60)
PredArgs patch_pred(Ctx* c, int dtype, const uint8_t* x, size_t n, bool outside) {
    MinMax mm = minmax(c, dtype, x, n);
    PredArgs p{};
    dispatch_native(dtype, [&](auto tag) {
        using T = decltype(tag);
        T mnT, mxT;
        decode_minmax(dtype, mm.host, &mnT, &mxT);
        volatile double mn = (double)mnT, mx = (double)mxT, factor = 0.1;
        const T dist = (T)(mx - mn);
        volatile double prod = factor * (double)dist;
        if (!outside) {
            p.kind = PRED_LOWER;
            p.a_bits = tbits((T)(mx - prod));
        } else {
            p.kind = PRED_INSIDE;
            p.a_bits = tbits((T)(mn + prod));
            p.b_bits = tbits((T)(mx - prod));
        }
    });
    kernel_bytes(c, "k_minmax", (double)(n * dtype_size(dtype)));
    return p;
}

std::vector<Part> patch_encode(Ctx* c, int dtype, DView in, bool outside) {
    const size_t es = dtype_size(dtype), n = rows(dtype, in.n);
    if (n == 0) return {empty(), empty(), empty()};
    Part x = aligned_view(c, in, es);
    const PredArgs p = patch_pred(c, dtype, x.p, n, outside);
    DevBuf st(c, 1 + (n + 7) / 8), yes(c, n * es), no(c, n * es), cnt(c, 8);
    // integer outputs carry their min/max (from the split's counting pass)
    const bool integral = dtype != DDJ_D_FLOAT && dtype != DDJ_D_DOUBLE;
    auto mm_yes = integral ? std::make_shared<DevBuf>(c, sizeof(MinMaxResult)) : nullptr;
    auto mm_no = integral ? std::make_shared<DevBuf>(c, sizeof(MinMaxResult)) : nullptr;
    launch_split(c, dtype, x.p, n, p, st.bytes_ptr(), yes.get(), no.get(), nullptr, cnt.get<uint64_t>(),
                 integral ? mm_yes->get<MinMaxResult>() : nullptr, integral ? mm_no->get<MinMaxResult>() : nullptr);
    // the count and both outputs' min/max in one read-back: the FL encoders
    // of the outputs then need no pass and no read-back of their own
    uint64_t m = 0;
    auto hy = std::make_shared<MinMaxResult>(), hn = std::make_shared<MinMaxResult>();
    if (integral) {
        const ReadBack rb[3] = {{&m, cnt.get(), 8},
                                {hy.get(), mm_yes->get(), sizeof(MinMaxResult)},
                                {hn.get(), mm_no->get(), sizeof(MinMaxResult)}};
        d2h_many(c, rb, 3);
    } else {
        d2h(c, &m, cnt.get(), 8);
    }
    yes.set_size(m * es);
    no.set_size((n - m) * es);
    kernel_bytes(c, "k_split", (double)(2 * n * es + 1 + (n + 7) / 8));
    Part py = own(std::move(yes)), pn = own(std::move(no));
    py.minmax = mm_yes;
    pn.minmax = mm_no;
    if (integral) {
        py.minmax_host = hy;
        pn.minmax_host = hn;
    }
    return {own(std::move(st)), py, pn};
}
Part patch_decode(Ctx* c, int dtype, const std::vector<DView>& in) {
    const size_t es = dtype_size(dtype);
    if (in[1].n == 0 && in[2].n == 0) return empty();
    const size_t n = stencil_rows(c, in[0]);
    const DView y = in[1], z = in[2];  // the merge reads rows at any byte address
    if (y.n / es + z.n / es < n) fail(DDJ_ECORRUPT, "patch outputs shorter than the stencil");
    DevBuf out(c, n * es);
    launch_merge(c, dtype, in[0].p + 1, n, y.p, y.n / es, z.p, z.n / es, false, 0, out.get());
    return own(std::move(out));
}

// Patch whose two outputs go to FL leaves (patch -> (afl -> none), (afl ->
// none), integer columns): the counting pass brings back the yes count and
// both sides' min/max, so the FL widths and every size are known on the
// host; the writing pass runs when the blob is assembled and writes the
// stencil and both FL streams straight into it (launch_split_fl).  The
// blob is byte-identical to patch_encode followed by the two AFL encoders.
struct PatchFlJob {
    SplitPlan plan;
    Part x;
    int by = 1, bn = 1;
    uint64_t sbytes = 0, ybytes = 0, nbytes = 0;
    uint8_t* dst[3] = {nullptr, nullptr, nullptr};  // stencil, yes stream, no stream
    int expected = 0, seen = 0;
    std::shared_ptr<DevBuf> count;
    void place(Ctx* c, int which, uint8_t* p) {
        dst[which] = p;
        if (++seen < expected) return;
        launch_split_fl(c, plan, dst[0], FlSide{dst[1], ybytes, by, nullptr}, FlSide{dst[2], nbytes, bn, nullptr},
                        count->get<uint64_t>());
    }
};

// ------------------------------------------------------------------ const
// ConstEncoding (compression/const/const_encoding.cu:43-131)
std::vector<Part> const_encode(Ctx* c, int dtype, DView in) {
    const size_t es = dtype_size(dtype), n = rows(dtype, in.n);
    if (n == 0) return {empty(), empty()};
    Part x = aligned_view(c, in, es);
    uint64_t key[4];
    topk_most_frequent(c, dtype, x.p, n, 1, key, nullptr);
    DevBuf meta(c, es + 1 + (n + 7) / 8), data(c, n * es), cnt(c, 8);
    h2d(c, meta.get(), key, es);
    PredArgs p{};
    p.kind = PRED_NOT_EQUAL;
    p.a_bits = key[0];
    launch_split(c, dtype, x.p, n, p, meta.bytes_ptr() + es, data.get(), nullptr, nullptr, cnt.get<uint64_t>());
    uint64_t m = 0;
    d2h(c, &m, cnt.get(), 8);
    data.set_size(m * es);
    return {own(std::move(meta)), own(std::move(data))};
}
Part const_decode(Ctx* c, int dtype, const std::vector<DView>& in) {
    const size_t es = dtype_size(dtype);
    if (in[0].n == 0) return empty();
    if (in[0].n < es + 1) fail(DDJ_ECORRUPT, "const metadata too short");
    uint64_t cv = 0;
    read_small(c, &cv, in[0].p, es);
    const size_t n = stencil_rows(c, DView{in[0].p + es, in[0].n - es});
    const DView y = in[1];  // the merge reads rows at any byte address
    DevBuf out(c, n * es);
    launch_merge(c, dtype, in[0].p + es + 1, n, y.p, y.n / es, nullptr, 0, true, cv, out.get());
    return own(std::move(out));
}

// ------------------------------------------------------------------- dict
// DictEncoding (compression/dict/dict_encoding.cu:66-130) with K = 4
// (dict_encoding.hpp:30) and UniqueEncoding::CompressUnique
// (unique_encoding.cu:88-131): out1 = [size_t K][size_t m][K x T][codes].
constexpr int kDictK = 4;
std::vector<Part> dict_encode(Ctx* c, int dtype, DView in, bool defer = false) {
    const size_t es = dtype_size(dtype), n = rows(dtype, in.n);
    if (n == 0) return {empty(), empty(), empty()};
    Part x = aligned_view(c, in, es);
    uint64_t keys[4];
    topk_most_frequent(c, dtype, x.p, n, kDictK, keys, nullptr);
    PredArgs p{};
    p.kind = PRED_IN_SET;
    p.set_n = kDictK;
    for (int i = 0; i < kDictK; i++) p.set_bits[i] = keys[i];
    // freq = [size_t K][size_t m][K keys][ceil(m / 16) words of 2-bit codes]
    // (ALT_BITLEN(K - 1) = 2); the split packs the codes in place, into words
    // sized for every row being frequent (it writes every word of the m codes)
    const size_t head = 16 + kDictK * es;
    if (defer) {
        // counting pass now, the writing pass at blob assembly: the other
        // rows then go straight to their blob range
        auto job = std::make_shared<SplitJob>();
        job->plan = launch_split_count(c, dtype, x.p, n, p);
        uint64_t m = 0;
        d2h(c, &m, job->plan.ts->prefix + job->plan.ts->tiles, 8);
        DevBuf st(c, 1 + (n + 7) / 8), f(c, head + 4 * ((m + 15) / 16));
        kernel_bytes(c, "k_hist_count", (double)(n * es));
        kernel_bytes(c, "k_split", (double)(n * es + 1 + (n + 7) / 8 + (m + 15) / 16 * 4 + (n - m) * es));
        Part freq = empty();
        if (m > 0) {
            std::vector<uint8_t> h(head);
            size_t sizes[2] = {(size_t)kDictK, (size_t)m};
            memcpy(h.data(), sizes, 16);
            for (int i = 0; i < kDictK; i++) memcpy(h.data() + 16 + i * es, &keys[i], es);
            h2d(c, f.get(), h.data(), head);
            job->codes = f.bytes_ptr() + head;
            freq = own(std::move(f));
        }
        Part stp = own(std::move(st));
        job->x = x;
        job->stencil = stp;
        job->freq = freq;
        job->count = std::make_shared<DevBuf>(c, 8);
        Part others;  // with no rows (every row frequent) the job still runs: it writes the stencil and codes
        others.n = (size_t)(n - m) * es;
        others.split = job;
        return {stp, freq, others};
    }
    DevBuf st(c, 1 + (n + 7) / 8), f(c, head + 4 * ((n + 15) / 16)), others(c, n * es), cnt(c, 8);
    launch_split(c, dtype, x.p, n, p, st.bytes_ptr(), nullptr, others.get(), f.bytes_ptr() + head,
                 cnt.get<uint64_t>());
    uint64_t m = 0;
    d2h(c, &m, cnt.get(), 8);
    others.set_size((n - m) * es);
    kernel_bytes(c, "k_hist_count", (double)(n * es));
    kernel_bytes(c, "k_split", (double)(n * es + 1 + (n + 7) / 8 + (m + 15) / 16 * 4 + (n - m) * es));
    Part freq = empty();
    if (m > 0) {
        std::vector<uint8_t> h(head);
        size_t sizes[2] = {(size_t)kDictK, (size_t)m};
        memcpy(h.data(), sizes, 16);
        for (int i = 0; i < kDictK; i++) memcpy(h.data() + 16 + i * es, &keys[i], es);
        h2d(c, f.get(), h.data(), head);
        f.set_size(head + 4 * ((m + 15) / 16));
        freq = own(std::move(f));
    }
    return {own(std::move(st)), freq, own(std::move(others))};
}
// DecompressUnique (unique_encoding.cu:215-259) then Splitter::Merge
Part dict_decode(Ctx* c, int dtype, const std::vector<DView>& in) {
    const size_t es = dtype_size(dtype);
    if (in[1].n == 0 && in[2].n == 0) return empty();
    const size_t n = stencil_rows(c, in[0]);
    const DView z = in[2];  // the merge reads rows at any byte address
    DevBuf out(c, n * es);
    if (in[1].n > 0) {
        if (in[1].n < 16) fail(DDJ_ECORRUPT, "dict header too short");
        size_t sizes[2];
        read_small(c, sizes, in[1].p, 16);
        const size_t K = sizes[0], m = sizes[1];
        if (K == 0 || K > (1u << 30) || 16 + K * es > in[1].n) fail(DDJ_ECORRUPT, "dict key count out of range");
        int bits = 1;
        while ((1ull << bits) < K) bits++;  // ALT_BITLEN(K - 1) for K >= 2; 1 for K == 1
        if (K == 1) bits = 1;
        const int P = 32 / bits;
        const size_t off = 16 + K * es;
        if ((m + P - 1) / P * 4 > in[1].n - off) fail(DDJ_ECORRUPT, "dict codes shorter than their count");
        if (m + z.n / es < n) fail(DDJ_ECORRUPT, "dict outputs shorter than the stencil");
        // _decompressUniqueKernel (unique_encoding.cu:183-204) folded into the merge
        launch_merge_codes(c, dtype, in[0].p + 1, n, in[1].p + off, m, bits, in[1].p + 16, (uint32_t)K, z.p,
                           z.n / es, out.get());
    } else {
        if (z.n / es < n) fail(DDJ_ECORRUPT, "dict outputs shorter than the stencil");
        launch_merge(c, dtype, in[0].p + 1, n, nullptr, 0, z.p, z.n / es, false, 0, out.get());
    }
    return own(std::move(out));
}

// ------------------------------------------------------------- floatToInt
// FloatEncoding (compression/float/float_encoding.cu:17-88)
std::vector<Part> float_encode(Ctx* c, int dtype, DView in) {
    const size_t es = dtype_size(dtype), n = rows(dtype, in.n);
    if (n == 0) return {empty(), empty()};
    Part x = aligned_view(c, in, es);
    int precision = 0;
    MinMaxResult mmr;
    if (dtype == DDJ_D_FLOAT || dtype == DDJ_D_DOUBLE) {
        DevBuf s(c, sizeof(FloatScanResult));
        launch_float_scan(c, dtype, x.p, n, s.get<FloatScanResult>());
        FloatScanResult h;
        d2h(c, &h, s.get(), sizeof h);
        precision = (int)h.precision;
        mmr = h.mm;
    } else {
        mmr = minmax(c, dtype, x.p, n).host;  // float(integer) always has precision 0
    }
    bool transform = false;
    if (precision < 10) {
        static const double pw[11] = {1e0, 1e1, 1e2, 1e3, 1e4, 1e5, 1e6, 1e7, 1e8, 1e9, 1e10};
        const int scale = (int)pw[precision];
        dispatch_native(dtype, [&](auto tag) {
            using T = decltype(tag);
            T mn, mx;
            decode_minmax(dtype, mmr, &mn, &mx);
            bool ok;
            if constexpr (std::is_floating_point<T>::value) {
                volatile T lo = mn * (T)scale, hi = mx * (T)scale;
                ok = lo > (T)std::numeric_limits<int32_t>::min() && hi < (T)std::numeric_limits<int32_t>::max();
            } else {
                long long lo = (long long)mn * scale, hi = (long long)mx * scale;
                if (sizeof(T) == 4) {  // int * int wraps in the reference
                    lo = (int32_t)(uint32_t)(uint64_t)lo;
                    hi = (int32_t)(uint32_t)(uint64_t)hi;
                }
                ok = lo > std::numeric_limits<int32_t>::min() && hi < std::numeric_limits<int32_t>::max();
            }
            transform = ok;
        });
        if (!transform) precision = 10;
    }
    Part data;
    if (transform) {
        DevBuf y(c, 4 * n);
        launch_float_to_int(c, dtype, x.p, n, precision, y.get<int32_t>());
        data = own(std::move(y));
    } else {
        DevBuf y(c, n * es);
        d2d(c, y.get(), x.p, n * es);
        data = own(std::move(y));
    }
    int32_t p32 = precision;
    return {host_part(&p32, 4), data};
}
Part float_decode(Ctx* c, int dtype, const std::vector<DView>& in) {
    const size_t es = dtype_size(dtype);
    if (in[1].n == 0) return empty();
    if (in[0].n < 4) fail(DDJ_ECORRUPT, "floatToInt metadata shorter than 4 bytes");
    const int32_t p = read_dev<int32_t>(c, in[0].p);
    if (p < 10) {
        if (p < 0) fail(DDJ_ECORRUPT, "negative precision");
        const size_t m = in[1].n / 4;
        Part y = aligned_view(c, in[1], 4);
        DevBuf out(c, m * es);
        launch_int_to_float(c, dtype, y.p, m, p, out.get());
        return own(std::move(out));
    }
    DevBuf out(c, in[1].n);
    d2d(c, out.get(), in[1].p, in[1].n);
    return own(std::move(out));
}

// -------------------------------------------------------------------- gfc
// GfcEncoding (compression/gfc/gfc_encoding.cpp:16-166): float/double; the
// integer types go through AFL plus an empty third output.
std::vector<Part> gfc_encode_any(Ctx* c, int dtype, DView in, bool defer) {
    if (dtype != DDJ_D_FLOAT && dtype != DDJ_D_DOUBLE) {
        auto r = afl_encode_int(c, dtype, in);
        r.push_back(empty());
        return r;
    }
    const size_t es = dtype_size(dtype), n = rows(dtype, in.n);
    if (n == 0) return {empty(), empty(), empty()};
    if (n > (size_t)std::numeric_limits<int32_t>::max() / 9) fail(DDJ_EUNSUPPORTED, "GFC offsets are int32");
    Part x = aligned_view(c, in, es);
    DevBuf meta, offs;
    if (!defer) {
        DevBuf pay;
        gfc_encode(c, dtype == DDJ_D_DOUBLE, x.p, n, meta, offs, pay);
        return {own(std::move(meta)), own(std::move(offs)), own(std::move(pay))};
    }
    // the payload's size is known after the size pass; its chunks are
    // written straight into the blob
    GfcPlan plan = gfc_plan(c, dtype == DDJ_D_DOUBLE, x.p, n, meta, offs);
    Part data;
    data.n = plan.payload_bytes;
    auto job = std::make_shared<FlJob>();
    job->dtype = dtype;
    job->n = n;
    job->x = x;
    job->write = [plan](Ctx* c, uint8_t* out, size_t bytes) {
        if (bytes != plan.payload_bytes) fail(DDJ_EHIP, "GFC payload size mismatch");
        gfc_write(c, plan, out);
    };
    data.fl = std::move(job);
    return {own(std::move(meta)), own(std::move(offs)), data};
}
Part gfc_decode_any(Ctx* c, int dtype, const std::vector<DView>& in) {
    if (dtype != DDJ_D_FLOAT && dtype != DDJ_D_DOUBLE) return afl_decode_int(c, dtype, in);
    if (in[2].n == 0) return empty();
    DevBuf out;
    gfc_decode(c, dtype == DDJ_D_DOUBLE, in[0].p, in[0].n, in[1].p, in[1].n, in[2].p, in[2].n, out);
    return own(std::move(out));
}

}  // namespace

PatchAflParts patch_afl_encode(Ctx* c, int dtype, DView in, bool outside) {
    const size_t es = dtype_size(dtype), n = rows(dtype, in.n);
    if (n == 0 || dtype == DDJ_D_FLOAT || dtype == DDJ_D_DOUBLE) fail(DDJ_EINVAL, "fused patch: integer rows only");
    Part x = aligned_view(c, in, es);
    const PredArgs p = patch_pred(c, dtype, x.p, n, outside);
    DevBuf mmb(c, 2 * sizeof(MinMaxResult));
    MinMaxResult* mmy = mmb.get<MinMaxResult>();
    auto job = std::make_shared<PatchFlJob>();
    job->plan = launch_split_count(c, dtype, x.p, n, p, mmy, mmy + 1);
    job->x = x;
    job->count = std::make_shared<DevBuf>(c, 8);
    uint64_t m = 0;
    MinMaxResult h[2];
    const ReadBack rb[2] = {{&m, job->plan.ts->prefix + job->plan.ts->tiles, 8}, {h, mmb.get(), sizeof h}};
    d2h_many(c, rb, 2);
    const uint64_t sides[2] = {m, n - m};
    PatchAflParts out;
    uint64_t bytes[2];
    int bits[2];
    for (int k = 0; k < 2; k++) {
        bits[k] = sides[k] ? bits_of(dtype, h[k]) : 1;
        bytes[k] = sides[k] ? fl_bytes(es, bits[k], sides[k]) : 0;
        out.side_in[k].n = sides[k] * es;  // never materialised: only its size is read
        if (sides[k]) {
            const uint8_t meta[4] = {(uint8_t)bits[k], (uint8_t)((bytes[k] * 8 - sides[k] * (uint64_t)bits[k]) & 0xff),
                                     0, 0};
            out.side_meta[k] = host_part(meta, 4);
        }
    }
    job->by = bits[0];
    job->bn = bits[1];
    job->sbytes = 1 + (n + 7) / 8;
    job->ybytes = bytes[0];
    job->nbytes = bytes[1];
    job->expected = 1 + (bytes[0] ? 1 : 0) + (bytes[1] ? 1 : 0);
    auto deferred = [&](int which, uint64_t nb) {
        Part part;
        part.n = (size_t)nb;
        if (!nb) return part;
        auto fj = std::make_shared<FlJob>();
        fj->dtype = dtype;
        fj->write = [job, which](Ctx* cc, uint8_t* dst, size_t) { job->place(cc, which, dst); };
        part.fl = std::move(fj);
        return part;
    };
    out.stencil = deferred(0, job->sbytes);
    out.side_data[0] = deferred(1, bytes[0]);
    out.side_data[1] = deferred(2, bytes[1]);
    return out;
}

Part delta_afl_decode(Ctx* c, int dtype, DView dmeta, DView ameta, DView data) {
    const size_t es = dtype_size(dtype);
    if (dmeta.n < es) fail(DDJ_ECORRUPT, "delta metadata shorter than one value");
    if (ameta.n < 2) fail(DDJ_ECORRUPT, "AFL metadata shorter than 2 bytes");
    int8_t mb[2];
    read_small(c, mb, ameta.p, 2);
    const int b = mb[0], rest = mb[1];
    const int W = 8 * (int)es;
    if (b <= 0 || b > W) fail(DDJ_ECORRUPT, "AFL bit width out of range");
    const long long bits = (long long)data.n * 8 - rest;
    if (bits < 0) fail(DDJ_ECORRUPT, "AFL rest larger than the data");
    const size_t m = (size_t)(bits / b);
    DevBuf out(c, (m + 1) * es);
    launch_fl_unpack_delta(c, (int)es, dmeta.p, data.p, data.n, m, b, out.get());
    kernel_bytes(c, "k_fl_unpack_delta_scan", (double)((m * b + 7) / 8 + (m + 1) * es));
    return own(std::move(out));
}

namespace {
// ----------------------------------------------------------------- unique
// UniqueEncoding::Encode / Decode (compression/unique/unique_encoding.cu:133-181,
// :261-300): [size_t U][size_t n][U values] and ceil(n / (32 / b)) code words.
std::vector<Part> unique_encode(Ctx* c, int dtype, DView in) {
    const size_t es = dtype_size(dtype), n = rows(dtype, in.n);
    if (n == 0) return {empty(), empty()};
    Part x = aligned_view(c, in, es);
    DevBuf meta, data;
    unique_encode_dev(c, dtype, x.p, n, meta, data);
    return {own(std::move(meta)), own(std::move(data))};
}
Part unique_decode(Ctx* c, int dtype, const std::vector<DView>& in) {
    if (in[1].n == 0) return empty();
    if (in[0].n < 16) fail(DDJ_ECORRUPT, "unique metadata shorter than its two sizes");
    uint64_t head[2];
    read_small(c, head, in[0].p, sizeof head);
    const size_t es = dtype_size(dtype);
    const uint64_t U = head[0], n = head[1];
    if (U == 0 || U > (in[0].n - 16) / es) fail(DDJ_ECORRUPT, "unique metadata shorter than its values");
    if (n > ((uint64_t)1 << 40)) fail(DDJ_ECORRUPT, "unique row count out of range");
    Part words = aligned_view(c, in[1], 4);
    DevBuf out(c, n * es);
    unique_decode_dev(c, dtype, in[0].p + 16, U, n, words.p, in[1].n / 4, out.get());
    return own(std::move(out));
}
}  // namespace

namespace {
// Width and value count of an integer AFL stream from its metadata
// (DecodeAfl, afl_encoding.cu:74-87: n = (8 * bytes - rest) / b).
size_t afl_stream_rows(Ctx* c, int dtype, DView meta, size_t data_bytes, int* b_out) {
    if (data_bytes == 0) {
        *b_out = 0;
        return 0;
    }
    if (meta.n < 2) fail(DDJ_ECORRUPT, "AFL metadata shorter than 2 bytes");
    int8_t mb[2];
    read_small(c, mb, meta.p, 2);
    const int b = mb[0], rest = mb[1];
    if (b <= 0 || b > 8 * (int)dtype_size(dtype)) fail(DDJ_ECORRUPT, "AFL bit width out of range");
    const long long bits = (long long)data_bytes * 8 - rest;
    if (bits < 0) fail(DDJ_ECORRUPT, "AFL rest larger than the data");
    *b_out = b;
    return (size_t)(bits / b);
}
}  // namespace

// PatchEncoding::Decode over two AflEncoding::Decode (lower_patch_encoding.cu,
// afl_encoding.cu:74-113): the merge reads both FL streams directly.
Part patch_afl_decode(Ctx* c, int dtype, DView stencil, DView ymeta, DView ydata, DView nmeta, DView ndata) {
    int by = 0, bn = 0;
    const size_t ny = afl_stream_rows(c, dtype, ymeta, ydata.n, &by);
    const size_t nn = afl_stream_rows(c, dtype, nmeta, ndata.n, &bn);
    if (ny == 0 && nn == 0) return empty();
    const size_t n = stencil_rows(c, stencil);
    if (ny + nn < n) fail(DDJ_ECORRUPT, "patch outputs shorter than the stencil");
    DevBuf out(c, n * dtype_size(dtype));
    launch_merge_fl(c, dtype, stencil.p + 1, n, ydata.p, ydata.n, ny, by, ndata.p, ndata.n, nn, bn, out.get());
    return own(std::move(out));
}

Part afl_decode_rows(Ctx* c, int dtype, DView meta, DView data, size_t begin, size_t count) {
    int b = 0;
    const size_t n = afl_stream_rows(c, dtype, meta, data.n, &b);
    if (begin > n || count > n - begin) fail(DDJ_EINVAL, "rows beyond the end of the column");
    const size_t es = dtype_size(dtype);
    DevBuf out(c, count * es);
    if (count) {
        launch_fl_unpack(c, dtype, data.p, data.n, count, b, out.get(), begin);
        kernel_bytes(c, "k_fl_unpack", (double)((count * b + 7) / 8 + count * es));
    }
    return own(std::move(out));
}

Part delta_afl_decode_rows(Ctx* c, int dtype, DView dmeta, DView ameta, DView data, size_t begin, size_t count) {
    const size_t es = dtype_size(dtype);
    if (dmeta.n < es) fail(DDJ_ECORRUPT, "delta metadata shorter than one value");
    int b = 0;
    const size_t m = afl_stream_rows(c, dtype, ameta, data.n, &b);
    if (begin > m + 1 || count > m + 1 - begin) fail(DDJ_EINVAL, "rows beyond the end of the column");
    DevBuf out(c, count * es);
    if (count == 0) return own(std::move(out));
    if (m == 0) {  // the first row alone
        d2d(c, out.get(), dmeta.p, es);
        return own(std::move(out));
    }
    launch_fl_unpack_delta_rows(c, (int)es, dmeta.p, data.p, data.n, b, begin, count, out.get());
    kernel_bytes(c, "k_fl_unpack_delta_scan", (double)(((count - 1) * b + 7) / 8 + count * es));
    return own(std::move(out));
}

std::vector<Part> delta_afl_encode(Ctx* c, int dtype, DView in, bool defer) {
    const size_t es = dtype_size(dtype), n = rows(dtype, in.n);
    if (n < 2) fail(DDJ_EINVAL, "fused delta/FL needs two rows");
    Part x = aligned_view(c, in, es);
    const size_t m = n - 1;
    DevBuf mmb;
    MinMaxResult h;
    if (defer) {  // the width only, folded on the host
        h = minmax_to_host(c, dtype, x.p, n, true);
    } else {
        mmb = DevBuf(c, sizeof(MinMaxResult));
        launch_minmax_delta(c, (int)es, x.p, n, mmb.get<MinMaxResult>());
    }
    kernel_bytes(c, "k_minmax_delta", (double)(n * es));
    const uint64_t W = 8 * es;
    Part data;
    if (defer) {  // the width first; the pack runs when the blob is assembled
        auto job = std::make_shared<FlJob>();
        job->dtype = dtype;
        job->delta = true;
        job->bits = bits_of(dtype, h);
        job->n = m;
        job->x = x;
        data.n = (size_t)fl_bytes(es, job->bits, m);
        data.fl = std::move(job);
    } else {
        DevBuf buf(c, (size_t)((W * m + W - 1) / W) * es);  // worst case b = W
        launch_fl_pack_delta(c, (int)es, x.p, m, mmb.get<MinMaxResult>(), buf.get());
        d2h(c, &h, mmb.get(), sizeof h);
        buf.set_size((size_t)fl_bytes(es, bits_of(dtype, h), m));
        kernel_bytes(c, "k_fl_pack", (double)(n * es + buf.size()));
        data = own(std::move(buf));
    }
    const int b = bits_of(dtype, h);
    const uint64_t bytes = data.n;
    uint8_t meta[4] = {(uint8_t)b, (uint8_t)((bytes * 8 - (uint64_t)m * b) & 0xff), 0, 0};
    return {borrow(DView{x.p, es}, x.owner), host_part(meta, 4), data};
}

void run_split_job(Ctx* c, const SplitJob& j, void* no_rows) {
    launch_split_write(c, j.plan, const_cast<uint8_t*>(j.stencil.p), nullptr, static_cast<uint8_t*>(no_rows), j.codes,
                       j.count->get<uint64_t>());
}

void run_fl_job(Ctx* c, const FlJob& j, void* out, size_t bytes) {
    if (j.write) {
        if (bytes) j.write(c, static_cast<uint8_t*>(out), bytes);
        return;
    }
    const size_t es = dtype_size(j.dtype);
    if (j.n == 0 || bytes == 0) return;
    if (bytes != fl_bytes(es, j.bits, j.n)) fail(DDJ_EHIP, "FL job size mismatch");
    if (j.delta)
        launch_fl_pack_delta(c, (int)es, j.x.p, j.n, nullptr, out, j.bits, bytes);
    else
        launch_fl_pack(c, j.dtype, j.x.p, j.n, nullptr, j.bits, out, bytes);
    kernel_bytes(c, "k_fl_pack", (double)((j.n + (j.delta ? 1 : 0)) * es + fl_bytes(es, j.bits, j.n)));
}

bool enc_ok(int enc) {
    return (enc >= DDJ_ENC_DELTA && enc <= DDJ_ENC_GFC) || enc == DDJ_ENC_OUTSIDE_PATCH;
}

int n_results(int enc) {
    switch (enc) {
    case DDJ_ENC_NONE: return 0;
    case DDJ_ENC_DELTA:
    case DDJ_ENC_SCALE:
    case DDJ_ENC_AFL:
    case DDJ_ENC_UNIQUE:
    case DDJ_ENC_CONST:
    case DDJ_ENC_FLOAT_TO_INT: return 1;
    case DDJ_ENC_RLE:
    case DDJ_ENC_DICT:
    case DDJ_ENC_PATCH:
    case DDJ_ENC_OUTSIDE_PATCH:
    case DDJ_ENC_GFC: return 2;
    default: fail(DDJ_EINVAL, "unknown encoding " + std::to_string(enc));
    }
}

void return_types(int enc, int dtype, int out[2]) {
    out[0] = dtype;
    out[1] = dtype;
    if (enc == DDJ_ENC_RLE || enc == DDJ_ENC_FLOAT_TO_INT) out[0] = DDJ_D_INT;  // rle_encoding.hpp:35, float_encoding.hpp:31
}

std::vector<Part> encode(Ctx* c, int enc, int dtype, DView in, const MinMaxResult* minmax, bool defer_fl,
                         const MinMaxResult* minmax_host, bool defer_split) {
    if (!dtype_ok(dtype)) fail(DDJ_EINVAL, "unknown data type");
    if (in.n && !in.p) fail(DDJ_EINVAL, "encoder input has no device bytes");
    switch (enc) {
    case DDJ_ENC_NONE: return none_encode(c, in);
    case DDJ_ENC_DELTA: return delta_encode(c, int_view(dtype), in);
    case DDJ_ENC_SCALE: return scale_encode(c, dtype, in);
    case DDJ_ENC_AFL:
        if (dtype == DDJ_D_DOUBLE) fail(DDJ_EUNSUPPORTED, "AFL on double returns no outputs (afl_encoding.cu:268-271)");
        if (dtype == DDJ_D_FLOAT) return afl_encode_float(c, in);
        return afl_encode_int(c, dtype, in, minmax, defer_fl, minmax_host);
    case DDJ_ENC_RLE: return rle_encode(c, dtype, in);
    case DDJ_ENC_DICT: return dict_encode(c, dtype, in, defer_split);
    case DDJ_ENC_CONST: return const_encode(c, dtype, in);
    case DDJ_ENC_PATCH: return patch_encode(c, dtype, in, false);
    case DDJ_ENC_OUTSIDE_PATCH: return patch_encode(c, dtype, in, true);
    case DDJ_ENC_FLOAT_TO_INT: return float_encode(c, dtype, in);
    case DDJ_ENC_GFC: return gfc_encode_any(c, dtype, in, defer_fl);
    case DDJ_ENC_UNIQUE: return unique_encode(c, dtype, in);
    default: fail(DDJ_EINVAL, "unknown encoding " + std::to_string(enc));
    }
}

Part decode(Ctx* c, int enc, int dtype, const std::vector<DView>& in) {
    if (!dtype_ok(dtype)) fail(DDJ_EINVAL, "unknown data type");
    const size_t need = 1 + (size_t)n_results(enc);
    if (in.size() < std::max<size_t>(need, 2)) fail(DDJ_EINVAL, "too few inputs for decode");
    switch (enc) {
    case DDJ_ENC_NONE: {
        DevBuf out(c, in[1].n);
        d2d(c, out.get(), in[1].p, in[1].n);
        return own(std::move(out));
    }
    case DDJ_ENC_DELTA: return delta_decode(c, int_view(dtype), in);
    case DDJ_ENC_SCALE: return scale_decode(c, int_view(dtype), in);
    case DDJ_ENC_AFL:
        if (dtype == DDJ_D_DOUBLE) fail(DDJ_EUNSUPPORTED, "AFL on double");
        if (dtype == DDJ_D_FLOAT) return afl_decode_float(c, in);
        return afl_decode_int(c, dtype, in);
    case DDJ_ENC_RLE: return rle_decode(c, dtype, in);
    case DDJ_ENC_DICT: return dict_decode(c, dtype, in);
    case DDJ_ENC_CONST: return const_decode(c, dtype, in);
    case DDJ_ENC_PATCH:
    case DDJ_ENC_OUTSIDE_PATCH: return patch_decode(c, dtype, in);
    case DDJ_ENC_FLOAT_TO_INT: return float_decode(c, dtype, in);
    case DDJ_ENC_GFC: return gfc_decode_any(c, dtype, in);
    case DDJ_ENC_UNIQUE: return unique_decode(c, dtype, in);
    default: fail(DDJ_EINVAL, "unknown encoding " + std::to_string(enc));
    }
}

// GetMetadataSize / GetCompressedSize of each encoding (its .hpp / .cu)
void predict_sizes(Ctx* c, int enc, int dtype, DView in, size_t* meta, size_t* data) {
    const size_t es = dtype_size(dtype), n = rows(dtype, in.n);
    auto sum_outs = [&](const std::vector<Part>& o, size_t from) {
        size_t s = 0;
        for (size_t i = from; i < o.size(); i++) s += o[i].n;
        return s;
    };
    switch (enc) {
    case DDJ_ENC_NONE: *meta = sizeof(size_t); *data = in.n; return;
    case DDJ_ENC_DELTA: *meta = n ? es : 0; *data = n ? in.n - es : 0; return;
    case DDJ_ENC_SCALE: *meta = n ? es : 0; *data = n ? in.n : 0; return;
    case DDJ_ENC_FLOAT_TO_INT: *meta = n ? 4 : 0; *data = in.n; return;
    case DDJ_ENC_GFC: *meta = 12; *data = 0; return;
    case DDJ_ENC_PATCH:
    case DDJ_ENC_OUTSIDE_PATCH: *meta = n ? (n + 7) / 8 + 1 : 0; *data = in.n; return;
    case DDJ_ENC_RLE:
        *meta = n ? 4 : 0;
        if (!n) { *data = 0; return; }
        if (dtype != DDJ_D_INT && dtype != DDJ_D_FLOAT) fail(DDJ_EUNSUPPORTED, "RLE size predictor: int/float only");
        *data = sum_outs(encode(c, enc, dtype, in), 1);
        return;
    case DDJ_ENC_CONST:
        *meta = n ? (n + 7) / 8 + es + 1 : 0;
        if (!n) { *data = 0; return; }
        if (dtype != DDJ_D_INT && dtype != DDJ_D_FLOAT) fail(DDJ_EUNSUPPORTED, "const size predictor: int/float only");
        *data = sum_outs(encode(c, enc, dtype, in), 1);
        return;
    case DDJ_ENC_DICT: {
        *meta = (n + 7) / 8 + 1;
        if (!n) { *data = 0; return; }
        if (dtype != DDJ_D_INT && dtype != DDJ_D_FLOAT) fail(DDJ_EUNSUPPORTED, "dict size predictor: int/float only");
        auto o = encode(c, enc, dtype, in);
        size_t m = 0;
        if (o[1].n >= 16) d2h(c, &m, o[1].p + 8, 8);
        *data = 16 + 4 * es + (m + 15) / 16 * 4 + o[2].n;  // dict_encoding.cu:156-172
        return;
    }
    case DDJ_ENC_AFL: {
        if (!n) { *meta = 0; *data = 0; return; }
        if (dtype != DDJ_D_INT && dtype != DDJ_D_FLOAT) fail(DDJ_EUNSUPPORTED, "AFL size predictor: int/float only");
        auto o = encode(c, enc, dtype, in);
        *meta = dtype == DDJ_D_INT ? 4 : sizeof(size_t) + 1;
        *data = o[1].n;
        return;
    }
    case DDJ_ENC_UNIQUE: {  // unique_encoding.cu:307-328: FindUnique(data, type) exists for int and float
        if (!n) { *meta = 0; *data = 0; return; }
        if (dtype != DDJ_D_INT && dtype != DDJ_D_FLOAT) fail(DDJ_EUNSUPPORTED, "unique size predictor: int/float only");
        Part x = aligned_view(c, in, es);
        const uint64_t U = unique_count(c, dtype, x.p, n);
        uint64_t v = U - 1;
        int b = 1;
        while (v >>= 1) b++;
        const size_t P = 32 / (size_t)b;
        *meta = 16 + (size_t)U * es;
        *data = (n + P - 1) / P * 4;
        return;
    }
    default: fail(DDJ_EINVAL, "unknown encoding");
    }
}

ddj_stats_t stats(Ctx* c, int dtype, DView in) {
    ddj_stats_t s;
    memset(&s, 0, sizeof s);
    const size_t es = dtype_size(dtype), n = rows(dtype, in.n);
    s.size = n;
    if (n == 0) return s;
    Part x = aligned_view(c, in, es);
    DevBuf d(c, sizeof(DevStats));
    launch_stats(c, dtype, x.p, n, d.get<DevStats>());
    DevStats h;
    d2h(c, &h, d.get(), sizeof h);
    dispatch_native(dtype, [&](auto tag) {
        using T = decltype(tag);
        MinMaxResult r{h.min_key, h.max_key};
        T mn, mx;
        decode_minmax(dtype, r, &mn, &mx);
        s.min = (double)mn;
        s.max = (double)mx;
        if constexpr (std::is_floating_point<T>::value) {
            s.mean = (double)(T)(h.sum_f / (double)n);
        } else {
            using U = std::make_unsigned_t<T>;
            T sum = (T)(U)h.sum_bits;
            s.mean = (double)(T)(sum / (double)n);
        }
    });
    s.minBitCnt = 64;  // BITLEN(double, double) (core/cuda_macros.cuh:187-189)
    s.precision = (int32_t)h.precision;
    s.sorted = h.unsorted ? 0 : 1;
    if (n <= 2)
        s.rlMetric = 1.0f;
    else
        s.rlMetric = (float)((double)((n - 2) + h.rl_eq) / (double)(n - 2));
    s.runs = h.runs;
    for (int i = 0; i < 65; i++) s.bit_hist[i] = h.bit_hist[i];
    s.delta_bits_max = h.delta_bits_max;
    // 64-register HyperLogLog estimate with the small-range correction
    double sum = 0;
    int zeros = 0;
    for (int i = 0; i < 64; i++) {
        sum += std::ldexp(1.0, -(int)h.hll[i]);
        zeros += h.hll[i] == 0;
    }
    double est = 0.709 * 64.0 * 64.0 / sum;
    if (est <= 2.5 * 64 && zeros > 0) est = 64.0 * std::log(64.0 / zeros);
    s.distinct_est = (uint64_t)std::llround(std::min(est, (double)n));
    return s;
}

}  // namespace ddj
