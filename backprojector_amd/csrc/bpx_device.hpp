// Device-side building blocks shared by the gfx950 kernels: exact IEEE arithmetic of the
// reference path-length formula, monotone-threshold search, buffer-resource sample loads.
//
// The translation units that include this header are compiled with -ffp-contract=off: every
// product and sum of the reference formula rounds separately.  sqrtf and __fdiv_rn lower to
// correctly rounded instruction sequences (HIP's default
// -fhip-fp32-correctly-rounded-divide-sqrt); fused multiply-adds appear only where written
// as __builtin_fmaf, and only in approximations whose error is bounded on the host.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bpx {

// ------------------------------------------------------------------ exact arithmetic
__device__ __forceinline__ float exact_sqrt(float x) { return __builtin_sqrtf(x); }
__device__ __forceinline__ float exact_div(float a, float b) { return __fdiv_rn(a, b); }
// C99 round(): half away from zero (the reference's CUDA round(float) and C++ round()).
__device__ __forceinline__ float round_away(float x) { return __builtin_roundf(x); }

// Squared distance ((a*a + b*b) + c*c) with every operation rounded (the expression of the
// reference's distance(), backproject_cuda.cu:25-29, with norm3df replaced by it).
__device__ __forceinline__ float sq3(float a, float b, float c) {
    return (a * a + b * b) + c * c;
}

// The single-tap bin of the GPU path, backproject_cuda.cu:155 with footprint 0:
// round((total - t0) / deltaT), total = ((lw + d1) + d2) + cw (:45).
__device__ __forceinline__ float tap_round(float total, float t0, float dT) {
    return round_away(exact_div(total - t0, dT));
}

// ------------------------------------------------------------------ ordered float keys
// key(a) < key(b) iff a < b for non-NaN floats; -0 and +0 are adjacent keys.
__device__ __forceinline__ uint32_t fkey(float x) {
    uint32_t u = __float_as_uint(x);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float fkey_inv(uint32_t k) {
    uint32_t u = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
    return __uint_as_float(u);
}

// Smallest key k in [klo, khi] with pred(k) true, for a predicate that is monotone
// (false ... false true ... true) on that range and true at khi.  Gallops outward from a
// guess, then bisects; at most ~64 predicate evaluations, typically 3-6 for a good guess.
template <class Pred>
__device__ uint32_t min_true_key(const Pred& pred, uint32_t kguess, uint32_t klo, uint32_t khi) {
    kguess = kguess < klo ? klo : (kguess > khi ? khi : kguess);
    uint32_t lo_false, hi_true;
    if (pred(kguess)) {
        hi_true = kguess;
        uint32_t step = 1;
        for (;;) {
            if (hi_true == klo) return klo;
            const uint32_t c = (hi_true - klo > step) ? hi_true - step : klo;
            if (pred(c)) {
                hi_true = c;
                step <<= 1;
            } else {
                lo_false = c;
                break;
            }
        }
    } else {
        lo_false = kguess;
        uint32_t step = 1;
        for (;;) {
            const uint32_t c = (khi - lo_false > step) ? lo_false + step : khi;
            if (c == khi || pred(c)) {
                hi_true = c;
                break;
            }
            lo_false = c;
            step <<= 1;
        }
    }
    while (hi_true - lo_false > 1u) {
        const uint32_t mid = lo_false + (hi_true - lo_false) / 2u;
        if (pred(mid)) hi_true = mid;
        else lo_false = mid;
    }
    return hi_true;
}

// The same search for a guess usually within 8 keys of the answer, in straight-line code: the
// predicate at kguess - 8 (false) and kguess + 8 (true) brackets the answer, and four fixed
// bisection steps find it (6 evaluations, no loop, no divergent branch).  Lanes whose answer
// lies outside the bracket, or whose bracket leaves [klo, khi], take min_true_key.
template <class Pred>
__device__ __forceinline__ uint32_t min_true_key_near(const Pred& pred, uint32_t kguess, uint32_t klo, uint32_t khi) {
    kguess = kguess < klo ? klo : (kguess > khi ? khi : kguess);
    const bool inside = kguess - klo >= 8u && khi - kguess >= 8u;
    uint32_t lo = inside ? kguess - 8u : klo, hi = inside ? kguess + 8u : khi;
    const bool bracketed = inside && !pred(lo) && pred(hi);
#pragma unroll
    for (int step = 0; step < 4; ++step) {
        const uint32_t mid = lo + ((hi - lo) >> 1);
        const bool t = pred(mid);
        hi = t ? mid : hi;
        lo = t ? lo : mid;
    }
    if (!bracketed) hi = min_true_key(pred, kguess, klo, khi);
    return hi;
}

// ------------------------------------------------------------------ buffer resources
// Raw buffer descriptor (gfx9 dword3 = 0x00020000: 32-bit data format, no swizzle).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void* base, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ float buf_f32(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 0));
}

// Byte offsets of row elements from a biased quotient.  With E-byte elements and
// kBias = 2^23 / E, y = q + kBias lies in [kBias, 2 kBias) for 0 <= q < kBias, where a unit in
// the last place is 1/E: the mantissa bits of y are E (y - kBias), and masking their low
// log2(E) bits leaves E floor(q), the offset of element floor(q) -- no float-to-int
// conversion and no multiply.  Forming y from a quotient rounds to 1/E (twice: the bias add
// and the fma that adds the distance term), which the callers' error budgets absorb.
template <uint32_t E>
struct ElemOffset {
    static_assert(E == 8 || E == 16, "8- or 16-byte elements");
    static constexpr float kBias = E == 8 ? 1048576.0f : 524288.0f;
    static constexpr uint32_t kMask = 0x7fffffu & ~(E - 1u);
    // offset of element floor(y - kBias), for kBias <= y < 2 kBias
    __device__ static __forceinline__ uint32_t of(float y) { return __float_as_uint(y) & kMask; }
    // offset of element clamp(floor(y - kBias), 0, top - kBias), top = kBias + an integer < kBias
    __device__ static __forceinline__ uint32_t clamped(float y, float top) {
        return of(__builtin_amdgcn_fmed3f(y, kBias, top));
    }
};

// Transient sample formats.  `idx` is an element index inside a row that starts `soff`
// bytes into the buffer.  Real formats return float, complex formats float2.
struct SmpF32 {
    static constexpr int kBytes = 4;
    static constexpr bool kComplex = false;
    using acc_t = float;
    __device__ static __forceinline__ float load(__amdgpu_buffer_rsrc_t r, uint32_t idx, uint32_t soff) {
        return buf_f32(r, idx * 4u, soff);
    }
};
struct SmpF16 {
    static constexpr int kBytes = 2;
    static constexpr bool kComplex = false;
    using acc_t = float;
    __device__ static __forceinline__ float load(__amdgpu_buffer_rsrc_t r, uint32_t idx, uint32_t soff) {
        const unsigned short b = __builtin_amdgcn_raw_buffer_load_b16(r, (int)(idx * 2u), (int)soff, 0);
        return (float)__builtin_bit_cast(_Float16, b);
    }
};
struct SmpBF16 {
    static constexpr int kBytes = 2;
    static constexpr bool kComplex = false;
    using acc_t = float;
    __device__ static __forceinline__ float load(__amdgpu_buffer_rsrc_t r, uint32_t idx, uint32_t soff) {
        const unsigned short b = __builtin_amdgcn_raw_buffer_load_b16(r, (int)(idx * 2u), (int)soff, 0);
        return __uint_as_float((uint32_t)b << 16);
    }
};
struct SmpC64 {
    static constexpr int kBytes = 8;
    static constexpr bool kComplex = true;
    using acc_t = float2;
    __device__ static __forceinline__ float2 load(__amdgpu_buffer_rsrc_t r, uint32_t idx, uint32_t soff) {
        const auto v = __builtin_amdgcn_raw_buffer_load_b64(r, (int)(idx * 8u), (int)soff, 0);
        return make_float2(__uint_as_float(v[0]), __uint_as_float(v[1]));
    }
};
struct SmpCF16 {
    static constexpr int kBytes = 4;
    static constexpr bool kComplex = true;
    using acc_t = float2;
    __device__ static __forceinline__ float2 load(__amdgpu_buffer_rsrc_t r, uint32_t idx, uint32_t soff) {
        const uint32_t b = __builtin_amdgcn_raw_buffer_load_b32(r, (int)(idx * 4u), (int)soff, 0);
        return make_float2((float)__builtin_bit_cast(_Float16, (unsigned short)(b & 0xffffu)),
                           (float)__builtin_bit_cast(_Float16, (unsigned short)(b >> 16)));
    }
};
struct SmpCBF16 {
    static constexpr int kBytes = 4;
    static constexpr bool kComplex = true;
    using acc_t = float2;
    __device__ static __forceinline__ float2 load(__amdgpu_buffer_rsrc_t r, uint32_t idx, uint32_t soff) {
        const uint32_t b = __builtin_amdgcn_raw_buffer_load_b32(r, (int)(idx * 4u), (int)soff, 0);
        return make_float2(__uint_as_float(b << 16), __uint_as_float(b & 0xffff0000u));
    }
};

// Fused layouts: sample and threshold interleaved so that ONE load at element j returns
// the threshold S[j] and both candidate samples Hp[j] (bin j-1) and Hp[j+1] (bin j).
//   F32      element 8 B  (Hp[j], S[j])                 load 12 B at 8j: Hp[j], S[j], Hp[j+1]
//   F16/BF16 element 8 B  (S[j], Hp[j] | Hp[j+1] << 16) load  8 B at 8j
//   C64      element 16 B (re[j], im[j], S[j], 0)        loads 16 B at 16j and 8 B at 16(j+1)
//   CF16/CBF16 element 16 B (S[j], c[j], c[j+1], 0)      load 12 B at 16j
// `pick(r, off, soff, s)` returns the sample of bin j-1+[s >= S[j]] for the element j at byte
// offset off (= kElemBytes j) of the row at soff.
// kLdsBytes: bytes of an element as the LDS-staged kernel copies it (the leading bytes of the
// row element; its lds_pick reads element j at e and the next element at e + kLdsBytes).
struct FusedF32 {
    static constexpr int kElemBytes = 8;
    static constexpr int kLdsBytes = 8;
    static constexpr int kSOffset = 1;  // float slot of S inside the element
    using acc_t = float;
    __device__ static __forceinline__ float pick(__amdgpu_buffer_rsrc_t r, uint32_t off, uint32_t soff, float s) {
        const auto v = __builtin_amdgcn_raw_buffer_load_b96(r, (int)off, (int)soff, 0);
        return s >= __uint_as_float(v[1]) ? __uint_as_float(v[2]) : __uint_as_float(v[0]);
    }
    // the same from a copy of the row in LDS; e = element j
    __device__ static __forceinline__ float lds_pick(const char* e, float s) {
        const uint2 a = *reinterpret_cast<const uint2*>(e);
        const float h1 = *reinterpret_cast<const float*>(e + 8);
        return s >= __uint_as_float(a.y) ? h1 : __uint_as_float(a.x);
    }
};
template <bool BF>
struct FusedH16 {
    static constexpr int kElemBytes = 8;
    static constexpr int kLdsBytes = 8;
    static constexpr int kSOffset = 0;
    using acc_t = float;
    __device__ static __forceinline__ float cvt(uint32_t b) {
        return BF ? __uint_as_float(b << 16) : (float)__builtin_bit_cast(_Float16, (unsigned short)b);
    }
    __device__ static __forceinline__ float pick(__amdgpu_buffer_rsrc_t r, uint32_t off, uint32_t soff, float s) {
        const auto v = __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, (int)soff, 0);
        const uint32_t h = s >= __uint_as_float(v[0]) ? (v[1] >> 16) : (v[1] & 0xffffu);
        return cvt(h);
    }
    __device__ static __forceinline__ float lds_pick(const char* e, float s) {
        const uint2 v = *reinterpret_cast<const uint2*>(e);
        return cvt(s >= __uint_as_float(v.x) ? (v.y >> 16) : (v.y & 0xffffu));
    }
    // acc += the sample an element (S[j], w = Hp[j] | Hp[j+1] << 16) holds for s, in three VALU
    // instructions: the compare, a select that also moves the chosen half into place (SDWA
    // word selects), and one add.  fp16: the chosen half lands in the low word and
    // v_fma_mix_f32 adds it as h * 1 + acc (h converts to fp32 exactly, one rounding: the
    // convert-then-add sum).  bf16: it lands in the high word over a zero low word, which is
    // its fp32 value.  The select reads VCC two wait states after the compare writes it, as
    // gfx950 requires of a VALU-written lane mask.  (ALU-only asm: the compiler neither forms
    // these selects here nor keeps a written fma(h, 1, acc), which it folds back into
    // convert + add.)
    __device__ static __forceinline__ void acc_elem(float& acc, uint32_t S, uint32_t w, float s) {
        uint32_t h;
        if constexpr (BF) {
            asm("v_cmp_ge_f32_e32 vcc, %1, %2\n\t"
                "s_nop 1\n\t"
                "v_cndmask_b32_sdwa %0, %3, %3, vcc dst_sel:WORD_1 dst_unused:UNUSED_PAD "
                "src0_sel:WORD_0 src1_sel:WORD_1"
                : "=v"(h) : "v"(s), "v"(S), "v"(w) : "vcc");
            acc = acc + __uint_as_float(h);
        } else {
            asm("v_cmp_ge_f32_e32 vcc, %1, %2\n\t"
                "s_nop 1\n\t"
                "v_cndmask_b32_sdwa %0, %3, %3, vcc dst_sel:DWORD dst_unused:UNUSED_PAD "
                "src0_sel:DWORD src1_sel:WORD_1"
                : "=v"(h) : "v"(s), "v"(S), "v"(w) : "vcc");
            asm("v_fma_mix_f32 %0, %1, 1.0, %0 op_sel_hi:[1,0,0]" : "+v"(acc) : "v"(h));
        }
    }
    // acc += lds_pick(e, s)
    __device__ static __forceinline__ void lds_acc(float& acc, const char* e, float s) {
        const uint2 v = *reinterpret_cast<const uint2*>(e);
        acc_elem(acc, v.x, v.y, s);
    }
    // acc += pick(r, off, soff, s)
    __device__ static __forceinline__ void pick_acc(float& acc, __amdgpu_buffer_rsrc_t r, uint32_t off,
                                                    uint32_t soff, float s) {
        const auto v = __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, (int)soff, 0);
        acc_elem(acc, v[0], v[1], s);
    }
};
struct FusedC64 {
    static constexpr int kElemBytes = 16;
    static constexpr int kLdsBytes = 16;
    static constexpr int kSOffset = 2;
    using acc_t = float2;
    __device__ static __forceinline__ float2 pick(__amdgpu_buffer_rsrc_t r, uint32_t off, uint32_t soff, float s) {
        const auto a = __builtin_amdgcn_raw_buffer_load_b128(r, (int)off, (int)soff, 0);
        const auto b = __builtin_amdgcn_raw_buffer_load_b64(r, (int)(off + 16u), (int)soff, 0);
        const bool up = s >= __uint_as_float(a[2]);
        return make_float2(__uint_as_float(up ? b[0] : a[0]), __uint_as_float(up ? b[1] : a[1]));
    }
    __device__ static __forceinline__ float2 lds_pick(const char* e, float s) {
        const uint4 a = *reinterpret_cast<const uint4*>(e);
        const uint2 b = *reinterpret_cast<const uint2*>(e + 16);
        const bool up = s >= __uint_as_float(a.z);
        return make_float2(__uint_as_float(up ? b.x : a.x), __uint_as_float(up ? b.y : a.y));
    }
};
// One component (COMP 0 = real, 1 = imaginary) of complex64 fused rows, as the LDS-staged
// kernel stages it: (c[j], S[j]) 8-B elements, picked like fp32 rows.  Two passes (one per
// component) give the complex volume; each component's sum is the complex kernel's.
template <int COMP>
struct FusedC64Part {
    static constexpr int kElemBytes = 16;
    static constexpr int kLdsBytes = 8;
    static constexpr int kComponent = COMP;
    using acc_t = float;
    __device__ static __forceinline__ uint2 stage_load(__amdgpu_buffer_rsrc_t r, int off, int soff) {
        const auto t = __builtin_amdgcn_raw_buffer_load_b96(r, off, soff, 0);
        return make_uint2(COMP ? t[1] : t[0], t[2]);
    }
    __device__ static __forceinline__ float lds_pick(const char* e, float s) {
        const uint2 a = *reinterpret_cast<const uint2*>(e);
        const float h1 = *reinterpret_cast<const float*>(e + 8);
        return s >= __uint_as_float(a.y) ? h1 : __uint_as_float(a.x);
    }
};

template <bool BF>
struct FusedC16 {
    static constexpr int kElemBytes = 16;
    static constexpr int kLdsBytes = 8;    // (S[j], c[j]); c[j + 1] is the next element's
    static constexpr int kSOffset = 0;
    using acc_t = float2;
    __device__ static __forceinline__ float cvt(uint32_t b) {
        return BF ? __uint_as_float(b << 16) : (float)__builtin_bit_cast(_Float16, (unsigned short)b);
    }
    __device__ static __forceinline__ float2 pick(__amdgpu_buffer_rsrc_t r, uint32_t off, uint32_t soff, float s) {
        const auto v = __builtin_amdgcn_raw_buffer_load_b96(r, (int)off, (int)soff, 0);
        const uint32_t c = s >= __uint_as_float(v[0]) ? v[2] : v[1];
        return make_float2(cvt(c & 0xffffu), cvt(c >> 16));
    }
    __device__ static __forceinline__ float2 lds_pick(const char* e, float s) {
        const uint2 v = *reinterpret_cast<const uint2*>(e);
        const uint32_t c1 = *reinterpret_cast<const uint32_t*>(e + 12);
        const uint32_t c = s >= __uint_as_float(v.x) ? c1 : v.y;
        return make_float2(cvt(c & 0xffffu), cvt(c >> 16));
    }
    // acc += lds_pick(e, s).  fp16: each half added by one mixed-precision FMA that reads it in
    // place (h x 1 + acc: exact conversion, one rounding -- the convert-then-add sum; op_sel
    // picks the high half for the imaginary part), instead of two conversions and two adds.
    __device__ static __forceinline__ void lds_acc(float2& acc, const char* e, float s) {
        if constexpr (BF) {
            const float2 x = lds_pick(e, s);
            acc.x = acc.x + x.x;
            acc.y = acc.y + x.y;
        } else {
            const uint2 v = *reinterpret_cast<const uint2*>(e);
            const uint32_t c1 = *reinterpret_cast<const uint32_t*>(e + 12);
            const uint32_t c = s >= __uint_as_float(v.x) ? c1 : v.y;
            asm("v_fma_mix_f32 %0, %1, 1.0, %0 op_sel_hi:[1,0,0]" : "+v"(acc.x) : "v"(c));
            asm("v_fma_mix_f32 %0, %1, 1.0, %0 op_sel:[1,0,0] op_sel_hi:[1,0,0]" : "+v"(acc.y) : "v"(c));
        }
    }
};

// ------------------------------------------------------------------ accumulators
__device__ __forceinline__ void acc_add(float& a, float x) { a = a + x; }
__device__ __forceinline__ void acc_add(float2& a, float2 x) { a.x = a.x + x.x; a.y = a.y + x.y; }
// acc += Fz::lds_pick(e, s), through Fz::lds_acc where the layout has one
template <class Fz>
__device__ __forceinline__ auto lds_accumulate(typename Fz::acc_t& a, const char* e, float s, int)
    -> decltype(Fz::lds_acc(a, e, s)) {
    Fz::lds_acc(a, e, s);
}
template <class Fz>
__device__ __forceinline__ void lds_accumulate(typename Fz::acc_t& a, const char* e, float s, long) {
    acc_add(a, Fz::lds_pick(e, s));
}
// acc += Fz::pick(r, off, soff, s), through Fz::pick_acc where the layout has one
template <class Fz>
__device__ __forceinline__ auto pick_accumulate(typename Fz::acc_t& a, __amdgpu_buffer_rsrc_t r, uint32_t off,
                                                uint32_t soff, float s, int) -> decltype(Fz::pick_acc(a, r, off, soff, s)) {
    Fz::pick_acc(a, r, off, soff, s);
}
template <class Fz>
__device__ __forceinline__ void pick_accumulate(typename Fz::acc_t& a, __amdgpu_buffer_rsrc_t r, uint32_t off,
                                                uint32_t soff, float s, long) {
    acc_add(a, Fz::pick(r, off, soff, s));
}
__device__ __forceinline__ void acc_zero(float& a) { a = 0.0f; }
__device__ __forceinline__ void acc_zero(float2& a) { a = make_float2(0.0f, 0.0f); }

// Kahan-compensated add of a block partial `x` into (sum, comp); comp holds the negated
// low-order part lost so far.
__device__ __forceinline__ void kahan_add(float& sum, float& comp, float x) {
    const float y = x - comp;
    const float t = sum + y;
    comp = (t - sum) - y;
    sum = t;
}
__device__ __forceinline__ void kahan_add(float2& sum, float2& comp, float2 x) {
    kahan_add(sum.x, comp.x, x.x);
    kahan_add(sum.y, comp.y, x.y);
}
__device__ __forceinline__ float kahan_result(float sum, float comp) { return sum - comp; }
__device__ __forceinline__ float2 kahan_result(float2 sum, float2 comp) {
    return make_float2(sum.x - comp.x, sum.y - comp.y);
}

// Work-group index remap so that consecutive tiles land on the same XCD (dispatch places
// work-group i on XCD i mod 8): neighbouring voxel tiles read the same transient-row
// segments and share that XCD's L2.
__device__ __forceinline__ uint32_t xcd_remap(uint32_t bid, uint32_t nblocks) {
    constexpr uint32_t kXcd = 8;
    const uint32_t per = nblocks / kXcd, rem = nblocks % kXcd;
    const uint32_t xcd = bid % kXcd, pos = bid / kXcd;
    // XCDs with index < rem receive one extra block.
    return (xcd < rem ? xcd * (per + 1) : rem * (per + 1) + (xcd - rem) * per) + pos;
}

}  // namespace bpx
