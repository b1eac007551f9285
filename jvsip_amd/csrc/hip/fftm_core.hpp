// fftm_core.hpp -- device code of the multiple-FFT kernels for gfx950 (MI355X).
//
// One workgroup owns `vpb` whole vectors of the matrix view.  A transform of
// length n = R_0 * R_1 * ... is computed as a self-sorting Stockham sequence:
// stage s reads its butterfly inputs with stride n/R_s, applies the stage
// twiddles W_{Ns*R}^{r*k} (Ns = R_0*...*R_{s-1}, k = j mod Ns), runs an R-point
// DFT in registers and writes the outputs to (j - k)*R + k + r*Ns.  The first
// stage reads straight from HBM and the last stage writes straight to HBM, both
// with consecutive lanes on consecutive addresses (along the vector for unit
// element stride, across vectors for unit vector stride), so each element is
// read once and written once; the stages in between exchange data through LDS
// with a padded layout (one spare slot per 256 bytes) that keeps the strided
// butterfly writes free of bank conflicts.
//
// The inverse transform is conj(F(conj(x))): the sign of the imaginary part is
// flipped on load and on store, which is exact, so only forward twiddles exist.
// The scale factor is applied in the final store.
//
// Real transforms (rcfftmop/crfftmop) run the n = N/2 complex FFT on the packed
// sequence x[2m] + i x[2m+1] and combine Y[k] with Y[n-k] through the table
// e^{-/+ i pi k / n} (behaviour of VI_rcfftmop_f_def.h:77-115 and
// VI_crfftmop_f_def.h:74-124).
#pragma once
#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#include <stdint.h>
#endif
#include "vhip.h"
#include "dft_consts.h"

namespace vhipk {

template <typename T>
struct __attribute__((aligned(2 * sizeof(T)))) cpx {
    T x, y;
};

template <typename T>
__device__ __forceinline__ cpx<T> operator+(cpx<T> a, cpx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__device__ __forceinline__ cpx<T> operator-(cpx<T> a, cpx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T>
__device__ __forceinline__ cpx<T> cmul(cpx<T> a, cpx<T> b) {
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// multiply by the constant (c, s)
template <typename T>
__device__ __forceinline__ cpx<T> cmulc(cpx<T> a, T c, T s) {
    return {a.x * c - a.y * s, a.x * s + a.y * c};
}
// multiply by -i
template <typename T>
__device__ __forceinline__ cpx<T> mnegi(cpx<T> a) { return {a.y, -a.x}; }

// --------------------------------------------------------------------------
// Small forward DFTs in registers, natural order in and out.
// --------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ void dft2(cpx<T> &a, cpx<T> &b) {
    cpx<T> t = a;
    a = t + b;
    b = t - b;
}

template <typename T>
__device__ __forceinline__ void dft4(cpx<T> &x0, cpx<T> &x1, cpx<T> &x2, cpx<T> &x3) {
    cpx<T> t0 = x0 + x2, t1 = x0 - x2, t2 = x1 + x3, t3 = mnegi(x1 - x3);
    x0 = t0 + t2;
    x2 = t0 - t2;
    x1 = t1 + t3;
    x3 = t1 - t3;
}

template <typename T, int R>
struct Dft;

// length-1 transforms (the real transforms of length 2 use one)
template <typename T>
struct Dft<T, 1> {
    __device__ __forceinline__ static void run(cpx<T> *) {}
};

template <typename T>
struct Dft<T, 2> {
    __device__ __forceinline__ static void run(cpx<T> *x) { dft2(x[0], x[1]); }
};

template <typename T>
struct Dft<T, 4> {
    __device__ __forceinline__ static void run(cpx<T> *x) { dft4(x[0], x[1], x[2], x[3]); }
};

// 8 = 2 x 4: two 4-point DFTs on even/odd samples, twiddle W8^k, 2-point combine.
template <typename T>
struct Dft<T, 8> {
    __device__ __forceinline__ static void run(cpx<T> *x) {
        const T r = (T)0.70710678118654752440084436210485;
        cpx<T> a0 = x[0], a1 = x[2], a2 = x[4], a3 = x[6];
        cpx<T> b0 = x[1], b1 = x[3], b2 = x[5], b3 = x[7];
        dft4(a0, a1, a2, a3);
        dft4(b0, b1, b2, b3);
        b1 = {r * (b1.x + b1.y), r * (b1.y - b1.x)};
        b2 = mnegi(b2);
        b3 = {r * (b3.y - b3.x), -r * (b3.x + b3.y)};
        x[0] = a0 + b0; x[4] = a0 - b0;
        x[1] = a1 + b1; x[5] = a1 - b1;
        x[2] = a2 + b2; x[6] = a2 - b2;
        x[3] = a3 + b3; x[7] = a3 - b3;
    }
};

// 16 = 4 x 4: n = n1 + 4 n2, k = k1 + 4 k2.
template <typename T>
struct Dft<T, 16> {
    __device__ __forceinline__ static void run(cpx<T> *x) {
        const T c1 = (T)0.92387953251128675612818318939679;  // cos(pi/8)
        const T s1 = (T)0.38268343236508977172845998403040;  // sin(pi/8)
        const T r = (T)0.70710678118654752440084436210485;
        dft4(x[0], x[4], x[8], x[12]);
        dft4(x[1], x[5], x[9], x[13]);
        dft4(x[2], x[6], x[10], x[14]);
        dft4(x[3], x[7], x[11], x[15]);
        // z[n1][k1] = x[n1 + 4 k1] after the first pass; multiply by W16^{n1 k1}
        x[5] = cmulc(x[5], c1, -s1);     // n1=1,k1=1: W^1
        x[9] = cmulc(x[9], r, -r);       // n1=1,k1=2: W^2
        x[13] = cmulc(x[13], s1, -c1);   // n1=1,k1=3: W^3
        x[6] = cmulc(x[6], r, -r);       // n1=2,k1=1: W^2
        x[10] = mnegi(x[10]);            // n1=2,k1=2: W^4
        x[14] = cmulc(x[14], -r, -r);    // n1=2,k1=3: W^6
        x[7] = cmulc(x[7], s1, -c1);     // n1=3,k1=1: W^3
        x[11] = cmulc(x[11], -r, -r);    // n1=3,k1=2: W^6
        x[15] = cmulc(x[15], -c1, s1);   // n1=3,k1=3: W^9
        // second pass over n1 for each k1; output y[k1 + 4 k2]
        dft4(x[0], x[1], x[2], x[3]);
        dft4(x[4], x[5], x[6], x[7]);
        dft4(x[8], x[9], x[10], x[11]);
        dft4(x[12], x[13], x[14], x[15]);
        // now x[4 k1 + k2] holds y[k1 + 4 k2]: transpose the 4x4 index grid
        cpx<T> t;
        t = x[1]; x[1] = x[4]; x[4] = t;
        t = x[2]; x[2] = x[8]; x[8] = t;
        t = x[3]; x[3] = x[12]; x[12] = t;
        t = x[6]; x[6] = x[9]; x[9] = t;
        t = x[7]; x[7] = x[13]; x[13] = t;
        t = x[11]; x[11] = x[14]; x[14] = t;
    }
};

// Odd prime P: pair x_m with x_{P-m}; y_k = x_0 + sum_m a_m cos(2 pi m k/P)
//   -/+ i sum_m b_m sin(2 pi m k/P), with a_m = x_m + x_{P-m}, b_m = x_m - x_{P-m}.
template <typename T, int P>
struct DftOdd {
    __device__ __forceinline__ static void run(cpx<T> *x) {
        constexpr int H = (P - 1) / 2;
        cpx<T> a[H], b[H];
        const cpx<T> x0 = x[0];
        cpx<T> y0 = x0;
#pragma unroll
        for (int m = 1; m <= H; ++m) {
            a[m - 1] = x[m] + x[P - m];
            b[m - 1] = x[m] - x[P - m];
            y0 = y0 + a[m - 1];
        }
#pragma unroll
        for (int k = 1; k <= H; ++k) {
            cpx<T> re = x0, im = {(T)0, (T)0};
#pragma unroll
            for (int m = 1; m <= H; ++m) {
                const T c = (T)PrimeTrig<P>::c[(m * k) % P];
                const T s = (T)PrimeTrig<P>::s[(m * k) % P];
                re.x += a[m - 1].x * c;
                re.y += a[m - 1].y * c;
                im.x += b[m - 1].x * s;
                im.y += b[m - 1].y * s;
            }
            // forward: y_k = re - i im, y_{P-k} = re + i im
            x[k] = {re.x + im.y, re.y - im.x};
            x[P - k] = {re.x - im.y, re.y + im.x};
        }
        x[0] = y0;
    }
};
// Composite R = R1 * R2 from two smaller DFTs: n = R2 n1 + n2, k = k1 + R1 k2,
// y[k1 + R1 k2] = sum_n2 W_R2^{n2 k2} [W_R^{n2 k1} sum_n1 x[R2 n1 + n2] W_R1^{n1 k1}].
template <typename T, int R1, int R2, template <typename, int> class D>
struct DftComp {
    __device__ __forceinline__ static void run(cpx<T> *x) {
        constexpr int R = R1 * R2;
        cpx<T> t[R1], u[R2], y[R];
#pragma unroll
        for (int n2 = 0; n2 < R2; ++n2) {
#pragma unroll
            for (int n1 = 0; n1 < R1; ++n1) t[n1] = x[R2 * n1 + n2];
            D<T, R1>::run(t);
#pragma unroll
            for (int k1 = 0; k1 < R1; ++k1) {
                const int m = (n2 * k1) % R;
                const T c = (T)CompTrig<R>::c[m];
                const T s = -(T)CompTrig<R>::s[m];   // forward: e^{-2 pi i m / R}
                x[R2 * k1 + n2] = (m == 0) ? t[k1] : cmulc(t[k1], c, s);
            }
        }
#pragma unroll
        for (int k1 = 0; k1 < R1; ++k1) {
#pragma unroll
            for (int n2 = 0; n2 < R2; ++n2) u[n2] = x[R2 * k1 + n2];
            D<T, R2>::run(u);
#pragma unroll
            for (int k2 = 0; k2 < R2; ++k2) y[k1 + R1 * k2] = u[k2];
        }
#pragma unroll
        for (int k = 0; k < R; ++k) x[k] = y[k];
    }
};

template <typename T> struct Dft<T, 3> : DftOdd<T, 3> {};
template <typename T> struct Dft<T, 5> : DftOdd<T, 5> {};
template <typename T> struct Dft<T, 7> : DftOdd<T, 7> {};
template <typename T> struct Dft<T, 11> : DftOdd<T, 11> {};
template <typename T> struct Dft<T, 13> : DftOdd<T, 13> {};
template <typename T> struct Dft<T, 6> : DftComp<T, 2, 3, Dft> {};
template <typename T> struct Dft<T, 9> : DftComp<T, 3, 3, Dft> {};
template <typename T> struct Dft<T, 10> : DftComp<T, 2, 5, Dft> {};
template <typename T> struct Dft<T, 12> : DftComp<T, 4, 3, Dft> {};
template <typename T> struct Dft<T, 14> : DftComp<T, 2, 7, Dft> {};
template <typename T> struct Dft<T, 15> : DftComp<T, 3, 5, Dft> {};

// --------------------------------------------------------------------------
// indexing helpers
// --------------------------------------------------------------------------

// q = a / b, r = a % b for 0 <= a < 2^22, 1 <= b: the float estimate a * rcp(b)
// (v_rcp_f32, within 1 ulp) is off from a / b by less than 2^22 * 1.5 * 2^-23 < 1,
// so one correction step gives the exact quotient.  Every caller divides an
// index below the element count of one workgroup.
__device__ __forceinline__ void divmod(int a, int b, int &q, int &r) {
    int qq = (int)((float)a * __builtin_amdgcn_rcpf((float)b));
    int rr = a - qq * b;
    if (rr < 0) { qq -= 1; rr += b; }
    else if (rr >= b) { qq += 1; rr -= b; }
    q = qq;
    r = rr;
}

// LDS padding: one spare complex slot per 256 bytes of data
template <typename T> struct Pad;
template <> struct Pad<float> { static constexpr int shift = 5; };
template <> struct Pad<double> { static constexpr int shift = 4; };
template <typename T>
__device__ __forceinline__ int pad(int i) { return i + (i >> Pad<T>::shift); }

__device__ __forceinline__ long vec_off(const vhip_view &g, long v) {
    if (g.nv0 <= 0 || v < g.nv0) return v * g.vs0;
    long q = v / g.nv0;
    return (v - q * g.nv0) * g.vs0 + q * g.vs1;
}

// --------------------------------------------------------------------------
// raw buffer access: a descriptor over [base, base + bytes) and 32-bit
// offsets, one VGPR per access plus a wave-uniform SGPR offset
// --------------------------------------------------------------------------
typedef __amdgpu_buffer_rsrc_t rsrc_t;
// word 3 of a raw buffer descriptor on gfx950 (32-bit data format, no swizzle)
constexpr int kBufWord3 = 0x00020000;

__device__ __forceinline__ rsrc_t make_rsrc(const void *p, long bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)bytes, kBufWord3);
}
typedef decltype(__builtin_amdgcn_raw_buffer_load_b64(rsrc_t(), 0, 0, 0)) raw64_t;
typedef decltype(__builtin_amdgcn_raw_buffer_load_b128(rsrc_t(), 0, 0, 0)) raw128_t;
// Cache policy (the buffer instructions' aux field) of the data's global
// loads and stores.  Non-temporal (2, stores 18) streams contiguous rows
// faster (the power-of-two c2c kernel uses it), but these kernels also
// serve column layouts, whose workgroups write 32-byte pieces of lines that
// neighbouring workgroups complete in L2, and there it measured up to 1.75x
// slower (COL 1920 x 1000); they keep the default policy.
constexpr int kLdData = 0;
constexpr int kStData = 0;
template <typename T> struct Buf;
template <> struct Buf<float> {
    template <int AUX = 0>
    static __device__ __forceinline__ cpx<float> ldc(rsrc_t r, int v, int s) {
        return __builtin_bit_cast(cpx<float>, __builtin_amdgcn_raw_buffer_load_b64(r, v, s, AUX));
    }
    template <int AUX = 0>
    static __device__ __forceinline__ float lds(rsrc_t r, int v, int s) {
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, v, s, AUX));
    }
    template <int AUX = 0>
    static __device__ __forceinline__ void stc(rsrc_t r, int v, int s, cpx<float> x) {
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(raw64_t, x), r, v, s, AUX);
    }
    template <int AUX = 0>
    static __device__ __forceinline__ void sts(rsrc_t r, int v, int s, float x) {
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, x), r, v, s, AUX);
    }
};
template <> struct Buf<double> {
    template <int AUX = 0>
    static __device__ __forceinline__ cpx<double> ldc(rsrc_t r, int v, int s) {
        return __builtin_bit_cast(cpx<double>, __builtin_amdgcn_raw_buffer_load_b128(r, v, s, AUX));
    }
    template <int AUX = 0>
    static __device__ __forceinline__ double lds(rsrc_t r, int v, int s) {
        return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, v, s, AUX));
    }
    template <int AUX = 0>
    static __device__ __forceinline__ void stc(rsrc_t r, int v, int s, cpx<double> x) {
        __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(raw128_t, x), r, v, s, AUX);
    }
    template <int AUX = 0>
    static __device__ __forceinline__ void sts(rsrc_t r, int v, int s, double x) {
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(raw64_t, x), r, v, s, AUX);
    }
};

struct Shared {
    long *voff_in;
    long *voff_out;
    int *vidx0;
};

// --------------------------------------------------------------------------
// global loads / stores
// --------------------------------------------------------------------------
// Global loads and stores of the first and last stage.  Every view of a
// launch spans less than 2 GiB from its base with non-negative strides (the
// host guarantees it: vhip_view.span), so all addressing is 32-bit buffer
// offsets: vb is the vector's element offset and element j + leg of the
// vector is accessed (leg: the butterfly leg's offset r * n / R, the same in
// every lane).  SOFF puts leg in the scalar offset (one VGPR offset per
// butterfly; suits kernels that read their descriptor from memory) or in the
// vector offset (no scalar registers per leg; suits kernels whose descriptor
// arrives as an argument and already occupies many scalar registers).
template <typename T, int KIND, bool SOFF = false>
__device__ __forceinline__ cpx<T> load_in(const vhip_fftm_desc &d, int vb, int j, int leg, T isgn) {
    constexpr int ES = (int)sizeof(T);
    const int es = (int)d.in.es;
    const int lv = SOFF ? 0 : leg;     // leg in the vector offset, or
    const int ls = SOFF ? leg : 0;     // in the scalar offset
    cpx<T> v;
    if (KIND == VHIP_R2C) {
        const rsrc_t r = make_rsrc(d.in.re, d.in.span);
        if (d.in.contig) {
            v = Buf<T>::template ldc<kLdData>(r, vb * ES + (j + lv) * 2 * ES, ls * 2 * ES);
        } else {
            const int o = (vb + 2 * (j + lv) * es) * ES, so = 2 * ls * es * ES;
            v.x = Buf<T>::template lds<kLdData>(r, o, so);
            v.y = Buf<T>::template lds<kLdData>(r, o + es * ES, so);
        }
        if (!SOFF && d.tap) {   // fused taper: standalone launches only
            const long s0 = 2L * (j + leg) * d.tap_s;
            v.x *= ((const T *)d.tap)[s0];
            v.y *= ((const T *)d.tap)[s0 + d.tap_s];
        }
        return v;
    }
    if (d.in.split) {
        const int o = (vb + (j + lv) * es) * ES, so = ls * es * ES;
        v.x = Buf<T>::template lds<kLdData>(make_rsrc(d.in.re, d.in.span), o, so);
        v.y = Buf<T>::template lds<kLdData>(make_rsrc(d.in.im, d.in.span), o, so);
    } else {
        v = Buf<T>::template ldc<kLdData>(make_rsrc(d.in.re, d.in.span), (vb + (j + lv) * es) * 2 * ES, ls * es * 2 * ES);
    }
    if (!SOFF && d.tap) {   // fused taper: standalone launches only
        const T t = ((const T *)d.tap)[(long)(j + leg) * d.tap_s];
        v.x *= t;
        v.y *= t;
    }
    v.y *= isgn;
    return v;
}

template <typename T>
__device__ __forceinline__ void store_c(const vhip_view &g, long voff, int n, cpx<T> v) {
    if (g.contig) {
        ((cpx<T> *)g.re)[voff + n] = v;
    } else {
        const long o = voff + (long)n * g.es;
        if (g.split) {
            ((T *)g.re)[o] = v.x;
            ((T *)g.im)[o] = v.y;
        } else {
            ((cpx<T> *)g.re)[o] = v;
        }
    }
}

template <typename T, int KIND, bool SOFF = false>
__device__ __forceinline__ void store_out(const vhip_fftm_desc &d, int vb, int vi0, int j, int leg,
                                          cpx<T> v, T osgn, T scale) {
    constexpr int ES = (int)sizeof(T);
    const int es = (int)d.out.es;
    const int lv = SOFF ? 0 : leg;
    const int ls = SOFF ? leg : 0;
    if (d.ptw) {
        const int t = vi0 * (j + leg);
        // a global-memory pointer: global loads count only in vmcnt, where a
        // generic (flat) load would also hold up the LDS waits (lgkmcnt)
        typedef const __attribute__((address_space(1))) T *gptr;
        const gptr w = (gptr)d.ptw;
        const int i0 = 2 * (t & ((1 << d.ptw_shift) - 1)), i1 = 2 * ((1 << d.ptw_shift) + (t >> d.ptw_shift));
        const cpx<T> w0 = {w[i0], w[i0 + 1]}, w1 = {w[i1], w[i1 + 1]};
        v = cmul(v, cmul(w0, w1));
    }
    v.x *= scale;
    v.y *= scale * osgn;
    if (KIND == VHIP_C2R) {
        const rsrc_t r = make_rsrc(d.out.re, d.out.span);
        if (d.out.contig) {
            Buf<T>::template stc<kStData>(r, vb * ES + (j + lv) * 2 * ES, ls * 2 * ES, v);
        } else {
            const int o = (vb + 2 * (j + lv) * es) * ES, so = 2 * ls * es * ES;
            Buf<T>::template sts<kStData>(r, o, so, v.x);
            Buf<T>::template sts<kStData>(r, o + es * ES, so, v.y);
        }
        return;
    }
    if (!SOFF && d.acc) {           // power accumulation in place of the store (standalone launches only)
        T *a = (T *)d.acc + (vb + (j + leg) * es);
        *a += (T)d.acc_gain * (v.x * v.x + v.y * v.y);
        return;
    }
    if (d.out.split) {
        const int o = (vb + (j + lv) * es) * ES, so = ls * es * ES;
        Buf<T>::template sts<kStData>(make_rsrc(d.out.re, d.out.span), o, so, v.x);
        Buf<T>::template sts<kStData>(make_rsrc(d.out.im, d.out.span), o, so, v.y);
    } else {
        Buf<T>::template stc<kStData>(make_rsrc(d.out.re, d.out.span), (vb + (j + lv) * es) * 2 * ES, ls * es * 2 * ES, v);
    }
}

// --------------------------------------------------------------------------
// one Stockham stage
// --------------------------------------------------------------------------

// The batched kernel's form of the stages (BATCH = true): the stages after
// the first take their twiddles from a two-level table of W_n = e^{-2 pi i / n}
// in LDS, W_n^e = lo[e mod 2^lg] hi[e >> lg], instead of a global load per
// twiddle.  The stage twiddle W_{Ns R}^{r k} is W_n^e with e = r k n / (Ns R),
// which is below n.
template <typename T>
struct BatchCtx {
    const cpx<T> *lo;
    const cpx<T> *hi;
    int lg;
};

// butterflies one thread holds in a stage of radix R (the host planner uses the
// same rule): MAXE / R, at least one; radices other than 1-5, 7, 8 and 16
// (composites and the primes 11, 13) hold one
constexpr int stage_maxb(int R, int maxe) {
    return (R == 6 || R >= 9) && R != 16 ? 1 : ((maxe / R) > 0 ? maxe / R : 1);
}
template <typename T, int R, int MAXE, int KIND, bool SRC_G, bool DST_G, bool SOFF, bool BATCH = false>
__device__ __forceinline__ void stage(const vhip_fftm_desc &d, const Shared &sh, cpx<T> *lds,
                                      int s, int Ns, int nv, T isgn, T osgn, T scale,
                                      const BatchCtx<T> *bc = nullptr) {
    // butterflies per thread: the register class allows MAXE elements, except
    // that composite and large prime radices stay at one butterfly per thread
    constexpr int MAXB = stage_maxb(R, MAXE);
    const int n = d.n;
    const int nbv = n / R;          // butterflies per vector
    const int nbt = nv * nbv;       // butterflies in this workgroup
    const int nthr = blockDim.x;
    const int S = d.lds_stride;
    const int across = SRC_G ? (d.map_in == VHIP_MAP_ACROSS)
                             : (DST_G ? (d.map_out == VHIP_MAP_ACROSS) : 0);
    // this stage's twiddles: (R - 1) * Ns entries from tw_off[s] (BATCH: the
    // powers of W_n, in steps of n / (Ns R))
    const rsrc_t twr = make_rsrc((const cpx<T> *)d.tw + d.tw_off[s], (long)n * (long)sizeof(cpx<T>));
    const int tstride = BATCH ? n / (Ns * R) : 0;

    cpx<T> x[MAXB][R];
    int bv[MAXB], bj[MAXB];
    // Keep the per-stage index arithmetic inside the stage: hoisting it out of
    // the stage loop would pin one address register per butterfly leg.
    int tid = threadIdx.x;
    asm volatile("" : "+v"(tid));

#pragma unroll
    for (int i = 0; i < MAXB; ++i) {
        const int b = tid + i * nthr;
        bv[i] = -1;
        bj[i] = 0;
        if (b < nbt) {
            int v, j;
            if (nv == 1) { v = 0; j = b; }
            else if (across) divmod(b, nv, j, v);
            else divmod(b, nbv, v, j);
            bv[i] = v;
            bj[i] = j;
            if (SRC_G) {
                const int vo = (int)sh.voff_in[v];
#pragma unroll
                for (int r = 0; r < R; ++r) x[i][r] = load_in<T, KIND, SOFF>(d, vo, j, r * nbv, isgn);
            } else {
                const cpx<T> *src = lds + v * S;
#pragma unroll
                for (int r = 0; r < R; ++r) x[i][r] = src[pad<T>(j + r * nbv)];
            }
        }
    }
    if constexpr (BATCH) {
        // The batched kernel's items: twiddles from its LDS table after the
        // barrier, each butterfly stored as soon as it is computed.  The
        // standalone form below (twiddles ahead of the barrier, all stores
        // after all butterflies) measured slower here: C4 0.918-0.920 vs
        // 0.905-0.909 ms per replay on one box.
        if (!SRC_G) __syncthreads();
#pragma unroll
        for (int i = 0; i < MAXB; ++i) {
            if (bv[i] >= 0) {
                const int v = bv[i], j = bj[i];
                int k = 0;
                if (Ns > 1) {
                    int q;
                    divmod(j, Ns, q, k);
                    const int step = k * tstride, mask = (1 << bc->lg) - 1;
#pragma unroll
                    for (int r = 1; r < R; ++r) {
                        const int e = r * step;
                        x[i][r] = cmul(x[i][r], cmul(bc->lo[e & mask], bc->hi[e >> bc->lg]));
                    }
                }
                Dft<T, R>::run(x[i]);
                if (DST_G) {
                    const int vo = (int)sh.voff_out[v];
                    const int vi0 = d.ptw ? sh.vidx0[v] : 0;
#pragma unroll
                    for (int r = 0; r < R; ++r)
                        store_out<T, KIND, SOFF>(d, vo, vi0, j, r * nbv, x[i][r], osgn, scale);
                } else {
                    cpx<T> *dst = lds + v * S;
                    const int base = (j - k) * R + k;
#pragma unroll
                    for (int r = 0; r < R; ++r) dst[pad<T>(base + r * Ns)] = x[i][r];
                }
            }
        }
        if (!DST_G) __syncthreads();
        return;
    }
    // this stage's twiddles: from L2, fetched ahead of the barrier when they
    // fit the register class (up to 16 elements per lane), so their latency
    // overlaps it; with 32 elements per lane they are fetched after it, as
    // needed
    constexpr bool PRE_TW = MAXE <= 16 && R > 1;
    cpx<T> w[PRE_TW ? MAXB : 1][PRE_TW ? R - 1 : 1];
    int kk[MAXB];
#pragma unroll
    for (int i = 0; i < MAXB; ++i) {
        kk[i] = 0;
        if (bv[i] >= 0 && Ns > 1) {
            int q;
            divmod(bj[i], Ns, q, kk[i]);
            if (PRE_TW) {
#pragma unroll
                for (int r = 1; r < R; ++r)
                    w[i][r - 1] = Buf<T>::ldc(twr, (kk[i] + (r - 1) * Ns) * (int)sizeof(cpx<T>), 0);
            }
        }
    }
    if (!SRC_G) __syncthreads();

    // every butterfly of the lane first, then every write: a store's data
    // registers are then never reused while stores are in flight (the
    // compiler would wait for the stores before reusing them)
#pragma unroll
    for (int i = 0; i < MAXB; ++i) {
        if (bv[i] >= 0) {
            const int k = kk[i];
            if (Ns > 1) {
                if (PRE_TW) {
#pragma unroll
                    for (int r = 1; r < R; ++r) x[i][r] = cmul(x[i][r], w[i][r - 1]);
                } else {
#pragma unroll
                    for (int r = 1; r < R; ++r)
                        x[i][r] = cmul(x[i][r], Buf<T>::ldc(twr, (k + (r - 1) * Ns) * (int)sizeof(cpx<T>), 0));
                }
            }
            Dft<T, R>::run(x[i]);
        }
    }
#pragma unroll
    for (int i = 0; i < MAXB; ++i) {
        if (bv[i] >= 0) {
            const int v = bv[i], j = bj[i], k = kk[i];
            if (DST_G) {
                // last stage: Ns * R == n, outputs land at j + r * nbv
                const int vo = (int)sh.voff_out[v];
                const int vi0 = d.ptw ? sh.vidx0[v] : 0;
#pragma unroll
                for (int r = 0; r < R; ++r)
                    store_out<T, KIND, SOFF>(d, vo, vi0, j, r * nbv, x[i][r], osgn, scale);
            } else {
                cpx<T> *dst = lds + v * S;
                const int base = (j - k) * R + k;
#pragma unroll
                for (int r = 0; r < R; ++r) dst[pad<T>(base + r * Ns)] = x[i][r];
            }
        }
    }
    if (!DST_G) __syncthreads();
}

template <typename T, int R, int MAXE, int KIND, bool SOFF, bool BATCH = false>
__device__ __forceinline__ void stage_dispatch_pos(const vhip_fftm_desc &d, const Shared &sh,
                                                   cpx<T> *lds, int s, int Ns, int nv, bool first,
                                                   bool last, T isgn, T osgn, T scale,
                                                   const BatchCtx<T> *bc = nullptr) {
    if (first && last) stage<T, R, MAXE, KIND, true, true, SOFF, BATCH>(d, sh, lds, s, Ns, nv, isgn, osgn, scale, bc);
    else if (first) stage<T, R, MAXE, KIND, true, false, SOFF, BATCH>(d, sh, lds, s, Ns, nv, isgn, osgn, scale, bc);
    else if (last) stage<T, R, MAXE, KIND, false, true, SOFF, BATCH>(d, sh, lds, s, Ns, nv, isgn, osgn, scale, bc);
    else stage<T, R, MAXE, KIND, false, false, SOFF, BATCH>(d, sh, lds, s, Ns, nv, isgn, osgn, scale, bc);
}

template <typename T, int MAXE, int KIND, bool SOFF, bool BATCH = false>
__device__ __forceinline__ void run_stage(const vhip_fftm_desc &d, const Shared &sh, cpx<T> *lds,
                                          int s, int Ns, int nv, bool first, bool last, T isgn,
                                          T osgn, T scale, const BatchCtx<T> *bc = nullptr) {
    switch (d.radix[s]) {
    case 16:
        // radix 16 only in the float kernels (the planner never emits it for double)
        if constexpr (sizeof(T) == 4)
            stage_dispatch_pos<T, 16, MAXE, KIND, SOFF, BATCH>(d, sh, lds, s, Ns, nv, first, last, isgn, osgn, scale, bc);
        break;
    case 8: stage_dispatch_pos<T, 8, MAXE, KIND, SOFF, BATCH>(d, sh, lds, s, Ns, nv, first, last, isgn, osgn, scale, bc); break;
    case 4: stage_dispatch_pos<T, 4, MAXE, KIND, SOFF, BATCH>(d, sh, lds, s, Ns, nv, first, last, isgn, osgn, scale, bc); break;
    case 2: stage_dispatch_pos<T, 2, MAXE, KIND, SOFF, BATCH>(d, sh, lds, s, Ns, nv, first, last, isgn, osgn, scale, bc); break;
    case 3: stage_dispatch_pos<T, 3, MAXE, KIND, SOFF, BATCH>(d, sh, lds, s, Ns, nv, first, last, isgn, osgn, scale, bc); break;
    case 5: stage_dispatch_pos<T, 5, MAXE, KIND, SOFF, BATCH>(d, sh, lds, s, Ns, nv, first, last, isgn, osgn, scale, bc); break;
    case 7: stage_dispatch_pos<T, 7, MAXE, KIND, SOFF, BATCH>(d, sh, lds, s, Ns, nv, first, last, isgn, osgn, scale, bc); break;
    case 11: stage_dispatch_pos<T, 11, MAXE, KIND, SOFF, BATCH>(d, sh, lds, s, Ns, nv, first, last, isgn, osgn, scale, bc); break;
    case 13: stage_dispatch_pos<T, 13, MAXE, KIND, SOFF, BATCH>(d, sh, lds, s, Ns, nv, first, last, isgn, osgn, scale, bc); break;
    case 1: stage_dispatch_pos<T, 1, MAXE, KIND, SOFF, BATCH>(d, sh, lds, s, Ns, nv, first, last, isgn, osgn, scale, bc); break;
    // composite radices (the planner emits them for float only)
    case 6: if constexpr (sizeof(T) == 4) stage_dispatch_pos<T, 6, MAXE, KIND, SOFF, BATCH>(d, sh, lds, s, Ns, nv, first, last, isgn, osgn, scale, bc); break;
    case 9: if constexpr (sizeof(T) == 4) stage_dispatch_pos<T, 9, MAXE, KIND, SOFF, BATCH>(d, sh, lds, s, Ns, nv, first, last, isgn, osgn, scale, bc); break;
    case 10: if constexpr (sizeof(T) == 4) stage_dispatch_pos<T, 10, MAXE, KIND, SOFF, BATCH>(d, sh, lds, s, Ns, nv, first, last, isgn, osgn, scale, bc); break;
    case 12: if constexpr (sizeof(T) == 4) stage_dispatch_pos<T, 12, MAXE, KIND, SOFF, BATCH>(d, sh, lds, s, Ns, nv, first, last, isgn, osgn, scale, bc); break;
    case 14: if constexpr (sizeof(T) == 4) stage_dispatch_pos<T, 14, MAXE, KIND, SOFF, BATCH>(d, sh, lds, s, Ns, nv, first, last, isgn, osgn, scale, bc); break;
    case 15: if constexpr (sizeof(T) == 4) stage_dispatch_pos<T, 15, MAXE, KIND, SOFF, BATCH>(d, sh, lds, s, Ns, nv, first, last, isgn, osgn, scale, bc); break;
    default: break;  // the planner only emits the radices above
    }
}

// rcfftmop epilogue: Z[k] = 1/2 [(Y_k + conj Y_{n-k}) - i W^k (Y_k - conj Y_{n-k})],
// W^k = e^{-i pi k / n}, k = 0..n, with Y_n == Y_0 (VI_rcfftmop_f_def.h:81-115).
template <typename T>
__device__ __forceinline__ void r2c_post(const vhip_fftm_desc &d, const Shared &sh,
                                         const cpx<T> *lds, int nv, T scale) {
    const int n = d.n;
    const int per = n + 1;
    const int total = nv * per;
    const int S = d.lds_stride;
    const cpx<T> *eu = (const cpx<T> *)d.euler;
    const T half = (T)0.5 * scale;
    for (int t = threadIdx.x; t < total; t += blockDim.x) {
        int v, k;
        if (nv == 1) { v = 0; k = t; }
        else if (d.map_out == VHIP_MAP_ACROSS) divmod(t, nv, k, v);
        else divmod(t, per, v, k);
        const cpx<T> *y = lds + v * S;
        const cpx<T> F = y[pad<T>(k == n ? 0 : k)];
        const cpx<T> B = y[pad<T>(k == 0 ? 0 : n - k)];
        const cpx<T> Bc = {B.x, -B.y};
        const cpx<T> A = F + Bc;
        const cpx<T> WD = cmul(eu[k], F - Bc);
        cpx<T> z = {half * (A.x + WD.y), half * (A.y - WD.x)};
        store_c(d.out, sh.voff_out[v], k, z);
    }
}

// crfftmop prologue: Y_k = (Z_k + conj Z_{n-k}) + i W^k (Z_k - conj Z_{n-k}),
// W^k = e^{+i pi k / n}, k = 0..n-1, no 1/2 factor (VI_crfftmop_f_def.h:74-106).
// The conjugate of Y is written so the forward stages compute the inverse.
template <typename T>
__device__ __forceinline__ void c2r_pre(const vhip_fftm_desc &d, const Shared &sh, cpx<T> *lds,
                                        int nv) {
    const int n = d.n;
    const int total = nv * n;
    const int S = d.lds_stride;
    const cpx<T> *eu = (const cpx<T> *)d.euler;
    for (int t = threadIdx.x; t < total; t += blockDim.x) {
        int v, k;
        if (nv == 1) { v = 0; k = t; }
        else if (d.map_in == VHIP_MAP_ACROSS) divmod(t, nv, k, v);
        else divmod(t, n, v, k);
        const long vo = sh.voff_in[v];
        const cpx<T> F = load_in<T, VHIP_C2C>(d, (int)vo, k, 0, (T)1);
        const cpx<T> B = load_in<T, VHIP_C2C>(d, (int)vo, n - k, 0, (T)1);
        const cpx<T> Bc = {B.x, -B.y};
        const cpx<T> A = F + Bc;
        const cpx<T> WD = cmul(eu[k], F - Bc);
        // Y = A + i WD ; store conj(Y)
        const cpx<T> Y = {A.x - WD.y, A.y + WD.x};
        lds[v * S + pad<T>(k)] = {Y.x, -Y.y};
    }
}

template <typename T, int MAXE, int KIND, bool SOFF = false>
__device__ __forceinline__ void fft_body(const vhip_fftm_desc &d, long vbase, int nv,
                                         char *smem) {
    cpx<T> *lds = (cpx<T> *)smem;
    Shared sh;
    sh.voff_in = (long *)(smem + (size_t)d.vpb * d.lds_stride * sizeof(cpx<T>));
    sh.voff_out = sh.voff_in + d.vpb;
    sh.vidx0 = (int *)(sh.voff_out + d.vpb);
    for (int t = threadIdx.x; t < nv; t += blockDim.x) {
        const long v = vbase + t;
        sh.voff_in[t] = vec_off(d.in, v);
        sh.voff_out[t] = vec_off(d.out, v);
        sh.vidx0[t] = (int)(d.ptw_n > 0 ? v % d.ptw_n : v);
    }
    __syncthreads();

    const T scale = (T)d.scale;
    // C2C inverse: conjugate on load and on store (a four-step split does the
    // two in different passes).  C2R: the prologue writes the conjugate and the
    // store conjugates back.  R2C: forward only.
    const T isgn = (KIND == VHIP_C2C && d.conj_in) ? (T)-1 : (T)1;
    const T osgn = (KIND == VHIP_C2R || (KIND == VHIP_C2C && d.conj_out)) ? (T)-1 : (T)1;

    if (KIND == VHIP_C2R) {
        c2r_pre<T>(d, sh, lds, nv);
        __syncthreads();
    }
    int Ns = 1;
    for (int s = 0; s < d.nst; ++s) {
        const bool first = (s == 0) && KIND != VHIP_C2R;
        const bool last = (s == d.nst - 1) && KIND != VHIP_R2C;
        run_stage<T, MAXE, KIND, SOFF>(d, sh, lds, s, Ns, nv, first, last, isgn, osgn, scale);
        Ns *= d.radix[s];
    }
    if (KIND == VHIP_R2C) r2c_post<T>(d, sh, lds, nv, scale);
}

// The 32-element register class runs at most 512 lanes per workgroup, which
// leaves it 256 registers per lane; the smaller classes allow 1024 lanes.
// Workgroups are dealt to the 8 XCDs of the device in turn (workgroup b runs
// on XCD b % 8).  Renumbering them so that consecutive vector groups land on
// one XCD lets the groups that read neighbouring columns of the same cache
// lines (unit vector stride: fewer than 16 vectors per workgroup cover less
// than a 128-byte line) share that XCD's L2 instead of fetching the line once
// per XCD.
__device__ __forceinline__ unsigned xcd_group(unsigned b, unsigned g) {
    const unsigned x = b & 7u, i = b >> 3, per = g >> 3, rem = g & 7u;
    return x * per + (x < rem ? x : rem) + i;
}

template <typename T, int MAXE, int KIND>
__global__ void __launch_bounds__(MAXE > 16 ? 512 : 1024) fftm_kernel(const vhip_fftm_desc d) {
    extern __shared__ __attribute__((aligned(16))) char smem[];
    const bool across = d.map_in == VHIP_MAP_ACROSS || d.map_out == VHIP_MAP_ACROSS;
    const long vbase = (long)(across ? xcd_group(blockIdx.x, gridDim.x) : blockIdx.x) * d.vpb;
    long left = d.nvec - vbase;
    const int nv = left < d.vpb ? (int)left : d.vpb;
    fft_body<T, MAXE, KIND>(d, vbase, nv, smem);
}

typedef void (*kernel_fn)(const vhip_fftm_desc);
kernel_fn fftm_kernel_f16(int kind);
kernel_fn fftm_kernel_f32(int kind);
kernel_fn fftm_kernel_d8(int kind);
int pow2_try_launch(const vhip_fftm_desc *d, void *stream, int cus);
int pow2_set_lds_limit(int bytes);
int pow2_real_try_launch(const vhip_fftm_desc *d, void *stream, int cus);
int pow2_real_set_lds_limit(int bytes);
int batch_set_lds_limit(int bytes);

}  // namespace vhipk
