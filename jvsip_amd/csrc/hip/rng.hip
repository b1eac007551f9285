// rng.hip -- the VSIPL random number generators filling device views.
//
// Both generators are 32-bit linear congruential sequences, X <- a X + c
// (mod 2^32); the portable one subtracts a second sequence X1 <- a1 X1 + c1
// that skips one value whenever it meets its marker X2 (the host finds
// where, see vhip_rng_desc.kstar).  Element e of a fill takes draws
// D e + 1 ... D e + D of the launch (D = 1, 2, 12 or 6 for real uniform,
// complex uniform, real normal, complex normal), in the order the reference
// fills the view.  A lane reaches its first element by jump-ahead (the
// composed map of 2^i steps is X <- A_i X + C_i, A_{i+1} = A_i^2,
// C_{i+1} = C_i (A_i + 1)) and its later elements by one composed step of
// the grid stride.  The draws become values with the reference's rounding
// steps, so the output is the reference's bit for bit.
#include <hip/hip_runtime.h>
#include "vhip.h"

extern "C" int vhip_cu_count(int *cus);

namespace {

struct Lcg {
    unsigned a, c;
};

__device__ __forceinline__ unsigned step(Lcg g, unsigned x) { return g.a * x + g.c; }

// x after `k` steps (k < 2^32) through the 2^i-step table
__device__ __forceinline__ unsigned jump(const unsigned *ja, const unsigned *jc, unsigned long k, unsigned x) {
    for (int i = 0; i < 32 && k; ++i, k >>= 1)
        if (k & 1) x = ja[i] * x + jc[i];
    return x;
}

// a draw as a value in [0, 1): nonportable X / 2^32; portable t = X - X1
// either as (t >> 8 | 1) / 2^24 or as t / 2^32, whichever the filling
// function of the reference uses (vhip_rng_desc.p24); the conversion of a
// 32-bit value to float rounds to nearest, as the reference's does
template <typename T>
__device__ __forceinline__ T u32(unsigned x) { return (T)x / (T)4294967296.0; }
template <typename T>
__device__ __forceinline__ T u24(unsigned t) { return (T)((t >> 8) | 1u) / (T)16777216.0; }

template <typename T, int DIST, int CPLX>
__global__ void __launch_bounds__(256) rng_kernel(const vhip_rng_desc d) {
#pragma clang fp contract(off)
    constexpr int D = DIST == VHIP_RNG_UNIFORM ? (CPLX ? 2 : 1) : (CPLX ? 6 : 12);
    const long total = d.m * d.n;
    const long stride = (long)gridDim.x * blockDim.x;
    long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= total) return;
    const Lcg g0 = {d.a, d.c}, g1 = {d.a1, d.c1};
    const Lcg gs0 = {d.stride_a, d.stride_c}, gs1 = {d.stride_a1, d.stride_c1};
    unsigned x = jump(d.ja, d.jc, (unsigned long)(D * e), d.x0);
    unsigned x1 = 0;
    if (d.portable && !d.kstar) x1 = jump(d.ja1, d.jc1, (unsigned long)(D * e), d.x10);
    for (; e < total; e += stride) {
        const long base = D * e;
        if (d.portable && d.kstar) {   // the launch holds the skip of X1: place X1 from scratch
            x1 = base >= d.kstar ? jump(d.ja1, d.jc1, (unsigned long)(base - d.kstar), d.x2 + 1u)
                                 : jump(d.ja1, d.jc1, (unsigned long)base, d.x10);
        }
        T u[D];
        unsigned xs = x, x1s = x1;
#pragma unroll
        for (int q = 0; q < D; ++q) {
            xs = step(g0, xs);
            if (d.portable) {
                x1s = step(g1, x1s);
                const unsigned t = xs - x1s;
                if (d.kstar && base + q + 1 == d.kstar) x1s += 1u;
                u[q] = d.p24 ? u24<T>(t) : u32<T>(t);
            } else {
                u[q] = u32<T>(xs);
            }
        }
        T re, im = 0;
        if (DIST == VHIP_RNG_UNIFORM) {
            re = u[0];
            if (CPLX) im = u[1];
        } else if (CPLX) {
            T t1 = u[0];
            t1 += u[1];
            t1 += u[2];
            T t2 = u[3];
            t2 += u[4];
            t2 += u[5];
            im = t1 - t2;
            re = (T)3 - t2 - t1;
        } else {
            T s = 0;
#pragma unroll
            for (int q = 0; q < 12; ++q) s += u[q];
            re = d.portable ? (T)6 - s : s - (T)6;
        }
        long i, j;
        if (d.by_col) { j = e / d.m; i = e - j * d.m; }
        else { i = e / d.n; j = e - i * d.n; }
        const long o = i * d.r.cs + j * d.r.rs;
        if (!CPLX) {
            ((T *)d.r.re)[o] = re;
        } else if (d.r.split) {
            ((T *)d.r.re)[o] = re;
            ((T *)d.r.im)[o] = im;
        } else {
            ((T *)d.r.re)[2 * o] = re;
            ((T *)d.r.re)[2 * o + 1] = im;
        }
        x = step(gs0, x);
        if (d.portable && !d.kstar) x1 = step(gs1, x1);
    }
}

using rng_fn = void (*)(vhip_rng_desc);

template <typename T>
rng_fn pick(int dist, int cplx) {
    if (dist == VHIP_RNG_UNIFORM) return cplx ? rng_kernel<T, VHIP_RNG_UNIFORM, 1> : rng_kernel<T, VHIP_RNG_UNIFORM, 0>;
    return cplx ? rng_kernel<T, VHIP_RNG_NORMAL, 1> : rng_kernel<T, VHIP_RNG_NORMAL, 0>;
}

}  // namespace

// grid of a fill of `total` elements: lanes of one stride, the host composes
// the generators over it (vhip_rng_desc.stride_*)
extern "C" long vhip_rng_lanes(long total) {
    int cus = 256;
    vhip_cu_count(&cus);
    long g = (total + 255) / 256, cap = (long)(cus > 0 ? cus : 256) * 8;
    if (g > cap) g = cap;
    if (g < 1) g = 1;
    return g * 256;
}

extern "C" int vhip_rng_fill(const vhip_rng_desc *d, void *stream) {
    const long total = d->m * d->n;
    if (total <= 0) return 0;
    if (vhip_rng_lanes(total) != d->lanes) return -1;
    rng_fn f = d->prec == 4 ? pick<float>(d->dist, d->cplx) : d->prec == 8 ? pick<double>(d->dist, d->cplx) : nullptr;
    if (!f) return -1;
    hipLaunchKernelGGL(f, dim3((unsigned)(d->lanes / 256)), dim3(256), 0, (hipStream_t)stream, *d);
    return (int)hipGetLastError();
}
