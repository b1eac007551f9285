// fftm_staged.hip -- the element passes of staged transforms.
//
// A staged transform runs its complex FFT on contiguous scratch rows with the
// ordinary kernels and moves data between the caller's view and the scratch
// in two element passes, PRE and POST, with an optional MUL between two
// forward FFTs.  That covers what the single- and two-pass kernels do not:
//   - lengths with a prime factor above 13 (Bluestein's chirp-z: the n-point
//     DFT as a cyclic convolution of length m >= 2n - 1, computed with two
//     forward m-point FFTs and a pointwise product with the chirp filter's
//     spectrum; the reference instead runs an O(L^2) DFT on such a
//     remainder, VI_fftm_building_blocks_f.h:95-155);
//   - real transforms whose half length needs the two-pass split: the
//     even/odd packing and the Euler combine of VI_rcfftmop_f_def.h:53-115 /
//     VI_crfftmop_f_def.h:74-124 run in PRE/POST around the complex FFT.
// Every pass is one read and one write of each element; lanes walk along
// the vector for unit element strides and across vectors for unit vector
// strides, so the view side is coalesced in both layouts.
#include "fftm_core.hpp"

namespace vhipk {
namespace {

template <typename T>
__device__ __forceinline__ cpx<T> conjx(cpx<T> a) { return {a.x, -a.y}; }

template <typename T>
__device__ __forceinline__ cpx<T> ld_c(const vhip_view &g, long voff, long k) {
    const long o = voff + k * g.es;
    if (g.split) return {((const T *)g.re)[o], ((const T *)g.im)[o]};
    return ((const cpx<T> *)g.re)[o];
}

template <typename T>
__device__ __forceinline__ void st_c(const vhip_view &g, long voff, long k, cpx<T> v) {
    const long o = voff + k * g.es;
    if (g.split) {
        ((T *)g.re)[o] = v.x;
        ((T *)g.im)[o] = v.y;
    } else {
        ((cpx<T> *)g.re)[o] = v;
    }
}

// Y_k of the complex FFT result in scratch row w
template <typename T>
__device__ __forceinline__ cpx<T> post_get(const vhip_xform_desc &d, const cpx<T> *w, long k) {
    const long idx = (d.reverse && k) ? d.m - k : k;
    cpx<T> y = w[idx];
    if (d.chirp) y = cmul(y, ((const cpx<T> *)d.chirp)[k]);
    return y;
}

template <typename T, int STEP, int KIND>
__global__ void __launch_bounds__(256) xform_kernel(const vhip_xform_desc d) {
    const long dom = (STEP != VHIP_XF_POST) ? d.m : (KIND == VHIP_R2C ? d.n + 1 : d.n);
    const long total = d.nvec * dom;
    const long step = (long)gridDim.x * blockDim.x;
    const cpx<T> *eu = (const cpx<T> *)d.euler;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += step) {
        long v, k;
        if (d.across && STEP != VHIP_XF_MUL) {
            k = t / d.nvec;
            v = t - k * d.nvec;
        } else {
            v = t / dom;
            k = t - v * dom;
        }
        cpx<T> *w = (cpx<T> *)d.work + v * d.m;
        if (STEP == VHIP_XF_MUL) {
            w[k] = cmul(w[k], ((const cpx<T> *)d.table)[k]);
            continue;
        }
        const long vo = vec_off(d.view, d.vbase + v);
        if (STEP == VHIP_XF_PRE) {
            cpx<T> y = {(T)0, (T)0};
            if (k < d.n) {
                if (KIND == VHIP_R2C) {
                    const T *re = (const T *)d.view.re;
                    y.x = re[vo + 2 * k * d.view.es];
                    y.y = re[vo + (2 * k + 1) * d.view.es];
                } else if (KIND == VHIP_C2R) {
                    const cpx<T> F = ld_c<T>(d.view, vo, k);
                    const cpx<T> Bc = conjx(ld_c<T>(d.view, vo, d.n - k));
                    const cpx<T> A = F + Bc;
                    const cpx<T> WD = cmul(eu[k], F - Bc);
                    y = {A.x - WD.y, -(A.y + WD.x)};          // conj(A + i WD)
                } else {
                    y = ld_c<T>(d.view, vo, k);
                    if (d.conj) y.y = -y.y;
                }
                if (d.chirp) y = cmul(y, ((const cpx<T> *)d.chirp)[k]);
            }
            w[k] = y;
        } else {
            const T s = (T)d.scale;
            if (KIND == VHIP_R2C) {
                const cpx<T> F = post_get<T>(d, w, k == d.n ? 0 : k);
                const cpx<T> Bc = conjx(post_get<T>(d, w, k == 0 ? 0 : d.n - k));
                const cpx<T> A = F + Bc;
                const cpx<T> WD = cmul(eu[k], F - Bc);
                const T h = (T)0.5 * s;
                st_c<T>(d.view, vo, k, {h * (A.x + WD.y), h * (A.y - WD.x)});
            } else if (KIND == VHIP_C2R) {
                const cpx<T> y = post_get<T>(d, w, k);
                T *re = (T *)d.view.re;
                re[vo + 2 * k * d.view.es] = s * y.x;
                re[vo + (2 * k + 1) * d.view.es] = -s * y.y;
            } else {
                cpx<T> y = post_get<T>(d, w, k);
                y.x *= s;
                y.y *= d.conj ? -s : s;
                st_c<T>(d.view, vo, k, y);
            }
        }
    }
}

typedef void (*xform_fn)(const vhip_xform_desc);

template <typename T>
xform_fn pick(int step, int kind) {
    if (step == VHIP_XF_MUL) return xform_kernel<T, VHIP_XF_MUL, VHIP_C2C>;   // any kind: rows of w only
    if (step == VHIP_XF_PRE) {
        if (kind == VHIP_C2C) return xform_kernel<T, VHIP_XF_PRE, VHIP_C2C>;
        if (kind == VHIP_R2C) return xform_kernel<T, VHIP_XF_PRE, VHIP_R2C>;
        if (kind == VHIP_C2R) return xform_kernel<T, VHIP_XF_PRE, VHIP_C2R>;
    }
    if (step == VHIP_XF_POST) {
        if (kind == VHIP_C2C) return xform_kernel<T, VHIP_XF_POST, VHIP_C2C>;
        if (kind == VHIP_R2C) return xform_kernel<T, VHIP_XF_POST, VHIP_R2C>;
        if (kind == VHIP_C2R) return xform_kernel<T, VHIP_XF_POST, VHIP_C2R>;
    }
    return nullptr;
}

}  // namespace
}  // namespace vhipk

extern "C" int vhip_cu_count(int *cus);

extern "C" int vhip_xform_launch(const vhip_xform_desc *d, void *stream) {
    if (!d || d->nvec < 0 || d->n < 1 || d->m < d->n || !d->work) return -1;
    if (d->nvec == 0) return 0;
    if (d->step == VHIP_XF_MUL && !d->table) return -1;
    const bool needs_euler = (d->kind == VHIP_C2R && d->step == VHIP_XF_PRE) ||
                             (d->kind == VHIP_R2C && d->step == VHIP_XF_POST);
    if (needs_euler && !d->euler) return -1;
    vhipk::xform_fn f = d->prec == 4 ? vhipk::pick<float>(d->step, d->kind)
                      : d->prec == 8 ? vhipk::pick<double>(d->step, d->kind) : nullptr;
    if (!f) return -2;
    const long dom = d->step != VHIP_XF_POST ? d->m : (d->kind == VHIP_R2C ? d->n + 1 : d->n);
    const long total = d->nvec * dom;
    int cus = 256;
    vhip_cu_count(&cus);
    long grid = (total + 255) / 256;
    const long cap = (long)(cus > 0 ? cus : 256) * 32;
    if (grid > cap) grid = cap;
    hipLaunchKernelGGL(f, dim3((unsigned)grid), dim3(256), 0, (hipStream_t)stream, *d);
    return (int)hipGetLastError();
}
