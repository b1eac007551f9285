// corr_ops.hip -- the element passes of 1-D correlation around the library's
// FFTs (vsip_correlate1d_[fd], vsip_ccorrelate1d_[fd]).  The transform buffer
// holds two rows of N complex elements, N the smallest power of two above
// m + n - 1 (m the reference length, n the data length):
//   pack     row 0: N - n zeros, then x; row 1: h, then N - m zeros; real
//            views enter with zero imaginary parts
//            (c_VSIP_src/vsip_ccorrelate1d_f.c:135-153,
//             vsip_correlate1d_f.c:95-113)
//   product  after the forward transform of both rows:
//            row0[k] = conj(row0[k] conj(row1[k])) * (1/N)
//            (vsip_ccorrelate1d_f.c:157-159)
//   conv     convolution (vsip_convolve1d_[fd]) on a real row of even
//            length L around real-to-complex and complex-to-real
//            transforms: x then zeros; the half-spectrum product with the
//            kernel's; the convolution read at lags s + k D
//   window   after the forward transform of row 0: y[i] = w_i * row0[o + i],
//            o = N - (m + n - 1) plus 0 (FULL), (m - 1)/2 (SAME) or m - 1
//            (MIN), w_i the unbiasing weights of the support
//            (vsip_ccorrelate1d_f.c:24-120 and 164-193), the real part only
//            for real correlation
// Each pass is one lane per element over a grid-stride loop; the buffer is
// contiguous, the views take any stride and split or interleaved storage.
#include <hip/hip_runtime.h>
#include "vhip.h"

extern "C" int vhip_cu_count(int *cus);

namespace {

template <typename T>
struct Cx {
    T r, i;
};

// element j of a vector view: complex (split or interleaved) or real
template <typename T>
__device__ __forceinline__ Cx<T> vload(const vhip_mat &a, int cplx, long j) {
    const long o = j * a.rs;
    if (!cplx) return {((const T *)a.re)[o], (T)0};
    if (a.split) return {((const T *)a.re)[o], ((const T *)a.im)[o]};
    return {((const T *)a.re)[2 * o], ((const T *)a.re)[2 * o + 1]};
}

template <typename T>
__device__ __forceinline__ void vstore(const vhip_mat &a, int cplx, long j, Cx<T> v) {
    const long o = j * a.rs;
    if (!cplx) {
        ((T *)a.re)[o] = v.r;
    } else if (a.split) {
        ((T *)a.re)[o] = v.r;
        ((T *)a.im)[o] = v.i;
    } else {
        ((T *)a.re)[2 * o] = v.r;
        ((T *)a.re)[2 * o + 1] = v.i;
    }
}

template <typename T>
__global__ void __launch_bounds__(256) corr_pack_kernel(const vhip_mat x, int xc, long n, const vhip_mat h, int hc,
                                                        long m, Cx<T> *buf, long N) {
    const long total = 2 * N;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
        Cx<T> v = {(T)0, (T)0};
        if (t < N) {
            if (t >= N - n) v = vload<T>(x, xc, t - (N - n));
        } else if (t - N < m) {
            v = vload<T>(h, hc, t - N);
        }
        buf[t] = v;
    }
}

template <typename T>
__global__ void __launch_bounds__(256) corr_product_kernel(Cx<T> *buf, long N, T inv_n) {
#pragma clang fp contract(off)
    for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < N; k += (long)gridDim.x * blockDim.x) {
        const Cx<T> a = buf[k], b = buf[N + k];
        // a conj(b), conjugated, times 1/N
        const T re = a.r * b.r + a.i * b.i;
        const T im = a.i * b.r - a.r * b.i;
        buf[k] = {re * inv_n, -im * inv_n};
    }
}

// y[i] = w_i * buf[o + i]; the weights follow the reference's loops:
//   FULL unbiased (length m + n - 1): i < m: / (i + 1); then * (1/m) up to
//     the last min(m, n - 1) lags, which are / (m + n - 1 - i)
//   SAME unbiased (length n): i < m/2: / (i + m/2 + (m & 1)); i < n - m/2:
//     / m; else / (n + m/2 - i)
//   MIN unbiased: * (1/m)
template <typename T>
__global__ void __launch_bounds__(256) corr_window_kernel(const Cx<T> *buf, long o, const vhip_mat y, int yc, long len,
                                                          int support, int unbiased, long m, long n) {
#pragma clang fp contract(off)
    const long mn = m + n - 1;
    const long l3 = m < n - 1 ? m : n - 1;
    const long n0 = m / 2, n1 = n - n0;
    const T mceil = (T)(n0 + (m & 1)), nsame = (T)(n + n0), tm = (T)m, rm = (T)1 / (T)m;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < len; i += (long)gridDim.x * blockDim.x) {
        Cx<T> v = buf[o + i];
        if (unbiased) {
            if (support == VHIP_CORR_FULL) {
                if (i < m) { const T d = (T)(i + 1); v = {v.r / d, v.i / d}; }
                else if (i < mn - l3) v = {v.r * rm, v.i * rm};
                else { const T d = (T)(mn - i); v = {v.r / d, v.i / d}; }
            } else if (support == VHIP_CORR_SAME) {
                const T d = i < n0 ? (T)i + mceil : i < n1 ? tm : nsame - (T)i;
                v = {v.r / d, v.i / d};
            } else {
                v = {v.r * rm, v.i * rm};
            }
        }
        vstore<T>(y, yc, i, v);
    }
}

// real-input convolution: r = x then L - n zeros (real row), for the
// real-to-complex transform
template <typename T>
__global__ void __launch_bounds__(256) conv_pack_r_kernel(const vhip_mat x, long n, T *r, long L) {
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < L; t += (long)gridDim.x * blockDim.x)
        r[t] = t < n ? ((const T *)x.re)[t * x.rs] : (T)0;
}

// half spectra: s[k] = s[k] k[k] / L, k <= L/2 (no conjugate)
template <typename T>
__global__ void __launch_bounds__(256) conv_mul_half_kernel(Cx<T> *s, const Cx<T> *kspec, long nb, T inv_l) {
#pragma clang fp contract(off)
    for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < nb; k += (long)gridDim.x * blockDim.x) {
        const Cx<T> a = s[k], b = kspec[k];
        s[k] = {(a.r * b.r - a.i * b.i) * inv_l, (a.r * b.i + a.i * b.r) * inv_l};
    }
}

// y[k] = r[s + k D] from the real row
template <typename T>
__global__ void __launch_bounds__(256) conv_window_r_kernel(const T *r, long s, long D, const vhip_mat y, long len) {
    for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < len; k += (long)gridDim.x * blockDim.x)
        ((T *)y.re)[k * y.rs] = r[s + k * D];
}

long grid_for(long total) {
    int cus = 256;
    vhip_cu_count(&cus);
    long g = (total + 255) / 256, cap = (long)(cus > 0 ? cus : 256) * 16;
    if (g > cap) g = cap;
    return g < 1 ? 1 : g;
}

}  // namespace

extern "C" int vhip_corr_pack(int prec, const vhip_mat *x, int x_cplx, long n, const vhip_mat *h, int h_cplx, long m,
                              void *buf, long N, void *stream) {
    if (n < 1 || m < 1 || n > N || m > N) return -1;
    const long g = grid_for(2 * N);
    hipStream_t s = (hipStream_t)stream;
    if (prec == 4)
        hipLaunchKernelGGL(corr_pack_kernel<float>, dim3((unsigned)g), dim3(256), 0, s, *x, x_cplx, n, *h, h_cplx, m,
                           (Cx<float> *)buf, N);
    else if (prec == 8)
        hipLaunchKernelGGL(corr_pack_kernel<double>, dim3((unsigned)g), dim3(256), 0, s, *x, x_cplx, n, *h, h_cplx, m,
                           (Cx<double> *)buf, N);
    else
        return -1;
    return (int)hipGetLastError();
}

extern "C" int vhip_corr_product(int prec, void *buf, long N, void *stream) {
    if (N < 1) return -1;
    const long g = grid_for(N);
    hipStream_t s = (hipStream_t)stream;
    if (prec == 4)
        hipLaunchKernelGGL(corr_product_kernel<float>, dim3((unsigned)g), dim3(256), 0, s, (Cx<float> *)buf, N,
                           (float)1 / (float)N);
    else if (prec == 8)
        hipLaunchKernelGGL(corr_product_kernel<double>, dim3((unsigned)g), dim3(256), 0, s, (Cx<double> *)buf, N,
                           (double)1 / (double)N);
    else
        return -1;
    return (int)hipGetLastError();
}

extern "C" int vhip_corr_window(int prec, const void *buf, long N, long o, const vhip_mat *y, int y_cplx, long len,
                                int support, int unbiased, long m, long n, void *stream) {
    if (len < 1 || o < 0 || o + len > N || m < 1 || n < m) return -1;
    if (support != VHIP_CORR_FULL && support != VHIP_CORR_SAME && support != VHIP_CORR_MIN) return -1;
    const long g = grid_for(len);
    hipStream_t s = (hipStream_t)stream;
    if (prec == 4)
        hipLaunchKernelGGL(corr_window_kernel<float>, dim3((unsigned)g), dim3(256), 0, s, (const Cx<float> *)buf, o, *y,
                           y_cplx, len, support, unbiased, m, n);
    else if (prec == 8)
        hipLaunchKernelGGL(corr_window_kernel<double>, dim3((unsigned)g), dim3(256), 0, s, (const Cx<double> *)buf, o,
                           *y, y_cplx, len, support, unbiased, m, n);
    else
        return -1;
    return (int)hipGetLastError();
}

extern "C" int vhip_conv_pack_r(int prec, const vhip_mat *x, long n, void *r, long L, void *stream) {
    if (n < 1 || n > L) return -1;
    const long g = grid_for(L);
    hipStream_t s = (hipStream_t)stream;
    if (prec == 4)
        hipLaunchKernelGGL(conv_pack_r_kernel<float>, dim3((unsigned)g), dim3(256), 0, s, *x, n, (float *)r, L);
    else if (prec == 8)
        hipLaunchKernelGGL(conv_pack_r_kernel<double>, dim3((unsigned)g), dim3(256), 0, s, *x, n, (double *)r, L);
    else
        return -1;
    return (int)hipGetLastError();
}

extern "C" int vhip_conv_mul_half(int prec, void *spec, const void *kspec, long L, void *stream) {
    if (L < 2 || (L & 1)) return -1;
    const long nb = L / 2 + 1, g = grid_for(nb);
    hipStream_t s = (hipStream_t)stream;
    if (prec == 4)
        hipLaunchKernelGGL(conv_mul_half_kernel<float>, dim3((unsigned)g), dim3(256), 0, s, (Cx<float> *)spec,
                           (const Cx<float> *)kspec, nb, (float)1 / (float)L);
    else if (prec == 8)
        hipLaunchKernelGGL(conv_mul_half_kernel<double>, dim3((unsigned)g), dim3(256), 0, s, (Cx<double> *)spec,
                           (const Cx<double> *)kspec, nb, (double)1 / (double)L);
    else
        return -1;
    return (int)hipGetLastError();
}

extern "C" int vhip_conv_window_r(int prec, const void *r, long L, long s0, long D, const vhip_mat *y, long len,
                                  void *stream) {
    if (len < 1 || D < 1 || s0 < 0 || s0 + (len - 1) * D >= L) return -1;
    const long g = grid_for(len);
    hipStream_t s = (hipStream_t)stream;
    if (prec == 4)
        hipLaunchKernelGGL(conv_window_r_kernel<float>, dim3((unsigned)g), dim3(256), 0, s, (const float *)r, s0, D,
                           *y, len);
    else if (prec == 8)
        hipLaunchKernelGGL(conv_window_r_kernel<double>, dim3((unsigned)g), dim3(256), 0, s, (const double *)r, s0,
                           D, *y, len);
    else
        return -1;
    return (int)hipGetLastError();
}
