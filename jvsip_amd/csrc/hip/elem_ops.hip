// elem_ops.hip -- element operations on strided complex matrices: the scale,
// difference and mean-square reductions that VSIPL callers (and the
// reference's own tests of this path) apply to FFT operands and results.
// One lane per element over a grid-stride loop; the lane order follows the
// smaller of the two strides so either storage order coalesces.
#include <hip/hip_runtime.h>
#include "vhip.h"

extern "C" int vhip_cu_count(int *cus);
extern "C" int vhip_malloc(void **p, size_t bytes);
extern "C" int vhip_free(void *p);

namespace {

template <typename T>
struct Cx {
    T r, i;
};

template <typename T>
__device__ __forceinline__ Cx<T> ld(const vhip_mat &a, long o) {
    if (a.split) return {((const T *)a.re)[o], ((const T *)a.im)[o]};
    return {((const T *)a.re)[2 * o], ((const T *)a.re)[2 * o + 1]};
}

template <typename T>
__device__ __forceinline__ void st(const vhip_mat &a, long o, Cx<T> v) {
    if (a.split) {
        ((T *)a.re)[o] = v.r;
        ((T *)a.im)[o] = v.i;
    } else {
        ((T *)a.re)[2 * o] = v.r;
        ((T *)a.re)[2 * o + 1] = v.i;
    }
}

// element t of the m x n iteration space, in the order of the smaller stride
__device__ __forceinline__ void ij(long t, long m, long n, bool rows_inner, long &i, long &j) {
    if (rows_inner) { i = t / n; j = t - i * n; }
    else { j = t / m; i = t - j * m; }
}

template <typename T>
__global__ void __launch_bounds__(256) scale_kernel(const vhip_mat a, const vhip_mat r, long m, long n,
                                                    T alpha, bool rows_inner) {
    const long total = m * n;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
        long i, j;
        ij(t, m, n, rows_inner, i, j);
        Cx<T> v = ld<T>(a, i * a.cs + j * a.rs);
        st<T>(r, i * r.cs + j * r.rs, {alpha * v.r, alpha * v.i});
    }
}

template <typename T>
__global__ void __launch_bounds__(256) sub_kernel(const vhip_mat a, const vhip_mat b, const vhip_mat c, long m,
                                                  long n, bool rows_inner) {
    const long total = m * n;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
        long i, j;
        ij(t, m, n, rows_inner, i, j);
        const Cx<T> x = ld<T>(a, i * a.cs + j * a.rs), y = ld<T>(b, i * b.cs + j * b.rs);
        st<T>(c, i * c.cs + j * c.rs, {x.r - y.r, x.i - y.i});
    }
}

// one partial sum of |a|^2 per workgroup (double), reduced in LDS
template <typename T>
__global__ void __launch_bounds__(256) sumsq_kernel(const vhip_mat a, long m, long n, bool rows_inner,
                                                    double *partial) {
    __shared__ double red[256];
    const long total = m * n;
    double s = 0.0;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
        long i, j;
        ij(t, m, n, rows_inner, i, j);
        const Cx<T> v = ld<T>(a, i * a.cs + j * a.rs);
        s += (double)v.r * (double)v.r + (double)v.i * (double)v.i;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// real element operations (VHIP_EOP_*); products and sums round separately
// (no fused multiply-add), as the reference's scalar C loops do, so results
// are bit-identical to them
template <typename T>
__global__ void __launch_bounds__(256) real_kernel(int op, const vhip_mat a, const vhip_mat b, const vhip_mat r,
                                                   const T *vec, long vs, T alpha, long m, long n,
                                                   bool rows_inner) {
#pragma clang fp contract(off)
    const long total = m * n;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
        long i, j;
        ij(t, m, n, rows_inner, i, j);
        const long ro = i * r.cs + j * r.rs;
        T y;
        switch (op) {
        case VHIP_EOP_FILL: y = alpha; break;
        case VHIP_EOP_SMUL: y = alpha * ((const T *)a.re)[i * a.cs + j * a.rs]; break;
        case VHIP_EOP_ADD: y = ((const T *)a.re)[i * a.cs + j * a.rs] + ((const T *)b.re)[i * b.cs + j * b.rs]; break;
        case VHIP_EOP_VMUL_ROW: y = vec[j * vs] * ((const T *)a.re)[i * a.cs + j * a.rs]; break;
        case VHIP_EOP_VMUL_COL: y = vec[i * vs] * ((const T *)a.re)[i * a.cs + j * a.rs]; break;
        default: {   // VHIP_EOP_MAGSQ: a complex
            const Cx<T> x = ld<T>(a, i * a.cs + j * a.rs);
            y = x.r * x.r + x.i * x.i;
        }
        }
        ((T *)r.re)[ro] = y;
    }
}

long grid_for(long total) {
    int cus = 256;
    vhip_cu_count(&cus);
    long g = (total + 255) / 256, cap = (long)(cus > 0 ? cus : 256) * 16;
    if (g > cap) g = cap;
    return g < 1 ? 1 : g;
}

bool rows_inner_of(const vhip_mat *a) {
    const long rs = a->rs < 0 ? -a->rs : a->rs, cs = a->cs < 0 ? -a->cs : a->cs;
    return rs <= cs;
}

}  // namespace

extern "C" int vhip_elem_scale(int prec, const vhip_mat *a, const vhip_mat *r, long m, long n, double alpha,
                               void *stream) {
    if (m <= 0 || n <= 0) return 0;
    const long g = grid_for(m * n);
    const bool ri = rows_inner_of(r);
    hipStream_t s = (hipStream_t)stream;
    if (prec == 4) hipLaunchKernelGGL(scale_kernel<float>, dim3((unsigned)g), dim3(256), 0, s, *a, *r, m, n, (float)alpha, ri);
    else if (prec == 8) hipLaunchKernelGGL(scale_kernel<double>, dim3((unsigned)g), dim3(256), 0, s, *a, *r, m, n, alpha, ri);
    else return -1;
    return (int)hipGetLastError();
}

extern "C" int vhip_elem_sub(int prec, const vhip_mat *a, const vhip_mat *b, const vhip_mat *c, long m, long n,
                             void *stream) {
    if (m <= 0 || n <= 0) return 0;
    const long g = grid_for(m * n);
    const bool ri = rows_inner_of(c);
    hipStream_t s = (hipStream_t)stream;
    if (prec == 4) hipLaunchKernelGGL(sub_kernel<float>, dim3((unsigned)g), dim3(256), 0, s, *a, *b, *c, m, n, ri);
    else if (prec == 8) hipLaunchKernelGGL(sub_kernel<double>, dim3((unsigned)g), dim3(256), 0, s, *a, *b, *c, m, n, ri);
    else return -1;
    return (int)hipGetLastError();
}

extern "C" int vhip_elem_sumsq(int prec, const vhip_mat *a, long m, long n, double *out, void *stream) {
    *out = 0.0;
    if (m <= 0 || n <= 0) return 0;
    const long g = grid_for(m * n);
    const bool ri = rows_inner_of(a);
    hipStream_t s = (hipStream_t)stream;
    void *partial = nullptr;
    int rc = vhip_malloc(&partial, sizeof(double) * (size_t)g);
    if (rc) return rc;
    if (prec == 4) hipLaunchKernelGGL(sumsq_kernel<float>, dim3((unsigned)g), dim3(256), 0, s, *a, m, n, ri, (double *)partial);
    else if (prec == 8) hipLaunchKernelGGL(sumsq_kernel<double>, dim3((unsigned)g), dim3(256), 0, s, *a, m, n, ri, (double *)partial);
    else rc = -1;
    if (!rc) rc = (int)hipGetLastError();
    double *host = (double *)malloc(sizeof(double) * (size_t)g);
    if (!host) rc = rc ? rc : -1;
    if (!rc) rc = (int)hipMemcpyAsync(host, partial, sizeof(double) * (size_t)g, hipMemcpyDeviceToHost, s);
    if (!rc) rc = (int)hipStreamSynchronize(s);
    if (!rc) {
        double t = 0.0;
        for (long k = 0; k < g; ++k) t += host[k];
        *out = t;
    }
    free(host);
    vhip_free(partial);
    return rc;
}

extern "C" int vhip_elem_real(int prec, int op, const vhip_mat *a, const vhip_mat *b, const vhip_mat *r,
                              const void *vec, long vs, double alpha, long m, long n, void *stream) {
    if (m <= 0 || n <= 0) return 0;
    if (op < VHIP_EOP_FILL || op > VHIP_EOP_MAGSQ) return -1;
    const long g = grid_for(m * n);
    const bool ri = rows_inner_of(r);
    const vhip_mat z = {};
    const vhip_mat &aa = a ? *a : z, &bb = b ? *b : z;
    hipStream_t s = (hipStream_t)stream;
    if (prec == 4)
        hipLaunchKernelGGL(real_kernel<float>, dim3((unsigned)g), dim3(256), 0, s, op, aa, bb, *r,
                           (const float *)vec, vs, (float)alpha, m, n, ri);
    else if (prec == 8)
        hipLaunchKernelGGL(real_kernel<double>, dim3((unsigned)g), dim3(256), 0, s, op, aa, bb, *r,
                           (const double *)vec, vs, alpha, m, n, ri);
    else
        return -1;
    return (int)hipGetLastError();
}
