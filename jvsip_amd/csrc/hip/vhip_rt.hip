// vhip_rt.hip -- device memory, copies and the library stream behind vhip.h.
//
// All work of the plain VSIPL entry points is ordered on one library stream
// per process; host-visible completion happens at the synchronous copies
// (admit/release, get/put, copyto/copyfrom) and at vhip_sync().
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>
#include <atomic>
#include <mutex>
#include <unordered_map>
#include "vhip.h"

extern "C" int vhip_fftm_set_lds_limit(int bytes);

namespace {
hipStream_t g_stream = nullptr;
int g_lds_max = 65536;
int g_cus = 0;
bool g_ready = false;
std::mutex g_init_mu;
std::atomic<size_t> g_bytes{0};
std::atomic<size_t> g_budget{0};
std::atomic<int> g_reserved_cus{0};   // CUs left to a resident kernel

// sizes of live allocations, so frees are charged back against the budget
std::mutex g_alloc_mu;
std::unordered_map<void *, size_t> g_sizes;
}  // namespace

extern "C" int vhip_init(void) {
    std::lock_guard<std::mutex> lk(g_init_mu);
    if (g_ready) return 0;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) return (int)e;
    if (n < 1) return (int)hipErrorNoDevice;
    hipDeviceProp_t p;
    e = hipGetDeviceProperties(&p, 0);
    if (e != hipSuccess) return (int)e;
    g_cus = p.multiProcessorCount;
    int lds = 0;
    if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, 0) == hipSuccess && lds > 0)
        g_lds_max = lds;
    e = hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking);
    if (e != hipSuccess) return (int)e;
    int rc = vhip_fftm_set_lds_limit(g_lds_max);
    if (rc != 0) return rc;
    rc = vhip_noalloc_setup();         // this device's no-allocation scratch
    if (rc != 0) return rc;
    const char *b = getenv("VSIP_AMD_HBM_BUDGET");
    if (b) g_budget.store((size_t)strtoull(b, nullptr, 10));
    g_ready = true;
    return 0;
}

extern "C" int vhip_device_count(int *count) {
    hipError_t e = hipGetDeviceCount(count);
    return (int)e;
}

extern "C" int vhip_lds_max(int *bytes) {
    *bytes = g_lds_max;
    return 0;
}

extern "C" int vhip_cu_count(int *cus) {
    const int r = g_reserved_cus.load();
    *cus = (r > 0 && r < g_cus) ? g_cus - r : g_cus;
    return 0;
}

extern "C" int vhip_set_reserved_cus(int n) {
    if (n < 0) return -1;
    g_reserved_cus.store(n);
    return 0;
}

extern "C" int vhip_malloc(void **p, size_t bytes) {
    *p = nullptr;
    if (bytes == 0) bytes = 1;
    size_t cap = g_budget.load();
    size_t cur = g_bytes.fetch_add(bytes);
    if (cap && cur + bytes > cap) {
        g_bytes.fetch_sub(bytes);
        return (int)hipErrorOutOfMemory;
    }
    void *raw = nullptr;
    hipError_t e = hipMalloc(&raw, bytes);
    if (e != hipSuccess) {
        g_bytes.fetch_sub(bytes);
        (void)hipGetLastError();
        return (int)e;
    }
    {
        std::lock_guard<std::mutex> lk(g_alloc_mu);
        g_sizes[raw] = bytes;
    }
    *p = raw;
    return 0;
}

extern "C" int vhip_free(void *p) {
    if (!p) return 0;
    {
        std::lock_guard<std::mutex> lk(g_alloc_mu);
        auto it = g_sizes.find(p);
        if (it != g_sizes.end()) {
            g_bytes.fetch_sub(it->second);
            g_sizes.erase(it);
        }
    }
    return (int)hipFree(p);
}

extern "C" size_t vhip_bytes_in_use(void) { return g_bytes.load(); }
extern "C" void vhip_set_budget(size_t bytes) { g_budget.store(bytes); }

extern "C" int vhip_memcpy_h2d(void *dst, const void *src, size_t bytes) {
    if (!bytes) return 0;
    hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, g_stream);
    if (e == hipSuccess) e = hipStreamSynchronize(g_stream);
    return (int)e;
}

extern "C" int vhip_memcpy_d2h(void *dst, const void *src, size_t bytes) {
    if (!bytes) return 0;
    hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, g_stream);
    if (e == hipSuccess) e = hipStreamSynchronize(g_stream);
    return (int)e;
}

extern "C" int vhip_memcpy_d2d(void *dst, const void *src, size_t bytes) {
    if (!bytes) return 0;
    return (int)hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, g_stream);
}

extern "C" int vhip_memset(void *dst, int value, size_t bytes) {
    if (!bytes) return 0;
    hipError_t e = hipMemsetAsync(dst, value, bytes, g_stream);
    if (e == hipSuccess) e = hipStreamSynchronize(g_stream);
    return (int)e;
}

extern "C" int vhip_memcpy2d_h2d(void *dst, size_t dpitch, const void *src, size_t spitch,
                                 size_t width, size_t rows) {
    if (!width || !rows) return 0;
    hipError_t e = hipMemcpy2DAsync(dst, dpitch, src, spitch, width, rows, hipMemcpyHostToDevice, g_stream);
    if (e == hipSuccess) e = hipStreamSynchronize(g_stream);
    return (int)e;
}

extern "C" int vhip_memcpy2d_d2h(void *dst, size_t dpitch, const void *src, size_t spitch,
                                 size_t width, size_t rows) {
    if (!width || !rows) return 0;
    hipError_t e = hipMemcpy2DAsync(dst, dpitch, src, spitch, width, rows, hipMemcpyDeviceToHost, g_stream);
    if (e == hipSuccess) e = hipStreamSynchronize(g_stream);
    return (int)e;
}

extern "C" int vhip_sync(void) {
    if (!g_stream) return 0;
    return (int)hipStreamSynchronize(g_stream);
}

extern "C" void *vhip_stream(void) { return (void *)g_stream; }

extern "C" const char *vhip_error_string(int code) {
    if (code < 0) return "vhip: invalid argument";
    return hipGetErrorString((hipError_t)code);
}

// ---------------------------------------------------------------------------
// strided matrix copy (view <-> dense staging buffer), any element type
// ---------------------------------------------------------------------------
namespace {
template <typename T, bool CPLX>
__global__ void mat_copy_kernel(const vhip_mat src, const vhip_mat dst, long m, long n) {
    const long total = m * n;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (long)gridDim.x * blockDim.x) {
        const long i = t / n, j = t - (t / n) * n;
        const long so = i * src.cs + j * src.rs;
        const long dof = i * dst.cs + j * dst.rs;
        T re, im = (T)0;
        if (CPLX && src.split) {
            re = ((const T *)src.re)[so];
            im = ((const T *)src.im)[so];
        } else if (CPLX) {
            re = ((const T *)src.re)[2 * so];
            im = ((const T *)src.re)[2 * so + 1];
        } else {
            re = ((const T *)src.re)[so];
        }
        if (CPLX && dst.split) {
            ((T *)dst.re)[dof] = re;
            ((T *)dst.im)[dof] = im;
        } else if (CPLX) {
            ((T *)dst.re)[2 * dof] = re;
            ((T *)dst.re)[2 * dof + 1] = im;
        } else {
            ((T *)dst.re)[dof] = re;
        }
    }
}
}  // namespace

extern "C" int vhip_mat_copy(int prec, int cplx, const vhip_mat *src, const vhip_mat *dst,
                             long m, long n, void *stream) {
    if (m <= 0 || n <= 0) return 0;
    const long total = m * n;
    long blocks = (total + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipStream_t s = stream ? (hipStream_t)stream : g_stream;
    if (prec == 4 && cplx) hipLaunchKernelGGL((mat_copy_kernel<float, true>), dim3(blocks), dim3(256), 0, s, *src, *dst, m, n);
    else if (prec == 4) hipLaunchKernelGGL((mat_copy_kernel<float, false>), dim3(blocks), dim3(256), 0, s, *src, *dst, m, n);
    else if (prec == 8 && cplx) hipLaunchKernelGGL((mat_copy_kernel<double, true>), dim3(blocks), dim3(256), 0, s, *src, *dst, m, n);
    else if (prec == 8) hipLaunchKernelGGL((mat_copy_kernel<double, false>), dim3(blocks), dim3(256), 0, s, *src, *dst, m, n);
    else return -1;
    return (int)hipGetLastError();
}

// ---------------------------------------------------------------------------
// streams, graphs and event timing
// ---------------------------------------------------------------------------
extern "C" int vhip_stream_create(void **stream) {
    hipStream_t s = nullptr;
    hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    *stream = (void *)s;
    return (int)e;
}

extern "C" int vhip_stream_destroy(void *stream) {
    if (!stream) return 0;
    return (int)hipStreamDestroy((hipStream_t)stream);
}

extern "C" int vhip_graph_begin(void *stream) {
    return (int)hipStreamBeginCapture((hipStream_t)stream, hipStreamCaptureModeThreadLocal);
}

extern "C" int vhip_graph_end(void *stream, void **graph_exec) {
    hipGraph_t g = nullptr;
    hipGraphExec_t x = nullptr;
    *graph_exec = nullptr;
    hipError_t e = hipStreamEndCapture((hipStream_t)stream, &g);
    if (e != hipSuccess) return (int)e;
    e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return (int)e;
    *graph_exec = (void *)x;
    return 0;
}

extern "C" int vhip_graph_launch(void *graph_exec, void *stream) {
    return (int)hipGraphLaunch((hipGraphExec_t)graph_exec, stream ? (hipStream_t)stream : g_stream);
}

extern "C" int vhip_graph_destroy(void *graph_exec) {
    if (!graph_exec) return 0;
    return (int)hipGraphExecDestroy((hipGraphExec_t)graph_exec);
}

extern "C" int vhip_event_time(void *stream, int (*fn)(void *arg), void *arg, int reps, float *ms) {
    hipStream_t s = stream ? (hipStream_t)stream : g_stream;
    hipEvent_t a, b;
    *ms = 0.f;
    hipError_t e = hipEventCreate(&a);
    if (e != hipSuccess) return (int)e;
    e = hipEventCreate(&b);
    if (e != hipSuccess) { (void)hipEventDestroy(a); return (int)e; }
    int rc = fn(arg);  // warm-up
    if (!rc) rc = (int)hipEventRecord(a, s);
    for (int i = 0; i < reps && !rc; ++i) rc = fn(arg);
    if (!rc) rc = (int)hipEventRecord(b, s);
    if (!rc) rc = (int)hipEventSynchronize(b);
    float t = 0.f;
    if (!rc) rc = (int)hipEventElapsedTime(&t, a, b);
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    *ms = reps > 0 ? t / reps : 0.f;
    return rc;
}

// ---------------------------------------------------------------------------
// HBM reference kernels: in-place read-modify-write and read-only streams
// ---------------------------------------------------------------------------
namespace {
// In-place read+write in the FFT kernels' pattern, at the best cache policy
// measured for it: one 32 KiB row per workgroup of 256 lanes, each lane
// reads 16 8-byte values (non-temporal), the workgroup synchronises, each
// lane writes them back (non-temporal, sc1).  On MI355X this streams
// 6.3 TB/s, against 5.1-5.2 TB/s for a grid-stride float4 loop at the
// default policy.
typedef decltype(__builtin_amdgcn_raw_buffer_load_b64(__amdgpu_buffer_rsrc_t(), 0, 0, 0)) probe_raw64_t;
__global__ void __launch_bounds__(256) probe_rmw(float *p) {
    const __amdgpu_buffer_rsrc_t r =
        __builtin_amdgcn_make_buffer_rsrc(p + (long)blockIdx.x * 8192, (short)0, 32768, 0x00020000);
    float2 x[16];
    const int t = threadIdx.x;
#pragma unroll
    for (int k = 0; k < 16; ++k)
        x[k] = __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, (t + k * 256) * 8, 0, 2));
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        x[k].x *= 1.0000001f;
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(probe_raw64_t, x[k]), r, (t + k * 256) * 8, 0, 18);
    }
}
__global__ void probe_read(const float4 *p, long n, float *sink) {
    float acc = 0.f;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        float4 v = p[i];
        acc += v.x + v.y + v.z + v.w;
    }
    if (acc == 1.2345e30f) sink[0] = acc;  // never true for the probe's data; keeps the loads
}

// Access pattern of the column-mapped transforms with no transform in it:
// workgroup g takes 2^lg adjacent columns of a rows x cols complex FP32 view
// whose rows are ld_in elements apart, reads every row of them into
// registers (lane t: element t + 512 k of the group, adjacent lanes on
// adjacent columns), synchronises, and writes them to the same place of a
// view whose rows are ld_out apart.  Groups are numbered as the column
// kernels number them: consecutive groups on one XCD, so that groups
// sharing cache lines share that XCD's L2.
__device__ __forceinline__ unsigned probe_xcd_group(unsigned b, unsigned g) {
    const unsigned x = b & 7u, i = b >> 3, per = g >> 3, rem = g & 7u;
    return x * per + (x < rem ? x : rem) + i;
}
__global__ void __launch_bounds__(512) probe_cols(const float2 *in, float2 *out, int rows, int cols, long ld_in,
                                                   long ld_out, int lg) {
    // buffer ranges end at each view's last element: rows past the end read
    // 0 and their stores are dropped, with no per-element test
    const unsigned in_bytes = (unsigned)(((long)(rows - 1) * ld_in + cols) * 8);
    const unsigned out_bytes = (unsigned)(((long)(rows - 1) * ld_out + cols) * 8);
    const __amdgpu_buffer_rsrc_t ri = __builtin_amdgcn_make_buffer_rsrc((void *)in, (short)0, in_bytes, 0x00020000);
    const __amdgpu_buffer_rsrc_t ro = __builtin_amdgcn_make_buffer_rsrc(out, (short)0, out_bytes, 0x00020000);
    const int t = threadIdx.x;
    const int rs = 512 >> lg;                                   // rows per pass of the workgroup
    const long c = ((long)probe_xcd_group(blockIdx.x, gridDim.x) << lg) + (t & ((1 << lg) - 1));
    const int r0 = t >> lg;
    // a lane past the last column stays beyond both ranges on every pass
    const bool live = c < cols;
    const unsigned oi = live ? (unsigned)(((long)r0 * ld_in + c) * 8) : 0xfffffff0u;
    const unsigned oo = live ? (unsigned)(((long)r0 * ld_out + c) * 8) : 0xfffffff0u;
    const unsigned si = live ? (unsigned)(rs * ld_in * 8) : 0u, so = live ? (unsigned)(rs * ld_out * 8) : 0u;
    const int passes = (rows + rs - 1) / rs;                    // <= 32, uniform
    float2 x[32];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        if (j * 8 < passes) {
#pragma unroll
            for (int k = j * 8; k < j * 8 + 8; ++k)
                x[k] = __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(ri, oi + k * si, 0, 0));
        }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        if (j * 8 < passes) {
#pragma unroll
            for (int k = j * 8; k < j * 8 + 8; ++k) {
                x[k].x *= 1.0000001f;
                __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(probe_raw64_t, x[k]), ro, oo + k * so, 0, 0);
            }
        }
    }
}
}  // namespace

// one launch of probe_cols; the caller's buffers hold ((rows - 1) * ld + cols)
// complex FP32 elements on each side
extern "C" int vhip_colprobe_launch(const void *in, void *out, int rows, int cols, long ld_in, long ld_out, int lg,
                                    void *stream) {
    if (rows < 1 || cols < 1 || lg < 0 || lg > 4 || ((long)rows << lg) > 32L * 512 || ld_in < cols ||
        ld_out < cols)
        return -1;
    // 32-bit buffer offsets: a chunk of 8 passes may reach 8 passes past
    // the last row; those offsets stay below 2^31 bytes
    const long over = (long)rows + 8L * (512 >> lg);
    if (over * ld_in * 8 >= (1L << 31) || over * ld_out * 8 >= (1L << 31)) return -1;
    const long groups = ((long)cols + (1L << lg) - 1) >> lg;
    hipStream_t s = stream ? (hipStream_t)stream : g_stream;
    hipLaunchKernelGGL(probe_cols, dim3((unsigned)groups), dim3(512), 0, s, (const float2 *)in, (float2 *)out,
                       rows, cols, ld_in, ld_out, lg);
    return (int)hipGetLastError();
}

extern "C" int vhip_probe_launch(int kind, void *buf, size_t bytes, void *sink, void *stream) {
    const long n = (long)(bytes / 16);
    hipStream_t s = stream ? (hipStream_t)stream : g_stream;
    if (kind == 0) {
        const long rows = (long)(bytes / 32768);     // whole 32 KiB rows only
        if (rows < 1 || rows > 0x7fffffffL) return -1;
        hipLaunchKernelGGL(probe_rmw, dim3((unsigned)rows), dim3(256), 0, s, (float *)buf);
    }
    else hipLaunchKernelGGL(probe_read, dim3(32768), dim3(256), 0, s, (const float4 *)buf, n, (float *)sink);
    return (int)hipGetLastError();
}

// ---------------------------------------------------------------------------
// events, pinned host memory, async copies
// ---------------------------------------------------------------------------
extern "C" int vhip_event_create(void **ev) {
    hipEvent_t e = nullptr;
    hipError_t r = hipEventCreateWithFlags(&e, hipEventDisableTiming);
    *ev = (void *)e;
    return (int)r;
}

extern "C" int vhip_event_destroy(void *ev) {
    return ev ? (int)hipEventDestroy((hipEvent_t)ev) : 0;
}

extern "C" int vhip_event_record(void *ev, void *stream) {
    return (int)hipEventRecord((hipEvent_t)ev, stream ? (hipStream_t)stream : g_stream);
}

extern "C" int vhip_event_query(void *ev) {
    hipError_t r = hipEventQuery((hipEvent_t)ev);
    if (r == hipSuccess) return 0;
    if (r == hipErrorNotReady) return 1;
    return (int)r;
}

extern "C" int vhip_host_alloc(void **p, size_t bytes) {
    *p = nullptr;
    return (int)hipHostMalloc(p, bytes ? bytes : 1, hipHostMallocDefault);
}

extern "C" int vhip_host_alloc_coherent(void **p, size_t bytes) {
    *p = nullptr;
    return (int)hipHostMalloc(p, bytes ? bytes : 1, hipHostMallocCoherent | hipHostMallocMapped);
}

extern "C" int vhip_host_device_ptr(void **dev, void *host) {
    *dev = nullptr;
    return (int)hipHostGetDevicePointer(dev, host, 0);
}

extern "C" void vhip_host_publish(unsigned *p, unsigned v) {
    __atomic_store_n(p, v, __ATOMIC_RELEASE);
}

extern "C" unsigned vhip_host_peek(const unsigned *p) {
    return __atomic_load_n(p, __ATOMIC_ACQUIRE);
}

extern "C" int vhip_stream_sync(void *stream) {
    return (int)hipStreamSynchronize((hipStream_t)stream);
}

extern "C" int vhip_stream_done(void *stream) {
    const hipError_t e = hipStreamQuery((hipStream_t)stream);
    if (e == hipSuccess) return 0;
    if (e == hipErrorNotReady) return 1;
    return (int)e;
}

extern "C" int vhip_host_free(void *p) {
    return p ? (int)hipHostFree(p) : 0;
}

extern "C" int vhip_memcpy_h2d_async(void *dst, const void *src, size_t bytes, void *stream) {
    if (!bytes) return 0;
    return (int)hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice,
                               stream ? (hipStream_t)stream : g_stream);
}
