// fftm_batch.hip -- one persistent launch for many multiple-FFTs of uneven
// lengths and radices.
//
// The host lays the work out as a list of items (a group of vectors of one
// transform, see vhip_batch_item) and a table of per-transform descriptors.
// Every workgroup of the persistent grid takes the next item from an atomic
// head counter, runs the generic Stockham body on it with the schedule of
// that transform, and goes back for more until the list is empty, so the
// grid never waits for the slowest transform of a batch.
//
// Lengths longer than one workgroup's LDS run as two passes (n = n1 * n2);
// the second-pass items of a transform come after all first-pass items in
// the list and wait on a per-transform counter that the first-pass items
// release.  A workgroup only waits after it took a second-pass item, which
// happens only once every first-pass item has been taken by a running
// workgroup, so the wait always ends.  The counter handoff uses agent-scope
// release/acquire atomics, which write back and invalidate the per-XCD L2
// so the second pass sees the first pass's stores from any XCD.
#include "fftm_core.hpp"

namespace vhipk {
namespace {

__global__ void batch_reset(unsigned *ctl, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) ctl[i] = 0u;
}

template <int MAXE>
__global__ void __launch_bounds__(512)
    batch_kernel(const vhip_fftm_desc *descs, const vhip_batch_item *items, long nitems,
                 unsigned *ctl) {
    extern __shared__ __attribute__((aligned(16))) char smem[];
    __shared__ long s_item;
    for (;;) {
        __syncthreads();  // the previous item is done with LDS and s_item
        if (threadIdx.x == 0)
            s_item = (long)__hip_atomic_fetch_add(&ctl[0], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
        const long it = s_item;
        if (it >= nitems) break;
        const vhip_batch_item item = items[it];
        const int di = __builtin_amdgcn_readfirstlane(item.desc);
        const vhip_fftm_desc &d = descs[di];
        if (item.wait_slot >= 0) {
            if (threadIdx.x == 0) {
                unsigned *c = &ctl[1 + item.wait_slot];
                while (__hip_atomic_load(c, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) <
                       (unsigned)item.wait_count)
                    __builtin_amdgcn_s_sleep(2);
            }
            __syncthreads();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");  // every wave drops stale lines
        }
        const long left = d.nvec - item.vbase;
        const int nv = left < d.vpb ? (int)left : d.vpb;
        fft_body<float, MAXE, VHIP_C2C>(d, item.vbase, nv, smem);
        if (item.signal_slot >= 0) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");  // every wave's stores are visible
            __syncthreads();
            if (threadIdx.x == 0)
                __hip_atomic_fetch_add(&ctl[1 + item.signal_slot], 1u, __ATOMIC_RELEASE,
                                       __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

typedef void (*batch_fn)(const vhip_fftm_desc *, const vhip_batch_item *, long, unsigned *);

batch_fn pick_batch(int maxe) {
    return maxe <= 16 ? batch_kernel<16> : batch_kernel<32>;
}

}  // namespace

int batch_set_lds_limit(int bytes) {
    const void *f[2] = {(const void *)batch_kernel<16>, (const void *)batch_kernel<32>};
    for (int i = 0; i < 2; ++i) {
        // the kernel's own __shared__ word comes out of the same per-workgroup LDS
        hipError_t e = hipFuncSetAttribute(f[i], hipFuncAttributeMaxDynamicSharedMemorySize, bytes - 256);
        if (e != hipSuccess) return (int)e;
    }
    return 0;
}

}  // namespace vhipk

extern "C" int vhip_batch_max_grid(int threads, int maxe, int lds_bytes, int *grid) {
    int per_cu = 0, cus = 0, dev = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per_cu, (const void *)vhipk::pick_batch(maxe), threads, (size_t)lds_bytes);
    if (e != hipSuccess) return (int)e;
    e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) return (int)e;
    if (per_cu < 1) per_cu = 1;
    *grid = per_cu * cus;
    return 0;
}

extern "C" int vhip_batch_launch(const vhip_fftm_desc *descs, const vhip_batch_item *items, long nitems,
                                 unsigned *ctl, int nslots, int threads, int maxe, int lds_bytes,
                                 int grid, void *stream) {
    if (threads != 512 || grid < 1 || nitems < 0) return -1;
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(vhipk::batch_reset, dim3(1), dim3(256), 0, s, ctl, nslots + 1);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return (int)e;
    hipLaunchKernelGGL(vhipk::pick_batch(maxe), dim3(grid), dim3(threads), (size_t)lds_bytes, s,
                       descs, items, nitems, ctl);
    return (int)hipGetLastError();
}
