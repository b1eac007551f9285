// fftm_batch.hip -- persistent launches that run many multiple-FFTs of uneven
// lengths and radices.
//
// The host lays the work out as a list of items (a group of vectors of one
// transform, see vhip_batch_item) and a table of per-transform descriptors.
// Every workgroup of the persistent grid takes the next item from an atomic
// head counter, runs the generic Stockham body on it with the schedule of
// that transform, and goes back for more until the list is empty, so the
// grid never waits for the slowest transform of a batch.
//
// Power-of-two lengths from 256 to 8192 on contiguous rows run the
// specialised power-of-two stages (8192 / n rows per item); the rest run the
// generic body.  Each workgroup fetches its next item index while it
// transforms the current one.
//
// Lengths longer than one workgroup's LDS run as two passes (n = n1 * n2).
// Their second-pass items go to a second list drained by a second persistent
// launch on the same stream: the kernel boundary makes the first pass's
// stores visible to every XCD, where an in-kernel handoff would need an
// agent-scope release (a write-back of the whole L2 of the XCD) per item.
#include "fftm_pow2.hpp"

namespace vhipk {
namespace {

__global__ void batch_reset(unsigned *ctl, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) ctl[i] = 0u;
}

__device__ __forceinline__ unsigned next_index(unsigned *head) {
    return __hip_atomic_fetch_add(head, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ctl[0]: next item; ctl[1]: workgroups finished.  The last workgroup to
// finish clears both (every workgroup has made its final fetch by then), so
// the next launch on the stream starts from zero without a clearing kernel.
template <int MAXE>
__global__ void __launch_bounds__(512)
    batch_kernel(const vhip_fftm_desc *__restrict__ descs, const vhip_batch_item *__restrict__ items,
                 long nitems, unsigned *ctl) {
    // descriptors and items are read-only for the launch (restrict): their
    // fields load as scalars, so buffer descriptors built from them stay in
    // SGPRs
    extern __shared__ __attribute__((aligned(16))) char smem[];
    __shared__ long s_item;
    if (threadIdx.x == 0) s_item = (long)next_index(&ctl[0]);
    for (;;) {
        __syncthreads();  // s_item is set; the previous item is done with LDS
        const long it = s_item;
        if (it >= nitems) break;
        long nxt = 0;
        if (threadIdx.x == 0) nxt = (long)next_index(&ctl[0]);   // overlaps the item below
        const vhip_batch_item item = items[it];
        const int di = __builtin_amdgcn_readfirstlane(item.desc);
        const vhip_fftm_desc &d = descs[di];
        const int l2 = __builtin_amdgcn_readfirstlane(p2::pow2_batch_log2(d));
        switch (l2) {
        case 8: p2::pow2_batch_item<8>(d, item.vbase, smem); break;
        case 9: p2::pow2_batch_item<9>(d, item.vbase, smem); break;
        case 10: p2::pow2_batch_item<10>(d, item.vbase, smem); break;
        case 11: p2::pow2_batch_item<11>(d, item.vbase, smem); break;
        case 12: p2::pow2_batch_item<12>(d, item.vbase, smem); break;
        case 13: p2::pow2_batch_item<13>(d, item.vbase, smem); break;
        default: {
            const long left = d.nvec - item.vbase;
            const int nv = left < d.vpb ? (int)left : d.vpb;
            fft_body<float, MAXE, VHIP_C2C, true>(d, item.vbase, nv, smem);
        }
        }
        __syncthreads();  // every lane has read s_item
        if (threadIdx.x == 0) s_item = nxt;
    }
    if (threadIdx.x == 0) {
        const unsigned fin = __hip_atomic_fetch_add(&ctl[1], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (fin == gridDim.x - 1) {
            __hip_atomic_store(&ctl[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&ctl[1], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
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

// ---- the no-allocation batched entry point's scratch ------------------------
//
// One file-scope device array per GPU holds the control counters, the launch
// descriptors and the work items of vsip_amd_batch_run_noalloc_f.  Host code
// never uses the array's name as an address (in host code it names a host-side
// stand-in): at set-up, with each device current, hipGetSymbolAddress gives
// that device's own copy, and only that address is passed to kernels and
// copies.  Two events per device order the reuse of the host staging halves.
static __device__ __align__(256) unsigned char scratch[VHIP_NOALLOC_BYTES];

namespace {
constexpr int kMaxDev = 64;
struct NoallocDev {
    void *ptr;
    hipEvent_t ev[2];
    int recorded[2];
    int ready;
};
NoallocDev g_na[kMaxDev];
int g_na_dev = -1;   // the device the library's stream runs on
}  // namespace

extern "C" int vhip_noalloc_setup(void) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return (int)e;
    if (dev < 0 || dev >= kMaxDev) return (int)hipErrorInvalidDevice;
    NoallocDev &d = g_na[dev];
    if (!d.ready) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return (int)e;
        if (p == nullptr || ((size_t)p & 255) != 0) return (int)hipErrorInvalidSymbol;
        for (int i = 0; i < 2; ++i) {
            e = hipEventCreateWithFlags(&d.ev[i], hipEventDisableTiming);
            if (e != hipSuccess) return (int)e;
            d.recorded[i] = 0;
        }
        d.ptr = p;
        d.ready = 1;
    }
    if (g_na_dev < 0) g_na_dev = dev;
    return 0;
}

extern "C" int vhip_noalloc_scratch(void **ptr) {
    if (g_na_dev < 0 || !g_na[g_na_dev].ready) return (int)hipErrorNotInitialized;
    *ptr = g_na[g_na_dev].ptr;
    return 0;
}

extern "C" int vhip_noalloc_stage_wait(int slot) {
    if (g_na_dev < 0 || slot < 0 || slot > 1) return (int)hipErrorInvalidValue;
    NoallocDev &d = g_na[g_na_dev];
    if (!d.recorded[slot]) return 0;
    return (int)hipEventSynchronize(d.ev[slot]);
}

extern "C" int vhip_noalloc_stage_done(int slot, void *stream) {
    if (g_na_dev < 0 || slot < 0 || slot > 1) return (int)hipErrorInvalidValue;
    NoallocDev &d = g_na[g_na_dev];
    hipError_t e = hipEventRecord(d.ev[slot], (hipStream_t)stream);
    if (e == hipSuccess) d.recorded[slot] = 1;
    return (int)e;
}

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

extern "C" int vhip_batch_clear(unsigned *ctl, void *stream) {
    hipLaunchKernelGGL(vhipk::batch_reset, dim3(1), dim3(64), 0, (hipStream_t)stream, ctl, 4);
    return (int)hipGetLastError();
}

extern "C" int vhip_batch_launch(const vhip_fftm_desc *descs, const vhip_batch_item *items, long nitems1,
                                 long nitems2, unsigned *ctl, int threads, int maxe, int lds_bytes,
                                 int grid, void *stream) {
    if (threads != 512 || grid < 1 || nitems1 < 0 || nitems2 < 0) return -1;
    hipStream_t s = (hipStream_t)stream;
    hipError_t e = hipSuccess;
    long g1 = nitems1 < grid ? nitems1 : grid;
    if (g1 > 0)
        hipLaunchKernelGGL(vhipk::pick_batch(maxe), dim3((unsigned)g1), dim3(threads), (size_t)lds_bytes, s,
                           descs, items, nitems1, ctl);
    e = hipGetLastError();
    if (e != hipSuccess || nitems2 == 0) return (int)e;
    long g2 = nitems2 < grid ? nitems2 : grid;
    hipLaunchKernelGGL(vhipk::pick_batch(maxe), dim3((unsigned)g2), dim3(threads), (size_t)lds_bytes, s,
                       descs, items + nitems1, nitems2, ctl + 2);
    return (int)hipGetLastError();
}
