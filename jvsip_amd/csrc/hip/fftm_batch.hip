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
// generic body, whose stages take their twiddles from a table in LDS.  Each
// workgroup fetches its next item index while it transforms the current one.
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


// Twiddle table of the generic items, copied into LDS from the one the host
// appends to the schedule's stage twiddles (d.btw_off): W_n^e = lo[e mod
// 2^lg] hi[e >> lg], lg = ceil(ceil(log2 n) / 2).  Batched single-pass
// lengths are at most 8560 (the LDS of one workgroup), so 2^lg <= 128 and the
// table holds at most 128 + 67 entries.
constexpr int kTwCap = 256;

__device__ __forceinline__ int load_twiddles(cpx<float> *tab, const vhip_fftm_desc &d) {
    const int n = d.n;
    const int cl = n > 1 ? 32 - __builtin_clz((unsigned)(n - 1)) : 0;
    const int lg = (cl + 1) >> 1;
    const int cnt = (1 << lg) + ((n + (1 << lg) - 1) >> lg);
    const rsrc_t src = make_rsrc((const cpx<float> *)d.tw + d.btw_off, (long)cnt * (long)sizeof(cpx<float>));
    for (int t = threadIdx.x; t < cnt && t < kTwCap; t += blockDim.x) tab[t] = Buf<float>::ldc(src, t * 8, 0);
    return lg;
}

// One generic item: the Stockham body of fft_body (C2C) with the stages after
// the first reading their twiddles from the LDS table `tw` (lg as loaded).
template <int MAXE>
__device__ __forceinline__ void batch_generic(const vhip_fftm_desc &d, long vbase, int nv, char *smem,
                                              const cpx<float> *tw, int lg, unsigned &nxt) {
    cpx<float> *lds = (cpx<float> *)smem;
    Shared sh;
    sh.voff_in = (long *)(smem + (size_t)d.vpb * d.lds_stride * sizeof(cpx<float>));
    sh.voff_out = sh.voff_in + d.vpb;
    sh.vidx0 = (int *)(sh.voff_out + d.vpb);
    for (int t = threadIdx.x; t < nv; t += blockDim.x) {
        const long v = vbase + t;
        sh.voff_in[t] = vec_off(d.in, v);
        sh.voff_out[t] = vec_off(d.out, v);
        sh.vidx0[t] = (int)(d.ptw_n > 0 ? v % d.ptw_n : v);
    }
    __syncthreads();
    const float scale = (float)d.scale;
    const float isgn = d.conj_in ? -1.f : 1.f;
    const float osgn = d.conj_out ? -1.f : 1.f;
    BatchCtx<float> bc;
    bc.lo = tw;
    bc.hi = tw + (1 << lg);
    bc.lg = lg;
    const int nst = d.nst;
    // stage 0 outside the loop: the claim is taken on every path after it
    run_stage<float, MAXE, VHIP_C2C, true, true>(d, sh, lds, 0, 1, nv, true, nst == 1, isgn, osgn, scale, &bc);
    nxt = __builtin_amdgcn_readfirstlane(nxt);
    int Ns = d.radix[0];
    for (int s = 1; s < nst; ++s) {
        run_stage<float, MAXE, VHIP_C2C, true, true>(d, sh, lds, s, Ns, nv, false, s == nst - 1, isgn, osgn,
                                                     scale, &bc);
        Ns *= d.radix[s];
    }
}

// One work item: the power-of-two stages for rows of 256..8192 points on
// contiguous rows, else the generic stages with the LDS twiddle table `tw`
// (reloaded when the item length differs from *twn).
// nxt: lane 0's claim of the next item (issued before this item's loads);
// every wave takes its value right after the item's first stage, where the
// wait for it is already covered by the wait for the item's own data (memory
// operations retire in issue order) and no later access is held up by it.
// Item indices are 32-bit: with a 64-bit index the compiler gave the claim's
// register to the next item's index and waited for every store of the item
// at the loop head.
template <int MAXE>
__device__ __forceinline__ void run_item(const vhip_fftm_desc &d, long vbase, char *smem, cpx<float> *tw,
                                         int &twn, int &lg, unsigned &nxt) {
    const int l2 = __builtin_amdgcn_readfirstlane(p2::pow2_batch_log2(d));
    switch (l2) {
    case 8: p2::pow2_batch_item<8>(d, vbase, smem, nxt); break;
    case 9: p2::pow2_batch_item<9>(d, vbase, smem, nxt); break;
    case 10: p2::pow2_batch_item<10>(d, vbase, smem, nxt); break;
    case 11: p2::pow2_batch_item<11>(d, vbase, smem, nxt); break;
    case 12: p2::pow2_batch_item<12>(d, vbase, smem, nxt); break;
    case 13: p2::pow2_batch_item<13>(d, vbase, smem, nxt); break;
    default: {
        const long left = d.nvec - vbase;
        const int nv = left < d.vpb ? (int)left : d.vpb;
        // the previous item's last reads of the table came before the
        // caller's barrier; stage 0's barrier publishes the copy
        if (d.nst > 1 && d.n != twn) {
            lg = load_twiddles(tw, d);
            twn = d.n;
        }
        batch_generic<MAXE>(d, vbase, nv, smem, tw, lg, nxt);
    }
    }
}

// Read every 64-byte line of a descriptor at once: the stages read its fields
// as scalars in several dependent steps, which would otherwise each wait for
// a line of their own.
__device__ __forceinline__ void touch_desc(const vhip_fftm_desc *__restrict__ d) {
    const char *p = (const char *)d;
    static_assert(sizeof(vhip_fftm_desc) <= 512, "eight lines");
    unsigned t0 = *(const unsigned *)(p + 0), t1 = *(const unsigned *)(p + 64), t2 = *(const unsigned *)(p + 128),
             t3 = *(const unsigned *)(p + 192), t4 = *(const unsigned *)(p + 256),
             t5 = *(const unsigned *)(p + 320), t6 = *(const unsigned *)(p + 384),
             t7 = *(const unsigned *)(p + sizeof(vhip_fftm_desc) - 4);
    asm volatile("" ::"s"(t0), "s"(t1), "s"(t2), "s"(t3), "s"(t4), "s"(t5), "s"(t6), "s"(t7));
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
    __shared__ unsigned s_item;
    __shared__ cpx<float> s_tw[kTwCap];
    int twn = 0, lg = 0;
    if (threadIdx.x == 0) s_item = next_index(&ctl[0]);
    for (;;) {
        __syncthreads();  // s_item is set; the previous item is done with LDS
        const long it = (long)__builtin_amdgcn_readfirstlane(s_item);
        if (it >= nitems) break;
        unsigned nxt = 0;
        if (threadIdx.x == 0) nxt = next_index(&ctl[0]);   // overlaps the item below
        const vhip_batch_item item = items[it];
        const int di = __builtin_amdgcn_readfirstlane(item.desc);
        touch_desc(&descs[di]);
        run_item<MAXE>(descs[di], item.vbase, smem, s_tw, twn, lg, nxt);
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

// ---- the resident kernel of the server --------------------------------------
//
// A persistent grid that stays on the device for a whole request stream and
// takes work as the host publishes it, with no launch per request:
//   * items[0 .. nitems) lists every work item of the stream's resident
//     requests in publication order (item.desc = the request's slot); the
//     host publishes a prefix of it by raising *htail in coherent host memory,
//     after writing the request's descriptor into slot hdesc + slot *
//     VHIP_RES_DESC_STRIDE (also coherent host memory, written once per run,
//     so no cache holds an older copy of it);
//   * a workgroup claims the next item index from ctl[0] and waits until it is
//     published: ctl[1] mirrors *htail on the device, and only the waiter that
//     holds the oldest unpublished index (index == ctl[1]) reads *htail over
//     the bus, the others sleep and read ctl[1];
//   * after an item the workgroup's barrier waits for every wave's stores,
//     one lane counts the item off req_left[slot] with an agent-scope release
//     (one L2 write-back for the item), and the workgroup that finishes a
//     request's last item sets hdone[slot] (system scope);
//   * every wait is bounded: a waiter that spins spin_limit times, or sees
//     *hstop, sets ctl[2] (abort), and every workgroup then stops claiming, so
//     the grid always drains.  A request whose items were abandoned never gets
//     its done flag; the host sees the grid end first and reports an error.
// ctl = {next claim, published items seen, abort, 0}: zero before the launch.
__device__ __forceinline__ unsigned resident_claim(unsigned *ctl, const unsigned *htail, const unsigned *hstop,
                                                   unsigned nitems, unsigned spin_limit) {
    if (__hip_atomic_load(&ctl[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return nitems;
    const unsigned i = __hip_atomic_fetch_add(&ctl[0], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (i >= nitems) return nitems;
    // polls are relaxed; one acquire once the item is published
    for (unsigned spins = 0;; ++spins) {
        const unsigned t = __hip_atomic_load(&ctl[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (i < t) {
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            return i;
        }
        if (__hip_atomic_load(&ctl[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return nitems;
        if (spins >= spin_limit) {
            __hip_atomic_store(&ctl[2], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return nitems;
        }
        if (i == t) {
            if (__hip_atomic_load(hstop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM)) {
                __hip_atomic_store(&ctl[2], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                return nitems;
            }
            unsigned h = __hip_atomic_load(htail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            if (h > nitems) h = nitems;
            if (h > t) {
                // the descriptors the host wrote before raising htail live in
                // uncached host memory: no device cache holds them, so the
                // mirror needs no release of this workgroup's writes
                __hip_atomic_fetch_max(&ctl[1], h, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                continue;
            }
            __builtin_amdgcn_s_sleep(1);
        } else {
            __builtin_amdgcn_s_sleep(8);
        }
    }
}

template <int MAXE>
__global__ void __launch_bounds__(512)
    resident_kernel(const char *__restrict__ hdesc, const vhip_batch_item *__restrict__ items,
                    unsigned *req_left, unsigned *ctl, const unsigned *htail, const unsigned *hstop,
                    unsigned *hdone, unsigned nitems, unsigned spin_limit) {
    extern __shared__ __attribute__((aligned(16))) char smem[];
    __shared__ unsigned s_item;
    __shared__ cpx<float> s_tw[kTwCap];
    int twn = 0, lg = 0;
    for (;;) {
        if (threadIdx.x == 0) s_item = resident_claim(ctl, htail, hstop, nitems, spin_limit);
        __syncthreads();
        const unsigned it = __builtin_amdgcn_readfirstlane(s_item);
        if (it >= nitems) break;
        const vhip_batch_item item = items[it];
        const int slot = __builtin_amdgcn_readfirstlane(item.desc);
        const char *sd = hdesc + (size_t)slot * VHIP_RES_DESC_STRIDE;
        // one bus round trip for the whole descriptor: touch each of its
        // 64-byte lines before the stages read its fields
        {
            unsigned t0 = *(const unsigned *)(sd + 0), t1 = *(const unsigned *)(sd + 64),
                     t2 = *(const unsigned *)(sd + 128), t3 = *(const unsigned *)(sd + 192),
                     t4 = *(const unsigned *)(sd + 256), t5 = *(const unsigned *)(sd + 320),
                     t6 = *(const unsigned *)(sd + 384), t7 = *(const unsigned *)(sd + 448);
            asm volatile("" ::"s"(t0), "s"(t1), "s"(t2), "s"(t3), "s"(t4), "s"(t5), "s"(t6), "s"(t7));
        }
        unsigned none = 0;
        run_item<MAXE>(*(const vhip_fftm_desc *)sd, item.vbase, smem, s_tw, twn, lg, none);
        // every wave's stores have reached L2 at the barrier; the count's
        // agent-scope release then writes the XCD's L2 back once for all
        __syncthreads();
        if (threadIdx.x == 0) {
            const unsigned left =
                __hip_atomic_fetch_sub(&req_left[slot], 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            if (left == 1u) {
                // the request's other items were released into the count
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                __hip_atomic_store(&hdone[slot], 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
            }
        }
    }
}

typedef void (*batch_fn)(const vhip_fftm_desc *, const vhip_batch_item *, long, unsigned *);

// the batched kernel has one register class, 16 elements per lane
batch_fn pick_batch(int maxe) {
    return maxe == 16 ? batch_kernel<16> : nullptr;
}

}  // namespace

int batch_set_lds_limit(int bytes) {
    // the kernel's own __shared__ item indices and twiddle table come out of
    // the same per-workgroup LDS
    const int dyn = bytes - 256 - kTwCap * (int)sizeof(cpx<float>);
    hipError_t e = hipFuncSetAttribute((const void *)batch_kernel<16>, hipFuncAttributeMaxDynamicSharedMemorySize, dyn);
    if (e == hipSuccess)
        e = hipFuncSetAttribute((const void *)resident_kernel<16>, hipFuncAttributeMaxDynamicSharedMemorySize, dyn);
    return (int)e;
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
    if (!vhipk::pick_batch(maxe)) return -1;
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

extern "C" int vhip_resident_launch(const vhip_resident_args *a, int grid, int lds_bytes, void *stream) {
    if (!a || grid < 1 || !a->hdesc || !a->items || !a->req_left || !a->ctl || !a->htail || !a->hstop ||
        !a->hdone || ((size_t)a->hdesc & (VHIP_RES_DESC_STRIDE - 1)) != 0)
        return -1;
    static_assert(sizeof(vhip_fftm_desc) <= VHIP_RES_DESC_STRIDE, "a request's descriptor fits its slot");
    hipLaunchKernelGGL(vhipk::resident_kernel<16>, dim3((unsigned)grid), dim3(512), (size_t)lds_bytes,
                       (hipStream_t)stream, (const char *)a->hdesc, a->items, a->req_left, a->ctl, a->htail,
                       a->hstop, a->hdone, a->nitems, a->spin_limit);
    return (int)hipGetLastError();
}

extern "C" int vhip_batch_clear(unsigned *ctl, void *stream) {
    hipLaunchKernelGGL(vhipk::batch_reset, dim3(1), dim3(64), 0, (hipStream_t)stream, ctl, 4);
    return (int)hipGetLastError();
}

extern "C" int vhip_batch_launch(const vhip_fftm_desc *descs, const vhip_batch_item *items, long nitems1,
                                 long nitems2, unsigned *ctl, int threads, int maxe, int lds_bytes,
                                 int grid, void *stream) {
    // item indices are 32-bit (the claim counter and the kernel's index)
    if (threads != 512 || !vhipk::pick_batch(maxe) || grid < 1 || nitems1 < 0 || nitems2 < 0 ||
        nitems1 > 0x7fffffffL - grid || nitems2 > 0x7fffffffL - grid)
        return -1;
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
