// fftm_pow2.hip -- launches of the power-of-two row kernel (device code in fftm_pow2.hpp).
#include "fftm_pow2.hpp"

namespace vhipk {
namespace {
using namespace p2;

// variant 0: no prefetch; 1: prefetch with registers capped at 4 waves/SIMD;
// 2: prefetch, registers uncapped
template <int LOG2N, int VPB, bool PLAIN>
__global__ void __launch_bounds__(VPB * (1 << LOG2N) / 16)
    pow2_kernel_v0(const vhip_fftm_desc d, int per_wg) {
    extern __shared__ __attribute__((aligned(16))) char smem[];
    pow2_body<LOG2N, VPB, false, PLAIN>(d, per_wg, (cpx<float> *)smem);
}

template <int LOG2N, int VPB, bool PLAIN>
__global__ void __launch_bounds__(VPB * (1 << LOG2N) / 16, 4)
    pow2_kernel_v1(const vhip_fftm_desc d, int per_wg) {
    extern __shared__ __attribute__((aligned(16))) char smem[];
    pow2_body<LOG2N, VPB, true, PLAIN>(d, per_wg, (cpx<float> *)smem);
}

template <int LOG2N, int VPB, bool PLAIN>
__global__ void __launch_bounds__(VPB * (1 << LOG2N) / 16)
    pow2_kernel_v2(const vhip_fftm_desc d, int per_wg) {
    extern __shared__ __attribute__((aligned(16))) char smem[];
    pow2_body<LOG2N, VPB, true, PLAIN>(d, per_wg, (cpx<float> *)smem);
}

// the pipelined 4096-point kernel: at most 2 workgroups per CU, so the
// compiler may use up to 256 VGPRs
// row groups per workgroup from which the c2c pipelined kernels run
// (VSIP_AMD_P2_PIPE_MIN overrides the given default; read once per process)
static int p2_pipe_min(int dflt) {
    static const int env = getenv("VSIP_AMD_P2_PIPE_MIN") ? atoi(getenv("VSIP_AMD_P2_PIPE_MIN")) : 0;
    return env > 0 ? env : dflt;
}

template <bool PLAIN, bool DB>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2)))
    pow2_pipe_kernel(const vhip_fftm_desc d) {
    __shared__ __attribute__((aligned(16))) cpx<float> lds[(DB ? 2 : 1) * kPipeLdsv];
    pow2_pipe_body<PLAIN, DB>(d, lds);
}

// the 4096-point kernel with cross-lane (DPP) exchanges (device code in
// fftm_pow2.hpp): one LDS transpose per row
template <bool PLAIN>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2)))
    pow2_dpp_kernel(const vhip_fftm_desc d) {
    __shared__ __attribute__((aligned(16))) cpx<float> lds[2 * kDppLds];
    pow2_dpp_body<PLAIN>(d, lds);
}

// 4096-point rows on the pipelined kernel when every workgroup of a grid of
// 2 per CU gets at least 4 rows (measured: 8 rows per workgroup beat one row
// per workgroup by 7% at 4096 x 4096, 4 rows by 6% at 2048 x 4096; from 2
// rows on it gains nothing more)
int p2_pipe_launch(const vhip_fftm_desc *d, void *stream, int cus) {
    static const bool off = getenv("VSIP_AMD_P2_PIPE") && atoi(getenv("VSIP_AMD_P2_PIPE")) == 0;
    const long grid = 2L * cus;
    if (off || d->nvec < p2_pipe_min(4) * grid || d->nvec > 0x7fffffffL) return -100;
    if (d->in.vs0 < 4096 || d->out.vs0 < 4096) return -100;   // rows must not overlap
    // two LDS rows per workgroup (stages 0 and 1 write different rows, two
    // barriers per row instead of four); VSIP_AMD_P2_PIPE_DB=0 keeps one
    static const bool db = !(getenv("VSIP_AMD_P2_PIPE_DB") && atoi(getenv("VSIP_AMD_P2_PIPE_DB")) == 0);
    const bool plain = !d->conj_in && !d->conj_out && d->scale == 1.0;
    hipStream_t st = (hipStream_t)stream;
    // VSIP_AMD_P2_DPP=1: the kernel with cross-lane exchanges (read per call)
    const char *dpp = getenv("VSIP_AMD_P2_DPP");
    if (dpp && atoi(dpp) == 1) {
        if (plain) hipLaunchKernelGGL((pow2_dpp_kernel<true>), dim3((unsigned)grid), dim3(256), 0, st, *d);
        else hipLaunchKernelGGL((pow2_dpp_kernel<false>), dim3((unsigned)grid), dim3(256), 0, st, *d);
        return (int)hipGetLastError();
    }
    if (plain && db) hipLaunchKernelGGL((pow2_pipe_kernel<true, true>), dim3((unsigned)grid), dim3(256), 0, st, *d);
    else if (db) hipLaunchKernelGGL((pow2_pipe_kernel<false, true>), dim3((unsigned)grid), dim3(256), 0, st, *d);
    else if (plain) hipLaunchKernelGGL((pow2_pipe_kernel<true, false>), dim3((unsigned)grid), dim3(256), 0, st, *d);
    else hipLaunchKernelGGL((pow2_pipe_kernel<false, false>), dim3((unsigned)grid), dim3(256), 0, st, *d);
    return (int)hipGetLastError();
}

// the pipelined kernel for rows of 256..2048 (groups of 4096 / N rows)
template <int LOG2N, bool PLAIN>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2)))
    pow2_pipe_gen_kernel(const vhip_fftm_desc d) {
    __shared__ __attribute__((aligned(16))) cpx<float> lds[p2_vpb<LOG2N>() * P2<LOG2N>::template ldsv<kPipePS>()];
    pow2_pipe_gen_body<LOG2N, PLAIN>(d, lds);
}

// the pipelined 8192- and 16384-point kernels: 2 workgroups of 512 lanes or
// 1 of 1024 per CU, 128 VGPRs per lane; their LDS (a padded row and the
// stage-1 twiddles, 70.5 or 137.9 KiB) is dynamic
template <int LOG2N, bool PLAIN>
__global__ void __launch_bounds__(1 << (LOG2N - 4)) __attribute__((amdgpu_waves_per_eu(4, 4)))
    pow2_pipe_big_kernel(const vhip_fftm_desc d) {
    extern __shared__ __attribute__((aligned(16))) char smem[];
    pow2_pipe_big_body<LOG2N, PLAIN>(d, (cpx<float> *)smem);
}

// 8192- / 16384-point rows on the pipelined kernel when every workgroup of
// a grid of 2 / 1 per CU gets at least 4 rows
template <int LOG2N>
int p2_pipe_big_launch(const vhip_fftm_desc *d, void *stream, int cus) {
    static const bool off = getenv("VSIP_AMD_P2_PIPE") && atoi(getenv("VSIP_AMD_P2_PIPE")) == 0;
    constexpr int N = 1 << LOG2N;
    const long grid = (LOG2N == 13 ? 2L : 1L) * cus;
    if (off || d->nvec < p2_pipe_min(4) * grid || d->nvec > 0x7fffffffL) return -100;
    if (d->in.vs0 < N || d->out.vs0 < N) return -100;   // rows must not overlap
    const size_t lds = (size_t)P2B<LOG2N>::LDS * 8;
    const bool plain = !d->conj_in && !d->conj_out && d->scale == 1.0;
    hipStream_t st = (hipStream_t)stream;
    if (plain)
        hipLaunchKernelGGL((pow2_pipe_big_kernel<LOG2N, true>), dim3((unsigned)grid), dim3(N / 16), lds, st, *d);
    else
        hipLaunchKernelGGL((pow2_pipe_big_kernel<LOG2N, false>), dim3((unsigned)grid), dim3(N / 16), lds, st, *d);
    return (int)hipGetLastError();
}

// row groups on the pipelined kernel when every workgroup of a grid of 2 per
// CU gets at least 4 groups (the caller has checked radices, strides, spans)
template <int LOG2N>
int p2_pipe_gen_launch(const vhip_fftm_desc *d, void *stream, int cus) {
    static const bool off = getenv("VSIP_AMD_P2_PIPE") && atoi(getenv("VSIP_AMD_P2_PIPE")) == 0;
    constexpr int VPB = p2_vpb<LOG2N>();
    const long grid = 2L * cus;
    const long ngroups = (d->nvec + VPB - 1) / VPB;
    if (off || ngroups < p2_pipe_min(4) * grid || ngroups > 0x7fffffffL) return -100;
    const bool plain = !d->conj_in && !d->conj_out && d->scale == 1.0;
    hipStream_t st = (hipStream_t)stream;
    if (plain) hipLaunchKernelGGL((pow2_pipe_gen_kernel<LOG2N, true>), dim3((unsigned)grid), dim3(256), 0, st, *d);
    else hipLaunchKernelGGL((pow2_pipe_gen_kernel<LOG2N, false>), dim3((unsigned)grid), dim3(256), 0, st, *d);
    return (int)hipGetLastError();
}

template <int LOG2N>
int p2_launch(const vhip_fftm_desc *d, void *stream, int cus) {
    constexpr int VPB = p2_vpb<LOG2N>();
    constexpr int T = VPB * (1 << LOG2N) / 16;
    if (!p2_radices_match<LOG2N>(d)) return -100;
    // buffer offsets are 32-bit: a group of rows must span less than 2 GiB
    const long span_in = ((long)(VPB - 1) * d->in.vs0 + (1 << LOG2N)) * 8;
    const long span_out = ((long)(VPB - 1) * d->out.vs0 + (1 << LOG2N)) * 8;
    if (d->in.vs0 < (1 << LOG2N) || d->out.vs0 < (1 << LOG2N)) return -100;   // rows must not overlap
    if (span_in >= (1L << 31) || span_out >= (1L << 31)) return -100;
    if constexpr (LOG2N >= 8 && LOG2N <= 11) {
        const int rc = p2_pipe_gen_launch<LOG2N>(d, stream, cus);
        if (rc != -100) return rc;
    }
    const long ngroups = (d->nvec + VPB - 1) / VPB;
    const size_t lds = (size_t)VPB * P2<LOG2N>::LDSV * 8;
    static const int venv = getenv("VSIP_AMD_P2_VARIANT") ? atoi(getenv("VSIP_AMD_P2_VARIANT")) : 1;
    int variant = venv;
    if (T >= 1024) variant = 0;
    // one group per workgroup (measured best up to 8192); beyond that a
    // persistent grid of as many workgroups as the LDS lets every CU hold
    int per_wg = 1;
    if (LOG2N >= 14) {
        int per_cu = (int)(163840 / lds);
        if (per_cu < 1) per_cu = 1;
        if (per_cu > 2048 / T) per_cu = 2048 / T;
        const long wgs = (long)cus * per_cu;
        per_wg = (int)((ngroups + wgs - 1) / wgs);
        if (per_wg < 1) per_wg = 1;
    }
    static const int per_wg_env = getenv("VSIP_AMD_P2_PER_WG") ? atoi(getenv("VSIP_AMD_P2_PER_WG")) : 0;
    if (per_wg_env > 0) per_wg = per_wg_env;
    const long grid = (ngroups + per_wg - 1) / per_wg;
    const bool plain = !d->conj_in && !d->conj_out && d->scale == 1.0;
    const dim3 gd((unsigned)grid), bd(T);
    hipStream_t st = (hipStream_t)stream;
    if (plain) {
        if (variant == 1) hipLaunchKernelGGL((pow2_kernel_v1<LOG2N, VPB, true>), gd, bd, lds, st, *d, per_wg);
        else if (variant == 2) hipLaunchKernelGGL((pow2_kernel_v2<LOG2N, VPB, true>), gd, bd, lds, st, *d, per_wg);
        else hipLaunchKernelGGL((pow2_kernel_v0<LOG2N, VPB, true>), gd, bd, lds, st, *d, per_wg);
    } else {
        if (variant == 1) hipLaunchKernelGGL((pow2_kernel_v1<LOG2N, VPB, false>), gd, bd, lds, st, *d, per_wg);
        else if (variant == 2) hipLaunchKernelGGL((pow2_kernel_v2<LOG2N, VPB, false>), gd, bd, lds, st, *d, per_wg);
        else hipLaunchKernelGGL((pow2_kernel_v0<LOG2N, VPB, false>), gd, bd, lds, st, *d, per_wg);
    }
    return (int)hipGetLastError();
}

template <int LOG2N>
int p2_set_lds(int bytes) {
    constexpr int VPB = p2_vpb<LOG2N>();
    const void *fns[6] = {(const void *)pow2_kernel_v0<LOG2N, VPB, true>, (const void *)pow2_kernel_v1<LOG2N, VPB, true>,
                          (const void *)pow2_kernel_v2<LOG2N, VPB, true>, (const void *)pow2_kernel_v0<LOG2N, VPB, false>,
                          (const void *)pow2_kernel_v1<LOG2N, VPB, false>, (const void *)pow2_kernel_v2<LOG2N, VPB, false>};
    for (int i = 0; i < 6; ++i) {
        hipError_t e = hipFuncSetAttribute(fns[i], hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
        if (e != hipSuccess) return (int)e;
    }
    return 0;
}

}  // namespace

// Launch the power-of-two row kernel when it applies; returns -100 when it
// does not (the caller then uses the generic kernel).
int pow2_try_launch(const vhip_fftm_desc *d, void *stream, int cus) {
    if (d->prec != 4 || d->ptw || d->tap || d->acc) return -100;
    if (d->kind != VHIP_C2C) return pow2_real_try_launch(d, stream, cus);
    if (!d->in.contig || !d->out.contig || d->in.split || d->out.split) return -100;
    if (d->map_in != VHIP_MAP_ALONG || d->map_out != VHIP_MAP_ALONG) return -100;
    if (d->in.nv0 < d->nvec || d->out.nv0 < d->nvec) return -100;   // single-level vectors only
    static const bool no_pow2 = getenv("VSIP_AMD_NO_POW2") != nullptr;   // read once per process
    if (no_pow2) return -100;
    switch (d->n) {
    case 256: return p2_launch<8>(d, stream, cus);
    case 512: return p2_launch<9>(d, stream, cus);
    case 1024: return p2_launch<10>(d, stream, cus);
    case 2048: return p2_launch<11>(d, stream, cus);
    case 4096: {
        if (!p2_radices_match<12>(d)) return -100;
        const int rc = p2_pipe_launch(d, stream, cus);
        return rc == -100 ? p2_launch<12>(d, stream, cus) : rc;
    }
    case 8192: {
        if (!p2_radices_match<13>(d)) return -100;
        const int rc = p2_pipe_big_launch<13>(d, stream, cus);
        return rc == -100 ? p2_launch<13>(d, stream, cus) : rc;
    }
    case 16384: {
        if (!p2_radices_match<14>(d)) return -100;
        const int rc = p2_pipe_big_launch<14>(d, stream, cus);
        return rc == -100 ? p2_launch<14>(d, stream, cus) : rc;
    }
    }
    return -100;
}

// the test p2_pipe_gen_launch makes, for the planner
static int p2_pipelined(int n, long nvec) {
    static const bool off = getenv("VSIP_AMD_P2_PIPE") && atoi(getenv("VSIP_AMD_P2_PIPE")) == 0;
    if (off || n < 256 || n > 2048 || (n & (n - 1)) != 0) return 0;
    int cus = 256;
    vhip_cu_count(&cus);
    if (cus < 1) cus = 256;
    const long vpb = 4096 / n;
    return (nvec + vpb - 1) / vpb >= (long)p2_pipe_min(4) * 2L * cus;
}

int pow2_set_lds_limit(int bytes) {
    int rc = 0;
    for (const void *f : {(const void *)pow2_pipe_big_kernel<13, true>, (const void *)pow2_pipe_big_kernel<13, false>,
                          (const void *)pow2_pipe_big_kernel<14, true>, (const void *)pow2_pipe_big_kernel<14, false>}) {
        const hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
        if (e != hipSuccess) return (int)e;
    }
    rc = rc ? rc : p2_set_lds<8>(bytes);
    rc = rc ? rc : p2_set_lds<9>(bytes);
    rc = rc ? rc : p2_set_lds<10>(bytes);
    rc = rc ? rc : p2_set_lds<11>(bytes);
    rc = rc ? rc : p2_set_lds<12>(bytes);
    rc = rc ? rc : p2_set_lds<13>(bytes);
    rc = rc ? rc : p2_set_lds<14>(bytes);
    return rc;
}

}  // namespace vhipk

extern "C" int vhip_p2_pipelined(int n, long nvec) { return vhipk::p2_pipelined(n, nvec); }
