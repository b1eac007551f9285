// fftm_pow2_real.hip -- launches of the power-of-two real-row kernels
// (rcfftmop / crfftmop on contiguous FP32 rows, half length 256 ... 16384;
// device code pow2_real_body in fftm_pow2.hpp).
#include "fftm_pow2.hpp"

namespace vhipk {
namespace {
using namespace p2;

template <int LOG2N, int VPB, int KIND>
__global__ void __launch_bounds__(VPB * (1 << LOG2N) / 16)
    pow2_real_kernel(const vhip_fftm_desc d) {
    extern __shared__ __attribute__((aligned(16))) char smem[];
    pow2_real_body<LOG2N, VPB, KIND>(d, (cpx<float> *)smem);
}

// the pipelined R2C kernel: 2 workgroups per CU, up to 256 VGPRs per lane
template <int LOG2N>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2)))
    pow2_r2c_pipe_kernel(const vhip_fftm_desc d) {
    __shared__ __attribute__((aligned(16))) cpx<float> lds[p2_vpb<LOG2N>() * P2<LOG2N>::template ldsv<kPipePS>()];
    pow2_r2c_pipe_body<LOG2N>(d, lds);
}

// Real rows (R2C / C2R) of half length n = 2^LOG2N, one group of rows per
// workgroup; R2C with half length 256..4096 and at least 4 groups per
// workgroup of a grid of 2 per CU on the pipelined kernel instead (faster at
// every half length measured, 256..4096).  C2R stays on one group per
// workgroup: pipelined forms measured slower at every length (8192 x 2048:
// 26.7-29.4 us against 24.3).  Row strides of the real side are counted in
// reals.
template <int LOG2N, int KIND>
int p2_real_launch(const vhip_fftm_desc *d, void *stream, int cus) {
    constexpr int VPB = p2_vpb<LOG2N>();
    constexpr int N = 1 << LOG2N;
    constexpr int T = VPB * N / 16;
    if (!p2_radices_match<LOG2N>(d) || !d->euler) return -100;
    const long ivs = KIND == VHIP_R2C ? d->in.vs0 / 2 : d->in.vs0;
    const long ovs = KIND == VHIP_C2R ? d->out.vs0 / 2 : d->out.vs0;
    const long ilen = KIND == VHIP_R2C ? N : N + 1, olen = KIND == VHIP_R2C ? N + 1 : N;
    if (d->nvec > 1 && (ivs < ilen || ovs < olen)) return -100;          // rows must not overlap
    if (((long)(VPB - 1) * ivs + ilen) * 8 >= (1L << 31) || ((long)(VPB - 1) * ovs + olen) * 8 >= (1L << 31))
        return -100;
    const long grid = (d->nvec + VPB - 1) / VPB;
    if constexpr (KIND == VHIP_R2C && LOG2N >= 8 && LOG2N <= 12) {
        static const bool off = getenv("VSIP_AMD_P2_PIPE") && atoi(getenv("VSIP_AMD_P2_PIPE")) == 0;
        const long pgrid = 2L * cus;
        if (!off && grid >= 4 * pgrid && grid <= 0x7fffffffL) {
            hipLaunchKernelGGL((pow2_r2c_pipe_kernel<LOG2N>), dim3((unsigned)pgrid), dim3(T), 0,
                               (hipStream_t)stream, *d);
            return (int)hipGetLastError();
        }
    }
    const size_t lds = (size_t)VPB * P2<LOG2N>::template ldsv<kPipePS>() * 8;   // rows padded one element per 16
    hipLaunchKernelGGL((pow2_real_kernel<LOG2N, VPB, KIND>), dim3((unsigned)grid), dim3(T), lds,
                       (hipStream_t)stream, *d);
    return (int)hipGetLastError();
}

template <int LOG2N>
int p2_real_dispatch(const vhip_fftm_desc *d, void *stream, int cus) {
    return d->kind == VHIP_R2C ? p2_real_launch<LOG2N, VHIP_R2C>(d, stream, cus)
                               : p2_real_launch<LOG2N, VHIP_C2R>(d, stream, cus);
}

template <int LOG2N>
int p2_real_set_lds(int bytes) {
    constexpr int VPB = p2_vpb<LOG2N>();
    const void *rf[2] = {(const void *)pow2_real_kernel<LOG2N, VPB, VHIP_R2C>,
                         (const void *)pow2_real_kernel<LOG2N, VPB, VHIP_C2R>};
    for (int i = 0; i < 2; ++i) {
        hipError_t e = hipFuncSetAttribute(rf[i], hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
        if (e != hipSuccess) return (int)e;
    }
    return 0;
}

}  // namespace

// Launch the real-row kernel when it applies; -100 when it does not.
int pow2_real_try_launch(const vhip_fftm_desc *d, void *stream, int cus) {
    if (d->prec != 4 || d->ptw || (d->kind != VHIP_R2C && d->kind != VHIP_C2R)) return -100;
    if (!d->in.contig || !d->out.contig || d->in.split || d->out.split) return -100;
    if (d->map_in != VHIP_MAP_ALONG || d->map_out != VHIP_MAP_ALONG) return -100;
    if (d->in.nv0 < d->nvec || d->out.nv0 < d->nvec) return -100;   // single-level vectors only
    static const bool no_pow2 = getenv("VSIP_AMD_NO_POW2") != nullptr;   // read once per process
    if (no_pow2) return -100;
    switch (d->n) {
    case 256: return p2_real_dispatch<8>(d, stream, cus);
    case 512: return p2_real_dispatch<9>(d, stream, cus);
    case 1024: return p2_real_dispatch<10>(d, stream, cus);
    case 2048: return p2_real_dispatch<11>(d, stream, cus);
    case 4096: return p2_real_dispatch<12>(d, stream, cus);
    case 8192: return p2_real_dispatch<13>(d, stream, cus);
    case 16384: return p2_real_dispatch<14>(d, stream, cus);
    }
    return -100;
}

int pow2_real_set_lds_limit(int bytes) {
    int rc = p2_real_set_lds<8>(bytes);
    rc = rc ? rc : p2_real_set_lds<9>(bytes);
    rc = rc ? rc : p2_real_set_lds<10>(bytes);
    rc = rc ? rc : p2_real_set_lds<11>(bytes);
    rc = rc ? rc : p2_real_set_lds<12>(bytes);
    rc = rc ? rc : p2_real_set_lds<13>(bytes);
    rc = rc ? rc : p2_real_set_lds<14>(bytes);
    return rc;
}

}  // namespace vhipk
