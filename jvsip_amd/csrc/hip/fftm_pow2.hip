// fftm_pow2.hip -- specialised multiple-FFT kernel for power-of-two row
// transforms of contiguous interleaved complex FP32 data (the common case and
// the 4096 x 4096 headline).
//
// Same Stockham schedule as the generic kernel (radices 16/8/4/2 with the
// power-of-two bits split evenly over ceil(log2 N / 4) stages, stage twiddles
// W_{Ns R}^{r k} from the plan's table), but with the length, the radices and
// the vectors per workgroup fixed at compile time:
//   - every thread holds 16 elements in every stage (16/R butterflies of
//     radix R) and all index arithmetic reduces to shifts and masks;
//   - HBM and twiddle accesses are buffer loads/stores with one VGPR offset per
//     butterfly and the leg offsets r * (n/R) * 8 in SGPRs, so the 16 loads of
//     a thread need one address register (the buffer range also bounds every
//     access to the rows of the current group);
//   - LDS accesses are a per-thread base plus immediate offsets: the padded
//     index i + (i >> 5) splits over the disjoint bit fields of the butterfly
//     index and the leg;
//   - a workgroup runs `per_wg` consecutive groups of VPB rows (persistent) and
//     may issue the HBM loads of its next group before transforming the
//     current one.
#include <cstdlib>
#include "fftm_core.hpp"

namespace vhipk {
namespace {

typedef __amdgpu_buffer_rsrc_t rsrc_t;
typedef decltype(__builtin_amdgcn_raw_buffer_load_b64(rsrc_t(), 0, 0, 0)) raw64_t;

// word 3 of a raw buffer descriptor on gfx950 (32-bit data format, no swizzle)
constexpr int kBufWord3 = 0x00020000;

__device__ __forceinline__ rsrc_t make_rsrc(const void *p, unsigned bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)bytes, kBufWord3);
}
__device__ __forceinline__ cpx<float> bload(rsrc_t r, int voff, int soff) {
    return __builtin_bit_cast(cpx<float>, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}
__device__ __forceinline__ void bstore(rsrc_t r, int voff, int soff, cpx<float> x) {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(raw64_t, x), r, voff, soff, 0);
}

template <int LOG2N>
struct P2 {
    static constexpr int N = 1 << LOG2N;
    static constexpr int NST = (LOG2N + 3) / 4;
    static constexpr int bits(int s) { return LOG2N / NST + (s < LOG2N % NST ? 1 : 0); }
    static constexpr int log2ns(int s) { return s == 0 ? 0 : log2ns(s - 1) + bits(s - 1); }
    // padded LDS vector stride (odd)
    static constexpr int S0 = N + (N >> 5);
    static constexpr int LDSV = (S0 % 2 == 0) ? S0 + 1 : S0;
};

__host__ __device__ constexpr int padc(int i) { return i + (i >> 5); }

// Stage S: 16 elements per thread in x (already loaded for S == 0), twiddle,
// DFT, write to LDS (or HBM for the last stage).
template <int LOG2N, int VPB, int S, bool PLAIN>
__device__ __forceinline__ void p2_stage(cpx<float> (&x)[16], cpx<float> *lds, rsrc_t tw,
                                         const int *tw_off, rsrc_t out, long out_vs, float osgn,
                                         float scale) {
    using P = P2<LOG2N>;
    constexpr int N = P::N;
    constexpr int LR = P::bits(S);
    constexpr int R = 1 << LR;
    constexpr int B = 16 / R;                 // butterflies per thread
    constexpr int LNB = LOG2N - LR;           // log2 of butterflies per vector
    constexpr int NB = 1 << LNB;
    constexpr int LNS = P::log2ns(S);
    constexpr int NS = 1 << LNS;
    constexpr bool LAST = (S == P::NST - 1);
    constexpr int T = VPB * N / 16;
    int t = threadIdx.x;
    asm volatile("" : "+v"(t));

    if (S > 0) {
#pragma unroll
        for (int i = 0; i < B; ++i) {
            const int b = t + i * T;
            const int v = b >> LNB, j = b & (NB - 1);
            const cpx<float> *src = lds + v * P::LDSV + padc(j);
#pragma unroll
            for (int r = 0; r < R; ++r) x[i * R + r] = src[padc(r * NB)];
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < B; ++i) {
        const int b = t + i * T;
        const int v = b >> LNB, j = b & (NB - 1);
        const int k = j & (NS - 1);
        if (S > 0) {
            const int toff = tw_off[S] * 8;
#pragma unroll
            for (int r = 1; r < R; ++r)
                x[i * R + r] = cmul(x[i * R + r], bload(tw, k * 8, toff + (r - 1) * NS * 8));
        }
        Dft<float, R>::run(&x[i * R]);
        if (LAST) {
            const int vo = (int)(v * out_vs + j) * 8;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const cpx<float> y = x[i * R + r];
                if (PLAIN) bstore(out, vo, r * NB * 8, y);
                else bstore(out, vo, r * NB * 8, {y.x * scale, y.y * (scale * osgn)});
            }
        } else {
            const int base = ((j >> LNS) << (LNS + LR)) + k;
            cpx<float> *dst = lds + v * P::LDSV + padc(base);
#pragma unroll
            for (int r = 0; r < R; ++r) dst[padc(r * NS)] = x[i * R + r];
        }
    }
    if (!LAST) __syncthreads();
}

// HBM loads of one group of VPB rows in first-stage order
template <int LOG2N, int VPB, bool PLAIN>
__device__ __forceinline__ void p2_load(cpx<float> (&x)[16], rsrc_t in, long in_vs, float isgn) {
    constexpr int N = 1 << LOG2N;
    constexpr int LR = P2<LOG2N>::bits(0);
    constexpr int R = 1 << LR;
    constexpr int B = 16 / R;
    constexpr int LNB = LOG2N - LR;
    constexpr int NB = 1 << LNB;
    constexpr int T = VPB * N / 16;
    int t = threadIdx.x;
    asm volatile("" : "+v"(t));
#pragma unroll
    for (int i = 0; i < B; ++i) {
        const int b = t + i * T;
        const int v = b >> LNB, j = b & (NB - 1);
        const int vo = (int)(v * in_vs + j) * 8;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const cpx<float> y = bload(in, vo, r * NB * 8);   // rows past the group end read 0
            if (PLAIN) x[i * R + r] = y;
            else x[i * R + r] = {y.x, y.y * isgn};
        }
    }
}

// buffer range of the group starting at row g*VPB: its rows, clipped to nvec
template <int LOG2N, int VPB>
__device__ __forceinline__ unsigned group_bytes(long nvec, long g, long vs) {
    long nv = nvec - g * VPB;
    if (nv > VPB) nv = VPB;
    return (unsigned)(((nv - 1) * vs + (1 << LOG2N)) * 8);
}

template <int LOG2N, int VPB, bool PF, bool PLAIN>
__device__ __forceinline__ void pow2_body(const vhip_fftm_desc &d, int per_wg, cpx<float> *lds) {
    using P = P2<LOG2N>;
    const cpx<float> *in = (const cpx<float> *)d.in.re;
    cpx<float> *out = (cpx<float> *)d.out.re;
    const rsrc_t tw = make_rsrc(d.tw, (unsigned)(d.tw_off[P::NST - 1] + 15 * (P::N / 16) + 16) * 8);
    const float isgn = d.conj_in ? -1.f : 1.f;
    const float osgn = d.conj_out ? -1.f : 1.f;
    const float scale = (float)d.scale;
    const long ivs = d.in.vs0, ovs = d.out.vs0;
    const long ngroups = (d.nvec + VPB - 1) / VPB;
    long g = (long)blockIdx.x * per_wg;
    long gend = g + per_wg;
    if (gend > ngroups) gend = ngroups;
    if (g >= gend) return;

    cpx<float> x[16];
    p2_load<LOG2N, VPB, PLAIN>(x, make_rsrc(in + g * VPB * ivs, group_bytes<LOG2N, VPB>(d.nvec, g, ivs)), ivs, isgn);
    for (; g < gend; ++g) {
        cpx<float> nx[PF ? 16 : 1];
        const long gn = g + 1;
        if (PF && gn < gend)
            p2_load<LOG2N, VPB, PLAIN>(*reinterpret_cast<cpx<float>(*)[16]>(nx),
                                make_rsrc(in + gn * VPB * ivs, group_bytes<LOG2N, VPB>(d.nvec, gn, ivs)),
                                ivs, isgn);
        const rsrc_t o = make_rsrc(out + g * VPB * ovs, group_bytes<LOG2N, VPB>(d.nvec, g, ovs));
        p2_stage<LOG2N, VPB, 0, PLAIN>(x, lds, tw, d.tw_off, o, ovs, osgn, scale);
        if constexpr (P::NST > 1) p2_stage<LOG2N, VPB, (P::NST > 1 ? 1 : 0), PLAIN>(x, lds, tw, d.tw_off, o, ovs, osgn, scale);
        if constexpr (P::NST > 2) p2_stage<LOG2N, VPB, (P::NST > 2 ? 2 : 0), PLAIN>(x, lds, tw, d.tw_off, o, ovs, osgn, scale);
        if constexpr (P::NST > 3) p2_stage<LOG2N, VPB, (P::NST > 3 ? 3 : 0), PLAIN>(x, lds, tw, d.tw_off, o, ovs, osgn, scale);
        if (PF) {
#pragma unroll
            for (int e = 0; e < 16; ++e) x[e] = nx[PF ? e : 0];
        } else if (gn < gend) {
            p2_load<LOG2N, VPB, PLAIN>(x, make_rsrc(in + gn * VPB * ivs, group_bytes<LOG2N, VPB>(d.nvec, gn, ivs)),
                                ivs, isgn);
        }
    }
}

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

template <int LOG2N>
constexpr int p2_vpb() { return LOG2N >= 12 ? 1 : (1 << (12 - LOG2N)); }

template <int LOG2N>
bool p2_radices_match(const vhip_fftm_desc *d) {
    using P = P2<LOG2N>;
    if (d->nst != P::NST) return false;
    for (int s = 0; s < P::NST; ++s)
        if (d->radix[s] != (1 << P::bits(s))) return false;
    return true;
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
    const long ngroups = (d->nvec + VPB - 1) / VPB;
    const size_t lds = (size_t)VPB * P2<LOG2N>::LDSV * 8;
    const char *venv = getenv("VSIP_AMD_P2_VARIANT");
    int variant = venv ? atoi(venv) : 1;
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
    const char *env = getenv("VSIP_AMD_P2_PER_WG");
    if (env && atoi(env) > 0) per_wg = atoi(env);
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
    if (d->prec != 4 || d->kind != VHIP_C2C || d->ptw) return -100;
    if (!d->in.contig || !d->out.contig || d->in.split || d->out.split) return -100;
    if (d->map_in != VHIP_MAP_ALONG || d->map_out != VHIP_MAP_ALONG) return -100;
    if (d->in.nv0 < d->nvec || d->out.nv0 < d->nvec) return -100;   // single-level vectors only
    if (getenv("VSIP_AMD_NO_POW2")) return -100;
    switch (d->n) {
    case 256: return p2_launch<8>(d, stream, cus);
    case 512: return p2_launch<9>(d, stream, cus);
    case 1024: return p2_launch<10>(d, stream, cus);
    case 2048: return p2_launch<11>(d, stream, cus);
    case 4096: return p2_launch<12>(d, stream, cus);
    case 8192: return p2_launch<13>(d, stream, cus);
    case 16384: return p2_launch<14>(d, stream, cus);
    }
    return -100;
}

int pow2_set_lds_limit(int bytes) {
    int rc = 0;
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
