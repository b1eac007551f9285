// fftm_static.hpp -- the Stockham transform of fftm_core.hpp with its whole
// schedule fixed at compile time: length N, radices R0..R7 (NST of them),
// vectors per workgroup VPB and lanes NT.  The library compiles one such
// kernel per plan schedule at plan creation (hipRTC, in-process cache), for
// contiguous rows of interleaved complex FP32 data; every butterfly index,
// twiddle index and LDS offset then folds to constants, shifts and
// multiply-highs instead of the runtime divisions and 64-bit address
// arithmetic of the schedule-generic kernel.
//
// Stage s (radix R = R_s, Ns = R_0 ... R_{s-1}): butterfly b of the group
// is vector v = b / (N/R), index j = b mod (N/R); it reads elements
// j + r N/R (from HBM in stage 0, else from LDS), multiplies leg r by
// W_{Ns R}^{r k}, k = j mod Ns (stage twiddles at d.tw + d.tw_off[s]), runs
// the R-point DFT and writes leg r to (j - k) R + k + r Ns (to HBM, scaled,
// in the last stage).  Inverse transforms conjugate on the way in and out.
#pragma once
#include "fftm_core.hpp"

namespace vhipk {
namespace st {

template <int NST_, int R0, int R1, int R2, int R3, int R4, int R5, int R6, int R7>
struct Sched {
    static constexpr int NST = NST_;
    static constexpr int rad(int s) {
        return s == 0 ? R0 : s == 1 ? R1 : s == 2 ? R2 : s == 3 ? R3 : s == 4 ? R4 : s == 5 ? R5 : s == 6 ? R6 : R7;
    }
    static constexpr int ns(int s) { return s == 0 ? 1 : ns(s - 1) * rad(s - 1); }
};

// padded LDS index (one spare slot per 32 complex FP32 values), as the host
// planner sizes it
__device__ __forceinline__ constexpr int lpad(int i) { return i + (i >> 5); }

struct Io {
    rsrc_t in, out, tw;
    long ivs, ovs;
    float isgn, osgn, scale;
};

template <int N, int VPB, int NT, int S, class SC>
__device__ __forceinline__ void sstage(cpx<float> *lds, const vhip_fftm_desc &d, const Io &io) {
    constexpr int R = SC::rad(S);
    constexpr int NS = SC::ns(S);
    constexpr int NBV = N / R;
    constexpr int NBT = VPB * NBV;
    constexpr int B = (NBT + NT - 1) / NT;
    constexpr bool FIRST = (S == 0), LAST = (S == SC::NST - 1);
    constexpr int LS = (lpad(N) % 2 == 0) ? lpad(N) + 1 : lpad(N);
    int t = threadIdx.x;
    asm volatile("" : "+v"(t));
    cpx<float> x[B][R];
#pragma unroll
    for (int i = 0; i < B; ++i) {
        const int b = t + i * NT;
        if (B * NT <= NBT || b < NBT) {
            const int v = b / NBV, j = b - v * NBV;
            if (FIRST) {
                const int vo = (int)(v * io.ivs + j) * 8;
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    cpx<float> y = Buf<float>::ldc(io.in, vo, r * NBV * 8);
                    y.y *= io.isgn;
                    x[i][r] = y;
                }
            } else {
                const cpx<float> *src = lds + v * LS;
#pragma unroll
                for (int r = 0; r < R; ++r) x[i][r] = src[lpad(j + r * NBV)];
            }
        }
    }
    if (!FIRST) __syncthreads();
#pragma unroll
    for (int i = 0; i < B; ++i) {
        const int b = t + i * NT;
        if (B * NT <= NBT || b < NBT) {
            const int v = b / NBV, j = b - v * NBV;
            const int k = (NS > 1) ? j % NS : 0;
            if (NS > 1) {
                const int toff = d.tw_off[S];
#pragma unroll
                for (int r = 1; r < R; ++r)
                    x[i][r] = cmul(x[i][r], Buf<float>::ldc(io.tw, (toff + k + (r - 1) * NS) * 8, 0));
            }
            Dft<float, R>::run(x[i]);
            if (LAST) {
                const int vo = (int)(v * io.ovs + j) * 8;
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const cpx<float> y = x[i][r];
                    Buf<float>::stc(io.out, vo, r * NBV * 8, {y.x * io.scale, y.y * (io.scale * io.osgn)});
                }
            } else {
                const int base = (j - k) * R + k;
                cpx<float> *dst = lds + v * LS;
#pragma unroll
                for (int r = 0; r < R; ++r) dst[lpad(base + r * NS)] = x[i][r];
            }
        }
    }
    if (!LAST) __syncthreads();
}

template <int N, int VPB, int NT, int S, class SC>
__device__ __forceinline__ void sstages(cpx<float> *lds, const vhip_fftm_desc &d, const Io &io) {
    sstage<N, VPB, NT, S, SC>(lds, d, io);
    if constexpr (S + 1 < SC::NST) sstages<N, VPB, NT, S + 1, SC>(lds, d, io);
}

// One workgroup per group of VPB rows (contiguous interleaved complex FP32,
// single-level vector stride d.in.vs0 / d.out.vs0 in complex elements);
// rows past the end of the view read zeros and their stores fall outside the
// group's buffer range.
template <int N, int VPB, int NT, int NST, int R0, int R1, int R2, int R3, int R4, int R5, int R6, int R7>
__global__ void __launch_bounds__(NT) fftm_static_kernel(const vhip_fftm_desc d) {
    extern __shared__ __attribute__((aligned(16))) char smem[];
    using SC = Sched<NST, R0, R1, R2, R3, R4, R5, R6, R7>;
    const long g = blockIdx.x;
    long nvl = d.nvec - g * VPB;
    const long nv = nvl < VPB ? nvl : VPB;
    Io io;
    io.ivs = d.in.vs0;
    io.ovs = d.out.vs0;
    io.in = make_rsrc((const cpx<float> *)d.in.re + g * VPB * io.ivs, ((nv - 1) * io.ivs + N) * 8);
    io.out = make_rsrc((cpx<float> *)d.out.re + g * VPB * io.ovs, ((nv - 1) * io.ovs + N) * 8);
    io.tw = make_rsrc(d.tw, (long)N * 8);
    io.isgn = d.conj_in ? -1.f : 1.f;
    io.osgn = d.conj_out ? -1.f : 1.f;
    io.scale = (float)d.scale;
    sstages<N, VPB, NT, 0, SC>((cpx<float> *)smem, d, io);
}

}  // namespace st
}  // namespace vhipk
