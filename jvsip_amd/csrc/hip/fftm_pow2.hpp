// fftm_pow2.hpp -- specialised multiple-FFT kernel for power-of-two row
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
// For large batches the pipelined forms further down (pow2_pipe_body,
// pow2_pipe_gen_body, pow2_pipe_big_body, pow2_r2c_pipe_body) keep the next group's loads in
// flight through the whole transform of the current one, with no other
// vector-memory access in between.
#pragma once
#include <cstdlib>
#include "fftm_core.hpp"

namespace vhipk {
namespace p2 {

typedef __amdgpu_buffer_rsrc_t rsrc_t;
typedef decltype(__builtin_amdgcn_raw_buffer_load_b64(rsrc_t(), 0, 0, 0)) raw64_t;

// word 3 of a raw buffer descriptor on gfx950 (32-bit data format, no swizzle)
constexpr int kBufWord3 = 0x00020000;

__device__ __forceinline__ rsrc_t make_rsrc(const void *p, unsigned bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)bytes, kBufWord3);
}
// Cache policy of the data's HBM loads and stores (the buffer
// instructions' aux field).  NT = true marks them non-temporal (loads 2 =
// nt, stores 18 = nt | sc1).  Measured on MI355X: the c2c kernel streams
// both (4096 x 4096 in place: 4.80 -> 5.24 TB/s); C2R streams its stores
// (8192 x 2048: 5.04 -> 5.53 TB/s) but not its loads, which read each
// element twice; R2C keeps the default for both (streamed loads 4.57 ->
// 3.95 TB/s; its n + 1 outputs per row share lines between rows); the
// batched items keep the default.
__device__ __forceinline__ cpx<float> bload(rsrc_t r, int voff, int soff) {
    return __builtin_bit_cast(cpx<float>, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}
template <bool NT = false>
__device__ __forceinline__ cpx<float> dload(rsrc_t r, int voff, int soff) {
    return __builtin_bit_cast(cpx<float>, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, NT ? 2 : 0));
}
template <bool NT = false>
__device__ __forceinline__ void bstore(rsrc_t r, int voff, int soff, cpx<float> x) {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(raw64_t, x), r, voff, soff, NT ? 18 : 0);
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
    // the same for a padding of one element per 2^PS
    template <int PS>
    static constexpr int ldsv() { return (N + (N >> PS)) % 2 == 0 ? N + (N >> PS) + 1 : N + (N >> PS); }
};

__host__ __device__ constexpr int padc(int i) { return i + (i >> 5); }
// padding of one element per 2^PS (padc is PS = 5; the pipelined kernels use
// 4, which leaves stage 0's stride-16 writes free of bank conflicts)
template <int PS>
__host__ __device__ constexpr int padx(int i) { return i + (i >> PS); }

// Stage S: 16 elements per thread in x (already loaded for S == 0), twiddle,
// DFT, write to LDS (or HBM for the last stage).
template <int LOG2N, int VPB, int S, bool PLAIN, bool LAST_TO_LDS = false, bool NT = false, int PS = 5>
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
    constexpr bool LAST = (S == P::NST - 1) && !LAST_TO_LDS;   // the last stage stores to HBM
    constexpr int T = VPB * N / 16;
    int t = threadIdx.x;
    asm volatile("" : "+v"(t));

    if (S > 0) {
#pragma unroll
        for (int i = 0; i < B; ++i) {
            const int b = t + i * T;
            const int v = b >> LNB, j = b & (NB - 1);
            const cpx<float> *src = lds + v * P::template ldsv<PS>() + padx<PS>(j);
#pragma unroll
            for (int r = 0; r < R; ++r) x[i * R + r] = src[padx<PS>(r * NB)];
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
                if (PLAIN) bstore<NT>(out, vo, r * NB * 8, y);
                else bstore<NT>(out, vo, r * NB * 8, {y.x * scale, y.y * (scale * osgn)});
            }
        } else {
            const int base = ((j >> LNS) << (LNS + LR)) + k;
            cpx<float> *dst = lds + v * P::template ldsv<PS>() + padx<PS>(base);
#pragma unroll
            for (int r = 0; r < R; ++r) dst[padx<PS>(r * NS)] = x[i * R + r];
        }
    }
    if (!LAST) __syncthreads();
}

// rows per workgroup of the power-of-two kernels: 4096 elements per group
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

// HBM loads of one group of VPB rows in first-stage order
template <int LOG2N, int VPB, bool PLAIN, bool NT = false>
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
            const cpx<float> y = dload<NT>(in, vo, r * NB * 8);   // rows past the group end read 0
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
    p2_load<LOG2N, VPB, PLAIN, true>(x, make_rsrc(in + g * VPB * ivs, group_bytes<LOG2N, VPB>(d.nvec, g, ivs)), ivs, isgn);
    for (; g < gend; ++g) {
        cpx<float> nx[PF ? 16 : 1];
        const long gn = g + 1;
        if (PF && gn < gend)
            p2_load<LOG2N, VPB, PLAIN, true>(*reinterpret_cast<cpx<float>(*)[16]>(nx),
                                make_rsrc(in + gn * VPB * ivs, group_bytes<LOG2N, VPB>(d.nvec, gn, ivs)),
                                ivs, isgn);
        const rsrc_t o = make_rsrc(out + g * VPB * ovs, group_bytes<LOG2N, VPB>(d.nvec, g, ovs));
        p2_stage<LOG2N, VPB, 0, PLAIN, false, true>(x, lds, tw, d.tw_off, o, ovs, osgn, scale);
        if constexpr (P::NST > 1) p2_stage<LOG2N, VPB, (P::NST > 1 ? 1 : 0), PLAIN, false, true>(x, lds, tw, d.tw_off, o, ovs, osgn, scale);
        if constexpr (P::NST > 2) p2_stage<LOG2N, VPB, (P::NST > 2 ? 2 : 0), PLAIN, false, true>(x, lds, tw, d.tw_off, o, ovs, osgn, scale);
        if constexpr (P::NST > 3) p2_stage<LOG2N, VPB, (P::NST > 3 ? 3 : 0), PLAIN, false, true>(x, lds, tw, d.tw_off, o, ovs, osgn, scale);
        if (PF) {
#pragma unroll
            for (int e = 0; e < 16; ++e) x[e] = nx[PF ? e : 0];
        } else if (gn < gend) {
            p2_load<LOG2N, VPB, PLAIN, true>(x, make_rsrc(in + gn * VPB * ivs, group_bytes<LOG2N, VPB>(d.nvec, gn, ivs)),
                                ivs, isgn);
        }
    }
}

// ---------------------------------------------------------------------------
// Pipelined persistent form of the 4096-point row kernel (radices 16/16/16,
// 256 lanes, one row per pass).  Workgroup b of a grid of G transforms rows
// b, b + G, b + 2G, ...; while it transforms one row the 16 loads of the
// next are in flight, so HBM stays busy through the LDS stages.  Two things
// keep the loads of the next row from stalling the current one:
//   - no vector-memory access inside a row besides the row's own loads and
//     stores: the stage twiddles of lane t, W_256^{r (t mod 16)} and
//     W_4096^{rt} (r = 1..15), are loaded once into 60 registers (the 2
//     workgroups per CU leave 256 per lane), ahead of the first row;
//   - the prefetch is unconditional (rows past the end use a zero-length
//     buffer range and read 0), so the wait before stage 0 always counts
//     exactly the loads and stores issued after the row's own loads.
// Measured on MI355X, 4096 x 4096 in place: 48.6 us vs 51.5-52.5 us for one
// row per workgroup (2 workgroups per CU, 8 rows each); issuing the second
// row's loads before the first row is transformed gained 1.4%.

// The pipelined kernels pad their LDS rows by one element per 16
// (padx<kPipePS>, i + (i >> 4)) rather than per 32: stage 0 writes each
// lane's 16 outputs at a stride of 16 elements, which the 32-element padding
// maps to a 132-byte lane stride and two-way bank conflicts; a 136-byte
// stride has none (0.4% faster on 4096 x 4096).  Every index split below
// (base plus leg offset) adds a multiple of 16 to a value below 16, so the
// padding still splits.
constexpr int kPipePS = 4;
constexpr int kPipeLdsv = P2<12>::ldsv<kPipePS>();   // > padx<kPipePS>(4095), odd

// Stage 0 of the 4096-point row from registers: DFT, lane t's outputs to
// elements 16 t + r.  With a double buffer (DB) stage 0 writes the first LDS
// row and stage 1 the second, so no stage waits for the other lanes to finish
// reading before it writes: the row's two barriers (after each write) also
// order the next row's writes after this row's reads.
__device__ __forceinline__ void p2p_stage0(cpx<float> (&x)[16], cpx<float> *lds) {
    int t = threadIdx.x;
    asm volatile("" : "+v"(t));
    Dft<float, 16>::run(x);
    cpx<float> *dst = lds + padx<kPipePS>(16 * t);
#pragma unroll
    for (int r = 0; r < 16; ++r) dst[r] = x[r];
    __syncthreads();
}

// stage 1 or 2 of the 4096-point row with the twiddles w[r-1] of this lane,
// reading LDS row `in` and (stage 1) writing `out_lds`
template <int S, bool PLAIN, bool DB = false>
__device__ __forceinline__ void p2p_stage(cpx<float> (&x)[16], const cpx<float> *in, cpx<float> *out_lds,
                                          const cpx<float> (&w)[15], rsrc_t out, float osgn, float scale) {
    constexpr int R = 16, NB = 256;
    constexpr int LNS = 4 * S, NS = 1 << LNS;
    constexpr bool LAST = (S == 2);
    int t = threadIdx.x;
    asm volatile("" : "+v"(t));
    const int j = t;
    {
        const cpx<float> *src = in + padx<kPipePS>(j);
#pragma unroll
        for (int r = 0; r < R; ++r) x[r] = src[padx<kPipePS>(r * NB)];
        if (!DB) __syncthreads();
    }
#pragma unroll
    for (int r = 1; r < R; ++r) x[r] = cmul(x[r], w[r - 1]);
    Dft<float, R>::run(x);
    if (LAST) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const cpx<float> y = x[r];
            if (PLAIN) bstore<true>(out, j * 8, r * NB * 8, y);
            else bstore<true>(out, j * 8, r * NB * 8, {y.x * scale, y.y * (scale * osgn)});
        }
    } else {
        const int base = ((j >> LNS) << (LNS + 4)) + (j & (NS - 1));
        cpx<float> *dst = out_lds + padx<kPipePS>(base);
#pragma unroll
        for (int r = 0; r < R; ++r) dst[padx<kPipePS>(r * NS)] = x[r];
        __syncthreads();
    }
}

// one 4096-point row held in x (stage 0 from registers, stages 1-2 as above)
template <bool PLAIN, bool DB>
__device__ __forceinline__ void p2p_row(cpx<float> (&x)[16], cpx<float> *lds, const cpx<float> (&w1)[15],
                                        const cpx<float> (&w2)[15], rsrc_t out, float isgn, float osgn,
                                        float scale) {
    cpx<float> *lds2 = DB ? lds + kPipeLdsv : lds;
    // the input conjugation is applied here rather than at the load, which
    // would make the prefetch of the next row wait for its data
    if (!PLAIN) {
#pragma unroll
        for (int e = 0; e < 16; ++e) x[e].y *= isgn;
    }
    p2p_stage0(x, lds);
    p2p_stage<1, PLAIN, DB>(x, lds, lds2, w1, out, osgn, scale);
    p2p_stage<2, PLAIN, DB>(x, lds2, lds2, w2, out, osgn, scale);
}

// lds: kPipeLdsv complex elements (twice that with the double buffer DB)
template <bool PLAIN, bool DB = false>
__device__ __forceinline__ void pow2_pipe_body(const vhip_fftm_desc &d, cpx<float> *lds) {
    constexpr int N = 4096;
    const cpx<float> *tw = (const cpx<float> *)d.tw;
    const cpx<float> *in = (const cpx<float> *)d.in.re;
    cpx<float> *out = (cpx<float> *)d.out.re;
    const int t = threadIdx.x;
    const long ivs = d.in.vs0, ovs = d.out.vs0;
    const float isgn = d.conj_in ? -1.f : 1.f;
    const float osgn = d.conj_out ? -1.f : 1.f;
    const float scale = (float)d.scale;
    const int G = gridDim.x;
    const int nvec = __builtin_amdgcn_readfirstlane((int)d.nvec);
    int g = blockIdx.x;
    if (g >= nvec) return;
    auto rin = [&](int row) {
        return make_rsrc(in + (long)(row < nvec ? row : 0) * ivs, row < nvec ? (unsigned)(N * 8) : 0u);
    };
    // twiddles first: loads return in order, and the row waits for both
    cpx<float> w1[15], w2[15];
#pragma unroll
    for (int r = 1; r < 16; ++r) {
        w1[r - 1] = tw[d.tw_off[1] + (r - 1) * 16 + (t & 15)];
        w2[r - 1] = tw[d.tw_off[2] + (r - 1) * 256 + t];
    }
    cpx<float> a[16], b[16];
    // the first row is transformed with the second already in flight; from
    // then on each row's prefetch is issued right after the previous row's
    // stores, so every wait before stage 0 counts the same 32 later
    // accesses, on every path into it
    p2_load<12, 1, true, true>(a, rin(g), ivs, 1.f);
    p2_load<12, 1, true, true>(b, rin(g + G), ivs, 1.f);
    p2p_row<PLAIN, DB>(a, lds, w1, w2, make_rsrc(out + (long)g * ovs, (unsigned)(N * 8)), isgn, osgn, scale);
    g += G;
    if (g >= nvec) return;
    p2_load<12, 1, true, true>(a, rin(g + G), ivs, 1.f);
    for (;;) {
        p2p_row<PLAIN, DB>(b, lds, w1, w2, make_rsrc(out + (long)g * ovs, (unsigned)(N * 8)), isgn, osgn, scale);
        g += G;
        if (g >= nvec) break;
        p2_load<12, 1, true, true>(b, rin(g + G), ivs, 1.f);
        p2p_row<PLAIN, DB>(a, lds, w1, w2, make_rsrc(out + (long)g * ovs, (unsigned)(N * 8)), isgn, osgn, scale);
        g += G;
        if (g >= nvec) break;
        p2_load<12, 1, true, true>(a, rin(g + G), ivs, 1.f);
    }
}

// ---- 4096-point rows with cross-lane (DPP) exchanges ----------------------
//
// 4096 = 64 x 64.  Each 64-point transform is held by a quad of 4 adjacent
// lanes, 16 elements per lane: a 16-point DFT in registers, the twiddles
// W_64^{q k1} of the lane's quad position q, then a 4-point DFT across the
// quad as two radix-2 butterflies whose partner values come from DPP
// quad_perm moves (lane q ^ 2, then q ^ 1) -- no LDS.  Lane q ends with the
// outputs k1 + 16 brev(q) (brev = 0, 2, 1, 3).  A row takes two such
// super-stages:
//   1. lane (g, q) = (L >> 2, L & 3) holds x[g + 64 (q + 4 m)], m = 0..15
//      (each load instruction reads four whole 128-byte lines); after the
//      quad DFT it holds S[g][k], k = k1 + 16 brev(q), times W_4096^{g k};
//   2. after one LDS transpose, lane (h, p) holds S[p + 4 m][h] and the same
//      steps give X[h + 64 (j1 + 16 brev(p))] (whole 128-byte lines again).
// One LDS round trip per row instead of two, and one barrier: the rows
// alternate between two LDS buffers, so a row's barrier also orders the next
// row's writes after the reads of the row before.  The per-lane twiddles are
// computed once in double precision (sincospi) and rounded to float.
// LDS layout of the transpose: element (k, g) at 68 k + g + 8 (k >> 4), which
// spreads both the write (g contiguous, four k blocks) and the read (h = k
// across lanes) over all banks.
constexpr int kDppLds = 68 * 63 + 63 + 8 * 3 + 1;   // elements per buffer

__device__ __forceinline__ constexpr int dpp_a(int k, int g) { return 68 * k + g + 8 * (k >> 4); }

template <int CTRL>
__device__ __forceinline__ float dpp_f(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, 0xF, 0xF, false));
}

// forward 4-point DFT over the 4 lanes of each quad, for 16 registers; lane
// q = lane & 3 keeps output brev(q)
__device__ __forceinline__ void quad_dft4(cpx<float> (&x)[16], int q) {
    const float sa = (q < 2) ? 1.f : -1.f;      // lanes 0,1: y + y[q^2]; lanes 2,3: y[q^2] - y
    const float sb = (q & 1) ? -1.f : 1.f;      // even: t + t[q^1]; odd: t[q^1] - t
    const bool rot = (q == 3);                  // times -i (W_4^1) before the second butterfly
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const float px = dpp_f<0x4E>(x[r].x), py = dpp_f<0x4E>(x[r].y);
        float tx = px + sa * x[r].x, ty = py + sa * x[r].y;
        const float ux = rot ? ty : tx, uy = rot ? -tx : ty;
        const float qx = dpp_f<0xB1>(ux), qy = dpp_f<0xB1>(uy);
        x[r] = {qx + sb * ux, qy + sb * uy};
    }
}

template <bool PLAIN>
__device__ __forceinline__ void p2d_row(cpx<float> (&x)[16], cpx<float> *buf, const cpx<float> (&wa)[16],
                                        const cpx<float> (&wb)[16], rsrc_t out, int g, int q, int bq, float isgn,
                                        float osgn, float scale) {
    if (!PLAIN) {
#pragma unroll
        for (int e = 0; e < 16; ++e) x[e].y *= isgn;
    }
    // super-stage 1: DFT over n_hi = q + 4 m
    Dft<float, 16>::run(x);
#pragma unroll
    for (int k1 = 1; k1 < 16; ++k1) x[k1] = cmul(x[k1], wa[k1]);
    quad_dft4(x, q);
    {
        cpx<float> *dst = buf + dpp_a(16 * bq, g);
#pragma unroll
        for (int k1 = 0; k1 < 16; ++k1) dst[68 * k1] = cmul(x[k1], wb[k1]);
    }
    __syncthreads();
    // super-stage 2: lane (h, p) = (g, q) reads S[p + 4 m][h]
    {
        const cpx<float> *src = buf + dpp_a(g, q);
#pragma unroll
        for (int m = 0; m < 16; ++m) x[m] = src[4 * m];
    }
    Dft<float, 16>::run(x);
#pragma unroll
    for (int j1 = 1; j1 < 16; ++j1) x[j1] = cmul(x[j1], wa[j1]);
    quad_dft4(x, q);
    const int vo = (g + 1024 * bq) * 8;
#pragma unroll
    for (int j1 = 0; j1 < 16; ++j1) {
        const cpx<float> y = x[j1];
        if (PLAIN) bstore<true>(out, vo, j1 * 512, y);
        else bstore<true>(out, vo, j1 * 512, {y.x * scale, y.y * (scale * osgn)});
    }
}

// loads of one row in super-stage 1 order: lane (g, q), register m reads
// element g + 64 q + 256 m
__device__ __forceinline__ void p2d_load(cpx<float> (&x)[16], rsrc_t in, int g, int q) {
    const int vo = (g + 64 * q) * 8;
#pragma unroll
    for (int m = 0; m < 16; ++m) x[m] = dload<true>(in, vo, m * 2048);
}

// lds: 2 * kDppLds complex elements; 256 lanes, rows b, b + G, ... as the
// pipelined kernel
template <bool PLAIN>
__device__ __forceinline__ void pow2_dpp_body(const vhip_fftm_desc &d, cpx<float> *lds) {
    constexpr int N = 4096;
    const cpx<float> *in = (const cpx<float> *)d.in.re;
    cpx<float> *out = (cpx<float> *)d.out.re;
    const int t = threadIdx.x, g = t >> 2, q = t & 3;
    const int bq = (q == 1) ? 2 : (q == 2) ? 1 : q;
    const long ivs = d.in.vs0, ovs = d.out.vs0;
    const float isgn = d.conj_in ? -1.f : 1.f;
    const float osgn = d.conj_out ? -1.f : 1.f;
    const float scale = (float)d.scale;
    const int G = gridDim.x;
    const int nvec = __builtin_amdgcn_readfirstlane((int)d.nvec);
    int row = blockIdx.x;
    if (row >= nvec) return;
    auto rin = [&](int r) {
        return make_rsrc(in + (long)(r < nvec ? r : 0) * ivs, r < nvec ? (unsigned)(N * 8) : 0u);
    };
    cpx<float> a[16], b[16];
    p2d_load(a, rin(row), g, q);
    p2d_load(b, rin(row + G), g, q);
    // per-lane twiddles in double, rounded once: W_64^{q k1} and
    // W_4096^{g (k1 + 16 brev(q))}
    cpx<float> wa[16], wb[16];
#pragma unroll
    for (int k1 = 0; k1 < 16; ++k1) {
        double sn, cs;
        sincospi(-2.0 * (double)((q * k1) & 63) / 64.0, &sn, &cs);
        wa[k1] = {(float)cs, (float)sn};
        sincospi(-2.0 * (double)((g * (k1 + 16 * bq)) & 4095) / 4096.0, &sn, &cs);
        wb[k1] = {(float)cs, (float)sn};
    }
    int parity = 0;
    p2d_row<PLAIN>(a, lds, wa, wb, make_rsrc(out + (long)row * ovs, (unsigned)(N * 8)), g, q, bq, isgn, osgn, scale);
    row += G;
    if (row >= nvec) return;
    p2d_load(a, rin(row + G), g, q);
    for (;;) {
        parity ^= 1;
        p2d_row<PLAIN>(b, lds + parity * kDppLds, wa, wb, make_rsrc(out + (long)row * ovs, (unsigned)(N * 8)), g, q,
                       bq, isgn, osgn, scale);
        row += G;
        if (row >= nvec) break;
        p2d_load(b, rin(row + G), g, q);
        parity ^= 1;
        p2d_row<PLAIN>(a, lds + parity * kDppLds, wa, wb, make_rsrc(out + (long)row * ovs, (unsigned)(N * 8)), g, q,
                       bq, isgn, osgn, scale);
        row += G;
        if (row >= nvec) break;
        p2d_load(a, rin(row + G), g, q);
    }
}

// buffer range of `len` complex elements per row for the group at row g*VPB
template <int VPB>
__device__ __forceinline__ unsigned group_bytes_len(long nvec, long g, long vs, long len) {
    long nv = nvec - g * VPB;
    if (nv > VPB) nv = VPB;
    return (unsigned)(((nv - 1) * vs + len) * 8);
}

// Real transforms of contiguous FP32 rows whose half length n = 2^LOG2N runs
// on the power-of-two stages (behaviour of VI_rcfftmop_f_def.h:53-115 and
// VI_crfftmop_f_def.h:74-124, as in the generic kernel's r2c_post/c2r_pre):
//   R2C: the real row read as n complex values x[2m] + i x[2m+1] (its own
//        interleaved layout), the n-point stages with the last one writing
//        LDS, then z_k = scale/2 [(Y_k + conj Y_{n-k}) - i e_k (Y_k - conj Y_{n-k})],
//        k = 0..n, e_k = e^{-i pi k/n}, stored to the complex output row;
//   C2R: conj(Y_k), Y_k = (Z_k + conj Z_{n-k}) + i e_k (Z_k - conj Z_{n-k}),
//        e_k = e^{+i pi k/n}, formed in registers from the complex input row,
//        the n-point stages, and scale * conj(result) stored as the real
//        row's interleaved pairs.
// Strides d.in.vs0 / d.out.vs0 of a real row are counted in reals (even, so
// rows stay 8-byte aligned); they are halved to complex strides here.
template <int LOG2N, int VPB, int KIND>
__device__ __forceinline__ void pow2_real_body(const vhip_fftm_desc &d, cpx<float> *lds) {
    using P = P2<LOG2N>;
    constexpr int N = P::N;
    constexpr int T = VPB * N / 16;
    const rsrc_t tw = make_rsrc(d.tw, (unsigned)(d.tw_off[P::NST - 1] + 15 * (N / 16) + 16) * 8);
    const cpx<float> *eu = (const cpx<float> *)d.euler;
    const float scale = (float)d.scale;
    const long ivs = KIND == VHIP_R2C ? d.in.vs0 / 2 : d.in.vs0;
    const long ovs = KIND == VHIP_C2R ? d.out.vs0 / 2 : d.out.vs0;
    const long g = blockIdx.x;
    const cpx<float> *in = (const cpx<float> *)d.in.re + g * VPB * ivs;
    cpx<float> *out = (cpx<float> *)d.out.re + g * VPB * ovs;
    long nvl = d.nvec - g * VPB;
    const int nv = nvl > VPB ? VPB : (int)nvl;
    cpx<float> x[16];
    if (KIND == VHIP_R2C) {
        p2_load<LOG2N, VPB, true>(x, make_rsrc(in, group_bytes<LOG2N, VPB>(d.nvec, g, ivs)), ivs, 1.f);
    } else {
        constexpr int LR = P::bits(0);
        constexpr int R = 1 << LR;
        constexpr int B = 16 / R;
        constexpr int LNB = LOG2N - LR;
        constexpr int NB = 1 << LNB;
        const rsrc_t ri = make_rsrc(in, group_bytes_len<VPB>(d.nvec, g, ivs, N + 1));
        int t = threadIdx.x;
        asm volatile("" : "+v"(t));
#pragma unroll
        for (int i = 0; i < B; ++i) {
            const int b = t + i * T;
            const int v = b >> LNB, j = b & (NB - 1);
            const int vo = (int)(v * ivs) * 8;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int e = j + r * NB;
                const cpx<float> F = bload(ri, vo + e * 8, 0);
                const cpx<float> B2 = bload(ri, vo + (N - e) * 8, 0);
                const cpx<float> Bc = {B2.x, -B2.y};
                const cpx<float> A = F + Bc;
                const cpx<float> WD = cmul(eu[e], F - Bc);
                x[i * R + r] = {A.x - WD.y, -(A.y + WD.x)};       // conj(A + i WD)
            }
        }
    }
    const rsrc_t o = make_rsrc(out, group_bytes<LOG2N, VPB>(d.nvec, g, ovs));
    constexpr bool TOL = (KIND == VHIP_R2C);
    const float osgn = KIND == VHIP_C2R ? -1.f : 1.f;
    p2_stage<LOG2N, VPB, 0, false, TOL, KIND == VHIP_C2R, kPipePS>(x, lds, tw, d.tw_off, o, ovs, osgn, scale);
    if constexpr (P::NST > 1) p2_stage<LOG2N, VPB, (P::NST > 1 ? 1 : 0), false, TOL, KIND == VHIP_C2R, kPipePS>(x, lds, tw, d.tw_off, o, ovs, osgn, scale);
    if constexpr (P::NST > 2) p2_stage<LOG2N, VPB, (P::NST > 2 ? 2 : 0), false, TOL, KIND == VHIP_C2R, kPipePS>(x, lds, tw, d.tw_off, o, ovs, osgn, scale);
    if constexpr (P::NST > 3) p2_stage<LOG2N, VPB, (P::NST > 3 ? 3 : 0), false, TOL, KIND == VHIP_C2R, kPipePS>(x, lds, tw, d.tw_off, o, ovs, osgn, scale);
    if (KIND == VHIP_R2C) {
        // The n + 1 outputs of each row from Y in LDS (natural order), in
        // mirror pairs: item m < n/2 of row v reads Y_m and Y_{n-m} once and
        // stores z_m and z_{n-m} (m = 0 gives z_0 and z_n, both from Y_0,
        // and also z_{n/2}).  VPB n/2 items per group = 8 per lane, so the
        // item index needs no division.  Rows past nv are skipped (with a
        // single row the launch accepts any row stride, so their offsets may
        // fall inside the store range).
        const float half = 0.5f * scale;
        const rsrc_t ro = make_rsrc(out, group_bytes_len<VPB>(d.nvec, g, ovs, N + 1));
        auto z = [&](const cpx<float> F, const cpx<float> B2, int k) {
            const cpx<float> Bc = {B2.x, -B2.y};
            const cpx<float> A = F + Bc;
            const cpx<float> WD = cmul(eu[k], F - Bc);
            return cpx<float>{half * (A.x + WD.y), half * (A.y - WD.x)};
        };
        static_assert((VPB * (N / 2)) % T == 0, "mirror items must divide over the lanes");
#pragma unroll
        for (int i = 0; i < VPB * (N / 2) / T; ++i) {
            const int q = threadIdx.x + i * T;
            const int v = q / (N / 2), m = q & (N / 2 - 1);
            if (v >= nv) continue;
            const cpx<float> *y = lds + v * P::template ldsv<kPipePS>();
            const cpx<float> F = y[padx<kPipePS>(m)];
            const cpx<float> B2 = y[padx<kPipePS>(m == 0 ? 0 : N - m)];
            const int vo = (int)(v * ovs) * 8;
            bstore(ro, vo + m * 8, 0, z(F, B2, m));
            bstore(ro, vo + (N - m) * 8, 0, z(B2, F, N - m));
            if (m == 0) {
                const cpx<float> H = y[padx<kPipePS>(N / 2)];
                bstore(ro, vo + (N / 2) * 8, 0, z(H, H, N / 2));
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Pipelined persistent form of the R2C real-row kernel (the same arithmetic
// as pow2_real_body<LOG2N, VPB, VHIP_R2C>, VPB n = 4096, 256 lanes):
// workgroup b of a grid of G transforms row groups b, b + G, ... with the
// loads of its next group in flight, as the 4096-point pipelined kernel does.
// What keeps vector memory out of the group's own work:
//   - every stage's twiddles are the same for all butterflies of a lane
//     (k = t mod NS, as 256 is a multiple of every NS here), loaded once into
//     registers; so are the Euler factors of the lane's mirror items;
//   - every store of the combine is issued by every lane; a lane with
//     nothing to write there stores at an offset past the buffer range,
//     which the hardware drops, so no store is conditional and the wait
//     before the next group's stage 0 counts a fixed number of accesses.

// buffer offset no range here reaches (ranges stay below 2 GiB)
constexpr int kDropOffset = (int)0x80000000u;

// stage S >= 1 with the lane's twiddles w[r-1] (the same for its B butterflies)
template <int LOG2N, int VPB, int S, int PS = 5>
__device__ __forceinline__ void p2r_stage(cpx<float> (&x)[16], cpx<float> *lds, const cpx<float> *w) {
    using P = P2<LOG2N>;
    constexpr int N = P::N;
    constexpr int LR = P::bits(S);
    constexpr int R = 1 << LR;
    constexpr int B = 16 / R;
    constexpr int LNB = LOG2N - LR;
    constexpr int NB = 1 << LNB;
    constexpr int LNS = P::log2ns(S);
    constexpr int NS = 1 << LNS;
    constexpr int T = VPB * N / 16;
    static_assert(T % NS == 0, "a lane's twiddles must not depend on its butterfly");
    int t = threadIdx.x;
    asm volatile("" : "+v"(t));
#pragma unroll
    for (int i = 0; i < B; ++i) {
        const int b = t + i * T;
        const int v = b >> LNB, j = b & (NB - 1);
        const cpx<float> *src = lds + v * P::template ldsv<PS>() + padx<PS>(j);
#pragma unroll
        for (int r = 0; r < R; ++r) x[i * R + r] = src[padx<PS>(r * NB)];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < B; ++i) {
        const int b = t + i * T;
        const int v = b >> LNB, j = b & (NB - 1);
        const int k = j & (NS - 1);
#pragma unroll
        for (int r = 1; r < R; ++r) x[i * R + r] = cmul(x[i * R + r], w[r - 1]);
        Dft<float, R>::run(&x[i * R]);
        const int base = ((j >> LNS) << (LNS + LR)) + k;
        cpx<float> *dst = lds + v * P::template ldsv<PS>() + padx<PS>(base);
#pragma unroll
        for (int r = 0; r < R; ++r) dst[padx<PS>(r * NS)] = x[i * R + r];
    }
    __syncthreads();
}

template <int LOG2N>
struct P2R {
    static constexpr int N = 1 << LOG2N;
    static constexpr int VPB = p2_vpb<LOG2N>();
    static constexpr int T = VPB * N / 16;
    static constexpr int NST = P2<LOG2N>::NST;
    static constexpr int NIT = VPB * (N / 2) / T;                      // mirror items per lane
    static constexpr int NE = (N / 2) > T ? (N / 2) / T : 1;            // distinct m per lane
    static_assert(NST <= 3 && NIT == 8 && NE <= NIT, "shape of the pipelined R2C kernel");
};

// one group: the row transforms from x (stage 0 from registers), the
// combine from LDS, the stores to the group's output range `ro`
template <int LOG2N>
__device__ __forceinline__ void p2r_group(cpx<float> (&x)[16], cpx<float> *lds, const cpx<float> (&w)[2][15],
                                          const cpx<float> (&em)[P2R<LOG2N>::NE], const cpx<float> (&en)[P2R<LOG2N>::NE],
                                          cpx<float> eh, rsrc_t ro, long ovs, int nv, float half) {
    using P = P2<LOG2N>;
    using Q = P2R<LOG2N>;
    constexpr int N = Q::N, VPB = Q::VPB, T = Q::T;
    p2_stage<LOG2N, VPB, 0, true, true, false, kPipePS>(x, lds, make_rsrc(nullptr, 0u), nullptr, ro, 0, 1.f, 1.f);
    if constexpr (Q::NST > 1) p2r_stage<LOG2N, VPB, (Q::NST > 1 ? 1 : 0), kPipePS>(x, lds, w[0]);
    if constexpr (Q::NST > 2) p2r_stage<LOG2N, VPB, (Q::NST > 2 ? 2 : 0), kPipePS>(x, lds, w[1]);
    auto z = [&](const cpx<float> F, const cpx<float> B2, cpx<float> e) {
        const cpx<float> Bc = {B2.x, -B2.y};
        const cpx<float> A = F + Bc;
        const cpx<float> WD = cmul(e, F - Bc);
        return cpx<float>{half * (A.x + WD.y), half * (A.y - WD.x)};
    };
    const int t = threadIdx.x;
#pragma unroll
    for (int i = 0; i < Q::NIT; ++i) {
        const int q = t + i * T;
        const int v = q / (N / 2), m = q & (N / 2 - 1);
        const cpx<float> *y = lds + v * P::template ldsv<kPipePS>();
        const cpx<float> F = y[padx<kPipePS>(m)];
        const cpx<float> B2 = y[padx<kPipePS>(m == 0 ? 0 : N - m)];
        const int vo = (int)(v * ovs) * 8;
        const bool live = v < nv;
        bstore(ro, live ? vo + m * 8 : kDropOffset, 0, z(F, B2, em[i % Q::NE]));
        bstore(ro, live ? vo + (N - m) * 8 : kDropOffset, 0, z(B2, F, en[i % Q::NE]));
    }
    // z_{n/2} of row t, for the first VPB lanes
    {
        const int v = t < VPB ? t : 0;
        const cpx<float> H = lds[v * P::template ldsv<kPipePS>() + padx<kPipePS>(N / 2)];
        bstore(ro, (t < VPB && t < nv) ? (int)(v * ovs) * 8 + (N / 2) * 8 : kDropOffset, 0, z(H, H, eh));
    }
    __syncthreads();   // the next group's stage 0 overwrites the LDS rows
}

template <int LOG2N>
__device__ __forceinline__ void pow2_r2c_pipe_body(const vhip_fftm_desc &d, cpx<float> *lds) {
    using P = P2<LOG2N>;
    using Q = P2R<LOG2N>;
    constexpr int N = Q::N, VPB = Q::VPB, T = Q::T;
    const cpx<float> *tw = (const cpx<float> *)d.tw;
    const cpx<float> *eu = (const cpx<float> *)d.euler;
    const cpx<float> *in = (const cpx<float> *)d.in.re;
    cpx<float> *out = (cpx<float> *)d.out.re;
    const long ivs = d.in.vs0 / 2, ovs = d.out.vs0;   // the real side's stride is in reals
    const float half = 0.5f * (float)d.scale;
    const int G = gridDim.x;
    const int ngroups = __builtin_amdgcn_readfirstlane((int)((d.nvec + VPB - 1) / VPB));
    const long nvec = d.nvec;
    int g = blockIdx.x;
    if (g >= ngroups) return;
    const int t = threadIdx.x;
    // per-lane constants first (loads return in order)
    cpx<float> w[2][15];
#pragma unroll
    for (int s = 1; s < Q::NST; ++s) {
        const int R = 1 << P::bits(s), NS = 1 << P::log2ns(s);
#pragma unroll
        for (int r = 1; r < 16; ++r)
            if (r < R) w[s - 1][r - 1] = tw[d.tw_off[s] + (r - 1) * NS + (t & (NS - 1))];
    }
    cpx<float> em[Q::NE], en[Q::NE];
#pragma unroll
    for (int i = 0; i < Q::NE; ++i) {
        const int m = (t + i * T) & (N / 2 - 1);
        em[i] = eu[m];
        en[i] = eu[N - m];
    }
    const cpx<float> eh = eu[N / 2];
    auto rin = [&](int gi) {
        return make_rsrc(in + (long)(gi < ngroups ? gi : 0) * VPB * ivs,
                         gi < ngroups ? group_bytes<LOG2N, VPB>(nvec, gi, ivs) : 0u);
    };
    auto rout = [&](int gi) { return make_rsrc(out + (long)gi * VPB * ovs, group_bytes_len<VPB>(nvec, gi, ovs, N + 1)); };
    auto nvg = [&](int gi) { const long r = nvec - (long)gi * VPB; return r > VPB ? VPB : (int)r; };
    cpx<float> a[16], b[16];
    p2_load<LOG2N, VPB, true>(a, rin(g), ivs, 1.f);
    p2_load<LOG2N, VPB, true>(b, rin(g + G), ivs, 1.f);
    p2r_group<LOG2N>(a, lds, w, em, en, eh, rout(g), ovs, nvg(g), half);
    g += G;
    if (g >= ngroups) return;
    p2_load<LOG2N, VPB, true>(a, rin(g + G), ivs, 1.f);
    for (;;) {
        p2r_group<LOG2N>(b, lds, w, em, en, eh, rout(g), ovs, nvg(g), half);
        g += G;
        if (g >= ngroups) break;
        p2_load<LOG2N, VPB, true>(b, rin(g + G), ivs, 1.f);
        p2r_group<LOG2N>(a, lds, w, em, en, eh, rout(g), ovs, nvg(g), half);
        g += G;
        if (g >= ngroups) break;
        p2_load<LOG2N, VPB, true>(a, rin(g + G), ivs, 1.f);
    }
}

// ---------------------------------------------------------------------------
// Pipelined persistent c2c kernel for the shorter power-of-two rows
// (LOG2N 8..11: groups of VPB = 4096 / N rows, 256 lanes), the design of
// the 4096-point one: workgroup b transforms groups b, b + G, ... with the
// next group's loads in flight; stage twiddles per lane in registers; the
// input conjugation applied at use.

// the last stage of a c2c group, stored to HBM (streamed)
template <int LOG2N, int VPB, int S, bool PLAIN, int PS = 5>
__device__ __forceinline__ void p2g_last(cpx<float> (&x)[16], cpx<float> *lds, const cpx<float> *w, rsrc_t out,
                                         long out_vs, float osgn, float scale) {
    using P = P2<LOG2N>;
    constexpr int N = P::N;
    constexpr int LR = P::bits(S);
    constexpr int R = 1 << LR;
    constexpr int B = 16 / R;
    constexpr int LNB = LOG2N - LR;
    constexpr int NB = 1 << LNB;
    constexpr int NS = 1 << P::log2ns(S);
    constexpr int T = VPB * N / 16;
    static_assert(S == P::NST - 1 && T % NS == 0, "last stage with per-lane twiddles");
    int t = threadIdx.x;
    asm volatile("" : "+v"(t));
#pragma unroll
    for (int i = 0; i < B; ++i) {
        const int b = t + i * T;
        const int v = b >> LNB, j = b & (NB - 1);
        const cpx<float> *src = lds + v * P::template ldsv<PS>() + padx<PS>(j);
#pragma unroll
        for (int r = 0; r < R; ++r) x[i * R + r] = src[padx<PS>(r * NB)];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < B; ++i) {
        const int b = t + i * T;
        const int v = b >> LNB, j = b & (NB - 1);
#pragma unroll
        for (int r = 1; r < R; ++r) x[i * R + r] = cmul(x[i * R + r], w[r - 1]);
        Dft<float, R>::run(&x[i * R]);
        const int vo = (int)(v * out_vs + j) * 8;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const cpx<float> y = x[i * R + r];
            if (PLAIN) bstore<true>(out, vo, r * NB * 8, y);
            else bstore<true>(out, vo, r * NB * 8, {y.x * scale, y.y * (scale * osgn)});
        }
    }
}

template <int LOG2N, bool PLAIN>
__device__ __forceinline__ void p2g_group(cpx<float> (&x)[16], cpx<float> *lds, const cpx<float> (&w)[2][15],
                                          rsrc_t out, long ovs, float isgn, float osgn, float scale) {
    using P = P2<LOG2N>;
    constexpr int VPB = p2_vpb<LOG2N>();
    static_assert(P::NST == 2 || P::NST == 3, "two or three stages");
    if (!PLAIN) {
#pragma unroll
        for (int e = 0; e < 16; ++e) x[e].y *= isgn;
    }
    p2_stage<LOG2N, VPB, 0, PLAIN, false, true, kPipePS>(x, lds, make_rsrc(nullptr, 0u), nullptr, out, ovs, osgn, scale);
    if constexpr (P::NST == 3) {
        p2r_stage<LOG2N, VPB, 1, kPipePS>(x, lds, w[0]);
        p2g_last<LOG2N, VPB, 2, PLAIN, kPipePS>(x, lds, w[1], out, ovs, osgn, scale);
    } else {
        p2g_last<LOG2N, VPB, 1, PLAIN, kPipePS>(x, lds, w[0], out, ovs, osgn, scale);
    }
}

template <int LOG2N, bool PLAIN>
__device__ __forceinline__ void pow2_pipe_gen_body(const vhip_fftm_desc &d, cpx<float> *lds) {
    using P = P2<LOG2N>;
    constexpr int VPB = p2_vpb<LOG2N>();
    static_assert(LOG2N >= 8 && LOG2N <= 11 && VPB * P::N == 4096, "256-lane groups");
    const cpx<float> *tw = (const cpx<float> *)d.tw;
    const cpx<float> *in = (const cpx<float> *)d.in.re;
    cpx<float> *out = (cpx<float> *)d.out.re;
    const long ivs = d.in.vs0, ovs = d.out.vs0;
    const float isgn = d.conj_in ? -1.f : 1.f;
    const float osgn = d.conj_out ? -1.f : 1.f;
    const float scale = (float)d.scale;
    const int G = gridDim.x;
    const long nvec = d.nvec;
    const int ngroups = __builtin_amdgcn_readfirstlane((int)((nvec + VPB - 1) / VPB));
    int g = blockIdx.x;
    if (g >= ngroups) return;
    const int t = threadIdx.x;
    cpx<float> w[2][15];
#pragma unroll
    for (int s = 1; s < P::NST; ++s) {
        const int R = 1 << P::bits(s), NS = 1 << P::log2ns(s);
#pragma unroll
        for (int r = 1; r < 16; ++r)
            if (r < R) w[s - 1][r - 1] = tw[d.tw_off[s] + (r - 1) * NS + (t & (NS - 1))];
    }
    auto rin = [&](int gi) {
        return make_rsrc(in + (long)(gi < ngroups ? gi : 0) * VPB * ivs,
                         gi < ngroups ? group_bytes<LOG2N, VPB>(nvec, gi, ivs) : 0u);
    };
    auto rout = [&](int gi) { return make_rsrc(out + (long)gi * VPB * ovs, group_bytes<LOG2N, VPB>(nvec, gi, ovs)); };
    cpx<float> a[16], b[16];
    p2_load<LOG2N, VPB, true, true>(a, rin(g), ivs, 1.f);
    p2_load<LOG2N, VPB, true, true>(b, rin(g + G), ivs, 1.f);
    p2g_group<LOG2N, PLAIN>(a, lds, w, rout(g), ovs, isgn, osgn, scale);
    g += G;
    if (g >= ngroups) return;
    p2_load<LOG2N, VPB, true, true>(a, rin(g + G), ivs, 1.f);
    for (;;) {
        p2g_group<LOG2N, PLAIN>(b, lds, w, rout(g), ovs, isgn, osgn, scale);
        g += G;
        if (g >= ngroups) break;
        p2_load<LOG2N, VPB, true, true>(b, rin(g + G), ivs, 1.f);
        p2g_group<LOG2N, PLAIN>(a, lds, w, rout(g), ovs, isgn, osgn, scale);
        g += G;
        if (g >= ngroups) break;
        p2_load<LOG2N, VPB, true, true>(a, rin(g + G), ivs, 1.f);
    }
}

// ---------------------------------------------------------------------------
// Pipelined persistent kernels for 8192- and 16384-point rows (radices
// 16/8/8/8 and 16/16/8/8; 512 or 1024 lanes; 2 or 1 workgroups per CU, so
// 128 VGPRs per lane either way), the design of the shorter ones with the
// twiddles fitted into that budget:
//   - stage 1 (k = t mod 16): from a 16 (R1 - 1)-entry LDS table;
//   - stage 2 (k = t mod NS2, a divisor of the lane count): 7 per lane in
//     registers;
//   - stage 3 (k = t + T i for the lane's butterflies i = 0, 1, T = N / 16):
//     the 7 of i = 0 in registers, those of i = 1 as the same times W_16^r
//     (W_N^{r (t + T)} = W_N^{rt} W_16^r), the 7 constants loaded once from
//     the plan's table (uniform, so into SGPRs).
template <int LOG2N>
struct P2B {
    using P = P2<LOG2N>;
    static constexpr int N = 1 << LOG2N;
    static constexpr int T = N / 16;
    static constexpr int R1 = 1 << P::bits(1);
    static constexpr int NS2 = 1 << P::log2ns(2);
    static constexpr int NS3 = 1 << P::log2ns(3);
    static constexpr int TW1 = 16 * (R1 - 1);                  // stage-1 table entries
    static constexpr int LDS = P::template ldsv<kPipePS>() + TW1;
    static_assert(P::NST == 4 && P::bits(0) == 4 && P::bits(2) == 3 && P::bits(3) == 3 && NS3 * 8 == N &&
                      T % NS2 == 0 && TW1 <= T,
                  "8192 = 16 8 8 8, 16384 = 16 16 8 8");
};

template <int LOG2N, bool PLAIN>
__device__ __forceinline__ void p2e_last(cpx<float> (&x)[16], cpx<float> *lds, const cpx<float> (&w)[7],
                                         const cpx<float> (&c)[7], rsrc_t out, float osgn, float scale) {
    constexpr int R = 8, NB = P2B<LOG2N>::N / 8, T = P2B<LOG2N>::T;
    int t = threadIdx.x;
    asm volatile("" : "+v"(t));
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const cpx<float> *src = lds + padx<kPipePS>(t + i * T);
#pragma unroll
        for (int r = 0; r < R; ++r) x[i * R + r] = src[padx<kPipePS>(r * NB)];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const int j = t + i * T;
#pragma unroll
        for (int r = 1; r < R; ++r) x[i * R + r] = cmul(x[i * R + r], i == 0 ? w[r - 1] : cmul(w[r - 1], c[r - 1]));
        Dft<float, R>::run(&x[i * R]);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const cpx<float> y = x[i * R + r];
            if (PLAIN) bstore<true>(out, j * 8, r * NB * 8, y);
            else bstore<true>(out, j * 8, r * NB * 8, {y.x * scale, y.y * (scale * osgn)});
        }
    }
}

// lds: P2<LOG2N>::ldsv<kPipePS>() row elements, then the stage-1 twiddles
template <int LOG2N, bool PLAIN>
__device__ __forceinline__ void pow2_pipe_big_body(const vhip_fftm_desc &d, cpx<float> *lds) {
    using P = P2<LOG2N>;
    using Q = P2B<LOG2N>;
    constexpr int N = Q::N, T = Q::T, R1 = Q::R1;
    cpx<float> *tw1 = lds + P::template ldsv<kPipePS>();
    const cpx<float> *tw = (const cpx<float> *)d.tw;
    const cpx<float> *in = (const cpx<float> *)d.in.re;
    cpx<float> *out = (cpx<float> *)d.out.re;
    const int t = threadIdx.x;
    const long ivs = d.in.vs0, ovs = d.out.vs0;
    const float isgn = d.conj_in ? -1.f : 1.f;
    const float osgn = d.conj_out ? -1.f : 1.f;
    const float scale = (float)d.scale;
    const int G = gridDim.x;
    const int nvec = __builtin_amdgcn_readfirstlane((int)d.nvec);
    int g = blockIdx.x;
    if (g >= nvec) return;
    auto rin = [&](int row) {
        return make_rsrc(in + (long)(row < nvec ? row : 0) * ivs, row < nvec ? (unsigned)(N * 8) : 0u);
    };
    // stage 1 (NS 16): tw1[k (R1 - 1) + r - 1] = W_{16 R1}^{rk}, loaded here by
    // every lane (lanes past the table load entry 0) and written to LDS with
    // the first row, whose wait covers the load; stages 2 and 3 per lane;
    // the stage-3 constants W_16^r
    cpx<float> t1;
    {
        const int e = t < Q::TW1 ? t : 0, k = e / (R1 - 1), r = e - (R1 - 1) * k + 1;
        t1 = tw[d.tw_off[1] + (r - 1) * 16 + k];
    }
    cpx<float> w2[7], w3[7], c3[7];
#pragma unroll
    for (int r = 1; r < 8; ++r) {
        w2[r - 1] = tw[d.tw_off[2] + (r - 1) * Q::NS2 + (t & (Q::NS2 - 1))];
        w3[r - 1] = tw[d.tw_off[3] + (r - 1) * Q::NS3 + t];
        c3[r - 1] = tw[d.tw_off[3] + (r - 1) * Q::NS3 + T];
    }
    const cpx<float> *w1 = tw1 + (t & 15) * (R1 - 1);
    auto row = [&](cpx<float> (&x)[16], int gg) {
        if (!PLAIN) {
#pragma unroll
            for (int e = 0; e < 16; ++e) x[e].y *= isgn;
        }
        const rsrc_t o = make_rsrc(out + (long)gg * ovs, (unsigned)(N * 8));
        p2_stage<LOG2N, 1, 0, PLAIN, false, true, kPipePS>(x, lds, make_rsrc(nullptr, 0u), nullptr, o, 0, osgn, scale);
        p2r_stage<LOG2N, 1, 1, kPipePS>(x, lds, w1);
        p2r_stage<LOG2N, 1, 2, kPipePS>(x, lds, w2);
        p2e_last<LOG2N, PLAIN>(x, lds, w3, c3, o, osgn, scale);
    };
    cpx<float> a[16], b[16];
    p2_load<LOG2N, 1, true, true>(a, rin(g), ivs, 1.f);
    p2_load<LOG2N, 1, true, true>(b, rin(g + G), ivs, 1.f);
    if (t < Q::TW1) tw1[t] = t1;   // read from stage 1 on, after stage 0's barrier
    row(a, g);
    g += G;
    if (g >= nvec) return;
    p2_load<LOG2N, 1, true, true>(a, rin(g + G), ivs, 1.f);
    for (;;) {
        row(b, g);
        g += G;
        if (g >= nvec) break;
        p2_load<LOG2N, 1, true, true>(b, rin(g + G), ivs, 1.f);
        row(a, g);
        g += G;
        if (g >= nvec) break;
        p2_load<LOG2N, 1, true, true>(a, rin(g + G), ivs, 1.f);
    }
}

// One work item of the batched kernel: VPB = 8192 / N rows starting at
// vector vbase of the transform d, with the batched workgroup of 512 lanes.
// nxt: the batched kernel's claim of its next item, made uniform after stage 0
template <int LOG2N>
__device__ __forceinline__ void pow2_batch_item(const vhip_fftm_desc &d, long vbase, char *smem, unsigned &nxt) {
    constexpr int VPB = 8192 >> LOG2N;
    using P = P2<LOG2N>;
    const cpx<float> *in = (const cpx<float> *)d.in.re;
    cpx<float> *out = (cpx<float> *)d.out.re;
    cpx<float> *lds = (cpx<float> *)smem;
    const rsrc_t tw = make_rsrc(d.tw, (unsigned)(d.tw_off[P::NST - 1] + 15 * (P::N / 16) + 16) * 8);
    const float isgn = d.conj_in ? -1.f : 1.f;
    const float osgn = d.conj_out ? -1.f : 1.f;
    const float scale = (float)d.scale;
    const long ivs = d.in.vs0, ovs = d.out.vs0;
    const long g = vbase / VPB;
    cpx<float> x[16];
    p2_load<LOG2N, VPB, false>(x, make_rsrc(in + vbase * ivs, group_bytes<LOG2N, VPB>(d.nvec, g, ivs)), ivs, isgn);
    const rsrc_t o = make_rsrc(out + vbase * ovs, group_bytes<LOG2N, VPB>(d.nvec, g, ovs));
    p2_stage<LOG2N, VPB, 0, false>(x, lds, tw, d.tw_off, o, ovs, osgn, scale);
    nxt = __builtin_amdgcn_readfirstlane(nxt);
    if constexpr (P::NST > 1) p2_stage<LOG2N, VPB, (P::NST > 1 ? 1 : 0), false>(x, lds, tw, d.tw_off, o, ovs, osgn, scale);
    if constexpr (P::NST > 2) p2_stage<LOG2N, VPB, (P::NST > 2 ? 2 : 0), false>(x, lds, tw, d.tw_off, o, ovs, osgn, scale);
    if constexpr (P::NST > 3) p2_stage<LOG2N, VPB, (P::NST > 3 ? 3 : 0), false>(x, lds, tw, d.tw_off, o, ovs, osgn, scale);
}

// Does descriptor d run on the power-of-two path of the batched kernel?
__device__ __forceinline__ int pow2_batch_log2(const vhip_fftm_desc &d) {
    const int n = d.n;
    if ((n & (n - 1)) || n < 256 || n > 8192) return 0;
    if (!d.in.contig || !d.out.contig || d.ptw || d.map_in != VHIP_MAP_ALONG || d.map_out != VHIP_MAP_ALONG)
        return 0;
    if (d.vpb != 8192 / n || d.in.nv0 < d.nvec || d.out.nv0 < d.nvec) return 0;
    return 31 - __builtin_clz(n);
}

}  // namespace p2
}  // namespace vhipk
