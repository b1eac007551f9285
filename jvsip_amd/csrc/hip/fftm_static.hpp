// fftm_static.hpp -- the Stockham transform of fftm_core.hpp with its whole
// schedule fixed at compile time: element type T (float or double), length
// N, radices R0..R7 (NST of them), vectors per workgroup VPB, lanes NT and
// the layout of the global load and store (MIN, MOUT).  The library compiles
// one such kernel per plan schedule at plan creation (hipRTC, in-process
// cache) for interleaved complex data; every butterfly index, twiddle index
// and LDS offset then folds to
// constants, shifts and multiply-highs instead of the runtime divisions and
// address arithmetic of the schedule-generic kernel.
//
// Stage s (radix R = R_s, Ns = R_0 ... R_{s-1}): butterfly b of the group
// is vector v, index j; it reads elements j + r N/R (from HBM in stage 0,
// else from LDS), multiplies leg r by W_{Ns R}^{r k}, k = j mod Ns (stage
// twiddles at d.tw + d.tw_off[s]), runs the R-point DFT and writes leg r to
// (j - k) R + k + r Ns (to HBM, scaled, in the last stage).  Inverse
// transforms conjugate on the way in and out.
//
// Global layouts (bit 0 of MIN / MOUT):
//   ST_CONTIG  rows of unit element stride, one vector stride (vs0); lanes
//              walk along the vector (v = b / (N/R))
//   ST_ACROSS  any element stride, two-level vector offsets
//              (v % nv0) vs0 + (v / nv0) vs1; lanes walk across the group's
//              vectors (v = b mod VPB), for unit vector stride (columns)
// Middle stages always walk along the vector in LDS.  Bit 1 of MIN fuses a
// real taper into the load (element j times d.tap[j * d.tap_s]); bit 1 of
// MOUT replaces the store by d.acc[o] += d.acc_gain |y|^2 (vhip_fftm_desc).
// Bit 2 of MIN makes the transform real to complex (rcfftmop): the load
// reads real samples 2m, 2m + 1 as complex element m, the last stage leaves
// Y in LDS, and the n + 1 outputs are combined from Y[k] and Y[n - k] with
// the Euler table (r2c_post of the generic kernel).  Bit 4 of MIN makes it
// complex to real (crfftmop): the n + 1 inputs of each vector are combined
// into conj(Y) in LDS (c2r_pre of the generic kernel), stage 0 reads LDS,
// and the last stage conjugates back and stores complex element m as real
// samples 2m, 2m + 1 (offsets of the output side in real elements).  Bit 5
// of MIN / bit 4 of MOUT make that side split complex: the real and
// imaginary parts in two arrays at the same element offsets, addressed per
// vector offset and element stride whatever the lane mapping (plain c2c
// only).
#pragma once
#include "fftm_core.hpp"

#define ST_CONTIG 0
#define ST_ACROSS 1
#define ST_FUSED 2
#define ST_REAL 4
#define ST_PTW 8    // store side: the split passes' post-twiddle (vhip_fftm_desc.ptw)
#define ST_C2R 16   // load side: complex to real (crfftmop)
#define ST_SPLIT_IN 32   // load side: split complex (real and imaginary arrays)
#define ST_SPLIT_OUT 16  // store side: split complex

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

// padded LDS index (one spare slot per 32 complex FP32 or 16 FP64 values),
// as the host planner sizes it
template <typename T>
__device__ __forceinline__ constexpr int lpad(int i) { return i + (i >> (sizeof(T) == 4 ? 5 : 4)); }
template <typename T>
constexpr int lstride(int n) { return (lpad<T>(n) % 2 == 0) ? lpad<T>(n) + 1 : lpad<T>(n); }

template <typename T>
struct Io {
    rsrc_t in, out, tw;
    rsrc_t in_im, out_im;  // split sides: the imaginary arrays
    long ivs, ovs;         // ST_CONTIG vector strides
    int ies, oes;          // ST_ACROSS element strides
    int nv;                // vectors in this group
    const int *voff_in;    // ST_ACROSS: element offset of each vector of the group
    const int *voff_out;
    const int *vidx;       // post-twiddled launches: index of each vector within its row
    long vbase;            // first vector of the group
    T isgn, osgn, scale;
};

template <typename T, int N, int VPB, int NT, int MIN, int MOUT, int S, class SC>
__device__ __forceinline__ void sstage(cpx<T> *lds, const vhip_fftm_desc &d, const Io<T> &io) {
    constexpr int ES = (int)sizeof(cpx<T>);
    constexpr int R = SC::rad(S);
    constexpr int NS = SC::ns(S);
    constexpr int NBV = N / R;
    constexpr int NBT = VPB * NBV;
    constexpr int B = (NBT + NT - 1) / NT;
    constexpr bool FIRST = (S == 0), LAST = (S == SC::NST - 1);
    constexpr int LIN = MIN & 1, LOUT = MOUT & 1;
    constexpr bool TAP = (MIN & ST_FUSED) != 0, ACC = (MOUT & ST_FUSED) != 0;
    constexpr bool REAL = (MIN & ST_REAL) != 0;
    constexpr bool C2R = (MIN & ST_C2R) != 0;
    constexpr bool PTW = (MOUT & ST_PTW) != 0;
    constexpr bool SPIN = (MIN & ST_SPLIT_IN) != 0, SPOUT = (MOUT & ST_SPLIT_OUT) != 0;
    constexpr int RS = (int)sizeof(T);
    // stage 0 of a complex-to-real transform reads the combined input in LDS
    constexpr bool GLOAD = FIRST && !C2R;
    // lane mapping of this stage: the global layout's for the first and last
    // stage (the load's when one stage does both), along the vector otherwise
    constexpr int MODE = GLOAD ? LIN : (LAST ? LOUT : ST_CONTIG);
    constexpr int LS = lstride<T>(N);
    int t = threadIdx.x;
    asm volatile("" : "+v"(t));
    cpx<T> x[B][R];
    // this stage's twiddles, loaded (from L2) ahead of the LDS reads and the
    // barrier, so their latency overlaps both instead of following them
    cpx<T> w[B][R > 1 ? R - 1 : 1];
    int vv[B], jj[B];
    bool ok[B];
#pragma unroll
    for (int i = 0; i < B; ++i) {
        const int b = t + i * NT;
        int v, j;
        if (MODE == ST_ACROSS) { v = b % VPB; j = b / VPB; }
        else { v = b / NBV; j = b - v * NBV; }
        vv[i] = v;
        jj[i] = j;
        ok[i] = (B * NT <= NBT || b < NBT) && v < io.nv;
        if (NS > 1 && ok[i]) {
            const int k = j % NS;
            const int toff = d.tw_off[S];
#pragma unroll
            for (int r = 1; r < R; ++r) w[i][r - 1] = Buf<T>::template ldc<0>(io.tw, (toff + k + (r - 1) * NS) * ES, 0);
        }
    }
#pragma unroll
    for (int i = 0; i < B; ++i) {
        const int v = vv[i], j = jj[i];
        if (ok[i]) {
            if (GLOAD) {
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    cpx<T> y;
                    if (REAL) {             // real samples 2m, 2m + 1 (offsets in real elements)
                        const int m2 = 2 * (j + r * NBV);
                        if (LIN == ST_CONTIG) {
                            y = Buf<T>::template ldc<kLdData>(io.in, (int)(v * io.ivs + m2) * RS, 0);
                        } else {
                            const int o = (io.voff_in[v] + m2 * io.ies) * RS;
                            y.x = Buf<T>::template lds<kLdData>(io.in, o, 0);
                            y.y = Buf<T>::template lds<kLdData>(io.in, o + io.ies * RS, 0);
                        }
                        if (TAP) {
                            const long s0 = (long)m2 * d.tap_s;
                            y.x *= ((const T *)d.tap)[s0];
                            y.y *= ((const T *)d.tap)[s0 + d.tap_s];
                        }
                    } else if (SPIN) {
                        const int o = (io.voff_in[v] + (j + r * NBV) * io.ies) * RS;
                        y.x = Buf<T>::template lds<kLdData>(io.in, o, 0);
                        y.y = Buf<T>::template lds<kLdData>(io.in_im, o, 0) * io.isgn;
                    } else {
                        if (LIN == ST_CONTIG && (d.stream & 1))
                            y = Buf<T>::template ldc<2>(io.in, (int)(v * io.ivs + j) * ES, r * NBV * ES);
                        else if (LIN == ST_CONTIG)
                            y = Buf<T>::template ldc<kLdData>(io.in, (int)(v * io.ivs + j) * ES, r * NBV * ES);
                        else
                            y = Buf<T>::template ldc<kLdData>(io.in, (io.voff_in[v] + (j + r * NBV) * io.ies) * ES, 0);
                        if (TAP) {
                            const T tv = ((const T *)d.tap)[(long)(j + r * NBV) * d.tap_s];
                            y.x *= tv;
                            y.y *= tv;
                        }
                        y.y *= io.isgn;
                    }
                    x[i][r] = y;
                }
            } else {
                const cpx<T> *src = lds + v * LS;
#pragma unroll
                for (int r = 0; r < R; ++r) x[i][r] = src[lpad<T>(j + r * NBV)];
            }
        }
    }
    if (!GLOAD) __syncthreads();
    // every butterfly of the lane first, then every write: a store's data
    // registers are then never reused while stores are in flight (the
    // compiler would wait for the stores before reusing them)
#pragma unroll
    for (int i = 0; i < B; ++i) {
        if (ok[i]) {
            if (NS > 1) {
#pragma unroll
                for (int r = 1; r < R; ++r) x[i][r] = cmul(x[i][r], w[i][r - 1]);
            }
            Dft<T, R>::run(x[i]);
        }
    }
#pragma unroll
    for (int i = 0; i < B; ++i) {
        if (ok[i]) {
            const int v = vv[i], j = jj[i];
            const int k = (NS > 1) ? j % NS : 0;
            if (LAST && REAL) {     // Y in natural order, for the combine
                cpx<T> *dst = lds + v * LS;
#pragma unroll
                for (int r = 0; r < R; ++r) dst[lpad<T>(j + r * NBV)] = x[i][r];
            } else if (LAST && C2R) {   // complex element m as real samples 2m, 2m + 1
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const cpx<T> z = {x[i][r].x * io.scale, -x[i][r].y * io.scale};
                    const int m2 = 2 * (j + r * NBV);
                    if (LOUT == ST_CONTIG) {
                        Buf<T>::template stc<kStData>(io.out, (int)(v * io.ovs + m2) * RS, 0, z);
                    } else {
                        const int o = (io.voff_out[v] + m2 * io.oes) * RS;
                        Buf<T>::template sts<kStData>(io.out, o, 0, z.x);
                        Buf<T>::template sts<kStData>(io.out, o + io.oes * RS, 0, z.y);
                    }
                }
            } else if (LAST) {
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    cpx<T> y = x[i][r];
                    if (PTW) {    // split passes: e^{-2 pi i t / N} from the two-level table
                        const int t = io.vidx[v] * (j + r * NBV);
                        const cpx<T> *w = (const cpx<T> *)d.ptw;
                        y = cmul(y, cmul(w[t & ((1 << d.ptw_shift) - 1)], w[(1 << d.ptw_shift) + (t >> d.ptw_shift)]));
                    }
                    const cpx<T> z = {y.x * io.scale, y.y * (io.scale * io.osgn)};
                    if (SPOUT) {
                        const int o = (io.voff_out[v] + (j + r * NBV) * io.oes) * RS;
                        Buf<T>::template sts<kStData>(io.out, o, 0, z.x);
                        Buf<T>::template sts<kStData>(io.out_im, o, 0, z.y);
                    } else if (ACC) {
                        const long o = LOUT == ST_CONTIG ? (io.vbase + v) * io.ovs + j + r * NBV
                                                         : (long)io.voff_out[v] + (long)(j + r * NBV) * io.oes;
                        T *a = (T *)d.acc + o;
                        *a += (T)d.acc_gain * (z.x * z.x + z.y * z.y);
                    } else if (LOUT == ST_CONTIG && (d.stream & 2))
                        Buf<T>::template stc<18>(io.out, (int)(v * io.ovs + j) * ES, r * NBV * ES, z);
                    else if (LOUT == ST_CONTIG)
                        Buf<T>::template stc<kStData>(io.out, (int)(v * io.ovs + j) * ES, r * NBV * ES, z);
                    else Buf<T>::template stc<kStData>(io.out, (io.voff_out[v] + (j + r * NBV) * io.oes) * ES, 0, z);
                }
            } else {
                const int base = (j - k) * R + k;
                cpx<T> *dst = lds + v * LS;
#pragma unroll
                for (int r = 0; r < R; ++r) dst[lpad<T>(base + r * NS)] = x[i][r];
            }
        }
    }
    if (!LAST) __syncthreads();
}

template <typename T, int N, int VPB, int NT, int MIN, int MOUT, int S, class SC>
__device__ __forceinline__ void sstages(cpx<T> *lds, const vhip_fftm_desc &d, const Io<T> &io) {
    sstage<T, N, VPB, NT, MIN, MOUT, S, SC>(lds, d, io);
    if constexpr (S + 1 < SC::NST) sstages<T, N, VPB, NT, MIN, MOUT, S + 1, SC>(lds, d, io);
}

// One workgroup per group of VPB vectors.  ST_CONTIG sides address the
// group's rows through a buffer range that ends at the last row of the view
// (rows past it read zeros, their stores are dropped); ST_ACROSS sides
// address the whole view (vhip_view.span) and skip vectors past its end.
// Column-mapped launches take the XCD-aware group numbering of the generic
// kernel (xcd_group).
template <typename T, int N, int VPB, int NT, int MIN, int MOUT, int NST, int R0, int R1, int R2, int R3, int R4,
          int R5, int R6, int R7>
__global__ void __launch_bounds__(NT) fftm_static_kernel(const vhip_fftm_desc d) {
    using SC = Sched<NST, R0, R1, R2, R3, R4, R5, R6, R7>;
    constexpr int LS = lstride<T>(N);
    constexpr int ES = (int)sizeof(cpx<T>);
    // static LDS: the size is known here, and a static allocation may use
    // all of the workgroup's LDS without a launch attribute
    __shared__ __attribute__((aligned(16))) cpx<T> slds[VPB * LS];
    __shared__ int voff[3 * VPB];
    constexpr bool ACROSS = ((MIN & 1) == ST_ACROSS || (MOUT & 1) == ST_ACROSS);
    const long g = ACROSS ? (long)xcd_group(blockIdx.x, gridDim.x) : (long)blockIdx.x;
    const long vbase = g * VPB;
    long nvl = d.nvec - vbase;
    Io<T> io;
    io.nv = nvl < VPB ? (int)nvl : VPB;
    io.ivs = d.in.vs0;
    io.ovs = d.out.vs0;
    io.ies = (int)d.in.es;
    io.oes = (int)d.out.es;
    io.voff_in = voff;
    io.voff_out = voff + VPB;
    io.vidx = voff + 2 * VPB;
    constexpr bool PTW = (MOUT & ST_PTW) != 0;
    constexpr bool SPIN = (MIN & ST_SPLIT_IN) != 0, SPOUT = (MOUT & ST_SPLIT_OUT) != 0;
    if (ACROSS || PTW || SPIN || SPOUT) {
        for (int t = threadIdx.x; t < io.nv; t += NT) {
            const long v = vbase + t;
            voff[t] = (int)vec_off(d.in, v);
            voff[VPB + t] = (int)vec_off(d.out, v);
            voff[2 * VPB + t] = (int)(d.ptw_n > 0 ? v % d.ptw_n : v);
        }
        __syncthreads();
    }
    io.vbase = vbase;
    constexpr bool REAL = (MIN & ST_REAL) != 0;
    constexpr bool C2R = (MIN & ST_C2R) != 0;
    constexpr int RS = (int)sizeof(T);
    if (SPIN) {
        io.in = make_rsrc(d.in.re, d.in.span);
        io.in_im = make_rsrc(d.in.im, d.in.span);
    } else if ((MIN & 1) == ST_CONTIG && REAL)
        io.in = make_rsrc((const T *)d.in.re + vbase * io.ivs, ((io.nv - 1) * io.ivs + 2 * N) * RS);
    else if ((MIN & 1) == ST_CONTIG)
        io.in = make_rsrc((const cpx<T> *)d.in.re + vbase * io.ivs, ((io.nv - 1) * io.ivs + N + (C2R ? 1 : 0)) * ES);
    else
        io.in = make_rsrc(d.in.re, d.in.span);
    if (SPOUT) {
        io.out = make_rsrc(d.out.re, d.out.span);
        io.out_im = make_rsrc(d.out.im, d.out.span);
    } else if ((MOUT & 1) == ST_CONTIG && C2R)
        io.out = make_rsrc((T *)d.out.re + vbase * io.ovs, ((io.nv - 1) * io.ovs + 2 * N) * RS);
    else if ((MOUT & 1) == ST_CONTIG)
        io.out = make_rsrc((cpx<T> *)d.out.re + vbase * io.ovs, ((io.nv - 1) * io.ovs + N + (REAL ? 1 : 0)) * ES);
    else
        io.out = make_rsrc(d.out.re, d.out.span);
    io.tw = make_rsrc(d.tw, (long)N * ES);
    io.isgn = d.conj_in ? (T)-1 : (T)1;
    io.osgn = d.conj_out ? (T)-1 : (T)1;
    io.scale = (T)d.scale;
    if constexpr (C2R) {
        // Y_k = (Z_k + conj Z_{n-k}) + i W^k (Z_k - conj Z_{n-k}), W the Euler
        // table e^{+i pi k / n}, k = 0..n-1; conj(Y) goes to LDS, so the
        // forward stages compute the inverse
        const cpx<T> *eu = (const cpx<T> *)d.euler;
        for (int t = threadIdx.x; t < io.nv * N; t += NT) {
            int v, k;
            if ((MIN & 1) == ST_ACROSS) { v = t % io.nv; k = t / io.nv; }
            else { v = t / N; k = t - v * N; }
            cpx<T> F, B;
            if ((MIN & 1) == ST_CONTIG) {
                F = Buf<T>::template ldc<kLdData>(io.in, (int)(v * io.ivs + k) * ES, 0);
                B = Buf<T>::template ldc<kLdData>(io.in, (int)(v * io.ivs + N - k) * ES, 0);
            } else {
                F = Buf<T>::template ldc<kLdData>(io.in, (io.voff_in[v] + k * io.ies) * ES, 0);
                B = Buf<T>::template ldc<kLdData>(io.in, (io.voff_in[v] + (N - k) * io.ies) * ES, 0);
            }
            const cpx<T> Bc = {B.x, -B.y};
            const cpx<T> A = F + Bc;
            const cpx<T> WD = cmul(eu[k], F - Bc);
            slds[v * LS + lpad<T>(k)] = {A.x - WD.y, -(A.y + WD.x)};
        }
        __syncthreads();
    }
    sstages<T, N, VPB, NT, MIN, MOUT, 0, SC>(slds, d, io);
    if constexpr (REAL) {
        // the n + 1 outputs of each vector: Z_k = (A + i^-1 W_k D) / 2 with
        // A = Y_k + conj(Y_{n-k}), D = Y_k - conj(Y_{n-k}), W_k the Euler table
        __syncthreads();
        constexpr int PER = N + 1;
        const cpx<T> *eu = (const cpx<T> *)d.euler;
        const T half = (T)0.5 * io.scale;
        for (int t = threadIdx.x; t < io.nv * PER; t += NT) {
            int v, k;
            if ((MOUT & 1) == ST_ACROSS) { v = t % io.nv; k = t / io.nv; }
            else { v = t / PER; k = t - v * PER; }
            const cpx<T> *y = slds + v * LS;
            const cpx<T> F = y[lpad<T>(k == N ? 0 : k)];
            const cpx<T> Bv = y[lpad<T>(k == 0 ? 0 : N - k)];
            const cpx<T> Bc = {Bv.x, -Bv.y};
            const cpx<T> A = F + Bc;
            const cpx<T> WD = cmul(eu[k], F - Bc);
            const cpx<T> z = {half * (A.x + WD.y), half * (A.y - WD.x)};
            if ((MOUT & 1) == ST_CONTIG) Buf<T>::template stc<kStData>(io.out, (int)(v * io.ovs + k) * ES, 0, z);
            else Buf<T>::template stc<kStData>(io.out, (io.voff_out[v] + k * io.oes) * ES, 0, z);
        }
    }
}

}  // namespace st
}  // namespace vhipk
