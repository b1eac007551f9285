/*
 * vhip.h -- the thin C ABI between the C89 VSIPL host layer and the HIP code.
 *
 * Only plain C types cross this boundary: pointers, integers, doubles and the
 * POD descriptors below.  No HIP type appears here, so the host layer compiles
 * with any C89 compiler and links against the hipcc-built object.
 *
 * Every function returns 0 on success and a nonzero code on failure; the code
 * is the hipError_t value when a HIP call failed (negative values are this
 * layer's own argument errors).
 */
#ifndef VHIP_H
#define VHIP_H

#ifndef __HIPCC_RTC__
#include <stddef.h>
#endif

#ifdef __cplusplus
extern "C" {
#endif

#define VHIP_MAX_STAGES 24

/* transform kinds */
#define VHIP_C2C 0
#define VHIP_R2C 1
#define VHIP_C2R 2

/* how a workgroup's butterflies map onto the vectors it owns at the global
 * load/store: consecutive lanes walk along the vector (unit element stride,
 * ROW-like) or across vectors (unit vector stride, COL-like). */
#define VHIP_MAP_ALONG 0
#define VHIP_MAP_ACROSS 1

/*
 * A strided set of vectors in device memory.  Element n of vector v sits at
 *     base + (v % nv0) * vs0 + (v / nv0) * vs1 + n * es
 * counted in elements of the view (complex elements for complex views, real
 * elements for real views).  For interleaved complex data `re` points at the
 * complex element and `im` is NULL; for split complex data `re`/`im` point at
 * the real and imaginary arrays; for real data `re` is the real array.
 */
typedef struct {
    void *re;
    void *im;
    long es;
    long vs0;
    long vs1;
    long nv0;
    int split;
    int contig;   /* set by the host: interleaved/real with es == 1 and aligned */
    long span;    /* set by the host: bytes from re (and im) that cover every
                     element a launch touches, when all strides are >= 0 and
                     that is below 2 GiB (kernels then use 32-bit buffer
                     offsets); 0 otherwise */
} vhip_view;

/* One launch of the multiple-FFT kernel: nvec independent transforms. */
typedef struct {
    int prec;                      /* 4 = float, 8 = double */
    int kind;                      /* VHIP_C2C / VHIP_R2C / VHIP_C2R */
    int n;                         /* complex FFT length */
    int nst;                       /* number of Stockham stages */
    int radix[VHIP_MAX_STAGES];
    int tw_off[VHIP_MAX_STAGES];   /* stage twiddle offset (complex elements) */
    long nvec;                     /* vectors in this launch */
    int vpb;                       /* vectors per workgroup */
    int threads;                   /* workgroup size */
    int lds_stride;                /* padded vector stride in LDS, complex elements */
    int maxe;                      /* register budget class: 16 or 32 elements */
    int map_in, map_out;           /* VHIP_MAP_ALONG / VHIP_MAP_ACROSS */
    int conj_in;                   /* C2C: conjugate on load (inverse transforms) */
    int conj_out;                  /* C2C/C2R: conjugate on store */
    double scale;                  /* applied in the final store */
    const void *tw;                /* stage twiddles, forward sign */
    const void *euler;             /* R2C: e^{-i pi k/n}; C2R: e^{+i pi k/n}; k = 0..n */
    const void *ptw;               /* post-twiddle table, or NULL: output element k of
                                      vector v is multiplied by w(t) = e^{-2 pi i t / N},
                                      t = (v mod ptw_n) k, stored in two levels:
                                      w(t) = ptw[t mod L] ptw[L + t / L], L = 2^ptw_shift */
    long ptw_n;
    int ptw_shift;
    int stream;                    /* bit 0: load, bit 1: store the data non-temporally
                                      (set by the host for dense, line-aligned rows;
                                      honoured by the compiled-schedule kernels) */
    int compiled;                  /* the plan may run on its compiled-schedule kernels */
    int btw_off;                   /* batched kernel: offset (complex elements) in tw of
                                      the powers of W_n = e^{-2 pi i / n} in two levels,
                                      lo[t] = W_n^t (t < 2^lg), then hi[m] = W_n^{m 2^lg}
                                      (m < ceil(n / 2^lg)), lg = ceil(ceil(log2 n) / 2);
                                      -1 when the schedule has one stage */
    vhip_view in, out;
    /* fused element work of a pipeline (NULL when unused; only the
     * standalone generic kernel implements it, not the batched one):
     *   tap   real weights applied on load, by element position within the
     *         vector: R2C real sample s times tap[s * tap_s]; C2C element j
     *         times tap[j * tap_s]
     *   acc   C2C: instead of storing y, acc[o] += acc_gain * |y|^2 where o
     *         is y's element offset in `out` (acc is real, laid out with the
     *         element and vector strides of `out`) */
    const void *tap;
    long tap_s;
    void *acc;
    double acc_gain;
} vhip_fftm_desc;

/* runtime */
int vhip_init(void);
int vhip_device_count(int *count);
int vhip_lds_max(int *bytes);           /* max dynamic LDS per workgroup */
int vhip_cu_count(int *cus);
int vhip_malloc(void **p, size_t bytes);
int vhip_free(void *p);
int vhip_memcpy_h2d(void *dst, const void *src, size_t bytes);
int vhip_memcpy_d2h(void *dst, const void *src, size_t bytes);
int vhip_memcpy_d2d(void *dst, const void *src, size_t bytes);
int vhip_memset(void *dst, int value, size_t bytes);   /* synchronous, outside graphs only */
/* 2-D strided copies: `rows` rows of `width` bytes, row pitches in bytes */
int vhip_memcpy2d_h2d(void *dst, size_t dpitch, const void *src, size_t spitch,
                      size_t width, size_t rows);
int vhip_memcpy2d_d2h(void *dst, size_t dpitch, const void *src, size_t spitch,
                      size_t width, size_t rows);
int vhip_sync(void);
void *vhip_stream(void);                /* the library stream (hipStream_t) */
const char *vhip_error_string(int code);
size_t vhip_bytes_in_use(void);         /* device bytes held by this library */
void vhip_set_budget(size_t bytes);     /* 0 = no cap; allocations beyond it fail */

/* A strided m x n matrix in device memory: element (i, j) at
 * base + i * cs + j * rs, in elements (complex elements when complex). */
typedef struct {
    void *re;
    void *im;        /* split complex imaginary array, else NULL */
    long rs;         /* stride between consecutive j (VSIPL row_stride) */
    long cs;         /* stride between consecutive i (VSIPL col_stride) */
    int split;
    int pad_;
} vhip_mat;

/* One work item of a batched launch: vectors [vbase, vbase + vpb) of the
 * transform described by descs[desc]. */
typedef struct {
    int desc;
    int pad_;
    long vbase;
} vhip_batch_item;

/*
 * Element passes of a staged transform, which runs its c2c part on
 * contiguous scratch rows (row v of the launch at work + v * m complex
 * elements) between a PRE pass from the caller's view and a POST pass back:
 *   PRE  C2C  w_k = chirp_k * (conj?) x_k, k < n; 0 for n <= k < m
 *        R2C  w_k = chirp_k * (x_2k + i x_2k+1)
 *        C2R  w_k = chirp_k * conj(Y_k), Y_k = (Z_k + conj Z_{n-k})
 *                   + i e_k (Z_k - conj Z_{n-k}), e_k = euler[k]
 *   MUL  w_k *= table[k], k < m (every row)
 *   POST with Y_k = chirp_k * w_{rev(k)}, rev(k) = (m - k) mod m when
 *        reverse, else k:
 *        C2C  y_k = scale * (conj?) Y_k, k < n
 *        R2C  z_k = scale / 2 * [(Y_k + conj Y_{n-k}) - i e_k (Y_k - conj Y_{n-k})],
 *             k = 0..n, Y_n = Y_0
 *        C2R  x_2k + i x_2k+1 = scale * conj(Y_k)
 * chirp == NULL means chirp_k = 1.  The view's vector v of a launch is
 * vector vbase + v of the descriptor's view.
 */
#define VHIP_XF_PRE 0
#define VHIP_XF_MUL 1
#define VHIP_XF_POST 2

typedef struct {
    int prec;
    int kind;              /* VHIP_C2C / VHIP_R2C / VHIP_C2R */
    int step;              /* VHIP_XF_PRE / VHIP_XF_MUL / VHIP_XF_POST */
    int conj;              /* C2C: conjugate on the way in (PRE) or out (POST) */
    int reverse;           /* POST: read w at (m - k) mod m */
    int across;            /* lanes walk across vectors (unit vector stride) */
    long n;                /* complex transform length */
    long m;                /* scratch row length, m >= n */
    long nvec;             /* vectors in this launch */
    long vbase;            /* first vector of the view */
    double scale;          /* POST */
    const void *chirp;     /* n entries or NULL */
    const void *table;     /* MUL: m entries */
    const void *euler;     /* R2C: e^{-i pi k/n}; C2R: e^{+i pi k/n}; k = 0..n */
    vhip_view view;
    void *work;
} vhip_xform_desc;

/* Element operations on strided complex m x n matrices (vhip_mat), in the
 * working precision `prec`:
 *   scale   r = alpha * a (real alpha)
 *   sub     c = a - b
 *   sumsq   *out = sum |a_ij|^2, accumulated in double (synchronous) */
int vhip_elem_scale(int prec, const vhip_mat *a, const vhip_mat *r, long m, long n, double alpha,
                    void *stream);
int vhip_elem_sub(int prec, const vhip_mat *a, const vhip_mat *b, const vhip_mat *c, long m, long n,
                  void *stream);
int vhip_elem_sumsq(int prec, const vhip_mat *a, long m, long n, double *out, void *stream);
/* Real m x n element operations r = op(...), r real:
 *   FILL      r = alpha
 *   SMUL      r = alpha * a
 *   ADD       r = a + b
 *   VMUL_ROW  r_ij = vec[j * vs] * a_ij      (vector along each row)
 *   VMUL_COL  r_ij = vec[i * vs] * a_ij      (vector down each column)
 *   MAGSQ     r = |a|^2, a complex
 * Products and sums are rounded separately (no fused multiply-add). */
#define VHIP_EOP_FILL 0
#define VHIP_EOP_SMUL 1
#define VHIP_EOP_ADD 2
#define VHIP_EOP_VMUL_ROW 3
#define VHIP_EOP_VMUL_COL 4
#define VHIP_EOP_MAGSQ 5
int vhip_elem_real(int prec, int op, const vhip_mat *a, const vhip_mat *b, const vhip_mat *r, const void *vec,
                   long vs, double alpha, long m, long n, void *stream);

/* 1-D correlation passes (corr_ops.hip) on a transform buffer of two rows
 * of N complex elements (interleaved, precision prec):
 *   pack     row 0 = N - n zeros then x (n elements), row 1 = h (m elements)
 *            then zeros; x_cplx / h_cplx 0 for real vector views
 *   product  row0[k] = conj(row0[k] conj(row1[k])) / N
 *   window   y[i] = w_i row0[o + i], i < len, w_i the unbiasing weights of
 *            the support (VHIP_CORR_*) when unbiased; y_cplx 0 stores the
 *            real part to a real view
 * x, h, y are vector views: element j at re (+ im) + j * rs. */
#define VHIP_CORR_FULL 0
#define VHIP_CORR_SAME 1
#define VHIP_CORR_MIN 2
int vhip_corr_pack(int prec, const vhip_mat *x, int x_cplx, long n, const vhip_mat *h, int h_cplx, long m,
                   void *buf, long N, void *stream);
int vhip_corr_product(int prec, void *buf, long N, void *stream);
/* convolution on a real row of even length L around real-to-complex and
 * complex-to-real transforms: pack_r r = x (n reals) then zeros; mul_half
 * spec[k] = spec[k] kspec[k] / L for k <= L/2; window_r y[k] = r[s + k D]
 * for k < len */
int vhip_conv_pack_r(int prec, const vhip_mat *x, long n, void *r, long L, void *stream);
int vhip_conv_mul_half(int prec, void *spec, const void *kspec, long L, void *stream);
int vhip_conv_window_r(int prec, const void *r, long L, long s, long D, const vhip_mat *y, long len, void *stream);
int vhip_corr_window(int prec, const void *buf, long N, long o, const vhip_mat *y, int y_cplx, long len,
                     int support, int unbiased, long m, long n, void *stream);

/* One fill of a strided m x n view by a VSIPL generator (rng.hip).  Element
 * e (column-major order when by_col, else row-major) takes draws
 * D e + 1 .. D e + D of this launch.  x0 / x10 are the generator states
 * before the first draw; (ja, jc) and (ja1, jc1) the maps of 2^i steps;
 * stride_* the maps of `lanes` elements' draws (one grid stride); kstar the
 * 1-based draw at which X1 meets its marker x2 (then X1 continues from
 * x2 + 1), 0 when that does not happen in this launch. */
#define VHIP_RNG_UNIFORM 0
#define VHIP_RNG_NORMAL 1
typedef struct {
    int prec, dist, cplx, portable, by_col;
    int p24;          /* portable draw t as (t >> 8 | 1) / 2^24, else t / 2^32 */
    long m, n, lanes, kstar;
    unsigned a, c, a1, c1, x0, x10, x2;
    unsigned stride_a, stride_c, stride_a1, stride_c1;
    unsigned ja[32], jc[32], ja1[32], jc1[32];
    vhip_mat r;
} vhip_rng_desc;
long vhip_rng_lanes(long total);
int vhip_rng_fill(const vhip_rng_desc *d, void *stream);

/* Kernels compiled at plan creation for one schedule (hipRTC): prepare
 * compiles and loads the kernel of (n, radices, vpb, threads, load layout,
 * store layout) once per process (0 when available); try_launch runs a
 * descriptor on it when it is loaded and the descriptor's layouts match,
 * else returns -100.  Layouts: VHIP_JIT_CONTIG (interleaved rows of unit
 * element stride, single-level vector stride), VHIP_JIT_ACROSS (interleaved,
 * lanes across vectors, any strides within vhip_view.span); + VHIP_JIT_FUSED
 * on the load side fuses vhip_fftm_desc.tap, on the store side .acc;
 * + VHIP_JIT_PTW on the store side applies .ptw (compiled in, so a kernel
 * without it has no post-twiddle branch at its stores). */
#define VHIP_JIT_CONTIG 0
#define VHIP_JIT_ACROSS 1
#define VHIP_JIT_FUSED 2
#define VHIP_JIT_REAL 4     /* load side: real input, real-to-complex transform */
#define VHIP_JIT_PTW 8      /* store side: post-twiddle (the split's passes before the last) */
#define VHIP_JIT_C2R 16     /* load side: complex-to-real transform (n + 1 inputs, real output) */
#define VHIP_JIT_SPLIT_IN 32   /* load side: split complex input (plain c2c) */
#define VHIP_JIT_SPLIT_OUT 16  /* store side: split complex output (plain c2c) */
int vhip_jit_prepare(int prec, int n, int nst, const int *radix, int vpb, int threads, int layout_in,
                     int layout_out);
int vhip_jit_ready(int prec, int n, int nst, const int *radix, int vpb, int threads, int layout_in, int layout_out);
int vhip_jit_try_launch(const vhip_fftm_desc *d, void *stream);
/* 0: launches never compile a layout the plan did not prepare (while a
 * resident kernel waits for the host, which no module load may block); 1: they
 * may (the default) */
void vhip_jit_launch_compiles(int on);
int vhip_jit_prepare_desc(const vhip_fftm_desc *d);
int vhip_jit_compile_check(int prec, int n, int nst, const int *radix, int vpb, int threads, int layout_in,
                           int layout_out, const char *arch, size_t *code_bytes);
const char *vhip_jit_last_error(void);

/* would complex FP32 rows of n points (256..2048), nvec of them, take the
 * pipelined power-of-two kernel (1) or its one-group form (0) */
int vhip_p2_pipelined(int n, long nvec);

/* kernels */
int vhip_xform_launch(const vhip_xform_desc *d, void *stream);
int vhip_mat_copy(int prec, int cplx, const vhip_mat *src, const vhip_mat *dst,
                  long m, long n, void *stream);
int vhip_fftm_launch(const vhip_fftm_desc *d, void *stream);

/* Batched launch: a persistent kernel of `threads` lanes per workgroup and
 * `lds_bytes` of LDS drains items[0 .. nitems1), then (when nitems2 > 0) a
 * second one drains items[nitems1 .. nitems1 + nitems2), the second passes of
 * two-pass transforms.  ctl holds 4 counters that must be zero before the
 * first launch (vhip_batch_clear); every launch leaves them zero again.  All
 * launches go to `stream` (which may be capturing a graph).  threads must be
 * 512 and maxe 16 (the batched kernel's only shape). */
int vhip_batch_clear(unsigned *ctl, void *stream);
int vhip_batch_launch(const vhip_fftm_desc *descs, const vhip_batch_item *items, long nitems1,
                      long nitems2, unsigned *ctl, int threads, int maxe, int lds_bytes,
                      int grid, void *stream);
int vhip_batch_max_grid(int threads, int maxe, int lds_bytes, int *grid);

/* The server's resident kernel: a persistent grid of the batched kernel's
 * shape (512 lanes, 16 elements per lane, lds_bytes of LDS) that takes the
 * work items of a request stream as the host publishes them, with no launch
 * per request.
 *   hdesc      coherent host memory: one descriptor per request slot,
 *              VHIP_RES_DESC_STRIDE bytes apart (the base aligned to that);
 *              slot s is written once, before any item of it is published
 *   items      device: every item of the stream in publication order,
 *              items[i].desc = the request's slot
 *   req_left   device: per slot, the number of its items (counted down)
 *   ctl        device: 4 counters, zero before the launch
 *   htail      coherent host memory: items published so far (only grows)
 *   hstop      coherent host memory: nonzero makes waiting workgroups stop
 *   hdone      coherent host memory: per slot, set to 1 (system scope, after
 *              the request's stores are written back) when its last item is
 *              done
 * The grid ends once it has claimed every one of nitems items, or when a
 * wait exceeds spin_limit polls or hstop is set (ctl[2] then reads 1 and the
 * unfinished requests never get their done flag). */
#define VHIP_RES_DESC_STRIDE 512
typedef struct {
    const vhip_fftm_desc *hdesc;
    const vhip_batch_item *items;
    unsigned *req_left;
    unsigned *ctl;
    const unsigned *htail;
    const unsigned *hstop;
    unsigned *hdone;
    unsigned nitems;
    unsigned spin_limit;
} vhip_resident_args;
int vhip_resident_launch(const vhip_resident_args *a, int grid, int lds_bytes, void *stream);
/* CUs that persistent launches leave to a resident kernel (their grids are
 * sized to the remaining CUs while it runs) */
int vhip_set_reserved_cus(int n);
/* coherent pinned host memory (the device reads and writes it over the bus,
 * uncached), and its device address */
int vhip_host_alloc_coherent(void **p, size_t bytes);
int vhip_host_device_ptr(void **dev, void *host);
int vhip_stream_done(void *stream);   /* 0 idle, 1 work pending, else error */
int vhip_stream_sync(void *stream);
/* host-side release store / acquire load of a word the device reads or
 * writes (coherent host memory) */
void vhip_host_publish(unsigned *p, unsigned v);
unsigned vhip_host_peek(const unsigned *p);

/* The no-allocation batched entry point's scratch: a static device array of
 * VHIP_NOALLOC_BYTES per GPU (control counters at offset 0, descriptors and
 * work items from VHIP_NOALLOC_HEAD on).  vhip_noalloc_setup (called by
 * vhip_init with the library's device current) looks up this device's copy;
 * vhip_noalloc_scratch returns it.  The stage calls order the reuse of the
 * caller's two host staging halves: wait before refilling half `slot`, mark
 * it after queueing its upload on `stream`. */
#define VHIP_NOALLOC_BYTES (4u << 20)
#define VHIP_NOALLOC_HEAD 256u
int vhip_noalloc_setup(void);
int vhip_noalloc_scratch(void **ptr);
int vhip_noalloc_stage_wait(int slot);
int vhip_noalloc_stage_done(int slot, void *stream);

/* graphs: capture everything issued to `stream` between begin and end */
int vhip_stream_create(void **stream);
int vhip_stream_destroy(void *stream);
int vhip_graph_begin(void *stream);
int vhip_graph_end(void *stream, void **graph_exec);
int vhip_graph_launch(void *graph_exec, void *stream);
int vhip_graph_destroy(void *graph_exec);
/* events, pinned host memory and async copies on a given stream */
int vhip_event_create(void **ev);
int vhip_event_destroy(void *ev);
int vhip_event_record(void *ev, void *stream);
int vhip_event_query(void *ev);          /* 0 done, 1 not yet, else error */
int vhip_host_alloc(void **p, size_t bytes);
int vhip_host_free(void *p);
int vhip_memcpy_h2d_async(void *dst, const void *src, size_t bytes, void *stream);

/* HBM reference kernels over `bytes` of device memory: kind 0 in-place
 * read-modify-write, kind 1 read-only (sink: 4 writable bytes) */
int vhip_probe_launch(int kind, void *buf, size_t bytes, void *sink, void *stream);
int vhip_colprobe_launch(const void *in, void *out, int rows, int cols, long ld_in, long ld_out, int lg,
                         void *stream);
/* time `reps` runs of fn(arg) on stream with events; returns milliseconds per run */
int vhip_event_time(void *stream, int (*fn)(void *arg), void *arg, int reps, float *ms);
size_t vhip_fftm_lds_bytes(const vhip_fftm_desc *d);   /* dynamic LDS of one launch */

#ifdef __cplusplus
}
#endif
#endif /* VHIP_H */
