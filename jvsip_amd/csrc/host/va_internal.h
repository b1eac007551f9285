/*
 * va_internal.h -- private structures of the MI355X VSIPL FFTM library.
 *
 * The public handles of vsip.h are incomplete struct types; this file gives
 * them their layout.  One block structure (VA_block) and one matrix-view
 * structure (VA_mview) serve all four element types (real/complex x
 * float/double), the public struct of each type wraps it as its only member.
 *
 * Storage (reference behaviour: vsip_cblockattributes_f.h:42-67,
 * VI_cblockbind_f_di.h:17-57, VI_cblockadmit_f_di.h:16-29,
 * VI_cblockrelease_f_di.h:16-34, vsip_cblockcreate_f.c:20-86):
 *   - user blocks (bind) start released and keep the caller's arrays;
 *   - library blocks (create) are born admitted and own device memory only;
 *   - complex user blocks bound with Ip == NULL are interleaved (cstride 2),
 *     with Ip != NULL split (cstride 1); library complex blocks are
 *     interleaved.
 * While admitted, a block's data lives in HBM (dev_r/dev_i).  Admit with
 * update copies host -> device, release with update copies device -> host.
 *
 * C89.
 */
#ifndef VA_INTERNAL_H
#define VA_INTERNAL_H

#include <stddef.h>
#include "vsip.h"
#include "vhip.h"

#define VA_USER_BLOCK 1
#define VA_LIB_BLOCK 0
#define VA_DERIVED_BLOCK 2   /* real or imaginary part of a complex block */
#define VA_VALID 0x5555
#define VA_FREED 0xAAAA

typedef struct VA_block {
    int prec;              /* sizeof the real scalar: 4 or 8 */
    int cplx;              /* 1 complex, 0 real */
    int kind;              /* VA_USER_BLOCK or VA_LIB_BLOCK */
    int admitted;          /* 1 admitted, 0 released */
    int split;             /* complex only: 1 split, 0 interleaved */
    int markings;
    vsip_length size;      /* elements (complex elements for complex blocks) */
    void *user_r;          /* user data: real / interleaved / split real */
    void *user_i;          /* user data: split imaginary, else NULL */
    void *dev_r;           /* device data, same layout as the user data */
    void *dev_i;
    vsip_memory_hint hint;
    /* derived blocks (the real and imaginary parts of a complex block, as
     * vsip_mrealview/mimagview see them; vsip_cblockattributes_f.h:42-67 R
     * and I): element e of the derived block is scalar e * rstride of
     * user_r/dev_r, which the parent keeps pointing into its own data */
    int rstride;                   /* 0 or 1: unit; 2: part of an interleaved block */
    struct VA_block *parent;       /* derived blocks: the complex block */
    struct VA_block *part[2];      /* complex blocks: real and imaginary parts, made on demand */
} VA_block;

/* scalar stride of a block's elements within its arrays */
#define VA_RS(b) ((b)->rstride > 1 ? (long)(b)->rstride : 1L)

typedef struct VA_mview {
    VA_block *block;
    vsip_offset offset;
    vsip_stride row_stride;
    vsip_length row_length;
    vsip_stride col_stride;
    vsip_length col_length;
    int markings;
} VA_mview;

struct vsip_blockattributes_f { VA_block b; };
struct vsip_blockattributes_d { VA_block b; };
struct vsip_cblockattributes_f { VA_block b; };
struct vsip_cblockattributes_d { VA_block b; };
struct vsip_mviewattributes_f { VA_mview v; };
struct vsip_mviewattributes_d { VA_mview v; };
struct vsip_cmviewattributes_f { VA_mview v; };
struct vsip_cmviewattributes_d { VA_mview v; };
/* a vector view is the 1 x length matrix view (offset, stride, length) */
struct vsip_vviewattributes_f { VA_mview v; };
struct vsip_vviewattributes_d { VA_mview v; };
struct vsip_cvviewattributes_f { VA_mview v; };
struct vsip_cvviewattributes_d { VA_mview v; };

/* ---- FFT schedules ------------------------------------------------------ */

/* launch shape of one schedule for one lane-to-data mapping */
typedef struct {
    int vpb;        /* vectors per workgroup */
    int threads;    /* workgroup size */
    int maxe;       /* register budget class */
} VA_launch;

/* a single-pass transform of length n */
typedef struct {
    int n;
    int nst;
    int radix[VHIP_MAX_STAGES];
    int tw_off[VHIP_MAX_STAGES];
    int tw_len;              /* complex entries of the stage twiddle table */
    int btw_off;             /* the batched kernel's two-level table in dev_tw, or -1 */
    int lds_stride;          /* padded LDS vector stride */
    VA_launch along;         /* for unit element stride (rows) */
    VA_launch across;        /* for unit vector stride (columns) */
    void *dev_tw;            /* device twiddles */
} VA_sched;

#define VA_CCFFTIP 0
#define VA_CCFFTOP 1
#define VA_RCFFTOP 2
#define VA_CRFFTOP 3

typedef struct VA_fftm {
    int prec;
    int type;                /* VA_CCFFTIP ... VA_CRFFTOP */
    vsip_length M, N;        /* matrix shape given at create */
    vsip_length MN;          /* complex transform length along the major axis */
    vsip_length mN;          /* number of transforms */
    vsip_major major;
    int dir;                 /* -1 forward, +1 inverse */
    double scale;
    unsigned int ntimes;
    vsip_alg_hint hint;
    int npass;               /* 1 single pass, 2 or 3 split passes */
    vsip_length n1, n2, n3;  /* split factors, MN = n1 * n2 (* n3) */
    VA_sched pass[3];        /* npass == 1: pass[0] of length MN;
                                npass == 2: lengths n1, n2; npass == 3: n1, n2, n3 */
    void *dev_euler;         /* real transforms: e^{-/+ i pi k/MN}, k = 0..MN */
    void *dev_ptw;           /* split: e^{-2 pi i m / MN}, m = 0..MN-1 */
    void *dev_ptw2;          /* three passes: e^{-2 pi i m / (n2 n3)} */
    void *dev_work;          /* split: mN * MN complex scratch */
    void *dev_work2;         /* three passes: a second one */
    /* staged transforms (va_stage.c): PRE pass view -> scratch rows of length
     * sm, a forward c2c plan on them (twice, with the chirp filter product
     * between, for Bluestein), POST pass scratch -> view; `chunk` vectors per
     * slab */
    int jit;                 /* 1: schedules compiled at creation may run this plan */
    int jit_worth;           /* 1: the plan's uses repay compiling (also fused launches' schedules) */
    int staged;              /* 0 none, 1 real transform around a long c2c, 2 Bluestein */
    vsip_length sm;
    vsip_length chunk;
    struct VA_fftm *inner;   /* forward ccfftmip, ROW, chunk x sm */
    void *dev_chirp;         /* Bluestein: e^{-i pi k^2 / MN}, k < MN */
    void *dev_filt;          /* Bluestein: sm-point DFT of the conjugate chirp, / sm */
    void *dev_stage;         /* chunk * sm complex scratch */
    int markings;
} VA_fftm;

struct vsip_fftmattributes_f { VA_fftm p; };
struct vsip_fftmattributes_d { VA_fftm p; };
/* a 1-D FFT is the multiple FFT of one ROW vector */
struct vsip_fftattributes_f { VA_fftm p; };
struct vsip_fftattributes_d { VA_fftm p; };

/* Workgroup shape of the batched (persistent) kernel: every transform of a
 * batch runs with these lanes, register class and LDS bytes per workgroup;
 * single-pass lengths up to 8192 fit, longer ones run in two passes.  The
 * LDS size is what 8192/n rows of every power-of-two n from 256 to 8192 need
 * (the largest is 32 rows of 256, padded), two workgroups per CU. */
#define VA_BATCH_THREADS 512
#define VA_BATCH_MAXE 16
#define VA_BATCH_LDS 68480

/* launch descriptors of one plan execution: up to three passes */
#define VA_MAX_DESC 3

/* ---- shared helpers ------------------------------------------------------ */
int va_ensure_init(void);
void va_error(const char *where, const char *what, int code);

/* blocks */
VA_block *va_block_bind(int prec, int cplx, void *r, void *i, vsip_length n, vsip_memory_hint h);
VA_block *va_block_create(int prec, int cplx, vsip_length n, vsip_memory_hint h);
int va_block_admit(VA_block *b, int update);
int va_block_release(VA_block *b, int update, void **r, void **i);
void va_block_destroy(VA_block *b);
VA_block *va_block_part(VA_block *b, int imag);

/* views */
VA_mview *va_mview_bind(VA_block *b, vsip_offset o, vsip_stride cs, vsip_length cl,
                        vsip_stride rs, vsip_length rl);
void va_mview_destroy(VA_mview *v);
int va_elem_dev(const VA_mview *v, vsip_index i, vsip_index j, void **r, void **im);
int va_mview_copy_host(const VA_mview *v, int to_device, void *r, void *im, vsip_major major);
void va_get(const VA_mview *v, vsip_index i, vsip_index j, void *re, void *im);
void va_view_mat(const VA_mview *v, int dev, vhip_mat *m);
void va_put(const VA_mview *v, vsip_index i, vsip_index j, const void *re, const void *im);

/* planning and execution */
int va_sched_shape(VA_sched *s, int prec, int n, vsip_length nvec, int lds_max);
int va_sched_plan(VA_sched *s, int prec, int n, vsip_length nvec);
int va_sched_shape_fixed(VA_sched *s, int prec, int n, const int *radix, int nst,
                         int threads, int maxe, int lds_max, int vpb_max);
int va_sched_upload(VA_sched *s, int prec);
int va_factor_variant(int n, int prec, int maxbits, int odd_first, int *radix);
int va_factor_pack(int n, int *radix);
int va_tune_get(int prec, int n, const VA_sched **out);
int va_tune_lookup(int prec, int n, const VA_sched **out);
void va_sched_free(VA_sched *s);
int va_fftm_init(VA_fftm *p, int prec, int type, vsip_length M, vsip_length N, double scale,
                 int dir, vsip_major major, unsigned int ntimes, vsip_alg_hint hint);
int va_fftm_init_inner(VA_fftm *p, int prec, vsip_length M, vsip_length N);
void va_fftm_release(VA_fftm *p);
int va_jit_ready(const VA_fftm *p);
int va_ptw_shift(long n);
int va_ptw_upload(void **dev, int prec, long n);
int va_fftm_exec(const VA_fftm *p, const VA_mview *in, const VA_mview *out, void *stream);
int va_staged_init(VA_fftm *p, int bluestein);
void va_staged_release(VA_fftm *p);
int va_staged_exec(const VA_fftm *p, const VA_mview *in, const VA_mview *out, void *stream);
int va_exec_views(const VA_fftm *p, const VA_mview *in, const VA_mview *out, vhip_view *gi,
                  vhip_view *go, long *es_i, long *vs_i, long *es_o, long *vs_o);
long va_bluestein_length(long n);
void va_host_fft(long n, double *x);
int va_fftm_descs(const VA_fftm *p, const VA_mview *in, const VA_mview *out, const VA_sched *sA,
                  const VA_sched *sB, void *work, const void *ptw, vhip_fftm_desc *d);

#endif /* VA_INTERNAL_H */
