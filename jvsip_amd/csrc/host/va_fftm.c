/*
 * va_fftm.c -- the vsip_fftm_{f,d} plan objects and the multiple-FFT entry
 * points vsip_ccfftmip, vsip_ccfftmop, vsip_rcfftmop, vsip_crfftmop.
 *
 * Reference behaviour kept (file:line in rrjudd/jvsip c_VSIP_src/):
 *   plan fields / lengths   VI_ccfftmip_create_f_def.h:35-39 (MN = N for ROW,
 *                           M for COL; mN the other), VI_rcfftmop_create_f_def.h:38-40
 *                           (real: MN = N/2 or M/2)
 *   ccfftmip                VI_ccfftmip_f_def.h:17-23, VI_fftm_building_blocks_f.h:1064-1088
 *                           (DFT with sign dir, then multiply by scale)
 *   ccfftmop                VI_ccfftmop_f_def.h:20-29 (copy x -> y, transform y)
 *   rcfftmop                VI_rcfftmop_f_def.h:24-120 (output MN+1 per vector)
 *   crfftmop                VI_crfftmop_f_def.h:25-126 (unnormalised inverse,
 *                           times scale)
 *   getattr                 vsip_fftm_getattr_f.c:18-69
 *   destroy                 vsip_fftm_destroy_f.c:20-40 (returns 0)
 * create returns NULL on allocation failure as the reference does
 * (VI_ccfftmip_create_f_def.h:34,48-53), and also for lengths this
 * implementation does not transform: odd N for the real transforms
 * (undefined in the reference, which truncates N/2) and lengths with a prime
 * factor above 13.  Length 1 (undefined in the reference, whose factoriser
 * then reads pF[-1]) is the identity times scale here.
 *
 * Execution is one kernel launch per call on the library stream
 * (two for lengths beyond one workgroup's LDS, split as MN = n1 * n2 with
 * the twiddle e^{-2 pi i n2 k1 / MN} applied between the passes).  The
 * plan holds no per-call state, so a plan may be used from several threads.
 *
 * C89.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "va_internal.h"

#define VA_PI 3.14159265358979323846264338327950288

static int va_upload_table(void **dev, int prec, long count, int sign, double denom, double mul)
{
    /* entries e^{sign i * mul * pi * k / denom}, k = 0..count-1 */
    size_t esz = (size_t)(2 * prec);
    char *host = (char *)malloc(esz * (size_t)count);
    long k;
    int rc;
    if (!host) return -1;
    for (k = 0; k < count; k++) {
        double a = (double)sign * mul * VA_PI * (double)k / denom;
        double c = cos(a), s = sin(a);
        if (prec == 4) {
            ((float *)host)[2 * k] = (float)c;
            ((float *)host)[2 * k + 1] = (float)s;
        } else {
            ((double *)host)[2 * k] = c;
            ((double *)host)[2 * k + 1] = s;
        }
    }
    rc = vhip_malloc(dev, esz * (size_t)count);
    if (!rc) rc = vhip_memcpy_h2d(*dev, host, esz * (size_t)count);
    free(host);
    if (rc && *dev) { vhip_free(*dev); *dev = NULL; }
    return rc;
}

/* low-table size of the split's post-twiddle for modulus n: 2^ceil(log2(n)/2) */
int va_ptw_shift(long n)
{
    int b = 0;
    while ((1L << b) < n) b++;
    return (b + 1) / 2;
}

/* the post-twiddle e^{-2 pi i t / n}, t < n, as two tables whose product
 * gives it (vhip_fftm_desc.ptw): lo[r] = w(r), r < L; hi[q] = w(q L),
 * q < ceil(n / L); angles from exact integer residues, in double */
int va_ptw_upload(void **dev, int prec, long n)
{
    const int sh = va_ptw_shift(n);
    const long L = 1L << sh, H = (n + L - 1) / L;
    size_t esz = (size_t)(2 * prec), cnt = (size_t)(L + H);
    char *host = (char *)malloc(esz * cnt);
    long k;
    int rc;
    if (!host) return -1;
    for (k = 0; k < L + H; k++) {
        long t = k < L ? k : ((k - L) * L) % n;
        double a = -2.0 * VA_PI * (double)t / (double)n;
        if (prec == 4) {
            ((float *)host)[2 * k] = (float)cos(a);
            ((float *)host)[2 * k + 1] = (float)sin(a);
        } else {
            ((double *)host)[2 * k] = cos(a);
            ((double *)host)[2 * k + 1] = sin(a);
        }
    }
    *dev = NULL;
    rc = vhip_malloc(dev, esz * cnt);
    if (!rc) rc = vhip_memcpy_h2d(*dev, host, esz * cnt);
    free(host);
    if (rc && *dev) { vhip_free(*dev); *dev = NULL; }
    return rc;
}

void va_fftm_release(VA_fftm *p)
{
    if (!p) return;
    vhip_sync();
    va_sched_free(&p->pass[0]);
    va_sched_free(&p->pass[1]);
    va_sched_free(&p->pass[2]);
    va_staged_release(p);
    if (p->dev_euler) vhip_free(p->dev_euler);
    if (p->dev_ptw) vhip_free(p->dev_ptw);
    if (p->dev_ptw2) vhip_free(p->dev_ptw2);
    if (p->dev_work) vhip_free(p->dev_work);
    if (p->dev_work2) vhip_free(p->dev_work2);
    p->dev_euler = p->dev_ptw = p->dev_ptw2 = p->dev_work = p->dev_work2 = NULL;
    p->markings = VA_FREED;
}

/* choose MN = n1 * n2 with both factors single-pass; returns 0 on success */
static int va_plan_split(VA_fftm *p)
{
    long MN = (long)p->MN, d, best = 0;
    long r = (long)sqrt((double)MN);
    for (d = r + 1; d >= 2; d--) {
        VA_sched a, b;
        if (MN % d) continue;
        if (va_sched_plan(&a, p->prec, (int)(MN / d), p->mN * (vsip_length)d)) { va_sched_free(&a); continue; }
        if (va_sched_plan(&b, p->prec, (int)d, p->mN * (vsip_length)(MN / d))) {
            va_sched_free(&a);
            va_sched_free(&b);
            continue;
        }
        /* pass A: length n1 = MN / d over n2 = d interleaved vectors; pass B: length n2 */
        p->pass[0] = a;
        p->pass[1] = b;
        p->n1 = (vsip_length)(MN / d);
        p->n2 = (vsip_length)d;
        best = d;
        break;
    }
    return best ? 0 : -1;
}

/* three passes for lengths from VSIP_AMD_SPLIT3_MIN (default 2^21): every
 * factor small enough that the passes' column-mapped sides take 16 vectors
 * per workgroup (128-byte pieces of each line); the factors as even as the
 * length allows, the largest last (the pass that reads contiguous rows) */
static int va_plan_split3(VA_fftm *p)
{
    const char *e = getenv("VSIP_AMD_SPLIT3_MIN");
    long MN = (long)p->MN, lim = e ? atol(e) : (1L << 21), a, b, best = -1, ba = 0, bb = 0;
    int lds_max = 65536;
    if (MN < lim || lim <= 0) return -1;
    vhip_lds_max(&lds_max);
    for (a = 2; a * a * a <= MN; a++) {
        long rest;
        if (MN % a) continue;
        rest = MN / a;
        for (b = a; b * b <= rest; b++) {
            long c;
            VA_sched t;
            if (rest % b) continue;
            c = rest / b;                   /* a <= b <= c */
            if (c > 1024 || (best >= 0 && c >= best)) continue;
            memset(&t, 0, sizeof(t));
            if (va_sched_shape(&t, p->prec, (int)a, p->mN, lds_max) ||
                va_sched_shape(&t, p->prec, (int)b, p->mN, lds_max) ||
                va_sched_shape(&t, p->prec, (int)c, p->mN, lds_max))
                continue;                   /* a factor without a schedule */
            best = c; ba = a; bb = b;
        }
    }
    if (best < 0) return -1;
    {
        VA_sched s[3];
        long f[3];
        int k;
        f[0] = ba; f[1] = bb; f[2] = MN / (ba * bb);
        memset(s, 0, sizeof(s));
        for (k = 0; k < 3; k++) {
            vsip_length others = p->mN * (vsip_length)(MN / f[k]);
            if (va_sched_plan(&s[k], p->prec, (int)f[k], others)) {
                int q;
                for (q = 0; q <= k; q++) va_sched_free(&s[q]);
                return -1;
            }
        }
        for (k = 0; k < 3; k++) p->pass[k] = s[k];
        p->n1 = (vsip_length)f[0];
        p->n2 = (vsip_length)f[1];
        p->n3 = (vsip_length)f[2];
    }
    return 0;
}

/* the compiled kernels hold their vectors and offsets in static LDS */
static int va_jit_lds_ok(const VA_sched *s, const VA_launch *L, int prec)
{
    int lds_max = 65536;
    vhip_lds_max(&lds_max);
    return (long)L->vpb * s->lds_stride * 2 * prec + (long)L->vpb * 12 <= lds_max;
}

/* Does a plan's expected use repay compiling its schedules?  A compile
 * takes 0.3-1 s once per schedule and process (or one file read with
 * VSIP_AMD_JIT_CACHE_DIR); a compiled kernel saves about half of a small
 * launch (MI355X: 16 x 1000 rows 7.3 -> 4.0 us, 1 x 1000 6.8 -> 3.6 us, the
 * 1-D 2^16 split 21.7 -> 10.5 us; the generic kernel's launch takes ~12x
 * the instruction-cache misses, its paths being chosen at run time).  VSIPL
 * gives the expected number of uses at creation: ntimes 0 stands for an
 * unbounded number, and such plans compile from 2^12 elements; a plan with
 * ntimes > 0 compiles when ntimes x elements >= 2^30.  VSIP_AMD_JIT_MIN_ELEMS
 * replaces the rule by a plain floor on elements. */
static int va_jit_worth(const VA_fftm *p)
{
    const char *e = getenv("VSIP_AMD_JIT_MIN_ELEMS");
    const long elems = (long)p->mN * (long)p->MN;
    if (e) return elems >= atol(e);
    if (elems < (1L << 12)) return 0;
    if (p->ntimes == 0) return 1;
    return (double)p->ntimes * (double)elems >= (double)(1L << 30);
}

/* Complex, real-to-complex and complex-to-real plans (FP32, FP64) whose passes the
 * schedule-generic kernel would run (every pass except the FP32
 * power-of-two kernels' contiguous rows; complex rows of 256..2048 in
 * batches too small for the pipelined kernel compile as well) and whose use
 * repays it
 * (va_jit_worth) get their schedules compiled now, for the layouts the
 * plan's views normally have: contiguous rows (ROW), lanes across columns
 * (COL), and the passes of the splits.  Execution finds the compiled kernel
 * when the layout matches. */
static void va_jit_prepare(VA_fftm *p)
{
    long n = (long)p->MN;
    int k, real = (p->type == VA_RCFFTOP) ? VHIP_JIT_REAL : (p->type == VA_CRFFTOP) ? VHIP_JIT_C2R : 0;
    if (p->staged) return;
    if (!va_jit_worth(p)) return;
    if (getenv("VSIP_AMD_JIT") && getenv("VSIP_AMD_JIT")[0] == '0') return;
    p->jit_worth = 1;
    if (p->npass == 1) {
        const VA_sched *s = &p->pass[0];
        if (p->major == VSIP_ROW) {
            /* the power-of-two kernels, except rows of 256..2048 complex
             * points (c2c, or the half length of r2c and c2r) in batches
             * below the pipelined kernels': there the compiled schedule beats
             * the one-group kernel (MI355X, per launch: 64 x 1024 5.4 -> 3.5 us,
             * 4096 x 1024 14.2 -> 12.0, 128 x 512 5.2 -> 3.0, 2048 x 2048
             * 13.8 -> 11.9; from 4096 points even) */
            if (p->prec == 4 && (n & (n - 1)) == 0 && n >= 256 && n <= 16384 && !getenv("VSIP_AMD_NO_POW2") &&
                (n > 2048 || vhip_p2_pipelined((int)n, (long)p->mN)))
                return;                                   /* p->jit stays 0: never a compiled kernel */
            p->jit = 1;
            if (va_jit_lds_ok(s, &s->along, p->prec))
                vhip_jit_prepare(p->prec, s->n, s->nst, s->radix, s->along.vpb, s->along.threads,
                                 VHIP_JIT_CONTIG | real, VHIP_JIT_CONTIG);
        } else {
            p->jit = 1;
            if (va_jit_lds_ok(s, &s->across, p->prec))
                vhip_jit_prepare(p->prec, s->n, s->nst, s->radix, s->across.vpb, s->across.threads,
                                 VHIP_JIT_ACROSS | real, VHIP_JIT_ACROSS);
        }
        return;
    }
    p->jit = 1;
    /* split: every pass maps lanes across vectors at its global sides except
     * the last one's load, which reads scratch rows */
    for (k = 0; k < p->npass; k++) {
        const VA_sched *a = &p->pass[k];
        if (va_jit_lds_ok(a, &a->across, p->prec))
            vhip_jit_prepare(p->prec, a->n, a->nst, a->radix, a->across.vpb, a->across.threads,
                             k == p->npass - 1 ? VHIP_JIT_CONTIG : VHIP_JIT_ACROSS,
                             k == p->npass - 1 ? VHIP_JIT_ACROSS : VHIP_JIT_ACROSS | VHIP_JIT_PTW);
    }
}

/* does the plan have compiled kernels for its expected layouts */
int va_jit_ready(const VA_fftm *p)
{
    const VA_sched *s = &p->pass[0];
    int k, real = (p->type == VA_RCFFTOP) ? VHIP_JIT_REAL : (p->type == VA_CRFFTOP) ? VHIP_JIT_C2R : 0;
    if (p->npass == 1) {
        if (p->major == VSIP_ROW)
            return vhip_jit_ready(p->prec, s->n, s->nst, s->radix, s->along.vpb, s->along.threads,
                                  VHIP_JIT_CONTIG | real, VHIP_JIT_CONTIG);
        return vhip_jit_ready(p->prec, s->n, s->nst, s->radix, s->across.vpb, s->across.threads,
                              VHIP_JIT_ACROSS | real, VHIP_JIT_ACROSS);
    }
    if (p->npass < 2) return 0;
    for (k = 0; k < p->npass; k++) {
        const VA_sched *a = &p->pass[k];
        if (!vhip_jit_ready(p->prec, a->n, a->nst, a->radix, a->across.vpb, a->across.threads,
                            k == p->npass - 1 ? VHIP_JIT_CONTIG : VHIP_JIT_ACROSS,
                            k == p->npass - 1 ? VHIP_JIT_ACROSS : VHIP_JIT_ACROSS | VHIP_JIT_PTW))
            return 0;
    }
    return 1;
}

/* allow_staged == 0 gives up where a staged transform would be needed (the
 * inner plans of staged transforms, whose lengths are always smooth); 2 makes
 * a staged plan for every length (views the kernels cannot address directly) */
static int va_fftm_init2(VA_fftm *p, int prec, int type, vsip_length M, vsip_length N, double scale,
                         int dir, vsip_major major, unsigned int ntimes, vsip_alg_hint hint,
                         int allow_staged)
{
    int rc, real = (type == VA_RCFFTOP || type == VA_CRFFTOP);
    memset(p, 0, sizeof(*p));
    if (va_ensure_init()) return -1;
    p->prec = prec;
    p->type = type;
    p->M = M;
    p->N = N;
    p->major = major;
    p->scale = scale;
    p->dir = dir;
    p->ntimes = ntimes;
    p->hint = hint;
    p->markings = VA_VALID;
    if (real) {
        vsip_length L = (major == VSIP_ROW) ? N : M;
        if (L < 2 || (L % 2) != 0) return -1;
        p->MN = L / 2;
        p->mN = (major == VSIP_ROW) ? M : N;
    } else {
        p->MN = (major == VSIP_COL) ? M : N;
        p->mN = (major == VSIP_COL) ? N : M;
    }
    if (p->MN < 1) return -1;
    if (p->MN > 0x7fffffffUL) return -1;
    p->npass = 1;
    rc = va_sched_plan(&p->pass[0], prec, (int)p->MN, p->mN);
    if (allow_staged == 2 && (rc == 0 || rc == -2)) {
        va_sched_free(&p->pass[0]);
        rc = -4;                      /* staged around the plain schedule */
    }
    if (rc == -2 && !real && !va_plan_split3(p)) {
        long n23 = (long)(p->n2 * p->n3);
        p->npass = 3;
        if (va_ptw_upload(&p->dev_ptw, prec, (long)p->MN)) goto fail;
        if (va_ptw_upload(&p->dev_ptw2, prec, n23)) goto fail;
        if (vhip_malloc(&p->dev_work, (size_t)p->mN * p->MN * 2 * (size_t)prec)) goto fail;
        if (vhip_malloc(&p->dev_work2, (size_t)p->mN * p->MN * 2 * (size_t)prec)) goto fail;
    } else if (rc == -2 && !real && !va_plan_split(p)) {
        p->npass = 2;
        if (va_ptw_upload(&p->dev_ptw, prec, (long)p->MN)) goto fail;
        if (vhip_malloc(&p->dev_work, (size_t)p->mN * p->MN * 2 * (size_t)prec)) goto fail;
    } else if (allow_staged && (rc == -1 || rc == -2 || rc == -4)) {
        /* a prime factor without a butterfly (-1), or a length beyond one
         * workgroup that the split cannot serve: staged transform; Bluestein
         * unless only the real transform's length is the obstacle (or the
         * staging is asked for, -4) */
        int blue = (rc == -1) || (rc == -2 && !real);
        va_sched_free(&p->pass[0]);
        memset(&p->pass[0], 0, sizeof(p->pass[0]));
        p->npass = 0;
        if (blue && va_staged_init(p, 1)) goto fail;
        if (!blue && va_staged_init(p, 0)) goto fail;
    } else if (rc) {
        va_sched_free(&p->pass[0]);
        return -1;
    }
    va_jit_prepare(p);
    if (real) {
        /* rcfftmop: e^{-i pi k / MN}; crfftmop: e^{+i pi k / MN}; k = 0..MN */
        int sign = (type == VA_RCFFTOP) ? -1 : 1;
        if (va_upload_table(&p->dev_euler, prec, (long)p->MN + 1, sign, (double)p->MN, 1.0)) goto fail;
    }
    return 0;
fail:
    va_fftm_release(p);
    return -1;
}

int va_fftm_init(VA_fftm *p, int prec, int type, vsip_length M, vsip_length N, double scale,
                 int dir, vsip_major major, unsigned int ntimes, vsip_alg_hint hint)
{
    return va_fftm_init2(p, prec, type, M, N, scale, dir, major, ntimes, hint, 1);
}

int va_fftm_init_inner(VA_fftm *p, int prec, vsip_length M, vsip_length N)
{
    return va_fftm_init2(p, prec, VA_CCFFTIP, M, N, 1.0, VSIP_FFT_FWD, VSIP_ROW, 0, VSIP_ALG_TIME, 0);
}

/* transform axis of a view: element stride, vector stride, length, count */
static void va_axis(const VA_mview *v, vsip_major major, long *es, long *vs,
                    vsip_length *n, vsip_length *nvec)
{
    /* strides in scalars of the block's arrays: x VA_RS for the real or
     * imaginary part of an interleaved complex block */
    const long rs = VA_RS(v->block);
    if (major == VSIP_ROW) {
        *es = v->row_stride * rs; *vs = v->col_stride * rs; *n = v->row_length; *nvec = v->col_length;
    } else {
        *es = v->col_stride * rs; *vs = v->row_stride * rs; *n = v->col_length; *nvec = v->row_length;
    }
}

static int va_view_dev(const VA_mview *v, vhip_view *g)
{
    VA_block *b = v->block;
    size_t p;
    if (!b || !b->admitted || !b->dev_r) return -1;
    p = (size_t)b->prec;
    memset(g, 0, sizeof(*g));
    if (b->cplx && !b->split) {
        g->re = (char *)b->dev_r + (size_t)v->offset * 2 * p;
    } else if (b->cplx) {
        g->re = (char *)b->dev_r + (size_t)v->offset * p;
        g->im = (char *)b->dev_i + (size_t)v->offset * p;
        g->split = 1;
    } else {
        g->re = (char *)b->dev_r + (size_t)v->offset * (size_t)VA_RS(b) * p;
    }
    return 0;
}

static long va_labs(long x) { return x < 0 ? -x : x; }

/* bytes from the view base covering `nvec` vectors of `len` elements (of
 * `esz` bytes); 0 when a stride is negative or the span reaches 2 GiB, so the
 * kernels fall back to 64-bit addressing */
static long va_span(const vhip_view *g, long nvec, long len, long esz)
{
    long nv0 = g->nv0 > 0 ? g->nv0 : nvec, last;
    if (g->es < 0 || g->vs0 < 0 || g->vs1 < 0 || nvec < 1 || len < 1) return 0;
    last = (nvec > nv0 ? nv0 - 1 : nvec - 1) * g->vs0 + ((nvec - 1) / nv0) * g->vs1 + (len - 1) * g->es;
    if (last < 0 || (last + 1) > ((1L << 31) - 64) / esz) return 0;
    return (last + 1) * esz;
}

/* element size of a view's arrays: complex interleaved 2*prec, split or real prec */
static long va_esz(const vhip_view *g, int prec, int cplx)
{
    return (cplx && !g->split) ? 2L * prec : (long)prec;
}

static int va_map(long es, long vs, vsip_length nvec)
{
    return (nvec > 1 && va_labs(vs) < va_labs(es)) ? VHIP_MAP_ACROSS : VHIP_MAP_ALONG;
}

static void va_desc_common(vhip_fftm_desc *d, const VA_fftm *p, const VA_sched *s, int kind,
                           long nvec, int map_in, int map_out)
{
    const VA_launch *L = (map_in == VHIP_MAP_ACROSS || map_out == VHIP_MAP_ACROSS)
                             ? &s->across : &s->along;
    int i;
    memset(d, 0, sizeof(*d));
    d->prec = p->prec;
    d->kind = kind;
    d->n = s->n;
    d->nst = s->nst;
    for (i = 0; i < s->nst; i++) {
        d->radix[i] = s->radix[i];
        d->tw_off[i] = s->tw_off[i];
    }
    d->nvec = nvec;
    d->vpb = L->vpb;
    d->threads = L->threads;
    d->lds_stride = s->lds_stride;
    d->maxe = L->maxe;
    d->map_in = map_in;
    d->map_out = map_out;
    d->tw = s->dev_tw;
    d->btw_off = s->btw_off;
    d->scale = 1.0;
    d->compiled = p->jit;
}

/*
 * Launch descriptors of one execution of plan p on (in, out): one for a
 * single pass, two for the four-step split.  sA/sB, work and ptw replace the
 * plan's own schedules, scratch and twiddle table when given (the batched
 * launch uses schedules tuned for its fixed workgroup size); sB != NULL
 * selects the split with n1 = sA->n, n2 = sB->n.  Returns the descriptor
 * count, or -1 after reporting an error.
 */
/*
 * Check that (in, out) fit plan p and give their device views with the
 * transform axis' element and vector strides.  Returns the transform kind,
 * or -1 after reporting an error.
 */
int va_exec_views(const VA_fftm *p, const VA_mview *in, const VA_mview *out, vhip_view *gi,
                  vhip_view *go, long *es_i, long *vs_i, long *es_o, long *vs_o)
{
    vsip_length n_i, nv_i, n_o, nv_o, want_in, want_out;
    int kind;
    if (!p || p->markings != VA_VALID) { va_error("fftm", "invalid plan", -1); return -1; }
    if (!in || !out) { va_error("fftm", "NULL view", -1); return -1; }
    if (in->block->prec != p->prec || out->block->prec != p->prec) {
        va_error("fftm", "view precision does not match the plan", -1);
        return -1;
    }
    va_axis(in, p->major, es_i, vs_i, &n_i, &nv_i);
    va_axis(out, p->major, es_o, vs_o, &n_o, &nv_o);
    kind = (p->type == VA_RCFFTOP) ? VHIP_R2C : (p->type == VA_CRFFTOP) ? VHIP_C2R : VHIP_C2C;
    want_in = (kind == VHIP_R2C) ? 2 * p->MN : (kind == VHIP_C2R) ? p->MN + 1 : p->MN;
    want_out = (kind == VHIP_R2C) ? p->MN + 1 : (kind == VHIP_C2R) ? 2 * p->MN : p->MN;
    if (n_i < want_in || nv_i != p->mN || n_o < want_out || nv_o != p->mN) {
        va_error("fftm", "view shape does not match the plan", -1);
        return -1;
    }
    if (va_view_dev(in, gi) || va_view_dev(out, go)) {
        va_error("fftm", "block is not admitted", -1);
        return -1;
    }
    return kind;
}

/* rows that a launch may stream past the caches (vhip_fftm_desc.stream): a
 * contiguous, single-level, interleaved side whose rows are dense
 * (vector stride = length) and start on 128-byte lines, so no line is
 * shared with data the launch does not write.  Used for the scratch rows
 * that the last pass of the three-pass split reads (2^24: 245 -> 197 us);
 * dense single-pass rows measured slower streamed on the compiled kernels
 * (FP64 2048 x 2048: 25.1 -> 28.4 us) and keep the default policy. */
static int va_dense_rows(const vhip_view *g, long n, long nvec, int prec)
{
    const long esz = 2L * prec;
    if (!g->contig || g->split || g->es != 1 || g->nv0 < nvec) return 0;
    if (nvec > 1 && g->vs0 != n) return 0;
    return ((size_t)g->re % 128 == 0) && ((n * esz) % 128 == 0);
}

/*
 * The three-pass split, MN = n1 n2 n3, input element j = j1 n2 n3 + j2 n3 + j3
 * and output element k = k1 + n1 k2 + n1 n2 k3, on scratch T1, T2 (mN x MN):
 *   pass 1  length n1 over j1 for each (row, m = j2 n3 + j3), times
 *           e^{-2 pi i m k1 / MN}; T1[row][k1][m]
 *   pass 2  length n2 over j2 for each (row, k1, j3), times
 *           e^{-2 pi i j3 k2 / (n2 n3)}; T2[row][k2][k1][j3]
 *   pass 3  length n3 over j3 for each (row, k2, k1) -- T2's rows of n3 in
 *           order -- into X[row][k1 + n1 k2 + n1 n2 k3]
 * Vectors are numbered so that consecutive ones are adjacent in memory on
 * every column-mapped side (unit vector stride): j3 fastest in passes 1 and
 * 2, k1 fastest in pass 3.
 */
static int va_descs3(const VA_fftm *p, const vhip_view *gin, const vhip_view *gout, long es_i, long vs_i,
                     long es_o, long vs_o, int inv, vhip_fftm_desc *d)
{
    long n1 = (long)p->n1, n2 = (long)p->n2, n3 = (long)p->n3, MN = (long)p->MN, mN = (long)p->mN;
    long n23 = n2 * n3;
    vhip_view gi = *gin, go = *gout, t1, t2, u1, u2;
    memset(&t1, 0, sizeof(t1));
    memset(&t2, 0, sizeof(t2));
    t1.re = p->dev_work;
    t2.re = p->dev_work2;
    u1 = t1;
    u2 = t2;
    /* pass 1 */
    gi.es = n23 * es_i; gi.vs0 = es_i; gi.nv0 = n23; gi.vs1 = vs_i; gi.contig = 0;
    gi.span = va_span(&gi, mN * n23, n1, va_esz(&gi, p->prec, 1));
    t1.es = n23; t1.vs0 = 1; t1.nv0 = n23; t1.vs1 = MN;
    t1.span = va_span(&t1, mN * n23, n1, va_esz(&t1, p->prec, 1));
    va_desc_common(&d[0], p, &p->pass[0], VHIP_C2C, mN * n23, VHIP_MAP_ACROSS, VHIP_MAP_ACROSS);
    d[0].in = gi;
    d[0].out = t1;
    d[0].conj_in = inv;
    d[0].scale = 1.0;
    d[0].ptw = p->dev_ptw;
    d[0].ptw_n = n23;                       /* m */
    d[0].ptw_shift = va_ptw_shift(MN);
    /* pass 2: vector v = j3 + n3 (k1 + n1 row) */
    u1.es = n3; u1.vs0 = 1; u1.nv0 = n3; u1.vs1 = n23;
    u1.span = va_span(&u1, mN * n1 * n3, n2, va_esz(&u1, p->prec, 1));
    t2.es = n1 * n3; t2.vs0 = 1; t2.nv0 = n1 * n3; t2.vs1 = MN;
    t2.span = va_span(&t2, mN * n1 * n3, n2, va_esz(&t2, p->prec, 1));
    va_desc_common(&d[1], p, &p->pass[1], VHIP_C2C, mN * n1 * n3, VHIP_MAP_ACROSS, VHIP_MAP_ACROSS);
    d[1].in = u1;
    d[1].out = t2;
    d[1].scale = 1.0;
    d[1].ptw = p->dev_ptw2;
    d[1].ptw_n = n3;                        /* j3 */
    d[1].ptw_shift = va_ptw_shift(n23);
    /* pass 3: vector v = k1 + n1 (k2 + n2 row), T2 row v */
    u2.es = 1; u2.vs0 = n3; u2.nv0 = mN * n1 * n2; u2.vs1 = 0; u2.contig = 1;
    u2.span = va_span(&u2, mN * n1 * n2, n3, va_esz(&u2, p->prec, 1));
    go.es = n1 * n2 * es_o; go.vs0 = es_o; go.nv0 = n1 * n2; go.vs1 = vs_o; go.contig = 0;
    go.span = va_span(&go, mN * n1 * n2, n3, va_esz(&go, p->prec, 1));
    va_desc_common(&d[2], p, &p->pass[2], VHIP_C2C, mN * n1 * n2, VHIP_MAP_ALONG, VHIP_MAP_ACROSS);
    d[2].in = u2;
    d[2].out = go;
    d[2].conj_out = inv;
    d[2].scale = p->scale;
    d[2].stream = va_dense_rows(&u2, n3, mN * n1 * n2, p->prec) ? 1 : 0;
    if (!gi.span || !go.span || !t1.span || !t2.span || !u1.span || !u2.span) {
        va_error("fftm", "a view of one launch spans 2 GiB or more or has a negative stride", -1);
        return -1;
    }
    return 3;
}

int va_fftm_descs(const VA_fftm *p, const VA_mview *in, const VA_mview *out, const VA_sched *sA,
                  const VA_sched *sB, void *work, const void *ptw, vhip_fftm_desc *d)
{
    long es_i, vs_i, es_o, vs_o;
    vhip_view gi, go;
    int kind, inv, two;
    kind = va_exec_views(p, in, out, &gi, &go, &es_i, &vs_i, &es_o, &vs_o);
    if (kind < 0) return -1;
    if (p->staged) { va_error("fftm", "a staged plan has no single launch descriptor", -1); return -1; }
    inv = (kind == VHIP_C2C && p->dir > 0);
    if (!sA && p->npass == 3) {
        if (kind != VHIP_C2C) { va_error("fftm", "three-pass real transform", -1); return -1; }
        return va_descs3(p, &gi, &go, es_i, vs_i, es_o, vs_o, inv, d);
    }
    /* schedules given by the caller run in the batched kernel, whose stages
     * after the first read the two-level twiddle table of the upload */
    if (sA && ((sA->nst > 1 && sA->btw_off < 0) || (sB && sB->nst > 1 && sB->btw_off < 0))) {
        va_error("fftm", "batched schedule without its twiddle table", sA->n);
        return -1;
    }
    if (!sA) {
        sA = &p->pass[0];
        sB = (p->npass == 2) ? &p->pass[1] : NULL;
        work = p->dev_work;
        ptw = p->dev_ptw;
    }
    two = (sB != NULL);
    if (two && kind != VHIP_C2C) { va_error("fftm", "two-pass real transform", -1); return -1; }

    if (!two) {
        int map_in = va_map(es_i, vs_i, p->mN), map_out = va_map(es_o, vs_o, p->mN);
        gi.es = es_i; gi.vs0 = vs_i; gi.nv0 = (long)p->mN; gi.vs1 = 0;
        go.es = es_o; go.vs0 = vs_o; go.nv0 = (long)p->mN; go.vs1 = 0;
        if (kind == VHIP_R2C)
            gi.contig = (es_i == 1 && ((long)in->offset * VA_RS(in->block)) % 2 == 0 &&
                         (p->mN <= 1 || vs_i % 2 == 0));
        else
            gi.contig = (!gi.split && es_i == 1);
        if (kind == VHIP_C2R)
            go.contig = (es_o == 1 && ((long)out->offset * VA_RS(out->block)) % 2 == 0 &&
                         (p->mN <= 1 || vs_o % 2 == 0));
        else
            go.contig = (!go.split && es_o == 1);
        gi.span = va_span(&gi, (long)p->mN, (long)(kind == VHIP_R2C ? 2 * p->MN : kind == VHIP_C2R ? p->MN + 1 : p->MN),
                          va_esz(&gi, p->prec, kind != VHIP_R2C));
        go.span = va_span(&go, (long)p->mN, (long)(kind == VHIP_C2R ? 2 * p->MN : kind == VHIP_R2C ? p->MN + 1 : p->MN),
                          va_esz(&go, p->prec, kind != VHIP_C2R));
        if (!gi.span || !go.span) {
            va_error("fftm", "a view of one launch spans 2 GiB or more or has a negative stride", -1);
            return -1;
        }
        va_desc_common(&d[0], p, sA, kind, (long)p->mN, map_in, map_out);
        d[0].in = gi;
        d[0].out = go;
        d[0].conj_in = inv;
        d[0].conj_out = inv;
        d[0].scale = p->scale;
        d[0].euler = p->dev_euler;
        return 1;
    }

    /* four-step: MN = n1 * n2, element n = n2 * j1 + j2 */
    {
        long n1 = sA->n, n2 = sB->n, MN = (long)p->MN, mN = (long)p->mN;
        vhip_view t;
        if (n1 * n2 != MN) { va_error("fftm", "two-pass split does not match the length", -1); return -1; }
        memset(&t, 0, sizeof(t));
        t.re = work;
        /* pass A: vectors (row, j2), elements j1 at stride n2*es; out T[row][k1][j2] */
        gi.es = n2 * es_i; gi.vs0 = es_i; gi.nv0 = n2; gi.vs1 = vs_i;
        gi.contig = 0;
        gi.span = va_span(&gi, mN * n2, n1, va_esz(&gi, p->prec, 1));
        t.es = n2; t.vs0 = 1; t.nv0 = n2; t.vs1 = MN;
        t.span = va_span(&t, mN * n2, n1, va_esz(&t, p->prec, 1));
        va_desc_common(&d[0], p, sA, VHIP_C2C, mN * n2, VHIP_MAP_ACROSS, VHIP_MAP_ACROSS);
        d[0].in = gi;
        d[0].out = t;
        d[0].conj_in = inv;
        d[0].conj_out = 0;
        d[0].scale = 1.0;
        d[0].ptw = ptw;
        d[0].ptw_n = n2;                 /* row index j2 of the vector */
        d[0].ptw_shift = va_ptw_shift(MN);
        /* pass B: vectors (row, k1), elements j2 contiguous; out X[row][k1 + n1*k2] */
        /* vector (row, k1) starts at row * MN + k1 * n2 = v * n2: one level */
        t.es = 1; t.vs0 = n2; t.nv0 = mN * n1; t.vs1 = 0; t.contig = 1;
        t.span = va_span(&t, mN * n1, n2, va_esz(&t, p->prec, 1));
        go.es = n1 * es_o; go.vs0 = es_o; go.nv0 = n1; go.vs1 = vs_o;
        go.contig = 0;
        go.span = va_span(&go, mN * n1, n2, va_esz(&go, p->prec, 1));
        if (!gi.span || !go.span || !t.span) {
            va_error("fftm", "a view of one launch spans 2 GiB or more or has a negative stride", -1);
            return -1;
        }
        va_desc_common(&d[1], p, sB, VHIP_C2C, mN * n1, VHIP_MAP_ALONG, VHIP_MAP_ACROSS);
        d[1].in = t;
        d[1].out = go;
        d[1].conj_in = 0;
        d[1].conj_out = inv;
        d[1].scale = p->scale;
        return 2;
    }
}

/* vectors per launch that keep every view of the launch within 2 GiB of its
 * base (32-bit buffer offsets in the kernels); 0 when not even one fits */
static long va_exec_chunk(const VA_fftm *p, const VA_mview *in, const VA_mview *out)
{
    const long lim = (1L << 31) - 64;
    long chunk = (long)p->mN, es, vs, len, esz, c;
    vsip_length n, nv;
    int k;
    for (k = 0; k < 2; k++) {
        const VA_mview *v = k ? out : in;
        int real_side = (p->type == VA_RCFFTOP && k == 0) || (p->type == VA_CRFFTOP && k == 1);
        va_axis(v, p->major, &es, &vs, &n, &nv);
        len = real_side ? 2L * (long)p->MN : (p->type == VA_RCFFTOP || p->type == VA_CRFFTOP) ? (long)p->MN + 1
                                                                                              : (long)p->MN;
        esz = (!real_side && !v->block->split) ? 2L * p->prec : (long)p->prec;
        if ((len - 1) * es + 1 > lim / esz) return 0;
        c = vs > 0 ? (lim / esz - 1 - (len - 1) * es) / vs + 1 : chunk;
        if (c < chunk) chunk = c;
    }
    if (p->npass >= 2) {                        /* the split's scratch rows */
        c = lim / ((long)p->MN * 2L * p->prec);
        if (c < chunk) chunk = c;
    }
    return chunk;
}

/* a view with a negative stride */
static int va_view_negative(const VA_mview *v)
{
    return v->row_stride < 0 || v->col_stride < 0;
}

/* views the launches cannot address from one base (a negative stride, or a
 * single vector spanning 2 GiB): the plan's transform staged through
 * contiguous scratch rows, with a temporary staged plan of the same shape */
static int va_exec_via_copy(const VA_fftm *p, const VA_mview *in, const VA_mview *out, void *stream)
{
    VA_fftm t;
    int rc;
    if (va_fftm_init2(&t, p->prec, p->type, p->M, p->N, p->scale, p->dir, p->major, p->ntimes, p->hint, 2)) {
        va_error("fftm", "no staged plan for this view", -1);
        return -1;
    }
    rc = va_staged_exec(&t, in, out, stream);
    va_fftm_release(&t);             /* synchronises before freeing */
    return rc;
}

int va_fftm_exec(const VA_fftm *p, const VA_mview *in, const VA_mview *out, void *stream)
{
    vhip_fftm_desc d[VA_MAX_DESC];
    int n, k, rc = 0;
    long chunk, v0, mN;
    if (p && p->mN == 0) return 0;
    if (!p || p->markings != VA_VALID || !in || !out || !in->block || !out->block) {
        va_error("fftm", "invalid plan or view", -1);
        return -1;
    }
    if (p->staged) return va_staged_exec(p, in, out, stream);
    if (va_view_negative(in) || va_view_negative(out)) return va_exec_via_copy(p, in, out, stream);
    mN = (long)p->mN;
    chunk = va_exec_chunk(p, in, out);
    if (chunk < 1) return va_exec_via_copy(p, in, out, stream);
    for (v0 = 0; v0 < mN && !rc; v0 += chunk) {
        long cnt = mN - v0 < chunk ? mN - v0 : chunk;
        VA_fftm q = *p;
        VA_mview vi = *in, vo = *out;
        q.mN = (vsip_length)cnt;
        if (p->major == VSIP_ROW) {
            vi.offset += (vsip_offset)(v0 * in->col_stride);
            vo.offset += (vsip_offset)(v0 * out->col_stride);
            vi.col_length = vo.col_length = (vsip_length)cnt;
        } else {
            vi.offset += (vsip_offset)(v0 * in->row_stride);
            vo.offset += (vsip_offset)(v0 * out->row_stride);
            vi.row_length = vo.row_length = (vsip_length)cnt;
        }
        n = va_fftm_descs(&q, &vi, &vo, NULL, NULL, NULL, NULL, d);
        if (n < 0) return -1;
        for (k = 0; k < n && !rc; k++) {
            rc = vhip_fftm_launch(&d[k], stream);
            if (rc) va_error("fftm", vhip_error_string(rc), rc);
        }
    }
    return rc;
}

/* ---- public API -------------------------------------------------------- */

static VA_fftm *va_new_plan(int prec, int type, vsip_length M, vsip_length N, double scale,
                            int dir, vsip_major major, unsigned int ntimes, vsip_alg_hint hint)
{
    VA_fftm *p = (VA_fftm *)malloc(sizeof(VA_fftm));
    if (!p) return NULL;
    if (va_fftm_init(p, prec, type, M, N, scale, dir, major, ntimes, hint)) {
        va_fftm_release(p);
        free(p);
        return NULL;
    }
    return p;
}

vsip_fftm_f *vsip_ccfftmip_create_f(vsip_length M, vsip_length N, vsip_scalar_f scale,
                                    vsip_fft_dir dir, vsip_major major, unsigned int ntimes,
                                    vsip_alg_hint hint)
{
    return (vsip_fftm_f *)va_new_plan(4, VA_CCFFTIP, M, N, scale, dir, major, ntimes, hint);
}

vsip_fftm_f *vsip_ccfftmop_create_f(vsip_length M, vsip_length N, vsip_scalar_f scale,
                                    vsip_fft_dir dir, vsip_major major, unsigned int ntimes,
                                    vsip_alg_hint hint)
{
    return (vsip_fftm_f *)va_new_plan(4, VA_CCFFTOP, M, N, scale, dir, major, ntimes, hint);
}

vsip_fftm_f *vsip_rcfftmop_create_f(vsip_length M, vsip_length N, vsip_scalar_f scale,
                                    vsip_major major, unsigned int ntimes, vsip_alg_hint hint)
{
    return (vsip_fftm_f *)va_new_plan(4, VA_RCFFTOP, M, N, scale, VSIP_FFT_FWD, major, ntimes, hint);
}

vsip_fftm_f *vsip_crfftmop_create_f(vsip_length M, vsip_length N, vsip_scalar_f scale,
                                    vsip_major major, unsigned int ntimes, vsip_alg_hint hint)
{
    return (vsip_fftm_f *)va_new_plan(4, VA_CRFFTOP, M, N, scale, VSIP_FFT_INV, major, ntimes, hint);
}

vsip_fftm_d *vsip_ccfftmip_create_d(vsip_length M, vsip_length N, vsip_scalar_d scale,
                                    vsip_fft_dir dir, vsip_major major, unsigned int ntimes,
                                    vsip_alg_hint hint)
{
    return (vsip_fftm_d *)va_new_plan(8, VA_CCFFTIP, M, N, scale, dir, major, ntimes, hint);
}

vsip_fftm_d *vsip_ccfftmop_create_d(vsip_length M, vsip_length N, vsip_scalar_d scale,
                                    vsip_fft_dir dir, vsip_major major, unsigned int ntimes,
                                    vsip_alg_hint hint)
{
    return (vsip_fftm_d *)va_new_plan(8, VA_CCFFTOP, M, N, scale, dir, major, ntimes, hint);
}

vsip_fftm_d *vsip_rcfftmop_create_d(vsip_length M, vsip_length N, vsip_scalar_d scale,
                                    vsip_major major, unsigned int ntimes, vsip_alg_hint hint)
{
    return (vsip_fftm_d *)va_new_plan(8, VA_RCFFTOP, M, N, scale, VSIP_FFT_FWD, major, ntimes, hint);
}

vsip_fftm_d *vsip_crfftmop_create_d(vsip_length M, vsip_length N, vsip_scalar_d scale,
                                    vsip_major major, unsigned int ntimes, vsip_alg_hint hint)
{
    return (vsip_fftm_d *)va_new_plan(8, VA_CRFFTOP, M, N, scale, VSIP_FFT_INV, major, ntimes, hint);
}

static void va_check_type(const VA_fftm *p, int type, const char *fn)
{
    if (p && p->type != type) va_error(fn, "plan was created for another transform", -1);
}

void vsip_ccfftmip_f(const vsip_fftm_f *plan, const vsip_cmview_f *X)
{
    va_check_type(&plan->p, VA_CCFFTIP, "vsip_ccfftmip_f");
    va_fftm_exec(&plan->p, &X->v, &X->v, vhip_stream());
}

void vsip_ccfftmip_d(const vsip_fftm_d *plan, const vsip_cmview_d *X)
{
    va_check_type(&plan->p, VA_CCFFTIP, "vsip_ccfftmip_d");
    va_fftm_exec(&plan->p, &X->v, &X->v, vhip_stream());
}

void vsip_ccfftmop_f(const vsip_fftm_f *plan, const vsip_cmview_f *X, const vsip_cmview_f *Y)
{
    va_check_type(&plan->p, VA_CCFFTOP, "vsip_ccfftmop_f");
    va_fftm_exec(&plan->p, &X->v, &Y->v, vhip_stream());
}

void vsip_ccfftmop_d(const vsip_fftm_d *plan, const vsip_cmview_d *X, const vsip_cmview_d *Y)
{
    va_check_type(&plan->p, VA_CCFFTOP, "vsip_ccfftmop_d");
    va_fftm_exec(&plan->p, &X->v, &Y->v, vhip_stream());
}

void vsip_rcfftmop_f(const vsip_fftm_f *plan, const vsip_mview_f *X, const vsip_cmview_f *Z)
{
    va_check_type(&plan->p, VA_RCFFTOP, "vsip_rcfftmop_f");
    va_fftm_exec(&plan->p, &X->v, &Z->v, vhip_stream());
}

void vsip_rcfftmop_d(const vsip_fftm_d *plan, const vsip_mview_d *X, const vsip_cmview_d *Z)
{
    va_check_type(&plan->p, VA_RCFFTOP, "vsip_rcfftmop_d");
    va_fftm_exec(&plan->p, &X->v, &Z->v, vhip_stream());
}

void vsip_crfftmop_f(const vsip_fftm_f *plan, const vsip_cmview_f *Z, const vsip_mview_f *X)
{
    va_check_type(&plan->p, VA_CRFFTOP, "vsip_crfftmop_f");
    va_fftm_exec(&plan->p, &Z->v, &X->v, vhip_stream());
}

void vsip_crfftmop_d(const vsip_fftm_d *plan, const vsip_cmview_d *Z, const vsip_mview_d *X)
{
    va_check_type(&plan->p, VA_CRFFTOP, "vsip_crfftmop_d");
    va_fftm_exec(&plan->p, &Z->v, &X->v, vhip_stream());
}

int vsip_fftm_destroy_f(vsip_fftm_f *plan)
{
    if (plan) {
        va_fftm_release(&plan->p);
        free(plan);
    }
    return 0;
}

int vsip_fftm_destroy_d(vsip_fftm_d *plan)
{
    if (plan) {
        va_fftm_release(&plan->p);
        free(plan);
    }
    return 0;
}

/* vsip_fftm_getattr_f.c:18-69 */
static void va_getattr(const VA_fftm *p, vsip_scalar_mi *in, vsip_scalar_mi *out,
                       vsip_fft_place *place, vsip_fft_dir *dir, vsip_major *major)
{
    in->r = p->M; in->c = p->N;
    out->r = p->M; out->c = p->N;
    *place = (p->type == VA_CCFFTIP) ? VSIP_FFT_IP : VSIP_FFT_OP;
    *dir = (p->dir < 0) ? VSIP_FFT_FWD : VSIP_FFT_INV;
    *major = p->major;
    if (p->type == VA_RCFFTOP) {
        out->r = (p->major == VSIP_ROW) ? p->mN : p->MN + 1;
        out->c = (p->major == VSIP_ROW) ? p->MN + 1 : p->mN;
    } else if (p->type == VA_CRFFTOP) {
        in->r = (p->major == VSIP_ROW) ? p->mN : p->MN + 1;
        in->c = (p->major == VSIP_ROW) ? p->MN + 1 : p->mN;
    }
}

void vsip_fftm_getattr_f(const vsip_fftm_f *plan, vsip_fftm_attr_f *a)
{
    va_getattr(&plan->p, &a->input, &a->output, &a->place, &a->dir, &a->major);
    a->scale = (vsip_scalar_f)plan->p.scale;
}

void vsip_fftm_getattr_d(const vsip_fftm_d *plan, vsip_fftm_attr_d *a)
{
    va_getattr(&plan->p, &a->input, &a->output, &a->place, &a->dir, &a->major);
    a->scale = plan->p.scale;
}
