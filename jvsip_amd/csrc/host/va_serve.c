/*
 * va_serve.c -- a continuous-batching server for multiple-FFT requests.
 *
 * Requests (rows x n complex FP32 transforms, forward or inverse, with a
 * scale) arrive over time.  The server keeps every schedule and twiddle table
 * resident on the device (the autotuner's per-length cache) and runs this
 * loop until every request is done:
 *   1. requests whose arrival time has passed join a FIFO;
 *   2. finished batches are retired: their requests get their completion
 *      time and their output/scratch memory returns to the HBM arena;
 *   3. while a launch lane (a stream with its own descriptor tables) is free,
 *      requests are admitted from the FIFO head as long as their output and
 *      scratch fit the HBM budget and the batch has room; on the lane's
 *      stream, requests of at least VA_DIRECT_ELEMS elements launch their
 *      own length-specialised kernel, and the smaller ones share one
 *      persistent batched kernel (when there are more than VA_DIRECT_MAX of
 *      them: fewer launch directly too, which needs no descriptor upload and
 *      keeps the latency of a lightly loaded server at that of a single
 *      call).
 * Requests that arrive while others run are admitted as lanes and memory
 * free up; batches on different lanes overlap on the device.  Inputs are read
 * from a device-resident input pool (the request names an offset); outputs go
 * to a caller-given output pool offset, or to memory the server allocates
 * from its arena (released when the request completes, as if the client
 * consumed it at once).
 *
 * vsip_amd_serve_plain_f runs the same request stream as sequential plain
 * calls (one vsip_ccfftmop_f-equivalent launch per request followed by a
 * synchronisation), the baseline the server is compared with.
 *
 * C89 with POSIX clock_gettime.
 */
#define _POSIX_C_SOURCE 199309L
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include "va_internal.h"
#include "vsip_amd.h"

/* batches of at most this many requests launch each request's own kernel */
#define VA_DIRECT_MAX 2
/* requests of at least this many elements launch their own (plain,
 * length-specialised) kernel on the lane's stream; smaller ones share the
 * lane's persistent batched launch.  VSIP_AMD_SERVE_DIRECT overrides. */
#define VA_DIRECT_ELEMS 32768L

static long va_direct_elems(void)
{
    const char *e = getenv("VSIP_AMD_SERVE_DIRECT");
    if (e) return atol(e);
    return VA_DIRECT_ELEMS;
}

static double va_now(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* ---- HBM arena: first-fit free list over one device allocation ------------ */

typedef struct {
    size_t off, len;
} VA_span;

typedef struct {
    char *base;
    size_t size, used, peak;
    VA_span *free_list;
    int nfree, cap;
} VA_arena;

static int va_arena_init(VA_arena *a, size_t size)
{
    memset(a, 0, sizeof(*a));
    a->cap = 1024;
    a->free_list = (VA_span *)malloc(sizeof(VA_span) * (size_t)a->cap);
    if (!a->free_list) return -1;
    if (vhip_malloc((void **)&a->base, size)) {
        free(a->free_list);
        a->free_list = NULL;
        return -1;
    }
    a->size = size;
    a->free_list[0].off = 0;
    a->free_list[0].len = size;
    a->nfree = 1;
    return 0;
}

static void va_arena_fini(VA_arena *a)
{
    if (a->base) vhip_free(a->base);
    free(a->free_list);
    memset(a, 0, sizeof(*a));
}

/* offset of `len` bytes (256-aligned), or (size_t)-1 when the budget is full */
static size_t va_arena_alloc(VA_arena *a, size_t len)
{
    int i;
    len = (len + 255) & ~(size_t)255;
    for (i = 0; i < a->nfree; i++) {
        if (a->free_list[i].len >= len) {
            size_t off = a->free_list[i].off;
            a->free_list[i].off += len;
            a->free_list[i].len -= len;
            if (a->free_list[i].len == 0) {
                memmove(&a->free_list[i], &a->free_list[i + 1], sizeof(VA_span) * (size_t)(a->nfree - i - 1));
                a->nfree--;
            }
            a->used += len;
            if (a->used > a->peak) a->peak = a->used;
            return off;
        }
    }
    return (size_t)-1;
}

static void va_arena_free(VA_arena *a, size_t off, size_t len)
{
    int i = 0;
    len = (len + 255) & ~(size_t)255;
    a->used -= len;
    while (i < a->nfree && a->free_list[i].off < off) i++;
    if (a->nfree == a->cap) {
        VA_span *t = (VA_span *)realloc(a->free_list, sizeof(VA_span) * (size_t)(2 * a->cap));
        if (!t) return;   /* the span is lost to this run; the budget only shrinks */
        a->free_list = t;
        a->cap *= 2;
    }
    memmove(&a->free_list[i + 1], &a->free_list[i], sizeof(VA_span) * (size_t)(a->nfree - i));
    a->free_list[i].off = off;
    a->free_list[i].len = len;
    a->nfree++;
    if (i + 1 < a->nfree && a->free_list[i].off + a->free_list[i].len == a->free_list[i + 1].off) {
        a->free_list[i].len += a->free_list[i + 1].len;
        memmove(&a->free_list[i + 1], &a->free_list[i + 2], sizeof(VA_span) * (size_t)(a->nfree - i - 2));
        a->nfree--;
    }
    if (i > 0 && a->free_list[i - 1].off + a->free_list[i - 1].len == a->free_list[i].off) {
        a->free_list[i - 1].len += a->free_list[i].len;
        memmove(&a->free_list[i], &a->free_list[i + 1], sizeof(VA_span) * (size_t)(a->nfree - i - 1));
        a->nfree--;
    }
}

/* ---- one request as launch descriptors ------------------------------------ */

/* twiddle tables e^{-2 pi i m / n} for two-pass lengths, kept for the run */
typedef struct {
    long n;
    void *dev;
} VA_ptwc;

typedef struct {
    VA_ptwc tab[64];
    int n;
} VA_ptw_cache;

static const void *va_ptw_get(VA_ptw_cache *c, long n)
{
    int i;
    void *dev = NULL;
    for (i = 0; i < c->n; i++)
        if (c->tab[i].n == n) return c->tab[i].dev;
    if (c->n >= 64) return NULL;
    if (va_ptw_upload(&dev, 4, n)) return NULL;
    c->tab[c->n].n = n;
    c->tab[c->n].dev = dev;
    c->n++;
    return dev;
}

static void va_ptw_clear(VA_ptw_cache *c)
{
    int i;
    for (i = 0; i < c->n; i++) vhip_free(c->tab[i].dev);
    c->n = 0;
}

/* a dense rows x n matrix at device address p, as a view for va_fftm_descs */
static void va_fake_view(VA_block *blk, VA_mview *v, void *p, long rows, long n)
{
    memset(blk, 0, sizeof(*blk));
    blk->prec = 4;
    blk->cplx = 1;
    blk->kind = VA_LIB_BLOCK;
    blk->admitted = 1;
    blk->markings = VA_VALID;
    blk->size = (vsip_length)(rows * n);
    blk->dev_r = p;
    memset(v, 0, sizeof(*v));
    v->block = blk;
    v->row_stride = 1;
    v->row_length = (vsip_length)n;
    v->col_stride = n;
    v->col_length = (vsip_length)rows;
    v->markings = VA_VALID;
}

static void va_fake_plan(VA_fftm *p, const vsip_amd_request *r)
{
    memset(p, 0, sizeof(*p));
    p->prec = 4;
    p->type = VA_CCFFTOP;
    p->M = (vsip_length)r->rows;
    p->N = (vsip_length)r->n;
    p->MN = (vsip_length)r->n;
    p->mN = (vsip_length)r->rows;
    p->major = VSIP_ROW;
    p->dir = r->dir;
    p->scale = r->scale;
    p->markings = VA_VALID;
}

/* batched-shape descriptors of request r; *work_bytes gets the two-pass scratch size */
static int va_req_descs(const vsip_amd_request *r, void *in, void *out, void *work,
                        VA_ptw_cache *pc, vhip_fftm_desc *d)
{
    const VA_sched *sA = NULL, *sB = NULL;
    const void *ptw = NULL;
    VA_fftm p;
    VA_block bi, bo;
    VA_mview vi, vo;
    int t = va_tune_get(4, (int)r->n, &sA);
    if (t == -2) {
        long dd, rt = 1;
        while ((rt + 1) * (rt + 1) <= r->n) rt++;
        for (dd = rt + 1; dd >= 2; dd--) {
            if (r->n % dd) continue;
            if (va_tune_get(4, (int)(r->n / dd), &sA)) continue;
            if (va_tune_get(4, (int)dd, &sB)) continue;
            break;
        }
        if (dd < 2) return -1;
        ptw = va_ptw_get(pc, r->n);
        if (!ptw) return -1;
    } else if (t) {
        return -1;
    }
    va_fake_plan(&p, r);
    va_fake_view(&bi, &vi, in, r->rows, r->n);
    va_fake_view(&bo, &vo, out, r->rows, r->n);
    return va_fftm_descs(&p, &vi, &vo, sA, sB, work, ptw, d);
}

static int va_req_two_pass(const vsip_amd_request *r)
{
    const VA_sched *s = NULL;
    return va_tune_get(4, (int)r->n, &s) == -2;
}

/* ---- launch lanes ---------------------------------------------------------- */

typedef struct {
    void *stream, *event;
    vhip_fftm_desc *desc_dev, *desc_host;
    vhip_batch_item *items_host;
    unsigned *ctl_dev;
    int cap_desc;
    long cap_items;
    long *reqs;
    char *direct;          /* per admitted request: 1 = own kernel, 0 = batched */
    int nreqs, busy, seq;
    int blocked;           /* its stream waits behind the resident kernel */
} VA_lane;

static int va_lane_init(VA_lane *l, int max_jobs)
{
    int rc;
    memset(l, 0, sizeof(*l));
    l->cap_desc = 2 * max_jobs;
    l->cap_items = (long)max_jobs * 1200;
    l->reqs = (long *)malloc(sizeof(long) * (size_t)max_jobs);
    l->direct = (char *)malloc((size_t)max_jobs);
    if (!l->reqs || !l->direct) return -1;
    rc = vhip_stream_create(&l->stream);
    if (!rc) rc = vhip_event_create(&l->event);
    /* descriptors and items travel in one upload: desc_host/desc_dev hold both */
    if (!rc) rc = vhip_malloc((void **)&l->desc_dev, sizeof(vhip_fftm_desc) * (size_t)l->cap_desc +
                              sizeof(vhip_batch_item) * (size_t)l->cap_items);
    if (!rc) rc = vhip_malloc((void **)&l->ctl_dev, sizeof(unsigned) * 4);
    if (!rc) rc = vhip_memset(l->ctl_dev, 0, sizeof(unsigned) * 4);
    if (!rc) rc = vhip_host_alloc((void **)&l->desc_host, sizeof(vhip_fftm_desc) * (size_t)l->cap_desc +
                                  sizeof(vhip_batch_item) * (size_t)l->cap_items);
    if (!rc) rc = vhip_host_alloc((void **)&l->items_host, sizeof(vhip_batch_item) * (size_t)l->cap_items);
    return rc;
}

static void va_lane_fini(VA_lane *l)
{
    if (l->stream) vhip_stream_destroy(l->stream);
    if (l->event) vhip_event_destroy(l->event);
    if (l->desc_dev) vhip_free(l->desc_dev);
    if (l->ctl_dev) vhip_free(l->ctl_dev);
    if (l->desc_host) vhip_host_free(l->desc_host);
    if (l->items_host) vhip_host_free(l->items_host);
    free(l->reqs);
    free(l->direct);
}

typedef struct {
    size_t out_off, work_off, out_len, work_len;
    int own_out, own_work;
    void *work, *work2;    /* scratch matrices inside the arena span work_off */
} VA_reqmem;

static void *va_pool_addr(const vsip_cblock_f *pool, long off)
{
    return (char *)pool->b.dev_r + (size_t)off * 8;
}

/* plain schedules by length, for batches small enough to launch directly */
typedef struct {
    VA_fftm *p;
    int n;
} VA_plan_cache;

static const VA_fftm *va_plan_get(VA_plan_cache *c, long n)
{
    int i;
    for (i = 0; i < c->n; i++)
        if ((long)c->p[i].MN == n) return &c->p[i];
    return NULL;
}

static int va_plan_cache_fill(VA_plan_cache *c, const vsip_amd_request *reqs, long nreq)
{
    long i;
    c->p = (VA_fftm *)calloc(64, sizeof(VA_fftm));
    c->n = 0;
    if (!c->p) return -1;
    for (i = 0; i < nreq; i++) {
        long j, rows = 0;
        if (reqs[i].rows < 1 || reqs[i].n < 1) return -1;
        if (va_plan_get(c, reqs[i].n)) continue;
        if (c->n == 64) return -1;
        /* the plan's split scratch must hold the largest request of this length */
        for (j = i; j < nreq; j++)
            if (reqs[j].n == reqs[i].n && reqs[j].rows > rows) rows = reqs[j].rows;
        if (va_fftm_init(&c->p[c->n], 4, VA_CCFFTOP, (vsip_length)rows, (vsip_length)reqs[i].n, 1.0, -1, VSIP_ROW,
                         0, VSIP_ALG_TIME))
            return -1;
        c->n++;
    }
    return 0;
}

static void va_plan_cache_free(VA_plan_cache *c)
{
    int i;
    for (i = 0; i < c->n; i++) va_fftm_release(&c->p[i]);
    free(c->p);
    c->p = NULL;
    c->n = 0;
}

/* scratch matrices (rows x n each) of length n's plain schedule: 0, or 1
 * or 2 for the two- and three-pass splits */
static int va_plain_nwork(const VA_plan_cache *pc, long n)
{
    const VA_fftm *base = va_plan_get((VA_plan_cache *)pc, n);
    return (base && base->npass >= 2) ? base->npass - 1 : 0;
}

/* launch request r directly with its length's plain schedule; a split
 * schedule runs on the request's own scratch (work, work2), never on the
 * cached plan's, since requests of one length run on several lanes at once */
static int va_direct_launch(const VA_plan_cache *pc, const vsip_amd_request *r, void *in, void *out,
                            void *work, void *work2, void *stream)
{
    const VA_fftm *base = va_plan_get((VA_plan_cache *)pc, r->n);
    VA_fftm p;
    VA_block bi, bo;
    VA_mview vi, vo;
    vhip_fftm_desc d[VA_MAX_DESC];
    int nd, k, rc = 0;
    if (!base || base->staged || base->npass < 1) return -1;
    if (base->npass >= 2 && !work) return -1;
    if (base->npass == 3 && !work2) return -1;
    p = *base;
    p.mN = (vsip_length)r->rows;
    p.M = (vsip_length)r->rows;
    p.dir = r->dir;
    p.scale = r->scale;
    p.dev_work = work;
    p.dev_work2 = work2;
    va_fake_view(&bi, &vi, in, r->rows, r->n);
    va_fake_view(&bo, &vo, out, r->rows, r->n);
    nd = va_fftm_descs(&p, &vi, &vo, NULL, NULL, NULL, NULL, d);
    if (nd < 0) return -1;
    for (k = 0; k < nd && !rc; k++) rc = vhip_fftm_launch(&d[k], stream);
    return rc;
}

/* work items of request q's batched launch (descriptors built on dummy
 * addresses: only their vector counts are used), or -1 */
static long va_req_items(const vsip_amd_request *q, void *dummy, VA_ptw_cache *pc)
{
    vhip_fftm_desc d[VA_MAX_DESC];
    long items = 0;
    int k, nd = va_req_descs(q, dummy, dummy, dummy, pc, d);
    if (nd < 0) return -1;
    for (k = 0; k < nd; k++) items += (d[k].nvec + d[k].vpb - 1) / d[k].vpb;
    return items;
}

/* ---- the resident path: one persistent kernel for the whole stream ------- */

/* requests of at most this many elements whose length runs in one pass of
 * the batched workgroup go to the resident kernel; VSIP_AMD_SERVE_RESIDENT
 * overrides (0 turns the path off) */
#define VA_RES_ELEMS 131072L
/* resident workgroups (one per CU; persistent launches of the other requests
 * size their grids to the remaining CUs); VSIP_AMD_SERVE_RESIDENT_WGS */
#define VA_RES_WGS 32
/* polls after which a waiting resident workgroup gives up (a bound, not a
 * timeout the stream is expected to reach); VSIP_AMD_SERVE_SPIN_LIMIT */
#define VA_RES_SPIN (1u << 22)

static long va_env_long(const char *name, long dflt)
{
    const char *e = getenv(name);
    return e ? atol(e) : dflt;
}

typedef struct {
    long nslots, nitems, grid;
    long *slot_of;          /* per request: its slot, or -1 */
    long *req_of;           /* per slot: the request */
    unsigned *item_end;     /* per slot: items published once it is */
    char *hblk;             /* coherent host memory: slots, then htail, hstop, hdone[] */
    vhip_fftm_desc *hdesc;  /* host addresses */
    unsigned *htail, *hstop, *hdone;
    vhip_resident_args args;   /* device addresses */
    void *stream;
    int launched;
    int unreserved;         /* its CUs were handed back to the lanes */
    int lanes_back;         /* the lanes on its queue were taken back */
} VA_res;

/* Start the resident kernel.  From here until the stream ends, the host must
 * make no call that can wait for the device (allocation, free, a synchronous
 * copy, a module load): the kernel waits for the host to publish work, so
 * such a call would only return when a waiting workgroup gives up.  Every
 * schedule, twiddle table and buffer of the stream exists before this. */
static int va_res_start(VA_res *rs)
{
    int rc;
    if (rs->nslots == 0) return 0;
    /* the persistent launches of the other requests leave these CUs alone */
    rc = vhip_set_reserved_cus((int)rs->grid);
    /* and no launch compiles a kernel (a module load) while it runs */
    if (!rc) vhip_jit_launch_compiles(0);
    if (!rc) rc = vhip_resident_launch(&rs->args, (int)rs->grid, VA_BATCH_LDS, rs->stream);
    if (!rc) rs->launched = 1;
    else {
        vhip_set_reserved_cus(0);
        vhip_jit_launch_compiles(1);
    }
    return rc;
}

/* wait up to `seconds` for a stream to go idle: 0 idle, 1 still busy, else error */
static int va_stream_wait(void *stream, double seconds)
{
    double t0 = va_now();
    struct timespec ts;
    ts.tv_sec = 0;
    ts.tv_nsec = 20000;
    for (;;) {
        int q = vhip_stream_done(stream);
        if (q != 1) return q;
        if (va_now() - t0 > seconds) return 1;
        nanosleep(&ts, NULL);
    }
}

/* stop and release the resident kernel: 0, or -1 when it did not end within
 * the bound (its memory is then left allocated, since it may still read it) */
static int va_res_fini(VA_res *rs)
{
    if (rs->launched) {
        vhip_host_publish(rs->hstop, 1u);   /* waiting workgroups stop at once */
        if (va_stream_wait(rs->stream, 10.0) != 0) {
            va_error("serve", "the resident kernel did not end", -1);
            memset(rs, 0, sizeof(*rs));
            return -1;
        }
        vhip_set_reserved_cus(0);
        vhip_jit_launch_compiles(1);
    }
    if (rs->stream) vhip_stream_destroy(rs->stream);
    if (rs->hblk) vhip_host_free(rs->hblk);
    if (rs->args.items) vhip_free((void *)rs->args.items);
    if (rs->args.req_left) vhip_free(rs->args.req_left);
    if (rs->args.ctl) vhip_free(rs->args.ctl);
    free(rs->slot_of);
    free(rs->req_of);
    free(rs->item_end);
    memset(rs, 0, sizeof(*rs));
    return 0;
}

/* route requests to the resident kernel and lay out every item of the
 * stream in arrival order (the kernel starts later, in va_res_start) */
static int va_res_setup(VA_res *rs, const vsip_amd_request *reqs, long nreq, void *dummy, VA_ptw_cache *pc,
                        int grid_max)
{
    long max_elems = va_env_long("VSIP_AMD_SERVE_RESIDENT", VA_RES_ELEMS);
    long wgs = va_env_long("VSIP_AMD_SERVE_RESIDENT_WGS", VA_RES_WGS);
    unsigned spin = (unsigned)va_env_long("VSIP_AMD_SERVE_SPIN_LIMIT", (long)VA_RES_SPIN);
    long i, k, ni = 0;
    size_t slots_bytes, ctl_off, blk_bytes;
    vhip_batch_item *items = NULL;
    unsigned *left = NULL;
    void *dev = NULL;
    int rc = 0;
    memset(rs, 0, sizeof(*rs));
    rs->slot_of = (long *)malloc(sizeof(long) * (size_t)nreq);
    rs->req_of = (long *)malloc(sizeof(long) * (size_t)nreq);
    rs->item_end = (unsigned *)malloc(sizeof(unsigned) * (size_t)nreq);
    if (!rs->slot_of || !rs->req_of || !rs->item_end) return -1;
    for (i = 0; i < nreq; i++) {
        const vsip_amd_request *q = &reqs[i];
        vhip_fftm_desc d[VA_MAX_DESC];
        int nd;
        rs->slot_of[i] = -1;
        if (max_elems <= 0 || wgs <= 0 || (long)q->rows * q->n > max_elems || va_req_two_pass(q)) continue;
        nd = va_req_descs(q, dummy, dummy, NULL, pc, d);
        if (nd != 1) continue;
        ni += (d[0].nvec + d[0].vpb - 1) / d[0].vpb;
        if (ni >= 0x7fffffffL) return -1;
        rs->slot_of[i] = rs->nslots;
        rs->req_of[rs->nslots] = i;
        rs->item_end[rs->nslots] = (unsigned)ni;
        rs->nslots++;
    }
    rs->nitems = ni;
    if (rs->nslots == 0) return 0;
    /* coherent host block: descriptor slots (aligned to their stride), then
     * htail and hstop on a line of their own, then the done flags */
    slots_bytes = (size_t)rs->nslots * VHIP_RES_DESC_STRIDE;
    ctl_off = slots_bytes;
    blk_bytes = ctl_off + 128 + sizeof(unsigned) * (size_t)rs->nslots;
    rc = vhip_host_alloc_coherent((void **)&rs->hblk, blk_bytes);
    if (!rc && ((size_t)rs->hblk & (VHIP_RES_DESC_STRIDE - 1)) != 0) rc = -1;
    if (rc) return rc ? rc : -1;
    memset(rs->hblk, 0, blk_bytes);
    rs->hdesc = (vhip_fftm_desc *)rs->hblk;
    rs->htail = (unsigned *)(rs->hblk + ctl_off);
    rs->hstop = rs->htail + 1;
    rs->hdone = (unsigned *)(rs->hblk + ctl_off + 128);
    rc = vhip_host_device_ptr(&dev, rs->hblk);
    if (rc) return rc;
    rs->args.hdesc = (const vhip_fftm_desc *)dev;
    rs->args.htail = (const unsigned *)((char *)dev + ctl_off);
    rs->args.hstop = rs->args.htail + 1;
    rs->args.hdone = (unsigned *)((char *)dev + ctl_off + 128);
    /* every item of the stream, slot by slot, and each slot's item count */
    items = (vhip_batch_item *)malloc(sizeof(vhip_batch_item) * (size_t)ni);
    left = (unsigned *)malloc(sizeof(unsigned) * (size_t)rs->nslots);
    if (!items || !left) { rc = -1; goto done; }
    for (k = 0; k < rs->nslots; k++) {
        unsigned first = k ? rs->item_end[k - 1] : 0u, g;
        vhip_fftm_desc d[VA_MAX_DESC];
        if (va_req_descs(&reqs[rs->req_of[k]], dummy, dummy, NULL, pc, d) != 1) { rc = -1; goto done; }
        left[k] = rs->item_end[k] - first;
        for (g = 0; g < left[k]; g++) {
            items[first + g].desc = (int)k;
            items[first + g].pad_ = 0;
            items[first + g].vbase = (long)g * d[0].vpb;
        }
    }
    rc = vhip_malloc((void **)&rs->args.items, sizeof(vhip_batch_item) * (size_t)ni);
    if (!rc) rc = vhip_malloc((void **)&rs->args.req_left, sizeof(unsigned) * (size_t)rs->nslots);
    if (!rc) rc = vhip_malloc((void **)&rs->args.ctl, sizeof(unsigned) * 4);
    if (!rc) rc = vhip_memcpy_h2d((void *)rs->args.items, items, sizeof(vhip_batch_item) * (size_t)ni);
    if (!rc) rc = vhip_memcpy_h2d(rs->args.req_left, left, sizeof(unsigned) * (size_t)rs->nslots);
    if (!rc) rc = vhip_memset(rs->args.ctl, 0, sizeof(unsigned) * 4);
    if (!rc) rc = vhip_stream_create(&rs->stream);
    if (rc) goto done;
    rs->args.nitems = (unsigned)ni;
    rs->args.spin_limit = spin ? spin : 1u;
    rs->grid = wgs < grid_max ? wgs : grid_max;
    if (rs->grid > ni) rs->grid = ni;
done:
    free(items);
    free(left);
    return rc;
}

int vsip_amd_serve_f(const vsip_amd_request *reqs, long nreq, const vsip_cblock_f *in_pool,
                     const vsip_cblock_f *out_pool, size_t hbm_budget, int nstreams, int max_jobs,
                     vsip_amd_result *res, char *stats, size_t stats_len)
{
    VA_arena arena;
    VA_lane *lanes = NULL;
    VA_reqmem *mem = NULL;
    VA_ptw_cache pc;
    VA_plan_cache plc;
    void **req_in = NULL, **req_out = NULL;
    char *req_direct = NULL;        /* 1: the request launches its own kernel */
    signed char *req_nwork = NULL;  /* scratch matrices the request needs */
    long *fifo = NULL, head = 0, tail = 0, next = 0, done = 0, i, direct = 0;
    long batches = 0, jobs_launched = 0, max_inflight = 0, batched_launches = 0, batched_reqs = 0;
    /* resident requests: arrived (rfifo), published and not yet done (rlive) */
    long *rfifo = NULL, rhead = 0, rtail = 0, *rlive = NULL, nlive = 0, res_left = 0, polls = 0;
    long last_done = 0;
    double t_prog;
    VA_res rs;
    /* one lane and one job per group: every request runs alone, in arrival
     * order, on either path */
    const int serial = (nstreams == 1 && max_jobs == 1);
    int L, rc = 0, grid_max = 0, lanes_usable = nstreams;
    long direct_elems = va_direct_elems();
    double t0, tend;
    if (va_ensure_init()) return -1;
    if (!in_pool || !in_pool->b.dev_r || nreq <= 0 || nstreams < 1 || max_jobs < 1) return -1;
    memset(&pc, 0, sizeof(pc));
    memset(&plc, 0, sizeof(plc));
    memset(&rs, 0, sizeof(rs));
    if (va_arena_init(&arena, hbm_budget)) { va_error("serve", "HBM budget allocation failed", -1); return -1; }
    lanes = (VA_lane *)calloc((size_t)nstreams, sizeof(VA_lane));
    mem = (VA_reqmem *)calloc((size_t)nreq, sizeof(VA_reqmem));
    fifo = (long *)malloc(sizeof(long) * (size_t)nreq);
    req_in = (void **)malloc(sizeof(void *) * (size_t)nreq);
    req_out = (void **)malloc(sizeof(void *) * (size_t)nreq);
    req_direct = (char *)malloc((size_t)nreq);
    req_nwork = (signed char *)malloc((size_t)nreq);
    rfifo = (long *)malloc(sizeof(long) * (size_t)nreq);
    rlive = (long *)malloc(sizeof(long) * (size_t)nreq);
    if (!lanes || !mem || !fifo || !req_in || !req_out || !req_direct || !req_nwork || !rfifo || !rlive) {
        rc = -1;
        goto out;
    }
    for (L = 0; L < nstreams && !rc; L++) rc = va_lane_init(&lanes[L], max_jobs);
    if (!rc) rc = vhip_batch_max_grid(VA_BATCH_THREADS, VA_BATCH_MAXE, VA_BATCH_LDS, &grid_max);
    /* warm the schedule caches for every length before the clock starts */
    for (i = 0; i < nreq && !rc; i++) {
        const VA_sched *s;
        if (reqs[i].rows < 1 || reqs[i].n < 1 || va_tune_get(4, (int)reqs[i].n, &s) == -1) rc = -1;
        if (reqs[i].out_offset >= 0 && !out_pool) rc = -1;
    }
    if (!rc && va_plan_cache_fill(&plc, reqs, nreq)) rc = -1;
    /* the small single-pass requests go to the resident kernel, started here */
    if (!rc) rc = va_res_setup(&rs, reqs, nreq, in_pool->b.dev_r, &pc, grid_max);
    res_left = rs.nslots;
    /* route every request: its own kernel for large requests and for any
     * whose batched items would not fit one lane's item table (so the FIFO
     * head can always be admitted by an idle lane), else the lane's batched
     * launch; refuse requests that could never fit the budget */
    for (i = 0; i < nreq && !rc; i++) {
        const vsip_amd_request *q = &reqs[i];
        const VA_fftm *pl = va_plan_get(&plc, q->n);
        size_t len = (size_t)q->rows * (size_t)q->n * 8;
        int nb = 0, np;
        if (!pl || pl->staged || pl->npass < 1) { rc = -1; break; }
        if (rs.slot_of && rs.slot_of[i] >= 0) {
            req_direct[i] = 0;
            req_nwork[i] = 0;
            if (len + 1024 > hbm_budget) rc = -1;
            continue;
        }
        np = va_plain_nwork(&plc, q->n);
        req_direct[i] = ((long)q->rows * q->n >= direct_elems);
        if (!req_direct[i]) {
            long items = va_req_items(q, in_pool->b.dev_r, &pc);
            if (items < 0) { rc = -1; break; }
            if (items > lanes[0].cap_items) req_direct[i] = 1;
            nb = va_req_two_pass(q) ? 1 : 0;
        }
        /* a batched request may still run on its plain schedule (small groups) */
        req_nwork[i] = (signed char)(req_direct[i] ? np : (nb > np ? nb : np));
        if (len * (size_t)(1 + req_nwork[i]) + 1024 > hbm_budget) rc = -1;
    }
    if (rc) { va_error("serve", "setup failed", rc); goto out; }
    for (i = 0; i < nreq; i++) {
        res[i].t_admit = res[i].t_done = -1.0;
        res[i].batch = -1;
        res[i].stream = -1;
    }
    rc = va_res_start(&rs);
    if (rc) { va_error("serve", "resident kernel launch failed", rc); goto out; }
    /* Streams beyond the hardware queues of the process share them, and a
     * launch on a queue shared with the resident kernel would wait for the
     * whole stream.  A tiny kernel on every lane finds such lanes, which then
     * stay unused (their probe runs when the resident kernel ends). */
    if (rs.launched) {
        long lane_reqs = 0;
        int usable = 0;
        for (L = 0; L < nstreams && !rc; L++) {
            rc = vhip_batch_clear(lanes[L].ctl_dev, lanes[L].stream);
            if (!rc) rc = vhip_event_record(lanes[L].event, lanes[L].stream);
        }
        if (rc) { va_error("serve", vhip_error_string(rc), rc); goto out; }
        for (L = 0; L < nstreams; L++) {
            double tq = va_now();
            int q;
            while ((q = vhip_event_query(lanes[L].event)) == 1 && va_now() - tq < 0.02) {
            }
            if (q != 0 && q != 1) { rc = q; va_error("serve", vhip_error_string(q), q); goto out; }
            lanes[L].blocked = (q == 1);
            usable += !lanes[L].blocked;
        }
        for (i = 0; i < nreq; i++) lane_reqs += (rs.slot_of[i] < 0);
        lanes_usable = usable;
        if (usable == 0 && lane_reqs > 0) {
            rc = -1;
            va_error("serve", "every lane shares a hardware queue with the resident kernel "
                     "(GPU_MAX_HW_QUEUES too small)", nstreams);
            goto out;
        }
    }

    t0 = va_now();
    t_prog = t0;
    while (done < nreq) {
        double now = va_now() - t0;
        int inflight = 0;
        while (next < nreq && reqs[next].t_arrive <= now) {
            if (rs.nslots && rs.slot_of[next] >= 0) rfifo[rtail++] = next++;
            else fifo[tail++] = next++;
        }
        /* resident requests: retire the finished, then publish the arrived
         * ones in order while their output fits the budget */
        for (i = 0; i < nlive; i++) {
            long r = rlive[i];
            if (!vhip_host_peek(&rs.hdone[rs.slot_of[r]])) continue;
            res[r].t_done = va_now() - t0;
            if (mem[r].own_out) va_arena_free(&arena, mem[r].out_off, mem[r].out_len);
            done++;
            res_left--;
            rlive[i--] = rlive[--nlive];
        }
        while (rhead < rtail) {
            long r = rfifo[rhead], slot = rs.slot_of[r];
            const vsip_amd_request *q = &reqs[r];
            vhip_fftm_desc *sd = (vhip_fftm_desc *)(rs.hblk + (size_t)slot * VHIP_RES_DESC_STRIDE);
            void *out;
            if (serial && (nlive > 0 || lanes[0].busy || (head < tail && fifo[head] < r))) break;
            mem[r].out_len = (size_t)q->rows * (size_t)q->n * 8;
            mem[r].own_out = (q->out_offset < 0);
            mem[r].own_work = 0;
            if (mem[r].own_out) {
                mem[r].out_off = va_arena_alloc(&arena, mem[r].out_len);
                if (mem[r].out_off == (size_t)-1) break;
            }
            out = mem[r].own_out ? (void *)(arena.base + mem[r].out_off) : va_pool_addr(out_pool, q->out_offset);
            if (va_req_descs(q, va_pool_addr(in_pool, q->in_offset), out, NULL, &pc, sd) != 1) {
                rc = -1;
                va_error("serve", "resident request has no single-pass schedule", (int)q->n);
                goto out;
            }
            vhip_host_publish(rs.htail, rs.item_end[slot]);
            res[r].t_admit = va_now() - t0;
            res[r].batch = -1;
            res[r].stream = nstreams;   /* the resident kernel, after the lanes */
            rlive[nlive++] = r;
            rhead++;
        }
        /* every resident request published and done: the kernel is ending,
         * and the lanes' persistent launches may use every CU again */
        if (rs.launched && !rs.unreserved && rhead == rs.nslots && nlive == 0) {
            vhip_set_reserved_cus(0);
            rs.unreserved = 1;
        }
        /* once it has ended, the lanes that shared its queue run again */
        if (rs.unreserved && !rs.lanes_back && (polls & 63) == 0 && vhip_stream_done(rs.stream) == 0) {
            int still = 0;
            for (L = 0; L < nstreams; L++) {
                if (lanes[L].blocked && vhip_event_query(lanes[L].event) == 0) lanes[L].blocked = 0;
                still += lanes[L].blocked;
            }
            rs.lanes_back = (still == 0);
        }
        /* the resident kernel only ends once it has claimed every item, so it
         * must not be idle while resident requests are unfinished */
        ++polls;
        if (res_left > 0 && (polls & 1023) == 0) {
            int q = vhip_stream_done(rs.stream);
            if (q == 0) {
                for (i = 0; i < nlive; i++)
                    if (!vhip_host_peek(&rs.hdone[rs.slot_of[rlive[i]]])) break;
                if (i < nlive || rhead < rs.nslots) {
                    rc = -1;
                    va_error("serve", "the resident kernel stopped with requests unfinished", (int)res_left);
                    goto out;
                }
            } else if (q != 1) {
                rc = q;
                va_error("serve", vhip_error_string(q), q);
                goto out;
            }
        }
        /* retire */
        for (L = 0; L < nstreams; L++) {
            VA_lane *l = &lanes[L];
            int q;
            if (!l->busy || l->blocked) continue;
            q = vhip_event_query(l->event);
            if (q == 1) continue;
            if (q != 0) { rc = q; va_error("serve", vhip_error_string(q), q); goto out; }
            now = va_now() - t0;
            for (i = 0; i < l->nreqs; i++) {
                long r = l->reqs[i];
                res[r].t_done = now;
                if (mem[r].own_out) va_arena_free(&arena, mem[r].out_off, mem[r].out_len);
                if (mem[r].own_work) va_arena_free(&arena, mem[r].work_off, mem[r].work_len);
                done++;
            }
            l->busy = 0;
            l->nreqs = 0;
        }
        /* admit and launch */
        for (L = 0; L < nstreams && head < tail; L++) {
            VA_lane *l = &lanes[L];
            int nd = 0, nbatched = 0;
            long ni = 0, n1 = 0, n2b = 0, g;
            if (serial && (nlive > 0 || (rhead < rtail && rfifo[rhead] < fifo[head]))) break;
            if (l->busy || l->blocked) continue;
            while (head < tail && l->nreqs < max_jobs) {
                long r = fifo[head];
                const vsip_amd_request *q = &reqs[r];
                void *in, *out;
                int k, n2;
                long need_items = 0;
                mem[r].out_len = (size_t)q->rows * (size_t)q->n * 8;
                mem[r].own_out = (q->out_offset < 0);
                mem[r].own_work = (req_nwork[r] > 0);
                mem[r].work_len = mem[r].out_len * (size_t)req_nwork[r];
                mem[r].work = mem[r].work2 = NULL;
                if (mem[r].own_out) {
                    mem[r].out_off = va_arena_alloc(&arena, mem[r].out_len);
                    if (mem[r].out_off == (size_t)-1) break;
                }
                if (mem[r].own_work) {
                    mem[r].work_off = va_arena_alloc(&arena, mem[r].work_len);
                    if (mem[r].work_off == (size_t)-1) {
                        if (mem[r].own_out) va_arena_free(&arena, mem[r].out_off, mem[r].out_len);
                        break;
                    }
                    mem[r].work = arena.base + mem[r].work_off;
                    if (req_nwork[r] == 2) mem[r].work2 = (char *)mem[r].work + mem[r].out_len;
                }
                in = va_pool_addr(in_pool, q->in_offset);
                out = mem[r].own_out ? (void *)(arena.base + mem[r].out_off) : va_pool_addr(out_pool, q->out_offset);
                if (req_direct[r]) {
                    req_in[r] = in;
                    req_out[r] = out;
                    res[r].t_admit = va_now() - t0;
                    res[r].batch = (int)batches;
                    res[r].stream = L;
                    l->direct[l->nreqs] = 1;
                    l->reqs[l->nreqs++] = r;
                    head++;
                    continue;
                }
                n2 = va_req_descs(q, in, out, mem[r].work, &pc, &l->desc_host[nd]);
                if (n2 < 0) { rc = -1; va_error("serve", "request has no schedule", (int)q->n); goto out; }
                for (k = 0; k < n2; k++)
                    need_items += (l->desc_host[nd + k].nvec + l->desc_host[nd + k].vpb - 1) / l->desc_host[nd + k].vpb;
                if (ni + need_items > l->cap_items) {
                    /* the lane is full; routing guarantees an idle lane takes it */
                    if (mem[r].own_out) va_arena_free(&arena, mem[r].out_off, mem[r].out_len);
                    if (mem[r].own_work) va_arena_free(&arena, mem[r].work_off, mem[r].work_len);
                    break;
                }
                for (k = 0; k < n2; k++) {
                    const vhip_fftm_desc *d = &l->desc_host[nd + k];
                    long ng = (d->nvec + d->vpb - 1) / d->vpb;
                    /* first lists fill from the front, second passes from the back */
                    for (g = 0; g < ng; g++) {
                        vhip_batch_item *it = (k == 0) ? &l->items_host[n1++]
                                                       : &l->items_host[l->cap_items - 1 - (n2b++)];
                        it->desc = nd + k;
                        it->pad_ = 0;
                        it->vbase = g * d->vpb;
                    }
                    ni += ng;
                }
                nd += n2;
                req_in[r] = in;
                req_out[r] = out;
                res[r].t_admit = va_now() - t0;
                res[r].batch = (int)batches;
                res[r].stream = L;
                l->direct[l->nreqs] = 0;
                l->reqs[l->nreqs++] = r;
                nbatched++;
                head++;
            }
            if (l->nreqs == 0) break;   /* the FIFO head does not fit the budget yet */
            /* requests of their own kernel, and a batch of at most VA_DIRECT_MAX
             * requests (plain kernels need no descriptor upload) */
            for (g = 0; g < l->nreqs && !rc; g++) {
                long r = l->reqs[g];
                if (!l->direct[g] && nbatched > VA_DIRECT_MAX) continue;
                rc = va_direct_launch(&plc, &reqs[r], req_in[r], req_out[r], mem[r].work, mem[r].work2,
                                      l->stream);
                direct++;
            }
            if (!rc && nbatched > VA_DIRECT_MAX) {
                /* one upload: the descriptors, then the first list, then the second
                 * list (stacked from the back of the item area while building) */
                char *hb = (char *)l->desc_host, *db = (char *)l->desc_dev;
                size_t dbytes = sizeof(vhip_fftm_desc) * (size_t)nd;
                vhip_batch_item *hi = (vhip_batch_item *)(hb + dbytes);
                memmove(hi, l->items_host, sizeof(vhip_batch_item) * (size_t)n1);
                for (g = 0; g < n2b; g++) hi[n1 + g] = l->items_host[l->cap_items - n2b + g];
                rc = vhip_memcpy_h2d_async(db, hb, dbytes + sizeof(vhip_batch_item) * (size_t)ni, l->stream);
                if (!rc) rc = vhip_batch_launch(l->desc_dev, (const vhip_batch_item *)(db + dbytes), n1, n2b,
                                                l->ctl_dev, VA_BATCH_THREADS, VA_BATCH_MAXE, VA_BATCH_LDS,
                                                grid_max, l->stream);
                batched_launches++;
                batched_reqs += nbatched;
            }
            if (!rc) rc = vhip_event_record(l->event, l->stream);
            if (rc) { va_error("serve", vhip_error_string(rc), rc); goto out; }
            l->busy = 1;
            batches++;
            jobs_launched += l->nreqs;
        }
        for (inflight = 0, L = 0; L < nstreams; L++) inflight += lanes[L].busy;
        if (inflight > max_inflight) max_inflight = inflight;
        /* a bound, not a schedule: work in flight with no completion for
         * this long means the device stopped making progress */
        if (done != last_done || (inflight == 0 && nlive == 0)) {
            last_done = done;
            t_prog = va_now();
        } else if (va_now() - t_prog > 10.0) {
            rc = -1;
            va_error("serve", "no request completed for 10 s", (int)(nreq - done));
            goto out;
        }
    }
    tend = va_now() - t0;
    if (stats && stats_len) {
        char tmp[1280];
        double bytes = 0.0;
        for (i = 0; i < nreq; i++) bytes += 2.0 * (double)reqs[i].rows * (double)reqs[i].n * 8.0;
        sprintf(tmp, "{\"requests\": %ld, \"seconds\": %.6f, \"req_per_s\": %.1f, \"gb_per_s\": %.2f, "
                "\"batches\": %ld, \"jobs_per_batch\": %.2f, \"lanes\": %d, \"max_jobs\": %d, "
                "\"max_lanes_in_flight\": %ld, \"hbm_budget\": %lu, \"hbm_peak\": %lu, "
                "\"direct_launched\": %ld, \"batched_requests\": %ld, \"batched_launches\": %ld, "
                "\"requests_per_batched_launch\": %.2f, \"direct_threshold_elems\": %ld, "
                "\"resident_requests\": %ld, \"resident_items\": %ld, \"resident_workgroups\": %ld, "
                "\"resident_max_elems\": %ld, \"lanes_usable\": %d}",
                nreq, tend, (double)nreq / tend, bytes / tend / 1e9, batches,
                batches ? (double)jobs_launched / (double)batches : 0.0, nstreams, max_jobs,
                max_inflight, (unsigned long)hbm_budget, (unsigned long)arena.peak, direct,
                batched_reqs, batched_launches,
                batched_launches ? (double)batched_reqs / (double)batched_launches : 0.0, direct_elems,
                rs.nslots, rs.nitems, rs.grid, va_env_long("VSIP_AMD_SERVE_RESIDENT", VA_RES_ELEMS), lanes_usable);
        strncpy(stats, tmp, stats_len - 1);
        stats[stats_len - 1] = '\0';
    }
out:
    if (va_res_fini(&rs) && !rc) rc = -1;
    vhip_sync();
    for (L = 0; lanes && L < nstreams; L++) {
        if (lanes[L].stream) vhip_stream_destroy(lanes[L].stream);
        lanes[L].stream = NULL;
        va_lane_fini(&lanes[L]);
    }
    va_ptw_clear(&pc);
    va_plan_cache_free(&plc);
    va_arena_fini(&arena);
    free(lanes);
    free(mem);
    free(fifo);
    free(req_in);
    free(req_out);
    free(req_direct);
    free(req_nwork);
    free(rfifo);
    free(rlive);
    return rc;
}

/* the same request stream as plain calls in arrival order on the library
 * stream: with sync_each, each call is followed by a synchronisation (a
 * caller that waits for every result); without, the calls go back to back
 * and an event per request gives its completion time */
#define VA_PLAIN_EVENTS 256

int vsip_amd_serve_plain_f(const vsip_amd_request *reqs, long nreq, const vsip_cblock_f *in_pool,
                           const vsip_cblock_f *out_pool, size_t hbm_budget, int sync_each,
                           vsip_amd_result *res, char *stats, size_t stats_len)
{
    VA_arena arena;
    VA_fftm *plans = NULL;
    void *ev[VA_PLAIN_EVENTS];
    size_t *offs = NULL;
    int *plan_of = NULL;     /* each request's plan, resolved before the clock starts */
    int nplans = 0, cap = 64, rc = 0, e;
    long i, oldest = 0;
    double t0, tend;
    if (va_ensure_init()) return -1;
    if (!in_pool || nreq <= 0) return -1;
    memset(ev, 0, sizeof(ev));
    if (va_arena_init(&arena, hbm_budget)) { va_error("serve", "HBM budget allocation failed", -1); return -1; }
    plans = (VA_fftm *)calloc((size_t)cap, sizeof(VA_fftm));
    offs = (size_t *)malloc(sizeof(size_t) * (size_t)nreq);
    plan_of = (int *)malloc(sizeof(int) * (size_t)nreq);
    if (!plans || !offs || !plan_of) { rc = -1; goto out; }
    for (e = 0; e < VA_PLAIN_EVENTS && !rc; e++) rc = vhip_event_create(&ev[e]);
    /* plans are made before the clock starts, as a server would keep them */
    for (i = 0; i < nreq && !rc; i++) {
        int k, found = 0;
        if (reqs[i].out_offset >= 0 && !out_pool) { rc = -1; break; }
        for (k = 0; k < nplans && !found; k++)
            found = (plans[k].mN == (vsip_length)reqs[i].rows && plans[k].MN == (vsip_length)reqs[i].n &&
                     plans[k].dir == reqs[i].dir && plans[k].scale == (double)reqs[i].scale);
        if (found) { plan_of[i] = k - 1; continue; }
        plan_of[i] = nplans;
        if (nplans == cap) {
            VA_fftm *t = (VA_fftm *)realloc(plans, sizeof(VA_fftm) * (size_t)(2 * cap));
            if (!t) { rc = -1; break; }
            plans = t;
            cap *= 2;
        }
        rc = va_fftm_init(&plans[nplans], 4, VA_CCFFTOP, (vsip_length)reqs[i].rows, (vsip_length)reqs[i].n,
                          reqs[i].scale, reqs[i].dir, VSIP_ROW, 0, VSIP_ALG_TIME);
        if (!rc) nplans++;
    }
    if (rc) { va_error("serve_plain", "plan creation failed", rc); goto out; }
    t0 = va_now();
    for (i = 0; i <= nreq && !rc; i++) {
        const vsip_amd_request *q = (i < nreq) ? &reqs[i] : NULL;
        VA_block bi, bo;
        VA_mview vi, vo;
        size_t len = q ? (size_t)q->rows * (size_t)q->n * 8 : 0;
        void *out;
        /* retire finished requests (in order: one stream); wait for the
         * request's arrival, for a free event, or at the end for everyone */
        for (;;) {
            int need = (i == nreq && oldest < i) || (i - oldest >= VA_PLAIN_EVENTS);
            while (oldest < i) {
                int qr = vhip_event_query(ev[oldest % VA_PLAIN_EVENTS]);
                if (qr == 1) break;
                if (qr != 0) { rc = qr; va_error("serve_plain", vhip_error_string(qr), qr); goto out; }
                res[oldest].t_done = va_now() - t0;
                if (reqs[oldest].out_offset < 0) va_arena_free(&arena, offs[oldest], (size_t)reqs[oldest].rows *
                                                                (size_t)reqs[oldest].n * 8);
                oldest++;
            }
            if (need && (i == nreq ? oldest < i : i - oldest >= VA_PLAIN_EVENTS)) continue;
            if (q && va_now() - t0 < q->t_arrive) continue;
            if (q && q->out_offset < 0) {
                offs[i] = va_arena_alloc(&arena, len);
                if (offs[i] == (size_t)-1) {
                    if (oldest == i) { rc = -1; va_error("serve_plain", "HBM budget exhausted", -1); goto out; }
                    continue;            /* wait for an earlier request's memory */
                }
            }
            break;
        }
        if (!q) break;
        res[i].t_admit = va_now() - t0;
        out = (q->out_offset < 0) ? (void *)(arena.base + offs[i]) : va_pool_addr(out_pool, q->out_offset);
        va_fake_view(&bi, &vi, va_pool_addr(in_pool, q->in_offset), q->rows, q->n);
        va_fake_view(&bo, &vo, out, q->rows, q->n);
        rc = va_fftm_exec(&plans[plan_of[i]], &vi, &vo, vhip_stream());
        if (!rc) rc = vhip_event_record(ev[i % VA_PLAIN_EVENTS], vhip_stream());
        if (!rc && sync_each) rc = vhip_sync();
        res[i].batch = -1;
        res[i].stream = 0;
    }
    tend = va_now() - t0;
    if (!rc && stats && stats_len) {
        char tmp[256];
        double bytes = 0.0;
        for (i = 0; i < nreq; i++) bytes += 2.0 * (double)reqs[i].rows * (double)reqs[i].n * 8.0;
        sprintf(tmp, "{\"requests\": %ld, \"seconds\": %.6f, \"req_per_s\": %.1f, \"gb_per_s\": %.2f, "
                "\"plans\": %d, \"sync_each\": %d}",
                nreq, tend, (double)nreq / tend, bytes / tend / 1e9, nplans, sync_each ? 1 : 0);
        strncpy(stats, tmp, stats_len - 1);
        stats[stats_len - 1] = '\0';
    }
out:
    vhip_sync();
    for (e = 0; e < VA_PLAIN_EVENTS; e++)
        if (ev[e]) vhip_event_destroy(ev[e]);
    for (i = 0; i < nplans; i++) va_fftm_release(&plans[i]);
    free(plans);
    free(offs);
    free(plan_of);
    va_arena_fini(&arena);
    return rc;
}
