/*
 * va_batch.c -- many multiple-FFTs of uneven lengths in one persistent
 * launch, captured in a HIP graph.
 *
 * vsip_amd_batch_create_f takes njobs (plan, input view, output view) triples
 * (ccfftmip plans run in place, ccfftmop plans out of place) and lays out:
 *   - one launch descriptor per transform pass, with the schedule the
 *     autotuner picked for that length and the batched workgroup shape;
 *     lengths that do not fit one workgroup run as two passes
 *     (n = n1 * n2, twiddle e^{-2 pi i n2 k1 / n} between them, scratch owned
 *     by the batch);
 *   - two work-item lists: the single-pass items and the first-pass items of
 *     two-pass transforms (longest transforms first, items of one length
 *     together so its twiddles stay in cache), then the second-pass items;
 *   - a graph of [persistent kernel over list 1, persistent kernel over
 *     list 2] captured on a private stream (each launch leaves its work
 *     counters cleared for the next replay).
 * vsip_amd_batch_run_f replays the graph on the library stream, so it orders
 * with the plain vsip_* calls and costs one graph launch per batch.  The
 * views' blocks must stay admitted and their device data must not move while
 * the batch exists (their device addresses are baked into the descriptors).
 *
 * C89.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "va_internal.h"
#include "vsip_amd.h"

#define VA_PI 3.14159265358979323846264338327950288

typedef struct {
    long n;
    void *dev;
} VA_ptw;

struct vsip_amd_batch_f {
    int njobs;
    int ndesc;
    long nitems;
    long nitems2;
    int grid;
    vhip_fftm_desc *desc_dev;
    vhip_batch_item *items_dev;
    unsigned *ctl_dev;
    void *work_dev;
    VA_ptw *ptw;
    int nptw;
    void *graph;
    void *cap_stream;
    long total_elems;
    int two_pass_jobs;
};

static const void *va_batch_ptw(struct vsip_amd_batch_f *b, long n, int prec)
{
    int i;
    void *dev = NULL;
    for (i = 0; i < b->nptw; i++)
        if (b->ptw[i].n == n) return b->ptw[i].dev;
    if (va_ptw_upload(&dev, prec, n)) return NULL;
    b->ptw[b->nptw].n = n;
    b->ptw[b->nptw].dev = dev;
    b->nptw++;
    return dev;
}

/* two-pass split n = n1 * n2 with both factors tuned; 0 on success */
static int va_batch_split(int prec, long n, const VA_sched **a, const VA_sched **bb)
{
    long d, r = (long)sqrt((double)n);
    for (d = r + 1; d >= 2; d--) {
        if (n % d) continue;
        if (va_tune_get(prec, (int)(n / d), a)) continue;
        if (va_tune_get(prec, (int)d, bb)) continue;
        return 0;
    }
    return -1;
}

typedef struct {
    long len;       /* length of the item's transform pass */
    vhip_batch_item it;
    int phase;      /* 0 single or first pass, 1 second pass */
} VA_item_sort;

static int va_item_cmp(const void *x, const void *y)
{
    const VA_item_sort *a = (const VA_item_sort *)x, *b = (const VA_item_sort *)y;
    if (a->phase != b->phase) return a->phase - b->phase;
    if (a->len != b->len) return a->len > b->len ? -1 : 1;
    if (a->it.desc != b->it.desc) return a->it.desc - b->it.desc;
    return a->it.vbase < b->it.vbase ? -1 : (a->it.vbase > b->it.vbase);
}

void vsip_amd_batch_destroy_f(vsip_amd_batch_f *b);

vsip_amd_batch_f *vsip_amd_batch_create_f(vsip_length njobs, vsip_fftm_f *const *plans,
                                          const vsip_cmview_f *const *ins,
                                          const vsip_cmview_f *const *outs)
{
    struct vsip_amd_batch_f *b;
    vhip_fftm_desc *descs = NULL;
    VA_item_sort *items = NULL;
    vhip_batch_item *flat = NULL;
    long ni = 0, cap_items = 0, work_elems = 0, work_off = 0;
    int j, rc = 0;
    if (va_ensure_init() || njobs == 0) return NULL;
    b = (struct vsip_amd_batch_f *)calloc(1, sizeof(*b));
    if (!b) return NULL;
    b->njobs = (int)njobs;
    descs = (vhip_fftm_desc *)calloc(2 * njobs, sizeof(vhip_fftm_desc));
    b->ptw = (VA_ptw *)calloc(njobs, sizeof(VA_ptw));
    if (!descs || !b->ptw) goto fail;

    /* pass 1: scratch size of two-pass transforms */
    for (j = 0; j < (int)njobs; j++) {
        const VA_fftm *p = &plans[j]->p;
        const VA_sched *s;
        if (p->prec != 4 || (p->type != VA_CCFFTIP && p->type != VA_CCFFTOP)) {
            va_error("batch", "only complex FP32 ccfftmip/ccfftmop plans batch", -1);
            goto fail;
        }
        if (va_tune_get(4, (int)p->MN, &s) == -2) work_elems += (long)(p->MN * p->mN);
    }
    if (work_elems && vhip_malloc(&b->work_dev, (size_t)work_elems * 8)) goto fail;

    /* pass 2: descriptors and items */
    for (j = 0; j < (int)njobs; j++) {
        const VA_fftm *p = &plans[j]->p;
        const VA_mview *in = &ins[j]->v;
        const VA_mview *out = (outs && outs[j]) ? &outs[j]->v : in;
        const VA_sched *sA = NULL, *sB = NULL;
        const void *ptw = NULL;
        void *work = NULL;
        int nd, k, t = va_tune_get(4, (int)p->MN, &sA);
        if (t == -2) {
            if (va_batch_split(4, (long)p->MN, &sA, &sB)) {
                va_error("batch", "no two-pass split for a length", (int)p->MN);
                goto fail;
            }
            ptw = va_batch_ptw(b, (long)p->MN, 4);
            work = (char *)b->work_dev + work_off * 8;
            work_off += (long)(p->MN * p->mN);
            if (!ptw) goto fail;
            b->two_pass_jobs++;
        } else if (t) {
            va_error("batch", "length not supported", (int)p->MN);
            goto fail;
        }
        nd = va_fftm_descs(p, in, out, sA, sB, work, ptw, &descs[b->ndesc]);
        if (nd < 0) goto fail;
        b->total_elems += (long)(p->MN * p->mN);
        for (k = 0; k < nd; k++) {
            const vhip_fftm_desc *d = &descs[b->ndesc + k];
            long g, ng = (d->nvec + d->vpb - 1) / d->vpb;
            if (ni + ng > cap_items) {
                long nc = 2 * (ni + ng) + 1024;
                VA_item_sort *tmp = (VA_item_sort *)realloc(items, sizeof(VA_item_sort) * (size_t)nc);
                if (!tmp) goto fail;
                items = tmp;
                cap_items = nc;
            }
            for (g = 0; g < ng; g++) {
                VA_item_sort *it = &items[ni++];
                it->it.desc = b->ndesc + k;
                it->it.pad_ = 0;
                it->it.vbase = g * d->vpb;
                it->len = d->n;
                it->phase = (k == 1);
                if (k == 1) b->nitems2++;
            }
        }
        b->ndesc += nd;
    }
    qsort(items, (size_t)ni, sizeof(VA_item_sort), va_item_cmp);
    flat = (vhip_batch_item *)malloc(sizeof(vhip_batch_item) * (size_t)(ni ? ni : 1));
    if (!flat) goto fail;
    for (j = 0; j < ni; j++) flat[j] = items[j].it;
    b->nitems = ni - b->nitems2;

    rc = vhip_malloc((void **)&b->desc_dev, sizeof(vhip_fftm_desc) * (size_t)b->ndesc);
    if (!rc) rc = vhip_malloc((void **)&b->items_dev, sizeof(vhip_batch_item) * (size_t)(ni ? ni : 1));
    if (!rc) rc = vhip_malloc((void **)&b->ctl_dev, sizeof(unsigned) * 4);
    if (!rc) rc = vhip_memset(b->ctl_dev, 0, sizeof(unsigned) * 4);
    if (!rc) rc = vhip_memcpy_h2d(b->desc_dev, descs, sizeof(vhip_fftm_desc) * (size_t)b->ndesc);
    if (!rc && ni) rc = vhip_memcpy_h2d(b->items_dev, flat, sizeof(vhip_batch_item) * (size_t)ni);
    if (!rc) rc = vhip_batch_max_grid(VA_BATCH_THREADS, VA_BATCH_MAXE, VA_BATCH_LDS, &b->grid);
    if (!rc) rc = vhip_stream_create(&b->cap_stream);
    if (!rc) rc = vhip_graph_begin(b->cap_stream);
    if (!rc) {
        int lr = vhip_batch_launch(b->desc_dev, b->items_dev, b->nitems, b->nitems2, b->ctl_dev,
                                   VA_BATCH_THREADS, VA_BATCH_MAXE, VA_BATCH_LDS, b->grid, b->cap_stream);
        rc = vhip_graph_end(b->cap_stream, &b->graph);
        if (!rc) rc = lr;
    }
    if (rc) { va_error("batch", vhip_error_string(rc), rc); goto fail; }
    free(descs);
    free(items);
    free(flat);
    return b;
fail:
    free(descs);
    free(items);
    free(flat);
    vsip_amd_batch_destroy_f(b);
    return NULL;
}

int vsip_amd_batch_run_f(vsip_amd_batch_f *b)
{
    int rc;
    if (!b || !b->graph) return -1;
    rc = vhip_graph_launch(b->graph, vhip_stream());
    if (rc) va_error("batch", vhip_error_string(rc), rc);
    return rc;
}

void vsip_amd_batch_destroy_f(vsip_amd_batch_f *b)
{
    int i;
    if (!b) return;
    vhip_sync();
    if (b->graph) vhip_graph_destroy(b->graph);
    if (b->cap_stream) vhip_stream_destroy(b->cap_stream);
    if (b->desc_dev) vhip_free(b->desc_dev);
    if (b->items_dev) vhip_free(b->items_dev);
    if (b->ctl_dev) vhip_free(b->ctl_dev);
    if (b->work_dev) vhip_free(b->work_dev);
    for (i = 0; i < b->nptw; i++) vhip_free(b->ptw[i].dev);
    free(b->ptw);
    free(b);
}

int vsip_amd_batch_info_f(const vsip_amd_batch_f *b, char *buf, size_t len)
{
    char tmp[512];
    int pos;
    if (!b) return -1;
    pos = sprintf(tmp, "{\"jobs\": %d, \"descriptors\": %d, \"items\": %ld, \"second_pass_items\": %ld, "
                  "\"two_pass_jobs\": %d, \"grid\": %d, \"threads\": %d, \"lds_bytes\": %d, \"elements\": %ld}",
                  b->njobs, b->ndesc, b->nitems + b->nitems2, b->nitems2, b->two_pass_jobs, b->grid,
                  VA_BATCH_THREADS, VA_BATCH_LDS, b->total_elems);
    if (buf && len) {
        strncpy(buf, tmp, len - 1);
        buf[len - 1] = '\0';
    }
    return pos;
}

/* ---- one-shot batch that allocates nothing -------------------------------
 *
 * The same persistent batched kernel as a created batch, for callers that may
 * not allocate (code under graph capture, tight serving loops): the launch
 * descriptors, work items and control counters go to the library's static
 * device scratch (VHIP_NOALLOC_BYTES per GPU, looked up at library set-up)
 * instead of memory the library allocates.  They are built in one of two
 * static host staging halves, uploaded with one copy on the library stream,
 * and the kernel drains them after a small kernel has zeroed the counters.
 * Every job must be single pass in the batched shape with its schedule
 * already in the autotuner's cache (a lookup that misses would tune, and
 * tuning allocates).  All calls share the device scratch: they are ordered
 * on the library stream, and calls from several host threads at once race.
 */
typedef union {
    double align_;
    char b[VHIP_NOALLOC_BYTES];
} VA_stage;
static VA_stage va_na_stage[2];
static int va_na_turn = 0;
static int va_na_grid = 0;

/* items sorted longest transform first, then by descriptor and first vector
 * (pad_ carries the length while sorting); a shell sort, which needs no memory */
static int va_na_before(const vhip_batch_item *a, const vhip_batch_item *b)
{
    if (a->pad_ != b->pad_) return a->pad_ > b->pad_;
    if (a->desc != b->desc) return a->desc < b->desc;
    return a->vbase < b->vbase;
}

static void va_na_sort(vhip_batch_item *it, long n)
{
    static const long gaps[] = {88573, 29524, 9841, 3280, 1093, 364, 121, 40, 13, 4, 1};
    int g;
    for (g = 0; g < (int)(sizeof(gaps) / sizeof(gaps[0])); g++) {
        long h = gaps[g], i, j;
        if (h >= n) continue;
        for (i = h; i < n; i++) {
            vhip_batch_item t = it[i];
            for (j = i; j >= h && va_na_before(&t, &it[j - h]); j -= h) it[j] = it[j - h];
            it[j] = t;
        }
    }
}

int vsip_amd_batch_run_noalloc_f(vsip_length njobs, vsip_fftm_f *const *plans,
                                 const vsip_cmview_f *const *ins, const vsip_cmview_f *const *outs)
{
    char *dev = NULL, *host;
    vhip_fftm_desc *descs;
    vhip_batch_item *items;
    size_t max_desc, used;
    long ni = 0, j, cap_items;
    int slot, rc;
    if (va_ensure_init()) return -1;
    if (njobs == 0) return 0;
    if (!plans || !ins) { va_error("batch_noalloc", "no plans or input views", -1); return -1; }
    rc = vhip_noalloc_scratch((void **)&dev);
    if (rc) { va_error("batch_noalloc", "no device scratch", rc); return -1; }
    max_desc = (VHIP_NOALLOC_BYTES - VHIP_NOALLOC_HEAD) / sizeof(vhip_fftm_desc);
    if (njobs > max_desc) {
        va_error("batch_noalloc", "more jobs than the scratch holds descriptors for", (int)max_desc);
        return -1;
    }
    if (!va_na_grid) {
        rc = vhip_batch_max_grid(VA_BATCH_THREADS, VA_BATCH_MAXE, VA_BATCH_LDS, &va_na_grid);
        if (rc) { va_na_grid = 0; va_error("batch_noalloc", vhip_error_string(rc), rc); return -1; }
    }
    /* this staging half was last uploaded two calls ago: wait for that copy */
    slot = va_na_turn;
    rc = vhip_noalloc_stage_wait(slot);
    if (rc) { va_error("batch_noalloc", vhip_error_string(rc), rc); return -1; }
    host = va_na_stage[slot].b;
    descs = (vhip_fftm_desc *)(host + VHIP_NOALLOC_HEAD);
    items = (vhip_batch_item *)(descs + njobs);
    cap_items = (long)((VHIP_NOALLOC_BYTES - VHIP_NOALLOC_HEAD - njobs * sizeof(vhip_fftm_desc)) /
                       sizeof(vhip_batch_item));
    for (j = 0; j < (long)njobs; j++) {
        const VA_fftm *p;
        const VA_mview *in, *out;
        const VA_sched *s = NULL;
        long g, ng;
        int t;
        if (!plans[j] || !ins[j]) { va_error("batch_noalloc", "missing plan or view", (int)j); return -1; }
        p = &plans[j]->p;
        in = &ins[j]->v;
        out = (outs && outs[j]) ? &outs[j]->v : in;
        if (p->markings != VA_VALID || p->prec != 4 || (p->type != VA_CCFFTIP && p->type != VA_CCFFTOP)) {
            va_error("batch_noalloc", "only complex FP32 ccfftmip/ccfftmop plans batch", (int)j);
            return -1;
        }
        t = va_tune_lookup(4, (int)p->MN, &s);
        if (t == -4) {
            va_error("batch_noalloc", "length has no cached batched schedule (vsip_amd_tune it first)", (int)p->MN);
            return -1;
        }
        if (t) {
            va_error("batch_noalloc", "length is not single pass in the batched shape", (int)p->MN);
            return -1;
        }
        if (va_fftm_descs(p, in, out, s, NULL, NULL, NULL, &descs[j]) != 1) return -1;
        ng = (descs[j].nvec + descs[j].vpb - 1) / descs[j].vpb;
        if (ni + ng > cap_items) {
            va_error("batch_noalloc", "more work items than the scratch holds", (int)cap_items);
            return -1;
        }
        for (g = 0; g < ng; g++) {
            items[ni].desc = (int)j;
            items[ni].pad_ = (int)p->MN;
            items[ni].vbase = g * descs[j].vpb;
            ni++;
        }
    }
    va_na_sort(items, ni);
    for (j = 0; j < ni; j++) items[j].pad_ = 0;
    used = njobs * sizeof(vhip_fftm_desc) + (size_t)ni * sizeof(vhip_batch_item);
    rc = vhip_memcpy_h2d_async(dev + VHIP_NOALLOC_HEAD, host + VHIP_NOALLOC_HEAD, used, vhip_stream());
    if (!rc) rc = vhip_noalloc_stage_done(slot, vhip_stream());
    va_na_turn ^= 1;
    /* the counters start at zero: cleared by a kernel, not a memset node */
    if (!rc) rc = vhip_batch_clear((unsigned *)dev, vhip_stream());
    if (!rc && ni)
        rc = vhip_batch_launch((const vhip_fftm_desc *)(dev + VHIP_NOALLOC_HEAD),
                               (const vhip_batch_item *)(dev + VHIP_NOALLOC_HEAD + njobs * sizeof(vhip_fftm_desc)),
                               ni, 0, (unsigned *)dev, VA_BATCH_THREADS, VA_BATCH_MAXE, VA_BATCH_LDS, va_na_grid,
                               vhip_stream());
    if (rc) va_error("batch_noalloc", vhip_error_string(rc), rc);
    return rc;
}
