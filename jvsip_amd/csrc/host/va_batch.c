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
