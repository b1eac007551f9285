/*
 * va_tune.c -- autotuner for the batched multiple-FFT kernel.
 *
 * The batched launch runs every transform of a batch with one fixed
 * workgroup shape (VA_BATCH_THREADS lanes, VA_BATCH_MAXE register class,
 * VA_BATCH_LDS bytes of LDS).  For each transform length the tuner times the
 * candidate schedules on the device and keeps the fastest:
 *   - factorisations: all factors packed into radices <= 16; the power-of-two
 *     part in stages of at most 16, 8 or 4 points with the odd stages before
 *     or after it;
 *   - vectors per workgroup: the most that fit, and half of that.
 * Each candidate runs as a one-transform batch over about 64 MiB of scratch
 * data, timed with HIP events.  Results live in a per-process table and, when
 * the environment variable VSIP_AMD_TUNE_CACHE names a file, in that file
 * (one line per length: "prec n vpb nst r0 r1 ...").  Lengths that do not
 * fit one workgroup return -2; the batch then splits them in two passes and
 * tunes each factor.
 *
 * C89.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "va_internal.h"

#define VA_TUNE_MAX 4096

typedef struct {
    int prec, n, rc;
    float ms;
    VA_sched s;          /* with device twiddles when rc == 0 */
} VA_tuned;

static VA_tuned *va_tuned_tab = NULL;
static int va_tuned_n = 0;
static int va_disk_loaded = 0;

typedef struct {
    vhip_fftm_desc *desc_dev;
    vhip_batch_item *items_dev;
    unsigned *ctl;
    long nitems;
    int grid;
    void *stream;
} VA_trial;

static int va_trial_run(void *arg)
{
    VA_trial *t = (VA_trial *)arg;
    return vhip_batch_launch(t->desc_dev, t->items_dev, t->nitems, 0, t->ctl, VA_BATCH_THREADS,
                             VA_BATCH_MAXE, VA_BATCH_LDS, t->grid, t->stream);
}

/* time schedule s (uploaded) on `nvec` vectors of scratch; returns ms or -1 */
static float va_time_sched(const VA_sched *s, int prec, void *scratch, long nvec, void *stream)
{
    vhip_fftm_desc d;
    vhip_batch_item *items;
    VA_trial tr;
    long ni, k;
    float ms = -1.f;
    int rc, i;
    memset(&d, 0, sizeof(d));
    memset(&tr, 0, sizeof(tr));
    d.prec = prec;
    d.kind = VHIP_C2C;
    d.n = s->n;
    d.nst = s->nst;
    for (i = 0; i < s->nst; i++) { d.radix[i] = s->radix[i]; d.tw_off[i] = s->tw_off[i]; }
    d.nvec = nvec;
    d.vpb = s->along.vpb;
    d.threads = VA_BATCH_THREADS;
    d.lds_stride = s->lds_stride;
    d.maxe = VA_BATCH_MAXE;
    d.map_in = d.map_out = VHIP_MAP_ALONG;
    d.scale = 1.0;
    d.tw = s->dev_tw;
    d.btw_off = s->btw_off;
    d.in.re = d.out.re = scratch;
    d.in.es = d.out.es = 1;
    d.in.vs0 = d.out.vs0 = s->n;
    d.in.nv0 = d.out.nv0 = nvec;
    d.in.contig = d.out.contig = 1;
    if (nvec * s->n * 2L * prec < (1L << 31) - 64) d.in.span = d.out.span = nvec * s->n * 2L * prec;
    ni = (nvec + d.vpb - 1) / d.vpb;
    items = (vhip_batch_item *)malloc(sizeof(vhip_batch_item) * (size_t)ni);
    if (!items) return -1.f;
    for (k = 0; k < ni; k++) {
        items[k].desc = 0;
        items[k].pad_ = 0;
        items[k].vbase = k * d.vpb;
    }
    rc = vhip_malloc((void **)&tr.desc_dev, sizeof(d));
    if (!rc) rc = vhip_malloc((void **)&tr.items_dev, sizeof(vhip_batch_item) * (size_t)ni);
    if (!rc) rc = vhip_malloc((void **)&tr.ctl, 64);
    if (!rc) rc = vhip_memset(tr.ctl, 0, 64);
    if (!rc) rc = vhip_memcpy_h2d(tr.desc_dev, &d, sizeof(d));
    if (!rc) rc = vhip_memcpy_h2d(tr.items_dev, items, sizeof(vhip_batch_item) * (size_t)ni);
    if (!rc) rc = vhip_batch_max_grid(VA_BATCH_THREADS, VA_BATCH_MAXE, VA_BATCH_LDS, &tr.grid);
    tr.nitems = ni;
    tr.stream = stream;
    if (!rc) rc = vhip_event_time(stream, va_trial_run, &tr, 3, &ms);
    if (rc) ms = -1.f;
    vhip_sync();
    if (tr.desc_dev) vhip_free(tr.desc_dev);
    if (tr.items_dev) vhip_free(tr.items_dev);
    if (tr.ctl) vhip_free(tr.ctl);
    free(items);
    return ms;
}

static void va_disk_load(void)
{
    const char *path = getenv("VSIP_AMD_TUNE_CACHE");
    FILE *f;
    int prec, n, vpb, nst, r[VHIP_MAX_STAGES], i, ok;
    va_disk_loaded = 1;
    if (!path || !(f = fopen(path, "r"))) return;
    while (fscanf(f, "%d %d %d %d", &prec, &n, &vpb, &nst) == 4) {
        VA_tuned *t;
        ok = (nst >= 1 && nst <= VHIP_MAX_STAGES);
        for (i = 0; ok && i < nst; i++) ok = (fscanf(f, "%d", &r[i]) == 1);
        if (!ok || va_tuned_n >= VA_TUNE_MAX) break;
        t = &va_tuned_tab[va_tuned_n];
        memset(t, 0, sizeof(*t));
        if (va_sched_shape_fixed(&t->s, prec, n, r, nst, VA_BATCH_THREADS, VA_BATCH_MAXE,
                                 VA_BATCH_LDS, vpb) || va_sched_upload(&t->s, prec)) {
            va_sched_free(&t->s);
            continue;
        }
        t->prec = prec;
        t->n = n;
        t->rc = 0;
        t->ms = 0.f;
        va_tuned_n++;
    }
    fclose(f);
}

static void va_disk_append(const VA_tuned *t)
{
    const char *path = getenv("VSIP_AMD_TUNE_CACHE");
    FILE *f;
    int i;
    if (!path || !(f = fopen(path, "a"))) return;
    fprintf(f, "%d %d %d %d", t->prec, t->n, t->s.along.vpb, t->s.nst);
    for (i = 0; i < t->s.nst; i++) fprintf(f, " %d", t->s.radix[i]);
    fprintf(f, "\n");
    fclose(f);
}

/*
 * Tuned batched schedule for length n: 0 and *out (owned by the tuner),
 * -1 unsupported length, -2 does not fit one workgroup, -3 device error.
 */
int va_tune_get(int prec, int n, const VA_sched **out)
{
    int i, mb, of, h, nc = 0, rad[VHIP_MAX_STAGES], nst, best = -1, vmax, rc;
    static const int bits_f[3] = {4, 3, 2};
    VA_sched cand[16];
    float ms[16];
    void *scratch = NULL, *stream = NULL;
    long nvec;
    VA_tuned *t;
    *out = NULL;
    if (!va_tuned_tab) {
        va_tuned_tab = (VA_tuned *)calloc(VA_TUNE_MAX, sizeof(VA_tuned));
        if (!va_tuned_tab) return -3;
    }
    if (!va_disk_loaded) va_disk_load();
    for (i = 0; i < va_tuned_n; i++) {
        if (va_tuned_tab[i].prec == prec && va_tuned_tab[i].n == n) {
            if (va_tuned_tab[i].rc) return va_tuned_tab[i].rc;
            *out = &va_tuned_tab[i].s;
            return 0;
        }
    }
    if (va_tuned_n >= VA_TUNE_MAX) return -3;
    t = &va_tuned_tab[va_tuned_n];
    memset(t, 0, sizeof(*t));
    t->prec = prec;
    t->n = n;

    /* power-of-two lengths 256..8192 take the batched kernel's specialised
     * path, whose schedule is fixed: even radix-16/8/4/2 split, 8192/n rows */
    if (prec == 4 && n >= 256 && n <= 8192 && (n & (n - 1)) == 0) {
        int r2[VHIP_MAX_STAGES], k2 = va_factor_variant(n, prec, 4, 0, r2);
        if (va_sched_shape_fixed(&t->s, prec, n, r2, k2, VA_BATCH_THREADS, VA_BATCH_MAXE, VA_BATCH_LDS,
                                 8192 / n) == 0 && t->s.along.vpb == 8192 / n && va_sched_upload(&t->s, prec) == 0) {
            t->rc = 0;
            va_tuned_n++;
            *out = &t->s;
            return 0;
        }
        va_sched_free(&t->s);
        memset(&t->s, 0, sizeof(t->s));
    }

    /* candidates */
    vmax = (VA_BATCH_THREADS * VA_BATCH_MAXE) / n;
    if (vmax < 1) vmax = 1;
    for (mb = -1; mb < 3; mb++) {
        for (of = 0; of < 2; of++) {
            int dup = 0, c;
            if (mb < 0 && (of || prec != 4)) continue;
            nst = (mb < 0) ? va_factor_pack(n, rad) : va_factor_variant(n, prec, bits_f[mb], of, rad);
            if (nst < 0) { t->rc = -1; va_tuned_n++; return -1; }
            for (c = 0; c < nc && !dup; c++) {
                int same = (cand[c].nst == nst), q;
                for (q = 0; same && q < nst; q++) same = (cand[c].radix[q] == rad[q]);
                dup = same;
            }
            if (dup) continue;
            for (h = 0; h < 2; h++) {
                int vcap = h ? vmax / 2 : vmax;
                if (vcap < 1) continue;
                if (va_sched_shape_fixed(&cand[nc], prec, n, rad, nst, VA_BATCH_THREADS, VA_BATCH_MAXE,
                                         VA_BATCH_LDS, vcap))
                    continue;
                if (h && cand[nc].along.vpb == cand[nc - 1].along.vpb) continue;
                nc++;
            }
        }
    }
    if (nc == 0) { t->rc = -2; va_tuned_n++; return -2; }

    /* time them */
    nvec = (8L << 20) / n;          /* 64 MiB of complex float */
    if (nvec < 1) nvec = 1;
    rc = vhip_malloc(&scratch, (size_t)nvec * n * 2 * prec);
    if (!rc) rc = vhip_stream_create(&stream);
    if (!rc) rc = vhip_memset(scratch, 0, (size_t)nvec * n * 2 * prec);
    for (i = 0; i < nc && !rc; i++) {
        ms[i] = -1.f;
        if (va_sched_upload(&cand[i], prec)) continue;
        ms[i] = va_time_sched(&cand[i], prec, scratch, nvec, stream);
        if (ms[i] >= 0.f && (best < 0 || ms[i] < ms[best])) best = i;
    }
    for (i = 0; i < nc; i++)
        if (i != best) va_sched_free(&cand[i]);
    if (stream) vhip_stream_destroy(stream);
    if (scratch) vhip_free(scratch);
    if (best < 0) {
        /* no device timing possible: keep the default factorisation */
        if (va_sched_shape_fixed(&t->s, prec, n, NULL, 0, VA_BATCH_THREADS, VA_BATCH_MAXE,
                                 VA_BATCH_LDS, vmax) || va_sched_upload(&t->s, prec)) {
            va_sched_free(&t->s);
            return -3;
        }
    } else {
        t->s = cand[best];
        t->ms = ms[best];
    }
    t->rc = 0;
    va_tuned_n++;
    va_disk_append(t);
    *out = &t->s;
    return 0;
}

/*
 * The cached batched schedule of length n, without tuning or uploading
 * anything: 0 and *out, -2 cached as not fitting one workgroup, -1 cached as
 * unsupported, -4 not in the cache yet.
 */
int va_tune_lookup(int prec, int n, const VA_sched **out)
{
    int i;
    *out = NULL;
    for (i = 0; va_tuned_tab && i < va_tuned_n; i++) {
        if (va_tuned_tab[i].prec == prec && va_tuned_tab[i].n == n) {
            if (va_tuned_tab[i].rc) return va_tuned_tab[i].rc;
            *out = &va_tuned_tab[i].s;
            return 0;
        }
    }
    return -4;
}

/* describe the tuned schedule of length n as JSON */
int vsip_amd_tune(int prec, long n, char *buf, size_t len)
{
    const VA_sched *s = NULL;
    char tmp[512];
    int pos = 0, i, rc;
    if (va_ensure_init()) return -3;
    rc = va_tune_get(prec, (int)n, &s);
    pos += sprintf(tmp + pos, "{\"rc\": %d", rc);
    if (!rc) {
        pos += sprintf(tmp + pos, ", \"n\": %d, \"vpb\": %d, \"threads\": %d, \"radix\": [",
                       s->n, s->along.vpb, s->along.threads);
        for (i = 0; i < s->nst && pos < 400; i++) pos += sprintf(tmp + pos, "%s%d", i ? ", " : "", s->radix[i]);
        pos += sprintf(tmp + pos, "]");
        for (i = 0; i < va_tuned_n; i++)
            if (va_tuned_tab[i].prec == prec && va_tuned_tab[i].n == n)
                pos += sprintf(tmp + pos, ", \"ms\": %.5f", va_tuned_tab[i].ms);
    }
    pos += sprintf(tmp + pos, "}");
    if (buf && len) {
        strncpy(buf, tmp, len - 1);
        buf[len - 1] = '\0';
    }
    return rc;
}
