/*
 * va_lib.c -- library lifecycle, error reporting and the vsip_amd_*
 * extensions.
 *
 * vsip_init / vsip_finalize are no-ops in the reference (vsip_init.c:17-25,
 * vsip_finalize.c:16-26); here vsip_init brings up the device and every
 * constructor calls it lazily, so programs that never call it still work.
 *
 * C89.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "va_internal.h"
#include "vsip_amd.h"

static int va_ready = 0;
static int va_init_rc = 0;
static char va_last[256] = "";

int va_ensure_init(void)
{
    if (va_ready) return va_init_rc;
    va_init_rc = vhip_init();
    va_ready = 1;
    if (va_init_rc) va_error("vsip_init", vhip_error_string(va_init_rc), va_init_rc);
    return va_init_rc;
}

void va_error(const char *where, const char *what, int code)
{
    sprintf(va_last, "%.60s: %.150s (code %d)", where, what ? what : "", code);
    if (getenv("VSIP_AMD_QUIET") == NULL) fprintf(stderr, "vsip_amd: %s\n", va_last);
}

int vsip_init(void *ptr)
{
    (void)ptr;
    return va_ensure_init() ? 1 : 0;
}

int vsip_finalize(void *ptr)
{
    (void)ptr;
    if (va_ready && !va_init_rc) vhip_sync();
    return 0;
}

void *vsip_amd_stream(void)
{
    if (va_ensure_init()) return NULL;
    return vhip_stream();
}

int vsip_amd_sync(void)
{
    if (va_ensure_init()) return va_init_rc;
    return vhip_sync();
}

const char *vsip_amd_last_error(void) { return va_last; }
void vsip_amd_clear_error(void) { va_last[0] = '\0'; }
size_t vsip_amd_bytes_in_use(void) { return vhip_bytes_in_use(); }
void vsip_amd_set_hbm_budget(size_t bytes) { vhip_set_budget(bytes); }

void *vsip_amd_cblock_device_f(const vsip_cblock_f *b, void **im)
{
    if (im) *im = b ? b->b.dev_i : NULL;
    return b ? b->b.dev_r : NULL;
}

void *vsip_amd_cblock_device_d(const vsip_cblock_d *b, void **im)
{
    if (im) *im = b ? b->b.dev_i : NULL;
    return b ? b->b.dev_r : NULL;
}

void *vsip_amd_block_device_f(const vsip_block_f *b) { return b ? b->b.dev_r : NULL; }
void *vsip_amd_block_device_d(const vsip_block_d *b) { return b ? b->b.dev_r : NULL; }

static int va_describe_passes(const VA_fftm *p, char *tmp)
{
    int pos = 0, k, s;
    pos += sprintf(tmp + pos, "\"npass\": %d, \"passes\": [", p->npass);
    for (k = 0; k < p->npass; k++) {
        const VA_sched *sc = &p->pass[k];
        pos += sprintf(tmp + pos, "%s{\"n\": %d, \"radix\": [", k ? ", " : "", sc->n);
        for (s = 0; s < sc->nst; s++)
            pos += sprintf(tmp + pos, "%s%d", s ? ", " : "", sc->radix[s]);
        pos += sprintf(tmp + pos, "], \"along\": [%d, %d, %d], \"across\": [%d, %d, %d], "
                       "\"lds_stride\": %d}", sc->along.vpb, sc->along.threads, sc->along.maxe,
                       sc->across.vpb, sc->across.threads, sc->across.maxe, sc->lds_stride);
    }
    pos += sprintf(tmp + pos, "]");
    return pos;
}

static int va_describe(const VA_fftm *p, char *buf, size_t len)
{
    /* at most 2 + 2 passes of <= 24 stages each: well under 4096 bytes */
    char tmp[4096];
    int pos = 0;
    static const char *types[4] = {"ccfftmip", "ccfftmop", "rcfftmop", "crfftmop"};
    pos += sprintf(tmp + pos, "{\"type\": \"%s\", \"prec\": %d, \"MN\": %lu, \"mN\": %lu, ",
                   types[p->type & 3], p->prec, (unsigned long)p->MN, (unsigned long)p->mN);
    pos += va_describe_passes(p, tmp + pos);
    if (p->npass >= 1)
        pos += sprintf(tmp + pos, ", \"compiled_schedule\": %s", (p->jit && va_jit_ready(p)) ? "true" : "false");
    if (p->staged && p->inner) {
        pos += sprintf(tmp + pos, ", \"staged\": \"%s\", \"scratch_length\": %lu, \"chunk\": %lu, "
                       "\"inner\": {", p->staged == 2 ? "bluestein" : "real", (unsigned long)p->sm,
                       (unsigned long)p->chunk);
        pos += va_describe_passes(p->inner, tmp + pos);
        pos += sprintf(tmp + pos, "}");
    }
    pos += sprintf(tmp + pos, "}");
    if (buf && len) {
        strncpy(buf, tmp, len - 1);
        buf[len - 1] = '\0';
    }
    return pos;
}

typedef struct {
    void *buf[4], *out[4];
    size_t bytes;
    int kind, i;
    void *sink;
    long rows, cols, ld_in, ld_out;   /* kind 2: the column probe */
    int lg;
} VA_probe;

static int va_probe_step(void *arg)
{
    VA_probe *p = (VA_probe *)arg;
    int rc;
    if (p->kind == 2)
        rc = vhip_colprobe_launch(p->buf[p->i & 3], p->out[p->i & 3], (int)p->rows, (int)p->cols, p->ld_in,
                                  p->ld_out, p->lg, vhip_stream());
    else rc = vhip_probe_launch(p->kind, p->buf[p->i & 3], p->bytes, p->sink, vhip_stream());
    p->i++;
    return rc;
}

int vsip_amd_col_probe(long rows, long cols, long in_row_stride, long out_row_stride, int cols_per_group,
                       int reps, double *gbs)
{
    VA_probe p;
    float ms = 0.f;
    int k, rc = 0, lg = 0;
    size_t bin, bout;
    *gbs = 0.0;
    while (lg < 5 && (1 << lg) < cols_per_group) lg++;
    if (rows < 1 || cols < 1 || rows > 0x7fffffffL || cols > 0x7fffffffL || lg > 4 || (1 << lg) != cols_per_group ||
        (rows << lg) > 32L * 512 || in_row_stride < cols || out_row_stride < cols || reps < 1)
        return -1;
    if (va_ensure_init()) return va_init_rc;
    bin = (size_t)((rows - 1) * in_row_stride + cols) * 8;
    bout = (size_t)((rows - 1) * out_row_stride + cols) * 8;
    memset(&p, 0, sizeof(p));
    p.kind = 2;
    p.rows = rows;
    p.cols = cols;
    p.ld_in = in_row_stride;
    p.ld_out = out_row_stride;
    p.lg = lg;
    for (k = 0; k < 4 && !rc; k++) {
        rc = vhip_malloc(&p.buf[k], bin);
        if (!rc) rc = vhip_memset(p.buf[k], 0, bin);
        if (!rc) rc = vhip_malloc(&p.out[k], bout);
    }
    if (!rc) rc = vhip_event_time(vhip_stream(), va_probe_step, &p, reps, &ms);
    if (!rc && ms > 0.f) *gbs = 2.0 * 8.0 * (double)rows * (double)cols / (ms * 1e-3) / 1e9;
    vhip_sync();
    for (k = 0; k < 4; k++) {
        if (p.buf[k]) vhip_free(p.buf[k]);
        if (p.out[k]) vhip_free(p.out[k]);
    }
    return rc;
}

int vsip_amd_hbm_probe(size_t bytes, int reps, double *rmw_gbs, double *read_gbs)
{
    VA_probe p;
    float ms = 0.f;
    int k, rc = 0;
    *rmw_gbs = *read_gbs = 0.0;
    bytes -= bytes % 32768;               /* the in-place probe moves whole 32 KiB rows */
    if (bytes == 0) return -1;
    if (va_ensure_init()) return va_init_rc;
    memset(&p, 0, sizeof(p));
    p.bytes = bytes;
    for (k = 0; k < 4 && !rc; k++) {
        rc = vhip_malloc(&p.buf[k], bytes);
        if (!rc) rc = vhip_memset(p.buf[k], 0, bytes);
    }
    if (!rc) rc = vhip_malloc(&p.sink, 256);
    if (!rc) {
        p.kind = 0;
        rc = vhip_event_time(vhip_stream(), va_probe_step, &p, reps, &ms);
        if (!rc && ms > 0.f) *rmw_gbs = 2.0 * (double)bytes / (ms * 1e-3) / 1e9;
    }
    if (!rc) {
        p.kind = 1;
        rc = vhip_event_time(vhip_stream(), va_probe_step, &p, reps, &ms);
        if (!rc && ms > 0.f) *read_gbs = (double)bytes / (ms * 1e-3) / 1e9;
    }
    vhip_sync();
    for (k = 0; k < 4; k++)
        if (p.buf[k]) vhip_free(p.buf[k]);
    if (p.sink) vhip_free(p.sink);
    return rc;
}

int vsip_amd_schedule(int prec, long n, long nvec, int lds_max, char *buf, size_t len)
{
    VA_sched sc;
    char tmp[1024];
    int pos = 0, s, rc;
    rc = va_sched_shape(&sc, prec, (int)n, (vsip_length)nvec, lds_max);
    pos += sprintf(tmp + pos, "{\"rc\": %d", rc);
    if (rc == 0) {
        pos += sprintf(tmp + pos, ", \"n\": %d, \"radix\": [", sc.n);
        for (s = 0; s < sc.nst && pos < 900; s++)
            pos += sprintf(tmp + pos, "%s%d", s ? ", " : "", sc.radix[s]);
        pos += sprintf(tmp + pos, "], \"along\": [%d, %d, %d], \"across\": [%d, %d, %d], "
                       "\"lds_stride\": %d, \"tw_len\": %d", sc.along.vpb, sc.along.threads,
                       sc.along.maxe, sc.across.vpb, sc.across.threads, sc.across.maxe,
                       sc.lds_stride, sc.tw_len);
    }
    pos += sprintf(tmp + pos, "}");
    if (buf && len) {
        strncpy(buf, tmp, len - 1);
        buf[len - 1] = '\0';
    }
    return rc;
}

int vsip_amd_fftm_describe_f(const vsip_fftm_f *plan, char *buf, size_t len)
{
    return va_describe(&plan->p, buf, len);
}

int vsip_amd_fftm_describe_d(const vsip_fftm_d *plan, char *buf, size_t len)
{
    return va_describe(&plan->p, buf, len);
}
