/*
 * va_komega.c -- the k-omega beamformer step of the reference's example
 * (examples/komegaExamples/kw_c/kw.c:7-25) as two launches:
 *
 *   reference, seven passes over the data:
 *     vmmul(ts_taper, data, ROW)     taper every row (time)
 *     vmmul(array_taper, data, COL)  taper every column (sensors)
 *     rcfftmop ROW                   data -> freq
 *     ccfftmip COL                   freq in place
 *     mcmagsq -> real view of freq   |freq|^2
 *     smmul(1 / Navg)
 *     madd into gram
 *
 *   here:
 *     launch 1  rcfftmop ROW, each real sample times ts_taper on load
 *     launch 2  ccfftm COL over freq, each element times array_taper on
 *               load (the sensor taper commutes with the row transform),
 *               gram += gain * |y|^2 in place of the store
 *
 * The data is read once and never written (the reference tapers it in
 * place); freq holds the row transforms of the tapered data afterwards (the
 * reference leaves the powers in its real part).  gram is the result both
 * compute.
 *
 * C89.
 */
#include "va_internal.h"

static int va_vec_dev(const VA_mview *v, vsip_length len, const void **p, long *s)
{
    VA_block *b = v ? v->block : NULL;
    if (!b || !b->admitted || !b->dev_r || b->cplx || v->row_length != len) return -1;
    *p = (const char *)b->dev_r + (size_t)v->offset * (size_t)VA_RS(b) * (size_t)b->prec;
    *s = v->row_stride * VA_RS(b);
    return 0;
}

int vsip_amd_komega_f(const vsip_fftm_f *rcfftm, const vsip_fftm_f *ccfftm, const vsip_vview_f *ts_taper,
                      const vsip_vview_f *array_taper, const vsip_mview_f *data, const vsip_cmview_f *freq,
                      const vsip_mview_f *gram, vsip_scalar_f gain)
{
    const char *fn = "vsip_amd_komega_f";
    const VA_fftm *pr, *pc;
    const VA_mview *g;
    vhip_fftm_desc d[VA_MAX_DESC];
    int n, rc;
    if (va_ensure_init()) return -1;
    if (!rcfftm || !ccfftm || !ts_taper || !array_taper || !data || !freq || !gram) {
        va_error(fn, "null argument", -1);
        return -1;
    }
    pr = &rcfftm->p;
    pc = &ccfftm->p;
    g = &gram->v;
    if (pr->markings != VA_VALID || pc->markings != VA_VALID || pr->type != VA_RCFFTOP ||
        (pc->type != VA_CCFFTIP && pc->type != VA_CCFFTOP) || pr->major != VSIP_ROW || pc->major != VSIP_COL) {
        va_error(fn, "needs an rcfftmop ROW plan and a ccfftm COL plan", -1);
        return -1;
    }
    if (pr->npass != 1 || pc->npass != 1 || pr->staged || pc->staged) {
        va_error(fn, "a transform of this length runs in several passes; use the VSIPL calls", -1);
        return -1;
    }
    if (data->v.col_length != pr->M || data->v.row_length != pr->N || freq->v.col_length != pc->M ||
        freq->v.row_length != pc->N || pc->M != pr->M || pc->N != pr->N / 2 + 1) {
        va_error(fn, "view shapes do not match the plans", -1);
        return -1;
    }
    if (!g->block || g->block->cplx || VA_RS(g->block) != 1 || !g->block->admitted || !g->block->dev_r ||
        g->col_length != freq->v.col_length || g->row_length != freq->v.row_length ||
        g->col_stride != freq->v.col_stride || g->row_stride != freq->v.row_stride || freq->v.block->split) {
        va_error(fn, "gram must be an admitted real view with the strides of freq (interleaved)", -1);
        return -1;
    }

    /* launch 1: tapered rows -> freq */
    n = va_fftm_descs(pr, &data->v, &freq->v, NULL, NULL, NULL, NULL, d);
    if (n != 1) return -1;
    if (va_vec_dev(&ts_taper->v, pr->N, &d[0].tap, &d[0].tap_s)) {
        va_error(fn, "ts_taper must be an admitted vector of the row length", -1);
        return -1;
    }
    if (pr->jit_worth && vhip_jit_prepare_desc(&d[0]) == 0) d[0].compiled = 1;
    rc = vhip_fftm_launch(&d[0], vhip_stream());
    if (rc) { va_error(fn, vhip_error_string(rc), rc); return -1; }

    /* launch 2: tapered columns -> power accumulated into gram */
    n = va_fftm_descs(pc, &freq->v, &freq->v, NULL, NULL, NULL, NULL, d);
    if (n != 1) return -1;
    if (va_vec_dev(&array_taper->v, pc->M, &d[0].tap, &d[0].tap_s)) {
        va_error(fn, "array_taper must be an admitted vector of the column length", -1);
        return -1;
    }
    d[0].acc = (char *)g->block->dev_r + (size_t)g->offset * sizeof(float);
    d[0].acc_gain = gain;
    /* the column transform's schedule compiled with the taper and the power
     * accumulation (once per process; the generic kernel otherwise) */
    if (pc->jit_worth && vhip_jit_prepare_desc(&d[0]) == 0) d[0].compiled = 1;
    rc = vhip_fftm_launch(&d[0], vhip_stream());
    if (rc) { va_error(fn, vhip_error_string(rc), rc); return -1; }
    return 0;
}
