/*
 * va_fft1d.c -- the VSIPL 1-D FFT: vsip_ccfftip, vsip_ccfftop,
 * vsip_rcfftop, vsip_crfftop (_f and _d).
 *
 * The reference computes the 1-D transforms with the same DFT conventions as
 * its multiple FFT (VI_ccfftip_f_def.h, VI_ccfftop_f_def.h,
 * VI_rcfftop_f_def.h, VI_crfftop_f_def.h: sign dir, times scale, no other
 * normalisation; rcfftop returns N/2 + 1 bins of an even-length real
 * vector; crfftop applies the same unnormalised Hermitian pre-combine as
 * crfftmop, VI_crfftop_f_def.h:47-80).  A 1-D plan here is the multiple-FFT
 * plan of one ROW vector, so every length, view stride and precision the
 * multiple FFT handles (single pass, two-pass split, staged Bluestein, real
 * wrap) serves the 1-D calls through the same kernels.
 *   create     vsip.h:1801-1808 (NULL on failure, odd N refused for rc/cr)
 *   execute    vsip.h:712-719, 828-829, 1185-1186
 *   destroy    vsip.h:423-424 (returns 0)
 *   getattr    vsip_fft_getattr_f.c:18-55 (input/output lengths per type)
 *
 * C89.
 */
#include <stdlib.h>
#include "va_internal.h"

static VA_fftm *va_new_fft(int prec, int type, vsip_length n, double scale, int dir,
                           unsigned int ntimes, vsip_alg_hint hint)
{
    VA_fftm *p = (VA_fftm *)malloc(sizeof(VA_fftm));
    if (!p) return NULL;
    if (va_fftm_init(p, prec, type, 1, n, scale, dir, VSIP_ROW, ntimes, hint)) {
        va_fftm_release(p);
        free(p);
        return NULL;
    }
    return p;
}

static void va_fft_check(const VA_fftm *p, int type, const char *fn)
{
    if (p && p->type != type) va_error(fn, "plan was created for another transform", -1);
}

/* ---- create ------------------------------------------------------------- */

vsip_fft_f *vsip_ccfftip_create_f(vsip_length n, vsip_scalar_f scale, vsip_fft_dir dir,
                                  unsigned int ntimes, vsip_alg_hint hint)
{
    return (vsip_fft_f *)va_new_fft(4, VA_CCFFTIP, n, scale, dir, ntimes, hint);
}

vsip_fft_d *vsip_ccfftip_create_d(vsip_length n, vsip_scalar_d scale, vsip_fft_dir dir,
                                  unsigned int ntimes, vsip_alg_hint hint)
{
    return (vsip_fft_d *)va_new_fft(8, VA_CCFFTIP, n, scale, dir, ntimes, hint);
}

vsip_fft_f *vsip_ccfftop_create_f(vsip_length n, vsip_scalar_f scale, vsip_fft_dir dir,
                                  unsigned int ntimes, vsip_alg_hint hint)
{
    return (vsip_fft_f *)va_new_fft(4, VA_CCFFTOP, n, scale, dir, ntimes, hint);
}

vsip_fft_d *vsip_ccfftop_create_d(vsip_length n, vsip_scalar_d scale, vsip_fft_dir dir,
                                  unsigned int ntimes, vsip_alg_hint hint)
{
    return (vsip_fft_d *)va_new_fft(8, VA_CCFFTOP, n, scale, dir, ntimes, hint);
}

vsip_fft_f *vsip_rcfftop_create_f(vsip_length n, vsip_scalar_f scale, unsigned int ntimes,
                                  vsip_alg_hint hint)
{
    return (vsip_fft_f *)va_new_fft(4, VA_RCFFTOP, n, scale, VSIP_FFT_FWD, ntimes, hint);
}

vsip_fft_d *vsip_rcfftop_create_d(vsip_length n, vsip_scalar_d scale, unsigned int ntimes,
                                  vsip_alg_hint hint)
{
    return (vsip_fft_d *)va_new_fft(8, VA_RCFFTOP, n, scale, VSIP_FFT_FWD, ntimes, hint);
}

vsip_fft_f *vsip_crfftop_create_f(vsip_length n, vsip_scalar_f scale, unsigned int ntimes,
                                  vsip_alg_hint hint)
{
    return (vsip_fft_f *)va_new_fft(4, VA_CRFFTOP, n, scale, VSIP_FFT_INV, ntimes, hint);
}

vsip_fft_d *vsip_crfftop_create_d(vsip_length n, vsip_scalar_d scale, unsigned int ntimes,
                                  vsip_alg_hint hint)
{
    return (vsip_fft_d *)va_new_fft(8, VA_CRFFTOP, n, scale, VSIP_FFT_INV, ntimes, hint);
}

/* ---- execute ------------------------------------------------------------ */

void vsip_ccfftip_f(const vsip_fft_f *plan, const vsip_cvview_f *x)
{
    va_fft_check(&plan->p, VA_CCFFTIP, "vsip_ccfftip_f");
    va_fftm_exec(&plan->p, &x->v, &x->v, vhip_stream());
}

void vsip_ccfftip_d(const vsip_fft_d *plan, const vsip_cvview_d *x)
{
    va_fft_check(&plan->p, VA_CCFFTIP, "vsip_ccfftip_d");
    va_fftm_exec(&plan->p, &x->v, &x->v, vhip_stream());
}

void vsip_ccfftop_f(const vsip_fft_f *plan, const vsip_cvview_f *x, const vsip_cvview_f *y)
{
    va_fft_check(&plan->p, VA_CCFFTOP, "vsip_ccfftop_f");
    va_fftm_exec(&plan->p, &x->v, &y->v, vhip_stream());
}

void vsip_ccfftop_d(const vsip_fft_d *plan, const vsip_cvview_d *x, const vsip_cvview_d *y)
{
    va_fft_check(&plan->p, VA_CCFFTOP, "vsip_ccfftop_d");
    va_fftm_exec(&plan->p, &x->v, &y->v, vhip_stream());
}

void vsip_rcfftop_f(const vsip_fft_f *plan, const vsip_vview_f *x, const vsip_cvview_f *z)
{
    va_fft_check(&plan->p, VA_RCFFTOP, "vsip_rcfftop_f");
    va_fftm_exec(&plan->p, &x->v, &z->v, vhip_stream());
}

void vsip_rcfftop_d(const vsip_fft_d *plan, const vsip_vview_d *x, const vsip_cvview_d *z)
{
    va_fft_check(&plan->p, VA_RCFFTOP, "vsip_rcfftop_d");
    va_fftm_exec(&plan->p, &x->v, &z->v, vhip_stream());
}

void vsip_crfftop_f(const vsip_fft_f *plan, const vsip_cvview_f *z, const vsip_vview_f *x)
{
    va_fft_check(&plan->p, VA_CRFFTOP, "vsip_crfftop_f");
    va_fftm_exec(&plan->p, &z->v, &x->v, vhip_stream());
}

void vsip_crfftop_d(const vsip_fft_d *plan, const vsip_cvview_d *z, const vsip_vview_d *x)
{
    va_fft_check(&plan->p, VA_CRFFTOP, "vsip_crfftop_d");
    va_fftm_exec(&plan->p, &z->v, &x->v, vhip_stream());
}

/* ---- destroy / getattr ---------------------------------------------------- */

int vsip_fft_destroy_f(vsip_fft_f *plan)
{
    if (plan) {
        va_fftm_release(&plan->p);
        free(plan);
    }
    return 0;
}

int vsip_fft_destroy_d(vsip_fft_d *plan)
{
    if (plan) {
        va_fftm_release(&plan->p);
        free(plan);
    }
    return 0;
}

static void va_fft_attr(const VA_fftm *p, vsip_scalar_vi *in, vsip_scalar_vi *out, vsip_fft_place *place,
                        vsip_fft_dir *dir)
{
    *in = *out = p->MN;
    *place = (p->type == VA_CCFFTIP) ? VSIP_FFT_IP : VSIP_FFT_OP;
    *dir = (p->dir < 0) ? VSIP_FFT_FWD : VSIP_FFT_INV;
    if (p->type == VA_RCFFTOP) {
        *in = 2 * p->MN;
        *out = p->MN + 1;
    } else if (p->type == VA_CRFFTOP) {
        *in = p->MN + 1;
        *out = 2 * p->MN;
    }
}

void vsip_fft_getattr_f(const vsip_fft_f *plan, vsip_fft_attr_f *a)
{
    va_fft_attr(&plan->p, &a->input, &a->output, &a->place, &a->dir);
    a->scale = (vsip_scalar_f)plan->p.scale;
}

void vsip_fft_getattr_d(const vsip_fft_d *plan, vsip_fft_attr_d *a)
{
    va_fft_attr(&plan->p, &a->input, &a->output, &a->place, &a->dir);
    a->scale = plan->p.scale;
}
