// fftm_launch.hip -- C ABI entry points that launch the multiple-FFT kernels.
#include "fftm_core.hpp"

extern "C" int vhip_cu_count(int *cus);

using vhipk::kernel_fn;

static kernel_fn pick_kernel(const vhip_fftm_desc *d) {
    if (d->prec == 4) return d->maxe <= 16 ? vhipk::fftm_kernel_f16(d->kind) : vhipk::fftm_kernel_f32(d->kind);
    if (d->prec == 8) return vhipk::fftm_kernel_d8(d->kind);
    return nullptr;
}

extern "C" size_t vhip_fftm_lds_bytes(const vhip_fftm_desc *d) {
    return (size_t)d->vpb * d->lds_stride * (size_t)(2 * d->prec) + (size_t)d->vpb * (8 + 8 + 4);
}

extern "C" int vhip_fftm_set_lds_limit(int bytes) {
    // allow every kernel variant to take up to `bytes` of dynamic LDS
    const int precs[2] = {4, 8};
    const int kinds[3] = {VHIP_C2C, VHIP_R2C, VHIP_C2R};
    const int maxes[2] = {8, 32};
    for (int p = 0; p < 2; ++p)
        for (int k = 0; k < 3; ++k)
            for (int m = 0; m < 2; ++m) {
                vhip_fftm_desc d;
                d.prec = precs[p];
                d.kind = kinds[k];
                d.maxe = maxes[m];
                kernel_fn f = pick_kernel(&d);
                if (!f) continue;
                hipError_t e = hipFuncSetAttribute((const void *)f,
                                                   hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
                if (e != hipSuccess) return (int)e;
            }
    int rc = vhipk::pow2_set_lds_limit(bytes);
    rc = rc ? rc : vhipk::pow2_real_set_lds_limit(bytes);
    return rc ? rc : vhipk::batch_set_lds_limit(bytes);
}

extern "C" int vhip_fftm_launch(const vhip_fftm_desc *d, void *stream) {
    if (!d || d->nvec <= 0 || d->vpb <= 0 || d->threads <= 0 || d->threads > 1024) return -1;
    int cus = 256;
    vhip_cu_count(&cus);
    // a plan that compiled its schedule runs it first (small power-of-two
    // rows compile because their compiled schedule beats the one-group
    // power-of-two kernel); the power-of-two kernels take the rest they fit
    int rc = d->compiled ? vhip_jit_try_launch(d, stream) : -100;
    if (rc != -100) return rc;
    rc = vhipk::pow2_try_launch(d, stream, cus > 0 ? cus : 256);
    if (rc != -100) return rc;
    kernel_fn f = pick_kernel(d);
    if (!f) return -2;
    const long grid = (d->nvec + d->vpb - 1) / d->vpb;
    if (grid > 0x7fffffffL) return -3;
    const size_t lds = vhip_fftm_lds_bytes(d);
    hipLaunchKernelGGL(f, dim3((unsigned)grid), dim3((unsigned)d->threads), lds,
                       (hipStream_t)stream, *d);
    return (int)hipGetLastError();
}
