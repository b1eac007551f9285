/*
 * vsip.h -- VSIPL 1.3 C API, multiple-FFT subset, for the MI355X implementation.
 *
 * Types, enum values and prototypes keep the VSIPL C API surface of the
 * reference header c_VSIP_src/vsip.h so that callers recompile unchanged
 * (the multiple FFT, its storage calls, and the 1-D FFT with its vector
 * views):
 *   scalar typedefs            vsip.h:124-135
 *   enums                      vsip.h:138-153  (vsip_memory_hint, vsip_major,
 *                                               vsip_fft_dir, vsip_fft_place,
 *                                               vsip_alg_hint)
 *   opaque handles             vsip.h:165-208
 *   vsip_fftm_attr_{f,d}        vsip.h:277-278
 *   attribute structs          vsip.h:242-245 (vsip_mattr_*, vsip_cmattr_*)
 * Each prototype below names the reference declaration it stands for.
 *
 * Storage model: a block's data lives in HBM while the block is admitted.
 * Binding user memory (vsip_*blockbind_*) records the user's arrays; admitting
 * with update = true copies them to the device, releasing with update = true
 * copies the device data back.  Blocks made by vsip_*blockcreate_* (and the
 * views made by vsip_*mcreate_*) live on the device only and are born
 * admitted.  FFT execution is asynchronous on the library stream; release,
 * get/put and copyto/copyfrom wait for it.
 *
 * The header is C89.
 */
#ifndef _vsip_h
#define _vsip_h 1

#ifdef __cplusplus
extern "C" {
#endif

/* ---- scalar types (vsip.h:124-135) ---------------------------------- */
typedef signed int vsip_scalar_bl;
typedef vsip_scalar_bl vsip_bool;
typedef float vsip_scalar_f;
typedef double vsip_scalar_d;
typedef unsigned long int vsip_scalar_vi;
typedef vsip_scalar_vi vsip_offset;
typedef vsip_scalar_vi vsip_length;
typedef vsip_scalar_vi vsip_index;
typedef signed long int vsip_stride;
typedef struct { vsip_scalar_vi r, c; } vsip_scalar_mi;
typedef struct { vsip_scalar_d r, i; } vsip_cscalar_d;
typedef struct { vsip_scalar_f r, i; } vsip_cscalar_f;
typedef unsigned int vsip_scalar_ue32;                                   /* vsip.h:123 */

#define VSIP_TRUE 1
#define VSIP_FALSE 0

/* ---- enums (vsip.h:138-153) ----------------------------------------- */
typedef enum { VSIP_MEM_NONE = 0, VSIP_MEM_RDONLY = 1, VSIP_MEM_CONST = 2,
               VSIP_MEM_SHARED = 3, VSIP_MEM_SHARED_RDONLY = 4,
               VSIP_MEM_SHARED_CONST = 5 } vsip_memory_hint;
typedef enum { VSIP_CMPLX_INTERLEAVED, VSIP_CMPLX_SPLIT, VSIP_CMPLX_NONE } vsip_cmplx_mem;
typedef enum { VSIP_ROW = 0, VSIP_COL = 1 } vsip_major;
typedef enum { VSIP_FFT_FWD = -1, VSIP_FFT_INV = +1 } vsip_fft_dir;
typedef enum { VSIP_FFT_IP = 0, VSIP_FFT_OP = 1 } vsip_fft_place;
typedef enum { VSIP_ALG_SPACE = 0, VSIP_ALG_TIME = 1, VSIP_ALG_NOISE = 2 } vsip_alg_hint;
typedef enum { VSIP_PRNG = 0, VSIP_NPRNG = 1 } vsip_rng;                  /* vsip.h:159 */
typedef enum { VSIP_SUPPORT_FULL = 0, VSIP_SUPPORT_SAME = 1, VSIP_SUPPORT_MIN = 2 } vsip_support_region; /* vsip.h:155 */
typedef enum { VSIP_BIASED = 0, VSIP_UNBIASED = 1 } vsip_bias;          /* vsip.h:156 */
typedef enum { VSIP_NONSYM = 0, VSIP_SYM_EVEN_LEN_ODD = 1, VSIP_SYM_EVEN_LEN_EVEN = 2 } vsip_symmetry; /* vsip.h:154 */

/* ---- opaque handles (vsip.h:165-208) -------------------------------- */
typedef struct vsip_blockattributes_d vsip_block_d;
typedef struct vsip_cblockattributes_d vsip_cblock_d;
typedef struct vsip_blockattributes_f vsip_block_f;
typedef struct vsip_cblockattributes_f vsip_cblock_f;
typedef struct vsip_mviewattributes_d vsip_mview_d;
typedef struct vsip_mviewattributes_f vsip_mview_f;
typedef struct vsip_cmviewattributes_d vsip_cmview_d;
typedef struct vsip_cmviewattributes_f vsip_cmview_f;
typedef struct vsip_fftmattributes_f vsip_fftm_f;
typedef struct vsip_fftmattributes_d vsip_fftm_d;
typedef struct vsip_vviewattributes_d vsip_vview_d;
typedef struct vsip_vviewattributes_f vsip_vview_f;
typedef struct vsip_cvviewattributes_d vsip_cvview_d;
typedef struct vsip_cvviewattributes_f vsip_cvview_f;
typedef struct vsip_fftattributes_f vsip_fft_f;
typedef struct vsip_fftattributes_d vsip_fft_d;
typedef struct vsip_randobject vsip_randstate;                           /* vsip.h:191 */
typedef struct vsip_corr1dattributes_f vsip_corr1d_f;                    /* vsip.h:215 */
typedef struct vsip_ccorr1dattributes_f vsip_ccorr1d_f;                  /* vsip.h:216 */
typedef struct vsip_corr1dattributes_d vsip_corr1d_d;                    /* vsip.h:217 */
typedef struct vsip_ccorr1dattributes_d vsip_ccorr1d_d;                  /* vsip.h:218 */
typedef struct vsip_conv1dattributes_f vsip_conv1d_f;                    /* vsip.h:219 */
typedef struct vsip_conv1dattributes_d vsip_conv1d_d;                    /* vsip.h:220 */

/* ---- attribute structs (vsip.h:242-245, 277-278) -------------------- */
typedef struct { vsip_offset offset; vsip_stride row_stride; vsip_length row_length;
                 vsip_stride col_stride; vsip_length col_length; vsip_block_d* block; } vsip_mattr_d;
typedef struct { vsip_offset offset; vsip_stride row_stride; vsip_length row_length;
                 vsip_stride col_stride; vsip_length col_length; vsip_block_f* block; } vsip_mattr_f;
typedef struct { vsip_offset offset; vsip_stride row_stride; vsip_length row_length;
                 vsip_stride col_stride; vsip_length col_length; vsip_cblock_d* block; } vsip_cmattr_d;
typedef struct { vsip_offset offset; vsip_stride row_stride; vsip_length row_length;
                 vsip_stride col_stride; vsip_length col_length; vsip_cblock_f* block; } vsip_cmattr_f;
typedef struct { vsip_scalar_mi input; vsip_scalar_mi output; vsip_fft_place place;
                 vsip_scalar_d scale; vsip_fft_dir dir; vsip_major major; } vsip_fftm_attr_d;
typedef struct { vsip_scalar_mi input; vsip_scalar_mi output; vsip_fft_place place;
                 vsip_scalar_f scale; vsip_fft_dir dir; vsip_major major; } vsip_fftm_attr_f;
/* vsip.h:231-234, 275-276 */
typedef struct { vsip_offset offset; vsip_stride stride; vsip_length length; vsip_block_d* block; } vsip_vattr_d;
typedef struct { vsip_offset offset; vsip_stride stride; vsip_length length; vsip_block_f* block; } vsip_vattr_f;
typedef struct { vsip_offset offset; vsip_stride stride; vsip_length length; vsip_cblock_d* block; } vsip_cvattr_d;
typedef struct { vsip_offset offset; vsip_stride stride; vsip_length length; vsip_cblock_f* block; } vsip_cvattr_f;
typedef struct { vsip_scalar_vi input; vsip_scalar_vi output; vsip_fft_place place;
                 vsip_scalar_d scale; vsip_fft_dir dir; } vsip_fft_attr_d;
typedef struct { vsip_scalar_vi input; vsip_scalar_vi output; vsip_fft_place place;
                 vsip_scalar_f scale; vsip_fft_dir dir; } vsip_fft_attr_f;
/* vsip.h:282 */
typedef struct { vsip_scalar_vi ref_len; vsip_scalar_vi data_len; vsip_support_region support;
                 vsip_scalar_vi lag_len; } vsip_corr1d_attr;
/* vsip.h:283 */
typedef struct { vsip_scalar_vi kernel_len; vsip_symmetry symm; vsip_scalar_vi data_len;
                 vsip_support_region support; vsip_scalar_vi out_len; int decimation; } vsip_conv1d_attr;

/* ---- library lifecycle ---------------------------------------------- */
int vsip_init(void *ptr);                                         /* vsip.h:432 */
int vsip_finalize(void *ptr);                                     /* vsip.h:427 */

/* ---- real blocks ----------------------------------------------------- */
vsip_block_f* vsip_blockbind_f(vsip_scalar_f* const, vsip_length, vsip_memory_hint);  /* vsip.h:300 */
vsip_block_d* vsip_blockbind_d(vsip_scalar_d* const, vsip_length, vsip_memory_hint);  /* vsip.h:292 */
vsip_block_f* vsip_blockcreate_f(vsip_length, vsip_memory_hint);                     /* vsip.h:301 */
vsip_block_d* vsip_blockcreate_d(vsip_length, vsip_memory_hint);                     /* vsip.h:293 */
int vsip_blockadmit_f(vsip_block_f*, vsip_scalar_bl);                                 /* vsip.h:364 */
int vsip_blockadmit_d(vsip_block_d*, vsip_scalar_bl);                                 /* vsip.h:363 */
vsip_scalar_f* vsip_blockrelease_f(vsip_block_f*, vsip_scalar_bl);                    /* vsip.h:1835 */
vsip_scalar_d* vsip_blockrelease_d(vsip_block_d*, vsip_scalar_bl);                    /* vsip.h:1832 */
void vsip_blockdestroy_f(vsip_block_f*);                                              /* vsip.h:697 */
void vsip_blockdestroy_d(vsip_block_d*);                                              /* vsip.h:696 */
vsip_scalar_f* vsip_blockfind_f(const vsip_block_f*);                                 /* vsip.h:1833 */
vsip_scalar_d* vsip_blockfind_d(const vsip_block_d*);                                 /* vsip.h:1830 */
vsip_scalar_f* vsip_blockrebind_f(vsip_block_f*, vsip_scalar_f* const);               /* vsip.h:1834 */
vsip_scalar_d* vsip_blockrebind_d(vsip_block_d*, vsip_scalar_d* const);               /* vsip.h:1831 */

/* ---- complex blocks -------------------------------------------------- */
vsip_cblock_f* vsip_cblockbind_f(vsip_scalar_f* const, vsip_scalar_f* const,
                                 vsip_length, vsip_memory_hint);                      /* vsip.h:356 */
vsip_cblock_d* vsip_cblockbind_d(vsip_scalar_d* const, vsip_scalar_d* const,
                                 vsip_length, vsip_memory_hint);                      /* vsip.h:348 */
vsip_cblock_f* vsip_cblockcreate_f(vsip_length, vsip_memory_hint);                   /* vsip.h:357 */
vsip_cblock_d* vsip_cblockcreate_d(vsip_length, vsip_memory_hint);                   /* vsip.h:349 */
int vsip_cblockadmit_f(vsip_cblock_f*, vsip_scalar_bl);                               /* vsip.h:372 */
int vsip_cblockadmit_d(vsip_cblock_d*, vsip_scalar_bl);                               /* vsip.h:371 */
void vsip_cblockrelease_f(vsip_cblock_f*, vsip_scalar_bl,
                          vsip_scalar_f**, vsip_scalar_f**);                          /* vsip.h:711 */
void vsip_cblockrelease_d(vsip_cblock_d*, vsip_scalar_bl,
                          vsip_scalar_d**, vsip_scalar_d**);                          /* vsip.h:710 */
void vsip_cblockdestroy_f(vsip_cblock_f*);                                            /* vsip.h:705 */
void vsip_cblockdestroy_d(vsip_cblock_d*);                                            /* vsip.h:704 */
void vsip_cblockfind_f(const vsip_cblock_f*, vsip_scalar_f**, vsip_scalar_f**);       /* vsip.h:707 */
void vsip_cblockfind_d(const vsip_cblock_d*, vsip_scalar_d**, vsip_scalar_d**);       /* vsip.h:706 */
void vsip_cblockrebind_f(vsip_cblock_f*, vsip_scalar_f* const, vsip_scalar_f* const,
                         vsip_scalar_f**, vsip_scalar_f**);                           /* vsip.h:709 */
void vsip_cblockrebind_d(vsip_cblock_d*, vsip_scalar_d* const, vsip_scalar_d* const,
                         vsip_scalar_d**, vsip_scalar_d**);                           /* vsip.h:708 */

/* ---- complex matrix views -------------------------------------------- */
vsip_cmview_f* vsip_cmbind_f(const vsip_cblock_f*, vsip_offset, vsip_stride, vsip_length,
                             vsip_stride, vsip_length);                               /* vsip.h:455 */
vsip_cmview_d* vsip_cmbind_d(const vsip_cblock_d*, vsip_offset, vsip_stride, vsip_length,
                             vsip_stride, vsip_length);                               /* vsip.h:454 */
vsip_cmview_f* vsip_cmcreate_f(vsip_length, vsip_length, vsip_major, vsip_memory_hint); /* vsip.h:495 */
vsip_cmview_d* vsip_cmcreate_d(vsip_length, vsip_length, vsip_major, vsip_memory_hint); /* vsip.h:494 */
vsip_cmview_f* vsip_cmsubview_f(const vsip_cmview_f*, vsip_index, vsip_index,
                                vsip_length, vsip_length);                            /* vsip.h:1959 */
vsip_cmview_d* vsip_cmsubview_d(const vsip_cmview_d*, vsip_index, vsip_index,
                                vsip_length, vsip_length);                            /* vsip.h:1958 */
vsip_cmview_f* vsip_cmtransview_f(const vsip_cmview_f*);                              /* vsip.h:1902 */
vsip_cmview_d* vsip_cmtransview_d(const vsip_cmview_d*);                              /* vsip.h:1901 */
vsip_cblock_f* vsip_cmdestroy_f(vsip_cmview_f*);                                      /* vsip.h:358 */
vsip_cblock_d* vsip_cmdestroy_d(vsip_cmview_d*);                                      /* vsip.h:350 */
void vsip_cmalldestroy_f(vsip_cmview_f*);                                             /* vsip.h:745 */
void vsip_cmalldestroy_d(vsip_cmview_d*);                                             /* vsip.h:744 */
vsip_cscalar_f vsip_cmget_f(const vsip_cmview_f*, vsip_index, vsip_index);            /* vsip.h:552 */
vsip_cscalar_d vsip_cmget_d(const vsip_cmview_d*, vsip_index, vsip_index);            /* vsip.h:523 */
void vsip_cmput_f(const vsip_cmview_f*, vsip_index, vsip_index, vsip_cscalar_f);      /* vsip.h:795 */
void vsip_cmput_d(const vsip_cmview_d*, vsip_index, vsip_index, vsip_cscalar_d);      /* vsip.h:794 */
void vsip_cmgetattrib_f(const vsip_cmview_f*, vsip_cmattr_f*);              /* vsip.h:767 */
void vsip_cmgetattrib_d(const vsip_cmview_d*, vsip_cmattr_d*);              /* vsip.h:766 */
vsip_cblock_f* vsip_cmgetblock_f(const vsip_cmview_f*);                               /* vsip.h:359 */
vsip_cblock_d* vsip_cmgetblock_d(const vsip_cmview_d*);                               /* vsip.h:351 */
void vsip_cmcopyfrom_user_f(vsip_scalar_f* const, vsip_scalar_f* const, vsip_major,
                            const vsip_cmview_f*);                                    /* vsip.h:753 */
void vsip_cmcopyfrom_user_d(vsip_scalar_d* const, vsip_scalar_d* const, vsip_major,
                            const vsip_cmview_d*);                                    /* vsip.h:752 */
void vsip_cmcopyto_user_f(const vsip_cmview_f*, vsip_major, vsip_scalar_f* const,
                          vsip_scalar_f* const);                                      /* vsip.h:755 */
void vsip_cmcopyto_user_d(const vsip_cmview_d*, vsip_major, vsip_scalar_d* const,
                          vsip_scalar_d* const);                                      /* vsip.h:754 */

/* ---- real matrix views ----------------------------------------------- */
vsip_mview_f* vsip_mbind_f(const vsip_block_f*, vsip_offset, vsip_stride, vsip_length,
                           vsip_stride, vsip_length);                                 /* vsip.h:457 */
vsip_mview_d* vsip_mbind_d(const vsip_block_d*, vsip_offset, vsip_stride, vsip_length,
                           vsip_stride, vsip_length);                                 /* vsip.h:456 */
vsip_mview_f* vsip_mcreate_f(vsip_length, vsip_length, vsip_major, vsip_memory_hint);  /* vsip.h:493 */
vsip_mview_d* vsip_mcreate_d(vsip_length, vsip_length, vsip_major, vsip_memory_hint);  /* vsip.h:492 */
vsip_mview_f* vsip_msubview_f(const vsip_mview_f*, vsip_index, vsip_index,
                              vsip_length, vsip_length);                              /* vsip.h:1868 */
vsip_mview_d* vsip_msubview_d(const vsip_mview_d*, vsip_index, vsip_index,
                              vsip_length, vsip_length);                              /* vsip.h:1867 */
vsip_block_f* vsip_mdestroy_f(vsip_mview_f*);                                         /* vsip.h:302 */
vsip_block_d* vsip_mdestroy_d(vsip_mview_d*);                                         /* vsip.h:294 */
void vsip_malldestroy_f(vsip_mview_f*);                                               /* vsip.h:967 */
void vsip_malldestroy_d(vsip_mview_d*);                                               /* vsip.h:966 */
vsip_scalar_f vsip_mget_f(const vsip_mview_f*, vsip_index, vsip_index);               /* vsip.h:623 */
vsip_scalar_d vsip_mget_d(const vsip_mview_d*, vsip_index, vsip_index);               /* vsip.h:590 */
void vsip_mput_f(const vsip_mview_f*, vsip_index, vsip_index, vsip_scalar_f);         /* vsip.h:1138 */
void vsip_mput_d(const vsip_mview_d*, vsip_index, vsip_index, vsip_scalar_d);         /* vsip.h:1137 */
void vsip_mgetattrib_f(const vsip_mview_f*, vsip_mattr_f*);                  /* vsip.h:1072 */
void vsip_mgetattrib_d(const vsip_mview_d*, vsip_mattr_d*);                  /* vsip.h:1071 */
void vsip_mcopyfrom_user_f(vsip_scalar_f* const, vsip_major, const vsip_mview_f*);     /* vsip.h:1040 */
void vsip_mcopyfrom_user_d(vsip_scalar_d* const, vsip_major, const vsip_mview_d*);     /* vsip.h:1039 */
void vsip_mcopyto_user_f(const vsip_mview_f*, vsip_major, vsip_scalar_f* const);       /* vsip.h:1042 */
void vsip_mcopyto_user_d(const vsip_mview_d*, vsip_major, vsip_scalar_d* const);       /* vsip.h:1041 */
/* real and imaginary parts of a complex matrix view (views of the complex
 * block's derived real blocks) */
vsip_mview_f* vsip_mrealview_f(const vsip_cmview_f*);                                  /* vsip.h:1914 */
vsip_mview_d* vsip_mrealview_d(const vsip_cmview_d*);                                  /* vsip.h:1913 */
vsip_mview_f* vsip_mimagview_f(const vsip_cmview_f*);                                  /* vsip.h:1924 */
vsip_mview_d* vsip_mimagview_d(const vsip_cmview_d*);                                  /* vsip.h:1923 */

/* ---- multiple FFT ------------------------------------------------------ */
vsip_fftm_f* vsip_ccfftmip_create_f(vsip_length M, vsip_length N, vsip_scalar_f scale,
                                    vsip_fft_dir dir, vsip_major major,
                                    unsigned int ntimes, vsip_alg_hint hint); /* vsip.h:1813 */
vsip_fftm_f* vsip_ccfftmop_create_f(vsip_length M, vsip_length N, vsip_scalar_f scale,
                                    vsip_fft_dir dir, vsip_major major,
                                    unsigned int ntimes, vsip_alg_hint hint); /* vsip.h:1814 */
vsip_fftm_f* vsip_rcfftmop_create_f(vsip_length M, vsip_length N, vsip_scalar_f scale,
                                    vsip_major major, unsigned int ntimes,
                                    vsip_alg_hint hint);                      /* vsip.h:1816 */
vsip_fftm_f* vsip_crfftmop_create_f(vsip_length M, vsip_length N, vsip_scalar_f scale,
                                    vsip_major major, unsigned int ntimes,
                                    vsip_alg_hint hint);                      /* vsip.h:1815 */
vsip_fftm_d* vsip_ccfftmip_create_d(vsip_length M, vsip_length N, vsip_scalar_d scale,
                                    vsip_fft_dir dir, vsip_major major,
                                    unsigned int ntimes, vsip_alg_hint hint); /* vsip.h:1809 */
vsip_fftm_d* vsip_ccfftmop_create_d(vsip_length M, vsip_length N, vsip_scalar_d scale,
                                    vsip_fft_dir dir, vsip_major major,
                                    unsigned int ntimes, vsip_alg_hint hint); /* vsip.h:1810 */
vsip_fftm_d* vsip_rcfftmop_create_d(vsip_length M, vsip_length N, vsip_scalar_d scale,
                                    vsip_major major, unsigned int ntimes,
                                    vsip_alg_hint hint);                      /* vsip.h:1812 */
vsip_fftm_d* vsip_crfftmop_create_d(vsip_length M, vsip_length N, vsip_scalar_d scale,
                                    vsip_major major, unsigned int ntimes,
                                    vsip_alg_hint hint);                      /* vsip.h:1811 */

void vsip_ccfftmip_f(const vsip_fftm_f*, const vsip_cmview_f*);                       /* vsip.h:715 */
void vsip_ccfftmip_d(const vsip_fftm_d*, const vsip_cmview_d*);                       /* vsip.h:714 */
void vsip_ccfftmop_f(const vsip_fftm_f*, const vsip_cmview_f*, const vsip_cmview_f*); /* vsip.h:717 */
void vsip_ccfftmop_d(const vsip_fftm_d*, const vsip_cmview_d*, const vsip_cmview_d*); /* vsip.h:716 */
void vsip_rcfftmop_f(const vsip_fftm_f*, const vsip_mview_f*, const vsip_cmview_f*);  /* vsip.h:1184 */
void vsip_rcfftmop_d(const vsip_fftm_d*, const vsip_mview_d*, const vsip_cmview_d*);  /* vsip.h:1183 */
void vsip_crfftmop_f(const vsip_fftm_f*, const vsip_cmview_f*, const vsip_mview_f*);  /* vsip.h:827 */
void vsip_crfftmop_d(const vsip_fftm_d*, const vsip_cmview_d*, const vsip_mview_d*);  /* vsip.h:826 */
int vsip_fftm_destroy_f(vsip_fftm_f*);                                                /* vsip.h:426 */
int vsip_fftm_destroy_d(vsip_fftm_d*);                                                /* vsip.h:425 */
void vsip_fftm_getattr_f(const vsip_fftm_f*, vsip_fftm_attr_f*);                      /* vsip.h:947 */
void vsip_fftm_getattr_d(const vsip_fftm_d*, vsip_fftm_attr_d*);                      /* vsip.h:946 */

/* ---- vector views (the 1-D FFT's operands) ----------------------------- */
vsip_vview_f* vsip_vbind_f(const vsip_block_f*, vsip_offset, vsip_stride, vsip_length); /* vsip.h:491 */
vsip_vview_d* vsip_vbind_d(const vsip_block_d*, vsip_offset, vsip_stride, vsip_length); /* vsip.h:503 */
vsip_vview_f* vsip_vcreate_f(vsip_length, vsip_memory_hint);                          /* vsip.h:488 */
vsip_vview_d* vsip_vcreate_d(vsip_length, vsip_memory_hint);                          /* vsip.h:501 */
vsip_block_f* vsip_vdestroy_f(vsip_vview_f*);                                         /* vsip.h:304 */
vsip_block_d* vsip_vdestroy_d(vsip_vview_d*);                                         /* vsip.h:296 */
void vsip_valldestroy_f(vsip_vview_f*);                                               /* vsip.h:1292 */
void vsip_valldestroy_d(vsip_vview_d*);                                               /* vsip.h:1291 */
vsip_scalar_f vsip_vget_f(const vsip_vview_f*, vsip_index);                           /* vsip.h:638 */
vsip_scalar_d vsip_vget_d(const vsip_vview_d*, vsip_index);                           /* vsip.h:605 */
void vsip_vput_f(const vsip_vview_f*, vsip_index, vsip_scalar_f);                     /* vsip.h:1556 */
void vsip_vput_d(const vsip_vview_d*, vsip_index, vsip_scalar_d);                     /* vsip.h:1555 */
void vsip_vgetattrib_f(const vsip_vview_f*, vsip_vattr_f*);                           /* vsip.h:1422 */
void vsip_vgetattrib_d(const vsip_vview_d*, vsip_vattr_d*);                           /* vsip.h:1421 */
void vsip_vcopyfrom_user_f(vsip_scalar_f* const, const vsip_vview_f*);                /* vsip.h:1378 */
void vsip_vcopyfrom_user_d(vsip_scalar_d* const, const vsip_vview_d*);                /* vsip.h:1377 */
void vsip_vcopyto_user_f(const vsip_vview_f*, vsip_scalar_f* const);                  /* vsip.h:1384 */
void vsip_vcopyto_user_d(const vsip_vview_d*, vsip_scalar_d* const);                  /* vsip.h:1383 */

vsip_cvview_f* vsip_cvbind_f(const vsip_cblock_f*, vsip_offset, vsip_stride, vsip_length); /* vsip.h:514 */
vsip_cvview_d* vsip_cvbind_d(const vsip_cblock_d*, vsip_offset, vsip_stride, vsip_length); /* vsip.h:517 */
vsip_cvview_f* vsip_cvcreate_f(vsip_length, vsip_memory_hint);                        /* vsip.h:512 */
vsip_cvview_d* vsip_cvcreate_d(vsip_length, vsip_memory_hint);                        /* vsip.h:515 */
vsip_cblock_f* vsip_cvdestroy_f(vsip_cvview_f*);                                      /* vsip.h:360 */
vsip_cblock_d* vsip_cvdestroy_d(vsip_cvview_d*);                                      /* vsip.h:352 */
void vsip_cvalldestroy_f(vsip_cvview_f*);                                             /* vsip.h:865 */
void vsip_cvalldestroy_d(vsip_cvview_d*);                                             /* vsip.h:864 */
vsip_cscalar_f vsip_cvget_f(const vsip_cvview_f*, vsip_index);                        /* vsip.h:568 */
vsip_cscalar_d vsip_cvget_d(const vsip_cvview_d*, vsip_index);                        /* vsip.h:539 */
void vsip_cvput_f(const vsip_cvview_f*, vsip_index, vsip_cscalar_f);                  /* vsip.h:915 */
void vsip_cvput_d(const vsip_cvview_d*, vsip_index, vsip_cscalar_d);                  /* vsip.h:914 */
void vsip_cvgetattrib_f(const vsip_cvview_f*, vsip_cvattr_f*);                        /* vsip.h:889 */
void vsip_cvgetattrib_d(const vsip_cvview_d*, vsip_cvattr_d*);                        /* vsip.h:888 */
void vsip_cvcopyfrom_user_f(vsip_scalar_f* const, vsip_scalar_f* const,
                            const vsip_cvview_f*);                                    /* vsip.h:875 */
void vsip_cvcopyfrom_user_d(vsip_scalar_d* const, vsip_scalar_d* const,
                            const vsip_cvview_d*);                                    /* vsip.h:874 */
void vsip_cvcopyto_user_f(const vsip_cvview_f*, vsip_scalar_f* const,
                          vsip_scalar_f* const);                                      /* vsip.h:877 */
void vsip_cvcopyto_user_d(const vsip_cvview_d*, vsip_scalar_d* const,
                          vsip_scalar_d* const);                                      /* vsip.h:876 */

/* ---- 1-D FFT ------------------------------------------------------------ */
vsip_fft_f* vsip_ccfftip_create_f(vsip_length N, vsip_scalar_f scale, vsip_fft_dir dir,
                                  unsigned int ntimes, vsip_alg_hint hint);           /* vsip.h:1805 */
vsip_fft_d* vsip_ccfftip_create_d(vsip_length N, vsip_scalar_d scale, vsip_fft_dir dir,
                                  unsigned int ntimes, vsip_alg_hint hint);           /* vsip.h:1801 */
vsip_fft_f* vsip_ccfftop_create_f(vsip_length N, vsip_scalar_f scale, vsip_fft_dir dir,
                                  unsigned int ntimes, vsip_alg_hint hint);           /* vsip.h:1806 */
vsip_fft_d* vsip_ccfftop_create_d(vsip_length N, vsip_scalar_d scale, vsip_fft_dir dir,
                                  unsigned int ntimes, vsip_alg_hint hint);           /* vsip.h:1802 */
vsip_fft_f* vsip_rcfftop_create_f(vsip_length N, vsip_scalar_f scale, unsigned int ntimes,
                                  vsip_alg_hint hint);                                /* vsip.h:1808 */
vsip_fft_d* vsip_rcfftop_create_d(vsip_length N, vsip_scalar_d scale, unsigned int ntimes,
                                  vsip_alg_hint hint);                                /* vsip.h:1804 */
vsip_fft_f* vsip_crfftop_create_f(vsip_length N, vsip_scalar_f scale, unsigned int ntimes,
                                  vsip_alg_hint hint);                                /* vsip.h:1807 */
vsip_fft_d* vsip_crfftop_create_d(vsip_length N, vsip_scalar_d scale, unsigned int ntimes,
                                  vsip_alg_hint hint);                                /* vsip.h:1803 */
void vsip_ccfftip_f(const vsip_fft_f*, const vsip_cvview_f*);                         /* vsip.h:713 */
void vsip_ccfftip_d(const vsip_fft_d*, const vsip_cvview_d*);                         /* vsip.h:712 */
void vsip_ccfftop_f(const vsip_fft_f*, const vsip_cvview_f*, const vsip_cvview_f*);   /* vsip.h:719 */
void vsip_ccfftop_d(const vsip_fft_d*, const vsip_cvview_d*, const vsip_cvview_d*);   /* vsip.h:718 */
void vsip_rcfftop_f(const vsip_fft_f*, const vsip_vview_f*, const vsip_cvview_f*);    /* vsip.h:1186 */
void vsip_rcfftop_d(const vsip_fft_d*, const vsip_vview_d*, const vsip_cvview_d*);    /* vsip.h:1185 */
void vsip_crfftop_f(const vsip_fft_f*, const vsip_cvview_f*, const vsip_vview_f*);    /* vsip.h:829 */
void vsip_crfftop_d(const vsip_fft_d*, const vsip_cvview_d*, const vsip_vview_d*);    /* vsip.h:828 */
int vsip_fft_destroy_f(vsip_fft_f*);                                                  /* vsip.h:424 */
int vsip_fft_destroy_d(vsip_fft_d*);                                                  /* vsip.h:423 */
void vsip_fft_getattr_f(const vsip_fft_f*, vsip_fft_attr_f*);                         /* vsip.h:945 */
void vsip_fft_getattr_d(const vsip_fft_d*, vsip_fft_attr_d*);                         /* vsip.h:944 */

/* ---- 1-D correlation (reference length M, data length N, M <= N) ---------
 * y = the correlation of reference h with data x over the support region
 * (FULL: M + N - 1 lags, SAME: N, MIN: N - M + 1), unbiased on request.  The
 * reference computes it with zero-padded FFTs of a power-of-two length
 * above M + N - 1 (vsip_ccorrelate1d_f.c:123-196), and so does this library
 * on the device.  Operands are views of admitted blocks.  create returns
 * NULL for M < 1 or M > N.  Calls on one object share its transform buffer,
 * so one object serves one thread at a time. */
vsip_corr1d_f* vsip_corr1d_create_f(vsip_length M, vsip_length N, vsip_support_region support,
                                    unsigned int ntimes, vsip_alg_hint hint);          /* vsip.h:1798 */
vsip_corr1d_d* vsip_corr1d_create_d(vsip_length M, vsip_length N, vsip_support_region support,
                                    unsigned int ntimes, vsip_alg_hint hint);          /* vsip.h:1797 */
vsip_ccorr1d_f* vsip_ccorr1d_create_f(vsip_length M, vsip_length N, vsip_support_region support,
                                      unsigned int ntimes, vsip_alg_hint hint);        /* vsip.h:1785 */
vsip_ccorr1d_d* vsip_ccorr1d_create_d(vsip_length M, vsip_length N, vsip_support_region support,
                                      unsigned int ntimes, vsip_alg_hint hint);        /* vsip.h:1784 */
int vsip_corr1d_destroy_f(vsip_corr1d_f*);                                             /* vsip.h:408 */
int vsip_corr1d_destroy_d(vsip_corr1d_d*);                                             /* vsip.h:407 */
int vsip_ccorr1d_destroy_f(vsip_ccorr1d_f*);                                           /* vsip.h:380 */
int vsip_ccorr1d_destroy_d(vsip_ccorr1d_d*);                                           /* vsip.h:379 */
void vsip_corr1d_getattr_f(const vsip_corr1d_f*, vsip_corr1d_attr*);                   /* vsip.h:821 */
void vsip_corr1d_getattr_d(const vsip_corr1d_d*, vsip_corr1d_attr*);                   /* vsip.h:820 */
void vsip_ccorr1d_getattr_f(const vsip_ccorr1d_f*, vsip_corr1d_attr*);                 /* vsip.h:723 */
void vsip_ccorr1d_getattr_d(const vsip_ccorr1d_d*, vsip_corr1d_attr*);                 /* vsip.h:722 */
void vsip_correlate1d_f(const vsip_corr1d_f*, vsip_bias, const vsip_vview_f *h, const vsip_vview_f *x,
                        const vsip_vview_f *y);                                         /* vsip.h:823 */
void vsip_correlate1d_d(const vsip_corr1d_d*, vsip_bias, const vsip_vview_d *h, const vsip_vview_d *x,
                        const vsip_vview_d *y);                                         /* vsip.h:822 */
void vsip_ccorrelate1d_f(const vsip_ccorr1d_f*, vsip_bias, const vsip_cvview_f *h, const vsip_cvview_f *x,
                         const vsip_cvview_f *y);                                       /* vsip.h:725 */
void vsip_ccorrelate1d_d(const vsip_ccorr1d_d*, vsip_bias, const vsip_cvview_d *h, const vsip_cvview_d *x,
                         const vsip_cvview_d *y);                                       /* vsip.h:724 */

/* ---- 1-D convolution (real, decimated) -------------------------------------
 * y[k] = z[k D + s] of z = kernel * x, the kernel expanded by its symmetry
 * (length M), s = 0 (FULL), M/2 (SAME), M - 1 (MIN).  The kernel view is read
 * at create.  Every lag is computed (the reference's sectioned method loses
 * the lags past its last full section).  create returns NULL for D < 1 or a
 * MIN support with no output.  Calls on one object share its buffers, so one
 * object serves one thread at a time. */
vsip_conv1d_f* vsip_conv1d_create_f(const vsip_vview_f *h, vsip_symmetry symm, vsip_length N, int D,
                                    vsip_support_region support, unsigned int ntimes,
                                    vsip_alg_hint hint);                                /* vsip.h:1796 */
vsip_conv1d_d* vsip_conv1d_create_d(const vsip_vview_d *h, vsip_symmetry symm, vsip_length N, int D,
                                    vsip_support_region support, unsigned int ntimes,
                                    vsip_alg_hint hint);                                /* vsip.h:1795 */
int vsip_conv1d_destroy_f(vsip_conv1d_f*);                                             /* vsip.h:406 */
int vsip_conv1d_destroy_d(vsip_conv1d_d*);                                             /* vsip.h:405 */
void vsip_conv1d_getattr_f(const vsip_conv1d_f*, vsip_conv1d_attr*);                   /* vsip.h:817 */
void vsip_conv1d_getattr_d(const vsip_conv1d_d*, vsip_conv1d_attr*);                   /* vsip.h:816 */
void vsip_convolve1d_f(const vsip_conv1d_f*, const vsip_vview_f *x, const vsip_vview_f *y); /* vsip.h:819 */
void vsip_convolve1d_d(const vsip_conv1d_d*, const vsip_vview_d *x, const vsip_vview_d *y); /* vsip.h:818 */

/* ---- element operations on transform operands ---------------------------- */
void vsip_rscvmul_f(vsip_scalar_f, const vsip_cvview_f*, const vsip_cvview_f*);        /* vsip.h:1216 */
void vsip_rscvmul_d(vsip_scalar_d, const vsip_cvview_d*, const vsip_cvview_d*);        /* vsip.h:1215 */
void vsip_rscmmul_f(vsip_scalar_f, const vsip_cmview_f*, const vsip_cmview_f*);        /* vsip.h:1208 */
void vsip_rscmmul_d(vsip_scalar_d, const vsip_cmview_d*, const vsip_cmview_d*);        /* vsip.h:1207 */
void vsip_cvsub_f(const vsip_cvview_f*, const vsip_cvview_f*, const vsip_cvview_f*);   /* vsip.h:941 */
void vsip_cvsub_d(const vsip_cvview_d*, const vsip_cvview_d*, const vsip_cvview_d*);   /* vsip.h:940 */
void vsip_cmsub_f(const vsip_cmview_f*, const vsip_cmview_f*, const vsip_cmview_f*);   /* vsip.h:805 */
void vsip_cmsub_d(const vsip_cmview_d*, const vsip_cmview_d*, const vsip_cmview_d*);   /* vsip.h:804 */
vsip_scalar_f vsip_cvmeansqval_f(const vsip_cvview_f*);                                /* vsip.h:619 */
vsip_scalar_d vsip_cvmeansqval_d(const vsip_cvview_d*);                                /* vsip.h:586 */
vsip_scalar_f vsip_cmmeansqval_f(const vsip_cmview_f*);                                /* vsip.h:618 */
vsip_scalar_d vsip_cmmeansqval_d(const vsip_cmview_d*);                                /* vsip.h:585 */

/* ---- element operations of the k-omega beamformer (kw.c:7-25) ------------ */
void vsip_vmmul_f(const vsip_vview_f*, const vsip_mview_f*, vsip_major, const vsip_mview_f*);  /* vsip.h:1520 */
void vsip_vmmul_d(const vsip_vview_d*, const vsip_mview_d*, vsip_major, const vsip_mview_d*);  /* vsip.h:1519 */
void vsip_mcmagsq_f(const vsip_cmview_f*, const vsip_mview_f*);                        /* vsip.h:1013 */
void vsip_mcmagsq_d(const vsip_cmview_d*, const vsip_mview_d*);                        /* vsip.h:1012 */
void vsip_smmul_f(vsip_scalar_f, const vsip_mview_f*, const vsip_mview_f*);            /* vsip.h:1226 */
void vsip_smmul_d(vsip_scalar_d, const vsip_mview_d*, const vsip_mview_d*);            /* vsip.h:1225 */
void vsip_madd_f(const vsip_mview_f*, const vsip_mview_f*, const vsip_mview_f*);       /* vsip.h:961 */
void vsip_madd_d(const vsip_mview_d*, const vsip_mview_d*, const vsip_mview_d*);       /* vsip.h:960 */
void vsip_mfill_f(vsip_scalar_f, const vsip_mview_f*);                                 /* vsip.h:1063 */
void vsip_mfill_d(vsip_scalar_d, const vsip_mview_d*);                                 /* vsip.h:1062 */
vsip_vview_f* vsip_vcreate_hanning_f(vsip_length, vsip_memory_hint);                   /* vsip.h:477 */
vsip_vview_d* vsip_vcreate_hanning_d(vsip_length, vsip_memory_hint);                   /* vsip.h:473 */

/* ---- random numbers (state on the host, vector/matrix fills on the device) */
vsip_randstate* vsip_randcreate(vsip_index, vsip_index, vsip_index, vsip_rng);          /* vsip.h:1826 */
int vsip_randdestroy(vsip_randstate*);                                                 /* vsip.h:451 */
vsip_scalar_f vsip_randu_f(vsip_randstate*);                                           /* vsip.h:633 */
vsip_scalar_d vsip_randu_d(vsip_randstate*);                                           /* vsip.h:600 */
vsip_scalar_f vsip_randn_f(vsip_randstate*);                                           /* vsip.h:632 */
vsip_scalar_d vsip_randn_d(vsip_randstate*);                                           /* vsip.h:599 */
vsip_cscalar_f vsip_crandu_f(vsip_randstate*);                                         /* vsip.h:560 */
vsip_cscalar_d vsip_crandu_d(vsip_randstate*);                                         /* vsip.h:531 */
vsip_cscalar_f vsip_crandn_f(vsip_randstate*);                                         /* vsip.h:559 */
vsip_cscalar_d vsip_crandn_d(vsip_randstate*);                                         /* vsip.h:530 */
void vsip_vrandu_f(vsip_randstate*, const vsip_vview_f*);                              /* vsip.h:1573 */
void vsip_vrandu_d(vsip_randstate*, const vsip_vview_d*);                              /* vsip.h:1572 */
void vsip_vrandn_f(vsip_randstate*, const vsip_vview_f*);                              /* vsip.h:1571 */
void vsip_vrandn_d(vsip_randstate*, const vsip_vview_d*);                              /* vsip.h:1570 */
void vsip_cvrandu_f(vsip_randstate*, const vsip_cvview_f*);                            /* vsip.h:919 */
void vsip_cvrandu_d(vsip_randstate*, const vsip_cvview_d*);                            /* vsip.h:918 */
void vsip_cvrandn_f(vsip_randstate*, const vsip_cvview_f*);                            /* vsip.h:917 */
void vsip_cvrandn_d(vsip_randstate*, const vsip_cvview_d*);                            /* vsip.h:916 */
void vsip_mrandu_f(vsip_randstate*, const vsip_mview_f*);                              /* vsip.h:1577 */
void vsip_mrandu_d(vsip_randstate*, const vsip_mview_d*);                              /* vsip.h:1576 */
void vsip_mrandn_f(vsip_randstate*, const vsip_mview_f*);                              /* vsip.h:1575 */
void vsip_mrandn_d(vsip_randstate*, const vsip_mview_d*);                              /* vsip.h:1574 */
void vsip_cmrandu_f(vsip_randstate*, const vsip_cmview_f*);                            /* vsip.h:923 */
void vsip_cmrandu_d(vsip_randstate*, const vsip_cmview_d*);                            /* vsip.h:922 */
void vsip_cmrandn_f(vsip_randstate*, const vsip_cmview_f*);                            /* vsip.h:921 */
void vsip_cmrandn_d(vsip_randstate*, const vsip_cmview_d*);                            /* vsip.h:920 */

#ifdef __cplusplus
}
#endif

#endif /* _vsip_h */
