/* ipfg_types.h -- the value types of libipfg's C ABI (pixel types, border
 * modes, precision modes, status codes, the image descriptor and the stream
 * handle). Included by ipfg.h; the kernels include only this file. */
#ifndef IPFG_TYPES_H
#define IPFG_TYPES_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define IPFG_ABI_VERSION 1

/* Pixel types: ipf's documented aliases ImageU8/U16/I16/I32/F/D (core.hpp:403-408). */
typedef enum ipfg_dtype {
    IPFG_U8 = 0,
    IPFG_U16 = 1,
    IPFG_I16 = 2,
    IPFG_I32 = 3,
    IPFG_F32 = 4,
    IPFG_F64 = 5
} ipfg_dtype;

/* Same order and meaning as ipf::BorderMode (core.hpp:183). */
typedef enum ipfg_border {
    IPFG_BORDER_CONSTANT = 0,
    IPFG_BORDER_REPLICATE = 1,
    IPFG_BORDER_REFLECT = 2,
    IPFG_BORDER_REFLECT101 = 3,
    IPFG_BORDER_WRAP = 4
} ipfg_border;

/* Arithmetic used by the linear filters and the bilateral filter. */
typedef enum ipfg_precision {
    /* FP64 fused multiply-add accumulation, certified: every output whose
     * FP64 value lies within the accumulated error bound of a rounding
     * boundary of the output type (an integer's half-way point, a float's
     * midpoint) is recomputed in FP64 in the reference's operation order by
     * the same launch. Integer and float32 outputs are bit-identical to the
     * reference; float64 outputs take EXACT. Box filters on float inputs use
     * window sums instead of the reference's running sums and carry FP64
     * rounding error. */
    IPFG_PREC_DEFAULT = 0,
    /* FP64 in the reference's operation order for every sample, no fused
     * multiply-add: bit-identical to the reference for every pixel type. */
    IPFG_PREC_EXACT = 1,
    /* FP32 arithmetic without certification (integer outputs may differ from
     * the reference by one unit at rounding ties). */
    IPFG_PREC_FP32 = 2,
    /* filter2D only: MFMA implicit GEMM with split-precision weights
     * (hi + lo), FP32 accumulation. */
    IPFG_PREC_BF16X3 = 3,
    IPFG_PREC_FP16X2 = 4
} ipfg_precision;

typedef enum ipfg_status {
    IPFG_OK = 0,
    IPFG_ERR_ARG = 1,         /* reference argument check failed (ipf::Error) */
    IPFG_ERR_DEVICE = 2,      /* HIP runtime error */
    IPFG_ERR_NOMEM = 3,       /* device or host allocation failed */
    IPFG_ERR_UNSUPPORTED = 4, /* combination not provided by this build */
    IPFG_ERR_STATE = 5        /* library not initialised / shut down */
} ipfg_status;

typedef struct ipfg_image {
    void* data;
    int width;
    int height;
    int channels;
    int dtype;               /* ipfg_dtype */
    size_t row_stride_bytes; /* 0: packed rows (width*channels*sizeof) */
    int on_device;           /* 0: host memory, 1: device memory */
} ipfg_image;

typedef void* ipfg_stream; /* a hipStream_t */

#ifdef __cplusplus
}
#endif

#endif /* IPFG_TYPES_H */
