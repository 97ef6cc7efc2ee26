/* ipfg.h -- C ABI of libipfg.so, the MI355X (gfx950) implementation of ipf's
 * filtering path.
 *
 * The reference (ipf, a header-only C++17 library) has no FFI layer: its
 * boundary is the set of function templates in namespace ipf declared in
 * include/ipf/filter.hpp. Each entry point below replaces one of them; the
 * header-only front end include/ipf/filter.hpp of this repository keeps the
 * reference's C++ signatures and calls these functions. Citations are
 * file:line in the reference source tree.
 *
 * Conventions
 *  - Every function returns an ipfg_status; no C++ exception crosses the ABI.
 *    On failure ipfg_last_error() returns the message for the calling thread.
 *    Argument errors carry the reference's exact "ipf: ..." text and are raised
 *    before any device work, in the reference's check order.
 *  - Images are interleaved HWC, row-major: sample (x, y, c) lives at byte
 *    y*row_stride + (x*channels + c)*sizeof(dtype)   (core.hpp:220-222).
 *  - If src and dst are both device-resident (on_device != 0) the call is
 *    asynchronous on `stream` (NULL: the calling thread's own stream).
 *    If either is host memory the call stages through device memory and
 *    returns when dst holds the result (the reference's synchronous contract).
 *  - dst must not alias src.
 */
#ifndef IPFG_H
#define IPFG_H

#include <stddef.h>
#include <stdint.h>

#include "ipfg_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- runtime */
/* Selects the device and creates the library state. Called implicitly by
 * every other entry point with device 0 if not called before. */
int ipfg_init(int device);
int ipfg_shutdown(void);
const char* ipfg_last_error(void);
int ipfg_abi_version(void);
/* Number of samples recomputed in FP64 by the last certified call of this
 * thread (device-resident calls: after the stream has been synchronised). */
int ipfg_last_fixups(long long* count);

/* Launch-shape autotuner (no counterpart in the reference, whose CPU loops
 * have no launch shape). enable != 0: the first call of each shape class
 * with at least 2^20 samples times the candidate launch shapes on its stream
 * (synchronising it once) and keeps the fastest; 0: built-in choices, except
 * for classes already in the table. Default on unless IPFG_AUTOTUNE=0. */
int ipfg_autotune(int enable);
/* Writes the table as text (one "<op> <tin> <tout> <cn> <kw> <kh> <mode>
 * <log2 w> <log2 h> <value>" line per class) into buf (capacity bytes, NUL
 * terminated when it fits); *needed receives the full length + 1. */
int ipfg_tune_export(char* buf, size_t capacity, size_t* needed);
/* Adds the lines of text to the table; *entries receives their number. */
int ipfg_tune_import(const char* text, int* entries);
/* Empties the table. */
int ipfg_tune_clear(void);

/* Thread count for host-side staging; same contract as ipf::setNumThreads /
 * getNumThreads (parallel.hpp:38-51). GPU work does not depend on it. */
void ipfg_set_num_threads(int n);
int ipfg_get_num_threads(void);

/* Device memory and streams for resident images. */
int ipfg_malloc(void** ptr, size_t bytes);
int ipfg_free(void* ptr);
int ipfg_memcpy_h2d(void* dst, const void* src, size_t bytes, ipfg_stream stream);
int ipfg_memcpy_d2h(void* dst, const void* src, size_t bytes, ipfg_stream stream);
int ipfg_memset(void* dst, int value, size_t bytes, ipfg_stream stream);
int ipfg_stream_create(ipfg_stream* stream);
int ipfg_stream_destroy(ipfg_stream stream);
int ipfg_stream_synchronize(ipfg_stream stream);
/* No reference counterpart. Waits for `stream` and detaches the library's
 * cached state from it (events of parameter uploads, scratch blocks and
 * batch buffers last recorded on it), so that the caller may then destroy a
 * stream it created itself; the HIP runtime consults an event's recording
 * stream when the event is later synchronised. ipfg_stream_destroy does this
 * for streams from ipfg_stream_create. */
int ipfg_stream_forget(ipfg_stream stream);
int ipfg_device_synchronize(void);

/* ---------------------------------------------------------------- filters */

/* gaussianKernel1D (filter.hpp:129-145): writes min(n, capacity) weights to
 * out and returns n (> 0), or -IPFG_ERR_ARG. Host-only. */
int ipfg_gaussian_kernel_1d(double sigma, int ksize, double* out, int capacity);

/* sepFilter2D / detail::sepFilterImpl (filter.hpp:34-90, 196-203): horizontal
 * correlation with kx (anchor ax) then vertical with ky (anchor ay), plus
 * delta; src/dst types may differ (TOut). */
int ipfg_sep_filter(const ipfg_image* src, const ipfg_image* dst, const double* kx, int kw, const double* ky,
                    int kh, int ax, int ay, int border, double delta, double border_value, int precision,
                    ipfg_stream stream);

/* No counterpart in the reference. ipfg_sep_filter for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the same arguments and results, but it calls no
 * hipMalloc, hipFree or other allocating routine. Weights, the certification
 * counters and the fix-up report live in a fixed 4 KB device array per GPU
 * (held until the process ends). Limits: device-resident src and dst that do
 * not overlap, at most 128 taps per axis, precision DEFAULT or FP32 and types
 * the fast kernels take (IPFG_ERR_UNSUPPORTED otherwise); the strip height is
 * the built-in rule (no autotuning); a NULL stream is the device's null
 * stream; ipfg_last_fixups does not report these calls. All calls on one GPU
 * share the array, so calls that overlap in time on one GPU race. */
int ipfg_sep_filter_noalloc(const ipfg_image* src, const ipfg_image* dst, const double* kx, int kw, const double* ky,
                            int kh, int ax, int ay, int border, double delta, double border_value, int precision,
                            ipfg_stream stream);

/* gaussianBlur (filter.hpp:261-267): sigmaY <= 0 uses sigmaX; ksize <= 0
 * derives the window from sigma. */
int ipfg_gaussian_blur(const ipfg_image* src, const ipfg_image* dst, double sigma_x, double sigma_y, int ksize,
                       int border, int precision, ipfg_stream stream);

/* boxFilter (filter.hpp:205-253); blur(src, k, border) is
 * ipfg_box_filter(src, dst, k, k, 1, border, 0.0, ...) (filter.hpp:255-259). */
int ipfg_box_filter(const ipfg_image* src, const ipfg_image* dst, int kw, int kh, int normalize, int border,
                    double border_value, int precision, ipfg_stream stream);

/* filter2D (filter.hpp:157-194): kernel is kw*kh row-major float weights,
 * anchor (-1,-1) = centre; zero weights do not take part. */
int ipfg_filter2d(const ipfg_image* src, const ipfg_image* dst, const float* kernel, int kw, int kh, int ax, int ay,
                  int border, double delta, double border_value, int precision, ipfg_stream stream);

/* medianBlur (filter.hpp:300-332): exact median of each ksize x ksize window
 * (ksize odd); Constant border is rejected. Always bit-exact. */
int ipfg_median(const ipfg_image* src, const ipfg_image* dst, int ksize, int border, ipfg_stream stream);

/* bilateralFilter (filter.hpp:334-387). */
int ipfg_bilateral(const ipfg_image* src, const ipfg_image* dst, int d, double sigma_color, double sigma_space,
                   int border, int precision, ipfg_stream stream);

/* ---------------------------------------------------------------- derivatives and sharpening
 * filter.hpp:388-567: compositions of the filters above with elementwise
 * steps (run on the device). Derivative, Laplacian, magnitude and phase
 * images are float32 (dst->dtype == IPFG_F32); unsharpMask, sharpen, emboss
 * and adaptiveThreshold keep the source type. Argument errors carry the
 * reference's messages. */
/* sobel(src, dx, dy, ksize, scale, border), filter.hpp:392-406 */
int ipfg_sobel(const ipfg_image* src, const ipfg_image* dst, int dx, int dy, int ksize, double scale, int border,
               int precision, ipfg_stream stream);
/* scharr(src, dx, dy, scale, border), filter.hpp:408-413 */
int ipfg_scharr(const ipfg_image* src, const ipfg_image* dst, int dx, int dy, double scale, int border, int precision,
                ipfg_stream stream);
/* prewitt(src, dx, dy, border), filter.hpp:415-421 */
int ipfg_prewitt(const ipfg_image* src, const ipfg_image* dst, int dx, int dy, int border, int precision,
                 ipfg_stream stream);
/* laplacian(src, ksize, scale, border), filter.hpp:423-436 */
int ipfg_laplacian(const ipfg_image* src, const ipfg_image* dst, int ksize, double scale, int border, int precision,
                   ipfg_stream stream);
/* magnitude(gx, gy), filter.hpp:438-444; phase(gx, gy, degrees), 446-458.
 * magnitude is bit-identical to the reference (FP64 sqrt(x^2 + y^2) rounded
 * to float); phase takes atan2 in FP64 rounded to float, within one ulp of
 * the host's atan2f before the degree scaling. */
int ipfg_magnitude(const ipfg_image* gx, const ipfg_image* gy, const ipfg_image* dst, ipfg_stream stream);
int ipfg_phase(const ipfg_image* gx, const ipfg_image* gy, const ipfg_image* dst, int degrees, ipfg_stream stream);
/* unsharpMask(src, sigma, amount, threshold), filter.hpp:472-483 */
int ipfg_unsharp_mask(const ipfg_image* src, const ipfg_image* dst, double sigma, double amount, double threshold,
                      int precision, ipfg_stream stream);
/* sharpen(src, strength), filter.hpp:485-490; emboss(src), 492-496 */
int ipfg_sharpen(const ipfg_image* src, const ipfg_image* dst, double strength, int precision, ipfg_stream stream);
int ipfg_emboss(const ipfg_image* src, const ipfg_image* dst, int precision, ipfg_stream stream);
/* adaptiveThreshold(src, maxval, method (0 Mean, 1 Gaussian), type (0
 * Binary, 1 BinaryInv), blockSize, C), filter.hpp:545-565 */
int ipfg_adaptive_threshold(const ipfg_image* src, const ipfg_image* dst, double maxval, int method, int type,
                            int block_size, double c, int precision, ipfg_stream stream);

/* ---------------------------------------------------------------- pyramids
 * pyrDown(src), geometry.hpp:501-512: the 5-tap binomial filter (Reflect101)
 * sampled at even (x, y); dst is (w + 1) / 2 x (h + 1) / 2.
 * pyrUp(src), geometry.hpp:514-524: the zero-inserted 2w x 2h image filtered
 * with the binomial kernel x 4; dst is 2w x 2h.
 * Same pixel type and channels as src, any pixel type, bit-identical to the
 * reference. buildPyramid (geometry.hpp:526-535) is a loop of pyrDown calls in
 * the front ends. */
int ipfg_pyr_down(const ipfg_image* src, const ipfg_image* dst, ipfg_stream stream);
int ipfg_pyr_up(const ipfg_image* src, const ipfg_image* dst, ipfg_stream stream);

/* ---------------------------------------------------------------- ssim
 * ssim(a, b, ssimMap), metrics.hpp:46-78: 11 x 11 Gaussian window (sigma 1.5,
 * Reflect101), K1 = 0.01, K2 = 0.03, dynamic range PixelTraits<T>::max()
 * (core.hpp:57-64). *result receives the mean of the map over the valid
 * region (windows fully inside the image; the whole image when a side is at
 * most 10). ssim_map may be NULL or a float32 image of a's shape; it receives
 * the full-size map, bit-identical to the reference's. The mean differs from
 * the reference's only by summation order (a fixed-order tree sum on the
 * device). Blocks until *result is written. */
int ipfg_ssim(const ipfg_image* a, const ipfg_image* b, const ipfg_image* ssim_map, double* result,
              ipfg_stream stream);

/* houghCircles(gray, minRadius, maxRadius, cannyHigh, centerVotes, minDist,
 * minCoverage), features.hpp:388-504: one-channel src of any pixel type.
 * Writes up to `capacity` circles as (cx, cy, radius, votes) quadruples of
 * doubles into `circles`, strongest first, and their number into *count
 * (call again with a larger capacity when *count exceeds it). canny, the
 * smoothed gradient, the votes and every peak's refinement run on the
 * device; the sums of the least-squares fits follow the reference's point
 * order, so the circles are the reference's bit for bit unless a distance or
 * an alignment falls exactly on one of the reference's thresholds (distances
 * use a device hypot whose last bit can differ from glibc's). Blocks until
 * the result is on the host. maxRadius above 15000: IPFG_ERR_UNSUPPORTED. */
int ipfg_hough_circles(const ipfg_image* src, int min_radius, int max_radius, double canny_high, int center_votes,
                       double min_dist, double min_coverage, double* circles, int capacity, int* count,
                       ipfg_stream stream);

/* ---------------------------------------------------------------- ingest
 * Netpbm (P1-P6) and PFM decoding, codec/pnm.hpp:58-107 (decodePNM) and
 * 143-176 (decodePFM), for request pipelines that decode straight into device
 * memory. ipfg_pnm_info parses the header of the file in `bytes` (host
 * memory) and returns its size, channel count (1 or 3) and whether it is a
 * PFM. ipfg_decode_pnm writes the image into dst (device or host memory,
 * uint8 or uint16 for P1-P6 with the reference's maxval rescaling, float32
 * for PFM). Headers and ASCII rasters (P1-P3) are parsed on the host; binary
 * rasters (P4-P6, PFM) are copied to the device as stored and unpacked
 * there. Results are bit-identical to the reference's decoders; errors carry
 * their messages. Blocks until `bytes` may be reused (and, for a host dst,
 * until dst holds the image). */
int ipfg_pnm_info(const uint8_t* bytes, size_t n, int* width, int* height, int* channels, int* is_pfm);
int ipfg_decode_pnm(const uint8_t* bytes, size_t n, const ipfg_image* dst, ipfg_stream stream);

/* ---------------------------------------------------------------- integral images
 * integralImage(src, squared), filter.hpp:522-538: dst is a float64
 * (W + 1) x (H + 1) image with src's channels, I(x, y) = sum of src (or of
 * its squares) over [0, x) x [0, y); row 0 and column 0 are zero. The running
 * sums follow the reference's order (each row left to right, then down each
 * column), so every entry is bit-identical to the reference's, float inputs
 * included. integralSum (filter.hpp:540-543) is host arithmetic on the table
 * (the front ends). */
int ipfg_integral_image(const ipfg_image* src, const ipfg_image* dst, int squared, ipfg_stream stream);

/* ---------------------------------------------------------------- filter2DFFT
 * filter2DFFT(src, kernel, border, delta, borderValue), fft.hpp:196-216:
 * filter2D with the anchor at the kernel's centre, computed as the reference
 * does through FP64 radix-2 FFTs of the padded image and the kernel
 * (2^ceil(log2(W + kw - 1)) x 2^ceil(log2(H + kh - 1)) points per channel);
 * every butterfly and twiddle as in the reference, so results are
 * bit-identical to it for finite inputs. kernel: kh rows of kw float32
 * weights (host memory). dst has src's shape, any pixel type. Padded sizes
 * above 2^28 points return IPFG_ERR_UNSUPPORTED. */
int ipfg_filter2d_fft(const ipfg_image* src, const ipfg_image* dst, const float* kernel, int kw, int kh, int border,
                      double delta, double border_value, ipfg_stream stream);

/* ---------------------------------------------------------------- features
 * canny(src, low, high, ksize, l2Gradient), features.hpp:26-90: dst is a
 * one-channel uint8 image of src's size (255 on edges, 0 elsewhere); any pixel
 * type and channel count. Sobel gradients (Replicate borders) as the
 * reference computes them, the strongest channel per pixel, non-maximum
 * suppression and hysteresis; bit-identical to the reference.
 * cannyAuto(src, blurSigma), features.hpp:92-101: toGray (1-4 channels),
 * gaussianBlur<float>, the high threshold at the 70th percentile of the L2
 * Sobel magnitude (selected on the device; no host round trip), low = 0.4
 * high. Same output and parity as ipfg_canny. */
int ipfg_canny(const ipfg_image* src, const ipfg_image* dst, double low_threshold, double high_threshold, int ksize,
               int l2_gradient, ipfg_stream stream);
int ipfg_canny_auto(const ipfg_image* src, const ipfg_image* dst, double blur_sigma, ipfg_stream stream);

/* cornerHarris(src, blockSize, ksize, k), features.hpp:128-139, and
 * cornerMinEigenVal(src, blockSize, ksize), features.hpp:141-153: src one
 * channel of any pixel type, dst a one-channel float32 image of its size.
 * The structure tensor's unnormalised box sums follow the reference's running
 * sums in FP64 (boxFilter, filter.hpp:205-253); bit-identical results. */
int ipfg_corner_harris(const ipfg_image* src, const ipfg_image* dst, int block_size, int ksize, double k,
                       ipfg_stream stream);
int ipfg_corner_min_eigen_val(const ipfg_image* src, const ipfg_image* dst, int block_size, int ksize,
                              ipfg_stream stream);

/* ---------------------------------------------------------------- batches
 * No counterpart in the reference (its filters run one image per call).
 * A batch runs many uint8 images (1, 3 or 4 channels, any sizes) through
 * gaussianBlur, medianBlur (3 or 5) and bilateralFilter (radius <= 5) with
 * one persistent tile-queue launch per operation class; results are the
 * single-image calls' results bit for bit (default precision). */
typedef enum ipfg_op { IPFG_OP_GAUSSIAN = 0, IPFG_OP_MEDIAN = 1, IPFG_OP_BILATERAL = 2 } ipfg_op;

typedef struct ipfg_job {
    ipfg_image src;     /* uint8; device memory for batches, host or device for servers */
    ipfg_image dst;     /* same shape as src */
    int op;             /* ipfg_op */
    int border;         /* ipfg_border (median: not Constant) */
    double sigma;       /* gaussian: sigmaX = sigmaY */
    int ksize;          /* gaussian window (0: from sigma) or median window (3 or 5) */
    int d;              /* bilateral diameter (<= 0: from sigma_space) */
    double sigma_color; /* bilateral */
    double sigma_space; /* bilateral */
} ipfg_job;

typedef void* ipfg_batch;
/* Creates an executor for batches of up to max_jobs jobs. use_graph != 0
 * captures the launches of a batch in a hipGraph once and replays it. */
int ipfg_batch_create(ipfg_batch* batch, int max_jobs, int use_graph);
/* Enqueues one batch on stream (device-resident images only); asynchronous. */
int ipfg_batch_run(ipfg_batch batch, const ipfg_job* jobs, int njobs, ipfg_stream stream);
int ipfg_batch_destroy(ipfg_batch batch);

/* ---------------------------------------------------------------- serving
 * No counterpart in the reference. A request loop with continuous batching:
 * the server allocates one device block of device_budget_bytes at creation;
 * a worker thread admits submitted requests in arrival order while their
 * device buffers can be placed in that block, uploads host images, groups the
 * admitted requests into batches of at most max_batch jobs and keeps two
 * batches in flight on two streams (requests that arrive while a batch runs
 * join the next one), downloads host results and returns the buffers' space
 * to the block when their batch completes. The server never holds more
 * request-buffer memory than the budget. */
typedef void* ipfg_server;
int ipfg_server_create(ipfg_server* server, size_t device_budget_bytes, int max_batch, int use_graph);
/* Queues a request (the buffers must stay valid until it completes); *id
 * receives its number (0, 1, ...). Host images are uploaded / downloaded by
 * the server; a dst with on_device = 1 and data = NULL is a server-owned
 * device output that is dropped when the request completes (for measuring
 * the device-side rate). Host buffers and server-owned outputs count against
 * the budget; requests larger than the budget fail. */
int ipfg_server_submit(ipfg_server server, const ipfg_job* request, long long* id);
/* Queues n requests in order (ids first_id, first_id + 1, ...). */
int ipfg_server_submit_many(ipfg_server server, const ipfg_job* requests, int n, long long* first_id);
/* Completed requests since the last poll: ids and latency (submit to result
 * in place, milliseconds); *count receives how many were written. */
int ipfg_server_poll(ipfg_server server, long long* ids, double* latency_ms, int capacity, int* count);
/* Blocks until every submitted request has completed. */
int ipfg_server_drain(ipfg_server server);
/* Batches launched so far and the largest number of jobs in one batch. */
int ipfg_server_stats(ipfg_server server, long long* batches, int* max_jobs_seen);
/* Device memory of the server's request buffers: *held_bytes is what the
 * server holds for them (one block of the budget, rounded down to 256 bytes,
 * allocated at creation), *peak_in_use_bytes the most any set of admitted
 * requests occupied at once (never above *held_bytes). Either may be NULL. */
int ipfg_server_memory(ipfg_server server, size_t* held_bytes, size_t* peak_in_use_bytes);
int ipfg_server_destroy(ipfg_server server);

#ifdef __cplusplus
}
#endif

#endif /* IPFG_H */
