// Launch interface of houghCircles (hough.hip) for the C-ABI runtime.
#pragma once

#include "common.hpp"

namespace ipfg {

// An edge pixel: position, unit gradient (refinement) and the gradient with its
// length (voting).
struct HoughPoint {
    double x, y, nx, ny;
};

struct HoughScratch {
    HoughPoint* pts;        // up to W * H points, row-major order
    double* grad;           // 3 doubles per point: gx, gy, |g|
    unsigned* counts;       // one per 1024-pixel chunk, then the scanned offsets (same size)
    unsigned* total;        // [0]: number of points, [1]: number of candidates
    float* acc;             // W x H votes
    float* sm;              // W x H 3 x 3 box sums of the votes
    int* cand_xy;           // up to W * H (x, y) pairs of accumulator peaks, row-major
    float* cand_v;          // their box sums
};

// Chunks of the W x H image the compaction works in.
size_t hough_chunks(int w, int h);
// Edge points of `edges` (uint8, nonzero = edge) with |(dx, dy)| >= 1e-9,
// in row-major order, and their votes into acc; then the 3 x 3 box sums and
// the accumulator peaks above `threshold` (row-major). total[0..1] receive the
// counts (device memory).
hipError_t launch_hough_votes(const DevImage& edges, const DevImage& dx, const DevImage& dy, int min_r, int max_r,
                              float threshold, const HoughScratch& sc, hipStream_t s);

// Result of one peak's refinement: centre, radius, support and whether the
// circle is accepted.
struct HoughFit {
    double cx, cy, r;
    int support, accepted;
};
// Refines each of the n peaks (x, y pairs) against the npts edge points: one
// workgroup per peak. Radii above kHoughMaxRadius are not supported.
constexpr int kHoughMaxRadius = 15000;
hipError_t launch_hough_refine(const HoughPoint* pts, unsigned npts, const float* acc, int w, int h, const int* peaks,
                               int n, int min_r, int max_r, double min_coverage, HoughFit* out, hipStream_t s);

}  // namespace ipfg
