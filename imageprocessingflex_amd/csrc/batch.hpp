// Batched execution: many images of different sizes, channel counts and
// filter parameters in one launch per operation class.
//
// A batch is a table of jobs in device memory. Each job is cut into tiles
// (strips for the separable filter, TX x TH tiles for median and bilateral);
// tiles of all jobs of a class are numbered consecutively (job.tile0 is the
// first tile of the job). A persistent kernel of a fixed grid (the number of
// workgroups the device holds at once) pops tile numbers from a queue counter
// and runs the same per-tile code as the single-image kernels; a workgroup
// exits when the counter passes the class's tile count, and the last one to
// finish resets the counter for the next launch. Because the grids and the
// table addresses never change, a batch executor captures these launches
// once in a hipGraph and replays it for every batch.
#pragma once

#include "common.hpp"

namespace ipfg {

// Separable jobs are split by channel count and kernel size so that each
// class's LDS (and so its occupancy) fits its own jobs: classes 0..5 are
// (cn 1, 3, 4) x (kernel <= 17, kernel <= 49).
enum BatchClass { BCLS_SEP = 0, BCLS_MEDIAN = 6, BCLS_BILATERAL = 7, BCLS_COUNT = 8 };
constexpr int kBatchSepSmallK = 17;
constexpr int kBatchSepLargeK = 49;
__host__ __device__ inline int batch_sep_class(int cn, int k) {
    return 2 * (cn == 1 ? 0 : cn == 3 ? 1 : 2) + (k <= kBatchSepSmallK ? 0 : 1);
}
__host__ __device__ inline int batch_sep_cn(int cls) { return cls / 2 == 0 ? 1 : cls / 2 == 1 ? 3 : 4; }
__host__ __device__ inline int batch_sep_kmax(int cls) { return cls % 2 ? kBatchSepLargeK : kBatchSepSmallK; }
constexpr int kBatchTileRows = 32;  // TH of median and bilateral tiles

struct BatchJob {
    DevImage src, dst;  // uint8, 1, 3 or 4 channels, device-resident
    int border;
    int tiles_x, tiles_y, tile0;
    int rows;  // strip height (separable) or tile rows (median, bilateral)
    // separable: FP64 weights with 16 zeros of padding, certification bound
    const double* kx;
    const double* ky;
    int kw, kh;
    double bcoef, bdelta;
    // median
    int k;
    // bilateral: taps, FP64 space weights and range table
    const int* taps;
    const double* spaced;
    const double* lutd;
    int ntaps, lut_len, r;
    double cc, bound;
};

// Per class: number of jobs and of tiles, followed in memory by the jobs.
struct BatchClassHeader {
    int njobs, ntiles;
    int pad[2];
};

// Queue state of one class (persistent across launches, zero between them).
struct BatchQueue {
    unsigned int next;
    unsigned int done;
};

// Runs tile(job, tx, ty) for every tile of the class, each workgroup popping
// tile numbers until they run out. All threads of the workgroup call it.
template <class Tile>
__device__ inline void batch_loop(const BatchClassHeader* __restrict__ hdr, const BatchJob* __restrict__ jobs,
                                  BatchQueue* q, Tile tile) {
    __shared__ int s_tile, s_job;
    const int njobs = hdr->njobs, ntiles = hdr->ntiles;
    for (;;) {
        __syncthreads();  // the previous tile is done with the LDS and s_tile
        if (threadIdx.x == 0) {
            const int t = static_cast<int>(atomicAdd(&q->next, 1u));
            int lo = 0, hi = njobs - 1;  // last job with tile0 <= t
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (jobs[mid].tile0 <= t) lo = mid;
                else hi = mid - 1;
            }
            s_tile = t;
            s_job = lo;
        }
        __syncthreads();
        const int t = __builtin_amdgcn_readfirstlane(s_tile);
        if (t >= ntiles) break;
        const int jb = __builtin_amdgcn_readfirstlane(s_job);
        const BatchJob& J = jobs[jb];
        const int local = t - J.tile0;
        tile(J, local % J.tiles_x, local / J.tiles_x);
    }
    // the last workgroup to leave resets the queue for the next launch
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned int nblk = gridDim.x * gridDim.y * gridDim.z;
        if (atomicAdd(&q->done, 1u) == nblk - 1) {
            __hip_atomic_store(&q->next, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&q->done, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

// Launchers of the persistent kernels (one per class). grid = workgroups;
// lds = dynamic LDS bytes (the class maximum).
hipError_t launch_sep_batch(int cls, const BatchClassHeader* hdr, const BatchJob* jobs, BatchQueue* q,
                            FixList fix, int grid, size_t lds, hipStream_t s);
hipError_t launch_median_batch(const BatchClassHeader* hdr, const BatchJob* jobs, BatchQueue* q, int grid,
                               size_t lds, hipStream_t s);
hipError_t launch_bilateral_batch(const BatchClassHeader* hdr, const BatchJob* jobs, BatchQueue* q, FixList fix,
                                  int grid, size_t lds, hipStream_t s);
// Occupancy (workgroups per CU at the given LDS) and LDS of the class kernels.
int sep_batch_occupancy(int cls, size_t lds);
size_t sep_batch_lds(int cn, int kmax);
int median_batch_occupancy(size_t lds);
size_t median_batch_lds();
int bilateral_batch_occupancy(size_t lds);
size_t bilateral_batch_lds();
// Dynamic LDS of each class and the persistent grid the device holds at once.
size_t batch_lds_bytes(int cls);
int batch_grid(int cls, size_t lds);
// Tiles of one job and its rows per tile.
void batch_tiles(int cls, int w, int h, int cn, int k, int* tiles_x, int* tiles_y, int* rows);

}  // namespace ipfg
