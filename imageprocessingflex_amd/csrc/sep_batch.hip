// Persistent batched separable filter: certified uint8 -> uint8 strips of any
// kernel size for 1-, 3- and 4-channel images, popped from a tile queue.
#include "batch.hpp"
#include "sep_kernels.hpp"

namespace ipfg {
namespace sepk {
namespace {

// SJ: stager samples per lane for the class's widest rows (PTX + kmax - 1)
// x CN, so the small-kernel classes keep their registers for occupancy.
template <int CN, int SJ>
__global__ __launch_bounds__(kBlock) void sep_batch_kernel(const BatchClassHeader* __restrict__ hdr,
                                                           const BatchJob* __restrict__ jobs, BatchQueue* q,
                                                           FixList fix) {
    extern __shared__ __align__(16) unsigned char sep_smem[];
    batch_loop(hdr, jobs, q, [&](const BatchJob& J, int tx, int ty) {
        SepArgs a{};
        a.src = J.src;
        a.dst = J.dst;
        a.kx = J.kx;
        a.ky = J.ky;
        a.kxd = J.kx;
        a.kyd = J.ky;
        a.kw = J.kw;
        a.kh = J.kh;
        a.ax = J.kw / 2;
        a.ay = J.kh / 2;
        a.border = J.border;
        a.delta = 0.0;
        a.bval = 0.0;
        a.bcoef = J.bcoef;
        a.bdelta = J.bdelta;
        a.xbound = 255.0;
        a.SH = J.rows;
        a.oscale = 1.0;
        a.fix = fix;
        sep_plane_strip<uint8_t, uint8_t, double, float, CN, SJ, true, 0>(a, J.kx, J.ky, tx, ty, sep_smem);
    });
    fix_finish(fix);
}

}  // namespace
}  // namespace sepk

template <int CN, int KMAX>
const void* class_kernel() {
    return reinterpret_cast<const void*>(sepk::sep_batch_kernel<CN, sepk::plane_sj(CN, KMAX)>);
}

const void* sep_class_kernel(int cls) {
    switch (cls) {
        case 0: return class_kernel<1, kBatchSepSmallK>();
        case 1: return class_kernel<1, kBatchSepLargeK>();
        case 2: return class_kernel<3, kBatchSepSmallK>();
        case 3: return class_kernel<3, kBatchSepLargeK>();
        case 4: return class_kernel<4, kBatchSepSmallK>();
        default: return class_kernel<4, kBatchSepLargeK>();
    }
}

template <int CN, int KMAX>
static hipError_t launch_ck(const BatchClassHeader* hdr, const BatchJob* jobs, BatchQueue* q, FixList fix, int grid,
                            size_t lds, hipStream_t s) {
    static size_t raised = 0;
    auto kern = sepk::sep_batch_kernel<CN, sepk::plane_sj(CN, KMAX)>;
    if (lds > 64 * 1024 && lds > raised) {
        const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                                 hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
        if (e != hipSuccess) return e;
        raised = lds;
    }
    hipLaunchKernelGGL(kern, dim3(grid), dim3(kBlock), lds, s, hdr, jobs, q, fix);
    return hipGetLastError();
}

hipError_t launch_sep_batch(int cls, const BatchClassHeader* hdr, const BatchJob* jobs, BatchQueue* q, FixList fix,
                            int grid, size_t lds, hipStream_t s) {
    switch (cls) {
        case 0: return launch_ck<1, kBatchSepSmallK>(hdr, jobs, q, fix, grid, lds, s);
        case 1: return launch_ck<1, kBatchSepLargeK>(hdr, jobs, q, fix, grid, lds, s);
        case 2: return launch_ck<3, kBatchSepSmallK>(hdr, jobs, q, fix, grid, lds, s);
        case 3: return launch_ck<3, kBatchSepLargeK>(hdr, jobs, q, fix, grid, lds, s);
        case 4: return launch_ck<4, kBatchSepSmallK>(hdr, jobs, q, fix, grid, lds, s);
        default: return launch_ck<4, kBatchSepLargeK>(hdr, jobs, q, fix, grid, lds, s);
    }
}

int sep_batch_occupancy(int cls, size_t lds) {
    const void* k = sep_class_kernel(cls);
    if (lds > 64 * 1024)  // the attribute must allow the LDS before occupancy is asked
        (void)hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k, kBlock, lds) != hipSuccess) return 1;
    return n > 0 ? n : 1;
}

size_t sep_batch_lds(int cn, int kmax) { return sepk::plane_lds_bytes(cn, kmax, kmax, sizeof(float)); }

}  // namespace ipfg
