// Separable-filter kernel templates shared by the translation units that
// instantiate them (sep_u8.hip, sep_f32.hip, sep_generic.hip); see sep.hip for
// the algorithm.
#pragma once

#include "kernels.hpp"
#include "stage.hpp"

namespace ipfg {
namespace sepk {

constexpr int P = 4;    // horizontal outputs per thread
constexpr int M = 4;    // vertical outputs per thread
constexpr int CH = 16;  // rows per step: staged input rows = h rows = output rows

template <int CN>
constexpr int tile_x() {
    return CN == 1 ? 64 : 32;
}

struct SepTile {
    int XS;  // staged row length in samples: (TX + kw - 1) * cn
    int XP;  // staged row pitch (floats), multiple of 4
    int HP;  // ring row pitch (elements)
    int R;   // ring rows (power of two >= (K0 + 1) * CH)
    int K0;  // steps before the first output rows: ceil((kh - 1) / CH)
};

__host__ __device__ inline SepTile sep_tile(int TX, int cn, int kw, int kh) {
    SepTile t;
    t.XS = (TX + kw - 1) * cn;
    t.XP = ((t.XS + cn + P + 3) / 4) * 4;
    t.HP = TX * cn;  // rows are written and read along the row: no padding needed
    t.K0 = (kh - 1 + CH - 1) / CH;
    int r = CH;
    while (r < (t.K0 + 1) * CH) r *= 2;
    t.R = r;
    return t;
}

// Bytes of the ring, rounded up so the staged rows behind it start 16-byte aligned.
__host__ __device__ inline size_t sep_ring_bytes(const SepTile& t, size_t acc_size) {
    return ((static_cast<size_t>(t.R) * t.HP * acc_size + 15) / 16) * 16;
}

__host__ inline size_t sep_lds_bytes(int TX, int cn, int kw, int kh, size_t acc_size) {
    const SepTile t = sep_tile(TX, cn, kw, kh);
    return sep_ring_bytes(t, acc_size) + sizeof(float) * static_cast<size_t>(CH) * t.XP;
}

template <class Acc>
__device__ __forceinline__ Acc fma_acc(Acc a, Acc b, Acc c) {
    if constexpr (sizeof(Acc) == 8) return fma(a, b, c);
    else return fmaf(a, b, c);
}

// Loads 4 consecutive weights (the weight arrays carry 16 zeros of padding).
template <class Acc>
__device__ __forceinline__ void load4(const Acc* __restrict__ k, int i, Acc (&out)[4]) {
    out[0] = k[i];
    out[1] = k[i + 1];
    out[2] = k[i + 2];
    out[3] = k[i + 3];
}

// One output sample in the reference's FP64 operation order (filter.hpp:57-86:
// tmp = 0.0 + kx[0] * v[0] + kx[1] * v[1] + ... along the padded row, then
// acc = delta + ky[0] * tmp[0] + ... down the rows), computed by a whole wave:
// lane j forms the tmp value of row j (64 rows per pass) and the vertical sum
// runs over broadcast values. All 64 lanes must call it; all return the result.
template <class Ti>
__device__ inline double sep_exact_wave(const DevImage& src, int x, int y, int c, const double* __restrict__ kx,
                                        int kw, const double* __restrict__ ky, int kh, int ax, int ay, int border,
                                        double delta, double bval) {
    const int cn = src.cn;
    const int xl = x - ax;
    const bool inside_x = xl >= 0 && xl + kw <= src.w;
    const int lane = lane_id();
    double acc = delta;
    for (int j0 = 0; j0 < kh; j0 += 64) {
        const int j = j0 + lane;
        double t = 0.0;
        if (j < kh) {
            const int sy = border_index(y + j - ay, src.h, border);
            if (sy < 0) {
                for (int i = 0; i < kw; ++i) t = t + kx[i] * bval;
            } else {
                const Ti* row = reinterpret_cast<const Ti*>(static_cast<const char*>(src.data) +
                                                            static_cast<size_t>(sy) * src.stride) + c;
                if (inside_x) {
                    const Ti* p = row + static_cast<size_t>(xl) * cn;
#pragma unroll 8
                    for (int i = 0; i < kw; ++i) t = t + kx[i] * static_cast<double>(p[static_cast<size_t>(i) * cn]);
                } else {
                    for (int i = 0; i < kw; ++i) {
                        const int sx = border_index(xl + i, src.w, border);
                        const double v = sx < 0 ? bval : static_cast<double>(row[static_cast<size_t>(sx) * cn]);
                        t = t + kx[i] * v;
                    }
                }
            }
        }
        const int jn = min(64, kh - j0);
        for (int jj = 0; jj < jn; ++jj) acc = acc + ky[j0 + jj] * bcast_d(t, jj);
    }
    return acc;
}

struct SepArgs {
    DevImage src, dst;
    const void* kx;      // kw weights of type Acc, zero-padded
    const void* ky;      // kh weights of type Acc, zero-padded
    const double* kxd;   // FP64 weights for exact recomputation
    const double* kyd;
    int kw, kh, ax, ay, border;
    double delta, bval;
    double bcoef;        // certification bound = bcoef * X + bdelta, X = max |input| of the tile
    double bdelta;
    double xbound;       // fixed X for integer images (< 0: measure the tile)
    double oscale;       // results are multiplied by this (1 for sepFilter2D; 1 / (kw kh) for a mean box)
    int unit;            // every weight is 1 (box): the strip kernel slides its sums
    int SH;
    FixList fix;
    double kabs;         // sum|kx| * sum|ky| (bounds |result - delta| / X)
};

template <class Ti, class Acc, int CN, int SJ, bool CERT>
__global__ __launch_bounds__(kBlock) void sep_fast_kernel(SepArgs a, const Acc* __restrict__ kx,
                                                          const Acc* __restrict__ ky) {
    extern __shared__ __align__(16) unsigned char sep_smem[];
    constexpr int TX = tile_x<CN>();
    constexpr int G = TX / P;
    const int W = a.src.w, H = a.src.h;
    const int cn = CN > 0 ? CN : a.src.cn;
    const int kw = a.kw, kh = a.kh;
    const SepTile T = sep_tile(TX, cn, kw, kh);
    Acc* ring = reinterpret_cast<Acc*>(sep_smem);
    float* in_buf = reinterpret_cast<float*>(sep_smem + sep_ring_bytes(T, sizeof(Acc)));
    __shared__ unsigned int tile_max_bits;  // max |staged value| as float bits
    const Ti* __restrict__ src = static_cast<const Ti*>(a.src.data);
    const int tid = threadIdx.x;
    const int x0 = blockIdx.x * TX, y0 = blockIdx.y * a.SH;
    const int rows_out = min(a.SH, H - y0);
    const int nh = rows_out + kh - 1;
    const int nhstep = (nh + CH - 1) / CH;
    const int nstep = T.K0 + (rows_out + CH - 1) / CH;
    const float bvalf = static_cast<float>(a.bval);
    const bool measure = CERT && a.xbound < 0;
    if (tid == 0) tile_max_bits = 0u;
    float my_max = 0.f;
    Stager<Ti, SJ> st;
    auto conv = [&](Ti v) {
        const float f = static_cast<float>(v);
        if (measure) my_max = fmaxf(my_max, fabsf(f));  // NaN/inf inputs give NaN/inf sums: flagged anyway
        return f;
    };
    const int ncol = TX * cn;
    const int gx0 = x0 - a.ax, gy0 = y0 - a.ay;

    st.load(src, a.src.stride, W, H, cn, a.border, gx0, gy0, min(CH, nh), T.XS, tid);
    for (int k = 0; k < nstep; ++k) {
        const int r0 = k * CH;
        const int nr = min(CH, nh - r0);
        // Constant border: the stager left image samples at the outside positions
        const bool cpatch = nr > 0 && a.border == IPFG_BORDER_CONSTANT &&
                            chunk_outside(gx0, gy0 + r0, TX + kw - 1, nr, W, H);
        if (nr > 0) st.store(in_buf, T.XP, nr, T.XS, tid, conv);
        if (cpatch) {
            for (int e = tid; e < nr * T.XS; e += kBlock) {
                const int r = e / T.XS, s = e - r * T.XS;
                const int gy = gy0 + r0 + r, gx = gx0 + s / cn;
                if (gy < 0 || gy >= H || gx < 0 || gx >= W) in_buf[r * T.XP + s] = bvalf;
            }
        }
        if (measure) {
            if (cpatch) my_max = fmaxf(my_max, fabsf(bvalf));
            wave_max_into(&tile_max_bits, my_max);
        }
        __syncthreads();
        if (k + 1 < nhstep)
            st.load(src, a.src.stride, W, H, cn, a.border, gx0, gy0 + r0 + CH, min(CH, nh - r0 - CH), T.XS, tid);
        // ---- horizontal taps (reference order: i ascending from 0) into the ring
        if (nr > 0) {
            for (int t = tid; t < CH * G * cn; t += kBlock) {
                const int cg = t % (G * cn);
                const int r = t / (G * cn);
                if (r >= nr) continue;
                const int g = cg / cn, c = cg - g * cn;
                const float* base = in_buf + r * T.XP + g * P * cn + c;
                Acc acc[P], win[P];
                if (CN == 1) {
                    const float4 f = *reinterpret_cast<const float4*>(base);
                    win[0] = f.x;
                    win[1] = f.y;
                    win[2] = f.z;
                    win[3] = f.w;
                } else {
#pragma unroll
                    for (int m = 0; m < P; ++m) win[m] = static_cast<Acc>(base[m * cn]);
                }
#pragma unroll
                for (int m = 0; m < P; ++m) acc[m] = Acc(0);
                Acc kc[4];
                load4(kx, 0, kc);
                int i0 = 0;
                for (; i0 + P <= kw; i0 += P) {
                    Acc kn[4];
                    load4(kx, i0 + P, kn);
                    Acc nx[P];
                    if (CN == 1) {
                        const float4 f = *reinterpret_cast<const float4*>(base + i0 + P);
                        nx[0] = f.x;
                        nx[1] = f.y;
                        nx[2] = f.z;
                        nx[3] = f.w;
                    } else {
#pragma unroll
                        for (int ii = 0; ii < P; ++ii) nx[ii] = base[(i0 + ii + P) * cn];
                    }
#pragma unroll
                    for (int ii = 0; ii < P; ++ii) {
#pragma unroll
                        for (int m = 0; m < P; ++m) acc[m] = fma_acc(kc[ii], win[(ii + m) % P], acc[m]);
                        win[ii] = nx[ii];
                    }
#pragma unroll
                    for (int ii = 0; ii < 4; ++ii) kc[ii] = kn[ii];
                }
#pragma unroll
                for (int ii = 0; ii < P; ++ii) {
                    if (i0 + ii < kw) {
#pragma unroll
                        for (int m = 0; m < P; ++m) acc[m] = fma_acc(kc[ii], win[(ii + m) % P], acc[m]);
                        win[ii] = base[(i0 + ii + P) * cn];
                    }
                }
                Acc* hrow = ring + ((r0 + r) & (T.R - 1)) * T.HP + c;
                if (CN == 1) {  // P consecutive outputs: 16-byte stores, no bank conflicts
                    if constexpr (sizeof(Acc) == 8) {
                        reinterpret_cast<double2*>(hrow + g * P)[0] = make_double2(acc[0], acc[1]);
                        reinterpret_cast<double2*>(hrow + g * P)[1] = make_double2(acc[2], acc[3]);
                    } else {
                        *reinterpret_cast<float4*>(hrow + g * P) = make_float4(acc[0], acc[1], acc[2], acc[3]);
                    }
                } else {
#pragma unroll
                    for (int m = 0; m < P; ++m) hrow[(g * P + m) * cn] = acc[m];
                }
            }
        }
        __syncthreads();
        // ---- vertical taps for output rows [(k - K0) * CH, (k - K0 + 1) * CH)
        const int ob = (k - T.K0) * CH;
        if (ob < 0) continue;
        double bound = 0.0;
        if (CERT) bound = a.bcoef * (measure ? static_cast<double>(__uint_as_float(tile_max_bits)) : a.xbound) + a.bdelta;
        for (int t = tid; t < (CH / M) * ncol; t += kBlock) {
            const int col = t % ncol;
            const int v = t / ncol;
            const int yr = ob + v * M;  // first output row of this task (strip-relative)
            const int px = col / cn, c = col - px * cn;
            const int x = x0 + px;
            const size_t q = static_cast<size_t>(x) * cn + c;
            unsigned int dmask = 0u;  // doubtful outputs of this task
            if (yr < rows_out) {
                const Acc* colp = ring + col;
                const int msk = T.R - 1;
                Acc acc[M], win[M];
#pragma unroll
                for (int m = 0; m < M; ++m) {
                    acc[m] = static_cast<Acc>(a.delta);
                    win[m] = colp[((yr + m) & msk) * T.HP];
                }
                Acc kc[M];
                load4(ky, 0, kc);
                int j0 = 0;
                for (; j0 + M <= kh; j0 += M) {
                    Acc kn[M];
                    load4(ky, j0 + M, kn);
                    Acc nx[M];
#pragma unroll
                    for (int jj = 0; jj < M; ++jj) nx[jj] = colp[((yr + j0 + jj + M) & msk) * T.HP];
#pragma unroll
                    for (int jj = 0; jj < M; ++jj) {
#pragma unroll
                        for (int m = 0; m < M; ++m) acc[m] = fma_acc(kc[jj], win[(jj + m) % M], acc[m]);
                        win[jj] = nx[jj];
                    }
#pragma unroll
                    for (int jj = 0; jj < M; ++jj) kc[jj] = kn[jj];
                }
#pragma unroll
                for (int jj = 0; jj < M; ++jj) {
                    if (j0 + jj < kh) {
#pragma unroll
                        for (int m = 0; m < M; ++m) acc[m] = fma_acc(kc[jj], win[(jj + m) % M], acc[m]);
                        win[jj] = colp[((yr + j0 + jj + M) & msk) * T.HP];
                    }
                }
                if (x < W) {
#pragma unroll
                    for (int m = 0; m < M; ++m) {
                        if (yr + m >= rows_out) break;
                        const int y = y0 + yr + m;
                        char* drow = static_cast<char*>(a.dst.data) + static_cast<size_t>(y) * a.dst.stride;
                        const double r = static_cast<double>(acc[m]) * a.oscale;
                        bool doubt = false;
                        switch (a.dst.dt) {
                            case IPFG_U8:
                                reinterpret_cast<uint8_t*>(drow)[q] = saturate_from_double<uint8_t>(r);
                                if (CERT) doubt = needs_exact(r, bound, 0.0, 255.0);
                                break;
                            case IPFG_U16:
                                reinterpret_cast<uint16_t*>(drow)[q] = saturate_from_double<uint16_t>(r);
                                if (CERT) doubt = needs_exact(r, bound, 0.0, 65535.0);
                                break;
                            case IPFG_I16:
                                reinterpret_cast<int16_t*>(drow)[q] = saturate_from_double<int16_t>(r);
                                if (CERT) doubt = needs_exact(r, bound, -32768.0, 32767.0);
                                break;
                            case IPFG_I32:
                                reinterpret_cast<int32_t*>(drow)[q] = saturate_from_double<int32_t>(r);
                                if (CERT) doubt = needs_exact(r, bound, -2147483648.0, 2147483647.0);
                                break;
                            case IPFG_F32:
                                reinterpret_cast<float*>(drow)[q] = static_cast<float>(r);
                                if (CERT) doubt = needs_exact_f32(r, bound);
                                break;
                            default:
                                reinterpret_cast<double*>(drow)[q] = r;
                                break;
                        }
                        if (CERT && doubt) dmask |= 1u << m;
                    }
                }
            }
            if (CERT) {
                const size_t es = dtype_size(a.dst.dt);
                wave_fix(a.fix, dmask, [&](int l, int m) {
                    const int xl = bcast_i(x, l), cl = bcast_i(c, l);
                    const int yl = y0 + bcast_i(yr, l) + m;
                    const double e = sep_exact_wave<Ti>(a.src, xl, yl, cl, a.kxd, kw, a.kyd, kh, a.ax, a.ay, a.border,
                                                        a.delta, a.bval) * a.oscale;
                    if (lane_id() == l)
                        store_sample(a.dst.data, static_cast<size_t>(yl) * a.dst.stride + (static_cast<size_t>(xl) * cn + cl) * es,
                                     a.dst.dt, e);
                });
            }
        }
    }
    if (CERT) fix_finish(a.fix);
}

// ------------------------------------------------------------------ planar kernel
// Channel counts 1, 3 and 4: staged rows are de-interleaved into one LDS plane
// per channel, so every plane is a contiguous row of pixels and the
// horizontal pass reads its windows with 16-byte LDS loads and writes four
// results with one store. Ring rows are addressed once per group of M rows
// (groups never straddle the ring's wrap) with immediate offsets.
//
// KS > 0 specialises the kernel for kw = kh = KS: every tap loop is unrolled,
// the tile geometry is constant and weights become scalar loads at fixed
// offsets; KS = 0 reads the sizes at run time.
//
// Hr is the type of the horizontally filtered rows kept in the ring. For
// integer outputs a certified launch keeps them as float: the FP64 sums are
// rounded once (relative error 2^-24), which the caller adds to the bound,
// and the ring takes half the LDS.
constexpr int PTX = 64;  // tile width in pixels (per plane)

struct PlaneTile {
    int XS;     // staged samples per row: (PTX + kw - 1) * cn
    int XPP;    // staged plane row pitch (floats), multiple of 4, >= PTX + kw - 1 + P
    int R;      // ring rows per plane (power of two)
    int K0;     // steps before the first output rows
    int RPL;    // ring plane size (elements): R * PTX + pad
};

__host__ __device__ constexpr PlaneTile plane_tile(int cn, int kw, int kh) {
    PlaneTile t{};
    t.XS = (PTX + kw - 1) * cn;
    t.XPP = ((PTX + kw - 1 + P + 3) / 4) * 4;
    t.K0 = (kh - 1 + CH - 1) / CH;
    int r = CH;
    while (r < (t.K0 + 1) * CH) r *= 2;
    t.R = r;
    t.RPL = r * PTX + 4;
    return t;
}

__host__ inline size_t plane_lds_bytes(int cn, int kw, int kh, size_t ring_elem) {
    const PlaneTile t = plane_tile(cn, kw, kh);
    const size_t ring = ((static_cast<size_t>(cn) * t.RPL * ring_elem + 15) / 16) * 16;
    return ring + sizeof(float) * static_cast<size_t>(cn) * CH * t.XPP;
}

// Samples per lane of the stager for a kernel width (0: run time).
constexpr int plane_sj(int cn, int ks) { return ks ? ((PTX + ks - 1) * cn + 63) / 64 : 0; }

template <class To>
__device__ __forceinline__ void put(char* p, double r) {
    if constexpr (std::is_same<To, float>::value) *reinterpret_cast<float*>(p) = static_cast<float>(r);
    else if constexpr (std::is_same<To, double>::value) *reinterpret_cast<double*>(p) = r;
    else *reinterpret_cast<To*>(p) = saturate_from_double<To>(r);
}

template <class To>
__device__ __forceinline__ bool doubtful(double r, double bound) {
    if constexpr (std::is_same<To, float>::value) return needs_exact_f32(r, bound);
    else if constexpr (std::is_same<To, double>::value) return false;
    else
        return needs_exact(r, bound, static_cast<double>(std::numeric_limits<To>::lowest()),
                           static_cast<double>(std::numeric_limits<To>::max()));
}

template <class Hr, class Acc>
__device__ __forceinline__ void store_h4(Hr* p, const Acc (&v)[4]) {
    if constexpr (sizeof(Hr) == 8) {
        reinterpret_cast<double2*>(p)[0] = make_double2(v[0], v[1]);
        reinterpret_cast<double2*>(p)[1] = make_double2(v[2], v[3]);
    } else {
        *reinterpret_cast<float4*>(p) = make_float4(static_cast<float>(v[0]), static_cast<float>(v[1]),
                                                    static_cast<float>(v[2]), static_cast<float>(v[3]));
    }
}

// One strip (bx, by): PTX pixels wide, a.SH rows tall. The workgroup's LDS is
// sep_smem. Certified launches recompute doubtful samples in place
// (wave_fix); the caller runs fix_finish once per workgroup.
template <class Ti, class To, class Acc, class Hr, int CN, int SJ, bool CERT, int KS>
__device__ __forceinline__ void sep_plane_strip(const SepArgs& a, const Acc* __restrict__ kx,
                                                const Acc* __restrict__ ky, int bx, int by,
                                                unsigned char* __restrict__ sep_smem) {
    constexpr int G = PTX / P;                  // horizontal tasks per row
    constexpr int U = KS ? 2 : 1;              // unroll of the tap loops (more hoists weights into VGPRs)
    static_assert(CH * G == kBlock && (CH / M) * PTX == kBlock, "one task per thread per plane");
    const int W = a.src.w, H = a.src.h;
    const int kw = KS ? KS : a.kw, kh = KS ? KS : a.kh;
    const PlaneTile T = plane_tile(CN, kw, kh);
    Hr* ring = reinterpret_cast<Hr*>(sep_smem);
    float* in_buf = reinterpret_cast<float*>(sep_smem + ((static_cast<size_t>(CN) * T.RPL * sizeof(Hr) + 15) / 16) * 16);
    // Per-wave running maximum of |staged value| (float bits): each wave
    // rewrites its own word every chunk, readers take the maximum of the four.
    // The words only grow, so a reader that sees a later chunk's value still
    // gets an upper bound of every sample it uses.
    __shared__ __align__(16) unsigned int wave_max_bits[4];
    const Ti* __restrict__ src = static_cast<const Ti*>(a.src.data);
    // the wave index is uniform: keep it (and what derives from it) in scalar registers
    const int tid = threadIdx.x, wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
    const int x0 = bx * PTX, y0 = by * a.SH;
    const int rows_out = min(a.SH, H - y0);
    const int nh = rows_out + kh - 1;
    const int nhstep = (nh + CH - 1) / CH;
    const int nstep = T.K0 + (rows_out + CH - 1) / CH;
    const float bvalf = static_cast<float>(a.bval);
    const bool measure = CERT && a.xbound < 0;
    if (lane == 0) wave_max_bits[wave] = 0u;
    unsigned int my_max = 0u;  // float bits of max |staged value| of this lane
    const int gx0 = x0 - a.ax, gy0 = y0 - a.ay;
    const int plane_in = CH * T.XPP;
    const int msk = T.R - 1;
    // horizontal task of this thread: row hr of the chunk, pixels [hg * P, hg * P + P)
    const int hr = tid / G, hg = tid % G;
    // vertical task: column vx, rows [vv * M, vv * M + M) of the chunk (vv is wave-uniform)
    const int vx = lane, vv = wave;

    // One-channel float images are staged with 16-byte loads: slot j of this
    // thread covers chunk row vrow[j], samples [4 vseg[j], 4 vseg[j] + 4) of
    // the LR = ceil(XS / 4) segments of a row (SJ slots cover 16 rows). Other
    // images use the sample-per-lane Stager and de-interleave into planes.
    constexpr bool VEC = CN == 1 && std::is_same<Ti, float>::value;
    constexpr int NV = VEC ? SJ : 1;
    const int LR = (T.XS + 3) / 4;
    int vrow[NV], vseg[NV];
#pragma unroll
    for (int j = 0; j < NV; ++j) {
        const int q = tid + kBlock * j;
        vrow[j] = q / LR;
        vseg[j] = q - vrow[j] * LR;
    }
    struct VStage {
        float4 v[NV];
    };
    using Stage = typename std::conditional<VEC, VStage, Stager<Ti, SJ>>::type;
    // Strips that reach past the left or right edge map their columns once,
    // into LDS (kept out of registers: the map is read only by border chunks).
    __shared__ __align__(16) int colmap[VEC ? 4 * 128 : 1];
    const bool hin = gx0 >= 0 && gx0 + 4 * LR <= W;
    if constexpr (VEC) {
        if (!hin) {
            for (int px = tid; px < 4 * LR; px += kBlock)
                colmap[px] = max(border_index(gx0 + min(px, T.XS - 1), W, a.border), 0);  // Constant: overwritten later
            __syncthreads();
        }
    }

    // Two stagers alternate so that the loads of the next two chunks are in
    // flight while a chunk is filtered.
    Stage sa, sb;
    auto issue = [&](Stage& st, int k) {
        const int gyc = gy0 + k * CH, nr = min(CH, nh - k * CH);
        if constexpr (VEC) {
            const char* csrc = reinterpret_cast<const char*>(src);
            if (hin && gyc >= 0 && gyc + nr <= H) {
#pragma unroll
                for (int j = 0; j < NV; ++j) {  // slots past the chunk re-read its last row (never stored)
                    const int r = min(vrow[j], nr - 1);
                    const float* p = reinterpret_cast<const float*>(csrc + static_cast<size_t>(gyc + r) * a.src.stride) +
                                     gx0 + 4 * vseg[j];
                    __builtin_memcpy(&st.v[j], p, sizeof(float4));
                }
            } else {  // border chunk (Constant: any sample, overwritten at the store)
#pragma unroll
                for (int j = 0; j < NV; ++j) {
                    const int gy = gyc + min(vrow[j], nr - 1);
                    const int sy = (gy >= 0 && gy < H) ? gy : max(border_index(gy, H, a.border), 0);
                    const float* rp = reinterpret_cast<const float*>(csrc + static_cast<size_t>(sy) * a.src.stride);
                    if (hin) {
                        __builtin_memcpy(&st.v[j], rp + gx0 + 4 * vseg[j], sizeof(float4));
                    } else {
                        const int4 sx = *reinterpret_cast<const int4*>(colmap + 4 * vseg[j]);
                        st.v[j] = make_float4(rp[sx.x], rp[sx.y], rp[sx.z], rp[sx.w]);
                    }
                }
            }
        } else {
            st.load(src, a.src.stride, W, H, CN, a.border, gx0, gyc, nr, T.XS, tid);
        }
    };
    issue(sa, 0);
    if (nhstep > 1) issue(sb, 1);
    auto step = [&](Stage& st, const int k) {
        const int r0 = k * CH;
        const int nr = min(CH, nh - r0);
        if (nr > 0) {
            // Constant border: the stager left image samples at the outside positions
            const bool cpatch = a.border == IPFG_BORDER_CONSTANT && chunk_outside(gx0, gy0 + r0, T.XS / CN, nr, W, H);
            if constexpr (VEC) {
#pragma unroll
                for (int j = 0; j < NV; ++j) {
                    if (vrow[j] >= nr) continue;
                    float e[4] = {st.v[j].x, st.v[j].y, st.v[j].z, st.v[j].w};
                    if (cpatch) {
                        const int gy = gy0 + r0 + vrow[j];
#pragma unroll
                        for (int i = 0; i < 4; ++i) {
                            const int gx = gx0 + 4 * vseg[j] + i;
                            if (gy < 0 || gy >= H || gx < 0 || gx >= W) e[i] = bvalf;
                        }
                    }
                    if (measure) {  // |x| as bits; a NaN's bits exceed +inf's, so it poisons the bound
                        const unsigned int m01 = max(__float_as_uint(e[0]) & 0x7fffffffu, __float_as_uint(e[1]) & 0x7fffffffu);
                        const unsigned int m23 = max(__float_as_uint(e[2]) & 0x7fffffffu, __float_as_uint(e[3]) & 0x7fffffffu);
                        my_max = max(my_max, max(m01, m23));
                    }
                    *reinterpret_cast<float4*>(in_buf + vrow[j] * T.XPP + 4 * vseg[j]) = make_float4(e[0], e[1], e[2], e[3]);
                }
            } else {
                // de-interleave: sample s of staged row r -> plane s % CN, pixel s / CN
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int r = wave + 4 * i;
                    if (r >= nr) continue;
#pragma unroll
                    for (int j = 0; j < SJ; ++j) {
                        const int s = lane + 64 * j;
                        if (s < T.XS) {
                            const int px = s / CN, c = s - px * CN;
                            float f = static_cast<float>(st.val[i][j]);
                            if (cpatch) {
                                const int gy = gy0 + r0 + r, gx = gx0 + px;
                                if (gy < 0 || gy >= H || gx < 0 || gx >= W) f = bvalf;
                            }
                            if (measure) my_max = max(my_max, __float_as_uint(f) & 0x7fffffffu);
                            in_buf[c * plane_in + r * T.XPP + px] = f;
                        }
                    }
                }
            }
            if (measure) {
                const unsigned int wm = wave_max_u32(my_max);
                if (lane == 0) wave_max_bits[wave] = wm;
            }
        }
        __syncthreads();
        if (k + 2 < nhstep) issue(st, k + 2);  // two chunks in flight: this one's registers take chunk k + 2
        // ---- horizontal taps (i ascending from 0) into the ring
        if (nr > 0 && hr < nr) {
#pragma unroll 1
            for (int c = 0; c < CN; ++c) {
                const float* base = in_buf + c * plane_in + hr * T.XPP + hg * P;
                Acc acc[P], win[P];
                if (a.unit) {  // unit weights: one window sum, then slide it across the P outputs
                    Acc sum = Acc(0);
#pragma unroll U
                    for (int i = 0; i < kw; ++i) sum += static_cast<Acc>(base[i]);
                    acc[0] = sum;
#pragma unroll
                    for (int m = 1; m < P; ++m) {
                        sum += static_cast<Acc>(base[m - 1 + kw]) - static_cast<Acc>(base[m - 1]);
                        acc[m] = sum;
                    }
                    store_h4(ring + c * T.RPL + ((r0 + hr) & msk) * PTX + hg * P, acc);
                    continue;
                }
                {
                    const float4 f = *reinterpret_cast<const float4*>(base);
                    win[0] = f.x;
                    win[1] = f.y;
                    win[2] = f.z;
                    win[3] = f.w;
                }
#pragma unroll
                for (int m = 0; m < P; ++m) acc[m] = Acc(0);
                Acc kc[4];
                load4(kx, 0, kc);
                int i0 = 0;
#pragma unroll U
                for (; i0 + P <= kw; i0 += P) {
                    Acc kn[4];
                    load4(kx, i0 + P, kn);
                    const float4 f = *reinterpret_cast<const float4*>(base + i0 + P);
                    const Acc nx[P] = {static_cast<Acc>(f.x), static_cast<Acc>(f.y), static_cast<Acc>(f.z),
                                       static_cast<Acc>(f.w)};
#pragma unroll
                    for (int ii = 0; ii < P; ++ii) {
#pragma unroll
                        for (int m = 0; m < P; ++m) acc[m] = fma_acc(kc[ii], win[(ii + m) % P], acc[m]);
                        win[ii] = nx[ii];
                    }
#pragma unroll
                    for (int ii = 0; ii < 4; ++ii) kc[ii] = kn[ii];
                }
                if (i0 < kw) {
                    const float4 f = *reinterpret_cast<const float4*>(base + i0 + P);
                    const Acc nx[P] = {static_cast<Acc>(f.x), static_cast<Acc>(f.y), static_cast<Acc>(f.z),
                                       static_cast<Acc>(f.w)};
#pragma unroll
                    for (int ii = 0; ii < P; ++ii) {
                        if (i0 + ii < kw) {
#pragma unroll
                            for (int m = 0; m < P; ++m) acc[m] = fma_acc(kc[ii], win[(ii + m) % P], acc[m]);
                            win[ii] = nx[ii];
                        }
                    }
                }
                store_h4(ring + c * T.RPL + ((r0 + hr) & msk) * PTX + hg * P, acc);
            }
        }
        __syncthreads();
        // ---- vertical taps for output rows [(k - K0) * CH, (k - K0 + 1) * CH)
        const int ob = (k - T.K0) * CH;
        if (ob < 0) return;
        const int yr = ob + vv * M;  // first output row of this task (strip-relative), wave-uniform
        if (yr >= rows_out) return;
        double bound = 0.0;
        if (CERT) {
            double xb = a.xbound;
            if (measure) {
                const uint4 m4 = *reinterpret_cast<const uint4*>(wave_max_bits);
                const unsigned int mb = __builtin_amdgcn_readfirstlane(max(max(m4.x, m4.y), max(m4.z, m4.w)));
                xb = static_cast<double>(__uint_as_float(mb));
            }
            bound = a.bcoef * xb + a.bdelta;
        }
        const int xg = x0 + vx;
#pragma unroll 1
        for (int c = 0; c < CN; ++c) {
            const Hr* plane = ring + c * T.RPL + vx;
            Acc acc[M], win[M];
            if (a.unit) {  // unit weights: one window sum, then slide it down the M rows
                Acc sum = static_cast<Acc>(a.delta);
#pragma unroll U
                for (int j = 0; j < kh; ++j) sum += static_cast<Acc>(plane[((yr + j) & msk) * PTX]);
                acc[0] = sum;
#pragma unroll
                for (int m = 1; m < M; ++m) {
                    sum += static_cast<Acc>(plane[((yr + m - 1 + kh) & msk) * PTX]) -
                           static_cast<Acc>(plane[((yr + m - 1) & msk) * PTX]);
                    acc[m] = sum;
                }
            } else {
            {
                const Hr* g0 = plane + (yr & msk) * PTX;  // rows yr .. yr + M - 1 do not wrap
#pragma unroll
                for (int m = 0; m < M; ++m) {
                    acc[m] = static_cast<Acc>(a.delta);
                    win[m] = static_cast<Acc>(g0[m * PTX]);
                }
            }
            Acc kc[M];
            load4(ky, 0, kc);
            int j0 = 0;
#pragma unroll U
            for (; j0 + M <= kh; j0 += M) {
                Acc kn[M];
                load4(ky, j0 + M, kn);
                const Hr* gp = plane + ((yr + j0 + M) & msk) * PTX;
                Acc nx[M];
#pragma unroll
                for (int jj = 0; jj < M; ++jj) nx[jj] = static_cast<Acc>(gp[jj * PTX]);
#pragma unroll
                for (int jj = 0; jj < M; ++jj) {
#pragma unroll
                    for (int m = 0; m < M; ++m) acc[m] = fma_acc(kc[jj], win[(jj + m) % M], acc[m]);
                    win[jj] = nx[jj];
                }
#pragma unroll
                for (int jj = 0; jj < M; ++jj) kc[jj] = kn[jj];
            }
            if (j0 < kh) {
                const Hr* gp = plane + ((yr + j0 + M) & msk) * PTX;
#pragma unroll
                for (int jj = 0; jj < M; ++jj) {
                    if (j0 + jj < kh) {
#pragma unroll
                        for (int m = 0; m < M; ++m) acc[m] = fma_acc(kc[jj], win[(jj + m) % M], acc[m]);
                        win[jj] = static_cast<Acc>(gp[jj * PTX]);
                    }
                }
            }
            }
            unsigned int dmask = 0u;  // doubtful outputs of this task
            if (xg < W) {
                char* op = static_cast<char*>(a.dst.data) + static_cast<size_t>(y0 + yr) * a.dst.stride +
                           (static_cast<size_t>(xg) * CN + c) * sizeof(To);
#pragma unroll
                for (int m = 0; m < M; ++m) {
                    if (yr + m >= rows_out) break;
                    const double rv = static_cast<double>(acc[m]) * a.oscale;
                    put<To>(op, rv);
                    if (CERT && doubtful<To>(rv, bound)) dmask |= 1u << m;
                    op += a.dst.stride;
                }
            }
            if (CERT) {  // yr and c are wave-uniform: only the column comes from lane l
                wave_fix(a.fix, dmask, [&](int l, int m) {
                    const int xl = bcast_i(xg, l);
                    const int yl = y0 + yr + m;
                    const double e = sep_exact_wave<Ti>(a.src, xl, yl, c, a.kxd, kw, a.kyd, kh, a.ax, a.ay, a.border,
                                                        a.delta, a.bval) * a.oscale;
                    if (lane == l)
                        put<To>(static_cast<char*>(a.dst.data) + static_cast<size_t>(yl) * a.dst.stride +
                                    (static_cast<size_t>(xl) * CN + c) * sizeof(To),
                                e);
                });
            }
        }
    };
    for (int k = 0; k < nstep; k += 2) {
        step(sa, k);
        if (k + 1 < nstep) step(sb, k + 1);
    }
}

template <class Ti, class To, class Acc, class Hr, int CN, int SJ, bool CERT, int KS>
__global__ __launch_bounds__(kBlock) void sep_plane_kernel(SepArgs a, const Acc* __restrict__ kx,
                                                           const Acc* __restrict__ ky) {
    extern __shared__ __align__(16) unsigned char sep_smem[];
    sep_plane_strip<Ti, To, Acc, Hr, CN, SJ, CERT, KS>(a, kx, ky, blockIdx.x, blockIdx.y, sep_smem);
    if (CERT) fix_finish(a.fix);
}

template <class Ti, class Acc, int CN, int SJ, bool CERT>
hipError_t sep_fast_launch_v(const SepArgs& a, size_t lds, hipStream_t s) {
    constexpr int TX = tile_x<CN>();
    auto kern = sep_fast_kernel<Ti, Acc, CN, SJ, CERT>;
    if (lds > 64 * 1024) {
        static size_t raised = 0;  // largest dynamic LDS granted to this instantiation
        if (lds > raised) {
            const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                                     hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
            if (e != hipSuccess) return e;
            raised = lds;
        }
    }
    const dim3 grid(static_cast<unsigned>((a.src.w + TX - 1) / TX), static_cast<unsigned>((a.src.h + a.SH - 1) / a.SH));
    hipLaunchKernelGGL(kern, grid, dim3(kBlock), lds, s, a, static_cast<const Acc*>(a.kx),
                       static_cast<const Acc*>(a.ky));
    return hipGetLastError();
}

template <class Ti, class Acc, int CN, bool CERT>
hipError_t sep_fast_launch(SepArgs a, hipStream_t s) {
    constexpr int TX = tile_x<CN>();
    const int cn = a.src.cn;
    const size_t lds = sep_lds_bytes(TX, cn, a.kw, a.kh, sizeof(Acc));
    const int xs = (TX + a.kw - 1) * cn;  // staged samples per row
    if (lds > 156 * 1024 || xs > 64 * 8) return hipErrorInvalidConfiguration;
    // Strip height: long strips amortise the kh - 1 history rows, but the
    // grid must still cover the 256 CUs several times over.
    if (a.SH <= 0) {
        const long long gx = (a.src.w + TX - 1) / TX;
        int SH = 512;
        while (SH > CH && gx * ((a.src.h + SH - 1) / SH) < 2048) SH /= 2;
        a.SH = SH;
    }
    if (xs <= 128) return sep_fast_launch_v<Ti, Acc, CN, 2, CERT>(a, lds, s);
    return sep_fast_launch_v<Ti, Acc, CN, 8, CERT>(a, lds, s);
}

template <class Ti, class To, class Acc, class Hr, int CN, int SJ, bool CERT, int KS>
hipError_t plane_launch_v(const SepArgs& a, size_t lds, hipStream_t s) {
    auto kern = sep_plane_kernel<Ti, To, Acc, Hr, CN, SJ, CERT, KS>;
    if (lds > 64 * 1024) {
        static size_t raised = 0;  // largest dynamic LDS granted to this instantiation
        if (lds > raised) {
            const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                                     hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
            if (e != hipSuccess) return e;
            raised = lds;
        }
    }
    const dim3 grid(static_cast<unsigned>((a.src.w + PTX - 1) / PTX), static_cast<unsigned>((a.src.h + a.SH - 1) / a.SH));
    hipLaunchKernelGGL(kern, grid, dim3(kBlock), lds, s, a, static_cast<const Acc*>(a.kx),
                       static_cast<const Acc*>(a.ky));
    return hipGetLastError();
}

// Ring element type of a planar launch: float for FP32 launches and for
// certified launches with integer outputs, double for float outputs.
template <class To, class Acc>
using plane_hr = typename std::conditional<sizeof(Acc) == 8 && !std::is_integral<To>::value, double, float>::type;

template <class Ti, class To, class Acc, int CN, bool CERT, int KS>
hipError_t plane_launch(SepArgs a, hipStream_t s) {
    using Hr = plane_hr<To, Acc>;
    const size_t lds = plane_lds_bytes(CN, a.kw, a.kh, sizeof(Hr));
    const int xs = (PTX + a.kw - 1) * CN;
    if (lds > 156 * 1024 || xs > 64 * 8) return hipErrorInvalidConfiguration;
    if (a.SH <= 0) {  // long strips amortise the kh - 1 halo rows; keep >= 2048 workgroups
        const long long gx = (a.src.w + PTX - 1) / PTX;
        int SH = 512;
        while (SH > CH && gx * ((a.src.h + SH - 1) / SH) < 2048) SH /= 2;
        a.SH = SH;
    }
    if constexpr (KS > 0) {
        return plane_launch_v<Ti, To, Acc, Hr, CN, plane_sj(CN, KS), CERT, KS>(a, lds, s);
    } else {
        if (xs <= 128) return plane_launch_v<Ti, To, Acc, Hr, CN, 2, CERT, 0>(a, lds, s);
        return plane_launch_v<Ti, To, Acc, Hr, CN, 8, CERT, 0>(a, lds, s);
    }
}

// Kernel widths with a specialised planar kernel (kw = kh): every odd size
// from 3 to 49, which covers gaussianBlur for sigma in [0.5, 8].
constexpr int kSpecMin = 3, kSpecMax = 49;

template <class Ti, class To, class Acc, int CN, bool CERT, int KS = kSpecMin>
hipError_t plane_launch_k(const SepArgs& a, hipStream_t s) {
    if constexpr (KS > kSpecMax) {
        return hipErrorNotSupported;
    } else {
        if (a.kw == KS) return plane_launch<Ti, To, Acc, CN, CERT, KS>(a, s);
        return plane_launch_k<Ti, To, Acc, CN, CERT, KS + 2>(a, s);
    }
}

// True when the certified launch of this call keeps float rows in the ring.
inline bool plane_ring_is_float(int dst_dt) { return dtype_is_int(dst_dt); }

// Instantiation groups (one per translation unit). The _k groups return
// hipErrorNotSupported for sizes without a specialised kernel.
hipError_t plane_u8(const SepArgs& a, bool cert, hipStream_t s);
hipError_t plane_f32(const SepArgs& a, bool cert, hipStream_t s);
hipError_t plane_u8_k1(const SepArgs& a, hipStream_t s);
hipError_t plane_u8_k3(const SepArgs& a, hipStream_t s);
hipError_t plane_u8_k4(const SepArgs& a, hipStream_t s);
hipError_t plane_u8f_k1(const SepArgs& a, hipStream_t s);  // certified uint8 -> float32, 1 channel, 3 / 5 / 7 taps
hipError_t plane_f32_k1(const SepArgs& a, hipStream_t s);
hipError_t generic_any(const SepArgs& a, bool cert, hipStream_t s);

// One-pass tile kernel (sep_tile.hip): certified uint8 -> uint8, 1 / 3 / 4
// channels, kw = kh odd in [3, 15], centred anchor. Launches that carry
// strip height kTileSH + rows run it with tiles of `rows` (16 or 32; 0 means
// 32) output rows.
constexpr int kTileSH = 1 << 22;
bool tile_u8_fits(const DevImage& src, const DevImage& dst, int kw, int kh, int ax, int ay, bool unit);
hipError_t tile_u8(const SepArgs& a, int rows, hipStream_t s);

}  // namespace sepk
}  // namespace ipfg
