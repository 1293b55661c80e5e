// The host steps of DTFE() around the tessellation, on the device.
//
// Reference (file:line relative to the reference root):
//   findParticlesInBox / insertParticlesInBox ... src/subpartition.h:79-154
//   averageDensity ............................... src/subpartition.h:437-449
//   vertex data set from the inserted particle ... src/CGAL_triangulation/triangulation.cpp:176-185
//
// Periodic images are a stream compaction over (offset, particle) pairs in
// the reference's order: offset-major over the 27 translates, particles in
// input order within an offset. Three launches: per-workgroup counts for each
// offset, one scan of the counts, and the write, which recomputes the
// membership masks instead of storing them. Every pass reads the positions
// once, coalesced; HBM-bound.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <limits>

#include "context.h"

namespace dtfe_front {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;
constexpr int kPer = 4;                    // particles per thread
constexpr int kTile = kThreads * kPer;     // particles per workgroup
constexpr int kOffsets = 27;
constexpr int kScanThreads = 1024;
constexpr int kScanWaves = kScanThreads / 64;
constexpr int kSumParts = 1024;

// The translated padded boxes, per axis: a box translated by offset
// (i1, i2, i3) spans [lo[0][i1], hi[0][i1]] x [lo[1][i2], hi[1][i2]] x
// [lo[2][i3], hi[2][i3]]; an image is written as position - off[axis][i].
template <typename T>
struct ImageArgs {
    const T *pos;  // [n][3]
    int64_t n;
    int64_t n_tiles;
    T lo[3][3];
    T hi[3][3];
    T off[3][3];
};

template <typename T>
__device__ inline uint32_t image_mask(const ImageArgs<T> &a, const T q[3]) {
    uint32_t ax[3];
    for (int d = 0; d < 3; ++d) {
        uint32_t m = 0;
        for (int i = 0; i < 3; ++i) m |= (uint32_t)(q[d] >= a.lo[d][i] && q[d] <= a.hi[d][i]) << i;
        ax[d] = m;
    }
    uint32_t mask = 0;
    for (int i1 = 0; i1 < 3; ++i1)
        for (int i2 = 0; i2 < 3; ++i2)
            for (int i3 = 0; i3 < 3; ++i3)
                mask |= (uint32_t)((ax[0] >> i1) & (ax[1] >> i2) & (ax[2] >> i3) & 1u) << (i1 * 9 + i2 * 3 + i3);
    return mask;
}

__device__ inline uint32_t wave_or(uint32_t v) {
    for (int s = 32; s > 0; s >>= 1) v |= (uint32_t)__shfl_xor((int)v, s, 64);
    return (uint32_t)__builtin_amdgcn_readfirstlane((int)v);
}

__device__ inline int lanes_below(uint64_t b) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(b >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)b, 0));
}

template <typename T>
__device__ inline uint32_t load_masks(const ImageArgs<T> &a, int64_t base, uint32_t mk[kPer], T q[kPer][3]) {
    uint32_t any = 0;
    for (int k = 0; k < kPer; ++k) {
        const int64_t p = base + (int64_t)k * kThreads + threadIdx.x;
        mk[k] = 0;
        if (p < a.n) {
            for (int d = 0; d < 3; ++d) q[k][d] = a.pos[3 * p + d];
            mk[k] = image_mask(a, q[k]);
        }
        any |= mk[k];
    }
    return any;
}

// counts[t * n_tiles + tile]: images of offset t among the tile's particles
template <typename T>
__global__ __launch_bounds__(kThreads) void k_image_count(ImageArgs<T> a, int *__restrict__ counts) {
    __shared__ int cnt[kOffsets];
    if (threadIdx.x < kOffsets) cnt[threadIdx.x] = 0;
    __syncthreads();
    uint32_t mk[kPer];
    T q[kPer][3];
    const uint32_t any = load_masks(a, (int64_t)blockIdx.x * kTile, mk, q);
    uint32_t todo = wave_or(any);
    while (todo) {
        const int t = __builtin_ctz(todo);
        todo &= todo - 1;
        int c = 0;
        for (int k = 0; k < kPer; ++k) c += __popcll(__ballot((mk[k] >> t) & 1u));
        if ((threadIdx.x & 63) == 0) atomicAdd(&cnt[t], c);
    }
    __syncthreads();
    if (threadIdx.x < kOffsets) counts[threadIdx.x * a.n_tiles + blockIdx.x] = cnt[threadIdx.x];
}

// offsets[i] = sum of counts[0..i), offsets[K] = total. One workgroup; wave w
// scans the w-th contiguous segment 64 entries at a time.
__global__ __launch_bounds__(kScanThreads) void k_image_scan(const int *__restrict__ counts, int64_t K,
                                                             int64_t *__restrict__ offsets) {
    __shared__ int64_t wsum[kScanWaves];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int64_t seg = ((K + kScanWaves - 1) / kScanWaves + 63) / 64 * 64;
    const int64_t lo = std::min<int64_t>(K, w * seg), hi = std::min<int64_t>(K, lo + seg);
    int64_t s = 0;
    for (int64_t i = lo + lane; i < hi; i += 64) s += counts[i];
    for (int d = 32; d > 0; d >>= 1) s += __shfl_xor(s, d, 64);
    if (lane == 0) wsum[w] = s;
    __syncthreads();
    int64_t carry = 0, total = 0;
    for (int v = 0; v < kScanWaves; ++v) {
        if (v < w) carry += wsum[v];
        total += wsum[v];
    }
    for (int64_t c0 = lo; c0 < hi; c0 += 64) {
        const int64_t i = c0 + lane;
        const int64_t x = i < hi ? counts[i] : 0;
        int64_t inc = x;
        for (int d = 1; d < 64; d <<= 1) {
            const int64_t y = __shfl_up(inc, d, 64);
            if (lane >= d) inc += y;
        }
        if (i < hi) offsets[i] = carry + inc - x;
        carry += __shfl(inc, 63, 64);
    }
    if (threadIdx.x == 0) offsets[K] = total;
}

template <typename T>
__global__ __launch_bounds__(kThreads) void k_image_write(ImageArgs<T> a, const int64_t *__restrict__ offsets,
                                                          T *__restrict__ image_pos, int32_t *__restrict__ image_of) {
    __shared__ uint32_t block_any;
    __shared__ int wc[kPer][kWaves];
    if (threadIdx.x == 0) block_any = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * kTile;
    uint32_t mk[kPer];
    T q[kPer][3];
    const uint32_t any = wave_or(load_masks(a, base, mk, q));
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) atomicOr(&block_any, any);
    __syncthreads();
    uint32_t todo = block_any;
    while (todo) {
        const int t = __builtin_ctz(todo);
        todo &= todo - 1;
        uint64_t b[kPer];
        for (int k = 0; k < kPer; ++k) {
            b[k] = __ballot((mk[k] >> t) & 1u);
            if (lane == 0) wc[k][wave] = __popcll(b[k]);
        }
        __syncthreads();
        int64_t run = offsets[(int64_t)t * a.n_tiles + blockIdx.x];
        const int i1 = t / 9, i2 = (t / 3) % 3, i3 = t % 3;
        for (int k = 0; k < kPer; ++k) {
            int before = 0, tot = 0;
            for (int v = 0; v < kWaves; ++v) {
                before += v < wave ? wc[k][v] : 0;
                tot += wc[k][v];
            }
            if ((mk[k] >> t) & 1u) {
                const int64_t idx = run + before + lanes_below(b[k]);
                image_pos[3 * idx + 0] = q[k][0] - a.off[0][i1];
                image_pos[3 * idx + 1] = q[k][1] - a.off[1][i2];
                image_pos[3 * idx + 2] = q[k][2] - a.off[2][i3];
                image_of[idx] = (int32_t)(base + (int64_t)k * kThreads + threadIdx.x);
            }
            run += tot;
        }
        __syncthreads();
    }
}

// Fixed-order FP64 sum: workgroup b sums its contiguous chunk thread-strided,
// then a tree in LDS; k_sum_final adds the partial sums the same way.
__device__ inline double block_sum(double s) {
    __shared__ double red[kThreads];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int h = kThreads / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    return red[0];
}

template <typename T>
__global__ __launch_bounds__(kThreads) void k_sum_partial(const T *__restrict__ w, int64_t n, int64_t chunk,
                                                          double *__restrict__ part) {
    const int64_t lo = (int64_t)blockIdx.x * chunk, hi = std::min<int64_t>(n, lo + chunk);
    double s = 0.0;
    for (int64_t i = lo + threadIdx.x; i < hi; i += kThreads) s += (double)w[i];
    const double r = block_sum(s);
    if (threadIdx.x == 0) part[blockIdx.x] = r;
}

__global__ __launch_bounds__(kThreads) void k_sum_final(const double *__restrict__ part, int parts,
                                                        double *__restrict__ out) {
    double s = 0.0;
    for (int i = threadIdx.x; i < parts; i += kThreads) s += part[i];
    const double r = block_sum(s);
    if (threadIdx.x == 0) *out = r;
}

// One row per thread; an index outside [0, n_src) zero-fills its row (gather)
// or is skipped (scatter) and raises *bad.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_gather_rows(const int32_t *__restrict__ index, int64_t n_rows,
                                                          int width, int64_t n_src, const T *__restrict__ src,
                                                          T *__restrict__ dst, int *__restrict__ bad) {
    const int64_t stride = (int64_t)gridDim.x * kThreads;
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n_rows; i += stride) {
        const int64_t j = index[i];
        const bool ok = j >= 0 && j < n_src;
        for (int c = 0; c < width; ++c) dst[i * width + c] = ok ? src[j * width + c] : T(0);
        if (!ok) atomicOr(bad, 1);
    }
}

template <typename T>
__global__ __launch_bounds__(kThreads) void k_scatter_add_rows(const int32_t *__restrict__ index, int64_t n_rows,
                                                               int width, int64_t n_dst, const T *__restrict__ src,
                                                               T *__restrict__ dst, int *__restrict__ bad) {
    const int64_t stride = (int64_t)gridDim.x * kThreads;
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n_rows; i += stride) {
        const int64_t j = index[i];
        if (j < 0 || j >= n_dst) {
            atomicOr(bad, 1);
            continue;
        }
        for (int c = 0; c < width; ++c) atomicAdd(&dst[j * width + c], src[i * width + c]);
    }
}

__global__ void k_zero_int(int *p) { *p = 0; }

}  // namespace dtfe_front

using namespace dtfe_front;

namespace {

// The translated boxes in T arithmetic, as the reference forms them in Real:
// L = full[2d+1] - full[2d], offset = T(i - 1) * L, box coordinate + offset.
template <typename T>
ImageArgs<T> image_args(const T *pos, int64_t n, const double full_box[6], const double padded_box[6], bool periodic) {
#pragma clang fp contract(off)
    ImageArgs<T> a{};
    a.pos = pos;
    a.n = n;
    a.n_tiles = (n + kTile - 1) / kTile;
    for (int d = 0; d < 3; ++d) {
        const T L = (T)full_box[2 * d + 1] - (T)full_box[2 * d];
        for (int i = 0; i < 3; ++i) {
            const T off = (T)(i - 1) * L;
            if (periodic || i == 1) {
                a.off[d][i] = periodic ? off : T(0);
                a.lo[d][i] = (T)padded_box[2 * d] + a.off[d][i];
                a.hi[d][i] = (T)padded_box[2 * d + 1] + a.off[d][i];
            } else {  // empty interval: only the untranslated box holds images
                a.off[d][i] = T(0);
                a.lo[d][i] = std::numeric_limits<T>::infinity();
                a.hi[d][i] = -std::numeric_limits<T>::infinity();
            }
        }
    }
    return a;
}

template <typename T>
int periodic_images(dtfe_ctx *ctx, int64_t n, const T *pos, const double *full_box, const double *padded_box,
                    bool periodic, int64_t capacity, T *image_pos, int32_t *image_of, int64_t *n_images) {
    ImageArgs<T> a = image_args<T>(pos, n, full_box, padded_box, periodic);
    const int64_t K = kOffsets * a.n_tiles;
    const size_t count_bytes = ((size_t)K * sizeof(int) + 255) / 256 * 256;
    int rc = ensure(ctx, ctx->front, count_bytes + (size_t)(K + 1) * sizeof(int64_t));
    if (rc) return rc;
    int *counts = (int *)ctx->front.p;
    int64_t *offsets = (int64_t *)((char *)ctx->front.p + count_bytes);
    if (a.n_tiles > 0)
        hipLaunchKernelGGL(k_image_count<T>, dim3((unsigned)a.n_tiles), dim3(kThreads), 0, ctx->stream, a, counts);
    hipLaunchKernelGGL(k_image_scan, dim3(1), dim3(kScanThreads), 0, ctx->stream, counts, K, offsets);
    HIP_TRY(hipGetLastError());
    int64_t total = 0;
    HIP_TRY(hipMemcpyAsync(&total, offsets + K, sizeof(total), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    *n_images = total;
    if (!image_pos && !image_of) return DTFE_OK;
    if (!image_pos || !image_of) return fail(ctx, DTFE_ERR_ARGUMENT, "image_pos and image_of go together");
    if (total > capacity)
        return fail(ctx, DTFE_ERR_ARGUMENT, "capacity " + std::to_string(capacity) + " below the " +
                                                std::to_string(total) + " images");
    if (a.n_tiles > 0)
        hipLaunchKernelGGL(k_image_write<T>, dim3((unsigned)a.n_tiles), dim3(kThreads), 0, ctx->stream, a, offsets,
                           image_pos, image_of);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return DTFE_OK;
}

template <typename T>
int average_density(dtfe_ctx *ctx, int64_t n, const T *w, double *rho_bar) {
    int rc = ensure(ctx, ctx->front, (kSumParts + 1) * sizeof(double));
    if (rc) return rc;
    double *part = (double *)ctx->front.p;
    const int64_t per = (int64_t)kThreads * 16;
    const int parts = (int)std::max<int64_t>(1, std::min<int64_t>(kSumParts, (n + per - 1) / per));
    const int64_t chunk = (n + parts - 1) / parts;
    hipLaunchKernelGGL(k_sum_partial<T>, dim3(parts), dim3(kThreads), 0, ctx->stream, w, n, chunk, part);
    hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(kThreads), 0, ctx->stream, part, parts, part + kSumParts);
    HIP_TRY(hipGetLastError());
    double total = 0.0;
    HIP_TRY(hipMemcpyAsync(&total, part + kSumParts, sizeof(total), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    // Box::volume() of the reference (src/box.h:141-148)
    *rho_bar = total / (35000. * 35000. * 35000.);
    return DTFE_OK;
}

template <typename T>
int rows_call(dtfe_ctx *ctx, bool scatter, int64_t n_rows, const int32_t *index, int64_t n_other, int width,
              const T *src, T *dst) {
    int rc = ensure(ctx, ctx->front, 256);
    if (rc) return rc;
    int *bad = (int *)ctx->front.p;
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(1), 0, ctx->stream, bad);
    const int grid = grid_for(n_rows, kThreads);
    if (n_rows > 0) {
        if (scatter)
            hipLaunchKernelGGL(k_scatter_add_rows<T>, dim3(grid), dim3(kThreads), 0, ctx->stream, index, n_rows,
                               width, n_other, src, dst, bad);
        else
            hipLaunchKernelGGL(k_gather_rows<T>, dim3(grid), dim3(kThreads), 0, ctx->stream, index, n_rows, width,
                               n_other, src, dst, bad);
    }
    HIP_TRY(hipGetLastError());
    int hbad = 0;
    HIP_TRY(hipMemcpyAsync(&hbad, bad, sizeof(hbad), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (hbad) return fail(ctx, DTFE_ERR_ARGUMENT, "row index outside [0, n)");
    return DTFE_OK;
}

int check_rows(dtfe_ctx *ctx, int64_t n_rows, const int32_t *index, int64_t n_other, int dtype, int width,
               const void *src, void *dst) {
    if (dtype != DTFE_F32 && dtype != DTFE_F64) return fail(ctx, DTFE_ERR_ARGUMENT, "bad dtype");
    if (n_rows < 0 || n_other < 0 || width < 1) return fail(ctx, DTFE_ERR_ARGUMENT, "bad sizes");
    if (n_other > INT32_MAX) return fail(ctx, DTFE_ERR_ARGUMENT, "more than 2^31 rows to index");
    if (n_rows > 0 && (!index || !src || !dst)) return fail(ctx, DTFE_ERR_ARGUMENT, "NULL array");
    return DTFE_OK;
}

}  // namespace

extern "C" {

int dtfe_hip_periodic_images(dtfe_ctx *ctx, int64_t n, int dtype, const void *pos, const double full_box[6],
                             const double padded_box[6], int periodic, int64_t capacity, void *image_pos,
                             int32_t *image_of, int64_t *n_images) {
    if (!ctx) return DTFE_ERR_ARGUMENT;
    if (dtype != DTFE_F32 && dtype != DTFE_F64) return fail(ctx, DTFE_ERR_ARGUMENT, "bad dtype");
    if (!full_box || !padded_box || !n_images) return fail(ctx, DTFE_ERR_ARGUMENT, "NULL argument");
    if (n < 0 || n > INT32_MAX) return fail(ctx, DTFE_ERR_ARGUMENT, "particle count outside [0, 2^31)");
    if (n > 0 && !pos) return fail(ctx, DTFE_ERR_ARGUMENT, "pos is NULL");
    if (capacity < 0) return fail(ctx, DTFE_ERR_ARGUMENT, "negative capacity");
    *n_images = 0;
    HIP_TRY(hipSetDevice(ctx->device));
    if (dtype == DTFE_F64)
        return periodic_images<double>(ctx, n, (const double *)pos, full_box, padded_box, periodic != 0, capacity,
                                       (double *)image_pos, image_of, n_images);
    return periodic_images<float>(ctx, n, (const float *)pos, full_box, padded_box, periodic != 0, capacity,
                                  (float *)image_pos, image_of, n_images);
}

int dtfe_hip_average_density(dtfe_ctx *ctx, int64_t n, int dtype, const void *weight, double *rho_bar) {
    if (!ctx) return DTFE_ERR_ARGUMENT;
    if (dtype != DTFE_F32 && dtype != DTFE_F64) return fail(ctx, DTFE_ERR_ARGUMENT, "bad dtype");
    if (!rho_bar || n < 0 || (n > 0 && !weight)) return fail(ctx, DTFE_ERR_ARGUMENT, "bad arguments");
    HIP_TRY(hipSetDevice(ctx->device));
    if (dtype == DTFE_F64) return average_density<double>(ctx, n, (const double *)weight, rho_bar);
    return average_density<float>(ctx, n, (const float *)weight, rho_bar);
}

int dtfe_hip_gather_rows(dtfe_ctx *ctx, int64_t n_rows, const int32_t *index, int64_t n_src, int dtype, int width,
                         const void *src, void *dst) {
    if (!ctx) return DTFE_ERR_ARGUMENT;
    int rc = check_rows(ctx, n_rows, index, n_src, dtype, width, src, dst);
    if (rc) return rc;
    HIP_TRY(hipSetDevice(ctx->device));
    if (dtype == DTFE_F64)
        return rows_call<double>(ctx, false, n_rows, index, n_src, width, (const double *)src, (double *)dst);
    return rows_call<float>(ctx, false, n_rows, index, n_src, width, (const float *)src, (float *)dst);
}

int dtfe_hip_scatter_add_rows(dtfe_ctx *ctx, int64_t n_rows, const int32_t *index, int64_t n_dst, int dtype,
                              int width, const void *src, void *dst) {
    if (!ctx) return DTFE_ERR_ARGUMENT;
    int rc = check_rows(ctx, n_rows, index, n_dst, dtype, width, src, dst);
    if (rc) return rc;
    HIP_TRY(hipSetDevice(ctx->device));
    if (dtype == DTFE_F64)
        return rows_call<double>(ctx, true, n_rows, index, n_dst, width, (const double *)src, (double *)dst);
    return rows_call<float>(ctx, true, n_rows, index, n_dst, width, (const float *)src, (float *)dst);
}

}  // extern "C"
