// Device work of index creation and incremental update (spf_index_update):
// entry arrays gathered from the previous mirror or from uploaded rows, and
// the arrays derived from them.
//
//   k_gather_tiled   tiled 16-byte units of each new entry from the old tiled
//                    array or from the tiled upload of changed postings
//   k_gather_rows    f32 rows (dpad stride) from old rows or uploaded rows
//   k_tiled_f16      tiled f16 copy of tiled f32 codes (round to nearest even)
//   k_post_bounds    per posting max |x - x16| and max |x16|^2 (the pruned
//                    rerank's inputs), in double, rounded up
//   k_entry_post     posting of every entry
//   k_id_entry       entry of every id (UINT32_MAX = none)
#include "spf_internal.h"

namespace spf {

__global__ void k_gather_tiled(const uint4 *old_t, const uint4 *up_t, const uint32_t *src,
                               uint32_t n, uint32_t units, uint4 *dst) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (uint64_t)n * units) return;
    const uint32_t e = (uint32_t)(i / units), u = (uint32_t)(i - (uint64_t)e * units);
    const uint32_t s = src[e];
    const uint32_t j = s & 0x7fffffffu;
    const uint4 *from = (s & 0x80000000u) ? up_t : old_t;
    dst[((size_t)(e >> 6) * units + u) * 64 + (e & 63)] = from[((size_t)(j >> 6) * units + u) * 64 + (j & 63)];
}

__global__ void k_gather_rows(const float *old_r, const float *up_r, const uint32_t *src,
                              uint32_t n, uint32_t dpad, uint32_t up_stride, uint32_t dim,
                              float *dst) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (uint64_t)n * dpad) return;
    const uint32_t e = (uint32_t)(i / dpad), c = (uint32_t)(i - (uint64_t)e * dpad);
    const uint32_t s = src[e], j = s & 0x7fffffffu;
    float v;
    if (s & 0x80000000u)
        v = c < dim ? up_r[(size_t)j * up_stride + c] : 0.0f;
    else
        v = old_r[(size_t)j * dpad + c];
    dst[(size_t)e * dpad + c] = v;
}

// unit u16 of an f16 entry (8 halves) = f32 units 2 u16 and 2 u16 + 1
__global__ void k_tiled_f16(const uint4 *f32t, uint32_t units32, uint32_t n, uint4 *f16t,
                            uint32_t units16) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (uint64_t)n * units16) return;
    const uint32_t e = (uint32_t)(i / units16), u = (uint32_t)(i - (uint64_t)e * units16);
    const uint4 *src = f32t + (size_t)(e >> 6) * units32 * 64 + (e & 63);
    const uint4 a = src[(size_t)(2 * u) * 64], b = src[(size_t)(2 * u + 1) * 64];
    const float f[8] = {__uint_as_float(a.x), __uint_as_float(a.y), __uint_as_float(a.z),
                        __uint_as_float(a.w), __uint_as_float(b.x), __uint_as_float(b.y),
                        __uint_as_float(b.z), __uint_as_float(b.w)};
    uint32_t w[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const uint32_t lo = __builtin_bit_cast(uint16_t, (_Float16)f[2 * k]);
        const uint32_t hi = __builtin_bit_cast(uint16_t, (_Float16)f[2 * k + 1]);
        w[k] = lo | (hi << 16);
    }
    f16t[((size_t)(e >> 6) * units16 + u) * 64 + (e & 63)] = make_uint4(w[0], w[1], w[2], w[3]);
}

__device__ __forceinline__ uint32_t posting_of(const uint32_t *post_off, uint32_t n_post, uint32_t e) {
    uint32_t lo = 0, hi = n_post;   // post_off[lo] <= e < post_off[lo + 1]
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (post_off[mid] <= e) lo = mid;
        else hi = mid;
    }
    return lo;
}

// 64 entries per workgroup: their postings found by 64 threads at once, then
// each wave takes 16 of them one after another, lane l over dims l, l + 64, ...
// (the row read coalesced, once): |x - x16| and |x16|^2 in double (x16 = the
// tiled f16 code when given, else the raw row rounded to f16), rounded up, max
// per posting by atomicMax on the bits (non-negative floats order as their bits).
__device__ __forceinline__ void max_bits(uint32_t *a, uint32_t v) {
    // a long posting's entries would queue on one address: the atomic only
    // when the value seen is below (a stale read just costs the atomic)
    if (__atomic_load_n(a, __ATOMIC_RELAXED) < v) atomicMax(a, v);
}
__global__ __launch_bounds__(256) void k_post_bounds(const float *raw_e, const uint4 *codes16,
                                                     uint32_t units16, uint32_t n, uint32_t dpad,
                                                     uint32_t dim, const uint32_t *post_off,
                                                     uint32_t n_post, uint32_t *post_ex,
                                                     uint32_t *post_xn) {
    __shared__ uint32_t s_post[64];
    const uint32_t e0 = blockIdx.x * 64, lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (threadIdx.x < 64 && e0 + threadIdx.x < n) s_post[threadIdx.x] = posting_of(post_off, n_post, e0 + threadIdx.x);
    __syncthreads();
    for (uint32_t k = 16 * wave; k < 16 * wave + 16; k++) {
        const uint32_t e = e0 + k;
        if (e >= n) break;   // whole wave
        const float *x = raw_e + (size_t)e * dpad;
        const uint4 *c = codes16 ? codes16 + (size_t)(e >> 6) * units16 * 64 + (e & 63) : nullptr;
        double s2 = 0.0, n2 = 0.0;
        for (uint32_t j = lane; j < dim; j += 64) {
            double h;
            if (c) {
                const uint4 v = c[(size_t)(j >> 3) * 64];
                const uint32_t w = (j & 7) < 4 ? ((j & 7) < 2 ? v.x : v.y) : ((j & 7) < 6 ? v.z : v.w);
                const uint16_t b = (uint16_t)((j & 1) ? (w >> 16) : (w & 0xffffu));
                h = (double)(float)__builtin_bit_cast(_Float16, b);
            } else {
                h = (double)(float)(_Float16)x[j];
            }
            const double t = (double)x[j] - h;
            s2 += t * t;
            n2 += h * h;
        }
        for (int off = 32; off > 0; off >>= 1) {
            s2 += __shfl_xor(s2, off, 64);
            n2 += __shfl_xor(n2, off, 64);
        }
        if (lane == 0) {
            const uint32_t p = s_post[k];
            const float ex = (float)(sqrt(s2) * (1.0 + 1e-12));
            const float xn = (float)(n2 * (1.0 + 1e-12));
            max_bits(post_ex + p, __float_as_uint(ex) + 1u);   // the next float up
            max_bits(post_xn + p, __float_as_uint(xn) + 1u);
        }
    }
}

__global__ void k_entry_post(const uint32_t *post_off, uint32_t n_post, uint32_t n, uint32_t *out) {
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    out[e] = posting_of(post_off, n_post, e);
}

__global__ void k_fill_u32(uint32_t *p, uint64_t n, uint32_t v) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_id_entry(const uint32_t *entry_id, uint32_t n, uint32_t *id_entry) {
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) id_entry[entry_id[e]] = e;
}

__global__ void k_iota_u64(uint64_t *p, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = i;
}

static inline uint32_t blocks_for(uint64_t n) { return (uint32_t)((n + 255) / 256); }

hipError_t launch_gather_tiled(const uint4 *old_t, const uint4 *up_t, const uint32_t *src,
                               uint32_t n, uint32_t units, uint4 *dst, hipStream_t s) {
    if (!n || !units) return hipSuccess;
    hipLaunchKernelGGL(k_gather_tiled, dim3(blocks_for((uint64_t)n * units)), dim3(256), 0, s,
                       old_t, up_t, src, n, units, dst);
    return hipGetLastError();
}

hipError_t launch_gather_rows(const float *old_r, const float *up_r, const uint32_t *src,
                              uint32_t n, uint32_t dpad, uint32_t up_stride, uint32_t dim,
                              float *dst, hipStream_t s) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_gather_rows, dim3(blocks_for((uint64_t)n * dpad)), dim3(256), 0, s,
                       old_r, up_r, src, n, dpad, up_stride, dim, dst);
    return hipGetLastError();
}

hipError_t launch_tiled_f16(const uint4 *f32t, uint32_t units32, uint32_t n, uint4 *f16t,
                            uint32_t units16, hipStream_t s) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_tiled_f16, dim3(blocks_for((uint64_t)n * units16)), dim3(256), 0, s,
                       f32t, units32, n, f16t, units16);
    return hipGetLastError();
}

hipError_t launch_post_bounds(const float *raw_e, const uint4 *codes16, uint32_t units16,
                              uint32_t n, uint32_t dpad, uint32_t dim, const uint32_t *post_off,
                              uint32_t n_post, uint32_t *post_ex, uint32_t *post_xn,
                              hipStream_t s) {
    if (!n || !n_post) return hipSuccess;
    hipLaunchKernelGGL(k_post_bounds, dim3((n + 63) / 64), dim3(256), 0, s, raw_e, codes16,
                       units16, n, dpad, dim, post_off, n_post, post_ex, post_xn);
    return hipGetLastError();
}

hipError_t launch_entry_post(const uint32_t *post_off, uint32_t n_post, uint32_t n, uint32_t *out,
                             hipStream_t s) {
    if (!n || !n_post) return hipSuccess;
    hipLaunchKernelGGL(k_entry_post, dim3(blocks_for(n)), dim3(256), 0, s, post_off, n_post, n, out);
    return hipGetLastError();
}

hipError_t launch_id_entry(const uint32_t *entry_id, uint32_t n, uint32_t *id_entry, uint64_t span,
                           hipStream_t s) {
    if (span) hipLaunchKernelGGL(k_fill_u32, dim3(blocks_for(span)), dim3(256), 0, s, id_entry, span,
                                 0xffffffffu);
    if (n) hipLaunchKernelGGL(k_id_entry, dim3(blocks_for(n)), dim3(256), 0, s, entry_id, n, id_entry);
    return hipGetLastError();
}

hipError_t launch_iota_u64(uint64_t *p, uint64_t n, hipStream_t s) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_iota_u64, dim3(blocks_for(n)), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

}  // namespace spf
