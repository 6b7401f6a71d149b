// Streaming kernels that measure the HBM rate this device reaches with plain
// 16-byte vector loads and stores: the practical ceiling against which the
// stencil passes' achieved bandwidth is read, and the reference kernel of
// known size for scaling the FETCH_SIZE / WRITE_SIZE counters.  No
// counterpart in the reference.
//
//   mode 0  copy   dst[i] = src[i]           (bytes read = bytes written = n)
//   mode 1  read   one word per block written (bytes read = n)
//   mode 2  write  dst[i] = constant         (bytes written = n)
//   + 4     with non-temporal loads and stores (plain ones otherwise)
//
// Each thread moves UNROLL 16-byte words per iteration of a grid-stride loop,
// all loads issued before the stores.
#include "common.h"

namespace {

constexpr int CAL_BLOCK = 256;
constexpr int CAL_UNROLL = 4;
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

template <bool NT>
__device__ __forceinline__ u32x4 ld(const u32x4 *p) {
    if constexpr (NT) return __builtin_nontemporal_load(p);
    else return *p;
}
template <bool NT>
__device__ __forceinline__ void st(u32x4 v, u32x4 *p) {
    if constexpr (NT) __builtin_nontemporal_store(v, p);
    else *p = v;
}

template <bool NT>
__global__ __launch_bounds__(CAL_BLOCK) void copy_kernel(const u32x4 *__restrict__ src, u32x4 *__restrict__ dst,
                                                         long long n16) {
    const long long stride = (long long)gridDim.x * CAL_BLOCK * CAL_UNROLL;
    for (long long base = (long long)blockIdx.x * CAL_BLOCK * CAL_UNROLL + threadIdx.x; base < n16;
         base += stride) {
        u32x4 v[CAL_UNROLL];
#pragma unroll
        for (int u = 0; u < CAL_UNROLL; ++u) {
            const long long i = base + (long long)u * CAL_BLOCK;
            if (i < n16) v[u] = ld<NT>(src + i);
        }
#pragma unroll
        for (int u = 0; u < CAL_UNROLL; ++u) {
            const long long i = base + (long long)u * CAL_BLOCK;
            if (i < n16) st<NT>(v[u], dst + i);
        }
    }
}

template <bool NT>
__global__ __launch_bounds__(CAL_BLOCK) void read_kernel(const u32x4 *__restrict__ src, unsigned *__restrict__ out,
                                                         long long n16) {
    const long long stride = (long long)gridDim.x * CAL_BLOCK * CAL_UNROLL;
    unsigned acc = 0u;
    for (long long base = (long long)blockIdx.x * CAL_BLOCK * CAL_UNROLL + threadIdx.x; base < n16;
         base += stride) {
        u32x4 v[CAL_UNROLL];
#pragma unroll
        for (int u = 0; u < CAL_UNROLL; ++u) {
            const long long i = base + (long long)u * CAL_BLOCK;
            v[u] = i < n16 ? ld<NT>(src + i) : u32x4{0u, 0u, 0u, 0u};
        }
#pragma unroll
        for (int u = 0; u < CAL_UNROLL; ++u) acc ^= v[u].x ^ v[u].y ^ v[u].z ^ v[u].w;
    }
    __shared__ unsigned red[CAL_BLOCK];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = CAL_BLOCK / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] ^= red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

template <bool NT>
__global__ __launch_bounds__(CAL_BLOCK) void write_kernel(u32x4 *__restrict__ dst, long long n16, unsigned val) {
    const long long stride = (long long)gridDim.x * CAL_BLOCK * CAL_UNROLL;
    const u32x4 v = {val, val, val, val};
    for (long long base = (long long)blockIdx.x * CAL_BLOCK * CAL_UNROLL + threadIdx.x; base < n16;
         base += stride) {
#pragma unroll
        for (int u = 0; u < CAL_UNROLL; ++u) {
            const long long i = base + (long long)u * CAL_BLOCK;
            if (i < n16) st<NT>(v, dst + i);
        }
    }
}

}  // namespace

extern "C" int pycles_stream(int mode, const void *src, void *dst, long long bytes, int blocks, void *stream) {
    const bool nt = (mode & 4) != 0;
    if (mode < 0 || mode > 6 || (mode & 3) == 3) {
        pycles_set_error("stream: bad mode %d", mode);
        return PYCLES_ERR_ARG;
    }
    if (bytes <= 0 || bytes % 16 != 0 || blocks < 1 || blocks > (1 << 20)) {
        pycles_set_error("stream: bad size %lld or grid %d", bytes, blocks);
        return PYCLES_ERR_ARG;
    }
    if ((mode != 2 && !src) || !dst || ((size_t)src % 16) || ((size_t)dst % 16)) {
        pycles_set_error("stream: NULL or unaligned pointer");
        return PYCLES_ERR_ARG;
    }
    mode &= 3;
    const long long n16 = bytes / 16;
    hipStream_t sm = (hipStream_t)stream;
    const dim3 grid(blocks), blk(CAL_BLOCK);
    if (mode == 0) {
        if (nt) hipLaunchKernelGGL(copy_kernel<true>, grid, blk, 0, sm, (const u32x4 *)src, (u32x4 *)dst, n16);
        else hipLaunchKernelGGL(copy_kernel<false>, grid, blk, 0, sm, (const u32x4 *)src, (u32x4 *)dst, n16);
    } else if (mode == 1) {
        if (nt) hipLaunchKernelGGL(read_kernel<true>, grid, blk, 0, sm, (const u32x4 *)src, (unsigned *)dst, n16);
        else hipLaunchKernelGGL(read_kernel<false>, grid, blk, 0, sm, (const u32x4 *)src, (unsigned *)dst, n16);
    } else {
        if (nt) hipLaunchKernelGGL(write_kernel<true>, grid, blk, 0, sm, (u32x4 *)dst, n16, 0x3f800000u);
        else hipLaunchKernelGGL(write_kernel<false>, grid, blk, 0, sm, (u32x4 *)dst, n16, 0x3f800000u);
    }
    PYCLES_CHECK_LAUNCH();
    return PYCLES_OK;
}
