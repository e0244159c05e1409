// Device helpers shared by the HIP translation units of the ED solver.
#pragma once
#include "edgpu_internal.h"

namespace edg {

// ---------------------------------------------------------------- helpers

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Sum NS per-thread values over the block; result valid in thread 0.
template <int NS, int BS = kBlock>
__device__ __forceinline__ void block_sums(double (&v)[NS], double (&out)[NS])
{
    __shared__ double red[NS][BS / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        double t = wave_sum(v[s]);
        if (lane == 0) red[s][w] = t;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            double t = 0.0;
#pragma unroll
            for (int k = 0; k < BS / 64; ++k) t += red[s][k];
            out[s] = t;
        }
    }
    __syncthreads();
}

// Write this block's NS partial sums to partial[blockIdx.x*NS + s].  A
// separate one-workgroup finisher kernel combines them in block order: the
// kernel boundary makes the partials visible across XCDs without the
// device-scope release fences a last-block-done scheme would need in every
// workgroup (each of which writes back the XCD's L2).
template <int NS, int BS = kBlock>
__device__ __forceinline__ void block_partial_at(double (&mine)[NS], double *partial, int slot)
{
    double blk[NS];
    block_sums<NS, BS>(mine, blk);
    if (threadIdx.x == 0)
#pragma unroll
        for (int s = 0; s < NS; ++s) partial[(long long)slot * NS + s] = blk[s];
}
template <int NS, int BS = kBlock>
__device__ __forceinline__ void block_partial(double (&mine)[NS], double *partial)
{
    block_partial_at<NS, BS>(mine, partial, blockIdx.x);
}

// Sum of n partials in a fixed order: thread t adds partials t, t+BS, ...,
// then a fixed block tree; every workgroup of size BS gets the same value.
// Maximum of v over the workgroup; s holds BS / 64 doubles of LDS (it may be
// reused by the next call).
template <int BS>
__device__ __forceinline__ double block_max(double v, double *s)
{
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
    __syncthreads();
    if ((threadIdx.x & 63) == 0) s[threadIdx.x / 64] = v;
    __syncthreads();
    double m = s[0];
    for (int i = 1; i < BS / 64; ++i) m = fmax(m, s[i]);
    return m;
}

template <int BS>
__device__ __forceinline__ double sum_partials(const double *part, int n)
{
    __shared__ double s_sum;
    double v[1] = {0.0}, out[1];
    for (int i = threadIdx.x; i < n; i += BS) v[0] += part[i];
    block_sums<1, BS>(v, out);
    if (threadIdx.x == 0) s_sum = out[0];
    __syncthreads();
    return s_sum;
}

// Lanczos scales of step j for the H.v kernel: returns false when the chain
// has broken down (beta_j < 1e-16); the lead workgroup records beta_j and s_j.
template <int BS>
__device__ __forceinline__ bool lanczos_scales(const HvArgs &a, double &sx, double &bsp,
                                               bool lead_block)
{
    if (a.step == 0) {
        sx = a.scale[0];
        bsp = 0.0;
        return true;
    }
    const double bj = sqrt(sum_partials<BS>(a.prev_part, a.prev_n));
    const bool first = lead_block && threadIdx.x == 0;
    if (bj < 1.0e-16) {
        if (first) {
            a.beta[a.step] = bj;
            *a.stop = 1;
            *a.nstep = a.step;
        }
        return false;
    }
    sx = 1.0 / bj;
    bsp = bj * a.scale[a.step - 1];
    if (first) {
        a.beta[a.step] = bj;
        a.scale[a.step] = sx;
    }
    return true;
}


template <typename T, int C>
struct alignas(sizeof(T) * C) Row {
    T v[C];
};

// Loads and stores through the global address space.  Pointers that a
// kernel reads out of a descriptor in memory (rather than receiving as
// arguments) are generic to the compiler, which then emits flat accesses
// with 64-bit vector addresses; these keep them global.
// Aggregates (hop records, rows) are loaded as 32-bit vectors: clang copies
// an aggregate through a generic pointer, which would undo the cast.
template <int A>
struct U32x2;
template <>
struct U32x2<4> {
    typedef uint32_t type __attribute__((ext_vector_type(2), aligned(4)));
};
template <>
struct U32x2<8> {
    typedef uint32_t type __attribute__((ext_vector_type(2)));
};
template <typename T>
__device__ __forceinline__ T gload(const T *p)
{
    if constexpr (__is_scalar(T)) {
        return *(const __attribute__((address_space(1))) T *)p;
    } else if constexpr (sizeof(T) % 16 == 0 && alignof(T) >= 16) {
        typedef uint32_t u4 __attribute__((ext_vector_type(4)));
        constexpr int n = (int)(sizeof(T) / 16);
        const __attribute__((address_space(1))) u4 *q = (const __attribute__((address_space(1))) u4 *)p;
        u4 w[n];
#pragma unroll
        for (int i = 0; i < n; ++i) w[i] = q[i];
        T r;
        __builtin_memcpy(&r, w, sizeof(T));
        return r;
    } else {
        static_assert(sizeof(T) % 8 == 0 && alignof(T) >= 4, "gload: 8-byte multiple aggregates");
        using u2 = typename U32x2<(alignof(T) < 8 ? alignof(T) : 8)>::type;
        constexpr int n = (int)(sizeof(T) / 8);
        const __attribute__((address_space(1))) u2 *q = (const __attribute__((address_space(1))) u2 *)p;
        u2 w[n];
#pragma unroll
        for (int i = 0; i < n; ++i) w[i] = q[i];
        T r;
        __builtin_memcpy(&r, w, sizeof(T));
        return r;
    }
}

// Raw buffer loads: the array's base lives in a 4-register scalar resource,
// each lane gives a 32-bit byte offset and the wave a 32-bit scalar byte
// offset.  The down-hop gathers of a tile read C x KH neighbour columns at
// wave-uniform column offsets and one per-lane row offset: as buffer loads
// they need no per-gather address arithmetic (a global load needs a 64-bit
// vector address or a vector add of the column offset).  Word 3 is the
// raw-buffer setting of CDNA (32-bit data format, no swizzle); accesses at
// or beyond `bytes` read 0.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void *p, long long bytes)
{
    const int n = bytes < 0x7FFFFFFFLL ? (int)bytes : 0x7FFFFFFF;
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, n, 0x00020000);
}
template <typename T>
__device__ __forceinline__ T bload(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff);
template <>
__device__ __forceinline__ float bload<float>(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 0));
}
template <>
__device__ __forceinline__ double bload<double>(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff)
{
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, (int)soff, 0));
}

template <typename T>
__device__ __forceinline__ void bstore(__amdgpu_buffer_rsrc_t r, uint32_t voff, T v);
template <>
__device__ __forceinline__ void bstore<float>(__amdgpu_buffer_rsrc_t r, uint32_t voff, float v)
{
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, (int)voff, 0, 0);
}
template <>
__device__ __forceinline__ void bstore<double>(__amdgpu_buffer_rsrc_t r, uint32_t voff, double v)
{
    typedef uint32_t u2 __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2, v), r, (int)voff, 0, 0);
}

// Fields of a 32-bit hop record (k_compress): the target rank in bits 0-15
// and the value-table index in bits 24-31, bits 16-23 zero, so that the LDS
// address of the target row is one 24-bit multiply-add and the value index
// one byte select.
__device__ __forceinline__ uint32_t rec_rank(uint32_t w) { return w & 0xFFFFFFu; }
__device__ __forceinline__ uint32_t rec_vidx(uint32_t w) { return w >> 24; }
// The staged row a record points at, in a row-interleaved LDS tile of C
// columns of T starting at `tile`: one 24-bit multiply-add of the record
// (the compiler turns a plain multiply by the row size into a shift, a mask
// and an add).  Arithmetic only; the load is an ordinary LDS read.
template <typename T, int C>
__device__ __forceinline__ Row<T, C> rec_row(const T *tile, uint32_t w)
{
    typedef const __attribute__((address_space(3))) T *LdsT;
    const uint32_t base = (uint32_t)(size_t)(LdsT)tile;
    uint32_t a;
    asm("v_mad_u32_u24 %0, %1, %2, %3" : "=v"(a) : "v"(w), "i"((int)(C * sizeof(T))), "v"(base));
    const LdsT p = (LdsT)(size_t)a;
    Row<T, C> r;
#pragma unroll
    for (int c = 0; c < C; ++c) r.v[c] = p[c];
    return r;
}

// A value or pointer the whole wave holds, moved to scalar registers.
__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }
template <typename T>
__device__ __forceinline__ T *uniform(T *p)
{
    const unsigned long long a = (unsigned long long)p;
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)a);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(a >> 32));
    return (T *)(((unsigned long long)hi << 32) | lo);
}
template <typename T>
__device__ __forceinline__ void gstore(T *p, const T &v)
{
    if constexpr (__is_scalar(T)) {
        *(__attribute__((address_space(1))) T *)p = v;
    } else {
        static_assert(sizeof(T) % 16 == 0 && alignof(T) >= 16, "gstore: 16-byte multiple aggregates");
        typedef uint32_t u4 __attribute__((ext_vector_type(4)));
        constexpr int n = (int)(sizeof(T) / 16);
        u4 w[n];
        __builtin_memcpy(w, &v, sizeof(T));
        __attribute__((address_space(1))) u4 *q = (__attribute__((address_space(1))) u4 *)p;
#pragma unroll
        for (int i = 0; i < n; ++i) q[i] = w[i];
    }
}



template <typename ACC>
__device__ __forceinline__ ACC energy_of(const SpinTab &t, int r);
template <>
__device__ __forceinline__ float energy_of<float>(const SpinTab &t, int r) { return gload(t.energyf + r); }
template <>
__device__ __forceinline__ double energy_of<double>(const SpinTab &t, int r) { return gload(t.energy + r); }

// XCD-aware group order.  Workgroups of one launch are dispatched round-robin
// over the 8 XCDs (workgroup b runs on XCD b % 8), each with its own 4 MiB L2.
// Giving XCD x a contiguous range of column groups keeps neighbour columns of
// the down hops in that XCD's L2: in increasing-pattern (colex) order the high
// bath bits change slowly, so a contiguous rank range is closed under most
// hops.  Bijective for any ngroups.
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_group(int b, int ngroups)
{
    const int q = ngroups / kXcds, r = ngroups - q * kXcds;
    const int x = b % kXcds, i = b / kXcds;
    return x * q + (x < r ? x : r) + i;
}

template <typename ACC>
struct HopOf;
template <>
struct HopOf<float> {
    using type = HopF;
    static __device__ __forceinline__ const HopF *tab(const SpinTab &t) { return t.hf; }
};
template <>
struct HopOf<double> {
    using type = HopD;
    static __device__ __forceinline__ const HopD *tab(const SpinTab &t) { return t.hd; }
};

}  // namespace edg
