// Device helpers shared by the kernels: error recording, deterministic
// wavefront/block reductions and the last-block-done pattern.
#pragma once

#include "ctx.hpp"

namespace cbsdev
{
    __device__ inline void raise_error(DevScalars* s, unsigned long long key)
    {
        atomicMin(&s->err, key);
    }

    // Butterfly over the 64-lane wavefront: every lane ends with the same sum
    // (each step adds the same two operands in every lane), so the result is
    // deterministic for a fixed launch shape.
    __device__ inline double wave_sum(double v)
    {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
        return v;
    }

    __device__ inline double wave_max(double v)
    {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1)
        {
            const double w = __shfl_xor(v, o);
            v = (v < w) ? w : v;
        }
        return v;
    }

    // Block-wide sum of NV values per thread (blockDim.x == kBlock).
    template <int NV>
    __device__ inline void block_sum(double (&v)[NV], double (*lds)[kBlock / 64])
    {
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
        for (int i = 0; i < NV; ++i)
        {
            v[i] = wave_sum(v[i]);
            if (lane == 0) lds[i][wave] = v[i];
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < NV; ++i)
        {
            double s = 0.0;
#pragma unroll
            for (int w = 0; w < kBlock / 64; ++w) s += lds[i][w];
            v[i] = s;
        }
        __syncthreads();
    }

    __device__ inline double block_max(double v, double* lds)
    {
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        v = wave_max(v);
        if (lane == 0) lds[wave] = v;
        __syncthreads();
        double m = lds[0];
#pragma unroll
        for (int w = 1; w < kBlock / 64; ++w) m = (m < lds[w]) ? lds[w] : m;
        __syncthreads();
        return m;
    }

    // Sum of `count` partials (stride `stride`, component `comp`) by one block
    // in a fixed order: thread t sums t, t+256, ..., then a block reduction.
    // The partials come from an earlier kernel on the same stream, so plain
    // loads see them: reductions are always split into a kernel that writes
    // per-block partials and a one-block kernel that combines them.  (A
    // last-block-done scheme inside one kernel needs an agent-scope release
    // fence per block, and on a multi-XCD part each such fence writes back
    // the XCD's L2: hundreds of microseconds per launch.)
    __device__ inline double reduce_partials(const double* part, int count, int stride, int comp,
                                             double (*lds)[kBlock / 64])
    {
        double v[1] = {0.0};
#pragma unroll 4
        for (int i = threadIdx.x; i < count; i += kBlock) v[0] += part[size_t(i) * stride + comp];
        block_sum<1>(v, lds);
        return v[0];
    }

    // reduce_partials of components 0 and 1 and max_partials of component 2
    // of one set of per-block partials (stride 4), in one pass: a thread's
    // loads of all three go out together.  Each result is the same sum (or
    // maximum) in the same order as the separate calls give.
    __device__ inline void reduce3_partials(const double* part, int count, double (*lds)[kBlock / 64],
                                            double* ldm, double& s0, double& s1, double& mx)
    {
        double v[2] = {0.0, 0.0};
        double m = 0.0;
#pragma unroll 4
        for (int i = threadIdx.x; i < count; i += kBlock)
        {
            const double a = part[size_t(i) * 4], b = part[size_t(i) * 4 + 1], w = part[size_t(i) * 4 + 2];
            v[0] += a;
            v[1] += b;
            m = (m < w) ? w : m;
        }
        block_sum<2>(v, lds);
        s0 = v[0];
        s1 = v[1];
        mx = block_max(m, ldm);
    }

    __device__ inline double max_partials(const double* part, int count, int stride, int comp, double* lds)
    {
        double m = 0.0;
        for (int i = threadIdx.x; i < count; i += kBlock)
        {
            const double w = part[size_t(i) * stride + comp];
            m = (m < w) ? w : m;
        }
        return block_max(m, lds);
    }

    // Block sum for any block size up to 1024 (multiple of 64); the result
    // is valid in thread 0 only.  Waves are combined in wave order.
    __device__ inline double block_sum_dyn(double v, double* lds /* >= 16 */)
    {
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        v = wave_sum(v);
        if (lane == 0) lds[wave] = v;
        __syncthreads();
        double s = 0.0;
        if (threadIdx.x == 0)
            for (int w = 0; w < int(blockDim.x >> 6); ++w) s += lds[w];
        return s;
    }

    // XCD-aware block index.  gfx950 deals a grid's workgroups to its 8 XCDs
    // round-robin (block b to XCD b % 8) and each XCD has its own L2.  The
    // remap gives XCD x one contiguous range of logical blocks (x*q +
    // min(x, r) onwards, q = n / 8, r = n % 8), so neighbouring blocks, which
    // share most of the element data a node gather reads, share one L2.  It
    // is a bijection of [0, n) for any n; only locality depends on it.
    __device__ inline int xcd_block(int n)
    {
        constexpr int kXcd = 8;
        const int b = int(blockIdx.x);
        const int q = n / kXcd, r = n % kXcd, x = b % kXcd;
        return x * q + (x < r ? x : r) + b / kXcd;
    }

    // A node the pressure solve updates: touches fluid, pressure not fixed.
    __device__ inline bool is_free(uint8_t f) { return (f & 1u) && !(f & 4u); }

    // Multi-column node vectors: R = 2^rshift columns (R <= kMultiMaxCols)
    // in blocks of 8 nodes, [N8/8][R][8] with N8 = N rounded up to a multiple
    // of 8: column c of node i is element ((i/8) R + c) 8 + i%8, and element
    // q is column (q/8) % R of node (q / 8R) 8 + q%8.  Eight consecutive
    // nodes of one column are contiguous, so a workgroup that handles a few
    // columns of a run of nodes reads and writes whole cache lines.  A
    // grid-stride loop over the elements with a stride that is a multiple of
    // 8R (256 threads per block) keeps each thread on one column.
    constexpr int kMultiMaxCols = 32;

    __host__ __device__ inline size_t mcol_at(int i, int c, int rshift)
    {
        return ((((size_t(i) >> 3) << rshift) + size_t(c)) << 3) | size_t(i & 7);
    }

    __host__ __device__ inline int mcol_node(size_t q, int rshift)
    {
        return int(((q >> (3 + rshift)) << 3) | (q & 7));
    }

    __host__ __device__ inline int mcol_col(size_t q, int rshift)
    {
        return int((q >> 3) & ((size_t(1) << rshift) - 1));
    }

    __host__ __device__ inline size_t mcol_size(int n, int rshift) { return size_t((n + 7) & ~7) << rshift; }

    inline int reduce_grid(int64_t n)
    {
        const int64_t b = (n + kBlock - 1) / kBlock;
        return int(b < kReduceBlocks ? (b < 1 ? 1 : b) : kReduceBlocks);
    }

    inline int grid_for(int64_t n)
    {
        const int64_t b = (n + kBlock - 1) / kBlock;
        return int(b < 1 ? 1 : b);
    }
}
