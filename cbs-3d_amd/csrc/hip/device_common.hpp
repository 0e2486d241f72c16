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

    // Returns true in every thread of the last block to finish.  Each block
    // must have written its partial results (by thread 0) before the call.
    __device__ inline bool last_block_done(unsigned int* counter)
    {
        __shared__ bool am_last;
        __threadfence();
        __syncthreads();
        if (threadIdx.x == 0)
        {
            const unsigned int prev = atomicAdd(counter, 1u);
            am_last = (prev == gridDim.x - 1);
        }
        __syncthreads();
        if (am_last) __threadfence();
        return am_last;
    }

    __device__ inline double load_coherent(const double* p)
    {
        return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }

    // Sum of `count` partials (stride `stride`, component `comp`) by one block
    // in a fixed order: thread t sums t, t+256, ..., then a block reduction.
    __device__ inline double reduce_partials(const double* part, int count, int stride, int comp,
                                             double (*lds)[kBlock / 64])
    {
        double v[1] = {0.0};
        for (int i = threadIdx.x; i < count; i += kBlock) v[0] += load_coherent(part + size_t(i) * stride + comp);
        block_sum<1>(v, lds);
        return v[0];
    }

    __device__ inline double max_partials(const double* part, int count, int stride, int comp, double* lds)
    {
        double m = 0.0;
        for (int i = threadIdx.x; i < count; i += kBlock)
        {
            const double w = load_coherent(part + size_t(i) * stride + comp);
            m = (m < w) ? w : m;
        }
        return block_max(m, lds);
    }

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
