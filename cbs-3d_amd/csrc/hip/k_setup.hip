// Device-side set-up: the element colouring of Coloring::build
// (include/cbs/parallel/Coloring.hpp:27-97) on the GPU, with the same
// result bit for bit.
//
// The reference colours elements greedily in index order: element e takes
// the smallest colour no earlier element sharing a node with it holds
// (first fit).  That colouring is the only solution of
//   colour(e) = mex { colour(f) : f < e, f shares a node with e },
// since colour(e) is fixed once the colours of its lower neighbours are.  So
// every element can be coloured as soon as all its lower neighbours are,
// in any order, and the result is the reference's.  The device does exactly
// that, level by level (a parallel topological order of the "lower
// neighbour" relation):
//  1. node -> element lists, each sorted ascending (fill, then a per-node
//     insertion sort);
//  2. pending(e) = the number of (node, lower element) incidences of e; the
//     elements with none form the first level;
//  3. one kernel per level colours its elements from their lower
//     neighbours' colours and decrements pending(g) of every higher
//     neighbour g once per shared node; g joins the next level when its
//     count reaches zero.  Levels are launched in batches without waiting
//     for the host; a level with no element returns at once.  At cavity160
//     (24.6 M tets) there are ~4,500 levels of ~5,500 elements;
//  4. colour-major order with elements ascending inside a colour
//     (Coloring.hpp:75-96): per-chunk colour counts, per-colour offsets,
//     then each chunk places its elements in index order.
// Colours are limited to kMaxColours (the reference has no limit; meshes
// that need more are reported as unsupported and stay with the host
// colouring).
#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "ctx.hpp"
#include "device_common.hpp"

namespace cbsdev
{
    namespace
    {
        constexpr int kMaxColours = 256;
        constexpr int kLevelGrid = 256;        // blocks of kBlock threads per level kernel
        constexpr int kLevelsPerBatch = 64;    // level kernels queued between host checks
        constexpr int kChunk = 4096;           // elements per chunk of the colour-major sort
        constexpr int kChunkThreads = 64;      // chunks per block (LDS: 64 x 256 counters)

        enum ColourError : int
        {
            kNodeOutOfRange = 1,
            kTooManyColours = 2,
        };

        __global__ void k_node_degree(int64_t n4, int N, const int32_t* __restrict__ intma, int32_t* __restrict__ deg,
                                      int* __restrict__ err)
        {
            const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
            if (i >= n4) return;
            const int32_t ip = intma[i];
            if (ip < 1 || ip > N)
            {
                atomicOr(err, kNodeOutOfRange);
                return;
            }
            atomicAdd(&deg[ip - 1], 1);
        }

        // inclusive prefix of the degrees into ptr[1..N] (ptr[0] = 0): one
        // block, each thread scans a contiguous range, then the range sums
        __global__ void k_node_scan(int N, const int32_t* __restrict__ deg, int32_t* __restrict__ ptr)
        {
            __shared__ int64_t part[kBlock];
            const int t = threadIdx.x;
            const int64_t per = (int64_t(N) + kBlock - 1) / kBlock;
            const int64_t b = std::min<int64_t>(int64_t(t) * per, N), e = std::min<int64_t>(b + per, N);
            int64_t s = 0;
            for (int64_t i = b; i < e; ++i) s += deg[i];
            part[t] = s;
            __syncthreads();
            if (t == 0)
            {
                int64_t run = 0;
                for (int k = 0; k < kBlock; ++k)
                {
                    const int64_t v = part[k];
                    part[k] = run;
                    run += v;
                }
                ptr[0] = 0;
            }
            __syncthreads();
            int64_t run = part[t];
            for (int64_t i = b; i < e; ++i)
            {
                run += deg[i];
                ptr[i + 1] = int32_t(run);
            }
        }

        __global__ void k_node_fill(int64_t n4, const int32_t* __restrict__ intma, const int32_t* __restrict__ ptr,
                                    int32_t* __restrict__ fill, int32_t* __restrict__ lst)
        {
            const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
            if (i >= n4) return;
            const int32_t node = intma[i] - 1;
            lst[ptr[node] + atomicAdd(&fill[node], 1)] = int32_t(i >> 2);
        }

        __global__ void k_node_sort(int N, const int32_t* __restrict__ ptr, int32_t* __restrict__ lst)
        {
            const int i = blockIdx.x * kBlock + threadIdx.x;
            if (i >= N) return;
            const int b = ptr[i], e = ptr[i + 1];
            for (int k = b + 1; k < e; ++k)
            {
                const int32_t v = lst[k];
                int j = k - 1;
                while (j >= b && lst[j] > v)
                {
                    lst[j + 1] = lst[j];
                    --j;
                }
                lst[j + 1] = v;
            }
        }

        // first position in the sorted list [b, e) holding a value >= v
        __device__ inline int lower_bound(const int32_t* __restrict__ lst, int b, int e, int32_t v)
        {
            while (b < e)
            {
                const int m = (b + e) >> 1;
                if (lst[m] < v) b = m + 1;
                else e = m;
            }
            return b;
        }

        __global__ void k_pending(int E, const int32_t* __restrict__ intma, const int32_t* __restrict__ ptr,
                                  const int32_t* __restrict__ lst, int32_t* __restrict__ pending,
                                  int32_t* __restrict__ level0, int32_t* __restrict__ cnt0)
        {
            const int e = blockIdx.x * kBlock + threadIdx.x;
            if (e >= E) return;
            int p = 0;
            for (int a = 0; a < 4; ++a)
            {
                const int node = intma[4 * int64_t(e) + a] - 1;
                p += lower_bound(lst, ptr[node], ptr[node + 1], e) - ptr[node];
            }
            pending[e] = p;
            if (p == 0) level0[atomicAdd(cnt0, 1)] = e;
        }

        // Level `level`: frontier buffer level % 3 (its count in cnt[level % 3])
        // is coloured; the next level is appended to buffer (level + 1) % 3;
        // buffer (level + 2) % 3, read by the previous level, is reset.
        __global__ __launch_bounds__(kBlock) void k_colour_level(int level, int E, const int32_t* __restrict__ intma,
                                                                 const int32_t* __restrict__ ptr,
                                                                 const int32_t* __restrict__ lst,
                                                                 int32_t* __restrict__ colour,
                                                                 int32_t* __restrict__ pending,
                                                                 int32_t* __restrict__ frontier,
                                                                 int32_t* __restrict__ cnt,
                                                                 unsigned long long* __restrict__ done,
                                                                 int* __restrict__ last_level,
                                                                 int* __restrict__ err)
        {
            const int cur = level % 3, nxt = (level + 1) % 3, old = (level + 2) % 3;
            const int n = cnt[cur];
            if (blockIdx.x == 0 && threadIdx.x == 0)
            {
                cnt[old] = 0;
                if (n > 0)
                {
                    atomicAdd(done, static_cast<unsigned long long>(n));
                    atomicMax(last_level, level);
                }
            }
            const int32_t* fr = frontier + int64_t(cur) * E;
            int32_t* fn = frontier + int64_t(nxt) * E;
            const int stride = gridDim.x * kBlock;
            for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += stride)
            {
                const int32_t e = fr[i];
                unsigned long long used[kMaxColours / 64] = {0ull, 0ull, 0ull, 0ull};
                int lo[4];
                for (int a = 0; a < 4; ++a)
                {
                    const int node = intma[4 * int64_t(e) + a] - 1;
                    const int b = ptr[node];
                    int k = b;
                    for (; k < ptr[node + 1]; ++k)
                    {
                        const int32_t f = lst[k];
                        if (f >= e) break;  // the list ascends
                        const int32_t c = colour[f];
                        if (c < kMaxColours) used[c >> 6] |= 1ull << (c & 63);
                    }
                    lo[a] = k;  // lst[k] == e
                }
                int c = 0;
                while (c < kMaxColours && ((used[c >> 6] >> (c & 63)) & 1ull)) ++c;
                if (c == kMaxColours) atomicOr(err, kTooManyColours);
                colour[e] = c;
                for (int a = 0; a < 4; ++a)
                {
                    const int node = intma[4 * int64_t(e) + a] - 1;
                    for (int k = lo[a] + 1; k < ptr[node + 1]; ++k)
                    {
                        const int32_t g = lst[k];
                        if (g == e) continue;  // an element listing a node twice
                        if (atomicSub(&pending[g], 1) == 1) fn[atomicAdd(&cnt[nxt], 1)] = g;
                    }
                }
            }
        }

        // per chunk of kChunk elements: how many of each colour
        __global__ void k_chunk_count(int E, int nchunks, const int32_t* __restrict__ colour,
                                      int32_t* __restrict__ counts)
        {
            __shared__ int32_t h[kChunkThreads][kMaxColours];
            const int t = threadIdx.x, ch = blockIdx.x * kChunkThreads + t;
            for (int c = 0; c < kMaxColours; ++c) h[t][c] = 0;
            if (ch < nchunks)
            {
                const int b = ch * kChunk, e = std::min(b + kChunk, E);
                for (int i = b; i < e; ++i) ++h[t][colour[i]];
                for (int c = 0; c < kMaxColours; ++c) counts[int64_t(ch) * kMaxColours + c] = h[t][c];
            }
        }

        // counts -> each chunk's first position per colour; color_ptr
        __global__ void k_colour_offsets(int nchunks, int32_t* __restrict__ counts, int32_t* __restrict__ color_ptr)
        {
            __shared__ int32_t tot[kMaxColours];
            const int c = threadIdx.x;
            int32_t s = 0;
            for (int ch = 0; ch < nchunks; ++ch) s += counts[int64_t(ch) * kMaxColours + c];
            tot[c] = s;
            __syncthreads();
            if (c == 0)
            {
                int32_t run = 0;
                for (int k = 0; k < kMaxColours; ++k)
                {
                    const int32_t v = tot[k];
                    tot[k] = run;
                    color_ptr[k] = run;
                    run += v;
                }
                color_ptr[kMaxColours] = run;
            }
            __syncthreads();
            int32_t off = tot[c];
            for (int ch = 0; ch < nchunks; ++ch)
            {
                const int64_t at = int64_t(ch) * kMaxColours + c;
                const int32_t v = counts[at];
                counts[at] = off;
                off += v;
            }
        }

        __global__ void k_chunk_place(int E, int nchunks, const int32_t* __restrict__ colour,
                                      const int32_t* __restrict__ offsets, int32_t* __restrict__ color_elem)
        {
            __shared__ int32_t o[kChunkThreads][kMaxColours];
            const int t = threadIdx.x, ch = blockIdx.x * kChunkThreads + t;
            if (ch >= nchunks) return;
            for (int c = 0; c < kMaxColours; ++c) o[t][c] = offsets[int64_t(ch) * kMaxColours + c];
            const int b = ch * kChunk, e = std::min(b + kChunk, E);
            for (int i = b; i < e; ++i) color_elem[o[t][colour[i]]++] = i + 1;  // 1-based ids
        }

        template <typename T>
        struct DevBuf
        {
            T* p = nullptr;
            explicit DevBuf(size_t n)
            {
                hip_check(hipMalloc(reinterpret_cast<void**>(&p), sizeof(T) * std::max<size_t>(n, 1)), "hipMalloc");
            }
            ~DevBuf()
            {
                if (p) (void)hipFree(p);
            }
            DevBuf(const DevBuf&) = delete;
            DevBuf& operator=(const DevBuf&) = delete;
        };

        // Device memory the colouring works in, beyond its input (intma) and
        // outputs (colour pointers, colour-major element ids).
        struct ColourWork
        {
            int32_t *deg, *ptr, *lst, *pending, *colour, *frontier, *cnt, *counts;
            unsigned long long* done;
            int *err, *last_level;
        };

        constexpr int64_t kAlignInts = 64;  // 256-byte regions
        constexpr int64_t aligned_ints(int64_t n) { return (n + kAlignInts - 1) / kAlignInts * kAlignInts; }
        constexpr int64_t nchunks_of(int64_t E) { return (E + kChunk - 1) / kChunk; }

        // 32-bit words of a ColourWork for E elements and N nodes, each
        // region starting on a 256-byte boundary
        constexpr int64_t colour_work_ints(int64_t E, int64_t N)
        {
            return aligned_ints(N) + aligned_ints(N + 1) + aligned_ints(4 * E) + 2 * aligned_ints(E) +
                   aligned_ints(3 * E) + aligned_ints(3) + aligned_ints(nchunks_of(E) * kMaxColours) +
                   3 * kAlignInts;
        }

        ColourWork carve_colour_work(int32_t* base, int64_t E, int64_t N)
        {
            ColourWork w{};
            int32_t* at = base;
            auto take = [&](int64_t n) {
                int32_t* r = at;
                at += aligned_ints(n);
                return r;
            };
            w.deg = take(N);
            w.ptr = take(N + 1);
            w.lst = take(4 * E);
            w.pending = take(E);
            w.colour = take(E);
            w.frontier = take(3 * E);
            w.cnt = take(3);
            w.counts = take(nchunks_of(E) * kMaxColours);
            w.done = reinterpret_cast<unsigned long long*>(take(2));
            w.err = take(1);
            w.last_level = take(1);
            return w;
        }

        // Coloring::build on device data.  intma (4, E) 1-based node ids on
        // the device; writes color_ptr (kMaxColours + 1 entries, 0-based
        // offsets) and color_elem (E, 1-based ids, colour-major) on the
        // device, and host_ptr (kMaxColours + 1) on the host.  Clears every
        // part of w it reads before writing it, so w may hold anything.
        int colour_on_device(int E, int N, const int32_t* intma, int32_t* color_ptr, int32_t* color_elem,
                             const ColourWork& w, hipStream_t s, int32_t* ncolor, int32_t* levels, int32_t* host_ptr,
                             std::string& msg)
        {
            const int64_t n4 = 4 * int64_t(E);
            hip_check(hipMemsetAsync(w.deg, 0, sizeof(int32_t) * size_t(N), s), "memset");
            hip_check(hipMemsetAsync(w.cnt, 0, sizeof(int32_t) * 3, s), "memset");
            hip_check(hipMemsetAsync(w.done, 0, sizeof(unsigned long long), s), "memset");
            hip_check(hipMemsetAsync(w.err, 0, sizeof(int), s), "memset");
            hip_check(hipMemsetAsync(w.last_level, 0, sizeof(int), s), "memset");
            hipLaunchKernelGGL(k_node_degree, dim3(grid_for(n4)), dim3(kBlock), 0, s, n4, N, intma, w.deg, w.err);
            int h_err = 0;
            hip_check(hipMemcpyAsync(&h_err, w.err, sizeof(int), hipMemcpyDeviceToHost, s), "download");
            hip_check(hipStreamSynchronize(s), "colouring");
            if (h_err & kNodeOutOfRange)
            {
                msg = "Coloring::build - element connectivity node out of range";
                return CBS_ERR_BAD_ARGUMENT;
            }
            hipLaunchKernelGGL(k_node_scan, dim3(1), dim3(kBlock), 0, s, N, w.deg, w.ptr);
            hip_check(hipMemsetAsync(w.deg, 0, sizeof(int32_t) * size_t(N), s), "memset");
            hipLaunchKernelGGL(k_node_fill, dim3(grid_for(n4)), dim3(kBlock), 0, s, n4, intma, w.ptr, w.deg, w.lst);
            hipLaunchKernelGGL(k_node_sort, dim3(grid_for(N)), dim3(kBlock), 0, s, N, w.ptr, w.lst);
            hipLaunchKernelGGL(k_pending, dim3(grid_for(E)), dim3(kBlock), 0, s, E, intma, w.ptr, w.lst, w.pending,
                               w.frontier, w.cnt);
            // levels, in batches; every level colours at least one element
            // while any is left, so the loop ends within E levels
            unsigned long long h_done = 0, prev = 0;
            int level = 0;
            for (;;)
            {
                for (int k = 0; k < kLevelsPerBatch; ++k, ++level)
                    hipLaunchKernelGGL(k_colour_level, dim3(kLevelGrid), dim3(kBlock), 0, s, level, E, intma, w.ptr,
                                       w.lst, w.colour, w.pending, w.frontier, w.cnt, w.done, w.last_level, w.err);
                hip_check(hipMemcpyAsync(&h_done, w.done, sizeof h_done, hipMemcpyDeviceToHost, s), "download");
                hip_check(hipStreamSynchronize(s), "colouring levels");
                if (h_done >= static_cast<unsigned long long>(E)) break;
                if (h_done == prev)
                {
                    msg = "Coloring::build - device colouring made no progress";
                    return CBS_ERR_HIP;
                }
                prev = h_done;
            }
            int h_last = 0;
            hip_check(hipMemcpyAsync(&h_last, w.last_level, sizeof(int), hipMemcpyDeviceToHost, s), "download");
            hip_check(hipMemcpyAsync(&h_err, w.err, sizeof(int), hipMemcpyDeviceToHost, s), "download");
            hip_check(hipStreamSynchronize(s), "colouring levels");
            if (h_err & kTooManyColours)
            {
                msg = "Coloring::build - the device colouring supports at most 256 colours";
                return CBS_ERR_UNSUPPORTED;
            }
            *levels = h_last + 1;  // the last batch may end with empty levels
            const int nchunks = int(nchunks_of(E));
            const int cgrid = (nchunks + kChunkThreads - 1) / kChunkThreads;
            hipLaunchKernelGGL(k_chunk_count, dim3(cgrid), dim3(kChunkThreads), 0, s, E, nchunks, w.colour, w.counts);
            hipLaunchKernelGGL(k_colour_offsets, dim3(1), dim3(kMaxColours), 0, s, nchunks, w.counts, color_ptr);
            hipLaunchKernelGGL(k_chunk_place, dim3(cgrid), dim3(kChunkThreads), 0, s, E, nchunks, w.colour, w.counts,
                               color_elem);
            hip_check(hipMemcpyAsync(host_ptr, color_ptr, sizeof(int32_t) * (kMaxColours + 1), hipMemcpyDeviceToHost,
                                     s),
                      "download");
            hip_check(hipStreamSynchronize(s), "colour order");
            int nc = 0;
            for (int c = 0; c < kMaxColours; ++c)
                if (host_ptr[c + 1] > host_ptr[c]) nc = c + 1;
            *ncolor = nc;
            return CBS_OK;
        }

        // ---- the variant that allocates nothing ----
        // Its workspace is this file-scope device array.  In host code the
        // name `scratch` is only a stand-in: its device address comes from
        // hipGetSymbolAddress, once per device (each GPU holds its own copy).
        constexpr int64_t kNoAllocMaxElems = int64_t(1) << 20;
        constexpr int64_t kNoAllocMaxNodes = int64_t(1) << 19;
        constexpr int64_t kNoAllocScratchInts = colour_work_ints(kNoAllocMaxElems, kNoAllocMaxNodes);
        static __device__ __align__(256) int32_t scratch[kNoAllocScratchInts];

        struct NoAllocDevice
        {
            int32_t* scratch = nullptr;  // device address of `scratch` on this device
            hipStream_t stream = nullptr;
        };
        constexpr int kNoAllocMaxDevices = 64;
        NoAllocDevice g_noalloc[kNoAllocMaxDevices];
        std::mutex g_noalloc_init;

        // The current device's scratch address and stream, set up on its
        // first call; nullptr (and msg) on failure.
        const NoAllocDevice* noalloc_device(std::string& msg)
        {
            int dev = -1;
            const hipError_t rd = hipGetDevice(&dev);
            if (rd != hipSuccess || dev < 0 || dev >= kNoAllocMaxDevices)
            {
                msg = std::string("cbs_hip_coloring_noalloc - no usable current HIP device (hipGetDevice: ") +
                      hipGetErrorString(rd) + ", device " + std::to_string(dev) + ")";
                return nullptr;
            }
            std::lock_guard<std::mutex> lock(g_noalloc_init);
            NoAllocDevice& d = g_noalloc[dev];
            if (!d.scratch)
            {
                void* p = nullptr;
                const hipError_t rc = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
                if (rc != hipSuccess || !p)
                {
                    msg = std::string("cbs_hip_coloring_noalloc - hipGetSymbolAddress(scratch): ") +
                          hipGetErrorString(rc);
                    return nullptr;
                }
                hipStream_t s = nullptr;
                const hipError_t rs = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
                if (rs != hipSuccess)
                {
                    msg = std::string("cbs_hip_coloring_noalloc - hipStreamCreate: ") + hipGetErrorString(rs);
                    return nullptr;
                }
                d.stream = s;
                d.scratch = static_cast<int32_t*>(p);
            }
            return &d;
        }
    }  // namespace

    // Returns 0, or a cbs_status with the message in `msg`.
    int colour_elements(int32_t nelem, int32_t npoin, const int32_t* intma_h, int32_t* ncolor, int32_t* color_ptr_h,
                        int32_t* color_elem_h, int32_t* levels, std::string& msg)
    {
        const int E = nelem, N = npoin;
        const int64_t n4 = 4 * int64_t(E);
        hipStream_t s = nullptr;
        hip_check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
        struct StreamGuard
        {
            hipStream_t s;
            ~StreamGuard() { (void)hipStreamDestroy(s); }
        } sg{s};
        DevBuf<int32_t> work{static_cast<size_t>(colour_work_ints(E, N))};
        DevBuf<int32_t> intma{static_cast<size_t>(n4)}, cptr{kMaxColours + 1}, celem{static_cast<size_t>(E)};
        hip_check(hipMemcpyAsync(intma.p, intma_h, sizeof(int32_t) * size_t(n4), hipMemcpyHostToDevice, s),
                  "upload intma");
        std::vector<int32_t> hp(kMaxColours + 1);
        const int rc = colour_on_device(E, N, intma.p, cptr.p, celem.p, carve_colour_work(work.p, E, N), s, ncolor,
                                        levels, hp.data(), msg);
        if (rc != CBS_OK) return rc;
        hip_check(hipMemcpyAsync(color_elem_h, celem.p, sizeof(int32_t) * size_t(E), hipMemcpyDeviceToHost, s),
                  "download");
        hip_check(hipStreamSynchronize(s), "colour order");
        for (int c = 0; c <= *ncolor; ++c) color_ptr_h[c] = hp[size_t(c)];
        return CBS_OK;
    }

    int colour_elements_noalloc(int32_t nelem, int32_t npoin, const int32_t* intma_d, int32_t* ncolor,
                                int32_t* color_ptr_d, int32_t* color_elem_d, int32_t* levels, std::string& msg)
    {
        if (nelem > kNoAllocMaxElems || npoin > kNoAllocMaxNodes)
        {
            msg = "cbs_hip_coloring_noalloc - mesh exceeds the fixed workspace (at most " +
                  std::to_string(kNoAllocMaxElems) + " elements and " + std::to_string(kNoAllocMaxNodes) +
                  " nodes; use cbs_hip_coloring)";
            return CBS_ERR_BAD_ARGUMENT;
        }
        const NoAllocDevice* d = noalloc_device(msg);
        if (!d) return CBS_ERR_HIP;
        int32_t hp[kMaxColours + 1];
        return colour_on_device(nelem, npoin, intma_d, color_ptr_d, color_elem_d,
                                carve_colour_work(d->scratch, nelem, npoin), d->stream, ncolor, levels, hp, msg);
    }

    int64_t colour_noalloc_scratch_bytes() { return kNoAllocScratchInts * int64_t(sizeof(int32_t)); }
}  // namespace cbsdev
