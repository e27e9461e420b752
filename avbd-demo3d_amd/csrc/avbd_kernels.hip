// avbd_kernels.hip -- gfx950 kernels of the resident AVBD step.
//
// One step (Solver::stepCpuReference, solver.cpp:1036-1041) is a fixed
// sequence of launches with every count kept on the device:
//   broadphase  axis choice -> 4-pass LSD radix sort of interval starts ->
//               chunked sweep (count, scan, emit) -> constrained-pair hash ->
//               ordered creation of new manifolds            (solver.cpp:488-577, 131-248)
//   init        box-box narrowphase + contact merge + warm start, joint warm
//               start, stable compaction                      (manifold.cpp:100-174, joint.cpp:35-54)
//   bodies      per-body force lists in descending creation serial, body warm
//               start                                         (solver.cpp:1134-1169)
//   colouring   index-priority Jones-Plassmann fixpoint == sequential greedy (avbd_gpu_solver.cpp:1812-1865)
//   iterate     per colour one primal launch, then one dual launch (solver.cpp:1172-1273)
//   finish      BDF1 velocities                               (solver.cpp:1275-1298)
//
// Arithmetic follows the reference operation by operation (see avbd_math.hpp);
// the build uses -ffp-contract=off.
#include "avbd_device.hpp"
#include "avbd_math.hpp"
#include "avbd_collide.hpp"
#include "avbd_round.hpp"

#include <float.h>

namespace avbd
{

constexpr float kPenaltyMin = 1.0f;
constexpr float kPenaltyMax = 10000000000.0f;
constexpr float kMargin = 0.01f;
constexpr float kStickThresh = 0.00001f;

// ------------------------------------------------------------------------------------------
// wave / block primitives (wave64, blockDim multiple of 64, <= 1024)
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ int wave_inclusive_scan(int v)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1)
    {
        int t = __shfl_up(v, off, 64);
        if (lane >= off)
            v += t;
    }
    return v;
}

__device__ __forceinline__ int wave_sum(int v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        v += __shfl_xor(v, off, 64);
    return v;
}

// Block-wide sums / maximum handed to one same-address device atomic per block. Device atomics
// on one address from many CUs (and XCDs) serialize, about 10 ns each, so one per wave costs
// tens of microseconds on large grids. Every thread of the block must call these (no early
// return before them); blocks of at most 1024 threads.
__device__ __forceinline__ void block_add2(int a, int *da, int b, int *db)
{
    __shared__ int part[2][16];
    a = wave_sum(a);
    b = wave_sum(b);
    if ((threadIdx.x & 63) == 0)
    {
        part[0][threadIdx.x >> 6] = a;
        part[1][threadIdx.x >> 6] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0)
    {
        int sa = 0, sb = 0;
        for (int w = 0; w < (int)(blockDim.x + 63) / 64; ++w)
        {
            sa += part[0][w];
            sb += part[1][w];
        }
        if (sa && da)
            atomicAdd(da, sa);
        if (sb && db)
            atomicAdd(db, sb);
    }
    __syncthreads();  // part[] may be reused by a later call in the same kernel
}

__device__ __forceinline__ void block_add(int a, int *da) { block_add2(a, da, 0, nullptr); }

// Block maximum of v, applied with atomicMax when above `floor`.
__device__ __forceinline__ void block_max(int v, int *dst, int floor)
{
    __shared__ int part[16];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        v = max(v, __shfl_xor(v, off, 64));
    if ((threadIdx.x & 63) == 0)
        part[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0)
    {
        int m = part[0];
        for (int w = 1; w < (int)(blockDim.x + 63) / 64; ++w)
            m = max(m, part[w]);
        if (m > floor)
            atomicMax(dst, m);
    }
    __syncthreads();  // part[] may be reused by a later call in the same kernel
}

// Exclusive scan over the block; lds needs blockDim/64 ints.
__device__ __forceinline__ int block_exclusive_scan(int v, int *lds, int &total)
{
    int incl = wave_inclusive_scan(v);
    int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 63)
        lds[wave] = incl;
    __syncthreads();
    int off = 0, tot = 0;
    int nw = blockDim.x >> 6;
    for (int w = 0; w < nw; ++w)
    {
        int s = lds[w];
        off += w < wave ? s : 0;
        tot += s;
    }
    __syncthreads();
    total = tot;
    return off + incl - v;
}

// Grid-wide barrier for a launch whose blocks are all resident. A spin limit turns a would-be
// hang into an error flag; once set, every later barrier returns at once.
__device__ __forceinline__ void grid_sync(GridBarrier *gb, unsigned nblocks, int *timeout)
{
    __syncthreads();
    if (threadIdx.x == 0)
    {
        __threadfence();
        unsigned g = __hip_atomic_load(&gb->gen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
        unsigned arrived = atomicAdd(&gb->count, 1u);
        if (arrived == nblocks - 1)
        {
            atomicExch(&gb->count, 0u);
            __threadfence();
            atomicAdd(&gb->gen, 1u);
        }
        else
        {
            unsigned spins = 0;
            while (__hip_atomic_load(&gb->gen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == g)
            {
                __builtin_amdgcn_s_sleep(2);
                if (++spins > (1u << 22) || __hip_atomic_load(timeout, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
                {
                    atomicOr(timeout, 1);
                    break;
                }
            }
        }
        __threadfence();
    }
    __syncthreads();
}

// ------------------------------------------------------------------------------------------
// scans: one block for short arrays with a host-known length, single-pass look-back otherwise
// ------------------------------------------------------------------------------------------
// Exclusive scan of up to kScanSmallMax ints in one block of 1024 threads: tiles of 4096
// (4 consecutive elements per thread, coalesced), each scanned across the block and carried
// into the next. One launch instead of three for the small histogram scans.
constexpr int kScanSmallMax = 8192;

// Stores a scan's total clamped to cap; a clamped total sets `bit` in *overflow.
__device__ __forceinline__ void put_total(int *total, int v, int cap, int *overflow, int bit)
{
    if (v > cap)
    {
        v = cap;
        atomicOr(overflow, bit);
    }
    *total = v;
}

__global__ __launch_bounds__(1024) void k_scan_small(const int *in, int *out, int n, int *total, int cap,
                                                     int *overflow, int bit)
{
    __shared__ int lds[16];
    int carry = 0;
    for (int base = 0; base < n; base += 4 * 1024)
    {
        int i0 = base + 4 * (int)threadIdx.x;
        int v[4];
#pragma unroll
        for (int q = 0; q < 4; ++q)
            v[q] = i0 + q < n ? in[i0 + q] : 0;
        int tot;
        int run = carry + block_exclusive_scan((v[0] + v[1]) + (v[2] + v[3]), lds, tot);
#pragma unroll
        for (int q = 0; q < 4; ++q)
        {
            if (i0 + q < n)
                out[i0 + q] = run;
            run += v[q];
        }
        carry += tot;
    }
    if (threadIdx.x == 0 && total)
        put_total(total, carry, cap, overflow, bit);
}

// Single-pass exclusive scan with decoupled look-back, for counts that live on the device.
// A fixed grid of blocks claims tiles of kLbTile ints in increasing order from a device counter;
// each tile publishes its own sum at once (flag 1) and, after looking back over its
// predecessors until one carries an inclusive prefix (flag 2), its inclusive prefix. A tile waits
// only for tiles claimed before it by blocks that are running, and those publish their sums
// without waiting on anything, so the look-back always ends; a spin limit turns a missing
// status into the timeout flag. Status words: epoch (30 bits) | flag (2) | value (32). The epoch
// is read from ctl[2] at the start of a launch and advanced by the launch's last block, which
// also resets the tile and exit counters (ctl[0], ctl[1]), so no host-side clearing (and no
// memset) is needed between scans and the launch can be replayed from a graph.
constexpr int kLbThreads = 256;
constexpr int kLbItems = 8;
constexpr int kLbTile = kLbThreads * kLbItems;
constexpr int kLbBlocks = 128;
constexpr int kLbSpin = 1 << 22;

// The status word carries its value, so publishing and reading it need no ordering with other
// memory: relaxed device-scope atomics, which skip the cache write-back and invalidation that
// release and acquire semantics cost.
__device__ __forceinline__ void lb_publish(unsigned long long *w, unsigned epoch, unsigned flag, int value)
{
    unsigned long long v = ((unsigned long long)epoch << 34) | ((unsigned long long)flag << 32) | (unsigned)value;
    __hip_atomic_store(w, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ __launch_bounds__(kLbThreads) void k_scan_lookback(const int *in, int *out, const int *nptr, int nconst,
                                                              int *total, unsigned long long *status, unsigned *ctl,
                                                              int maxTiles, int *timeout, int cap, int *overflow,
                                                              int bit)
{
    __shared__ int lds[kLbThreads / 64];
    __shared__ int sTile, sPrefix;
    const int n = nptr ? *nptr : nconst;
    const int numTiles = min((max(n, 0) + kLbTile - 1) / kLbTile, maxTiles);
    const unsigned epoch = __hip_atomic_load(&ctl[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & 0x3FFFFFFFu;
    if (numTiles == 0 && blockIdx.x == 0 && threadIdx.x == 0 && total)
        put_total(total, 0, cap, overflow, bit);
    const bool vec = ((((uintptr_t)in) | ((uintptr_t)out)) & 15) == 0;
    for (;;)
    {
        if (threadIdx.x == 0)
            sTile = (int)atomicAdd(&ctl[0], 1u);
        __syncthreads();
        const int t = sTile;
        __syncthreads();
        if (t >= numTiles)
            break;
        const int base = t * kLbTile + (int)threadIdx.x * kLbItems;
        int v[kLbItems];
        if (vec && base + kLbItems <= n)
        {
            int4 a = *(const int4 *)(in + base), b = *(const int4 *)(in + base + 4);
            v[0] = a.x, v[1] = a.y, v[2] = a.z, v[3] = a.w, v[4] = b.x, v[5] = b.y, v[6] = b.z, v[7] = b.w;
        }
        else
        {
#pragma unroll
            for (int q = 0; q < kLbItems; ++q)
                v[q] = base + q < n ? in[base + q] : 0;
        }
        int sum = 0;
#pragma unroll
        for (int q = 0; q < kLbItems; ++q)
            sum += v[q];
        int tot;
        const int ex = block_exclusive_scan(sum, lds, tot);
        if (threadIdx.x < 64)
        {
            // wave 0 looks back over 64 predecessors per round: lane j reads tile p - j
            const int lane = threadIdx.x;
            int prefix = 0;
            if (t == 0)
            {
                if (lane == 0)
                    lb_publish(&status[0], epoch, 2u, tot);
            }
            else
            {
                if (lane == 0)
                    lb_publish(&status[t], epoch, 1u, tot);
                int p = t - 1, spins = 0;
                for (;;)
                {
                    const int q = p - lane;
                    unsigned fl = 2u;
                    int val = 0;
                    if (q >= 0)
                    {
                        unsigned long long w = __hip_atomic_load(&status[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        fl = (unsigned)(w >> 34) == epoch ? (unsigned)(w >> 32) & 3u : 0u;
                        val = (int)(unsigned)w;
                    }
                    const unsigned long long incl = __ballot(fl == 2u);
                    const unsigned long long wait = __ballot(fl == 0u);
                    const int first = incl ? __ffsll((long long)incl) - 1 : 63;   // nearest inclusive prefix
                    const unsigned long long upto = first == 63 ? ~0ull : ((2ull << first) - 1ull);
                    if (wait & upto)
                    {
                        __builtin_amdgcn_s_sleep(1);
                        if (++spins > kLbSpin || __hip_atomic_load(timeout, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
                        {
                            if (lane == 0)
                                atomicOr(timeout, 2);
                            break;
                        }
                        continue;
                    }
                    prefix += wave_sum(lane <= first ? val : 0);
                    if (incl)
                        break;
                    p -= 64;
                }
                if (lane == 0)
                    lb_publish(&status[t], epoch, 2u, prefix + tot);
            }
            if (lane == 0)
            {
                sPrefix = prefix;
                if (t == numTiles - 1 && total)
                    put_total(total, prefix + tot, cap, overflow, bit);
            }
        }
        __syncthreads();
        int run = sPrefix + ex;
        int o[kLbItems];
#pragma unroll
        for (int q = 0; q < kLbItems; ++q)
        {
            o[q] = run;
            run += v[q];
        }
        if (vec && base + kLbItems <= n)
        {
            *(int4 *)(out + base) = make_int4(o[0], o[1], o[2], o[3]);
            *(int4 *)(out + base + 4) = make_int4(o[4], o[5], o[6], o[7]);
        }
        else
        {
#pragma unroll
            for (int q = 0; q < kLbItems; ++q)
                if (base + q < n)
                    out[base + q] = o[q];
        }
    }
    // A block's exit count follows its last tile claim, whose returned value it waited for, so
    // the last block to exit resets the counters after every claim of this launch. The scanned
    // output is published to later kernels by the launch boundary.
    if (threadIdx.x == 0)
    {
        if (atomicAdd(&ctl[1], 1u) == gridDim.x - 1)
        {
            atomicExch(&ctl[0], 0u);
            atomicExch(&ctl[1], 0u);
            atomicAdd(&ctl[2], 1u);
        }
    }
}

// ------------------------------------------------------------------------------------------
// stable LSD radix sort, 8 bits per pass, wave-ballot ranking
// ------------------------------------------------------------------------------------------
// One block of kRadixThreads per tile of kRadixTile keys (two rounds of 1024).
constexpr int kRadixThreads = 1024;
static_assert(kRadixTile % kRadixThreads == 0, "whole rounds per tile");

__global__ __launch_bounds__(kRadixThreads) void k_radix_hist(const unsigned *keys, int n, int shift, int *hist,
                                                              int tiles)
{
    __shared__ int cnt[256];
    if (threadIdx.x < 256)
        cnt[threadIdx.x] = 0;
    __syncthreads();
    int base = blockIdx.x * kRadixTile;
    for (int r = 0; r < kRadixTile / kRadixThreads; ++r)
    {
        int i = base + r * kRadixThreads + threadIdx.x;
        if (i < n)
            atomicAdd(&cnt[(keys[i] >> shift) & 255u], 1);
    }
    __syncthreads();
    if (threadIdx.x < 256)
        hist[threadIdx.x * tiles + blockIdx.x] = cnt[threadIdx.x];
}

// offs: the exclusive scan of the digit-major histogram, or, with fused != 0, the histogram
// itself: each block then forms its own offsets (digit d's total over all tiles and its count in
// the tiles before this one, then a block scan over the digits), which saves the separate scan
// launch when the tiles are few. Keys are ranked per round by (wave, lane), i.e. in index order,
// so the sort is stable.
__global__ __launch_bounds__(kRadixThreads) void k_radix_scatter(const unsigned *keys, const int *vals, int n,
                                                                 int shift, const int *offs, int tiles,
                                                                 unsigned *keysOut, int *valsOut, int fused)
{
    constexpr int W = kRadixThreads / 64;
    __shared__ int base[256];
    __shared__ int waveCnt[W][256];
    if (fused)
    {
        __shared__ int lds[W];
        int tot = 0, pre = 0;
        if (threadIdx.x < 256)
        {
            const int *h = offs + threadIdx.x * tiles;
            for (int t = 0; t < tiles; ++t)
            {
                const int c = h[t];
                tot += c;
                pre += t < (int)blockIdx.x ? c : 0;
            }
        }
        int all;
        const int ex = block_exclusive_scan(tot, lds, all);
        if (threadIdx.x < 256)
            base[threadIdx.x] = ex + pre;
    }
    else if (threadIdx.x < 256)
        base[threadIdx.x] = offs[threadIdx.x * tiles + blockIdx.x];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const unsigned long long lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
    for (int r = 0; r < kRadixTile / kRadixThreads; ++r)
    {
        for (int t = threadIdx.x; t < W * 256; t += kRadixThreads)
            (&waveCnt[0][0])[t] = 0;
        __syncthreads();
        int i = blockIdx.x * kRadixTile + r * kRadixThreads + threadIdx.x;
        bool valid = i < n;
        unsigned k = valid ? keys[i] : 0u;
        int v = valid ? vals[i] : 0;
        int d = (int)((k >> shift) & 255u);
        unsigned long long peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b)
        {
            int bit = (d >> b) & 1;
            unsigned long long m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        int rank = __popcll(peers & lt);
        if (valid && rank == 0)
            waveCnt[wave][d] = __popcll(peers);
        __syncthreads();
        if (threadIdx.x < 256)
        {
            int dd = threadIdx.x;
            int run = base[dd];
            for (int w = 0; w < W; ++w)
            {
                int c = waveCnt[w][dd];
                waveCnt[w][dd] = run;
                run += c;
            }
            base[dd] = run;
        }
        __syncthreads();
        if (valid)
        {
            int dst = waveCnt[wave][d] + rank;
            keysOut[dst] = k;
            valsOut[dst] = v;
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------------------------------------
// small utilities
// ------------------------------------------------------------------------------------------
__global__ void k_fill_int(int *p, int n, int v)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        p[i] = v;
}


__global__ void k_fill_u64(unsigned long long *p, int n, unsigned long long v)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        p[i] = v;
}

// Per-step counter reset (a kernel instead of a memset so the step can be graph-captured).
// Start of a step: the per-step counters, and the colouring's grid barrier.
__global__ void k_step_begin(Counters *c, GridBarrier *gb)
{
    gb->count = 0u;
    c->numChunks = 0;
    c->numHits = 0;
    c->numCreated = 0;
    c->pairChecks = 0;
    c->axisAmbiguous = 0;
    c->colorCount = 0;
    c->coloringRounds = 0;
    c->solvedCount = 0;
    c->contactCount = 0;
    c->contactEndpoints = 0;
    c->jointLive = 0;
    c->jointEndpoints = 0;
    c->maxStampCount = 0;
    c->npRedoCount = 0;
    c->colorChanged[0] = c->colorChanged[1] = c->colorChanged[2] = 0u;
}

// ------------------------------------------------------------------------------------------
// sweep axis: the reference picks the axis of largest centre variance from serial float sums
// (solver.cpp:499-515). The decision is taken in double; only when two variances lie within
// the float sums' error bound of each other is the exact serial float pass run.
// ------------------------------------------------------------------------------------------
// Sweep axis from double sums s = (sum x, sum y, sum z, sum x^2, sum y^2, sum z^2) of n centres:
// the axis of largest variance, and whether the reference's own float computation could pick
// another (solver.cpp:499-515: a serial float sum for the mean, then a serial float sum of the
// squared deviations, first maximum wins). Per axis, the float variance lies within
//   [V (1 - t) - e, (V + n d^2)(1 + t) + e]
// of the true one V: t = 4 (n + 16) 2^-24 bounds the rounding of the n-term serial sums (four
// times the textbook bound); d bounds the float mean's error, (n + 1) 2^-24 sum|c| / n with
// sum|c| <= sqrt(n sum c^2), and shifting every deviation by d adds n d^2 to their squares'
// sum (which matters for a world far from the origin, where sum c^2 >> V); e covers this double
// estimate's own rounding. Overlapping intervals make the decision ambiguous: the caller then
// runs the reference's float computation serially.
__device__ __forceinline__ int axis_decide(const double s[6], int n, bool &amb)
{
    const double u = 5.960464477539063e-08, ud = 1.1102230246251565e-16;
    const double tol = 4.0 * ((double)n + 16.0) * u + 1e-9;
    const double g = ((double)n + 1.0) * u * 1.01;
    double v[3], lo[3], hi[3];
    for (int k = 0; k < 3; ++k)
    {
        const double mean = s[k] / n;
        v[k] = s[k + 3] - n * mean * mean;
        const double e = ((double)n + 8.0) * 2.0 * ud * s[k + 3];
        lo[k] = (v[k] - e) * (1.0 - tol);
        hi[k] = (v[k] + e + g * g * s[k + 3]) * (1.0 + tol);
    }
    int best = 0;
    if (v[1] > v[best])
        best = 1;
    if (v[2] > v[best])
        best = 2;
    amb = false;
    for (int k = 0; k < 3; ++k)
        if (k != best && hi[k] >= lo[best])
            amb = true;
    return best;
}

__global__ void k_axis_partial(const float4 *pose, int n, double *partials)
{
    __shared__ double red[6][256];
    double a[6] = {0, 0, 0, 0, 0, 0};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    {
        float4 p = pose[4 * i + P_POS];
        a[0] += p.x;
        a[1] += p.y;
        a[2] += p.z;
        a[3] += (double)p.x * p.x;
        a[4] += (double)p.y * p.y;
        a[5] += (double)p.z * p.z;
    }
    for (int k = 0; k < 6; ++k)
        red[k][threadIdx.x] = a[k];
    __syncthreads();
    for (int st = blockDim.x / 2; st > 0; st >>= 1)
    {
        if ((int)threadIdx.x < st)
            for (int k = 0; k < 6; ++k)
                red[k][threadIdx.x] += red[k][threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x < 6)
        partials[blockIdx.x * 6 + threadIdx.x] = red[threadIdx.x][0];
}

// Decision on the summed statistics (axis_decide); flags the exact serial pass when the
// reference's float computation could pick another axis.
// One wave: lane b loads block b's six partial sums and the wave adds them by butterfly (any
// order of the double sums is within axis_decide's bound on this estimate's rounding).
__global__ void k_axis_decide(const double *partials, int nparts, int n, Counters *ctr)
{
    const int lane = threadIdx.x;
    double t[6];
#pragma unroll
    for (int k = 0; k < 6; ++k)
        t[k] = 0.0;
    for (int b = lane; b < nparts; b += 64)
#pragma unroll
        for (int k = 0; k < 6; ++k)
            t[k] += partials[b * 6 + k];
#pragma unroll
    for (int k = 0; k < 6; ++k)
        for (int off = 32; off > 0; off >>= 1)
            t[k] += __shfl_xor(t[k], off, 64);
    if (lane != 0)
        return;
    bool amb = false;
    const int best = axis_decide(t, n, amb);
    ctr->sweepAxis = best;
    ctr->axisAmbiguous = amb ? 1 : 0;
}

__global__ void k_axis_serial(const float4 *pose, int n, Counters *ctr)
{
    if (threadIdx.x != 0 || !ctr->axisAmbiguous)
        return;
    V3 mean{0.0f, 0.0f, 0.0f};
    for (int i = 0; i < n; ++i)
        mean += ld3(pose[4 * i + P_POS]);
    mean = mean / (float)n;
    V3 var{0.0f, 0.0f, 0.0f};
    for (int i = 0; i < n; ++i)
    {
        V3 d = ld3(pose[4 * i + P_POS]) - mean;
        var += V3{d.x * d.x, d.y * d.y, d.z * d.z};
    }
    int best = 0;
    if (var.y > comp(var, best))
        best = 1;
    if (var.z > comp(var, best))
        best = 2;
    ctr->sweepAxis = best;
    ctr->axisFallbacks += 1;
}

// Interval start keys: order-preserving float bits, -0 folded onto +0 so that ties fall to the
// body index exactly as the reference comparator does (solver.cpp:525-531).
__global__ void k_sap_keys(const float4 *pose, const float4 *sizeRad, int n, const Counters *ctr, unsigned *keys,
                           int *vals)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    int axis = ctr->sweepAxis;
    float c = comp(ld3(pose[4 * i + P_POS]), axis);
    float lo = c - sizeRad[i].w;
    if (lo == 0.0f)
        lo = 0.0f;
    unsigned u = __float_as_uint(lo);
    keys[i] = (u & 0x80000000u) ? ~u : (u | 0x80000000u);
    vals[i] = i;
}

__global__ void k_sap_gather(const int *sortedIdx, const float4 *pose, const float4 *sizeRad, int n, const Counters *ctr,
                             float *sLo, float *sHi, float4 *sPos, int *sIdx)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n)
        return;
    int i = sortedIdx[k];
    float4 p = pose[4 * i + P_POS];
    float r = sizeRad[i].w;
    float c = comp(ld3(p), ctr->sweepAxis);
    sLo[k] = c - r;
    sHi[k] = c + r;
    sPos[k] = make_float4(p.x, p.y, p.z, r);
    sIdx[k] = i;
}

// Sweep extent of each sorted interval: the first later interval whose start exceeds this one's
// end stops the reference's sweep (solver.cpp:555-556); starts are sorted, so a binary search finds it.
__global__ void k_sweep_len(const float *sLo, const float *sHi, int n, int *sweepEnd, int *chunkCount, Counters *ctr)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int len = 0;
    if (i < n)
    {
        float hi = sHi[i];
        int lo_ = i + 1, hi_ = n;
        while (lo_ < hi_)
        {
            int mid = (lo_ + hi_) >> 1;
            if (sLo[mid] > hi)
                hi_ = mid;
            else
                lo_ = mid + 1;
        }
        sweepEnd[i] = lo_;
        len = lo_ - i - 1;
        chunkCount[i] = (len + kSweepChunk - 1) / kSweepChunk;
    }
    block_add(len, &ctr->pairChecks);
}

// Batched worlds (each an independent reference Solver, bodies stored world by world): one
// block per world picks the world's sweep axis (solver.cpp:499-515, decided as k_axis_decide /
// k_axis_serial do), sorts the world's interval starts by (start, list index) in LDS and finds
// each interval's sweep extent inside the world (solver.cpp:525-556). Sorted arrays are
// world-major, so the sweep stages after it never pair bodies of two worlds.
constexpr int kWorldSapMax = 2048;

__global__ __launch_bounds__(256) void k_world_sap(const float4 *pose, const float4 *sizeRad, const int *worldStart,
                                                   float *sLo, float *sHi, float4 *sPos, int *sIdx, int *sweepEnd,
                                                   int *chunkCount, Counters *ctr)
{
    __shared__ unsigned long long key[kWorldSapMax];
    __shared__ float lo_s[kWorldSapMax];
    __shared__ float hi_s[kWorldSapMax];
    __shared__ double red[6][256];
    __shared__ int axis_s;
    const int w = blockIdx.x, t = threadIdx.x;
    const int ws = worldStart[w], n = worldStart[w + 1] - ws;
    double a[6] = {0, 0, 0, 0, 0, 0};
    // the centres are also staged in LDS (x in lo_s, y in hi_s, z over key, all free until the
    // sort) so that an exact serial pass below reads LDS instead of making n dependent loads
    float *zs = (float *)key;
    for (int l = t; l < n; l += 256)
    {
        float4 p = pose[4 * (ws + l) + P_POS];
        lo_s[l] = p.x;
        hi_s[l] = p.y;
        zs[l] = p.z;
        a[0] += p.x;
        a[1] += p.y;
        a[2] += p.z;
        a[3] += (double)p.x * p.x;
        a[4] += (double)p.y * p.y;
        a[5] += (double)p.z * p.z;
    }
    for (int k = 0; k < 6; ++k)
        red[k][t] = a[k];
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1)
    {
        if (t < st)
            for (int k = 0; k < 6; ++k)
                red[k][t] += red[k][t + st];
        __syncthreads();
    }
    if (t == 0)
    {
        double sums[6];
        for (int k = 0; k < 6; ++k)
            sums[k] = red[k][0];
        bool amb = false;
        int best = axis_decide(sums, n, amb);
        if (amb && n > 0)
        {
            V3 mean{0.0f, 0.0f, 0.0f};
            for (int l = 0; l < n; ++l)
                mean += V3{lo_s[l], hi_s[l], zs[l]};
            mean = mean / (float)n;
            V3 var{0.0f, 0.0f, 0.0f};
            for (int l = 0; l < n; ++l)
            {
                V3 d = V3{lo_s[l], hi_s[l], zs[l]} - mean;
                var += V3{d.x * d.x, d.y * d.y, d.z * d.z};
            }
            best = 0;
            if (var.y > comp(var, best))
                best = 1;
            if (var.z > comp(var, best))
                best = 2;
            atomicAdd(&ctr->axisFallbacks, 1);
        }
        axis_s = best;
    }
    __syncthreads();
    const int axis = axis_s;
    int m = 1;
    while (m < n)
        m <<= 1;
    for (int l = t; l < m; l += 256)
    {
        unsigned long long k = ~0ull;
        if (l < n)
        {
            int i = ws + l;
            float lo = comp(ld3(pose[4 * i + P_POS]), axis) - sizeRad[i].w;
            if (lo == 0.0f)
                lo = 0.0f;
            unsigned u = __float_as_uint(lo);
            unsigned ok = (u & 0x80000000u) ? ~u : (u | 0x80000000u);
            k = ((unsigned long long)ok << 32) | (unsigned)l;
        }
        key[l] = k;
    }
    __syncthreads();
    // bitonic sort of the m (power of two) composite keys, ascending
    for (int size = 2; size <= m; size <<= 1)
        for (int stride = size >> 1; stride > 0; stride >>= 1)
        {
            for (int x = t; x < (m >> 1); x += 256)
            {
                int i0 = 2 * x - (x & (stride - 1)), i1 = i0 + stride;
                bool asc = (i0 & size) == 0;
                unsigned long long k0 = key[i0], k1 = key[i1];
                if ((k0 > k1) == asc)
                {
                    key[i0] = k1;
                    key[i1] = k0;
                }
            }
            __syncthreads();
        }
    for (int k = t; k < n; k += 256)
    {
        int i = ws + (int)(key[k] & 0xFFFFFFFFull);
        float4 p = pose[4 * i + P_POS];
        float r = sizeRad[i].w;
        float c = comp(ld3(p), axis);
        lo_s[k] = c - r;
        hi_s[k] = c + r;
        sLo[ws + k] = c - r;
        sHi[ws + k] = c + r;
        sPos[ws + k] = make_float4(p.x, p.y, p.z, r);
        sIdx[ws + k] = i;
    }
    __syncthreads();
    int checks = 0;
    for (int k = t; k < n; k += 256)
    {
        float hi = hi_s[k];
        int lo_ = k + 1, hi_ = n;
        while (lo_ < hi_)
        {
            int mid = (lo_ + hi_) >> 1;
            if (lo_s[mid] > hi)
                hi_ = mid;
            else
                lo_ = mid + 1;
        }
        sweepEnd[ws + k] = ws + lo_;
        int len = lo_ - k - 1;
        chunkCount[ws + k] = (len + kSweepChunk - 1) / kSweepChunk;
        checks += len;
    }
    block_add(checks, &ctr->pairChecks);
}

// Bounding-sphere test of the reference sweep (solver.cpp:558-561).
__device__ __forceinline__ bool sphere_hit(float4 a, float4 b)
{
    V3 dp = V3{a.x, a.y, a.z} - V3{b.x, b.y, b.z};
    float r = a.w + b.w;
    return !(dot(dp, dp) > r * r);
}

// The sweep (solver.cpp:525-577) by chunks of 64 consecutive candidates. k_chunk_owner writes each
// chunk's interval (one wave per interval, lanes over its chunks). The count and emission kernels
// then give each wave kSweepGroup consecutive chunks: lanes 0..kSweepGroup-1 load their chunk's
// interval data at once, and for each chunk lane l tests candidate l (coalesced loads), the
// chunk's mask being the wave's ballot.
constexpr int kSweepGroup = 16;

__global__ void k_chunk_owner(const int *chunkStart, const int *chunkCount, int n, const Counters *ctr, int *owner)
{
    const int T = ctr->numChunks;
    const int lane = threadIdx.x & 63, wpb = blockDim.x >> 6;
    for (int i = blockIdx.x * wpb + (threadIdx.x >> 6); i < n; i += gridDim.x * wpb)
    {
        const int c0 = chunkStart[i], nc = chunkCount[i];
        for (int c = lane; c < nc && c0 + c < T; c += 64)
            owner[c0 + c] = i;
    }
}

__device__ __forceinline__ float4 shfl4(float4 v, int src)
{
    return make_float4(__shfl(v.x, src, 64), __shfl(v.y, src, 64), __shfl(v.z, src, 64), __shfl(v.w, src, 64));
}

// Also empties the constrained-pair hash (hashSize entries) that k_hash_insert fills next.
__global__ void k_sweep_count_g(const int *owner, const int *chunkStart, const int *sweepEnd, const Counters *ctr,
                                const float4 *sPos, int *hitCount, unsigned long long *hitMask,
                                unsigned long long *hash, int hashSize)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < hashSize; i += gridDim.x * blockDim.x)
        hash[i] = ~0ull;
    const int T = ctr->numChunks;
    const int lane = threadIdx.x & 63, wpb = blockDim.x >> 6;
    const int waves = gridDim.x * wpb;
    for (int g = blockIdx.x * wpb + (threadIdx.x >> 6); g * kSweepGroup < T; g += waves)
    {
        const int t0 = g * kSweepGroup;
        int pj0 = 0, pend = 0;
        float4 pa = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (lane < kSweepGroup && t0 + lane < T)
        {
            const int i = owner[t0 + lane];
            pj0 = i + 1 + (t0 + lane - chunkStart[i]) * kSweepChunk;
            pend = sweepEnd[i];
            pa = sPos[i];
        }
        const int kn = min(kSweepGroup, T - t0);
        for (int k0 = 0; k0 < kn; k0 += 4)
        {
            // four chunks' candidate loads in flight together
            float4 b[4];
            bool in[4];
            float4 a[4];
#pragma unroll
            for (int q = 0; q < 4; ++q)
            {
                const int k = min(k0 + q, kn - 1);
                const int j = __shfl(pj0, k, 64) + lane;
                a[q] = shfl4(pa, k);
                in[q] = k0 + q < kn && j < __shfl(pend, k, 64);
                b[q] = in[q] ? sPos[j] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            }
#pragma unroll
            for (int q = 0; q < 4; ++q)
            {
                const unsigned long long m = __ballot(in[q] && sphere_hit(a[q], b[q]));
                if (lane == 0 && k0 + q < kn)
                {
                    hitMask[t0 + k0 + q] = m;
                    hitCount[t0 + k0 + q] = __popcll(m);
                }
            }
        }
    }
}

__device__ __forceinline__ unsigned long long pair_key(int a, int b)
{
    unsigned lo = (unsigned)min(a, b), hi = (unsigned)max(a, b);
    return ((unsigned long long)lo << 32) | hi;
}

__device__ __forceinline__ unsigned hash_slot(unsigned long long k, int mask)
{
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdull;
    k ^= k >> 33;
    k *= 0xc4ceb9fe1a85ec53ull;
    k ^= k >> 33;
    return (unsigned)k & (unsigned)mask;
}

constexpr unsigned long long kHashEmpty = ~0ull;

__device__ __forceinline__ void hash_insert(unsigned long long *h, int mask, unsigned long long key, int *overflow)
{
    unsigned s = hash_slot(key, mask);
    for (int probe = 0; probe <= mask; ++probe)
    {
        unsigned long long prev = atomicCAS(&h[s], kHashEmpty, key);
        if (prev == kHashEmpty || prev == key)
            return;
        s = (s + 1) & (unsigned)mask;
    }
    atomicOr(overflow, 16);
}

__device__ __forceinline__ bool hash_contains(const unsigned long long *h, int mask, unsigned long long key)
{
    unsigned s = hash_slot(key, mask);
    for (int probe = 0; probe <= mask; ++probe)
    {
        unsigned long long v = h[s];
        if (v == key)
            return true;
        if (v == kHashEmpty)
            return false;
        s = (s + 1) & (unsigned)mask;
    }
    return false;
}

// Pairs already joined by any force are skipped by the broadphase (Rigid::constrainedTo, rigid.cpp:105-116).
__global__ void k_hash_insert(const int2 *mab, const Counters *ctr, const Joint *joints, int nj, const Spring *springs,
                              int ns, const int2 *ignores, int ni, unsigned long long *h, int mask, Counters *ctrw)
{
    int nm = ctr->numManifolds;
    int total = nm + nj + ns + ni;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < total; t += gridDim.x * blockDim.x)
    {
        int a, b;
        if (t < nm)
        {
            int2 p = mab[t];
            a = p.x;
            b = p.y;
        }
        else if (t < nm + nj)
        {
            const Joint &j = joints[t - nm];
            if (j.dead || j.a < 0)
                continue;
            a = j.a;
            b = j.b;
        }
        else if (t < nm + nj + ns)
        {
            const Spring &s = springs[t - nm - nj];
            a = s.a;
            b = s.b;
        }
        else
        {
            int2 p = ignores[t - nm - nj - ns];
            a = p.x;
            b = p.y;
        }
        hash_insert(h, mask, pair_key(a, b), &ctrw->overflow);
    }
}

#include "avbd_external.hip"

// Emission by the same chunk groups: the lanes whose bit is set in a chunk's mask write their pair
// at the chunk's start plus the count of set bits below them (the order of a serial walk over the
// mask), each checking the constrained-pair hash on its own.
__global__ void k_sweep_emit_g(const int *owner, const int *chunkStart, Counters *ctr, const unsigned long long *hitMask,
                               const int *sIdx, const int *hitStart, const unsigned long long *h, int mask, int2 *hits,
                               int *hitNew, int maxPairs)
{
    const int T = ctr->numChunks;
    const int lane = threadIdx.x & 63, wpb = blockDim.x >> 6;
    const int waves = gridDim.x * wpb;
    const unsigned long long below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    for (int g = blockIdx.x * wpb + (threadIdx.x >> 6); g * kSweepGroup < T; g += waves)
    {
        const int t0 = g * kSweepGroup;
        unsigned long long pbits = 0ull;
        int pj0 = 0, pstart = 0, pia = 0;
        if (lane < kSweepGroup && t0 + lane < T)
        {
            pbits = hitMask[t0 + lane];
            if (pbits)
            {
                const int i = owner[t0 + lane];
                pj0 = i + 1 + (t0 + lane - chunkStart[i]) * kSweepChunk;
                pstart = hitStart[t0 + lane];
                pia = sIdx[i];
            }
        }
        const int kn = min(kSweepGroup, T - t0);
        for (int k = 0; k < kn; ++k)
        {
            // every shuffle runs with the whole wave active; only the stores diverge
            const unsigned long long bits = __shfl(pbits, k, 64);
            if (!bits)
                continue;
            const int start = __shfl(pstart, k, 64), ia = __shfl(pia, k, 64), j0 = __shfl(pj0, k, 64);
            if (!((bits >> lane) & 1ull))
                continue;
            const int out = start + __popcll(bits & below);
            if (out < maxPairs)
            {
                const int ib = sIdx[j0 + lane];
                hits[out] = make_int2(ia, ib);
                hitNew[out] = hash_contains(h, mask, pair_key(ia, ib)) ? 0 : 1;
            }
            else
                atomicOr(&ctr->overflow, 1);
        }
    }
}

// New manifolds are appended in sweep order with consecutive creation serials (solver.cpp:243-247).
__global__ void k_create(const int2 *hits, const int *hitNew, const int *hitRank, Counters *ctr, Manifolds M,
                         int maxManifolds, int maxPairs)
{
    int H = min(ctr->numHits, maxPairs);
    int base = ctr->numManifolds;
    long long serial0 = ctr->nextSerial;
    for (int h = blockIdx.x * blockDim.x + threadIdx.x; h < H; h += gridDim.x * blockDim.x)
    {
        if (!hitNew[h])
            continue;
        int m = base + hitRank[h];
        if (m >= maxManifolds)
        {
            atomicOr(&ctr->overflow, 2);
            continue;
        }
        M.ab[m] = hits[h];
        M.serial[m] = serial0 + hitRank[h];
        M.count[m] = 0;
    }
}

__global__ void k_create_finish(Counters *ctr, int maxManifolds, int maxPairs)
{
    if (ctr->numHits > maxPairs)
        ctr->overflow |= 1;
    int total = ctr->numManifolds + ctr->numCreated;
    if (total > maxManifolds)
    {
        ctr->overflow |= 2;
        total = maxManifolds;
    }
    ctr->numTotal = total;
    ctr->nextSerial += ctr->numCreated;
}

// ------------------------------------------------------------------------------------------
// box-box narrowphase (collide.cpp:60-433, 898-955)
// ------------------------------------------------------------------------------------------
// (collide_boxes<T> in avbd_collide.hpp)

// Manifold::initialize (manifold.cpp:100-174). Reads the persistent buffer `cur` and writes the
// initialised manifold to the same slot of `tmp`. Box pairs take the box-box path; a pair with a
// sphere, capsule or cylinder goes through the shape dispatch, is frictionless when a sphere is
// in it, and re-takes its anchors and friction rows at every match.
// matchOld / stickUsed (adjoint recording, may be null): per new contact slot, the old contact it
// continues (-1: none) and whether its anchors were kept by the stick rule.
// Round: the world holds round bodies (worlds of boxes only run k_narrowphase_boxes below).
// list / listCount (may be null): process only the listed manifolds.
// X: the external contact table (none by default); a manifold with an entry takes the entry's
// basis and contacts instead of colliding (manifold.cpp:110-121): its unmatched contacts keep the
// entry's penalty, lambda and stick flag, where colliding starts them at zero.
template <bool Round>
__global__ void k_narrowphase(Manifolds cur, Manifolds tmp, Bodies B, Params P, Counters *ctr, int *keep,
                              int *matchOld, int *stickUsed, const int *list, const int *listCount,
                              ExtTable X = ExtTable())
{
    const int total = list ? *listCount : ctr->numTotal;
    const int old = ctr->numManifolds;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < total; k += gridDim.x * blockDim.x)
    {
        const int m = list ? list[k] : k;
        int2 ab = cur.ab[m];
        int nOld = m < old ? cur.count[m] : 0;
        V3 pA = ld3(B.pose[4 * (ab.x) + P_POS]), pB = ld3(B.pose[4 * (ab.y) + P_POS]);
        Q4 qA = ldq(B.pose[4 * (ab.x) + P_ROT]), qB = ldq(B.pose[4 * (ab.y) + P_ROT]);
        float4 sA4 = B.sizeRad[ab.x], sB4 = B.sizeRad[ab.y];
        float4 hA = B.shape[ab.x], hB = B.shape[ab.y];
        int shA = (int)hA.z, shB = (int)hB.z;
        float friction = sqrtf(B.friction[ab.x] * B.friction[ab.y]);
        if (shA == kShapeSphere || shB == kShapeSphere)
            friction = 0.0f;
        bool refresh = any_round(shA) || any_round(shB);
        NewContactsT<float> nc;
        V3T<float> t0{0, 0, 0}, t1{0, 0, 0}, t2{0, 0, 0};
        int nn;
        const avbd_external_contact *xe = ext_find(X, ab.x, ab.y);
        if (xe)
        {
            nn = min(xe->num_contacts, kMaxContacts);
            const float *bs = xe->basis;
            t0 = tv(V3{bs[0], bs[1], bs[2]});
            t1 = tv(V3{bs[3], bs[4], bs[5]});
            t2 = tv(V3{bs[6], bs[7], bs[8]});
            for (int i = 0; i < nn; ++i)
            {
                const avbd_contact &xc = xe->contacts[i];
                nc.rA[i] = tv(V3{xc.r_a[0], xc.r_a[1], xc.r_a[2]});
                nc.rB[i] = tv(V3{xc.r_b[0], xc.r_b[1], xc.r_b[2]});
                nc.feature[i] = xc.feature;
            }
        }
        else if (!Round || (shA == kShapeBox && shB == kShapeBox))
            nn = collide_boxes<float>(tv(pA), tq(qA), tv(ld3(sA4)), tv(pB), tq(qB), tv(ld3(sB4)), nc, t0, t1, t2);
        else
        {
            ShapeBody A{tv(pA), tv(ld3(sA4)), tv(ld3(B.vel[ab.x])), tq(qA), B.massMoment[ab.x].x, hA.x, hA.y, shA};
            ShapeBody Bb{tv(pB), tv(ld3(sB4)), tv(ld3(B.vel[ab.y])), tq(qB), B.massMoment[ab.y].x, hB.x, hB.y, shB};
            nn = collide_pair(A, Bb, P.dt, nc, t0, t1, t2);
        }
        V3 b0 = fv(t0), b1 = fv(t1), b2 = fv(t2);
        if (nn <= 0)
        {
            keep[m] = 0;
            continue;
        }
        keep[m] = 1;
        tmp.ab[m] = ab;
        tmp.serial[m] = cur.serial[m];
        tmp.count[m] = nn;
        for (int i = 0; i < nn; ++i)
        {
            V3 rA = fv(nc.rA[i]), rB = fv(nc.rB[i]);
            float4 pen = make_float4(0, 0, 0, 0), lam = make_float4(0, 0, 0, 0);
            if (xe)
            {
                const avbd_contact &xc = xe->contacts[i];
                pen = make_float4(xc.penalty[0], xc.penalty[1], xc.penalty[2], 0.0f);
                lam = make_float4(xc.lambda[0], xc.lambda[1], xc.lambda[2], xc.stick ? 1.0f : 0.0f);
            }
            int match = -1, kept = 0;
            for (int j = 0; j < nOld; ++j)
            {
                if (cur.feature[m * kMaxContacts + j] == nc.feature[i])
                {
                    const float4 *o = cur.cr + (size_t)(m * kMaxContacts + j) * kCR;
                    pen = o[CR_PEN];
                    lam = o[CR_LAM];
                    match = j;
                    if (!refresh && lam.w != 0.0f)
                    {
                        rA = ld3(o[CR_RA]);
                        rB = ld3(o[CR_RB]);
                        kept = 1;
                    }
                    if (refresh)
                    {
                        lam = make_float4(lam.x, 0.0f, 0.0f, 0.0f);
                        pen = make_float4(pen.x, kPenaltyMin, kPenaltyMin, pen.w);
                    }
                    break;
                }
            }
            if (matchOld)
            {
                matchOld[m * kMaxContacts + i] = match;
                stickUsed[m * kMaxContacts + i] = kept;
            }
            int code = Round ? shape_code(shA, shB, nc.feature[i]) : 0;
            V3 xA, xB;
            if (code == 0)
            {
                xA = pA + rotate(qA, rA);
                xB = pB + rotate(qB, rB);
            }
            else
            {
                bool cap = (code & kCodeCap) != 0;
                xA = pA + contact_offset(shA, hA.x, hA.y, cap, qA, rA, b0, true);
                xB = pB + contact_offset(shB, hB.x, hB.y, cap, qB, rB, b0, false);
            }
            V3 d = xA - xB;
            V3 c0 = V3{dot(b0, d), dot(b1, d), dot(b2, d)} + V3{kMargin, 0.0f, 0.0f};
            V3 l = V3{lam.x, lam.y, lam.z} * P.alpha * P.gamma;
            V3 k = V3{pen.x, pen.y, pen.z} * P.gamma;
            k = V3{clampr(k.x, kPenaltyMin, kPenaltyMax), clampr(k.y, kPenaltyMin, kPenaltyMax),
                   clampr(k.z, kPenaltyMin, kPenaltyMax)};
            int sl = m * kMaxContacts + i;
            tmp.feature[sl] = nc.feature[i];
            float4 *o = tmp.cr + (size_t)sl * kCR;
            // shape code and the two radii / half lengths ride in the free w slots
            o[CR_RA] = st3(rA, friction);
            o[CR_RB] = st3(rB, __int_as_float(code));
            o[CR_C0] = st3(c0, hB.y);
            o[CR_N] = st3(b0, hA.x);
            o[CR_T1] = st3(b1, hB.x);
            o[CR_T2] = st3(b2, hA.y);
            o[CR_PEN] = st3(k, 0.0f);
            o[CR_LAM] = st3(l, lam.w);
        }
    }
}

// ---- box-only narrowphase (worlds without round bodies) ----
// The same Manifold::initialize as k_narrowphase<false>, with no private arrays: each contact's
// record is written as collide_boxes finds it (only the merge midpoints stay, in registers), and
// the two clip polygons sit in LDS. Private arrays would go to scratch memory, whose latency and
// per-wave allocation throttled the kernel.
constexpr int kNpBoxThreads = 64;
// LDS vertices per clip polygon. A quad clipped by four planes keeps at most 8 vertices; a
// polygon that would grow past this (rounding near a plane can add vertices; the reference
// allows 16) raises the lane's overflow flag, and the manifold is redone by the full-capacity
// narrowphase.
constexpr int kNpPolyCap = 8;

// One lane's clip polygon in LDS, component-major with the lane index fastest.
struct LdsPoly
{
    float *p;
    int *ovf;  // the lane's overflow flag (LDS)
    int cap;   // vertices kept (at most kNpPolyCap)
    AVBD_HD V3T<float> get(int i) const
    {
        i = min(i, cap - 1);  // only read past cap after an overflow, whose result is discarded
        return {p[(3 * i) * kNpBoxThreads], p[(3 * i + 1) * kNpBoxThreads], p[(3 * i + 2) * kNpBoxThreads]};
    }
    AVBD_HD void set(int i, V3T<float> x)
    {
        if (i >= cap)
        {
            *ovf = 1;
            return;
        }
        p[(3 * i) * kNpBoxThreads] = x.x;
        p[(3 * i + 1) * kNpBoxThreads] = x.y;
        p[(3 * i + 2) * kNpBoxThreads] = x.z;
    }
};

// Contact sink of the box-only narrowphase: the manifold's context and the contacts' midpoints.
struct BoxRecordSink
{
    int count;
    V3T<float> mid[kMaxNewContacts];
    V3T<float> b0, b1, b2; // the manifold basis, set by collide_boxes before the first contact
    float *poly;           // this lane's two clip polygons in LDS
    int *ovf;              // this lane's polygon overflow flag in LDS
    int polyCap;           // vertices per polygon kept in LDS
    int m, nOld;
    bool refresh;
    float friction, alpha, gamma, hAx, hAy, hBx, hBy;
    V3 pA, pB;
    Q4 qA, qB;
    const int *curFeature;
    const float4 *curCr;
    int *tmpFeature;
    float4 *tmpCr;
    int *matchOld, *stickUsed;
};

AVBD_HD LdsPoly make_poly(BoxRecordSink &nc, int k)
{
    return LdsPoly{nc.poly + k * 3 * kNpPolyCap * kNpBoxThreads, nc.ovf, nc.polyCap};
}

// ContactBuilder::add (collide.cpp:30-58), then the contact's part of Manifold::initialize
// (manifold.cpp:100-174): warm start from the old contact with the same feature key, the stick
// rule, and the record. The merge test and the cap run over the midpoints as add_contact's do.
AVBD_HD void add_contact(BoxRecordSink &nc, V3T<float> posA, Q4T<float> rotA, V3T<float> posB, Q4T<float> rotB,
                         V3T<float> xA, V3T<float> xB, int key)
{
    V3T<float> mp = (xA + xB) * 0.5f;
    bool merged = false;
#pragma unroll
    for (int i = 0; i < kMaxNewContacts; ++i)
        if (i < nc.count && lensq(mp - nc.mid[i]) < kMergeDistSq)
            merged = true;
    if (merged || nc.count >= kMaxNewContacts)
        return;
    const int c = nc.count;
#pragma unroll
    for (int i = 0; i < kMaxNewContacts; ++i)
        if (i == c)
            nc.mid[i] = mp;
    nc.count = c + 1;
    V3 rA = fv(rotate(qconj(rotA), xA - posA)), rB = fv(rotate(qconj(rotB), xB - posB));
    float4 pen = make_float4(0, 0, 0, 0), lam = make_float4(0, 0, 0, 0);
    int match = -1, kept = 0;
    for (int j = 0; j < nc.nOld; ++j)
    {
        if (nc.curFeature[nc.m * kMaxContacts + j] == key)
        {
            const float4 *o = nc.curCr + (size_t)(nc.m * kMaxContacts + j) * kCR;
            pen = o[CR_PEN];
            lam = o[CR_LAM];
            match = j;
            if (!nc.refresh && lam.w != 0.0f)
            {
                rA = ld3(o[CR_RA]);
                rB = ld3(o[CR_RB]);
                kept = 1;
            }
            if (nc.refresh)
            {
                lam = make_float4(lam.x, 0.0f, 0.0f, 0.0f);
                pen = make_float4(pen.x, kPenaltyMin, kPenaltyMin, pen.w);
            }
            break;
        }
    }
    const int sl = nc.m * kMaxContacts + c;
    if (nc.matchOld)
    {
        nc.matchOld[sl] = match;
        nc.stickUsed[sl] = kept;
    }
    V3 b0 = fv(nc.b0), b1 = fv(nc.b1), b2 = fv(nc.b2);
    V3 wA = nc.pA + rotate(nc.qA, rA);
    V3 wB = nc.pB + rotate(nc.qB, rB);
    V3 d = wA - wB;
    V3 c0 = V3{dot(b0, d), dot(b1, d), dot(b2, d)} + V3{kMargin, 0.0f, 0.0f};
    V3 l = V3{lam.x, lam.y, lam.z} * nc.alpha * nc.gamma;
    V3 k = V3{pen.x, pen.y, pen.z} * nc.gamma;
    k = V3{clampr(k.x, kPenaltyMin, kPenaltyMax), clampr(k.y, kPenaltyMin, kPenaltyMax),
           clampr(k.z, kPenaltyMin, kPenaltyMax)};
    nc.tmpFeature[sl] = key;
    float4 *o = nc.tmpCr + (size_t)sl * kCR;
    o[CR_RA] = st3(rA, nc.friction);
    o[CR_RB] = st3(rB, __int_as_float(0));
    o[CR_C0] = st3(c0, nc.hBy);
    o[CR_N] = st3(b0, nc.hAx);
    o[CR_T1] = st3(b1, nc.hBx);
    o[CR_T2] = st3(b2, nc.hAy);
    o[CR_PEN] = st3(k, 0.0f);
    o[CR_LAM] = st3(l, lam.w);
}

// Launched with kNpBoxThreads threads per block, grid-stride over the manifolds. A manifold whose
// clip polygon overflowed is listed in redo (its partial records are rewritten there).
__global__ __launch_bounds__(kNpBoxThreads) void k_narrowphase_boxes(Manifolds cur, Manifolds tmp, Bodies B,
                                                                     Params P, Counters *ctr, int *keep,
                                                                     int *matchOld, int *stickUsed, int *redo,
                                                                     int polyCap)
{
    __shared__ float polyLds[2 * 3 * kNpPolyCap * kNpBoxThreads];
    __shared__ int polyOvf[kNpBoxThreads];
    const int total = ctr->numTotal;
    const int old = ctr->numManifolds;
    for (int m = blockIdx.x * kNpBoxThreads + threadIdx.x; m < total; m += gridDim.x * kNpBoxThreads)
    {
        int2 ab = cur.ab[m];
        BoxRecordSink nc;
        nc.count = 0;
        nc.poly = polyLds + threadIdx.x;
        nc.ovf = polyOvf + threadIdx.x;
        nc.polyCap = polyCap;
        polyOvf[threadIdx.x] = 0;
        nc.m = m;
        nc.nOld = m < old ? cur.count[m] : 0;
        nc.pA = ld3(B.pose[4 * (ab.x) + P_POS]);
        nc.pB = ld3(B.pose[4 * (ab.y) + P_POS]);
        nc.qA = ldq(B.pose[4 * (ab.x) + P_ROT]);
        nc.qB = ldq(B.pose[4 * (ab.y) + P_ROT]);
        float4 sA4 = B.sizeRad[ab.x], sB4 = B.sizeRad[ab.y];
        float4 hA = B.shape[ab.x], hB = B.shape[ab.y];
        int shA = (int)hA.z, shB = (int)hB.z;
        nc.friction = sqrtf(B.friction[ab.x] * B.friction[ab.y]);
        if (shA == kShapeSphere || shB == kShapeSphere)
            nc.friction = 0.0f;
        nc.refresh = any_round(shA) || any_round(shB);
        nc.alpha = P.alpha;
        nc.gamma = P.gamma;
        nc.hAx = hA.x;
        nc.hAy = hA.y;
        nc.hBx = hB.x;
        nc.hBy = hB.y;
        nc.curFeature = cur.feature;
        nc.curCr = cur.cr;
        nc.tmpFeature = tmp.feature;
        nc.tmpCr = tmp.cr;
        nc.matchOld = matchOld;
        nc.stickUsed = stickUsed;
        nc.b0 = nc.b1 = nc.b2 = V3T<float>{0, 0, 0};
        int nn = collide_boxes<float>(tv(nc.pA), tq(nc.qA), tv(ld3(sA4)), tv(nc.pB), tq(nc.qB), tv(ld3(sB4)), nc,
                                      nc.b0, nc.b1, nc.b2);
        if (polyOvf[threadIdx.x])
        {
            redo[atomicAdd(&ctr->npRedoCount, 1)] = m;
            continue;
        }
        if (nn == 0)
        {
            keep[m] = 0;
            continue;
        }
        keep[m] = 1;
        tmp.ab[m] = ab;
        tmp.serial[m] = cur.serial[m];
        tmp.count[m] = nn;
    }
}

// srcSlot (adjoint recording, may be null): the pre-compaction slot of each kept manifold.
// Eight lanes per manifold: lane q copies float4 q of each live contact record (and feature q),
// so a manifold's records move as whole 128-byte rows.
// Also clears the per-body force-list counts and cursors (deg, cursor: nBodies each), which the
// force lists of this step are built from after the compaction.
__global__ void k_compact_manifolds(Manifolds tmp, Manifolds cur, const int *keep, const int *keepRank,
                                    const float4 *massMoment, Counters *ctr, int *srcSlot, int *deg, int *cursor,
                                    int nBodies)
{
    static_assert(kCR == 8 && kMaxContacts == 8, "one lane per record float4 and per feature slot");
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nBodies; i += gridDim.x * blockDim.x)
    {
        deg[i] = 0;
        cursor[i] = 0;
    }
    const int total = ctr->numTotal;
    const int q = threadIdx.x & 7;
    int endpoints = 0;
    for (int m = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 3); m < total;
         m += (int)((gridDim.x * blockDim.x) >> 3))
    {
        if (!keep[m])
            continue;
        const int d = keepRank[m];
        const int n = tmp.count[m];
        if (q == 0)
        {
            int2 ab = tmp.ab[m];
            if (srcSlot)
                srcSlot[d] = m;
            endpoints += n * ((massMoment[ab.x].x > 0.0f ? 1 : 0) + (massMoment[ab.y].x > 0.0f ? 1 : 0));
            cur.ab[d] = ab;
            cur.serial[d] = tmp.serial[m];
            cur.count[d] = n;
        }
        if (q < n)
            cur.feature[d * kMaxContacts + q] = tmp.feature[m * kMaxContacts + q];
        for (int i = 0; i < n; ++i)
            cur.cr[(size_t)(d * kMaxContacts + i) * kCR + q] = tmp.cr[(size_t)(m * kMaxContacts + i) * kCR + q];
    }
    block_add(endpoints, &ctr->contactEndpoints);
}

// Joint::initialize (joint.cpp:35-54); a broken joint is removed (dead) here.
// One joint of k_joint_init; live / ends receive the joint's live count and solved endpoints.
__device__ __forceinline__ void joint_init_one(Joint *joints, int k, Bodies &B, const Params &P, int &live, int &ends)
{
    Joint j = joints[k];
    if (j.dead)
        return;
    V3 pb = rotate(ldq(B.pose[4 * (j.b) + P_ROT]), ld3(j.rB)) + ld3(B.pose[4 * (j.b) + P_POS]);
    V3 pa = j.a >= 0 ? rotate(ldq(B.pose[4 * (j.a) + P_ROT]), ld3(j.rA)) + ld3(B.pose[4 * (j.a) + P_POS]) : ld3(j.rA);
    V3 c0l = pa - pb;
    Q4 qa = j.a >= 0 ? ldq(B.pose[4 * (j.a) + P_ROT]) : Q4{0.0f, 0.0f, 0.0f, 1.0f};
    V3 c0a = qdiff(qa, ldq(B.pose[4 * (j.b) + P_ROT])) * j.c0Lin.w;
    V3 ll = ld3(j.lamLin) * P.alpha * P.gamma;
    V3 la = ld3(j.lamAng) * P.alpha * P.gamma;
    V3 pl = ld3(j.penLin) * P.gamma;
    V3 pa2 = ld3(j.penAng) * P.gamma;
    float kl = j.rA.w, ka = j.rB.w;
    pl = V3{clampr(pl.x, kPenaltyMin, kPenaltyMax), clampr(pl.y, kPenaltyMin, kPenaltyMax),
            clampr(pl.z, kPenaltyMin, kPenaltyMax)};
    pa2 = V3{clampr(pa2.x, kPenaltyMin, kPenaltyMax), clampr(pa2.y, kPenaltyMin, kPenaltyMax),
             clampr(pa2.z, kPenaltyMin, kPenaltyMax)};
    pl = V3{fminr(pl.x, kl), fminr(pl.y, kl), fminr(pl.z, kl)};
    pa2 = V3{fminr(pa2.x, ka), fminr(pa2.y, ka), fminr(pa2.z, ka)};
    joints[k].c0Lin = st3(c0l, j.c0Lin.w);
    joints[k].c0Ang = st3(c0a, j.c0Ang.w);
    joints[k].lamLin = st3(ll, 0.0f);
    joints[k].lamAng = st3(la, 0.0f);
    joints[k].penLin = st3(pl, 0.0f);
    joints[k].penAng = st3(pa2, 0.0f);
    if (j.broken)
        joints[k].dead = 1;
    else
    {
        int e = (j.a >= 0 && B.massMoment[j.a].x > 0.0f ? 1 : 0) + (B.massMoment[j.b].x > 0.0f ? 1 : 0);
        live = 1;
        ends = e;
    }
}

__global__ void k_joint_init(Joint *joints, int nj, Bodies B, Params P, Counters *ctr)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    int live = 0, ends = 0;
    if (k < nj)
        joint_init_one(joints, k, B, P, live, ends);
    block_add2(live, &ctr->jointLive, ends, &ctr->jointEndpoints);
}

// Algorithmic HBM bytes of this step from its own counts, with the reference's
// logical record sizes (body pose 28 B, contact 108 B read / 28 B dual write,
// joint 124 B, spring 40 B; per-endpoint reads of the partner poses):
//   iteration = Ns*(100+28) + sum_c [e_c*(108+56) + (108+2*56+28)]
//             + sum_j [e_j*(124+28) + (124+2*28+52)] + sum_s e_s*(40+28)
//   step      = iters*iteration + N*(148+92+16+64+8) + hits*(80+8) + 36*M + 130*contacts
__device__ void step_bytes(Counters *c, int nBodies, int iterations)
{
    double ns = c->solvedCount;
    double primal = ns * 128.0 + (double)c->contactEndpoints * 164.0 + (double)c->jointEndpoints * 152.0 +
                    (double)c->springEndpoints * 68.0;
    double dual = (double)c->contactCount * 248.0 + (double)c->jointLive * 232.0;
    double iter = primal + dual;
    double it = (double)iterations * iter;
    c->cumPrimalBytes += (double)iterations * primal;
    c->cumDualBytes += (double)iterations * dual;
    double step = it + (double)nBodies * 328.0 + (double)c->numHits * 88.0 + 36.0 * (double)c->numManifolds +
                  130.0 * (double)c->contactCount;
    c->lastIterBytes = it;
    c->lastStepBytes = step;
    c->cumIterBytes += it;
    c->cumStepBytes += step;
    c->cumSteps += 1;
}

// Byte accounting of the lazy-dual form, applied to the step's figures once it is chosen: in
// iterations 1.. the contact duals ride on the primal's loads and add only their penalty/lambda
// stores (32 B per contact point) to the primal, and one dual launch remains.
__global__ void k_lazy_bytes(Counters *c, int iterations)
{
    double moved = (double)(iterations - 1) * (double)c->contactCount;
    double saved = moved * (248.0 - 32.0);
    c->cumPrimalBytes += moved * 32.0;
    c->cumDualBytes -= moved * 248.0;
    c->lastIterBytes -= saved;
    c->lastStepBytes -= saved;
    c->cumIterBytes -= saved;
    c->cumStepBytes -= saved;
}

// ------------------------------------------------------------------------------------------
// per-body force lists, descending creation serial (the order force.cpp:25-36 leaves them in)
// ------------------------------------------------------------------------------------------
template <bool FILL>
__global__ void k_adj(Manifolds M, const Counters *ctr, const Joint *joints, int nj, const Spring *springs, int ns,
                      const float4 *massMoment, int *deg, const int *adjStart, int *cursor, int *adj,
                      long long *adjSerial, int *adjPartner, int maxAdj, int *overflow)
{
    int nm = ctr->numManifolds;
    int total = nm + nj + ns;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < total; t += gridDim.x * blockDim.x)
    {
        int a, b, kind, idx;
        long long serial;
        if (t < nm)
        {
            int2 p = M.ab[t];
            a = p.x;
            b = p.y;
            kind = kAdjManifold;
            idx = t;
            serial = M.serial[t];
        }
        else if (t < nm + nj)
        {
            const Joint &j = joints[t - nm];
            if (j.dead)
                continue;
            a = j.a;
            b = j.b;
            kind = kAdjJoint;
            idx = t - nm;
            serial = j.serial;
        }
        else
        {
            const Spring &s = springs[t - nm - nj];
            a = s.a;
            b = s.b;
            kind = kAdjSpring;
            idx = t - nm - nj;
            serial = s.serial;
        }
        int ends[2] = {a, b};
        for (int e = 0; e < 2; ++e)
        {
            int body = ends[e];
            if (body < 0 || !(massMoment[body].x > 0.0f))
                continue;
            if (!FILL)
                atomicAdd(&deg[body], 1);
            else
            {
                int slot = adjStart[body] + atomicAdd(&cursor[body], 1);
                if (slot < maxAdj)
                {
                    adj[slot] = adjEncode(kind, idx);
                    adjSerial[slot] = serial;
                    adjPartner[slot] = ends[1 - e];
                }
                else
                    atomicOr(overflow, 4);
            }
        }
    }
}

// One wave per body: its force list in descending creation serial. With at most 64 entries, lane l
// holds entry l and moves it to the number of entries with a larger serial (earlier entries first
// among equal serials, as a stable insertion sort leaves them); longer lists are insertion-sorted
// by lane 0.
__global__ void k_adj_sort(const int *deg, const int *adjStart, int *adj, long long *adjSerial, int *adjPartner, int n,
                           int maxAdj)
{
    const int lane = threadIdx.x & 63, wpb = blockDim.x >> 6;
    for (int i = blockIdx.x * wpb + (threadIdx.x >> 6); i < n; i += gridDim.x * wpb)
    {
        const int b = adjStart[i], e = min(b + deg[i], maxAdj);
        const int d = e - b;
        if (d <= 1)
            continue;
        if (d <= 64)
        {
            long long s = 0;
            int v = 0, pv = 0;
            if (lane < d)
            {
                s = adjSerial[b + lane];
                v = adj[b + lane];
                pv = adjPartner[b + lane];
            }
            int rank = 0;
            for (int m = 0; m < d; ++m)
            {
                const long long sm = __shfl(s, m, 64);
                rank += (sm > s || (sm == s && m < lane)) ? 1 : 0;
            }
            if (lane < d)
            {
                adjSerial[b + rank] = s;
                adj[b + rank] = v;
                adjPartner[b + rank] = pv;
            }
            continue;
        }
        if (lane == 0)
            for (int x = b + 1; x < e; ++x)
            {
                long long s = adjSerial[x];
                int v = adj[x], pv = adjPartner[x];
                int y = x - 1;
                while (y >= b && adjSerial[y] < s)
                {
                    adjSerial[y + 1] = adjSerial[y];
                    adj[y + 1] = adj[y];
                    adjPartner[y + 1] = adjPartner[y];
                    --y;
                }
                adjSerial[y + 1] = s;
                adj[y + 1] = v;
                adjPartner[y + 1] = pv;
            }
    }
}

// Body warm start (solver.cpp:1134-1169).
// ignoreDeg (may be null): per body, its IgnoreCollision forces; they make a body constrained.
// One body of k_warmstart; returns 1 when the body is solved this step.
__device__ __forceinline__ int warmstart_one(Bodies &B, const Params &P, const int *deg, const int *ignoreDeg, int i,
                                             int *color)
{
    float mass = B.massMoment[i].x;
    V3 pos = ld3(B.pose[4 * (i) + P_POS]), vel = ld3(B.vel[i]), w = ld3(B.angVel[i]);
    Q4 rot = ldq(B.pose[4 * (i) + P_ROT]);
    float dt = P.dt, g = P.gravity;
    V3 iner = pos + vel * dt;
    if (mass > 0.0f)
        iner += V3{0.0f, 0.0f, g} * (dt * dt);
    Q4 inerA = qintegrate(rot, w * dt);
    V3 accel = (vel - ld3(B.prevVel[i])) / dt;
    float accelExt = accel.z * signr(g);
    float aw = clampr(accelExt / fabsf(g), 0.0f, 1.0f);
    if (!isfinite(aw))
        aw = 0.0f;
    B.pose[4 * (i) + P_ILIN] = st3(pos, 0.0f);
    B.pose[4 * (i) + P_IANG] = stq(rot);
    B.inerLin[i] = st3(iner, 0.0f);
    B.inerAng[i] = stq(inerA);
    bool solved = mass > 0.0f && (deg[i] > 0 || (ignoreDeg && ignoreDeg[i] > 0));
    if (mass > 0.0f)
    {
        V3 np = pos + vel * dt + V3{0.0f, 0.0f, g} * (aw * dt * dt);
        Q4 nq = qintegrate(rot, w * dt);
        if (!solved)
        {
            np = iner;
            nq = inerA;
        }
        B.pose[4 * (i) + P_POS] = st3(np, 0.0f);
        B.pose[4 * (i) + P_ROT] = stq(nq);
    }
    if (!solved)
        color[i] = -1;
    return solved ? 1 : 0;
}

__global__ void k_warmstart(Bodies B, Params P, const int *deg, const int *ignoreDeg, int n, int *color, Counters *ctr)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int solved = i < n ? warmstart_one(B, P, deg, ignoreDeg, i, color) : 0;
    block_add(solved, &ctr->solvedCount);
}

// ------------------------------------------------------------------------------------------
// colouring: c(i) = mex{ c(j) : j < i adjacent, solved } is the sequential greedy colouring
// (avbd_gpu_solver.cpp:1812-1865); iterating it from any start converges to that unique fixpoint
// in at most (longest index-increasing path) rounds. Warm-started from the previous step's colours.
// ------------------------------------------------------------------------------------------
// One launch of a few large blocks (grid barrier between rounds). Block b evaluates the bodies
// b*blockDim + t (+ k*gridDim*blockDim): in a global round it re-evaluates its own bodies up to
// kColorLocalPasses times, with block barriers only, until none of them changes, so colour
// changes that chain through neighbouring indices settle inside the block instead of costing a
// grid-wide round each. The fixpoint is unique (each body's colour depends only on lower-index
// neighbours), so the order of evaluation does not change the result; the colouring ends after
// a global round in which no block changed any colour.
constexpr int kColorLocalPasses = 8;
constexpr int kColorCachedPartners = 16;

// Marks in `used` the colours (< 256) of body i's partners below i in adjPartner[b, e), reading
// them in batches of 8: all partner indices, then all their colours, are in flight together.
__device__ __forceinline__ void color_used_list(const int *adjPartner, int b, int e, int i, const int *color,
                                                unsigned long long used[4])
{
    for (int x0 = b; x0 < e; x0 += 8)
    {
        int o[8], c[8];
#pragma unroll
        for (int q = 0; q < 8; ++q)
            o[q] = x0 + q < e ? adjPartner[x0 + q] : -1;
#pragma unroll
        for (int q = 0; q < 8; ++q)
            c[q] = (o[q] >= 0 && o[q] < i) ? __hip_atomic_load(&color[o[q]], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                                           : -1;
#pragma unroll
        for (int q = 0; q < 8; ++q)
            if (c[q] >= 0 && c[q] < 256)
                used[c[q] >> 6] |= 1ull << (c[q] & 63);
    }
}

// Smallest colour not in `used` (256 when all 256 are taken).
__device__ __forceinline__ int color_mex(const unsigned long long used[4])
{
    int c = 0;
    for (int w = 0; w < 4; ++w)
    {
        if (~used[w])
        {
            c = w * 64 + __ffsll((long long)~used[w]) - 1;
            break;
        }
        c = 256;
    }
    return c;
}

// When every thread holds at most one body (n <= grid threads), the body's status, its own
// colour and its lower-index partners (up to kColorCachedPartners) are loaded once and kept in
// registers through all passes and rounds; a pass then costs one round trip for the partners'
// colours. Only the owning thread writes a body's colour, so the register copy stays exact.
__global__ void k_color(const int *deg, const int *ignoreDeg, const int *adjStart, const int *adjPartner,
                        const float4 *massMoment, int n, int *color, Counters *ctr, GridBarrier *gb, int maxRounds)
{
    const unsigned nblocks = gridDim.x;
    __shared__ int localChange, roundChange;
    const int i0 = blockIdx.x * blockDim.x + threadIdx.x;
    const bool oneEach = (long long)n <= (long long)nblocks * blockDim.x;
    bool cAct = false, cFit = false;
    int cOwn = 0, cb = 0, ce = 0;
    int cp[kColorCachedPartners];
#pragma unroll
    for (int q = 0; q < kColorCachedPartners; ++q)
        cp[q] = -1;
    if (oneEach && i0 < n)
    {
        cAct = (massMoment[i0].x > 0.0f) && !(deg[i0] == 0 && !(ignoreDeg && ignoreDeg[i0] > 0));
        if (cAct)
        {
            cb = adjStart[i0];
            ce = cb + deg[i0];
            cFit = ce - cb <= kColorCachedPartners;
#pragma unroll
            for (int q = 0; q < kColorCachedPartners; ++q)
            {
                const int o = (cFit && cb + q < ce) ? adjPartner[cb + q] : -1;
                cp[q] = (o >= 0 && o < i0) ? o : -1;
            }
            cOwn = color[i0];
        }
    }
    int rounds = 0;
    for (int r = 0; r < maxRounds; ++r)
    {
        if (threadIdx.x == 0)
        {
            roundChange = 0;
            if (blockIdx.x == 0)
                ctr->colorChanged[(r + 1) % 3] = 0u;
        }
        for (int pass = 0; pass < kColorLocalPasses; ++pass)
        {
            if (threadIdx.x == 0)
                localChange = 0;
            __syncthreads();
            int changed = 0;
            if (oneEach)
            {
                if (cAct)
                {
                    unsigned long long used[4] = {0ull, 0ull, 0ull, 0ull};
                    if (cFit)
                    {
                        int c[kColorCachedPartners];
#pragma unroll
                        for (int q = 0; q < kColorCachedPartners; ++q)
                            c[q] = cp[q] >= 0 ? __hip_atomic_load(&color[cp[q]], __ATOMIC_RELAXED,
                                                                  __HIP_MEMORY_SCOPE_AGENT)
                                              : -1;
#pragma unroll
                        for (int q = 0; q < kColorCachedPartners; ++q)
                            if (c[q] >= 0 && c[q] < 256)
                                used[c[q] >> 6] |= 1ull << (c[q] & 63);
                    }
                    else
                        color_used_list(adjPartner, cb, ce, i0, color, used);
                    const int c = color_mex(used);
                    if (c >= kMaxColors)
                        atomicOr(&ctr->overflow, 8);
                    if (cOwn != c)
                    {
                        __hip_atomic_store(&color[i0], c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        cOwn = c;
                        changed = 1;
                    }
                }
            }
            else
            {
                for (int i = i0; i < n; i += nblocks * blockDim.x)
                {
                    if (!(massMoment[i].x > 0.0f) || (deg[i] == 0 && !(ignoreDeg && ignoreDeg[i] > 0)))
                        continue;
                    unsigned long long used[4] = {0ull, 0ull, 0ull, 0ull};
                    color_used_list(adjPartner, adjStart[i], adjStart[i] + deg[i], i, color, used);
                    const int c = color_mex(used);
                    if (c >= kMaxColors)
                        atomicOr(&ctr->overflow, 8);
                    if (color[i] != c)
                    {
                        __hip_atomic_store(&color[i], c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        changed = 1;
                    }
                }
            }
            if (changed)
            {
                localChange = 1;
                roundChange = 1;
            }
            __syncthreads();
            const bool settled = localChange == 0;
            __syncthreads();
            if (settled)
                break;
        }
        if (threadIdx.x == 0 && roundChange)
            atomicOr(&ctr->colorChanged[r % 3], 1u);
        grid_sync(gb, nblocks, &ctr->barrierTimeout);
        rounds = r + 1;
        unsigned any = __hip_atomic_load(&ctr->colorChanged[r % 3], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (!any || ctr->barrierTimeout)
            break;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0)
    {
        ctr->coloringRounds = rounds;
        // still changing after the last allowed round: k_color_worlds finishes the colouring
        ctr->colorPending = (rounds == maxRounds && !ctr->barrierTimeout &&
                             __hip_atomic_load(&ctr->colorChanged[(rounds - 1) % 3], __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_AGENT) != 0u)
                                ? 1
                                : 0;
    }
}

// Batched worlds: each world (at most kWorldSapMax bodies, no force across worlds) is coloured by
// one workgroup of one wave. Its bodies' colours, force-list starts, degrees (-1 for a body that
// is not solved) and, when it fits, its partner slice are staged in LDS; then the bodies are
// visited in index order and each takes the smallest colour none of its lower-index partners
// holds (the lanes read the partners' colours, ballots find the smallest free colour). That is
// the sequential greedy colouring itself, the fixpoint k_color converges to, with no rounds: a
// colour change that runs along a long chain costs one visit per body instead of a grid round.
constexpr int kColorWorldParts = 4096;
// onlyIfPending: run only when k_color left the colouring unfinished (a small single world whose
// round limit was reached); the greedy rule does not depend on the colours it starts from.
__global__ __launch_bounds__(64) void k_color_worlds(const int *worldStart, const int *deg, const int *ignoreDeg,
                                                     const int *adjStart, const int *adjPartner,
                                                     const float4 *massMoment, int *color, Counters *ctr,
                                                     int onlyIfPending, int cap)
{
    if (onlyIfPending && !ctr->colorPending)
        return;
    // dynamic LDS sized for the batch's largest world (cap bodies): force-list starts (int), then
    // colours and degrees (short), then the partner slice as world-local indices (short)
    extern __shared__ int colorLds[];
    int *st = colorLds;
    short *col = (short *)(st + cap);
    short *dg = col + cap;
    short *part = dg + cap;
    const int lane = threadIdx.x;
    const int ws = worldStart[blockIdx.x], n = worldStart[blockIdx.x + 1] - ws;
    if (n > cap)
    {
        // cap is the batch's largest world, so this cannot happen; report it rather than skip
        if (lane == 0)
            atomicOr(&ctr->overflow, 128);
        return;
    }
    if (n <= 0)
        return;
    const int pb = adjStart[ws];
    for (int l = lane; l < n; l += 64)
    {
        const int i = ws + l;
        const bool solved = massMoment[i].x > 0.0f && (deg[i] > 0 || (ignoreDeg && ignoreDeg[i] > 0));
        col[l] = (short)color[i];
        st[l] = adjStart[i] - pb;
        dg[l] = (short)(solved ? min(deg[i], 32767) : -1);
    }
    const int np = adjStart[ws + n - 1] + deg[ws + n - 1] - pb;
    const bool fits = np <= kColorWorldParts;
    if (fits)
        for (int x = lane; x < np; x += 64)
        {
            const int o = adjPartner[pb + x];
            part[x] = (short)(o >= 0 ? o - ws : -1);
        }
    __syncthreads();
    for (int l = 0; l < n; ++l)
    {
        const int d = dg[l];
        if (d < 0)
            continue;
        const int i = ws + l, b = st[l];
        int c = 0;
        if (d <= 64)
        {
            int pc = -1;
            if (lane < d)
            {
                const int o = fits ? (part[b + lane] >= 0 ? ws + part[b + lane] : -1) : adjPartner[pb + b + lane];
                if (o >= 0 && o < i)
                    pc = col[o - ws];
            }
            // smallest colour in [0, 256) that no lane holds
            c = 256;
            for (int k = 0; k < 256; ++k)
                if (__ballot(pc == k) == 0ull)
                {
                    c = k;
                    break;
                }
        }
        else if (lane == 0)
        {
            unsigned long long used[4] = {0ull, 0ull, 0ull, 0ull};
            for (int x = 0; x < d; ++x)
            {
                const int o = fits ? (part[b + x] >= 0 ? ws + part[b + x] : -1) : adjPartner[pb + b + x];
                if (o < 0 || o >= i)
                    continue;
                const int pcx = col[o - ws];
                if (pcx >= 0 && pcx < 256)
                    used[pcx >> 6] |= 1ull << (pcx & 63);
            }
            c = 256;
            for (int w = 0; w < 4; ++w)
                if (~used[w])
                {
                    c = w * 64 + __ffsll((long long)~used[w]) - 1;
                    break;
                }
        }
        if (lane == 0)
        {
            if (c >= kMaxColors)
                atomicOr(&ctr->overflow, 8);
            col[l] = c;
        }
        __syncthreads();
    }
    for (int l = lane; l < n; l += 64)
        color[ws + l] = col[l];
    if (blockIdx.x == 0 && lane == 0 && !onlyIfPending)
        ctr->coloringRounds = 1;
}

// Sort keys of the colour buckets: colour in bits 8-15, and 255 - stamp count in bits 0-7, so
// that inside a colour the bodies with the most stamps come first (the order inside a colour
// does not change any result: its bodies share no force).
__global__ void k_color_keys(const int *color, const int *stampCount, int n, unsigned *keys, int *vals, Counters *ctr)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int mx = 0;  // colour count seen: max colour + 1, capped
    if (i < n)
    {
        int c = color[i];
        unsigned col = c < 0 ? 255u : (unsigned)min(c, kMaxColors - 1);
        keys[i] = (col << 8) | (unsigned)(255 - min(stampCount[i], 255));
        vals[i] = i;
        mx = c >= 0 ? min(c + 1, kMaxColors) : 0;
    }
    block_max(mx, &ctr->colorCount, 0);
}

// colorOffsets[0..256]: the colour buckets' starts (and n); colorOffsets[kSegBase + 3 c + q]: where
// the bodies of colour c with at most seg_stamps(q) stamps start (k_tasks writes the boundaries
// that exist; a colour with no such body keeps its end, set here).
// Also the step's algorithmic bytes (its counts are final once the warm start has run).
__global__ void k_color_offsets(const int *scannedHist, int tiles, int n, int *colorOffsets, Counters *c,
                                int nBodies, int iterations)
{
    int d = threadIdx.x;
    if (d == 0)
        step_bytes(c, nBodies, iterations);
    if (d < 256)
    {
        colorOffsets[d] = scannedHist[d * tiles];
        const int end = d + 1 < 256 ? scannedHist[(d + 1) * tiles] : n;
#pragma unroll
        for (int q = 0; q < kSegCount; ++q)
            colorOffsets[kSegBase + kSegCount * d + q] = end;
    }
    if (d == 0)
        colorOffsets[256] = n;
}

// ------------------------------------------------------------------------------------------
// primal / dual (solver.cpp:1177-1271)
//
// A body's primal update is its inertia term plus the Hessian/gradient stamps of its forces,
// taken in its force order, followed by a 6x6 LDL^T solve. The stamps of one body are
// independent of each other, so a group of kGroupLanes lanes computes them in parallel (one
// 27-float stamp record per lane, staged in LDS); then each lane owns one or two of the 27
// system entries and adds the records in their sequential order. Every entry therefore sees
// the same sequence of FP32 additions as the reference's single-threaded loop.
// ------------------------------------------------------------------------------------------
constexpr int kRec = 27; // lhsLin lower 6 | lhsAng lower 6 | lhsCross 9 | rhsLin 3 | rhsAng 3
enum
{
    R_L00 = 0, R_L10, R_L11, R_L20, R_L21, R_L22,
    R_A00, R_A10, R_A11, R_A20, R_A21, R_A22,
    R_C = 12, R_RL = 21, R_RA = 24
};

// Contact force with normal clamp and friction cone (manifold.cpp:203-216).
struct ContactEval
{
    float C0, C1, C2, F0, F1, F2, fs, bounds;
};

__device__ __forceinline__ ContactEval contact_eval(V3 c0, V3 pen, V3 lam, float friction, float oma, V3 b0, V3 b1,
                                                    V3 b2, V3 nb0, V3 nb1, V3 nb2, V3 ja0, V3 ja1, V3 ja2, V3 jb0,
                                                    V3 jb1, V3 jb2, V3 dqAL, V3 dqBL, V3 dqAA, V3 dqBA)
{
    ContactEval r;
    r.C0 = (((c0.x * oma + dot(b0, dqAL)) + dot(nb0, dqBL)) + dot(ja0, dqAA)) + dot(jb0, dqBA);
    r.C1 = (((c0.y * oma + dot(b1, dqAL)) + dot(nb1, dqBL)) + dot(ja1, dqAA)) + dot(jb1, dqBA);
    r.C2 = (((c0.z * oma + dot(b2, dqAL)) + dot(nb2, dqBL)) + dot(ja2, dqAA)) + dot(jb2, dqBA);
    float F0 = pen.x * r.C0 + lam.x;
    float F1 = pen.y * r.C1 + lam.y;
    float F2 = pen.z * r.C2 + lam.z;
    F0 = fminr(F0, 0.0f);
    float bounds = fabsf(F0) * friction;
    float fs = sqrtf(F1 * F1 + F2 * F2);
    if (fs > bounds && fs > 0.0f)
    {
        float q = bounds / fs;
        F1 *= q;
        F2 *= q;
    }
    r.F0 = F0;
    r.F1 = F1;
    r.F2 = F2;
    r.fs = fs;
    r.bounds = bounds;
    return r;
}

// Stamp record of one Jacobian block: rows J_L (l0..l2), J_A (g0..g2), K = diag(k), force F.
// lhsLin += (J_L^T K) J_L, lhsAng += (J_A^T K) J_A, lhsCross += (J_A^T K) J_L,
// rhsLin += J_L^T F, rhsAng += J_A^T F  (manifold.cpp:218-228)
__device__ __forceinline__ void rows_record(float *r, V3 l0, V3 l1, V3 l2, V3 g0, V3 g1, V3 g2, V3 k, float F0,
                                            float F1, float F2)
{
    float T[3][3] = {{l0.x * k.x, l1.x * k.y, l2.x * k.z}, {l0.y * k.x, l1.y * k.y, l2.y * k.z},
                     {l0.z * k.x, l1.z * k.y, l2.z * k.z}};
    float TA[3][3] = {{g0.x * k.x, g1.x * k.y, g2.x * k.z}, {g0.y * k.x, g1.y * k.y, g2.y * k.z},
                      {g0.z * k.x, g1.z * k.y, g2.z * k.z}};
    float L[3][3] = {{l0.x, l0.y, l0.z}, {l1.x, l1.y, l1.z}, {l2.x, l2.y, l2.z}};
    float G[3][3] = {{g0.x, g0.y, g0.z}, {g1.x, g1.y, g1.z}, {g2.x, g2.y, g2.z}};
#define AVBD_PROD(A, i, B, j) ((A[i][0] * B[0][j] + A[i][1] * B[1][j]) + A[i][2] * B[2][j])
    r[R_L00] = AVBD_PROD(T, 0, L, 0);
    r[R_L10] = AVBD_PROD(T, 1, L, 0);
    r[R_L11] = AVBD_PROD(T, 1, L, 1);
    r[R_L20] = AVBD_PROD(T, 2, L, 0);
    r[R_L21] = AVBD_PROD(T, 2, L, 1);
    r[R_L22] = AVBD_PROD(T, 2, L, 2);
    r[R_A00] = AVBD_PROD(TA, 0, G, 0);
    r[R_A10] = AVBD_PROD(TA, 1, G, 0);
    r[R_A11] = AVBD_PROD(TA, 1, G, 1);
    r[R_A20] = AVBD_PROD(TA, 2, G, 0);
    r[R_A21] = AVBD_PROD(TA, 2, G, 1);
    r[R_A22] = AVBD_PROD(TA, 2, G, 2);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            r[R_C + 3 * i + j] = AVBD_PROD(TA, i, L, j);
#undef AVBD_PROD
#pragma unroll
    for (int i = 0; i < 3; ++i)
    {
        r[R_RL + i] = (L[0][i] * F0 + L[1][i] * F1) + L[2][i] * F2;
        r[R_RA + i] = (G[0][i] * F0 + G[1][i] * F1) + G[2][i] * F2;
    }
}

// One contact of a manifold as seen by body i (Manifold::updatePrimal, manifold.cpp:176-229).
// Returns the contact's branch flags (kFlag*), which the adjoint's tape keeps.
constexpr int kFlagClamp0 = 1, kFlagScaled = 2, kFlagUpd0 = 4, kFlagUpd12 = 8, kFlagSat0 = 16;
__device__ __forceinline__ int contact_flags(const ContactEval &ev, V3 pen, V3 lam)
{
    float f0 = pen.x * ev.C0 + lam.x;
    return (!(f0 < 0.0f) ? kFlagClamp0 : 0) | ((ev.fs > ev.bounds && ev.fs > 0.0f) ? kFlagScaled : 0);
}

// Agent-scope (device-coherent across the XCDs' L2s) 16-byte load / store of a pose row, as two
// 64-bit relaxed atomics: the dataflow primal reads poses other workgroups wrote in the same launch.
__device__ __forceinline__ float4 ld_coherent(const float4 *p)
{
    unsigned long long *q = reinterpret_cast<unsigned long long *>(const_cast<float4 *>(p));
    unsigned long long a = __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    unsigned long long b = __hip_atomic_load(q + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    float4 r;
    r.x = __uint_as_float((unsigned)a);
    r.y = __uint_as_float((unsigned)(a >> 32));
    r.z = __uint_as_float((unsigned)b);
    r.w = __uint_as_float((unsigned)(b >> 32));
    return r;
}
__device__ __forceinline__ void st_coherent(float4 *p, float4 v)
{
    unsigned long long *q = reinterpret_cast<unsigned long long *>(p);
    unsigned long long a = (unsigned long long)__float_as_uint(v.x) | ((unsigned long long)__float_as_uint(v.y) << 32);
    unsigned long long b = (unsigned long long)__float_as_uint(v.z) | ((unsigned long long)__float_as_uint(v.w) << 32);
    __hip_atomic_store(q, a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(q + 1, b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// World offsets of a contact's two anchors (manifold.cpp:29-90). Box pairs carry shape code 0
// in CR_RB.w and rotate their local anchors; other pairs read the shapes, radii (CR_N.w, CR_T1.w)
// and half lengths (CR_T2.w, CR_C0.w) the narrowphase stored with the record.
__device__ __forceinline__ void contact_anchors(float4 rA4, float4 rB4, float4 c04, float4 bb0, float4 bb1, float4 bb2,
                                                Q4 qA, Q4 qB, V3 &rAW, V3 &rBW)
{
    int code = __float_as_int(rB4.w);
    if (code == 0)
    {
        rAW = rotate(qA, ld3(rA4));
        rBW = rotate(qB, ld3(rB4));
        return;
    }
    bool cap = (code & kCodeCap) != 0;
    V3 n = ld3(bb0);
    rAW = contact_offset(code & 3, bb0.w, bb2.w, cap, qA, ld3(rA4), n, true);
    rBW = contact_offset((code >> 2) & 3, bb1.w, c04.w, cap, qB, ld3(rB4), n, false);
}

// What one contact stamp reads: the contact record's rows and both bodies' pose records.
// Co: the current poses and the penalty / lambda row are read coherently (dataflow primal); the
// constant rows (x-) are ordinary loads.
struct ContactIn
{
    float4 rA4, rB4, c04, bb0, bb1, bb2, pen4, lam4, pA4, pB4, qA4, qB4, iA4, iB4, jA4, jB4;
};

template <bool Co = false>
__device__ __forceinline__ ContactIn contact_load(int i, int partner, bool isA, int sl, const Manifolds &M,
                                                  const Bodies &B)
{
    ContactIn in;
    const int2 ab = isA ? make_int2(i, partner) : make_int2(partner, i);
    const float4 *cr = M.cr + (size_t)sl * kCR;
    in.rA4 = cr[CR_RA];
    in.rB4 = cr[CR_RB];
    in.c04 = cr[CR_C0];
    in.bb0 = cr[CR_N];
    in.bb1 = cr[CR_T1];
    in.bb2 = cr[CR_T2];
    in.pen4 = Co ? ld_coherent(&cr[CR_PEN]) : cr[CR_PEN];
    in.lam4 = Co ? ld_coherent(&cr[CR_LAM]) : cr[CR_LAM];
    const float4 *pa = B.pose + 4 * ab.x, *pb = B.pose + 4 * ab.y;
    in.pA4 = Co ? ld_coherent(pa + P_POS) : pa[P_POS];
    in.pB4 = Co ? ld_coherent(pb + P_POS) : pb[P_POS];
    in.qA4 = Co ? ld_coherent(pa + P_ROT) : pa[P_ROT];
    in.qB4 = Co ? ld_coherent(pb + P_ROT) : pb[P_ROT];
    in.iA4 = B.pose[4 * ab.x + P_ILIN];
    in.iB4 = B.pose[4 * ab.y + P_ILIN];
    in.jA4 = B.pose[4 * ab.x + P_IANG];
    in.jB4 = B.pose[4 * ab.y + P_IANG];
    return in;
}

// One contact of a manifold as seen by one of its bodies, from its loaded rows `in`.
// dualP (lazy-dual form, null otherwise): this body owns the contact point's dual update, which
// is applied first (and stored to contact slot sl).
template <bool Co = false>
__device__ __forceinline__ int contact_record(float *r, const ContactIn &in, bool isA, int sl, const Manifolds &M,
                                              float oma, const Params *dualP = nullptr)
{
    float4 *cr = M.cr + (size_t)sl * kCR;
    const float4 rA4 = in.rA4, rB4 = in.rB4, c04 = in.c04, bb0 = in.bb0, bb1 = in.bb1, bb2 = in.bb2;
    float4 lam4 = in.lam4;
    const float4 pen4 = in.pen4;
    Q4 qA = ldq(in.qA4), qB = ldq(in.qB4);
    V3 dqAL = ld3(in.pA4) - ld3(in.iA4);
    V3 dqAA = qdiff(qA, ldq(in.jA4));
    V3 dqBL = ld3(in.pB4) - ld3(in.iB4);
    V3 dqBA = qdiff(qB, ldq(in.jB4));
    V3 b0 = ld3(bb0), b1 = ld3(bb1), b2 = ld3(bb2);
    V3 nb0 = -b0, nb1 = -b1, nb2 = -b2;
    V3 rAW, rBW;
    contact_anchors(rA4, rB4, c04, bb0, bb1, bb2, qA, qB, rAW, rBW);
    V3 ja0 = cross(rAW, b0), ja1 = cross(rAW, b1), ja2 = cross(rAW, b2);
    V3 jb0 = cross(rBW, nb0), jb1 = cross(rBW, nb1), jb2 = cross(rBW, nb2);
    V3 pen = ld3(pen4);
    ContactEval ev = contact_eval(ld3(c04), pen, ld3(lam4), rA4.w, oma, b0, b1, b2, nb0, nb1, nb2, ja0, ja1, ja2, jb0,
                                  jb1, jb2, dqAL, dqBL, dqAA, dqBA);
    if (dualP)
    {
        // the previous iteration's dual update of this contact point (Manifold::updateDual,
        // manifold.cpp:231-280), done here because neither body has moved since that iteration
        // ended; then the record from the updated penalty and lambda
        const Params &P = *dualP;
        float stick = lam4.w;
        if (ev.F0 < 0.0f)
            pen.x = fminr(pen.x + P.betaLin * fabsf(ev.C0), kPenaltyMax);
        if (ev.fs <= ev.bounds)
        {
            float k1 = pen.y + P.betaLin * fabsf(ev.C1), k2 = pen.z + P.betaLin * fabsf(ev.C2);
            pen.y = fminr(k1, kPenaltyMax);
            pen.z = fminr(k2, kPenaltyMax);
            stick = sqrtf(ev.C1 * ev.C1 + ev.C2 * ev.C2) < kStickThresh ? 1.0f : 0.0f;
        }
        lam4 = make_float4(ev.F0, ev.F1, ev.F2, stick);
        if (Co)
        {
            st_coherent(&cr[CR_LAM], lam4);
            st_coherent(&cr[CR_PEN], st3(pen, 0.0f));
        }
        else
        {
            cr[CR_LAM] = lam4;
            cr[CR_PEN] = st3(pen, 0.0f);
        }
        ev = contact_eval(ld3(c04), pen, ld3(lam4), rA4.w, oma, b0, b1, b2, nb0, nb1, nb2, ja0, ja1, ja2, jb0, jb1, jb2,
                          dqAL, dqBL, dqAA, dqBA);
    }
    if (isA)
        rows_record(r, b0, b1, b2, ja0, ja1, ja2, pen, ev.F0, ev.F1, ev.F2);
    else
        rows_record(r, nb0, nb1, nb2, jb0, jb1, jb2, pen, ev.F0, ev.F1, ev.F2);
    return contact_flags(ev, pen, ld3(lam4));
}

// Dense 3x3 helpers for the joint stamp (joint.cpp:56-123), kept literal.
struct D3
{
    float m[3][3];
};
__device__ __forceinline__ D3 d3_diag(float a, float b, float c)
{
    D3 r = {{{a, 0.0f, 0.0f}, {0.0f, b, 0.0f}, {0.0f, 0.0f, c}}};
    return r;
}
__device__ __forceinline__ D3 d3_mul(const D3 &a, const D3 &b)
{
    D3 r;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            r.m[i][j] = (a.m[i][0] * b.m[0][j] + a.m[i][1] * b.m[1][j]) + a.m[i][2] * b.m[2][j];
    return r;
}
__device__ __forceinline__ D3 d3_t(const D3 &a)
{
    D3 r;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            r.m[i][j] = a.m[j][i];
    return r;
}
__device__ __forceinline__ D3 d3_scale(const D3 &a, float s)
{
    D3 r;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            r.m[i][j] = a.m[i][j] * s;
    return r;
}
__device__ __forceinline__ V3 d3_mv(const D3 &a, V3 v)
{
    return V3{(a.m[0][0] * v.x + a.m[0][1] * v.y) + a.m[0][2] * v.z, (a.m[1][0] * v.x + a.m[1][1] * v.y) + a.m[1][2] * v.z,
              (a.m[2][0] * v.x + a.m[2][1] * v.y) + a.m[2][2] * v.z};
}
__device__ __forceinline__ D3 d3_skew(V3 r)
{
    D3 o = {{{0.0f, -r.z, r.y}, {r.z, 0.0f, -r.x}, {-r.y, r.x, 0.0f}}};
    return o;
}
// Entry (a, b) of the ball joint's geometric-stiffness block G_k(v) (joint.cpp:96-104): the
// diagonal -v_k is written first, then v is added to column k.
template <int k, int a, int b> __device__ __forceinline__ float gs_entry(V3 v)
{
    float d = a == b ? -comp(v, k) : 0.0f;
    return b == k ? d + comp(v, a) : d;
}
__device__ __forceinline__ void rec_lin(float *r, const D3 &a)
{
    r[R_L00] = a.m[0][0];
    r[R_L10] = a.m[1][0];
    r[R_L11] = a.m[1][1];
    r[R_L20] = a.m[2][0];
    r[R_L21] = a.m[2][1];
    r[R_L22] = a.m[2][2];
}
__device__ __forceinline__ void rec_ang(float *r, const D3 &a)
{
    r[R_A00] = a.m[0][0];
    r[R_A10] = a.m[1][0];
    r[R_A11] = a.m[1][1];
    r[R_A20] = a.m[2][0];
    r[R_A21] = a.m[2][1];
    r[R_A22] = a.m[2][2];
}

// A joint adds to lhsAng three times in sequence (linear row K term, its geometric-stiffness
// diagonal, angular row) and to rhsAng twice, so it is split into three ordered records:
// part 0 linear row, part 1 the diagonal of H, part 2 the angular row (joint.cpp:58-122).
__device__ __forceinline__ void joint_record(float *r, int i, const Joint &j, int part, const Bodies &B, float alpha)
{
    bool isA = j.a == i;
    Q4 qB = ldq(B.pose[4 * (j.b) + P_ROT]);
    V3 pB = ld3(B.pose[4 * (j.b) + P_POS]);
    V3 penL = ld3(j.penLin), penA = ld3(j.penAng);
    if (part < 2)
    {
        if (!(lensq(penL) > 0.0f))
            return;
        V3 rAW = j.a >= 0 ? rotate(ldq(B.pose[4 * (j.a) + P_ROT]), ld3(j.rA)) : V3{0.0f, 0.0f, 0.0f};
        V3 rBW = rotate(qB, ld3(j.rB));
        D3 K = d3_diag(penL.x, penL.y, penL.z);
        V3 C = (j.a >= 0 ? ld3(B.pose[4 * (j.a) + P_POS]) + rAW : ld3(j.rA)) - (pB + rBW);
        if (isinf(j.rA.w))
            C -= ld3(j.c0Lin) * alpha;
        V3 F = d3_mv(K, C) + ld3(j.lamLin);
        if (part == 0)
        {
            D3 jL = isA ? d3_diag(1.0f, 1.0f, 1.0f) : d3_diag(-1.0f, -1.0f, -1.0f);
            D3 jA = isA ? d3_skew(-rAW) : d3_skew(rBW);
            D3 jLT = d3_t(jL), jAT = d3_t(jA);
            D3 jATk = d3_mul(jAT, K);
            rec_lin(r, d3_mul(d3_mul(jLT, K), jL));
            rec_ang(r, d3_mul(jATk, jA));
            D3 cr = d3_mul(jATk, jL);
#pragma unroll
            for (int a = 0; a < 3; ++a)
#pragma unroll
                for (int b = 0; b < 3; ++b)
                    r[R_C + 3 * a + b] = cr.m[a][b];
            V3 rl = d3_mv(jLT, F), ra = d3_mv(jAT, F);
            r[R_RL] = rl.x;
            r[R_RL + 1] = rl.y;
            r[R_RL + 2] = rl.z;
            r[R_RA] = ra.x;
            r[R_RA + 1] = ra.y;
            r[R_RA + 2] = ra.z;
        }
        else
        {
            V3 rr = isA ? rAW : -rBW;
            // H = sum_k F_k G_k(rr), entry by entry (straight-line, so the record stays in registers)
#define AVBD_GS_H(a, b) ((gs_entry<0, a, b>(rr) * F.x + gs_entry<1, a, b>(rr) * F.y) + gs_entry<2, a, b>(rr) * F.z)
            V3 c0{AVBD_GS_H(0, 0), AVBD_GS_H(1, 0), AVBD_GS_H(2, 0)};
            V3 c1{AVBD_GS_H(0, 1), AVBD_GS_H(1, 1), AVBD_GS_H(2, 1)};
            V3 c2{AVBD_GS_H(0, 2), AVBD_GS_H(1, 2), AVBD_GS_H(2, 2)};
#undef AVBD_GS_H
            rec_ang(r, d3_diag(len(c0), len(c1), len(c2)));
        }
        return;
    }
    if (!(lensq(penA) > 0.0f))
        return;
    D3 K = d3_diag(penA.x, penA.y, penA.z);
    Q4 qa = j.a >= 0 ? ldq(B.pose[4 * (j.a) + P_ROT]) : Q4{0.0f, 0.0f, 0.0f, 1.0f};
    V3 C = qdiff(qa, qB) * j.c0Lin.w;
    if (isinf(j.rB.w))
        C -= ld3(j.c0Ang) * alpha;
    V3 F = d3_mv(K, C) + ld3(j.lamAng);
    D3 jA = d3_scale(isA ? d3_diag(1.0f, 1.0f, 1.0f) : d3_diag(-1.0f, -1.0f, -1.0f), j.c0Lin.w);
    D3 jAT = d3_t(jA);
    rec_ang(r, d3_mul(d3_mul(jAT, K), jA));
    V3 ra = d3_mv(jAT, F);
    r[R_RA] = ra.x;
    r[R_RA + 1] = ra.y;
    r[R_RA + 2] = ra.z;
}

// Spring::updatePrimal (spring.cpp:27-69)
__device__ __forceinline__ void spring_record(float *r, int i, const Spring &sp, const Bodies &B)
{
    Q4 qA = ldq(B.pose[4 * (sp.a) + P_ROT]), qB = ldq(B.pose[4 * (sp.b) + P_ROT]);
    V3 d = (rotate(qA, ld3(sp.rA)) + ld3(B.pose[4 * (sp.a) + P_POS])) - (rotate(qB, ld3(sp.rB)) + ld3(B.pose[4 * (sp.b) + P_POS]));
    float dl = len(d);
    if (dl <= 1.0e-6f)
        return;
    V3 n = d / dl;
    float k = sp.rB.w;
    float f = k * (dl - sp.rA.w);
    V3 jL, jA;
    if (i == sp.a)
    {
        V3 rW = rotate(qA, ld3(sp.rA));
        jL = n;
        jA = cross(rW, n);
    }
    else
    {
        V3 rW = rotate(qB, ld3(sp.rB));
        jL = -n;
        jA = -cross(rW, n);
    }
    // outer(a, b) * k: entry (r, c) = (b[c] * a[r]) * k
    float gl[3] = {jL.x, jL.y, jL.z}, ga[3] = {jA.x, jA.y, jA.z};
    r[R_L00] = (gl[0] * gl[0]) * k;
    r[R_L10] = (gl[0] * gl[1]) * k;
    r[R_L11] = (gl[1] * gl[1]) * k;
    r[R_L20] = (gl[0] * gl[2]) * k;
    r[R_L21] = (gl[1] * gl[2]) * k;
    r[R_L22] = (gl[2] * gl[2]) * k;
    r[R_A00] = (ga[0] * ga[0]) * k;
    r[R_A10] = (ga[0] * ga[1]) * k;
    r[R_A11] = (ga[1] * ga[1]) * k;
    r[R_A20] = (ga[0] * ga[2]) * k;
    r[R_A21] = (ga[1] * ga[2]) * k;
    r[R_A22] = (ga[2] * ga[2]) * k;
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b)
            r[R_C + 3 * a + b] = (gl[b] * ga[a]) * k;
    r[R_RL] = jL.x * f;
    r[R_RL + 1] = jL.y * f;
    r[R_RL + 2] = jL.z * f;
    r[R_RA] = jA.x * f;
    r[R_RA + 1] = jA.y * f;
    r[R_RA + 2] = jA.z * f;
}

// Per-body stamp lists, built once per step from the force lists: each manifold contributes
// one record per contact, each joint three, each spring one, in force order.
__global__ void k_stamp_count(const int *deg, const int *adjStart, const int *adj, const int *mcount, int n,
                              int *stampCount, int *maxStampCount)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int s = 0;
    if (i < n)
    {
        int b = adjStart[i], e = b + deg[i];
        // entries in batches of 8, then their contact counts, each batch in flight together
        for (int x0 = b; x0 < e; x0 += 8)
        {
            int en[8];
#pragma unroll
            for (int q = 0; q < 8; ++q)
                en[q] = x0 + q < e ? adj[x0 + q] : -1;
#pragma unroll
            for (int q = 0; q < 8; ++q)
                if (x0 + q < e)
                {
                    const int kind = adjKind(en[q]);
                    s += kind == kAdjManifold ? mcount[adjIndex(en[q])] : (kind == kAdjJoint ? 3 : 1);
                }
        }
        stampCount[i] = s;
    }
    block_max(s, maxStampCount, 0);
}

// Stamps in colour-sorted task order at a fixed stride S per task (S a power of two at least the
// largest stamp count): the primal then reads task k's stamps at k * S + t, an address that does
// not wait for the task record. Entries past a task's count are never read. A contact stamp gets
// kStampDualOwner when its body updates the contact's dual in the lazy-dual form: the
// lower-colour body of the pair, or the only solved one (colour -1 marks an unsolved partner).
constexpr int kStampDualOwner = 1 << 17;
__global__ void k_pad_stamps(const int4 *tasks, const int4 *stamps, const int *color, int numTasks, int shiftS,
                             int4 *padded)
{
    const int S = 1 << shiftS;
    for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < (long long)numTasks << shiftS;
         idx += (long long)gridDim.x * blockDim.x)
    {
        int k = (int)(idx >> shiftS), t = (int)(idx & (S - 1));
        int4 task = tasks[k];
        if (t < task.z)
        {
            int4 st = stamps[task.y + t];
            if (adjKind(st.x) == kAdjManifold)
            {
                int cp = color[st.z], ci = color[task.x];
                if (cp < 0 || ci < cp)
                    st.y |= kStampDualOwner;
            }
            padded[idx] = st;
        }
    }
}

__global__ void k_stamp_fill(const int *deg, const int *adjStart, const int *adj, const int *adjPartner,
                             const int *mcount, const int2 *mab, int n, const int *stampStart, int4 *stamps,
                             int maxStamps, int *overflow)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    int o = stampStart[i];
    int b = adjStart[i], e = b + deg[i];
    // force entries in batches of 8: the entries and partners, then their manifolds' contact
    // counts and body pairs, are in flight together; stamps are written in force order
    for (int x0 = b; x0 < e; x0 += 8)
    {
        int en[8], partner[8], parts[8], isA[8];
#pragma unroll
        for (int q = 0; q < 8; ++q)
        {
            en[q] = x0 + q < e ? adj[x0 + q] : -1;
            partner[q] = x0 + q < e ? adjPartner[x0 + q] : -1;
        }
#pragma unroll
        for (int q = 0; q < 8; ++q)
        {
            parts[q] = 0;
            isA[q] = 0;
            if (x0 + q < e)
            {
                const int kind = adjKind(en[q]), idx = adjIndex(en[q]);
                parts[q] = kind == kAdjManifold ? mcount[idx] : (kind == kAdjJoint ? 3 : 1);
                isA[q] = kind == kAdjManifold ? (mab[idx].x == i ? 1 : 0) : 0;
            }
        }
#pragma unroll
        for (int q = 0; q < 8; ++q)
        {
            const int kind = adjKind(en[q]), idx = adjIndex(en[q]);
            for (int p = 0; p < parts[q]; ++p, ++o)
            {
                if (o < maxStamps)
                    stamps[o] = make_int4(en[q], p | (isA[q] << 16), partner[q],
                                          kind == kAdjManifold ? idx * kMaxContacts + p : -1);
                else
                    atomicOr(overflow, 64);
            }
        }
    }
}

// Colour-sorted primal tasks: (body, first stamp, stamp count).
// done: the flow primal's per-body count of finished iterations, zeroed for every solved body.
// Inside a colour the tasks run by descending stamp count; the task where the count first drops
// to seg_stamps(q) or below starts segment q + 1 of the adaptive primal's lane counts.
__global__ void k_tasks(const int *colorBodies, const int *stampStart, const int *stampCount, const int *color, int n,
                        int4 *tasks, int *done, int *colorOffsets)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n)
        return;
    int i = colorBodies[k];
    const int cnt = stampCount[i];
    tasks[k] = make_int4(i, stampStart[i], cnt, 0);
    done[i] = 0;
    const int c = color[i];
    if (c < 0 || c >= kMaxColors)
        return;
    int cp = -1, cntp = 0;
    if (k > 0)
    {
        const int ip = colorBodies[k - 1];
        cp = color[ip];
        cntp = stampCount[ip];
    }
#pragma unroll
    for (int q = 0; q < kSegCount; ++q)
        if (cnt <= seg_stamps(q) && (cp != c || cntp > seg_stamps(q)))
            colorOffsets[kSegBase + kSegCount * c + q] = k;
}

__device__ __forceinline__ float init_entry(int e, float ml, float ax, float ay, float az, V3 dl, V3 da)
{
    switch (e)
    {
    case R_L00:
    case R_L11:
    case R_L22: return ml;
    case R_A00: return ax;
    case R_A11: return ay;
    case R_A22: return az;
    case R_RL: return ml * dl.x;
    case R_RL + 1: return ml * dl.y;
    case R_RL + 2: return ml * dl.z;
    case R_RA: return ax * da.x;
    case R_RA + 1: return ay * da.y;
    case R_RA + 2: return az * da.z;
    default: return 0.0f;
    }
}

// Offset in a body's 27-float system record of entry (r, k), k <= r, of the 6x6 matrix solve6
// factors (rows 0-2 linear, 3-5 angular; the coupling block is the record's lhsCross).
__device__ __forceinline__ int sys_entry(int r, int k)
{
    return r < 3 ? R_L00 + r * (r + 1) / 2 + k
                 : (k < 3 ? R_C + 3 * (r - 3) + k : R_A00 + (r - 3) * (r - 2) / 2 + (k - 3));
}

// Value of lane j (0..15) of this lane's row of 16 lanes (a DPP row broadcast). Every lane of the
// wave must be active; j must fold to a constant (it selects the DPP control code).
__device__ __forceinline__ float row_bcast(float x, int j)
{
    const int v = __float_as_int(x);
    int out = v;
    switch (j)
    {
#define AVBD_ROW_BCAST(n)                                                                                  \
    case n: out = __builtin_amdgcn_update_dpp(0, v, 0x150 + n, 0xF, 0xF, false); break;
        AVBD_ROW_BCAST(0) AVBD_ROW_BCAST(1) AVBD_ROW_BCAST(2) AVBD_ROW_BCAST(3) AVBD_ROW_BCAST(4)
        AVBD_ROW_BCAST(5) AVBD_ROW_BCAST(6) AVBD_ROW_BCAST(7) AVBD_ROW_BCAST(8) AVBD_ROW_BCAST(9)
        AVBD_ROW_BCAST(10) AVBD_ROW_BCAST(11) AVBD_ROW_BCAST(12) AVBD_ROW_BCAST(13) AVBD_ROW_BCAST(14)
        AVBD_ROW_BCAST(15)
#undef AVBD_ROW_BCAST
    default: break;
    }
    return __int_as_float(out);
}

// Value of lane j of this lane's group of L lanes (groups are aligned to L): a row broadcast for
// L >= 16; for L = 8, two groups share a row and each half takes its own broadcast.
template <int L>
__device__ __forceinline__ float group_bcast(float x, int j)
{
    static_assert(L >= 8 && L % 8 == 0, "groups of 8 lanes or more");
    if constexpr (L >= 16)
        return row_bcast(x, j);
    else
    {
        const float lo = row_bcast(x, j), hi = row_bcast(x, 8 + j);
        return (threadIdx.x & 8) ? hi : lo;
    }
}

// LDS hand-off between the lanes of one wave: a group's 16 lanes always sit in one wave, so
// ordering the wave's own LDS writes before its reads is enough (no block barrier).
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Primal update of the bodies of one colour (solver.cpp:1184-1240): L lanes per body, T threads
// per block. Lane l computes the stamp records l, l+L, l+2L, ... of its body; the records of one
// round are then added entry by entry in stamp order, so every (L, T) gives the same sums.
// Bodies of a colour are sorted by stamp count, so the groups of a wave need similar round counts.
// sysTape / flagTape (adjoint recording, null otherwise): per body the accumulated system and
// dx (9 float4), per stamp the contact's branch flags.
// One colour-sorted task (a body and its stamps) of the primal, for L lanes per body and T
// threads per block: the stamps' 27-float records go straight into the lanes' LDS rows (odd
// stride), the owner lanes add each system entry in stamp order, lane 0 solves and stores.
// Every (L, T) gives the same sums. Co: the dataflow form (coherent partner poses and pose
// stores).
// pad (may be null): the stamps at stride 2^padShift per task, task k = kpos; st0 is this lane's
// first padded stamp, loaded by the caller next to the task record.
template <int L, int T, bool Co>
__device__ __forceinline__ void primal_task(float *rec, int g, int lane, bool valid, int4 task, const Bodies &B,
                                            const Manifolds &M, const Joint *joints, const Spring *springs,
                                            const int4 *stamps, const Params &P, float4 *sysTape,
                                            unsigned char *flagTape, const int4 *pad = nullptr, int padShift = 0,
                                            int kpos = 0, int4 st0 = make_int4(0, 0, 0, 0), bool lazyDual = false)
{
    static_assert(L >= 4 && L <= 64 && (64 % L) == 0 && (T % 64) == 0, "a body's lanes must sit in one wave");
    constexpr int E = (kRec + L - 1) / L;  // accumulated entries per lane
    constexpr int kRow = kRec + 2;         // odd row stride: a wave's row writes hit distinct banks
    const int i = task.x;
    const int s0 = task.y;
    const int ns = valid ? task.z : 0;
    const float4 mm = B.massMoment[i];
    const float4 pos4 = Co ? ld_coherent(&B.pose[4 * i + P_POS]) : B.pose[4 * i + P_POS];
    const float4 rot4 = Co ? ld_coherent(&B.pose[4 * i + P_ROT]) : B.pose[4 * i + P_ROT];
    const float4 il4 = B.inerLin[i], ia4 = B.inerAng[i];
    V3 pos = ld3(pos4);
    Q4 rot = ldq(rot4);
    const float oma = 1.0f - P.alpha;
    // Round r's records (stamps r L + lane) straight into this lane's LDS row; a contact writes
    // all of it, the other kinds write part of it and the row is cleared first.
    auto writeRound = [&](int r) {
        int t = r * L + lane;
        float *row = rec + (g * L + lane) * kRow;
        if (valid && t < ns)
        {
            // the flow form (Co) always has task-ordered stamps and only contacts
            int4 st;
            if (!Co && !pad)
                st = stamps[s0 + t];
            else if (r == 0)
                st = st0;
            else
                st = pad[((size_t)kpos << padShift) + t];
            int kind = adjKind(st.x), idx = adjIndex(st.x);
            if (Co || kind == kAdjManifold)
            {
                const bool owner = lazyDual && (st.y & kStampDualOwner) != 0;
                const bool isA = ((st.y >> 16) & 1) != 0;
                const ContactIn in = contact_load<Co>(i, st.z, isA, st.w, M, B);
                int fl = contact_record<Co>(row, in, isA, st.w, M, oma, owner ? &P : nullptr);
                if (flagTape)
                    flagTape[s0 + t] = (unsigned char)fl;
            }
            else
            {
#pragma unroll
                for (int q = 0; q < kRec; ++q)
                    row[q] = 0.0f;
                if (kind == kAdjJoint)
                    joint_record(row, i, joints[idx], st.y & 0xFFFF, B, P.alpha);
                else
                    spring_record(row, i, springs[idx], B);
            }
        }
        else
        {
#pragma unroll
            for (int q = 0; q < kRec; ++q)
                row[q] = 0.0f;
        }
    };
    // The records of round r added entry by entry in stamp order.
    float acc[E];
    auto sumRound = [&](int r) {
        wave_lds_sync();
        int cnt = valid ? min(L, ns - r * L) : 0;
        float v[E][L];
#pragma unroll
        for (int e = 0; e < E; ++e)
#pragma unroll
            for (int q = 0; q < L; ++q)
                v[e][q] = lane + e * L < kRec ? rec[(g * L + q) * kRow + lane + e * L] : 0.0f;
#pragma unroll
        for (int q = 0; q < L; ++q)
        {
            if (q < cnt)
            {
#pragma unroll
                for (int e = 0; e < E; ++e)
                    acc[e] += v[e][q];
            }
        }
        wave_lds_sync();
    };
    // Round 0 first: its stamp loads go out next to the body's own rows, and the inertia terms
    // (divisions by dt^2, the quaternion difference) are formed while they are in flight.
    writeRound(0);
    {
        float dt2 = P.dt * P.dt;
        float ml = mm.x / dt2, ax = mm.y / dt2, ay = mm.z / dt2, az = mm.w / dt2;
        V3 dl = pos - ld3(il4);
        V3 da = qdiff(rot, ldq(ia4));
#pragma unroll
        for (int q = 0; q < E; ++q)
            acc[q] = lane + q * L < kRec ? init_entry(lane + q * L, ml, ax, ay, az, dl, da) : 0.0f;
    }
    sumRound(0);
    int rounds = (ns + L - 1) / L;
#pragma unroll
    for (int off = 32; off >= L; off >>= 1)
        rounds = max(rounds, __shfl_xor(rounds, off, 64));
    for (int r = 1; r < rounds; ++r)
    {
        writeRound(r);
        sumRound(r);
    }
#pragma unroll
    for (int e = 0; e < E; ++e)
        if (lane + e * L < kRec)
            rec[g * L * kRow + lane + e * L] = acc[e];
    wave_lds_sync();
    // The 6x6 LDL^T solve of solve6 (maths.h:481-547) spread over lanes 0..5 of the group: lane r
    // owns row r + 1, the columns are factored in order (the column's owner broadcasts its
    // diagonal and its row of L, every lower row forms its entry with solve6's own expression),
    // the forward substitution broadcasts each y_k, each lane divides its own z; lane 0 then runs
    // the back substitution from L and z broadcast to it. Broadcasts inside the group are DPP row
    // broadcasts (a group's lanes 0..5 sit in one row of 16). Every value is computed by the same
    // operations in the same order as solve6, so the result is bit-identical. All lanes of the
    // wave run this part (invalid groups on junk, never stored): the broadcasts need every lane.
    V3 dxL, dxA;
    {
        const float *a = rec + g * L * kRow;
        const int r = lane < 6 ? lane : 5;
        float arow[6], lrow[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f}, d[6];
#pragma unroll
        for (int k = 0; k < 6; ++k)
            arow[k] = k <= r ? a[sys_entry(r, k)] : 0.0f;
        float myD = 0.0f;
        d[0] = group_bcast<L>(arow[0], 0);  // D1 = A11
        myD = r == 0 ? d[0] : myD;
        lrow[0] = arow[0] / d[0];           // L_i1 = A_i1 / A11
#pragma unroll
        for (int j = 1; j < 6; ++j)
        {
            // row j + 1 of L, from its owner
            float lj[6];
#pragma unroll
            for (int k = 0; k < j; ++k)
                lj[k] = group_bcast<L>(lrow[k], j);
            // D_j = A_jj - (sum_k (L_jk L_jk) D_k), summed left to right (computed by its owner)
            float sum = (lrow[0] * lrow[0]) * d[0];
#pragma unroll
            for (int k = 1; k < j; ++k)
                sum = sum + (lrow[k] * lrow[k]) * d[k];
            d[j] = group_bcast<L>(arow[j] - sum, j);
            myD = r == j ? d[j] : myD;
            // L_ij = (A_ij - L_j1 L_i1 D_1 - ... ) / D_j for the rows below
            float t = arow[j];
#pragma unroll
            for (int k = 0; k < j; ++k)
                t = t - (lj[k] * lrow[k]) * d[k];
            if (r > j)
                lrow[j] = t / d[j];
        }
        // forward substitution: y_i = b_i - L_i1 y_1 - ... - L_i(i-1) y_(i-1)
        float y = -(r < 3 ? a[R_RL + r] : a[R_RA + r - 3]);
#pragma unroll
        for (int k = 0; k < 5; ++k)
        {
            const float yk = group_bcast<L>(y, k);
            if (r > k)
                y = y - lrow[k] * yk;
        }
        const float z = y / myD;
        // L (1-based L_ik = lrow[k - 1] of lane i - 1) and z for lane 0
        float Lg[6][6], zg[6];
#pragma unroll
        for (int i = 0; i < 6; ++i)
        {
            zg[i] = group_bcast<L>(z, i);
#pragma unroll
            for (int k = 0; k < i; ++k)
                Lg[i][k] = group_bcast<L>(lrow[k], i);
        }
        if (lane == 0)
        {
            auto Lm = [&](int i, int k) { return Lg[i - 1][k - 1]; };  // L_ik, 1-based
            const float z1 = zg[0], z2 = zg[1], z3 = zg[2], z4 = zg[3], z5 = zg[4], z6 = zg[5];
            dxA.z = z6;
            dxA.y = z5 - Lm(6, 5) * dxA.z;
            dxA.x = z4 - Lm(5, 4) * dxA.y - Lm(6, 4) * dxA.z;
            dxL.z = z3 - Lm(4, 3) * dxA.x - Lm(5, 3) * dxA.y - Lm(6, 3) * dxA.z;
            dxL.y = z2 - Lm(3, 2) * dxL.z - Lm(4, 2) * dxA.x - Lm(5, 2) * dxA.y - Lm(6, 2) * dxA.z;
            dxL.x = z1 - Lm(2, 1) * dxL.y - Lm(3, 1) * dxL.z - Lm(4, 1) * dxA.x - Lm(5, 1) * dxA.y -
                    Lm(6, 1) * dxA.z;
        }
    }
    if (valid && lane == 0)
    {
        const float *a = rec + g * L * kRow;
        float4 np = st3(pos + dxL, 0.0f), nq = stq(qintegrate(rot, dxA));
        if (Co)
        {
            st_coherent(&B.pose[4 * i + P_POS], np);
            st_coherent(&B.pose[4 * i + P_ROT], nq);
        }
        else
        {
            B.pose[4 * i + P_POS] = np;
            B.pose[4 * i + P_ROT] = nq;
        }
        if (sysTape)
        {
            float4 *o = sysTape + 9 * (size_t)i;
#pragma unroll
            for (int q = 0; q < 6; ++q)
                o[q] = make_float4(a[4 * q], a[4 * q + 1], a[4 * q + 2], a[4 * q + 3]);
            o[6] = make_float4(a[24], a[25], a[26], dxL.x);
            o[7] = make_float4(dxL.y, dxL.z, dxA.x, dxA.y);
            o[8] = make_float4(dxA.z, 0.0f, 0.0f, 0.0f);
        }
    }
}

// Primal update of the bodies of one colour (solver.cpp:1184-1240): one launch per colour.
// Bodies of a colour are sorted by stamp count, so the groups of a wave need similar round counts.
// sysTape / flagTape (adjoint recording, null otherwise): per body the accumulated system and
// dx (9 float4), per stamp the contact's branch flags.
// pad (may be null): the stamps in task order at stride 2^padShift (k_pad_stamps); a lane's first
// stamp is then loaded together with the task record instead of after it. lazyDual (with pad only):
// a stamp marked kStampDualOwner applies its contact point's dual update of the previous iteration
// before building its record.
template <int L, int T>
__global__ __launch_bounds__(T) void k_primal_color(Bodies B, Manifolds M, const Joint *joints, const Spring *springs,
                                                    const int4 *stamps, const int4 *tasks, int begin, int end, Params P,
                                                    float4 *sysTape, unsigned char *flagTape, const int4 *pad,
                                                    int padShift, int lazyDual)
{
    constexpr int G = T / L;
    __shared__ float rec[T * (kRec + 2)];
    const int lane = threadIdx.x % L, g = threadIdx.x / L;
    const int k = begin + blockIdx.x * G + g;
    const bool valid = k < end;
    const int4 task = valid ? tasks[k] : make_int4(0, 0, 0, 0);
    const int4 st0 = (valid && pad) ? pad[((size_t)k << padShift) + lane] : make_int4(0, 0, 0, 0);
    primal_task<L, T, false>(rec, g, lane, valid, task, B, M, joints, springs, stamps, P, sysTape, flagTape, pad,
                             padShift, k, st0, lazyDual != 0);
}

// Adaptive lanes per body (launch config primal_lanes = 0): each body gets the fewest of 64, 32,
// 16 or 8 lanes that hold all its stamps in one round, so no wave waits on a second round of
// stamp loads and records. A colour's tasks (descending stamp count) are cut into four segments
// at counts 32, 16 and 8 (k_tasks); the launch's blocks serve the segments in order, block b of
// segment q running T / L_q bodies. Every body's sums are the same as with any fixed L.
template <int L, int T>
__device__ __forceinline__ void primal_segment(float *rec, int bb, int begin, int end, const Bodies &B,
                                               const Manifolds &M, const Joint *joints, const Spring *springs,
                                               const int4 *stamps, const int4 *tasks, const Params &P,
                                               float4 *sysTape, unsigned char *flagTape, const int4 *pad,
                                               int padShift, int lazyDual)
{
    constexpr int G = T / L;
    const int lane = threadIdx.x % L, g = threadIdx.x / L;
    const int k = begin + bb * G + g;
    const bool valid = k < end;
    const int4 task = valid ? tasks[k] : make_int4(0, 0, 0, 0);
    const int4 st0 = (valid && pad) ? pad[((size_t)k << padShift) + lane] : make_int4(0, 0, 0, 0);
    primal_task<L, T, false>(rec, g, lane, valid, task, B, M, joints, springs, stamps, P, sysTape, flagTape, pad,
                             padShift, k, st0, lazyDual != 0);
}

// seg: the starts of the 64-, 32-, 16- and 8-lane segments of the colour; blk: the first block
// of the 32-, 16- and 8-lane segments and the block count.
template <int T>
__global__ __launch_bounds__(T) void k_primal_adaptive(Bodies B, Manifolds M, const Joint *joints,
                                                       const Spring *springs, const int4 *stamps, const int4 *tasks,
                                                       int4 seg, int end, int4 blk, Params P, float4 *sysTape,
                                                       unsigned char *flagTape, const int4 *pad, int padShift,
                                                       int lazyDual)
{
    __shared__ float rec[T * (kRec + 2)];
    const int b = blockIdx.x;
    if (b < blk.x)
        primal_segment<64, T>(rec, b, seg.x, seg.y, B, M, joints, springs, stamps, tasks, P, sysTape, flagTape, pad,
                              padShift, lazyDual);
    else if (b < blk.y)
        primal_segment<32, T>(rec, b - blk.x, seg.y, seg.z, B, M, joints, springs, stamps, tasks, P, sysTape,
                              flagTape, pad, padShift, lazyDual);
    else if (b < blk.z)
        primal_segment<16, T>(rec, b - blk.y, seg.z, seg.w, B, M, joints, springs, stamps, tasks, P, sysTape,
                              flagTape, pad, padShift, lazyDual);
    else
        primal_segment<8, T>(rec, b - blk.z, seg.w, end, B, M, joints, springs, stamps, tasks, P, sysTape, flagTape,
                             pad, padShift, lazyDual);
}

// Flow form of a step's solver iterations (contact-only worlds, lazy dual): ONE launch runs every
// iteration and colour. Instead of a kernel boundary after each colour (70 per step at 7 colours),
// each body waits only for the partners it reads: done[j] counts the iterations body j has finished
// in this step (k_tasks sets it to 0), and body i of colour c starts iteration `it` once every
// partner j of a lower colour has done[j] >= it + 1 and every partner of a higher colour has
// done[j] >= it (its pose and the dual updates i applies are then those the per-colour sweep sees).
// The dependencies are exactly the coloured Gauss-Seidel order, so results equal the per-colour
// launches bit for bit.
//
// Schedule: the colour-sorted tasks of each colour are cut into batches of one wave (64 / L
// bodies); batch b of the step's batch list goes to wave b mod NW, every wave visits its batches in
// (iteration, colour) order. The grid is sized so that every wave is resident at once (the
// occupancy limit), so the wave holding the earliest unfinished batch always has its dependencies
// met and the grid drains; a wait that outlives kFlowSpin polls raises the timeout flag and every
// later wait gives up at once (the step's results are then void and the host reports the error).
//
// Coherence without cache flushes: poses and each contact's penalty/lambda (the data one body
// writes and another reads) are stored and loaded with agent-scope atomics, which bypass the
// per-XCD L2 slices; a body publishes done[i] only after s_waitcnt vmcnt(0) has seen its wave's
// stores complete. Everything else the primal reads is fixed for the step and cached normally.
constexpr int kFlowSpin = 1 << 18;
// s_waitcnt with vmcnt = 0 and expcnt / lgkmcnt at their maxima (GFX9 encoding): wait until every
// vector memory operation of the wave has completed.
constexpr int kWaitVmcnt0 = 0x0F70;

__device__ __forceinline__ bool flow_timed_out(Counters *ctr)
{
    return __hip_atomic_load(&ctr->barrierTimeout, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
}

// Waits until body `i` (task position k, colour c) may run iteration `it`: lane l checks the
// partners of stamps l, l + L, ...
template <int L>
__device__ __forceinline__ void flow_wait(int lane, int4 task, int k, int c, int it, const int4 *pad, int padShift,
                                          const int *color, const int *done, Counters *ctr)
{
    for (int t = lane; t < task.z; t += L)
    {
        const int j = pad[((size_t)k << padShift) + t].z;
        if (j < 0)
            continue;
        const int cj = color[j];
        if (cj < 0)
            continue;  // never solved in this step: its pose does not change
        const int need = cj < c ? it + 1 : it;
        int spins = 0;
        while (__hip_atomic_load(&done[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < need)
        {
            __builtin_amdgcn_s_sleep(1);
            if (++spins >= kFlowSpin || ((spins & 255) == 0 && flow_timed_out(ctr)))
            {
                atomicOr(&ctr->barrierTimeout, 1);
                break;
            }
        }
    }
    // the partner loads that follow are issued after the polls returned (no compiler reordering)
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

template <int L, int T>
__global__ __launch_bounds__(T) void k_primal_flow(Bodies B, Manifolds M, const int4 *tasks, const int *colorOffsets,
                                                   int colors, const int *color, int *done, Params P, const int4 *pad,
                                                   int padShift, Counters *ctr)
{
    constexpr int G = 64 / L;  // bodies per wave batch
    __shared__ float rec[T * (kRec + 2)];
    const int lane = threadIdx.x % L, g = threadIdx.x / L, gw = (int)(threadIdx.x & 63) / L;
    const int wave = (int)blockIdx.x * (T / 64) + (int)(threadIdx.x >> 6);
    const int numWaves = (int)gridDim.x * (T / 64);
    for (int it = 0; it < P.iterations; ++it)
    {
        int batchBase = 0;  // batches of the colours before c
        for (int c = 0; c < colors; ++c)
        {
            const int cb = colorOffsets[c], ce = colorOffsets[c + 1];
            const int nb = (ce - cb + G - 1) / G;
            for (int m = ((wave - batchBase) % numWaves + numWaves) % numWaves; m < nb; m += numWaves)
            {
                const int k = cb + m * G + gw;
                const bool valid = k < ce;
                const int4 task = valid ? tasks[k] : make_int4(0, 0, 0, 0);
                const int4 st0 = valid ? pad[((size_t)k << padShift) + lane] : make_int4(0, 0, 0, 0);
                if (valid && !flow_timed_out(ctr))
                    flow_wait<L>(lane, task, k, c, it, pad, padShift, color, done, ctr);
                primal_task<L, T, true>(rec, g, lane, valid, task, B, M, nullptr, nullptr, nullptr, P, nullptr, nullptr,
                                        pad, padShift, k, st0, it > 0);
                // publish only once this wave's pose and penalty/lambda stores have completed
                __atomic_signal_fence(__ATOMIC_SEQ_CST);
                __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
                __atomic_signal_fence(__ATOMIC_SEQ_CST);
                if (valid && lane == 0)
                    __hip_atomic_store(&done[task.x], it + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            batchBase += nb;
        }
    }
}

// Flattened contact records (manifold, contact, body A, body B) for the per-contact dual pass.
__global__ void k_contact_records(const int *mcount, const int2 *mab, const int *contactStart, const Counters *ctr,
                                  int4 *rec)
{
    int nm = ctr->numManifolds;
    for (int m = blockIdx.x * blockDim.x + threadIdx.x; m < nm; m += gridDim.x * blockDim.x)
    {
        int o = contactStart[m], n = mcount[m];
        int2 ab = mab[m];
        for (int c = 0; c < n; ++c)
            rec[o + c] = make_int4(m, c, ab.x, ab.y);
    }
}

// Manifold::updateDual for one contact (manifold.cpp:231-280)
__device__ __forceinline__ int dual_contact(int4 cr, const Manifolds &M, const Bodies &B, const Params &P)
{
    int2 ab = make_int2(cr.z, cr.w);
    int sl = cr.x * kMaxContacts + cr.y;
    float4 *rec = M.cr + (size_t)sl * kCR;
    float4 rA4 = rec[CR_RA], rB4 = rec[CR_RB], c04 = rec[CR_C0], bb0 = rec[CR_N], bb1 = rec[CR_T1], bb2 = rec[CR_T2];
    float4 pen4 = rec[CR_PEN], lam4 = rec[CR_LAM];
    float4 pA4 = B.pose[4 * (ab.x) + P_POS], pB4 = B.pose[4 * (ab.y) + P_POS], qA4 = B.pose[4 * (ab.x) + P_ROT], qB4 = B.pose[4 * (ab.y) + P_ROT];
    float4 iA4 = B.pose[4 * (ab.x) + P_ILIN], iB4 = B.pose[4 * (ab.y) + P_ILIN], jA4 = B.pose[4 * (ab.x) + P_IANG], jB4 = B.pose[4 * (ab.y) + P_IANG];
    Q4 qA = ldq(qA4), qB = ldq(qB4);
    V3 dqAL = ld3(pA4) - ld3(iA4);
    V3 dqAA = qdiff(qA, ldq(jA4));
    V3 dqBL = ld3(pB4) - ld3(iB4);
    V3 dqBA = qdiff(qB, ldq(jB4));
    V3 b0 = ld3(bb0), b1 = ld3(bb1), b2 = ld3(bb2);
    V3 nb0 = -b0, nb1 = -b1, nb2 = -b2;
    V3 rAW, rBW;
    contact_anchors(rA4, rB4, c04, bb0, bb1, bb2, qA, qB, rAW, rBW);
    V3 ja0 = cross(rAW, b0), ja1 = cross(rAW, b1), ja2 = cross(rAW, b2);
    V3 jb0 = cross(rBW, nb0), jb1 = cross(rBW, nb1), jb2 = cross(rBW, nb2);
    V3 pen = ld3(pen4);
    ContactEval ev = contact_eval(ld3(c04), pen, ld3(lam4), rA4.w, 1.0f - P.alpha, b0, b1, b2, nb0, nb1, nb2, ja0,
                                  ja1, ja2, jb0, jb1, jb2, dqAL, dqBL, dqAA, dqBA);
    float stick = lam4.w;
    int fl = contact_flags(ev, pen, ld3(lam4));
    if (ev.F0 < 0.0f)
    {
        float k0 = pen.x + P.betaLin * fabsf(ev.C0);
        fl |= kFlagUpd0 | (!(k0 < kPenaltyMax) ? kFlagSat0 : 0);
        pen.x = fminr(k0, kPenaltyMax);
    }
    if (ev.fs <= ev.bounds)
    {
        float k1 = pen.y + P.betaLin * fabsf(ev.C1), k2 = pen.z + P.betaLin * fabsf(ev.C2);
        fl |= kFlagUpd12 | (!(k1 < kPenaltyMax) ? kFlagSat0 << 1 : 0) | (!(k2 < kPenaltyMax) ? kFlagSat0 << 2 : 0);
        pen.y = fminr(k1, kPenaltyMax);
        pen.z = fminr(k2, kPenaltyMax);
        stick = sqrtf(ev.C1 * ev.C1 + ev.C2 * ev.C2) < kStickThresh ? 1.0f : 0.0f;
    }
    rec[CR_LAM] = make_float4(ev.F0, ev.F1, ev.F2, stick);
    rec[CR_PEN] = st3(pen, 0.0f);
    return fl;
}

// Joint::updateDual (joint.cpp:125-182)
__device__ void dual_joint(Joint *jp, const Bodies &B, const Params &P)
{
    Joint j = *jp;
    if (j.dead)
        return;
    Q4 qB = ldq(B.pose[4 * (j.b) + P_ROT]);
    V3 pB = ld3(B.pose[4 * (j.b) + P_POS]);
    V3 penL = ld3(j.penLin), penA = ld3(j.penAng), lamL = ld3(j.lamLin), lamA = ld3(j.lamAng);
    if (lensq(penL) > 0.0f)
    {
        D3 K = d3_diag(penL.x, penL.y, penL.z);
        V3 pa = j.a >= 0 ? rotate(ldq(B.pose[4 * (j.a) + P_ROT]), ld3(j.rA)) + ld3(B.pose[4 * (j.a) + P_POS]) : ld3(j.rA);
        V3 C = pa - (rotate(qB, ld3(j.rB)) + pB);
        if (isinf(j.rA.w))
        {
            C -= ld3(j.c0Lin) * P.alpha;
            lamL = d3_mv(K, C) + lamL;
        }
        float cap = fminr(j.rA.w, kPenaltyMax);
        V3 t = penL + V3{fabsf(C.x), fabsf(C.y), fabsf(C.z)} * P.betaLin;
        penL = V3{fminr(t.x, cap), fminr(t.y, cap), fminr(t.z, cap)};
    }
    if (lensq(penA) > 0.0f)
    {
        D3 K = d3_diag(penA.x, penA.y, penA.z);
        Q4 qa = j.a >= 0 ? ldq(B.pose[4 * (j.a) + P_ROT]) : Q4{0.0f, 0.0f, 0.0f, 1.0f};
        V3 C = qdiff(qa, qB) * j.c0Lin.w;
        if (isinf(j.rB.w))
        {
            C -= ld3(j.c0Ang) * P.alpha;
            lamA = d3_mv(K, C) + lamA;
        }
        float cap = fminr(j.rB.w, kPenaltyMax);
        V3 t = penA + V3{fabsf(C.x), fabsf(C.y), fabsf(C.z)} * P.betaAng;
        penA = V3{fminr(t.x, cap), fminr(t.y, cap), fminr(t.z, cap)};
    }
    int broken = j.broken;
    if (lensq(lamA) > j.c0Ang.w * j.c0Ang.w)
    {
        penL = penA = lamL = lamA = V3{0.0f, 0.0f, 0.0f};
        broken = 1;
    }
    jp->penLin = st3(penL, 0.0f);
    jp->penAng = st3(penA, 0.0f);
    jp->lamLin = st3(lamL, 0.0f);
    jp->lamAng = st3(lamA, 0.0f);
    jp->broken = broken;
}

// Dual pass of one iteration: one lane per contact, then one per joint.
// flagTape (adjoint recording, may be null): per contact slot the dual's branch flags.
// staticRepeat (lazy-dual form's closing launch, 1 otherwise): a contact between two bodies that
// are never solved (both static) has no owner stamp to carry its per-iteration dual updates; its
// bodies never move, so its updates of all iterations are applied here one after another.
// The adjoint tape's slot after this iteration, written by the dual launch (null pointers when no
// step is taped): each contact's penalty and lambda after its update, and threads past the joints
// copy the bodies' poses (the primal's results; the dual does not move them).
struct DualTape
{
    float4 *pose = nullptr, *contact = nullptr;
    int nb = 0;
};

__global__ void k_dual(Bodies B, Manifolds M, Joint *joints, int nj, const int4 *contactRec, int nc, Params P,
                       unsigned char *flagTape, int staticRepeat, DualTape tape)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < nc)
    {
        const int4 cr = contactRec[t];
        int reps = 1;
        if (staticRepeat > 1 && !(B.massMoment[cr.z].x > 0.0f) && !(B.massMoment[cr.w].x > 0.0f))
            reps = staticRepeat;
        int fl = 0;
        for (int r = 0; r < reps; ++r)
            fl = dual_contact(cr, M, B, P);
        const int sl = cr.x * kMaxContacts + cr.y;
        if (flagTape)
            flagTape[sl] = (unsigned char)fl;
        if (tape.contact)
        {
            tape.contact[2 * sl] = M.cr[(size_t)sl * kCR + CR_PEN];
            tape.contact[2 * sl + 1] = M.cr[(size_t)sl * kCR + CR_LAM];
        }
    }
    else if (t - nc < nj)
        dual_joint(&joints[t - nc], B, P);
    else if (tape.pose && t - nc - nj < tape.nb)
    {
        const int i = t - nc - nj;
        tape.pose[2 * i] = B.pose[4 * i + P_POS];
        tape.pose[2 * i + 1] = B.pose[4 * i + P_ROT];
    }
}

// Adjoint tape of the step being reversed: poses (pos, rot per body) and contact penalty/lambda
// (per contact slot) after the warm start / force init and after every iteration's primal and dual.
// One launch: threads [0, n) the poses, the rest (grid-stride) the contacts.
__global__ void k_tape(const float4 *pose, int n, float4 *poseDst, const int4 *contactRec, const Counters *ctr,
                       const float4 *cr, float4 *contactDst)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n)
    {
        poseDst[2 * t] = pose[4 * t + P_POS];
        poseDst[2 * t + 1] = pose[4 * t + P_ROT];
        return;
    }
    const int nc = ctr->contactCount;
    for (int c = t - n; c < nc; c += gridDim.x * blockDim.x - n)
    {
        int4 r = contactRec[c];
        int sl = r.x * kMaxContacts + r.y;
        contactDst[2 * sl] = cr[(size_t)sl * kCR + CR_PEN];
        contactDst[2 * sl + 1] = cr[(size_t)sl * kCR + CR_LAM];
    }
}

// Device-to-device copies in one launch (the adjoint's checkpoints and step stores): entry k
// copies words[k] 4-byte words from src[k] to dst[k]; start[k] is the running word total.
constexpr int kCopyMax = 32;
struct CopyList
{
    const unsigned *src[kCopyMax];
    unsigned *dst[kCopyMax];
    long long start[kCopyMax + 1];
    int n;
};

__global__ void k_copy_many(CopyList L)
{
    const long long total = L.start[L.n];
    for (long long w = (long long)blockIdx.x * blockDim.x + threadIdx.x; w < total;
         w += (long long)gridDim.x * blockDim.x)
    {
        int k = 0;
        while (w >= L.start[k + 1])
            ++k;
        const long long o = w - L.start[k];
        L.dst[k][o] = L.src[k][o];
    }
}


// Solver::finishStep for boxes (solver.cpp:1275-1288).
// applySphereRollingFriction (solver.cpp:628-660), per dynamic sphere after k_finish: its
// manifolds in force-list order (the adjacency is sorted newest first) against static bodies,
// where this sphere is the pair's chosen sphere (A if A is a sphere), blend the contact-point
// slip out of the velocity and pull the spin toward rolling, once per contact point.
__global__ void k_rolling(Bodies B, Manifolds M, const int *deg, const int *adjStart, const int *adj,
                          const int *adjPartner, int n, int maxAdj)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    float4 hs = B.shape[i];
    if ((int)hs.z != kShapeSphere || !(B.massMoment[i].x > 0.0f))
        return;
    int b = adjStart[i], e = min(b + deg[i], maxAdj);
    if (b >= e)
        return;
    V3 v = ld3(B.vel[i]), w = ld3(B.angVel[i]);
    float r = hs.x, fi = B.friction[i];
    for (int x = b; x < e; ++x)
    {
        int ent = adj[x];
        if (adjKind(ent) != kAdjManifold)
            continue;
        int m = adjIndex(ent), o = adjPartner[x];
        int2 ab = M.ab[m];
        bool isA = ab.x == i;
        if (!isA && (int)B.shape[ab.x].z == kShapeSphere)
            continue;
        int nc = M.count[m];
        if (B.massMoment[o].x > 0.0f || nc <= 0)
            continue;
        float mu = sqrtf(fi * B.friction[o]);
        if (mu <= 0.0f)
            continue;
        float blend = clampr(mu * 0.08f, 0.0f, 0.18f);
        V3 n0 = ld3(M.cr[(size_t)m * kMaxContacts * kCR + CR_N]);
        V3 nrm = isA ? n0 : -n0;
        for (int c = 0; c < nc; ++c)
        {
            V3 rW = -nrm * r;
            V3 cv = v + cross(w, rW);
            V3 slip = cv - nrm * dot(cv, nrm);
            v = v - slip * blend;
            V3 roll = cross(nrm, v) / r;
            w = w + (roll - w) * blend;
        }
    }
    B.vel[i] = st3(v, 0.0f);
    B.angVel[i] = st3(w, 0.0f);
}

// AVBS body records (viewer_bridge.cpp:796-828): 64 bytes per body at its rank in ascending
// dense id; id, shape, material (viewer_bridge.cpp:475-493), dynamic, position, rotation, size,
// shape radius, half length. Four 16-byte stores per body.
__global__ void k_avbs_pack(Bodies B, const int *rank, int n, uint4 *out)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    float4 p = B.pose[4 * i + P_POS], q = B.pose[4 * i + P_ROT], sr = B.sizeRad[i], sh = B.shape[i];
    float mass = B.massMoment[i].x;
    int shape = (int)sh.z;
    unsigned material = 0u;
    if (mass > 0.0f)
        material = shape == kShapeSphere ? 3u : shape == kShapeCylinder ? 4u : shape == kShapeCapsule ? 5u : 1u;
    uint4 *o = out + 4 * (size_t)rank[i];
    o[0] = make_uint4((unsigned)B.denseId[i], (unsigned)shape, material, mass > 0.0f ? 1u : 0u);
    o[1] = make_uint4(__float_as_uint(p.x), __float_as_uint(p.y), __float_as_uint(p.z), __float_as_uint(q.x));
    o[2] = make_uint4(__float_as_uint(q.y), __float_as_uint(q.z), __float_as_uint(q.w), __float_as_uint(sr.x));
    o[3] = make_uint4(__float_as_uint(sr.y), __float_as_uint(sr.z), __float_as_uint(sh.x), __float_as_uint(sh.y));
}

__global__ void k_finish(Bodies B, Params P, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    float4 v = B.vel[i];
    B.prevVel[i] = v;
    if (B.massMoment[i].x > 0.0f)
    {
        V3 nv = (ld3(B.pose[4 * (i) + P_POS]) - ld3(B.pose[4 * (i) + P_ILIN])) / P.dt;
        V3 nw = qdiff(ldq(B.pose[4 * (i) + P_ROT]), ldq(B.pose[4 * (i) + P_IANG])) / P.dt;
        B.vel[i] = st3(nv, 0.0f);
        B.angVel[i] = st3(nw, 0.0f);
    }
}

} // namespace avbd

