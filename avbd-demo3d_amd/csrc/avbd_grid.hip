// avbd_grid.hip -- the uniform-grid and all-pairs broadphases on the GPU
// (broadphaseUniformGrid, solver.cpp:284-486; broadphaseAllPairs, solver.cpp:176-185).
//
// Grid. Cells are twice the 95th-percentile bounding radius (at least 1e-3 m); bodies no larger
// than half a cell are binned by centre, the larger ones carry a world AABB (boxes) or their
// bounding cube (round shapes) plus a |v| dt slack. Emission order is the reference's: binned
// bodies sorted by (cell key, list index), each with its hits in its own cell, then in the 13
// forward neighbour cells, then against the large bodies in list order; after them the large
// bodies' pairs, in list order, by the plain bounding-sphere test.
//
// The launch sequence:
//   k_grid_select   one block: the rank-pct radius by an 8-bit radix select over its float bits
//   k_grid_bounds   min/max of the binned centres, binned count
//   (host)          cell size and dimensions from those, with the host's cbrtf when the grid has
//                   to be coarsened below 2^22 cells, exactly as the reference computes them
//   k_grid_keys     cell key per body (large bodies: a sentinel after every cell key)
//   radix sort      3 passes of 8 bits, stable, so ties keep list order
//   k_grid_gather   sorted binned entries (centre, radius, slack), first entry of each cell,
//                   and the large bodies' AABBs, centres and radii
//   k_grid_pairs<0> hits and checks per entry (binned entries, then large bodies)
//   scan, hash      start of each entry's hits; the constrained-pair hash (shared with SAP)
//   k_grid_pairs<1> the hits in emission order, with their new-manifold flag
//   k_grid_clear    the cell-start table back to empty for the next step
// All pairs uses the same pair kernels with every body a large body, in list order.
//
// Every float operation follows the reference's expression and order, so the pair emission,
// the check count and therefore the manifold creation order are bit-exact.
#include <cstring>

namespace avbd
{

constexpr unsigned kGridSentinel = 0xFFFFFFu;  // cell key of a large body (keys are below 2^22)
constexpr unsigned long long kGridMaxCells = 1ull << 22;
constexpr int kGridTile = 512;                 // large bodies staged in LDS per round
constexpr float kGridReachMargin = 4.0f * 0.01f;  // 4 * COLLISION_MARGIN (solver.h:23)

// Filled on the device after the radius select and the bounds, read by the host.
struct GridInfo
{
    float binnedRadius;
    unsigned bmin[3], bmax[3];  // order-preserving encodings of the binned centres' bounds
    int numBinned;
};

// Cell parameters, computed on the host and passed to the kernels by value.
struct GridParams
{
    float binnedRadius, cellSize;
    float bminx, bminy, bminz;
    int nx, ny, nz;
    int numBinned, numLarge;
};

__host__ __device__ __forceinline__ unsigned grid_ord(float f)
{
    unsigned u;
    std::memcpy(&u, &f, sizeof(u));
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__host__ __device__ __forceinline__ float grid_unord(unsigned u)
{
    const unsigned v = (u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u;
    float f;
    std::memcpy(&f, &v, sizeof(f));
    return f;
}

// maths.h min/max/clamp (maths.h:292-310): the first operand wins ties and unordered compares.
__device__ __forceinline__ float ref_min(float a, float b) { return a < b ? a : b; }
__device__ __forceinline__ float ref_max(float a, float b) { return a > b ? a : b; }
__device__ __forceinline__ float ref_clamp(float x, float a, float b) { return ref_max(a, ref_min(b, x)); }

// The radius of rank `rank` in ascending order (std::nth_element at rank, solver.cpp:296-301):
// four passes of 8 bits over the order-preserving encodings, a 256-bin histogram in LDS each.
// Also resets the bounds and the binned count for k_grid_bounds.
__global__ __launch_bounds__(1024) void k_grid_select(const float4 *sizeRad, int n, int rank, GridInfo *info)
{
    __shared__ unsigned hist[256];
    __shared__ unsigned shPrefix;
    __shared__ int shRank;
    unsigned prefix = 0u, mask = 0u;
    int k = rank;
    for (int shift = 24; shift >= 0; shift -= 8)
    {
        for (int b = threadIdx.x; b < 256; b += blockDim.x)
            hist[b] = 0u;
        __syncthreads();
        for (int i = threadIdx.x; i < n; i += blockDim.x)
        {
            const unsigned u = grid_ord(sizeRad[i].w);
            if ((u & mask) == prefix)
                atomicAdd(&hist[(u >> shift) & 255u], 1u);
        }
        __syncthreads();
        if (threadIdx.x == 0)
        {
            int acc = 0, b = 0;
            for (; b < 255; ++b)
            {
                if (acc + (int)hist[b] > k)
                    break;
                acc += (int)hist[b];
            }
            shPrefix = prefix | ((unsigned)b << shift);
            shRank = k - acc;
        }
        __syncthreads();
        prefix = shPrefix;
        k = shRank;
        mask |= 255u << shift;
        __syncthreads();
    }
    if (threadIdx.x == 0)
    {
        const float r = grid_unord(prefix);
        info->binnedRadius = ref_max(r, 1e-3f);
        for (int a = 0; a < 3; ++a)
        {
            info->bmin[a] = grid_ord(FLT_MAX);
            info->bmax[a] = grid_ord(-FLT_MAX);
        }
        info->numBinned = 0;
    }
}

// Bounds of the binned bodies' centres (solver.cpp:323-326) and their count.
__global__ __launch_bounds__(256) void k_grid_bounds(const float4 *pose, const float4 *sizeRad, int n, GridInfo *info)
{
    const float br = info->binnedRadius;
    float mn[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, mx[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    int cnt = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    {
        if (sizeRad[i].w > br)
            continue;
        const float4 c = pose[4 * i + P_POS];
        mn[0] = fminf(mn[0], c.x), mn[1] = fminf(mn[1], c.y), mn[2] = fminf(mn[2], c.z);
        mx[0] = fmaxf(mx[0], c.x), mx[1] = fmaxf(mx[1], c.y), mx[2] = fmaxf(mx[2], c.z);
        ++cnt;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
    {
        for (int a = 0; a < 3; ++a)
        {
            mn[a] = fminf(mn[a], __shfl_xor(mn[a], off, 64));
            mx[a] = fmaxf(mx[a], __shfl_xor(mx[a], off, 64));
        }
        cnt += __shfl_xor(cnt, off, 64);
    }
    if ((threadIdx.x & 63) == 0 && cnt > 0)
    {
        for (int a = 0; a < 3; ++a)
        {
            atomicMin(&info->bmin[a], grid_ord(mn[a]));
            atomicMax(&info->bmax[a], grid_ord(mx[a]));
        }
        atomicAdd(&info->numBinned, cnt);
    }
}

// Cell key per body (solver.cpp:346-352); large bodies get the sentinel, so the stable sort puts
// them after every binned body, in list order.
__global__ void k_grid_keys(const float4 *pose, const float4 *sizeRad, int n, GridParams g, unsigned *keys, int *vals)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    unsigned key = kGridSentinel;
    if (!(sizeRad[i].w > g.binnedRadius))
    {
        const float4 c = pose[4 * i + P_POS];
        const int ix = (int)ref_clamp((c.x - g.bminx) / g.cellSize, 0.0f, (float)(g.nx - 1));
        const int iy = (int)ref_clamp((c.y - g.bminy) / g.cellSize, 0.0f, (float)(g.ny - 1));
        const int iz = (int)ref_clamp((c.z - g.bminz) / g.cellSize, 0.0f, (float)(g.nz - 1));
        key = (unsigned)(ix + g.nx * (iy + g.ny * iz));
    }
    keys[i] = key;
    vals[i] = i;
}

// Sorted position e: binned entries (centre, radius; slack; first entry of each cell), large
// bodies (AABB with the slack in lo.w; centre and radius). order: the sorted body indices, or
// null for list order (all pairs, where every body is large).
__global__ void k_grid_gather(const int *order, const unsigned *keys, int n, GridParams g, Bodies B, float dt,
                              float4 *entryC, float *entrySlack, unsigned *cellStart, float4 *largeLo,
                              float4 *largeHi, float4 *largeC)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n)
        return;
    const int i = order ? order[e] : e;
    const float4 c4 = B.pose[4 * i + P_POS];
    const float r = B.sizeRad[i].w;
    const V3 v = ld3(B.vel[i]);
    const float slack = len(v) * dt;
    if (e < g.numBinned)
    {
        entryC[e] = make_float4(c4.x, c4.y, c4.z, r);
        entrySlack[e] = slack;
        if (e == 0 || keys[e - 1] != keys[e])
            cellStart[keys[e]] = (unsigned)e;
        return;
    }
    const int l = e - g.numBinned;
    V3 h{r, r, r};
    if ((int)B.shape[i].z == kShapeBox)
    {
        const float4 s4 = B.sizeRad[i];
        const V3 ex = V3{s4.x, s4.y, s4.z} * 0.5f;
        const Q4 q = ldq(B.pose[4 * i + P_ROT]);
        const V3 cx = rotate(q, V3{1.0f, 0.0f, 0.0f}), cy = rotate(q, V3{0.0f, 1.0f, 0.0f}),
                 cz = rotate(q, V3{0.0f, 0.0f, 1.0f});
        h = V3{fabsf(cx.x) * ex.x + fabsf(cy.x) * ex.y + fabsf(cz.x) * ex.z,
               fabsf(cx.y) * ex.x + fabsf(cy.y) * ex.y + fabsf(cz.y) * ex.z,
               fabsf(cx.z) * ex.x + fabsf(cy.z) * ex.y + fabsf(cz.z) * ex.z};
    }
    const V3 c{c4.x, c4.y, c4.z};
    largeLo[l] = st3(c - h, slack);
    largeHi[l] = st3(c + h, 0.0f);
    largeC[l] = make_float4(c4.x, c4.y, c4.z, r);
}

// Emit = false: hits per entry (hitCount) and the check count. Emit = true: the hits in emission
// order from hitStart, each with its new-manifold flag (no force joins the pair yet). Entries:
// the sorted binned bodies, then the large bodies (their large-large pairs). idx: the sorted body
// indices (null: list order).
template <bool Emit>
__global__ __launch_bounds__(256) void k_grid_pairs(GridParams g, const float4 *entryC, const float *entrySlack,
                                                   const unsigned *keys, const int *idx, const unsigned *cellStart,
                                                   const float4 *largeLo, const float4 *largeHi, const float4 *largeC,
                                                   int *hitCount, const int *hitStart, const unsigned long long *hash,
                                                   int hashMask, int2 *hits, int *hitNew, int maxPairs, Counters *ctr)
{
    __shared__ float4 sLo[kGridTile], sHi[kGridTile], sC[kGridTile];
    __shared__ int sIdx[kGridTile];
    const int numEntries = g.numBinned + g.numLarge;
    // the forward half of the 27-cell neighbourhood (solver.cpp:373-379), in the reference's order
    const int fwd[13][3] = {{1, 0, 0},  {-1, 1, 0}, {0, 1, 0}, {1, 1, 0}, {-1, -1, 1}, {0, -1, 1}, {1, -1, 1},
                            {-1, 0, 1}, {0, 0, 1},  {1, 0, 1}, {-1, 1, 1}, {0, 1, 1},  {1, 1, 1}};
    for (int base = blockIdx.x * blockDim.x; base < numEntries; base += gridDim.x * blockDim.x)
    {
        const int e = base + (int)threadIdx.x;
        const bool valid = e < numEntries;
        const bool binned = valid && e < g.numBinned;
        int nh = 0, checks = 0, out = 0;
        if (Emit && valid)
            out = hitStart[e];
        auto emit = [&](int a, int b) {
            if (Emit)
            {
                if (out < maxPairs)
                {
                    hits[out] = make_int2(a, b);
                    hitNew[out] = hash_contains(hash, hashMask, pair_key(a, b)) ? 0 : 1;
                }
                else
                    atomicOr(&ctr->overflow, 1);
                ++out;
            }
            ++nh;
        };
        float4 a = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        float aSlack = 0.0f;
        int ia = -1;
        if (binned)
        {
            a = entryC[e];
            aSlack = entrySlack[e];
            ia = idx[e];
            const unsigned key = keys[e];
            auto test = [&](int j) {
                ++checks;
                const float4 b = entryC[j];
                const V3 dp{a.x - b.x, a.y - b.y, a.z - b.z};
                const float r = a.w + b.w;
                if (!(dot(dp, dp) > r * r))
                    emit(ia, idx[j]);
            };
            for (int j = e + 1; j < g.numBinned && keys[j] == key; ++j)
                test(j);
            const int ix = (int)(key % (unsigned)g.nx);
            const unsigned t = key / (unsigned)g.nx;
            const int iy = (int)(t % (unsigned)g.ny), iz = (int)(t / (unsigned)g.ny);
            for (int f = 0; f < 13; ++f)
            {
                const int cx = ix + fwd[f][0], cy = iy + fwd[f][1], cz = iz + fwd[f][2];
                if (cx < 0 || cx >= g.nx || cy < 0 || cy >= g.ny || cz < 0 || cz >= g.nz)
                    continue;
                const unsigned nkey = (unsigned)(cx + g.nx * (cy + g.ny * cz));
                const unsigned s0 = cellStart[nkey];
                if (s0 == 0xFFFFFFFFu)
                    continue;
                for (int j = (int)s0; j < g.numBinned && keys[j] == nkey; ++j)
                    test(j);
            }
        }
        const int l0 = valid && !binned ? e - g.numBinned : -1;  // this entry's place in the large list
        float4 c0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        int i0 = -1;
        if (l0 >= 0)
        {
            c0 = largeC[l0];
            i0 = idx ? idx[e] : l0;
        }
        for (int t0 = 0; t0 < g.numLarge; t0 += kGridTile)
        {
            const int tn = min(kGridTile, g.numLarge - t0);
            __syncthreads();
            for (int q = threadIdx.x; q < tn; q += blockDim.x)
            {
                sLo[q] = largeLo[t0 + q];
                sHi[q] = largeHi[t0 + q];
                sC[q] = largeC[t0 + q];
                sIdx[q] = idx ? idx[g.numBinned + t0 + q] : t0 + q;
            }
            __syncthreads();
            if (binned)
            {
                for (int q = 0; q < tn; ++q)
                {
                    ++checks;
                    const float4 lo = sLo[q], hi = sHi[q];
                    const V3 dp{a.x - ref_clamp(a.x, lo.x, hi.x), a.y - ref_clamp(a.y, lo.y, hi.y),
                                a.z - ref_clamp(a.z, lo.z, hi.z)};
                    const float reach = a.w + aSlack + lo.w + kGridReachMargin;
                    if (!(dot(dp, dp) > reach * reach))
                        emit(ia, sIdx[q]);
                }
            }
            else if (l0 >= 0)
            {
                for (int q = max(0, l0 + 1 - t0); q < tn; ++q)
                {
                    ++checks;
                    const float4 b = sC[q];
                    const V3 dp{c0.x - b.x, c0.y - b.y, c0.z - b.z};
                    const float r = c0.w + b.w;
                    if (dot(dp, dp) <= r * r)
                        emit(i0, sIdx[q]);
                }
            }
        }
        if (!Emit)
        {
            if (valid)
                hitCount[e] = nh;
            block_add(checks, &ctr->pairChecks);
        }
    }
}

// The cell-start table back to empty (the reference restores it after each frame, solver.cpp:481-483).
__global__ void k_grid_clear(const unsigned *keys, int numBinned, unsigned *cellStart)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < numBinned && (e == 0 || keys[e - 1] != keys[e]))
        cellStart[keys[e]] = 0xFFFFFFFFu;
}

} // namespace avbd
