// avbd_external.hip -- external pair and contact sources (Solver::externalBroadphasePairs and the
// external contact table, solver.h:336-343, 361-365): the pair list that replaces the broadphase
// (broadphaseExternalPairs, solver.cpp:579-626) and the contact table Manifold::initialize reads
// instead of colliding (manifold.cpp:110-121).
//
// Entries: the valid external pairs in their order, then the valid contact-table entries (both
// bodies in range, ascending dense id), as (a, b, flags, 0). The host builds the list when the
// sources or the bodies change; the flags carry what the serial reference learns from the manifolds
// earlier entries of the same step created (a new manifold joins its two bodies at once, so a
// later entry for the same two bodies is a constrained hit):
//   kExtContact     the entry is a contact-table entry: a hit with no bounding-sphere test;
//   kExtAfterPair   a pair entry for the same two bodies comes before it (pair entries: they hit
//                   exactly when this one does, the test being symmetric in the two bodies;
//                   contact entries: they hit when the two bodies' spheres overlap);
//   kExtAfterEntry  a contact entry for the same two bodies comes before it (it hit).
// Per step: k_ext_hits writes each entry's hit flag and the check count, the hit flags are
// scanned, and k_ext_emit writes the hits in entry order with their new-manifold flags, which then
// go through the same ordered creation as the broadphase's hits.
//
// A fragment of avbd_kernels.hip (included inside its namespace, after the constrained-pair hash).

constexpr int kExtContact = 1, kExtAfterPair = 2, kExtAfterEntry = 4;

// generatePair's bounding-sphere test (solver.cpp:166-174)
__device__ __forceinline__ bool ext_spheres_overlap(const float4 *pose, const float4 *sizeRad, int a, int b)
{
    const float4 pa = pose[4 * a + P_POS], pb = pose[4 * b + P_POS];
    const V3 dp{pa.x - pb.x, pa.y - pb.y, pa.z - pb.z};
    const float r = sizeRad[a].w + sizeRad[b].w;
    return dot(dp, dp) <= r * r;
}

__global__ void k_ext_hits(const int4 *entries, int n, const float4 *pose, const float4 *sizeRad, int *hit,
                           Counters *ctr)
{
    if (blockIdx.x == 0 && threadIdx.x == 0)
        ctr->pairChecks = n;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x)
    {
        const int4 e = entries[k];
        hit[k] = (e.z & kExtContact) ? 1 : (ext_spheres_overlap(pose, sizeRad, e.x, e.y) ? 1 : 0);
    }
}

__global__ void k_ext_emit(const int4 *entries, int n, const float4 *pose, const float4 *sizeRad, const int *hit,
                           const int *hitStart, const unsigned long long *hash, int hashMask, int2 *hits, int *hitNew,
                           int maxPairs, Counters *ctr)
{
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x)
    {
        if (!hit[k])
            continue;
        const int4 e = entries[k];
        const int out = hitStart[k];
        if (out >= maxPairs)
        {
            atomicOr(&ctr->overflow, 1);
            continue;
        }
        bool earlier = (e.z & kExtAfterEntry) != 0;
        if (e.z & kExtAfterPair)
            earlier = earlier || (e.z & kExtContact) == 0 || ext_spheres_overlap(pose, sizeRad, e.x, e.y);
        hits[out] = make_int2(e.x, e.y);
        hitNew[out] = earlier || hash_contains(hash, hashMask, pair_key(e.x, e.y)) ? 0 : 1;
    }
}

// The contact table's lookup (Solver::findExternalManifoldContact, solver.cpp:976-987): an
// open-addressed table of ordered pair keys (a << 32 | b), each holding the index of the last table
// entry with that key; the manifold (a, b) takes the entry for (a, b), else the one for (b, a).
struct ExtTable
{
    const avbd_external_contact *entries = nullptr;
    const unsigned long long *keys = nullptr;
    const int *index = nullptr;
    int mask = -1;  // -1: no table
};

__host__ __device__ inline unsigned long long ext_key(int a, int b)
{
    return ((unsigned long long)(unsigned)a << 32) | (unsigned)b;
}

__host__ __device__ inline unsigned ext_slot(unsigned long long k, int mask)
{
    k ^= k >> 29;
    k *= 0xbf58476d1ce4e5b9ull;
    k ^= k >> 32;
    return (unsigned)k & (unsigned)mask;
}

__device__ inline int ext_find_key(const ExtTable &X, unsigned long long key)
{
    unsigned s = ext_slot(key, X.mask);
    for (int probe = 0; probe <= X.mask; ++probe)
    {
        const unsigned long long v = X.keys[s];
        if (v == key)
            return X.index[s];
        if (v == kHashEmpty)
            return -1;
        s = (s + 1) & (unsigned)X.mask;
    }
    return -1;
}

__device__ inline const avbd_external_contact *ext_find(const ExtTable &X, int a, int b)
{
    if (X.mask < 0)
        return nullptr;
    int k = ext_find_key(X, ext_key(a, b));
    if (k < 0)
        k = ext_find_key(X, ext_key(b, a));
    return k < 0 ? nullptr : X.entries + k;
}
