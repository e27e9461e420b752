// avbd_capi.hip -- C ABI of the resident MI355X AVBD step (declared in include/avbd_hip.h).
//
// Owns the device buffers of one solver instance, converts the POD state tables
// of the ABI to the structure-of-arrays layout the kernels use, and enqueues
// the per-step launch sequence on the handle's stream. Every count produced
// inside a step stays on the device, so a step is a fixed, host-sync-free
// sequence of launches.
#include "../../include/avbd_hip.h"
#include "avbd_kernels.hip"
#include "avbd_grid.hip"

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <unordered_set>
#include <vector>

using namespace avbd;

// Staging of the allocation-free snapshot (avbd_snapshot_avbs_static): the 64-byte AVBS body
// records of at most kStaticSnapshotBodies bodies (4 MiB). A file-scope device array, so each GPU
// holds its own copy, at its own address, from the module's load until the process ends. In host
// code `scratch` names a host-side stand-in, never the device memory: host code uses only the
// address hipGetSymbolAddress returns for the current device (g_scratchPtr, bound in avbd_create).
constexpr int kStaticSnapshotBodies = 65536;
static __device__ __align__(256) uint4 scratch[4 * kStaticSnapshotBodies];

namespace
{

thread_local std::string g_lastError;

int fail(int code, const std::string &msg)
{
    g_lastError = msg;
    return code;
}

#define HIP_TRY(expr)                                                                                      \
    do                                                                                                     \
    {                                                                                                      \
        hipError_t e_ = (expr);                                                                            \
        if (e_ != hipSuccess)                                                                              \
            return fail(AVBD_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));                  \
    } while (0)

int ceil_div(long long a, long long b) { return (int)((a + b - 1) / b); }

// Per device ordinal: the device address of that GPU's copy of `scratch` (null until bound).
constexpr int kMaxStaticDevices = 64;
std::mutex g_scratchMutex;
uint4 *g_scratchPtr[kMaxStaticDevices] = {};

// With `device` current: looks up the address of its copy of `scratch` once and keeps it.
int bind_static_scratch(int device)
{
    std::lock_guard<std::mutex> lock(g_scratchMutex);
    if (device < 0 || device >= kMaxStaticDevices || g_scratchPtr[device])
        return AVBD_OK;
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fail(AVBD_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratchPtr[device] = (uint4 *)p;
    return AVBD_OK;
}

uint4 *static_scratch(int device)
{
    std::lock_guard<std::mutex> lock(g_scratchMutex);
    return device >= 0 && device < kMaxStaticDevices ? g_scratchPtr[device] : nullptr;
}

} // namespace

struct avbd_handle
{
    int device = 0;
    hipStream_t stream = nullptr;
    avbd_config cfg{};
    Params P{};
    int nb = 0, nj = 0, ns = 0;
    int maxBodies = 0, maxJoints = 0, maxSprings = 0, maxManifolds = 0, maxPairs = 0, maxChunks = 0, maxAdj = 0;
    int hashSize = 0;
    int coopColor = 0;
    int4 *padStamps = nullptr;   // stamps in task order, kPadStampsMax per task at most
    unsigned long long *scanStatus = nullptr;
    unsigned *scanCtl = nullptr;
    int scanMaxTiles = 0;
    int numRound = 0, numSpheres = 0; // bodies of the uploaded state that are not boxes / are spheres
    int *avbsRank = nullptr;          // per body: its rank in ascending dense id (AVBS snapshot order)
    uint4 *avbsBuf = nullptr;         // AVBS body records (64 B each), allocated on first use
    int2 *ignorePairs = nullptr;      // IgnoreCollision pairs (list indices)
    int numIgnore = 0, ignoreCap = 0;
    int *ignoreDeg = nullptr;         // per body: its IgnoreCollision forces
    int injectFault = 0;
    Bodies B{};
    Manifolds cur{}, tmp{};
    Joint *joints = nullptr;
    Spring *springs = nullptr;
    Scratch S{};
    Counters *ctr = nullptr;
    GridBarrier *gb = nullptr;
    int *sweepEnd = nullptr;
    int *radixHistScan = nullptr;
    int *adjTotal = nullptr;
    int maxStamps = 0;
    // pinned host copy of the counters and colour offsets, read once per step before the
    // iteration launches (their grid sizes depend on the colour and contact counts)
    Counters *hostCtr = nullptr;
    int *hostOffsets = nullptr;
    std::vector<void *> allocs;
    hipEvent_t ev[8] = {};
    // timing ring: per step, events at step start / before iterate / after iterate / step end
    static constexpr int kRing = 2048;
    // per step: 0 step start, 1 iterate start, 2 iterate end, 3 step end, 4-6 the last
    // iteration's primal start / dual start / dual end
    static constexpr int kSlots = 7;
    std::vector<hipEvent_t> ring;
    std::vector<int> ringIters;                    // iterations marked in each ring step
    std::vector<int> ringPrimalLaunches;           // primal launches in each ring step
    std::vector<int> ringPrimalIters;              // solver iterations those launches cover
    long long primalLaunches = 0, dualLaunches = 0;
    long long ringSteps = 0;   // steps recorded into the ring since the last reset
    avbd_stats stats{};
    long long stepsTaken = 0;
    bool timing = false;
    // batched worlds: world w owns bodies [worldStart[w], worldStart[w + 1])
    int numWorlds = 1;
    int *worldStart = nullptr;
    std::vector<int> worldStartHost;
    int maxWorldBodies = 0;  // largest world of the batch (sizes k_color_worlds' LDS)
    int npPolyCap = kNpPolyCap;  // clip-polygon vertices the box-only narrowphase keeps in LDS
    // launch shapes of the iteration kernels, and the iteration graph built from them
    // default: adaptive lanes per body in 64-thread blocks (the fastest shape measured on the settled
    // 50,000-box pile, the batched worlds and the 1,089-body projectile workload)
    avbd_launch_config launch{0, 64, 256, 0, 0, 1};
    // broadphase mode (Solver::broadphaseMode, solver.h:366) and the grid's buffers, allocated
    // when a mode other than sweep and prune is first set
    int broadphaseMode = AVBD_BROADPHASE_SAP;
    GridInfo *gridInfo = nullptr, *hostGridInfo = nullptr;
    unsigned *cellStart = nullptr;  // kGridMaxCells entries, 0xFFFFFFFF when empty
    float4 *largeLo = nullptr, *largeHi = nullptr, *largeC = nullptr;
    // external pair and contact sources (avbd_set_external_pairs / _contacts): the host copies, the
    // bodies' dense ids (from the last upload) and the device lists built from them before the next
    // step once either changed (extDirty); these buffers are freed by avbd_destroy
    bool extPairsOn = false, extDirty = false;
    std::vector<int2> extPairs;
    std::vector<avbd_external_contact> extContacts;
    std::vector<int> hostDenseIds;
    int numExtEntries = 0, extEntryCap = 0, extTableCap = 0, extMask = -1;
    int4 *extEntries = nullptr;            // (a, b, flags, 0): valid pairs, then valid table entries
    int *extHit = nullptr, *extHitStart = nullptr;
    avbd_external_contact *extTable = nullptr;
    unsigned long long *extKeys = nullptr;  // ordered pair keys of the table, kHashEmpty when free
    int *extIdx = nullptr;                  // per key: the last table entry with that key
    int *dfDone = nullptr;  // flow primal: iterations each solved body has finished in this step
    int numCUs = 0;
    int flowBlocks[3][3] = {};  // flow primal grid per (lanes, threads) shape: every wave resident
    hipGraph_t graph = nullptr;
    hipGraphExec_t graphExec = nullptr;
    std::vector<hipGraphNode_t> graphPrimal;
    hipGraphNode_t graphDual = nullptr;
    avbd_launch_config graphLaunch{};
    // adjoint: recording outputs of the narrowphase/compaction (null unless a step is being taped),
    // the tape of the step being reversed and the rest of the adjoint state (avbd_adjoint.hip)
    int *recMatch = nullptr, *recStick = nullptr, *recSrc = nullptr;
    float4 *tapePose = nullptr, *tapeContact = nullptr, *tapeSys = nullptr;
    unsigned char *tapeStampFlags = nullptr, *tapeDualFlags = nullptr;
    Joint *tapeJoint = nullptr;  // joint states after init (slot 0) and after each iteration's dual
    struct AdjointState *adj = nullptr;
    int lastNumManifolds = 0;
    // FP64 body inputs and outputs in host or device memory (avbd_set_body_inputs,
    // avbd_get_body_state_f64): a staging buffer, and per body {dynamic, moment / mass} of the
    // state as last uploaded (taken at the first avbd_set_body_inputs after an upload)
    double *io64 = nullptr, *massRatio = nullptr;
    bool ratioValid = false;
    unsigned uploadEpoch = 0;  // counts uploads: a recorded forward belongs to one of them

    template <class T>
    int alloc(T *&p, size_t count)
    {
        void *q = nullptr;
        size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
        hipError_t e = hipMalloc(&q, bytes);
        if (e != hipSuccess)
            return fail(AVBD_ERR_HIP, std::string("hipMalloc failed: ") + hipGetErrorString(e));
        allocs.push_back(q);
        p = (T *)q;
        return AVBD_OK;
    }
};

namespace
{

Params to_params(const avbd_params &p)
{
    Params o;
    o.dt = p.dt;
    o.gravity = p.gravity;
    o.alpha = p.alpha;
    o.betaLin = p.beta_lin;
    o.betaAng = p.beta_ang;
    o.gamma = p.gamma;
    o.iterations = p.iterations;
    return o;
}

int alloc_manifolds(avbd_handle *h, Manifolds &m, int cap)
{
    int r;
    size_t cs = (size_t)cap * kMaxContacts;
    if ((r = h->alloc(m.ab, cap)) || (r = h->alloc(m.serial, cap)) || (r = h->alloc(m.count, cap)) ||
        (r = h->alloc(m.feature, cs)) || (r = h->alloc(m.cr, cs * kCR)))
        return r;
    return AVBD_OK;
}

int grid_for(int n) { return std::max(1, std::min(ceil_div(n, 256), 4096)); }

using PrimalKernel = void (*)(Bodies, Manifolds, const Joint *, const Spring *, const int4 *, const int4 *, int, int,
                              Params, float4 *, unsigned char *, const int4 *, int, int);

// The primal kernel's launch shapes: lanes per body x threads per block.
PrimalKernel primal_kernel(int lanes, int threads)
{
#define AVBD_PRIMAL_SHAPE(l, t)                                                                            \
    if (lanes == l && threads == t)                                                                        \
        return &k_primal_color<l, t>;
    AVBD_PRIMAL_SHAPE(8, 64)
    AVBD_PRIMAL_SHAPE(8, 128)
    AVBD_PRIMAL_SHAPE(8, 256)
    AVBD_PRIMAL_SHAPE(16, 64)
    AVBD_PRIMAL_SHAPE(16, 128)
    AVBD_PRIMAL_SHAPE(16, 256)
    AVBD_PRIMAL_SHAPE(32, 64)
    AVBD_PRIMAL_SHAPE(32, 128)
    AVBD_PRIMAL_SHAPE(32, 256)
#undef AVBD_PRIMAL_SHAPE
    return nullptr;
}

using FlowKernel = void (*)(Bodies, Manifolds, const int4 *, const int *, int, const int *, int *, Params,
                            const int4 *, int, Counters *);

FlowKernel flow_kernel(int lanes, int threads)
{
#define AVBD_DATAFLOW_SHAPE(l, t)                                                                          \
    if (lanes == l && threads == t)                                                                        \
        return &k_primal_flow<l, t>;
    AVBD_DATAFLOW_SHAPE(8, 64)
    AVBD_DATAFLOW_SHAPE(8, 128)
    AVBD_DATAFLOW_SHAPE(8, 256)
    AVBD_DATAFLOW_SHAPE(16, 64)
    AVBD_DATAFLOW_SHAPE(16, 128)
    AVBD_DATAFLOW_SHAPE(16, 256)
    AVBD_DATAFLOW_SHAPE(32, 64)
    AVBD_DATAFLOW_SHAPE(32, 128)
    AVBD_DATAFLOW_SHAPE(32, 256)
#undef AVBD_DATAFLOW_SHAPE
    return nullptr;
}

using AdaptiveKernel = void (*)(Bodies, Manifolds, const Joint *, const Spring *, const int4 *, const int4 *, int4,
                                int, int4, Params, float4 *, unsigned char *, const int4 *, int, int);

// primal_lanes = 0: lanes per body chosen per body from its stamp count (k_primal_adaptive).
AdaptiveKernel adaptive_kernel(int threads)
{
    return threads == 64 ? &k_primal_adaptive<64>
           : threads == 128 ? &k_primal_adaptive<128>
           : threads == 256 ? &k_primal_adaptive<256>
                            : nullptr;
}

bool valid_launch(const avbd_launch_config &c)
{
    auto threads = [](int t) { return t == 64 || t == 128 || t == 256; };
    const bool primal = c.primal_lanes == 0 ? adaptive_kernel(c.primal_threads) != nullptr
                                            : primal_kernel(c.primal_lanes, c.primal_threads) != nullptr;
    return primal && threads(c.dual_threads) &&
           (c.use_graph == 0 || c.use_graph == 1) && (c.dataflow == 0 || c.dataflow == 1) &&
           (c.lazy_dual == 0 || c.lazy_dual == 1);
}

// Kernel argument blocks (hipLaunchKernel / graph kernel nodes take an array of argument pointers).
struct PrimalArgs
{
    Bodies B;
    Manifolds M;
    const Joint *joints;
    const Spring *springs;
    const int4 *stamps;
    const int4 *tasks;
    int begin, end;
    Params P;
    float4 *sysTape;
    unsigned char *flagTape;
    const int4 *pad;
    int padShift;
    int lazyDual;
    void *ptr[14];
    PrimalArgs(const avbd_handle *h, int b, int e, float4 *sys = nullptr, unsigned char *flags = nullptr,
               const int4 *pad_ = nullptr, int padShift_ = 0, int lazyDual_ = 0);
};

// Adaptive-lanes launch of one colour: its segments (from the per-step readback of the colour
// offsets) and the blocks each one takes.
struct AdaptiveArgs
{
    Bodies B;
    Manifolds M;
    const Joint *joints;
    const Spring *springs;
    const int4 *stamps;
    const int4 *tasks;
    int4 seg;
    int end;
    int4 blk;
    Params P;
    float4 *sysTape;
    unsigned char *flagTape;
    const int4 *pad;
    int padShift;
    int lazyDual;
    void *ptr[15];
    AdaptiveArgs(const avbd_handle *h, int color, int threads, float4 *sys = nullptr, unsigned char *flags = nullptr,
                 const int4 *pad_ = nullptr, int padShift_ = 0, int lazyDual_ = 0);
    int grid() const { return blk.w; }
};

struct DualArgs
{
    Bodies B;
    Manifolds M;
    Joint *joints;
    int nj;
    const int4 *contactRec;
    int nc;
    Params P;
    unsigned char *flagTape;
    int staticRepeat;
    DualTape tape;
    void *ptr[10];
    DualArgs(const avbd_handle *h, int nc_, unsigned char *flags = nullptr, int nj_ = -1, int staticRepeat_ = 1,
             DualTape tape_ = DualTape());
};

// Colouring runs as one launch of a few large blocks (at most 64, one per compute unit, every
// one of them resident at once on 256 CUs): every round ends in a grid barrier whose cost grows
// with the block count, and whose spin limit turns a block that never arrives into an error.
// The launch is an ordinary one: a cooperative launch (hipLaunchCooperativeKernel) is refused
// with hipErrorIllegalState in a second process while another process uses cooperative launches
// on the same GPU, which would keep two programs from stepping on one device.
constexpr int kColorThreads = 1024;
constexpr int kColorSmallRounds = 6;  // round limit of the fixpoint for a world of at most kWorldSapMax bodies
constexpr long long kMaxSweepItems = 1LL << 26;  // sweep work-item capacity for the all-pairs worst case

// Exclusive scan with the count either on the device (nptr) or given (nconst).
// total (may be null) receives the sum, clamped to cap (a clamp sets capBit in the overflow flags).
void scan(avbd_handle *h, const int *in, int *out, const int *nptr, int nconst, int *total, int cap = INT_MAX,
          int capBit = 0)
{
    hipStream_t s = h->stream;
    if (!nptr && nconst <= kScanSmallMax)
    {
        k_scan_small<<<1, 1024, 0, s>>>(in, out, nconst, total, cap, &h->ctr->overflow, capBit);
        return;
    }
    k_scan_lookback<<<kLbBlocks, kLbThreads, 0, s>>>(in, out, nptr, nconst, total, h->scanStatus, h->scanCtl,
                                                    h->scanMaxTiles, &h->ctr->barrierTimeout, cap,
                                                    &h->ctr->overflow, capBit);
}

// Stable LSD radix sort of (key, value); `passes` 8-bit digits from bit 0. Result in keys/vals.
// Up to this many tiles, each scatter block sums the histogram itself (no scan launch).
constexpr int kRadixFusedTiles = 64;

void radix_sort(avbd_handle *h, unsigned *&keys, int *&vals, unsigned *&keysAlt, int *&valsAlt, int n, int passes)
{
    int tiles = ceil_div(n, kRadixTile);
    for (int p = 0; p < passes; ++p)
    {
        int shift = 8 * p;
        k_radix_hist<<<tiles, kRadixThreads, 0, h->stream>>>(keys, n, shift, h->S.radixHist, tiles);
        if (tiles <= kRadixFusedTiles)
            k_radix_scatter<<<tiles, kRadixThreads, 0, h->stream>>>(keys, vals, n, shift, h->S.radixHist, tiles, keysAlt,
                                                          valsAlt, 1);
        else
        {
            scan(h, h->S.radixHist, h->radixHistScan, nullptr, 256 * tiles, nullptr);
            k_radix_scatter<<<tiles, kRadixThreads, 0, h->stream>>>(keys, vals, n, shift, h->radixHistScan, tiles, keysAlt,
                                                          valsAlt, 0);
        }
        std::swap(keys, keysAlt);
        std::swap(vals, valsAlt);
    }
}



// Bodies at `dev` -> host records (fields the body table carries).
constexpr int kBodyRows = 12; // the rows bodies_to_host reads

void bodies_to_host(const std::vector<float4> *v, const std::vector<float> &fr, const std::vector<int> &ids, int n,
                    avbd_body *out)
{
    // v[kBodyRows]: pos rot initLin initAng inerLin inerAng vel angVel prevVel massMoment sizeRad shape
    for (int i = 0; i < n; ++i)
    {
        avbd_body &b = out[i];
        std::memset(&b, 0, sizeof(b));
        auto c3 = [&](float *d, const float4 &s) { d[0] = s.x; d[1] = s.y; d[2] = s.z; };
        auto c4 = [&](float *d, const float4 &s) { d[0] = s.x; d[1] = s.y; d[2] = s.z; d[3] = s.w; };
        c3(b.pos, v[0][i]);
        c4(b.rot, v[1][i]);
        c3(b.initial_lin, v[2][i]);
        c4(b.initial_ang, v[3][i]);
        c3(b.inertial_lin, v[4][i]);
        c4(b.inertial_ang, v[5][i]);
        c3(b.vel, v[6][i]);
        c3(b.ang_vel, v[7][i]);
        c3(b.prev_vel, v[8][i]);
        b.mass = v[9][i].x;
        b.moment[0] = v[9][i].y;
        b.moment[1] = v[9][i].z;
        b.moment[2] = v[9][i].w;
        c3(b.size, v[10][i]);
        b.radius = v[10][i].w;
        b.friction = fr[i];
        b.shape_radius = v[11][i].x;
        b.half_length = v[11][i].y;
        b.shape = (int32_t)v[11][i].z;
        b.dense_id = ids[i];
    }
}

int check_device_errors(avbd_handle *h)
{
    Counters c;
    HIP_TRY(hipMemcpyAsync(&c, h->ctr, sizeof(Counters), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    if (c.barrierTimeout)
        return fail(AVBD_ERR_HIP, (c.barrierTimeout & 2) ? "a scan's look-back timed out"
                                                           : "a grid barrier or dataflow wait timed out");
    if (c.overflow)
    {
        const char *what = (c.overflow & 1) ? "broadphase pair buffer (max_pairs)"
                           : (c.overflow & 2) ? "manifold buffer (max_manifolds)"
                           : (c.overflow & 4) ? "adjacency buffer"
                           : (c.overflow & 8) ? "colour count (more than 254 colours)"
                           : (c.overflow & 16) ? "constrained-pair hash"
                           : (c.overflow & 64) ? "primal stamp buffer"
                           : (c.overflow & 128) ? "per-world LDS (a world larger than the batch's largest)"
                                                : "sweep work-item buffer";
        return fail((c.overflow & 8) ? AVBD_ERR_UNSUPPORTED : AVBD_ERR_CAPACITY,
                    std::string("capacity exceeded: ") + what);
    }
    return AVBD_OK;
}

// Uniform-grid or all-pairs broadphase of a single world (solver.cpp:176-185, 284-486): the hits in
// the reference's emission order into S.hits / S.hitNew, their count into ctr->numHits and the
// checks into ctr->pairChecks, like the sweep-and-prune launches. The grid's cell size needs
// the host: the radius select and the bounds are read back, and the cell size, dimensions and
// the coarsening loop (with the host's cbrtf) are computed here exactly as the reference does.
int enqueue_grid_pairs(avbd_handle *h)
{
    hipStream_t s = h->stream;
    const int N = h->nb;
    Scratch &S = h->S;
    Counters *ctr = h->ctr;
    const bool grid = h->broadphaseMode == AVBD_BROADPHASE_GRID;
    GridParams g{};
    unsigned *keys = S.keys, *keysAlt = S.keysAlt;
    int *vals = S.vals, *valsAlt = S.valsAlt;
    if (grid)
    {
        const int rank = (int)std::min<long long>((long long)N - 1, (long long)N * 95 / 100);
        k_grid_select<<<1, 1024, 0, s>>>(h->B.sizeRad, N, rank, h->gridInfo);
        k_grid_bounds<<<std::max(1, std::min(ceil_div(N, 256), 1024)), 256, 0, s>>>(h->B.pose, h->B.sizeRad, N,
                                                                                   h->gridInfo);
        HIP_TRY(hipMemcpyAsync(h->hostGridInfo, h->gridInfo, sizeof(GridInfo), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        const GridInfo &info = *h->hostGridInfo;
        g.binnedRadius = info.binnedRadius;
        float cellSize = info.binnedRadius * 2.0f;
        int nx = 1, ny = 1, nz = 1;
        g.numBinned = info.numBinned;
        if (g.numBinned > 0)
        {
            const float lo[3] = {grid_unord(info.bmin[0]), grid_unord(info.bmin[1]), grid_unord(info.bmin[2])};
            const float hi[3] = {grid_unord(info.bmax[0]), grid_unord(info.bmax[1]), grid_unord(info.bmax[2])};
            const float ex = hi[0] - lo[0], ey = hi[1] - lo[1], ez = hi[2] - lo[2];
            auto dims = [&](float cs) -> unsigned long long {
                nx = (int)(ex / cs) + 1;
                ny = (int)(ey / cs) + 1;
                nz = (int)(ez / cs) + 1;
                return (unsigned long long)nx * (unsigned long long)ny * (unsigned long long)nz;
            };
            unsigned long long cells = dims(cellSize);
            while (cells > kGridMaxCells)
            {
                const float grow = std::cbrt((float)cells / (float)kGridMaxCells);
                cellSize *= grow > 1.1f ? grow : 1.1f;
                cells = dims(cellSize);
            }
            g.bminx = lo[0];
            g.bminy = lo[1];
            g.bminz = lo[2];
        }
        g.cellSize = cellSize;
        g.nx = nx;
        g.ny = ny;
        g.nz = nz;
        g.numLarge = N - g.numBinned;
        k_grid_keys<<<ceil_div(N, 256), 256, 0, s>>>(h->B.pose, h->B.sizeRad, N, g, keys, vals);
        radix_sort(h, keys, vals, keysAlt, valsAlt, N, 3);
        k_grid_gather<<<ceil_div(N, 256), 256, 0, s>>>(vals, keys, N, g, h->B, h->P.dt, S.sPos, S.sLo, h->cellStart,
                                                      h->largeLo, h->largeHi, h->largeC);
    }
    else
    {
        // all pairs: every body is a large body, in list order, tested by the bounding spheres
        g.numBinned = 0;
        g.numLarge = N;
        g.nx = g.ny = g.nz = 1;
        k_grid_gather<<<ceil_div(N, 256), 256, 0, s>>>(nullptr, keys, N, g, h->B, h->P.dt, S.sPos, S.sLo, h->cellStart,
                                                      h->largeLo, h->largeHi, h->largeC);
    }
    const int *order = grid ? vals : nullptr;
    const int blocks = std::max(1, std::min(ceil_div(N, 256), 8192));
    k_grid_pairs<false><<<blocks, 256, 0, s>>>(g, S.sPos, S.sLo, keys, order, h->cellStart, h->largeLo, h->largeHi,
                                               h->largeC, S.hitCount, nullptr, nullptr, 0, nullptr, nullptr, 0, ctr);
    scan(h, S.hitCount, S.hitStart, nullptr, N, &ctr->numHits, h->maxPairs, 1);
    k_fill_u64<<<grid_for(h->hashSize), 256, 0, s>>>(S.hash, h->hashSize, ~0ull);
    k_hash_insert<<<1024, 256, 0, s>>>(h->cur.ab, ctr, h->joints, h->nj, h->springs, h->ns, h->ignorePairs, h->numIgnore,
                                       S.hash, S.hashMask, ctr);
    k_grid_pairs<true><<<blocks, 256, 0, s>>>(g, S.sPos, S.sLo, keys, order, h->cellStart, h->largeLo, h->largeHi,
                                              h->largeC, nullptr, S.hitStart, S.hash, S.hashMask, S.hits, S.hitNew,
                                              h->maxPairs, ctr);
    if (grid && g.numBinned > 0)
        k_grid_clear<<<ceil_div(g.numBinned, 256), 256, 0, s>>>(keys, g.numBinned, h->cellStart);
    HIP_TRY(hipGetLastError());
    return AVBD_OK;
}

// (Re)allocates an external-source buffer to hold `count` items when it holds fewer.
template <class T>
int ext_buffer(T *&p, int &cap, int count)
{
    if (p && cap >= count)
        return AVBD_OK;
    if (p)
        (void)hipFree(p);
    p = nullptr;
    cap = std::max(count, 1);
    hipError_t e = hipMalloc((void **)&p, sizeof(T) * (size_t)cap);
    if (e != hipSuccess)
    {
        p = nullptr;
        cap = 0;
        return fail(AVBD_ERR_HIP, std::string("hipMalloc failed: ") + hipGetErrorString(e));
    }
    return AVBD_OK;
}

// Device lists of the external sources (avbd_external.hip) from the host copies and the bodies'
// dense ids: the entries of broadphaseExternalPairs (solver.cpp:579-626) with their flags, and the
// contact table with its ordered-key index, where a later entry replaces an earlier one with the
// same key (Solver::setExternalManifoldContacts, solver.cpp:958-967). Table entries naming a body
// outside the state never match a manifold and are left out of the index.
int ext_rebuild(avbd_handle *h)
{
    const int N = h->nb;
    auto inRange = [&](int i) { return i >= 0 && i < N; };
    auto unordered = [](int a, int b) {
        return ((unsigned long long)(unsigned)std::min(a, b) << 32) | (unsigned)std::max(a, b);
    };
    std::vector<int4> entries;
    if (h->extPairsOn)
    {
        std::unordered_set<unsigned long long> pairSeen, entrySeen;
        for (const int2 &p : h->extPairs)
        {
            if (!inRange(p.x) || !inRange(p.y) || p.x == p.y)
                continue;
            const unsigned long long k = unordered(p.x, p.y);
            entries.push_back(make_int4(p.x, p.y, pairSeen.count(k) ? kExtAfterPair : 0, 0));
            pairSeen.insert(k);
        }
        for (const avbd_external_contact &e : h->extContacts)
        {
            const int a = e.body_a, b = e.body_b;
            if (!inRange(a) || !inRange(b) || (unsigned)h->hostDenseIds[a] >= (unsigned)h->hostDenseIds[b])
                continue;
            const unsigned long long k = unordered(a, b);
            const int flags =
                kExtContact | (pairSeen.count(k) ? kExtAfterPair : 0) | (entrySeen.count(k) ? kExtAfterEntry : 0);
            entries.push_back(make_int4(a, b, flags, 0));
            entrySeen.insert(k);
        }
    }
    if ((long long)entries.size() > h->maxPairs)
        return fail(AVBD_ERR_CAPACITY, "external pairs and contact entries exceed the handle's pair capacity");
    const int n = (int)entries.size();
    int hitCap = h->extEntryCap, startCap = h->extEntryCap, r;
    if ((r = ext_buffer(h->extEntries, h->extEntryCap, n)) || (r = ext_buffer(h->extHit, hitCap, n)) ||
        (r = ext_buffer(h->extHitStart, startCap, n)))
        return r;
    hipStream_t s = h->stream;
    if (n)
        HIP_TRY(hipMemcpyAsync(h->extEntries, entries.data(), sizeof(int4) * n, hipMemcpyHostToDevice, s));
    h->numExtEntries = n;
    const int nt = (int)h->extContacts.size();
    h->extMask = -1;
    if (nt > 0)
    {
        int size = 16;
        while (size < 2 * nt)
            size *= 2;
        std::vector<unsigned long long> keys(size, ~0ull);
        std::vector<int> idx(size, -1);
        for (int k = 0; k < nt; ++k)
        {
            const avbd_external_contact &e = h->extContacts[k];
            if (!inRange(e.body_a) || !inRange(e.body_b))
                continue;
            const unsigned long long key = ext_key(e.body_a, e.body_b);
            unsigned slot = ext_slot(key, size - 1);
            while (keys[slot] != ~0ull && keys[slot] != key)
                slot = (slot + 1) & (unsigned)(size - 1);
            keys[slot] = key;
            idx[slot] = k;
        }
        int tableCap = h->extTableCap, keyCap = h->extTableCap, idxCap = h->extTableCap;
        if ((r = ext_buffer(h->extTable, tableCap, nt)) || (r = ext_buffer(h->extKeys, keyCap, size)) ||
            (r = ext_buffer(h->extIdx, idxCap, size)))
            return r;
        h->extTableCap = std::min({tableCap, keyCap, idxCap});
        HIP_TRY(hipMemcpyAsync(h->extTable, h->extContacts.data(), sizeof(avbd_external_contact) * nt,
                               hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemcpyAsync(h->extKeys, keys.data(), sizeof(unsigned long long) * size, hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemcpyAsync(h->extIdx, idx.data(), sizeof(int) * size, hipMemcpyHostToDevice, s));
        h->extMask = size - 1;
    }
    // the host vectors above are copied before they go out of scope
    HIP_TRY(hipStreamSynchronize(s));
    h->extDirty = false;
    return AVBD_OK;
}

// External pairs instead of the broadphase (broadphaseExternalPairs, solver.cpp:579-626): hit flags,
// their scan, then the hits in entry order with their new-manifold flags, into S.hits / S.hitNew and
// ctr->numHits like the other broadphases.
int enqueue_external_pairs(avbd_handle *h)
{
    hipStream_t s = h->stream;
    Scratch &S = h->S;
    Counters *ctr = h->ctr;
    const int n = h->numExtEntries;
    const int blocks = std::max(1, std::min(ceil_div(n, 256), 1024));
    k_ext_hits<<<blocks, 256, 0, s>>>(h->extEntries, n, h->B.pose, h->B.sizeRad, h->extHit, ctr);
    scan(h, h->extHit, h->extHitStart, nullptr, n, &ctr->numHits, h->maxPairs, 1);
    k_fill_u64<<<grid_for(h->hashSize), 256, 0, s>>>(S.hash, h->hashSize, ~0ull);
    k_hash_insert<<<1024, 256, 0, s>>>(h->cur.ab, ctr, h->joints, h->nj, h->springs, h->ns, h->ignorePairs, h->numIgnore,
                                       S.hash, S.hashMask, ctr);
    k_ext_emit<<<blocks, 256, 0, s>>>(h->extEntries, n, h->B.pose, h->B.sizeRad, h->extHit, h->extHitStart, S.hash,
                                      S.hashMask, S.hits, S.hitNew, h->maxPairs, ctr);
    HIP_TRY(hipGetLastError());
    return AVBD_OK;
}

ExtTable ext_table(const avbd_handle *h)
{
    ExtTable X;
    if (h->extMask >= 0)
    {
        X.entries = h->extTable;
        X.keys = h->extKeys;
        X.index = h->extIdx;
        X.mask = h->extMask;
    }
    return X;
}

int enqueue_prepare(avbd_handle *h)
{
    hipStream_t s = h->stream;
    const int N = h->nb;
    Scratch &S = h->S;
    Counters *ctr = h->ctr;
    const int G = 1024;
    if (h->timing)
    {
        HIP_TRY(hipEventRecord(h->ev[0], s));
        HIP_TRY(hipEventRecord(h->ring[avbd_handle::kSlots * (h->ringSteps % avbd_handle::kRing) + 0], s));
    }
    k_step_begin<<<1, 1, 0, s>>>(ctr, h->gb);

    // ---- broadphase: sweep and prune (solver.cpp:488-577), or the grid / all pairs, or the
    // external pairs (solver.cpp:1075-1076) ----
    if (h->extDirty)
    {
        int r = ext_rebuild(h);
        if (r)
            return r;
    }
    if (N >= 2 && h->extPairsOn)
    {
        int r = enqueue_external_pairs(h);
        if (r)
            return r;
        scan(h, S.hitNew, S.hitRank, &ctr->numHits, 0, &ctr->numCreated);
        k_create<<<G, 256, 0, s>>>(S.hits, S.hitNew, S.hitRank, ctr, h->cur, h->maxManifolds, h->maxPairs);
    }
    else if (N >= 2 && h->broadphaseMode != AVBD_BROADPHASE_SAP)
    {
        int r = enqueue_grid_pairs(h);
        if (r)
            return r;
        scan(h, S.hitNew, S.hitRank, &ctr->numHits, 0, &ctr->numCreated);
        k_create<<<G, 256, 0, s>>>(S.hits, S.hitNew, S.hitRank, ctr, h->cur, h->maxManifolds, h->maxPairs);
    }
    else if (N >= 2)
    {
        if (h->numWorlds > 1)
        {
            // batched worlds: per-world axis, sort and sweep extents in one launch
            k_world_sap<<<h->numWorlds, 256, 0, s>>>(h->B.pose, h->B.sizeRad, h->worldStart, S.sLo, S.sHi, S.sPos,
                                                     S.sIdx, h->sweepEnd, S.chunkCount, ctr);
        }
        else
        {
            k_axis_partial<<<64, 256, 0, s>>>(h->B.pose, N, S.axisSums);
            k_axis_decide<<<1, 64, 0, s>>>(S.axisSums, 64, N, ctr);
            k_axis_serial<<<1, 64, 0, s>>>(h->B.pose, N, ctr);
            k_sap_keys<<<ceil_div(N, 256), 256, 0, s>>>(h->B.pose, h->B.sizeRad, N, ctr, S.keys, S.vals);
            unsigned *keys = S.keys, *keysAlt = S.keysAlt;
            int *vals = S.vals, *valsAlt = S.valsAlt;
            radix_sort(h, keys, vals, keysAlt, valsAlt, N, 4);
            k_sap_gather<<<ceil_div(N, 256), 256, 0, s>>>(vals, h->B.pose, h->B.sizeRad, N, ctr, S.sLo, S.sHi, S.sPos,
                                                          S.sIdx);
            k_sweep_len<<<ceil_div(N, 1024), 1024, 0, s>>>(S.sLo, S.sHi, N, h->sweepEnd, S.chunkCount, ctr);
        }
        scan(h, S.chunkCount, S.chunkStart, nullptr, N, &ctr->numChunks, h->maxChunks, 32);
        // chunk owners (one wave per interval), then groups of chunks per wave
        k_chunk_owner<<<std::max(1, std::min(ceil_div(N, 4), 16384)), 256, 0, s>>>(S.chunkStart, S.chunkCount, N, ctr,
                                                                                   S.chunkOwner);
        const int sweepBlocks = std::max(1, std::min(ceil_div(h->maxChunks, 4 * kSweepGroup), 4096));
        k_sweep_count_g<<<sweepBlocks, 256, 0, s>>>(S.chunkOwner, S.chunkStart, h->sweepEnd, ctr, S.sPos, S.hitCount,
                                                    S.hitMask, S.hash,
                                                    h->hashSize);
        // numHits is clamped to the pair capacity here (k_sweep_emit_g bounds its writes itself)
        scan(h, S.hitCount, S.hitStart, &ctr->numChunks, 0, &ctr->numHits, h->maxPairs, 1);
        k_hash_insert<<<G, 256, 0, s>>>(h->cur.ab, ctr, h->joints, h->nj, h->springs, h->ns, h->ignorePairs,
                                        h->numIgnore, S.hash, S.hashMask, ctr);
        k_sweep_emit_g<<<sweepBlocks, 256, 0, s>>>(S.chunkOwner, S.chunkStart, ctr, S.hitMask, S.sIdx, S.hitStart,
                                                   S.hash, S.hashMask, S.hits, S.hitNew, h->maxPairs);
        scan(h, S.hitNew, S.hitRank, &ctr->numHits, 0, &ctr->numCreated);
        k_create<<<G, 256, 0, s>>>(S.hits, S.hitNew, S.hitRank, ctr, h->cur, h->maxManifolds, h->maxPairs);
    }
    k_create_finish<<<1, 1, 0, s>>>(ctr, h->maxManifolds, h->maxPairs);
    if (h->timing)
        HIP_TRY(hipEventRecord(h->ev[1], s));

    // ---- force initialisation (manifold.cpp:100-174, joint.cpp:35-54) ----
    if (h->numRound > 0 || h->extMask >= 0)
        k_narrowphase<true><<<G, 256, 0, s>>>(h->cur, h->tmp, h->B, h->P, ctr, S.keep, h->recMatch, h->recStick,
                                              nullptr, nullptr, ext_table(h));
    else
    {
        k_narrowphase_boxes<<<4 * G, kNpBoxThreads, 0, s>>>(h->cur, h->tmp, h->B, h->P, ctr, S.keep, h->recMatch,
                                                            h->recStick, S.npRedo, h->npPolyCap);
        // the rare manifolds whose clip polygon outgrew the LDS capacity, at full capacity
        k_narrowphase<true><<<64, 256, 0, s>>>(h->cur, h->tmp, h->B, h->P, ctr, S.keep, h->recMatch, h->recStick,
                                               S.npRedo, &ctr->npRedoCount);
    }
    // the kept count is the new manifold count (nothing until the compaction reads the old one)
    scan(h, S.keep, S.keepRank, &ctr->numTotal, 0, &ctr->numManifolds);
    k_compact_manifolds<<<G / 2, 1024, 0, s>>>(h->tmp, h->cur, S.keep, S.keepRank, h->B.massMoment, ctr, h->recSrc,
                                               S.deg, S.adjCursor, N);
    scan(h, h->cur.count, S.contactStart, &ctr->numManifolds, 0, &ctr->contactCount);
    k_contact_records<<<G, 256, 0, s>>>(h->cur.count, h->cur.ab, S.contactStart, ctr, S.contactRec);
    if (h->nj > 0)
        k_joint_init<<<ceil_div(h->nj, 1024), 1024, 0, s>>>(h->joints, h->nj, h->B, h->P, ctr);
    if (h->timing)
        HIP_TRY(hipEventRecord(h->ev[2], s));

    // ---- per-body force lists, body warm start (solver.cpp:1134-1169) ----
    k_adj<false><<<G, 256, 0, s>>>(h->cur, ctr, h->joints, h->nj, h->springs, h->ns, h->B.massMoment, S.deg, S.adjStart,
                                   S.adjCursor, S.adj, S.adjSerial, S.adjPartner, h->maxAdj, &ctr->overflow);
    scan(h, S.deg, S.adjStart, nullptr, N, h->adjTotal);
    k_adj<true><<<G, 256, 0, s>>>(h->cur, ctr, h->joints, h->nj, h->springs, h->ns, h->B.massMoment, S.deg, S.adjStart,
                                  S.adjCursor, S.adj, S.adjSerial, S.adjPartner, h->maxAdj, &ctr->overflow);
    k_adj_sort<<<std::max(1, std::min(ceil_div(N, 4), 16384)), 256, 0, s>>>(S.deg, S.adjStart, S.adj, S.adjSerial,
                                                                          S.adjPartner, N, h->maxAdj);
    k_stamp_count<<<ceil_div(N, 1024), 1024, 0, s>>>(S.deg, S.adjStart, S.adj, h->cur.count, N, S.stampCount,
                                                   &ctr->maxStampCount);
    scan(h, S.stampCount, S.stampStart, nullptr, N, &ctr->numStamps);
    k_stamp_fill<<<ceil_div(N, 256), 256, 0, s>>>(S.deg, S.adjStart, S.adj, S.adjPartner, h->cur.count, h->cur.ab, N,
                                                  S.stampStart, S.stamps, h->maxStamps, &ctr->overflow);
    const int *ignoreDeg = h->numIgnore > 0 ? h->ignoreDeg : nullptr;
    k_warmstart<<<ceil_div(N, 1024), 1024, 0, s>>>(h->B, h->P, S.deg, ignoreDeg, N, S.color, ctr);

    // ---- colouring (avbd_gpu_solver.cpp:1812-1865) ----
    {
        // batched worlds: the sequential greedy colouring per world, one wave each. One world: the
        // parallel fixpoint; a world small enough for one workgroup gets kColorSmallRounds rounds,
        // after which (colours still changing along a long chain) the sequential rule finishes.
        // k_color_worlds' LDS: starts (int) + colours and degrees (short) per body of the largest
        // world, and the partner slice (short)
        const int cap = std::max(1, std::min(h->maxWorldBodies, kWorldSapMax));
        const size_t worldLds = (size_t)cap * 8 + (size_t)kColorWorldParts * 2;
        if (h->numWorlds > 1)
            k_color_worlds<<<h->numWorlds, 64, worldLds, s>>>(h->worldStart, S.deg, ignoreDeg, S.adjStart,
                                                              S.adjPartner, h->B.massMoment, S.color, ctr, 0, cap);
        else
        {
            const bool small = N <= kWorldSapMax;
            k_color<<<h->coopColor, kColorThreads, 0, s>>>(S.deg, ignoreDeg, S.adjStart, S.adjPartner, h->B.massMoment, N,
                                                        S.color, ctr, h->gb, small ? kColorSmallRounds : N + 2);
            if (small)
                k_color_worlds<<<1, 64, worldLds, s>>>(h->worldStart, S.deg, ignoreDeg, S.adjStart, S.adjPartner,
                                                       h->B.massMoment, S.color, ctr, 1, cap);
        }
        HIP_TRY(hipGetLastError());
    }
    {
        unsigned *keys = S.keys, *keysAlt = S.keysAlt;
        int *vals = S.vals, *valsAlt = S.valsAlt;
        k_color_keys<<<ceil_div(N, 1024), 1024, 0, s>>>(S.color, S.stampCount, N, keys, vals, ctr);
        int tiles = ceil_div(N, kRadixTile);
        k_radix_hist<<<tiles, kRadixThreads, 0, s>>>(keys, N, 0, S.radixHist, tiles);
        if (tiles <= kRadixFusedTiles)
            k_radix_scatter<<<tiles, kRadixThreads, 0, s>>>(keys, vals, N, 0, S.radixHist, tiles, keysAlt, valsAlt, 1);
        else
        {
            scan(h, S.radixHist, h->radixHistScan, nullptr, 256 * tiles, nullptr);
            k_radix_scatter<<<tiles, kRadixThreads, 0, s>>>(keys, vals, N, 0, h->radixHistScan, tiles, keysAlt, valsAlt, 0);
        }
        // the second pass keeps its scanned histogram: the colour offsets are read from it
        k_radix_hist<<<tiles, kRadixThreads, 0, s>>>(keysAlt, N, 8, S.radixHist, tiles);
        scan(h, S.radixHist, h->radixHistScan, nullptr, 256 * tiles, nullptr);
        k_radix_scatter<<<tiles, kRadixThreads, 0, s>>>(keysAlt, valsAlt, N, 8, h->radixHistScan, tiles, keys, S.colorBodies,
                                              0);
        k_color_offsets<<<1, 256, 0, s>>>(h->radixHistScan, tiles, N, S.colorOffsets, ctr, N, h->P.iterations);
        k_tasks<<<ceil_div(N, 256), 256, 0, s>>>(S.colorBodies, S.stampStart, S.stampCount, S.color, N, S.tasks, h->dfDone,
                                                 S.colorOffsets);
    }
    if (h->timing)
    {
        HIP_TRY(hipEventRecord(h->ev[3], s));
        HIP_TRY(hipEventRecord(h->ring[avbd_handle::kSlots * (h->ringSteps % avbd_handle::kRing) + 1], s));
    }
    return AVBD_OK;
}

PrimalArgs::PrimalArgs(const avbd_handle *h, int b, int e, float4 *sys, unsigned char *flags, const int4 *pad_,
                       int padShift_, int lazyDual_)
    : B(h->B), M(h->cur), joints(h->joints), springs(h->springs), stamps(h->S.stamps), tasks(h->S.tasks), begin(b),
      end(e), P(h->P), sysTape(sys), flagTape(flags), pad(pad_), padShift(padShift_), lazyDual(lazyDual_)
{
    void *p[14] = {&B, &M, &joints, &springs, &stamps, &tasks, &begin, &end, &P, &sysTape, &flagTape, &pad, &padShift,
                   &lazyDual};
    std::memcpy(ptr, p, sizeof(p));
}

AdaptiveArgs::AdaptiveArgs(const avbd_handle *h, int color, int threads, float4 *sys, unsigned char *flags,
                           const int4 *pad_, int padShift_, int lazyDual_)
    : B(h->B), M(h->cur), joints(h->joints), springs(h->springs), stamps(h->S.stamps), tasks(h->S.tasks), P(h->P),
      sysTape(sys), flagTape(flags), pad(pad_), padShift(padShift_), lazyDual(lazyDual_)
{
    const int *o = h->hostOffsets;
    const int cb = o[color], ce = o[color + 1];
    const int *b = o + kSegBase + kSegCount * color;
    const int b32 = std::min(std::max(b[0], cb), ce), b16 = std::min(std::max(b[1], b32), ce),
              b8 = std::min(std::max(b[2], b16), ce);
    seg = make_int4(cb, b32, b16, b8);
    end = ce;
    const int n64 = ceil_div(b32 - cb, threads / 64), n32 = ceil_div(b16 - b32, threads / 32),
              n16 = ceil_div(b8 - b16, threads / 16), n8 = ceil_div(ce - b8, threads / 8);
    blk = make_int4(n64, n64 + n32, n64 + n32 + n16, n64 + n32 + n16 + n8);
    void *p[15] = {&B, &M, &joints, &springs, &stamps, &tasks, &seg, &end, &blk, &P, &sysTape, &flagTape, &pad,
                   &padShift, &lazyDual};
    std::memcpy(ptr, p, sizeof(p));
}

DualArgs::DualArgs(const avbd_handle *h, int nc_, unsigned char *flags, int nj_, int staticRepeat_, DualTape tape_)
    : B(h->B), M(h->cur), joints(h->joints), nj(nj_ < 0 ? h->nj : nj_), contactRec(h->S.contactRec), nc(nc_), P(h->P),
      flagTape(flags), staticRepeat(staticRepeat_), tape(tape_)
{
    void *p[10] = {&B, &M, &joints, &nj, &contactRec, &nc, &P, &flagTape, &staticRepeat, &tape};
    std::memcpy(ptr, p, sizeof(p));
}

void destroy_graph(avbd_handle *h)
{
    if (h->graphExec)
        (void)hipGraphExecDestroy(h->graphExec);
    if (h->graph)
        (void)hipGraphDestroy(h->graph);
    h->graphExec = nullptr;
    h->graph = nullptr;
    h->graphPrimal.clear();
    h->graphDual = nullptr;
}

hipKernelNodeParams node_params(const void *fn, int grid, int threads, void **args)
{
    hipKernelNodeParams kp{};
    kp.func = const_cast<void *>(fn);
    kp.gridDim = dim3(std::max(grid, 1));
    kp.blockDim = dim3(threads);
    kp.sharedMemBytes = 0;
    kp.kernelParams = args;
    kp.extra = nullptr;
    return kp;
}

// One solver iteration as a graph: the colour launches in order, then the dual launch. The
// graph is rebuilt when its node count or launch shape changes; otherwise only the nodes'
// grids and ranges are updated for the step.
int iterate_graph(avbd_handle *h, const std::vector<std::pair<int, int>> &ranges, const std::vector<int> &colours,
                  int nc, int dualLanes)
{
    const avbd_launch_config &lc = h->launch;
    const bool adaptive = lc.primal_lanes == 0;
    const void *pfn = adaptive ? (const void *)adaptive_kernel(lc.primal_threads)
                               : (const void *)primal_kernel(lc.primal_lanes, lc.primal_threads);
    const int groups = adaptive ? 1 : lc.primal_threads / lc.primal_lanes;
    // the kernel node of colour range k (its argument block lives for the call that copies it)
    auto withNode = [&](size_t k, auto &&call) -> hipError_t {
        if (adaptive)
        {
            AdaptiveArgs a(h, colours[k], lc.primal_threads);
            hipKernelNodeParams kp = node_params(pfn, a.grid(), lc.primal_threads, a.ptr);
            return call(kp);
        }
        PrimalArgs a(h, ranges[k].first, ranges[k].second);
        hipKernelNodeParams kp = node_params(pfn, ceil_div(ranges[k].second - ranges[k].first, groups),
                                             lc.primal_threads, a.ptr);
        return call(kp);
    };
    const bool dual = dualLanes > 0;
    bool rebuild = !h->graphExec || h->graphPrimal.size() != ranges.size() || (h->graphDual != nullptr) != dual ||
                   h->graphLaunch.primal_lanes != lc.primal_lanes ||
                   h->graphLaunch.primal_threads != lc.primal_threads || h->graphLaunch.dual_threads != lc.dual_threads;
    if (rebuild)
    {
        destroy_graph(h);
        HIP_TRY(hipGraphCreate(&h->graph, 0));
        hipGraphNode_t prev = nullptr;
        for (size_t k = 0; k < ranges.size(); ++k)
        {
            hipGraphNode_t node;
            HIP_TRY(withNode(k, [&](hipKernelNodeParams &kp) {
                return hipGraphAddKernelNode(&node, h->graph, prev ? &prev : nullptr, prev ? 1 : 0, &kp);
            }));
            h->graphPrimal.push_back(node);
            prev = node;
        }
        if (dual)
        {
            DualArgs a(h, nc);
            hipKernelNodeParams kp = node_params((const void *)k_dual, ceil_div(dualLanes, lc.dual_threads),
                                                 lc.dual_threads, a.ptr);
            HIP_TRY(hipGraphAddKernelNode(&h->graphDual, h->graph, prev ? &prev : nullptr, prev ? 1 : 0, &kp));
        }
        HIP_TRY(hipGraphInstantiate(&h->graphExec, h->graph, nullptr, nullptr, 0));
        h->graphLaunch = lc;
    }
    else
    {
        for (size_t k = 0; k < ranges.size(); ++k)
            HIP_TRY(withNode(k, [&](hipKernelNodeParams &kp) {
                return hipGraphExecKernelNodeSetParams(h->graphExec, h->graphPrimal[k], &kp);
            }));
        if (dual)
        {
            DualArgs a(h, nc);
            hipKernelNodeParams kp = node_params((const void *)k_dual, ceil_div(dualLanes, lc.dual_threads),
                                                 lc.dual_threads, a.ptr);
            HIP_TRY(hipGraphExecKernelNodeSetParams(h->graphExec, h->graphDual, &kp));
        }
    }
    return AVBD_OK;
}

// Solver iterations: per colour one primal launch, then one dual launch. The colour and
// contact counts are read back once per step to size the launches.
int enqueue_iterate(avbd_handle *h)
{
    hipStream_t s = h->stream;
    HIP_TRY(hipMemcpyAsync(h->hostCtr, h->ctr, sizeof(Counters), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(h->hostOffsets, h->S.colorOffsets, sizeof(int) * kOffsetsLen, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    const Counters &c = *h->hostCtr;
    if (c.overflow || c.barrierTimeout)
        return check_device_errors(h);
    const int colors = std::min(c.colorCount, kMaxColors);
    const int dualLanes = c.contactCount + h->nj;
    // Timing marks (when enabled) bracket the last iteration's primal launches and its dual
    // launch only. By then the queue holds the earlier iterations' launches, so the marks see
    // the kernels back to back (the first iteration after the per-step readback would also
    // time the host's enqueueing), and a few marks per step keep the measured step close to
    // the unmeasured one.
    hipEvent_t *mark = h->timing ? &h->ring[avbd_handle::kSlots * (h->ringSteps % avbd_handle::kRing)] : nullptr;
    std::vector<std::pair<int, int>> ranges;
    std::vector<int> rangeColours;
    for (int col = 0; col < colors; ++col)
        if (h->hostOffsets[col + 1] > h->hostOffsets[col])
        {
            ranges.emplace_back(h->hostOffsets[col], h->hostOffsets[col + 1]);
            rangeColours.push_back(col);
        }
    const int primalPerIter = (int)ranges.size();
    const avbd_launch_config &lc = h->launch;
    h->lastNumManifolds = c.numManifolds;
    // adjoint tape (only while a step is being recomputed for its reversal): slot 0 holds the
    // state after warm start and force init, slot it+1 the state after iteration it
    const bool tape = h->tapePose != nullptr;
    const size_t tapePoseStride = 2 * (size_t)h->maxBodies, tapeContactStride = 2 * (size_t)kMaxContacts * h->maxManifolds;
    auto tapeAt = [&](int slot) {
        const int blocks = ceil_div(h->nb, 256) + std::max(1, std::min(ceil_div(c.contactCount, 256), 256));
        k_tape<<<blocks, 256, 0, s>>>(h->B.pose, h->nb, h->tapePose + slot * tapePoseStride, h->S.contactRec, h->ctr,
                                      h->cur.cr, h->tapeContact + slot * tapeContactStride);
    };
    // joint tape slot k: the joint records the primal of iteration k reads
    auto tapeJoints = [&](int slot) -> hipError_t {
        return h->nj > 0 && h->tapeJoint ? hipMemcpyAsync(h->tapeJoint + (size_t)slot * h->maxJoints, h->joints,
                                                          sizeof(Joint) * h->nj, hipMemcpyDeviceToDevice, s)
                                         : hipSuccess;
    };
    if (tape)
    {
        tapeAt(0);
        HIP_TRY(tapeJoints(0));
    }
    if (h->timing)
    {
        h->ringPrimalLaunches[h->ringSteps % avbd_handle::kRing] = primalPerIter;
        h->ringPrimalIters[h->ringSteps % avbd_handle::kRing] = 1;
    }
    if (lc.use_graph && !tape && h->P.iterations > 0 && (primalPerIter > 0 || dualLanes > 0))
    {
        int r = iterate_graph(h, ranges, rangeColours, c.contactCount, dualLanes);
        if (r)
            return r;
        // in graph form the last iteration is timed as a whole (its primal figure includes the dual)
        const int last = h->P.iterations - 1;
        for (int it = 0; it < h->P.iterations; ++it)
        {
            if (mark && it == last)
                HIP_TRY(hipEventRecord(mark[4], s));
            HIP_TRY(hipGraphLaunch(h->graphExec, s));
            if (mark && it == last)
            {
                HIP_TRY(hipEventRecord(mark[5], s));
                HIP_TRY(hipEventRecord(mark[6], s));
            }
        }
    }
    else
    {
        const bool adaptive = lc.primal_lanes == 0;
        const void *pfn = adaptive ? (const void *)adaptive_kernel(lc.primal_threads)
                                   : (const void *)primal_kernel(lc.primal_lanes, lc.primal_threads);
        const int groups = adaptive ? 1 : lc.primal_threads / lc.primal_lanes;
        const int last = h->P.iterations - 1;
        const int numTasks = h->hostOffsets[colors];
        // task-ordered stamps at a power-of-two stride (direct launches)
        int padShift = -1;
        if (numTasks > 0 && c.maxStampCount <= kPadStampsMax)
        {
            padShift = 0;
            // the stride holds every lane of a body's group: adaptive groups have 8 lanes or more
            while ((1 << padShift) < std::max(c.maxStampCount, adaptive ? 8 : lc.primal_lanes))
                ++padShift;
            const long long work = (long long)numTasks << padShift;
            k_pad_stamps<<<(int)std::min<long long>(ceil_div(work, 256LL), 8192), 256, 0, s>>>(
                h->S.tasks, h->S.stamps, h->S.color, numTasks, padShift, h->padStamps);
        }
        // lazy dual: contact duals move into the next iteration's primal (owner stamps), joints keep
        // their own dual launch per iteration, and one contact dual follows the last iteration
        // (a taped step keeps one dual launch per iteration: the tape records each dual's flags)
        const bool lazy = lc.lazy_dual && !tape && padShift >= 0 && c.contactCount > 0 && h->P.iterations > 0;
        if (lazy)
            k_lazy_bytes<<<1, 1, 0, s>>>(h->ctr, h->P.iterations);
        // flow form: one primal launch for every iteration and colour (contact-only worlds with the
        // lazy dual), its grid the number of blocks that are resident at once
        int flowGrid = 0;
        const int flowLanes = adaptive ? 16 : lc.primal_lanes;  // the flow form has fixed lanes per body
        if (lc.dataflow && lazy && h->nj == 0 && h->ns == 0 && primalPerIter > 0)
        {
            int &fb = h->flowBlocks[flowLanes == 8 ? 0 : flowLanes == 16 ? 1 : 2]
                                   [lc.primal_threads == 64 ? 0 : lc.primal_threads == 128 ? 1 : 2];
            if (fb == 0)
            {
                int per = 0;
                HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(
                    &per, (const void *)flow_kernel(flowLanes, lc.primal_threads), lc.primal_threads, 0));
                fb = std::max(per, 0) * h->numCUs;
                if (fb <= 0)
                    fb = -1;  // does not fit: per-colour launches
            }
            flowGrid = std::max(fb, 0);
        }
        if (flowGrid > 0)
        {
            if (mark)
                HIP_TRY(hipEventRecord(mark[4], s));
            flow_kernel(flowLanes, lc.primal_threads)<<<flowGrid, lc.primal_threads, 0, s>>>(
                h->B, h->cur, h->S.tasks, h->S.colorOffsets, colors, h->S.color, h->dfDone, h->P, h->padStamps,
                padShift, h->ctr);
            HIP_TRY(hipGetLastError());
            if (mark)
                HIP_TRY(hipEventRecord(mark[5], s));
            DualArgs a(h, c.contactCount, nullptr, 0, h->P.iterations);
            HIP_TRY(hipLaunchKernel((const void *)k_dual, dim3(ceil_div(c.contactCount, lc.dual_threads)),
                                    dim3(lc.dual_threads), a.ptr, 0, s));
            if (mark)
                HIP_TRY(hipEventRecord(mark[6], s));
        }
        else
            for (int it = 0; it < h->P.iterations; ++it)
            {
                if (mark && it == last)
                    HIP_TRY(hipEventRecord(mark[4], s));
                float4 *sys = tape ? h->tapeSys + (size_t)it * 9 * h->maxBodies : nullptr;
                unsigned char *sfl = tape ? h->tapeStampFlags + (size_t)it * h->maxStamps : nullptr;
                unsigned char *dfl = tape ? h->tapeDualFlags + (size_t)it * kMaxContacts * h->maxManifolds : nullptr;
                for (size_t q = 0; q < ranges.size(); ++q)
                {
                    const int4 *pad = padShift >= 0 ? h->padStamps : nullptr;
                    const int lz = lazy && it > 0 ? 1 : 0;
                    if (adaptive)
                    {
                        AdaptiveArgs a(h, rangeColours[q], lc.primal_threads, sys, sfl, pad, std::max(padShift, 0), lz);
                        HIP_TRY(hipLaunchKernel(pfn, dim3(a.grid()), dim3(lc.primal_threads), a.ptr, 0, s));
                    }
                    else
                    {
                        PrimalArgs a(h, ranges[q].first, ranges[q].second, sys, sfl, pad, std::max(padShift, 0), lz);
                        HIP_TRY(hipLaunchKernel(pfn, dim3(ceil_div(ranges[q].second - ranges[q].first, groups)),
                                                dim3(lc.primal_threads), a.ptr, 0, s));
                    }
                }
                if (mark && it == last)
                    HIP_TRY(hipEventRecord(mark[5], s));
                if (lazy)
                {
                    if (h->nj > 0)
                    {
                        DualArgs a(h, 0, nullptr);
                        HIP_TRY(hipLaunchKernel((const void *)k_dual, dim3(ceil_div(h->nj, lc.dual_threads)),
                                                dim3(lc.dual_threads), a.ptr, 0, s));
                    }
                    if (it == last)
                    {
                        DualArgs a(h, c.contactCount, nullptr, 0, h->P.iterations);
                        HIP_TRY(hipLaunchKernel((const void *)k_dual, dim3(ceil_div(c.contactCount, lc.dual_threads)),
                                                dim3(lc.dual_threads), a.ptr, 0, s));
                    }
                }
                else if (dualLanes > 0 || tape)
                {
                    // a taped step's tape slot after this iteration is written by its dual launch
                    DualTape dt;
                    if (tape)
                    {
                        dt.pose = h->tapePose + (it + 1) * tapePoseStride;
                        dt.contact = h->tapeContact + (it + 1) * tapeContactStride;
                        dt.nb = h->nb;
                    }
                    DualArgs a(h, c.contactCount, dfl, -1, 1, dt);
                    HIP_TRY(hipLaunchKernel((const void *)k_dual,
                                            dim3(std::max(1, ceil_div(dualLanes + dt.nb, lc.dual_threads))),
                                            dim3(lc.dual_threads), a.ptr, 0, s));
                    if (tape)
                        HIP_TRY(tapeJoints(it + 1));
                }
                if (mark && it == last)
                    HIP_TRY(hipEventRecord(mark[6], s));
            }
        if (h->timing && flowGrid > 0)
        {
            h->ringPrimalLaunches[h->ringSteps % avbd_handle::kRing] = 1;
            h->ringPrimalIters[h->ringSteps % avbd_handle::kRing] = h->P.iterations;
        }
    }
    if (h->timing)
    {
        HIP_TRY(hipEventRecord(h->ev[4], s));
        HIP_TRY(hipEventRecord(mark[2], s));
        h->ringIters[h->ringSteps % avbd_handle::kRing] = h->P.iterations > 0 ? 1 : 0;
    }
    return AVBD_OK;
}

int enqueue_finish(avbd_handle *h)
{
    if (h->nb > 0)
        k_finish<<<ceil_div(h->nb, 256), 256, 0, h->stream>>>(h->B, h->P, h->nb);
    if (h->numSpheres > 0)
        k_rolling<<<ceil_div(h->nb, 256), 256, 0, h->stream>>>(h->B, h->cur, h->S.deg, h->S.adjStart, h->S.adj,
                                                               h->S.adjPartner, h->nb, h->maxAdj);
    if (h->timing)
    {
        HIP_TRY(hipEventRecord(h->ev[5], h->stream));
        HIP_TRY(hipEventRecord(h->ring[avbd_handle::kSlots * (h->ringSteps % avbd_handle::kRing) + 3], h->stream));
        h->ringSteps++;
    }
    h->stepsTaken++;
    return AVBD_OK;
}

int pick_device(avbd_handle *h) { HIP_TRY(hipSetDevice(h->device)); return AVBD_OK; }

} // namespace

extern "C" {

const char *avbd_last_error(void) { return g_lastError.c_str(); }

int avbd_default_params(avbd_params *out)
{
    if (!out)
        return fail(AVBD_ERR_INVALID_ARGUMENT, "avbd_default_params: null output");
    out->dt = 1.0f / 60.0f;
    out->gravity = -10.0f;
    out->iterations = 10;
    out->alpha = 0.99f;
    out->beta_lin = 10000.0f;
    out->beta_ang = 100.0f;
    out->gamma = 0.999f;
    return AVBD_OK;
}

int avbd_make_box_body(const float size[3], float density, float friction, const float pos[3], const float rot[4],
                       const float vel[3], const float ang_vel[3], avbd_body *out)
{
    if (!size || !pos || !out)
        return fail(AVBD_ERR_INVALID_ARGUMENT, "avbd_make_box_body: null argument");
    static const float qid[4] = {0, 0, 0, 1}, zero[3] = {0, 0, 0};
    const float *q = rot ? rot : qid;
    const float *v = vel ? vel : zero;
    const float *w = ang_vel ? ang_vel : zero;
    std::memset(out, 0, sizeof(*out));
    float sx = size[0], sy = size[1], sz = size[2];
    float mass = sx * sy * sz * density;
    out->mass = mass;
    out->moment[0] = (sy * sy + sz * sz) / 12.0f * mass;
    out->moment[1] = (sx * sx + sz * sz) / 12.0f * mass;
    out->moment[2] = (sx * sx + sy * sy) / 12.0f * mass;
    float hx = sx * 0.5f, hy = sy * 0.5f, hz = sz * 0.5f;
    out->radius = sqrtf(hx * hx + hy * hy + hz * hz);
    out->friction = friction;
    for (int k = 0; k < 3; ++k)
    {
        out->size[k] = size[k];
        out->pos[k] = pos[k];
        out->initial_lin[k] = pos[k];
        out->inertial_lin[k] = pos[k];
        out->vel[k] = v[k];
        out->prev_vel[k] = v[k];
        out->ang_vel[k] = w[k];
    }
    for (int k = 0; k < 4; ++k)
        out->rot[k] = out->initial_ang[k] = out->inertial_ang[k] = q[k];
    out->shape = AVBD_SHAPE_BOX;
    out->shape_radius = out->radius;
    return AVBD_OK;
}

int avbd_make_round_body(int32_t shape, float radius, float half_length, float density, float friction,
                         const float pos[3], const float rot[4], const float vel[3], const float ang_vel[3],
                         avbd_body *out)
{
    if (!pos || !out)
        return fail(AVBD_ERR_INVALID_ARGUMENT, "avbd_make_round_body: null argument");
    if (shape != AVBD_SHAPE_SPHERE && shape != AVBD_SHAPE_CAPSULE && shape != AVBD_SHAPE_CYLINDER)
        return fail(AVBD_ERR_INVALID_ARGUMENT, "avbd_make_round_body: shape must be sphere, capsule or cylinder");
    if (!(radius > 0.0f) || !(half_length >= 0.0f))
        return fail(AVBD_ERR_INVALID_ARGUMENT, "avbd_make_round_body: radius > 0 and half_length >= 0 required");
    // the box constructor's record of the bounding size, then the shape's own mass properties
    // (rigid.cpp:44-86: sphere 4/3 pi r^3 rho, I = 2/5 m r^2; capsule = cylinder + two caps;
    // cylinder pi r^2 h rho, I = m r^2 / 2 about z and m (3 r^2 + h^2) / 12 across)
    const float pi = 3.14159265358979323846f;
    float d = radius * 2.0f;
    if (shape == AVBD_SHAPE_SPHERE)
        half_length = 0.0f;
    float len = half_length * 2.0f;
    float size[3] = {d, d, shape == AVBD_SHAPE_CAPSULE ? len + d : shape == AVBD_SHAPE_CYLINDER ? len : d};
    int r = avbd_make_box_body(size, density, friction, pos, rot, vel, ang_vel, out);
    if (r)
        return r;
    auto ball = [&](float rr) { return 4.0f / 3.0f * pi * rr * rr * rr * density; };
    if (shape == AVBD_SHAPE_SPHERE)
    {
        out->mass = ball(radius);
        float in = 2.0f / 5.0f * out->mass * radius * radius;
        out->moment[0] = out->moment[1] = out->moment[2] = in;
        out->radius = radius;
    }
    else if (shape == AVBD_SHAPE_CAPSULE)
    {
        float mc = pi * radius * radius * len * density;
        float mb = ball(radius);
        out->mass = mc + mb;
        float axial = 0.5f * mc * radius * radius + 2.0f / 5.0f * mb * radius * radius;
        float across = (3.0f * radius * radius + len * len) / 12.0f * mc + 2.0f / 5.0f * mb * radius * radius +
                       mb * half_length * half_length;
        out->moment[0] = out->moment[1] = across;
        out->moment[2] = axial;
        out->radius = half_length + radius;
    }
    else
    {
        out->mass = pi * radius * radius * len * density;
        float axial = 0.5f * out->mass * radius * radius;
        float across = (3.0f * radius * radius + len * len) / 12.0f * out->mass;
        out->moment[0] = out->moment[1] = across;
        out->moment[2] = axial;
        out->radius = sqrtf(radius * radius + half_length * half_length);
    }
    out->shape = shape;
    out->shape_radius = radius;
    out->half_length = half_length;
    return AVBD_OK;
}

int avbd_make_joint(const avbd_body *body_a, const avbd_body *body_b, int32_t index_a, int32_t index_b, int64_t serial,
                    const float r_a[3], const float r_b[3], float stiffness_lin, float stiffness_ang, float fracture,
                    avbd_joint *out)
{
    if (!body_b || !r_a || !r_b || !out)
        return fail(AVBD_ERR_INVALID_ARGUMENT, "avbd_make_joint: null argument");
    std::memset(out, 0, sizeof(*out));
    out->body_a = body_a ? index_a : -1;
    out->body_b = index_b;
    out->serial = serial;
    for (int k = 0; k < 3; ++k)
    {
        out->r_a[k] = r_a[k];
        out->r_b[k] = r_b[k];
    }
    out->stiffness_lin = stiffness_lin;
    out->stiffness_ang = stiffness_ang;
    out->fracture = fracture;
    float s0 = (body_a ? body_a->size[0] : 0.0f) + body_b->size[0];
    float s1 = (body_a ? body_a->size[1] : 0.0f) + body_b->size[1];
    float s2 = (body_a ? body_a->size[2] : 0.0f) + body_b->size[2];
    out->torque_arm = s0 * s0 + s1 * s1 + s2 * s2;
    return AVBD_OK;
}

int avbd_make_spring(const avbd_body *body_a, const avbd_body *body_b, int32_t index_a, int32_t index_b, int64_t serial,
                     const float r_a[3], const float r_b[3], float stiffness, float rest, avbd_spring *out)
{
    if (!body_a || !body_b || !r_a || !r_b || !out)
        return fail(AVBD_ERR_INVALID_ARGUMENT, "avbd_make_spring: null argument");
    std::memset(out, 0, sizeof(*out));
    out->body_a = index_a;
    out->body_b = index_b;
    out->serial = serial;
    for (int k = 0; k < 3; ++k)
    {
        out->r_a[k] = r_a[k];
        out->r_b[k] = r_b[k];
    }
    out->stiffness = stiffness;
    if (rest < 0.0f)
    {
        V3 pa = rotate(Q4{body_a->rot[0], body_a->rot[1], body_a->rot[2], body_a->rot[3]}, V3{r_a[0], r_a[1], r_a[2]}) +
                V3{body_a->pos[0], body_a->pos[1], body_a->pos[2]};
        V3 pb = rotate(Q4{body_b->rot[0], body_b->rot[1], body_b->rot[2], body_b->rot[3]}, V3{r_b[0], r_b[1], r_b[2]}) +
                V3{body_b->pos[0], body_b->pos[1], body_b->pos[2]};
        rest = len(pa - pb);
    }
    out->rest = rest;
    return AVBD_OK;
}

int avbd_create(const avbd_config *config, avbd_handle **out)
{
    if (!config || !out)
        return fail(AVBD_ERR_INVALID_ARGUMENT, "avbd_create: null argument");
    if (config->max_bodies <= 0)
        return fail(AVBD_ERR_INVALID_ARGUMENT, "avbd_create: max_bodies must be positive");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return fail(AVBD_ERR_NO_DEVICE, "avbd_create: no HIP device");
    if (config->device < 0 || config->device >= ndev)
        return fail(AVBD_ERR_INVALID_ARGUMENT, "avbd_create: device ordinal out of range");
    avbd_handle *h = new avbd_handle();
    h->cfg = *config;
    h->device = config->device;
    h->P = to_params(config->params);
    h->timing = config->enable_timing != 0;
    int N = config->max_bodies;
    h->maxBodies = N;
    h->maxJoints = std::max(config->max_joints, 0);
    h->maxSprings = std::max(config->max_springs, 0);
    h->maxManifolds = config->max_manifolds > 0 ? config->max_manifolds : 12 * N + 1024;
    h->maxPairs = config->max_pairs > 0 ? config->max_pairs : 16 * N + 4096;
    // One work item per 64 interval visits. The sweep axis is the axis of largest centre variance,
    // so a single body far from the rest (one that fell through the ground, say) can turn it to an
    // axis along which every interval overlaps every other: the sweep then visits all n(n-1)/2
    // pairs, as the reference's does. Room for that worst case, up to kMaxSweepItems work items
    // (20 bytes each); beyond it ~2560 checks per body and the capacity error.
    {
        const long long worst = (long long)N * (N - 1) / 2 / kSweepChunk + N + 1;
        const long long typical = 41LL * N + h->maxPairs / 8 + 65536;
        h->maxChunks = (int)std::max(typical, std::min(worst, kMaxSweepItems));
    }
    h->maxAdj = 2 * (h->maxManifolds + h->maxJoints + h->maxSprings) + 16;
    h->maxStamps = 2 * kMaxContacts * h->maxManifolds + 6 * h->maxJoints + 2 * h->maxSprings + 16;
    int hs = 1024;
    while (hs < 2 * (h->maxManifolds + h->maxJoints + h->maxSprings))
        hs <<= 1;
    h->hashSize = hs;
    int r = AVBD_OK;
    auto bail = [&](int code) {
        avbd_destroy(h);
        return code;
    };
    if ((r = pick_device(h)) || (r = bind_static_scratch(h->device)))
        return bail(r);
    hipError_t e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
    if (e != hipSuccess)
        return bail(fail(AVBD_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e)));
    Bodies &B = h->B;
    if ((r = h->alloc(B.pose, 4 * (size_t)N)) || (r = h->alloc(B.inerLin, N)) || (r = h->alloc(B.inerAng, N)) ||
        (r = h->alloc(B.vel, N)) || (r = h->alloc(B.angVel, N)) || (r = h->alloc(B.prevVel, N)) ||
        (r = h->alloc(B.massMoment, N)) || (r = h->alloc(B.sizeRad, N)) || (r = h->alloc(B.shape, N)) ||
        (r = h->alloc(h->avbsRank, N)) || (r = h->alloc(h->ignoreDeg, N)) || (r = h->alloc(B.friction, N)) ||
        (r = h->alloc(B.denseId, N)))
        return bail(r);
    if ((r = alloc_manifolds(h, h->cur, h->maxManifolds)) || (r = alloc_manifolds(h, h->tmp, h->maxManifolds)))
        return bail(r);
    if ((r = h->alloc(h->joints, h->maxJoints)) || (r = h->alloc(h->springs, h->maxSprings)))
        return bail(r);
    Scratch &S = h->S;
    int tiles = ceil_div(N, kRadixTile);
    // look-back scan status: one word per tile of the longest scanned array
    h->scanMaxTiles = ceil_div(std::max({N, h->maxChunks, h->maxPairs, h->maxManifolds, 256 * tiles}), kLbTile);
    if ((r = h->alloc(S.axisSums, 64 * 6)) || (r = h->alloc(S.adjPartner, h->maxAdj)) ||
        (r = h->alloc(S.stampCount, N)) || (r = h->alloc(S.stampStart, N)) || (r = h->alloc(S.stamps, h->maxStamps)) ||
        (r = h->alloc(S.tasks, N)) || (r = h->alloc(S.contactStart, h->maxManifolds)) ||
        (r = h->alloc(S.contactRec, (size_t)kMaxContacts * h->maxManifolds)) || (r = h->alloc(S.keys, N)) || (r = h->alloc(S.keysAlt, N)) ||
        (r = h->alloc(S.vals, N)) || (r = h->alloc(S.valsAlt, N)) || (r = h->alloc(S.radixHist, 256 * tiles)) ||
        (r = h->alloc(h->radixHistScan, 256 * tiles)) || (r = h->alloc(S.sLo, N)) || (r = h->alloc(S.sHi, N)) ||
        (r = h->alloc(S.sPos, N)) || (r = h->alloc(S.sIdx, N)) || (r = h->alloc(S.chunkCount, N)) ||
        (r = h->alloc(S.chunkStart, N)) || (r = h->alloc(h->sweepEnd, N)) || (r = h->alloc(S.hitCount, h->maxChunks)) ||
        (r = h->alloc(S.hitStart, h->maxChunks)) || (r = h->alloc(S.hitMask, h->maxChunks)) ||
        (r = h->alloc(S.chunkOwner, h->maxChunks)) || (r = h->alloc(S.hits, h->maxPairs)) ||
        (r = h->alloc(S.hitNew, h->maxPairs)) || (r = h->alloc(S.hitRank, h->maxPairs)) ||
        (r = h->alloc(S.hash, h->hashSize)) || (r = h->alloc(S.keep, h->maxManifolds)) || (r = h->alloc(S.npRedo, h->maxManifolds)) ||
        (r = h->alloc(S.keepRank, h->maxManifolds)) || (r = h->alloc(S.deg, N)) || (r = h->alloc(S.adjStart, N)) ||
        (r = h->alloc(S.adjCursor, N)) || (r = h->alloc(S.adj, h->maxAdj)) || (r = h->alloc(S.adjSerial, h->maxAdj)) ||
        (r = h->alloc(S.color, N)) || (r = h->alloc(S.colorOffsets, kOffsetsLen)) || (r = h->alloc(S.colorBodies, N)) ||
        (r = h->alloc(h->ctr, 1)) || (r = h->alloc(h->gb, 1)) ||
        (r = h->alloc(h->scanStatus, (size_t)h->scanMaxTiles)) || (r = h->alloc(h->scanCtl, 4)) ||
        (r = h->alloc(h->padStamps, (size_t)N * kPadStampsMax)) ||
        (r = h->alloc(h->adjTotal, 1)) || (r = h->alloc(h->worldStart, (size_t)N + 1)) || (r = h->alloc(h->dfDone, N)))
        return bail(r);
    S.hashMask = h->hashSize - 1;
    if ((e = hipHostMalloc((void **)&h->hostCtr, sizeof(Counters))) != hipSuccess ||
        (e = hipHostMalloc((void **)&h->hostOffsets, sizeof(int) * kOffsetsLen)) != hipSuccess)
        return bail(fail(AVBD_ERR_HIP, std::string("hipHostMalloc: ") + hipGetErrorString(e)));
    if ((e = hipMemsetAsync(h->ctr, 0, sizeof(Counters), h->stream)) != hipSuccess ||
        (e = hipMemsetAsync(h->gb, 0, sizeof(GridBarrier), h->stream)) != hipSuccess ||
        (e = hipMemsetAsync(h->scanStatus, 0, sizeof(unsigned long long) * h->scanMaxTiles, h->stream)) != hipSuccess ||
        (e = hipMemsetAsync(h->scanCtl, 0, sizeof(unsigned) * 4, h->stream)) != hipSuccess ||
        (e = hipMemsetAsync(S.color, 0xFF, sizeof(int) * N, h->stream)) != hipSuccess ||
        (e = hipMemsetAsync(h->dfDone, 0, sizeof(int) * N, h->stream)) != hipSuccess)
        return bail(fail(AVBD_ERR_HIP, std::string("hipMemsetAsync: ") + hipGetErrorString(e)));
    // colouring grid: co-resident blocks of kColorThreads threads
    hipDeviceProp_t prop;
    if ((e = hipGetDeviceProperties(&prop, h->device)) != hipSuccess)
        return bail(fail(AVBD_ERR_HIP, std::string("hipGetDeviceProperties: ") + hipGetErrorString(e)));
    int perCuCol = 0;
    if ((e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCuCol, (const void *)k_color, kColorThreads, 0)) !=
        hipSuccess)
        return bail(fail(AVBD_ERR_HIP, std::string("hipOccupancy: ") + hipGetErrorString(e)));
    if (perCuCol < 1)
        return bail(fail(AVBD_ERR_HIP, "the colouring kernel does not fit on a compute unit"));
    h->coopColor = std::min(prop.multiProcessorCount, 64);
    h->numCUs = prop.multiProcessorCount;
    for (int k = 0; k < 8; ++k)
        if ((e = hipEventCreate(&h->ev[k])) != hipSuccess)
            return bail(fail(AVBD_ERR_HIP, std::string("hipEventCreate: ") + hipGetErrorString(e)));
    if (h->timing)
    {
        h->ring.assign((size_t)avbd_handle::kSlots * avbd_handle::kRing, nullptr);
        h->ringIters.assign(avbd_handle::kRing, 0);
        h->ringPrimalLaunches.assign(avbd_handle::kRing, 0);
        h->ringPrimalIters.assign(avbd_handle::kRing, 0);
        for (hipEvent_t &ev : h->ring)
            if ((e = hipEventCreate(&ev)) != hipSuccess)
                return bail(fail(AVBD_ERR_HIP, std::string("hipEventCreate: ") + hipGetErrorString(e)));
    }
    if ((e = hipStreamSynchronize(h->stream)) != hipSuccess)
        return bail(fail(AVBD_ERR_HIP, std::string("hipStreamSynchronize: ") + hipGetErrorString(e)));
    *out = h;
    return AVBD_OK;
}

int avbd_destroy(avbd_handle *h)
{
    if (!h)
        return AVBD_OK;
    (void)hipSetDevice(h->device);
    if (h->stream)
        (void)hipStreamSynchronize(h->stream);
    destroy_graph(h);
    for (void *p : h->allocs)
        (void)hipFree(p);
    for (void *p : {(void *)h->extEntries, (void *)h->extHit, (void *)h->extHitStart, (void *)h->extTable,
                    (void *)h->extKeys, (void *)h->extIdx})
        if (p)
            (void)hipFree(p);
    if (h->hostCtr)
        (void)hipHostFree(h->hostCtr);
    if (h->hostOffsets)
        (void)hipHostFree(h->hostOffsets);
    if (h->hostGridInfo)
        (void)hipHostFree(h->hostGridInfo);
    for (int k = 0; k < 8; ++k)
        if (h->ev[k])
            (void)hipEventDestroy(h->ev[k]);
    for (hipEvent_t ev : h->ring)
        if (ev)
            (void)hipEventDestroy(ev);
    if (h->stream)
        (void)hipStreamDestroy(h->stream);
    delete h;
    return AVBD_OK;
}

int avbd_set_params(avbd_handle *h, const avbd_params *params)
{
    if (!h || !params)
        return fail(AVBD_ERR_INVALID_ARGUMENT, "avbd_set_params: null argument");
    if (params->iterations < 0 || !(params->dt > 0.0f))
        return fail(AVBD_ERR_INVALID_ARGUMENT, "avbd_set_params: iterations must be >= 0 and dt > 0");
    h->cfg.params = *params;
    h->P = to_params(*params);
    return AVBD_OK;
}

int avbd_upload_state(avbd_handle *h, const avbd_body *bodies, int32_t num_bodies, const avbd_joint *joints,
                      int32_t num_joints, const avbd_spring *springs, int32_t num_springs,
                      const avbd_manifold *manifolds, int32_t num_manifolds, int64_t next_serial)
{
    if (!h)
        return fail(AVBD_ERR_INVALID_ARGUMENT, "avbd_upload_state: null handle");
    if (num_bodies < 0 || num_joints < 0 || num_springs < 0 || num_manifolds < 0 ||
        (num_bodies && !bodies) || (num_joints && !joints) || (num_springs && !springs) ||
        (num_manifolds && !manifolds))
        return fail(AVBD_ERR_INVALID_ARGUMENT, "avbd_upload_state: bad table pointer or count");
    if (num_bodies > h->maxBodies || num_joints > h->maxJoints || num_springs > h->maxSprings ||
        num_manifolds > h->maxManifolds)
        return fail(AVBD_ERR_CAPACITY, "avbd_upload_state: state exceeds the handle's capacities");
    int numRound = 0, numSpheres = 0;
    for (int i = 0; i < num_bodies; ++i)
    {
        int sh = bodies[i].shape;
        if (sh < AVBD_SHAPE_BOX || sh > AVBD_SHAPE_CYLINDER)
            return fail(AVBD_ERR_UNSUPPORTED, "avbd_upload_state: unknown shape code");
        if (sh != AVBD_SHAPE_BOX && !(bodies[i].shape_radius > 0.0f && bodies[i].half_length >= 0.0f))
            return fail(AVBD_ERR_INVALID_ARGUMENT, "avbd_upload_state: round shape needs radius > 0 and half_length >= 0");
        numRound += sh != AVBD_SHAPE_BOX;
        numSpheres += sh == AVBD_SHAPE_SPHERE;
    }
    auto badBody = [&](int b, bool allowWorld) { return (b < 0 && !(allowWorld && b == -1)) || b >= num_bodies; };
    for (int i = 0; i < num_joints; ++i)
        if (badBody(joints[i].body_a, true) || badBody(joints[i].body_b, false))
            return fail(AVBD_ERR_INVALID_ARGUMENT, "avbd_upload_state: joint body index out of range");
    for (int i = 0; i < num_springs; ++i)
        if (badBody(springs[i].body_a, false) || badBody(springs[i].body_b, false))
            return fail(AVBD_ERR_INVALID_ARGUMENT, "avbd_upload_state: spring body index out of range");
    for (int i = 0; i < num_manifolds; ++i)
        if (badBody(manifolds[i].body_a, false) || badBody(manifolds[i].body_b, false) ||
            manifolds[i].num_contacts < 0 || manifolds[i].num_contacts > AVBD_MAX_CONTACTS)
            return fail(AVBD_ERR_INVALID_ARGUMENT, "avbd_upload_state: manifold out of range");
    int r = pick_device(h);
    if (r)
        return r;
    h->ratioValid = false;
    ++h->uploadEpoch;
    int N = num_bodies;
    std::vector<float4> pos(N), rot(N), il(N), ia(N), nl(N), na(N), vel(N), av(N), pv(N), mm(N), sr(N), sh(N);
    std::vector<float> fr(N);
    std::vector<int> ids(N);
    for (int i = 0; i < N; ++i)
    {
        const avbd_body &b = bodies[i];
        pos[i] = make_float4(b.pos[0], b.pos[1], b.pos[2], 0.0f);
        rot[i] = make_float4(b.rot[0], b.rot[1], b.rot[2], b.rot[3]);
        il[i] = make_float4(b.initial_lin[0], b.initial_lin[1], b.initial_lin[2], 0.0f);
        ia[i] = make_float4(b.initial_ang[0], b.initial_ang[1], b.initial_ang[2], b.initial_ang[3]);
        nl[i] = make_float4(b.inertial_lin[0], b.inertial_lin[1], b.inertial_lin[2], 0.0f);
        na[i] = make_float4(b.inertial_ang[0], b.inertial_ang[1], b.inertial_ang[2], b.inertial_ang[3]);
        vel[i] = make_float4(b.vel[0], b.vel[1], b.vel[2], 0.0f);
        av[i] = make_float4(b.ang_vel[0], b.ang_vel[1], b.ang_vel[2], 0.0f);
        pv[i] = make_float4(b.prev_vel[0], b.prev_vel[1], b.prev_vel[2], 0.0f);
        mm[i] = make_float4(b.mass, b.moment[0], b.moment[1], b.moment[2]);
        sr[i] = make_float4(b.size[0], b.size[1], b.size[2], b.radius);
        sh[i] = make_float4(b.shape_radius, b.half_length, (float)b.shape, 0.0f);
        fr[i] = b.friction;
        ids[i] = b.dense_id;
    }
    // AVBS snapshot order: bodies by ascending dense id (ties by list index); ids are fixed while
    // the state is resident, so the ranks are computed once here
    std::vector<int> byId(N), avbsRank(N);
    for (int i = 0; i < N; ++i)
        byId[i] = i;
    std::stable_sort(byId.begin(), byId.end(), [&](int x, int y) { return ids[x] < ids[y]; });
    for (int k = 0; k < N; ++k)
        avbsRank[byId[k]] = k;
    hipStream_t s = h->stream;
    auto up = [&](void *dst, const void *src, size_t bytes) -> hipError_t {
        return bytes ? hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s) : hipSuccess;
    };
    size_t f4 = sizeof(float4) * N;
    std::vector<float4> pose(4 * (size_t)N);
    for (int i = 0; i < N; ++i)
    {
        pose[4 * i + P_POS] = pos[i];
        pose[4 * i + P_ROT] = rot[i];
        pose[4 * i + P_ILIN] = il[i];
        pose[4 * i + P_IANG] = ia[i];
    }
    hipError_t e;
    if ((e = up(h->B.pose, pose.data(), 4 * f4)) || (e = up(h->B.inerLin, nl.data(), f4)) ||
        (e = up(h->B.inerAng, na.data(), f4)) || (e = up(h->B.vel, vel.data(), f4)) || (e = up(h->B.angVel, av.data(), f4)) ||
        (e = up(h->B.prevVel, pv.data(), f4)) || (e = up(h->B.massMoment, mm.data(), f4)) ||
        (e = up(h->B.sizeRad, sr.data(), f4)) || (e = up(h->B.shape, sh.data(), f4)) ||
        (e = up(h->avbsRank, avbsRank.data(), sizeof(int) * N)) ||
        (e = up(h->B.friction, fr.data(), sizeof(float) * N)) ||
        (e = up(h->B.denseId, ids.data(), sizeof(int) * N)))
        return fail(AVBD_ERR_HIP, std::string("upload bodies: ") + hipGetErrorString(e));
    h->hostDenseIds = ids;
    h->extDirty = true;
    // joints / springs
    std::vector<Joint> jt(num_joints);
    for (int i = 0; i < num_joints; ++i)
    {
        const avbd_joint &j = joints[i];
        Joint &o = jt[i];
        std::memset(&o, 0, sizeof(o));
        o.a = j.body_a;
        o.b = j.body_b;
        o.broken = j.broken;
        o.dead = 0;
        o.serial = j.serial;
        o.rA = make_float4(j.r_a[0], j.r_a[1], j.r_a[2], j.stiffness_lin);
        o.rB = make_float4(j.r_b[0], j.r_b[1], j.r_b[2], j.stiffness_ang);
        o.c0Lin = make_float4(j.c0_lin[0], j.c0_lin[1], j.c0_lin[2], j.torque_arm);
        o.c0Ang = make_float4(j.c0_ang[0], j.c0_ang[1], j.c0_ang[2], j.fracture);
        o.penLin = make_float4(j.penalty_lin[0], j.penalty_lin[1], j.penalty_lin[2], 0.0f);
        o.penAng = make_float4(j.penalty_ang[0], j.penalty_ang[1], j.penalty_ang[2], 0.0f);
        o.lamLin = make_float4(j.lambda_lin[0], j.lambda_lin[1], j.lambda_lin[2], 0.0f);
        o.lamAng = make_float4(j.lambda_ang[0], j.lambda_ang[1], j.lambda_ang[2], 0.0f);
    }
    std::vector<Spring> st(num_springs);
    for (int i = 0; i < num_springs; ++i)
    {
        const avbd_spring &sp = springs[i];
        Spring &o = st[i];
        std::memset(&o, 0, sizeof(o));
        o.a = sp.body_a;
        o.b = sp.body_b;
        o.serial = sp.serial;
        o.rA = make_float4(sp.r_a[0], sp.r_a[1], sp.r_a[2], sp.rest);
        o.rB = make_float4(sp.r_b[0], sp.r_b[1], sp.r_b[2], sp.stiffness);
    }
    if ((e = up(h->joints, jt.data(), sizeof(Joint) * num_joints)) ||
        (e = up(h->springs, st.data(), sizeof(Spring) * num_springs)))
        return fail(AVBD_ERR_HIP, std::string("upload forces: ") + hipGetErrorString(e));
    // manifolds in ascending serial
    std::vector<int> order(num_manifolds);
    for (int i = 0; i < num_manifolds; ++i)
        order[i] = i;
    std::stable_sort(order.begin(), order.end(),
                     [&](int x, int y) { return manifolds[x].serial < manifolds[y].serial; });
    int M = num_manifolds;
    std::vector<int2> ab(M);
    std::vector<long long> serial(M);
    std::vector<int> cnt(M);
    size_t CS = (size_t)M * kMaxContacts;
    std::vector<int> feat(CS, 0);
    std::vector<float4> cr(CS * kCR, make_float4(0, 0, 0, 0));
    for (int k = 0; k < M; ++k)
    {
        const avbd_manifold &m = manifolds[order[k]];
        ab[k] = make_int2(m.body_a, m.body_b);
        serial[k] = m.serial;
        cnt[k] = m.num_contacts;
        for (int c = 0; c < kMaxContacts; ++c)
        {
            const avbd_contact &ct = m.contacts[c];
            size_t sl = (size_t)k * kMaxContacts + c;
            feat[sl] = ct.feature;
            float4 *o = &cr[sl * kCR];
            // the shape code and radii / half lengths of the pair, as k_narrowphase stores them
            const avbd_body &ba = bodies[m.body_a], &bb = bodies[m.body_b];
            int code = shape_code(ba.shape, bb.shape, ct.feature);
            int codeBits;
            std::memcpy(&codeBits, &code, sizeof(int));
            float codeF;
            std::memcpy(&codeF, &codeBits, sizeof(float));
            o[CR_RA] = make_float4(ct.r_a[0], ct.r_a[1], ct.r_a[2], m.friction);
            o[CR_RB] = make_float4(ct.r_b[0], ct.r_b[1], ct.r_b[2], codeF);
            o[CR_C0] = make_float4(ct.c0[0], ct.c0[1], ct.c0[2], bb.half_length);
            o[CR_N] = make_float4(m.basis[0], m.basis[1], m.basis[2], ba.shape_radius);
            o[CR_T1] = make_float4(m.basis[3], m.basis[4], m.basis[5], bb.shape_radius);
            o[CR_T2] = make_float4(m.basis[6], m.basis[7], m.basis[8], ba.half_length);
            o[CR_PEN] = make_float4(ct.penalty[0], ct.penalty[1], ct.penalty[2], 0.0f);
            o[CR_LAM] = make_float4(ct.lambda[0], ct.lambda[1], ct.lambda[2], ct.stick ? 1.0f : 0.0f);
        }
    }
    Manifolds &C = h->cur;
    if ((e = up(C.ab, ab.data(), sizeof(int2) * M)) || (e = up(C.serial, serial.data(), sizeof(long long) * M)) ||
        (e = up(C.count, cnt.data(), sizeof(int) * M)) || (e = up(C.feature, feat.data(), sizeof(int) * CS)) ||
        (e = up(C.cr, cr.data(), sizeof(float4) * CS * kCR)))
        return fail(AVBD_ERR_HIP, std::string("upload manifolds: ") + hipGetErrorString(e));
    Counters c;
    std::memset(&c, 0, sizeof(c));
    c.numManifolds = M;
    c.nextSerial = next_serial;
    for (int i = 0; i < num_springs; ++i)
        c.springEndpoints += (bodies[springs[i].body_a].mass > 0.0f ? 1 : 0) + (bodies[springs[i].body_b].mass > 0.0f ? 1 : 0);
    if ((e = up(h->ctr, &c, sizeof(c))))
        return fail(AVBD_ERR_HIP, std::string("upload counters: ") + hipGetErrorString(e));
    std::vector<int> negs(std::max(N, 1), -1);
    if ((e = up(h->S.color, negs.data(), sizeof(int) * N)))
        return fail(AVBD_ERR_HIP, std::string("upload colours: ") + hipGetErrorString(e));
    const int oneWorld[2] = {0, N};  // the whole state is world 0
    if ((e = up(h->worldStart, oneWorld, sizeof(oneWorld))))
        return fail(AVBD_ERR_HIP, std::string("upload world range: ") + hipGetErrorString(e));
    if ((e = hipStreamSynchronize(s)) != hipSuccess)
        return fail(AVBD_ERR_HIP, std::string("upload sync: ") + hipGetErrorString(e));
    h->nb = N;
    h->nj = num_joints;
    h->ns = num_springs;
    h->numRound = numRound;
    h->numSpheres = numSpheres;
    h->numIgnore = 0;
    h->lastNumManifolds = M;
    h->numWorlds = 1;
    h->worldStartHost.assign({0, N});
    h->maxWorldBodies = N;
    return AVBD_OK;
}

int avbd_state_counts(avbd_handle *h, int32_t *num_bodies, int32_t *num_joints, int32_t *num_springs,
                      int32_t *num_manifolds, int64_t *next_serial)
{
    if (!h)
        return fail(AVBD_ERR_INVALID_ARGUMENT, "avbd_state_counts: null handle");
    int r = pick_device(h);
    if (r)
        return r;
    Counters c;
    HIP_TRY(hipMemcpyAsync(&c, h->ctr, sizeof(c), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    std::vector<Joint> jt(h->nj);
    if (h->nj)
        HIP_TRY(hipMemcpy(jt.data(), h->joints, sizeof(Joint) * h->nj, hipMemcpyDeviceToHost));
    int live = 0;
    for (const Joint &j : jt)
        live += j.dead ? 0 : 1;
    if (num_bodies)
        *num_bodies = h->nb;
    if (num_joints)
        *num_joints = live;
    if (num_springs)
        *num_springs = h->ns;
    if (num_manifolds)
        *num_manifolds = c.numManifolds;
    if (next_serial)
        *next_serial = c.nextSerial;
    return AVBD_OK;
}

int avbd_download_state(avbd_handle *h, avbd_body *bodies, avbd_joint *joints, avbd_spring *springs,
                        avbd_manifold *manifolds)
{
    if (!h)
        return fail(AVBD_ERR_INVALID_ARGUMENT, "avbd_download_state: null handle");
    int r = pick_device(h);
    if (r)
        return r;
    HIP_TRY(hipStreamSynchronize(h->stream));
    if ((r = check_device_errors(h)))
        return r;
    int N = h->nb;
    if (bodies && N)
    {
        std::vector<float4> v[kBodyRows];
        std::vector<float4> pose(4 * (size_t)N);
        HIP_TRY(hipMemcpy(pose.data(), h->B.pose, sizeof(float4) * 4 * N, hipMemcpyDeviceToHost));
        for (int k = 0; k < 4; ++k)
        {
            v[k].resize(N);
            for (int i = 0; i < N; ++i)
                v[k][i] = pose[4 * i + k];
        }
        float4 *src[kBodyRows] = {nullptr, nullptr, nullptr, nullptr, h->B.inerLin, h->B.inerAng, h->B.vel,
                           h->B.angVel, h->B.prevVel, h->B.massMoment, h->B.sizeRad, h->B.shape};
        for (int k = 4; k < kBodyRows; ++k)
        {
            v[k].resize(N);
            HIP_TRY(hipMemcpy(v[k].data(), src[k], sizeof(float4) * N, hipMemcpyDeviceToHost));
        }
        std::vector<float> fr(N);
        std::vector<int> ids(N);
        HIP_TRY(hipMemcpy(fr.data(), h->B.friction, sizeof(float) * N, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(ids.data(), h->B.denseId, sizeof(int) * N, hipMemcpyDeviceToHost));
        bodies_to_host(v, fr, ids, N, bodies);
    }
    if (joints && h->nj)
    {
        std::vector<Joint> jt(h->nj);
        HIP_TRY(hipMemcpy(jt.data(), h->joints, sizeof(Joint) * h->nj, hipMemcpyDeviceToHost));
        int k = 0;
        for (const Joint &j : jt)
        {
            if (j.dead)
                continue;
            avbd_joint &o = joints[k++];
            std::memset(&o, 0, sizeof(o));
            o.body_a = j.a;
            o.body_b = j.b;
            o.serial = j.serial;
            o.broken = j.broken;
            const float4 *f[8] = {&j.rA, &j.rB, &j.c0Lin, &j.c0Ang, &j.penLin, &j.penAng, &j.lamLin, &j.lamAng};
            float *d[8] = {o.r_a, o.r_b, o.c0_lin, o.c0_ang, o.penalty_lin, o.penalty_ang, o.lambda_lin, o.lambda_ang};
            for (int q = 0; q < 8; ++q)
            {
                d[q][0] = f[q]->x;
                d[q][1] = f[q]->y;
                d[q][2] = f[q]->z;
            }
            o.stiffness_lin = j.rA.w;
            o.stiffness_ang = j.rB.w;
            o.torque_arm = j.c0Lin.w;
            o.fracture = j.c0Ang.w;
        }
    }
    if (springs && h->ns)
    {
        std::vector<Spring> st(h->ns);
        HIP_TRY(hipMemcpy(st.data(), h->springs, sizeof(Spring) * h->ns, hipMemcpyDeviceToHost));
        for (int i = 0; i < h->ns; ++i)
        {
            avbd_spring &o = springs[i];
            std::memset(&o, 0, sizeof(o));
            o.body_a = st[i].a;
            o.body_b = st[i].b;
            o.serial = st[i].serial;
            o.r_a[0] = st[i].rA.x;
            o.r_a[1] = st[i].rA.y;
            o.r_a[2] = st[i].rA.z;
            o.r_b[0] = st[i].rB.x;
            o.r_b[1] = st[i].rB.y;
            o.r_b[2] = st[i].rB.z;
            o.rest = st[i].rA.w;
            o.stiffness = st[i].rB.w;
        }
    }
    if (manifolds)
    {
        Counters c;
        HIP_TRY(hipMemcpy(&c, h->ctr, sizeof(c), hipMemcpyDeviceToHost));
        int M = c.numManifolds;
        if (M)
        {
            size_t CS = (size_t)M * kMaxContacts;
            std::vector<int2> ab(M);
            std::vector<long long> serial(M);
            std::vector<int> cnt(M), feat(CS);
            std::vector<float4> cr(CS * kCR);
            const Manifolds &C = h->cur;
            HIP_TRY(hipMemcpy(ab.data(), C.ab, sizeof(int2) * M, hipMemcpyDeviceToHost));
            HIP_TRY(hipMemcpy(serial.data(), C.serial, sizeof(long long) * M, hipMemcpyDeviceToHost));
            HIP_TRY(hipMemcpy(cnt.data(), C.count, sizeof(int) * M, hipMemcpyDeviceToHost));
            HIP_TRY(hipMemcpy(feat.data(), C.feature, sizeof(int) * CS, hipMemcpyDeviceToHost));
            HIP_TRY(hipMemcpy(cr.data(), C.cr, sizeof(float4) * CS * kCR, hipMemcpyDeviceToHost));
            for (int k = 0; k < M; ++k)
            {
                avbd_manifold &o = manifolds[k];
                std::memset(&o, 0, sizeof(o));
                o.body_a = ab[k].x;
                o.body_b = ab[k].y;
                o.serial = serial[k];
                o.num_contacts = cnt[k];
                if (cnt[k] > 0)
                {
                    const float4 *r0 = &cr[(size_t)k * kMaxContacts * kCR];
                    o.friction = r0[CR_RA].w;
                    const float4 *bs[3] = {&r0[CR_N], &r0[CR_T1], &r0[CR_T2]};
                    for (int q = 0; q < 3; ++q)
                    {
                        o.basis[3 * q] = bs[q]->x;
                        o.basis[3 * q + 1] = bs[q]->y;
                        o.basis[3 * q + 2] = bs[q]->z;
                    }
                }
                for (int q = 0; q < cnt[k]; ++q)
                {
                    size_t sl = (size_t)k * kMaxContacts + q;
                    const float4 *r = &cr[sl * kCR];
                    avbd_contact &ct = o.contacts[q];
                    ct.feature = feat[sl];
                    const float4 *f[5] = {&r[CR_RA], &r[CR_RB], &r[CR_C0], &r[CR_PEN], &r[CR_LAM]};
                    float *d[5] = {ct.r_a, ct.r_b, ct.c0, ct.penalty, ct.lambda};
                    for (int z = 0; z < 5; ++z)
                    {
                        d[z][0] = f[z]->x;
                        d[z][1] = f[z]->y;
                        d[z][2] = f[z]->z;
                    }
                    ct.stick = r[CR_LAM].w != 0.0f ? 1 : 0;
                }
            }
        }
    }
    return AVBD_OK;
}

int avbd_prepare(avbd_handle *h)
{
    if (!h)
        return fail(AVBD_ERR_INVALID_ARGUMENT, "avbd_prepare: null handle");
    if (h->injectFault)
        return fail(AVBD_ERR_INJECTED, "avbd_prepare: injected fault");
    int r = pick_device(h);
    return r ? r : enqueue_prepare(h);
}

int avbd_iterate(avbd_handle *h)
{
    if (!h)
        return fail(AVBD_ERR_INVALID_ARGUMENT, "avbd_iterate: null handle");
    int r = pick_device(h);
    return r ? r : enqueue_iterate(h);
}

int avbd_finish(avbd_handle *h)
{
    if (!h)
        return fail(AVBD_ERR_INVALID_ARGUMENT, "avbd_finish: null handle");
    int r = pick_device(h);
    return r ? r : enqueue_finish(h);
}

int avbd_step(avbd_handle *h, int32_t num_steps)
{
    if (!h)
        return fail(AVBD_ERR_INVALID_ARGUMENT, "avbd_step: null handle");
    if (num_steps < 0)
        return fail(AVBD_ERR_INVALID_ARGUMENT, "avbd_step: negative step count");
    if (h->injectFault)
        return fail(AVBD_ERR_INJECTED, "avbd_step: injected fault");
    int r = pick_device(h);
    if (r)
        return r;
    for (int k = 0; k < num_steps; ++k)
    {
        if ((r = enqueue_prepare(h)) || (r = enqueue_iterate(h)) || (r = enqueue_finish(h)))
            return r;
    }
    return AVBD_OK;
}

int avbd_synchronize(avbd_handle *h)
{
    if (!h)
        return fail(AVBD_ERR_INVALID_ARGUMENT, "avbd_synchronize: null handle");
    int r = pick_device(h);
    if (r)
        return r;
    HIP_TRY(hipStreamSynchronize(h->stream));
    return check_device_errors(h);
}

int avbd_get_stats(avbd_handle *h, avbd_stats *out)
{
    if (!h || !out)
        return fail(AVBD_ERR_INVALID_ARGUMENT, "avbd_get_stats: null argument");
    int r = pick_device(h);
    if (r)
        return r;
    Counters c;
    HIP_TRY(hipMemcpyAsync(&c, h->ctr, sizeof(c), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    std::vector<float4> mm(h->nb);
    if (h->nb)
        HIP_TRY(hipMemcpy(mm.data(), h->B.massMoment, sizeof(float4) * h->nb, hipMemcpyDeviceToHost));
    int active = 0;
    for (const float4 &m : mm)
        active += m.x > 0.0f ? 1 : 0;
    std::memset(out, 0, sizeof(*out));
    out->body_count = h->nb;
    out->active_body_count = active;
    out->solved_body_count = c.solvedCount;
    out->pair_checks = c.pairChecks;
    out->sphere_hits = c.numHits;
    out->constrained_hits = c.numHits - c.numCreated;
    out->manifolds_created = c.numCreated;
    out->manifold_count = c.numManifolds;
    out->contact_count = c.contactCount;
    out->joint_count = h->nj;
    out->spring_count = h->ns;
    out->color_count = c.colorCount;
    out->coloring_rounds = c.coloringRounds;
    out->sweep_axis = c.sweepAxis;
    out->axis_fallbacks = c.axisFallbacks;
    out->steps_taken = h->stepsTaken;
    if (h->timing && h->stepsTaken > 0)
    {
        float t[5] = {0, 0, 0, 0, 0};
        for (int k = 0; k < 5; ++k)
            if (hipEventElapsedTime(&t[k], h->ev[k], h->ev[k + 1]) != hipSuccess)
                t[k] = 0.0f;
        out->broadphase_ms = t[0];
        out->narrowphase_ms = t[1];
        out->coloring_ms = t[2];
        out->iterate_ms = t[3];
        out->finish_ms = t[4];
        float total = 0.0f;
        if (hipEventElapsedTime(&total, h->ev[0], h->ev[5]) == hipSuccess)
            out->step_ms = total;
    }
    out->iterate_bytes = c.lastIterBytes;
    out->step_bytes = c.lastStepBytes;
    out->iterate_bytes_total = c.cumIterBytes;
    out->primal_bytes_total = c.cumPrimalBytes;
    out->dual_bytes_total = c.cumDualBytes;
    out->step_bytes_total = c.cumStepBytes;
    out->counted_steps = c.cumSteps;
    if (h->timing && h->ringSteps > 0)
    {
        long long n = std::min<long long>(h->ringSteps, avbd_handle::kRing);
        double itTot = 0.0, stTot = 0.0, prTot = 0.0, duTot = 0.0;
        long long prLaunch = 0, duLaunch = 0, prIters = 0;
        for (long long k = h->ringSteps - n; k < h->ringSteps; ++k)
        {
            hipEvent_t *e = &h->ring[avbd_handle::kSlots * (k % avbd_handle::kRing)];
            float a = 0.0f, b = 0.0f;
            HIP_TRY(hipEventElapsedTime(&a, e[1], e[2]));
            HIP_TRY(hipEventElapsedTime(&b, e[0], e[3]));
            itTot += a;
            stTot += b;
            int iters = h->ringIters[k % avbd_handle::kRing];
            if (iters > 0)
            {
                // the last iteration: e[4] before its primal launches, e[5] before its dual, e[6] after
                float p = 0.0f, d = 0.0f;
                HIP_TRY(hipEventElapsedTime(&p, e[4], e[5]));
                HIP_TRY(hipEventElapsedTime(&d, e[5], e[6]));
                prTot += p;
                duTot += d;
            }
            prLaunch += h->ringPrimalLaunches[k % avbd_handle::kRing];
            if (iters > 0)
                prIters += h->ringPrimalIters[k % avbd_handle::kRing];
            duLaunch += iters;
        }
        out->iterate_ms_total = itTot;
        out->step_ms_total = stTot;
        out->timed_steps = n;
        out->primal_ms_total = prTot;
        out->dual_ms_total = duTot;
        out->primal_launches = prLaunch;
        out->primal_iterations = prIters;
        out->dual_launches = duLaunch;
    }
    if (c.barrierTimeout || c.overflow)
        return check_device_errors(h);
    return AVBD_OK;
}

__global__ void k_reset_cumulative(Counters *c)
{
    c->cumSteps = 0;
    c->cumIterBytes = 0.0;
    c->cumStepBytes = 0.0;
    c->cumPrimalBytes = 0.0;
    c->cumDualBytes = 0.0;
}

int avbd_reset_stats(avbd_handle *h)
{
    if (!h)
        return fail(AVBD_ERR_INVALID_ARGUMENT, "avbd_reset_stats: null handle");
    int r = pick_device(h);
    if (r)
        return r;
    k_reset_cumulative<<<1, 1, 0, h->stream>>>(h->ctr);
    HIP_TRY(hipStreamSynchronize(h->stream));
    h->ringSteps = 0;
    return AVBD_OK;
}

int avbd_get_colors(avbd_handle *h, int32_t *colors, int32_t capacity)
{
    if (!h || (!colors && capacity > 0))
        return fail(AVBD_ERR_INVALID_ARGUMENT, "avbd_get_colors: null argument");
    int r = pick_device(h);
    if (r)
        return r;
    HIP_TRY(hipStreamSynchronize(h->stream));
    int n = std::min(capacity, h->nb);
    if (n > 0)
        HIP_TRY(hipMemcpy(colors, h->S.color, sizeof(int) * n, hipMemcpyDeviceToHost));
    return AVBD_OK;
}

int avbd_set_ignore_pairs(avbd_handle *h, const int32_t *pairs, int32_t count)
{
    if (!h || count < 0 || (count > 0 && !pairs))
        return fail(AVBD_ERR_INVALID_ARGUMENT, "avbd_set_ignore_pairs: null handle or bad pair table");
    for (int k = 0; k < count; ++k)
    {
        int a = pairs[2 * k], b = pairs[2 * k + 1];
        if (a < 0 || b < 0 || a >= h->nb || b >= h->nb || a == b)
            return fail(AVBD_ERR_INVALID_ARGUMENT, "avbd_set_ignore_pairs: body index out of range");
    }
    int r = pick_device(h);
    if (r)
        return r;
    HIP_TRY(hipStreamSynchronize(h->stream));
    // the constrained-pair hash holds manifolds, joints, springs and these pairs at most half full
    int64_t need = 2 * ((int64_t)h->maxManifolds + h->maxJoints + h->maxSprings + count);
    if (need > h->hashSize)
    {
        int hs = h->hashSize;
        while (hs < need)
            hs <<= 1;
        unsigned long long *nh = nullptr;
        if ((r = h->alloc(nh, (size_t)hs)))
            return r;
        h->S.hash = nh;
        h->hashSize = hs;
        h->S.hashMask = hs - 1;
    }
    if (count > h->ignoreCap)
    {
        if ((r = h->alloc(h->ignorePairs, (size_t)count)))
            return r;
        h->ignoreCap = count;
    }
    std::vector<int> degs(std::max(h->nb, 1), 0);
    for (int k = 0; k < count; ++k)
    {
        ++degs[pairs[2 * k]];
        ++degs[pairs[2 * k + 1]];
    }
    if (count > 0)
        HIP_TRY(hipMemcpy(h->ignorePairs, pairs, sizeof(int2) * count, hipMemcpyHostToDevice));
    if (h->nb > 0)
        HIP_TRY(hipMemcpy(h->ignoreDeg, degs.data(), sizeof(int) * h->nb, hipMemcpyHostToDevice));
    h->numIgnore = count;
    return AVBD_OK;
}

int64_t avbd_snapshot_avbs_size(avbd_handle *h, int32_t scene_bytes)
{
    if (!h || scene_bytes < 0)
        return fail(AVBD_ERR_INVALID_ARGUMENT, "avbd_snapshot_avbs_size: null handle or negative scene length");
    return AVBD_AVBS_HEADER_BYTES + (int64_t)scene_bytes + 64 * (int64_t)h->nb;
}

int64_t avbd_snapshot_avbs(avbd_handle *h, uint64_t frame, const char *scene, void *out, int64_t capacity)
{
    if (!h || !out)
        return fail(AVBD_ERR_INVALID_ARGUMENT, "avbd_snapshot_avbs: null argument");
    size_t sb = scene ? std::strlen(scene) : 0;
    if (sb > 0xFFFFFFFFu)
        return fail(AVBD_ERR_INVALID_ARGUMENT, "avbd_snapshot_avbs: scene name too long");
    int64_t need = AVBD_AVBS_HEADER_BYTES + (int64_t)sb + 64 * (int64_t)h->nb;
    if (capacity < need)
        return fail(AVBD_ERR_CAPACITY, "avbd_snapshot_avbs: output buffer smaller than avbd_snapshot_avbs_size");
    int r = pick_device(h);
    if (r)
        return r;
    if (!h->avbsBuf && (r = h->alloc(h->avbsBuf, 4 * (size_t)h->maxBodies)))
        return r;
    // header (viewer_bridge.cpp:807-817): magic, version, frame, scene bytes, body count, name
    unsigned char *o = (unsigned char *)out;
    uint32_t magic = AVBD_AVBS_MAGIC, version = 1, sceneBytes = (uint32_t)sb, bodies = (uint32_t)h->nb;
    std::memcpy(o, &magic, 4);
    std::memcpy(o + 4, &version, 4);
    std::memcpy(o + 8, &frame, 8);
    std::memcpy(o + 16, &sceneBytes, 4);
    std::memcpy(o + 20, &bodies, 4);
    if (sb)
        std::memcpy(o + AVBD_AVBS_HEADER_BYTES, scene, sb);
    if (h->nb > 0)
    {
        k_avbs_pack<<<ceil_div(h->nb, 256), 256, 0, h->stream>>>(h->B, h->avbsRank, h->nb, h->avbsBuf);
        HIP_TRY(hipMemcpyAsync(o + AVBD_AVBS_HEADER_BYTES + sb, h->avbsBuf, 64 * (size_t)h->nb, hipMemcpyDeviceToHost,
                               h->stream));
    }
    return need;
}

int64_t avbd_snapshot_avbs_static(avbd_handle *h, uint64_t frame, const char *scene, void *out, int64_t capacity)
{
    if (!h || !out)
        return fail(AVBD_ERR_INVALID_ARGUMENT, "avbd_snapshot_avbs_static: null argument");
    size_t sb = scene ? std::strlen(scene) : 0;
    if (sb > 0xFFFFFFFFu)
        return fail(AVBD_ERR_INVALID_ARGUMENT, "avbd_snapshot_avbs_static: scene name too long");
    if (h->nb > kStaticSnapshotBodies)
        return fail(AVBD_ERR_CAPACITY, "avbd_snapshot_avbs_static: " + std::to_string(h->nb) +
                                           " bodies, more than the static staging holds (" +
                                           std::to_string(kStaticSnapshotBodies) + "); use avbd_snapshot_avbs");
    int64_t need = AVBD_AVBS_HEADER_BYTES + (int64_t)sb + 64 * (int64_t)h->nb;
    if (capacity < need)
        return fail(AVBD_ERR_CAPACITY, "avbd_snapshot_avbs_static: output buffer smaller than avbd_snapshot_avbs_size");
    int r = pick_device(h);
    if (r)
        return r;
    uint4 *stage = static_scratch(h->device);
    if (!stage)
        return fail(AVBD_ERR_UNSUPPORTED, "avbd_snapshot_avbs_static: no static staging on device ordinal " +
                                              std::to_string(h->device));
    // the copy out of the staging must not need a runtime staging buffer of its own: pinned host memory
    hipPointerAttribute_t attr{};
    hipError_t e = hipPointerGetAttributes(&attr, out);
    if (e != hipSuccess)
        (void)hipGetLastError();
    if (e != hipSuccess || attr.type != hipMemoryTypeHost)
        return fail(AVBD_ERR_INVALID_ARGUMENT, "avbd_snapshot_avbs_static: out must be pinned host memory (hipHostMalloc)");
    unsigned char *o = (unsigned char *)out;
    uint32_t magic = AVBD_AVBS_MAGIC, version = 1, sceneBytes = (uint32_t)sb, bodies = (uint32_t)h->nb;
    std::memcpy(o, &magic, 4);
    std::memcpy(o + 4, &version, 4);
    std::memcpy(o + 8, &frame, 8);
    std::memcpy(o + 16, &sceneBytes, 4);
    std::memcpy(o + 20, &bodies, 4);
    if (sb)
        std::memcpy(o + AVBD_AVBS_HEADER_BYTES, scene, sb);
    if (h->nb > 0)
    {
        // every record in [0, nb) is written (the ranks are a permutation), so nothing is cleared
        k_avbs_pack<<<ceil_div(h->nb, 256), 256, 0, h->stream>>>(h->B, h->avbsRank, h->nb, stage);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(o + AVBD_AVBS_HEADER_BYTES + sb, stage, 64 * (size_t)h->nb, hipMemcpyDeviceToHost,
                               h->stream));
    }
    return need;
}

int avbd_get_hits(avbd_handle *h, int32_t *triples, int32_t capacity, int32_t *count)
{
    if (!h || !count)
        return fail(AVBD_ERR_INVALID_ARGUMENT, "avbd_get_hits: null argument");
    int r = pick_device(h);
    if (r)
        return r;
    Counters c;
    HIP_TRY(hipMemcpyAsync(&c, h->ctr, sizeof(c), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    int H = h->nb >= 2 ? std::min(c.numHits, h->maxPairs) : 0;
    *count = H;
    int n = std::min(H, capacity);
    if (n > 0 && triples)
    {
        std::vector<int2> hits(n);
        std::vector<int> fresh(n);
        HIP_TRY(hipMemcpy(hits.data(), h->S.hits, sizeof(int2) * n, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(fresh.data(), h->S.hitNew, sizeof(int) * n, hipMemcpyDeviceToHost));
        for (int k = 0; k < n; ++k)
        {
            triples[3 * k] = hits[k].x;
            triples[3 * k + 1] = hits[k].y;
            triples[3 * k + 2] = fresh[k];
        }
    }
    return AVBD_OK;
}

int avbd_batch_upload(avbd_handle *h, const int32_t *world_body_offsets, int32_t num_worlds,
                      const avbd_body *bodies, int32_t num_bodies, const avbd_joint *joints, int32_t num_joints,
                      const avbd_spring *springs, int32_t num_springs, const avbd_manifold *manifolds,
                      int32_t num_manifolds, int64_t next_serial)
{
    if (!h || !world_body_offsets)
        return fail(AVBD_ERR_INVALID_ARGUMENT, "avbd_batch_upload: null argument");
    if (num_worlds < 1 || num_worlds > std::max(num_bodies, 1) || num_bodies < 0)
        return fail(AVBD_ERR_INVALID_ARGUMENT, "avbd_batch_upload: world count must be in [1, num_bodies]");
    if (world_body_offsets[0] != 0 || world_body_offsets[num_worlds] != num_bodies)
        return fail(AVBD_ERR_INVALID_ARGUMENT, "avbd_batch_upload: offsets must run from 0 to num_bodies");
    if (num_worlds > 1 && h->broadphaseMode != AVBD_BROADPHASE_SAP)
        return fail(AVBD_ERR_UNSUPPORTED, "avbd_batch_upload: batched worlds use sweep and prune "
                                          "(avbd_set_broadphase(h, AVBD_BROADPHASE_SAP) first)");
    if (num_worlds > 1 && (h->extPairsOn || !h->extContacts.empty()))
        return fail(AVBD_ERR_UNSUPPORTED, "avbd_batch_upload: external pairs or contacts are for a single world "
                                          "(clear them first)");
    std::vector<int> worldOf(std::max(num_bodies, 1), 0);
    for (int w = 0; w < num_worlds; ++w)
    {
        int b = world_body_offsets[w], e = world_body_offsets[w + 1];
        if (e <= b)
            return fail(AVBD_ERR_INVALID_ARGUMENT, "avbd_batch_upload: every world needs at least one body");
        if (num_worlds > 1 && e - b > AVBD_MAX_WORLD_BODIES)
            return fail(AVBD_ERR_UNSUPPORTED, "avbd_batch_upload: a batched world holds at most "
                                              "AVBD_MAX_WORLD_BODIES bodies");
        for (int i = b; i < e; ++i)
            worldOf[i] = w;
    }
    auto inRange = [&](int b) { return b >= 0 && b < num_bodies; };
    for (int i = 0; i < num_joints; ++i)
    {
        int a = joints ? joints[i].body_a : -1, b = joints ? joints[i].body_b : -1;
        if (inRange(a) && inRange(b) && worldOf[a] != worldOf[b])
            return fail(AVBD_ERR_INVALID_ARGUMENT, "avbd_batch_upload: joint spans two worlds");
    }
    for (int i = 0; i < num_springs; ++i)
    {
        int a = springs ? springs[i].body_a : -1, b = springs ? springs[i].body_b : -1;
        if (inRange(a) && inRange(b) && worldOf[a] != worldOf[b])
            return fail(AVBD_ERR_INVALID_ARGUMENT, "avbd_batch_upload: spring spans two worlds");
    }
    for (int i = 0; i < num_manifolds; ++i)
    {
        int a = manifolds ? manifolds[i].body_a : -1, b = manifolds ? manifolds[i].body_b : -1;
        if (inRange(a) && inRange(b) && worldOf[a] != worldOf[b])
            return fail(AVBD_ERR_INVALID_ARGUMENT, "avbd_batch_upload: manifold spans two worlds");
    }
    int r = avbd_upload_state(h, bodies, num_bodies, joints, num_joints, springs, num_springs, manifolds,
                              num_manifolds, next_serial);
    if (r)
        return r;
    HIP_TRY(hipMemcpy(h->worldStart, world_body_offsets, sizeof(int) * ((size_t)num_worlds + 1),
                      hipMemcpyHostToDevice));
    h->numWorlds = num_worlds;
    h->worldStartHost.assign(world_body_offsets, world_body_offsets + num_worlds + 1);
    h->maxWorldBodies = 0;
    for (int w = 0; w < num_worlds; ++w)
        h->maxWorldBodies = std::max(h->maxWorldBodies, world_body_offsets[w + 1] - world_body_offsets[w]);
    return AVBD_OK;
}

int avbd_set_broadphase(avbd_handle *h, int32_t mode)
{
    if (!h)
        return fail(AVBD_ERR_INVALID_ARGUMENT, "avbd_set_broadphase: null handle");
    if (mode != AVBD_BROADPHASE_ALL_PAIRS && mode != AVBD_BROADPHASE_GRID && mode != AVBD_BROADPHASE_SAP)
        return fail(AVBD_ERR_INVALID_ARGUMENT, "avbd_set_broadphase: unknown mode " + std::to_string(mode));
    if (mode != AVBD_BROADPHASE_SAP && h->numWorlds > 1)
        return fail(AVBD_ERR_UNSUPPORTED, "avbd_set_broadphase: batched worlds use sweep and prune");
    int r = pick_device(h);
    if (r)
        return r;
    if (mode != AVBD_BROADPHASE_SAP && !h->gridInfo)
    {
        if ((r = h->alloc(h->gridInfo, 1)) || (r = h->alloc(h->cellStart, (size_t)kGridMaxCells)) ||
            (r = h->alloc(h->largeLo, h->maxBodies)) || (r = h->alloc(h->largeHi, h->maxBodies)) ||
            (r = h->alloc(h->largeC, h->maxBodies)))
            return r;
        HIP_TRY(hipHostMalloc((void **)&h->hostGridInfo, sizeof(GridInfo)));
        HIP_TRY(hipMemsetAsync(h->cellStart, 0xFF, sizeof(unsigned) * kGridMaxCells, h->stream));
        HIP_TRY(hipStreamSynchronize(h->stream));
    }
    h->broadphaseMode = mode;
    return AVBD_OK;
}

int avbd_set_external_pairs(avbd_handle *h, const int32_t *pairs, int32_t count)
{
    if (!h || count < 0 || (count > 0 && !pairs))
        return fail(AVBD_ERR_INVALID_ARGUMENT, "avbd_set_external_pairs: null handle, bad count or null pairs");
    if (h->numWorlds > 1)
        return fail(AVBD_ERR_UNSUPPORTED, "avbd_set_external_pairs: external pairs are for a single world");
    if (count > h->maxPairs)
        return fail(AVBD_ERR_CAPACITY, "avbd_set_external_pairs: more pairs than the handle's pair capacity");
    h->extPairs.resize(count);
    for (int k = 0; k < count; ++k)
        h->extPairs[k] = make_int2(pairs[2 * k], pairs[2 * k + 1]);
    h->extPairsOn = true;
    h->extDirty = true;
    return AVBD_OK;
}

int avbd_clear_external_pairs(avbd_handle *h)
{
    if (!h)
        return fail(AVBD_ERR_INVALID_ARGUMENT, "avbd_clear_external_pairs: null handle");
    h->extPairs.clear();
    h->extPairsOn = false;
    h->extDirty = true;
    return AVBD_OK;
}

int avbd_set_external_contacts(avbd_handle *h, const avbd_external_contact *contacts, int32_t count)
{
    if (!h || count < 0 || (count > 0 && !contacts))
        return fail(AVBD_ERR_INVALID_ARGUMENT, "avbd_set_external_contacts: null handle, bad count or null table");
    if (count > 0 && h->numWorlds > 1)
        return fail(AVBD_ERR_UNSUPPORTED, "avbd_set_external_contacts: external contacts are for a single world");
    if (count > h->maxPairs)
        return fail(AVBD_ERR_CAPACITY, "avbd_set_external_contacts: more entries than the handle's pair capacity");
    h->extContacts.assign(contacts, contacts + count);
    h->extDirty = true;
    return AVBD_OK;
}

int avbd_get_broadphase(avbd_handle *h, int32_t *mode)
{
    if (!h || !mode)
        return fail(AVBD_ERR_INVALID_ARGUMENT, "avbd_get_broadphase: null argument");
    *mode = h->broadphaseMode;
    return AVBD_OK;
}

int avbd_batch_num_worlds(avbd_handle *h, int32_t *num_worlds)
{
    if (!h || !num_worlds)
        return fail(AVBD_ERR_INVALID_ARGUMENT, "avbd_batch_num_worlds: null argument");
    *num_worlds = h->numWorlds;
    return AVBD_OK;
}

int avbd_set_launch_config(avbd_handle *h, const avbd_launch_config *cfg)
{
    if (!h || !cfg)
        return fail(AVBD_ERR_INVALID_ARGUMENT, "avbd_set_launch_config: null argument");
    if (!valid_launch(*cfg))
        return fail(AVBD_ERR_INVALID_ARGUMENT, "avbd_set_launch_config: primal lanes must be 8/16/32, thread "
                                               "counts 64/128/256 and use_graph 0 or 1");
    h->launch = *cfg;
    return AVBD_OK;
}

int avbd_get_launch_config(avbd_handle *h, avbd_launch_config *out)
{
    if (!h || !out)
        return fail(AVBD_ERR_INVALID_ARGUMENT, "avbd_get_launch_config: null argument");
    *out = h->launch;
    return AVBD_OK;
}

int avbd_autotune(avbd_handle *h, int32_t trial_steps, int32_t flags, avbd_launch_config *out)
{
    if (!h)
        return fail(AVBD_ERR_INVALID_ARGUMENT, "avbd_autotune: null handle");
    if (trial_steps < 1)
        return fail(AVBD_ERR_INVALID_ARGUMENT, "avbd_autotune: trial_steps must be at least 1");
    int r = avbd_synchronize(h);
    if (r)
        return r;
    // snapshot of the resident state through the public tables, restored after the trials
    int32_t nb = 0, nj = 0, ns = 0, nm = 0;
    int64_t next = 0;
    if ((r = avbd_state_counts(h, &nb, &nj, &ns, &nm, &next)))
        return r;
    std::vector<avbd_body> bodies(std::max(nb, 1));
    std::vector<avbd_joint> joints(std::max(nj, 1));
    std::vector<avbd_spring> springs(std::max(ns, 1));
    std::vector<avbd_manifold> manifolds(std::max(nm, 1));
    if ((r = avbd_download_state(h, bodies.data(), joints.data(), springs.data(), manifolds.data())))
        return r;
    const std::vector<int> worlds = h->worldStartHost;
    const int numWorlds = h->numWorlds;
    auto restore = [&]() {
        return avbd_batch_upload(h, worlds.data(), numWorlds, bodies.data(), nb, joints.data(), nj, springs.data(), ns,
                                 manifolds.data(), nm, next);
    };
    const avbd_launch_config start = h->launch;
    hipEvent_t t0 = h->ev[6], t1 = h->ev[7];
    // Trials continue from the state the previous trial left (every shape steps to the same bits);
    // the snapshot is restored once at the end. Restoring before each trial would put a
    // host-to-device upload, and an idle GPU, in front of every timed window.
    auto trial = [&](const avbd_launch_config &c, float &ms) -> int {
        int rr;
        h->launch = c;
        if ((rr = avbd_step(h, 1)) || (rr = avbd_synchronize(h)))
            return rr;
        HIP_TRY(hipEventRecord(t0, h->stream));
        if ((rr = avbd_step(h, trial_steps)))
            return rr;
        HIP_TRY(hipEventRecord(t1, h->stream));
        if ((rr = avbd_synchronize(h)))
            return rr;
        HIP_TRY(hipEventElapsedTime(&ms, t0, t1));
        return AVBD_OK;
    };
    const bool allowGraph = (flags & AVBD_TUNE_NO_GRAPH) == 0;
    avbd_launch_config best = start;
    float bestMs = 0.0f;
    if ((r = trial(best, bestMs)))
        return r;
    // primal shape first (with the dual at the current shape), then the dual, then the graph form.
    // Every primal shape is timed twice, the second pass in reverse order, and keeps its faster
    // time, so that a clock change during the sweep does not decide the shape.
    struct Shape
    {
        int lanes, threads;
        float ms;
    };
    std::vector<Shape> shapes;
    for (int lanes : {0, 8, 16, 32})  // 0: lanes per body by stamp count
        for (int threads : {64, 128, 256})
            shapes.push_back({lanes, threads, 3.0e38f});
    for (int pass = 0; pass < 2; ++pass)
        for (size_t q = 0; q < shapes.size(); ++q)
        {
            Shape &sh = shapes[pass == 0 ? q : shapes.size() - 1 - q];
            avbd_launch_config c = best;
            c.primal_lanes = sh.lanes;
            c.primal_threads = sh.threads;
            float ms = 0.0f;
            if ((r = trial(c, ms)))
                return r;
            sh.ms = std::min(sh.ms, ms);
        }
    {
        const avbd_launch_config base = best;
        for (const Shape &sh : shapes)
            if (sh.ms < bestMs)
            {
                bestMs = sh.ms;
                best = base;
                best.primal_lanes = sh.lanes;
                best.primal_threads = sh.threads;
            }
    }
    for (int threads : {64, 128, 256})
    {
        avbd_launch_config c = best;
        c.dual_threads = threads;
        float ms = 0.0f;
        if ((r = trial(c, ms)))
            return r;
        if (ms < bestMs)
        {
            bestMs = ms;
            best = c;
        }
    }
    for (int lz : {0, 1})
    {
        avbd_launch_config c = best;
        c.lazy_dual = lz;
        float ms = 0.0f;
        if ((r = trial(c, ms)))
            return r;
        if (ms < bestMs)
        {
            bestMs = ms;
            best = c;
        }
    }
    // the flow form (one primal launch per step) for contact-only worlds, at the tuned shape and at
    // two shapes that suit its resident grid
    if (h->nj == 0 && h->ns == 0 && best.lazy_dual == 1)
    {
        const avbd_launch_config base = best;
        const int flowShapes[3][2] = {{base.primal_lanes, base.primal_threads}, {16, 64}, {16, 128}};
        for (int q = 0; q < 3; ++q)
        {
            if (q > 0 && flowShapes[q][0] == base.primal_lanes && flowShapes[q][1] == base.primal_threads)
                continue;
            avbd_launch_config c = base;
            c.dataflow = 1;
            c.primal_lanes = flowShapes[q][0];
            c.primal_threads = flowShapes[q][1];
            float ms = 0.0f;
            if ((r = trial(c, ms)))
                return r;
            if (ms < bestMs)
            {
                bestMs = ms;
                best = c;
            }
        }
    }
    for (int g : {0, 1})
    {
        if (g == 1 && (!allowGraph || best.dataflow))
            continue;
        avbd_launch_config c = best;
        c.use_graph = g;
        float ms = 0.0f;
        if ((r = trial(c, ms)))
            return r;
        if (ms < bestMs)
        {
            bestMs = ms;
            best = c;
        }
    }
    if ((r = restore()))
        return r;
    h->launch = best;
    h->stepsTaken = 0;
    if ((r = avbd_reset_stats(h)))
        return r;
    if (out)
        *out = best;
    return AVBD_OK;
}

int avbd_debug_inject_fault(avbd_handle *h, int32_t code)
{
    if (!h)
        return fail(AVBD_ERR_INVALID_ARGUMENT, "avbd_debug_inject_fault: null handle");
    h->injectFault = code;
    return AVBD_OK;
}

int avbd_debug_narrowphase_capacity(avbd_handle *h, int32_t vertices)
{
    if (!h)
        return fail(AVBD_ERR_INVALID_ARGUMENT, "avbd_debug_narrowphase_capacity: null handle");
    if (vertices < 1 || vertices > kNpPolyCap)
        return fail(AVBD_ERR_INVALID_ARGUMENT, "avbd_debug_narrowphase_capacity: vertices must be in 1..8");
    h->npPolyCap = vertices;
    return AVBD_OK;
}

int avbd_sizeof(int which)
{
    switch (which)
    {
    case 0: return (int)sizeof(avbd_body);
    case 1: return (int)sizeof(avbd_contact);
    case 2: return (int)sizeof(avbd_manifold);
    case 3: return (int)sizeof(avbd_joint);
    case 4: return (int)sizeof(avbd_spring);
    case 5: return (int)sizeof(avbd_params);
    case 6: return (int)sizeof(avbd_launch_config);
    case 7: return (int)sizeof(avbd_stats);
    case 8: return (int)sizeof(avbd_config);
    case 9: return (int)sizeof(avbd_external_contact);
    default: return -1;
    }
}

} // extern "C"

#include "avbd_adjoint.hip"
