// avbd_device.hpp -- resident device layout of the AVBD step (structure of arrays in HBM).
//
// Bodies are indexed by body-list position (the reference's Solver::bodies
// order). Manifolds live in two ping-pong buffers kept in ascending creation
// serial; each manifold owns 8 contact slots at [m * 8, m * 8 + 8).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace avbd
{

constexpr int kMaxContacts = 8;
constexpr int kMaxColors = 255;      // colour 255 is the "unsolved" bucket of the colour sort
// Adaptive primal lanes: segment q + 1 of a colour holds the bodies with at most seg_stamps(q)
// (32, 16, 8) stamps; the segments' starts follow the 257 colour offsets in one device array.
constexpr int kSegCount = 3;
__host__ __device__ constexpr int seg_stamps(int q) { return 32 >> q; }
constexpr int kSegBase = 257;
constexpr int kOffsetsLen = kSegBase + kSegCount * 256;
constexpr int kSweepChunk = 64;      // sorted intervals visited per sweep work item
constexpr int kPadStampsMax = 64;  // largest per-task stride of the task-ordered stamp copy
constexpr int kRadixTile = 2048;     // keys per radix-sort block (256 threads x 8)
constexpr int kAdjManifold = 0;
constexpr int kAdjJoint = 1;
constexpr int kAdjSpring = 2;

// Solver parameters as the kernels read them.
struct Params
{
    float dt, gravity, alpha, betaLin, betaAng, gamma;
    int iterations;
};

// Gathered per-body pose record: the four float4 the solver iterations read for a
// neighbour sit in one 64-byte record, pose[4 i + k].
enum
{
    P_POS = 0,   // positionLin
    P_ROT = 1,   // positionAng (x, y, z, w)
    P_ILIN = 2,  // x- linear
    P_IANG = 3   // x- angular
};

struct Bodies
{
    float4 *pose;     // [4 N] records of P_POS, P_ROT, P_ILIN, P_IANG
    float4 *inerLin;  // x~
    float4 *inerAng;
    float4 *vel;
    float4 *angVel;
    float4 *prevVel;
    float4 *massMoment; // mass, moment xyz
    float4 *sizeRad;    // full widths xyz, bounding radius
    float4 *shape;      // RigidShape radius, halfLength, type (as float), 0
    float *friction;
    int *denseId;
};

// One 128-byte record per contact slot (8 float4): everything a primal stamp or a
// dual update of the contact reads, in one cache line. The manifold's basis and
// friction are repeated in each of its contacts.
constexpr int kCR = 8;
enum
{
    CR_RA = 0,   // rA (A's local frame), w = manifold friction
    CR_RB = 1,   // rB (B's local frame)
    CR_C0 = 2,   // C0 = basis (xA - xB) + (margin, 0, 0) at x-
    CR_N = 3,    // basis row 0: normal (B to A)
    CR_T1 = 4,   // basis row 1
    CR_T2 = 5,   // basis row 2
    CR_PEN = 6,  // penalty
    CR_LAM = 7   // lambda, w = stick (0/1)
};

struct Manifolds
{
    int2 *ab;            // body A, body B
    long long *serial;
    int *count;          // live contacts
    int *feature;        // [cap * 8]
    float4 *cr;          // [cap * 8 * kCR] contact records
};

struct Joint
{
    int a, b;            // a = -1: world anchor
    int broken, dead;    // broken: fractured in a dual pass; dead: removed at initialisation
    long long serial;
    float4 rA;           // w = stiffnessLin
    float4 rB;           // w = stiffnessAng
    float4 c0Lin;        // w = torqueArm
    float4 c0Ang;        // w = fracture
    float4 penLin, penAng, lamLin, lamAng;
};

struct Spring
{
    int a, b;
    long long serial;
    float4 rA;           // w = rest length
    float4 rB;           // w = stiffness
};

// Device-side counters of the current step (one instance per handle).
struct Counters
{
    int numManifolds;    // persistent manifolds before this step's creation
    int numTotal;        // persistent + created candidates
    int numKept;         // survivors after initialisation (next numManifolds)
    int numChunks;
    int numHits;
    int numCreated;
    int colorCount;
    int coloringRounds;
    int sweepAxis;
    int axisAmbiguous;
    int axisFallbacks;
    int pairChecks;
    int overflow;        // bit 0 pairs, bit 1 manifolds, bit 2 adjacency, bit 3 colours, bit 4 hash
    int barrierTimeout;
    int solvedCount;
    int contactCount;
    int contactEndpoints;   // sum over contacts of their dynamic endpoints
    int numStamps;
    int maxStampCount;      // the most stamps of one body this step
    int colorPending;       // k_color stopped at its round limit with colours still changing
    int npRedoCount;      // box-only narrowphase: manifolds whose clip polygon outgrew LDS
    int jointLive;
    int jointEndpoints;
    int springEndpoints;    // fixed by the uploaded springs
    long long nextSerial;
    long long cumSteps;
    double lastIterBytes;
    double lastStepBytes;
    double cumIterBytes;
    double cumStepBytes;
    double cumPrimalBytes;
    double cumDualBytes;
    unsigned int colorChanged[3];
    unsigned int pad_;
};

struct GridBarrier
{
    unsigned int count;
    unsigned int gen;
};

// All per-step scratch arrays.
struct Scratch
{
    // sweep and prune
    double *axisSums;       // 6 doubles
    unsigned int *keys;     // [N] sort keys
    unsigned int *keysAlt;
    int *vals;              // [N] body indices
    int *valsAlt;
    int *radixHist;         // [256 * tiles]
    float *sLo;             // sorted interval low end
    float *sHi;             // sorted interval high end
    float4 *sPos;           // sorted centre, w = radius
    int *sIdx;              // sorted body index
    int *chunkCount;        // [N] work items per interval
    int *chunkStart;        // [N]
    int *hitCount;          // [maxChunks] hits per work item
    unsigned long long *hitMask; // [maxChunks] bit k: the k-th interval visited by the work item is a hit
    int *chunkOwner;             // [maxChunks] the interval each sweep chunk belongs to
    int *npRedo;          // their indices, for the full-capacity narrowphase
    int *hitStart;
    int2 *hits;             // [maxPairs]
    int *hitNew;            // [maxPairs] 1 = unconstrained
    int *hitRank;
    // constrained-pair hash set
    unsigned long long *hash;
    int hashMask;
    // manifold initialisation
    int *keep;              // [maxManifolds]
    int *keepRank;
    // adjacency
    int *deg;               // [N]
    int *adjStart;          // [N]
    int *adjCursor;         // [N]
    int *adj;               // [maxAdj] (kind << 30) | index
    long long *adjSerial;   // [maxAdj]
    int *adjPartner;        // [maxAdj] the force's other body (-1: world)
    // primal stamp lists and the per-contact dual index
    int *stampCount;        // [N]
    int *stampStart;        // [N]
    int4 *stamps;           // [maxStamps] (adj entry, part | isA << 16, partner body, contact slot)
    int4 *tasks;            // [N] colour-sorted (body, first stamp, stamp count, 0)
    int *contactStart;      // [maxManifolds]
    int4 *contactRec;       // [maxManifolds * 8] (manifold, contact, body A, body B)
    // colouring
    int *color;             // [N]
    int *colorOffsets;      // [256]
    int *colorBodies;       // [N]
    // scan partials
};

inline __host__ __device__ int adjKind(int e) { return (int)((unsigned)e >> 30); }
inline __host__ __device__ int adjIndex(int e) { return e & 0x3FFFFFFF; }
inline __host__ __device__ int adjEncode(int kind, int idx) { return (int)(((unsigned)kind << 30) | (unsigned)idx); }

} // namespace avbd
