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
constexpr int kSweepChunk = 64;      // sorted intervals visited per sweep work item
constexpr int kScanBlocks = 256;     // fixed block count of the device-count scans
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

struct Bodies
{
    float4 *pos;      // xyz
    float4 *rot;      // quaternion xyzw
    float4 *initLin;  // x-
    float4 *initAng;
    float4 *inerLin;  // x~
    float4 *inerAng;
    float4 *vel;
    float4 *angVel;
    float4 *prevVel;
    float4 *massMoment; // mass, moment xyz
    float4 *sizeRad;    // full widths xyz, bounding radius
    float *friction;
    int *denseId;
};

struct Manifolds
{
    int2 *ab;            // body A, body B
    long long *serial;
    int *count;          // live contacts
    float4 *basis0;      // normal, w = friction
    float4 *basis1;      // tangent 1
    float4 *basis2;      // tangent 2
    int *feature;        // [cap * 8]
    float4 *rA;          // [cap * 8] local offsets
    float4 *rB;
    float4 *c0;          // constraint at x- (+margin)
    float4 *pen;         // penalty
    float4 *lam;         // lambda xyz, w = stick (0/1)
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
    int2 *stamps;           // [maxStamps] (adj entry, part)
    int *contactStart;      // [maxManifolds]
    int *contactOwner;      // [maxManifolds * 8]
    // colouring
    int *color;             // [N]
    int *colorOffsets;      // [256]
    int *colorBodies;       // [N]
    // scan partials
    int *partial;           // [kScanBlocks + 1]
};

inline __host__ __device__ int adjKind(int e) { return (int)((unsigned)e >> 30); }
inline __host__ __device__ int adjIndex(int e) { return e & 0x3FFFFFFF; }
inline __host__ __device__ int adjEncode(int kind, int idx) { return (int)(((unsigned)kind << 30) | (unsigned)idx); }

} // namespace avbd
