// hip_avbd_backend.h -- "hip-avbd" PhysicsBackend for the reference Solver.
//
// Drop-in replacement for makeWebGpuAvbdPhysicsBackend (avbd_gpu_solver.h:15)
// and makeCpuReferencePhysicsBackend (solver.h:417): compile this file in the
// reference tree (it includes the reference's solver.h) and install it with
//     solver.physicsBackend = makeHipAvbdPhysicsBackend();
// Each Solver::step then packs the Solver's object graph into the C ABI of
// include/avbd_hip.h, runs one full resident step on the GPU (broadphase to
// velocity update) and writes the result back into the same objects
// (bodies, joints, manifolds created/updated/deleted in creation order).
//
// Like the reference GPU lane (avbd_gpu_solver.cpp:2112-2170) it never throws:
// on any device error it prints to stderr, marks itself broken and steps the
// CPU reference from then on, counting the fallbacks.
#pragma once

#include <memory>

#include "solver.h"
#include "avbd_hip.h"

// Counters of the last step (AvbdGpuSolverStats, avbd_gpu_solver.h:19-41).
struct HipAvbdSolverStats
{
    avbd_stats gpu;
    int cpuIterateFallbacks;
    int uploads;
    float packMs;
    float gpuMs;
    float applyMs;
};

std::unique_ptr<PhysicsBackend> makeHipAvbdPhysicsBackend(int device = 0);
const HipAvbdSolverStats &hipAvbdSolverStats();
