// The layered Path B's frontier on the device (no reference counterpart; the
// flood it replaces is full_field_path_b.cpp:148-412).
//
// Per layer: the accepted nodes next to open ones, sorted in the flood's pop
// order (corr, then node index); a breadth-first search from them through
// the open nodes up to `reach` steps, in which a node takes the seed of the
// first node the search would pop next to it; one task per reached node
// whose claim is closer than its last failed one; after the solve, the
// layer's results update the accepted / finished / tried flags and the
// accepted warps.
//
// The search's tie-breaking is the FIFO queue's: a node's parent is its
// neighbour of the previous level that the queue pops first, and its place in
// the next level follows its parent's place, then the direction (right, left,
// down, up) it was reached in. That order is the lexicographic order of the
// path (source rank, direction, direction, ...), which packs into 64 bits for
// reach <= 16: the source's rank in the high word, then two bits per step from
// bit 31 down. A node's key is the minimum over its previous-level neighbours
// of (their key | its direction at its level), taken with a 64-bit atomicMin,
// so the search needs no ordering pass.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "common.hpp"

namespace sdic {

constexpr int kFloodDeviceMaxReach = 16;

// Counter slots of a layer (device ints): sources, reached, tasks, then the
// layer results: points accepted, IC-GN iterations, Simplex calls, saved,
// dead, and the committed nodes (the end-of-flood list).
enum FloodCounter {
    kFcSources = 0,
    kFcReached,
    kFcTasks,
    kFcPoints,
    kFcIters,
    kFcCalls,
    kFcSaved,
    kFcDead,
    kFcCommitted,
    kFcCount
};

struct FloodDev {
    int gw = 0, gh = 0;
    uint8_t* ok = nullptr;        // accepted
    uint8_t* fin = nullptr;       // solved, skipped or finally failed
    int8_t* tried = nullptr;      // a claim must come closer than this
    float* w = nullptr;           // the accepted warps: 7 planes u v ux uy vx vy corr
    int* stamp = nullptr;         // the layer that reached the node (-1: none)
    int* dist = nullptr;          // its distance in that layer's search
    int* cand = nullptr;          // the (layer, level) that listed it as a candidate
    unsigned long long* key = nullptr;   // search keys (~0: unset)
    unsigned long long* skeys = nullptr; // the sources' sort keys, and their sorted copy
    unsigned long long* skeys_sorted = nullptr;
    int* sources = nullptr;       // the sorted sources of the layer
    int* level_a = nullptr;       // the search's frontiers
    int* level_b = nullptr;
    int* reached = nullptr;       // the layer's reached nodes
    int* tdist = nullptr;         // each task's distance
    Task* tasks = nullptr;
    SolveOut* commit_res = nullptr;   // per node: the last claim's result
    uint8_t* committed = nullptr;     // per node: a result is held
    int* commit_nodes = nullptr;      // the end-of-flood list
    SolveOut* commit_list = nullptr;
    int* counters = nullptr;          // kFcCount ints
};

// Starts a flood: ok = accepted, fin = solved, tried = reach + 1, the warps
// from the node planes, every search and commit flag cleared.
hipError_t launch_flood_init(const FloodDev& f, const uint8_t* accepted, const uint8_t* solved, const float* planes,
                             int reach, hipStream_t st);
// The layer's sources (unordered) and their sort keys; counters[kFcSources].
hipError_t launch_flood_sources(const FloodDev& f, hipStream_t st);
// Sorts the sources' keys (temp: sort_temp_bytes(n) bytes) and unpacks them
// into f.sources.
size_t flood_sort_temp_bytes(int n);
hipError_t launch_flood_sort(const FloodDev& f, int nsources, void* temp, size_t temp_bytes, hipStream_t st);
// The search (one workgroup) and the layer's tasks; counters[kFcReached],
// counters[kFcTasks]. tag: the layer's number, from 1.
hipError_t launch_flood_search(const FloodDev& f, int nsources, int reach, int tag, int rect_x, int rect_y, int step,
                               bool learned, hipStream_t st);
// The layer's results into the flags, warps, counters and the commit state.
hipError_t launch_flood_update(const FloodDev& f, const SolveOut* res, int ntasks, hipStream_t st);
// The committed nodes and their results as a list, counters[kFcCommitted].
hipError_t launch_flood_commit_list(const FloodDev& f, hipStream_t st);

}  // namespace sdic
