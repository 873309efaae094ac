// d3q27.h — D3Q27 velocity set, shared by host and device code.
//
// Ordering, opposite table and weights are those of initial.cu:276-287 (rest, 6 faces, 12 edges,
// 8 corners).  Everything is constexpr so that fully unrolled loops fold the velocity
// components into the instruction stream.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace rbtp {

constexpr int NPOP = 27;

struct D3Q27 {
    static constexpr int cx[NPOP] = {0, 1, 0, -1, 0, 0, 0, 1, -1, -1, 1, 1, 0, -1,
                                     0, 1, 0, -1, 0, 1, -1, -1, 1, 1, -1, -1, 1};
    static constexpr int cy[NPOP] = {0, 0, 1, 0, -1, 0, 0, 1, 1, -1, -1, 0, 1, 0,
                                     -1, 0, 1, 0, -1, 1, 1, -1, -1, 1, 1, -1, -1};
    static constexpr int cz[NPOP] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, 1,
                                     1, -1, -1, -1, -1, 1, 1, 1, 1, -1, -1, -1, -1};
    static constexpr int opp[NPOP] = {0, 3, 4, 1, 2, 6, 5, 9, 10, 7, 8, 17, 18, 15,
                                      16, 13, 14, 11, 12, 25, 26, 23, 24, 21, 22, 19, 20};
};

// The tables packed into integer constants, for a direction index that differs between lanes:
// a lookup is then a shift and a mask instead of a memory load.  Components are stored as c + 1
// in 2 bits; opposites in 5 bits, 12 to a word.
namespace packed {
constexpr uint64_t comp(const int (&c)[NPOP]) {
    uint64_t v = 0;
    for (int i = 0; i < NPOP; i++) v |= (uint64_t)(c[i] + 1) << (2 * i);
    return v;
}
constexpr uint64_t opp(int first) {
    uint64_t v = 0;
    for (int i = first; i < first + 12 && i < NPOP; i++) v |= (uint64_t)D3Q27::opp[i] << (5 * (i - first));
    return v;
}
constexpr uint64_t CX = comp(D3Q27::cx), CY = comp(D3Q27::cy), CZ = comp(D3Q27::cz);
constexpr uint64_t OPP0 = opp(0), OPP1 = opp(12), OPP2 = opp(24);
}  // namespace packed
__host__ __device__ inline int dir_cx(int i) { return (int)((packed::CX >> (2 * i)) & 3u) - 1; }
__host__ __device__ inline int dir_cy(int i) { return (int)((packed::CY >> (2 * i)) & 3u) - 1; }
__host__ __device__ inline int dir_cz(int i) { return (int)((packed::CZ >> (2 * i)) & 3u) - 1; }
__host__ __device__ inline int dir_opp(int i) {
    const uint64_t w = i < 12 ? packed::OPP0 : (i < 24 ? packed::OPP1 : packed::OPP2);
    const int sh = 5 * (i < 12 ? i : (i < 24 ? i - 12 : i - 24));
    return (int)((w >> sh) & 31u);
}

// Storage order of the populations.  f and g of one time level share a buffer (a "pair
// buffer").  Nodes are taken in runs of POP_RUN consecutive indices; a run keeps population 0
// of f for its nodes, then population 1, ..., population 26, then population 0 of g, ...,
// population 26 of g.  The g pointer of a pair is G_OFFSET elements after the f pointer, so
// pop_at(i, idx) addresses population i of either distribution relative to its own pointer.
// A workgroup's access to one population is one contiguous piece, and the 54 pieces a node reads
// sit together in HBM, instead of 27 streams N values apart in each of two arrays; the
// reference's order (pop[i*N + idx], initial.cu:276-279) is what the accessors and files see.
// The pull's access pattern alone, 320x320x160 on one MI355X: runs of 1024 nodes with f and g in
// separate arrays 5.29 TB/s (FP64) and 5.32 (16-bit storage); runs of 4096, separate 5.56 and
// 5.43; runs of 4096 with f and g interleaved 5.64 (5.78 with the FP64 row pull) and 5.54.
constexpr int POP_RUN_LOG2 = 12;
constexpr int POP_RUN = 1 << POP_RUN_LOG2;
constexpr int RUN_POPS = 2 * NPOP;
constexpr size_t G_OFFSET = (size_t)NPOP * POP_RUN;
__host__ __device__ inline size_t pop_at(int i, int idx) {
    return ((size_t)(idx >> POP_RUN_LOG2) * RUN_POPS + (size_t)i) * POP_RUN +
           (size_t)(idx & (POP_RUN - 1));
}
// Values in one pair buffer over n nodes (the last run padded).
__host__ __device__ inline size_t pop_buffer_len(size_t n) {
    return (size_t)RUN_POPS * ((n + POP_RUN - 1) / POP_RUN) * POP_RUN;
}

// Weights exactly as the reference evaluates them (8.0/27.0 etc. in double).
__host__ __device__ constexpr double weight(int i) {
    return i == 0 ? 8.0 / 27.0 : (i <= 6 ? 2.0 / 27.0 : (i <= 18 ? 1.0 / 54.0 : 1.0 / 216.0));
}

}  // namespace rbtp
