// d3q27.h — D3Q27 velocity set, shared by host and device code.
//
// Ordering, opposite table and weights are those of initial.cu:276-287 (rest, 6 faces, 12 edges,
// 8 corners).  Everything is constexpr so that fully unrolled loops fold the velocity
// components into the instruction stream.
#pragma once

#include <hip/hip_runtime.h>

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

// Weights exactly as the reference evaluates them (8.0/27.0 etc. in double).
__host__ __device__ constexpr double weight(int i) {
    return i == 0 ? 8.0 / 27.0 : (i <= 6 ? 2.0 / 27.0 : (i <= 18 ? 1.0 / 54.0 : 1.0 / 216.0));
}

}  // namespace rbtp
