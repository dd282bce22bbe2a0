// D3Q19 velocity set, weights and index maps for the river-reach solver.
//
// Direction order (floodsim/lbm3d.py:16-21): 0 rest; 1/2 +-x; 3/4 +-y; 5/6 +-z;
// 7..18 the twelve edge diagonals. OPP is the opposite direction, MIRZ the
// mirror image in a horizontal plane (used by the rigid lid).
// All tables are constexpr so that every loop over i unrolls to immediate
// operands: no constant-memory loads in the hot kernels.
#pragma once

#include <cstdint>

namespace wfs {

constexpr int Q = 19;
constexpr int kCX[Q] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
constexpr int kCY[Q] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1};
constexpr int kCZ[Q] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1};
constexpr int kOPP[Q] = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17};
constexpr int kMIRZ[Q] = {0, 1, 2, 3, 4, 6, 5, 7, 8, 9, 10, 13, 14, 11, 12, 17, 18, 15, 16};

// Lattice weight of direction i in precision T (1/3, 1/18, 1/36), each the
// correctly rounded quotient in T.
template <typename T>
__host__ __device__ constexpr T weight(int i) {
    return i == 0 ? T(1) / T(3) : (i < 7 ? T(1) / T(18) : T(1) / T(36));
}

// Centres of the 16-bit storage formats, passed by value to kernels. The
// centre of direction i depends only on its weight and c_ix, so there are
// seven: class 0 rest; 1..3 axis directions with c_x = 0, +1, -1; 4..6 edge
// directions with c_x = 0, +1, -1. c = e, n2 = -2 e (exact).
constexpr int kCenterClasses = 7;
__host__ __device__ constexpr int center_class(int i) {
    return i == 0 ? 0 : (i < 7 ? 1 : 4) + (kCX[i] == 0 ? 0 : (kCX[i] > 0 ? 1 : 2));
}
struct Center {
    float c[kCenterClasses];
    float n2[kCenterClasses];
};

// Per-cell flag byte.
enum : uint8_t {
    FLAG_SOLID = 1u << 0,     // solid cell: never updated
    FLAG_NEARWALL = 1u << 1,  // at least one pull source is a wall (bounce-back link)
    FLAG_CF = 1u << 2,        // wall-model friction coefficient > 0
    FLAG_FORCED = 1u << 3,    // actuator-disk body force present
    FLAG_BEDWALL = 1u << 4,   // bed or bank cell of the wall model (takes the column speed as uref)
};

}  // namespace wfs
