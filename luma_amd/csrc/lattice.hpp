// lattice.hpp -- D2Q9 and D3Q19 velocity sets in LUMA's population order.
//
// The order, weights and opposites restate LUMA/src/stdafx.cpp:81-153 and
// LUMA/src/GridUtils.cpp:54-67 (dir_opposites).  Keeping LUMA's order is what lets a host
// caller hand its AoS f[v + Q*id] arrays across the C ABI unchanged.
#pragma once

namespace lumagpu {

template <int Q>
struct Lattice;

template <>
struct Lattice<9> {
    static constexpr int D = 2;
    static constexpr int Q = 9;
    __host__ __device__ static constexpr int c(int v, int a) {
        constexpr int t[9][3] = {{1, 0, 0},  {-1, 0, 0}, {0, 1, 0},  {0, -1, 0}, {1, 1, 0},
                                 {-1, -1, 0}, {1, -1, 0}, {-1, 1, 0}, {0, 0, 0}};
        return t[v][a];
    }
    __host__ __device__ static constexpr int opp(int v) {
        constexpr int t[9] = {1, 0, 3, 2, 5, 4, 7, 6, 8};
        return t[v];
    }
    __host__ __device__ static constexpr double w(int v) {
        constexpr double t[9] = {1.0 / 9.0,  1.0 / 9.0,  1.0 / 9.0,  1.0 / 9.0, 1.0 / 36.0,
                                 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 4.0 / 9.0};
        return t[v];
    }
};

template <>
struct Lattice<19> {
    static constexpr int D = 3;
    static constexpr int Q = 19;
    __host__ __device__ static constexpr int c(int v, int a) {
        constexpr int t[19][3] = {{1, 0, 0},   {-1, 0, 0}, {0, 1, 0},   {0, -1, 0}, {0, 0, 1},
                                  {0, 0, -1},  {1, 1, 0},  {-1, -1, 0}, {1, -1, 0}, {-1, 1, 0},
                                  {0, 1, 1},   {0, -1, -1}, {0, 1, -1}, {0, -1, 1}, {1, 0, 1},
                                  {-1, 0, -1}, {-1, 0, 1}, {1, 0, -1},  {0, 0, 0}};
        return t[v][a];
    }
    __host__ __device__ static constexpr int opp(int v) {
        constexpr int t[19] = {1, 0, 3, 2, 5, 4, 7, 6, 9, 8, 11, 10, 13, 12, 15, 14, 17, 16, 18};
        return t[v];
    }
    __host__ __device__ static constexpr double w(int v) {
        constexpr double t[19] = {1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0,
                                  1.0 / 18.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0,
                                  1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0,
                                  1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 3.0};
        return t[v];
    }
};

// LUMA's eType numbering (LUMA/inc/Enumerations.h:84-96); the C ABI takes labels in it.
enum SiteType : int {
    kSolid = 0,
    kFluid = 1,
    kRefined = 2,
    kTL2C = 3,   // eTransitionToCoarser
    kTL2F = 4,   // eTransitionToFiner
    kBFL = 5,
    kVelocity = 6,
    kPressure = 7,
    kSlip = 8,
    kExtrapRight = 9,
};

// Per-cell flag byte used by the masked kernels.
//   bits 0-3 : SiteType
//   bit  4   : cell carries a body force (force_xyz is read)
//   bits 5-7 : inward wall normal of a regularised velocity site: 1..6 = +x,-x,+y,-y,+z,-z
constexpr unsigned kTypeMask = 0x0F;
constexpr unsigned kForcedBit = 0x10;
constexpr int kNormalShift = 5;
// wall-normal code 7: a regularised velocity site on an edge or corner, updated by its own
// kernel after the step kernel (the density comes from two inward neighbours)
constexpr int kNormalEdge = 7;

}  // namespace lumagpu
