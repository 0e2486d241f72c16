// Element-level helpers shared by the step kernels and their adjoints:
// the reference's constant factors, local face numbering and packed face
// flags.
#pragma once

#include "device_common.hpp"

namespace cbsdev
{
    constexpr double FCON1 = 1.0 / 24.0, FCON2 = 1.0 / 12.0, FCON3 = 1.0 / 6.0, FCON4 = 1.0 / 3.0;
    constexpr double FDIF1 = 1.0 / 6.0, FDIF2 = 1.0 / 3.0;
    constexpr double MASS_FACTOR = 1.0 / 24.0;

    // Local face f (1-based) is opposite local node f; node fn (1-based) of
    // face f.  A constexpr function so that fully unrolled loops index the
    // element arrays with constants (registers, not scratch).
    __device__ constexpr int ippn1(int is, int fn)
    {
        constexpr int t[5][4] = {{0, 0, 0, 0}, {0, 3, 2, 4}, {0, 3, 4, 1}, {0, 1, 4, 2}, {0, 1, 2, 3}};
        return t[is][fn];
    }

    __device__ inline double dmin(double a, double b) { return (b < a) ? b : a; }
    __device__ inline double dmax(double a, double b) { return (a < b) ? b : a; }

    __device__ inline int conn_at(const int4& n, int a)  // a 1-based
    {
        return a == 1 ? n.x : a == 2 ? n.y : a == 3 ? n.z : n.w;
    }

    __device__ inline int fedge_at(uint32_t f, int is)  // is 1-based
    {
        return int((f >> (8 * (is - 1))) & 0xffu);
    }

    // Flags of the velocity boundary programme (one entry per BC node).
    enum BcFlag : int32_t { BCV_PRESCRIBED = 1, BCV_SOLID = 2 };
}
