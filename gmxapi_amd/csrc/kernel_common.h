// Device helpers shared by the pair kernel (kernel.hip) and its adjoint
// (adjoint.hip): wave-uniform loads, the i- and j-atom records, the
// analytical Ewald correction's fitted coefficients and the wave reductions.
#pragma once

#include "nbx_internal.h"

namespace nbx
{
namespace
{

constexpr int kWave = 64;

// 1/12 and 1/6 split into a float and its float remainder
constexpr float kTwelfthHi = static_cast<float>(1.0 / 12.0);
constexpr float kTwelfthLo = static_cast<float>(1.0 / 12.0 - static_cast<double>(static_cast<float>(1.0 / 12.0)));
constexpr float kSixthHi   = static_cast<float>(1.0 / 6.0);
constexpr float kSixthLo   = static_cast<float>(1.0 / 6.0 - static_cast<double>(static_cast<float>(1.0 / 6.0)));

// Analytical Ewald correction, fitted by fit_ewald_correction.py (relative
// error below 1e-7 in exact arithmetic, about 4e-7 in float evaluation):
//   force   F_corr(r) / r = beta^3 P_g(s) / Q_g(s), s = (beta r)^2,
//           closed form beta^3 (erf z - 2 z exp(-z^2) / sqrt(pi)) / z^3
//   energy  erfc(beta r) / r = exp(-s) P_e(z) / Q_e(z) / r, z = beta r,
//           closed form erfc(z) exp(z^2) for P_e / Q_e (the energy of a
//           non-excluded pair, 1/r - erf(beta r)/r, taken without the
//           cancellation of the two terms)
constexpr float kGp0 = 0.7522528767585754f, kGp1 = -0.01339578628540039f, kGp2 = 0.016685180366039276f,
                kGp3 = 0.0003194923046976328f, kGp4 = 3.5602472053142264e-05f, kGp5 = -3.4331239362472843e-07f;
constexpr float kGq1 = 0.5821957588195801f, kGq2 = 0.15719348192214966f, kGq3 = 0.025581449270248413f,
                kGq4 = 0.002643152140080929f, kGq5 = 0.0001805289793992415f;
constexpr float kEp0 = 1.0f, kEp1 = 0.8535063862800598f, kEp2 = 0.35605478286743164f, kEp3 = 0.059257231652736664f,
                kEp4 = 1.4358666703628842e-05f;
constexpr float kEq1 = 1.9818848371505737f, kEq2 = 1.5923840999603271f, kEq3 = 0.6263630390167236f,
                kEq4 = 0.10554185509681702f;
constexpr float kAnaMaxS = 3.6f * 3.6f;
constexpr float kLog2e   = 1.4426950408889634f;

// 32-bit byte offset from a wave-uniform base: lets the compiler use the
// scalar-base + vector-offset form of global loads
template<typename T>
__device__ __forceinline__ T ldg(const T* base, unsigned index)
{
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + index * static_cast<unsigned>(sizeof(T)));
}

// readfirstlane moves 32-bit integers: floats go through their bit pattern
__device__ __forceinline__ float uniformFloat(float v)
{
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}

__device__ __forceinline__ int uniformInt(int v)
{
    return __builtin_amdgcn_readfirstlane(v);
}

// The four i-atoms of the current i-entry, wave-uniform (scalar registers):
// coordinates and charge as stored, and LJ table row offsets.
struct IAtoms
{
    float x[4], y[4], z[4], q[4];
    int   tOff[4];
};

// This lane's j-atom for one tile: coordinates minus the shift, w = epsfac
// q_j, and the LJ type.
struct JAtom
{
    float x, y, z, w;
    int   type;
};

// One butterfly step: lanes with the mask bit set keep b and send a, the
// others keep a and send b; each adds what its partner sent.
__device__ __forceinline__ float butterfly(float a, float b, bool hi, int m)
{
    const float keep = hi ? b : a;
    const float send = hi ? a : b;
    return keep + __shfl_xor(send, m, kWave);
}

// Reduce-scatter of the 12 i-force components over the wavefront: afterwards
// lane L holds the wave sum of component L >> 2 (lanes 48..63 hold zeros).
// Written on named scalars: with an indexed array the compiler folds the
// lane-dependent selects into a dynamically indexed register array and
// expands that into long compare/select chains.
__device__ __forceinline__ float waveReduceScatter12(const float* f, int lane)
{
    const bool  h32 = (lane & 32) != 0;
    const bool  h16 = (lane & 16) != 0;
    const bool  h8  = (lane & 8) != 0;
    const bool  h4  = (lane & 4) != 0;
    // component k + 8 for k >= 4 is zero
    const float a0  = butterfly(f[0], f[8], h32, 32);
    const float a1  = butterfly(f[1], f[9], h32, 32);
    const float a2  = butterfly(f[2], f[10], h32, 32);
    const float a3  = butterfly(f[3], f[11], h32, 32);
    const float a4  = butterfly(f[4], 0.0f, h32, 32);
    const float a5  = butterfly(f[5], 0.0f, h32, 32);
    const float a6  = butterfly(f[6], 0.0f, h32, 32);
    const float a7  = butterfly(f[7], 0.0f, h32, 32);
    const float b0  = butterfly(a0, a4, h16, 16);
    const float b1  = butterfly(a1, a5, h16, 16);
    const float b2  = butterfly(a2, a6, h16, 16);
    const float b3  = butterfly(a3, a7, h16, 16);
    const float c0  = butterfly(b0, b2, h8, 8);
    const float c1  = butterfly(b1, b3, h8, 8);
    float       r   = butterfly(c0, c1, h4, 4);
    r += __shfl_xor(r, 2, kWave);
    r += __shfl_xor(r, 1, kWave);
    return r;
}

__device__ __forceinline__ double waveSumD(double v)
{
#pragma unroll
    for (int m = 1; m < kWave; m *= 2)
    {
        v += __shfl_xor(v, m, kWave);
    }
    return v;
}

} // namespace
} // namespace nbx
