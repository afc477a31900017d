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
    int   lj[4]; // the pair kernel: whether the atom's type has LJ at all
};

// This lane's j-atom for one tile: coordinates minus the shift, w = epsfac
// q_j, and the LJ type.
struct JAtom
{
    float x, y, z, w;
    int   type;
};

// Butterfly steps of a reduce-scatter: lanes with the step's lane bit set
// keep b, the others keep a, and each adds its partner's copy of the value it
// keeps.  Lane bits 5 and 4 use CDNA4's lane-swap instructions: swapping the
// upper half of a with the lower half of b (bit 5), or the odd 16-lane rows
// of a with the even rows of b (bit 4), leaves every lane with its kept value
// in one register and the partner's in the other, so the step is the swap and
// one add.  Bits 3 and 2 pair lane l with its mirror image in its row of 16
// (l ^ 15) or half-row of 8 (l ^ 7), bits 1 and 0 with l ^ 2 and l ^ 1 (DPP
// row_mirror, row_half_mirror and quad permutations): no LDS traffic.
__device__ __forceinline__ float swapStep32(float a, float b)
{
    const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(a), __float_as_uint(b), false, false);
    return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}

__device__ __forceinline__ float swapStep16(float a, float b)
{
    const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(a), __float_as_uint(b), false, false);
    return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}

constexpr int kDppRowMirror     = 0x140;
constexpr int kDppRowHalfMirror = 0x141;
constexpr int kDppQuadXor2      = 0x4e; // quad_perm [2, 3, 0, 1]
constexpr int kDppQuadXor1      = 0xb1; // quad_perm [1, 0, 3, 2]

template<int kCtrl>
__device__ __forceinline__ float dppMove(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), kCtrl, 0xf, 0xf, false));
}

template<int kCtrl>
__device__ __forceinline__ float dppStep(float a, float b, bool hi)
{
    const float keep = hi ? b : a;
    const float send = hi ? a : b;
    return keep + dppMove<kCtrl>(send);
}

// The last steps of a reduce-scatter over 16 lane groups of four: c0..c3
// hold the partial sums of four components after the steps on lane bits 5
// and 4; afterwards lane L holds the wave sum of component (L >> 2) & 3 of
// its row pair.
__device__ __forceinline__ float quadTail(float c0, float c1, float c2, float c3, int lane)
{
    const float d0 = dppStep<kDppRowMirror>(c0, c2, (lane & 8) != 0);
    const float d1 = dppStep<kDppRowMirror>(c1, c3, (lane & 8) != 0);
    float       r  = dppStep<kDppRowHalfMirror>(d0, d1, (lane & 4) != 0);
    r += dppMove<kDppQuadXor2>(r);
    r += dppMove<kDppQuadXor1>(r);
    return r;
}

// Reduce-scatter of the 12 i-force components over the wavefront: afterwards
// lane L holds the wave sum of component L >> 2 (lanes 48..63 hold zeros).
// The whole wavefront must be active.  Written on named scalars: with an
// indexed array the compiler folds the lane-dependent selects into a
// dynamically indexed register array and expands that into long
// compare/select chains.
__device__ __forceinline__ float waveReduceScatter12(const float* f, int lane)
{
    // component k + 8 for k >= 4 is zero
    const float a0 = swapStep32(f[0], f[8]);
    const float a1 = swapStep32(f[1], f[9]);
    const float a2 = swapStep32(f[2], f[10]);
    const float a3 = swapStep32(f[3], f[11]);
    const float a4 = swapStep32(f[4], 0.0f);
    const float a5 = swapStep32(f[5], 0.0f);
    const float a6 = swapStep32(f[6], 0.0f);
    const float a7 = swapStep32(f[7], 0.0f);
    return quadTail(swapStep16(a0, a4), swapStep16(a1, a5), swapStep16(a2, a6), swapStep16(a3, a7), lane);
}

// The same for 16 components (4 i-atoms x {x, y, z, q}).
__device__ __forceinline__ float waveReduceScatter16(const float* f, int lane)
{
    const float a0 = swapStep32(f[0], f[8]);
    const float a1 = swapStep32(f[1], f[9]);
    const float a2 = swapStep32(f[2], f[10]);
    const float a3 = swapStep32(f[3], f[11]);
    const float a4 = swapStep32(f[4], f[12]);
    const float a5 = swapStep32(f[5], f[13]);
    const float a6 = swapStep32(f[6], f[14]);
    const float a7 = swapStep32(f[7], f[15]);
    return quadTail(swapStep16(a0, a4), swapStep16(a1, a5), swapStep16(a2, a6), swapStep16(a3, a7), lane);
}

// v in fixed point with kFixBits fraction bits, as the two's-complement bits
// of a 64-bit integer (for integer atomics), and back
__device__ __forceinline__ unsigned long long toFixed(float v)
{
    return static_cast<unsigned long long>(static_cast<long long>(rintf(v * static_cast<float>(1 << kFixBits))));
}

__device__ __forceinline__ float fromFixed(unsigned long long u)
{
    return static_cast<float>(static_cast<double>(static_cast<long long>(u)) * (1.0 / static_cast<double>(1 << kFixBits)));
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
