// Cluster-pair LJ + Coulomb (Ewald real space or reaction-field) kernel for
// CDNA4 (gfx950, wave64).
//
// Physics and masking follow the reference's plain-C kernel
// (src/gromacs/nbnxm/kernels_reference/kernel_ref_outer.h:85-409 and
// kernel_ref_inner.h:44-375, the "LJ + Coulomb with exclusion check" loop):
//   skipmask = 0 beyond rcoulomb and for j <= i on the central diagonal;
//   rsq is floored at 3.82e-7 before 1/sqrt; LJ uses nbfp = {6*C6, 12*C12}
//   with potential-shift energies; excluded pairs inside the cut-off keep the
//   Ewald (or RF) correction; the self term -1/2 q_i^2 epsfac V(0) is added
//   once per central i-entry; shift forces are the sum of i-forces per shift.
//
// Mapping to the hardware: a wavefront owns one i-entry (4 i-atoms, one
// shift) and walks its j-list four j-clusters at a time, so one instruction
// covers 4 i x 4 j x 4 j-clusters = 64 pairs; lane = i + 4*j + 16*k.
//   * j coordinates+charge arrive as one float4 load per lane (16 atoms x 16 B
//     per instruction, contiguous per cluster), types as one int load;
//   * the i-force stays in registers for the whole entry and is reduced over
//     the 16 lanes that share i with cross-lane shuffles once per entry;
//   * the j-force of each tile is summed over the 4 i-lanes of a quad with two
//     DPP adds (quad_perm) and scattered with one float atomic per component,
//     lane i = 0,1,2 carrying x,y,z;
//   * LJ parameters and the Ewald F/D/V table sit in LDS;
//   * energies accumulate per lane in float over one entry, then in double;
//     per-workgroup double partials are summed by a separate kernel, which
//     keeps the energies deterministic.
// Pair arithmetic is FP32 on the vector ALUs; no matrix instructions are
// used (the pair interaction is not a dense contraction).
#include "nbx_internal.h"

namespace nbx
{

namespace
{

constexpr int kWave       = 64;
constexpr int kWavesBlock = 4;
constexpr int kThreads    = kWave * kWavesBlock;

// 1/12 and 1/6 split into a float and its float remainder
constexpr float kTwelfthHi = static_cast<float>(1.0 / 12.0);
constexpr float kTwelfthLo = static_cast<float>(1.0 / 12.0 - static_cast<double>(static_cast<float>(1.0 / 12.0)));
constexpr float kSixthHi   = static_cast<float>(1.0 / 6.0);
constexpr float kSixthLo   = static_cast<float>(1.0 / 6.0 - static_cast<double>(static_cast<float>(1.0 / 6.0)));

__device__ __forceinline__ float dppQuadXor1(float v)
{
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0xB1, 0xF, 0xF, true));
}

__device__ __forceinline__ float dppQuadXor2(float v)
{
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x4E, 0xF, 0xF, true));
}

__device__ __forceinline__ float quadSum(float v)
{
    v = v + dppQuadXor1(v);
    return v + dppQuadXor2(v);
}

// 32-bit byte offset from a wave-uniform base: lets the compiler use the
// scalar-base + vector-offset form of global loads
template<typename T>
__device__ __forceinline__ T ldg(const T* base, unsigned index)
{
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + index * static_cast<unsigned>(sizeof(T)));
}

// Per i-entry constants of one lane (i-atom ii of the entry's cluster).
struct IEntry
{
    float xi, yi, zi, qi;
    int   tiOff;
    int   ci;
    bool  central;
};

// One 4 i x 4 j x 4 j-cluster tile.  kChecks: exclusion bits, the central
// diagonal (j <= i skipped) and lanes past the end of the j-list; without it
// the tile is known to be full and free of masks.
template<int kElec, bool kEnergy, bool kChecks>
__device__ __forceinline__ void pairTile(const IEntry& I, const uint2 jent, const float4 xqj, const int tj,
                                         const bool valid, const unsigned sj, const int ii, const int jj,
                                         const InteractionConst& ic, const float2* __restrict__ sNbfp,
                                         const float4* __restrict__ sTab, const int tabN, float4* __restrict__ fnb,
                                         float& fix, float& fiy, float& fiz, float& vlj, float& vc)
{
    const float dx  = I.xi - xqj.x;
    const float dy  = I.yi - xqj.y;
    const float dz  = I.zi - xqj.z;
    float       rsq = dx * dx + dy * dy + dz * dz;
    float interact  = 1.0f;
    float skipmask;
    if (kChecks)
    {
        interact = static_cast<float>((jent.y >> (ii * 4 + jj)) & 1u);
        const bool skip = !valid || (I.central && static_cast<int>(jent.x) == I.ci && jj <= ii) || rsq >= ic.rcoulomb2;
        skipmask        = skip ? 0.0f : 1.0f;
    }
    else
    {
        skipmask = (rsq >= ic.rcoulomb2) ? 0.0f : 1.0f;
    }
    rsq        = fmaxf(rsq, kMinRsq);
    float rinv = __builtin_amdgcn_rsqf(rsq);
    if (kEnergy)
    {
        // one Newton-Raphson step: removes the systematic bias of the
        // hardware estimate, which r^-12 would amplify in the energy sum
        const float hx = 0.5f * rsq;
        rinv           = rinv * fmaf(-hx * rinv, rinv, 1.5f);
    }
    rinv               = rinv * skipmask;
    const float rinvsq = rinv * rinv;

    const float2 c       = sNbfp[I.tiOff + tj];
    const float  rinv6   = kChecks ? interact * rinvsq * rinvsq * rinvsq : rinvsq * rinvsq * rinvsq;
    const float  FrLJ6   = c.x * rinv6;
    const float  FrLJ12  = c.y * rinv6 * rinv6;
    const float  frLJ    = FrLJ12 - FrLJ6;
    if (kEnergy)
    {
        // (FrLJ12 + c12 cpot_rep)/12 - (FrLJ6 + c6 cpot_disp)/6 with the
        // divisions as x*h + x*l (h + l = 1/12 or 1/6 to 2^-48), so the
        // quotient is rounded once and without bias
        const float v12 = fmaf(FrLJ12, kTwelfthHi, FrLJ12 * kTwelfthLo) + c.y * ic.cpot_rep12;
        const float v6  = fmaf(FrLJ6, kSixthHi, FrLJ6 * kSixthLo) + c.x * ic.cpot_disp6;
        vlj += kChecks ? (v12 - v6) * interact * skipmask : (v12 - v6) * skipmask;
    }

    const float qq = skipmask * I.qi * xqj.w;
    float       fcoul;
    if (kElec == kElecRF)
    {
        fcoul = qq * (interact * rinv * rinvsq - 2.0f * ic.k_rf);
        if (kEnergy)
        {
            vc += qq * (interact * rinv + ic.k_rf * rsq - ic.c_rf);
        }
    }
    else
    {
        const float  rs    = rsq * rinv * ic.tab_scale;
        // clamped so that no input (NaN, Inf) can index past the table
        const int    ri    = min(static_cast<int>(rs), tabN - 2);
        const float  frac  = rs - static_cast<float>(ri);
        const float4 tab   = sTab[ri];
        const float  fexcl = tab.x + frac * tab.y;
        fcoul              = interact * rinvsq - fexcl;
        if (kEnergy)
        {
            const float halfsp = 0.5f / ic.tab_scale;
            vc += qq * (interact * (rinv - ic.sh_ewald) - (tab.z - halfsp * frac * (tab.x + fexcl)));
        }
        fcoul *= qq * rinv;
    }

    const float fscal = frLJ * rinvsq + fcoul;
    const float fx    = fscal * dx;
    const float fy    = fscal * dy;
    const float fz    = fscal * dz;
    fix += fx;
    fiy += fy;
    fiz += fz;
    // j-force: sum over the four i-lanes of the quad; lane i = 0,1,2 scatters
    // component x,y,z
    const float gx   = quadSum(fx);
    const float gy   = quadSum(fy);
    const float gz   = quadSum(fz);
    const float comp = (ii == 0) ? gx : ((ii == 1) ? gy : gz);
    if ((!kChecks || valid) && ii < 3)
    {
        unsafeAtomicAdd(reinterpret_cast<float*>(reinterpret_cast<char*>(fnb) + sj * 16u) + ii, -comp);
    }
}

template<int kElec, bool kEnergy, bool kShift>
__global__ __launch_bounds__(kThreads) void nonbondedKernel(const KernelArgs a)
{
    extern __shared__ float4 smem4[];
    const int ntype  = a.ntype;
    const int ntype2 = ntype * ntype;
    float2*   sNbfp  = reinterpret_cast<float2*>(smem4);
    float4*   sTab   = smem4 + ((ntype2 + 1) / 2);
    const int tabN   = (kElec == kElecEwaldTable) ? a.ic.tab_n : 0;
    float*    sFsh   = reinterpret_cast<float*>(sTab + tabN);
    double*   sEner  = reinterpret_cast<double*>(sFsh + (kShift ? kNumShifts * 3 + 1 : 0));

    for (int t = threadIdx.x; t < ntype2; t += kThreads)
    {
        sNbfp[t] = a.nbfp[t];
    }
    if (kElec == kElecEwaldTable)
    {
        for (int t = threadIdx.x; t < tabN; t += kThreads)
        {
            sTab[t] = a.tabFDV0[t];
        }
    }
    if (kShift)
    {
        for (int t = threadIdx.x; t < kNumShifts * 3; t += kThreads)
        {
            sFsh[t] = 0.0f;
        }
    }
    __syncthreads();

    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    const int ii   = lane & 3;
    const int jj   = (lane >> 2) & 3;
    const int kk   = lane >> 4;

    const InteractionConst ic = a.ic;
    const float selfFac       = 0.5f * ic.self_v0;

    double eVdw = 0.0, eCoul = 0.0;

    for (int e = blockIdx.x * kWavesBlock + wave; e < a.nci; e += gridDim.x * kWavesBlock)
    {
        const int4 ie = a.ientry[__builtin_amdgcn_readfirstlane(e)];
        const int  ci = __builtin_amdgcn_readfirstlane(ie.x);
        const int  sh = __builtin_amdgcn_readfirstlane(ie.y);
        const int  js = __builtin_amdgcn_readfirstlane(ie.z);
        const int  je = __builtin_amdgcn_readfirstlane(ie.w);

        IEntry I;
        const float4 sv  = a.shiftVec[sh];
        const float4 xqi = ldg(a.xq, static_cast<unsigned>(ci * kClusterSize + ii));
        I.xi             = xqi.x + sv.x;
        I.yi             = xqi.y + sv.y;
        I.zi             = xqi.z + sv.z;
        I.qi             = ic.epsfac * xqi.w;
        I.tiOff          = ldg(a.type, static_cast<unsigned>(ci * kClusterSize + ii)) * ntype;
        I.ci             = ci;
        I.central        = (sh == kCentral);

        float fix = 0.0f, fiy = 0.0f, fiz = 0.0f;
        float vlj = 0.0f, vc = 0.0f;
        if (kEnergy && I.central && jj == 0 && kk == 0)
        {
            // kernel_ref_outer.h:286-326, Vc -= qi * q_i * V(0)/2
            vc -= I.qi * xqi.w * selfFac;
        }

        // software pipeline: j-entries two tiles ahead, j-atom data one ahead
        const int ntile = (je - js + 3) >> 2;
        auto entryOf    = [&](int t) {
            const int j = js + 4 * t + kk;
            return ldg(a.jentry, static_cast<unsigned>(j < je ? j : js));
        };
        uint2  eNext  = entryOf(0);
        uint2  eNext2 = entryOf(1);
        unsigned sNext = eNext.x * kClusterSize + jj;
        float4 xNext  = ldg(a.xq, sNext);
        int    tNext  = ldg(a.type, sNext);
        for (int t = 0; t < ntile; t++)
        {
            const uint2    jent = eNext;
            const float4   xqj  = xNext;
            const int      tj   = tNext;
            const unsigned sj   = sNext;
            eNext               = eNext2;
            eNext2              = entryOf(t + 2);
            sNext               = eNext.x * kClusterSize + jj;
            xNext               = ldg(a.xq, sNext);
            tNext               = ldg(a.type, sNext);

            const bool valid   = js + 4 * t + kk < je;
            const bool special = jent.y != kMaskAll || (I.central && static_cast<int>(jent.x) == ci) || !valid;
            if (__ballot(special) == 0ull)
            {
                pairTile<kElec, kEnergy, false>(I, jent, xqj, tj, true, sj, ii, jj, ic, sNbfp, sTab, tabN, a.fnb,
                                                fix, fiy, fiz, vlj, vc);
            }
            else
            {
                pairTile<kElec, kEnergy, true>(I, jent, xqj, tj, valid, sj, ii, jj, ic, sNbfp, sTab, tabN, a.fnb,
                                               fix, fiy, fiz, vlj, vc);
            }
        }

        // i-force: reduce over the 16 lanes sharing ii (lane bits 2..5)
#pragma unroll
        for (int m = 4; m < kWave; m *= 2)
        {
            fix += __shfl_xor(fix, m, kWave);
            fiy += __shfl_xor(fiy, m, kWave);
            fiz += __shfl_xor(fiz, m, kWave);
        }
        if (kk == 0 && jj < 3)
        {
            const float comp = (jj == 0) ? fix : ((jj == 1) ? fiy : fiz);
            unsafeAtomicAdd(reinterpret_cast<float*>(a.fnb + ci * kClusterSize + ii) + jj, comp);
        }
        if (kShift)
        {
            const float sx = quadSum(fix);
            const float sy = quadSum(fiy);
            const float sz = quadSum(fiz);
            if (lane < 3)
            {
                const float comp = (lane == 0) ? sx : ((lane == 1) ? sy : sz);
                atomicAdd(sFsh + sh * 3 + lane, comp);
            }
        }
        if (kEnergy)
        {
            eVdw += static_cast<double>(vlj);
            eCoul += static_cast<double>(vc);
        }
    }

    if (kEnergy)
    {
#pragma unroll
        for (int m = 1; m < kWave; m *= 2)
        {
            eVdw += __shfl_xor(eVdw, m, kWave);
            eCoul += __shfl_xor(eCoul, m, kWave);
        }
        if (lane == 0)
        {
            sEner[2 * wave]     = eVdw;
            sEner[2 * wave + 1] = eCoul;
        }
    }
    __syncthreads();
    if (kEnergy && threadIdx.x < 2)
    {
        double s = 0.0;
        for (int w = 0; w < kWavesBlock; w++)
        {
            s += sEner[2 * w + threadIdx.x];
        }
        a.energyPartial[2 * blockIdx.x + threadIdx.x] = s;
    }
    if (kShift)
    {
        for (int t = threadIdx.x; t < kNumShifts * 3; t += kThreads)
        {
            a.fshiftPartial[blockIdx.x * kNumShifts * 3 + t] = sFsh[t];
        }
    }
}

template<int kElec>
void launchElec(hipStream_t s, const KernelArgs& a, bool energy, bool shiftForces, int numBlocks, int lds)
{
    if (energy && shiftForces)
    {
        nonbondedKernel<kElec, true, true><<<numBlocks, kThreads, lds, s>>>(a);
    }
    else if (energy)
    {
        nonbondedKernel<kElec, true, false><<<numBlocks, kThreads, lds, s>>>(a);
    }
    else if (shiftForces)
    {
        nonbondedKernel<kElec, false, true><<<numBlocks, kThreads, lds, s>>>(a);
    }
    else
    {
        nonbondedKernel<kElec, false, false><<<numBlocks, kThreads, lds, s>>>(a);
    }
}

} // namespace

int nonbondedLdsBytes(int ntype, int tabN, int elec, bool /*energy*/)
{
    int bytes = ((ntype * ntype + 1) / 2) * 16;
    if (elec == kElecEwaldTable)
    {
        bytes += tabN * 16;
    }
    bytes += (kNumShifts * 3 + 1) * 4;
    bytes = (bytes + 15) & ~15;
    bytes += kWavesBlock * 2 * 8;
    return bytes;
}

void launchNonbonded(hipStream_t s, const KernelArgs& a, int elec, bool energy, bool shiftForces, int numBlocks)
{
    if (a.nci == 0 || numBlocks == 0)
    {
        return;
    }
    const int lds = nonbondedLdsBytes(a.ntype, a.ic.tab_n, elec, energy);
    if (elec == kElecRF)
    {
        launchElec<kElecRF>(s, a, energy, shiftForces, numBlocks, lds);
    }
    else
    {
        launchElec<kElecEwaldTable>(s, a, energy, shiftForces, numBlocks, lds);
    }
}

} // namespace nbx
