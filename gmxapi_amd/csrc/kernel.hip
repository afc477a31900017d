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
// Mapping to the hardware.
//   * A workgroup of 4 wavefronts owns a group of G consecutive i-clusters
//     (z-neighbours in one grid column) and the union of their j-clusters.
//     Every j-force goes into an LDS window (one float3 per atom per
//     distinct j-cluster); the window is flushed with one global atomic per
//     atom component when the group is done, which cuts global atomics to
//     about a third of a per-tile scatter.
//   * Wavefronts take work chunks of the group (pieces of i-entries) from an
//     LDS counter.  In a chunk the four i-atoms are wave-uniform (scalar
//     loads) and each lane owns one j-atom: lane = j + 4*k covers 16
//     j-clusters, and loops over the four i-atoms, so one instruction does
//     64 pairs and a tile of 16 cluster pairs is 256 pairs.  The lane keeps
//     its j-force in registers for the tile (no cross-lane reduction) and
//     adds it to the window with three LDS atomics; the i-forces stay in
//     registers for the chunk and are reduce-scattered across the wave once.
//   * j-entries are loaded two tiles ahead, j-atom data one tile ahead.
//   * Masked j-entries come first in every i-entry, so a chunk's leading
//     tiles run a loop with exclusion and diagonal checks and the remaining
//     tiles a loop with none; lanes past the end of a list read a dummy
//     cluster of far-away fillers instead of carrying a validity mask.
//   * LJ parameters and the Ewald F/D/V table are in LDS; energies are
//     summed per lane in float over a chunk, then in double; per-workgroup
//     double partials are summed by a separate kernel in a fixed order.
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

// Per chunk, wave-uniform (scalar registers): the four i-atoms as stored
// and their LJ table row offsets.  The periodic shift is applied to the
// j-atom instead (x_i + s - x_j = x_i - (x_j - s)), once per tile, so the
// i-data never has to leave the scalar unit.
struct IAtoms
{
    float x[4], y[4], z[4], q[4];
    int   tOff[4];
};

// One j-atom (this lane) against i-atom kI.  kChecks adds the exclusion
// bit and the central diagonal.
template<int kElec, bool kEnergy, bool kChecks, int kI>
__device__ __forceinline__ void pairI(const IAtoms& I, const float4 xqj, const int tj, const unsigned mbits,
                                      const bool diag, const int jj, const InteractionConst& ic,
                                      const float2* __restrict__ sNbfp, const float4* __restrict__ sTab,
                                      const int tabN, float* fi, float& fjx, float& fjy, float& fjz, float& vlj,
                                      float& vc)
{
    // xqj: the j-atom minus the shift, w = epsfac * q_j
    const float dx       = I.x[kI] - xqj.x;
    const float dy       = I.y[kI] - xqj.y;
    const float dz       = I.z[kI] - xqj.z;
    float       rsq      = dx * dx + dy * dy + dz * dz;
    float       interact = 1.0f;
    bool        skip;
    if (kChecks)
    {
        interact = static_cast<float>((mbits >> (4 * kI)) & 1u);
        skip     = (diag && jj <= kI) || !(rsq < ic.rcoulomb2);
    }
    else
    {
        skip = !(rsq < ic.rcoulomb2);
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
    rinv               = skip ? 0.0f : rinv;
    const float rinvsq = rinv * rinv;

    const float2 c      = sNbfp[I.tOff[kI] + tj];
    const float  rinv6  = kChecks ? interact * rinvsq * rinvsq * rinvsq : rinvsq * rinvsq * rinvsq;
    const float  FrLJ6  = c.x * rinv6;
    const float  FrLJ12 = c.y * rinv6 * rinv6;
    const float  frLJ   = FrLJ12 - FrLJ6;
    const float  smask  = skip ? 0.0f : 1.0f;
    if (kEnergy)
    {
        // (FrLJ12 + c12 cpot_rep)/12 - (FrLJ6 + c6 cpot_disp)/6, each quotient
        // as x*h + x*l (h + l = 1/12 or 1/6 to 2^-48): rounded once, unbiased
        const float v12 = fmaf(FrLJ12, kTwelfthHi, FrLJ12 * kTwelfthLo) + c.y * ic.cpot_rep12;
        const float v6  = fmaf(FrLJ6, kSixthHi, FrLJ6 * kSixthLo) + c.x * ic.cpot_disp6;
        vlj += kChecks ? (v12 - v6) * interact * smask : (v12 - v6) * smask;
    }

    // beyond the cut-off rinv = 0 zeroes the Ewald force; energies and the
    // reaction-field constant term need qq = 0 there as well
    const float qq = (kEnergy || kElec == kElecRF) ? smask * I.q[kI] * xqj.w : I.q[kI] * xqj.w;
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
    fi[3 * kI + 0] += fx;
    fi[3 * kI + 1] += fy;
    fi[3 * kI + 2] += fz;
    fjx -= fx;
    fjy -= fy;
    fjz -= fz;
}

template<int kElec, bool kEnergy, bool kChecks>
__device__ __forceinline__ void tile4(const IAtoms& I, const float4 xqj, const int tj, const unsigned mbits,
                                      const bool diag, const int jj, const InteractionConst& ic,
                                      const float2* __restrict__ sNbfp, const float4* __restrict__ sTab,
                                      const int tabN, float* fi, float& fjx, float& fjy, float& fjz, float& vlj,
                                      float& vc)
{
    // barriers between the i-atoms keep the scheduler from interleaving all
    // four pair evaluations, which costs registers (occupancy) for little overlap
    pairI<kElec, kEnergy, kChecks, 0>(I, xqj, tj, mbits, diag, jj, ic, sNbfp, sTab, tabN, fi, fjx, fjy, fjz, vlj, vc);
    __builtin_amdgcn_sched_barrier(0);
    pairI<kElec, kEnergy, kChecks, 1>(I, xqj, tj, mbits, diag, jj, ic, sNbfp, sTab, tabN, fi, fjx, fjy, fjz, vlj, vc);
    __builtin_amdgcn_sched_barrier(0);
    pairI<kElec, kEnergy, kChecks, 2>(I, xqj, tj, mbits, diag, jj, ic, sNbfp, sTab, tabN, fi, fjx, fjy, fjz, vlj, vc);
    __builtin_amdgcn_sched_barrier(0);
    pairI<kElec, kEnergy, kChecks, 3>(I, xqj, tj, mbits, diag, jj, ic, sNbfp, sTab, tabN, fi, fjx, fjy, fjz, vlj, vc);
}

// One tile: 16 j-entries x 4 i-atoms, this lane's j-atom against the four
// i-atoms.  The j-force goes to the entry's LDS window slot; energies are
// summed in float over the tile's four pairs, then in double.
template<int kElec, bool kEnergy>
__device__ __forceinline__ void tile(const IAtoms& I, const uint2 e, const float4 xj, const int tj, const float4 sv,
                                     const bool checks, const bool central, const int ci, const int jj,
                                     const InteractionConst& ic, const float2* __restrict__ sNbfp,
                                     const float4* __restrict__ sTab, const int tabN, float* __restrict__ sWin,
                                     float* fi, double& eVdw, double& eCoul)
{
    // the j-atom minus the shift, w = epsfac * q_j
    const float4 xqj = make_float4(xj.x - sv.x, xj.y - sv.y, xj.z - sv.z, ic.epsfac * xj.w);
    float        fjx = 0.0f, fjy = 0.0f, fjz = 0.0f, vlj = 0.0f, vc = 0.0f;
    if (checks)
    {
        const bool diag = central && static_cast<int>(e.x) == ci;
        tile4<kElec, kEnergy, true>(I, xqj, tj, (e.y & 0xffffu) >> jj, diag, jj, ic, sNbfp, sTab, tabN, fi, fjx, fjy,
                                    fjz, vlj, vc);
    }
    else
    {
        tile4<kElec, kEnergy, false>(I, xqj, tj, 0u, false, jj, ic, sNbfp, sTab, tabN, fi, fjx, fjy, fjz, vlj, vc);
    }
    float* w = sWin + (e.y >> 16) * 12 + jj * 3;
    atomicAdd(w + 0, fjx);
    atomicAdd(w + 1, fjy);
    atomicAdd(w + 2, fjz);
    if (kEnergy)
    {
        eVdw += static_cast<double>(vlj);
        eCoul += static_cast<double>(vc);
    }
}

// The tiles of one chunk.  Work entries are read two tiles ahead and j-atom
// data one tile ahead, in a loop unrolled by two so that no loaded value is
// copied between registers (a copy would wait for its load).  Chunks are
// padded to whole tiles with a dummy far-away cluster, and reads past the
// last tile are clamped to it, so no lane needs a bounds check.
template<int kElec, bool kEnergy>
__device__ __forceinline__ void chunkTiles(const KernelArgs& a, const IAtoms& I, const float4 sv, const int ci,
                                           const bool central, const unsigned ws, const int nt, const int nspecial,
                                           const int jj, const int kk, const InteractionConst& ic,
                                           const float2* __restrict__ sNbfp, const float4* __restrict__ sTab,
                                           const int tabN, float* __restrict__ sWin, float* fi, double& eVdw,
                                           double& eCoul)
{
    auto entryAt = [&](int t) { return ldg(a.work, ws + static_cast<unsigned>(kTileJ * min(t, nt - 1) + kk)); };
    uint2  eA = entryAt(0);
    uint2  eB = entryAt(1);
    float4 xA = ldg(a.xq, eA.x * kClusterSize + jj);
    int    tA = ldg(a.type, eA.x * kClusterSize + jj);
    for (int t = 0; t < nt; t += 2)
    {
        const float4 xB  = ldg(a.xq, eB.x * kClusterSize + jj);
        const int    tB  = ldg(a.type, eB.x * kClusterSize + jj);
        const uint2  eA2 = entryAt(t + 2);
        tile<kElec, kEnergy>(I, eA, xA, tA, sv, t < nspecial, central, ci, jj, ic, sNbfp, sTab, tabN, sWin, fi, eVdw,
                             eCoul);
        if (t + 1 >= nt)
        {
            break;
        }
        const uint2 eB2 = entryAt(t + 3);
        xA              = ldg(a.xq, eA2.x * kClusterSize + jj);
        tA              = ldg(a.type, eA2.x * kClusterSize + jj);
        tile<kElec, kEnergy>(I, eB, xB, tB, sv, t + 1 < nspecial, central, ci, jj, ic, sNbfp, sTab, tabN, sWin, fi,
                             eVdw, eCoul);
        eA = eA2;
        eB = eB2;
    }
}

// Reduce-scatter of 16 per-lane values over the wavefront: afterwards lane L
// holds the wave sum of value (L >> 2).
__device__ __forceinline__ float waveReduceScatter16(float* v, int lane)
{
#pragma unroll
    for (int step = 0; step < 4; step++)
    {
        const int  m    = 32 >> step;
        const int  half = 8 >> step;
        const bool hi   = (lane & m) != 0;
#pragma unroll
        for (int k = 0; k < half; k++)
        {
            const float keep = hi ? v[k + half] : v[k];
            const float send = hi ? v[k] : v[k + half];
            v[k]             = keep + __shfl_xor(send, m, kWave);
        }
    }
    float r = v[0];
    r += __shfl_xor(r, 2, kWave);
    r += __shfl_xor(r, 1, kWave);
    return r;
}

template<int kElec, bool kEnergy, bool kShift>
__global__ __launch_bounds__(kThreads) void nonbondedKernel(const KernelArgs a)
{
    extern __shared__ float4 smem4[];
    __shared__ int sNext;
    const int ntype  = a.ntype;
    const int ntype2 = ntype * ntype;
    float2*   sNbfp  = reinterpret_cast<float2*>(smem4);
    float4*   sTab   = smem4 + ((ntype2 + 1) / 2);
    const int tabN   = (kElec == kElecEwaldTable) ? a.ic.tab_n : 0;
    float*    sWin   = reinterpret_cast<float*>(sTab + tabN);
    float*    sFsh   = sWin + a.windowSlots * 12;
    double*   sEner  = reinterpret_cast<double*>(reinterpret_cast<float4*>(sFsh) + (kNumShifts * 3 + 4) / 4);

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
    for (int t = threadIdx.x; t < kNumShifts * 3; t += kThreads)
    {
        sFsh[t] = 0.0f;
    }

    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    const int jj   = lane & 3;
    const int kk   = lane >> 2;

    const InteractionConst ic = a.ic;
    const int              G  = a.groupSize;

    double eVdw = 0.0, eCoul = 0.0;

    for (int g = blockIdx.x; g < a.ngroups; g += gridDim.x)
    {
        const int4 grp    = a.groups[g]; // chunk range, union offset, union size
        const int  c0     = g * G;
        const int  nslots = G + grp.w;
        for (int t = threadIdx.x; t < nslots * 12; t += kThreads)
        {
            sWin[t] = 0.0f;
        }
        if (threadIdx.x == 0)
        {
            sNext = grp.x;
        }
        __syncthreads();

        for (;;)
        {
            int c = 0;
            if (lane == 0)
            {
                c = atomicAdd(&sNext, 1);
            }
            c = __builtin_amdgcn_readfirstlane(__shfl(c, 0, kWave));
            if (c >= grp.y)
            {
                break;
            }
            const int4 ch      = a.chunks[c];
            const int  ci      = __builtin_amdgcn_readfirstlane(ch.x);
            const int  flags   = __builtin_amdgcn_readfirstlane(ch.y);
            const int  sh      = flags & 0xff;
            const bool self    = ((flags >> 8) & 1) != 0;
            const int  nt      = __builtin_amdgcn_readfirstlane(ch.w);
            const int  nspec   = flags >> 9;
            const bool central = (sh == kCentral);

            IAtoms       I;
            const float4 sv    = a.shiftVec[sh];
            float        qSelf = 0.0f;
#pragma unroll
            for (int i = 0; i < 4; i++)
            {
                // a wave-uniform load, moved to scalar registers
                const float4 p = a.xq[ci * kClusterSize + i];
                I.x[i]         = uniformFloat(p.x);
                I.y[i]         = uniformFloat(p.y);
                I.z[i]         = uniformFloat(p.z);
                I.q[i]         = uniformFloat(p.w);
                I.tOff[i]      = __builtin_amdgcn_readfirstlane(a.type[ci * kClusterSize + i]) * ntype;
                qSelf          = (lane == i) ? p.w : qSelf;
            }
            float fi[12];
#pragma unroll
            for (int k = 0; k < 12; k++)
            {
                fi[k] = 0.0f;
            }
            if (kEnergy && self && lane < 4)
            {
                // kernel_ref_outer.h:286-326, Vc -= 1/2 epsfac q_i^2 V(0)
                eCoul -= a.selfCoef * static_cast<double>(qSelf) * static_cast<double>(qSelf);
            }
            chunkTiles<kElec, kEnergy>(a, I, sv, ci, central, static_cast<unsigned>(__builtin_amdgcn_readfirstlane(ch.z)),
                                       nt, nspec, jj, kk, ic, sNbfp, sTab, tabN, sWin, fi, eVdw, eCoul);

            // i-forces: lane L ends up with component (L >> 2) = 3*i + d
            float v16[16];
#pragma unroll
            for (int k = 0; k < 12; k++)
            {
                v16[k] = fi[k];
            }
#pragma unroll
            for (int k = 12; k < 16; k++)
            {
                v16[k] = 0.0f;
            }
            const float r   = waveReduceScatter16(v16, lane);
            const int   idx = lane >> 2;
            if ((lane & 3) == 0 && idx < 12)
            {
                atomicAdd(sWin + (ci - c0) * 12 + idx, r);
            }
            if (kShift)
            {
                // sum over the four i-atoms of component d: values d, 3+d, 6+d, 9+d
                const int   d  = lane < 3 ? lane : 0;
                const float s0 = __shfl(r, 4 * d, kWave);
                const float s1 = __shfl(r, 4 * (3 + d), kWave);
                const float s2 = __shfl(r, 4 * (6 + d), kWave);
                const float s3 = __shfl(r, 4 * (9 + d), kWave);
                if (lane < 3)
                {
                    atomicAdd(sFsh + sh * 3 + lane, (s0 + s1) + (s2 + s3));
                }
            }
        }
        __syncthreads();

        // flush the window: one atomic per nonzero atom component
        for (int t = threadIdx.x; t < nslots * 12; t += kThreads)
        {
            const float v = sWin[t];
            if (v != 0.0f)
            {
                const int slot = t / 12;
                const int r    = t - slot * 12;
                const int atom = r / 3;
                const int comp = r - atom * 3;
                const int cl   = slot < G ? c0 + slot : a.unionList[grp.z + slot - G];
                if (cl < a.ncluster)
                {
                    unsafeAtomicAdd(reinterpret_cast<float*>(a.fnb + cl * kClusterSize + atom) + comp, v);
                }
            }
        }
        __syncthreads();
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

int nonbondedLdsBytes(int ntype, int tabN, int elec, int windowSlots)
{
    int bytes = ((ntype * ntype + 1) / 2) * 16;
    if (elec == kElecEwaldTable)
    {
        bytes += tabN * 16;
    }
    bytes += windowSlots * 12 * 4;
    bytes += ((kNumShifts * 3 + 4) / 4) * 16;
    bytes += kWavesBlock * 2 * 8;
    return bytes;
}

void launchNonbonded(hipStream_t s, const KernelArgs& a, int elec, bool energy, bool shiftForces, int numBlocks)
{
    if (a.ngroups == 0 || numBlocks == 0)
    {
        return;
    }
    const int lds = nonbondedLdsBytes(a.ntype, a.ic.tab_n, elec, a.windowSlots);
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
