// Dynamic pruning of the pair list: the reference's rolling prune
// (nonbonded_verlet_t::dispatchPruneKernelCpu, nbnxm.h:312-313, with the
// plain-C kernel nbnxn_kernel_prune_ref, kernels_reference/kernel_ref_prune.cpp:39-123).
//
// The list as searched at rlist (the outer list) is kept; each prune writes
// the inner list: the outer j-entries of which at least one atom pair
// (fillers included, i-atom plus shift minus j-atom) has squared distance
// below rlistInner^2 with the current coordinates, in the outer order and
// with their exclusion masks; i-entries left without j-entries are dropped.
// Distances use the search's canonical rounding
// fl(fl(fl(dx*dx) + fl(dy*dy)) + fl(dz*dz)).
//
// Like the search it runs as a count pass, a device scan over clusters and a
// fill pass; a wavefront per i-cluster walks the cluster's outer i-entries in
// order, a lane per j-entry, with a ballot for the output positions.
#include "nbx_internal.h"

namespace nbx
{

namespace
{

constexpr int kPruneWaves = 4;

__device__ __forceinline__ float pruneD2(float dx, float dy, float dz)
{
    return __fadd_rn(__fadd_rn(__fmul_rn(dx, dx), __fmul_rn(dy, dy)), __fmul_rn(dz, dz));
}

template<bool kFill>
__global__ __launch_bounds__(64 * kPruneWaves) void pruneKernel(const PruneArgs a)
{
    const int lane = threadIdx.x & 63;
    const int c    = blockIdx.x * kPruneWaves + (threadIdx.x >> 6);
    if (c >= a.ncluster)
    {
        return;
    }
    const int iOut = kFill ? a.offI[c] : 0;
    const int jOut = kFill ? a.offJ[c] : 0;
    int       nI = 0, nJ = 0;
    for (int e = a.offIOuter[c]; e < a.offIOuter[c + 1]; e++)
    {
        const int4   ie = a.ientryOuter[e];
        const float4 sv = a.shiftVec[ie.y];
        float        xs[kClusterSize], ys[kClusterSize], zs[kClusterSize];
#pragma unroll
        for (int i = 0; i < kClusterSize; i++)
        {
            const float4 p = a.xq[ie.x * kClusterSize + i];
            xs[i]          = __fadd_rn(p.x, sv.x);
            ys[i]          = __fadd_rn(p.y, sv.y);
            zs[i]          = __fadd_rn(p.z, sv.z);
        }
        const int jFirst = nJ;
        for (int k0 = ie.z; k0 < ie.w; k0 += 64)
        {
            const int k    = k0 + lane;
            bool      keep = false;
            uint2     ent  = make_uint2(0u, 0u);
            if (k < ie.w)
            {
                ent          = a.jentryOuter[k];
                const int cj = static_cast<int>(ent.x);
#pragma unroll
                for (int j = 0; j < kClusterSize; j++)
                {
                    const float4 p = a.xq[cj * kClusterSize + j];
#pragma unroll
                    for (int i = 0; i < kClusterSize; i++)
                    {
                        keep = keep || pruneD2(__fsub_rn(xs[i], p.x), __fsub_rn(ys[i], p.y), __fsub_rn(zs[i], p.z)) < a.rlist2;
                    }
                }
            }
            const unsigned long long ballot = __ballot(keep);
            if (kFill && keep)
            {
                a.jentry[jOut + nJ + __popcll(ballot & ((1ull << lane) - 1ull))] = ent;
            }
            nJ += __popcll(ballot);
        }
        if (nJ > jFirst)
        {
            if (kFill && lane == 0)
            {
                a.ientry[iOut + nI] = make_int4(ie.x, ie.y, jOut + jFirst, jOut + nJ);
            }
            nI++;
        }
    }
    if (!kFill && lane == 0)
    {
        a.countI[c] = nI;
        a.countJ[c] = nJ;
    }
}

} // namespace

void launchPrune(hipStream_t s, const PruneArgs& a, bool fill)
{
    if (a.ncluster == 0)
    {
        return;
    }
    const int blocks = (a.ncluster + kPruneWaves - 1) / kPruneWaves;
    if (fill)
    {
        pruneKernel<true><<<blocks, 64 * kPruneWaves, 0, s>>>(a);
    }
    else
    {
        pruneKernel<false><<<blocks, 64 * kPruneWaves, 0, s>>>(a);
    }
}

} // namespace nbx
