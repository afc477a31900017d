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
#include <algorithm>

#include "nbx_internal.h"

namespace nbx
{

namespace
{

constexpr int kPruneWaves = 4;

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
                        keep = keep || canonicalD2(__fsub_rn(xs[i], p.x), __fsub_rn(ys[i], p.y), __fsub_rn(zs[i], p.z)) < a.rlist2;
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

// The chunk -> wavefront and chunk -> group maps of a layout (one thread
// per wavefront or group walks its contiguous chunk range).
__global__ void layoutPruneSetupKernel(const LayoutPruneArgs a)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < a.numWaves)
    {
        for (int k = a.waveStart[t].x; k < a.waveStart[t + 1].x; k++)
        {
            a.chunkWave[k] = t;
        }
    }
    if (t < a.ngroups)
    {
        const int4 g = a.groups[t];
        for (int k = g.x; k < g.y; k++)
        {
            a.chunkGroup[k] = t;
        }
    }
}

// One wavefront per chunk of the selected part: each outer work entry (a
// lane each, 64 at a time) is kept when any of its 16 atom pairs (fillers
// included, as the reference's prune kernel compares them; padding entries
// never) is closer than rlistInner, in the search's canonical rounding.
// Writes a 16-bit keep mask per outer tile and the chunk's kept and kept
// masked counts.
__global__ __launch_bounds__(64 * kPruneWaves) void layoutPruneEvalKernel(const LayoutPruneArgs a)
{
    const int lane = threadIdx.x & 63;
    const int k    = blockIdx.x * kPruneWaves + (threadIdx.x >> 6);
    if (k >= a.nchunk || a.chunkWave[k] % a.numParts != a.part)
    {
        return;
    }
    const int4   cd = a.chunksOuter[k];
    const float4 sv = a.chunkShift[k];
    float        xs[kClusterSize], ys[kClusterSize], zs[kClusterSize];
#pragma unroll
    for (int i = 0; i < kClusterSize; i++)
    {
        const float4 p = a.xq[cd.x * kClusterSize + i];
        xs[i]          = __fadd_rn(p.x, sv.x);
        ys[i]          = __fadd_rn(p.y, sv.y);
        zs[i]          = __fadd_rn(p.z, sv.z);
    }
    const int n    = cd.w * kTileJ;
    int       kept = 0, keptM = 0;
    for (int b = 0; b < n; b += 64)
    {
        bool keep = false, masked = false;
        if (b + lane < n)
        {
            const uint2 ent = a.workOuter[cd.z + b + lane];
            const int   cj  = static_cast<int>(ent.x);
            if (cj != a.ncluster)
            {
#pragma unroll
                for (int j = 0; j < kClusterSize; j++)
                {
                    const float4 p = a.xq[cj * kClusterSize + j];
#pragma unroll
                    for (int i = 0; i < kClusterSize; i++)
                    {
                        keep = keep
                               || canonicalD2(__fsub_rn(xs[i], p.x), __fsub_rn(ys[i], p.y), __fsub_rn(zs[i], p.z))
                                          < a.rlist2;
                    }
                }
                masked = (ent.y & 0xffffu) != 0xffffu;
            }
        }
        const unsigned long long ballot = __ballot(keep);
        kept += __popcll(ballot);
        keptM += __popcll(__ballot(keep && masked));
        // lanes 0, 16, 32, 48 write their tile's 16 bits
        if ((lane & 15) == 0 && b + lane < n)
        {
            a.tileKeep[cd.z / kTileJ + (b + lane) / kTileJ] = static_cast<unsigned short>((ballot >> lane) & 0xffffu);
        }
    }
    if (lane == 0)
    {
        a.keptCount[k]  = kept;
        a.keptMasked[k] = keptM;
    }
}

// One wavefront per pair-kernel wavefront of the selected part, its lanes
// over the chunks of that wavefront's run: the pruned chunks laid out back
// to back from the start of its outer region (a wave prefix sum of their
// padded lengths), each with at least one tile, and their descriptors
// (tiles, masked tiles).
__global__ __launch_bounds__(64 * kPruneWaves) void layoutPruneOffsetsKernel(const LayoutPruneArgs a)
{
    const int lane = threadIdx.x & 63;
    const int w    = blockIdx.x * kPruneWaves + (threadIdx.x >> 6);
    if (w >= a.numWaves || w % a.numParts != a.part)
    {
        return;
    }
    const int k0 = a.waveStart[w].x, k1 = a.waveStart[w + 1].x;
    if (k0 >= k1)
    {
        return;
    }
    const int start0 = a.chunksOuter[k0].z;
    int       base   = 0;
    for (int kb = k0; kb < k1; kb += 64)
    {
        const int k     = kb + lane;
        int4      cd    = make_int4(0, 0, 0, 0);
        int       tiles = 0, special = 0;
        if (k < k1)
        {
            cd      = a.chunksOuter[k];
            tiles   = max(1, (a.keptCount[k] + kTileJ - 1) / kTileJ);
            special = (a.keptMasked[k] + kTileJ - 1) / kTileJ;
        }
        const int len  = tiles * kTileJ;
        int       incl = len;
#pragma unroll
        for (int off = 1; off < 64; off *= 2)
        {
            const int t = __shfl_up(incl, off, 64);
            if (lane >= off)
            {
                incl += t;
            }
        }
        if (k < k1)
        {
            a.chunks[k] = make_int4(cd.x, (cd.y & 0x1ff) | (special << 9), start0 + base + incl - len, tiles);
        }
        base += __shfl(incl, 63, 64);
    }
}

// One wavefront per chunk of the selected part: the kept outer entries, in
// their order (masked ones stay first), at the chunk's new start, then
// padding entries (the dummy cluster in the group's spare slot) to whole
// tiles.
__global__ __launch_bounds__(64 * kPruneWaves) void layoutPruneCompactKernel(const LayoutPruneArgs a)
{
    const int lane = threadIdx.x & 63;
    const int k    = blockIdx.x * kPruneWaves + (threadIdx.x >> 6);
    if (k >= a.nchunk || a.chunkWave[k] % a.numParts != a.part)
    {
        return;
    }
    const int4 cd   = a.chunksOuter[k];
    const int4 nd   = a.chunks[k];
    const int  n    = cd.w * kTileJ;
    int        done = 0;
    for (int b = 0; b < n; b += 64)
    {
        bool  keep = false;
        uint2 ent  = make_uint2(0u, 0u);
        if (b + lane < n)
        {
            const unsigned bits = a.tileKeep[cd.z / kTileJ + (b + lane) / kTileJ];
            keep                = ((bits >> (lane & 15)) & 1u) != 0;
            if (keep)
            {
                ent = a.workOuter[cd.z + b + lane];
            }
        }
        const unsigned long long ballot = __ballot(keep);
        if (keep)
        {
            a.work[nd.z + done + __popcll(ballot & ((1ull << lane) - 1ull))] = ent;
        }
        done += __popcll(ballot);
    }
    const int  total = nd.w * kTileJ;
    const int4 g     = a.groups[a.chunkGroup[k]];
    const uint2 pad  = make_uint2(static_cast<unsigned>(a.ncluster),
                                 (static_cast<unsigned>(a.groupSize + g.w) << 16) | 0xffffu);
    for (int t = done + lane; t < total; t += 64)
    {
        a.work[nd.z + t] = pad;
    }
}

} // namespace

void launchLayoutPruneSetup(hipStream_t s, const LayoutPruneArgs& a)
{
    const int n = std::max(a.numWaves, a.ngroups);
    if (n > 0)
    {
        layoutPruneSetupKernel<<<(n + 255) / 256, 256, 0, s>>>(a);
    }
}

void launchLayoutPrune(hipStream_t s, const LayoutPruneArgs& a)
{
    if (a.nchunk == 0)
    {
        return;
    }
    const int blocks = (a.nchunk + kPruneWaves - 1) / kPruneWaves;
    layoutPruneEvalKernel<<<blocks, 64 * kPruneWaves, 0, s>>>(a);
    layoutPruneOffsetsKernel<<<(a.numWaves + kPruneWaves - 1) / kPruneWaves, 64 * kPruneWaves, 0, s>>>(a);
    layoutPruneCompactKernel<<<blocks, 64 * kPruneWaves, 0, s>>>(a);
}

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
