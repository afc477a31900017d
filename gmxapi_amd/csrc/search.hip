// Cluster-pair search on the device.
//
// Same list contract as the reference's simple search
// (src/gromacs/nbnxm/pairlist.cpp:3076-3582, nbnxn_make_pairlist_part):
//   * i-entries are (i-cluster, shift) with shift index <= CENTRAL (22), so a
//     pair across a periodic boundary is listed once, from the atom whose
//     shift is "backward" (c_pbcShiftBackward, pairlist.cpp:94,3339-3342);
//   * the central shift keeps only cj >= ci and the diagonal entry carries
//     the mask 0x08ce (j > i only, get_imask :907-910);
//   * exclusion bits 4*i + j are cleared for topology exclusions found in the
//     entry (setExclusionsForIEntry :1367-1435), skipping j <= i centrally.
// The candidate clusters are enumerated per column and z range like the
// reference (:3398-3481), but every candidate is then tested individually:
// accept when the float bounding-box distance is below rlist^2 and at least
// one atom pair has canonical squared distance ((dx*dx + dy*dy) + dz*dz,
// round-to-nearest, no fused multiply-add) below rlist^2.  The bounding-box
// test is a lower bound of every pair distance under monotone rounding, so
// the set of atom pairs within rlist equals the reference's exactly; the
// list itself is shorter (the reference keeps interior clusters unchecked).
// The diagonal entry is always kept, as in the reference, where it is the
// first j entry of every central i-entry.
//
// Work decomposition: one wavefront per i-cluster; lanes take one candidate
// cluster each, hits are compacted in candidate order with a ballot, so the
// output is deterministic.  The same kernel runs twice: a count pass (per
// i-cluster j and i counts, scanned on the device) and a fill pass.
#include <cstdint>

#include "nbx_internal.h"

namespace nbx
{

namespace
{

constexpr int kWave          = 64;
constexpr int kSearchWaves   = 4;

__device__ __forceinline__ int xyz2is(int x, int y, int z)
{
    return 5 * (3 * (z + 1) + y + 1) + x + 2;
}

// clusterBoundingBoxDistance2 (pairlist.cpp:351-372) with the same rounding order.
__device__ __forceinline__ float bbDist2(float4 ilo, float4 ihi, float4 jlo, float4 jhi)
{
    float d[3];
    const float il[3] = { ilo.x, ilo.y, ilo.z }, ih[3] = { ihi.x, ihi.y, ihi.z };
    const float jl[3] = { jlo.x, jlo.y, jlo.z }, jh[3] = { jhi.x, jhi.y, jhi.z };
#pragma unroll
    for (int k = 0; k < 3; k++)
    {
        const float dl = __fsub_rn(il[k], jh[k]);
        const float dh = __fsub_rn(jl[k], ih[k]);
        d[k]           = fmaxf(fmaxf(dl, dh), 0.0f);
    }
    return canonicalD2(d[0], d[1], d[2]);
}

__device__ __forceinline__ int wavePrefixExclusive(int v, int lane, int* total)
{
    int x = v;
#pragma unroll
    for (int off = 1; off < kWave; off *= 2)
    {
        const int t = __shfl_up(x, off, kWave);
        if (lane >= off)
        {
            x += t;
        }
    }
    *total = __shfl(x, kWave - 1, kWave);
    return x - v;
}

// The first m in [s, e) with pred(m) true (e if none) for a predicate that
// is false then true along the range: up to three probes stepping from the
// guess (an exact guess settles in two), then bisection of what is left.
template<typename P>
__device__ __forceinline__ int firstTrue(int s, int e, int guess, P pred)
{
    int lo = s, hi = e; // the answer lies in [lo, hi]
    int m  = min(max(guess, s), e - 1);
    for (int it = 0; it < 3 && lo < hi && m >= lo && m < hi; it++)
    {
        if (pred(m))
        {
            hi = m;
            m--;
        }
        else
        {
            lo = m + 1;
            m++;
        }
    }
    while (lo < hi)
    {
        const int mid = (lo + hi) >> 1;
        if (pred(mid))
        {
            hi = mid;
        }
        else
        {
            lo = mid + 1;
        }
    }
    return lo;
}

template<bool kFill>
__global__ __launch_bounds__(kWave* kSearchWaves) void searchKernel(SearchArgs a)
{
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    const int ci   = blockIdx.x * kSearchWaves + wave;
    if (ci >= a.ncluster)
    {
        return;
    }
    const GridDims& g = a.g;
    const float4 bbLo = a.bb[2 * ci];
    const float4 bbHi = a.bb[2 * ci + 1];
    float4 xi[kClusterSize];
    int    iIndexReal[kClusterSize];
#pragma unroll
    for (int i = 0; i < kClusterSize; i++)
    {
        xi[i]         = a.xq[ci * kClusterSize + i];
        iIndexReal[i] = a.atomIndices[ci * kClusterSize + i];
    }
    const float rl     = a.rlist;
    const float rlist2 = a.rlist2;

    int jpos = 0, ipos = 0;
    // fill pass: the i-cluster's topology exclusions (other than an atom
    // with itself), one per lane: the excluded atom's slot and the i-atom's
    // index in the cluster; numExcl is wave-uniform.  Every hit's mask is
    // then built from these with shuffles instead of walking the exclusion
    // lists per candidate.  More than a wavefront's worth (not seen for
    // bonded exclusions) falls back to the walk.
    int numExcl = 0, exSlot = -1, exI = 0;
    if (kFill)
    {
        // capJ > 0: a fill without a count pass, into this cluster's own
        // region of capJ j-entries and kMaxIEntriesPerCluster i-entries
        // (searchCompactKernel moves them together after the scans)
        jpos = a.capJ > 0 ? ci * a.capJ : a.offJ[ci];
        ipos = a.capJ > 0 ? ci * kMaxIEntriesPerCluster : a.offI[ci];
        int cnt[kClusterSize], beg[kClusterSize];
#pragma unroll
        for (int i = 0; i < kClusterSize; i++)
        {
            const int iAtom = iIndexReal[i];
            beg[i]          = iAtom >= 0 ? a.exclOff[iAtom] : 0;
            cnt[i]          = iAtom >= 0 ? a.exclOff[iAtom + 1] - beg[i] : 0;
            numExcl += cnt[i];
        }
        if (numExcl <= kWave)
        {
            int base = 0;
#pragma unroll
            for (int i = 0; i < kClusterSize; i++)
            {
                if (lane >= base && lane < base + cnt[i])
                {
                    const int jAtom = a.exclIdx[beg[i] + lane - base];
                    exSlot          = jAtom == iIndexReal[i] ? -1 : a.cells[jAtom];
                    exI             = i;
                }
                base += cnt[i];
            }
        }
    }
    int countJ = 0, countI = 0;

    for (int tz = -1; tz <= 1; tz++)
    {
        for (int ty = -1; ty <= 1; ty++)
        {
            for (int tx = -1; tx <= 1; tx++)
            {
                const int shift = xyz2is(tx, ty, tz);
                if (shift > kCentral)
                {
                    continue;
                }
                const bool central = (shift == kCentral);
                const float shx = static_cast<float>(tx) * g.box[0];
                const float shy = static_cast<float>(ty) * g.box[1];
                const float shz = static_cast<float>(tz) * g.box[2];
                // set_icell_bb_simple: shifted i bounding box
                const float4 lo4 = make_float4(__fadd_rn(bbLo.x, shx), __fadd_rn(bbLo.y, shy), __fadd_rn(bbLo.z, shz), 0);
                const float4 hi4 = make_float4(__fadd_rn(bbHi.x, shx), __fadd_rn(bbHi.y, shy), __fadd_rn(bbHi.z, shz), 0);
                // conservative column and z ranges (margin 1% of a cell)
                if (lo4.x - rl > g.box[0] || hi4.x + rl < 0 || lo4.y - rl > g.box[1] || hi4.y + rl < 0
                    || lo4.z - rl > g.box[2] || hi4.z + rl < 0)
                {
                    continue;
                }
                const int cxf = max(0, static_cast<int>(floorf((lo4.x - rl) * g.invCellSize[0] - 0.01f)));
                const int cxl = min(g.ncx - 1, static_cast<int>(floorf((hi4.x + rl) * g.invCellSize[0] + 0.01f)));
                const int cyf = max(0, static_cast<int>(floorf((lo4.y - rl) * g.invCellSize[1] - 0.01f)));
                const int cyl = min(g.ncy - 1, static_cast<int>(floorf((hi4.y + rl) * g.invCellSize[1] + 0.01f)));
                if (cxf > cxl || cyf > cyl)
                {
                    continue;
                }
                const float zlo = lo4.z - rl - 1e-3f;
                const float zhi = hi4.z + rl + 1e-3f;
                float xs[kClusterSize], ys[kClusterSize], zs[kClusterSize];
#pragma unroll
                for (int i = 0; i < kClusterSize; i++)
                {
                    xs[i] = __fadd_rn(xi[i].x, shx);
                    ys[i] = __fadd_rn(xi[i].y, shy);
                    zs[i] = __fadd_rn(xi[i].z, shz);
                }
                const int nyc      = cyl - cyf + 1;
                const int numCols  = (cxl - cxf + 1) * nyc;
                int       entryJ   = 0; // j entries of this (ci, shift)
                const int entryPos = jpos + countJ;
                for (int c0 = 0; c0 < numCols; c0 += kWave)
                {
                    // per-lane column: z range of candidate clusters
                    const int cc = c0 + lane;
                    int       k0 = 0, n = 0;
                    if (cc < numCols)
                    {
                        const int cx  = cxf + cc / nyc;
                        const int cy  = cyf + cc % nyc;
                        const int col = cx * g.ncy + cy;
                        int       s   = a.cxy_ind[col];
                        int       e   = a.cxy_ind[col + 1];
                        // the column's clusters are sorted by z (both bounds
                        // non-decreasing): the first with upper z >= zlo and
                        // the first with lower z > zhi, each searched from
                        // where an even spread over the box height puts it
                        const float perZ  = static_cast<float>(e - s) * g.invBoxZ;
                        const int   first = firstTrue(s, e, s + static_cast<int>(zlo * perZ),
                                                      [&](int m) { return a.bb[2 * m + 1].z >= zlo; });
                        int last = firstTrue(first, e, s + static_cast<int>(zhi * perZ) + 1,
                                             [&](int m) { return a.bb[2 * m].z > zhi; }); // exclusive
                        k0       = first;
                        if (central)
                        {
                            k0 = max(k0, ci);
                        }
                        n = max(0, last - k0);
                    }
                    int total;
                    const int pref     = wavePrefixExclusive(n, lane, &total);
                    const int ncolHere = min(kWave, numCols - c0);
                    for (int t0 = 0; t0 < total; t0 += kWave)
                    {
                        const int t = t0 + lane;
                        // column slot holding candidate t: the last slot whose
                        // prefix is <= t (uniform trip count, all lanes shuffle)
                        int lo = 0, hiIdx = ncolHere - 1;
#pragma unroll
                        for (int it = 0; it < 7; it++)
                        {
                            const int m  = (lo + hiIdx + 1) >> 1;
                            const int pm = __shfl(pref, m, kWave);
                            if (lo < hiIdx)
                            {
                                if (pm <= t)
                                {
                                    lo = m;
                                }
                                else
                                {
                                    hiIdx = m - 1;
                                }
                            }
                        }
                        const int colStart = __shfl(k0, lo, kWave);
                        const int colPref  = __shfl(pref, lo, kWave);
                        bool      hit      = false;
                        int       cj       = -1;
                        if (t < total)
                        {
                            cj = colStart + (t - colPref);
                            if (central && cj == ci)
                            {
                                hit = true;
                            }
                            else
                            {
                                const float4 jlo = a.bb[2 * cj];
                                const float4 jhi = a.bb[2 * cj + 1];
                                if (bbDist2(lo4, hi4, jlo, jhi) < rlist2)
                                {
#pragma unroll
                                    for (int j = 0; j < kClusterSize; j++)
                                    {
                                        const float4 xj = a.xq[cj * kClusterSize + j];
#pragma unroll
                                        for (int i = 0; i < kClusterSize; i++)
                                        {
                                            const float d2 = canonicalD2(__fsub_rn(xs[i], xj.x), __fsub_rn(ys[i], xj.y),
                                                                     __fsub_rn(zs[i], xj.z));
                                            hit = hit || (iIndexReal[i] >= 0 && d2 < rlist2);
                                        }
                                    }
                                }
                            }
                        }
                        const unsigned long long ballot = __ballot(hit);
                        const int before = __popcll(ballot & ((1ull << lane) - 1ull));
                        // exclusion bits 4 i + j of the exclusions in this
                        // j-cluster (setExclusionsForIEntry,
                        // pairlist.cpp:1367-1435), every lane taking part in
                        // the shuffles; centrally a pair is excluded only in
                        // the entry of its lower slot
                        unsigned exclMask = 0u;
                        if (kFill && ballot != 0 && numExcl <= kWave)
                        {
                            for (int k = 0; k < numExcl; k++)
                            {
                                const int js = __shfl(exSlot, k, kWave);
                                const int ii = __shfl(exI, k, kWave);
                                if (js >= 0 && (js >> 2) == cj && !(central && js <= ci * kClusterSize + ii))
                                {
                                    exclMask |= 1u << (ii * kClusterSize + (js & 3));
                                }
                            }
                        }
                        // a capped region past its end: not written, the
                        // count (above capJ) sends the host to the counted path
                        const bool room = a.capJ == 0 || entryPos + entryJ + before < jpos + a.capJ;
                        if (kFill && hit && room && numExcl <= kWave)
                        {
                            const unsigned mask = ((central && cj == ci) ? kMaskDiag : kMaskAll) & ~exclMask;
                            a.jentry[entryPos + entryJ + before] = make_uint2(static_cast<unsigned>(cj), mask);
                        }
                        else if (kFill && hit && room)
                        {
                            unsigned mask = (central && cj == ci) ? kMaskDiag : kMaskAll;
                            for (int i = 0; i < kClusterSize; i++)
                            {
                                const int iAtom = iIndexReal[i];
                                if (iAtom < 0)
                                {
                                    continue;
                                }
                                const int iIndex = ci * kClusterSize + i;
                                for (int k = a.exclOff[iAtom]; k < a.exclOff[iAtom + 1]; k++)
                                {
                                    const int jAtom = a.exclIdx[k];
                                    if (jAtom == iAtom)
                                    {
                                        continue;
                                    }
                                    const int jIndex = a.cells[jAtom];
                                    if (central && jIndex <= iIndex)
                                    {
                                        continue;
                                    }
                                    if ((jIndex >> 2) == cj)
                                    {
                                        mask &= ~(1u << (i * kClusterSize + (jIndex & 3)));
                                    }
                                }
                            }
                            a.jentry[entryPos + entryJ + before] = make_uint2(static_cast<unsigned>(cj), mask);
                        }
                        entryJ += __popcll(ballot);
                    }
                }
                if (entryJ > 0)
                {
                    if (kFill && lane == 0)
                    {
                        a.ientry[ipos + countI] = make_int4(ci, shift, entryPos, entryPos + entryJ);
                    }
                    countJ += entryJ;
                    countI += 1;
                }
            }
        }
    }
    if ((!kFill || a.capJ > 0) && lane == 0)
    {
        a.countJ[ci] = countJ;
        a.countI[ci] = countI;
    }
}

// A capped fill's lists moved to the scanned offsets (one wavefront per
// cluster), the i-entries' j ranges rebased.
__global__ __launch_bounds__(256) void searchCompactKernel(int ncluster, int capJ, const int* __restrict__ countJ,
                                                           const int* __restrict__ countI,
                                                           const int* __restrict__ offJ, const int* __restrict__ offI,
                                                           const uint2* __restrict__ jentryTmp,
                                                           const int4* __restrict__ ientryTmp,
                                                           uint2* __restrict__ jentry, int4* __restrict__ ientry)
{
    const int lane = threadIdx.x & 63;
    const int ci   = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (ci >= ncluster)
    {
        return;
    }
    const int nJ = countJ[ci], nI = countI[ci];
    const int src = ci * capJ, dst = offJ[ci];
    for (int t = lane; t < nJ; t += 64)
    {
        jentry[dst + t] = jentryTmp[src + t];
    }
    if (lane < nI)
    {
        int4 ie = ientryTmp[ci * kMaxIEntriesPerCluster + lane];
        ie.z    = dst + (ie.z - src);
        ie.w    = dst + (ie.w - src);
        ientry[offI[ci] + lane] = ie;
    }
}

// Exclusive scan of n ints by one workgroup (the lists scanned here hold
// one entry per cluster or group, at most a few 100k): rounds of 32768
// elements, each thread summing 32 consecutive ones (int4 loads), a
// wavefront shuffle scan over the threads and a scan over the 16 wavefront
// totals; out[n] and *total receive the sum.
constexpr int kScanThreads = 1024;
constexpr int kScanVec     = 8;              // int4 loads per thread per round
constexpr int kScanPer     = 4 * kScanVec;   // elements per thread per round
constexpr int kScanRound   = kScanThreads * kScanPer;

// One round's elements of this thread: [first, first + kScanPer) of in,
// zero past n; whole int4 loads when the run lies inside and in is 16-byte
// aligned.
__device__ __forceinline__ void scanLoad(const int* __restrict__ in, int n, int first, bool aligned, int4* v)
{
    if (aligned && first + kScanPer <= n)
    {
        const int4* p = reinterpret_cast<const int4*>(in + first);
#pragma unroll
        for (int k = 0; k < kScanVec; k++)
        {
            v[k] = p[k];
        }
    }
    else
    {
#pragma unroll
        for (int k = 0; k < kScanVec; k++)
        {
            const int e = first + 4 * k;
            v[k]        = make_int4(e < n ? in[e] : 0, e + 1 < n ? in[e + 1] : 0, e + 2 < n ? in[e + 2] : 0,
                                    e + 3 < n ? in[e + 3] : 0);
        }
    }
}

// Up to three independent scans of n elements each, one per workgroup.
struct ScanSet
{
    const int* in[3];
    int*       out[3];
    int*       total[3];
};

__global__ __launch_bounds__(kScanThreads) void exclusiveScanKernel(int n, ScanSet set)
{
    const int* __restrict__ in    = set.in[blockIdx.x];
    int* __restrict__       out   = set.out[blockIdx.x];
    int* __restrict__       total = set.total[blockIdx.x];
    __shared__ int sWave[kScanThreads / 64];
    const int      lane     = threadIdx.x & 63;
    const int      wave     = threadIdx.x >> 6;
    const bool     alignIn  = (reinterpret_cast<uintptr_t>(in) & 15) == 0;
    const bool     alignOut = (reinterpret_cast<uintptr_t>(out) & 15) == 0;
    int            carry    = 0;
    // rounds of kScanRound elements, each thread a contiguous run of
    // kScanPer; the next round's elements are loaded before this round's
    // scan, so the load latency overlaps the shuffles and barriers
    int4 cur[kScanVec];
    scanLoad(in, n, threadIdx.x * kScanPer, alignIn, cur);
    for (int base = 0; base < n; base += kScanRound)
    {
        const int first = base + threadIdx.x * kScanPer;
        int4      nxt[kScanVec];
        scanLoad(in, n, first + kScanRound, alignIn, nxt);
        int sum = 0;
#pragma unroll
        for (int k = 0; k < kScanVec; k++)
        {
            sum += (cur[k].x + cur[k].y) + (cur[k].z + cur[k].w);
        }
        // inclusive scan of the thread sums within the wavefront
        int incl = sum;
#pragma unroll
        for (int off = 1; off < 64; off *= 2)
        {
            const int t = __shfl_up(incl, off, 64);
            incl += (lane >= off) ? t : 0;
        }
        if (lane == 63)
        {
            sWave[wave] = incl;
        }
        __syncthreads();
        int waveBase = 0, roundTotal = 0;
#pragma unroll
        for (int w = 0; w < kScanThreads / 64; w++)
        {
            const int t = sWave[w];
            waveBase += (w < wave) ? t : 0;
            roundTotal += t;
        }
        int run = carry + waveBase + incl - sum;
        int4 res[kScanVec];
#pragma unroll
        for (int k = 0; k < kScanVec; k++)
        {
            res[k].x = run;
            run += cur[k].x;
            res[k].y = run;
            run += cur[k].y;
            res[k].z = run;
            run += cur[k].z;
            res[k].w = run;
            run += cur[k].w;
        }
        if (alignOut && first + kScanPer <= n)
        {
            int4* p = reinterpret_cast<int4*>(out + first);
#pragma unroll
            for (int k = 0; k < kScanVec; k++)
            {
                p[k] = res[k];
            }
        }
        else
        {
#pragma unroll
            for (int k = 0; k < kScanVec; k++)
            {
                const int e = first + 4 * k;
                if (e < n)
                {
                    out[e] = res[k].x;
                }
                if (e + 1 < n)
                {
                    out[e + 1] = res[k].y;
                }
                if (e + 2 < n)
                {
                    out[e + 2] = res[k].z;
                }
                if (e + 3 < n)
                {
                    out[e + 3] = res[k].w;
                }
            }
        }
        carry += roundTotal;
#pragma unroll
        for (int k = 0; k < kScanVec; k++)
        {
            cur[k] = nxt[k];
        }
        __syncthreads(); // sWave is rewritten next round
    }
    if (threadIdx.x == 0)
    {
        out[n] = carry;
        *total = carry;
    }
}

// Tiled scans: a workgroup of kTileThreads per tile of kScanTile elements,
// each thread four int4 loads at (k * kTileThreads + thread) * 4 within the
// tile, so every load instruction of a wavefront reads 1 KiB contiguously.
constexpr int kTileThreads = 256;
constexpr int kTileVec     = kScanTile / (4 * kTileThreads);
static_assert(kTileVec * 4 * kTileThreads == kScanTile, "tile of whole int4 loads");

__device__ __forceinline__ int4 tileLoad(const int* __restrict__ in, int n, int first, bool aligned)
{
    if (aligned && first + 4 <= n)
    {
        return *reinterpret_cast<const int4*>(in + first);
    }
    return make_int4(first < n ? in[first] : 0, first + 1 < n ? in[first + 1] : 0, first + 2 < n ? in[first + 2] : 0,
                     first + 3 < n ? in[first + 3] : 0);
}

__device__ __forceinline__ int waveSum(int v)
{
#pragma unroll
    for (int off = 32; off > 0; off /= 2)
    {
        v += __shfl_xor(v, off, 64);
    }
    return v;
}

// sums[tile] = the sum of the tile's elements (sums = out + n + 1)
__global__ __launch_bounds__(kTileThreads) void scanTileSumKernel(int n, ScanSet set)
{
    const int* __restrict__ in = set.in[blockIdx.y];
    int* __restrict__ sums     = set.out[blockIdx.y] + n + 1;
    __shared__ int  sWave[kTileThreads / 64];
    const bool      aligned = (reinterpret_cast<uintptr_t>(in) & 15) == 0;
    const int       base    = blockIdx.x * kScanTile;
    int             sum     = 0;
#pragma unroll
    for (int k = 0; k < kTileVec; k++)
    {
        const int4 v = tileLoad(in, n, base + (k * kTileThreads + threadIdx.x) * 4, aligned);
        sum += (v.x + v.y) + (v.z + v.w);
    }
    sum = waveSum(sum);
    if ((threadIdx.x & 63) == 0)
    {
        sWave[threadIdx.x >> 6] = sum;
    }
    __syncthreads();
    if (threadIdx.x == 0)
    {
        int t = 0;
#pragma unroll
        for (int w = 0; w < kTileThreads / 64; w++)
        {
            t += sWave[w];
        }
        sums[blockIdx.x] = t;
    }
}

// The tile's exclusive scan after the sum of the tiles before it; the last
// tile writes out[n] and *total.  Reads the tile sums, writes out[0, n].
__global__ __launch_bounds__(kTileThreads) void scanTileKernel(int n, ScanSet set)
{
    const int* __restrict__ in   = set.in[blockIdx.y];
    int* __restrict__       out  = set.out[blockIdx.y];
    const int* __restrict__ sums = out + n + 1;
    __shared__ int  sPrefix[kTileThreads / 64];
    __shared__ int  sWave[kTileVec][kTileThreads / 64];
    const int       lane     = threadIdx.x & 63;
    const int       wave     = threadIdx.x >> 6;
    const bool      alignIn  = (reinterpret_cast<uintptr_t>(in) & 15) == 0;
    const bool      alignOut = (reinterpret_cast<uintptr_t>(out) & 15) == 0;
    const int       base     = blockIdx.x * kScanTile;
    int4            v[kTileVec];
#pragma unroll
    for (int k = 0; k < kTileVec; k++)
    {
        v[k] = tileLoad(in, n, base + (k * kTileThreads + threadIdx.x) * 4, alignIn);
    }
    int prefix = 0;
    for (int t = threadIdx.x; t < static_cast<int>(blockIdx.x); t += kTileThreads)
    {
        prefix += sums[t];
    }
    prefix = waveSum(prefix);
    // per load k: the thread's sum and its inclusive scan over the wavefront
    int sum[kTileVec], incl[kTileVec];
#pragma unroll
    for (int k = 0; k < kTileVec; k++)
    {
        sum[k]  = (v[k].x + v[k].y) + (v[k].z + v[k].w);
        incl[k] = sum[k];
    }
#pragma unroll
    for (int off = 1; off < 64; off *= 2)
    {
#pragma unroll
        for (int k = 0; k < kTileVec; k++)
        {
            const int t = __shfl_up(incl[k], off, 64);
            incl[k] += (lane >= off) ? t : 0;
        }
    }
    if (lane == 0)
    {
        sPrefix[wave] = prefix;
    }
    if (lane == 63)
    {
#pragma unroll
        for (int k = 0; k < kTileVec; k++)
        {
            sWave[k][wave] = incl[k];
        }
    }
    __syncthreads();
    int run = 0;
#pragma unroll
    for (int w = 0; w < kTileThreads / 64; w++)
    {
        run += sPrefix[w];
    }
    // load k covers elements [k * 4 * kTileThreads, (k + 1) * 4 * kTileThreads)
    // of the tile: its base is the running sum of the loads before it
#pragma unroll
    for (int k = 0; k < kTileVec; k++)
    {
        int before = 0, total = 0;
#pragma unroll
        for (int w = 0; w < kTileThreads / 64; w++)
        {
            const int t = sWave[k][w];
            before += (w < wave) ? t : 0;
            total += t;
        }
        int        r = run + before + incl[k] - sum[k];
        int4       res;
        res.x = r;
        r += v[k].x;
        res.y = r;
        r += v[k].y;
        res.z = r;
        r += v[k].z;
        res.w        = r;
        const int e0 = base + (k * kTileThreads + threadIdx.x) * 4;
        if (alignOut && e0 + 4 <= n)
        {
            *reinterpret_cast<int4*>(out + e0) = res;
        }
        else
        {
            if (e0 < n)
            {
                out[e0] = res.x;
            }
            if (e0 + 1 < n)
            {
                out[e0 + 1] = res.y;
            }
            if (e0 + 2 < n)
            {
                out[e0 + 2] = res.z;
            }
            if (e0 + 3 < n)
            {
                out[e0 + 3] = res.w;
            }
        }
        run += total;
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0)
    {
        out[n]                 = run;
        *set.total[blockIdx.y] = run;
    }
}

} // namespace

void launchSearch(hipStream_t s, const SearchArgs& a, bool fill)
{
    if (a.ncluster == 0)
    {
        return;
    }
    const int blocks = (a.ncluster + kSearchWaves - 1) / kSearchWaves;
    if (fill)
    {
        searchKernel<true><<<blocks, kWave * kSearchWaves, 0, s>>>(a);
    }
    else
    {
        searchKernel<false><<<blocks, kWave * kSearchWaves, 0, s>>>(a);
    }
}

void launchSearchCompact(hipStream_t s, int ncluster, int capJ, const int* countJ, const int* countI,
                         const int* offJ, const int* offI, const uint2* jentryTmp, const int4* ientryTmp,
                         uint2* jentry, int4* ientry)
{
    if (ncluster > 0)
    {
        searchCompactKernel<<<(ncluster + 3) / 4, 256, 0, s>>>(ncluster, capJ, countJ, countI, offJ, offI, jentryTmp,
                                                              ientryTmp, jentry, ientry);
    }
}

void launchExclusiveScan(hipStream_t s, int n, const int* in, int* out, int* total)
{
    ScanSet set{ { in }, { out }, { total } };
    exclusiveScanKernel<<<1, kScanThreads, 0, s>>>(n, set);
}

void launchExclusiveScans(hipStream_t s, int n, int count, const int* const* in, int* const* out, int* const* total)
{
    ScanSet set{};
    for (int k = 0; k < count; k++)
    {
        set.in[k]    = in[k];
        set.out[k]   = out[k];
        set.total[k] = total[k];
    }
    exclusiveScanKernel<<<count, kScanThreads, 0, s>>>(n, set);
}

void launchExclusiveScansWide(hipStream_t s, int n, int count, const int* const* in, int* const* out,
                              int* const* total)
{
    if (n < kScanWideMin)
    {
        launchExclusiveScans(s, n, count, in, out, total);
        return;
    }
    ScanSet set{};
    for (int k = 0; k < count; k++)
    {
        set.in[k]    = in[k];
        set.out[k]   = out[k];
        set.total[k] = total[k];
    }
    const dim3 grid((n + kScanTile - 1) / kScanTile, count);
    scanTileSumKernel<<<grid, kTileThreads, 0, s>>>(n, set);
    scanTileKernel<<<grid, kTileThreads, 0, s>>>(n, set);
}

} // namespace nbx
