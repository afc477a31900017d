// Gridding on the device: the reference's simple (4-atom cluster) grid.
//
// The reference (src/gromacs/nbnxm/grid.cpp) bins atoms into x/y columns
// (calcColumnIndices :1173-1268), gives each column ceil(n/4) clusters
// (setCellIndices :1287-1445), sorts every column on z with a pigeonhole sort
// whose output order is ascending (z, atom index) (sort_atoms :292-427), pads
// the last cluster with fillers at -1e6 (atomdata.cpp:134-250) and stores a
// float bounding box per cluster over its real atoms (calc_bounding_box
// :454-484).  Here each step is a kernel: an atomic column histogram, a
// single-workgroup scan, a rank sort per column in LDS (one workgroup per
// column, total order on (z, index) so the result equals the reference's),
// and one thread per cluster for the fill.  The output is bit-identical to
// the reference's grid for the same float inputs.
#include "nbx_internal.h"

namespace nbx
{

namespace
{

__global__ void gridColumnsKernel(GridDims g, const float* __restrict__ x, int* __restrict__ atomCol,
                                  int* __restrict__ atomRank, int* __restrict__ cxy_na)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= g.natoms)
    {
        return;
    }
    // int((x - lowerCorner) * invCellSize), clamped above (grid.cpp:1207-1223);
    // the lower corner is 0 so x - 0 == x exactly.
    int cx = static_cast<int>(__fmul_rn(x[3 * a + 0], g.invCellSize[0]));
    int cy = static_cast<int>(__fmul_rn(x[3 * a + 1], g.invCellSize[1]));
    cx     = min(cx, g.ncx - 1);
    cy     = min(cy, g.ncy - 1);
    if (cx < 0 || cy < 0)
    {
        // below the grid: flagged in the extra counter slot, the host raises
        atomicAdd(&cxy_na[g.ncol], 1);
        atomCol[a]  = -1;
        atomRank[a] = 0;
        return;
    }
    const int col = cx * g.ncy + cy;
    atomCol[a]    = col;
    atomRank[a]   = atomicAdd(&cxy_na[col], 1);
}

// The same with the column counts aggregated per workgroup in LDS: each
// workgroup bins kColumnsAtomsPerBlock consecutive atoms with LDS atomics
// (an atom's rank within the workgroup), then takes one global atomic per
// column it touched for its base.  Ranks only place an atom in its column
// before the per-column sort by (z, index), so their order does not affect
// the grid.  For up to kColumnsLdsMax columns.
constexpr int kColumnsThreads       = 256;
constexpr int kColumnsAtomsPerBlock = 2048;
constexpr int kColumnsLdsMax        = 12288;

__global__ __launch_bounds__(kColumnsThreads) void gridColumnsLdsKernel(GridDims g, const float* __restrict__ x,
                                                                        int* __restrict__ atomCol,
                                                                        int* __restrict__ atomRank,
                                                                        int* __restrict__ cxy_na)
{
    __shared__ int hist[kColumnsLdsMax];
    for (int c = threadIdx.x; c < g.ncol; c += kColumnsThreads)
    {
        hist[c] = 0;
    }
    __syncthreads();
    constexpr int kPer = kColumnsAtomsPerBlock / kColumnsThreads;
    const int     a0   = blockIdx.x * kColumnsAtomsPerBlock + threadIdx.x;
    int           col[kPer], rank[kPer];
#pragma unroll
    for (int k = 0; k < kPer; k++)
    {
        const int a = a0 + k * kColumnsThreads;
        col[k]      = -2; // no atom
        rank[k]     = 0;
        if (a < g.natoms)
        {
            int cx = static_cast<int>(__fmul_rn(x[3 * a + 0], g.invCellSize[0]));
            int cy = static_cast<int>(__fmul_rn(x[3 * a + 1], g.invCellSize[1]));
            cx     = min(cx, g.ncx - 1);
            cy     = min(cy, g.ncy - 1);
            if (cx < 0 || cy < 0)
            {
                // below the grid: flagged in the extra counter slot, the host raises
                atomicAdd(&cxy_na[g.ncol], 1);
                col[k] = -1;
            }
            else
            {
                col[k]  = cx * g.ncy + cy;
                rank[k] = atomicAdd(&hist[col[k]], 1);
            }
        }
    }
    __syncthreads();
    for (int c = threadIdx.x; c < g.ncol; c += kColumnsThreads)
    {
        const int n = hist[c];
        hist[c]     = n > 0 ? atomicAdd(&cxy_na[c], n) : 0;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kPer; k++)
    {
        const int a = a0 + k * kColumnsThreads;
        if (col[k] >= 0)
        {
            atomCol[a]  = col[k];
            atomRank[a] = hist[col[k]] + rank[k];
        }
        else if (col[k] == -1)
        {
            atomCol[a]  = -1;
            atomRank[a] = 0;
        }
    }
}

// Exclusive scan of ceil(na/4) over the columns in one workgroup of 1024.
__global__ __launch_bounds__(1024) void gridScanKernel(int ncol, const int* __restrict__ cxy_na,
                                                       int* __restrict__ cxy_ind, int* __restrict__ total)
{
    __shared__ int s[1024];
    __shared__ int sMax;
    if (threadIdx.x == 0)
    {
        sMax = 0;
    }
    __syncthreads();
    int carry = 0;
    for (int base = 0; base < ncol; base += 1024)
    {
        const int i = base + threadIdx.x;
        const int v = (i < ncol) ? (cxy_na[i] + kClusterSize - 1) / kClusterSize : 0;
        if (i < ncol)
        {
            atomicMax(&sMax, cxy_na[i]);
        }
        s[threadIdx.x] = v;
        __syncthreads();
        for (int off = 1; off < 1024; off *= 2)
        {
            const int t = (threadIdx.x >= off) ? s[threadIdx.x - off] : 0;
            __syncthreads();
            s[threadIdx.x] += t;
            __syncthreads();
        }
        if (i < ncol)
        {
            cxy_ind[i] = carry + s[threadIdx.x] - v;
        }
        const int blockTotal = s[1023];
        __syncthreads();
        carry += blockTotal;
    }
    __syncthreads();
    if (threadIdx.x == 0)
    {
        cxy_ind[ncol] = carry;
        total[0]      = carry;
        total[1]      = sMax;
    }
}

__global__ void gridScatterKernel(int natoms, const int* __restrict__ atomCol, const int* __restrict__ atomRank,
                                  const int* __restrict__ cxy_ind, int* __restrict__ bucket)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= natoms || atomCol[a] < 0)
    {
        return;
    }
    bucket[cxy_ind[atomCol[a]] * kClusterSize + atomRank[a]] = a;
}

// One workgroup per column: rank every atom by (z, index) and place it.
constexpr int kSortThreads = 256;
constexpr int kSortLds     = 2048;

__global__ __launch_bounds__(kSortThreads) void gridSortKernel(GridDims g, const float* __restrict__ x,
                                                               const int* __restrict__ cxy_na,
                                                               const int* __restrict__ cxy_ind,
                                                               const int* __restrict__ bucket,
                                                               int* __restrict__ atomIndices, int* __restrict__ cells)
{
    __shared__ float sz[kSortLds];
    __shared__ int   sa[kSortLds];
    const int col  = blockIdx.x;
    const int na   = cxy_na[col];
    const int off  = cxy_ind[col] * kClusterSize;
    const int nslt = (cxy_ind[col + 1] - cxy_ind[col]) * kClusterSize;
    const bool inLds = na <= kSortLds;
    if (inLds)
    {
        for (int t = threadIdx.x; t < na; t += kSortThreads)
        {
            const int a = bucket[off + t];
            sa[t]       = a;
            sz[t]       = x[3 * a + 2];
        }
        __syncthreads();
    }
    for (int t = threadIdx.x; t < na; t += kSortThreads)
    {
        const int   a  = inLds ? sa[t] : bucket[off + t];
        const float za = inLds ? sz[t] : x[3 * a + 2];
        int         rank = 0;
        for (int u = 0; u < na; u++)
        {
            const int   b  = inLds ? sa[u] : bucket[off + u];
            const float zb = inLds ? sz[u] : x[3 * b + 2];
            rank += (zb < za || (zb == za && b < a)) ? 1 : 0;
        }
        atomIndices[off + rank] = a;
        cells[a]                = off + rank;
    }
    for (int t = na + threadIdx.x; t < nslt; t += kSortThreads)
    {
        atomIndices[off + t] = -1;
    }
}

// One thread per cluster: nbat coordinates + charge, type, bounding box.
__global__ void gridFillKernel(int ncluster, const float* __restrict__ x, const int* __restrict__ atomIndices,
                               const int* __restrict__ type, const float* __restrict__ q, int paddingType,
                               float4* __restrict__ xq, int* __restrict__ nbtype, float4* __restrict__ bb)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncluster)
    {
        // one extra cluster of far-away fillers after the real ones (the
        // pair kernel points lanes past the end of a j-list at it)
        if (c == ncluster)
        {
            for (int k = 0; k < kClusterSize; k++)
            {
                xq[c * kClusterSize + k]     = make_float4(kFarAway, kFarAway, kFarAway, 0.0f);
                nbtype[c * kClusterSize + k] = paddingType;
            }
        }
        return;
    }
    float lo[3] = { 0, 0, 0 }, hi[3] = { 0, 0, 0 };
    for (int k = 0; k < kClusterSize; k++)
    {
        const int s = c * kClusterSize + k;
        const int a = atomIndices[s];
        if (a >= 0)
        {
            const float p[3] = { x[3 * a + 0], x[3 * a + 1], x[3 * a + 2] };
            xq[s]            = make_float4(p[0], p[1], p[2], q[a]);
            nbtype[s]        = type[a];
            for (int d = 0; d < 3; d++)
            {
                lo[d] = (k == 0) ? p[d] : fminf(lo[d], p[d]);
                hi[d] = (k == 0) ? p[d] : fmaxf(hi[d], p[d]);
            }
        }
        else
        {
            xq[s]     = make_float4(kFarAway, kFarAway, kFarAway, 0.0f);
            nbtype[s] = paddingType;
        }
    }
    bb[2 * c]     = make_float4(lo[0], lo[1], lo[2], 0.0f);
    bb[2 * c + 1] = make_float4(hi[0], hi[1], hi[2], 0.0f);
}

} // namespace

void launchGridColumns(hipStream_t s, const GridDims& g, const float* x, int* atomCol, int* atomRank, int* cxy_na)
{
    if (g.natoms == 0)
    {
        return;
    }
    if (g.ncol <= kColumnsLdsMax)
    {
        gridColumnsLdsKernel<<<(g.natoms + kColumnsAtomsPerBlock - 1) / kColumnsAtomsPerBlock, kColumnsThreads, 0, s>>>(
                g, x, atomCol, atomRank, cxy_na);
    }
    else
    {
        gridColumnsKernel<<<(g.natoms + 255) / 256, 256, 0, s>>>(g, x, atomCol, atomRank, cxy_na);
    }
}

void launchGridScan(hipStream_t s, int ncol, const int* cxy_na, int* cxy_ind, int* ncluster_out)
{
    gridScanKernel<<<1, 1024, 0, s>>>(ncol, cxy_na, cxy_ind, ncluster_out);
}

void launchGridSortColumns(hipStream_t s, const GridDims& g, const float* x, const int* atomCol, const int* atomRank,
                           const int* cxy_na, const int* cxy_ind, int* bucket, int* atomIndices, int* cells)
{
    if (g.natoms == 0)
    {
        return;
    }
    gridScatterKernel<<<(g.natoms + 255) / 256, 256, 0, s>>>(g.natoms, atomCol, atomRank, cxy_ind, bucket);
    gridSortKernel<<<g.ncol, kSortThreads, 0, s>>>(g, x, cxy_na, cxy_ind, bucket, atomIndices, cells);
}

void launchGridFill(hipStream_t s, int ncluster, const float* x, const int* atomIndices, const int* type, const float* q,
                    int paddingType, float4* xq, int* nbtype, float4* bb)
{
    gridFillKernel<<<(ncluster + 1 + 127) / 128, 128, 0, s>>>(ncluster, x, atomIndices, type, q, paddingType, xq,
                                                              nbtype, bb);
}

} // namespace nbx
