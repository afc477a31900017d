// Coordinate and force buffer operations and output reductions.
//
//   convertX    nbnxn_atomdata_copy_x_to_nbat_x (atomdata.cpp:1018-1073):
//               gather x[atomIndices[slot]] into the float4 {x,y,z,q} slots,
//               fillers untouched
//   reduceF     atomdata_add_nbat_f_to_f after nblib's zero fill
//               (gmxcalculator.cpp:76-82, atomdata.cpp:1181-1228):
//               f[a] = fnb[cells[a]]
//   finalize    energy partials (reduce_energies_over_lists,
//               kernel_common.cpp:109-133) and shift-force partials
//               (nbnxn_atomdata_add_nbat_fshift_to_fshift,
//               atomdata.cpp:1471-1487) summed in double in a fixed order
//   virial      -1/2 sum x (x) f - 1/2 sum shift (x) fshift in double
//               (calcvir.cpp:103-148, sim_util.cpp:157-177)
//   exclusion   the Coulomb energy terms the pair kernel leaves out, in
//   energies    double (exclusion.h), per atom for batches; a single
//               system's pass runs in the pair kernel's launch
// All are HBM-bandwidth bound: one coalesced pass over the arrays.
#include <algorithm>

#include "exclusion.h"
#include "kernel_common.h"

namespace nbx
{

namespace
{

// also resets the kernel's work-group counter (a fill kernel rather than a
// memset, so that the step can be captured in a graph)
__global__ void clearF4Kernel(float4* __restrict__ p, int n, double* __restrict__ d, int nd)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
    {
        p[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    if (i < nd)
    {
        d[i] = 0.0;
    }
}

// A step's set-up in one launch: coordinates into the cluster layout (when
// x is given), the cluster force buffer and the energy slots cleared.
__global__ void prepareStepKernel(int nslot, const int* __restrict__ atomIndices, const float* __restrict__ x,
                                  float4* __restrict__ xq, float4* __restrict__ fnb, double* __restrict__ d, int nd,
                                  unsigned long long* __restrict__ fixF)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nslot)
    {
        if (fnb)
        {
            fnb[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        }
        if (fixF)
        {
            fixF[3 * i]     = 0ull;
            fixF[3 * i + 1] = 0ull;
            fixF[3 * i + 2] = 0ull;
        }
        if (x)
        {
            const int a = atomIndices[i];
            if (a >= 0)
            {
                float4 v = xq[i];
                v.x      = x[3 * a + 0];
                v.y      = x[3 * a + 1];
                v.z      = x[3 * a + 2];
                xq[i]    = v;
            }
        }
    }
    if (i < nd)
    {
        d[i] = 0.0;
    }
}

__global__ void convertXKernel(int nslot, const int* __restrict__ atomIndices, const float* __restrict__ x,
                               float4* __restrict__ xq)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nslot)
    {
        return;
    }
    const int a = atomIndices[s];
    if (a >= 0)
    {
        float4 v = xq[s];
        v.x      = x[3 * a + 0];
        v.y      = x[3 * a + 1];
        v.z      = x[3 * a + 2];
        xq[s]    = v;
    }
}

__global__ void setChargesKernel(int nslot, const int* __restrict__ atomIndices, const float* __restrict__ q,
                                 float4* __restrict__ xq)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < nslot)
    {
        const int a = atomIndices[s];
        if (a >= 0)
        {
            xq[s].w = q[a];
        }
    }
}

__global__ void reduceFKernel(int natoms, const int* __restrict__ cells, const float4* __restrict__ fnb,
                              float* __restrict__ f)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= natoms)
    {
        return;
    }
    const float4 v = fnb[cells[a]];
    f[3 * a + 0]   = v.x;
    f[3 * a + 1]   = v.y;
    f[3 * a + 2]   = v.z;
}

constexpr int kFinalizeThreads = 1024;

// Sum of v[k * stride + comp] over k in [0, n) by this thread's share of a
// workgroup, with four independent accumulators (the loads of one pass are
// in flight together); the order is fixed, so the result is reproducible.
template<typename T>
__device__ __forceinline__ double strideSum(const T* __restrict__ v, int n, int stride, int comp)
{
    double    a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    const int t  = kFinalizeThreads;
    int       k  = threadIdx.x;
    for (; k + 3 * t < n; k += 4 * t)
    {
        a0 += static_cast<double>(v[static_cast<size_t>(k) * stride + comp]);
        a1 += static_cast<double>(v[static_cast<size_t>(k + t) * stride + comp]);
        a2 += static_cast<double>(v[static_cast<size_t>(k + 2 * t) * stride + comp]);
        a3 += static_cast<double>(v[static_cast<size_t>(k + 3 * t) * stride + comp]);
    }
    // at most three left, loaded independently
    if (k < n)
    {
        a0 += static_cast<double>(v[static_cast<size_t>(k) * stride + comp]);
    }
    if (k + t < n)
    {
        a1 += static_cast<double>(v[static_cast<size_t>(k + t) * stride + comp]);
    }
    if (k + 2 * t < n)
    {
        a2 += static_cast<double>(v[static_cast<size_t>(k + 2 * t) * stride + comp]);
    }
    return (a0 + a1) + (a2 + a3);
}

__global__ __launch_bounds__(kFinalizeThreads) void finalizeKernel(int numEnergyPartials,
                                                                   const double* __restrict__ ePart,
                                                                   int numExclPartials,
                                                                   const double* __restrict__ exclPart,
                                                                   int numShiftPartials,
                                                                   const float* __restrict__ fshPart, int doShift,
                                                                   double* __restrict__ eOut,
                                                                   double* __restrict__ fshOut, int numVirPartials,
                                                                   const double* __restrict__ virPart,
                                                                   const float4* __restrict__ shiftVec,
                                                                   double* __restrict__ virOut)
{
    // blocks 0, 1: the energies; 2 .. 2 + 135 (with shift forces): one
    // shift-force component each; the last 9 (with the virial): one virial
    // component each, -1/2 sum x (x) f plus -1/2 sum_s s (x) fshift_s
    __shared__ double s[kFinalizeThreads / 64];
    __shared__ float  sShift[kNumShifts];
    const int comp = blockIdx.x;
    double    acc  = 0.0;
    const int vir0 = doShift ? 2 + kNumShifts * 3 : 2;
    if (virOut && comp >= vir0)
    {
        const int k = comp - vir0, d1 = k / 3, d2 = k % 3;
        acc         = strideSum(virPart + static_cast<size_t>(k) * numVirPartials, numVirPartials, 1, 0);
        // the shift-force partials as one array of elements i = 45 b + s
        // (component d2 at 3 i + d2): consecutive threads read consecutive
        // elements, eight independent loads per thread and round
        if (threadIdx.x < kNumShifts)
        {
            const float4 sv       = shiftVec[threadIdx.x];
            sShift[threadIdx.x]   = d1 == 0 ? sv.x : (d1 == 1 ? sv.y : sv.z);
        }
        __syncthreads();
        const int n    = numShiftPartials * kNumShifts;
        double    t[8] = { 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0 };
        for (int i0 = threadIdx.x; i0 < n; i0 += 8 * kFinalizeThreads)
        {
#pragma unroll
            for (int u = 0; u < 8; u++)
            {
                const int i = i0 + u * kFinalizeThreads;
                if (i < n)
                {
                    t[u] += static_cast<double>(sShift[i % kNumShifts]) * fshPart[3 * static_cast<size_t>(i) + d2];
                }
            }
        }
        acc -= 0.5 * (((t[0] + t[1]) + (t[2] + t[3])) + ((t[4] + t[5]) + (t[6] + t[7])));
    }
    else if (comp < 2)
    {
        acc = strideSum(ePart, numEnergyPartials, 2, comp);
        if (comp == 1)
        {
            acc += strideSum(exclPart, numExclPartials, 1, 0);
        }
    }
    else
    {
        acc = strideSum(fshPart, numShiftPartials, kNumShifts * 3, comp - 2);
    }
    // wavefront sums, then the wavefront sums in order
    acc = waveSumD(acc);
    if ((threadIdx.x & 63) == 0)
    {
        s[threadIdx.x / 64] = acc;
    }
    __syncthreads();
    if (threadIdx.x == 0)
    {
        double t = 0.0;
        for (int w = 0; w < kFinalizeThreads / 64; w++)
        {
            t += s[w];
        }
        if (virOut && comp >= vir0)
        {
            virOut[comp - vir0] = t;
        }
        else if (comp < 2)
        {
            eOut[comp] = t;
        }
        else
        {
            fshOut[comp - 2] = t;
        }
    }
    (void)doShift;
}

constexpr int kExclThreads = 256;

// Atom i's self energy and the Coulomb energies of its excluded partners
// j > i inside the cut-off, in double.
__device__ double atomExclusionEnergy(const ExclusionArgs& a, int i)
{
    // the atom's system box (a batch keeps one per system)
    const float* box = a.atomSystem ? a.boxes + 3 * a.atomSystem[i] : a.box;
    const double L[3] = { box[0], box[1], box[2] };
    const float4 pi   = a.xq[a.cells[i]];
    const double qi   = pi.w;
    double       ei   = -a.selfCoef * qi * qi;
    for (int k = a.exclOff[i]; k < a.exclOff[i + 1]; k++)
    {
        const int j = a.exclIdx[k];
        if (j <= i)
        {
            continue;
        }
        const float4 pj = a.xq[a.cells[j]];
        double       d[3] = { static_cast<double>(pi.x) - pj.x, static_cast<double>(pi.y) - pj.y,
                        static_cast<double>(pi.z) - pj.z };
        double       r2   = 0.0;
#pragma unroll
        for (int c = 0; c < 3; c++)
        {
            d[c] -= L[c] * rint(d[c] / L[c]);
            r2 += d[c] * d[c];
        }
        if (r2 < a.rc2)
        {
            const double qq = a.epsfacD * qi * static_cast<double>(pj.w);
            if (a.elec == kElecRF)
            {
                ei += qq * (static_cast<double>(a.ic.k_rf) * r2 - static_cast<double>(a.ic.c_rf));
            }
            else
            {
                ei -= qq * ewaldCorrectionD(sqrt(r2), a);
            }
        }
    }
    return ei;
}

__global__ __launch_bounds__(kExclThreads) void exclusionEnergyKernel(const ExclusionArgs a)
{
    __shared__ double s[kExclThreads / 64];
    double            e = 0.0;
    for (int i = blockIdx.x * kExclThreads + threadIdx.x; i < a.natoms; i += gridDim.x * kExclThreads)
    {
        const double ei = atomExclusionEnergy(a, i);
        if (a.atomEnergy)
        {
            a.atomEnergy[i] = ei;
        }
        e += ei;
    }
    if (a.atomEnergy)
    {
        return;
    }
#pragma unroll
    for (int m = 1; m < 64; m *= 2)
    {
        e += __shfl_xor(e, m, 64);
    }
    if ((threadIdx.x & 63) == 0)
    {
        s[threadIdx.x / 64] = e;
    }
    __syncthreads();
    if (threadIdx.x == 0)
    {
        double t = 0.0;
        for (int w = 0; w < kExclThreads / 64; w++)
        {
            t += s[w];
        }
        a.partial[blockIdx.x] = t;
    }
}

// Per-system energies of a batch: block r sums system r's group energies
// and per-atom exclusion energies in a fixed order.
__global__ __launch_bounds__(kFinalizeThreads) void batchFinalizeKernel(const int* __restrict__ slotOff,
                                                                        const double* __restrict__ ePart,
                                                                        const int* __restrict__ atomOff,
                                                                        const double* __restrict__ atomEnergy,
                                                                        double* __restrict__ eOut)
{
    __shared__ double s[2][kFinalizeThreads];
    const int r = blockIdx.x;
    double    v = 0.0, c = 0.0;
    for (int g = slotOff[r] + threadIdx.x; g < slotOff[r + 1]; g += kFinalizeThreads)
    {
        v += ePart[2 * g];
        c += ePart[2 * g + 1];
    }
    for (int i = atomOff[r] + threadIdx.x; i < atomOff[r + 1]; i += kFinalizeThreads)
    {
        c += atomEnergy[i];
    }
    s[0][threadIdx.x] = v;
    s[1][threadIdx.x] = c;
    __syncthreads();
    for (int off = kFinalizeThreads / 2; off > 0; off >>= 1)
    {
        if (threadIdx.x < off)
        {
            s[0][threadIdx.x] += s[0][threadIdx.x + off];
            s[1][threadIdx.x] += s[1][threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
    {
        eOut[2 * r]     = s[0][0];
        eOut[2 * r + 1] = s[1][0];
    }
}


// Forces of each cluster from the staged windows that hold it: a wavefront
// per cluster, lane l adding copies l, l + 64, ... (12 floats each), then a
// fixed butterfly over the lanes, so the sum order is the same every step;
// added to the cluster force buffer (which holds the contributions of
// clusters that spilled out of a window; with atom-order output it may be
// NULL when nothing spilled).  Wavefronts stride over the
// clusters.  With virPart the -1/2 x (x) f part of the virial is summed in
// double from the cluster-layout coordinates and the final forces, one
// partial per workgroup in a fixed order.
constexpr int kGatherThreads = 256;

template<bool kToAtoms>
__global__ __launch_bounds__(kGatherThreads) void gatherForcesKernel(int ncluster, const int* __restrict__ gposOff,
                                                                     const int* __restrict__ gpos,
                                                                     const float4* __restrict__ staging,
                                                                     float4* __restrict__ fnb,
                                                                     const int* __restrict__ atomIndices,
                                                                     float* __restrict__ fOut, int addToF,
                                                                     const float4* __restrict__ xq,
                                                                     double* __restrict__ virPart,
                                                                     const unsigned long long* __restrict__ fixF)
{
    constexpr int kWaves = kGatherThreads / 64;
    __shared__ double sVir[kWaves][12][3];
    const int         lane = threadIdx.x & 63;
    const int         wave = threadIdx.x >> 6;
    const int         idx  = lane >> 2; // component 3 * atom + d of this lane's sum
    double            vir[3] = { 0.0, 0.0, 0.0 };
    // The dispatcher deals workgroups round robin over the XCDs, each with
    // its own L2: with the grid a multiple of kGatherXcds, the workgroups of
    // one XCD take consecutive clusters, so a cache line that neighbouring
    // clusters' staged slots share is fetched into one L2 only.
    const int lb = gridDim.x % kGatherXcds == 0
                       ? (blockIdx.x % kGatherXcds) * (gridDim.x / kGatherXcds) + blockIdx.x / kGatherXcds
                       : blockIdx.x;
    for (int c = lb * kWaves + wave; c < ncluster; c += gridDim.x * kWaves)
    {
        float f[12];
#pragma unroll
        for (int k = 0; k < 12; k++)
        {
            f[k] = 0.0f;
        }
        for (int i = gposOff[c] + lane; i < gposOff[c + 1]; i += 64)
        {
            const int    p  = gpos[i];
            const float4 v0 = staging[p];
            const float4 v1 = staging[p + 1];
            const float4 v2 = staging[p + 2];
            f[0] += v0.x;
            f[1] += v0.y;
            f[2] += v0.z;
            f[3] += v0.w;
            f[4] += v1.x;
            f[5] += v1.y;
            f[6] += v1.z;
            f[7] += v1.w;
            f[8] += v2.x;
            f[9] += v2.y;
            f[10] += v2.z;
            f[11] += v2.w;
        }
        // lane L ends up with component L >> 2 = 3 * atom + d
        const float r = waveReduceScatter12(f, lane);
        if ((lane & 3) == 0 && idx < 12)
        {
            const int slot = c * kClusterSize + idx / 3;
            float*    dst  = fnb ? reinterpret_cast<float*>(fnb + slot) + idx % 3 : nullptr;
            // the fixed-point sums of the window overflow, if any
            const float over = fixF ? fromFixed(fixF[3 * slot + idx % 3]) : 0.0f;
            if (kToAtoms)
            {
                // straight to the caller's atom-order forces (nbat-f -> f)
                const int a = atomIndices[slot];
                if (a >= 0)
                {
                    const float fv = (fnb ? *dst : 0.0f) + (r + over);
                    float*      fo = fOut + 3 * a + idx % 3;
                    *fo            = addToF ? *fo + fv : fv;
                    if (virPart)
                    {
                        const float4 p = xq[slot];
                        vir[0] += static_cast<double>(p.x) * fv;
                        vir[1] += static_cast<double>(p.y) * fv;
                        vir[2] += static_cast<double>(p.z) * fv;
                    }
                }
            }
            else
            {
                *dst += r + over;
            }
        }
    }
    if (kToAtoms && virPart)
    {
        if ((lane & 3) == 0 && idx < 12)
        {
            sVir[wave][idx][0] = vir[0];
            sVir[wave][idx][1] = vir[1];
            sVir[wave][idx][2] = vir[2];
        }
        __syncthreads();
        if (threadIdx.x < 9)
        {
            // component (d1, d2): x_d1 f_d2 over the workgroup's atoms
            const int d1 = threadIdx.x / 3, d2 = threadIdx.x % 3;
            double    t  = 0.0;
            for (int w = 0; w < kWaves; w++)
            {
                for (int at = 0; at < kClusterSize; at++)
                {
                    t += sVir[w][3 * at + d2][d1];
                }
            }
            // component-major, [9][workgroup], for the final sums
            virPart[static_cast<size_t>(threadIdx.x) * gridDim.x + blockIdx.x] = -0.5 * t;
        }
    }
}

// Excluded pairs as cluster slots, after each gridding.
__global__ void exclusionSlotsKernel(int npairs, const int2* __restrict__ pairs, const int* __restrict__ cells,
                                     int2* __restrict__ pairSlots)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < npairs)
    {
        const int2 p = pairs[t];
        pairSlots[t] = make_int2(cells[p.x], cells[p.y]);
    }
}

// Pair count only: per-thread counts, one atomic per workgroup.
__global__ __launch_bounds__(64) void countPairsKernel(const int4* __restrict__ ientry, int nci,
                                                       const uint2* __restrict__ jentry,
                                                       const float4* __restrict__ xq,
                                                       const int* __restrict__ atomIndices,
                                                       const float4* __restrict__ shiftVec, float r2,
                                                       unsigned long long* counter)
{
    const int e = blockIdx.x;
    unsigned  n = 0;
    if (e < nci)
    {
        const int4   ie = ientry[e];
        const float4 sv = shiftVec[ie.y];
        for (int k = ie.z + threadIdx.x; k < ie.w; k += blockDim.x)
        {
            const int cj = static_cast<int>(jentry[k].x);
            for (int i = 0; i < kClusterSize; i++)
            {
                if (atomIndices[ie.x * kClusterSize + i] < 0)
                {
                    continue;
                }
                const float4 xi = xq[ie.x * kClusterSize + i];
                const float  xs = __fadd_rn(xi.x, sv.x), ys = __fadd_rn(xi.y, sv.y), zs = __fadd_rn(xi.z, sv.z);
                for (int j = 0; j < kClusterSize; j++)
                {
                    if (atomIndices[cj * kClusterSize + j] < 0 || (ie.y == kCentral && cj == ie.x && j <= i))
                    {
                        continue;
                    }
                    const float4 xj = xq[cj * kClusterSize + j];
                    const float  dx = __fsub_rn(xs, xj.x), dy = __fsub_rn(ys, xj.y), dz = __fsub_rn(zs, xj.z);
                    const float  d2 = canonicalD2(dx, dy, dz);
                    n += (d2 < r2) ? 1u : 0u;
                }
            }
        }
    }
    for (int m = 1; m < 64; m *= 2)
    {
        n += __shfl_xor(n, m, 64);
    }
    if (threadIdx.x == 0 && n > 0)
    {
        atomicAdd(counter, static_cast<unsigned long long>(n));
    }
}

// One thread per j-entry: the 16 atom pairs with canonical d^2 < r2.
__global__ void expandPairsKernel(const int4* __restrict__ ientry, int nci, const uint2* __restrict__ jentry,
                                  const float4* __restrict__ xq, const int* __restrict__ atomIndices,
                                  const float4* __restrict__ shiftVec, float r2, unsigned long long* counter,
                                  long long* __restrict__ out, long long cap)
{
    const int e = blockIdx.x;
    if (e >= nci)
    {
        return;
    }
    const int4   ie = ientry[e];
    const float4 sv = shiftVec[ie.y];
    for (int k = ie.z + threadIdx.x; k < ie.w; k += blockDim.x)
    {
        const int cj = static_cast<int>(jentry[k].x);
        for (int i = 0; i < kClusterSize; i++)
        {
            const int ai = atomIndices[ie.x * kClusterSize + i];
            if (ai < 0)
            {
                continue;
            }
            const float4 xi = xq[ie.x * kClusterSize + i];
            const float  xs = __fadd_rn(xi.x, sv.x), ys = __fadd_rn(xi.y, sv.y), zs = __fadd_rn(xi.z, sv.z);
            for (int j = 0; j < kClusterSize; j++)
            {
                const int aj = atomIndices[cj * kClusterSize + j];
                if (aj < 0 || (ie.y == kCentral && cj == ie.x && j <= i))
                {
                    continue;
                }
                const float4 xj = xq[cj * kClusterSize + j];
                const float  dx = __fsub_rn(xs, xj.x), dy = __fsub_rn(ys, xj.y), dz = __fsub_rn(zs, xj.z);
                const float  d2 = canonicalD2(dx, dy, dz);
                if (d2 < r2)
                {
                    const unsigned long long n = atomicAdd(counter, 1ull);
                    if (static_cast<long long>(n) < cap)
                    {
                        const long long lo = ai < aj ? ai : aj, hi = ai < aj ? aj : ai;
                        out[n]             = (lo << 32) | hi;
                    }
                }
            }
        }
    }
}


} // namespace

void launchClearF4(hipStream_t s, float4* p, int n, double* d, int nd)
{
    const int m = std::max(n, d ? nd : 0);
    if (m > 0)
    {
        clearF4Kernel<<<(m + 255) / 256, 256, 0, s>>>(p, n, d, d ? nd : 0);
    }
}

void launchPrepareStep(hipStream_t s, int nslot, const int* atomIndices, const float* x, float4* xq, float4* fnb,
                       double* d, int nd, unsigned long long* fixF)
{
    const int m = std::max(nslot, d ? nd : 0);
    if (m > 0)
    {
        prepareStepKernel<<<(m + 255) / 256, 256, 0, s>>>(nslot, atomIndices, x, xq, fnb, d, d ? nd : 0, fixF);
    }
}

void launchConvertX(hipStream_t s, int nslot, const int* atomIndices, const float* x, float4* xq)
{
    if (nslot > 0)
    {
        convertXKernel<<<(nslot + 255) / 256, 256, 0, s>>>(nslot, atomIndices, x, xq);
    }
}

int launchGatherForces(hipStream_t s, int ncluster, const int* gposOff, const int* gpos, const float4* staging,
                       float4* fnb, const int* atomIndices, float* f, bool addToF, const float4* xq, double* virPart,
                       const unsigned long long* fixF)
{
    if (ncluster == 0)
    {
        return 0;
    }
    // a multiple of kGatherXcds (kGatherMaxBlocks is one)
    const int blocks = std::min(kGatherMaxBlocks, ((ncluster + 3) / 4 + kGatherXcds - 1) / kGatherXcds * kGatherXcds);
    if (f)
    {
        gatherForcesKernel<true><<<blocks, kGatherThreads, 0, s>>>(ncluster, gposOff, gpos, staging, fnb, atomIndices,
                                                                    f, addToF ? 1 : 0, xq, virPart, fixF);
    }
    else
    {
        gatherForcesKernel<false><<<blocks, kGatherThreads, 0, s>>>(ncluster, gposOff, gpos, staging, fnb, nullptr,
                                                                     nullptr, 0, nullptr, nullptr, fixF);
    }
    return blocks;
}


void launchSetCharges(hipStream_t s, int nslot, const int* atomIndices, const float* q, float4* xq)
{
    if (nslot > 0)
    {
        setChargesKernel<<<(nslot + 255) / 256, 256, 0, s>>>(nslot, atomIndices, q, xq);
    }
}

void launchReduceF(hipStream_t s, int natoms, const int* cells, const float4* fnb, float* f)
{
    if (natoms > 0)
    {
        reduceFKernel<<<(natoms + 255) / 256, 256, 0, s>>>(natoms, cells, fnb, f);
    }
}

void launchFinalize(hipStream_t s, int numEnergyPartials, const double* energyPartial, int numExclPartials,
                    const double* exclPartial, int numShiftPartials, const float* fshiftPartial, bool shiftForces,
                    double* energyOut, double* fshiftOut, int numVirPartials, const double* virPartial,
                    const float4* shiftVec, double* virialOut)
{
    const int blocks = (shiftForces ? 2 + kNumShifts * 3 : 2) + (virialOut ? 9 : 0);
    finalizeKernel<<<blocks, kFinalizeThreads, 0, s>>>(numEnergyPartials, energyPartial, numExclPartials, exclPartial,
                                                       numShiftPartials, fshiftPartial, shiftForces ? 1 : 0, energyOut,
                                                       fshiftOut, numVirPartials, virPartial, shiftVec, virialOut);
}

void launchExclusionSlots(hipStream_t s, int npairs, const int2* pairs, const int* cells, int2* pairSlots)
{
    if (npairs > 0)
    {
        exclusionSlotsKernel<<<(npairs + 255) / 256, 256, 0, s>>>(npairs, pairs, cells, pairSlots);
    }
}

void launchBatchFinalize(hipStream_t s, int numSystems, const int* slotOff, const double* energyPartial,
                         const int* atomOff, const double* atomEnergy, double* energyOut)
{
    if (numSystems > 0)
    {
        batchFinalizeKernel<<<numSystems, kFinalizeThreads, 0, s>>>(slotOff, energyPartial, atomOff, atomEnergy,
                                                                    energyOut);
    }
}

int launchExclusionEnergy(hipStream_t s, const ExclusionArgs& a)
{
    const int blocks = std::max(1, std::min(kExclMaxBlocks, (a.natoms + kExclThreads - 1) / kExclThreads));
    exclusionEnergyKernel<<<blocks, kExclThreads, 0, s>>>(a);
    return blocks;
}


void launchExpandPairs(hipStream_t s, const int4* ientry, int nci, const uint2* jentry, const float4* xq,
                       const int* atomIndices, const float4* shiftVec, float r2, unsigned long long* counter,
                       long long* out, long long cap)
{
    if (nci > 0 && out == nullptr)
    {
        countPairsKernel<<<nci, 64, 0, s>>>(ientry, nci, jentry, xq, atomIndices, shiftVec, r2, counter);
    }
    else if (nci > 0)
    {
        expandPairsKernel<<<nci, 64, 0, s>>>(ientry, nci, jentry, xq, atomIndices, shiftVec, r2, counter, out, cap);
    }
}

} // namespace nbx
