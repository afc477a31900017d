// nbx_list_pairs_noalloc: the atom pairs of the cluster list within a radius,
// written to a caller's device array without any device allocation, for
// callers that cannot allocate (code under HIP graph capture, serving loops
// that forbid hipMalloc).
//
// nbx_list_pairs (api.hip) allocates a counter and a staging array on every
// call and appends pairs with a global atomic counter (unordered).  This
// variant keeps its per-i-entry counts and offsets in `scratch`, a static
// device array of fixed size, and writes the pairs in list order (i-entry,
// j-entry, i-atom, j-atom): a count pass (one wavefront per i-entry), an
// exclusive scan of the counts, and a fill pass in which each wavefront
// places its pairs at its entry's offset with a wave prefix sum.  Every count
// is written by the count pass and every offset and the total by the scan,
// so the array needs no clearing between calls.
//
// The array holds 2 * kNoAllocMaxEntries + 1 ints (2 MiB) and stays resident
// on every device the entry point runs on until the process ends.  All calls
// on one device share it: calls that overlap on one GPU (several host threads
// or streams at once) race.
#include <mutex>

#include "context.h"

namespace nbx
{

namespace
{

constexpr int kNoAllocMaxEntries = 1 << 18;
constexpr int kNoAllocMaxDevices = 64;

// counts[kNoAllocMaxEntries], offsets[kNoAllocMaxEntries], total
static __device__ __align__(256) int scratch[2 * kNoAllocMaxEntries + 1];

// The 16 pair bits (4 * i + j) of j-entry k of i-entry ie: atom pairs of
// real atoms with canonical squared distance below r2, the central
// diagonal's j <= i left out.
__device__ __forceinline__ unsigned entryPairBits(const int4 ie, const uint2* __restrict__ jentry, int k,
                                                  const float4* __restrict__ xq, const int* __restrict__ atomIndices,
                                                  const float4 sv, float r2)
{
    const int cj   = static_cast<int>(jentry[k].x);
    unsigned  bits = 0;
#pragma unroll
    for (int i = 0; i < kClusterSize; i++)
    {
        if (atomIndices[ie.x * kClusterSize + i] < 0)
        {
            continue;
        }
        const float4 xi = xq[ie.x * kClusterSize + i];
        const float  xs = xi.x + sv.x, ys = xi.y + sv.y, zs = xi.z + sv.z;
#pragma unroll
        for (int j = 0; j < kClusterSize; j++)
        {
            if (atomIndices[cj * kClusterSize + j] < 0 || (ie.y == kCentral && cj == ie.x && j <= i))
            {
                continue;
            }
            const float4 xj = xq[cj * kClusterSize + j];
            if (canonicalD2(xs - xj.x, ys - xj.y, zs - xj.z) < r2)
            {
                bits |= 1u << (4 * i + j);
            }
        }
    }
    return bits;
}

__global__ __launch_bounds__(64) void entryPairCountKernel(const int4* __restrict__ ientry, const uint2* __restrict__ jentry,
                                                           const float4* __restrict__ xq,
                                                           const int* __restrict__ atomIndices,
                                                           const float4* __restrict__ shiftVec, float r2,
                                                           int* __restrict__ counts)
{
    const int4   ie = ientry[blockIdx.x];
    const float4 sv = shiftVec[ie.y];
    int          n  = 0;
    for (int k = ie.z + static_cast<int>(threadIdx.x); k < ie.w; k += 64)
    {
        n += __popc(entryPairBits(ie, jentry, k, xq, atomIndices, sv, r2));
    }
#pragma unroll
    for (int m = 1; m < 64; m *= 2)
    {
        n += __shfl_xor(n, m, 64);
    }
    if (threadIdx.x == 0)
    {
        counts[blockIdx.x] = n;
    }
}

__global__ __launch_bounds__(64) void entryPairFillKernel(const int4* __restrict__ ientry, const uint2* __restrict__ jentry,
                                                          const float4* __restrict__ xq,
                                                          const int* __restrict__ atomIndices,
                                                          const float4* __restrict__ shiftVec, float r2,
                                                          const int* __restrict__ offsets, long long* __restrict__ out,
                                                          long long cap)
{
    const int    lane = static_cast<int>(threadIdx.x);
    const int4   ie   = ientry[blockIdx.x];
    const float4 sv   = shiftVec[ie.y];
    long long    base = offsets[blockIdx.x];
    for (int k0 = ie.z; k0 < ie.w; k0 += 64)
    {
        const int      k    = k0 + lane;
        const unsigned bits = k < ie.w ? entryPairBits(ie, jentry, k, xq, atomIndices, sv, r2) : 0u;
        const int      n    = __popc(bits);
        // exclusive prefix of n over the lanes: this lane's first output
        int incl = n;
#pragma unroll
        for (int off = 1; off < 64; off *= 2)
        {
            const int t = __shfl_up(incl, off, 64);
            if (lane >= off)
            {
                incl += t;
            }
        }
        const int total = __shfl(incl, 63, 64);
        long long pos   = base + (incl - n);
        if (bits)
        {
            const int cj = static_cast<int>(jentry[k].x);
            for (unsigned b = bits; b; b &= b - 1, pos++)
            {
                const int bit = __ffs(b) - 1;
                const int ai  = atomIndices[ie.x * kClusterSize + (bit >> 2)];
                const int aj  = atomIndices[cj * kClusterSize + (bit & 3)];
                if (pos < cap)
                {
                    const long long lo = ai < aj ? ai : aj, hi = ai < aj ? aj : ai;
                    out[pos]           = (lo << 32) | hi;
                }
            }
        }
        base += total;
    }
}

// The device address of `scratch` on each device, looked up once per device
// with that device current (each device holds its own copy of the array).
int* scratchOnDevice(int device)
{
    static std::mutex lock;
    static int*       ptr[kNoAllocMaxDevices] = {};
    if (device < 0 || device >= kNoAllocMaxDevices)
    {
        throw InputError("device index out of range for the allocation-free entry points");
    }
    std::lock_guard<std::mutex> guard(lock);
    if (!ptr[device])
    {
        void* p = nullptr;
        check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        ptr[device] = static_cast<int*>(p);
    }
    return ptr[device];
}

} // namespace

} // namespace nbx

using namespace nbx;

int64_t nbx_list_pairs_noalloc(nbx_context* c, double r, int64_t* out, int64_t cap)
{
    if (!c)
    {
        return -fail(NBX_ERROR_INPUT, "NULL context");
    }
    int64_t result = 0;
    const int rc = guarded([&] {
        if (!c->listValid)
        {
            throw StateError("no pair list: call nbx_construct_pairlist after nbx_put_on_grid");
        }
        if (cap < 0 || (cap > 0 && !out))
        {
            throw InputError("the output needs cap >= 0 and a device array when cap > 0");
        }
        if (c->nci > kNoAllocMaxEntries)
        {
            throw InputError("nbx_list_pairs_noalloc takes lists of at most " + std::to_string(kNoAllocMaxEntries)
                             + " i-entries (this list has " + std::to_string(c->nci)
                             + "); use nbx_list_pairs");
        }
        if (!(r > 0))
        {
            throw InputError("the radius must be positive");
        }
        if (static_cast<int64_t>(c->ncj) * kClusterSize * kClusterSize > INT32_MAX)
        {
            throw InputError("the list may hold more than 2^31 atom pairs; use nbx_list_pairs");
        }
        check(hipSetDevice(c->device), "hipSetDevice");
        if (cap > 0)
        {
            hipPointerAttribute_t attr{};
            if (hipPointerGetAttributes(&attr, out) != hipSuccess
                || (attr.type != hipMemoryTypeDevice && attr.type != hipMemoryTypeManaged))
            {
                (void)hipGetLastError();
                throw InputError("the output of nbx_list_pairs_noalloc must be device memory");
            }
        }
        int* const s       = scratchOnDevice(c->device);
        int* const counts  = s;
        int* const offsets = s + kNoAllocMaxEntries;
        int* const total   = s + 2 * kNoAllocMaxEntries;
        if (c->nci == 0)
        {
            result = 0;
            return;
        }
        const float rf = static_cast<float>(r);
        const float r2 = rf * rf;
        entryPairCountKernel<<<c->nci, 64, 0, c->stream>>>(c->ientry.p, c->jentry.p, c->xq.p, c->atomIndices.p,
                                                          c->shiftVec.p, r2, counts);
        check(hipGetLastError(), "count pairs");
        launchExclusiveScan(c->stream, c->nci, counts, offsets, total);
        check(hipGetLastError(), "scan pair counts");
        if (cap > 0)
        {
            entryPairFillKernel<<<c->nci, 64, 0, c->stream>>>(c->ientry.p, c->jentry.p, c->xq.p, c->atomIndices.p,
                                                             c->shiftVec.p, r2, offsets,
                                                             reinterpret_cast<long long*>(out), cap);
            check(hipGetLastError(), "write pairs");
        }
        int n = 0;
        check(hipMemcpyAsync(&n, total, sizeof(int), hipMemcpyDeviceToHost, c->stream), "read pair count");
        check(hipStreamSynchronize(c->stream), "sync");
        result = n;
    });
    return rc == NBX_SUCCESS ? result : -rc;
}
