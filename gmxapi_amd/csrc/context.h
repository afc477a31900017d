// Host-side internals shared by the C ABI translation units (api.hip,
// batch.hip): error types, the device buffer holder and the context.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/nbx.h"
#include "nbx_internal.h"

namespace nbx
{

// the message behind nbx_last_error() (per thread)
std::string& lastError();

struct InputError : std::runtime_error
{
    using std::runtime_error::runtime_error;
};

struct StateError : std::runtime_error
{
    using std::runtime_error::runtime_error;
};

struct DeviceError : std::runtime_error
{
    using std::runtime_error::runtime_error;
};

inline void check(hipError_t e, const char* what)
{
    if (e != hipSuccess)
    {
        throw DeviceError(std::string(what) + ": " + hipGetErrorString(e));
    }
}

// The calling thread's pinned host buffer (kPinnedInts ints, allocated on
// first use and kept for the thread's life) for the list build's small
// device-to-host reads: a copy into pageable memory waits for its own staging
// copy, one into pinned memory is queued, so several reads share one
// stream synchronisation.  Callers use disjoint ranges within one wait.
constexpr int kPinnedInts = 1024;
inline int* pinnedInts()
{
    thread_local int* p = nullptr;
    if (!p)
    {
        void* q = nullptr;
        check(hipHostMalloc(&q, sizeof(int) * kPinnedInts, hipHostMallocDefault), "hipHostMalloc");
        p = static_cast<int*>(q);
    }
    return p;
}

template<typename T>
struct DevBuf
{
    T*     p   = nullptr;
    size_t cap = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    void swap(DevBuf& o)
    {
        std::swap(p, o.p);
        std::swap(cap, o.cap);
    }
    void   ensure(size_t n)
    {
        if (n <= cap && p)
        {
            return;
        }
        if (p)
        {
            check(hipFree(p), "hipFree");
            p = nullptr;
        }
        size_t want = std::max<size_t>(n, 1);
        check(hipMalloc(reinterpret_cast<void**>(&p), want * sizeof(T)), "hipMalloc");
        cap = want;
    }
    void release()
    {
        if (p)
        {
            (void)hipFree(p);
        }
        p   = nullptr;
        cap = 0;
    }
    ~DevBuf() { release(); }
};

// The pair kernel's staged j-force windows and the lists that gather them
// (one per context, one per batch).
struct WindowGather
{
    DevBuf<int>    occCount, occOff, segSize, segBase, gposCount, gposOff, gpos;
    DevBuf<int2>   occ;
    DevBuf<int4>   segInfo;
    DevBuf<float4> staging;
};

// Occurrence lists of the clusters in the groups' windows (once per list).
// A group lists its own clusters and its windowed union, so the lists hold
// at most ncluster + (union entries of all groups) = maxTotal occurrences:
// with maxTotal given the host does not wait for the exact count.
inline void buildOccurrences(hipStream_t s, WindowGather& w, int ngroups, const int4* groups, const int4* groupInfo,
                             const int* unionList, int ncluster, int maxTotal = -1)
{
    w.occCount.ensure(ncluster + 1);
    w.occOff.ensure(scanOutSize(ncluster));
    if (ncluster == 0)
    {
        check(hipMemsetAsync(w.occOff.p, 0, sizeof(int), s), "clear occurrences");
        w.occ.ensure(1);
        return;
    }
    check(hipMemsetAsync(w.occCount.p, 0, sizeof(int) * ncluster, s), "clear occurrences");
    launchOccurrences(s, ngroups, groups, groupInfo, unionList, ncluster, w.occCount.p, nullptr, nullptr, false);
    launchExclusiveScanWide(s, ncluster, w.occCount.p, w.occOff.p, w.occOff.p + ncluster);
    int total = maxTotal;
    if (total < 0)
    {
        int* const h = pinnedInts();
        check(hipMemcpyAsync(h, w.occOff.p + ncluster, sizeof(int), hipMemcpyDeviceToHost, s), "read occurrences");
        check(hipStreamSynchronize(s), "occurrence sync");
        total = h[0];
    }
    w.occ.ensure(std::max(total, 1));
    check(hipMemsetAsync(w.occCount.p, 0, sizeof(int) * ncluster, s), "clear occurrences");
    launchOccurrences(s, ngroups, groups, groupInfo, unionList, ncluster, w.occCount.p, w.occOff.p, w.occ.p, true);
    check(hipGetLastError(), "occurrence kernels");
}

// Staging layout of a schedule (whenever the schedule changes).
// slotF4: float4 per window cluster slot (3: forces, 4: the adjoint);
// occurrences come from occ (the same groups' lists; w itself by default)
inline void buildSegments(hipStream_t s, WindowGather& w, int ngroups, const int4* groups, const int4* groupInfo,
                          int numWaves, const int2* waveStart, int ncluster, int slotF4 = 3,
                          const WindowGather* occ = nullptr)
{
    const WindowGather& o = occ ? *occ : w;
    w.segSize.ensure(ngroups + 1);
    w.segBase.ensure(ngroups + 1);
    w.segInfo.ensure(std::max(ngroups, 1));
    int total = 0;
    if (ngroups > 0)
    {
        launchSegments(s, ngroups, groups, groupInfo, numWaves, waveStart, slotF4, w.segSize.p, w.segBase.p,
                       w.segBase.p + ngroups, w.segInfo.p);
        int* const h = pinnedInts();
        check(hipMemcpyAsync(h, w.segBase.p + ngroups, sizeof(int), hipMemcpyDeviceToHost, s), "read segments");
        check(hipStreamSynchronize(s), "segment sync");
        total = h[0];
    }
    w.staging.ensure(std::max(total, 1));
    // the staging positions each cluster's gather reads
    w.gposCount.ensure(ncluster + 1);
    w.gposOff.ensure(scanOutSize(ncluster));
    // every staged slot is read by at most one cluster's gather (a group's
    // segment holds G + union slots per wavefront, its occurrences own +
    // union entries), so the positions number at most total / slotF4
    const int npos = total / slotF4;
    if (ncluster > 0)
    {
        launchGatherList(s, ncluster, o.occOff.p, o.occ.p, w.segInfo.p, slotF4, w.gposCount.p, nullptr, nullptr,
                         false);
        launchExclusiveScanWide(s, ncluster, w.gposCount.p, w.gposOff.p, w.gposOff.p + ncluster);
    }
    else
    {
        check(hipMemsetAsync(w.gposOff.p, 0, sizeof(int), s), "clear gather");
    }
    w.gpos.ensure(std::max(npos, 1));
    launchGatherList(s, ncluster, o.occOff.p, o.occ.p, w.segInfo.p, slotF4, w.gposCount.p, w.gposOff.p, w.gpos.p,
                     true);
    check(hipGetLastError(), "segment kernels");
}

// Work groups, chunks, work list, static schedule and window staging of one
// kernel over the pair list (groups.hip).
struct WorkLayout
{
    int            groupSize = 0, ngroups = 0, nchunk = 0, windowSlots = 0, nwork = 0;
    int            numBlocks = 0; // workgroups of kKernelWaves wavefronts (the schedule's wavefronts / 4)
    bool           spill     = false; // some union clusters lie beyond their window (global atomics)
    DevBuf<int>    unionCount, chunkCount, workCount, overflow, unionOff, chunkOff, workOff, unionList;
    DevBuf<int4>   chunks, groups, groupInfo;
    // the group pass's output at upper-bound bases, before compaction
    DevBuf<int>    unionTmp;
    DevBuf<int4>   chunksTmp;
    DevBuf<uint2>  workTmp;
    DevBuf<float4> chunkShift;
    DevBuf<uint2>  work;
    DevBuf<int2>   waveStart; // static schedule: each wavefront's first chunk and its group
    DevBuf<int>    chunkCost, costOff; // the schedule's per-chunk costs (and their total) and their scan
    WindowGather   gather;
};

inline int fail(int code, const std::string& msg)
{
    lastError() = msg;
    return code;
}

template<typename F>
int guarded(F&& f)
{
    try
    {
        f();
        lastError().clear();
        return NBX_SUCCESS;
    }
    catch (const InputError& e)
    {
        return fail(NBX_ERROR_INPUT, e.what());
    }
    catch (const StateError& e)
    {
        return fail(NBX_ERROR_STATE, e.what());
    }
    catch (const DeviceError& e)
    {
        return fail(NBX_ERROR_DEVICE, e.what());
    }
    catch (const std::exception& e)
    {
        return fail(NBX_ERROR_DEVICE, e.what());
    }
}

} // namespace nbx

// one system on one device (the state behind an nbx_context handle)
struct nbx_context
{
    nbx_options      opt{};
    int              device     = 0;
    hipStream_t      ownStream  = nullptr;
    hipStream_t      stream     = nullptr;
    // optional HIP events around the pair-kernel launches (benchmarking)
    std::vector<hipEvent_t> timingEvents; // start, stop pairs
    int                     timingRecorded = 0;
    int              numCU      = 256;
    int              elec       = nbx::kElecEwaldTable;
    nbx::InteractionConst ic{};
    double           selfCoef = 0.0;
    float            rlist = 1.0f, rcoulomb = 1.0f;
    bool             twinCut = false; // rvdw < rcoulomb

    // parameters
    int            ntypeUser = -1;
    int            ntype     = 0; // incl. padding
    nbx::DevBuf<float2> nbfp;
    nbx::DevBuf<float4> tabFDV0;
    nbx::DevBuf<double> tabFDV0d;
    nbx::DevBuf<float4> shiftVec;

    // topology
    int         natoms = -1;
    nbx::DevBuf<int> type;
    nbx::DevBuf<float> charge;
    nbx::DevBuf<int> exclOff, exclIdx;
    // excluded pairs (i < j as listed for i, duplicates kept), in atom
    // indices and, after each search, as cluster slots
    nbx::DevBuf<int2> exclPairs, exclPairSlots;
    int               numExclPairs = 0;

    // grid
    nbx::GridDims g{};
    bool        gridValid = false;
    int         ncluster  = 0;
    nbx::DevBuf<float> x;          // device copy of coordinates in atom order
    nbx::DevBuf<int>   cxy_na, cxy_ind, atomCol, atomRank, bucket, atomIndices, cells, nbtype, counter;
    nbx::DevBuf<float4> xq, bb, fnb;
    nbx::DevBuf<unsigned long long> fixF; // window-overflow forces in fixed point, 3 per atom slot

    // list
    bool           listValid = false;
    int            nci = 0, ncj = 0;
    // the fill-only search: j-entries per cluster region (0 before the
    // first search) and the regions
    int                 searchCapJ = 0;
    nbx::DevBuf<uint2>  searchJTmp;
    nbx::DevBuf<int4>   searchITmp;
    nbx::DevBuf<int>    countJ, countI, offJ, offI, totals;
    nbx::DevBuf<int4>   ientry;
    nbx::DevBuf<uint2>  jentry, jentryTmp;
    // the list as searched, kept by the first prune after a search (every
    // prune starts from it); the list above is then the pruned one
    bool                haveOuter = false;
    int                 nciOuter = 0, ncjOuter = 0;
    nbx::DevBuf<int>    offIOuter;
    nbx::DevBuf<int4>   ientryOuter;
    nbx::DevBuf<uint2>  jentryOuter;
    // kernel work layouts: the pair kernel's, and the adjoint's (its own
    // group size, sized for its larger window per cluster slot; built on
    // the first adjoint call after a search)
    nbx::DevBuf<int>    statsPartial; // per-workgroup maxima (launchPartialMax)
    nbx::DevBuf<int>    typeLJ; // per type (padding type included): 1 if it has LJ with any type
    bool                anyTypeWithoutLJ = false; // among the user's types
    // LJ-PME: the grid's 6 C6 per type pair; the host's 6 C6 diagonal and
    // the per-type atom counts give the self energy
    nbx::DevBuf<float>  c6grid;
    std::vector<float>  c6diag;
    std::vector<long>   typeCount;
    nbx::WorkLayout     lay;
    // pruning inside the pair kernel's layout: its chunks and work entries
    // as built (the outer layout), the chunk -> wavefront and group maps,
    // and per-prune scratch; valid until the layout is rebuilt
    bool                        layOuterValid = false;
    nbx::DevBuf<int4>           chunksOuter;
    nbx::DevBuf<uint2>          workOuter;
    nbx::DevBuf<int>            chunkWave, chunkGroup, keptCount, keptMasked;
    nbx::DevBuf<unsigned short> tileKeep;
    nbx::WorkLayout     adj;
    bool                adjValid = false;

    // outputs
    nbx::DevBuf<double>  energyPartial, energy, fshift, virialPartial, virial;
    nbx::DevBuf<double>  exclPartial;
    nbx::DevBuf<float>   fshiftPartial, fOut;
    // adjoint
    nbx::DevBuf<float4>  gq, gradXnb;
    nbx::DevBuf<double>  gradQnb, paramPartial, paramOut, gradQOut;
    nbx::DevBuf<float>   gIn, gradXOut;
    bool            haveEnergy = false, haveShift = false, haveVirial = false;
};

// LJ-PME self energy, 0.5 C6_ii beta_lj^6 / 6 per atom
// (kernel_ref_outer.h:315-322): a constant of the topology and parameters,
// added when the energies are read; 0 without LJ-PME
inline double ljPmeSelfEnergy(const nbx_context* c)
{
    double e = 0.0;
    if (c->opt.vdw_type != NBX_VDW_PME)
    {
        return e;
    }
    for (size_t t = 0; t < c->typeCount.size() && t < c->c6diag.size(); t++)
    {
        e += static_cast<double>(c->typeCount[t]) * 0.5 * static_cast<double>(c->c6diag[t]) / 6.0
             * static_cast<double>(c->ic.lje_coeff6_6);
    }
    return e;
}
