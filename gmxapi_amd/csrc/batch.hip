// Batched ensemble: many independent systems computed by one launch of the
// pair kernel (include/nbx.h, nbx_batch_*).
//
// Each system is set up as an ordinary context (grid, search and work
// groups run per system, nbx_construct_pairlist), then the systems' kernel
// inputs are concatenated into one set of device arrays: cluster indices,
// work-list offsets and exclusion indices are shifted by the system's
// offsets, all padding entries point at one shared far-away cluster, and
// every work group carries its system's shift-vector table and energy slot
// in its groupInfo record.  One launch then covers every group of every
// system: the kernel's static schedule cuts the concatenated chunk list into
// runs of equal cost, so small systems do not leave the GPU idle and one
// wavefront may pass from one system to the next.  Energies are summed per
// system in a fixed order.
// A step (coordinate conversion, force clearing, pair kernel, exclusion
// energies, per-system sums, force reduction) is launch-bound for small
// systems and can be recorded once into a HIP graph and replayed.
#include <cstring>
#include <numeric>

#include "context.h"

using namespace nbx;

struct nbx_batch
{
    nbx_options               opt{};
    int                       ntypes = 0;
    std::vector<double>       c6, c12;
    std::vector<nbx_context*> sys;
    std::vector<int>          natoms;
    std::vector<float>        boxHost;
    int                       device = 0, numCU = 256;
    hipStream_t               ownStream = nullptr, stream = nullptr;

    bool built = false;
    int  totAtoms = 0, totClusters = 0, ngroups = 0, nchunk = 0, nwork = 0, windowSlots = 0, numBlocks = 0;
    bool spill = false; // some member's union clusters lie beyond their group's window (fixed-point atomics)
    int  groupSize = 0, blocksPerCU = 0, chunkWeight = kChunkWeight;
    long long ncj = 0;

    DevBuf<float4> xq, fnb, shiftVec, chunkShift;
    DevBuf<unsigned long long> fixF; // window-overflow forces in fixed point, 3 per atom slot
    DevBuf<int>    nbtype, atomIndices, cells, exclOff, exclIdx, unionList, atomSystem, groupOff, atomOff, slotOff;
    DevBuf<int2>   waveStart;
    DevBuf<float>  boxes;
    WindowGather   gather;
    DevBuf<uint2>  work;
    DevBuf<int4>   chunks, groups, groupInfo;
    DevBuf<double> energyPartial, atomEnergy, energy;
    bool           haveEnergy = false;

    hipGraph_t     graph     = nullptr;
    hipGraphExec_t graphExec = nullptr;

    void dropGraph()
    {
        if (graphExec)
        {
            (void)hipGraphExecDestroy(graphExec);
            graphExec = nullptr;
        }
        if (graph)
        {
            (void)hipGraphDestroy(graph);
            graph = nullptr;
        }
    }
};

namespace
{

void statusCheck(int rc, const char* what)
{
    if (rc == NBX_SUCCESS)
    {
        return;
    }
    const std::string msg = std::string(what) + ": " + lastError();
    if (rc == NBX_ERROR_INPUT)
    {
        throw InputError(msg);
    }
    if (rc == NBX_ERROR_STATE)
    {
        throw StateError(msg);
    }
    throw DeviceError(msg);
}

template<typename T>
std::vector<T> download(const T* p, size_t n)
{
    std::vector<T> h(n);
    if (n)
    {
        check(hipMemcpy(h.data(), p, n * sizeof(T), hipMemcpyDeviceToHost), "batch gather");
    }
    return h;
}

template<typename T>
void upload(DevBuf<T>& d, const std::vector<T>& h)
{
    d.ensure(std::max<size_t>(h.size(), 1));
    if (!h.empty())
    {
        check(hipMemcpy(d.p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice), "batch upload");
    }
}

// The kernel grid (blocksPerCU workgroups per CU, at most what LDS allows)
// and its static schedule, with chunkWeight work entries per chunk.
void scheduleBatch(nbx_batch* b, int blocksPerCU, int chunkWeight)
{
    nbx_context* c0    = b->sys[0];
    const int    lds   = nonbondedLdsBytes(c0->ntype, c0->ic.tab_n, c0->elec, b->windowSlots,
                                           c0->opt.vdw_type == NBX_VDW_PME);
    const int    maxBpc = blocksPerCU > 0 ? blocksPerCU : 4;
    b->blocksPerCU     = std::max(1, std::min(maxBpc, (160 * 1024) / lds));
    b->numBlocks       = std::max(1, std::min((b->nchunk + kKernelWaves - 1) / kKernelWaves, b->blocksPerCU * b->numCU));
    b->chunkWeight     = chunkWeight;
    const int numWaves = b->numBlocks * kKernelWaves;
    const int R        = static_cast<int>(b->sys.size());
    b->waveStart.ensure(numWaves + 1);
    b->slotOff.ensure(R + 1);
    b->energyPartial.ensure(2 * static_cast<size_t>(b->ngroups + numWaves));
    launchWaveSchedule(b->stream, numWaves, b->nchunk, b->chunks.p, b->nwork, chunkWeight, b->ngroups, b->groups.p,
                       b->waveStart.p, nullptr, nullptr, b->numCU * kKernelWaves);
    launchSystemSlots(b->stream, R, b->groupOff.p, b->ngroups, b->groups.p, numWaves, b->waveStart.p, b->slotOff.p);
    check(hipGetLastError(), "schedule kernels");
    buildSegments(b->stream, b->gather, b->ngroups, b->groups.p, b->groupInfo.p, numWaves, b->waveStart.p,
                  b->totClusters);
    check(hipStreamSynchronize(b->stream), "schedule sync");
}

// Regroups every system with group size G (0 = each system's own choice)
// and concatenates the kernel inputs.
void buildBatch(nbx_batch* b, int G, int blocksPerCU, int chunkWeight)
{
    b->dropGraph();
    const int R = static_cast<int>(b->sys.size());
    if (R == 0)
    {
        throw StateError("the batch holds no systems");
    }
    for (nbx_context* c : b->sys)
    {
        c->opt.kernel_group_size = G;
        statusCheck(nbx_construct_pairlist(c), "nbx_construct_pairlist");
    }
    check(hipSetDevice(b->device), "hipSetDevice");

    std::vector<float4> hXq;
    std::vector<int>    hType, hAi, hCells, hExOff{ 0 }, hExIdx, hUnion, hAtomSys, hGroupOff{ 0 }, hAtomOff{ 0 };
    std::vector<uint2>  hWork;
    std::vector<int4>   hChunks, hGroups, hInfo;
    std::vector<float4> hShift, hChunkShift;
    int                 clusterOff = 0, atomOffset = 0, maxSlots = 0;
    bool                anySpill = false;
    long long           ncj        = 0;
    int                 totClusters = 0;
    for (nbx_context* c : b->sys)
    {
        totClusters += c->ncluster;
    }
    for (int r = 0; r < R; r++)
    {
        nbx_context* c  = b->sys[r];
        const int    nc = c->ncluster;
        const int    na = c->natoms;
        const int    wo = static_cast<int>(hWork.size());
        const int    co = static_cast<int>(hChunks.size());
        const int    uo = static_cast<int>(hUnion.size());
        const int    go = static_cast<int>(hGroups.size());

        std::vector<float4> xq = download(c->xq.p, 4 * static_cast<size_t>(nc) + 4);
        std::vector<int>    ty = download(c->nbtype.p, 4 * static_cast<size_t>(nc) + 4);
        std::vector<int>    ai = download(c->atomIndices.p, 4 * static_cast<size_t>(nc));
        std::vector<int>    ce = download(c->cells.p, na);
        std::vector<uint2>  wk = download(c->lay.work.p, c->lay.nwork);
        std::vector<int4>   ch = download(c->lay.chunks.p, c->lay.nchunk);
        std::vector<float4> cs = download(c->lay.chunkShift.p, c->lay.nchunk);
        std::vector<int4>   gr = download(c->lay.groups.p, c->lay.ngroups);
        std::vector<int4>   gi = download(c->lay.groupInfo.p, c->lay.ngroups);
        const int4          last = gr.back();
        std::vector<int>    un = download(c->lay.unionList.p, last.z + last.w);
        std::vector<float4> sv = download(c->shiftVec.p, kNumShifts);
        std::vector<int>    eo = download(c->exclOff.p, na + 1);
        std::vector<int>    ei = download(c->exclIdx.p, eo[na]);

        hXq.insert(hXq.end(), xq.begin(), xq.begin() + 4 * nc);
        hType.insert(hType.end(), ty.begin(), ty.begin() + 4 * nc);
        if (r == R - 1)
        {
            // the shared padding cluster: far-away fillers of the padding type
            hXq.insert(hXq.end(), xq.begin() + 4 * nc, xq.end());
            hType.insert(hType.end(), ty.begin() + 4 * nc, ty.end());
        }
        for (int v : ai)
        {
            hAi.push_back(v >= 0 ? v + atomOffset : -1);
        }
        for (int v : ce)
        {
            hCells.push_back(v + 4 * clusterOff);
        }
        for (uint2 e : wk)
        {
            e.x = (static_cast<int>(e.x) == nc) ? static_cast<unsigned>(totClusters)
                                                 : e.x + static_cast<unsigned>(clusterOff);
            hWork.push_back(e);
        }
        for (int4 v : ch)
        {
            hChunks.push_back(make_int4(v.x + clusterOff, v.y, v.z + wo, v.w));
        }
        for (size_t g = 0; g < gr.size(); g++)
        {
            const int4 v = gr[g];
            hGroups.push_back(make_int4(v.x + co, v.y + co, v.z + uo, v.w));
            hInfo.push_back(make_int4(gi[g].x + clusterOff, gi[g].y, kNumShifts * r, go + static_cast<int>(g)));
        }
        for (int v : un)
        {
            hUnion.push_back(v + clusterOff);
        }
        hShift.insert(hShift.end(), sv.begin(), sv.end());
        hChunkShift.insert(hChunkShift.end(), cs.begin(), cs.end());
        for (int k = 0; k < na; k++)
        {
            hExOff.push_back(hExOff.back() + (eo[k + 1] - eo[k]));
            hAtomSys.push_back(r);
        }
        for (int v : ei)
        {
            hExIdx.push_back(v + atomOffset);
        }
        maxSlots = std::max(maxSlots, c->lay.windowSlots);
        anySpill = anySpill || c->lay.spill;
        ncj += c->ncj;
        clusterOff += nc;
        atomOffset += na;
        hGroupOff.push_back(static_cast<int>(hGroups.size()));
        hAtomOff.push_back(atomOffset);
    }

    upload(b->xq, hXq);
    upload(b->nbtype, hType);
    upload(b->atomIndices, hAi);
    upload(b->cells, hCells);
    upload(b->work, hWork);
    upload(b->chunks, hChunks);
    upload(b->groups, hGroups);
    upload(b->groupInfo, hInfo);
    upload(b->unionList, hUnion);
    upload(b->shiftVec, hShift);
    upload(b->chunkShift, hChunkShift);
    upload(b->exclOff, hExOff);
    upload(b->exclIdx, hExIdx);
    upload(b->atomSystem, hAtomSys);
    upload(b->groupOff, hGroupOff);
    upload(b->atomOff, hAtomOff);
    upload(b->boxes, b->boxHost);
    b->fnb.ensure(4 * static_cast<size_t>(totClusters) + 4);
    b->fixF.ensure(3 * (4 * static_cast<size_t>(totClusters) + 4));
    b->atomEnergy.ensure(std::max(atomOffset, 1));
    b->energy.ensure(2 * static_cast<size_t>(R));

    b->totAtoms    = atomOffset;
    b->totClusters = totClusters;
    b->ngroups     = static_cast<int>(hGroups.size());
    b->nchunk      = static_cast<int>(hChunks.size());
    b->nwork       = static_cast<int>(hWork.size());
    b->windowSlots = maxSlots;
    b->spill       = anySpill;
    b->ncj         = ncj;
    b->groupSize   = G > 0 ? G : b->sys[0]->lay.groupSize;
    buildOccurrences(b->stream, b->gather, b->ngroups, b->groups.p, b->groupInfo.p, b->unionList.p, totClusters,
                     totClusters + static_cast<int>(hUnion.size()));
    scheduleBatch(b, blocksPerCU, chunkWeight);
    b->built      = true;
    b->haveEnergy = false;
}

void enqueueCompute(nbx_batch* b, const float* x, float* f, unsigned flags)
{
    nbx_context* c0     = b->sys[0];
    const bool   energy = (flags & NBX_COMPUTE_ENERGY) != 0;
    const int numSlots = b->ngroups + b->numBlocks * kKernelWaves;
    // the cluster force buffer when the forces stay in the cluster layout
    // or windows overflow with float atomics; the fixed-point buffer when
    // they overflow with NBX_COMPUTE_REPRODUCIBLE
    const bool                reproducible = (flags & NBX_COMPUTE_REPRODUCIBLE) != 0;
    float4* const             fnb = (!f || (b->spill && !reproducible)) ? b->fnb.p : nullptr;
    unsigned long long* const fix = b->spill && reproducible ? b->fixF.p : nullptr;
    launchPrepareStep(b->stream, 4 * b->totClusters, b->atomIndices.p, x, b->xq.p, fnb,
                      energy ? b->energyPartial.p : nullptr, 2 * numSlots, fix);
    KernelArgs a{};
    a.chunks       = b->chunks.p;
    a.groups       = b->groups.p;
    a.groupInfo    = b->groupInfo.p;
    a.ngroups      = b->ngroups;
    a.groupSize    = b->groupSize;
    a.ncluster     = b->totClusters;
    a.windowSlots  = b->windowSlots;
    a.unionList    = b->unionList.p;
    a.work         = b->work.p;
    a.xq           = b->xq.p;
    a.type         = b->nbtype.p;
    a.nbfp         = c0->nbfp.p;
    a.typeLJ       = c0->typeLJ.p;
    a.skipLJ       = c0->anyTypeWithoutLJ ? 1 : 0;
    a.vdwModifier  = c0->opt.vdw_modifier;
    a.ljPme        = c0->opt.vdw_type == NBX_VDW_PME ? 1 : 0;
    a.twinCut      = c0->twinCut ? 1 : 0;
    a.windowSpill  = b->spill ? 1 : 0;
    a.c6grid       = c0->c6grid.p;
    a.ntype        = c0->ntype;
    a.tabFDV0      = c0->tabFDV0.p;
    a.chunkShift   = b->chunkShift.p;
    a.ic           = c0->ic;
    a.waveStart    = b->waveStart.p;
    a.segInfo      = b->gather.segInfo.p;
    a.staging      = b->gather.staging.p;
    a.fnb          = b->fnb.p;
    a.fixF         = fix;
    a.energyPartial = b->energyPartial.p;
    a.fshiftPartial = nullptr;
    check(launchNonbonded(b->stream, a, c0->elec, energy, false, b->numBlocks), "nonbonded kernel");
    launchGatherForces(b->stream, b->totClusters, b->gather.gposOff.p, b->gather.gpos.p, b->gather.staging.p,
                       fnb, b->atomIndices.p, f, (flags & NBX_COMPUTE_ADD_FORCES) != 0, nullptr, nullptr, fix);
    if (energy)
    {
        ExclusionArgs e{};
        e.natoms     = b->totAtoms;
        e.cells      = b->cells.p;
        e.xq         = b->xq.p;
        e.exclOff    = b->exclOff.p;
        e.exclIdx    = b->exclIdx.p;
        e.elec       = c0->elec;
        e.ic         = c0->ic;
        e.rc2        = static_cast<double>(c0->ic.rcoulomb2);
        e.epsfacD    = static_cast<double>(c0->ic.epsfac);
        e.selfCoef   = c0->selfCoef;
        e.tabD       = c0->tabFDV0d.p;
        e.atomSystem = b->atomSystem.p;
        e.boxes      = b->boxes.p;
        e.atomEnergy = b->atomEnergy.p;
        launchExclusionEnergy(b->stream, e);
        launchBatchFinalize(b->stream, static_cast<int>(b->sys.size()), b->slotOff.p, b->energyPartial.p,
                            b->atomOff.p, b->atomEnergy.p, b->energy.p);
    }
}

void requireBuilt(const nbx_batch* b)
{
    if (!b->built)
    {
        throw StateError("call nbx_batch_construct_pairlists first");
    }
}

void checkFlags(unsigned flags)
{
    if (flags & (NBX_COMPUTE_SHIFT_FORCES | NBX_COMPUTE_VIRIAL))
    {
        throw InputError("a batch computes forces and energies; shift forces and the virial are per-system calls");
    }
}

} // namespace

extern "C" {

int nbx_batch_create(const nbx_options* options, int numTypes, const double* c6, const double* c12, nbx_batch** batch)
{
    if (!options || !batch || !c6 || !c12 || numTypes <= 0)
    {
        return fail(NBX_ERROR_INPUT, "nbx_batch_create: NULL argument or no types");
    }
    *batch = nullptr;
    auto* b = new nbx_batch();
    int   rc = guarded([&] {
        b->opt    = *options;
        b->ntypes = numTypes;
        b->c6.assign(c6, c6 + numTypes * numTypes);
        b->c12.assign(c12, c12 + numTypes * numTypes);
        // validates the options and the device like a single context would
        nbx_context* probe = nullptr;
        statusCheck(nbx_create(options, &probe), "nbx_create");
        b->numCU = probe->numCU;
        nbx_destroy(probe);
        b->device = options->device;
        check(hipSetDevice(b->device), "hipSetDevice");
        check(hipStreamCreateWithFlags(&b->ownStream, hipStreamNonBlocking), "hipStreamCreate");
        b->stream = b->ownStream;
    });
    if (rc != NBX_SUCCESS)
    {
        nbx_batch_destroy(b);
        return rc;
    }
    *batch = b;
    return NBX_SUCCESS;
}

int nbx_batch_destroy(nbx_batch* b)
{
    if (!b)
    {
        return NBX_SUCCESS;
    }
    (void)hipSetDevice(b->device);
    if (b->stream)
    {
        (void)hipStreamSynchronize(b->stream);
    }
    b->dropGraph();
    for (nbx_context* c : b->sys)
    {
        nbx_destroy(c);
    }
    if (b->ownStream)
    {
        (void)hipStreamDestroy(b->ownStream);
    }
    delete b;
    return NBX_SUCCESS;
}

int nbx_batch_set_stream(nbx_batch* b, void* stream)
{
    if (!b)
    {
        return fail(NBX_ERROR_INPUT, "NULL batch");
    }
    return guarded([&] {
        b->stream = static_cast<hipStream_t>(stream);
        for (nbx_context* c : b->sys)
        {
            statusCheck(nbx_set_stream(c, stream), "nbx_set_stream");
        }
        b->dropGraph();
    });
}

int nbx_batch_add_system(nbx_batch* b, int numAtoms, const int* types, const float* charges, const int* exclOffsets,
                         const int* exclIndices, const float* box, const float* x, int* index)
{
    if (!b || !box || !x)
    {
        return fail(NBX_ERROR_INPUT, "nbx_batch_add_system: NULL argument");
    }
    return guarded([&] {
        nbx_context* c = nullptr;
        statusCheck(nbx_create(&b->opt, &c), "nbx_create");
        b->sys.push_back(c);
        statusCheck(nbx_set_stream(c, b->stream), "nbx_set_stream");
        statusCheck(nbx_set_nonbonded_parameters(c, b->ntypes, b->c6.data(), b->c12.data()),
                    "nbx_set_nonbonded_parameters");
        statusCheck(nbx_set_topology(c, numAtoms, types, charges, exclOffsets, exclIndices), "nbx_set_topology");
        statusCheck(nbx_put_on_grid(c, box, x, 0), "nbx_put_on_grid");
        b->natoms.push_back(numAtoms);
        b->boxHost.insert(b->boxHost.end(), box, box + 3);
        b->built = false;
        b->dropGraph();
        if (index)
        {
            *index = static_cast<int>(b->sys.size()) - 1;
        }
    });
}

int nbx_batch_construct_pairlists(nbx_batch* b)
{
    if (!b)
    {
        return fail(NBX_ERROR_INPUT, "NULL batch");
    }
    return guarded([&] { buildBatch(b, b->opt.kernel_group_size, b->opt.kernel_blocks_per_cu, kChunkWeight); });
}

int nbx_batch_get_info(nbx_batch* b, int64_t* info)
{
    if (!b || !info)
    {
        return fail(NBX_ERROR_INPUT, "NULL argument");
    }
    info[0] = static_cast<int64_t>(b->sys.size());
    info[1] = b->totAtoms;
    info[2] = b->totClusters;
    info[3] = b->ncj;
    info[4] = b->ngroups;
    info[5] = b->groupSize;
    info[6] = b->blocksPerCU;
    return NBX_SUCCESS;
}

int nbx_batch_compute(nbx_batch* b, const float* x, float* f, unsigned flags)
{
    if (!b || !x || !f)
    {
        return fail(NBX_ERROR_INPUT, "nbx_batch_compute: NULL argument");
    }
    return guarded([&] {
        requireBuilt(b);
        checkFlags(flags);
        check(hipSetDevice(b->device), "hipSetDevice");
        enqueueCompute(b, x, f, flags);
        check(hipGetLastError(), "batch kernels");
        b->haveEnergy = (flags & NBX_COMPUTE_ENERGY) != 0;
    });
}

int nbx_batch_capture_graph(nbx_batch* b, const float* x, float* f, unsigned flags)
{
    if (!b || !x || !f)
    {
        return fail(NBX_ERROR_INPUT, "nbx_batch_capture_graph: NULL argument");
    }
    return guarded([&] {
        requireBuilt(b);
        checkFlags(flags);
        if (!b->stream)
        {
            throw InputError("graph capture needs a non-default stream");
        }
        check(hipSetDevice(b->device), "hipSetDevice");
        b->dropGraph();
        check(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
        enqueueCompute(b, x, f, flags);
        hipGraph_t g = nullptr;
        check(hipStreamEndCapture(b->stream, &g), "hipStreamEndCapture");
        b->graph = g;
        check(hipGraphInstantiate(&b->graphExec, b->graph, nullptr, nullptr, 0), "hipGraphInstantiate");
        check(hipGraphLaunch(b->graphExec, b->stream), "hipGraphLaunch");
        b->haveEnergy = (flags & NBX_COMPUTE_ENERGY) != 0;
    });
}

int nbx_batch_launch_graph(nbx_batch* b)
{
    if (!b)
    {
        return fail(NBX_ERROR_INPUT, "NULL batch");
    }
    return guarded([&] {
        if (!b->graphExec)
        {
            throw StateError("no graph: call nbx_batch_capture_graph first");
        }
        check(hipGraphLaunch(b->graphExec, b->stream), "hipGraphLaunch");
    });
}

int nbx_batch_get_energies(nbx_batch* b, double* energies)
{
    if (!b || !energies)
    {
        return fail(NBX_ERROR_INPUT, "NULL argument");
    }
    return guarded([&] {
        if (!b->haveEnergy)
        {
            throw StateError("the last batch compute did not request energies");
        }
        check(hipSetDevice(b->device), "hipSetDevice");
        check(hipStreamSynchronize(b->stream), "sync");
        check(hipMemcpy(energies, b->energy.p, 2 * b->sys.size() * sizeof(double), hipMemcpyDeviceToHost),
              "read energies");
        for (size_t r = 0; r < b->sys.size(); r++)
        {
            energies[2 * r] += ljPmeSelfEnergy(b->sys[r]);
        }
    });
}

int nbx_batch_autotune(nbx_batch* b, const float* x, float* f, unsigned flags, int repeats, int* chosen)
{
    if (!b || !x || !f || repeats <= 0)
    {
        return fail(NBX_ERROR_INPUT, "nbx_batch_autotune: NULL argument or no repeats");
    }
    return guarded([&] {
        checkFlags(flags);
        check(hipSetDevice(b->device), "hipSetDevice");
        hipEvent_t e0, e1;
        check(hipEventCreate(&e0), "hipEventCreate");
        check(hipEventCreate(&e1), "hipEventCreate");
        float best = -1.0f;
        int   bestG = 0, bestBpc = 0, bestWeight = 0;
        for (int G : { 16, 8, 4, 2, 1 })
        {
            try
            {
                buildBatch(b, G, 4, kChunkWeight);
            }
            catch (const InputError&)
            {
                continue; // this group size does not fit: try the next
            }
            const int fitBpc = b->blocksPerCU;
            for (int bpc = fitBpc; bpc >= 2; bpc--)
            {
                for (int weight : { kChunkWeight, 4 * kChunkWeight })
                {
                    scheduleBatch(b, bpc, weight);
                    enqueueCompute(b, x, f, flags); // warm-up
                    check(hipEventRecord(e0, b->stream), "hipEventRecord");
                    for (int k = 0; k < repeats; k++)
                    {
                        enqueueCompute(b, x, f, flags);
                    }
                    check(hipEventRecord(e1, b->stream), "hipEventRecord");
                    check(hipEventSynchronize(e1), "hipEventSynchronize");
                    float ms = 0.0f;
                    check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
                    if (best < 0 || ms < best)
                    {
                        best       = ms;
                        bestG      = G;
                        bestBpc    = b->blocksPerCU;
                        bestWeight = weight;
                    }
                }
            }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (best < 0)
        {
            throw InputError("no kernel schedule fits this batch");
        }
        buildBatch(b, bestG, bestBpc, bestWeight);
        if (chosen)
        {
            chosen[0] = bestG;
            chosen[1] = b->blocksPerCU;
            chosen[2] = bestWeight;
        }
    });
}

int nbx_batch_get_system(nbx_batch* b, int index, nbx_context** context)
{
    if (!b || !context || index < 0 || index >= static_cast<int>(b->sys.size()))
    {
        return fail(NBX_ERROR_INPUT, "nbx_batch_get_system: bad argument");
    }
    *context = b->sys[index];
    return NBX_SUCCESS;
}

} // extern "C"
