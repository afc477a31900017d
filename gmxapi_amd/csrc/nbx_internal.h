// Internal declarations shared by the HIP translation units of libnbx.
//
// Data layout in HBM (one context = one system on one device):
//   grid      cxy_na[ncol], cxy_ind[ncol+1] (first cluster of each column),
//             atomIndices[nslot] (-1 = filler), cells[natoms] (atom -> slot),
//             bb[2*ncluster] (float4 lower, float4 upper, w unused)
//   atoms     xq[nslot] float4 {x, y, z, q}, type[nslot] int
//   list      ientry[nci] int4 {ci, shift, cj_start, cj_end},
//             jentry[ncj] uint2 {cj, exclusion mask (bit 4*i + j)}
//   work      chunks int4, work[] uint2 {cj, window slot << 16 | mask}: the
//             list regrouped for the kernel, chunks padded to 16 entries
//             with a dummy far-away cluster (index ncluster)
//   forces    fnb[nslot] float4 accumulation buffer, f[natoms*3] output
//   params    nbfp[(T+1)^2] float2 {6*C6, 12*C12}, Ewald FDV0 table float4,
//             shift vectors float4[45]
// A slot is one atom position of a 4-atom cluster; nslot = 4 * ncluster.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

namespace nbx
{

constexpr int kClusterSize = 4;
constexpr int kNumShifts   = 45;
constexpr int kCentral     = 22;
constexpr unsigned kMaskAll  = 0xffffffffu;
constexpr unsigned kMaskDiag = 0x08ceu;
constexpr float kFarAway     = -1000000.0f;
// LJ types held in LDS by the kernel: (63 + 1)^2 float2 = 32 KiB
constexpr int kMaxTypes = 63;
// the per-column rank sort is quadratic in the column population
constexpr int kMaxColumnAtoms = 65536;
// kernel work groups: at most this many consecutive i-clusters share a
// wavefront's LDS j-force window; a group's j-entries are sorted in LDS by
// the group pass (at most kGroupMaxJ)
constexpr int kGroupSizeMax = 16;

// The canonical squared distance fl(fl(fl(dx*dx) + fl(dy*dy)) + fl(dz*dz)),
// the rounding sequence that decides which pairs are within rlist.  hipcc
// contracts a multiply feeding an add into one fused multiply-add by default
// (and __fmul_rn/__fadd_rn are plain operators in its headers), which rounds
// dx*dx + dy*dy once instead of three times and moves pairs at d^2 == rlist^2
// across the cut; contraction is switched off for this expression.
__device__ __forceinline__ float canonicalD2(float dx, float dy, float dz)
{
#pragma clang fp contract(off)
    const float a = dx * dx;
    const float b = dy * dy;
    const float c = dz * dz;
    return (a + b) + c;
}
// candidate group sizes: kGroupSizeMax, kGroupSizeMax / 2, ..., 1
constexpr int kGroupSizeCount = 5;
constexpr int kGroupMaxJ    = 4096;
// j-clusters per kernel tile, j-entries per work chunk (a multiple of it;
// large enough that a chunk is a whole i-entry)
constexpr int kTileJ        = 16;
constexpr int kChunkEntries = 1 << 20;
// cost of one chunk (i-entry) in work entries for the static schedule: its
// i-data set-up and i-force reduction
constexpr int kChunkWeight = 8;
// chunk cost model of the static schedule (1/32 tile) and the extra share of
// the first resident generation of workgroups (1/1000)
constexpr int kCostTile            = 32;
constexpr int kCostChunk           = 27;
constexpr int kGenerationSkewMilli = 45;
// wavefronts per nonbonded workgroup, and the LDS a workgroup may use (two
// resident workgroups per CU of 160 KiB, eight wavefronts); the adjoint may
// use a CU's whole LDS
constexpr int kKernelWaves      = 4;
constexpr int kKernelLdsBudget  = 80 * 1024;
constexpr int kAdjointLdsBudget = 160 * 1024;
// LDS per two-wavefront adjoint workgroup that its group size is chosen for
// (three resident workgroups per CU)
constexpr int kAdjointLayoutBudget = 48 * 1024;
// copies of the adjoint's per-workgroup C6/C12 accumulators in LDS (lanes
// spread over them, fewer lanes per address in each LDS atomic); the layout
// budget grows by the copies beyond the first
#ifndef NBX_ADJ_ACC_COPIES
#define NBX_ADJ_ACC_COPIES 2
#endif
__host__ __device__ inline int adjointAccCopies(int ntype)
{
    return ntype <= 24 ? NBX_ADJ_ACC_COPIES : 1;
}
__host__ __device__ inline int adjointLayoutBudget(int ntype)
{
    return kAdjointLayoutBudget + (adjointAccCopies(ntype) - 1) * ntype * ntype * 16;
}
// rsq floor for float (pairlist.h:146): r > 6.2e-4 nm
constexpr float kMinRsq = 3.82e-07f;

// Coulomb flavours of the kernel
enum Elec : int
{
    kElecRF         = 0, // reaction-field, also plain cut-off with eps_rf = 1
    kElecEwaldTable = 1, // quadratic-spline correction table (plain-C reference)
    kElecEwaldAna   = 2, // analytical correction
};

// LJ flavours of the pair kernel: the potential shift with every type
// having LJ, the potential shift testing each i-atom's LJ flag (a type with
// no LJ contributes exact zeros, so its terms are skipped), and the force
// and potential switches (which test the flag too)
enum LJFlavour : int
{
    kLJShift       = 0,
    kLJShiftSkip   = 1,
    kLJForceSwitch = 2,
    kLJPotSwitch   = 3,
    kLJEwald       = 4, // LJ-PME real space (potential shift; tests the flag)
    // flag bit on the flavour: the twin cut-off, LJ masked beyond rvdw
    // (VDW_CUTOFF_CHECK, kernel_ref_inner.h:254-271)
    kLJTwin        = 8,
    kLJFlavourMask = 7,
};

struct InteractionConst
{
    float rlist2;
    float rcoulomb2;
    float rvdw2; // the LJ cut-off squared (below rcoulomb2 with the twin cut-off)
    float epsfac;
    float k_rf;
    float c_rf;
    float beta;
    float sh_ewald;
    float cpot_disp;
    float cpot_rep;
    float cpot_disp6; // cpot_disp / 6, rounded once
    float cpot_rep12; // cpot_rep / 12, rounded once
    float tab_scale;
    int   tab_n;
    float self_v0; // V(0) of the Ewald correction (or c_rf for RF)
    float beta2;   // beta^2 and beta^3 for the analytical correction
    float beta3;
    // LJ modifiers: the switch radius; force switch: force coefficients
    // c2, c3 and 12-times-energy coefficients 4 c2, 3 c3 for dispersion (d)
    // and repulsion (r); potential switch: sw coefficients c3..c5 and its
    // derivative's 3 c3, 4 c4, 5 c5
    float rvdw_switch;
    float fsw_c2d, fsw_c3d, fsw_c2r, fsw_c3r;
    float fsw_e2d, fsw_e3d, fsw_e2r, fsw_e3r;
    float sw_c3, sw_c4, sw_c5, sw_f3, sw_f4, sw_f5;
    // LJ-PME: beta_lj^2, beta_lj^6 / 6 and the grid part's potential shift
    float lje_coeff2, lje_coeff6_6, sh_lj_ewald;
};

struct GridDims
{
    int   natoms;
    int   ncx, ncy, ncol;
    float box[3];
    float cellSize[2];
    float invCellSize[2];
    float invBoxZ;
};

// -- prune.hip: the outer list (as searched) pruned to rlist2 with the
// current coordinates xq, as a count pass (countI, countJ per cluster) and,
// after exclusive scans into offI, offJ, a fill pass
struct PruneArgs
{
    int           ncluster;
    float         rlist2;
    const float4* xq;
    const float4* shiftVec;
    const int*    offIOuter;
    const int4*   ientryOuter;
    const uint2*  jentryOuter;
    int*          countI;
    int*          countJ;
    const int*    offI;
    const int*    offJ;
    int4*         ientry;
    uint2*        jentry;
};
void launchPrune(hipStream_t s, const PruneArgs& a, bool fill);

// Pruning inside the pair kernel's work layout (prune.hip): the layout's
// chunks and work entries as built from the searched list are kept (the
// outer layout); a prune rewrites, for the chunks of the wavefronts in the
// selected part, the kept entries compacted at the front of each
// wavefront's region of the work list, every chunk padded to whole tiles
// (at least one), with its tile and masked-tile counts.  Groups, window
// slots, the static schedule and the staging layout stay as they are.
struct LayoutPruneArgs
{
    int            nchunk;
    int            numWaves;
    int            ngroups;
    int            part, numParts;  // wavefronts w with w % numParts == part
    int            groupSize;       // the layout's G (padding entries use slot G + windowed union)
    int            ncluster;        // the far-away dummy cluster of the padding entries
    float          rlist2;
    const float4*  xq;
    const float4*  chunkShift;      // per chunk: its i-entry's shift vector
    const int2*    waveStart;       // [numWaves + 1]
    const int4*    groups;          // [ngroups]: chunk range, union offset, windowed union size
    const int4*    chunksOuter;     // the layout as built
    const uint2*   workOuter;
    int*           chunkWave;       // set-up: each chunk's wavefront
    int*           chunkGroup;      // set-up: each chunk's group
    int*           keptCount;       // per chunk: kept entries
    int*           keptMasked;      // per chunk: kept entries with exclusion or diagonal bits
    unsigned short* tileKeep;       // per outer tile: one bit per kept entry
    int4*          chunks;          // the pruned layout
    uint2*         work;
};
void launchLayoutPruneSetup(hipStream_t s, const LayoutPruneArgs& a);
void launchLayoutPrune(hipStream_t s, const LayoutPruneArgs& a);

// -- grid.hip
void launchGridColumns(hipStream_t s, const GridDims& g, const float* x, int* atomCol, int* atomRank, int* cxy_na);
void launchGridScan(hipStream_t s, int ncol, const int* cxy_na, int* cxy_ind, int* ncluster_out);
void launchGridSortColumns(hipStream_t s, const GridDims& g, const float* x, const int* atomCol, const int* atomRank,
                           const int* cxy_na, const int* cxy_ind, int* bucket, int* atomIndices, int* cells);
void launchGridFill(hipStream_t s, int ncluster, const float* x, const int* atomIndices, const int* type, const float* q,
                    int paddingType, float4* xq, int* nbtype, float4* bb);

// -- search.hip
struct SearchArgs
{
    GridDims      g;
    float         rlist;
    float         rlist2;
    int           ncluster;
    const int*    cxy_ind;
    const float4* bb;
    const float4* xq;
    const int*    atomIndices;
    const int*    cells;
    const int*    exclOff;
    const int*    exclIdx;
    // outputs (count pass writes counts; fill pass writes the lists)
    int*   countJ;   // per ci: number of j entries
    int*   countI;   // per ci: number of i entries
    const int* offJ; // per ci: first j entry (exclusive scan of countJ)
    const int* offI; // per ci: first i entry
    int4*  ientry;
    uint2* jentry;
    // fill pass without a count pass (capJ > 0): cluster ci's entries at
    // jentry[ci * capJ] and ientry[ci * kMaxIEntriesPerCluster], counts
    // written, entries past capJ dropped (the host sees the count)
    int    capJ;
};
// shift indices at or below the central one: at most 14 i-entries per cluster
constexpr int kMaxIEntriesPerCluster = 14;
void launchSearch(hipStream_t s, const SearchArgs& a, bool fill);
void launchSearchCompact(hipStream_t s, int ncluster, int capJ, const int* countJ, const int* countI,
                         const int* offJ, const int* offI, const uint2* jentryTmp, const int4* ientryTmp,
                         uint2* jentry, int4* ientry);
void launchExclusiveScan(hipStream_t s, int n, const int* in, int* out, int* total);
// count (at most 3) independent scans of n elements each in one launch, one
// workgroup each
void launchExclusiveScans(hipStream_t s, int n, int count, const int* const* in, int* const* out, int* const* total);
// The same scans over many workgroups for long lists (n >= kScanWideMin):
// one launch sums tiles of kScanTile elements, a second scans each tile
// after the sum of the tiles before it.  The tile sums are kept in the
// output's tail, after out[n]: every out holds scanOutSize(n) ints.
constexpr int kScanTile    = 4096;
constexpr int kScanWideMin = 16384;
inline size_t scanOutSize(size_t n)
{
    return n + 1 + (n + kScanTile - 1) / kScanTile;
}
void launchExclusiveScansWide(hipStream_t s, int n, int count, const int* const* in, int* const* out,
                              int* const* total);
inline void launchExclusiveScanWide(hipStream_t s, int n, const int* in, int* out, int* total)
{
    launchExclusiveScansWide(s, n, 1, &in, &out, &total);
}

// -- groups.hip
struct GroupArgs
{
    int          ncluster;
    int          groupSize;
    int          ngroups;
    int          chunkEntries;
    const int*   offI;      // per cluster: first i-entry (ncluster + 1)
    const int*   offJ;      // per cluster: first j-entry (ncluster + 1)
    const int4*  ientry;    // search output {ci, shift, cj_start, cj_end}
    const uint2* jentryIn;  // search output {cj, mask}
    uint2*       jentryOut; // the same entries, masked entries first
    int*         unionCount; // count pass, per group
    int*         chunkCount; // count pass, per group
    int*         workCount;  // count pass, per group: padded work entries
    int*         overflow;   // count pass, per group
    const int*   unionOff;   // fill pass: exclusive scan of unionCount
    const int*   chunkOff;   // fill pass: exclusive scan of chunkCount
    const int*   workOff;    // fill pass: exclusive scan of workCount
    int*         unionList;
    uint2*       work;       // {cj, slot << 16 | mask}, each chunk padded to whole tiles
    int4*        chunks;     // {ci, shift | self << 8 | special << 9, work_start, tiles}
    int4*        groups;     // {chunk_start, chunk_end, union_offset, union_size}
    int4*        groupInfo;  // {first own cluster, own count | G << 16, shift-table offset, group index}
    int          windowCap;  // fill pass: window slots (own, union, spare); union beyond -> kSlotOverflow
    // fill pass without a count pass: writes the per-group counts and the
    // layout at upper-bound bases (union at offJ[c0], chunks at offI[c0],
    // work at offJ[c0] + (kTileJ - 1) offI[c0]) for launchGroupCompact
    int          upperBound;
    // 1: every group's union by a bitonic sort of its keys (the path taken
    // by groups whose j-clusters span more than the bitmap covers), for tests
    int          unionBySort;
};
// window slot of a j-entry whose cluster did not fit the group's window:
// the kernel adds its forces with global atomics
constexpr unsigned kSlotOverflow = 0xffffu;
// Forces of union clusters beyond a window are summed in fixed point with
// kFixBits fraction bits in 64-bit integers: integer addition does not depend
// on the order of the atomics, so the sums are reproducible.  Resolution
// 2^-20 (about 1e-6 kJ mol^-1 nm^-1), range +-2^43 (about 8.8e12).
constexpr int kFixBits = 20;
void launchBuildGroups(hipStream_t s, const GroupArgs& a, bool fill);
// moves an upper-bound layout (unionTmp, workTmp, chunksTmp) to a.unionOff,
// a.workOff, a.chunkOff in a.unionList, a.work, a.chunks; rebases a.groups
void launchGroupCompact(hipStream_t s, const GroupArgs& a, const int* unionTmp, const uint2* workTmp,
                        const int4* chunksTmp);
// Maxima for buffer sizing, as per-workgroup partials that the host reads
// and reduces (reducePartialMax): of nv (at most 2) arrays of n non-negative
// ints, and with offJ (the scanned j-entry counts of n clusters) of the
// j-entry span of every group of each candidate size (kGroupSizeMax, ...,
// 1, in that order, after the arrays').  partial holds kPartialMaxBlocks *
// stride ints; launchPartialMax returns the number of workgroups.
constexpr int kPartialMaxThreads = 256;
constexpr int kPartialMaxBlocks  = 128;
struct PartialMaxArgs
{
    const int* v[2];
    int        nv;
    const int* offJ;
    int        n;
    int*       partial;
};
__host__ __device__ constexpr int partialMaxStride(int nv, bool spans)
{
    return nv + (spans ? kGroupSizeCount : 0);
}
int  launchPartialMax(hipStream_t s, const PartialMaxArgs& a);
void reducePartialMax(const int* partial, int blocks, int stride, int* out);
// Occurrences of every cluster in the groups' j-force windows (own slot and
// windowed-union slots) as a CSR list of {group, slot}, sorted per cluster.
// Call once with fill = false (count must be zeroed; it returns the counts),
// scan the counts into occOff[ncluster + 1], zero count, call with fill = true
// (count holds ncluster + 1 ints; the fill's sort overwrites it).
void launchOccurrences(hipStream_t s, int ngroups, const int4* groups, const int4* groupInfo, const int* unionList,
                       int ncluster, int* count, int* occOff, int2* occ, bool fill);
// The pair kernel's window staging layout for a schedule: per group
// segInfo = {first float4 of its segments, first wavefront, wavefronts,
// window float4 (slotF4 per cluster slot)}; wavefront w of the group writes
// its window at x + (w - y) * w_size.  total receives the staging size in
// float4.
void launchSegments(hipStream_t s, int ngroups, const int4* groups, const int4* groupInfo, int numWaves,
                    const int2* waveStart, int slotF4, int* size, int* base, int* total, int4* segInfo);
// Per cluster, the staging positions of all its window copies (CSR:
// gposOff[ncluster + 1], gpos): count pass (fill = false) into count, scan,
// then fill.
void launchGatherList(hipStream_t s, int ncluster, const int* occOff, const int2* occ, const int4* segInfo,
                      int slotF4, int* count, const int* gposOff, int* gpos, bool fill);
// Per chunk, the shift vector of its i-entry (shiftVec[chunk shift index]),
// so that the pair kernel loads it together with the chunk descriptor.
void launchChunkShifts(hipStream_t s, int nchunk, const int4* chunks, const float4* shiftVec, float4* out);
// Static schedule of the pair kernel: the chunk list (in group order) cut into
// numWaves contiguous runs of about equal cost (the exclusive scan costOff of
// launchChunkCosts' per-chunk costs, total *costTotal, when costOff is given;
// else work entries plus chunkWeight per chunk).  With genWaves > 0 and
// numWaves = 2 * genWaves (two workgroups per CU, genWaves wavefronts per
// resident generation), the first generation takes a larger share.
// out[w] = {first chunk, group of that chunk} for w in [0, numWaves],
// out[numWaves] = {nchunk, ngroups}; runs are non-empty except for the
// trailing wavefronts when there are more wavefronts than distinct starts.
void launchWaveSchedule(hipStream_t s, int numWaves, int nchunk, const int4* chunks, int nwork, int chunkWeight,
                        int ngroups, const int4* groups, int2* out, const int* costOff = nullptr,
                        const int* costTotal = nullptr, int genWaves = 0);
// Per-chunk pair-kernel cost (1/32 tile units) from its tiles and the LJ
// i-atoms of its cluster: type per cluster slot, typeLJ per type.
void launchChunkCosts(hipStream_t s, int nchunk, const int4* chunks, const int* type, const int* typeLJ, int* cost);
// Energy-slot range of each system of a batch: slotOff[k] for k in [0, R]
// (the kernel's slot g + w of the first chunk of the system's first group).
void launchSystemSlots(hipStream_t s, int numSystems, const int* groupOff, int ngroups, const int4* groups,
                       int numWaves, const int2* waveStart, int* slotOff);

// -- kernel.hip
// the self energy and excluded-pair Coulomb energies in double
struct ExclusionArgs
{
    int              natoms;
    const int*       cells;
    const float4*    xq;
    const int*       exclOff;
    const int*       exclIdx;
    float            box[3];
    int              elec;
    InteractionConst ic;
    double           rc2;      // the float rcoulomb^2, as double
    double           epsfacD;  // the float epsfac, as double
    double           selfCoef; // epsfac V(0) / 2
    const double*    tabD;     // FDV0 table in double (table flavour)
    double*          partial;  // one per block
    // batches: the atom's system (boxes[3 * system]) and per-atom energies
    // instead of block partials; nullptr for a single system
    const int*       atomSystem;
    const float*     boxes;
    double*          atomEnergy;
    // single system, pair form: nslot cluster slots (self terms) and npairs
    // excluded pairs as cluster slots
    int              nslot;
    int              npairs;
    const int2*      pairSlots;
};

struct KernelArgs
{
    const int4*   chunks;
    const int4*   groups;
    const int4*   groupInfo;
    int           ngroups;
    int           groupSize;
    int           ncluster;
    int           windowSlots; // LDS window capacity in clusters (G + max union)
    const int*    unionList;
    const uint2*  work;
    const float4* xq;
    const int*    type;
    const float2* nbfp;
    const int*    typeLJ; // per type: 1 if it has LJ with any type
    int           skipLJ; // host: some type has no LJ (launch the variant that skips it)
    int           vdwModifier; // host: NBX_VDW_* modifier
    int           ljPme;       // host: LJ-PME (c6grid is read)
    int           twinCut;     // host: rvdw < rcoulomb (launch the twin-cut-off variant)
    int           windowSpill; // some union clusters lie beyond their window (fixed-point atomics)
    const float*  c6grid;      // LJ-PME: the grid's 6 C6 per type pair
    int           ntype;  // including the padding type
    const float4* tabFDV0;
    const float4* chunkShift; // per chunk: its i-entry's shift vector
    InteractionConst ic;
    const int2* waveStart; // [4 * gridDim.x + 1]: each wavefront's first chunk and its group
    const int4* segInfo;   // per group: window staging layout (launchSegments)
    float4*     staging;   // the wavefronts' j-force windows, gathered after the kernel
    // union clusters beyond their group's window: added to fnb with float
    // atomics, or (fixF given) to 3 fixed-point sums (kFixBits fraction
    // bits) per atom slot with integer atomics, independent of their order
    float4*             fnb;
    unsigned long long* fixF;
    // zeroed before the launch: with waveEnergy, [4 * gridDim.x][2], the
    // energies of wavefront w (its groups in order) in slot w; otherwise
    // [ngroups + 4 * gridDim.x][2], those of wavefront w in group g in slot
    // g + w (per-system sums of a batch)
    double*  energyPartial;
    int      waveEnergy;
    float*   fshiftPartial; // [gridDim.x - exclBlocks][45*3]
    // with energies: the last exclBlocks workgroups of the launch run the
    // single-system exclusion pass instead (partials excl.partial[0,
    // exclBlocks)), in the tail the pair workgroups leave; 0 for none
    int           exclBlocks;
    ExclusionArgs excl;
};
// Lets kernel fn use lds bytes of dynamic LDS (above the 64 KiB default) on
// the current device.  hipFuncSetAttribute is a host call that can hold up
// the stream, so it runs once per kernel and device (and again only for a
// larger size).
hipError_t allowDynamicLds(const void* fn, int lds);
// start/stop (either may be NULL) record the launch's own start and end
hipError_t launchNonbonded(hipStream_t s, const KernelArgs& a, int elec, bool energy, bool shiftForces, int numBlocks,
                           hipEvent_t start = nullptr, hipEvent_t stop = nullptr);
int  nonbondedLdsBytes(int ntype, int tabN, int elec, int windowSlots, bool ljPme = false);

// -- adjoint.hip
struct AdjointArgs
{
    const int4*   chunks;
    const int4*   groups;
    const int4*   groupInfo;
    int           ncluster;
    int           windowSlots;
    const int*    unionList;
    const uint2*  work;
    const float4* xq;
    const float4* gq; // dL/dF per cluster-layout atom (w unused)
    const int*    type;
    const float2* nbfp;
    int           ntype;
    const float4* tabFDV0;
    const float4* chunkShift;
    InteractionConst ic;
    const int2* waveStart; // the forward kernel's static schedule
    const int4* segInfo;   // the adjoint's window staging layout (4 float4 per cluster slot)
    float4*     staging;
    const int*  gposOff;   // per cluster: the staging positions of its window copies
    const int*  gpos;
    float4*     gradX;     // dL/dx per cluster-layout atom, zeroed by launchAdjointPrepare
    double*     gradQ;     // dL/dq per cluster-layout atom, zeroed by launchAdjointPrepare
    double*     paramPartial; // [adjoint workgroups][ntype^2][2]
};
// the adjoint's type-pair accumulators live in LDS: at most this many types
// (including the padding type)
constexpr int kAdjointMaxTypes = 32;
int  adjointLdsBytes(int ntype, int tabN, int elec, int windowSlots);
int  adjointBlocks(int forwardBlocks);
// launches the adjoint kernel on its layout's schedule, then writes dL/dx
// (natoms x 3 floats) and dL/dq (natoms doubles) in atom order (either may
// be NULL) and dL/dC6, dL/dC12 as numTypes x numTypes matrices (the padding
// type dropped; either may be NULL, both NULL skip the C6/C12 sums)
void launchAdjoint(hipStream_t s, const AdjointArgs& a, int elec, int forwardBlocks, const int* atomIndices,
                   float* gradXOut, double* gradQOut, int numTypes, double* gradC6Out, double* gradC12Out);
// coordinates (x may be NULL) and g into the cluster layout, spill
// accumulators cleared: xq, gq hold nslot + 4 entries (the padding cluster)
void launchAdjointPrepare(hipStream_t s, int nslot, const int* atomIndices, const float* x, const float* g, float4* xq,
                          float4* gq, float4* gradX, double* gradQ);

// -- buffers.hip
// zeroes n float4 and, when d is not NULL, nd doubles
void launchClearF4(hipStream_t s, float4* p, int n, double* d, int nd);
void launchConvertX(hipStream_t s, int nslot, const int* atomIndices, const float* x, float4* xq);
// after the pair kernel: the staged windows added to the cluster force
// buffer, or (f given) written with it to the atom-order forces f (added to
// them with addToF); with f
// and virPart, also the -1/2 x (x) f virial term from the cluster-layout
// coordinates xq, 9 doubles per workgroup.  Returns the workgroup count
// (at most kGatherMaxBlocks).
// 2048 workgroups of four wavefronts fill the 256 CUs once at full
// occupancy (measured: 1024, 4096 and 8192 slower on the 288k water step)
constexpr int kGatherMaxBlocks = 2048;
// MI355X: 8 XCDs of 32 CUs, each with its own L2; the gathers map
// consecutive clusters to one XCD's workgroups
constexpr int kGatherXcds = 8;
static_assert(kGatherMaxBlocks % kGatherXcds == 0, "gather grid a multiple of the XCDs");
int launchGatherForces(hipStream_t s, int ncluster, const int* gposOff, const int* gpos, const float4* staging,
                       float4* fnb, const int* atomIndices, float* f, bool addToF, const float4* xq,
                       double* virPart, const unsigned long long* fixF = nullptr);
// coordinates into the cluster layout (x may be NULL), cluster forces and
// nd doubles at d cleared, in one launch
void launchPrepareStep(hipStream_t s, int nslot, const int* atomIndices, const float* x, float4* xq, float4* fnb,
                       double* d, int nd, unsigned long long* fixF = nullptr);
// charges (atom order) into the w component of the cluster layout
void launchSetCharges(hipStream_t s, int nslot, const int* atomIndices, const float* q, float4* xq);
void launchReduceF(hipStream_t s, int natoms, const int* cells, const float4* fnb, float* f);
// energies, shift forces and (virialOut given) the virial from their
// partial sums, in fixed order
void launchFinalize(hipStream_t s, int numEnergyPartials, const double* energyPartial, int numExclPartials,
                    const double* exclPartial, int numShiftPartials, const float* fshiftPartial, bool shiftForces,
                    double* energyOut, double* fshiftOut, int numVirPartials, const double* virPartial,
                    const float4* shiftVec, double* virialOut);
constexpr int kExclMaxBlocks = 1024;
// per-atom form (batches: per-atom energies or block partials)
int launchExclusionEnergy(hipStream_t s, const ExclusionArgs& a);
// the excluded pairs (atom indices, each pair once) as cluster slots
void launchExclusionSlots(hipStream_t s, int npairs, const int2* pairs, const int* cells, int2* pairSlots);
void launchBatchFinalize(hipStream_t s, int numSystems, const int* slotOff, const double* energyPartial,
                         const int* atomOff, const double* atomEnergy, double* energyOut);
void launchExpandPairs(hipStream_t s, const int4* ientry, int nci, const uint2* jentry, const float4* xq,
                       const int* atomIndices, const float4* shiftVec, float r2, unsigned long long* counter,
                       long long* out, long long cap);

} // namespace nbx
