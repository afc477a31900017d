// C ABI of libnbx: context management and the host-side sequencing of the
// grid, search and kernel launches.  See include/nbx.h for the contract.
#include <algorithm>
#include <climits>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "context.h"

using namespace nbx;

namespace
{

thread_local std::string g_lastError;

} // namespace

std::string& nbx::lastError()
{
    return g_lastError;
}

namespace
{

// ---------------------------------------------------------------------------
// Interaction constants, restated from the reference in single precision
// ---------------------------------------------------------------------------

// calc_ewaldcoeff_q (ewald/ewald_utils.cpp:46-74) in float
float ewaldCoeff(float rc, float rtol)
{
    float beta = 5, low, high;
    int   n, i = 0;
    do
    {
        i++;
        beta *= 2;
    } while (std::erfc(beta * rc) > rtol);
    n    = i + 60;
    low  = 0;
    high = beta;
    for (i = 0; i < n; i++)
    {
        beta = (low + high) / 2;
        if (std::erfc(beta * rc) > rtol)
        {
            low = beta;
        }
        else
        {
            high = beta;
        }
    }
    return beta;
}

// calc_ewaldcoeff_lj (ewald/ewald_utils.cpp:76-115): the same doubling and
// bisection as ewaldCoeff on g(x) = exp(-x^2) (1 + x^2 + x^4 / 2), the
// dispersion's share beyond x = beta rc
float ljGridTail(float beta, float rc)
{
    const float x2 = (beta * rc) * (beta * rc);
    return std::exp(-x2) * (1 + x2 + x2 * x2 / 2.0f);
}

float ewaldCoeffLJ(float rc, float rtol)
{
    float beta = 5, low, high;
    int   n, i = 0;
    do
    {
        i++;
        beta *= 2;
    } while (ljGridTail(beta, rc) > rtol);
    n    = i + 60;
    low  = 0;
    high = beta;
    for (i = 0; i < n; i++)
    {
        beta = (low + high) / 2;
        if (ljGridTail(beta, rc) > rtol)
        {
            low = beta;
        }
        else
        {
            high = beta;
        }
    }
    return beta;
}

// ewald_spline3_table_scale, Coulomb part (tables/forcetable.cpp:304-358)
float ewaldTableScale(float beta, float rc)
{
    const double erf_x_d3  = 1.0522;
    const double etol      = 0.1 * std::erfc(beta * rc);
    const double deriv_tol = FLT_EPSILON;
    const double x_scale   = beta;
    const double sc_deriv  = std::sqrt(erf_x_d3 / (6 * 4 * deriv_tol * x_scale)) * x_scale;
    const double func_tol  = std::max(etol, static_cast<double>(FLT_EPSILON));
    const double sc_func   = std::cbrt(erf_x_d3 / (6 * 12 * std::sqrt(3.0) * func_tol)) * x_scale;
    const float  sc_q      = static_cast<float>(std::max(sc_deriv, sc_func));
    return std::max(0.0f, sc_q);
}

double vEwaldLr(double beta, double r)
{
    return r == 0 ? beta * 2 / std::sqrt(M_PI) : std::erf(beta * r) / r;
}

// generateEwaldCorrectionTables (tables/forcetable.cpp:144-298): quadratic
// spline of erf(beta r)/r; returns FDV0 quadruplets {F_i, F_i+1 - F_i, V_i, 0}
// stored in Real (float: the mixed-precision build's table; double: the
// double build's, used for the excluded-pair energies).
template<typename Real>
std::vector<Real> ewaldTableFDV0(int numPoints, double tableScaling, float beta)
{
    std::vector<Real> tf(numPoints, Real(0)), tv(numPoints, Real(0));
    const double dx       = 1 / tableScaling;
    const Real   tabMax   = static_cast<Real>(FLT_MAX * 0.0001);
    bool         outRange = false;
    int          iIn      = numPoints;
    double       vIn = 0, dc = 0;
    for (int i = numPoints - 1; i >= 0; i--)
    {
        const double xr0 = i * dx;
        const double vr0 = vEwaldLr(beta, xr0);
        double       vi;
        if (!outRange)
        {
            iIn = i;
            vIn = vr0;
            vi  = vr0;
        }
        else
        {
            vi = vIn - dc * (i - iIn) * dx;
        }
        tv[i] = static_cast<Real>(vi);
        if (i == 0)
        {
            continue;
        }
        const double vr1 = vEwaldLr(beta, (i - 1) * dx);
        if (vr1 != vr1 || vr1 < -tabMax || vr1 > tabMax)
        {
            outRange = true;
        }
        if (!outRange)
        {
            const double a2dx = (vr0 + vr1 - 2 * vEwaldLr(beta, xr0 - 0.5 * dx)) / (0.25 * dx);
            dc                = (vr0 - vr1) / dx + 0.5 * a2dx;
        }
        if (i == numPoints - 1)
        {
            tf[i] = static_cast<Real>(-dc);
        }
        else
        {
            tf[i] += -0.5 * dc;
        }
        if (!outRange)
        {
            const double a1    = dc;
            const double a2dx  = (a1 * dx + vr1 - vr0) * 2 / dx;
            const double dcNew = a1 - a2dx;
            if (dcNew != dcNew || dcNew < -tabMax || dcNew > tabMax)
            {
                outRange = true;
            }
            else
            {
                dc = dcNew;
            }
        }
        tf[i - 1] = static_cast<Real>(-0.5 * dc);
    }
    tf[0] *= 2;
    std::vector<Real> fdv0(4 * numPoints);
    for (int i = 0; i < numPoints - 1; i++)
    {
        fdv0[4 * i]     = tf[i];
        fdv0[4 * i + 1] = tf[i + 1] - tf[i];
        fdv0[4 * i + 2] = tv[i];
        fdv0[4 * i + 3] = 0;
    }
    const int last       = numPoints - 1;
    fdv0[4 * last]     = tf[last];
    fdv0[4 * last + 1] = -tf[last];
    fdv0[4 * last + 2] = tv[last];
    fdv0[4 * last + 3] = 0;
    return fdv0;
}

double one4PiEps0()
{
    const double eCharge  = 1.602176634e-19;
    const double avogadro = 6.02214076e23;
    const double eps0SI   = 8.8541878128e-12;
    const double eps0     = (eps0SI * 1e-9 * 1e3) / (eCharge * eCharge * avogadro);
    return 1.0 / (4.0 * M_PI * eps0);
}

} // namespace


namespace
{

void setDevice(nbx_context* c)
{
    check(hipSetDevice(c->device), "hipSetDevice");
}

// interaction_const_t's LJ modifier constants (forcerec.cpp:780-818 and
// 858-873): for the force switch, per power p (6: dispersion, 12:
// repulsion) force/p = r^-(p+1) + c2 rsw^2 + c3 rsw^3 between rsw and rc,
// with the potential shifted to 0 at rc (cpot); for the potential switch
// sw = 1 + c3 rsw^3 + c4 rsw^4 + c5 rsw^5, which has sw(0) = 1, sw(rc - rsw)
// = 0 and zero first and second derivatives at both ends.  Computed in
// double from the float radii, rounded once.  The kernel's forms: the
// force-switch force term rsw^2 r (c12' (c2r + c3r rsw) - c6' (c2d + c3d
// rsw)) and 12 times its potential term rsw^3 (c6' (4 c2d + 3 c3d rsw) -
// c12' (4 c2r + 3 c3r rsw)), with c6' = 6 C6 and c12' = 12 C12 as stored.
void setVdwModifierConstants(InteractionConst& ic, int modifier, float rsw, float rc)
{
    ic.rvdw_switch = rsw;
    ic.fsw_c2d = ic.fsw_c3d = ic.fsw_c2r = ic.fsw_c3r = 0.0f;
    ic.fsw_e2d = ic.fsw_e3d = ic.fsw_e2r = ic.fsw_e3r = 0.0f;
    ic.sw_c3 = ic.sw_c4 = ic.sw_c5 = ic.sw_f3 = ic.sw_f4 = ic.sw_f5 = 0.0f;
    const double d = static_cast<double>(rc) - static_cast<double>(rsw);
    if (modifier == NBX_VDW_FORCE_SWITCH)
    {
        const double rcd = rc, rswd = rsw;
        auto c2 = [&](double p) { return ((p + 1) * rswd - (p + 4) * rcd) / (std::pow(rcd, p + 2) * d * d); };
        auto c3 = [&](double p) { return -((p + 1) * rswd - (p + 3) * rcd) / (std::pow(rcd, p + 2) * d * d * d); };
        auto cpot = [&](double p) {
            return -std::pow(rcd, -p) + p * c2(p) / 3 * d * d * d + p * c3(p) / 4 * d * d * d * d;
        };
        ic.fsw_c2d   = static_cast<float>(c2(6));
        ic.fsw_c3d   = static_cast<float>(c3(6));
        ic.fsw_c2r   = static_cast<float>(c2(12));
        ic.fsw_c3r   = static_cast<float>(c3(12));
        ic.fsw_e2d   = static_cast<float>(4 * c2(6));
        ic.fsw_e3d   = static_cast<float>(3 * c3(6));
        ic.fsw_e2r   = static_cast<float>(4 * c2(12));
        ic.fsw_e3r   = static_cast<float>(3 * c3(12));
        ic.cpot_disp = static_cast<float>(cpot(6));
        ic.cpot_rep  = static_cast<float>(cpot(12));
    }
    else if (modifier == NBX_VDW_POTENTIAL_SWITCH)
    {
        ic.sw_c3     = static_cast<float>(-10 / (d * d * d));
        ic.sw_c4     = static_cast<float>(15 / (d * d * d * d));
        ic.sw_c5     = static_cast<float>(-6 / (d * d * d * d * d));
        ic.sw_f3     = static_cast<float>(3 * (-10 / (d * d * d)));
        ic.sw_f4     = static_cast<float>(4 * (15 / (d * d * d * d)));
        ic.sw_f5     = static_cast<float>(5 * (-6 / (d * d * d * d * d)));
        ic.cpot_disp = 0.0f;
        ic.cpot_rep  = 0.0f;
    }
}

void requireGrid(const nbx_context* c)
{
    if (!c->gridValid)
    {
        throw StateError("no grid: call nbx_put_on_grid first");
    }
}

void requireList(const nbx_context* c)
{
    requireGrid(c);
    if (!c->listValid)
    {
        throw StateError("no pair list: call nbx_construct_pairlist after nbx_put_on_grid");
    }
}

const float* deviceCoordinates(nbx_context* c, const float* x, int onDevice)
{
    if (x == nullptr)
    {
        throw InputError("coordinates must not be NULL");
    }
    if (onDevice)
    {
        return x;
    }
    c->x.ensure(3 * static_cast<size_t>(c->natoms));
    check(hipMemcpyAsync(c->x.p, x, sizeof(float) * 3 * c->natoms, hipMemcpyHostToDevice, c->stream), "upload x");
    return c->x.p;
}

// The list build's sizing reads: kStatsHead ints of scan totals, then the
// per-workgroup maxima of launchPartialMax, read back in one copy.
constexpr int kStatsHead = 4;
static_assert(kStatsHead + kPartialMaxBlocks * partialMaxStride(2, true) <= kPinnedInts, "stats fit the pinned buffer");
static void ensureStats(nbx_context* c)
{
    c->statsPartial.ensure(kStatsHead + kPartialMaxBlocks * partialMaxStride(2, true));
}

// Groups of consecutive i-clusters for the kernel's LDS j-force window
// (groups.hip): the largest group size whose j-entries fit the LDS sort and
// whose window fits the kernel's LDS budget (the pair kernel's, or with
// adjoint the adjoint kernel's, whose window holds 4 floats per atom).
void buildWorkGroups(nbx_context* c, WorkLayout& L, bool adjoint, const int* spansKnown = nullptr)
{
    auto ldsBytes = [&](int slots) {
        return adjoint ? adjointLdsBytes(c->ntype, c->ic.tab_n, c->elec, slots)
                       : nonbondedLdsBytes(c->ntype, c->ic.tab_n, c->elec, slots, c->opt.vdw_type == NBX_VDW_PME);
    };
    const int nc = c->ncluster;
    int spans[kGroupSizeCount] = {};
    if (spansKnown)
    {
        std::copy(spansKnown, spansKnown + kGroupSizeCount, spans);
    }
    else
    {
        ensureStats(c);
        const PartialMaxArgs pm{ { nullptr, nullptr }, 0, c->offJ.p, nc, c->statsPartial.p + kStatsHead };
        const int            nb = launchPartialMax(c->stream, pm);
        int* const           hp = pinnedInts();
        check(hipMemcpyAsync(hp, c->statsPartial.p + kStatsHead, sizeof(int) * nb * kGroupSizeCount,
                             hipMemcpyDeviceToHost, c->stream),
              "read spans");
        check(hipStreamSynchronize(c->stream), "group sync");
        reducePartialMax(hp, nb, kGroupSizeCount, spans);
    }
    c->jentryTmp.ensure(c->ncj);
    // the largest group size (or the chosen one) whose j-force windows, one
    // per wavefront, fit the workgroup's LDS budget: larger groups share more
    // j-clusters per flush, which pays more than a third resident workgroup.
    // When no size fits with at most 2% of the union entries spilled to
    // global atomics (long cut-offs, many types), the last pass takes the
    // smallest (or the chosen) size with whatever spill its window leaves.
    const int forced = adjoint ? 0 : c->opt.kernel_group_size;
    // NBX_UNION_SORT=1 takes the group pass's sort path for every group
    // (otherwise only for groups wider than its bitmap): for tests
    const char* sortEnv     = std::getenv("NBX_UNION_SORT");
    const bool  unionBySort = sortEnv != nullptr && sortEnv[0] == '1';
    // a layout rebuilt for a new list starts from the size its last build
    // chose: the sizes above it did not fit then and, for a system that
    // moved by a few steps, do not fit now (saves a count pass per search)
    int first = 0;
    while (forced <= 0 && L.groupSize > 0 && first < kGroupSizeCount - 1 && (kGroupSizeMax >> first) > L.groupSize)
    {
        first++;
    }
    for (int k = first; k <= kGroupSizeCount; k++)
    {
        const bool fallback = (k == kGroupSizeCount);
        const int  G        = fallback ? (forced > 0 ? forced : 1) : kGroupSizeMax >> k;
        int        si       = 0; // spans[] index of G
        while ((kGroupSizeMax >> si) > G)
        {
            si++;
        }
        const int  budget   = adjoint ? adjointLayoutBudget(c->ntype) : kKernelLdsBudget;
        if (!fallback && forced > 0 && G != forced)
        {
            continue;
        }
        if (spans[si] > kGroupMaxJ && G > 1)
        {
            continue;
        }
        if (spans[si] > kGroupMaxJ)
        {
            throw InputError("a grid cluster has more than " + std::to_string(kGroupMaxJ)
                             + " cluster-pair entries: the pair-list cut-off is too long for this kernel");
        }
        const int ng = (nc + G - 1) / G;
        L.unionCount.ensure(ng + 1);
        L.chunkCount.ensure(ng + 1);
        L.workCount.ensure(ng + 1);
        L.overflow.ensure(ng + 1);
        L.unionOff.ensure(scanOutSize(ng));
        L.chunkOff.ensure(scanOutSize(ng));
        L.workOff.ensure(scanOutSize(ng));
        L.groups.ensure(ng);
        L.groupInfo.ensure(ng);
        // one pass writes the layout at upper-bound bases (no count pass);
        // the scans of its per-group counts give the compact offsets
        L.unionTmp.ensure(std::max(c->ncj, 1));
        L.chunksTmp.ensure(std::max(c->nci, 1));
        L.workTmp.ensure(std::max(c->ncj + (kTileJ - 1) * c->nci, 1));
        GroupArgs ga{};
        ga.ncluster     = nc;
        ga.groupSize    = G;
        ga.ngroups      = ng;
        ga.chunkEntries = kChunkEntries;
        ga.offI         = c->offI.p;
        ga.offJ         = c->offJ.p;
        ga.ientry       = c->ientry.p;
        ga.jentryIn     = c->jentry.p;
        ga.jentryOut    = c->jentryTmp.p;
        ga.unionCount   = L.unionCount.p;
        ga.chunkCount   = L.chunkCount.p;
        ga.workCount    = L.workCount.p;
        ga.overflow     = L.overflow.p;
        ga.unionOff     = L.unionOff.p;
        ga.chunkOff     = L.chunkOff.p;
        ga.workOff      = L.workOff.p;
        ga.unionList    = L.unionTmp.p;
        ga.work         = L.workTmp.p;
        ga.chunks       = L.chunksTmp.p;
        ga.groups       = L.groups.p;
        ga.groupInfo    = L.groupInfo.p;
        ga.windowCap    = 1 << 30; // uncapped slots; the compaction applies the window capacity
        ga.upperBound   = 1;
        ga.unionBySort  = unionBySort ? 1 : 0;
        launchBuildGroups(c->stream, ga, true);
        // the largest union and the overflow flag, as per-workgroup maxima
        // after the three scan totals (one read)
        ensureStats(c);
        const PartialMaxArgs pm{ { L.unionCount.p, L.overflow.p }, 2, nullptr, ng, c->statsPartial.p + kStatsHead };
        const int            statBlocks = launchPartialMax(c->stream, pm);
        {
            const int* in[3]  = { L.unionCount.p, L.chunkCount.p, L.workCount.p };
            int*       out[3] = { L.unionOff.p, L.chunkOff.p, L.workOff.p };
            int*       tot[3] = { c->statsPartial.p, c->statsPartial.p + 1, c->statsPartial.p + 2 };
            launchExclusiveScansWide(c->stream, ng, 3, in, out, tot);
        }
        int* const hp = pinnedInts();
        check(hipMemcpyAsync(hp, c->statsPartial.p, sizeof(int) * (kStatsHead + statBlocks * 2), hipMemcpyDeviceToHost,
                             c->stream),
              "read stats");
        check(hipStreamSynchronize(c->stream), "group sync");
        int mx[2];
        reducePartialMax(hp + kStatsHead, statBlocks, 2, mx);
        const int maxUnion = mx[0], overflow = mx[1], unionTotal = hp[0], chunkTotal = hp[1], workTotal = hp[2];
        if (overflow)
        {
            continue;
        }
        // window capacity: own clusters, the union, and the padding entries'
        // spare slot; when the largest union does not fit the budget, the
        // window holds the first union clusters of each group and the rest
        // take direct global atomics, which is accepted while that is at most
        // 2% of the union entries (any share in the last pass)
        int slots = G + maxUnion + 1;
        L.spill   = false;
        if (ldsBytes(slots) > budget)
        {
            while (slots > G + 2 && ldsBytes(slots) > budget)
            {
                slots--;
            }
            if (ldsBytes(slots) > budget)
            {
                continue;
            }
            std::vector<int> counts(ng);
            check(hipMemcpy(counts.data(), L.unionCount.p, sizeof(int) * ng, hipMemcpyDeviceToHost), "read unions");
            long long spilled = 0;
            for (int u : counts)
            {
                spilled += std::max(0, u - (slots - 1 - G));
            }
            if (!fallback && spilled * 50 > unionTotal)
            {
                continue;
            }
            L.spill = spilled > 0;
        }
        // the compaction gives union clusters beyond the window the overflow slot
        ga.windowCap = slots;
        L.unionList.ensure(std::max(unionTotal, 1));
        L.chunks.ensure(std::max(chunkTotal, 1));
        L.work.ensure(std::max(workTotal, 1));
        ga.unionList = L.unionList.p;
        ga.work      = L.work.p;
        ga.chunks    = L.chunks.p;
        launchGroupCompact(c->stream, ga, L.unionTmp.p, L.workTmp.p, L.chunksTmp.p);
        L.chunkShift.ensure(std::max(chunkTotal, 1));
        launchChunkShifts(c->stream, chunkTotal, L.chunks.p, c->shiftVec.p, L.chunkShift.p);
        check(hipGetLastError(), "group kernels");
        c->jentry.swap(c->jentryTmp);
        buildOccurrences(c->stream, L.gather, ng, L.groups.p, L.groupInfo.p, L.unionList.p, nc, nc + unionTotal);
        L.groupSize   = G;
        L.ngroups     = ng;
        L.nchunk      = chunkTotal;
        L.nwork       = workTotal;
        L.windowSlots = slots;
        return;
    }
    throw InputError("the j-cluster window of one cluster does not fit in LDS: the pair-list cut-off is too long, or "
                     "there are too many atom types, for this kernel");
}
// The layout's grid and static schedule (the workgroups that can be
// resident, every wavefront a contiguous run of chunks of about equal cost)
// and its window staging.  The pair kernel runs kKernelWaves wavefronts per
// workgroup; the adjoint two (its numBlocks counts groups of four
// wavefronts, as the schedule does).
void scheduleLayout(nbx_context* c, WorkLayout& L, bool adjoint)
{
    int numWaves;
    if (adjoint)
    {
        const int lds   = adjointLdsBytes(c->ntype, c->ic.tab_n, c->elec, L.windowSlots);
        const int perCU = std::max(1, std::min(8, (160 * 1024) / lds)); // two-wavefront workgroups
        numWaves        = std::max(kKernelWaves, std::min((L.nchunk + kKernelWaves - 1) / kKernelWaves * kKernelWaves,
                                                          perCU * c->numCU * 2));
        numWaves        = (numWaves + kKernelWaves - 1) / kKernelWaves * kKernelWaves;
    }
    else
    {
        const int lds      = nonbondedLdsBytes(c->ntype, c->ic.tab_n, c->elec, L.windowSlots,
                                               c->opt.vdw_type == NBX_VDW_PME);
        const int maxPerCU = c->opt.kernel_blocks_per_cu > 0 ? c->opt.kernel_blocks_per_cu : 4;
        const int perCU    = std::max(1, std::min(maxPerCU, (160 * 1024) / lds));
        numWaves = std::max(1, std::min((L.nchunk + kKernelWaves - 1) / kKernelWaves, perCU * c->numCU)) * kKernelWaves;
    }
    L.numBlocks = numWaves / kKernelWaves;
    L.waveStart.ensure(numWaves + 1);
    // the forward kernel's runs by the measured cost model (unless the cost
    // total could overflow an int), with the first of two resident
    // generations of workgroups taking the larger share
    const long long costBound = static_cast<long long>(L.nwork / kTileJ + 1) * (kCostTile + 6)
                                + static_cast<long long>(kCostChunk) * L.nchunk;
    if (!adjoint && costBound < INT_MAX)
    {
        L.chunkCost.ensure(L.nchunk + 2);
        L.costOff.ensure(scanOutSize(L.nchunk));
        launchChunkCosts(c->stream, L.nchunk, L.chunks.p, c->nbtype.p, c->typeLJ.p, L.chunkCost.p);
        launchExclusiveScanWide(c->stream, L.nchunk, L.chunkCost.p, L.costOff.p, L.chunkCost.p + L.nchunk + 1);
        launchWaveSchedule(c->stream, numWaves, L.nchunk, L.chunks.p, L.nwork, kChunkWeight, L.ngroups, L.groups.p,
                           L.waveStart.p, L.costOff.p, L.chunkCost.p + L.nchunk + 1, c->numCU * kKernelWaves);
    }
    else
    {
        launchWaveSchedule(c->stream, numWaves, L.nchunk, L.chunks.p, L.nwork, kChunkWeight, L.ngroups, L.groups.p,
                           L.waveStart.p);
    }
    check(hipGetLastError(), "schedule kernel");
    buildSegments(c->stream, L.gather, L.ngroups, L.groups.p, L.groupInfo.p, numWaves, L.waveStart.p, c->ncluster,
                  adjoint ? 4 : 3);
}


} // namespace

extern "C" {

const char* nbx_last_error(void)
{
    return g_lastError.c_str();
}

void nbx_default_options(nbx_options* o)
{
    if (!o)
    {
        return;
    }
    o->coulomb_type          = NBX_COULOMB_EWALD;
    o->ewald_correction      = NBX_EWALD_TABLE;
    o->pairlist_cutoff       = 1.0;
    o->coulomb_cutoff        = 0.0;
    o->ewald_rtol            = 1e-5;
    o->epsilon_r             = 1.0;
    o->epsilon_rf            = 1.0;
    o->ewald_potential_shift = 0;
    o->device                = 0;
    o->kernel_group_size     = 0;
    o->kernel_blocks_per_cu  = 0;
    o->vdw_modifier          = NBX_VDW_POTENTIAL_SHIFT;
    o->rvdw_switch           = 0.0;
    o->vdw_type              = NBX_VDW_CUTOFF;
    o->ljpme_comb_rule       = NBX_LJPME_GEOMETRIC;
    o->ewald_rtol_lj         = 1e-3;
    o->vdw_cutoff            = 0.0;
}

int nbx_create(const nbx_options* options, nbx_context** context)
{
    if (!options || !context)
    {
        return fail(NBX_ERROR_INPUT, "nbx_create: NULL argument");
    }
    *context = nullptr;
    auto* c  = new nbx_context();
    int   rc = guarded([&] {
        c->opt = *options;
        const nbx_options& o = c->opt;
        if (!(o.pairlist_cutoff > 0) || !std::isfinite(o.pairlist_cutoff))
        {
            throw InputError("pairlist cut-off must be positive");
        }
        const double rcd = o.coulomb_cutoff > 0 ? o.coulomb_cutoff : o.pairlist_cutoff;
        if (rcd > o.pairlist_cutoff)
        {
            throw InputError("the interaction cut-off cannot exceed the pair-list cut-off");
        }
        if (o.coulomb_type < 0 || o.coulomb_type > 2)
        {
            throw InputError("Unsupported electrostatic interaction");
        }
        if (o.ewald_correction != NBX_EWALD_TABLE && o.ewald_correction != NBX_EWALD_ANALYTICAL)
        {
            throw InputError("unknown Ewald correction type");
        }
        if (o.kernel_group_size != 0 && o.kernel_group_size != 1 && o.kernel_group_size != 2
            && o.kernel_group_size != 4 && o.kernel_group_size != 8 && o.kernel_group_size != 16)
        {
            throw InputError("kernel_group_size must be 0, 1, 2, 4, 8 or 16");
        }
        if (o.kernel_blocks_per_cu < 0 || o.kernel_blocks_per_cu > 8)
        {
            throw InputError("kernel_blocks_per_cu must be in 0..8");
        }
        if (o.vdw_modifier != NBX_VDW_POTENTIAL_SHIFT && o.vdw_modifier != NBX_VDW_FORCE_SWITCH
            && o.vdw_modifier != NBX_VDW_POTENTIAL_SWITCH)
        {
            throw InputError("unknown LJ modifier");
        }
        if (!(o.vdw_cutoff >= 0) || o.vdw_cutoff > rcd)
        {
            throw InputError("the LJ cut-off must be in (0, rcoulomb] (0 means rcoulomb)");
        }
        const double rvdwd = o.vdw_cutoff > 0 ? o.vdw_cutoff : rcd;
        if (rvdwd < rcd && (o.coulomb_type != NBX_COULOMB_EWALD || o.vdw_type != NBX_VDW_CUTOFF))
        {
            // grompp's condition (readir.cpp:386-398)
            throw InputError("rvdw < rcoulomb is supported with Ewald electrostatics and LJ cut-off only");
        }
        if (o.vdw_modifier != NBX_VDW_POTENTIAL_SHIFT && !(o.rvdw_switch >= 0 && o.rvdw_switch < rvdwd))
        {
            throw InputError("the LJ switch radius must be in [0, rvdw)");
        }
        if (o.vdw_type != NBX_VDW_CUTOFF && o.vdw_type != NBX_VDW_PME)
        {
            throw InputError("unknown LJ type");
        }
        if (o.vdw_type == NBX_VDW_PME)
        {
            if (o.vdw_modifier != NBX_VDW_POTENTIAL_SHIFT)
            {
                throw InputError("LJ-PME takes the potential-shift modifier");
            }
            if (o.ljpme_comb_rule != NBX_LJPME_GEOMETRIC && o.ljpme_comb_rule != NBX_LJPME_LORENTZ_BERTHELOT)
            {
                throw InputError("unknown LJ-PME combination rule");
            }
            if (!(o.ewald_rtol_lj > 0 && o.ewald_rtol_lj < 1))
            {
                throw InputError("ewald_rtol_lj must be in (0, 1)");
            }
        }
        int ndev = 0;
        check(hipGetDeviceCount(&ndev), "hipGetDeviceCount");
        if (o.device < 0 || o.device >= ndev)
        {
            throw InputError("no such HIP device");
        }
        c->device = o.device;
        setDevice(c);
        hipDeviceProp_t prop;
        check(hipGetDeviceProperties(&prop, c->device), "hipGetDeviceProperties");
        c->numCU = prop.multiProcessorCount;
        check(hipStreamCreateWithFlags(&c->ownStream, hipStreamNonBlocking), "hipStreamCreate");
        c->stream = c->ownStream;

        // setupInteractionConst (gmxsetup.cpp:226-284), float arithmetic
        c->rlist    = static_cast<float>(o.pairlist_cutoff);
        c->rcoulomb = static_cast<float>(rcd);
        // interaction_const_t::rvdw: the LJ potential shift and switches
        // act at rvdw (forcerec.cpp:780-818, 848-873)
        const float rvdw = o.vdw_cutoff > 0 ? static_cast<float>(o.vdw_cutoff) : c->rcoulomb;
        c->twinCut       = rvdw < c->rcoulomb;
        const float p3   = rvdw * (rvdw * rvdw);
        const float p6   = p3 * p3;
        const float p12  = p6 * p6;
        InteractionConst& ic = c->ic;
        ic.rlist2    = c->rlist * c->rlist;
        ic.rcoulomb2 = c->rcoulomb * c->rcoulomb;
        ic.rvdw2     = rvdw * rvdw;
        ic.cpot_disp = static_cast<float>(-1.0 / p6);
        ic.cpot_rep  = static_cast<float>(-1.0 / p12);
        setVdwModifierConstants(ic, o.vdw_modifier, static_cast<float>(o.rvdw_switch), rvdw);
        if (o.vdw_type == NBX_VDW_PME)
        {
            // forcerec.cpp:703-717 and the kernel's lje_coeff6_6
            // (kernel_ref_outer.h:182-186), in float
            const float betaLJ = ewaldCoeffLJ(rvdw, static_cast<float>(o.ewald_rtol_lj));
            const float crc2   = (betaLJ * rvdw) * (betaLJ * rvdw);
            ic.lje_coeff2      = betaLJ * betaLJ;
            ic.lje_coeff6_6    = ic.lje_coeff2 * ic.lje_coeff2 * ic.lje_coeff2 / 6.0f;
            ic.sh_lj_ewald     = (std::exp(-crc2) * (1 + crc2 + 0.5f * crc2 * crc2) - 1) / p6;
        }
        ic.cpot_disp6 = static_cast<float>(static_cast<double>(ic.cpot_disp) / 6.0);
        ic.cpot_rep12 = static_cast<float>(static_cast<double>(ic.cpot_rep) / 12.0);
        const float epsr = static_cast<float>(o.epsilon_r);
        ic.epsfac        = epsr != 0 ? static_cast<float>(one4PiEps0() / epsr) : 0.0f;
        if (o.coulomb_type != NBX_COULOMB_CUTOFF)
        {
            // calc_rffac (mdlib/rf_util.cpp:54-66), which nblib calls for
            // every Coulomb type (gmxsetup.cpp:266-270); only the
            // reaction-field kernels read k_rf and c_rf
            const float epsrf = static_cast<float>(o.epsilon_rf), Rc = c->rcoulomb;
            ic.k_rf = (epsrf == 0) ? 1 / (2 * Rc * Rc * Rc) : (epsrf - epsr) / (2 * epsrf + epsr) / (Rc * Rc * Rc);
            ic.c_rf = 1 / Rc + ic.k_rf * Rc * Rc;
        }
        else
        {
            // plain cut-off with the potential shift (forcerec.cpp:902-915)
            ic.k_rf = 0;
            ic.c_rf = 1 / c->rcoulomb;
        }
        if (o.coulomb_type == NBX_COULOMB_EWALD)
        {
            c->elec      = o.ewald_correction == NBX_EWALD_ANALYTICAL ? kElecEwaldAna : kElecEwaldTable;
            ic.beta      = ewaldCoeff(c->rcoulomb, static_cast<float>(o.ewald_rtol));
            if (!(ic.beta > 0))
            {
                throw InputError("Ewald coefficient should be > 0");
            }
            ic.beta2 = ic.beta * ic.beta;
            ic.beta3 = ic.beta2 * ic.beta;
            if (c->elec == kElecEwaldAna && ic.beta * c->rcoulomb > 3.6f)
            {
                // the fitted range of the analytical correction
                throw InputError("the analytical Ewald correction covers beta * rc <= 3.6 (ewald_rtol >= 3e-7); "
                                 "use the table correction");
            }
            ic.sh_ewald  = o.ewald_potential_shift ? std::erfc(ic.beta * c->rcoulomb) / c->rcoulomb : 0.0f;
            ic.tab_scale = ewaldTableScale(ic.beta, c->rcoulomb);
            ic.tab_n     = static_cast<int>(c->rcoulomb * ic.tab_scale) + 2;
            std::vector<float>  fdv0  = ewaldTableFDV0<float>(ic.tab_n, ic.tab_scale, ic.beta);
            std::vector<double> fdv0d = ewaldTableFDV0<double>(ic.tab_n, ic.tab_scale, ic.beta);
            ic.self_v0   = fdv0[2];
            // the pair kernel reads V + sh_ewald from the unused fourth entry
            for (int i = 0; i < ic.tab_n; i++)
            {
                fdv0[4 * i + 3] = static_cast<float>(static_cast<double>(fdv0[4 * i + 2]) + ic.sh_ewald);
            }
            // the self energy sums one large term per atom against a pair sum
            // of nearly the same size: V(0) = 2 beta / sqrt(pi) in double (of
            // the float beta every pair term uses) keeps its rounding out
            c->selfCoef = 0.5 * static_cast<double>(ic.epsfac) * vEwaldLr(ic.beta, 0.0);
            c->tabFDV0.ensure(ic.tab_n);
            check(hipMemcpy(c->tabFDV0.p, fdv0.data(), sizeof(float) * fdv0.size(), hipMemcpyHostToDevice), "upload table");
            c->tabFDV0d.ensure(fdv0d.size());
            check(hipMemcpy(c->tabFDV0d.p, fdv0d.data(), sizeof(double) * fdv0d.size(), hipMemcpyHostToDevice),
                  "upload table");
        }
        else
        {
            c->elec      = kElecRF;
            // no correction table: nblib leaves its scale unset (zero)
            ic.tab_scale = 0.0f;
            ic.tab_n     = 0;
            ic.self_v0   = ic.c_rf;
            c->selfCoef  = 0.5 * static_cast<double>(ic.epsfac) * static_cast<double>(ic.c_rf);
        }
        c->shiftVec.ensure(kNumShifts);
    });
    if (rc != NBX_SUCCESS)
    {
        nbx_destroy(c);
        return rc;
    }
    *context = c;
    return NBX_SUCCESS;
}

int nbx_destroy(nbx_context* c)
{
    if (!c)
    {
        return NBX_SUCCESS;
    }
    (void)hipSetDevice(c->device);
    if (c->ownStream)
    {
        (void)hipStreamSynchronize(c->ownStream);
        (void)hipStreamDestroy(c->ownStream);
    }
    for (hipEvent_t e : c->timingEvents)
    {
        (void)hipEventDestroy(e);
    }
    delete c;
    return NBX_SUCCESS;
}

int nbx_set_kernel_timing(nbx_context* c, int maxLaunches)
{
    if (!c || maxLaunches < 0)
    {
        return fail(NBX_ERROR_INPUT, "nbx_set_kernel_timing: NULL context or negative count");
    }
    return guarded([&] {
        setDevice(c);
        check(hipStreamSynchronize(c->stream), "sync");
        for (hipEvent_t e : c->timingEvents)
        {
            check(hipEventDestroy(e), "hipEventDestroy");
        }
        c->timingEvents.assign(2 * static_cast<size_t>(maxLaunches), nullptr);
        for (hipEvent_t& e : c->timingEvents)
        {
            check(hipEventCreate(&e), "hipEventCreate");
        }
        c->timingRecorded = 0;
    });
}

int nbx_get_kernel_timing(nbx_context* c, double* averageMs, int* launches)
{
    if (!c || !averageMs || !launches)
    {
        return fail(NBX_ERROR_INPUT, "NULL argument");
    }
    return guarded([&] {
        setDevice(c);
        double total = 0.0;
        for (int k = 0; k < c->timingRecorded; k++)
        {
            check(hipEventSynchronize(c->timingEvents[2 * k + 1]), "hipEventSynchronize");
            float ms = 0.0f;
            check(hipEventElapsedTime(&ms, c->timingEvents[2 * k], c->timingEvents[2 * k + 1]), "hipEventElapsedTime");
            total += ms;
        }
        *launches  = c->timingRecorded;
        *averageMs = c->timingRecorded > 0 ? total / c->timingRecorded : 0.0;
        c->timingRecorded = 0;
    });
}

int nbx_set_stream(nbx_context* c, void* stream)
{
    if (!c)
    {
        return fail(NBX_ERROR_INPUT, "NULL context");
    }
    c->stream = static_cast<hipStream_t>(stream);
    return NBX_SUCCESS;
}

int nbx_set_nonbonded_parameters(nbx_context* c, int numTypes, const double* c6, const double* c12)
{
    if (!c)
    {
        return fail(NBX_ERROR_INPUT, "NULL context");
    }
    return guarded([&] {
        if (numTypes <= 0 || !c6 || !c12)
        {
            throw InputError("need at least one particle type and C6/C12 tables");
        }
        if (numTypes > kMaxTypes)
        {
            throw InputError("at most " + std::to_string(kMaxTypes) + " particle types are supported");
        }
        setDevice(c);
        const int t = numTypes + 1;
        std::vector<float2> nb(static_cast<size_t>(t) * t, make_float2(0.0f, 0.0f));
        for (int i = 0; i < numTypes; i++)
        {
            for (int j = 0; j < numTypes; j++)
            {
                const double a = c6[i * numTypes + j], b = c12[i * numTypes + j];
                if (!std::isfinite(a) || !std::isfinite(b))
                {
                    throw InputError("non-finite LJ parameter");
                }
                if (a != c6[j * numTypes + i] || b != c12[j * numTypes + i])
                {
                    throw InputError("the C6/C12 tables must be symmetric");
                }
                // nbfp stores 6*C6 and 12*C12 in real (gmxsetup.cpp:153-166)
                nb[i * t + j] = make_float2(static_cast<float>(a) * 6.0f, static_cast<float>(b) * 12.0f);
            }
        }
        // the stream's queued kernels may still read the old table
        check(hipStreamSynchronize(c->stream), "sync");
        c->nbfp.ensure(nb.size());
        check(hipMemcpy(c->nbfp.p, nb.data(), sizeof(float2) * nb.size(), hipMemcpyHostToDevice), "upload nbfp");
        // per type: whether it has LJ with any type (water hydrogens do not;
        // the pair kernel skips their LJ terms, which are exactly zero)
        std::vector<int> lj(t, 0);
        for (int i = 0; i < t; i++)
        {
            for (int j = 0; j < t; j++)
            {
                lj[i] |= (nb[i * t + j].x != 0.0f || nb[i * t + j].y != 0.0f) ? 1 : 0;
            }
        }
        c->typeLJ.ensure(t);
        check(hipMemcpy(c->typeLJ.p, lj.data(), sizeof(int) * t, hipMemcpyHostToDevice), "upload LJ flags");
        c->anyTypeWithoutLJ = std::count(lj.begin(), lj.end() - 1, 0) > 0;
        c->c6diag.assign(t, 0.0f);
        for (int i = 0; i < t; i++)
        {
            c->c6diag[i] = nb[i * t + i].x;
        }
        if (c->opt.vdw_type == NBX_VDW_PME)
        {
            // the grid's per-type parameters from the nbfp diagonal
            // (atomdata.cpp:288-330): geometric sqrt(6 C6_ii); LB
            // 0.5 (c12 / c6)^(1/6) and sqrt(c6^2 / c12), which combine to
            // 6 C6 of the LB sigma and epsilon; the padding type gets zeros
            std::vector<float> ljc(2 * t, 0.0f);
            for (int i = 0; i < t; i++)
            {
                const float c6i = nb[i * t + i].x, c12i = nb[i * t + i].y;
                if (c->opt.ljpme_comb_rule == NBX_LJPME_GEOMETRIC)
                {
                    ljc[2 * i] = std::sqrt(c6i);
                }
                else if (c6i > 0 && c12i > 0)
                {
                    ljc[2 * i]     = 0.5f * static_cast<float>(std::pow(static_cast<double>(c12i / c6i), 1.0 / 6.0));
                    ljc[2 * i + 1] = std::sqrt(c6i * c6i / c12i);
                }
            }
            std::vector<float> grid(static_cast<size_t>(t) * t, 0.0f);
            for (int i = 0; i < t; i++)
            {
                for (int j = 0; j < t; j++)
                {
                    if (c->opt.ljpme_comb_rule == NBX_LJPME_GEOMETRIC)
                    {
                        grid[i * t + j] = ljc[2 * i] * ljc[2 * j];
                    }
                    else
                    {
                        const float sg = ljc[2 * i] + ljc[2 * j], sg2 = sg * sg;
                        grid[i * t + j] = ljc[2 * i + 1] * ljc[2 * j + 1] * sg2 * sg2 * sg2;
                    }
                }
            }
            c->c6grid.ensure(grid.size());
            check(hipMemcpy(c->c6grid.p, grid.data(), sizeof(float) * grid.size(), hipMemcpyHostToDevice),
                  "upload LJ-PME grid C6");
        }
        // new values for the same types leave the grid and the pair list
        // valid (they depend on positions only); a new type count does not
        if (t != c->ntype)
        {
            c->gridValid = false;
            c->listValid = false;
        }
        c->ntypeUser = numTypes;
        c->ntype     = t;
    });
}

int nbx_set_topology(nbx_context* c, int numAtoms, const int* types, const float* charges, const int* exclOff,
                     const int* exclIdx)
{
    if (!c)
    {
        return fail(NBX_ERROR_INPUT, "NULL context");
    }
    return guarded([&] {
        if (c->ntypeUser < 0)
        {
            throw StateError("set the non-bonded parameters before the topology");
        }
        if (numAtoms < 0 || (numAtoms > 0 && (!types || !charges || !exclOff)))
        {
            throw InputError("invalid topology arrays");
        }
        for (int a = 0; a < numAtoms; a++)
        {
            if (types[a] < 0 || types[a] >= c->ntypeUser)
            {
                throw InputError("Missing nonbonded interaction parameters for an atom type");
            }
            if (!std::isfinite(charges[a]))
            {
                throw InputError("non-finite charge");
            }
        }
        c->typeCount.assign(c->ntypeUser, 0);
        for (int a = 0; a < numAtoms; a++)
        {
            c->typeCount[types[a]]++;
        }
        if (exclOff[0] != 0)
        {
            throw InputError("exclusion offsets must start at 0");
        }
        for (int a = 0; a < numAtoms; a++)
        {
            if (exclOff[a + 1] < exclOff[a])
            {
                throw InputError("exclusion offsets must be non-decreasing");
            }
        }
        const int nnz = numAtoms > 0 ? exclOff[numAtoms] : 0;
        for (int k = 0; k < nnz; k++)
        {
            if (exclIdx[k] < 0 || exclIdx[k] >= numAtoms)
            {
                throw InputError("exclusion index out of range");
            }
        }
        setDevice(c);
        c->natoms = numAtoms;
        c->type.ensure(numAtoms);
        c->charge.ensure(numAtoms);
        c->exclOff.ensure(numAtoms + 1);
        c->exclIdx.ensure(std::max(nnz, 1));
        if (numAtoms > 0)
        {
            check(hipMemcpy(c->type.p, types, sizeof(int) * numAtoms, hipMemcpyHostToDevice), "upload types");
            check(hipMemcpy(c->charge.p, charges, sizeof(float) * numAtoms, hipMemcpyHostToDevice), "upload q");
        }
        check(hipMemcpy(c->exclOff.p, exclOff, sizeof(int) * (numAtoms + 1), hipMemcpyHostToDevice), "upload excl");
        if (nnz > 0)
        {
            check(hipMemcpy(c->exclIdx.p, exclIdx, sizeof(int) * nnz, hipMemcpyHostToDevice), "upload excl");
        }
        // each excluded pair once, from the list of its lower-indexed atom
        std::vector<int2> pairs;
        for (int a = 0; a < numAtoms; a++)
        {
            for (int k = exclOff[a]; k < exclOff[a + 1]; k++)
            {
                if (exclIdx[k] > a)
                {
                    pairs.push_back(make_int2(a, exclIdx[k]));
                }
            }
        }
        c->numExclPairs = static_cast<int>(pairs.size());
        c->exclPairs.ensure(std::max<size_t>(pairs.size(), 1));
        c->exclPairSlots.ensure(std::max<size_t>(pairs.size(), 1));
        if (!pairs.empty())
        {
            check(hipMemcpy(c->exclPairs.p, pairs.data(), sizeof(int2) * pairs.size(), hipMemcpyHostToDevice),
                  "upload excluded pairs");
        }
        c->gridValid = false;
        c->listValid = false;
    });
}

int nbx_put_on_grid(nbx_context* c, const float* box, const float* x, int xOnDevice)
{
    if (!c)
    {
        return fail(NBX_ERROR_INPUT, "NULL context");
    }
    return guarded([&] {
        if (c->natoms < 0)
        {
            throw StateError("set the topology before gridding");
        }
        if (!box)
        {
            throw InputError("box must not be NULL");
        }
        for (int d = 0; d < 3; d++)
        {
            if (!std::isfinite(box[d]) || !(box[d] > 0))
            {
                throw InputError("box edges must be positive and finite");
            }
            if (c->rlist > 0.5f * box[d])
            {
                throw InputError("the pair-list cut-off must be at most half the box edge");
            }
        }
        setDevice(c);
        GridDims& g = c->g;
        const int n = c->natoms;
        g.natoms    = n;
        for (int d = 0; d < 3; d++)
        {
            g.box[d] = box[d];
        }
        // Grid::setDimensions for the simple 4x4 geometry (grid.cpp:103-262)
        // with the density nblib passes (gmxcalculator.cpp:97): N / det(box)
        const float det     = box[0] * (box[1] * box[2]);
        const float density = static_cast<float>(n) / det;
        if (n > kClusterSize)
        {
            const float tlen = std::cbrt(static_cast<float>(kClusterSize) / density);
            g.ncx            = std::max(1, static_cast<int>(box[0] / tlen));
            g.ncy            = std::max(1, static_cast<int>(box[1] / tlen));
        }
        else
        {
            g.ncx = 1;
            g.ncy = 1;
        }
        g.cellSize[0]    = box[0] / g.ncx;
        g.cellSize[1]    = box[1] / g.ncy;
        g.invCellSize[0] = 1 / g.cellSize[0];
        g.invCellSize[1] = 1 / g.cellSize[1];
        g.invBoxZ        = static_cast<float>(1.0 / static_cast<double>(box[2]));
        g.ncol           = g.ncx * g.ncy;
        const size_t maxCluster = static_cast<size_t>(n) / kClusterSize + g.ncol + 1;
        const size_t maxSlot    = maxCluster * kClusterSize;

        // shift vectors (calc_shifts, pbc.cpp:1187-1202)
        std::vector<float4> sv(kNumShifts);
        for (int n_ = 0, m = -1; m <= 1; m++)
        {
            for (int l = -1; l <= 1; l++)
            {
                for (int k = -2; k <= 2; k++, n_++)
                {
                    sv[n_] = make_float4(k * box[0], l * box[1], m * box[2], 0.0f);
                }
            }
        }
        check(hipMemcpyAsync(c->shiftVec.p, sv.data(), sizeof(float4) * kNumShifts, hipMemcpyHostToDevice, c->stream),
              "upload shifts");

        const float* xd = deviceCoordinates(c, x, xOnDevice);
        if (xd != c->x.p)
        {
            c->x.ensure(3 * static_cast<size_t>(n));
            check(hipMemcpyAsync(c->x.p, xd, sizeof(float) * 3 * n, hipMemcpyDeviceToDevice, c->stream), "copy x");
        }
        c->cxy_na.ensure(g.ncol + 1);
        c->cxy_ind.ensure(g.ncol + 1);
        c->atomCol.ensure(n);
        c->atomRank.ensure(n);
        c->cells.ensure(n);
        c->bucket.ensure(maxSlot);
        c->atomIndices.ensure(maxSlot);
        c->counter.ensure(4);
        check(hipMemsetAsync(c->cxy_na.p, 0, sizeof(int) * (g.ncol + 1), c->stream), "clear counts");
        launchGridColumns(c->stream, g, c->x.p, c->atomCol.p, c->atomRank.p, c->cxy_na.p);
        launchGridScan(c->stream, g.ncol, c->cxy_na.p, c->cxy_ind.p, c->counter.p);
        int* const hp = pinnedInts();
        check(hipMemcpyAsync(hp, c->counter.p, 2 * sizeof(int), hipMemcpyDeviceToHost, c->stream), "read count");
        check(hipMemcpyAsync(hp + 2, c->cxy_na.p + g.ncol, sizeof(int), hipMemcpyDeviceToHost, c->stream),
              "read flag");
        check(hipStreamSynchronize(c->stream), "grid sync");
        const int hostCounts[3] = { hp[0], hp[1], hp[2] };
        if (hostCounts[2] != 0)
        {
            c->gridValid = false;
            throw InputError("atoms must lie inside the box [0, L) for gridding");
        }
        if (hostCounts[1] > kMaxColumnAtoms)
        {
            c->gridValid = false;
            throw InputError("more than " + std::to_string(kMaxColumnAtoms)
                             + " atoms fall in one grid column: the system is far from uniform");
        }
        c->ncluster = hostCounts[0];
        const int nslot = c->ncluster * kClusterSize;
        launchGridSortColumns(c->stream, g, c->x.p, c->atomCol.p, c->atomRank.p, c->cxy_na.p, c->cxy_ind.p, c->bucket.p,
                              c->atomIndices.p, c->cells.p);
        // one extra cluster of far-away fillers after the real ones: the
        // kernel points lanes past the end of a j-list at it
        c->xq.ensure(nslot + kClusterSize);
        c->nbtype.ensure(nslot + kClusterSize);
        c->bb.ensure(2 * static_cast<size_t>(c->ncluster));
        c->fnb.ensure(nslot + kClusterSize);
        c->fixF.ensure(3 * static_cast<size_t>(nslot + kClusterSize));
        // the fill also writes the far-away filler cluster after the real ones
        launchGridFill(c->stream, c->ncluster, c->x.p, c->atomIndices.p, c->type.p, c->charge.p, c->ntype - 1, c->xq.p,
                       c->nbtype.p, c->bb.p);
        check(hipGetLastError(), "grid kernels");
        c->gridValid = true;
        c->listValid = false;
    });
}

// After a search or a prune: the pair kernel's work layout and schedule and
// the per-step output buffers sized for it.
static void finishList(nbx_context* c, const int* spans = nullptr)
{
    buildWorkGroups(c, c->lay, false, spans);
    scheduleLayout(c, c->lay, false);
    const int numWaves = c->lay.numBlocks * kKernelWaves;
    c->energyPartial.ensure(2 * static_cast<size_t>(c->lay.ngroups + numWaves));
    c->exclPartial.ensure(kExclMaxBlocks);
    c->fshiftPartial.ensure(static_cast<size_t>(c->lay.numBlocks) * kNumShifts * 3);
    c->energy.ensure(2);
    c->fshift.ensure(kNumShifts * 3);
    c->virialPartial.ensure(9 * static_cast<size_t>(kGatherMaxBlocks));
    c->virial.ensure(9);
    c->listValid  = true;
    c->adjValid   = false;
    c->layOuterValid = false;
    c->haveEnergy = c->haveShift = c->haveVirial = false;
}

int nbx_construct_pairlist(nbx_context* c)
{
    if (!c)
    {
        return fail(NBX_ERROR_INPUT, "NULL context");
    }
    return guarded([&] {
        requireGrid(c);
        setDevice(c);
        const int nc = c->ncluster;
        c->countJ.ensure(nc + 1);
        c->countI.ensure(nc + 1);
        c->offJ.ensure(scanOutSize(nc));
        c->offI.ensure(scanOutSize(nc));
        c->totals.ensure(3);
        ensureStats(c);
        SearchArgs a{};
        a.g           = c->g;
        a.rlist       = c->rlist;
        a.rlist2      = c->rlist * c->rlist;
        a.ncluster    = nc;
        a.cxy_ind     = c->cxy_ind.p;
        a.bb          = c->bb.p;
        a.xq          = c->xq.p;
        a.atomIndices = c->atomIndices.p;
        a.cells       = c->cells.p;
        a.exclOff     = c->exclOff.p;
        a.exclIdx     = c->exclIdx.p;
        a.countJ      = c->countJ.p;
        a.countI      = c->countI.p;
        a.offJ        = c->offJ.p;
        a.offI        = c->offI.p;
        // the j- and i-entry counts scanned together (two workgroups, one launch)
        auto scanCounts = [&] {
            const int* in[2]  = { c->countJ.p, c->countI.p };
            int*       out[2] = { c->offJ.p, c->offI.p };
            int*       tot[2] = { c->statsPartial.p, c->statsPartial.p + 1 };
            launchExclusiveScansWide(c->stream, nc, 2, in, out, tot);
        };
        // the list's totals, its largest j-entry count and the group pass's
        // largest j-entry span per group size (per-workgroup maxima), read
        // with one wait
        int  spans[kGroupSizeCount] = {};
        int  statBlocks             = 0;
        auto readStats              = [&] {
            constexpr int kStride = partialMaxStride(1, true);
            const PartialMaxArgs pm{ { c->countJ.p, nullptr }, 1, c->offJ.p, nc, c->statsPartial.p + kStatsHead };
            statBlocks = launchPartialMax(c->stream, pm);
            check(hipMemcpyAsync(pinnedInts(), c->statsPartial.p, sizeof(int) * (kStatsHead + statBlocks * kStride),
                                 hipMemcpyDeviceToHost, c->stream),
                  "read stats");
        };
        // after the wait: j- and i-entry totals and the largest j-entry count
        // in tot, the spans in spans
        auto takeReads = [&](int* tot) {
            const int* hp = pinnedInts();
            int        st[partialMaxStride(1, true)];
            reducePartialMax(hp + kStatsHead, statBlocks, partialMaxStride(1, true), st);
            tot[0] = hp[0];
            tot[1] = hp[1];
            tot[2] = st[0];
            std::copy(st + 1, st + 1 + kGroupSizeCount, spans);
        };
        // After a first search, the fill runs without a count pass: each
        // cluster writes into a region sized from the largest count seen
        // (with a margin) and a compaction follows the scans; a cluster that
        // outgrows it sends the search through the counted path again.
        bool done = false;
        if (c->searchCapJ > 0)
        {
            const int cap = c->searchCapJ;
            c->searchJTmp.ensure(static_cast<size_t>(nc) * cap);
            c->searchITmp.ensure(static_cast<size_t>(nc) * kMaxIEntriesPerCluster);
            a.capJ   = cap;
            a.jentry = c->searchJTmp.p;
            a.ientry = c->searchITmp.p;
            launchSearch(c->stream, a, true);
            scanCounts();
            readStats();
            check(hipStreamSynchronize(c->stream), "search sync");
            int tot[3] = { 0, 0, 0 };
            takeReads(tot);
            if (tot[2] <= cap)
            {
                c->ncj = tot[0];
                c->nci = tot[1];
                c->ientry.ensure(c->nci);
                c->jentry.ensure(c->ncj);
                launchSearchCompact(c->stream, nc, cap, c->countJ.p, c->countI.p, c->offJ.p, c->offI.p,
                                    c->searchJTmp.p, c->searchITmp.p, c->jentry.p, c->ientry.p);
                c->searchCapJ = tot[2] + tot[2] / 4 + 32;
                done          = true;
            }
        }
        if (!done)
        {
            a.capJ   = 0;
            a.ientry = nullptr;
            a.jentry = nullptr;
            launchSearch(c->stream, a, false);
            scanCounts();
            readStats();
            check(hipStreamSynchronize(c->stream), "search sync");
            int tot[3] = { 0, 0, 0 };
            takeReads(tot);
            c->ncj = tot[0];
            c->nci = tot[1];
            c->ientry.ensure(c->nci);
            c->jentry.ensure(c->ncj);
            a.ientry = c->ientry.p;
            a.jentry = c->jentry.p;
            launchSearch(c->stream, a, true);
            c->searchCapJ = tot[2] + tot[2] / 4 + 32;
        }
        check(hipGetLastError(), "search kernels");
        launchExclusionSlots(c->stream, c->numExclPairs, c->exclPairs.p, c->cells.p, c->exclPairSlots.p);
        c->haveOuter = false;
        finishList(c, spans);
    });
}

int nbx_prune_pairlist(nbx_context* c, double rlistInner)
{
    if (!c)
    {
        return fail(NBX_ERROR_INPUT, "NULL context");
    }
    return guarded([&] {
        requireList(c);
        if (!(rlistInner >= c->rcoulomb && rlistInner <= c->rlist))
        {
            throw InputError("the inner pair-list cut-off must lie between the interaction cut-off and rlist");
        }
        setDevice(c);
        const int nc = c->ncluster;
        if (!c->haveOuter)
        {
            // keep the list as searched: every prune starts from it
            c->nciOuter = c->nci;
            c->ncjOuter = c->ncj;
            c->offIOuter.ensure(nc + 1);
            c->ientryOuter.ensure(std::max(c->nci, 1));
            c->jentryOuter.ensure(std::max(c->ncj, 1));
            check(hipMemcpyAsync(c->offIOuter.p, c->offI.p, sizeof(int) * (nc + 1), hipMemcpyDeviceToDevice,
                                 c->stream),
                  "keep outer list");
            check(hipMemcpyAsync(c->ientryOuter.p, c->ientry.p, sizeof(int4) * c->nci, hipMemcpyDeviceToDevice,
                                 c->stream),
                  "keep outer list");
            check(hipMemcpyAsync(c->jentryOuter.p, c->jentry.p, sizeof(uint2) * c->ncj, hipMemcpyDeviceToDevice,
                                 c->stream),
                  "keep outer list");
            c->haveOuter = true;
        }
        const float r = static_cast<float>(rlistInner);
        PruneArgs   a{};
        a.ncluster    = nc;
        a.rlist2      = r * r;
        a.xq          = c->xq.p;
        a.shiftVec    = c->shiftVec.p;
        a.offIOuter   = c->offIOuter.p;
        a.ientryOuter = c->ientryOuter.p;
        a.jentryOuter = c->jentryOuter.p;
        a.countI      = c->countI.p;
        a.countJ      = c->countJ.p;
        launchPrune(c->stream, a, false);
        {
            const int* in[2]  = { c->countJ.p, c->countI.p };
            int*       out[2] = { c->offJ.p, c->offI.p };
            int*       tot[2] = { c->totals.p, c->totals.p + 1 };
            launchExclusiveScansWide(c->stream, nc, 2, in, out, tot);
        }
        int tot[2] = { 0, 0 };
        check(hipMemcpyAsync(tot, c->totals.p, sizeof(int) * 2, hipMemcpyDeviceToHost, c->stream), "read totals");
        check(hipStreamSynchronize(c->stream), "prune sync");
        c->ncj = tot[0];
        c->nci = tot[1];
        c->ientry.ensure(std::max(c->nci, 1));
        c->jentry.ensure(std::max(c->ncj, 1));
        a.offI   = c->offI.p;
        a.offJ   = c->offJ.p;
        a.ientry = c->ientry.p;
        a.jentry = c->jentry.p;
        launchPrune(c->stream, a, true);
        check(hipGetLastError(), "prune kernels");
        finishList(c);
    });
}

int nbx_prune_rolling(nbx_context* c, double rlistInner, int part, int numParts)
{
    if (!c)
    {
        return fail(NBX_ERROR_INPUT, "NULL context");
    }
    return guarded([&] {
        requireList(c);
        if (!(rlistInner >= c->rcoulomb && rlistInner <= c->rlist))
        {
            throw InputError("the inner pair-list cut-off must lie between the interaction cut-off and rlist");
        }
        if (numParts < 1 || part < 0 || part >= numParts)
        {
            throw InputError("the rolling prune part must lie in [0, num_parts) with num_parts >= 1");
        }
        setDevice(c);
        WorkLayout& L        = c->lay;
        const int   numWaves = L.numBlocks * kKernelWaves;
        LayoutPruneArgs a{};
        a.nchunk    = L.nchunk;
        a.numWaves  = numWaves;
        a.ngroups   = L.ngroups;
        a.part      = part;
        a.numParts  = numParts;
        a.groupSize = L.groupSize;
        a.ncluster  = c->ncluster;
        const float r = static_cast<float>(rlistInner);
        a.rlist2      = r * r;
        if (!c->layOuterValid)
        {
            // keep the layout as built: every prune starts from it
            c->chunksOuter.ensure(std::max(L.nchunk, 1));
            c->workOuter.ensure(std::max(L.nwork, 1));
            c->chunkWave.ensure(std::max(L.nchunk, 1));
            c->chunkGroup.ensure(std::max(L.nchunk, 1));
            c->keptCount.ensure(std::max(L.nchunk, 1));
            c->keptMasked.ensure(std::max(L.nchunk, 1));
            c->tileKeep.ensure(std::max(L.nwork / kTileJ, 1));
            check(hipMemcpyAsync(c->chunksOuter.p, L.chunks.p, sizeof(int4) * L.nchunk, hipMemcpyDeviceToDevice,
                                 c->stream),
                  "keep outer layout");
            check(hipMemcpyAsync(c->workOuter.p, L.work.p, sizeof(uint2) * L.nwork, hipMemcpyDeviceToDevice,
                                 c->stream),
                  "keep outer layout");
        }
        a.xq          = c->xq.p;
        a.chunkShift  = L.chunkShift.p;
        a.waveStart   = L.waveStart.p;
        a.groups      = L.groups.p;
        a.chunksOuter = c->chunksOuter.p;
        a.workOuter   = c->workOuter.p;
        a.chunkWave   = c->chunkWave.p;
        a.chunkGroup  = c->chunkGroup.p;
        a.keptCount   = c->keptCount.p;
        a.keptMasked  = c->keptMasked.p;
        a.tileKeep    = c->tileKeep.p;
        a.chunks      = L.chunks.p;
        a.work        = L.work.p;
        if (!c->layOuterValid)
        {
            launchLayoutPruneSetup(c->stream, a);
            c->layOuterValid = true;
        }
        launchLayoutPrune(c->stream, a);
        check(hipGetLastError(), "rolling prune kernels");
    });
}

int nbx_convert_coordinates(nbx_context* c, const float* x, int xOnDevice)
{
    if (!c)
    {
        return fail(NBX_ERROR_INPUT, "NULL context");
    }
    return guarded([&] {
        requireGrid(c);
        setDevice(c);
        const float* xd = deviceCoordinates(c, x, xOnDevice);
        launchConvertX(c->stream, c->ncluster * kClusterSize, c->atomIndices.p, xd, c->xq.p);
        check(hipGetLastError(), "convert coordinates");
    });
}

// One step on the context's stream: set-up (coordinates when xd is given,
// clears), the pair kernel, the force gather (into the cluster buffer, or
// with fOut straight into the caller's atom-order forces), the atom pass
// (exclusion energies, virial) and the final sums.
static void enqueueStep(nbx_context* c, unsigned flags, const float* xd, float* fOut)
{
    const bool energy = (flags & NBX_COMPUTE_ENERGY) != 0;
    const bool shift  = (flags & (NBX_COMPUTE_SHIFT_FORCES | NBX_COMPUTE_VIRIAL)) != 0;
    // the virial needs the atom-order forces and coordinates of this step
    const bool virial = (flags & NBX_COMPUTE_VIRIAL) != 0 && fOut && xd;
    // one energy slot per wavefront
    const int numSlots = c->lay.numBlocks * kKernelWaves;
    // the cluster force buffer is the output when no atom-order forces are
    // asked for, and holds the float atomics of union clusters beyond their
    // window; with NBX_COMPUTE_REPRODUCIBLE those go to the fixed-point
    // buffer instead.  Buffers not needed are neither cleared nor read.
    const bool reproducible = (flags & NBX_COMPUTE_REPRODUCIBLE) != 0;
    const bool useFnb       = !fOut || (c->lay.spill && !reproducible);
    unsigned long long* const fix = c->lay.spill && reproducible ? c->fixF.p : nullptr;
    launchPrepareStep(c->stream, c->ncluster * kClusterSize, c->atomIndices.p, xd, c->xq.p,
                      useFnb ? c->fnb.p : nullptr, energy ? c->energyPartial.p : nullptr, 2 * numSlots, fix);
    KernelArgs a{};
    a.chunks        = c->lay.chunks.p;
    a.groups        = c->lay.groups.p;
    a.groupInfo     = c->lay.groupInfo.p;
    a.ngroups       = c->lay.ngroups;
    a.groupSize     = c->lay.groupSize;
    a.ncluster      = c->ncluster;
    a.windowSlots   = c->lay.windowSlots;
    a.unionList     = c->lay.unionList.p;
    a.work          = c->lay.work.p;
    a.xq            = c->xq.p;
    a.type          = c->nbtype.p;
    a.nbfp          = c->nbfp.p;
    a.typeLJ        = c->typeLJ.p;
    a.skipLJ        = c->anyTypeWithoutLJ ? 1 : 0;
    a.vdwModifier   = c->opt.vdw_modifier;
    a.ljPme         = c->opt.vdw_type == NBX_VDW_PME ? 1 : 0;
    a.twinCut       = c->twinCut ? 1 : 0;
    a.windowSpill   = c->lay.spill ? 1 : 0;
    a.c6grid        = c->c6grid.p;
    a.ntype         = c->ntype;
    a.tabFDV0       = c->tabFDV0.p;
    a.chunkShift    = c->lay.chunkShift.p;
    a.ic            = c->ic;
    a.waveStart     = c->lay.waveStart.p;
    a.segInfo       = c->lay.gather.segInfo.p;
    a.staging       = c->lay.gather.staging.p;
    a.fnb           = c->fnb.p;
    a.fixF          = fix;
    a.energyPartial = c->energyPartial.p;
    a.waveEnergy    = 1;
    a.fshiftPartial = c->fshiftPartial.p;
    if (energy)
    {
        ExclusionArgs& e = a.excl;
        e.natoms         = c->natoms;
        e.nslot          = c->ncluster * kClusterSize;
        e.npairs         = c->numExclPairs;
        e.pairSlots      = c->exclPairSlots.p;
        e.xq             = c->xq.p;
        e.box[0]         = c->g.box[0];
        e.box[1]         = c->g.box[1];
        e.box[2]         = c->g.box[2];
        e.elec           = c->elec;
        e.ic             = c->ic;
        e.rc2            = static_cast<double>(c->ic.rcoulomb2);
        e.epsfacD        = static_cast<double>(c->ic.epsfac);
        e.selfCoef       = c->selfCoef;
        e.tabD           = c->tabFDV0d.p;
        e.partial        = c->exclPartial.p;
        // one workgroup per CU: each takes a few items per thread
        const int items = std::max(e.nslot, e.npairs);
        a.exclBlocks    = std::max(1, std::min({ kExclMaxBlocks, c->numCU, (items + 255) / 256 }));
    }
    const bool timed = 2 * static_cast<size_t>(c->timingRecorded) < c->timingEvents.size();
    check(launchNonbonded(c->stream, a, c->elec, energy, shift, c->lay.numBlocks,
                          timed ? c->timingEvents[2 * c->timingRecorded] : nullptr,
                          timed ? c->timingEvents[2 * c->timingRecorded + 1] : nullptr),
          "nonbonded kernel");
    if (timed)
    {
        c->timingRecorded++;
    }
    // the gather's workgroups also sum the virial's x (x) f term
    const int numVir =
            launchGatherForces(c->stream, c->ncluster, c->lay.gather.gposOff.p, c->lay.gather.gpos.p,
                               c->lay.gather.staging.p, useFnb ? c->fnb.p : nullptr, c->atomIndices.p, fOut,
                               (flags & NBX_COMPUTE_ADD_FORCES) != 0, c->xq.p, virial ? c->virialPartial.p : nullptr,
                               fix);
    const int numExcl = energy ? a.exclBlocks : 0;
    if (energy || shift || virial)
    {
        launchFinalize(c->stream, energy ? numSlots : 0, c->energyPartial.p, energy ? numExcl : 0,
                       c->exclPartial.p, c->lay.numBlocks, c->fshiftPartial.p, shift, c->energy.p, c->fshift.p,
                       virial ? numVir : 0, c->virialPartial.p, c->shiftVec.p, virial ? c->virial.p : nullptr);
    }
    check(hipGetLastError(), "nonbonded kernel");
    c->haveEnergy = energy;
    c->haveShift  = shift;
    c->haveVirial = virial;
}

int nbx_dispatch_kernel(nbx_context* c, unsigned flags)
{
    if (!c)
    {
        return fail(NBX_ERROR_INPUT, "NULL context");
    }
    return guarded([&] {
        requireList(c);
        setDevice(c);
        enqueueStep(c, flags & ~static_cast<unsigned>(NBX_COMPUTE_VIRIAL), nullptr, nullptr);
    });
}

int nbx_add_nbat_f_to_f(nbx_context* c, float* f, int fOnDevice)
{
    if (!c)
    {
        return fail(NBX_ERROR_INPUT, "NULL context");
    }
    return guarded([&] {
        requireList(c);
        if (!f)
        {
            throw InputError("force output must not be NULL");
        }
        setDevice(c);
        float* fd = f;
        if (!fOnDevice)
        {
            c->fOut.ensure(3 * static_cast<size_t>(c->natoms));
            fd = c->fOut.p;
        }
        launchReduceF(c->stream, c->natoms, c->cells.p, c->fnb.p, fd);
        if (!fOnDevice)
        {
            check(hipMemcpyAsync(f, fd, sizeof(float) * 3 * c->natoms, hipMemcpyDeviceToHost, c->stream), "download f");
            check(hipStreamSynchronize(c->stream), "sync");
        }
    });
}

int nbx_compute(nbx_context* c, const float* x, float* f, unsigned flags, int onDevice)
{
    if (!c)
    {
        return fail(NBX_ERROR_INPUT, "NULL context");
    }
    return guarded([&] {
        requireList(c);
        if (!f)
        {
            throw InputError("force output must not be NULL");
        }
        setDevice(c);
        const float* xd = deviceCoordinates(c, x, onDevice);
        float*       fd = f;
        if (!onDevice)
        {
            c->fOut.ensure(3 * static_cast<size_t>(c->natoms));
            fd = c->fOut.p;
            if (flags & NBX_COMPUTE_ADD_FORCES)
            {
                check(hipMemcpyAsync(fd, f, sizeof(float) * 3 * c->natoms, hipMemcpyHostToDevice, c->stream),
                      "upload f");
            }
        }
        enqueueStep(c, flags, xd, fd);
        if (!onDevice)
        {
            check(hipMemcpyAsync(f, fd, sizeof(float) * 3 * c->natoms, hipMemcpyDeviceToHost, c->stream), "download f");
            check(hipStreamSynchronize(c->stream), "sync");
        }
    });
}

int nbx_set_charges(nbx_context* c, const float* charges, int onDevice)
{
    if (!c)
    {
        return fail(NBX_ERROR_INPUT, "NULL context");
    }
    return guarded([&] {
        requireGrid(c);
        if (!charges)
        {
            throw InputError("charges must not be NULL");
        }
        setDevice(c);
        if (onDevice)
        {
            check(hipMemcpyAsync(c->charge.p, charges, sizeof(float) * c->natoms, hipMemcpyDeviceToDevice, c->stream),
                  "copy q");
        }
        else
        {
            for (int a = 0; a < c->natoms; a++)
            {
                if (!std::isfinite(charges[a]))
                {
                    throw InputError("non-finite charge");
                }
            }
            check(hipMemcpyAsync(c->charge.p, charges, sizeof(float) * c->natoms, hipMemcpyHostToDevice, c->stream),
                  "upload q");
        }
        launchSetCharges(c->stream, c->ncluster * kClusterSize, c->atomIndices.p, c->charge.p, c->xq.p);
        check(hipGetLastError(), "set charges");
        if (!onDevice)
        {
            check(hipStreamSynchronize(c->stream), "sync");
        }
    });
}

int nbx_compute_adjoint(nbx_context* c, const float* x, const float* g, float* gradX, double* gradQ, double* gradC6,
                        double* gradC12, int onDevice)
{
    if (!c)
    {
        return fail(NBX_ERROR_INPUT, "NULL context");
    }
    return guarded([&] {
        requireList(c);
        if (!g)
        {
            throw InputError("the force cotangent g must not be NULL");
        }
        if (c->opt.vdw_modifier != NBX_VDW_POTENTIAL_SHIFT || c->opt.vdw_type != NBX_VDW_CUTOFF || c->twinCut)
        {
            throw InputError("the adjoint supports the LJ cut-off with the potential shift at rvdw = rcoulomb only");
        }
        if (c->ntype > kAdjointMaxTypes)
        {
            throw InputError("the adjoint supports at most " + std::to_string(kAdjointMaxTypes - 1) + " particle types");
        }
        setDevice(c);
        if (!c->adjValid)
        {
            // the adjoint's own work layout: groups sized for its window
            buildWorkGroups(c, c->adj, true);
            scheduleLayout(c, c->adj, true);
            c->adjValid = true;
        }
        const int lds = adjointLdsBytes(c->ntype, c->ic.tab_n, c->elec, c->adj.windowSlots);
        if (lds > kAdjointLdsBudget)
        {
            throw InputError("the adjoint's j-cluster window does not fit in LDS: choose a smaller kernel_group_size");
        }
        setDevice(c);
        const size_t na    = static_cast<size_t>(c->natoms);
        const int    nslot = c->ncluster * kClusterSize;
        const float* xd    = deviceCoordinates(c, x, onDevice);
        const float* gd    = g;
        if (!onDevice)
        {
            c->gIn.ensure(3 * na);
            check(hipMemcpyAsync(c->gIn.p, g, sizeof(float) * 3 * na, hipMemcpyHostToDevice, c->stream), "upload g");
            gd = c->gIn.p;
        }
        c->gq.ensure(nslot + kClusterSize);
        c->gradXnb.ensure(nslot + kClusterSize);
        c->gradQnb.ensure(nslot + kClusterSize);
        launchAdjointPrepare(c->stream, nslot, c->atomIndices.p, xd, gd, c->xq.p, c->gq.p, c->gradXnb.p, c->gradQnb.p);
        const int    blocks = adjointBlocks(c->adj.numBlocks);
        const size_t np     = 2 * static_cast<size_t>(c->ntype) * c->ntype;
        c->paramPartial.ensure(blocks * np);
        const int tu    = c->ntypeUser;
        float*    gxOut = gradX;
        double*   gqOut = gradQ;
        double*   g6Out = gradC6;
        double*   g12Out = gradC12;
        if (!onDevice)
        {
            c->gradXOut.ensure(3 * na);
            c->gradQOut.ensure(na);
            c->paramOut.ensure(2 * static_cast<size_t>(tu) * tu);
            gxOut  = gradX ? c->gradXOut.p : nullptr;
            gqOut  = gradQ ? c->gradQOut.p : nullptr;
            g6Out  = gradC6 ? c->paramOut.p : nullptr;
            g12Out = gradC12 ? c->paramOut.p + static_cast<size_t>(tu) * tu : nullptr;
        }
        AdjointArgs a{};
        a.chunks       = c->adj.chunks.p;
        a.groups       = c->adj.groups.p;
        a.groupInfo    = c->adj.groupInfo.p;
        a.ncluster     = c->ncluster;
        a.windowSlots  = c->adj.windowSlots;
        a.unionList    = c->adj.unionList.p;
        a.work         = c->adj.work.p;
        a.xq           = c->xq.p;
        a.gq           = c->gq.p;
        a.type         = c->nbtype.p;
        a.nbfp         = c->nbfp.p;
        a.ntype        = c->ntype;
        a.tabFDV0      = c->tabFDV0.p;
        a.chunkShift   = c->adj.chunkShift.p;
        a.ic           = c->ic;
        a.waveStart    = c->adj.waveStart.p;
        a.segInfo      = c->adj.gather.segInfo.p;
        a.staging      = c->adj.gather.staging.p;
        a.gposOff      = c->adj.gather.gposOff.p;
        a.gpos         = c->adj.gather.gpos.p;
        a.gradX        = c->gradXnb.p;
        a.gradQ        = c->gradQnb.p;
        a.paramPartial = c->paramPartial.p;
        launchAdjoint(c->stream, a, c->elec, c->adj.numBlocks, c->atomIndices.p, gxOut, gqOut, tu, g6Out, g12Out);
        check(hipGetLastError(), "adjoint kernels");
        if (!onDevice && gradX)
        {
            check(hipMemcpyAsync(gradX, gxOut, sizeof(float) * 3 * na, hipMemcpyDeviceToHost, c->stream), "download");
        }
        if (!onDevice && gradQ)
        {
            check(hipMemcpyAsync(gradQ, gqOut, sizeof(double) * na, hipMemcpyDeviceToHost, c->stream), "download");
        }
        const size_t nt = static_cast<size_t>(tu) * tu;
        if (!onDevice && gradC6)
        {
            check(hipMemcpyAsync(gradC6, g6Out, sizeof(double) * nt, hipMemcpyDeviceToHost, c->stream), "download");
        }
        if (!onDevice && gradC12)
        {
            check(hipMemcpyAsync(gradC12, g12Out, sizeof(double) * nt, hipMemcpyDeviceToHost, c->stream), "download");
        }
        if (!onDevice)
        {
            check(hipStreamSynchronize(c->stream), "sync");
        }
    });
}

int nbx_get_energies(nbx_context* c, double* vlj, double* vc)
{
    if (!c)
    {
        return fail(NBX_ERROR_INPUT, "NULL context");
    }
    return guarded([&] {
        if (!c->haveEnergy)
        {
            throw StateError("no energies: compute with NBX_COMPUTE_ENERGY first");
        }
        setDevice(c);
        double e[2];
        check(hipMemcpyAsync(e, c->energy.p, sizeof(e), hipMemcpyDeviceToHost, c->stream), "read energies");
        check(hipStreamSynchronize(c->stream), "sync");
        e[0] += ljPmeSelfEnergy(c);
        if (vlj)
        {
            *vlj = e[0];
        }
        if (vc)
        {
            *vc = e[1];
        }
    });
}

int nbx_get_shift_forces(nbx_context* c, double* fshift)
{
    if (!c)
    {
        return fail(NBX_ERROR_INPUT, "NULL context");
    }
    return guarded([&] {
        if (!c->haveShift)
        {
            throw StateError("no shift forces: compute with NBX_COMPUTE_SHIFT_FORCES first");
        }
        setDevice(c);
        check(hipMemcpyAsync(fshift, c->fshift.p, sizeof(double) * kNumShifts * 3, hipMemcpyDeviceToHost, c->stream),
              "read fshift");
        check(hipStreamSynchronize(c->stream), "sync");
    });
}

int nbx_get_virial(nbx_context* c, double* virial)
{
    if (!c)
    {
        return fail(NBX_ERROR_INPUT, "NULL context");
    }
    return guarded([&] {
        if (!c->haveVirial)
        {
            throw StateError("no virial: compute with NBX_COMPUTE_VIRIAL first");
        }
        setDevice(c);
        check(hipMemcpyAsync(virial, c->virial.p, sizeof(double) * 9, hipMemcpyDeviceToHost, c->stream), "read virial");
        check(hipStreamSynchronize(c->stream), "sync");
    });
}

int nbx_synchronize(nbx_context* c)
{
    if (!c)
    {
        return fail(NBX_ERROR_INPUT, "NULL context");
    }
    return guarded([&] {
        setDevice(c);
        check(hipStreamSynchronize(c->stream), "sync");
    });
}

int nbx_get_pairlist_info(nbx_context* c, nbx_pairlist_info* info)
{
    if (!c || !info)
    {
        return fail(NBX_ERROR_INPUT, "NULL argument");
    }
    return guarded([&] {
        info->num_atoms     = std::max(c->natoms, 0);
        info->num_clusters  = c->gridValid ? c->ncluster : 0;
        info->num_columns_x = c->gridValid ? c->g.ncx : 0;
        info->num_columns_y = c->gridValid ? c->g.ncy : 0;
        info->num_i_entries = c->listValid ? c->nci : 0;
        info->num_j_entries = c->listValid ? c->ncj : 0;
    });
}

int nbx_get_grid(nbx_context* c, int* atomIndices, int* cells)
{
    if (!c)
    {
        return fail(NBX_ERROR_INPUT, "NULL context");
    }
    return guarded([&] {
        requireGrid(c);
        setDevice(c);
        if (atomIndices)
        {
            check(hipMemcpyAsync(atomIndices, c->atomIndices.p, sizeof(int) * c->ncluster * kClusterSize,
                                 hipMemcpyDeviceToHost, c->stream),
                  "read grid");
        }
        if (cells)
        {
            check(hipMemcpyAsync(cells, c->cells.p, sizeof(int) * c->natoms, hipMemcpyDeviceToHost, c->stream),
                  "read cells");
        }
        check(hipStreamSynchronize(c->stream), "sync");
    });
}

int nbx_get_pairlist(nbx_context* c, int* iEntries, uint32_t* jEntries)
{
    if (!c)
    {
        return fail(NBX_ERROR_INPUT, "NULL context");
    }
    return guarded([&] {
        requireList(c);
        setDevice(c);
        if (iEntries)
        {
            check(hipMemcpyAsync(iEntries, c->ientry.p, sizeof(int4) * c->nci, hipMemcpyDeviceToHost, c->stream),
                  "read ientries");
        }
        if (jEntries)
        {
            check(hipMemcpyAsync(jEntries, c->jentry.p, sizeof(uint2) * c->ncj, hipMemcpyDeviceToHost, c->stream),
                  "read jentries");
        }
        check(hipStreamSynchronize(c->stream), "sync");
        if (jEntries)
        {
            // the kernel keeps its LDS window slot in the upper 16 bits
            for (int k = 0; k < c->ncj; k++)
            {
                jEntries[2 * k + 1] &= 0xffffu;
            }
        }
    });
}

int64_t nbx_list_pairs(nbx_context* c, double r, int64_t* out, int64_t cap)
{
    if (!c)
    {
        return -fail(NBX_ERROR_INPUT, "NULL context");
    }
    int64_t result = 0;
    int     rc     = guarded([&] {
        requireList(c);
        setDevice(c);
        DevBuf<unsigned long long> counter;
        counter.ensure(1);
        DevBuf<long long> buf;
        if (out && cap > 0)
        {
            buf.ensure(static_cast<size_t>(cap));
        }
        check(hipMemsetAsync(counter.p, 0, sizeof(unsigned long long), c->stream), "clear counter");
        const float rf = static_cast<float>(r);
        launchExpandPairs(c->stream, c->ientry.p, c->nci, c->jentry.p, c->xq.p, c->atomIndices.p, c->shiftVec.p,
                          rf * rf, counter.p, (out && cap > 0) ? buf.p : nullptr, cap);
        check(hipGetLastError(), "expand pairs");
        unsigned long long n = 0;
        check(hipMemcpyAsync(&n, counter.p, sizeof(n), hipMemcpyDeviceToHost, c->stream), "read count");
        check(hipStreamSynchronize(c->stream), "sync");
        result = static_cast<int64_t>(n);
        if (out && cap > 0)
        {
            const int64_t m = std::min<int64_t>(result, cap);
            check(hipMemcpy(out, buf.p, sizeof(long long) * m, hipMemcpyDeviceToHost), "read pairs");
        }
    });
    return rc == NBX_SUCCESS ? result : -rc;
}

int nbx_get_work_groups(nbx_context* c, int* info, int* groups, int* chunks, int* unionList, uint32_t* workOut)
{
    if (!c || !info)
    {
        return fail(NBX_ERROR_INPUT, "NULL argument");
    }
    return guarded([&] {
        requireList(c);
        setDevice(c);
        info[0] = c->lay.groupSize;
        info[1] = c->lay.ngroups;
        info[2] = c->lay.nchunk;
        info[3] = c->lay.windowSlots;
        info[4] = c->lay.nwork;
        if (groups)
        {
            check(hipMemcpyAsync(groups, c->lay.groups.p, sizeof(int4) * c->lay.ngroups, hipMemcpyDeviceToHost, c->stream),
                  "read groups");
        }
        if (chunks)
        {
            check(hipMemcpyAsync(chunks, c->lay.chunks.p, sizeof(int4) * c->lay.nchunk, hipMemcpyDeviceToHost, c->stream),
                  "read chunks");
        }
        if (unionList)
        {
            int4 last;
            check(hipMemcpyAsync(&last, c->lay.groups.p + (c->lay.ngroups - 1), sizeof(int4), hipMemcpyDeviceToHost, c->stream),
                  "read groups");
            check(hipStreamSynchronize(c->stream), "sync");
            check(hipMemcpyAsync(unionList, c->lay.unionList.p, sizeof(int) * (last.z + last.w), hipMemcpyDeviceToHost,
                                 c->stream),
                  "read union");
        }
        if (workOut)
        {
            check(hipMemcpyAsync(workOut, c->lay.work.p, sizeof(uint2) * c->lay.nwork, hipMemcpyDeviceToHost, c->stream),
                  "read work list");
        }
        check(hipStreamSynchronize(c->stream), "sync");
    });
}

int nbx_get_interaction_constants(nbx_context* c, double* out)
{
    if (!c || !out)
    {
        return fail(NBX_ERROR_INPUT, "NULL argument");
    }
    const InteractionConst& ic = c->ic;
    out[0] = ic.beta;
    out[1] = ic.tab_scale;
    out[2] = ic.tab_n;
    out[3] = ic.epsfac;
    out[4] = ic.k_rf;
    out[5] = ic.c_rf;
    out[6] = ic.sh_ewald;
    out[7] = ic.cpot_disp;
    out[8] = ic.cpot_rep;
    return NBX_SUCCESS;
}

int nbx_get_vdw_constants(nbx_context* c, double* out)
{
    if (!c || !out)
    {
        return fail(NBX_ERROR_INPUT, "NULL argument");
    }
    const InteractionConst& ic = c->ic;
    const double v[12] = { std::sqrt(static_cast<double>(ic.rvdw2)), ic.rvdw_switch, ic.fsw_c2d, ic.fsw_c3d,
                           ic.fsw_c2r, ic.fsw_c3r, ic.sw_c3, ic.sw_c4, ic.sw_c5, ic.lje_coeff2, ic.lje_coeff6_6,
                           ic.sh_lj_ewald };
    for (int k = 0; k < 12; k++)
    {
        out[k] = v[k];
    }
    return NBX_SUCCESS;
}

int nbx_get_ewald_table(nbx_context* c, int* n, float* fdv0)
{
    if (!c || !n)
    {
        return fail(NBX_ERROR_INPUT, "NULL argument");
    }
    return guarded([&] {
        const bool ewald = c->opt.coulomb_type == NBX_COULOMB_EWALD;
        *n               = ewald ? c->ic.tab_n : 0;
        if (fdv0 && ewald)
        {
            setDevice(c);
            check(hipMemcpy(fdv0, c->tabFDV0.p, sizeof(float4) * c->ic.tab_n, hipMemcpyDeviceToHost),
                  "download table");
        }
    });
}

} // extern "C"
