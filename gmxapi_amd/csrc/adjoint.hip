// Adjoint of the cluster-pair forces: the vector-Jacobian product of the
// pair kernel's forces with a per-atom input g = dL/dF (gfx950, wave64).
//
// The pair kernel (kernel.hip) adds f_ij = s(r^2; q, C6, C12) d_ij to atom i
// and subtracts it from atom j, with d_ij = x_i + shift - x_j and
//   s = 12 C12 r^-14 - 6 C6 r^-8 + epsfac q_i q_j h(r),
// h(r) = r^-3 (non-excluded pairs only) minus the Ewald correction's force
// over r (or the reaction-field constant), all zero beyond the cut-off.
// For L(F) with g_i = dL/dF_i and u_ij = (g_i - g_j) . d_ij:
//   dL/dx_i += s (g_i - g_j) + 2 s' u_ij d_ij,   dL/dx_j -= the same,
//   dL/dq_i += u_ij epsfac q_j h,                dL/dq_j += u_ij epsfac q_i h,
//   dL/dC6[ti][tj] += -6 u_ij r^-8,              dL/dC12[ti][tj] += 12 u_ij r^-14,
// where s' = ds/d(r^2).  The C6/C12 gradients are with respect to the
// matrices given to nbx_set_nonbonded_parameters.  The pair list is not
// stored per pair: the kernel walks the same work list, groups and static
// schedule as the forward kernel and recomputes every pair.
//
// Mapping: as in the forward kernel, four wave-uniform i-atoms per tile and
// one j-atom per lane.  A wavefront's LDS window holds {dL/dx, dL/dq} per
// j-cluster atom (float4); i-atom sums live in 16 registers per i-entry and
// are reduce-scattered once per entry.  Windows are staged with plain stores
// and gathered per cluster after the launch, as in the forward path (dL/dq
// summed in double there); the C6/C12 gradients are summed per workgroup in
// LDS in double (ds_add_f64) and the workgroup sums are added in a fixed
// order.
// Two wavefronts per workgroup keep the float4 window of a workgroup small.
#include "kernel_common.h"

namespace nbx
{

namespace
{

constexpr int kAdjWaves   = 2;
constexpr int kAdjThreads = kWave * kAdjWaves;

// LDS layout of an adjoint workgroup, in float4 units
struct AdjLayout
{
    int nbfp, tab, ix, ig, it, win, acc, end_;
    __host__ __device__ AdjLayout(int ntype, int tabN, int windowSlots)
    {
        nbfp = 0;
        tab  = nbfp + (ntype * ntype + 1) / 2;
        ix   = tab + tabN;
        ig   = ix + kAdjWaves * kGroupSizeMax * kClusterSize;
        it   = ig + kAdjWaves * kGroupSizeMax * kClusterSize;
        win  = it + kAdjWaves * kGroupSizeMax;
        acc  = win + kAdjWaves * windowSlots * 4;
        // {dC6, dC12} per type pair as two doubles, in adjointAccCopies copies
        end_ = acc + adjointAccCopies(ntype) * ntype * ntype;
    }
};

// The four i-atoms' g (wave-uniform)
struct IGrad
{
    float x[4], y[4], z[4];
    float qe[4]; // epsfac q_i
};

template<int kElec, bool kParams, bool kChecks, int kI>
__device__ __forceinline__ void adjointPairI(const IAtoms& I, const IGrad& Gi, const JAtom& J, const float4 gj,
                                             const unsigned mbits, const bool diag, const int jj,
                                             const InteractionConst& ic, const float2* __restrict__ sNbfp,
                                             const float4* __restrict__ sTab, double* __restrict__ sAcc,
                                             const int accStride, const int accCopies, float* ai, float4& aj)
{
    const float dx       = I.x[kI] - J.x;
    const float dy       = I.y[kI] - J.y;
    const float dz       = I.z[kI] - J.z;
    float       rsq      = dx * dx + dy * dy + dz * dz;
    float       interact = 1.0f;
    bool        skip;
    if (kChecks)
    {
        interact = static_cast<float>((mbits >> (4 * kI)) & 1u);
        skip     = (diag && jj <= kI) || !(rsq < ic.rcoulomb2);
    }
    else
    {
        skip = !(rsq < ic.rcoulomb2);
    }
    rsq           = fmaxf(rsq, kMinRsq);
    // the hardware estimate (within one ulp, nearly unbiased; see the pair
    // kernel)
    float rinv    = __builtin_amdgcn_rsqf(rsq);
    rinv          = skip ? 0.0f : rinv;
    const float smask  = skip ? 0.0f : 1.0f;
    const float rinvsq = rinv * rinv;
    const float rinv3  = rinv * rinvsq;

    // LJ: s = c12 r^-14 - c6 r^-8 (c6 = 6 C6, c12 = 12 C12), s' = -7 c12 r^-16 + 4 c6 r^-10
    const float2 c     = sNbfp[I.tOff[kI] + J.type];
    const float  rinv6 = interact * rinvsq * rinvsq * rinvsq;
    const float  r8    = rinv6 * rinvsq;
    const float  r14   = rinv6 * r8;
    const float  a12   = c.y * r14;
    const float  a6    = c.x * r8;
    const float  sLJ   = a12 - a6;
    const float  dsLJ  = fmaf(a12, -7.0f, 4.0f * a6) * rinvsq;

    // Coulomb per unit epsfac q_i q_j: h and h' = dh/d(r^2)
    float hC, dhC;
    if (kElec == kElecRF)
    {
        hC  = smask * (interact * rinv3 - 2.0f * ic.k_rf);
        dhC = -1.5f * interact * rinv3 * rinvsq;
    }
    else if (kElec == kElecEwaldTable)
    {
        // h = interact r^-3 - F(r)/r with the table's linear force F;
        // dh/d(r^2) = -1.5 interact r^-5 - (F'(r) r - F(r)) / (2 r^3)
        const float  rs    = rsq * rinv * ic.tab_scale;
        const int    ri    = min(static_cast<int>(rs), ic.tab_n - 2);
        const float  frac  = rs - static_cast<float>(ri);
        const float4 tab   = sTab[ri];
        const float  fexcl = tab.x + frac * tab.y;
        hC                 = interact * rinv3 - fexcl * rinv;
        dhC                = -1.5f * interact * rinv3 * rinvsq - 0.5f * rinvsq * (tab.y * ic.tab_scale - fexcl * rinv);
    }
    else
    {
        // h = interact r^-3 - beta^3 G(s), s = beta^2 r^2, G = P/Q;
        // dh/d(r^2) = -1.5 interact r^-5 - beta^5 G'(s)
        const float s   = fminf(rsq * ic.beta2, kAnaMaxS);
        const float pg  = fmaf(fmaf(fmaf(fmaf(fmaf(kGp5, s, kGp4), s, kGp3), s, kGp2), s, kGp1), s, kGp0);
        const float qg  = fmaf(fmaf(fmaf(fmaf(fmaf(kGq5, s, kGq4), s, kGq3), s, kGq2), s, kGq1), s, 1.0f);
        const float dpg = fmaf(fmaf(fmaf(fmaf(5.0f * kGp5, s, 4.0f * kGp4), s, 3.0f * kGp3), s, 2.0f * kGp2), s, kGp1);
        const float dqg = fmaf(fmaf(fmaf(fmaf(5.0f * kGq5, s, 4.0f * kGq4), s, 3.0f * kGq3), s, 2.0f * kGq2), s, kGq1);
        const float rq  = __builtin_amdgcn_rcpf(qg);
        const float G   = pg * rq;
        const float dG  = (dpg - G * dqg) * rq;
        hC              = smask * (interact * rinv3 - ic.beta3 * G);
        dhC             = smask * (-1.5f * interact * rinv3 * rinvsq - ic.beta3 * ic.beta2 * dG);
    }
    const float qq  = I.q[kI] * J.w; // epsfac q_i q_j
    const float s   = fmaf(qq, hC, sLJ);
    const float sp  = fmaf(qq, dhC, dsLJ);
    const float gdx = Gi.x[kI] - gj.x;
    const float gdy = Gi.y[kI] - gj.y;
    const float gdz = Gi.z[kI] - gj.z;
    const float u   = gdx * dx + gdy * dy + gdz * dz;
    const float t   = 2.0f * sp * u;
    const float wx  = fmaf(s, gdx, t * dx);
    const float wy  = fmaf(s, gdy, t * dy);
    const float wz  = fmaf(s, gdz, t * dz);
    const float uh  = u * hC;
    ai[4 * kI + 0] += wx;
    ai[4 * kI + 1] += wy;
    ai[4 * kI + 2] += wz;
    ai[4 * kI + 3] = fmaf(uh, J.w, ai[4 * kI + 3]);
    aj.x -= wx;
    aj.y -= wy;
    aj.z -= wz;
    aj.w = fmaf(uh, Gi.qe[kI], aj.w);
    if (kParams)
    {
        // lanes jj of each j-cluster spread over the accumulator copies
        double* a = sAcc + 2 * (I.tOff[kI] + J.type) + accStride * (jj & (accCopies - 1));
        atomicAdd(a, static_cast<double>(-6.0f * u * r8));
        atomicAdd(a + 1, static_cast<double>(12.0f * u * r14));
    }
}

template<int kElec, bool kParams, bool kChecks>
__device__ __forceinline__ void adjointTile(const IAtoms& I, const IGrad& Gi, const JAtom& J, const float4 gj,
                                            const unsigned mbits, const bool diag, const int jj,
                                            const InteractionConst& ic, const float2* __restrict__ sNbfp,
                                            const float4* __restrict__ sTab, double* __restrict__ sAcc,
                                            const int accStride, const int accCopies, float* ai, float4& aj)
{
    adjointPairI<kElec, kParams, kChecks, 0>(I, Gi, J, gj, mbits, diag, jj, ic, sNbfp, sTab, sAcc, accStride,
                                             accCopies, ai, aj);
    __builtin_amdgcn_sched_barrier(0);
    adjointPairI<kElec, kParams, kChecks, 1>(I, Gi, J, gj, mbits, diag, jj, ic, sNbfp, sTab, sAcc, accStride,
                                             accCopies, ai, aj);
    __builtin_amdgcn_sched_barrier(0);
    adjointPairI<kElec, kParams, kChecks, 2>(I, Gi, J, gj, mbits, diag, jj, ic, sNbfp, sTab, sAcc, accStride,
                                             accCopies, ai, aj);
    __builtin_amdgcn_sched_barrier(0);
    adjointPairI<kElec, kParams, kChecks, 3>(I, Gi, J, gj, mbits, diag, jj, ic, sNbfp, sTab, sAcc, accStride,
                                             accCopies, ai, aj);
}

template<int kElec, bool kParams>
__global__ __launch_bounds__(kAdjThreads) void adjointKernel(const AdjointArgs a)
{
    extern __shared__ float4 smem4[];
    const int       ntype  = a.ntype;
    const int       ntype2 = ntype * ntype;
    const int       tabN   = (kElec == kElecEwaldTable) ? a.ic.tab_n : 0;
    const AdjLayout L(ntype, tabN, a.windowSlots);
    float2*         sNbfp = reinterpret_cast<float2*>(smem4 + L.nbfp);
    float4*         sTab  = smem4 + L.tab;
    double*         sAcc  = reinterpret_cast<double*>(smem4 + L.acc);
    const int       accCopies = adjointAccCopies(ntype);
    const int       accStride = 2 * ntype2; // doubles per accumulator copy

    for (int t = threadIdx.x; t < ntype2; t += kAdjThreads)
    {
        sNbfp[t] = a.nbfp[t];
    }
    for (int t = threadIdx.x; t < tabN; t += kAdjThreads)
    {
        sTab[t] = a.tabFDV0[t];
    }
    for (int t = threadIdx.x; t < accCopies * accStride; t += kAdjThreads)
    {
        sAcc[t] = 0.0;
    }
    __syncthreads();

    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    const int jj   = lane & 3;
    const int kk   = lane >> 2;

    const InteractionConst ic = a.ic;
    float4* ix  = smem4 + L.ix + wave * kGroupSizeMax * kClusterSize;
    float4* ig  = smem4 + L.ig + wave * kGroupSizeMax * kClusterSize;
    int*    it  = reinterpret_cast<int*>(smem4 + L.it) + wave * kGroupSizeMax * kClusterSize;
    float*  win = reinterpret_cast<float*>(smem4 + L.win) + wave * a.windowSlots * 16;

    // the wavefront's run of chunks, as in the forward kernel (the same
    // static schedule: wavefront w of the forward launch is wavefront w here)
    const int  wid    = blockIdx.x * kAdjWaves + wave;
    const int2 ws     = a.waveStart[wid];
    const int  chAll0 = uniformInt(ws.x);
    const int  chAll1 = uniformInt(a.waveStart[wid + 1].x);
    if (chAll0 < chAll1)
    {
        const int4     cFirst = a.chunks[chAll0];
        const int4     cLast  = a.chunks[chAll1 - 1];
        const unsigned W0     = static_cast<unsigned>(uniformInt(cFirst.z));
        const int      T      = (uniformInt(cLast.z) - static_cast<int>(W0)) / kTileJ + uniformInt(cLast.w);

        int    g = uniformInt(ws.y), chGrpEnd = 0, U = 0, c0 = 0, nOwn = 0, G = 0, stage = -1;
        IAtoms I;
        IGrad  Gi;
        float4 sv;
        float  ai[16];
        int    ci = 0, sh = 0, nspec = 0, chStart = 0, chEnd = 0, chIdx = chAll0;
        bool   central = false;

        auto beginGroup = [&](int gg) {
            const int4 grp = a.groups[gg];
            const int4 gi  = a.groupInfo[gg];
            const int4 si  = a.segInfo[gg];
            chGrpEnd       = uniformInt(grp.y);
            U              = uniformInt(grp.w);
            stage = uniformInt(si.z) > 0 ? uniformInt(si.x) + (wid - uniformInt(si.y)) * uniformInt(si.w) : -1;
            c0             = uniformInt(gi.x);
            nOwn           = uniformInt(gi.y) & 0xffff;
            G              = uniformInt(gi.y) >> 16;
            for (int t = lane; t < (G + U + 1) * 4; t += kWave)
            {
                reinterpret_cast<float4*>(win)[t] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            }
            if (lane < nOwn * kClusterSize)
            {
                ix[lane] = a.xq[c0 * kClusterSize + lane];
                ig[lane] = a.gq[c0 * kClusterSize + lane];
                it[lane] = a.type[c0 * kClusterSize + lane] * ntype;
            }
        };
        auto endGroup = [&]() {
            // stage the window (4 float4 per cluster slot) with plain stores;
            // the adjoint's gather adds each cluster's copies in a fixed order
            if (stage >= 0)
            {
                const float4* w4 = reinterpret_cast<const float4*>(win);
                for (int t = lane; t < (G + U) * 4; t += kWave)
                {
                    a.staging[stage + t] = w4[t];
                }
            }
        };
        auto beginChunk = [&](const int4 cd, const float4 sf, int t) {
            ci              = uniformInt(cd.x);
            const int flags = uniformInt(cd.y);
            sh              = flags & 0xff;
            nspec           = flags >> 9;
            central         = (sh == kCentral);
            chStart         = t;
            chEnd           = t + uniformInt(cd.w);
            const int lc    = (ci - c0) * kClusterSize;
#pragma unroll
            for (int i = 0; i < 4; i++)
            {
                const float4 p = ix[lc + i];
                I.x[i]         = uniformFloat(p.x);
                I.y[i]         = uniformFloat(p.y);
                I.z[i]         = uniformFloat(p.z);
                I.q[i]         = uniformFloat(p.w);
                I.tOff[i]      = uniformInt(it[lc + i]);
                const float4 q = ig[lc + i];
                Gi.x[i]        = uniformFloat(q.x);
                Gi.y[i]        = uniformFloat(q.y);
                Gi.z[i]        = uniformFloat(q.z);
                Gi.qe[i]       = uniformFloat(p.w * ic.epsfac);
            }
            sv = make_float4(uniformFloat(sf.x), uniformFloat(sf.y), uniformFloat(sf.z), 0.0f);
#pragma unroll
            for (int k = 0; k < 16; k++)
            {
                ai[k] = 0.0f;
            }
        };
        auto endChunk = [&]() {
            const float r = waveReduceScatter16(ai, lane);
            if ((lane & 3) == 0)
            {
                win[(ci - c0) * 16 + (lane >> 2)] += r;
            }
        };
        auto doTile = [&](const uint2 e, const float4 xj, const int tj, const float4 gj, const bool checked) {
            const JAtom J{ xj.x - sv.x, xj.y - sv.y, xj.z - sv.z, ic.epsfac * xj.w, tj };
            float4      aj = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (checked)
            {
                const bool diag = central && static_cast<int>(e.x) == ci;
                adjointTile<kElec, kParams, true>(I, Gi, J, gj, (e.y & 0xffffu) >> jj, diag, jj, ic, sNbfp, sTab,
                                                  sAcc, accStride, accCopies, ai, aj);
            }
            else
            {
                adjointTile<kElec, kParams, false>(I, Gi, J, gj, 0u, false, jj, ic, sNbfp, sTab, sAcc, accStride,
                                                   accCopies, ai, aj);
            }
            const unsigned slot = e.y >> 16;
            if (slot != kSlotOverflow)
            {
                float4* w = reinterpret_cast<float4*>(win + slot * 16 + jj * 4);
                float4  v = *w;
                v.x += aj.x;
                v.y += aj.y;
                v.z += aj.z;
                v.w += aj.w;
                *w = v;
            }
            else
            {
                const int k   = static_cast<int>(e.x) * kClusterSize + jj;
                float*    dst = reinterpret_cast<float*>(a.gradX + k);
                unsafeAtomicAdd(dst, aj.x);
                unsafeAtomicAdd(dst + 1, aj.y);
                unsafeAtomicAdd(dst + 2, aj.z);
                unsafeAtomicAdd(a.gradQ + k, static_cast<double>(aj.w));
            }
        };
        int4   cdNext = a.chunks[min(chIdx + 1, chAll1 - 1)];
        float4 svNext = a.chunkShift[min(chIdx + 1, chAll1 - 1)];
        auto   advance = [&](int t) {
            endChunk();
            chIdx++;
            while (chIdx == chGrpEnd)
            {
                endGroup();
                g++;
                beginGroup(g);
            }
            beginChunk(cdNext, svNext, t);
            cdNext = a.chunks[min(chIdx + 1, chAll1 - 1)];
            svNext = a.chunkShift[min(chIdx + 1, chAll1 - 1)];
        };

        // the forward kernel's pipeline: work entries two tiles ahead, j-atom
        // data one tile ahead; reads past the last tile are clamped to it
        auto entryAt = [&](int t) {
            return ldg(a.work, W0 + static_cast<unsigned>(kTileJ * min(t, T - 1) + kk));
        };
        uint2  eA = entryAt(0);
        uint2  eB = entryAt(1);
        float4 xA = ldg(a.xq, eA.x * kClusterSize + jj);
        int    tA = ldg(a.type, eA.x * kClusterSize + jj);
        float4 gA = ldg(a.gq, eA.x * kClusterSize + jj);
        beginGroup(g);
        beginChunk(cFirst, a.chunkShift[chAll0], 0);
        // an odd tile count: the first tile alone, so that the loop below
        // always runs whole pairs of tiles (no exit between its two tiles:
        // the accumulators stay in their registers around the loop)
        int t0 = 0;
        if (T & 1)
        {
            if (chEnd == 0)
            {
                advance(0);
            }
            doTile(eA, xA, tA, gA, 0 - chStart < nspec);
            xA = ldg(a.xq, eB.x * kClusterSize + jj);
            tA = ldg(a.type, eB.x * kClusterSize + jj);
            gA = ldg(a.gq, eB.x * kClusterSize + jj);
            eA = eB;
            eB = entryAt(2);
            t0 = 1;
        }
        for (int t = t0; t < T; t += 2)
        {
            const float4 xB  = ldg(a.xq, eB.x * kClusterSize + jj);
            const int    tB  = ldg(a.type, eB.x * kClusterSize + jj);
            const float4 gB  = ldg(a.gq, eB.x * kClusterSize + jj);
            const uint2  eA2 = entryAt(t + 2);
            if (t == chEnd)
            {
                advance(t);
            }
            doTile(eA, xA, tA, gA, t - chStart < nspec);
            const uint2 eB2 = entryAt(t + 3);
            xA              = ldg(a.xq, eA2.x * kClusterSize + jj);
            tA              = ldg(a.type, eA2.x * kClusterSize + jj);
            gA              = ldg(a.gq, eA2.x * kClusterSize + jj);
            if (t + 1 == chEnd)
            {
                advance(t + 1);
            }
            doTile(eB, xB, tB, gB, t + 1 - chStart < nspec);
            eA = eA2;
            eB = eB2;
        }
        endChunk();
        endGroup();
    }

    // the workgroup's C6/C12 gradient sums
    if (kParams)
    {
        __syncthreads();
        for (int t = threadIdx.x; t < accStride; t += kAdjThreads)
        {
            double v = sAcc[t];
            for (int k = 1; k < accCopies; k++)
            {
                v += sAcc[k * accStride + t];
            }
            a.paramPartial[static_cast<size_t>(blockIdx.x) * accStride + t] = v;
        }
    }
}

// Sum of the workgroup partials in a fixed order (deterministic): one
// wavefront per (type pair, C6 or C12) component, lane l taking workgroups
// l, l + 64, ..., then a butterfly over the lanes.  The padding type's row
// and column are dropped: out6/out12 are numTypes x numTypes.
constexpr int kParamThreads = 256; // the finishing launch's workgroup size

__device__ void adjointParamBlock(int k, int numBlocks, int ntype, int numTypes, const double* __restrict__ partial,
                                  double* __restrict__ out6, double* __restrict__ out12)
{
    const int p  = k >> 1;
    const int ti = p / ntype;
    const int tj = p - ti * ntype;
    if (ti >= numTypes || tj >= numTypes)
    {
        return;
    }
    // four independent sums per thread (loads in flight together), then the
    // wavefronts' sums in order
    const int    n  = 2 * ntype * ntype;
    constexpr int T = kParamThreads;
    double       s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    int          b  = threadIdx.x;
    for (; b + 3 * T < numBlocks; b += 4 * T)
    {
        s0 += partial[static_cast<size_t>(b) * n + k];
        s1 += partial[static_cast<size_t>(b + T) * n + k];
        s2 += partial[static_cast<size_t>(b + 2 * T) * n + k];
        s3 += partial[static_cast<size_t>(b + 3 * T) * n + k];
    }
    if (b < numBlocks)
    {
        s0 += partial[static_cast<size_t>(b) * n + k];
    }
    if (b + T < numBlocks)
    {
        s1 += partial[static_cast<size_t>(b + T) * n + k];
    }
    if (b + 2 * T < numBlocks)
    {
        s2 += partial[static_cast<size_t>(b + 2 * T) * n + k];
    }
    double s = waveSumD((s0 + s1) + (s2 + s3));
    __shared__ double sWave[T / 64];
    if ((threadIdx.x & 63) == 0)
    {
        sWave[threadIdx.x >> 6] = s;
    }
    __syncthreads();
    double* out = (k & 1) ? out12 : out6;
    if (threadIdx.x == 0 && out)
    {
        double t = 0.0;
#pragma unroll
        for (int w = 0; w < T / 64; w++)
        {
            t += sWave[w];
        }
        out[ti * numTypes + tj] = t;
    }
}

// The staged adjoint windows of each cluster (a wavefront per cluster, as
// the force gather): dL/dx summed in float with a fixed butterfly, dL/dq in
// double; plus the cluster-layout buffers, which hold the contributions of
// clusters that spilled out of a window, written to the atom-order outputs
// (either may be NULL).
__device__ void adjointGatherStagedBlock(int blk, int nblk, int ncluster, const int* __restrict__ gposOff,
                                         const int* __restrict__ gpos, const float4* __restrict__ staging,
                                         const float4* __restrict__ gradX, const double* __restrict__ gradQ,
                                         const int* __restrict__ atomIndices, float* __restrict__ outX,
                                         double* __restrict__ outQ)
{
    const int lane = threadIdx.x & 63;
    // the workgroups of one XCD (dealt round robin; blk counts from a
    // multiple of kGatherXcds and nblk is one) take consecutive clusters, so
    // a cache line that neighbouring clusters' staged slots share is
    // fetched into one L2
    const int lb = (blk % kGatherXcds) * (nblk / kGatherXcds) + blk / kGatherXcds;
    const int c  = lb * 4 + (threadIdx.x >> 6);
    if (c >= ncluster)
    {
        return;
    }
    float  f[16];
    double q[4] = { 0.0, 0.0, 0.0, 0.0 };
#pragma unroll
    for (int k = 0; k < 16; k++)
    {
        f[k] = 0.0f;
    }
    for (int i = gposOff[c] + lane; i < gposOff[c + 1]; i += kWave)
    {
        const int p = gpos[i];
#pragma unroll
        for (int at = 0; at < 4; at++)
        {
            const float4 v = staging[p + at];
            f[4 * at + 0] += v.x;
            f[4 * at + 1] += v.y;
            f[4 * at + 2] += v.z;
            q[at] += static_cast<double>(v.w);
        }
    }
    // lane L ends up with component L >> 2 = 4 * atom + {x, y, z, q}
    const float r   = waveReduceScatter16(f, lane);
    const int   idx = lane >> 2;
    if (outX && (lane & 3) == 0 && (idx & 3) < 3)
    {
        const int k = c * kClusterSize + (idx >> 2);
        const int a = atomIndices[k];
        if (a >= 0)
        {
            outX[3 * a + (idx & 3)] = reinterpret_cast<const float*>(gradX + k)[idx & 3] + r;
        }
    }
#pragma unroll
    for (int at = 0; at < 4; at++)
    {
        const double t = waveSumD(q[at]);
        if (outQ && lane == at)
        {
            const int k = c * kClusterSize + at;
            const int a = atomIndices[k];
            if (a >= 0)
            {
                outQ[a] = gradQ[k] + t;
            }
        }
    }
}

// A backward step's set-up in one launch: coordinates (when x is given) and
// the cotangent g into the cluster layout (fillers and the padding cluster
// after the last one get g = 0), the spill accumulators cleared.
// The backward's last launch: the C6/C12 gradient sums (one workgroup per
// type pair and parameter, a multiple of kGatherXcds with idle ones at the
// end) first, then the staged dL/dx, dL/dq gather; they do not depend on
// each other, so the short sums run beside the gather.
__global__ __launch_bounds__(256) void adjointFinishKernel(int paramBlocks, int gatherBlocks, int ncluster,
                                                           const int* __restrict__ gposOff,
                                                           const int* __restrict__ gpos,
                                                           const float4* __restrict__ staging,
                                                           const float4* __restrict__ gradX,
                                                           const double* __restrict__ gradQ,
                                                           const int* __restrict__ atomIndices,
                                                           float* __restrict__ outX, double* __restrict__ outQ,
                                                           int numBlocks, int ntype, int numTypes,
                                                           const double* __restrict__ partial,
                                                           double* __restrict__ out6, double* __restrict__ out12)
{
    const int b = blockIdx.x;
    if (b < paramBlocks)
    {
        if (b < 2 * ntype * ntype)
        {
            adjointParamBlock(b, numBlocks, ntype, numTypes, partial, out6, out12);
        }
    }
    else
    {
        adjointGatherStagedBlock(b - paramBlocks, gatherBlocks, ncluster, gposOff, gpos, staging, gradX, gradQ,
                                 atomIndices, outX, outQ);
    }
}

__global__ void adjointPrepareKernel(int nslot, const int* __restrict__ atomIndices, const float* __restrict__ x,
                                     const float* __restrict__ g, float4* __restrict__ xq, float4* __restrict__ gq,
                                     float4* __restrict__ gradX, double* __restrict__ gradQ)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nslot + kClusterSize)
    {
        return;
    }
    const int a = s < nslot ? atomIndices[s] : -1;
    if (x && a >= 0)
    {
        float4 v = xq[s];
        v.x      = x[3 * a + 0];
        v.y      = x[3 * a + 1];
        v.z      = x[3 * a + 2];
        xq[s]    = v;
    }
    gq[s] = a >= 0 ? make_float4(g[3 * a + 0], g[3 * a + 1], g[3 * a + 2], 0.0f) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (s < nslot)
    {
        gradX[s] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        gradQ[s] = 0.0;
    }
}

// dynamic LDS beyond 64 KiB is requested explicitly (a CU has 160 KiB)
template<int kElec, bool kParams>
void launchOne(hipStream_t s, const AdjointArgs& a, int numBlocks, int lds)
{
    (void)allowDynamicLds(reinterpret_cast<const void*>(&adjointKernel<kElec, kParams>), lds);
    adjointKernel<kElec, kParams><<<numBlocks, kAdjThreads, lds, s>>>(a);
}

template<bool kParams>
void launchElec(hipStream_t s, const AdjointArgs& a, int elec, int numBlocks, int lds)
{
    if (elec == kElecRF)
    {
        launchOne<kElecRF, kParams>(s, a, numBlocks, lds);
    }
    else if (elec == kElecEwaldAna)
    {
        launchOne<kElecEwaldAna, kParams>(s, a, numBlocks, lds);
    }
    else
    {
        launchOne<kElecEwaldTable, kParams>(s, a, numBlocks, lds);
    }
}

} // namespace

int adjointLdsBytes(int ntype, int tabN, int elec, int windowSlots)
{
    const AdjLayout L(ntype, elec == kElecEwaldTable ? tabN : 0, windowSlots);
    return L.end_ * 16;
}

int adjointBlocks(int forwardBlocks)
{
    return forwardBlocks * kKernelWaves / kAdjWaves;
}

void launchAdjoint(hipStream_t s, const AdjointArgs& a, int elec, int forwardBlocks, const int* atomIndices,
                   float* gradXOut, double* gradQOut, int numTypes, double* gradC6Out, double* gradC12Out)
{
    const int numBlocks = adjointBlocks(forwardBlocks);
    const int lds = adjointLdsBytes(a.ntype, a.ic.tab_n, elec, a.windowSlots);
    const bool params = gradC6Out || gradC12Out;
    if (params)
    {
        launchElec<true>(s, a, elec, numBlocks, lds);
    }
    else
    {
        launchElec<false>(s, a, elec, numBlocks, lds);
    }
    // both parts a multiple of kGatherXcds workgroups, so that the gather's
    // workgroups keep their XCD order
    auto roundXcd          = [](int n) { return (n + kGatherXcds - 1) / kGatherXcds * kGatherXcds; };
    const int paramBlocks  = params ? roundXcd(2 * a.ntype * a.ntype) : 0;
    const int gatherBlocks = a.ncluster > 0 && (gradXOut || gradQOut) ? roundXcd((a.ncluster + 3) / 4) : 0;
    if (paramBlocks + gatherBlocks > 0)
    {
        adjointFinishKernel<<<paramBlocks + gatherBlocks, 256, 0, s>>>(
                paramBlocks, gatherBlocks, a.ncluster, a.gposOff, a.gpos, a.staging, a.gradX, a.gradQ, atomIndices,
                gradXOut, gradQOut, numBlocks, a.ntype, numTypes, a.paramPartial, gradC6Out, gradC12Out);
    }
}

void launchAdjointPrepare(hipStream_t s, int nslot, const int* atomIndices, const float* x, const float* g, float4* xq,
                          float4* gq, float4* gradX, double* gradQ)
{
    adjointPrepareKernel<<<(nslot + kClusterSize + 255) / 256, 256, 0, s>>>(nslot, atomIndices, x, g, xq, gq, gradX,
                                                                           gradQ);
}

} // namespace nbx
