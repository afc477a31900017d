// Device code of the single-system exclusion pass in double, shared by the
// pair kernel (whose last workgroups run it in the launch's tail) and the
// batch's per-atom pass (buffers.hip): the Coulomb energy terms the pair
// kernel leaves out, the self energy -1/2 epsfac q_i^2 V(0) of every atom
// (kernel_ref_outer.h:286-326) and, for every excluded pair inside the
// cut-off, -epsfac q_i q_j V_corr(r) (Ewald) or epsfac q_i q_j (k_rf r^2 -
// c_rf) (reaction field) (kernel_ref_inner.h: excluded pairs keep the
// correction).
#pragma once

#include "nbx_internal.h"

namespace nbx
{
namespace
{

// V_corr(r) = erf(beta r)/r in double: the double build's table (the pair
// kernel's interpolation formula) or the closed form
__device__ double ewaldCorrectionD(double r, const ExclusionArgs& a)
{
    if (a.elec == kElecEwaldTable)
    {
        const double rs    = r * static_cast<double>(a.ic.tab_scale);
        const int    ri    = min(static_cast<int>(rs), a.ic.tab_n - 2);
        const double frac  = rs - static_cast<double>(ri);
        const double f0    = a.tabD[4 * ri];
        const double fexcl = f0 + frac * a.tabD[4 * ri + 1];
        return a.tabD[4 * ri + 2] - (0.5 / static_cast<double>(a.ic.tab_scale)) * frac * (f0 + fexcl);
    }
    return erf(static_cast<double>(a.ic.beta) * r) / r;
}

// Item t of the single-system exclusion pass, in double: the self term
// -selfCoef q^2 of cluster slot t (fillers carry q = 0) and the energy of
// excluded pair t inside the cut-off (both atoms as cluster slots, each pair
// once).  Every load depends only on t or on one list entry.
__device__ double exclusionItemEnergy(const ExclusionArgs& a, int t)
{
    double e = 0.0;
    if (t < a.nslot)
    {
        const double q = a.xq[t].w;
        e -= a.selfCoef * q * q;
    }
    if (t < a.npairs)
    {
        const int2   ps   = a.pairSlots[t];
        const float4 pi   = a.xq[ps.x];
        const float4 pj   = a.xq[ps.y];
        double       d[3] = { static_cast<double>(pi.x) - pj.x, static_cast<double>(pi.y) - pj.y,
                        static_cast<double>(pi.z) - pj.z };
        double       r2   = 0.0;
#pragma unroll
        for (int c = 0; c < 3; c++)
        {
            const double L = a.box[c];
            d[c] -= L * rint(d[c] / L);
            r2 += d[c] * d[c];
        }
        if (r2 < a.rc2)
        {
            const double qq = a.epsfacD * static_cast<double>(pi.w) * static_cast<double>(pj.w);
            if (a.elec == kElecRF)
            {
                e += qq * (static_cast<double>(a.ic.k_rf) * r2 - static_cast<double>(a.ic.c_rf));
            }
            else
            {
                e -= qq * ewaldCorrectionD(sqrt(r2), a);
            }
        }
    }
    return e;
}

// One workgroup of the pair form: items blk * nthreads + threadIdx.x with
// stride nblk * nthreads, summed per thread, over the wavefronts with
// shuffles and over the workgroup's wavefronts in order through sWave (LDS,
// nthreads / 64 doubles); partial[blk] = the workgroup's sum.
__device__ void exclusionPairBlock(const ExclusionArgs& a, int blk, int nblk, int nthreads, double* sWave)
{
    double    e = 0.0;
    const int n = max(a.nslot, a.npairs);
    for (int t = blk * nthreads + static_cast<int>(threadIdx.x); t < n; t += nblk * nthreads)
    {
        e += exclusionItemEnergy(a, t);
    }
#pragma unroll
    for (int m = 1; m < 64; m *= 2)
    {
        e += __shfl_xor(e, m, 64);
    }
    if ((threadIdx.x & 63) == 0)
    {
        sWave[threadIdx.x / 64] = e;
    }
    __syncthreads();
    if (threadIdx.x == 0)
    {
        double t = 0.0;
        for (int w = 0; w < nthreads / 64; w++)
        {
            t += sWave[w];
        }
        a.partial[blk] = t;
    }
}

} // namespace
} // namespace nbx
