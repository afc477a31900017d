// Cluster-pair LJ + Coulomb (Ewald real space or reaction-field) kernel for
// CDNA4 (gfx950, wave64).
//
// Physics and masking follow the reference's plain-C kernel
// (src/gromacs/nbnxm/kernels_reference/kernel_ref_outer.h:85-409 and
// kernel_ref_inner.h:44-375, the "LJ + Coulomb with exclusion check" loop):
//   skipmask = 0 beyond rcoulomb and for j <= i on the central diagonal;
//   rsq is floored at 3.82e-7 before 1/sqrt; LJ uses nbfp = {6*C6, 12*C12}
//   with potential-shift energies; excluded pairs inside the cut-off keep the
//   Ewald (or RF) correction; the self term -1/2 q_i^2 epsfac V(0) is added
//   once per central i-entry; shift forces are the sum of i-forces per shift.
// The Ewald correction is either the reference's quadratic-spline table
// (kElecEwaldTable, what the plain-C kernel uses) or an analytical form
// (kElecEwaldAna, as the reference's SIMD kernels use: the force and
// potential corrections of simd_math.h:1609-1720) evaluated with this
// project's own rational fits of the closed forms (fit_ewald_correction.py).
//
// Mapping to the hardware.
//   * The i-clusters are cut into groups of G consecutive clusters
//     (z-neighbours in one grid column, which share most j-clusters).  The
//     list's i-entries (chunks), stored in group order, are cut into one
//     contiguous run of about equal cost per wavefront when the list is
//     built (a static schedule; a group may be split between two
//     wavefronts).  A wavefront works alone on one group at a time with a
//     private LDS window holding one float3 per atom of each distinct
//     j-cluster of the group and of its own G clusters: j-forces are added to
//     it with plain read-modify-write (no atomics, no barriers).  When the
//     wavefront moves to its next group the window is written with plain
//     coalesced stores to its own staging segment, and a gather kernel
//     (buffers.hip) adds each cluster's staged copies in a fixed order after
//     the launch: no float atomics, and bitwise reproducible forces.
//   * In a tile the four i-atoms are wave-uniform (scalar registers) and lane
//     l owns j-atom l & 3 of j-entry l >> 2: 16 j-clusters per tile, four
//     i-atoms per lane, so one VALU instruction evaluates 64 pairs.  The
//     periodic shift is applied to the j-atom (x_i + s - x_j = x_i - (x_j -
//     s)), so the i-data never leave the scalar unit.  A lane keeps its
//     j-force for the tile in registers; i-forces stay in 12 registers for an
//     i-entry and are reduce-scattered across the wave once per entry.
//   * A wavefront's groups are one contiguous run of the work list (each
//     i-entry padded to 16 entries with a far-away dummy cluster), walked by
//     one software pipeline across i-entry and group boundaries: work entries
//     are loaded two tiles ahead and j-atom data one tile ahead, in a loop
//     unrolled by two so that no loaded value is copied between registers.
//   * Masked entries come first in every i-entry: its leading tiles run with
//     exclusion and diagonal checks, the rest without.
//   * LJ parameters, the Ewald table (table flavour) and each group's
//     i-atoms are staged in LDS.  A group's first cluster, size and shift
//     table come from its groupInfo record, so one launch can cover groups of
//     several independent systems (the batched ensemble).
//   * Energies: float per lane over the four pairs of a tile, then double;
//     the sum of each (wavefront, group) goes to its own slot and the
//     finalize kernel sums the slots in order, so energies are bitwise
//     reproducible.  The self energy
//     and the energies of excluded pairs inside the cut-off (large terms that
//     nearly cancel) are left to a separate double-precision pass over the
//     exclusion lists (buffers.hip); here excluded pairs contribute forces
//     only.
// Pair arithmetic is FP32 on the vector ALUs; no matrix instructions are used
// (the pair interaction is not a dense contraction).
#include <hip/hip_ext.h>

#include <map>
#include <mutex>
#include <utility>

#include "../../include/nbx.h"
#include "exclusion.h"
#include "kernel_common.h"

namespace nbx
{

namespace
{

constexpr int kWavesBlock = kKernelWaves;
constexpr int kThreads    = kWave * kWavesBlock;

// One j-atom (this lane) against i-atom kI.  kChecks adds the exclusion
// bit and the central diagonal.  vlj accumulates 12 times the LJ energy
// (FrLJ12 - 2 FrLJ6 plus the shift: one rounding fewer than dividing each
// pair by 12 and 6, and no bias from a rounded 1/12); the wavefront's double
// sum is divided by 12 once per group.
template<int kElec, bool kEnergy, bool kChecks, int kLJ, int kI>
__device__ __forceinline__ void pairI(const IAtoms& I, const JAtom& J, const unsigned mbits, const bool diag,
                                      const int jj, const InteractionConst& ic, const float2* __restrict__ sNbfp,
                                      const float* __restrict__ sC6g, const float4* __restrict__ sTab, float* fi, float& fjx, float& fjy,
                                      float& fjz, float& vlj, float& vc)
{
    constexpr int  kFl   = kLJ & kLJFlavourMask;
    constexpr bool kTwin = (kLJ & kLJTwin) != 0;
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
    rsq        = fmaxf(rsq, kMinRsq);
    // the hardware estimate as it is: within one ulp of 1/sqrt(rsq) and
    // correctly rounded for about 89 % of inputs, with a mean relative error
    // of -7e-9 over rsq in [0.01, 2] nm^2 (measured on gfx950), which a float
    // Newton-Raphson step does not reduce (its own roundings give -1.3e-8)
    float rinv = __builtin_amdgcn_rsqf(rsq);
    // r for the table position
    const float r = rsq * rinv;
    rinv               = skip ? 0.0f : rinv;
    const float rinvsq = rinv * rinv;
    // energy mask: 0 beyond the cut-off, on the diagonal and (with checks)
    // for excluded pairs, whose energies the exclusion pass adds
    const float emask = skip ? 0.0f : interact;
    // twin cut-off: LJ forces and energies only within rvdw (the floored
    // r^2, as the reference compares it)
    const bool  inVdw  = !kTwin || rsq < ic.rvdw2;
    const float ljmask = kTwin ? (inVdw ? emask : 0.0f) : emask;

    // LJ, skipped (a wave-uniform branch) for an i-atom whose type has no
    // LJ with any type, where every term is exactly zero
    float frLJ = 0.0f;
    if (kFl == kLJShift || I.lj[kI])
    {
        const float2 c     = *reinterpret_cast<const float2*>(reinterpret_cast<const char*>(sNbfp) + (I.tOff[kI] + J.type));
        const float  rinv6 = kChecks ? interact * rinvsq * rinvsq * rinvsq : rinvsq * rinvsq * rinvsq;
        const float  FrLJ6 = c.x * rinv6;
        frLJ               = fmaf(c.y * rinv6, rinv6, -FrLJ6);
        if (kFl == kLJForceSwitch || kFl == kLJPotSwitch)
        {
            // the switches act on rsw = max(r - rvdw_switch, 0), with r from
            // the masked 1/r: 0 beyond the cut-off and on the diagonal (and
            // so for the far-away padding cluster); excluded pairs inside
            // the cut-off keep their r, as in the plain-C kernel
            const float rm  = rsq * rinv;
            const float rsw = fmaxf(rm - ic.rvdw_switch, 0.0f);
            const float rs2 = rsw * rsw;
            if (kFl == kLJForceSwitch)
            {
                if (kEnergy)
                {
                    // 12 V = FrLJ12 - 2 FrLJ6 + c12 cpot_rep - 2 c6 cpot_disp
                    //        + rsw^3 (c6 (4 c2d + 3 c3d rsw) - c12 (4 c2r + 3 c3r rsw)),
                    // from the unswitched frLJ = FrLJ12 - FrLJ6
                    const float vsw = rs2 * rsw
                                      * fmaf(c.x, fmaf(ic.fsw_e3d, rsw, ic.fsw_e2d),
                                             -c.y * fmaf(ic.fsw_e3r, rsw, ic.fsw_e2r));
                    const float v12 = fmaf(c.x, -2.0f * ic.cpot_disp, fmaf(c.y, ic.cpot_rep, frLJ - FrLJ6 + vsw));
                    vlj             = fmaf(v12, ljmask, vlj);
                }
                frLJ = fmaf(rs2 * rm,
                            fmaf(c.y, fmaf(ic.fsw_c3r, rsw, ic.fsw_c2r), -c.x * fmaf(ic.fsw_c3d, rsw, ic.fsw_c2d)),
                            frLJ);
            }
            else
            {
                // V (masked, no shift) scaled by sw; F = F sw - r V dsw
                const float v12 = (frLJ - FrLJ6) * (kChecks ? interact : 1.0f);
                const float sw  = fmaf(rs2 * rsw, fmaf(fmaf(ic.sw_c5, rsw, ic.sw_c4), rsw, ic.sw_c3), 1.0f);
                const float dsw = rs2 * fmaf(fmaf(ic.sw_f5, rsw, ic.sw_f4), rsw, ic.sw_f3);
                frLJ            = fmaf(frLJ, sw, -rm * (v12 * (1.0f / 12.0f)) * dsw);
                if (kEnergy)
                {
                    vlj = fmaf(v12 * sw, ljmask, vlj);
                }
            }
        }
        else if (kEnergy)
        {
            // 12 V = FrLJ12 - 2 FrLJ6 + c12 cpot_rep - 2 c6 cpot_disp
            const float v12 = fmaf(c.x, -2.0f * ic.cpot_disp, fmaf(c.y, ic.cpot_rep, frLJ - FrLJ6));
            vlj             = fmaf(v12, ljmask, vlj);
        }
        if (kTwin)
        {
            frLJ = inVdw ? frLJ : 0.0f;
        }
        if (kFl == kLJEwald)
        {
            // LJ-PME: subtract the grid's dispersion, c6g (r^-6 - e^-x (r^-6
            // (1 + x + x^2/2) + beta^6/6)) with x = beta^2 r^2, from r^-6
            // without the exclusion mask (excluded pairs inside the cut-off
            // keep the grid correction, as Ewald's Coulomb ones do); beyond
            // the cut-off r^-6 = 0 and the force's 1/r^2 = 0
            const float c6g     = sC6g[(I.tOff[kI] + J.type) >> 3];
            const float rinv6nm = rinvsq * rinvsq * rinvsq;
            const float cr2     = ic.lje_coeff2 * rsq;
            const float expm    = __builtin_amdgcn_exp2f(-cr2 * kLog2e);
            const float poly    = fmaf(0.5f * cr2, cr2, 1.0f + cr2);
            frLJ                = fmaf(c6g, rinv6nm - expm * fmaf(rinv6nm, poly, ic.lje_coeff6_6), frLJ);
            if (kEnergy)
            {
                // 12 (c6g / 6) (r^-6 (1 - e^-x poly) + sh_lj_ewald interact),
                // masked by the cut-off and the diagonal only
                const float vg = 2.0f * c6g * fmaf(rinv6nm, fmaf(-expm, poly, 1.0f), ic.sh_lj_ewald * interact);
                vlj            = skip ? vlj : vlj + vg;
            }
        }
    }

    float fcoul;
    if (kElec == kElecRF)
    {
        // the reaction-field constant needs qq = 0 beyond the cut-off
        const float qq = (skip ? 0.0f : I.q[kI]) * J.w;
        fcoul          = qq * (interact * rinv * rinvsq - 2.0f * ic.k_rf);
        if (kEnergy)
        {
            vc = fmaf(interact * qq, rinv + ic.k_rf * rsq - ic.c_rf, vc);
        }
    }
    else if (kElec == kElecEwaldTable)
    {
        // the table sits at entries 1..tab_n of the LDS copy, after an entry
        // of zeros that pairs beyond the cut-off and on the diagonal (skip)
        // read: with rinv = 0 they add neither force nor energy and need no
        // mask; every other pair has r < rc (up to the rounding of rsq *
        // rinv, one entry at most), so no input (NaN, Inf, the far-away
        // padding cluster) can read past the table
        const float  rs    = skip ? 0.0f : fmaf(r, ic.tab_scale, 1.0f);
        const int    ri    = static_cast<int>(rs);
        const float  frac  = __builtin_amdgcn_fractf(rs);
        const float4 tab   = sTab[ri];
        const float  fexcl = fmaf(frac, tab.y, tab.x);
        const float  qq    = I.q[kI] * J.w;
        fcoul              = (kChecks ? fmaf(interact, rinvsq, -fexcl) : rinvsq - fexcl) * (qq * rinv);
        if (kEnergy)
        {
            // rinv - sh_ewald - V_tab, V_tab = V - (h/2) frac (F + fexcl);
            // tab.w holds V + sh_ewald; excluded pairs' energies come from
            // the exclusion pass
            const float vpair = fmaf((0.5f / ic.tab_scale) * frac, tab.x + fexcl, rinv - tab.w);
            vc                = fmaf(kChecks ? qq * interact : qq, vpair, vc);
        }
    }
    else
    {
        // qq = 0 beyond the cut-off; the correction is clamped to the fitted
        // range: pairs beyond the cut-off may be far away (the padding
        // cluster), where s^5 would overflow to inf and inf / inf is NaN
        const float qq = (skip ? 0.0f : I.q[kI]) * J.w;
        const float s  = fminf(rsq * ic.beta2, kAnaMaxS);
        const float pg = fmaf(fmaf(fmaf(fmaf(fmaf(kGp5, s, kGp4), s, kGp3), s, kGp2), s, kGp1), s, kGp0);
        const float qg = fmaf(fmaf(fmaf(fmaf(fmaf(kGq5, s, kGq4), s, kGq3), s, kGq2), s, kGq1), s, 1.0f);
        fcoul          = qq * (interact * rinv * rinvsq - ic.beta3 * pg * __builtin_amdgcn_rcpf(qg));
        if (kEnergy)
        {
            // erfc(beta r)/r directly: 1/r - erf(beta r)/r would cancel;
            // z = beta r from rsq / sqrt(rsq) (0 beyond the cut-off, where
            // qq = 0)
            const float z  = ic.beta * (rsq * rinv);
            const float pe = fmaf(fmaf(fmaf(fmaf(kEp4, z, kEp3), z, kEp2), z, kEp1), z, kEp0);
            const float qe = fmaf(fmaf(fmaf(fmaf(kEq4, z, kEq3), z, kEq2), z, kEq1), z, 1.0f);
            const float vpair =
                    rinv * __builtin_amdgcn_exp2f(-s * kLog2e) * pe * __builtin_amdgcn_rcpf(qe) - ic.sh_ewald;
            vc = fmaf(kChecks ? interact * qq : qq, vpair, vc);
        }
    }

    const float fscal = fmaf(frLJ, rinvsq, fcoul);
    fi[3 * kI + 0]    = fmaf(fscal, dx, fi[3 * kI + 0]);
    fi[3 * kI + 1]    = fmaf(fscal, dy, fi[3 * kI + 1]);
    fi[3 * kI + 2]    = fmaf(fscal, dz, fi[3 * kI + 2]);
    fjx               = fmaf(-fscal, dx, fjx);
    fjy               = fmaf(-fscal, dy, fjy);
    fjz               = fmaf(-fscal, dz, fjz);
}

template<int kElec, bool kEnergy, bool kChecks, int kLJ>
__device__ __forceinline__ void tile4(const IAtoms& I, const JAtom& J, const unsigned mbits, const bool diag,
                                      const int jj, const InteractionConst& ic, const float2* __restrict__ sNbfp,
                                      const float* __restrict__ sC6g, const float4* __restrict__ sTab, float* fi, float& fjx, float& fjy,
                                      float& fjz, float& vlj, float& vc)
{
    // barriers between the i-atoms keep the scheduler from interleaving all
    // four pair evaluations, which costs registers (occupancy) for little overlap
    pairI<kElec, kEnergy, kChecks, kLJ, 0>(I, J, mbits, diag, jj, ic, sNbfp, sC6g, sTab, fi, fjx, fjy, fjz, vlj, vc);
    __builtin_amdgcn_sched_barrier(0);
    pairI<kElec, kEnergy, kChecks, kLJ, 1>(I, J, mbits, diag, jj, ic, sNbfp, sC6g, sTab, fi, fjx, fjy, fjz, vlj, vc);
    __builtin_amdgcn_sched_barrier(0);
    pairI<kElec, kEnergy, kChecks, kLJ, 2>(I, J, mbits, diag, jj, ic, sNbfp, sC6g, sTab, fi, fjx, fjy, fjz, vlj, vc);
    __builtin_amdgcn_sched_barrier(0);
    pairI<kElec, kEnergy, kChecks, kLJ, 3>(I, J, mbits, diag, jj, ic, sNbfp, sC6g, sTab, fi, fjx, fjy, fjz, vlj, vc);
}

// LDS layout of a workgroup, in float4 units
struct LdsLayout
{
    int nbfp, grid, tab, fsh, ix, it, win;
    __host__ __device__ LdsLayout(int ntype, int tabN, int windowSlots, bool ljPme)
    {
        nbfp  = 0;
        grid  = nbfp + (ntype * ntype + 1) / 2;
        tab   = grid + (ljPme ? (ntype * ntype + 3) / 4 : 0);
        fsh   = tab + tabN;
        ix    = fsh + (kWavesBlock * kNumShifts * 3 + 3) / 4;
        it    = ix + kWavesBlock * kGroupSizeMax * kClusterSize;
        win   = it + kWavesBlock * kGroupSizeMax;
        end_  = win + kWavesBlock * windowSlots * 3;
    }
    int end_;
};

template<int kElec, bool kEnergy, bool kShift, int kLJ, bool kFixed>
__global__ __launch_bounds__(kThreads) void nonbondedKernel(const KernelArgs a)
{
    extern __shared__ float4 smem4[];
    if (kEnergy && a.exclBlocks > 0 && static_cast<int>(blockIdx.x) >= static_cast<int>(gridDim.x) - a.exclBlocks)
    {
        // the single-system exclusion pass: these workgroups are dispatched
        // after the pair workgroups and run where those have finished
        const int first = static_cast<int>(gridDim.x) - a.exclBlocks;
        exclusionPairBlock(a.excl, static_cast<int>(blockIdx.x) - first, a.exclBlocks, kThreads,
                           reinterpret_cast<double*>(smem4));
        return;
    }
    const int       ntype  = a.ntype;
    const int       ntype2 = ntype * ntype;
    // a zero entry and the table
    const int       tabN   = (kElec == kElecEwaldTable) ? a.ic.tab_n + 1 : 0;
    const LdsLayout L(ntype, tabN, a.windowSlots, kLJ == kLJEwald);
    float2*         sNbfp  = reinterpret_cast<float2*>(smem4 + L.nbfp);
    float*          sC6g   = reinterpret_cast<float*>(smem4 + L.grid);
    float4*         sTab   = smem4 + L.tab;
    float*          sFsh   = reinterpret_cast<float*>(smem4 + L.fsh);

    for (int t = threadIdx.x; t < ntype2; t += kThreads)
    {
        sNbfp[t] = a.nbfp[t];
    }
    if (kLJ == kLJEwald)
    {
        for (int t = threadIdx.x; t < ntype2; t += kThreads)
        {
            sC6g[t] = a.c6grid[t];
        }
    }
    for (int t = threadIdx.x; t < tabN; t += kThreads)
    {
        sTab[t] = t > 0 ? a.tabFDV0[t - 1] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    // one shift-force array per wavefront: each is summed in the wavefront's
    // own (fixed) order, and the four are added in wavefront order at the end
    for (int t = threadIdx.x; t < kWavesBlock * kNumShifts * 3; t += kThreads)
    {
        sFsh[t] = 0.0f;
    }
    __syncthreads();

    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    const int jj   = lane & 3;
    const int kk   = lane >> 2;

    const InteractionConst ic = a.ic;
    // this wavefront's i-atom cache and j-force window
    float4* ix  = smem4 + L.ix + wave * kGroupSizeMax * kClusterSize;
    int*    it  = reinterpret_cast<int*>(smem4 + L.it) + wave * kGroupSizeMax * kClusterSize;
    float*  win = reinterpret_cast<float*>(smem4 + L.win) + wave * a.windowSlots * 12;

    // this wavefront's contiguous run of chunks (balanced by cost when the
    // list was built); chunks and work entries are stored in group order, so
    // one software pipeline runs over the whole run, across i-entry and group
    // boundaries
    const int  wid    = blockIdx.x * kWavesBlock + wave;
    const int2 ws     = a.waveStart[wid];
    const int  chAll0 = uniformInt(ws.x);
    const int  chAll1 = uniformInt(a.waveStart[wid + 1].x);
    if (chAll0 < chAll1)
    {
        const int4     cFirst = a.chunks[chAll0];
        const int4     cLast  = a.chunks[chAll1 - 1];
        const unsigned W0     = static_cast<unsigned>(uniformInt(cFirst.z));
        const int      T      = (uniformInt(cLast.z) - static_cast<int>(W0)) / kTileJ + uniformInt(cLast.w);

        // current group (the run starts inside the group holding its first
        // chunk); wavefront w writes the energies of group g to slot g + w
        int g = uniformInt(ws.y), chGrpEnd = 0, U = 0, c0 = 0, nOwn = 0, G = 0, stage = -1;
        // current chunk
        IAtoms I;
        float4 sv;
        float  fi[12];
        // the current i-entry's energies per lane, in float (12 times the LJ
        // energy), added to the wavefront's double sums at the entry's end
        float cVdw = 0.0f, cCoul = 0.0f;
        int    ci = 0, sh = 0, nspec = 0, chStart = 0, chEnd = 0, chIdx = chAll0;
        bool   central = false;
        double eVdw = 0.0, eCoul = 0.0;

        auto beginGroup = [&](int gg) {
            const int4 grp = a.groups[gg]; // chunk range, union offset, windowed union size
            const int4 gi  = a.groupInfo[gg]; // first own cluster, own count | G << 16
            const int4 si  = a.segInfo[gg];   // staging: first float4, first wavefront, wavefronts, size
            chGrpEnd       = uniformInt(grp.y);
            U              = uniformInt(grp.w);
            c0             = uniformInt(gi.x);
            nOwn           = uniformInt(gi.y) & 0xffff;
            G              = uniformInt(gi.y) >> 16;
            // this wavefront's staging segment (none for a group without chunks)
            stage = uniformInt(si.z) > 0 ? uniformInt(si.x) + (wid - uniformInt(si.y)) * uniformInt(si.w) : -1;
            // clear the window (own slots, union slots, the padding entries'
            // spare slot) and cache the group's i-atoms; this wavefront's LDS
            // accesses complete in order, so no barrier
            for (int t = lane; t < (G + U + 1) * 3; t += kWave)
            {
                reinterpret_cast<float4*>(win)[t] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            }
            if (lane < nOwn * kClusterSize)
            {
                // type row offset, with bit 30 set for types that have LJ
                const int ty = a.type[c0 * kClusterSize + lane];
                ix[lane]     = a.xq[c0 * kClusterSize + lane];
                it[lane]     = ty * ntype | (a.typeLJ[ty] << 30);
            }
        };
        auto endGroup = [&]() {
            if (kEnergy && !a.waveEnergy)
            {
                const double v = waveSumD(eVdw) * (1.0 / 12.0);
                const double c = waveSumD(eCoul);
                if (lane == 0)
                {
                    a.energyPartial[2 * (g + wid)]     = v;
                    a.energyPartial[2 * (g + wid) + 1] = c;
                }
                eVdw  = 0.0;
                eCoul = 0.0;
            }
            // stage the window (own and union slots, 3 float4 each) with
            // plain coalesced stores; the gather kernel adds the staged
            // windows of every cluster after the launch
            if (stage >= 0)
            {
                const float4* w4 = reinterpret_cast<const float4*>(win);
                for (int t = lane; t < (G + U) * 3; t += kWave)
                {
                    a.staging[stage + t] = w4[t];
                }
            }
        };
        auto beginChunk = [&](const int4 cd, const float4 s, int t) {
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
                const int ti   = uniformInt(it[lc + i]);
                // byte offset of the type's row in the LJ table
                I.tOff[i]      = (ti & 0x3fffffff) * static_cast<int>(sizeof(float2));
                I.lj[i]        = ti >> 30;
            }
            sv = make_float4(uniformFloat(s.x), uniformFloat(s.y), uniformFloat(s.z), 0.0f);
#pragma unroll
            for (int k = 0; k < 12; k++)
            {
                fi[k] = 0.0f;
            }
            cVdw  = 0.0f;
            cCoul = 0.0f;
        };
        auto endChunk = [&]() {
            if (kEnergy)
            {
                eVdw += static_cast<double>(cVdw);
                eCoul += static_cast<double>(cCoul);
            }
            // i-forces: lane L ends up with component (L >> 2) = 3*i + d
            const float r   = waveReduceScatter12(fi, lane);
            const int   idx = lane >> 2;
            if ((lane & 3) == 0 && idx < 12)
            {
                win[(ci - c0) * 12 + idx] += r;
            }
            if (kShift)
            {
                // sum over the four i-atoms of component d: values d, 3+d, 6+d, 9+d
                const int   d  = lane < 3 ? lane : 0;
                const float s0 = __shfl(r, 4 * d, kWave);
                const float s1 = __shfl(r, 4 * (3 + d), kWave);
                const float s2 = __shfl(r, 4 * (6 + d), kWave);
                const float s3 = __shfl(r, 4 * (9 + d), kWave);
                if (lane < 3)
                {
                    sFsh[(wave * kNumShifts + sh) * 3 + lane] += (s0 + s1) + (s2 + s3);
                }
            }
        };
        auto doTile = [&](const uint2 e, const float4 xj, const int tj, const bool checked) {
            // type as a byte offset into a row of the LJ table
            const JAtom J{ xj.x - sv.x, xj.y - sv.y, xj.z - sv.z, ic.epsfac * xj.w, tj * static_cast<int>(sizeof(float2)) };
            float       fjx = 0.0f, fjy = 0.0f, fjz = 0.0f;
            if (checked)
            {
                const bool diag = central && static_cast<int>(e.x) == ci;
                tile4<kElec, kEnergy, true, kLJ>(I, J, (e.y & 0xffffu) >> jj, diag, jj, ic, sNbfp, sC6g, sTab, fi, fjx, fjy,
                                            fjz, cVdw, cCoul);
            }
            else
            {
                tile4<kElec, kEnergy, false, kLJ>(I, J, 0u, false, jj, ic, sNbfp, sC6g, sTab, fi, fjx, fjy, fjz, cVdw,
                                                  cCoul);
            }
            // the 16 entries of a tile have distinct slots (padding entries
            // share the spare slot, adding zeros): plain read-modify-write;
            // clusters that did not fit the window take global atomics, in
            // float or (kFixed) on fixed-point sums, whose integer additions
            // do not depend on their order (a layout without such clusters
            // skips the test)
            const unsigned slot = e.y >> 16;
            if (!a.windowSpill || slot != kSlotOverflow)
            {
                float* w = win + slot * 12 + jj * 3;
                w[0] += fjx;
                w[1] += fjy;
                w[2] += fjz;
            }
            else if constexpr (kFixed)
            {
                unsigned long long* dst = a.fixF + 3 * (e.x * kClusterSize + jj);
                atomicAdd(dst, toFixed(fjx));
                atomicAdd(dst + 1, toFixed(fjy));
                atomicAdd(dst + 2, toFixed(fjz));
            }
            else
            {
                float* dst = reinterpret_cast<float*>(a.fnb + e.x * kClusterSize + jj);
                unsafeAtomicAdd(dst, fjx);
                unsafeAtomicAdd(dst + 1, fjy);
                unsafeAtomicAdd(dst + 2, fjz);
            }
        };
        // at a chunk boundary: the next chunk, and the next group holding
        // chunks when the chunk was its group's last (groups without
        // i-entries in between are passed through)
        // the next chunk's descriptor and shift vector are loaded one chunk ahead
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

        // reads past the last tile are clamped to it
        auto entryAt = [&](int t) {
            return ldg(a.work, W0 + static_cast<unsigned>(kTileJ * min(t, T - 1) + kk));
        };
        uint2  eA = entryAt(0);
        uint2  eB = entryAt(1);
        float4 xA = ldg(a.xq, eA.x * kClusterSize + jj);
        int    tA = ldg(a.type, eA.x * kClusterSize + jj);
        beginGroup(g);
        beginChunk(cFirst, a.chunkShift[chAll0], 0);
        // an odd tile count: the first tile alone, so that the loop below
        // always runs whole pairs of tiles (with no exit between its two
        // tiles the compiler keeps the accumulators in place: about 14 register
        // copies fewer per tile)
        int t0 = 0;
        if (T & 1)
        {
            if (chEnd == 0)
            {
                advance(0);
            }
            doTile(eA, xA, tA, 0 - chStart < nspec);
            xA = ldg(a.xq, eB.x * kClusterSize + jj);
            tA = ldg(a.type, eB.x * kClusterSize + jj);
            eA = eB;
            eB = entryAt(2);
            t0 = 1;
        }
        for (int t = t0; t < T; t += 2)
        {
            const float4 xB  = ldg(a.xq, eB.x * kClusterSize + jj);
            const int    tB  = ldg(a.type, eB.x * kClusterSize + jj);
            const uint2  eA2 = entryAt(t + 2);
            if (t == chEnd)
            {
                advance(t);
            }
            doTile(eA, xA, tA, t - chStart < nspec);
            const uint2 eB2 = entryAt(t + 3);
            xA              = ldg(a.xq, eA2.x * kClusterSize + jj);
            tA              = ldg(a.type, eA2.x * kClusterSize + jj);
            if (t + 1 == chEnd)
            {
                advance(t + 1);
            }
            doTile(eB, xB, tB, t + 1 - chStart < nspec);
            eA = eA2;
            eB = eB2;
        }
        endChunk();
        endGroup();
        if (kEnergy && a.waveEnergy)
        {
            const double v = waveSumD(eVdw) * (1.0 / 12.0);
            const double c = waveSumD(eCoul);
            if (lane == 0)
            {
                a.energyPartial[2 * wid]     = v;
                a.energyPartial[2 * wid + 1] = c;
            }
        }
    }

    if (kShift)
    {
        __syncthreads();
        for (int t = threadIdx.x; t < kNumShifts * 3; t += kThreads)
        {
            float sum = 0.0f;
#pragma unroll
            for (int w = 0; w < kWavesBlock; w++)
            {
                sum += sFsh[w * kNumShifts * 3 + t];
            }
            a.fshiftPartial[blockIdx.x * kNumShifts * 3 + t] = sum;
        }
    }
}

// dynamic LDS beyond 64 KiB is requested explicitly (a CU has 160 KiB).
// With events, the launch itself records the kernel's start and end (the
// dispatch packet's own timestamps): no marker packets in the stream, so
// timing the kernel leaves the step's timeline unchanged.
template<int kElec, bool kEnergy, bool kShift, int kLJ, bool kFixed>
hipError_t launchOne(hipStream_t s, const KernelArgs& a, int numBlocks, int lds, hipEvent_t start, hipEvent_t stop)
{
    const void* fn = reinterpret_cast<const void*>(&nonbondedKernel<kElec, kEnergy, kShift, kLJ, kFixed>);
    const hipError_t e = allowDynamicLds(fn, lds);
    if (e != hipSuccess)
    {
        return e;
    }
    // the exclusion pass's workgroups follow the pair workgroups
    numBlocks += kEnergy ? a.exclBlocks : 0;
    if (start || stop)
    {
        void* args[] = { const_cast<KernelArgs*>(&a) };
        return hipExtLaunchKernel(fn, dim3(numBlocks), dim3(kThreads), args, lds, s, start, stop, 0);
    }
    nonbondedKernel<kElec, kEnergy, kShift, kLJ, kFixed><<<numBlocks, kThreads, lds, s>>>(a);
    return hipGetLastError();
}

template<int kElec, int kLJ>
hipError_t launchFlags(hipStream_t s, const KernelArgs& a, bool energy, bool shiftForces, int numBlocks, int lds,
                       hipEvent_t start, hipEvent_t stop)
{
    // fixed-point overflow sums (reproducible mode with overflowing windows)
    // in kernels of their own, so the default kernels carry no code for them
    const bool fixed = a.windowSpill && a.fixF;
    if (energy && shiftForces)
    {
        return fixed ? launchOne<kElec, true, true, kLJ, true>(s, a, numBlocks, lds, start, stop)
                     : launchOne<kElec, true, true, kLJ, false>(s, a, numBlocks, lds, start, stop);
    }
    if (energy)
    {
        return fixed ? launchOne<kElec, true, false, kLJ, true>(s, a, numBlocks, lds, start, stop)
                     : launchOne<kElec, true, false, kLJ, false>(s, a, numBlocks, lds, start, stop);
    }
    if (shiftForces)
    {
        return fixed ? launchOne<kElec, false, true, kLJ, true>(s, a, numBlocks, lds, start, stop)
                     : launchOne<kElec, false, true, kLJ, false>(s, a, numBlocks, lds, start, stop);
    }
    return fixed ? launchOne<kElec, false, false, kLJ, true>(s, a, numBlocks, lds, start, stop)
                 : launchOne<kElec, false, false, kLJ, false>(s, a, numBlocks, lds, start, stop);
}

// the LJ flavour: the switches, or the potential shift with the variant
// that tests each i-atom's LJ flag only when some type has no LJ
template<int kElec>
hipError_t launchElec(hipStream_t s, const KernelArgs& a, bool energy, bool shiftForces, int numBlocks, int lds,
                      hipEvent_t start, hipEvent_t stop)
{
    if constexpr (kElec != kElecRF)
    {
        // the twin cut-off (Ewald with an LJ cut-off shorter than rcoulomb)
        if (a.twinCut)
        {
            if (a.vdwModifier == NBX_VDW_FORCE_SWITCH)
            {
                return launchFlags<kElec, kLJForceSwitch | kLJTwin>(s, a, energy, shiftForces, numBlocks, lds, start,
                                                                    stop);
            }
            if (a.vdwModifier == NBX_VDW_POTENTIAL_SWITCH)
            {
                return launchFlags<kElec, kLJPotSwitch | kLJTwin>(s, a, energy, shiftForces, numBlocks, lds, start,
                                                                  stop);
            }
            return launchFlags<kElec, kLJShiftSkip | kLJTwin>(s, a, energy, shiftForces, numBlocks, lds, start, stop);
        }
    }
    if (a.ljPme)
    {
        return launchFlags<kElec, kLJEwald>(s, a, energy, shiftForces, numBlocks, lds, start, stop);
    }
    if (a.vdwModifier == NBX_VDW_FORCE_SWITCH)
    {
        return launchFlags<kElec, kLJForceSwitch>(s, a, energy, shiftForces, numBlocks, lds, start, stop);
    }
    if (a.vdwModifier == NBX_VDW_POTENTIAL_SWITCH)
    {
        return launchFlags<kElec, kLJPotSwitch>(s, a, energy, shiftForces, numBlocks, lds, start, stop);
    }
    return a.skipLJ ? launchFlags<kElec, kLJShiftSkip>(s, a, energy, shiftForces, numBlocks, lds, start, stop)
                    : launchFlags<kElec, kLJShift>(s, a, energy, shiftForces, numBlocks, lds, start, stop);
}

} // namespace

hipError_t allowDynamicLds(const void* fn, int lds)
{
    if (lds <= 64 * 1024)
    {
        return hipSuccess;
    }
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess)
    {
        return e;
    }
    static std::mutex                                 lock;
    static std::map<std::pair<const void*, int>, int> allowed;
    std::lock_guard<std::mutex>                       guard(lock);
    int&                                              cur = allowed[{ fn, dev }];
    if (lds > cur)
    {
        e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
        if (e == hipSuccess)
        {
            cur = lds;
        }
    }
    return e;
}

int nonbondedLdsBytes(int ntype, int tabN, int elec, int windowSlots, bool ljPme)
{
    const LdsLayout L(ntype, elec == kElecEwaldTable ? tabN + 1 : 0, windowSlots, ljPme);
    return L.end_ * 16;
}

hipError_t launchNonbonded(hipStream_t s, const KernelArgs& a, int elec, bool energy, bool shiftForces, int numBlocks,
                           hipEvent_t start, hipEvent_t stop)
{
    if (a.ngroups == 0 || numBlocks == 0)
    {
        return hipSuccess;
    }
    const int lds = nonbondedLdsBytes(a.ntype, a.ic.tab_n, elec, a.windowSlots, a.ljPme != 0);
    if (elec == kElecRF)
    {
        return launchElec<kElecRF>(s, a, energy, shiftForces, numBlocks, lds, start, stop);
    }
    if (elec == kElecEwaldAna)
    {
        return launchElec<kElecEwaldAna>(s, a, energy, shiftForces, numBlocks, lds, start, stop);
    }
    return launchElec<kElecEwaldTable>(s, a, energy, shiftForces, numBlocks, lds, start, stop);
}

} // namespace nbx
