/* CPU baseline only (compiled into liboracle_bench_f32.so, never the
 * checker): the cluster-pair kernel vectorised with AVX-512 the way the
 * reference's SIMD kernels are (kernels_simd_2xmm/kernel_inner.h,
 * kernel_outer.h): whole cluster pairs per vector register, exclusion masks
 * as vector masks, the analytical Ewald correction instead of the table
 * (simd_math.h pmeForceCorrection / pmePotentialCorrection semantics),
 * geometric LJ combination from per-type parameters (VdwLJCombGeom),
 * energies and shift forces, OpenMP threads over the i-entries with
 * per-thread force buffers reduced afterwards (kerneldispatch.cpp:243-356,
 * atomdata.cpp:1425-1469).  One __m512 holds the 16 pairs of a 4x4 cluster
 * pair: lane l = 4 i + j.
 *
 * The Ewald correction uses this project's rational fits (the constants of
 * the GPU kernel's analytical flavour): the force correction
 * (erf z - 2 z e^-z^2 / sqrt(pi)) / z^3 as P_g(s) / Q_g(s), s = z^2, and the
 * potential erfc(z) / r as e^-s P_e(z) / Q_e(z) / r, z = beta r; e^-s from
 * 2^n 2^f with a degree-7 Taylor polynomial of 2^f on |f| <= 1/2 (relative
 * error below 1e-8).  Dispatched at run time on hosts with AVX-512F. */

#include <immintrin.h>

#define BL_ATTR __attribute__((target("avx512f,avx512dq,fma")))

/* the fits of the GPU kernel's analytical Ewald correction */
static const float bl_gp[6] = { 0.7522528767585754f, -0.01339578628540039f, 0.016685180366039276f,
                                0.0003194923046976328f, 3.5602472053142264e-05f, -3.4331239362472843e-07f };
static const float bl_gq[6] = { 1.0f, 0.5821957588195801f, 0.15719348192214966f, 0.025581449270248413f,
                                0.002643152140080929f, 0.0001805289793992415f };
static const float bl_ep[5] = { 1.0f, 0.8535063862800598f, 0.35605478286743164f, 0.059257231652736664f,
                                1.4358666703628842e-05f };
static const float bl_eq[5] = { 1.0f, 1.9818848371505737f, 1.5923840999603271f, 0.6263630390167236f,
                                0.10554185509681702f };

/* 2^x for x <= 0 (flushes to 0 below -126) */
BL_ATTR static inline __m512 bl_exp2(__m512 x)
{
    x               = _mm512_max_ps(x, _mm512_set1_ps(-126.0f));
    const __m512 n  = _mm512_roundscale_ps(x, _MM_FROUND_TO_NEAREST_INT | _MM_FROUND_NO_EXC);
    const __m512 f  = _mm512_sub_ps(x, n);
    const double l2 = 0.69314718055994530942;
    /* c_k = ln2^k / k! */
    __m512 p = _mm512_set1_ps((float)(l2 * l2 * l2 * l2 * l2 * l2 * l2 / 5040.0));
    p        = _mm512_fmadd_ps(p, f, _mm512_set1_ps((float)(l2 * l2 * l2 * l2 * l2 * l2 / 720.0)));
    p        = _mm512_fmadd_ps(p, f, _mm512_set1_ps((float)(l2 * l2 * l2 * l2 * l2 / 120.0)));
    p        = _mm512_fmadd_ps(p, f, _mm512_set1_ps((float)(l2 * l2 * l2 * l2 / 24.0)));
    p        = _mm512_fmadd_ps(p, f, _mm512_set1_ps((float)(l2 * l2 * l2 / 6.0)));
    p        = _mm512_fmadd_ps(p, f, _mm512_set1_ps((float)(l2 * l2 / 2.0)));
    p        = _mm512_fmadd_ps(p, f, _mm512_set1_ps((float)l2));
    p        = _mm512_fmadd_ps(p, f, _mm512_set1_ps(1.0f));
    return _mm512_scalef_ps(p, n);
}

/* Per-type square roots of the diagonal 6 C6 and 12 C12 when the table is
 * geometric (nbfp[ij] = sqrt(nbfp[ii] nbfp[jj]) to float rounding), as
 * nblib's combination is; returns 0 otherwise. */
static int bl_geometric(const orc_sys* s, float* s6, float* s12)
{
    const int T = s->numTypes;
    for (int t = 0; t < T; t++)
    {
        s6[t]  = sqrtf(s->nbfp[(t * T + t) * 2]);
        s12[t] = sqrtf(s->nbfp[(t * T + t) * 2 + 1]);
    }
    for (int a = 0; a < T; a++)
    {
        for (int b = 0; b < T; b++)
        {
            const float c6 = s6[a] * s6[b], c12 = s12[a] * s12[b];
            if (fabsf(c6 - s->nbfp[(a * T + b) * 2]) > 1e-5f * fabsf(c6) + 1e-30f
                || fabsf(c12 - s->nbfp[(a * T + b) * 2 + 1]) > 1e-5f * fabsf(c12) + 1e-30f)
            {
                return 0;
            }
        }
    }
    return 1;
}

/* The i-entries [e0, e1) into f (cluster layout, 3 floats per slot),
 * fshift and the energies. */
BL_ATTR static void bl_entries(const orc_sys* s, const float* s6, const float* s12, int energy, int e0, int e1,
                               float* f, float* fshift, double* pVvdw, double* pVc)
{
    const float* x    = s->xnb;
    const float* q    = s->nbq;
    const int*   typ  = s->nbtype;
    const __m512 rc2  = _mm512_set1_ps(s->rcoulomb * s->rcoulomb);
    const __m512 minr = _mm512_set1_ps(ORC_MIN_RSQ);
    const float  beta = s->beta;
    const __m512 vb2  = _mm512_set1_ps(beta * beta);
    const __m512 vb3  = _mm512_set1_ps(beta * beta * beta);
    const __m512 vbeta = _mm512_set1_ps(beta);
    const __m512 maxs = _mm512_set1_ps(3.6f * 3.6f);
    const __m512 shew = _mm512_set1_ps(s->sh_ewald);
    const __m512 cpd  = _mm512_set1_ps(s->cpot_disp);
    const __m512 cpr  = _mm512_set1_ps(s->cpot_rep);
    const __m512 sixth = _mm512_set1_ps(1.0f / 6.0f), twelfth = _mm512_set1_ps(1.0f / 12.0f);
    const __m512 nlog2e = _mm512_set1_ps(-1.44269504088896341f);
    /* lane l = 4 i + j: indices of the j-atom's x, y, z in a cluster's 12
     * coordinates, and of its charge and type among the cluster's 4 */
    const __m512i jx = _mm512_setr_epi32(0, 3, 6, 9, 0, 3, 6, 9, 0, 3, 6, 9, 0, 3, 6, 9);
    const __m512i jy = _mm512_add_epi32(jx, _mm512_set1_epi32(1));
    const __m512i jz = _mm512_add_epi32(jx, _mm512_set1_epi32(2));
    const __m512i j4 = _mm512_setr_epi32(0, 1, 2, 3, 0, 1, 2, 3, 0, 1, 2, 3, 0, 1, 2, 3);
    const __m512i i4 = _mm512_setr_epi32(0, 0, 0, 0, 1, 1, 1, 1, 2, 2, 2, 2, 3, 3, 3, 3);
    /* the central diagonal keeps j > i only */
    const __mmask16 upper = 0x08ce;
    const __mmask16 m12   = 0x0fff;
    /* the self term's V(0) = 2 beta / sqrt(pi) */
    const float v0 = 2.0f * beta / 1.77245385090551602730f;
    double      vvdw = 0, vc = 0;
    for (int e = e0; e < e1; e++)
    {
        const int ci = s->ci[4 * e], ish = s->ci[4 * e + 1], c0 = s->ci[4 * e + 2], c1 = s->ci[4 * e + 3];
        const int central = (ish == ORC_CENTRAL);
        /* i-data per lane */
        const __m512 xic = _mm512_maskz_loadu_ps(m12, x + ci * 12);
        const __m512 sh  = _mm512_setr_ps(s->shiftvec[ish][0], s->shiftvec[ish][1], s->shiftvec[ish][2], 0, 0, 0,
                                         0, 0, 0, 0, 0, 0, 0, 0, 0, 0);
        const __m512i ix = _mm512_mullo_epi32(i4, _mm512_set1_epi32(3));
        const __m512 xi  = _mm512_add_ps(_mm512_permutexvar_ps(ix, xic),
                                         _mm512_permutexvar_ps(_mm512_setzero_si512(), sh));
        const __m512 yi  = _mm512_add_ps(_mm512_permutexvar_ps(_mm512_add_epi32(ix, _mm512_set1_epi32(1)), xic),
                                         _mm512_permutexvar_ps(_mm512_set1_epi32(1), sh));
        const __m512 zi  = _mm512_add_ps(_mm512_permutexvar_ps(_mm512_add_epi32(ix, _mm512_set1_epi32(2)), xic),
                                         _mm512_permutexvar_ps(_mm512_set1_epi32(2), sh));
        const __m512 qc  = _mm512_maskz_loadu_ps(0x000f, q + ci * 4);
        const __m512 qi  = _mm512_mul_ps(_mm512_set1_ps(s->epsfac), _mm512_permutexvar_ps(i4, qc));
        float        a6[4], a12[4];
        for (int i = 0; i < 4; i++)
        {
            a6[i]  = s6[typ[ci * 4 + i]];
            a12[i] = s12[typ[ci * 4 + i]];
        }
        const __m512 c6i  = _mm512_permutexvar_ps(i4, _mm512_castps128_ps512(_mm_loadu_ps(a6)));
        const __m512 c12i = _mm512_permutexvar_ps(i4, _mm512_castps128_ps512(_mm_loadu_ps(a12)));
        __m512 fix = _mm512_setzero_ps(), fiy = _mm512_setzero_ps(), fiz = _mm512_setzero_ps();
        __m512 vlj = _mm512_setzero_ps(), vcl = _mm512_setzero_ps();
        if (energy && central && c0 < c1 && s->cj[c0] == ci)
        {
            /* kernel_ref_outer.h:286-326: -1/2 epsfac q_i^2 V(0) */
            for (int i = 0; i < 4; i++)
            {
                vc -= 0.5 * (double)s->epsfac * (double)q[ci * 4 + i] * (double)q[ci * 4 + i] * (double)v0;
            }
        }
        for (int k = c0; k < c1; k++)
        {
            const int      cj   = s->cj[k];
            const __mmask16 intr = (__mmask16)(s->excl[k] & 0xffffu);
            const __m512   xjc  = _mm512_maskz_loadu_ps(m12, x + cj * 12);
            const __m512   dx   = _mm512_sub_ps(xi, _mm512_permutexvar_ps(jx, xjc));
            const __m512   dy   = _mm512_sub_ps(yi, _mm512_permutexvar_ps(jy, xjc));
            const __m512   dz   = _mm512_sub_ps(zi, _mm512_permutexvar_ps(jz, xjc));
            __m512         rsq  = _mm512_fmadd_ps(dz, dz, _mm512_fmadd_ps(dy, dy, _mm512_mul_ps(dx, dx)));
            __mmask16      in   = _mm512_cmp_ps_mask(rsq, rc2, _CMP_LT_OQ);
            if (central && cj == ci)
            {
                in &= upper;
            }
            if (!in)
            {
                continue;
            }
            rsq = _mm512_max_ps(rsq, minr);
            /* 1/sqrt: 14-bit estimate and one Newton-Raphson step */
            __m512 rinv = _mm512_rsqrt14_ps(rsq);
            rinv        = _mm512_mul_ps(_mm512_mul_ps(_mm512_set1_ps(0.5f), rinv),
                                        _mm512_fnmadd_ps(_mm512_mul_ps(rsq, rinv), rinv, _mm512_set1_ps(3.0f)));
            rinv              = _mm512_maskz_mov_ps(in, rinv);
            const __m512 rinvsq = _mm512_mul_ps(rinv, rinv);
            /* LJ (geometric, potential shift), masked by interact */
            const __m512 qjc  = _mm512_maskz_loadu_ps(0x000f, q + cj * 4);
            float        b6[4], b12[4];
            for (int j = 0; j < 4; j++)
            {
                b6[j]  = s6[typ[cj * 4 + j]];
                b12[j] = s12[typ[cj * 4 + j]];
            }
            const __m512 c6   = _mm512_mul_ps(c6i, _mm512_permutexvar_ps(j4, _mm512_castps128_ps512(_mm_loadu_ps(b6))));
            const __m512 c12  = _mm512_mul_ps(c12i, _mm512_permutexvar_ps(j4, _mm512_castps128_ps512(_mm_loadu_ps(b12))));
            const __m512 r6   = _mm512_maskz_mov_ps(intr, _mm512_mul_ps(_mm512_mul_ps(rinvsq, rinvsq), rinvsq));
            const __m512 fr6  = _mm512_mul_ps(c6, r6);
            const __m512 fr12 = _mm512_mul_ps(_mm512_mul_ps(c12, r6), r6);
            const __m512 frlj = _mm512_sub_ps(fr12, fr6);
            /* Coulomb with the analytical Ewald correction */
            const __m512 qq = _mm512_mul_ps(qi, _mm512_permutexvar_ps(j4, qjc));
            const __m512 sv = _mm512_min_ps(_mm512_mul_ps(rsq, vb2), maxs);
            __m512       pg = _mm512_set1_ps(bl_gp[5]), qg = _mm512_set1_ps(bl_gq[5]);
            for (int t = 4; t >= 0; t--)
            {
                pg = _mm512_fmadd_ps(pg, sv, _mm512_set1_ps(bl_gp[t]));
                qg = _mm512_fmadd_ps(qg, sv, _mm512_set1_ps(bl_gq[t]));
            }
            const __m512 fcorr = _mm512_div_ps(_mm512_mul_ps(vb3, pg), qg);
            const __m512 rinv3 = _mm512_maskz_mov_ps(intr, _mm512_mul_ps(rinv, rinvsq));
            const __m512 fcoul = _mm512_maskz_mov_ps(in, _mm512_mul_ps(qq, _mm512_sub_ps(rinv3, fcorr)));
            const __m512 fscal = _mm512_fmadd_ps(frlj, rinvsq, fcoul);
            fix               = _mm512_fmadd_ps(fscal, dx, fix);
            fiy               = _mm512_fmadd_ps(fscal, dy, fiy);
            fiz               = _mm512_fmadd_ps(fscal, dz, fiz);
            /* j-forces: sum the four i-lanes of every j (lanes j, j+4, j+8, j+12) */
            __m512 fjx = _mm512_mul_ps(fscal, dx), fjy = _mm512_mul_ps(fscal, dy), fjz = _mm512_mul_ps(fscal, dz);
            fjx        = _mm512_add_ps(fjx, _mm512_shuffle_f32x4(fjx, fjx, 0x4e));
            fjy        = _mm512_add_ps(fjy, _mm512_shuffle_f32x4(fjy, fjy, 0x4e));
            fjz        = _mm512_add_ps(fjz, _mm512_shuffle_f32x4(fjz, fjz, 0x4e));
            fjx        = _mm512_add_ps(fjx, _mm512_shuffle_f32x4(fjx, fjx, 0xb1));
            fjy        = _mm512_add_ps(fjy, _mm512_shuffle_f32x4(fjy, fjy, 0xb1));
            fjz        = _mm512_add_ps(fjz, _mm512_shuffle_f32x4(fjz, fjz, 0xb1));
            /* interleave lanes 0..3 of x, y, z into x0 y0 z0 x1 ... z3 */
            const __m512 xy  = _mm512_permutex2var_ps(fjx, _mm512_setr_epi32(0, 16, 0, 1, 17, 1, 2, 18, 2, 3, 19, 3,
                                                                              0, 0, 0, 0),
                                                      fjy);
            const __m512 xyz = _mm512_mask_permutexvar_ps(xy, 0x0924, _mm512_setr_epi32(0, 0, 0, 0, 0, 1, 0, 0, 2, 0,
                                                                                      0, 3, 0, 0, 0, 0),
                                                          fjz);
            float* fj = f + cj * 12;
            _mm512_mask_storeu_ps(fj, m12, _mm512_sub_ps(_mm512_maskz_loadu_ps(m12, fj), xyz));
            if (energy)
            {
                /* LJ: (12C12 (r^-12 + cpot_rep)) / 12 - (6C6 (r^-6 + cpot_disp)) / 6 on interacting pairs */
                const __mmask16 ie = in & intr;
                const __m512    v12 = _mm512_mul_ps(_mm512_mul_ps(c12, _mm512_fmadd_ps(r6, r6, cpr)), twelfth);
                const __m512    v6  = _mm512_mul_ps(_mm512_mul_ps(c6, _mm512_add_ps(r6, cpd)), sixth);
                vlj                 = _mm512_mask_add_ps(vlj, ie, vlj, _mm512_sub_ps(v12, v6));
                /* erfc(beta r)/r = e^-s P_e(z) / Q_e(z) / r - sh_ewald for interacting
                 * pairs; excluded ones get -erf(beta r)/r = that - 1/r + sh */
                const __m512 z  = _mm512_mul_ps(vbeta, _mm512_mul_ps(rsq, rinv));
                __m512       pe = _mm512_set1_ps(bl_ep[4]), qe = _mm512_set1_ps(bl_eq[4]);
                for (int t = 3; t >= 0; t--)
                {
                    pe = _mm512_fmadd_ps(pe, z, _mm512_set1_ps(bl_ep[t]));
                    qe = _mm512_fmadd_ps(qe, z, _mm512_set1_ps(bl_eq[t]));
                }
                const __m512 ex  = bl_exp2(_mm512_mul_ps(sv, nlog2e));
                const __m512 vef = _mm512_div_ps(_mm512_mul_ps(_mm512_mul_ps(rinv, ex), pe), qe);
                const __m512 vp  = _mm512_mask_sub_ps(_mm512_sub_ps(vef, rinv), intr, vef, shew);
                vcl              = _mm512_mask3_fmadd_ps(qq, vp, vcl, in);
            }
        }
        /* i-forces: sum the four j-lanes of every i (lanes 4i .. 4i+3) */
        float fx[16], fy[16], fz[16];
        _mm512_storeu_ps(fx, fix);
        _mm512_storeu_ps(fy, fiy);
        _mm512_storeu_ps(fz, fiz);
        float* fi = f + ci * 12;
        for (int i = 0; i < 4; i++)
        {
            const float sx = (fx[4 * i] + fx[4 * i + 1]) + (fx[4 * i + 2] + fx[4 * i + 3]);
            const float sy = (fy[4 * i] + fy[4 * i + 1]) + (fy[4 * i + 2] + fy[4 * i + 3]);
            const float sz = (fz[4 * i] + fz[4 * i + 1]) + (fz[4 * i + 2] + fz[4 * i + 3]);
            fi[3 * i] += sx;
            fi[3 * i + 1] += sy;
            fi[3 * i + 2] += sz;
            fshift[ish * 3] += sx;
            fshift[ish * 3 + 1] += sy;
            fshift[ish * 3 + 2] += sz;
        }
        if (energy)
        {
            vvdw += _mm512_reduce_add_ps(vlj);
            vc += _mm512_reduce_add_ps(vcl);
        }
    }
    *pVvdw += vvdw;
    *pVc += vc;
}

/* 1 when this host runs the AVX-512 baseline and the system suits it
 * (Ewald, potential-shifted LJ, geometric parameters). */
int orc_baseline_simd_available(const orc_sys* s)
{
    __builtin_cpu_init();
    if (!__builtin_cpu_supports("avx512f") || !__builtin_cpu_supports("avx512dq"))
    {
        return 0;
    }
    if (s->coulomb != 1 || s->vdw_mod != 0 || s->lj_pme || s->rvdw < s->rcoulomb || s->numTypes > 256)
    {
        return 0;
    }
    float s6[256], s12[256];
    return bl_geometric(s, s6, s12);
}

/* convertCoordinates, the AVX-512 kernel over OpenMP threads with
 * per-thread outputs, the reduction and add_nbat_f_to_f.  Returns 0 when
 * the AVX-512 path does not apply. */
int orc_compute_simd(orc_sys* s, const float* x, float* f, int energy, int nthreads, double* out_e)
{
    if (!orc_baseline_simd_available(s))
    {
        return 0;
    }
    float s6[256], s12[256];
    bl_geometric(s, s6, s12);
    orc_convert_coordinates(s, x);
    const int nslot = s->ncells * ORC_NC;
    float**   fth   = (float**)calloc((size_t)nthreads, sizeof(float*));
    float*    fsh   = (float*)calloc((size_t)nthreads * ORC_SHIFTS * 3, sizeof(float));
    double*   vv    = (double*)calloc((size_t)nthreads * 2, sizeof(double));
#pragma omp parallel num_threads(nthreads)
    {
        const int th = omp_get_thread_num();
        fth[th]      = (float*)calloc((size_t)nslot * 3 + 16, sizeof(float));
#pragma omp for schedule(dynamic, 32)
        for (int e = 0; e < s->nci; e++)
        {
            bl_entries(s, s6, s12, energy, e, e + 1, fth[th], fsh + th * ORC_SHIFTS * 3, vv + 2 * th, vv + 2 * th + 1);
        }
#pragma omp for schedule(static)
        for (int k = 0; k < nslot * 3; k++)
        {
            float sum = 0;
            for (int t = 0; t < nthreads; t++)
            {
                sum += fth[t][k];
            }
            s->fnb[k] = sum;
        }
    }
    for (int k = 0; k < ORC_SHIFTS * 3; k++)
    {
        s->fshift[k] = 0;
        for (int t = 0; t < nthreads; t++)
        {
            s->fshift[k] += fsh[t * ORC_SHIFTS * 3 + k];
        }
    }
    double evdw = 0, ec = 0;
    for (int t = 0; t < nthreads; t++)
    {
        evdw += vv[2 * t];
        ec += vv[2 * t + 1];
        free(fth[t]);
    }
    free(fth);
    free(fsh);
    free(vv);
    s->Vvdw = (real)evdw;
    s->Vc   = (real)ec;
    orc_add_nbat_f_to_f(s, f);
    out_e[0] = evdw;
    out_e[1] = ec;
    return 1;
}
